// phl_mpnn_api.cpp -- C-ABI of the warm-start network: parameters, graph,
// training data, forward, gradient and Adam training on the device.
#include "../../include/phl.h"
#include "phl_mesh.h"
#include "phl_mpnn_internal.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <random>
#include <string>
#include <vector>

using phl::nn::kHidden;
using phl::nn::kIn;
using phl::nn::kOut;

struct phl_mpnn {
    int device = 0;
    hipStream_t stream = nullptr;
    int L = 0, nParams = 0, step = 0;
    // flat parameter vector: W0 [8,128], b0 [128], A [L,128,128], B [L,128,128],
    // c [L,128], Wo [128,5], bo [5]
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    size_t oW0 = 0, ob0 = 0, oA = 0, oB = 0, oc = 0, oWo = 0, obo = 0;
    // graph
    int N = 0, W = 0;
    int *nbr = nullptr;
    float *invDeg = nullptr;
    // data and activations
    float *X = nullptr, *T = nullptr;
    float *z0 = nullptr, *hs = nullptr, *ms = nullptr, *zs = nullptr, *y = nullptr, *dy = nullptr;
    float *dh = nullptr, *dh2 = nullptr, *dz = nullptr, *dm = nullptr, *part = nullptr;
    double *lossPart = nullptr, *lossOut = nullptr;
    std::vector<void *> act;  // activation allocations (re-made per graph)
    // BF16 inference form of the hidden layers (phl_mpnn_set_forward_precision)
    int fwdBF16 = 0;
    __bf16 *wt16 = nullptr;  // [L][2][128][128] transposed BF16 A_l, B_l
};

namespace {

phl_status fail(phl_status s, const std::string &m)
{
    phl::set_last_error(m);
    return s;
}

#define HIPTRY(expr)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail(PHL_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));  \
    } while (0)

template <typename T>
phl_status dev_alloc(std::vector<void *> &list, T **ptr, size_t n)
{
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
    if (e != hipSuccess) return fail(PHL_ERR_ALLOC, std::string("hipMalloc: ") + hipGetErrorString(e));
    list.push_back(q);
    *ptr = (T *)q;
    return PHL_OK;
}

#define TRY(expr)                      \
    do {                               \
        phl_status s_ = (expr);        \
        if (s_ != PHL_OK) return s_;   \
    } while (0)

void free_list(std::vector<void *> &list)
{
    for (void *q : list) hipFree(q);
    list.clear();
}

phl_status alloc_activations(phl_mpnn *n)
{
    free_list(n->act);
    const size_t NH = (size_t)n->N * kHidden;
    TRY(dev_alloc(n->act, &n->X, (size_t)n->N * kIn));
    TRY(dev_alloc(n->act, &n->T, (size_t)n->N * kOut));
    TRY(dev_alloc(n->act, &n->z0, NH));
    TRY(dev_alloc(n->act, &n->hs, NH * (n->L + 1)));
    TRY(dev_alloc(n->act, &n->ms, NH * n->L));
    TRY(dev_alloc(n->act, &n->zs, NH * n->L));
    TRY(dev_alloc(n->act, &n->y, (size_t)n->N * kOut));
    TRY(dev_alloc(n->act, &n->dy, (size_t)n->N * kOut));
    TRY(dev_alloc(n->act, &n->dh, NH));
    TRY(dev_alloc(n->act, &n->dh2, NH));
    TRY(dev_alloc(n->act, &n->dz, NH));
    TRY(dev_alloc(n->act, &n->dm, NH));
    TRY(dev_alloc(n->act, &n->part, (size_t)phl::nn::grad_parts(n->N) * kHidden * kHidden));
    TRY(dev_alloc(n->act, &n->lossPart, (size_t)phl::nn::loss_parts(n->N * kOut)));
    TRY(dev_alloc(n->act, &n->lossOut, 1));
    return PHL_OK;
}

void forward(phl_mpnn *n)
{
    using namespace phl::nn;
    hipStream_t s = n->stream;
    const int N = n->N;
    const size_t NH = (size_t)N * kHidden, HH = (size_t)kHidden * kHidden;
    dense(n->X, n->p + n->oW0, kIn, nullptr, nullptr, n->p + n->ob0, N, 1, nullptr, n->z0, n->hs, false, s);
    for (int l = 0; l < n->L; ++l) {
        float *h = n->hs + NH * l, *hn = n->hs + NH * (l + 1);
        neighbour_mean(n->nbr, n->W, N, n->invDeg, h, n->ms + NH * l, s);
        dense(h, n->p + n->oA + HH * l, kHidden, n->ms + NH * l, n->p + n->oB + HH * l, n->p + n->oc + kHidden * l, N,
              2, h, n->zs + NH * l, hn, false, s);
    }
    out_layer(n->hs + NH * n->L, n->p + n->oWo, n->p + n->obo, N, n->y, s);
}

// The inference pass with the hidden layers' products in BF16: the weights
// are converted once per call; z is not kept (no backward pass follows).
void forward_bf16(phl_mpnn *n)
{
    using namespace phl::nn;
    hipStream_t s = n->stream;
    const int N = n->N;
    const size_t NH = (size_t)N * kHidden, HH = (size_t)kHidden * kHidden;
    weights_bf16(n->p + n->oA, n->p + n->oB, n->L, n->wt16, s);
    dense(n->X, n->p + n->oW0, kIn, nullptr, nullptr, n->p + n->ob0, N, 1, nullptr, n->z0, n->hs, false, s);
    for (int l = 0; l < n->L; ++l) {
        float *h = n->hs + NH * l, *hn = n->hs + NH * (l + 1);
        neighbour_mean(n->nbr, n->W, N, n->invDeg, h, n->ms + NH * l, s);
        dense_bf16(h, n->wt16 + 2 * HH * l, n->ms + NH * l, n->wt16 + 2 * HH * l + HH, n->p + n->oc + kHidden * l, N,
                   hn, s);
    }
    out_layer(n->hs + NH * n->L, n->p + n->oWo, n->p + n->obo, N, n->y, s);
}

// gradient of the loss into n->g (forward must have run; dy holds dL/dy)
void backward(phl_mpnn *n)
{
    using namespace phl::nn;
    hipStream_t s = n->stream;
    const int N = n->N;
    const size_t NH = (size_t)N * kHidden, HH = (size_t)kHidden * kHidden;
    small_grad(n->hs + NH * n->L, kHidden, n->dy, kOut, N, n->part, n->g + n->oWo, s);
    small_grad(nullptr, 1, n->dy, kOut, N, n->part, n->g + n->obo, s);
    out_layer_back(n->dy, n->p + n->oWo, N, n->dh, s);
    for (int l = n->L - 1; l >= 0; --l) {
        relu_back(n->dh, n->zs + NH * l, NH, n->dz, s);
        wgrad(n->hs + NH * l, n->dz, N, n->part, n->g + n->oA + HH * l, s);
        wgrad(n->ms + NH * l, n->dz, N, n->part, n->g + n->oB + HH * l, s);
        small_grad(nullptr, 1, n->dz, kHidden, N, n->part, n->g + n->oc + kHidden * l, s);
        // dm = dz B^T;  dh' = dh + dz A^T + (mean)^T dm
        dense(n->dz, n->p + n->oB + HH * l, kHidden, nullptr, nullptr, nullptr, N, 0, nullptr, nullptr, n->dm, true,
              s);
        dense(n->dz, n->p + n->oA + HH * l, kHidden, nullptr, nullptr, nullptr, N, 3, n->dh, nullptr, n->dh2, true, s);
        neighbour_mean_transpose_add(n->nbr, n->W, N, n->invDeg, n->dm, n->dh2, s);
        std::swap(n->dh, n->dh2);
    }
    relu_back(n->dh, n->z0, NH, n->dz, s);
    small_grad(n->X, kIn, n->dz, kHidden, N, n->part, n->g + n->oW0, s);
    small_grad(nullptr, 1, n->dz, kHidden, N, n->part, n->g + n->ob0, s);
}

}  // namespace

extern "C" {

phl_status phl_mpnn_create(int layers, unsigned seed, int device, phl_mpnn **out)
{
    if (!out || layers < 1 || layers > 32) return fail(PHL_ERR_ARG, "phl_mpnn_create: 1..32 layers");
    phl_mpnn *n = new (std::nothrow) phl_mpnn;
    if (!n) return fail(PHL_ERR_ALLOC, "out of memory");
    n->device = device;
    n->L = layers;
    const size_t HH = (size_t)kHidden * kHidden;
    n->oW0 = 0;
    n->ob0 = n->oW0 + (size_t)kIn * kHidden;
    n->oA = n->ob0 + kHidden;
    n->oB = n->oA + HH * layers;
    n->oc = n->oB + HH * layers;
    n->oWo = n->oc + (size_t)kHidden * layers;
    n->obo = n->oWo + (size_t)kHidden * kOut;
    n->nParams = (int)(n->obo + kOut);
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking);
    const size_t bytes = sizeof(float) * (size_t)n->nParams;
    if (e == hipSuccess) e = hipMalloc(&n->p, bytes);
    if (e == hipSuccess) e = hipMalloc(&n->g, bytes);
    if (e == hipSuccess) e = hipMalloc(&n->m, bytes);
    if (e == hipSuccess) e = hipMalloc(&n->v, bytes);
    if (e == hipSuccess) e = hipMalloc(&n->wt16, sizeof(__bf16) * 2 * HH * layers);
    if (e != hipSuccess) {
        phl_mpnn_destroy(n);
        return fail(PHL_ERR_HIP, std::string("phl_mpnn_create: ") + hipGetErrorString(e));
    }
    // He-style initialisation (normal, FP32), zero biases
    std::vector<float> h((size_t)n->nParams, 0.f);
    std::mt19937 rng(seed);
    std::normal_distribution<float> nd(0.f, 1.f);
    for (size_t i = 0; i < (size_t)kIn * kHidden; ++i) h[n->oW0 + i] = nd(rng) * std::sqrt(2.f / kIn);
    for (size_t i = 0; i < HH * layers; ++i) h[n->oA + i] = nd(rng) * std::sqrt(1.f / kHidden) * 0.5f;
    for (size_t i = 0; i < HH * layers; ++i) h[n->oB + i] = nd(rng) * std::sqrt(1.f / kHidden) * 0.5f;
    for (size_t i = 0; i < (size_t)kHidden * kOut; ++i) h[n->oWo + i] = nd(rng) * std::sqrt(1.f / kHidden);
    phl_status st = phl_mpnn_set_params(n, h.data());
    if (st != PHL_OK) {
        phl_mpnn_destroy(n);
        return st;
    }
    *out = n;
    return PHL_OK;
}

void phl_mpnn_destroy(phl_mpnn *n)
{
    if (!n) return;
    hipSetDevice(n->device);
    if (n->stream) hipStreamSynchronize(n->stream);
    free_list(n->act);
    for (float *q : {n->p, n->g, n->m, n->v})
        if (q) hipFree(q);
    if (n->wt16) hipFree(n->wt16);
    if (n->nbr) hipFree(n->nbr);
    if (n->invDeg) hipFree(n->invDeg);
    if (n->stream) hipStreamDestroy(n->stream);
    delete n;
}

phl_status phl_mpnn_n_params(const phl_mpnn *n, int *count)
{
    if (!n || !count) return fail(PHL_ERR_ARG, "null argument");
    *count = n->nParams;
    return PHL_OK;
}

phl_status phl_mpnn_set_params(phl_mpnn *n, const float *flat)
{
    if (!n || !flat) return fail(PHL_ERR_ARG, "null argument");
    HIPTRY(hipSetDevice(n->device));
    const size_t bytes = sizeof(float) * (size_t)n->nParams;
    HIPTRY(hipMemcpy(n->p, flat, bytes, hipMemcpyHostToDevice));
    HIPTRY(hipMemset(n->m, 0, bytes));
    HIPTRY(hipMemset(n->v, 0, bytes));
    n->step = 0;
    return PHL_OK;
}

phl_status phl_mpnn_get_params(phl_mpnn *n, float *flat)
{
    if (!n || !flat) return fail(PHL_ERR_ARG, "null argument");
    HIPTRY(hipSetDevice(n->device));
    HIPTRY(hipStreamSynchronize(n->stream));
    HIPTRY(hipMemcpy(flat, n->p, sizeof(float) * (size_t)n->nParams, hipMemcpyDeviceToHost));
    return PHL_OK;
}

phl_status phl_mpnn_set_graph(phl_mpnn *n, int nNodes, int width, const int *nbr)
{
    if (!n || nNodes <= 0 || width <= 0 || !nbr) return fail(PHL_ERR_ARG, "phl_mpnn_set_graph: bad argument");
    // [nNodes][width] host lists -> compacted [width][nNodes] device lists
    std::vector<int> dl((size_t)width * nNodes, -1);
    std::vector<float> inv(nNodes);
    for (int i = 0; i < nNodes; ++i) {
        int d = 0;
        for (int k = 0; k < width; ++k) {
            const int j = nbr[(size_t)i * width + k];
            if (j < -1 || j >= nNodes) return fail(PHL_ERR_ARG, "phl_mpnn_set_graph: neighbour out of range");
            if (j >= 0) dl[(size_t)d++ * nNodes + i] = j;
        }
        inv[i] = 1.f / (float)std::max(d, 1);
    }
    // The backward pass applies the transpose of the neighbour mean as a
    // gather over each node's own list, which is the transpose only if the
    // lists are symmetric with equal multiplicity: j appears in i's list as
    // often as i appears in j's.
    for (int i = 0; i < nNodes; ++i)
        for (int k = 0; k < width; ++k) {
            const int j = nbr[(size_t)i * width + k];
            if (j < 0) continue;
            int cij = 0, cji = 0;
            for (int t = 0; t < width; ++t) {
                cij += nbr[(size_t)i * width + t] == j;
                cji += nbr[(size_t)j * width + t] == i;
            }
            if (cij != cji)
                return fail(PHL_ERR_ARG, "phl_mpnn_set_graph: neighbour lists are not symmetric (node " +
                                             std::to_string(i) + " lists " + std::to_string(j) + ")");
        }
    HIPTRY(hipSetDevice(n->device));
    HIPTRY(hipStreamSynchronize(n->stream));
    if (n->nbr) hipFree(n->nbr);
    if (n->invDeg) hipFree(n->invDeg);
    n->nbr = nullptr;
    n->invDeg = nullptr;
    HIPTRY(hipMalloc(&n->nbr, dl.size() * sizeof(int)));
    HIPTRY(hipMalloc(&n->invDeg, inv.size() * sizeof(float)));
    HIPTRY(hipMemcpy(n->nbr, dl.data(), dl.size() * sizeof(int), hipMemcpyHostToDevice));
    HIPTRY(hipMemcpy(n->invDeg, inv.data(), inv.size() * sizeof(float), hipMemcpyHostToDevice));
    n->N = nNodes;
    n->W = width;
    return alloc_activations(n);
}

phl_status phl_mpnn_set_graph_mesh(phl_mpnn *n, const phl_mesh *mesh)
{
    if (!n || !mesh) return fail(PHL_ERR_ARG, "null argument");
    const phl::Mesh &m = mesh->m;
    std::vector<int> deg(m.nCell, 0);
    for (int f = m.nBFace; f < m.nFace; ++f) {
        deg[m.left[f]]++;
        deg[m.right[f]]++;
    }
    const int W = std::max(1, *std::max_element(deg.begin(), deg.end()));
    std::vector<int> nb((size_t)m.nCell * W, -1);
    std::fill(deg.begin(), deg.end(), 0);
    for (int f = m.nBFace; f < m.nFace; ++f) {
        const int a = m.left[f], b = m.right[f];
        nb[(size_t)a * W + deg[a]++] = b;
        nb[(size_t)b * W + deg[b]++] = a;
    }
    return phl_mpnn_set_graph(n, m.nCell, W, nb.data());
}

phl_status phl_mpnn_set_data(phl_mpnn *n, const float *X, const float *T)
{
    if (!n || !X || n->N == 0) return fail(PHL_ERR_ARG, "phl_mpnn_set_data: no graph or no features");
    HIPTRY(hipSetDevice(n->device));
    HIPTRY(hipMemcpyAsync(n->X, X, sizeof(float) * (size_t)n->N * kIn, hipMemcpyHostToDevice, n->stream));
    if (T) HIPTRY(hipMemcpyAsync(n->T, T, sizeof(float) * (size_t)n->N * kOut, hipMemcpyHostToDevice, n->stream));
    HIPTRY(hipStreamSynchronize(n->stream));
    return PHL_OK;
}

phl_status phl_mpnn_forward(phl_mpnn *n, float *Y)
{
    if (!n || !Y || n->N == 0) return fail(PHL_ERR_ARG, "phl_mpnn_forward: no graph");
    HIPTRY(hipSetDevice(n->device));
    if (n->fwdBF16) forward_bf16(n);
    else forward(n);
    HIPTRY(hipGetLastError());
    HIPTRY(hipMemcpyAsync(Y, n->y, sizeof(float) * (size_t)n->N * kOut, hipMemcpyDeviceToHost, n->stream));
    HIPTRY(hipStreamSynchronize(n->stream));
    return PHL_OK;
}

phl_status phl_mpnn_set_forward_precision(phl_mpnn *n, int bf16)
{
    if (!n || (bf16 != 0 && bf16 != 1)) return fail(PHL_ERR_ARG, "phl_mpnn_set_forward_precision: 0 (FP32) or 1 (BF16)");
    n->fwdBF16 = bf16;
    return PHL_OK;
}

phl_status phl_mpnn_gradient(phl_mpnn *n, double *lossOut, float *grad)
{
    if (!n || n->N == 0) return fail(PHL_ERR_ARG, "phl_mpnn_gradient: no graph");
    HIPTRY(hipSetDevice(n->device));
    forward(n);
    phl::nn::loss(n->y, n->T, n->N * kOut, n->dy, n->lossPart, n->lossOut, n->stream);
    backward(n);
    HIPTRY(hipGetLastError());
    if (lossOut) HIPTRY(hipMemcpyAsync(lossOut, n->lossOut, sizeof(double), hipMemcpyDeviceToHost, n->stream));
    if (grad)
        HIPTRY(hipMemcpyAsync(grad, n->g, sizeof(float) * (size_t)n->nParams, hipMemcpyDeviceToHost, n->stream));
    HIPTRY(hipStreamSynchronize(n->stream));
    return PHL_OK;
}

phl_status phl_mpnn_train(phl_mpnn *n, int steps, float lr, double *lossHist)
{
    if (!n || n->N == 0 || steps < 0) return fail(PHL_ERR_ARG, "phl_mpnn_train: bad argument");
    HIPTRY(hipSetDevice(n->device));
    std::vector<double> h(std::max(steps, 1));
    for (int i = 0; i < steps; ++i) {
        forward(n);
        phl::nn::loss(n->y, n->T, n->N * kOut, n->dy, n->lossPart, n->lossOut, n->stream);
        backward(n);
        ++n->step;
        phl::nn::adam(n->p, n->g, n->m, n->v, n->nParams, lr, n->step, n->stream);
        if (lossHist) HIPTRY(hipMemcpyAsync(&h[i], n->lossOut, sizeof(double), hipMemcpyDeviceToHost, n->stream));
        if ((i & 63) == 63) HIPTRY(hipStreamSynchronize(n->stream));
    }
    HIPTRY(hipGetLastError());
    HIPTRY(hipStreamSynchronize(n->stream));
    if (lossHist) std::copy(h.begin(), h.begin() + steps, lossHist);
    return PHL_OK;
}

}  // extern "C"
