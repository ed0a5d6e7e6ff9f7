// phl_api.cpp -- C-ABI entry points, device context and the host driver
// (the LU-SGS branch of Controller::Relaxation, CFD/Solver/src/Controller.cpp:549-608).
#include "../../include/phl.h"
#include "phl_internal.h"
#include "phl_mesh.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <utility>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

struct phl_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    phl_params p{};
    phl::DevMesh m{};
    phl::Consts k{};
    phl::DevFlow f{};
    phl::Levels fwd{}, bwd{};
    std::vector<void *> allocs;
    std::vector<int> hfwdPtr, hbwdPtr;
    int normSlots = 0;
    double *hostNorm = nullptr;  // pinned
    int *hostCount = nullptr;    // pinned
    double *hostCfl = nullptr;   // pinned, CFL of each step of a chunk
    int *hostZero = nullptr;     // pinned constant 0
    // nLUSGSSweeps > 1: the pinned half-sweep sums and the record of the last
    // NS and SA LU-SGS calls (their steps run eagerly, decided on the host)
    phl::SweepLog nsLog, saLog;
    // isInIniting for hook-by-hook callers (phl_set_init_phase); phl_iterate
    // sets it per step from flow_init_step and restores it afterwards
    int hookInit = 0;
    double primInf[5];
    // batched contexts: one parameter block and freestream per case
    std::vector<phl_params> caseParams;
    std::vector<phl::CaseConsts> hostCases;
    std::vector<int> caseCellPtr, caseBFacePtr, caseWallPtr;
    // storage position of each mesh cell and the mesh cell at each position
    // (real cells only; ghosts keep their index)
    std::vector<int> slotOf, cellAt;
    unsigned char *hostFrozen = nullptr;  // pinned [nCases]
    // kernel timing (HIP events on `stream`)
    bool timing = false;
    std::vector<hipEvent_t> events;
    double kernelMs[PHL_N_KERNEL_CLASSES] = {0};
    int timedSteps = 0;
    // hipGraphs of kGraphSteps steps, one per assignment of the q / qAlt buffers
    int launchMode = 1;  // 0 eager, 1 graph, 2 eager with the turbulence sweeps on the side stream
    double *qBuf0 = nullptr;
    hipGraphExec_t stepGraph[2] = {nullptr, nullptr};
    // Spalart-Allmaras: the scalar sweeps of a step touch only nu~ and the
    // turbulence LU-SGS arrays, which the next step's NS kernels do not use,
    // so they run on `side` beside them; evJoin orders the next relaxation
    hipStream_t side = nullptr;
    hipEvent_t evFork = nullptr, evJoin = nullptr;
    bool turbPending = false;
    // Dataflow sweeps: the scalar sweeps of a step's turbulence relaxation
    // ride along with the next step's NS sweeps (same lists, same order);
    // saDeferred: a relaxation is waiting for its sweeps and update.
    bool saFused = false;
    bool saDeferred = false;
    // phl_solve / phl_solve_batch: switch a mixed context to FP64 once no
    // unconverged case has halved its residual for this many steps (0: never)
    int precisionSwitchStall = 0;
    bool coefF32Ok = false;  // the sweep-ordered kernels (at most 6 faces a side) can read FP32 coefficients
    // hook form: the gradient kernel also evaluates the limiter, so
    // phl_limiter right after phl_gradient on an unchanged state has nothing
    // left to do.  stateVersion counts the calls that change q.
    unsigned long long stateVersion = 0, gradVersion = ~0ull;
    // launch statistics: graph replays and eagerly launched steps
    long long graphReplays = 0, eagerSteps = 0;
    // ggnode: the node temperatures of an iteration come from the previous
    // iteration's t (the reference computes node values in PostSolve, before
    // the next ComputeGamaAndTemperature); after a new state the first node
    // pass takes the state's own t, as InitDependentVariables does
    bool tFresh = true;
};

namespace {

thread_local std::string g_err;

phl_status fail(phl_status s, const std::string &msg)
{
    g_err = msg;
    return s;
}

#define HIPTRY(expr)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail(PHL_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));  \
    } while (0)

template <typename T>
phl_status dalloc(phl_ctx *c, T **p, size_t n)
{
    void *ptr = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc(&ptr, n * sizeof(T));
    if (e != hipSuccess) return fail(PHL_ERR_ALLOC, std::string("hipMalloc: ") + hipGetErrorString(e));
    c->allocs.push_back(ptr);
    *p = (T *)ptr;
    return PHL_OK;
}

template <typename T>
phl_status upload(phl_ctx *c, const T **dst, const std::vector<T> &src)
{
    T *d = nullptr;
    phl_status s = dalloc(c, &d, src.size());
    if (s != PHL_OK) return s;
    if (!src.empty()) HIPTRY(hipMemcpy(d, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
    *dst = d;
    return PHL_OK;
}

#define TRY(expr)                          \
    do {                                   \
        phl_status s_ = (expr);            \
        if (s_ != PHL_OK) return s_;       \
    } while (0)

// Cells of a zone ordered into dependency levels of the sequential sweep:
// forward, a cell depends on its face neighbours with a lower index;
// backward, on those with a higher index.  Cells of one level are
// independent, so a level can be processed in parallel and the result equals
// the reference's cell-by-cell loop.
void build_levels(const phl::Mesh &m, const std::vector<int> &c2fPtr, const std::vector<int> &c2n,
                  bool forward, std::vector<int> &ptr, std::vector<int> &cells)
{
    const int nC = m.nCell;
    std::vector<int> level(nC, 0);
    int nL = 0;
    if (forward) {
        for (int c = 0; c < nC; ++c) {
            int lv = 0;
            for (int k = c2fPtr[c]; k < c2fPtr[c + 1]; ++k) {
                const int nb = c2n[k];
                if (nb < c) lv = std::max(lv, level[nb] + 1);
            }
            level[c] = lv;
            nL = std::max(nL, lv + 1);
        }
    } else {
        for (int c = nC - 1; c >= 0; --c) {
            int lv = 0;
            for (int k = c2fPtr[c]; k < c2fPtr[c + 1]; ++k) {
                const int nb = c2n[k];
                if (nb > c && nb < nC) lv = std::max(lv, level[nb] + 1);
            }
            level[c] = lv;
            nL = std::max(nL, lv + 1);
        }
    }
    ptr.assign(nL + 1, 0);
    for (int c = 0; c < nC; ++c) ptr[level[c] + 1]++;
    for (int l = 0; l < nL; ++l) ptr[l + 1] += ptr[l];
    cells.resize(nC);
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    for (int c = 0; c < nC; ++c) cells[fill[level[c]]++] = c;
}

// Position-major active-face lists of one sweep direction: for position p
// (cell cells[p]) the faces whose neighbour precedes the cell in the sweep
// (lower index forward; higher index or ghost backward), in the cell's face
// order, as coefficient rows 2*f + side and neighbour indices.
// The lists are padded to Wpad slots (3, 4 or 6: the widths the sweep
// kernels are compiled for) with row -1.
void build_sweep_order(int nC, const std::vector<int> &c2fPtr, const std::vector<int> &c2f,
                       const std::vector<int> &c2n, const std::vector<char> &isLeft, const std::vector<int> &cells,
                       bool forward, int &W, int &Wpad, std::vector<int> &row, std::vector<int> &nb)
{
    W = 0;
    for (int c = 0; c < nC; ++c) {
        int n = 0;
        for (int k = c2fPtr[c]; k < c2fPtr[c + 1]; ++k) n += forward ? (c2n[k] < c) : (c2n[k] > c);
        W = std::max(W, n);
    }
    Wpad = W <= 3 ? 3 : (W <= 4 ? 4 : (W <= 6 ? 6 : W));
    const int Wa = Wpad;
    row.assign((size_t)Wa * nC, -1);
    nb.assign((size_t)Wa * nC, 0);
    for (int p = 0; p < nC; ++p) {
        const int c = cells[p];
        int s = 0;
        for (int k = c2fPtr[c]; k < c2fPtr[c + 1]; ++k) {
            const int n = c2n[k];
            if (forward ? (n < c) : (n > c)) {
                row[(size_t)s * nC + p] = 2 * c2f[k] + (isLeft[k] ? 0 : 1);
                nb[(size_t)s * nC + p] = n;
                ++s;
            }
        }
    }
}

// The dataflow sweep's order (phl::FlowOrder) for one direction: each level
// of `cells` (mesh numbers, level order) padded to whole chunks of 64
// positions, the slot lists of build_sweep_order moved along (row, nb in mesh
// numbers; converted to storage here), and per chunk the distinct chunks
// holding its real neighbours, all of earlier levels.
phl_status build_flow_order(phl_ctx *c, int nC, const std::vector<int> &cells, const std::vector<int> &lptr,
                            const std::vector<int> &row, const std::vector<int> &nb, int Wa,
                            const std::vector<int> &slotOf, phl::FlowOrder &fo)
{
    int nCh = 0;
    for (size_t L = 0; L + 1 < lptr.size(); ++L) nCh += (lptr[L + 1] - lptr[L] + 63) / 64;
    const size_t Pp = (size_t)64 * nCh;
    std::vector<int> fc(Pp, -1), frow((size_t)Wa * Pp, -1), fnb((size_t)Wa * Pp, 0), chunkOf(nC, -1);
    std::vector<int> depPtr(1, 0), dep, d;
    int ch = 0;
    for (size_t L = 0; L + 1 < lptr.size(); ++L) {
        const int levelChunk0 = ch;
        for (int b = lptr[L]; b < lptr[L + 1]; b += 64, ++ch) {
            d.clear();
            for (int j = 0; j < 64 && b + j < lptr[L + 1]; ++j) {
                const int pos = b + j;
                const size_t q = (size_t)ch * 64 + j;
                fc[q] = slotOf[cells[pos]];
                chunkOf[cells[pos]] = ch;
                for (int s = 0; s < Wa; ++s) {
                    const int r = row[(size_t)s * nC + pos], n = nb[(size_t)s * nC + pos];
                    frow[(size_t)s * Pp + q] = r;
                    fnb[(size_t)s * Pp + q] = n < nC ? slotOf[n] : n;
                    if (r < 0 || n >= nC) continue;
                    // a neighbour swept earlier lies in a chunk of an earlier level
                    if (chunkOf[n] < 0 || chunkOf[n] >= levelChunk0)
                        return fail(PHL_ERR_ARG, "dataflow sweep order: a neighbour is not in an earlier level");
                    d.push_back(chunkOf[n]);
                }
            }
            std::sort(d.begin(), d.end());
            d.erase(std::unique(d.begin(), d.end()), d.end());
            dep.insert(dep.end(), d.begin(), d.end());
            depPtr.push_back((int)dep.size());
        }
    }
    if (dep.empty()) dep.push_back(0);  // keeps the device pointer valid
    // XCD groups (sweep_form 3): block ids rise with the level, as in launch
    // order, so the in-order dispatch argument for progress is unchanged; the
    // padding blocks of a level return at once
    std::vector<int> blockChunk;
    {
        int ch0 = 0;
        for (size_t L = 0; L + 1 < lptr.size(); ++L) {
            const int nL = (lptr[L + 1] - lptr[L] + 63) / 64, g = (nL + 7) / 8;
            for (int j = 0; j < g; ++j)
                for (int x = 0; x < 8; ++x) blockChunk.push_back(x * g + j < nL ? ch0 + x * g + j : -1);
            ch0 += nL;
        }
        if (blockChunk.empty()) blockChunk.push_back(-1);
    }
    std::vector<int> width(std::max(nCh, 1), 0);
    for (int k = 0; k < nCh; ++k)
        for (int j = 0; j < 64; ++j)
            for (int s = 0; s < Wa; ++s)
                if (frow[(size_t)s * Pp + (size_t)k * 64 + j] >= 0) width[k] = std::max(width[k], s + 1);
    const int *t;
    TRY(upload(c, &t, width)); fo.width = t;
    TRY(upload(c, &t, fc)); fo.cells = t;
    TRY(upload(c, &t, frow)); fo.row = t;
    TRY(upload(c, &t, fnb)); fo.nb = t;
    TRY(upload(c, &t, depPtr)); fo.depPtr = t;
    TRY(upload(c, &t, dep)); fo.dep = t;
    TRY(upload(c, &t, blockChunk)); fo.blockChunk = t;
    fo.nBlocks = (int)blockChunk.size();
    fo.nChunks = nCh;
    return PHL_OK;
}

double compute_cfl(const phl_params &p, int step)
{
    // NSSolver::ComputeCFL (CFD/NSSolver/src/NSSolver.cpp:595-668), CFLMethod 0.
    if (step >= p.cfl_vary_step) return p.cfl_end;
    return p.cfl_start + (p.cfl_end - p.cfl_start) * step / p.cfl_vary_step;
}

double ptotal(double mach, double gama)
{
    // PerfectGas::ComputePTotal (API/Physics/src/PerfectGas.cpp:1616-1641).
    const double R = 1.0 / (gama * mach * mach);
    double pref = R * 1.0 * 1.0 / 1.0;
    const double m12 = mach * mach;
    const double gamp1 = gama + 1.0, gamm1 = gama - 1.0;
    const double c1 = 2.0 / gamm1 + m12;
    const double c2 = 2.0 * gama / gamm1 * m12 - 1.0;
    const double cc = gama / (gama - 1.0);
    const double t0t = 1.0 + 0.5 * gamm1 * c1 / c2;
    if (mach > 1.0) pref *= (2.0 * gama / gamp1 * m12 - gamm1 / gamp1);
    return pref * std::pow(t0t, cc);
}

double entropy_coef(double mach, double scale)
{
    // Param_NSSolver::SetEntropyFixCoefficients (CFD/NSSolver/src/Param_NSSolver.cpp:225-266).
    double e;
    if (mach <= 0.6) e = 0.0001;
    else if (mach <= 2.0) {
        const double x0 = 0.6, x1 = 2.0, y0 = 0.0001, y1 = 0.01;
        e = y0 * (mach - x1) / (x0 - x1) + y1 * (mach - x0) / (x1 - x0);
    } else if (mach <= 5.0) {
        const double x0 = 2.0, x1 = 5.0, y0 = 0.01, y1 = 0.5;
        e = y0 * (mach - x1) / (x0 - x1) + y1 * (mach - x0) / (x1 - x0);
    } else e = 1.0;
    return e * scale;
}

phl_status check_params(const phl_params &p)
{
    if (!(p.mach > 0.0) || !(p.gamma > 1.0)) return fail(PHL_ERR_ARG, "mach and gamma must be positive (gamma > 1)");
    if (p.viscous != 0 && p.viscous != 1 && p.viscous != 3)
        return fail(PHL_ERR_UNSUPPORTED, "viscous must be 0 (inviscid), 1 (laminar) or 3 (Spalart-Allmaras)");
    if (p.viscous == 3 && !(p.turb_cfl_scale > 0.0)) return fail(PHL_ERR_ARG, "turb_cfl_scale must be positive");
    if (p.viscous && !(p.reynolds > 0.0)) return fail(PHL_ERR_ARG, "reynolds must be positive for viscous flow");
    if (p.gradient < 0 || p.gradient > 2)
        return fail(PHL_ERR_UNSUPPORTED, "gradient must be 0 (ggcell), 1 (lsq) or 2 (ggnode)");
    if (p.limiter < 0 || p.limiter > 2) return fail(PHL_ERR_UNSUPPORTED, "limiter must be 0 (vencat), 1 (nolim) or 2 (1st)");
    if (p.entropy_fix < 1 || p.entropy_fix > 6)
        return fail(PHL_ERR_UNSUPPORTED, "roeEntropyFixMethod must be 1 to 6");
    if (p.n_lusgs_sweeps < 1) return fail(PHL_ERR_ARG, "nLUSGSSweeps must be at least 1");
    if (p.flow_init_step < 0) return fail(PHL_ERR_ARG, "flowInitStep must not be negative");
    if (p.precision != 0 && p.precision != 1)
        return fail(PHL_ERR_UNSUPPORTED, "precision must be 0 (FP64) or 1 (mixed)");
    if (p.sweep_form < 0 || p.sweep_form > 3)
        return fail(PHL_ERR_ARG, "sweep_form must be 0 (dataflow, launch order), 1 (one launch per level), 2 "
                                 "(dataflow, ticket order) or 3 (dataflow, levels split over the XCDs)");
    if (p.cfl_vary_step <= 0) return fail(PHL_ERR_ARG, "cfl_vary_step must be positive");
    return PHL_OK;
}

// Options that depend on the grid: the least-squares weights are restated
// for 3-D grids only (the reference's 2-D branch, Geo_UnstructGrid.cpp:3879,
// inverts 2x2 matrices).
phl_status check_mesh_params(const phl::Mesh &m, const phl_params &p)
{
    if (m.dim == 2 && p.gradient == 1)
        return fail(PHL_ERR_UNSUPPORTED, "the lsq gradient is implemented for 3-D grids only");
    return PHL_OK;
}

phl_status check_launch(phl_ctx *c)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(PHL_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
    (void)c;
    return PHL_OK;
}

// Default bound of a dataflow sweep's waits: 0.2 s of the 100 MHz wall clock.
constexpr int kSweepWaitTicks = 20000000;

// Dataflow sweeps bound every wait; a wave that gave up leaves its cells'
// dq unfinished and is counted in sweepCtl[2].  Checked after the work is
// done (synchronises the stream).
phl_status check_sweep_waits(phl_ctx *c)
{
    if (c->k.sweepForm == 1) return PHL_OK;
    int gaveUp = 0;
    HIPTRY(hipMemcpyAsync(&gaveUp, c->f.sweepCtl + 2, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPTRY(hipStreamSynchronize(c->stream));
    if (gaveUp != 0)
        return fail(PHL_ERR_NUMERIC, "dataflow sweep: " + std::to_string(gaveUp) +
                                         " wavefronts exceeded the wait bound; the state is invalid");
    return PHL_OK;
}

// Real cells of one cell-indexed row from storage order to mesh order (in place).
void row_to_mesh(const phl_ctx *c, double *row, std::vector<double> &tmp)
{
    const int nC = (int)c->cellAt.size();
    tmp.assign(row, row + nC);
    for (int i = 0; i < nC; ++i) row[c->cellAt[i]] = tmp[i];
}

phl_status copy_field(phl_ctx *c, const double *dsrc, int rows, int rowLen, int outRowLen, double *out)
{
    // rows of rowLen device values (cell-indexed, storage order) into rows of
    // outRowLen host values in mesh order (zero padded)
    for (int r = 0; r < rows; ++r) {
        double *o = out + (size_t)r * outRowLen;
        if (rowLen > 0)
            HIPTRY(hipMemcpyAsync(o, dsrc + (size_t)r * rowLen, (size_t)rowLen * sizeof(double),
                                  hipMemcpyDeviceToHost, c->stream));
        for (int i = rowLen; i < outRowLen; ++i) o[i] = 0.0;
    }
    HIPTRY(hipStreamSynchronize(c->stream));
    std::vector<double> tmp;
    for (int r = 0; r < rows && rowLen >= c->m.nCell; ++r) row_to_mesh(c, out + (size_t)r * outRowLen, tmp);
    return PHL_OK;
}

// The five components of a [n][kRow] device array (n rows, storage order)
// into five host rows of outRowLen values in mesh order (zero padded).
phl_status copy_rows(phl_ctx *c, const double *dsrc, int n, int outRowLen, double *out)
{
    std::vector<double> h((size_t)phl::kRow * n);
    if (n > 0)
        HIPTRY(hipMemcpyAsync(h.data(), dsrc, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->stream));
    HIPTRY(hipStreamSynchronize(c->stream));
    for (int m = 0; m < 5; ++m) {
        double *o = out + (size_t)m * outRowLen;
        for (int i = 0; i < outRowLen; ++i) o[i] = i < n ? h[(size_t)i * phl::kRow + m] : 0.0;
    }
    std::vector<double> tmp;
    for (int m = 0; m < 5 && n >= c->m.nCell; ++m) row_to_mesh(c, out + (size_t)m * outRowLen, tmp);
    return PHL_OK;
}

// Host rows of rowLen cell-indexed values (mesh order) to the device (storage order).
phl_status upload_cells(phl_ctx *c, double *ddst, int rows, int rowLen, const double *in)
{
    const int nC = c->m.nCell;
    std::vector<double> h(in, in + (size_t)rows * rowLen);
    for (int r = 0; r < rows; ++r) {
        double *o = h.data() + (size_t)r * rowLen;
        const double *src = in + (size_t)r * rowLen;
        for (int i = 0; i < nC && i < rowLen; ++i) o[c->slotOf[i]] = src[i];
    }
    HIPTRY(hipMemcpyAsync(ddst, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, c->stream));
    HIPTRY(hipStreamSynchronize(c->stream));
    return PHL_OK;
}

// FP32 rows (mixed-form gradients) widened to FP64 on the host.
phl_status copy_field_f32(phl_ctx *c, const float *dsrc, int rows, int rowLen, int outRowLen, double *out)
{
    std::vector<float> tmp((size_t)rows * rowLen);
    if (!tmp.empty())
        HIPTRY(hipMemcpyAsync(tmp.data(), dsrc, tmp.size() * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPTRY(hipStreamSynchronize(c->stream));
    for (int r = 0; r < rows; ++r)
        for (int i = 0; i < outRowLen; ++i)
            out[(size_t)r * outRowLen + i] = i < rowLen ? (double)tmp[(size_t)r * rowLen + i] : 0.0;
    std::vector<double> t2;
    for (int r = 0; r < rows && rowLen >= c->m.nCell; ++r) row_to_mesh(c, out + (size_t)r * outRowLen, t2);
    return PHL_OK;
}

void swap_state(phl_ctx *c) { std::swap(c->f.q, c->f.qAlt); }

}  // namespace

void phl::set_last_error(const std::string &msg) { g_err = msg; }

extern "C" {

const char *phl_last_error(void) { return g_err.c_str(); }

void phl_params_default(phl_params *p)
{
    std::memset(p, 0, sizeof(*p));
    p->mach = 0.5;
    p->alpha_deg = 0.0;
    p->beta_deg = 0.0;
    p->gamma = 1.4;
    p->reynolds = 6.5e6;
    p->t_ref = 288.15;
    p->prandtl = 0.72;
    p->prandtl_turb = 0.9;
    p->t_wall = -1.0;
    p->venkat_k = 5.0;
    p->entropy_scale = 1.0;
    p->cfl_start = 0.01;
    p->cfl_end = 10.0;
    p->ktmax = 1.0e10;
    p->lusgs_tol = 0.01;
    p->force_ref_length = 1.0;
    p->force_ref_area = 1.0;
    p->force_ref_span = 1.0;
    p->sa_nu_inf = 3.0;
    p->sa_vist_inf = 1.0e-3;
    p->turb_cfl_scale = 1.0;
    p->turb_skew_angle = 0.0;
    p->viscous = 0;
    p->gradient = 0;
    p->limiter = 0;
    p->entropy_fix = 3;
    p->cfl_vary_step = 500;
    p->n_lusgs_sweeps = 1;
    p->precision = 0;
}

// ------------------------------------------------------------------ meshes
phl_status phl_mesh_create(int nNode, const double *x, const double *y, const double *z, int nFace,
                           const int *faceNodePtr, const int *faceNodes, const int *left, const int *right,
                           int nBFace, const int *bcType, int nCell, const int *cellNodePtr,
                           const int *cellNodes, phl_mesh **out)
{
    if (!out || nNode <= 0 || nFace <= 0 || nCell <= 0 || nBFace < 0 || nBFace > nFace)
        return fail(PHL_ERR_ARG, "phl_mesh_create: bad sizes");
    if (!x || !y || !z || !faceNodePtr || !faceNodes || !left || !right || !cellNodePtr || !cellNodes ||
        (nBFace > 0 && !bcType))
        return fail(PHL_ERR_ARG, "phl_mesh_create: null array");
    phl_mesh *pm = new (std::nothrow) phl_mesh;
    if (!pm) return fail(PHL_ERR_ALLOC, "phl_mesh_create: out of memory");
    phl::Mesh &m = pm->m;
    m.nNode = nNode; m.nCell = nCell; m.nFace = nFace; m.nBFace = nBFace;
    m.x.assign(x, x + nNode); m.y.assign(y, y + nNode); m.z.assign(z, z + nNode);
    m.faceNodePtr.assign(faceNodePtr, faceNodePtr + nFace + 1);
    m.faceNodes.assign(faceNodes, faceNodes + faceNodePtr[nFace]);
    m.cellNodePtr.assign(cellNodePtr, cellNodePtr + nCell + 1);
    m.cellNodes.assign(cellNodes, cellNodes + cellNodePtr[nCell]);
    m.left.assign(left, left + nFace);
    m.right.assign(right, right + nFace);
    m.bcType.assign(bcType, bcType + nBFace);
    for (int f = 0; f < nFace; ++f) {
        const bool ok = left[f] >= 0 && left[f] < nCell &&
                        (f < nBFace ? right[f] == nCell + f : (right[f] >= 0 && right[f] < nCell && right[f] != left[f]));
        if (!ok) {
            delete pm;
            return fail(PHL_ERR_ARG, "phl_mesh_create: face " + std::to_string(f) + " has bad left/right cells");
        }
    }
    // a planar grid: every face an edge (the reference's 2-D metric path)
    bool edges = true;
    for (int f = 0; f < nFace && edges; ++f) edges = faceNodePtr[f + 1] - faceNodePtr[f] == 2;
    m.dim = edges ? 2 : 3;
    m.computeMetrics();
    *out = pm;
    return PHL_OK;
}

phl_status phl_mesh_bump_hex(int nx, int ny, int nz, double bump, phl_mesh **out)
{
    if (!out) return fail(PHL_ERR_ARG, "null out");
    phl_mesh *pm = new (std::nothrow) phl_mesh;
    if (!pm) return fail(PHL_ERR_ALLOC, "out of memory");
    std::string err;
    if (!phl::makeBumpHex(nx, ny, nz, bump, pm->m, err)) {
        delete pm;
        return fail(PHL_ERR_ARG, err);
    }
    *out = pm;
    return PHL_OK;
}

phl_status phl_mesh_bump_hybrid(int nx, int nz, int n_prism, double h1, int n_tet, double jitter, unsigned seed,
                                double bump, phl_mesh **out)
{
    if (!out) return fail(PHL_ERR_ARG, "null out");
    phl_mesh *pm = new (std::nothrow) phl_mesh;
    if (!pm) return fail(PHL_ERR_ALLOC, "out of memory");
    std::string err;
    if (!phl::makeBumpHybrid(nx, nz, n_prism, h1, n_tet, jitter, seed, bump, pm->m, err)) {
        delete pm;
        return fail(PHL_ERR_ARG, err);
    }
    *out = pm;
    return PHL_OK;
}

phl_status phl_mesh_swept_wing(int n_chord, int n_span, int n_prism, double h1, int n_tet, phl_mesh **out)
{
    if (!out) return fail(PHL_ERR_ARG, "null out");
    phl_mesh *pm = new (std::nothrow) phl_mesh;
    if (!pm) return fail(PHL_ERR_ALLOC, "out of memory");
    std::string err;
    if (!phl::makeSweptWing(n_chord, n_span, n_prism, h1, n_tet, pm->m, err)) {
        delete pm;
        return fail(PHL_ERR_ARG, err);
    }
    *out = pm;
    return PHL_OK;
}

phl_status phl_mesh_renumber(const phl_mesh *in, int method, const int *perm, int *n_colours, phl_mesh **out)
{
    if (!in || !out) return fail(PHL_ERR_ARG, "null argument");
    if (in->m.nCases() > 1) return fail(PHL_ERR_UNSUPPORTED, "phl_mesh_renumber: batched mesh");
    std::vector<int> pv;
    int nCol = 0;
    if (method == 0) {
        pv = phl::colourOrder(in->m, nCol);
    } else if (method == 1) {
        if (!perm) return fail(PHL_ERR_ARG, "phl_mesh_renumber: null permutation");
        pv.assign(perm, perm + in->m.nCell);
    } else {
        return fail(PHL_ERR_ARG, "phl_mesh_renumber: unknown method");
    }
    phl_mesh *pm = new (std::nothrow) phl_mesh;
    if (!pm) return fail(PHL_ERR_ALLOC, "out of memory");
    std::string err;
    if (!phl::renumberCells(in->m, pv, pm->m, err)) {
        delete pm;
        return fail(PHL_ERR_ARG, err);
    }
    if (n_colours) *n_colours = nCol;
    *out = pm;
    return PHL_OK;
}

phl_status phl_mesh_concat(const phl_mesh *const *parts, int n, phl_mesh **out)
{
    if (!parts || n <= 0 || !out) return fail(PHL_ERR_ARG, "phl_mesh_concat: bad argument");
    std::vector<const phl::Mesh *> pv;
    for (int i = 0; i < n; ++i) {
        if (!parts[i]) return fail(PHL_ERR_ARG, "phl_mesh_concat: null mesh");
        pv.push_back(&parts[i]->m);
    }
    phl_mesh *pm = new (std::nothrow) phl_mesh;
    if (!pm) return fail(PHL_ERR_ALLOC, "out of memory");
    std::string err;
    if (!phl::concatMeshes(pv, pm->m, err)) {
        delete pm;
        return fail(PHL_ERR_ARG, err);
    }
    *out = pm;
    return PHL_OK;
}

phl_status phl_mesh_cases(const phl_mesh *pm, int *n_cases, int *cell_ptr, int *bface_ptr)
{
    if (!pm || !n_cases) return fail(PHL_ERR_ARG, "null argument");
    const phl::Mesh &m = pm->m;
    *n_cases = m.nCases();
    if (m.nCases() == 1) {
        if (cell_ptr) { cell_ptr[0] = 0; cell_ptr[1] = m.nCell; }
        if (bface_ptr) { bface_ptr[0] = 0; bface_ptr[1] = m.nBFace; }
        return PHL_OK;
    }
    if (cell_ptr) std::copy(m.caseCellPtr.begin(), m.caseCellPtr.end(), cell_ptr);
    if (bface_ptr) std::copy(m.caseBFacePtr.begin(), m.caseBFacePtr.end(), bface_ptr);
    return PHL_OK;
}

phl_status phl_mesh_sizes(const phl_mesh *pm, int *nNode, int *nCell, int *nFace, int *nBFace, int *nFaceNodes,
                          int *nCellNodes)
{
    if (!pm) return fail(PHL_ERR_ARG, "null mesh");
    const phl::Mesh &m = pm->m;
    if (nNode) *nNode = m.nNode;
    if (nCell) *nCell = m.nCell;
    if (nFace) *nFace = m.nFace;
    if (nBFace) *nBFace = m.nBFace;
    if (nFaceNodes) *nFaceNodes = (int)m.faceNodes.size();
    if (nCellNodes) *nCellNodes = (int)m.cellNodes.size();
    return PHL_OK;
}

phl_status phl_mesh_connectivity(const phl_mesh *pm, double *x, double *y, double *z, int *faceNodePtr,
                                 int *faceNodes, int *left, int *right, int *bcType, int *cellNodePtr,
                                 int *cellNodes)
{
    if (!pm) return fail(PHL_ERR_ARG, "null mesh");
    const phl::Mesh &m = pm->m;
    auto cp = [](auto *dst, const auto &v) {
        if (dst) std::copy(v.begin(), v.end(), dst);
    };
    cp(x, m.x); cp(y, m.y); cp(z, m.z);
    cp(faceNodePtr, m.faceNodePtr); cp(faceNodes, m.faceNodes);
    cp(left, m.left); cp(right, m.right); cp(bcType, m.bcType);
    cp(cellNodePtr, m.cellNodePtr); cp(cellNodes, m.cellNodes);
    return PHL_OK;
}

phl_status phl_mesh_metrics(const phl_mesh *pm, double *xfc, double *yfc, double *zfc, double *xfn, double *yfn,
                            double *zfn, double *area, double *xcc, double *ycc, double *zcc, double *vol,
                            double *skew)
{
    if (!pm) return fail(PHL_ERR_ARG, "null mesh");
    const phl::Mesh &m = pm->m;
    auto cp = [](double *dst, const std::vector<double> &v) {
        if (dst) std::copy(v.begin(), v.end(), dst);
    };
    cp(xfc, m.xfc); cp(yfc, m.yfc); cp(zfc, m.zfc);
    cp(xfn, m.xfn); cp(yfn, m.yfn); cp(zfn, m.zfn); cp(area, m.area);
    cp(xcc, m.xcc); cp(ycc, m.ycc); cp(zcc, m.zcc); cp(vol, m.vol); cp(skew, m.skew);
    return PHL_OK;
}

void phl_mesh_destroy(phl_mesh *m) { delete m; }

int phl_mesh_dim(const phl_mesh *m) { return m ? m->m.dim : 0; }

// ------------------------------------------------------------- lifecycle
phl_status phl_device_info(int device, char *name, int name_len, int *n_cu, double *hbm_bytes)
{
    hipDeviceProp_t prop;
    HIPTRY(hipGetDeviceProperties(&prop, device));
    if (name && name_len > 0) {
        std::snprintf(name, (size_t)name_len, "%s (%s)", prop.name, prop.gcnArchName);
    }
    if (n_cu) *n_cu = prop.multiProcessorCount;
    if (hbm_bytes) *hbm_bytes = (double)prop.totalGlobalMem;
    return PHL_OK;
}

void phl_destroy(phl_ctx *c)
{
    if (!c) return;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    if (c->side) hipStreamSynchronize(c->side);
    for (void *p : c->allocs) hipFree(p);
    if (c->hostNorm) hipHostFree(c->hostNorm);
    if (c->hostFrozen) hipHostFree(c->hostFrozen);
    if (c->hostCount) hipHostFree(c->hostCount);
    if (c->nsLog.host) hipHostFree(c->nsLog.host);
    if (c->hostCfl) hipHostFree(c->hostCfl);
    if (c->hostZero) hipHostFree(c->hostZero);
    for (hipEvent_t e : c->events) hipEventDestroy(e);
    for (hipGraphExec_t g : c->stepGraph)
        if (g) hipGraphExecDestroy(g);
    if (c->side) {
        hipStreamSynchronize(c->side);
        hipStreamDestroy(c->side);
    }
    if (c->evFork) hipEventDestroy(c->evFork);
    if (c->evJoin) hipEventDestroy(c->evJoin);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

// The freestream-dependent scalars of one case (Param_NSSolver and
// PerfectGas nondimensionalisation: rho = 1, |V| = 1, p = 1/(gamma M^2)).
static phl::CaseConsts case_consts(const phl_params &p, double avgVol, int dim)
{
    phl::CaseConsts cc{};
    cc.R = 1.0 / (p.gamma * p.mach * p.mach);
    cc.mach = p.mach;
    const double kPiD = 3.141592653589793;
    const double attack = p.alpha_deg * kPiD / 180.0, slide = p.beta_deg * kPiD / 180.0;
    cc.primInf[0] = 1.0;
    cc.primInf[1] = 1.0 * (std::cos(attack) * std::cos(slide));
    cc.primInf[2] = 1.0 * (std::sin(attack) * std::cos(slide));
    cc.primInf[3] = 1.0 * std::sin(slide);
    cc.primInf[4] = cc.R * 1.0 * 1.0 / 1.0;
    cc.entropyCoef = entropy_coef(p.mach, p.entropy_scale);
    double e = p.venkat_k * p.venkat_k * p.venkat_k;
    e *= 1.0 / std::pow(std::pow(avgVol, 1.0 / dim), 3.0);
    cc.epsTmp = e;
    const double pT = ptotal(p.mach, p.gamma);
    cc.pLower = 0.001 * pT;
    cc.pUpper = 0.02 * pT;
    cc.cp = 1.0 / ((p.gamma - 1.0) * (p.mach * p.mach));
    cc.densLimit = 1.0e-6 * cc.primInf[0];
    cc.presLimit = 1.0e-6 * cc.primInf[4];
    return cc;
}

// Storage order of the real cells: within each case, cells sorted by the
// Morton key of their centroid, each coordinate replaced by its rank among
// the case's cells (so thin wall layers and stretched farfield cells get as
// many key bits as any other region); ties keep the mesh order.  Contiguous
// blocks of this order are compact in space, so a workgroup's neighbour
// gathers mostly hit data its own XCD's L2 already holds.
static void locality_order(const phl::Mesh &m, std::vector<int> &cellAt)
{
    const int nC = m.nCell;
    cellAt.resize(nC);
    std::vector<int> ranges = m.caseCellPtr;
    if (ranges.empty()) ranges = {0, nC};
    std::vector<uint64_t> key(nC, 0);
    std::vector<int> idx;
    for (size_t r = 0; r + 1 < ranges.size(); ++r) {
        const int b = ranges[r], e = ranges[r + 1], n = e - b;
        if (n <= 0) continue;
        std::vector<uint32_t> q[3];
        const double *xyz[3] = {m.xcc.data(), m.ycc.data(), m.zcc.data()};
        for (int a = 0; a < 3; ++a) {
            idx.resize(n);
            for (int i = 0; i < n; ++i) idx[i] = b + i;
            std::stable_sort(idx.begin(), idx.end(), [&](int u, int v) { return xyz[a][u] < xyz[a][v]; });
            q[a].assign(n, 0);
            for (int i = 0; i < n; ++i)
                q[a][idx[i] - b] = (uint32_t)(((uint64_t)i << 21) / (uint64_t)n);  // 21-bit rank
        }
        for (int i = 0; i < n; ++i) {
            uint64_t k = 0;
            for (int bit = 20; bit >= 0; --bit)
                for (int a = 0; a < 3; ++a) k = (k << 1) | ((q[a][i] >> bit) & 1u);
            key[b + i] = k;
        }
        for (int i = 0; i < n; ++i) cellAt[b + i] = b + i;
        std::stable_sort(cellAt.begin() + b, cellAt.begin() + e, [&](int u, int v) { return key[u] < key[v]; });
    }
}

// Node-based Green-Gauss gradient (gradient 2): each node's contributors
// in the order ComputeNodeValue adds them (NSSolverUnstruct.cpp:4734-5081).
// One BC region per code, in increasing code order, faces in increasing
// index: solid-surface faces (for solid nodes), farfield faces (for farfield
// nodes that are not solid), the other non-symmetry codes (for nodes that
// are neither), then the real cells in order for every other node
// (symmetry-plane nodes included).  Weights are 1/distance from the node to
// the face centre or cell centre, and their sums are added in the same
// order (the reference's nodeWeight), so the device divides exactly as the
// reference does.  The face node lists follow the stored face order.
static phl_status upload_node_lists(phl_ctx *c, const phl::Mesh &m, const std::vector<int> &fslot)
{
    phl::DevMesh &dm = c->m;
    dm.nNode = 0;
    dm.nodeSrcPtr = nullptr; dm.nodeSrc = nullptr; dm.nodeW = nullptr; dm.nodeWsum = nullptr;
    dm.fnPtr = nullptr; dm.fnList = nullptr;
    if (c->p.gradient != 2) return PHL_OK;
    const int nN = m.nNode, nC = m.nCell, nB = m.nBFace, nF = m.nFace;
    const std::vector<int> &slotOf = c->slotOf;
    std::vector<int> nodeBC(nN, 0);
    auto nodesOf = [&](int f, auto &&fn) {
        for (int j = m.faceNodePtr[f]; j < m.faceNodePtr[f + 1]; ++j) fn(m.faceNodes[j]);
    };
    for (int f = 0; f < nB; ++f)
        if (m.bcType[f] == PHL_BC_SOLID_SURFACE) nodesOf(f, [&](int n) { nodeBC[n] = PHL_BC_SOLID_SURFACE; });
    for (int f = 0; f < nB; ++f)
        if (m.bcType[f] == PHL_BC_FARFIELD)
            nodesOf(f, [&](int n) { if (nodeBC[n] != PHL_BC_SOLID_SURFACE) nodeBC[n] = PHL_BC_FARFIELD; });
    for (int code = 1; code <= 6; ++code) {
        if (code == PHL_BC_SOLID_SURFACE || code == PHL_BC_FARFIELD || code == PHL_BC_SYMMETRY) continue;
        for (int f = 0; f < nB; ++f)
            if (m.bcType[f] == code)
                nodesOf(f, [&](int n) {
                    if (nodeBC[n] != PHL_BC_SOLID_SURFACE && nodeBC[n] != PHL_BC_FARFIELD) nodeBC[n] = code;
                });
    }
    std::vector<std::vector<std::pair<int2, double>>> src(nN);
    auto add = [&](int n, double px, double py, double pz, int2 s2) {
        const double dx = m.x[n] - px, dy = m.y[n] - py, dz = m.z[n] - pz;
        const double dist = std::sqrt(dx * dx + dy * dy + dz * dz);
        src[n].push_back({s2, 1.0 / dist});
    };
    for (int step = 0; step < 3; ++step)
        for (int code = 1; code <= 6; ++code) {
            if (code == PHL_BC_SYMMETRY) continue;
            if (step == 0 && code != PHL_BC_SOLID_SURFACE) continue;
            if (step == 1 && code != PHL_BC_FARFIELD) continue;
            if (step == 2 && (code == PHL_BC_SOLID_SURFACE || code == PHL_BC_FARFIELD)) continue;
            for (int f = 0; f < nB; ++f) {
                if (m.bcType[f] != code) continue;
                const int2 s2 = make_int2(slotOf[m.left[f]], nC + f);
                nodesOf(f, [&](int n) {
                    if (step == 1 && nodeBC[n] == PHL_BC_SOLID_SURFACE) return;
                    if (step == 2 && (nodeBC[n] == PHL_BC_SOLID_SURFACE || nodeBC[n] == PHL_BC_FARFIELD)) return;
                    add(n, m.xfc[f], m.yfc[f], m.zfc[f], s2);
                });
            }
        }
    for (int i = 0; i < nC; ++i)
        for (int j = m.cellNodePtr[i]; j < m.cellNodePtr[i + 1]; ++j) {
            const int n = m.cellNodes[j];
            if (nodeBC[n] != 0) continue;
            add(n, m.xcc[i], m.ycc[i], m.zcc[i], make_int2(slotOf[i], -1));
        }
    std::vector<int> ptr(nN + 1, 0);
    std::vector<int2> list;
    std::vector<double> w, wsum(nN, 0.0);
    for (int n = 0; n < nN; ++n) {
        for (const auto &e : src[n]) {
            list.push_back(e.first);
            w.push_back(e.second);
            wsum[n] += e.second;
        }
        ptr[n + 1] = (int)list.size();
    }
    if (list.empty()) { list.push_back(make_int2(0, -1)); w.push_back(0.0); }
    std::vector<int> fslotInv(nF), fptr(nF + 1, 0), flist;
    for (int f = 0; f < nF; ++f) fslotInv[fslot[f]] = f;
    for (int sf = 0; sf < nF; ++sf) {
        const int f = fslotInv[sf];
        for (int j = m.faceNodePtr[f]; j < m.faceNodePtr[f + 1]; ++j) flist.push_back(m.faceNodes[j]);
        fptr[sf + 1] = (int)flist.size();
    }
    dm.nNode = nN;
    TRY(upload(c, &dm.nodeSrcPtr, ptr));
    TRY(upload(c, &dm.nodeSrc, list));
    TRY(upload(c, &dm.nodeW, w));
    TRY(upload(c, &dm.nodeWsum, wsum));
    TRY(upload(c, &dm.fnPtr, fptr));
    TRY(upload(c, &dm.fnList, flist));
    TRY(dalloc(c, &c->f.nodeQ, (size_t)6 * nN));
    TRY(dalloc(c, &c->f.faceQ, (size_t)6 * nF));
    HIPTRY(hipMemset(c->f.nodeQ, 0, sizeof(double) * 6 * nN));
    return PHL_OK;
}

static phl_status create_impl(phl_ctx *c, const phl::Mesh &m)
{
    const phl_params &p = c->p;
    const int nC = m.nCell, nB = m.nBFace, nF = m.nFace, nT = m.nTotal();
    HIPTRY(hipSetDevice(c->device));
    HIPTRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));

    // storage order of the real cells; S() maps a mesh index (real or ghost)
    // to its storage index
    if (p.storage_order == 0) {
        locality_order(m, c->cellAt);
    } else {
        c->cellAt.resize(nC);
        for (int i = 0; i < nC; ++i) c->cellAt[i] = i;
    }
    c->slotOf.assign(nC, 0);
    for (int i = 0; i < nC; ++i) c->slotOf[c->cellAt[i]] = i;
    const std::vector<int> &slotOf = c->slotOf, &cellAt = c->cellAt;
    auto S = [&](int i) { return i < nC ? slotOf[i] : i; };
    auto permuteCells = [&](const std::vector<double> &v) {  // [nTotal] or [nCell], mesh -> storage order
        std::vector<double> o(v);
        for (int i = 0; i < nC; ++i) o[slotOf[i]] = v[i];
        return o;
    };

    // cell -> face (increasing face order) and cell -> neighbour
    std::vector<int> c2fPtr(nC + 1, 0);
    for (int f = 0; f < nF; ++f) {
        c2fPtr[m.left[f] + 1]++;
        if (f >= nB) c2fPtr[m.right[f] + 1]++;
    }
    for (int i = 0; i < nC; ++i) c2fPtr[i + 1] += c2fPtr[i];
    std::vector<int> c2f(c2fPtr[nC]), c2n(c2fPtr[nC]), fill(nC, 0);
    std::vector<char> isLeft(c2fPtr[nC]);
    for (int f = 0; f < nF; ++f) {
        const int le = m.left[f];
        int k = c2fPtr[le] + fill[le]++;
        c2f[k] = f; c2n[k] = m.right[f]; isLeft[k] = 1;
        if (f >= nB) {
            const int re = m.right[f];
            k = c2fPtr[re] + fill[re]++;
            c2f[k] = f; c2n[k] = le; isLeft[k] = 0;
        }
    }
    // Storage order of the interior faces: sorted by the storage positions of
    // their two cells, so that a run of faces reads a compact run of cells
    // and a cell's faces sit close together; boundary faces keep their
    // index (their ghost cell is nCell+f).  Each cell still accumulates its
    // faces in the mesh's face order: fslot only maps where a face's data live.
    std::vector<int> fslot(nF);
    for (int f = 0; f < nF; ++f) fslot[f] = f;
    if (p.storage_order == 0 && nF > nB) {
        std::vector<int> ord(nF - nB);
        for (int f = nB; f < nF; ++f) ord[f - nB] = f;
        auto key = [&](int f) {
            const int a = slotOf[m.left[f]], b = slotOf[m.right[f]];
            return std::make_pair(std::min(a, b), std::max(a, b));
        };
        std::stable_sort(ord.begin(), ord.end(), [&](int u, int v) { return key(u) < key(v); });
        for (int i = 0; i < nF - nB; ++i) fslot[ord[i]] = nB + i;
    }
    std::vector<int> c2fS(c2f.size());
    for (size_t k = 0; k < c2f.size(); ++k) c2fS[k] = fslot[c2f[k]];
    int W = 0;
    for (int i = 0; i < nC; ++i) W = std::max(W, c2fPtr[i + 1] - c2fPtr[i]);
    std::vector<int> ellEnc((size_t)W * nC, 0), ellNb((size_t)W * nC, 0);
    for (int i = 0; i < nC; ++i) {
        const int st = slotOf[i];
        for (int k = c2fPtr[i]; k < c2fPtr[i + 1]; ++k) {
            const int s = k - c2fPtr[i];
            ellEnc[(size_t)s * nC + st] = isLeft[k] ? (c2fS[k] + 1) : -(c2fS[k] + 1);
            ellNb[(size_t)s * nC + st] = S(c2n[k]);
        }
    }

    std::vector<phl::FaceRec> faces(nF);
    for (int f = 0; f < nF; ++f)
        faces[fslot[f]] = {m.xfn[f], m.yfn[f], m.zfn[f], m.area[f], m.xfc[f], m.yfc[f], m.zfc[f], 0.0};
    // Least-squares weights and inverse normal matrices (UnstructGrid::ComputeWeight,
    // Geo_UnstructGrid.cpp:3761-3940, 3-D, every face marked): per face
    // 1/(|x_R - x_L|^2 + SMALL), per cell the 3x3 normal matrix summed in
    // face order, inverted by cofactors (zero below SMALL).
    std::vector<double> lsqInvS;  // uploaded with the other cell arrays
    if (p.gradient == 1) {
        std::vector<double> inv((size_t)6 * nC, 0.0);
        double *ixx = inv.data(), *iyy = ixx + nC, *izz = iyy + nC, *ixy = izz + nC, *ixz = ixy + nC, *iyz = ixz + nC;
        for (int f = 0; f < nF; ++f) {
            const int le = m.left[f], re = m.right[f];
            const double dx = (m.xcc[re] - m.xcc[le]) * 1.0, dy = (m.ycc[re] - m.ycc[le]) * 1.0,
                         dz = (m.zcc[re] - m.zcc[le]) * 1.0;
            const double ods = 1.0 / (dx * dx + dy * dy + dz * dz + 1.0e-40);
            const double txx = ods * dx * dx, tyy = ods * dy * dy, tzz = ods * dz * dz;
            const double txy = ods * dx * dy, txz = ods * dx * dz, tyz = ods * dy * dz;
            faces[fslot[f]].lsqW = ods;
            ixx[le] += txx; iyy[le] += tyy; izz[le] += tzz; ixy[le] += txy; ixz[le] += txz; iyz[le] += tyz;
            if (f < nB) continue;
            ixx[re] += txx; iyy[re] += tyy; izz[re] += tzz; ixy[re] += txy; ixz[re] += txz; iyz[re] += tyz;
        }
        std::vector<double> invS((size_t)6 * nC);
        for (int i = 0; i < nC; ++i) {
            const double txx = ixx[i], tyy = iyy[i], tzz = izz[i], txy = ixy[i], txz = ixz[i], tyz = iyz[i];
            const double delt = (txx * tyy * tzz + 2.0 * txy * txz * tyz - txx * tyz * tyz - tyy * txz * txz -
                                 tzz * txy * txy);
            double o[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
            if (!(std::fabs(delt) < 1.0e-40)) {
                const double odelt = 1.0 / delt;
                o[0] = (tyy * tzz - tyz * tyz) * odelt;
                o[1] = (txx * tzz - txz * txz) * odelt;
                o[2] = (txx * tyy - txy * txy) * odelt;
                o[3] = (txz * tyz - txy * tzz) * odelt;
                o[4] = (txy * tyz - txz * tyy) * odelt;
                o[5] = (txy * txz - tyz * txx) * odelt;
            }
            for (int k = 0; k < 6; ++k) invS[(size_t)k * nC + slotOf[i]] = o[k];
        }
        lsqInvS = std::move(invS);
    }

    // Venkatakrishnan scale and CFL skewness factor (geometry only).
    std::vector<double> vf(nC), cflSkew(nC);
    for (int i = 0; i < nC; ++i) {
        vf[slotOf[i]] = std::pow(std::pow(m.vol[i], 1.0 / m.dim), 3.0);
        const double sk = m.skew[i];
        cflSkew[slotOf[i]] = (sk < 0.1) ? 0.1 : ((sk < 10.0) ? (1.0 - std::exp(0.0 - sk)) : 1.0);
    }
    std::vector<unsigned char> limZero(nC, 0);
    std::vector<double> limInit(nT, 0.0);
    for (int i = 0; i < nT; ++i) limInit[i] = (p.limiter == 2) ? 0.0 : 1.0;
    for (int f = 0; f < nB; ++f) {
        const int bc = m.bcType[f];
        const bool keep = bc == PHL_BC_SYMMETRY || bc == PHL_BC_SOLID_SURFACE || bc == PHL_BC_OUTFLOW;
        if (!keep && p.limiter == 0) {
            limZero[slotOf[m.left[f]]] = 1;
            limInit[nC + f] = 0.0;
        }
    }
    // Viscous face flags from the skewness angles (ComputeVisflux :13101-13113,13174,13209).
    std::vector<unsigned char> visFlag(nF, 0);
    const double kPi = 3.141592653589793, kSmall = 1.0e-40;
    for (int f = 0; f < nF; ++f) {
        const int le = m.left[f], re = m.right[f];
        const double dxL = m.xcc[le] - m.xfc[f], dyL = m.ycc[le] - m.yfc[f], dzL = m.zcc[le] - m.zfc[f];
        const double dxR = m.xcc[re] - m.xfc[f], dyR = m.ycc[re] - m.yfc[f], dzR = m.zcc[re] - m.zfc[f];
        const double dL = m.xfn[f] * dxL + m.yfn[f] * dyL + m.zfn[f] * dzL;
        const double dR = m.xfn[f] * dxR + m.yfn[f] * dyR + m.zfn[f] * dzR;
        double t = -dL / (std::sqrt(dxL * dxL + dyL * dyL + dzL * dzL) + kSmall);
        t = std::min(t, 1.0); t = std::max(t, -1.0);
        const double a1 = std::asin(t) * 180.0 / kPi;
        t = dR / (std::sqrt(dxR * dxR + dyR * dyR + dzR * dzR) + kSmall);
        t = std::min(t, 1.0); t = std::max(t, -1.0);
        const double a2 = std::asin(t) * 180.0 / kPi;
        unsigned char fl = 0;
        if (a1 > 60.0 && a2 > 60.0) fl |= 1;
        if (a1 > 0.0 && a2 > 0.0) fl |= 2;
        // turbulence diffusion (TurbSolverUnstr.cpp ComputeVisflux, turbSkewnessAngle)
        if (a1 > p.turb_skew_angle && a2 > p.turb_skew_angle) fl |= 4;
        visFlag[fslot[f]] = fl;
    }

    phl::DevMesh &dm = c->m;
    dm.nCell = nC; dm.nBFace = nB; dm.nFace = nF; dm.nTotal = nT; dm.ellW = W;
    // Tile-fused face loop (face_loop 1): the faces touching each tile of
    // kTileCells storage-ordered cells, and each cell's faces as slots of that
    // list.  Off when a tile's face values would not fit 64 KB of LDS.
    dm.tileStride = 0;
    dm.tileFacePtr = nullptr; dm.tileFaces = nullptr; dm.c2loc = nullptr;
    if (p.face_loop == 1) {
        const int TB = phl::kTileCells, nTiles = (nC + TB - 1) / TB;
        std::vector<int> tptr(nTiles + 1, 0), tfaces;
        std::vector<unsigned short> c2loc((size_t)W * nC, 0);
        std::vector<int> list;
        int maxF = 0;
        for (int t = 0; t < nTiles; ++t) {
            list.clear();
            const int c0 = t * TB, c1 = std::min(nC, c0 + TB);
            for (int i = c0; i < c1; ++i)
                for (int sl = 0; sl < W; ++sl) {
                    const int enc = ellEnc[(size_t)sl * nC + i];
                    if (enc == 0) break;
                    list.push_back((enc > 0 ? enc : -enc) - 1);
                }
            std::sort(list.begin(), list.end());
            list.erase(std::unique(list.begin(), list.end()), list.end());
            maxF = std::max(maxF, (int)list.size());
            for (int i = c0; i < c1; ++i)
                for (int sl = 0; sl < W; ++sl) {
                    const int enc = ellEnc[(size_t)sl * nC + i];
                    if (enc == 0) break;
                    const int f = (enc > 0 ? enc : -enc) - 1;
                    c2loc[(size_t)sl * nC + i] =
                        (unsigned short)(std::lower_bound(list.begin(), list.end(), f) - list.begin());
                }
            tfaces.insert(tfaces.end(), list.begin(), list.end());
            tptr[t + 1] = (int)tfaces.size();
        }
        const size_t ldsBytes = (size_t)maxF * 6 * sizeof(double) * (p.viscous ? 2 : 1);
        dm.tileStride = (p.face_loop == 1 && maxF < 65536 && ldsBytes <= 65536) ? maxF : 0;
        dm.tileFacePtr = nullptr; dm.tileFaces = nullptr; dm.c2loc = nullptr;
        if (dm.tileStride > 0) {
            TRY(upload(c, &dm.tileFacePtr, tptr));
            TRY(upload(c, &dm.tileFaces, tfaces));
            TRY(upload(c, &dm.c2loc, c2loc));
        }
    }
    TRY(upload(c, &dm.face, faces));
    TRY(upload(c, &dm.bcType, m.bcType));
    {
        std::vector<int> fle(nF), fre(nF);
        for (int f = 0; f < nF; ++f) { fle[fslot[f]] = S(m.left[f]); fre[fslot[f]] = S(m.right[f]); }
        TRY(upload(c, &dm.fle, fle));
        TRY(upload(c, &dm.fre, fre));
    }
    TRY(upload(c, &dm.c2fEnc, ellEnc));
    TRY(upload(c, &dm.c2n, ellNb));
    TRY(upload(c, &dm.xcc, permuteCells(m.xcc)));
    TRY(upload(c, &dm.ycc, permuteCells(m.ycc)));
    TRY(upload(c, &dm.zcc, permuteCells(m.zcc)));
    TRY(upload(c, &dm.vol, permuteCells(m.vol)));
    TRY(upload(c, &dm.logicalOf, cellAt));
    TRY(upload(c, &dm.cellAt, slotOf));
    TRY(upload(c, &dm.venkatVf, vf));
    dm.lsqInv = nullptr;
    if (!lsqInvS.empty()) TRY(upload(c, &dm.lsqInv, lsqInvS));
    TRY(upload(c, &dm.cflSkew, cflSkew));
    TRY(upload_node_lists(c, m, fslot));
    TRY(upload(c, &dm.limZero, limZero));
    dm.limGhost = nullptr;
    TRY(upload(c, &dm.visFlag, visFlag));
    std::vector<int> wall;
    c->caseWallPtr.assign(1, 0);
    for (int i = 0; i < m.nCases(); ++i) {
        const int b = m.nCases() > 1 ? m.caseBFacePtr[i] : 0, e = m.nCases() > 1 ? m.caseBFacePtr[i + 1] : nB;
        for (int f = b; f < e; ++f)
            if (m.bcType[f] == PHL_BC_SOLID_SURFACE) wall.push_back(f);
        c->caseWallPtr.push_back((int)wall.size());
    }
    dm.nWall = (int)wall.size();
    if (wall.empty()) wall.push_back(0);  // keeps the device pointer valid
    TRY(upload(c, &dm.wallFaces, wall));

    // LU-SGS level schedules
    std::vector<int> cellsF, cellsB;
    build_levels(m, c2fPtr, c2n, true, c->hfwdPtr, cellsF);
    build_levels(m, c2fPtr, c2n, false, c->hbwdPtr, cellsB);
    // Within a level, cells in storage order: the level kernels hand the
    // k-th eighth of a level to the k-th XCD, which then also holds (in its
    // L2) the same region's lower neighbours from the previous levels and the
    // residuals and diagonals its gather and diagonal kernels wrote.
    for (int dir = 0; dir < 2; ++dir) {
        std::vector<int> &cl = dir == 0 ? cellsF : cellsB;
        const std::vector<int> &ptr = dir == 0 ? c->hfwdPtr : c->hbwdPtr;
        for (size_t L = 0; L + 1 < ptr.size(); ++L)
            std::sort(cl.begin() + ptr[L], cl.begin() + ptr[L + 1], [&](int a, int b) { return slotOf[a] < slotOf[b]; });
    }
    std::vector<int> cellsFs(nC), cellsBs(nC);
    for (int i = 0; i < nC; ++i) { cellsFs[i] = slotOf[cellsF[i]]; cellsBs[i] = slotOf[cellsB[i]]; }
    const int *tmp;
    TRY(upload(c, &tmp, c->hfwdPtr)); c->fwd.ptr = const_cast<int *>(tmp);
    TRY(upload(c, &tmp, cellsFs)); c->fwd.cells = const_cast<int *>(tmp);
    TRY(upload(c, &tmp, c->hbwdPtr)); c->bwd.ptr = const_cast<int *>(tmp);
    TRY(upload(c, &tmp, cellsBs)); c->bwd.cells = const_cast<int *>(tmp);
    c->fwd.nLevels = (int)c->hfwdPtr.size() - 1; c->fwd.hptr = c->hfwdPtr.data();
    c->bwd.nLevels = (int)c->hbwdPtr.size() - 1; c->bwd.hptr = c->hbwdPtr.data();
    for (int dir = 0; dir < 2; ++dir) {
        phl::Levels &lv = dir == 0 ? c->fwd : c->bwd;
        std::vector<int> row, nbs;
        build_sweep_order(nC, c2fPtr, c2fS, c2n, isLeft, dir == 0 ? cellsF : cellsB, dir == 0, lv.so.W, lv.so.Wpad,
                          row, nbs);
        // the dataflow sweep addresses dq rows through a buffer descriptor
        // with a 32-bit byte range; larger meshes take the level form
        const bool rowsAddressable = (size_t)m.nTotal() * phl::kRow * sizeof(double) <= (size_t)INT32_MAX;
        if (p.sweep_form != 1 && lv.so.Wpad <= 6 && rowsAddressable)
            TRY(build_flow_order(c, nC, dir == 0 ? cellsF : cellsB, dir == 0 ? c->hfwdPtr : c->hbwdPtr, row, nbs,
                                 lv.so.Wpad, slotOf, lv.flow));
        for (int &v : nbs) v = S(v);  // unused slots hold 0, a valid storage index
        lv.so.cells = lv.cells;
        TRY(upload(c, &lv.so.row, row));
        TRY(upload(c, &lv.so.nb, nbs));
    }
    c->coefF32Ok = c->fwd.so.Wpad <= 6 && c->bwd.so.Wpad <= 6;
    c->saFused = p.viscous == 3 && p.n_lusgs_sweeps == 1 && c->fwd.flow.nChunks > 0 && c->bwd.flow.nChunks > 0;

    // constants: the scalars of every case, the first one also in k
    phl::Consts &k = c->k;
    const int nCases = m.nCases();
    if (c->caseParams.empty()) c->caseParams.assign(1, p);
    if ((int)c->caseParams.size() != nCases) return fail(PHL_ERR_ARG, "one parameter block per case is required");
    c->hostCases.resize(nCases);
    for (int i = 0; i < nCases; ++i)
        c->hostCases[i] = case_consts(c->caseParams[i], nCases > 1 ? m.caseAvgVol[i] : m.avgVol, m.dim);
    const phl::CaseConsts &c0 = c->hostCases[0];
    k.gamma = p.gamma;
    k.R = c0.R;
    k.mach = c0.mach;
    k.reynolds = p.reynolds;
    k.oRe = p.reynolds > 0.0 ? 1.0 / p.reynolds : 0.0;
    for (int i = 0; i < 5; ++i) k.primInf[i] = c->primInf[i] = c0.primInf[i];
    k.entropyCoef = c0.entropyCoef;
    k.epsTmp = c0.epsTmp;
    k.pLower = c0.pLower;
    k.pUpper = c0.pUpper;
    k.cflStart = p.cfl_start;
    k.ktmax = p.ktmax;
    k.tsuth = 110.4 / p.t_ref;
    k.cp = c0.cp;
    k.oprl = 1.0 / p.prandtl;
    k.oprt = 1.0 / p.prandtl_turb;
    k.densLimit = c0.densLimit;
    k.presLimit = c0.presLimit;
    k.nCases = nCases;
    k.cases = nullptr;
    dm.caseOf = nullptr;
    dm.caseCellPtr = nullptr;
    if (nCases > 1) {
        const phl::CaseConsts *dc = nullptr;
        TRY(upload(c, &dc, c->hostCases));
        k.cases = dc;
        std::vector<unsigned short> caseOf(nT);
        for (int i = 0; i < nCases; ++i) {
            // the storage order keeps every case inside its own range
            for (int cc = m.caseCellPtr[i]; cc < m.caseCellPtr[i + 1]; ++cc) caseOf[cc] = (unsigned short)i;
            for (int f = m.caseBFacePtr[i]; f < m.caseBFacePtr[i + 1]; ++f) caseOf[nC + f] = (unsigned short)i;
        }
        TRY(upload(c, &dm.caseOf, caseOf));
        TRY(upload(c, &dm.caseCellPtr, m.caseCellPtr));
        c->caseCellPtr = m.caseCellPtr;
        c->caseBFacePtr = m.caseBFacePtr;
    }
    k.viscous = p.viscous;
    k.limiter = p.limiter;
    k.entropyFix = p.entropy_fix;
    k.nSweeps = p.n_lusgs_sweeps;
    k.lusgsTol = p.lusgs_tol;
    k.twall = p.t_wall > 0.0 ? p.t_wall / p.t_ref : 0.0;
    k.mixed = p.precision;
    k.coefF32 = (k.mixed && c->coefF32Ok) ? 1 : 0;
    k.gradient = p.gradient;
    k.sweepForm = p.sweep_form;
    k.coefInSweep = (c->fwd.flow.nChunks > 0 && c->bwd.flow.nChunks > 0) ? 1 : 0;
    k.saNuInf = p.sa_nu_inf;
    k.saVistInf = p.sa_vist_inf;
    k.turbCflScale = p.turb_cfl_scale;

    // flow fields
    phl::DevFlow &fl = c->f;
    TRY(dalloc(c, &fl.q, (size_t)5 * nT));
    TRY(dalloc(c, &fl.qAlt, (size_t)5 * nT));
    c->qBuf0 = fl.q;
    TRY(dalloc(c, &fl.res, (size_t)5 * nC));
    TRY(dalloc(c, &fl.grad, (size_t)15 * nC));
    TRY(dalloc(c, &fl.gradT, (size_t)3 * nC));
    TRY(dalloc(c, &fl.gradF, (size_t)15 * nC));
    TRY(dalloc(c, &fl.gradTF, (size_t)3 * nC));
    TRY(dalloc(c, &fl.lim, (size_t)nT));
    // the pressure factor of the pressure-sensor entropy fixes (methods 2, 6)
    fl.rtem = nullptr;
    if (p.entropy_fix == 2 || p.entropy_fix == 6) {
        TRY(dalloc(c, &fl.rtem, (size_t)nT));
        HIPTRY(hipMemset(fl.rtem, 0, sizeof(double) * nT));
    }
    k.rtem = fl.rtem;
    TRY(dalloc(c, &fl.dq, (size_t)phl::kRow * nT));
    TRY(dalloc(c, &fl.dflux, (size_t)phl::kRow * nC));
    TRY(dalloc(c, &fl.specInv, (size_t)nC));
    TRY(dalloc(c, &fl.specVis, (size_t)nC));
    TRY(dalloc(c, &fl.dtLoc, (size_t)nC));
    TRY(dalloc(c, &fl.dt, (size_t)nC));
    TRY(dalloc(c, &fl.diag, (size_t)nC));
    k.fluxTile = 256;
    k.gradTile = 256;
    const int maxFluxBlocks = (nC + 63) / 64;  // smallest tile
    TRY(dalloc(c, &fl.normPart, (size_t)5 * maxFluxBlocks));
    c->normSlots = 1024;
    TRY(dalloc(c, &fl.normOut, (size_t)5 * c->normSlots * nCases));
    TRY(dalloc(c, &fl.dtMinPart, (size_t)maxFluxBlocks));
    TRY(dalloc(c, &fl.dtMin, (size_t)nCases));
    TRY(dalloc(c, &fl.frozen, (size_t)nCases));
    HIPTRY(hipMemset(fl.frozen, 0, (size_t)nCases));
    TRY(dalloc(c, &fl.stepCtr, 1));
    TRY(dalloc(c, &fl.cflArr, (size_t)c->normSlots));
    TRY(dalloc(c, &fl.negFlag, (size_t)nC));
    TRY(dalloc(c, &fl.negCount, 1));
    TRY(dalloc(c, &fl.negTotal, 1));
    // LU-SGS face coefficients, stored only for the level form (the dataflow
    // sweeps evaluate them)
    TRY(dalloc(c, &fl.coef, k.coefInSweep ? (size_t)16 : (size_t)nF * 2 * 16));
    const int nChunks = std::max({c->fwd.flow.nChunks, c->bwd.flow.nChunks, 1});
    TRY(dalloc(c, &fl.sweepFlag, (size_t)nChunks));
    HIPTRY(hipMemset(fl.sweepFlag, 0, sizeof(int) * nChunks));
    TRY(dalloc(c, &fl.casePart, (size_t)nCases * 32 * 6));
    TRY(dalloc(c, &fl.sweepCtl, 4));
    {
        const int ctl[4] = {0, 0, 0, kSweepWaitTicks};
        HIPTRY(hipMemcpy(fl.sweepCtl, ctl, sizeof(ctl), hipMemcpyHostToDevice));
    }
    if (p.n_lusgs_sweeps > 1) {
        TRY(dalloc(c, &fl.dqPrev, (size_t)5 * nC));
        TRY(dalloc(c, &fl.dqTPrev, (size_t)nC));
        TRY(dalloc(c, &fl.sweepParts, (size_t)(nC + 255) / 256));
        TRY(dalloc(c, &fl.sweepNorm, 2));
    }
    TRY(dalloc(c, &fl.negList, (size_t)phl::kFixList));
    TRY(dalloc(c, &fl.t, (size_t)nT));
    TRY(dalloc(c, &fl.visl, (size_t)nT));
    TRY(dalloc(c, &fl.forceOut, 15));
    TRY(dalloc(c, &fl.faceInv, (size_t)6 * nF));
    TRY(dalloc(c, &fl.faceVis, p.viscous ? (size_t)6 * nF : 1));
    TRY(dalloc(c, &fl.vist, (size_t)nT));
    HIPTRY(hipMemset(fl.vist, 0, sizeof(double) * nT));
    {
        // gather records: centroids and limiter start values; states and
        // gradients are written by the boundary and gradient kernels
        std::vector<phl::GatherRec> rec((size_t)nT);
        const std::vector<double> xs = permuteCells(m.xcc), ys = permuteCells(m.ycc), zs = permuteCells(m.zcc);
        for (int i = 0; i < nT; ++i) {
            phl::GatherRec &r = rec[i];
            std::memset(&r, 0, sizeof(r));
            r.x = xs[i]; r.y = ys[i]; r.z = zs[i];
            r.lim = limInit[i];
        }
        const phl::GatherRec *d = nullptr;
        TRY(upload(c, &d, rec));
        fl.rec = const_cast<phl::GatherRec *>(d);
    }
    dm.wallDist = nullptr;
    if (p.viscous == 3) {
        HIPTRY(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
        HIPTRY(hipEventCreateWithFlags(&c->evFork, hipEventDisableTiming));
        HIPTRY(hipEventCreateWithFlags(&c->evJoin, hipEventDisableTiming));
        TRY(dalloc(c, &fl.nut, (size_t)nT));
        TRY(dalloc(c, &fl.resT, (size_t)nT));
        TRY(dalloc(c, &fl.specT, (size_t)nC));
        TRY(dalloc(c, &fl.saMat, (size_t)2 * nF));
        TRY(dalloc(c, &fl.dqT, (size_t)nT));
        TRY(dalloc(c, &fl.dfluxT, (size_t)nC));
        TRY(dalloc(c, &fl.gradNut, (size_t)3 * nC));
        HIPTRY(hipMemset(fl.nut, 0, sizeof(double) * nT));
        HIPTRY(hipMemset(fl.resT, 0, sizeof(double) * nT));
        HIPTRY(hipMemset(fl.specT, 0, sizeof(double) * nC));
        HIPTRY(hipMemset(fl.saMat, 0, sizeof(double) * 2 * nF));
        HIPTRY(hipMemset(fl.dqT, 0, sizeof(double) * nT));
        HIPTRY(hipMemset(fl.dfluxT, 0, sizeof(double) * nC));
        HIPTRY(hipMemset(fl.gradNut, 0, sizeof(double) * 3 * nC));
        // wall distance: in 3-D the exact point-to-surface distance over the
        // wall faces, each split into (centre, node i, node i+1) triangles; in
        // 2-D the reference's bisection over each wall edge taken both ways
        // (start, centre, end), edges of squared length <= 1e-16 skipped
        // (Pre_WalldistCompute.cpp:437-474)
        std::vector<double> tri;
        for (int i = 0; i < dm.nWall; ++i) {
            const int f = wall[i];
            const int b = m.faceNodePtr[f], n = m.faceNodePtr[f + 1] - b;
            for (int j = 0; j < n; ++j) {
                const int n1 = m.faceNodes[b + j], n2 = m.faceNodes[b + (j + 1) % n];
                if (m.dim == 2) {
                    const double dx = m.x[n2] - m.x[n1], dy = m.y[n2] - m.y[n1], dz = m.z[n2] - m.z[n1];
                    if (dx * dx + dy * dy + dz * dz <= 1.0e-16) continue;
                    const double e9[9] = {m.x[n1], m.y[n1], m.z[n1], m.xfc[f], m.yfc[f], m.zfc[f],
                                          m.x[n2], m.y[n2], m.z[n2]};
                    tri.insert(tri.end(), e9, e9 + 9);
                    continue;
                }
                const double t9[9] = {m.xfc[f], m.yfc[f], m.zfc[f], m.x[n1], m.y[n1], m.z[n1],
                                      m.x[n2], m.y[n2], m.z[n2]};
                tri.insert(tri.end(), t9, t9 + 9);
            }
        }
        const int nTri = (int)(tri.size() / 9);
        if (tri.empty()) tri.assign(9, 0.0);
        const double *dTri = nullptr;
        TRY(upload(c, &dTri, tri));
        double *wd = nullptr;
        TRY(dalloc(c, &wd, (size_t)nC));
        phl::launch_wall_distance(dm.xcc, dm.ycc, dm.zcc, nC, dTri, nTri, m.dim == 2, wd, c->stream);
        HIPTRY(hipGetLastError());
        HIPTRY(hipStreamSynchronize(c->stream));
        dm.wallDist = wd;
    }
    HIPTRY(hipMemset(fl.q, 0, sizeof(double) * 5 * nT));
    HIPTRY(hipMemset(fl.qAlt, 0, sizeof(double) * 5 * nT));
    HIPTRY(hipMemset(fl.res, 0, sizeof(double) * 5 * nC));
    HIPTRY(hipMemset(fl.grad, 0, sizeof(double) * 15 * nC));
    HIPTRY(hipMemset(fl.gradT, 0, sizeof(double) * 3 * nC));
    HIPTRY(hipMemset(fl.dq, 0, sizeof(double) * phl::kRow * nT));
    HIPTRY(hipMemset(fl.dflux, 0, sizeof(double) * phl::kRow * nC));
    HIPTRY(hipMemset(fl.specVis, 0, sizeof(double) * nC));
    HIPTRY(hipMemset(fl.negFlag, 0, (size_t)nC));
    HIPTRY(hipMemset(fl.negCount, 0, sizeof(int)));
    HIPTRY(hipMemset(fl.negTotal, 0, sizeof(int)));
    HIPTRY(hipMemset(fl.stepCtr, 0, sizeof(int)));
    HIPTRY(hipMemcpy(fl.lim, limInit.data(), sizeof(double) * nT, hipMemcpyHostToDevice));  // ghosts only differ
    HIPTRY(hipHostMalloc((void **)&c->hostNorm, sizeof(double) * 5 * c->normSlots * nCases, hipHostMallocDefault));
    HIPTRY(hipHostMalloc((void **)&c->hostFrozen, (size_t)nCases, hipHostMallocDefault));
    std::fill(c->hostFrozen, c->hostFrozen + nCases, (unsigned char)0);
    HIPTRY(hipHostMalloc((void **)&c->hostCount, sizeof(int), hipHostMallocDefault));
    HIPTRY(hipHostMalloc((void **)&c->hostCfl, sizeof(double) * c->normSlots, hipHostMallocDefault));
    HIPTRY(hipHostMalloc((void **)&c->hostZero, sizeof(int), hipHostMallocDefault));
    *c->hostZero = 0;
    if (p.n_lusgs_sweeps > 1) {
        HIPTRY(hipHostMalloc((void **)&c->nsLog.host, 4 * sizeof(double), hipHostMallocDefault));
        c->saLog.host = c->nsLog.host + 2;
    }
    HIPTRY(hipDeviceSynchronize());
    return PHL_OK;
}

phl_status phl_create(const phl_mesh *pm, const phl_params *p, int device, phl_ctx **out)
{
    if (!pm || !p || !out) return fail(PHL_ERR_ARG, "phl_create: null argument");
    if (pm->m.nCases() > 1) return fail(PHL_ERR_ARG, "phl_create: batched mesh, use phl_create_batch");
    TRY(check_params(*p));
    TRY(check_mesh_params(pm->m, *p));
    int nDev = 0;
    HIPTRY(hipGetDeviceCount(&nDev));
    if (device < 0 || device >= nDev) return fail(PHL_ERR_ARG, "phl_create: no such device");
    phl_ctx *c = new (std::nothrow) phl_ctx;
    if (!c) return fail(PHL_ERR_ALLOC, "phl_create: out of memory");
    c->device = device;
    c->p = *p;
    phl_status s = create_impl(c, pm->m);
    if (s != PHL_OK) {
        std::string keep = g_err;
        phl_destroy(c);
        g_err = keep;
        return s;
    }
    *out = c;
    return PHL_OK;
}

phl_status phl_create_batch(const phl_mesh *pm, const phl_params *params, int device, phl_ctx **out)
{
    if (!pm || !params || !out) return fail(PHL_ERR_ARG, "phl_create_batch: null argument");
    const int n = pm->m.nCases();
    // the cases share every option; only the freestream (Mach, angles) varies
    for (int i = 0; i < n; ++i) {
        TRY(check_params(params[i]));
        TRY(check_mesh_params(pm->m, params[i]));
        if (params[i].viscous == 3)
            return fail(PHL_ERR_UNSUPPORTED, "phl_create_batch: the Spalart-Allmaras model runs on single-case contexts");
        if (params[i].n_lusgs_sweeps != 1)
            return fail(PHL_ERR_UNSUPPORTED, "phl_create_batch: several LU-SGS sweep pairs run on single-case contexts");
        phl_params a = params[0], b = params[i];
        a.mach = b.mach = 0.0;
        a.alpha_deg = b.alpha_deg = 0.0;
        a.beta_deg = b.beta_deg = 0.0;
        if (std::memcmp(&a, &b, sizeof(phl_params)) != 0)
            return fail(PHL_ERR_ARG, "phl_create_batch: cases may differ only in mach, alpha_deg, beta_deg");
    }
    int nDev = 0;
    HIPTRY(hipGetDeviceCount(&nDev));
    if (device < 0 || device >= nDev) return fail(PHL_ERR_ARG, "phl_create_batch: no such device");
    phl_ctx *c = new (std::nothrow) phl_ctx;
    if (!c) return fail(PHL_ERR_ALLOC, "phl_create_batch: out of memory");
    c->device = device;
    c->p = params[0];
    c->caseParams.assign(params, params + n);
    phl_status s = create_impl(c, pm->m);
    if (s != PHL_OK) {
        std::string keep = g_err;
        phl_destroy(c);
        g_err = keep;
        return s;
    }
    *out = c;
    return PHL_OK;
}

// ---------------------------------------------------------------- state
// A new state makes the context valid again after a sweep gave up waiting:
// the count of waves that gave up is cleared with the state.
static phl_status clear_gave_up(phl_ctx *c)
{
    HIPTRY(hipMemsetAsync(c->f.sweepCtl + 2, 0, sizeof(int), c->stream));
    HIPTRY(hipStreamSynchronize(c->stream));
    return PHL_OK;
}

phl_status phl_set_state(phl_ctx *c, const double *q)
{
    if (!c || !q) return fail(PHL_ERR_ARG, "null argument");
    HIPTRY(hipSetDevice(c->device));
    ++c->stateVersion;
    c->tFresh = true;
    TRY(upload_cells(c, c->f.q, 5, c->m.nTotal, q));
    return clear_gave_up(c);
}

// ggnode: before a node pass, the temperature of a state set since the last
// one (the state's ghosts as they are).
static void fresh_temperature(phl_ctx *c, hipStream_t s)
{
    if (c->k.gradient == 2 && c->tFresh) {
        phl::launch_gamma_t(c->m, c->k, c->f, s);
        c->tFresh = false;
    }
}

// ComputeNodeValue + ModifyNodeValue (NSSolverUnstruct.cpp:4734-5117;
// CallGPUComputeNodeValue :4738, CallGPUModifyNodeValue :5087).
phl_status phl_compute_node_value(phl_ctx *c)
{
    if (!c) return fail(PHL_ERR_ARG, "null argument");
    if (c->k.gradient != 2) return fail(PHL_ERR_ARG, "phl_compute_node_value: gradient is not ggnode");
    HIPTRY(hipSetDevice(c->device));
    fresh_temperature(c, c->stream);
    phl::launch_node_values(c->m, c->f, c->stream);
    return check_launch(c);
}

phl_status phl_get_node_values(phl_ctx *c, double *out)
{
    if (!c || !out) return fail(PHL_ERR_ARG, "null argument");
    if (c->k.gradient != 2) return fail(PHL_ERR_ARG, "phl_get_node_values: gradient is not ggnode");
    HIPTRY(hipSetDevice(c->device));
    HIPTRY(hipMemcpyAsync(out, c->f.nodeQ, sizeof(double) * 6 * c->m.nNode, hipMemcpyDeviceToHost, c->stream));
    HIPTRY(hipStreamSynchronize(c->stream));
    return PHL_OK;
}

phl_status phl_get_state(phl_ctx *c, double *q)
{
    if (!c || !q) return fail(PHL_ERR_ARG, "null argument");
    HIPTRY(hipSetDevice(c->device));
    return copy_field(c, c->f.q, 5, c->m.nTotal, c->m.nTotal, q);
}

phl_status phl_get_residual(phl_ctx *c, double *res)
{
    if (!c || !res) return fail(PHL_ERR_ARG, "null argument");
    HIPTRY(hipSetDevice(c->device));
    return copy_field(c, c->f.res, 5, c->m.nCell, c->m.nTotal, res);
}

phl_status phl_get_field(phl_ctx *c, int field, double *out)
{
    if (!c || !out) return fail(PHL_ERR_ARG, "null argument");
    HIPTRY(hipSetDevice(c->device));
    const int nC = c->m.nCell, nT = c->m.nTotal;
    phl::DevFlow &f = c->f;
    switch (field) {
    case PHL_FIELD_Q: return copy_field(c, f.q, 5, nT, nT, out);
    case PHL_FIELD_RES: return copy_field(c, f.res, 5, nC, nT, out);
    case PHL_FIELD_DQDX:
    case PHL_FIELD_DQDY:
    case PHL_FIELD_DQDZ: {
        const size_t off = (size_t)5 * (field - PHL_FIELD_DQDX) * nC;
        phl::launch_rec_grad(c->m, c->k, f, c->stream);
        if (c->k.mixed) return copy_field_f32(c, f.gradF + off, 5, nC, nT, out);
        return copy_field(c, f.grad + off, 5, nC, nT, out);
    }
    case PHL_FIELD_LIMIT: return copy_field(c, f.lim, 1, nT, nT, out);
    case PHL_FIELD_DQ: return copy_rows(c, f.dq, nT, nT, out);
    case PHL_FIELD_SPEC_INV: return copy_field(c, f.specInv, 1, nC, nT, out);
    case PHL_FIELD_SPEC_VIS: return copy_field(c, f.specVis, 1, nC, nT, out);
    case PHL_FIELD_DT: return copy_field(c, f.dt, 1, nC, nT, out);
    case PHL_FIELD_DIAG: return copy_field(c, f.diag, 1, nC, nT, out);
    case PHL_FIELD_T:
    case PHL_FIELD_VISL:
        phl::launch_gamma_t(c->m, c->k, f, c->stream);
        TRY(check_launch(c));
        return copy_field(c, field == PHL_FIELD_T ? f.t : f.visl, 1, nT, nT, out);
    case PHL_FIELD_DTDX:
    case PHL_FIELD_DTDY:
    case PHL_FIELD_DTDZ: {
        const size_t off = (size_t)(field - PHL_FIELD_DTDX) * nC;
        if (c->k.mixed) return copy_field_f32(c, f.gradTF + off, 1, nC, nT, out);
        return copy_field(c, f.gradT + off, 1, nC, nT, out);
    }
    case PHL_FIELD_DELTA_FLUX: return copy_rows(c, f.dflux, nC, nT, out);
    case PHL_FIELD_VIST: return copy_field(c, f.vist, 1, nT, nT, out);
    case PHL_FIELD_PRESSURE_FACTOR:
        if (!f.rtem) {
            std::fill(out, out + nT, 0.0);
            return PHL_OK;
        }
        return copy_field(c, f.rtem, 1, nT, nT, out);
    default: break;
    }
    if (field >= PHL_FIELD_NUT && field <= PHL_FIELD_DNUDZ) {
        if (c->k.viscous != 3) return fail(PHL_ERR_ARG, "phl_get_field: turbulence fields need viscous = 3");
        switch (field) {
        case PHL_FIELD_NUT: return copy_field(c, f.nut, 1, nT, nT, out);
        case PHL_FIELD_RES_T: return copy_field(c, f.resT, 1, nC, nT, out);
        case PHL_FIELD_SPEC_T: return copy_field(c, f.specT, 1, nC, nT, out);
        case PHL_FIELD_DQ_T: return copy_field(c, f.dqT, 1, nT, nT, out);
        case PHL_FIELD_WDIST: return copy_field(c, c->m.wallDist, 1, nC, nT, out);
        default:  // a relaxation does not need the gradient of nu~: compute it from the current field
            if (c->k.gradient != 0)
                return fail(PHL_ERR_UNSUPPORTED, "the gradient of nu~ is implemented for ggcell only");
            phl::launch_sa_gradient(c->m, c->k, f, c->stream);
            TRY(check_launch(c));
            return copy_field(c, f.gradNut + (size_t)(field - PHL_FIELD_DNUDX) * nC, 1, nC, nT, out);
        }
    }
    return fail(PHL_ERR_ARG, "phl_get_field: unknown field");
}

phl_status phl_set_field(phl_ctx *c, int field, const double *in)
{
    if (!c || !in) return fail(PHL_ERR_ARG, "null argument");
    HIPTRY(hipSetDevice(c->device));
    const int nC = c->m.nCell, nT = c->m.nTotal;
    double *dst = nullptr;
    size_t n = 0;
    switch (field) {
    case PHL_FIELD_Q: return phl_set_state(c, in);
    case PHL_FIELD_DT: dst = c->f.dt; n = (size_t)nC; break;
    case PHL_FIELD_VIST: dst = c->f.vist; n = (size_t)nT; ++c->stateVersion; break;
    case PHL_FIELD_NUT:
    case PHL_FIELD_WDIST:
        if (c->k.viscous != 3) return fail(PHL_ERR_ARG, "phl_set_field: turbulence fields need viscous = 3");
        dst = field == PHL_FIELD_NUT ? c->f.nut : const_cast<double *>(c->m.wallDist);
        n = field == PHL_FIELD_NUT ? (size_t)nT : (size_t)nC;
        break;
    default: return fail(PHL_ERR_ARG, "phl_set_field: field is not settable");
    }
    return upload_cells(c, dst, 1, (int)n, in);
}

phl_status phl_init_freestream(phl_ctx *c)
{
    if (!c) return fail(PHL_ERR_ARG, "null argument");
    HIPTRY(hipSetDevice(c->device));
    c->tFresh = true;
    const int nT = c->m.nTotal, nC = c->m.nCell;
    std::vector<double> h((size_t)5 * nT);
    if (c->k.nCases <= 1) {
        for (int m = 0; m < 5; ++m)
            std::fill(h.begin() + (size_t)m * nT, h.begin() + (size_t)(m + 1) * nT, c->primInf[m]);
    } else {
        for (int i = 0; i < c->k.nCases; ++i)
            for (int m = 0; m < 5; ++m) {
                double *row = h.data() + (size_t)m * nT;
                const double v = c->hostCases[i].primInf[m];
                std::fill(row + c->caseCellPtr[i], row + c->caseCellPtr[i + 1], v);
                std::fill(row + nC + c->caseBFacePtr[i], row + nC + c->caseBFacePtr[i + 1], v);
            }
    }
    HIPTRY(hipMemsetAsync(c->f.vist, 0, sizeof(double) * nT, c->stream));
    if (c->k.viscous == 3) {
        std::vector<double> nut((size_t)nT, c->k.saNuInf);
        HIPTRY(hipMemcpyAsync(c->f.nut, nut.data(), sizeof(double) * nT, hipMemcpyHostToDevice, c->stream));
        HIPTRY(hipStreamSynchronize(c->stream));
    }
    return phl_set_state(c, h.data());
}

// ----------------------------------------------------------------- path
phl_status phl_boundary(phl_ctx *c)
{
    if (!c) return fail(PHL_ERR_ARG, "null argument");
    HIPTRY(hipSetDevice(c->device));
    ++c->stateVersion;  // ghost states
    phl::launch_boundary(c->m, c->k, c->f, c->stream);
    return check_launch(c);
}

phl_status phl_gamma_t(phl_ctx *c)
{
    if (!c) return fail(PHL_ERR_ARG, "null argument");
    HIPTRY(hipSetDevice(c->device));
    phl::launch_gamma_t(c->m, c->k, c->f, c->stream);
    return check_launch(c);
}

phl_status phl_pressure_factor(phl_ctx *c)
{
    if (!c) return fail(PHL_ERR_ARG, "null argument");
    HIPTRY(hipSetDevice(c->device));
    phl::launch_pressure_factor(c->m, c->k, c->f, c->stream);
    return check_launch(c);
}

phl_status phl_gradient(phl_ctx *c)
{
    if (!c) return fail(PHL_ERR_ARG, "null argument");
    HIPTRY(hipSetDevice(c->device));
    phl::launch_grad_lim(c->m, c->k, c->f, c->stream);
    c->gradVersion = c->stateVersion;
    return check_launch(c);
}

// The gradient kernel evaluates the limiter in the same pass (both read the
// cell's faces and neighbours once); on a state phl_gradient has just seen,
// the limiter is already in place.
phl_status phl_limiter(phl_ctx *c)
{
    if (!c) return fail(PHL_ERR_ARG, "null argument");
    if (c->gradVersion == c->stateVersion) return PHL_OK;
    return phl_gradient(c);
}

static void reset_step_counter(phl_ctx *c)
{
    hipMemcpyAsync(c->f.stepCtr, c->hostZero, sizeof(int), hipMemcpyHostToDevice, c->stream);
}

phl_status phl_load_residuals(phl_ctx *c)
{
    if (!c) return fail(PHL_ERR_ARG, "null argument");
    HIPTRY(hipSetDevice(c->device));
    phl::launch_load_residuals(c->m, c->f, c->stream);
    return check_launch(c);
}

// Face passes of the hook form: they read the gather records that the
// boundary and gradient kernels refreshed and add to the residual in res.
static phl_status face_pass(phl_ctx *c, int parts)
{
    if (!c) return fail(PHL_ERR_ARG, "null argument");
    HIPTRY(hipSetDevice(c->device));
    phl::launch_face_pass(c->m, c->k, c->f, parts, c->stream);
    return check_launch(c);
}

phl_status phl_inviscid_flux(phl_ctx *c) { return face_pass(c, 1); }
phl_status phl_viscous_flux(phl_ctx *c) { return face_pass(c, 2); }
phl_status phl_update_residuals(phl_ctx *c) { return face_pass(c, 3); }

// The fused residual evaluation: LoadResiduals, Boundary, the gradient and
// limiter pass, and both face loops in one face kernel + one gather, with the
// residual sums of the result.
// ggnode, after the boundary conditions: the node values (the state and the
// previous iteration's temperature), then this iteration's temperature.
static void node_pass(phl_ctx *c, hipStream_t s)
{
    if (c->k.gradient != 2) return;
    fresh_temperature(c, s);
    phl::launch_node_values(c->m, c->f, s);
    phl::launch_gamma_t(c->m, c->k, c->f, s);
}

static void enqueue_residual(phl_ctx *c)
{
    reset_step_counter(c);
    phl::launch_boundary(c->m, c->k, c->f, c->stream);
    node_pass(c, c->stream);
    // UpdateIndependentField's laminar viscosity, which the turbulence hooks
    // read (the NS kernels evaluate it themselves)
    if (c->k.viscous == 3) phl::launch_gamma_t(c->m, c->k, c->f, c->stream);
    phl::launch_pressure_factor(c->m, c->k, c->f, c->stream);
    phl::launch_grad_lim(c->m, c->k, c->f, c->stream);
    phl::launch_flux(c->m, c->k, c->f, c->stream, false);
    phl::launch_norm_final(c->m, c->k, c->f, false, c->stream);
}

phl_status phl_residual(phl_ctx *c)
{
    if (!c) return fail(PHL_ERR_ARG, "null argument");
    HIPTRY(hipSetDevice(c->device));
    enqueue_residual(c);
    c->gradVersion = c->stateVersion;
    return check_launch(c);
}

phl_status phl_spectrum_radius(phl_ctx *c)
{
    if (!c) return fail(PHL_ERR_ARG, "null argument");
    HIPTRY(hipSetDevice(c->device));
    phl::launch_spectrum(c->m, c->k, c->f, c->stream);
    return check_launch(c);
}

phl_status phl_time_step(phl_ctx *c, int step)
{
    if (c && c->k.nCases > 1) return fail(PHL_ERR_UNSUPPORTED, "not available on a batched context");
    if (!c) return fail(PHL_ERR_ARG, "null argument");
    HIPTRY(hipSetDevice(c->device));
    phl::launch_time_step(c->m, c->k, c->f, compute_cfl(c->p, step), c->stream);
    phl::launch_dt_diag(c->m, c->k, c->f, c->stream);
    return check_launch(c);
}

// Controller::LUSGS with nLUSGSSweeps > 1 (Controller.cpp:1878-1961): dq
// and LUplusDQ start from zero; after every sweep pair the squared change of
// dq over both half-sweeps (the reference's sweepNormal, summed here in block
// order) is read back, and the pairs stop once it falls below
// LUSGSTolerance times the first pair's.  The host decides, so such steps
// run eagerly.
static void ns_sweep_pairs(phl_ctx *c, hipStream_t s)
{
    phl::DevFlow &f = c->f;
    const int nT = c->m.nTotal, nC = c->m.nCell;
    hipMemsetAsync(f.dq, 0, sizeof(double) * phl::kRow * nT, s);
    hipMemsetAsync(f.dflux, 0, sizeof(double) * phl::kRow * nC, s);
    hipMemsetAsync(f.dqPrev, 0, sizeof(double) * 5 * nC, s);
    double init = 0.0, eps = 1.1;
    c->nsLog.done = 0;
    for (int it = 0; it < c->k.nSweeps; ++it) {
        phl::launch_sweep(c->m, c->k, f, c->fwd, true, it == 0, s);
        phl::launch_sweep_change(f.dq, phl::kRow, 1, 5, nC, f.dqPrev, f.sweepParts, f.sweepNorm, s);
        phl::launch_ghost_dq(c->m, c->k, f, s);
        phl::launch_sweep(c->m, c->k, f, c->bwd, false, false, s);
        phl::launch_sweep_change(f.dq, phl::kRow, 1, 5, nC, f.dqPrev, f.sweepParts, f.sweepNorm + 1, s);
        hipMemcpyAsync(c->nsLog.host, f.sweepNorm, 2 * sizeof(double), hipMemcpyDeviceToHost, s);
        if (hipStreamSynchronize(s) != hipSuccess) return;  // reported by the caller's launch check
        const double sn = c->nsLog.host[0] + c->nsLog.host[1];
        if (it < 8) c->nsLog.normals[it] = sn;
        c->nsLog.done = it + 1;
        if (it == 0) init = sn;
        else eps = sn / init;
        if (eps < c->k.lusgsTol) break;
    }
}

static void enqueue_lusgs(phl_ctx *c)
{
    if (c->k.nSweeps > 1) {
        ns_sweep_pairs(c, c->stream);
    } else {
        phl::launch_sweep(c->m, c->k, c->f, c->fwd, true, true, c->stream);
        phl::launch_ghost_dq(c->m, c->k, c->f, c->stream);
        phl::launch_sweep(c->m, c->k, c->f, c->bwd, false, false, c->stream);
    }
    phl::launch_update(c->m, c->k, c->f, false, c->stream);
    swap_state(c);
}

phl_status phl_sweep_normals(const phl_ctx *c, int turb, int *n_done, double *normals8)
{
    if (!c || !n_done) return fail(PHL_ERR_ARG, "null argument");
    const phl::SweepLog &l = turb ? c->saLog : c->nsLog;
    *n_done = c->k.nSweeps > 1 ? l.done : 0;
    if (normals8)
        for (int i = 0; i < 8; ++i) normals8[i] = l.normals[i];
    return PHL_OK;
}

phl_status phl_set_init_phase(phl_ctx *c, int on)
{
    if (!c) return fail(PHL_ERR_ARG, "null argument");
    c->hookInit = on ? 1 : 0;
    c->k.initPhase = c->hookInit;
    ++c->stateVersion;  // the limiter a gradient call leaves depends on it
    return PHL_OK;
}

phl_status phl_lusgs(phl_ctx *c, int *n_fixed)
{
    if (c && c->k.nCases > 1) return fail(PHL_ERR_UNSUPPORTED, "not available on a batched context");
    if (!c) return fail(PHL_ERR_ARG, "null argument");
    HIPTRY(hipSetDevice(c->device));
    HIPTRY(hipMemsetAsync(c->f.negCount, 0, sizeof(int), c->stream));
    ++c->stateVersion;
    enqueue_lusgs(c);
    TRY(check_launch(c));
    TRY(check_sweep_waits(c));
    if (n_fixed) {
        HIPTRY(hipMemcpyAsync(c->hostCount, c->f.negCount, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPTRY(hipStreamSynchronize(c->stream));
        *n_fixed = *c->hostCount;
    }
    return PHL_OK;
}

// ----------------------------------------------- Spalart-Allmaras hooks
static phl_status sa_ctx(phl_ctx *c)
{
    if (!c) return fail(PHL_ERR_ARG, "null argument");
    if (c->k.viscous != 3) return fail(PHL_ERR_ARG, "the turbulence hooks need viscous = 3");
    HIPTRY(hipSetDevice(c->device));
    return PHL_OK;
}

phl_status phl_turb_boundary(phl_ctx *c)
{
    TRY(sa_ctx(c));
    phl::launch_sa_boundary(c->m, c->k, c->f, c->stream);
    return check_launch(c);
}

phl_status phl_turb_viscosity(phl_ctx *c)
{
    TRY(sa_ctx(c));
    phl::launch_sa_viscosity(c->m, c->k, c->f, c->stream);
    return check_launch(c);
}

phl_status phl_turb_gradient(phl_ctx *c)
{
    TRY(sa_ctx(c));
    // TurbSolverUnstr::GetGradientField (:8652-8700) with gradientName; the
    // NSMB formation reads no gradient of nu~, so only ggcell is restated
    if (c->k.gradient != 0) return fail(PHL_ERR_UNSUPPORTED, "the gradient of nu~ is implemented for ggcell only");
    phl::launch_sa_gradient(c->m, c->k, c->f, c->stream);
    return check_launch(c);
}

phl_status phl_turb_update_residuals(phl_ctx *c)
{
    TRY(sa_ctx(c));
    phl::launch_sa_residual(c->m, c->k, c->f, c->stream);
    return check_launch(c);
}

phl_status phl_turb_spectrum(phl_ctx *c)
{
    TRY(sa_ctx(c));
    phl::launch_sa_spectrum(c->m, c->k, c->f, c->stream);
    return check_launch(c);
}

phl_status phl_turb_lusgs(phl_ctx *c)
{
    TRY(sa_ctx(c));
    phl::launch_sa_lusgs(c->m, c->k, c->f, c->fwd, c->bwd, c->stream, &c->saLog);
    return check_launch(c);
}

phl_status phl_turb_relax(phl_ctx *c)
{
    TRY(sa_ctx(c));
    phl::launch_sa_relax(c->m, c->k, c->f, c->fwd, c->bwd, c->stream, &c->saLog);
    return check_launch(c);
}

phl_status phl_residual_norm(phl_ctx *c, double *sum5)
{
    if (!c || !sum5) return fail(PHL_ERR_ARG, "null argument");
    HIPTRY(hipSetDevice(c->device));
    reset_step_counter(c);
    phl::launch_residual_norm(c->m, c->k, c->f, c->stream);
    TRY(check_launch(c));
    HIPTRY(hipMemcpyAsync(c->hostNorm, c->f.normOut, sizeof(double) * 5, hipMemcpyDeviceToHost, c->stream));
    HIPTRY(hipStreamSynchronize(c->stream));
    for (int m = 0; m < 5; ++m) sum5[m] = c->hostNorm[m];
    return PHL_OK;
}

phl_status phl_residual_norm_noalloc(const double *res, int n_cell, int ld, double *sum5, void *stream)
{
    static_assert(phl::kNoAllocMaxCells == PHL_NOALLOC_MAX_CELLS, "scratch size and header disagree");
    if (!res || !sum5) return fail(PHL_ERR_ARG, "phl_residual_norm_noalloc: null argument");
    if (n_cell < 1 || ld < n_cell) return fail(PHL_ERR_ARG, "phl_residual_norm_noalloc: need 1 <= n_cell <= ld");
    if (n_cell > PHL_NOALLOC_MAX_CELLS)
        return fail(PHL_ERR_ARG, "phl_residual_norm_noalloc: " + std::to_string(n_cell) +
                                     " cells exceed the static scratch array's 16,777,216");
    double *part = nullptr;
    HIPTRY(phl::noalloc_scratch(&part));
    phl::launch_residual_norm_noalloc(res, n_cell, (size_t)ld, sum5, part, static_cast<hipStream_t>(stream));
    HIPTRY(hipGetLastError());
    return PHL_OK;
}

// Force::CompForce (API/Physics/src/Force.cpp:439-533) for the all-wall
// part: scale the sums by the reference area and lengths, then lift, drag,
// side force, pressure and friction drag, induced-drag estimate and centre
// of pressure.
static void force_coefficients(const phl_params &p, const double *sums, double *out)
{
    const double kPiD = 3.141592653589793;
    const double scf = 1.0 / p.force_ref_area;
    const double vcm = 1.0 / (p.force_ref_area * p.force_ref_length);
    const double sb = 1.0 / (p.force_ref_area * p.force_ref_span);
    for (int i = 0; i < 9; ++i) out[i] = sums[i] * scf;
    for (int i = 9; i < 13; ++i) out[i] = sums[i] * sb;
    for (int i = 13; i < 15; ++i) out[i] = sums[i] * vcm;
    const double attack = p.alpha_deg * kPiD / 180.0, sideslip = p.beta_deg * kPiD / 180.0;
    const double sina = std::sin(attack), cosa = std::cos(attack);
    const double sinb = std::sin(sideslip), cosb = std::cos(sideslip);
    const double cfx = out[0] + out[1], cfy = out[2] + out[3], cfz = out[4] + out[5];
    const double cmz = out[13] + out[14];
    const double cl = -sina * cfx + cosa * cfy;
    const double cd = cosa * cosb * cfx + sina * cosb * cfy + sinb * cfz;
    const double cdp = cosa * cosb * out[6] + sina * cosb * out[7] + sinb * out[8];
    const double AR = 9.5;
    out[PHL_FORCE_CL] = cl;
    out[PHL_FORCE_CD] = cd;
    out[PHL_FORCE_SIDE] = -cosa * sinb * cfx - sina * sinb * cfy + cosb * cfz;
    out[PHL_FORCE_CD_PR] = cdp;
    out[PHL_FORCE_CD_SF] = cd - cdp;
    out[PHL_FORCE_CD_CL2PA] = cd - (cl * cl) / (kPiD * AR);
    out[PHL_FORCE_XCP] = (cfy < 0.0 ? -1.0 : 1.0) * cmz / (std::fabs(cfy) + 1.0e-40);
}

phl_status phl_forces_case(phl_ctx *c, int cs, double *coef)
{
    if (!c || !coef) return fail(PHL_ERR_ARG, "null argument");
    if (cs < 0 || cs >= c->k.nCases) return fail(PHL_ERR_ARG, "phl_forces_case: no such case");
    HIPTRY(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    phl::launch_boundary(c->m, c->k, c->f, s);
    phl::launch_gamma_t(c->m, c->k, c->f, s);
    if (c->k.gradient == 2) {  // node values of the current state and its temperature
        c->tFresh = false;
        phl::launch_node_values(c->m, c->f, s);
    }
    phl::launch_grad_lim(c->m, c->k, c->f, s, true);
    // the case's wall faces are a contiguous run of the sorted wall list
    phl::DevMesh dm = c->m;
    phl::Consts k = c->k;
    const phl_params &pc = c->caseParams.empty() ? c->p : c->caseParams[cs];
    if (c->k.nCases > 1) {
        dm.wallFaces = c->m.wallFaces + c->caseWallPtr[cs];
        dm.nWall = c->caseWallPtr[cs + 1] - c->caseWallPtr[cs];
        for (int i = 0; i < 5; ++i) k.primInf[i] = c->hostCases[cs].primInf[i];
    }
    const double ref[3] = {pc.force_ref_x, pc.force_ref_y, pc.force_ref_z};
    phl::launch_forces(dm, k, c->f, ref, s);
    HIPTRY(hipGetLastError());
    HIPTRY(hipMemcpyAsync(c->hostNorm, c->f.forceOut, sizeof(double) * 15, hipMemcpyDeviceToHost, s));
    HIPTRY(hipStreamSynchronize(s));
    force_coefficients(pc, c->hostNorm, coef);
    return PHL_OK;
}

phl_status phl_forces(phl_ctx *c, double *coef) { return phl_forces_case(c, 0, coef); }

// ----------------------------------------------------------- host driver
static const int kTimedChunk = 64;

// Make the solver stream wait for turbulence sweeps still running on the
// side stream.
static void join_turb(phl_ctx *c)
{
    if (c->turbPending) {
        hipStreamWaitEvent(c->stream, c->evJoin, 0);
        c->turbPending = false;
    }
}

// Finish a turbulence relaxation whose sweeps were left for the next step
// (one launch per level, on the solver stream).
static void flush_sa(phl_ctx *c)
{
    if (c->saDeferred) {
        phl::launch_sa_lusgs(c->m, c->k, c->f, c->fwd, c->bwd, c->stream);
        c->saDeferred = false;
    }
}

// One outer iteration: LoadResiduals+UpdateResiduals, SpectrumRadius,
// TimeStep, LUSGS (Controller::Relaxation, Controller.cpp:567-608), then the
// turbulence relaxation.  With `ev` non-null an event is recorded before the
// first and after every kernel class (everything on the solver stream).
static void enqueue_step(phl_ctx *c, hipEvent_t *ev)
{
    hipStream_t s = c->stream;
    int e = 0;
    auto mark = [&]() {
        if (ev) hipEventRecord(ev[e++], s);
    };
    mark();
    phl::launch_boundary(c->m, c->k, c->f, s);
    node_pass(c, s);
    // the turbulence model reads the laminar viscosity of the step's state
    if (c->k.viscous == 3) phl::launch_gamma_t(c->m, c->k, c->f, s);
    phl::launch_pressure_factor(c->m, c->k, c->f, s);
    mark();
    phl::launch_grad_lim(c->m, c->k, c->f, s); mark();
    phl::launch_flux(c->m, c->k, c->f, s, true); mark();
    phl::launch_norm_final(c->m, c->k, c->f, true, s); mark();
    phl::launch_dt_diag(c->m, c->k, c->f, s); mark();
    // the previous step's turbulence sweeps, carried by this step's sweeps:
    // they read only the turbulence arrays, which this step's NS kernels so
    // far have not touched, and write only nu~'s increment
    const bool carry = c->saFused && c->saDeferred;
    if (c->k.nSweeps > 1) {
        ns_sweep_pairs(c, s); mark(); mark(); mark();
    } else {
        if (carry) phl::launch_sa_sweep_init(c->m, c->f, s);
        phl::launch_sweep(c->m, c->k, c->f, c->fwd, true, true, s, carry); mark();
        phl::launch_ghost_dq(c->m, c->k, c->f, s);
        if (carry) phl::launch_sa_ghost_dq(c->m, c->f, s);
        mark();
        phl::launch_sweep(c->m, c->k, c->f, c->bwd, false, false, s, carry); mark();
    }
    phl::launch_update(c->m, c->k, c->f, true, s);
    if (carry) {
        phl::launch_sa_update(c->m, c->f, s);
        c->saDeferred = false;
    }
    mark();
    swap_state(c);
    // Spalart-Allmaras relaxation on the updated state
    if (c->k.viscous == 3) {
        if (c->saFused) {
            phl::launch_sa_pre(c->m, c->k, c->f, s);
            c->saDeferred = true;
        } else if (!ev && c->launchMode != 0 && c->k.nSweeps == 1) {
            join_turb(c);  // the previous step's sweeps wrote nu~
            phl::launch_sa_pre(c->m, c->k, c->f, s);
            hipEventRecord(c->evFork, s);
            hipStreamWaitEvent(c->side, c->evFork, 0);
            phl::launch_sa_lusgs(c->m, c->k, c->f, c->fwd, c->bwd, c->side);
            hipEventRecord(c->evJoin, c->side);
            c->turbPending = true;
        } else {
            phl::launch_sa_relax(c->m, c->k, c->f, c->fwd, c->bwd, s, &c->saLog);
        }
    }
    mark();
}

// kGraphSteps consecutive steps captured as one hipGraph.  A step swaps q
// and qAlt, so after an even number of steps the buffers are back where they
// started and the graph can be replayed; one graph is kept for each
// assignment of the buffers.  The captured sequence holds kernels and the
// event edges between the solver and side streams; it starts and ends with
// the side stream joined.
static const int kGraphSteps = 8;
static phl_status step_graph(phl_ctx *c, hipGraphExec_t *out)
{
    const int parity = c->f.q == c->qBuf0 ? 0 : 1;
    if (!c->stepGraph[parity]) {
        hipGraph_t g = nullptr;
        HIPTRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        for (int i = 0; i < kGraphSteps; ++i) enqueue_step(c, nullptr);
        join_turb(c);
        HIPTRY(hipStreamEndCapture(c->stream, &g));
        const hipError_t e = hipGraphInstantiate(&c->stepGraph[parity], g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        if (e != hipSuccess) {
            c->stepGraph[parity] = nullptr;
            return fail(PHL_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        }
        // capturing swapped the buffers kGraphSteps times: back at the start
    }
    *out = c->stepGraph[parity];
    return PHL_OK;
}

phl_status phl_sweep_levels(const phl_ctx *c, int *fwd_levels, int *bwd_levels)
{
    if (!c) return fail(PHL_ERR_ARG, "null argument");
    if (fwd_levels) *fwd_levels = c->fwd.nLevels;
    if (bwd_levels) *bwd_levels = c->bwd.nLevels;
    return PHL_OK;
}

phl_status phl_set_precision(phl_ctx *c, int precision)
{
    if (!c || (precision != 0 && precision != 1)) return fail(PHL_ERR_ARG, "precision must be 0 or 1");
    if (c->k.mixed == precision) return PHL_OK;
    HIPTRY(hipSetDevice(c->device));
    HIPTRY(hipStreamSynchronize(c->stream));
    c->k.mixed = precision;
    c->k.coefF32 = (precision && c->coefF32Ok) ? 1 : 0;
    c->p.precision = precision;
    for (hipGraphExec_t &g : c->stepGraph) {  // the graphs hold the other kernels
        if (g) hipGraphExecDestroy(g);
        g = nullptr;
    }
    return PHL_OK;
}

// Autotuner: time the fused flux kernel and the gradient kernel at each
// workgroup tile on the current state (a few launches each, HIP events on
// the context's stream) and keep the fastest.  The tile changes only the
// grouping of the residual partial sums, never a cell's arithmetic.
phl_status phl_autotune(phl_ctx *c, int *flux_tile, int *grad_tile, double *ms_table)
{
    if (!c) return fail(PHL_ERR_ARG, "null argument");
    HIPTRY(hipSetDevice(c->device));
    static const int tiles[3] = {64, 128, 256};
    hipEvent_t e0, e1;
    HIPTRY(hipEventCreate(&e0));
    HIPTRY(hipEventCreate(&e1));
    const int reps = 5;
    double best[2] = {1e300, 1e300};
    int pick[2] = {256, 256};
    const phl::Consts keep = c->k;
    // the flux kernel reads the step's CFL from cflArr[*stepCtr]
    c->hostCfl[0] = compute_cfl(c->p, 1);
    HIPTRY(hipMemcpyAsync(c->f.cflArr, c->hostCfl, sizeof(double), hipMemcpyHostToDevice, c->stream));
    reset_step_counter(c);
    phl::launch_boundary(c->m, c->k, c->f, c->stream);
    for (int kind = 0; kind < 2; ++kind)
        for (int t = 0; t < 3; ++t) {
            phl::Consts k = keep;
            if (kind == 0) k.fluxTile = tiles[t]; else k.gradTile = tiles[t];
            float ms = 0.f;
            for (int r = 0; r <= reps; ++r) {  // the first launch warms up
                if (r == 1) HIPTRY(hipEventRecord(e0, c->stream));
                if (kind == 0) phl::launch_flux(c->m, k, c->f, c->stream, true);
                else phl::launch_grad_lim(c->m, k, c->f, c->stream);
            }
            HIPTRY(hipEventRecord(e1, c->stream));
            HIPTRY(hipEventSynchronize(e1));
            HIPTRY(hipEventElapsedTime(&ms, e0, e1));
            const double per = ms / reps;
            if (ms_table) ms_table[kind * 3 + t] = per;
            if (per < best[kind]) { best[kind] = per; pick[kind] = tiles[t]; }
        }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    TRY(check_launch(c));
    c->k.fluxTile = pick[0];
    c->k.gradTile = pick[1];
    for (hipGraphExec_t &g : c->stepGraph) {  // the graphs hold the old tiles
        if (g) hipGraphExecDestroy(g);
        g = nullptr;
    }
    if (flux_tile) *flux_tile = pick[0];
    if (grad_tile) *grad_tile = pick[1];
    return PHL_OK;
}

phl_status phl_set_launch_mode(phl_ctx *c, int mode)
{
    if (!c || mode < 0 || mode > 2)
        return fail(PHL_ERR_ARG, "launch mode must be 0 (eager), 1 (graph) or 2 (eager, turbulence sweeps overlapped)");
    c->launchMode = mode;
    return PHL_OK;
}

phl_status phl_iterate(phl_ctx *c, int first_step, int n_steps, double *rms_hist)
{
    if (!c || n_steps < 0) return fail(PHL_ERR_ARG, "bad argument");
    HIPTRY(hipSetDevice(c->device));
    ++c->stateVersion;
    if (c->k.gradient == 2 && c->tFresh) {  // a new state's own temperature, outside the captured steps
        phl::launch_boundary(c->m, c->k, c->f, c->stream);
        fresh_temperature(c, c->stream);
    }
    const int K = PHL_N_KERNEL_CLASSES;
    if (c->timing && c->events.empty()) {
        c->events.resize((size_t)kTimedChunk * (K + 1));
        for (hipEvent_t &e : c->events) HIPTRY(hipEventCreate(&e));
    }
    int done = 0;
    while (done < n_steps) {
        const int cap = c->timing ? kTimedChunk : c->normSlots;
        const int chunk = std::min(n_steps - done, cap);
        // the chunk's CFL schedule and step counter live in device memory so
        // that the kernel sequence of a step does not depend on the step
        HIPTRY(hipStreamSynchronize(c->stream));  // hostCfl may still be in use
        for (int i = 0; i < chunk; ++i) c->hostCfl[i] = compute_cfl(c->p, first_step + done + i);
        HIPTRY(hipMemcpyAsync(c->f.cflArr, c->hostCfl, sizeof(double) * chunk, hipMemcpyHostToDevice, c->stream));
        if (c->k.nCases > 1)
            HIPTRY(hipMemcpyAsync(c->f.frozen, c->hostFrozen, (size_t)c->k.nCases, hipMemcpyHostToDevice, c->stream));
        reset_step_counter(c);
        int i = 0;
        // iterations 1..flow_init_step: first order with the Van Leer flux,
        // launched eagerly (the captured graphs hold the regular steps)
        for (; i < chunk && first_step + done + i <= c->p.flow_init_step; ++i) {
            c->k.initPhase = 1;
            enqueue_step(c, c->timing ? &c->events[(size_t)i * (K + 1)] : nullptr);
            ++c->eagerSteps;
        }
        c->k.initPhase = 0;
        if (!c->timing && c->launchMode == 1 && c->k.nSweeps == 1) {
            // graphs are captured with a turbulence relaxation deferred at entry
            if (c->saFused && !c->saDeferred && i + kGraphSteps < chunk) enqueue_step(c, nullptr), ++i;
            for (; i + kGraphSteps <= chunk && c->saFused == c->saDeferred; i += kGraphSteps) {
                hipGraphExec_t g = nullptr;
                join_turb(c);  // a graph starts with the side stream joined
                TRY(step_graph(c, &g));
                HIPTRY(hipGraphLaunch(g, c->stream));
                ++c->graphReplays;
            }
        }
        c->eagerSteps += chunk - i;
        for (; i < chunk; ++i)
            enqueue_step(c, c->timing ? &c->events[(size_t)i * (K + 1)] : nullptr);
        c->k.initPhase = c->hookInit;
        TRY(check_launch(c));
        const int nCases = c->k.nCases;
        HIPTRY(hipMemcpyAsync(c->hostNorm, c->f.normOut, sizeof(double) * 5 * chunk * nCases, hipMemcpyDeviceToHost,
                              c->stream));
        HIPTRY(hipStreamSynchronize(c->stream));
        TRY(check_sweep_waits(c));
        if (c->timing) {
            for (int i = 0; i < chunk; ++i)
                for (int k = 0; k < K; ++k) {
                    float ms = 0.f;
                    HIPTRY(hipEventElapsedTime(&ms, c->events[(size_t)i * (K + 1) + k],
                                               c->events[(size_t)i * (K + 1) + k + 1]));
                    c->kernelMs[k] += ms;
                }
            c->timedSteps += chunk;
        }
        for (int i = 0; i < chunk; ++i)
            for (int k = 0; k < nCases; ++k) {
                const double s0 = c->hostNorm[((size_t)i * nCases + k) * 5];
                if (!std::isfinite(s0))
                    return fail(PHL_ERR_NUMERIC, "residual is not finite at step " +
                                                     std::to_string(first_step + done + i) + " of case " +
                                                     std::to_string(k));
                const int nCk = nCases > 1 ? c->caseCellPtr[k + 1] - c->caseCellPtr[k] : c->m.nCell;
                if (rms_hist) rms_hist[(size_t)(done + i) * nCases + k] = std::sqrt(s0 / nCk);
            }
        done += chunk;
    }
    flush_sa(c);
    join_turb(c);  // later work on the solver stream sees the final nu~
    TRY(check_launch(c));
    return PHL_OK;
}

phl_status phl_set_kernel_timing(phl_ctx *c, int enable)
{
    if (!c) return fail(PHL_ERR_ARG, "null argument");
    c->timing = enable != 0;
    for (double &v : c->kernelMs) v = 0.0;
    c->timedSteps = 0;
    return PHL_OK;
}

phl_status phl_get_kernel_timing(phl_ctx *c, double *ms_total, int *n_steps)
{
    if (!c) return fail(PHL_ERR_ARG, "null argument");
    if (ms_total)
        for (int k = 0; k < PHL_N_KERNEL_CLASSES; ++k) ms_total[k] = c->kernelMs[k];
    if (n_steps) *n_steps = c->timedSteps;
    return PHL_OK;
}

phl_status phl_hbm_copy(int device, double bytes, int iters, double *gbps)
{
    if (!(bytes > 0) || iters <= 0 || !gbps) return fail(PHL_ERR_ARG, "bad argument");
    HIPTRY(hipSetDevice(device));
    const size_t n = ((size_t)(bytes / 2 / sizeof(double)) / 2) * 2;  // doubles per buffer
    double *a = nullptr, *b = nullptr;
    HIPTRY(hipMalloc(&a, n * sizeof(double)));
    hipError_t e = hipMalloc(&b, n * sizeof(double));
    if (e != hipSuccess) { hipFree(a); return fail(PHL_ERR_ALLOC, "hipMalloc failed"); }
    hipStream_t s;
    hipEvent_t e0, e1;
    HIPTRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    HIPTRY(hipEventCreate(&e0));
    HIPTRY(hipEventCreate(&e1));
    HIPTRY(hipMemsetAsync(a, 0, n * sizeof(double), s));
    double best = 0.0;
    for (int i = 0; i < iters + 1; ++i) {
        hipEventRecord(e0, s);
        phl::launch_copy(a, b, n, s);
        hipEventRecord(e1, s);
        HIPTRY(hipEventSynchronize(e1));
        float ms = 0.f;
        HIPTRY(hipEventElapsedTime(&ms, e0, e1));
        if (i > 0) best = std::max(best, 2.0 * n * sizeof(double) / (ms * 1e-3) / 1e9);
    }
    hipEventDestroy(e0); hipEventDestroy(e1); hipStreamDestroy(s);
    hipFree(a); hipFree(b);
    *gbps = best;
    return PHL_OK;
}

// Batched solve: every case runs until its own density residual has dropped
// by `drop`; a converged case is frozen (its state is no longer updated) at
// the end of the 32-step chunk in which it converged.
// Progress of one case for the precision switch: the best residual so far
// and the step at which it last halved.
struct Progress {
    double mark = INFINITY, best = INFINITY;
    int since = 0;
    void add(double r, int step)
    {
        if (r < best) best = r;
        if (best < 0.5 * mark) {
            mark = best;
            since = step;
        }
    }
};

phl_status phl_set_precision_switch(phl_ctx *c, int stall_steps)
{
    if (!c || stall_steps < 0) return fail(PHL_ERR_ARG, "bad argument");
    c->precisionSwitchStall = stall_steps;
    return PHL_OK;
}

phl_status phl_get_precision(const phl_ctx *c, int *precision)
{
    if (!c || !precision) return fail(PHL_ERR_ARG, "null argument");
    *precision = c->k.mixed;
    return PHL_OK;
}

phl_status phl_solve_batch(phl_ctx *c, int max_steps, double drop, int *steps_done, double *rms_hist)
{
    if (!c || max_steps <= 0 || !(drop > 0.0) || !steps_done) return fail(PHL_ERR_ARG, "bad argument");
    const int n = c->k.nCases;
    std::vector<double> hist((size_t)max_steps * n);
    std::vector<double> r1(n, 0.0);
    std::vector<int> hit(n, -1);
    if (n > 1) std::fill(c->hostFrozen, c->hostFrozen + n, (unsigned char)0);
    const int chunk = 32;
    int done = 0, left = n;
    std::vector<Progress> prog(n);
    while (done < max_steps && left > 0) {
        const int m = std::min(chunk, max_steps - done);
        TRY(phl_iterate(c, 1 + done, m, hist.data() + (size_t)done * n));
        bool stalled = true;
        for (int k = 0; k < n; ++k) {
            if (done == 0) r1[k] = hist[k];
            if (hit[k] >= 0) continue;
            for (int i = done; i < done + m; ++i) {
                prog[k].add(hist[(size_t)i * n + k], i);
                if (hist[(size_t)i * n + k] <= drop * r1[k]) { hit[k] = i; break; }
            }
            if (hit[k] >= 0) {
                --left;
                if (n > 1) c->hostFrozen[k] = 1;
            } else if (done + m - prog[k].since < c->precisionSwitchStall) {
                stalled = false;
            }
        }
        done += m;
        if (c->k.mixed && c->precisionSwitchStall > 0 && left > 0 && stalled) TRY(phl_set_precision(c, 0));
    }
    for (int k = 0; k < n; ++k) steps_done[k] = hit[k] >= 0 ? hit[k] + 1 : -done;
    if (rms_hist) std::copy(hist.begin(), hist.begin() + (size_t)done * n, rms_hist);
    if (n > 1) std::fill(c->hostFrozen, c->hostFrozen + n, (unsigned char)0);
    return PHL_OK;
}

phl_status phl_set_frozen(phl_ctx *c, const unsigned char *frozen)
{
    if (!c || !frozen) return fail(PHL_ERR_ARG, "null argument");
    if (c->k.nCases <= 1) return fail(PHL_ERR_UNSUPPORTED, "phl_set_frozen: not a batched context");
    std::copy(frozen, frozen + c->k.nCases, c->hostFrozen);  // uploaded by the next phl_iterate
    return PHL_OK;
}

phl_status phl_batch_size(const phl_ctx *c, int *n_cases)
{
    if (!c || !n_cases) return fail(PHL_ERR_ARG, "null argument");
    *n_cases = c->k.nCases;
    return PHL_OK;
}

phl_status phl_solve(phl_ctx *c, int max_steps, double drop, int *steps_done, double *rms_hist)
{
    if (!c || max_steps <= 0 || !(drop > 0.0)) return fail(PHL_ERR_ARG, "bad argument");
    std::vector<double> hist(max_steps);
    const int chunk = 32;  // steps between host checks of the residual history
    int done = 0;
    double r1 = 0.0;
    int hit = -1;
    Progress prog;
    while (done < max_steps && hit < 0) {
        const int n = std::min(chunk, max_steps - done);
        TRY(phl_iterate(c, 1 + done, n, hist.data() + done));
        if (done == 0) r1 = hist[0];
        for (int i = done; i < done + n; ++i) {
            prog.add(hist[i], i);
            if (hist[i] <= drop * r1) { hit = i; break; }
        }
        done += n;
        if (hit < 0 && c->k.mixed && c->precisionSwitchStall > 0 && done - prog.since >= c->precisionSwitchStall)
            TRY(phl_set_precision(c, 0));
    }
    if (steps_done) *steps_done = done;
    if (rms_hist) std::copy(hist.begin(), hist.begin() + done, rms_hist);
    return PHL_OK;
}

phl_status phl_synchronize(phl_ctx *c)
{
    if (!c) return fail(PHL_ERR_ARG, "null argument");
    HIPTRY(hipSetDevice(c->device));
    HIPTRY(hipStreamSynchronize(c->stream));
    return PHL_OK;
}

phl_status phl_set_sweep_wait(phl_ctx *c, int ticks)
{
    if (!c || ticks < 0) return fail(PHL_ERR_ARG, "phl_set_sweep_wait: bad argument");
    HIPTRY(hipSetDevice(c->device));
    const int v = ticks > 0 ? ticks : 0;
    HIPTRY(hipMemcpyAsync(c->f.sweepCtl + 3, &v, sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPTRY(hipStreamSynchronize(c->stream));
    return PHL_OK;
}

phl_status phl_launch_counts(const phl_ctx *c, long long *graph_replays, long long *eager_steps)
{
    if (!c) return fail(PHL_ERR_ARG, "null argument");
    if (graph_replays) *graph_replays = c->graphReplays;
    if (eager_steps) *eager_steps = c->eagerSteps;
    return PHL_OK;
}

}  // extern "C"
