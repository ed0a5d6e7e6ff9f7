// phl_mpnn.hip -- the warm-start message-passing network on gfx950: forward,
// backward and Adam, hand-written, FP32 with the dense layers on MFMA
// (v_mfma_f32_32x32x2_f32).  The network (no reference counterpart):
//
//     h0 = relu(X W0 + b0)                     X [N, 8] cell features
//     l = 1..L:  m = mean_{j in N(i)} h_j       (face neighbours)
//                z = h A_l + m B_l + c_l ;  h = h + relu(z)
//     y  = h Wo + bo                           y [N, 5]
//
// Rows are cells, columns features, all row-major.  The hidden width is 128.
// Every reduction (weight gradients, bias gradients, loss) is a sum in a
// fixed order (no atomics), so training is deterministic.
#include "phl_mpnn_internal.h"

#include <hip/hip_runtime.h>

namespace phl {
namespace nn {

namespace {

typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int H = kHidden;    // 128
constexpr int KC = 32;        // K chunk staged in LDS
constexpr int TR = 128;       // rows per workgroup tile
constexpr int kThreads = 256; // 4 waves
constexpr int kRowChunk = 2048;  // rows per partial of a weight gradient
constexpr int kMaxDeg = 8;       // neighbour lists up to this width take the batched-load path

// D(32x32) += A(32x2) * B(2x32).  Lane l holds A[l % 32][l / 32] and
// B[l / 32][l % 32]; accumulator register r of lane l holds
// D[(r % 4) + 4 * (l / 32) + 8 * (r / 4)][l % 32].
__device__ __forceinline__ int acc_row(int r, int lane) { return (r & 3) + 4 * (lane >> 5) + 8 * (r >> 2); }

// C[N, 128] = A1[N, K1] W1[K1, 128] (+ A2[N, 128] W2[128, 128]) + bias,
// W given row-major [K, 128], or transposed ([128, K], TW) for the backward
// products dZ W^T.  Epilogue: z = C stored to Z (if non-null); out = z
// (mode 0), relu(z) (mode 1), Hres + relu(z) (mode 2) or Hres + z (mode 3).
// One thread's share of a K chunk, as float4: A [TR][KC] and W [KC][H] are
// 1024 float4 each, 4 per thread.  K is a multiple of 4 (kIn or kHidden), so
// a float4 is either wholly inside K or wholly past it; rows and weight
// offsets are multiples of 4 floats, so every float4 load is aligned.
static_assert(kIn % 4 == 0 && H % 4 == 0, "float4 staging needs K % 4 == 0");
constexpr int kVec = TR * KC / 4 / kThreads;  // 4
static_assert(TR * KC == KC * H && kVec * 4 * kThreads == TR * KC, "A and W chunks have the same size");

struct Chunk {
    float4 a[kVec], w[kVec];
};

template <bool TW>
__device__ __forceinline__ void load_chunk(Chunk &st, const float *__restrict__ A, const float *__restrict__ W,
                                           int N, int K, int row0, int k0)
{
#pragma unroll
    for (int i = 0; i < kVec; ++i) {
        const int q = threadIdx.x + i * kThreads;
        const int row = row0 + (q >> 3), k = k0 + 4 * (q & 7);  // A: row q/8, k 4(q%8)..+3
        st.a[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (row < N && k < K) st.a[i] = *(const float4 *)(A + (size_t)row * K + k);
        st.w[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (TW) {  // W^T stored [H][K]: column q/8, k 4(q%8)..+3
            if (k < K) st.w[i] = *(const float4 *)(W + (size_t)(q >> 3) * K + k);
        } else {   // W stored [K][H]: k q/32, columns 4(q%32)..+3
            const int kw = k0 + (q >> 5);
            if (kw < K) st.w[i] = *(const float4 *)(W + (size_t)kw * H + 4 * (q & 31));
        }
    }
}

template <bool TW>
__device__ __forceinline__ void store_chunk(const Chunk &st, float (*Al)[KC + 1], float (*Wl)[H + 4])
{
#pragma unroll
    for (int i = 0; i < kVec; ++i) {
        const int q = threadIdx.x + i * kThreads;
        const int r = q >> 3, kk = 4 * (q & 7);
        Al[r][kk] = st.a[i].x;
        Al[r][kk + 1] = st.a[i].y;
        Al[r][kk + 2] = st.a[i].z;
        Al[r][kk + 3] = st.a[i].w;
        if (TW) {
            Wl[kk][r] = st.w[i].x;
            Wl[kk + 1][r] = st.w[i].y;
            Wl[kk + 2][r] = st.w[i].z;
            Wl[kk + 3][r] = st.w[i].w;
        } else {
            const int kw = q >> 5, c = 4 * (q & 31);
            Wl[kw][c] = st.w[i].x;
            Wl[kw][c + 1] = st.w[i].y;
            Wl[kw][c + 2] = st.w[i].z;
            Wl[kw][c + 3] = st.w[i].w;
        }
    }
}

// The next chunk's loads are issued before the MFMAs of the current one and
// land in registers while they run; the LDS tiles are rewritten between two
// barriers.
template <bool TW>
__device__ __forceinline__ void accumulate(f32x16 (&acc)[4], float (*Al)[KC + 1], float (*Wl)[H + 4],
                                           const float *__restrict__ A, const float *__restrict__ W, int N, int K,
                                           int row0)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    Chunk st;
    load_chunk<TW>(st, A, W, N, K, row0, 0);
    for (int k0 = 0; k0 < K; k0 += KC) {
        __syncthreads();
        store_chunk<TW>(st, Al, Wl);
        __syncthreads();
        load_chunk<TW>(st, A, W, N, K, row0, k0 + KC);  // past K: zeros, no loads
        const int kmax = (K - k0 < KC) ? (K - k0) : KC;
        for (int kk = 0; kk < kmax; kk += 2) {
            const float a = Al[wave * 32 + (lane & 31)][kk + (lane >> 5)];
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const float w = Wl[kk + (lane >> 5)][b * 32 + (lane & 31)];
                acc[b] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, w, acc[b], 0, 0, 0);
            }
        }
    }
}

template <bool TW>
__global__ void __launch_bounds__(kThreads, 2) k_dense(const float *__restrict__ A1, const float *__restrict__ W1, int K1,
                                                    const float *__restrict__ A2, const float *__restrict__ W2,
                                                    const float *__restrict__ bias, int N, int mode,
                                                    const float *__restrict__ Hres, float *__restrict__ Z,
                                                    float *__restrict__ out)
{
    __shared__ float Al[TR][KC + 1];
    __shared__ float Wl[KC][H + 4];
    const int row0 = blockIdx.x * TR;
    f32x16 acc[4];
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[b][r] = 0.f;
    accumulate<TW>(acc, Al, Wl, A1, W1, N, K1, row0);
    if (A2) accumulate<TW>(acc, Al, Wl, A2, W2, N, H, row0);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int row = row0 + wave * 32 + acc_row(r, lane);
            const int col = b * 32 + (lane & 31);
            if (row >= N) continue;
            const size_t o = (size_t)row * H + col;
            const float z = acc[b][r] + (bias ? bias[col] : 0.f);
            if (Z) Z[o] = z;
            const float rz = z > 0.f ? z : 0.f;
            out[o] = mode == 0 ? z : (mode == 1 ? rz : (mode == 2 ? Hres[o] + rz : Hres[o] + z));
        }
}

// ---- BF16 form of the hidden layers' products (inference) ----
// z = h A + m B + c with h, m and the weights rounded to BF16 (round to
// nearest even, v_cvt_pk_bf16_f32) and FP32 accumulation on
// v_mfma_f32_32x32x16_bf16; h' = h + relu(z) in FP32.  The weights come
// transposed ([128 columns][128 k], BF16) from k_weights_bf16, so that each
// lane's eight consecutive k of a B operand are one 16-byte LDS read, as
// are the eight k of its A operand.  Lane l supplies A[l % 32][8 (l / 32) ..
// + 7] and B[8 (l / 32) .. + 7][l % 32] of a K = 16 step; the accumulator
// layout is the one of the FP32 instruction (acc_row).
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
constexpr int KB = 32;       // K chunk staged in LDS
constexpr int KBP = KB + 8;  // LDS row length: 80-byte rows keep every 16-byte read aligned
static_assert(kVec * 4 * kThreads == TR * KB && 2 * 8 * kThreads == H * KB, "BF16 chunk shares");

struct ChunkB {
    float4 a[kVec];  // A [TR][KB]: row q/8, k 4(q%8)..+3
    uint4 w[2];      // W^T [H][KB]: column q/4, k 8(q%4)..+7
};

__device__ __forceinline__ void load_chunk_bf16(ChunkB &st, const float *__restrict__ A,
                                                const __bf16 *__restrict__ Wt, int N, int row0, int k0)
{
#pragma unroll
    for (int i = 0; i < kVec; ++i) {
        const int q = threadIdx.x + i * kThreads;
        const int row = row0 + (q >> 3), k = k0 + 4 * (q & 7);
        st.a[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (row < N && k < H) st.a[i] = *(const float4 *)(A + (size_t)row * H + k);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int q = threadIdx.x + i * kThreads;
        const int k = k0 + 8 * (q & 3);
        st.w[i] = make_uint4(0u, 0u, 0u, 0u);
        if (k < H) st.w[i] = *(const uint4 *)(Wt + (size_t)(q >> 2) * H + k);
    }
}

__device__ __forceinline__ void store_chunk_bf16(const ChunkB &st, __bf16 (*Al)[KBP], __bf16 (*Wl)[KBP])
{
#pragma unroll
    for (int i = 0; i < kVec; ++i) {
        const int q = threadIdx.x + i * kThreads;
        bf16x4 v;
        v[0] = (__bf16)st.a[i].x;
        v[1] = (__bf16)st.a[i].y;
        v[2] = (__bf16)st.a[i].z;
        v[3] = (__bf16)st.a[i].w;
        *(bf16x4 *)&Al[q >> 3][4 * (q & 7)] = v;
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int q = threadIdx.x + i * kThreads;
        *(uint4 *)&Wl[q >> 2][8 * (q & 3)] = st.w[i];
    }
}

__device__ __forceinline__ void accumulate_bf16(f32x16 (&acc)[4], __bf16 (*Al)[KBP], __bf16 (*Wl)[KBP],
                                                const float *__restrict__ A, const __bf16 *__restrict__ Wt, int N,
                                                int row0)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    ChunkB st;
    load_chunk_bf16(st, A, Wt, N, row0, 0);
    for (int k0 = 0; k0 < H; k0 += KB) {
        __syncthreads();
        store_chunk_bf16(st, Al, Wl);
        __syncthreads();
        load_chunk_bf16(st, A, Wt, N, row0, k0 + KB);  // past K: zeros, no loads
#pragma unroll
        for (int ks = 0; ks < KB; ks += 16) {
            const bf16x8 a = *(const bf16x8 *)&Al[wave * 32 + (lane & 31)][ks + 8 * (lane >> 5)];
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const bf16x8 w = *(const bf16x8 *)&Wl[b * 32 + (lane & 31)][ks + 8 * (lane >> 5)];
                acc[b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, w, acc[b], 0, 0, 0);
            }
        }
    }
}

// out = H + relu(A1 W1 + A2 W2 + bias), A1 = h and A2 = m [N, 128] FP32, W1t,
// W2t the layer's transposed BF16 weights.
__global__ void __launch_bounds__(kThreads, 2) k_dense_bf16(const float *__restrict__ A1,
                                                         const __bf16 *__restrict__ W1t,
                                                         const float *__restrict__ A2,
                                                         const __bf16 *__restrict__ W2t,
                                                         const float *__restrict__ bias, int N,
                                                         const float *__restrict__ Hres, float *__restrict__ out)
{
    __shared__ __attribute__((aligned(16))) __bf16 Al[TR][KBP];  // 16-byte reads and stores
    __shared__ __attribute__((aligned(16))) __bf16 Wl[H][KBP];
    const int row0 = blockIdx.x * TR;
    f32x16 acc[4];
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[b][r] = 0.f;
    accumulate_bf16(acc, Al, Wl, A1, W1t, N, row0);
    accumulate_bf16(acc, Al, Wl, A2, W2t, N, row0);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int row = row0 + wave * 32 + acc_row(r, lane);
            const int col = b * 32 + (lane & 31);
            if (row >= N) continue;
            const size_t o = (size_t)row * H + col;
            const float z = acc[b][r] + bias[col];
            out[o] = Hres[o] + (z > 0.f ? z : 0.f);
        }
}

// The hidden layers' A_l and B_l [L][128 k][128 columns] FP32 to BF16,
// transposed: out [L][2][128 columns][128 k] (A_l, then B_l).
__global__ void __launch_bounds__(kThreads) k_weights_bf16(const float *__restrict__ A, const float *__restrict__ B,
                                                           int L, __bf16 *__restrict__ out)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t HH = (size_t)H * H;
    if (i >= (size_t)L * 2 * HH) return;
    const size_t l = i / (2 * HH), r = i % (2 * HH);
    const int which = (int)(r / HH), col = (int)((r % HH) / H), k = (int)(r % H);
    const float *src = which ? B : A;
    out[i] = (__bf16)src[l * HH + (size_t)k * H + col];
}

constexpr int kPer = 32 * H / kThreads;  // elements of a 32-row slab staged per thread

// Partial of G[K, 128] = A[rows, K]^T D[rows, 128] over one chunk of
// kRowChunk rows; K = 128.  Wave w owns G rows 32w..32w+31.
__global__ void __launch_bounds__(kThreads) k_wgrad_part(const float *__restrict__ A, const float *__restrict__ D,
                                                         int N, float *__restrict__ part)
{
    __shared__ float Al[32][H + 4];
    __shared__ float Dl[32][H + 4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r0 = blockIdx.x * kRowChunk;
    f32x16 acc[4];
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[b][r] = 0.f;
    for (int s = 0; s < kRowChunk; s += 32) {
        float va[kPer], vd[kPer];
#pragma unroll
        for (int i = 0; i < kPer; ++i) {
            const int e = threadIdx.x + i * kThreads;
            const int row = r0 + s + e / H;
            const bool in = row < N;
            va[i] = in ? A[(size_t)row * H + e % H] : 0.f;
            vd[i] = in ? D[(size_t)row * H + e % H] : 0.f;
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < kPer; ++i) {
            const int e = threadIdx.x + i * kThreads;
            Al[e / H][e % H] = va[i];
            Dl[e / H][e % H] = vd[i];
        }
        __syncthreads();
        for (int n = 0; n < 32; n += 2) {
            const float a = Al[n + (lane >> 5)][wave * 32 + (lane & 31)];
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const float d = Dl[n + (lane >> 5)][b * 32 + (lane & 31)];
                acc[b] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, d, acc[b], 0, 0, 0);
            }
        }
    }
    float *p = part + (size_t)blockIdx.x * H * H;
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int r = 0; r < 16; ++r) p[(size_t)(wave * 32 + acc_row(r, lane)) * H + b * 32 + (lane & 31)] = acc[b][r];
}

// Small products on the vector ALUs: partial of G[K, C] = A[rows, K]^T D[rows, C]
// for K, C <= 128 (input layer, output layer, bias rows with A = 1).
__global__ void __launch_bounds__(kThreads) k_small_grad_part(const float *__restrict__ A, int K,
                                                              const float *__restrict__ D, int C, int N,
                                                              float *__restrict__ part)
{
    const int r0 = blockIdx.x * kRowChunk;
    const int r1 = (r0 + kRowChunk < N) ? r0 + kRowChunk : N;
    float *p = part + (size_t)blockIdx.x * K * C;
    for (int e = threadIdx.x; e < K * C; e += kThreads) {
        const int k = e / C, c = e % C;
        float s = 0.f;
#pragma unroll 8
        for (int r = r0; r < r1; ++r) s += (A ? A[(size_t)r * K + k] : 1.f) * D[(size_t)r * C + c];
        p[e] = s;
    }
}

// out = sum of nPart partials of size n.  A workgroup owns 64 consecutive
// outputs (lane = output); wave w adds partials w, w + 4, w + 8, .. in order
// and the four wave sums are added in wave order: a fixed order, so the
// result is deterministic.
constexpr int kSumWaves = kThreads / 64;
__global__ void __launch_bounds__(kThreads) k_sum_parts(const float *__restrict__ part, int nPart, int n,
                                                        float *__restrict__ out)
{
    __shared__ float red[kSumWaves][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int e = blockIdx.x * 64 + lane;
    float s = 0.f;
    if (e < n)
        for (int i = wave; i < nPart; i += kSumWaves) s += part[(size_t)i * n + e];
    red[wave][lane] = s;
    __syncthreads();
    if (wave == 0 && e < n) {
        float t = red[0][lane];
        for (int w = 1; w < kSumWaves; ++w) t += red[w][lane];
        out[e] = t;
    }
}

// m_i = mean of h_j over the neighbours of i (scale = 1/deg_i), or, for the
// transpose in the backward pass, out_j += sum over neighbours i of
// w_i * h_i with w_i = 1/deg_i (TRANSPOSE).
// One thread per (cell, 4 columns), float4 rows; each column is summed in
// list order, as a scalar loop over the neighbours would sum it.
template <bool TRANSPOSE>
__device__ __forceinline__ void add_row(float4 &s, const float4 x, const float *__restrict__ invDeg, int j)
{
    if (TRANSPOSE) {
        const float w = invDeg[j];
        s.x += x.x * w;
        s.y += x.y * w;
        s.z += x.z * w;
        s.w += x.w * w;
    } else {
        s.x += x.x;
        s.y += x.y;
        s.z += x.z;
        s.w += x.w;
    }
}

template <bool TRANSPOSE>
__global__ void __launch_bounds__(kThreads) k_neighbour(const int *__restrict__ nbr, int W, int N,
                                                        const float *__restrict__ invDeg,
                                                        const float *__restrict__ h, float *__restrict__ out)
{
    constexpr int Q = H / 4;
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)N * Q) return;
    const int i = (int)(e / Q), c = 4 * (int)(e % Q);
    float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
    if (W <= kMaxDeg) {
        // every index, then every row, in flight together; the lists are
        // compacted (valid entries first), so the sum order is k = 0, 1, ..
        int j[kMaxDeg];
        float4 x[kMaxDeg];
#pragma unroll
        for (int k = 0; k < kMaxDeg; ++k) j[k] = k < W ? nbr[(size_t)k * N + i] : -1;
#pragma unroll
        for (int k = 0; k < kMaxDeg; ++k) {
            x[k] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (j[k] >= 0) x[k] = *(const float4 *)(h + (size_t)j[k] * H + c);
        }
#pragma unroll
        for (int k = 0; k < kMaxDeg; ++k) {
            if (j[k] < 0) break;
            add_row<TRANSPOSE>(s, x[k], invDeg, j[k]);
        }
    } else {
        for (int k = 0; k < W; ++k) {
            const int jk = nbr[(size_t)k * N + i];
            if (jk < 0) break;
            add_row<TRANSPOSE>(s, *(const float4 *)(h + (size_t)jk * H + c), invDeg, jk);
        }
    }
    float4 *o = (float4 *)(out + (size_t)i * H + c);
    if (TRANSPOSE) {
        float4 v = *o;
        v.x += s.x;
        v.y += s.y;
        v.z += s.z;
        v.w += s.w;
        *o = v;
    } else {
        const float w = invDeg[i];
        *o = make_float4(s.x * w, s.y * w, s.z * w, s.w * w);
    }
}

// dz = dh * (z > 0): thread t does elements 4t..4t+3 (float4 when all four
// are in range; the arrays are rows of 128 floats, so only a tail past n
// would not be), scalar for a tail.
__global__ void __launch_bounds__(kThreads) k_relu_back(const float *__restrict__ dh, const float *__restrict__ z,
                                                        size_t n, float *__restrict__ dz)
{
    const size_t e = 4 * ((size_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (e + 4 <= n) {
        const float4 zv = *(const float4 *)(z + e), hv = *(const float4 *)(dh + e);
        *(float4 *)(dz + e) = make_float4(zv.x > 0.f ? hv.x : 0.f, zv.y > 0.f ? hv.y : 0.f,
                                          zv.z > 0.f ? hv.z : 0.f, zv.w > 0.f ? hv.w : 0.f);
    } else {
        for (size_t k = e; k < n; ++k) dz[k] = z[k] > 0.f ? dh[k] : 0.f;
    }
}

__global__ void __launch_bounds__(kThreads) k_add(float *__restrict__ a, const float *__restrict__ b, size_t n)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) a[e] += b[e];
}

// Output layer: y = h Wo + bo, one thread per row (float4 reads of h), each
// output summed over k = 0..127 in order.
__global__ void __launch_bounds__(kThreads) k_out(const float *__restrict__ h, const float *__restrict__ Wo,
                                                  const float *__restrict__ bo, int N, float *__restrict__ y)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    float s[kOut];
#pragma unroll
    for (int o = 0; o < kOut; ++o) s[o] = bo[o];
    const float4 *hr = (const float4 *)(h + (size_t)i * H);
    for (int q = 0; q < H / 4; ++q) {
        const float4 x = hr[q];
        const float *w = Wo + 4 * q * kOut;
#pragma unroll
        for (int o = 0; o < kOut; ++o) s[o] += x.x * w[o];
#pragma unroll
        for (int o = 0; o < kOut; ++o) s[o] += x.y * w[kOut + o];
#pragma unroll
        for (int o = 0; o < kOut; ++o) s[o] += x.z * w[2 * kOut + o];
#pragma unroll
        for (int o = 0; o < kOut; ++o) s[o] += x.w * w[3 * kOut + o];
    }
#pragma unroll
    for (int o = 0; o < kOut; ++o) y[(size_t)i * kOut + o] = s[o];
}

// dh = dy Wo^T, one thread per (row, hidden column).
__global__ void __launch_bounds__(kThreads) k_out_back(const float *__restrict__ dy, const float *__restrict__ Wo,
                                                       int N, float *__restrict__ dh)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)N * H) return;
    const int i = (int)(e / H), k = (int)(e % H);
    float s = 0.f;
#pragma unroll
    for (int o = 0; o < kOut; ++o) s += dy[(size_t)i * kOut + o] * Wo[k * kOut + o];
    dh[e] = s;
}

// dy = 2 (y - t) / (N * 5) and per-block partial sums of (y - t)^2.
__global__ void __launch_bounds__(kThreads) k_loss(const float *__restrict__ y, const float *__restrict__ t, int n,
                                                   float scale, float *__restrict__ dy, double *__restrict__ part)
{
    __shared__ double red[kThreads];
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    double s = 0.0;
    if (e < n) {
        const float d = y[e] - t[e];
        dy[e] = 2.f * d * scale;
        s = (double)d * (double)d;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void k_loss_final(const double *__restrict__ part, int nPart, double scale, double *__restrict__ out)
{
    __shared__ double red[kThreads];
    double s = 0.0;
    for (int i = threadIdx.x; i < nPart; i += blockDim.x) s += part[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0] * scale;
}

// Adam with bias correction (Kingma & Ba 2015, Algorithm 1).
__global__ void __launch_bounds__(kThreads) k_adam(float *__restrict__ p, const float *__restrict__ g,
                                                   float *__restrict__ m, float *__restrict__ v, int n, float lr,
                                                   float b1, float b2, float c1, float c2, float eps)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const float gi = g[e];
    const float mi = b1 * m[e] + (1.f - b1) * gi;
    const float vi = b2 * v[e] + (1.f - b2) * gi * gi;
    m[e] = mi;
    v[e] = vi;
    p[e] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

inline int nb(size_t n, int b) { return (int)((n + b - 1) / b); }

}  // namespace

// ---------------------------------------------------------------- launchers
void dense(const float *A1, const float *W1, int K1, const float *A2, const float *W2, const float *bias, int N,
           int mode, const float *Hres, float *Z, float *out, bool transW, hipStream_t s)
{
    if (transW)
        hipLaunchKernelGGL(k_dense<true>, dim3(nb(N, TR)), dim3(kThreads), 0, s, A1, W1, K1, A2, W2, bias, N, mode,
                           Hres, Z, out);
    else
        hipLaunchKernelGGL(k_dense<false>, dim3(nb(N, TR)), dim3(kThreads), 0, s, A1, W1, K1, A2, W2, bias, N, mode,
                           Hres, Z, out);
}

void weights_bf16(const float *A, const float *B, int L, __bf16 *out, hipStream_t s)
{
    const size_t n = (size_t)L * 2 * H * H;
    hipLaunchKernelGGL(k_weights_bf16, dim3((unsigned)nb(n, kThreads)), dim3(kThreads), 0, s, A, B, L, out);
}

void dense_bf16(const float *h, const __bf16 *At, const float *m, const __bf16 *Bt, const float *bias, int N,
                float *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_dense_bf16, dim3(nb(N, TR)), dim3(kThreads), 0, s, h, At, m, Bt, bias, N, h, out);
}

int grad_parts(int N) { return nb(N, kRowChunk); }

void wgrad(const float *A, const float *D, int N, float *part, float *out, hipStream_t s)
{
    const int np = grad_parts(N);
    hipLaunchKernelGGL(k_wgrad_part, dim3(np), dim3(kThreads), 0, s, A, D, N, part);
    hipLaunchKernelGGL(k_sum_parts, dim3(nb(H * H, 64)), dim3(kThreads), 0, s, part, np, H * H, out);
}

void small_grad(const float *A, int K, const float *D, int C, int N, float *part, float *out, hipStream_t s)
{
    const int np = grad_parts(N);
    hipLaunchKernelGGL(k_small_grad_part, dim3(np), dim3(kThreads), 0, s, A, K, D, C, N, part);
    hipLaunchKernelGGL(k_sum_parts, dim3(nb(K * C, 64)), dim3(kThreads), 0, s, part, np, K * C, out);
}

void neighbour_mean(const int *nbr, int W, int N, const float *invDeg, const float *h, float *m, hipStream_t s)
{
    hipLaunchKernelGGL(k_neighbour<false>, dim3(nb((size_t)N * (H / 4), kThreads)), dim3(kThreads), 0, s, nbr, W, N,
                       invDeg, h, m);
}

void neighbour_mean_transpose_add(const int *nbr, int W, int N, const float *invDeg, const float *dm, float *dh,
                                  hipStream_t s)
{
    hipLaunchKernelGGL(k_neighbour<true>, dim3(nb((size_t)N * (H / 4), kThreads)), dim3(kThreads), 0, s, nbr, W, N,
                       invDeg, dm, dh);
}

void relu_back(const float *dh, const float *z, size_t n, float *dz, hipStream_t s)
{
    hipLaunchKernelGGL(k_relu_back, dim3(nb((n + 3) / 4, kThreads)), dim3(kThreads), 0, s, dh, z, n, dz);
}

void add(float *a, const float *b, size_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_add, dim3(nb(n, kThreads)), dim3(kThreads), 0, s, a, b, n);
}

void out_layer(const float *h, const float *Wo, const float *bo, int N, float *y, hipStream_t s)
{
    hipLaunchKernelGGL(k_out, dim3(nb((size_t)N, kThreads)), dim3(kThreads), 0, s, h, Wo, bo, N, y);
}

void out_layer_back(const float *dy, const float *Wo, int N, float *dh, hipStream_t s)
{
    hipLaunchKernelGGL(k_out_back, dim3(nb((size_t)N * H, kThreads)), dim3(kThreads), 0, s, dy, Wo, N, dh);
}

int loss_parts(int n) { return nb(n, kThreads); }

void loss(const float *y, const float *t, int n, float *dy, double *part, double *out, hipStream_t s)
{
    const int np = loss_parts(n);
    hipLaunchKernelGGL(k_loss, dim3(np), dim3(kThreads), 0, s, y, t, n, 1.f / (float)n, dy, part);
    hipLaunchKernelGGL(k_loss_final, dim3(1), dim3(kThreads), 0, s, part, np, 1.0 / (double)n, out);
}

void adam(float *p, const float *g, float *m, float *v, int n, float lr, int step, hipStream_t s)
{
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3(nb(n, kThreads)), dim3(kThreads), 0, s, p, g, m, v, n, lr, b1, b2, c1, c2, 1e-8f);
}

}  // namespace nn
}  // namespace phl
