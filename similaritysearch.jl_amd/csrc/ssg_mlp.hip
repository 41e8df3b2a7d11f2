// ssg_mlp.hip -- the learned entry-point model: an MLP D -> H1 -> H2 -> (L + NB) with
// ReLU, whose first L outputs score landmarks (entry points) and last NB outputs
// score beam-size buckets.  FP16 operands on MFMA with FP32 accumulation, FP32
// master weights updated by Adam.  No reference counterpart: this is the learned
// form added beside the plain beam search.
//
// Kernels:
//   gemm_f16: C[M x N] (FP32) = op(A)[M x K] . op(B)[K x N], FP16 operands given by
//     element strides (so forward, backward-data and backward-weight products are
//     one kernel), staged through LDS 32 k at a time.  A block of 4 waves owns a
//     128 x 128 (or, for small products, 64 x 64) tile; each wave a quarter of it in
//     v_mfma_f32_16x16x16_f16 accumulators.
//   softmax_xent (two heads, loss + gradient in one pass per row),
//   relu_grad, colsum (bias gradients), adam, and topk_heads (inference output).
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "ssg_device.h"
#include "ssg_handle.h"

namespace ssg {

typedef _Float16 half4 __attribute__((ext_vector_type(4)));
typedef float floatx4v __attribute__((ext_vector_type(4)));

constexpr int GK = 32;       // k per LDS stage
constexpr int GKP = GK + 4;  // padded LDS row (halves): fragment reads hit distinct banks

// A tile of GT rows x GK k of operand P (element (r, k) at P[r * sr + k * sk]) goes to
// LDS as S[r * GKP + k] in two steps, so the next tile can be loaded into registers while
// the current one is multiplied.  Each thread handles 8-half chunks: along k (mode 1,
// sk == 1) or along rows (mode 2, sr == 1) with 16-byte aligned vector loads; any other
// layout (mode 0) and the matrix edges load element by element.
template <int GT>
struct TileRegs {
    static constexpr int CH = GT * GK / 8;          // 8-half chunks per tile
    static constexpr int N = (CH + 255) / 256;      // chunks per thread
    uint4 v[N];
};

template <int GT>
__device__ __forceinline__ void fetch_tile(TileRegs<GT>& t, const _Float16* __restrict__ P, int64_t sr,
                                           int64_t sk, int r0, int R, int k0, int K, int mode, int tid)
{
#pragma unroll
    for (int i = 0; i < TileRegs<GT>::N; i++) {
        const int c = tid + 256 * i;
        _Float16 v[8];
        if (c >= TileRegs<GT>::CH) {
#pragma unroll
            for (int e = 0; e < 8; e++) v[e] = (_Float16)0.f;
        } else if (mode == 1) {
            const int r = c / (GK / 8), kc = (c % (GK / 8)) * 8;
            const int gr = r0 + r, gk = k0 + kc;
            if (gr < R && gk + 7 < K) {
                t.v[i] = *reinterpret_cast<const uint4*>(P + gr * sr + gk);
                continue;
            }
#pragma unroll
            for (int e = 0; e < 8; e++) v[e] = (gr < R && gk + e < K) ? P[gr * sr + (gk + e)] : (_Float16)0.f;
        } else if (mode == 2) {
            const int k = c / (GT / 8), rc = (c % (GT / 8)) * 8;
            const int gk = k0 + k, gr = r0 + rc;
            if (gk < K && gr + 7 < R) {
                t.v[i] = *reinterpret_cast<const uint4*>(P + gk * sk + gr);
                continue;
            }
#pragma unroll
            for (int e = 0; e < 8; e++) v[e] = (gk < K && gr + e < R) ? P[(gr + e) * sr + gk * sk] : (_Float16)0.f;
        } else {
            const int r = c / (GK / 8), kc = (c % (GK / 8)) * 8;
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const int gr = r0 + r, gk = k0 + kc + e;
                v[e] = (gr < R && gk < K) ? P[gr * sr + gk * sk] : (_Float16)0.f;
            }
        }
        t.v[i] = *reinterpret_cast<const uint4*>(v);
    }
}

template <int GT>
__device__ __forceinline__ void store_tile(_Float16* __restrict__ S, const TileRegs<GT>& t, int mode, int tid)
{
#pragma unroll
    for (int i = 0; i < TileRegs<GT>::N; i++) {
        const int c = tid + 256 * i;
        if (c >= TileRegs<GT>::CH) break;
        if (mode == 2) {
            const int k = c / (GT / 8), rc = (c % (GT / 8)) * 8;
            const _Float16* v = reinterpret_cast<const _Float16*>(&t.v[i]);
#pragma unroll
            for (int e = 0; e < 8; e++) S[(rc + e) * GKP + k] = v[e];
        } else {
            const int r = c / (GK / 8), kc = (c % (GK / 8)) * 8;
            uint2* d = reinterpret_cast<uint2*>(S + r * GKP + kc);
            d[0] = make_uint2(t.v[i].x, t.v[i].y);
            d[1] = make_uint2(t.v[i].z, t.v[i].w);
        }
    }
}

// C[M x N] (FP32) = A[M x K] . B[K x N] with A(m, k) = A[m * sam + k * sak] and
// B(k, n) = B[k * sbk + n * sbn]; the epilogue adds bias[n] when given and writes an FP16
// copy (ReLU'd when asked) with the same row stride when out16 is given.  A block of 4
// waves owns a GT x GT tile, each wave a (GT/2) x (GT/2) sub-tile of
// v_mfma_f32_16x16x16_f16 accumulators; k advances GK at a time through LDS, the next
// stage loaded into registers while the current one is multiplied.  amode / bmode: 1 = k contiguous, 2 = rows contiguous, 0 = general.
template <int GT>
__global__ __launch_bounds__(256) void gemm_f16_kernel(int M, int N, int K, const _Float16* A,
                                                       int64_t sam, int64_t sak, const _Float16* B,
                                                       int64_t sbk, int64_t sbn, float* C,
                                                       int64_t ldc, int accumulate, int amode, int bmode,
                                                       const float* bias, int relu, _Float16* out16)
{
    constexpr int WT = GT / 2, NT = WT / 16;
    __shared__ _Float16 As[GT * GKP];
    __shared__ _Float16 Bs[GT * GKP];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int bm = blockIdx.y * GT, bn = blockIdx.x * GT;
    const int wm = (w >> 1) * WT, wn = (w & 1) * WT;
    const int r16 = lane & 15, kq = (lane >> 4) * 4;
    floatx4v acc[NT][NT];
#pragma unroll
    for (int a = 0; a < NT; a++)
#pragma unroll
        for (int b = 0; b < NT; b++) acc[a][b] = floatx4v{0.f, 0.f, 0.f, 0.f};
    TileRegs<GT> ra, rb;
    fetch_tile<GT>(ra, A, sam, sak, bm, M, 0, K, amode, tid);
    fetch_tile<GT>(rb, B, sbn, sbk, bn, N, 0, K, bmode, tid);
    for (int k0 = 0; k0 < K; k0 += GK) {
        if (k0) __syncthreads();  // every wave is done with the previous stage
        store_tile<GT>(As, ra, amode, tid);
        store_tile<GT>(Bs, rb, bmode, tid);
        __syncthreads();
        if (k0 + GK < K) {  // next stage in flight while this one is multiplied
            fetch_tile<GT>(ra, A, sam, sak, bm, M, k0 + GK, K, amode, tid);
            fetch_tile<GT>(rb, B, sbn, sbk, bn, N, k0 + GK, K, bmode, tid);
        }
#pragma unroll
        for (int kk = 0; kk < GK; kk += 16) {
            half4 fa[NT], fb[NT];
#pragma unroll
            for (int a = 0; a < NT; a++)
                fa[a] = *reinterpret_cast<const half4*>(As + (wm + 16 * a + r16) * GKP + kk + kq);
#pragma unroll
            for (int b = 0; b < NT; b++)
                fb[b] = *reinterpret_cast<const half4*>(Bs + (wn + 16 * b + r16) * GKP + kk + kq);
#pragma unroll
            for (int a = 0; a < NT; a++)
#pragma unroll
                for (int b = 0; b < NT; b++)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x16f16(fa[a], fb[b], acc[a][b], 0, 0, 0);
        }
    }
#pragma unroll
    for (int a = 0; a < NT; a++)
#pragma unroll
        for (int b = 0; b < NT; b++) {
            const int n = bn + wn + 16 * b + r16;
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const int m = bm + wm + 16 * a + kq + t;
                if (m < M && n < N) {
                    float v = acc[a][b][t];
                    if (bias) v += bias[n];
                    if (C) {
                        float* p = C + (int64_t)m * ldc + n;
                        *p = accumulate ? *p + v : v;
                    }
                    if (out16) out16[(int64_t)m * ldc + n] = (_Float16)(relu ? fmaxf(v, 0.f) : v);
                }
            }
        }
}

// Two softmax heads per row: [0, L) and [L, L+NB).  Writes the mean-loss contribution
// and dlogits = (softmax - onehot) / B in FP16 (and FP32 when asked).
__global__ __launch_bounds__(256) void softmax_xent_kernel(const float* logits, int B, int L,
                                                           int NB, const int32_t* yl,
                                                           const int32_t* yb, float wb,
                                                           _Float16* dlog16, float* dlog32,
                                                           float* loss)
{
    __shared__ float red[256];
    const int row = blockIdx.x;
    const int O = L + NB;
    const float* z = logits + (int64_t)row * O;
    const int tid = threadIdx.x;
    for (int head = 0; head < 2; head++) {
        const int s = head == 0 ? 0 : L, e = head == 0 ? L : L + NB;
        const float wgt = head == 0 ? 1.f : wb;
        float mx = -__builtin_inff();
        for (int c = s + tid; c < e; c += 256) mx = fmaxf(mx, z[c]);
        red[tid] = mx;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if (tid < o) red[tid] = fmaxf(red[tid], red[tid + o]);
            __syncthreads();
        }
        mx = red[0];
        __syncthreads();
        float se = 0.f;
        for (int c = s + tid; c < e; c += 256) se += __expf(z[c] - mx);
        red[tid] = se;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if (tid < o) red[tid] += red[tid + o];
            __syncthreads();
        }
        se = red[0];
        __syncthreads();
        const int y = s + (head == 0 ? yl[row] : yb[row]);
        const float zy = z[y];
        __syncthreads();  // every thread has read z[y] before any gradient overwrites it
        for (int c = s + tid; c < e; c += 256) {
            const float p = __expf(z[c] - mx) / se;
            const float g = wgt * (p - (c == y ? 1.f : 0.f)) / (float)B;
            dlog16[(int64_t)row * O + c] = (_Float16)g;
            if (dlog32) dlog32[(int64_t)row * O + c] = g;
        }
        if (tid == 0 && loss) atomicAdd(loss, wgt * (logf(se) + mx - zy) / (float)B);
        __syncthreads();
    }
}

// dH = dH_raw * (pre > 0), FP16 copy for the next product
__global__ void relu_grad_kernel(const float* draw, const float* pre, int64_t n, float* d32,
                                 _Float16* d16)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = pre[i] > 0.f ? draw[i] : 0.f;
    if (d32) d32[i] = g;
    d16[i] = (_Float16)g;
}

// column sums of an M x N FP32 matrix (bias gradients), in two deterministic passes:
// block (x, y) sums rows y, y + CS_CHUNKS, ... of 64 columns with 4 interleaved row
// phases, then colsum_finish adds the CS_CHUNKS partials of each column in order.
constexpr int CS_CHUNKS = 16;

__global__ __launch_bounds__(256) void colsum_kernel(const float* G, int M, int N, float* part)
{
    __shared__ float red[4][64];
    const int c = threadIdx.x & 63, ph = threadIdx.x >> 6;
    const int n = blockIdx.x * 64 + c;
    float s = 0.f;
    if (n < N)
        for (int m = blockIdx.y + CS_CHUNKS * ph; m < M; m += 4 * CS_CHUNKS) s += G[(int64_t)m * N + n];
    red[ph][c] = s;
    __syncthreads();
    if (ph == 0 && n < N) part[(int64_t)blockIdx.y * N + n] = ((red[0][c] + red[1][c]) + red[2][c]) + red[3][c];
}

__global__ void colsum_finish_kernel(const float* part, int N, float* out)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    float s = 0.f;
    for (int y = 0; y < CS_CHUNKS; y++) s += part[(int64_t)y * N + n];
    out[n] = s;
}

__global__ void adam_kernel(float* w, const float* g, float* m1, float* m2, int64_t n, float lr,
                            float b1, float b2, float eps, float c1, float c2, _Float16* w16)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float a = b1 * m1[i] + (1.f - b1) * gi;
    const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    const float v = w[i] - lr * (a / c1) / (sqrtf(b / c2) + eps);
    w[i] = v;
    if (w16) w16[i] = (_Float16)v;
}

__global__ void f32_to_f16_kernel(const float* s, int64_t n, _Float16* d)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = (_Float16)s[i];
}

__global__ void gather_i32_kernel(const int32_t* src, const int32_t* idx, int B, int32_t* dst)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < B) dst[i] = src[idx[i]];
}

__global__ void gather_rows_f16_kernel(const float* X, int dim, const int32_t* idx, int B,
                                       _Float16* out)
{
    const int r = blockIdx.x;
    if (r >= B) return;
    const int64_t src = idx ? idx[r] : r;
    for (int c = threadIdx.x; c < dim; c += blockDim.x)
        out[(int64_t)r * dim + c] = (_Float16)X[src * dim + c];
}

// Inference head: per row, the h best landmarks (ids into `landmarks`, 1-based),
// the bucket (the most probable one, or with tau > 0 the smallest whose cumulative
// probability reaches tau), and the top landmark's softmax probability.  One wave per row.
__global__ __launch_bounds__(64) void topk_heads_kernel(const float* logits, int B, int L, int NB,
                                                        int h, float tau, const uint32_t* landmarks,
                                                        uint32_t* hints, int32_t* bucket,
                                                        float* conf)
{
    const int row = blockIdx.x;
    const int lane = threadIdx.x;
    const int O = L + NB;
    const float* z = logits + (int64_t)row * O;
    // one pass: running maximum with its first index and the softmax sum rescaled to it
    float mx = -__builtin_inff(), se = 0.f;
    int mi = 0x7fffffff;
    for (int c = lane; c < L; c += 64) {
        const float v = z[c];
        if (v > mx) {
            se = se * __expf(mx - v) + 1.f;
            mx = v;
            mi = c;
        } else {
            se += __expf(v - mx);
        }
    }
    for (int o = 32; o >= 1; o >>= 1) {
        const float om = __shfl_xor(mx, o), os = __shfl_xor(se, o);
        const int oi = __shfl_xor(mi, o);
        const float nm = fmaxf(mx, om);
        se = (mx == -__builtin_inff() ? 0.f : se * __expf(mx - nm)) + (om == -__builtin_inff() ? 0.f : os * __expf(om - nm));
        if (om > mx || (om == mx && oi < mi)) mi = oi;
        mx = nm;
    }
    if (lane == 0) {
        conf[row] = 1.f / se;
        hints[(int64_t)row * h] = landmarks[mi];
    }
    // further landmarks: rounds of argmax with exclusion (h is small)
    float last = mx;
    int lastidx = mi;
    for (int t = 1; t < h; t++) {
        float bv = -__builtin_inff();
        int bi = 0x7fffffff;
        for (int c = lane; c < L; c += 64) {
            const float v = z[c];
            const bool allowed = v < last || (v == last && c > lastidx);
            if (allowed && (v > bv || (v == bv && c < bi))) {
                bv = v;
                bi = c;
            }
        }
        for (int o = 32; o >= 1; o >>= 1) {
            const float ov = __shfl_xor(bv, o);
            const int oi = __shfl_xor(bi, o);
            if (ov > bv || (ov == bv && oi < bi)) {
                bv = ov;
                bi = oi;
            }
        }
        if (lane == 0) hints[(int64_t)row * h + t] = landmarks[bi];
        last = bv;
        lastidx = bi;
    }
    if (lane == 0) {
        float bb = -__builtin_inff();
        int bix = 0;
        for (int c = 0; c < NB; c++)
            if (z[L + c] > bb) {
                bb = z[L + c];
                bix = c;
            }
        if (tau > 0.f) {
            // smallest bucket whose cumulative predicted probability reaches tau
            float tot = 0.f;
            for (int c = 0; c < NB; c++) tot += __expf(z[L + c] - bb);
            float acc = 0.f;
            bix = NB - 1;
            for (int c = 0; c < NB; c++) {
                acc += __expf(z[L + c] - bb);
                if (acc >= tau * tot) {
                    bix = c;
                    break;
                }
            }
        }
        bucket[row] = bix;
    }
}

// Fused output layer of inference (hidden width K in {64, 128, 192, 256}, H <= 4 hints,
// NB <= 16 buckets): the logits z = A2 . W3^T + b3 of a block of 128 queries are formed 32
// output units at a time (W3 tiles staged in LDS, double-buffered) and folded at once into
// each query's H best landmark units (value descending, index ascending), the softmax sum
// of its landmark head and its bucket logits; none of the L + NB logits per query is
// stored.  The products are those of gemm_f16_kernel (v_mfma_f32_16x16x16f16 in ascending
// k, queries as the A operand and units as B, the bias added after), so the logits, the
// landmarks and the buckets equal the unfused path's; the softmax sum (the confidence) is
// added in another order.
//
// A lane holds, per 16 x 16 product, 4 queries (rows 4g .. 4g+3 of a query tile) of one
// unit (column `col`), so it folds units col, col + 16, ... of each of its queries; the 16
// lanes of a group are merged by butterfly at the end.  W3 rows are staged in LDS with the
// 4 k of each lane group and product placed next to each other, so a lane reads its share
// of a unit row as one run of K / 2 bytes.
constexpr int HT_TU = 32;  // output units per LDS tile
constexpr int HT_NQ = 2;   // 16-query tiles per wave: 32 queries per wave, 128 per block

template <int H>
__device__ __forceinline__ void top_insert(float (&tv)[H], int (&ti)[H], float v, int u)
{
#pragma unroll
    for (int j = 0; j < H; j++) {
        const bool sw = v > tv[j] || (v == tv[j] && u < ti[j]);
        const float tf = tv[j];
        const int tn = ti[j];
        tv[j] = sw ? v : tf;
        ti[j] = sw ? u : tn;
        v = sw ? tf : v;
        u = sw ? tn : u;
    }
}

template <int K, int H>
__global__ __launch_bounds__(256, 2) void head_top_kernel(const _Float16* __restrict__ A2, int B,
                                                          const _Float16* __restrict__ W,
                                                          const float* __restrict__ bias, int L, int NB,
                                                          float tau, const uint32_t* __restrict__ landmarks,
                                                          uint32_t* __restrict__ hints,
                                                          int32_t* __restrict__ bucket, float* __restrict__ conf)
{
    constexpr int KS = K / 16;         // products per 16 x 16 output tile
    constexpr int LB = K / 2;          // bytes of a lane group's share of a unit row
    constexpr int RS = 2 * K + 16;     // LDS bytes per unit row (16 bytes of padding)
    constexpr int CH = HT_TU * K / 8;  // 16-byte chunks per tile
    constexpr int PF = CH / 256;       // chunks per thread
    static_assert(K % 64 == 0 && K <= 256 && CH % 256 == 0 && PF <= 4, "hidden width");
    __shared__ __attribute__((aligned(16))) uint8_t s_w[2][HT_TU * RS];
    __shared__ float s_b[2][HT_TU];
    __shared__ float s_zb[4][16 * HT_NQ][16];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int col = lane & 15, g = lane >> 4;
    const int O = L + NB;
    const int ntiles = (O + HT_TU - 1) / HT_TU;
    const int64_t wq0 = (int64_t)blockIdx.x * (64 * HT_NQ) + (int64_t)w * (16 * HT_NQ);

    // the wave's queries as A operands: row col of query tile q, k = 16 s + 4 g .. + 3
    half4 aq[HT_NQ][KS];
#pragma unroll
    for (int q = 0; q < HT_NQ; q++) {
        const int64_t r = min(wq0 + 16 * q + col, (int64_t)B - 1);
        const _Float16* p = A2 + r * K + 4 * g;
#pragma unroll
        for (int s = 0; s < KS; s++) aq[q][s] = *reinterpret_cast<const half4*>(p + 16 * s);
    }
    // per (query tile, row i): the H best landmark units and the softmax sum
    float tv[HT_NQ][4][H], se[HT_NQ][4];
    int ti[HT_NQ][4][H];
#pragma unroll
    for (int q = 0; q < HT_NQ; q++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            se[q][i] = 0.f;
#pragma unroll
            for (int j = 0; j < H; j++) {
                tv[q][i][j] = -__builtin_inff();
                ti[q][i][j] = 0x7fffffff;
            }
        }

    uint4 pf[PF];
    float pfb = 0.f;
    auto prefetch = [&](int t) {
        const int u0 = t * HT_TU;
#pragma unroll
        for (int i = 0; i < PF; i++) {
            const int c = tid + 256 * i, ur = c / (K / 8), cr = c % (K / 8);
            const int64_t u = min(u0 + ur, O - 1);  // units past O: see the epilogue
            pf[i] = reinterpret_cast<const uint4*>(W + u * K)[cr];
        }
        if (tid < HT_TU) pfb = bias[min(u0 + tid, O - 1)];
    };
    // chunk cr holds k = 8 cr .. 8 cr + 7 of its row: product cr / 2, lane groups
    // 2 (cr % 2) (its first 4 k) and 2 (cr % 2) + 1 (the last 4)
    auto stage = [&](int b) {
#pragma unroll
        for (int i = 0; i < PF; i++) {
            const int c = tid + 256 * i, ur = c / (K / 8), cr = c % (K / 8);
            uint8_t* rp = s_w[b] + ur * RS + (cr >> 1) * 8;
            *reinterpret_cast<uint2*>(rp + (2 * (cr & 1)) * LB) = make_uint2(pf[i].x, pf[i].y);
            *reinterpret_cast<uint2*>(rp + (2 * (cr & 1) + 1) * LB) = make_uint2(pf[i].z, pf[i].w);
        }
        if (tid < HT_TU) s_b[b][tid] = pfb;
    };

    prefetch(0);
    stage(0);
    __syncthreads();
    int b = 0;
    for (int t = 0; t < ntiles; t++) {
        const bool more = t + 1 < ntiles;
        if (more) prefetch(t + 1);
#pragma unroll
        for (int us = 0; us < HT_TU / 16; us++) {
            const uint8_t* rp = s_w[b] + (us * 16 + col) * RS + g * LB;
            floatx4v acc[HT_NQ];
#pragma unroll
            for (int q = 0; q < HT_NQ; q++) acc[q] = floatx4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int s2 = 0; s2 < KS / 2; s2++) {
                const uint4 v = *reinterpret_cast<const uint4*>(rp + 16 * s2);
                const half4 b0 = __builtin_bit_cast(half4, make_uint2(v.x, v.y));
                const half4 b1 = __builtin_bit_cast(half4, make_uint2(v.z, v.w));
#pragma unroll
                for (int q = 0; q < HT_NQ; q++) {
                    acc[q] = __builtin_amdgcn_mfma_f32_16x16x16f16(aq[q][2 * s2], b0, acc[q], 0, 0, 0);
                    acc[q] = __builtin_amdgcn_mfma_f32_16x16x16f16(aq[q][2 * s2 + 1], b1, acc[q], 0, 0, 0);
                }
            }
            const int u = t * HT_TU + us * 16 + col;  // this lane's unit
            const float bu = s_b[b][us * 16 + col];
            // online softmax sum against the running maximum (e = exp(-|v - max|)), then the
            // top-H list; u exceeds every unit the lane has folded, so a tie keeps the entry
            auto fold = [&](int q, int i, float v) {
                const float m0 = tv[q][i][0];
                const bool up = v > m0;
                const float e = __expf(-__builtin_fabsf(v - m0));
                se[q][i] = up ? __builtin_fmaf(se[q][i], e, 1.f) : se[q][i] + e;
                if constexpr (H == 1) {
                    ti[q][i][0] = up ? u : ti[q][i][0];
                    tv[q][i][0] = up ? v : m0;
                } else {
                    top_insert<H>(tv[q][i], ti[q][i], v, u);
                }
            };
            if (H == 1 && t * HT_TU + us * 16 + 16 <= L) {  // every unit a landmark
#pragma unroll
                for (int q = 0; q < HT_NQ; q++)
#pragma unroll
                    for (int i = 0; i < 4; i++) fold(q, i, acc[q][i] + bu);
            } else {
                const bool land = u < L, buck = u >= L && u < O;
#pragma unroll
                for (int q = 0; q < HT_NQ; q++)
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        const float v = acc[q][i] + bu;
                        if (buck) s_zb[w][16 * q + 4 * g + i][u - L] = v;
                        if (land) fold(q, i, v);
                    }
            }
        }
        if (more) stage(b ^ 1);
        __syncthreads();
        b ^= 1;
    }
    // merge the 16 lanes of each group (units of every residue mod 16)
#pragma unroll
    for (int q = 0; q < HT_NQ; q++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
#pragma unroll
            for (int off = 1; off < 16; off <<= 1) {
                const float ose = __shfl_xor(se[q][i], off);
                float ov[H];
                int oi[H];
#pragma unroll
                for (int j = 0; j < H; j++) {
                    ov[j] = __shfl_xor(tv[q][i][j], off);
                    oi[j] = __shfl_xor(ti[q][i][j], off);
                }
                const float ma = tv[q][i][0], mb = ov[0], nm = fmaxf(ma, mb);
                se[q][i] = (ma == -__builtin_inff() ? 0.f : se[q][i] * __expf(ma - nm)) +
                           (mb == -__builtin_inff() ? 0.f : ose * __expf(mb - nm));
#pragma unroll
                for (int j = 0; j < H; j++) top_insert<H>(tv[q][i], ti[q][i], ov[j], oi[j]);
            }
        }
    // lane col < 4 of group g writes query row 4g + col of each query tile
#pragma unroll
    for (int q = 0; q < HT_NQ; q++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int64_t row = wq0 + 16 * q + 4 * g + i;
            if (col != i || row >= B) continue;
            conf[row] = 1.f / se[q][i];
#pragma unroll
            for (int j = 0; j < H; j++) hints[row * H + j] = landmarks[ti[q][i][j]];
            const float* z = s_zb[w][16 * q + 4 * g + i];
            float bb = -__builtin_inff();
            int bix = 0;
            for (int c = 0; c < NB; c++)
                if (z[c] > bb) {
                    bb = z[c];
                    bix = c;
                }
            if (tau > 0.f) {
                // smallest bucket whose cumulative predicted probability reaches tau
                float tot = 0.f;
                for (int c = 0; c < NB; c++) tot += __expf(z[c] - bb);
                float ac = 0.f;
                bix = NB - 1;
                for (int c = 0; c < NB; c++) {
                    ac += __expf(z[c] - bb);
                    if (ac >= tau * tot) {
                        bix = c;
                        break;
                    }
                }
            }
            bucket[row] = bix;
        }
}

struct BucketTable {
    int32_t v[16];
};

__global__ void plan_kernel(const int32_t* bucket, const float* conf, int64_t m, int32_t nh,
                            float min_conf, BucketTable table, int32_t nb, int32_t fallback,
                            int32_t* nhints, int32_t* bsize)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const bool ok = conf[i] >= min_conf;
    const int b = bucket[i];
    nhints[i] = ok ? nh : -1;
    bsize[i] = ok && b >= 0 && b < nb ? table.v[b] : fallback;
}

}  // namespace ssg

using namespace ssg;

extern "C" int ssg_learned_plan_dev(ssg_handle* h, const int32_t* dbucket, const float* dconf,
                                    int64_t m, int32_t nh, float min_conf,
                                    const int32_t* bucket_bsize, int32_t nb, int32_t fallback_bsize,
                                    int32_t* dnhints, int32_t* dbsize, void* stream)
{
    if (!h || !dbucket || !dconf || !dnhints || !dbsize || !bucket_bsize || m < 1 || nb < 1 || nb > 16)
        return ssg_fail(SSG_E_ARG, "bad plan arguments");
    BucketTable t{};
    for (int i = 0; i < nb; i++) {
        if (bucket_bsize[i] < 1) return ssg_fail(SSG_E_ARG, "bucket beam sizes must be >= 1");
        t.v[i] = bucket_bsize[i];
    }
    if (fallback_bsize < 1) return ssg_fail(SSG_E_ARG, "fallback bsize must be >= 1");
    HIPCHK(hipSetDevice(h->device), "hipSetDevice");
    hipStream_t s = stream ? (hipStream_t)stream : h->stream;
    hipLaunchKernelGGL(plan_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, dbucket,
                       dconf, m, nh, min_conf, t, nb, fallback_bsize, dnhints, dbsize);
    HIPCHK(hipGetLastError(), "plan");
    return SSG_OK;
}

struct ssg_mlp;
extern "C" int ssg_mlp_infer_dev(ssg_mlp* p, const float* dQ, int64_t m, const uint32_t* dlandmarks,
                                 int32_t h, float bucket_quantile, uint32_t* dhints, int32_t* dbucket,
                                 float* dconf, void* stream);

// Learned search: inference -> plan -> beam search, one stream, handle-owned scratch.
extern "C" int ssg_learned_search_dev(ssg_handle* h, ssg_mlp* mlp, const uint32_t* dlandmarks,
                                      const float* dQ, int64_t m, int32_t k, int32_t kstride,
                                      const ssg_beam_params* bp, int32_t nh, float min_conf,
                                      float bucket_quantile, const int32_t* bucket_bsize, int32_t nb,
                                      uint32_t* dids, float* ddists, int64_t* dcd, int64_t* dcb,
                                      void* stream)
{
    if (!h || !mlp || !dlandmarks || !dQ || !bp || !bucket_bsize) return ssg_fail(SSG_E_ARG, "null argument");
    if (m < 1 || nh < 1 || nh > 64) return ssg_fail(SSG_E_ARG, "bad learned-search arguments");
    hipStream_t s = stream ? (hipStream_t)stream : h->stream;
    {
        std::lock_guard<std::mutex> lk(h->mu);
        HIPCHK(hipSetDevice(h->device), "hipSetDevice");
        HIPCHK(h->lH.ensure((size_t)m * nh * 4), "alloc");
        HIPCHK(h->lB.ensure((size_t)m * 4), "alloc");
        HIPCHK(h->lC.ensure((size_t)m * 4), "alloc");
        HIPCHK(h->lN.ensure((size_t)m * 4), "alloc");
        HIPCHK(h->lS.ensure((size_t)m * 4), "alloc");
        HIPCHK(h->lO.ensure((size_t)m * 4), "alloc");
        HIPCHK(h->lCnt.ensure(64 * 4), "alloc");
    }
    int rc = ssg_mlp_infer_dev(mlp, dQ, m, dlandmarks, nh, bucket_quantile, h->lH.as<uint32_t>(),
                               h->lB.as<int32_t>(), h->lC.as<float>(), s);
    if (rc) return rc;
    rc = ssg_learned_plan_dev(h, h->lB.as<int32_t>(), h->lC.as<float>(), m, nh, min_conf,
                              bucket_bsize, nb, bp->bsize, h->lN.as<int32_t>(), h->lS.as<int32_t>(), s);
    if (rc) return rc;
    HIPCHK(launch_order_desc(h->lS.as<int32_t>(), m, h->lO.as<uint32_t>(), h->lCnt.as<unsigned int>(), s),
           "order");
    ssg_query_opts qo{nullptr, h->lS.as<int32_t>(), h->lH.as<uint32_t>(), h->lN.as<int32_t>(), nh,
                      h->lO.as<uint32_t>()};
    return ssg_beamsearch_dev(h, dQ, m, k, kstride, bp, &qo, dids, ddists, dcd, dcb, s);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct ssg_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    int D = 0, H1 = 0, H2 = 0, L = 0, NB = 0;
    // parameters: FP32 master, FP16 operand copy, Adam moments, gradients
    DevBuf W[3], Wh[3], b[3], mW[3], vW[3], mb[3], vb[3], gW[3], gb[3];
    // per-batch work
    DevBuf X16, C1, P1, A1, C2, P2, A2, Z, dZ16, dA32, dA16, loss;
    DevBuf xin, yl, yb, perm;  // training inputs of one batch / one epoch
    DevBuf colpart;            // column-sum partials
    DevBuf X16i, A1i, A2i;     // inference activations (fused output layer)
    int cap_b = 0, cap_i = 0;
    int64_t step = 0;
    std::vector<uint32_t> landmarks;
    DevBuf dland;
};

namespace {

int dims_out(const ssg_mlp* p, int layer) { return layer == 0 ? p->H1 : layer == 1 ? p->H2 : p->L + p->NB; }
int dims_in(const ssg_mlp* p, int layer) { return layer == 0 ? p->D : layer == 1 ? p->H1 : p->H2; }

hipError_t colsum(ssg_mlp* p, hipStream_t s, const float* G, int M, int N, float* out)
{
    hipError_t e = p->colpart.ensure((size_t)CS_CHUNKS * N * sizeof(float));
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(colsum_kernel, dim3((unsigned)((N + 63) / 64), CS_CHUNKS), dim3(256), 0, s, G, M, N,
                       p->colpart.as<float>());
    hipLaunchKernelGGL(colsum_finish_kernel, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s,
                       p->colpart.as<float>(), N, out);
    return hipGetLastError();
}

// operand staging mode: 1 when k is contiguous, 2 when rows are, with 16-byte aligned
// 8-half runs; 0 otherwise
int stage_mode(const _Float16* P, int64_t sr, int64_t sk)
{
    const bool al = ((uintptr_t)P & 15) == 0;
    if (al && sk == 1 && sr % 8 == 0) return 1;
    if (al && sr == 1 && sk % 8 == 0) return 2;
    return 0;
}

hipError_t gemm(hipStream_t s, int M, int N, int K, const _Float16* A, int64_t sam, int64_t sak,
                const _Float16* B, int64_t sbk, int64_t sbn, float* C, int64_t ldc, int acc = 0,
                const float* bias = nullptr, int relu = 0, _Float16* out16 = nullptr)
{
    const int am = stage_mode(A, sam, sak), bmd = stage_mode(B, sbn, sbk);
    const int64_t big = (int64_t)((N + 127) / 128) * ((M + 127) / 128);
    if (big >= 256) {
        const dim3 grid((unsigned)((N + 127) / 128), (unsigned)((M + 127) / 128));
        hipLaunchKernelGGL(gemm_f16_kernel<128>, grid, dim3(256), 0, s, M, N, K, A, sam, sak, B, sbk,
                           sbn, C, ldc, acc, am, bmd, bias, relu, out16);
    } else {
        const dim3 grid((unsigned)((N + 63) / 64), (unsigned)((M + 63) / 64));
        hipLaunchKernelGGL(gemm_f16_kernel<64>, grid, dim3(256), 0, s, M, N, K, A, sam, sak, B, sbk,
                           sbn, C, ldc, acc, am, bmd, bias, relu, out16);
    }
    return hipGetLastError();
}

unsigned blocks_for(int64_t n) { return (unsigned)((n + 255) / 256); }

int ensure_batch(ssg_mlp* p, int B)
{
    if (B <= p->cap_b) return SSG_OK;
    const int O = p->L + p->NB;
    HIPCHK(p->X16.ensure((size_t)B * p->D * 2), "alloc");
    HIPCHK(p->C1.ensure((size_t)B * p->H1 * 4), "alloc");
    HIPCHK(p->P1.ensure((size_t)B * p->H1 * 4), "alloc");
    HIPCHK(p->A1.ensure((size_t)B * p->H1 * 2), "alloc");
    HIPCHK(p->C2.ensure((size_t)B * p->H2 * 4), "alloc");
    HIPCHK(p->P2.ensure((size_t)B * p->H2 * 4), "alloc");
    HIPCHK(p->A2.ensure((size_t)B * p->H2 * 2), "alloc");
    HIPCHK(p->Z.ensure((size_t)B * O * 4), "alloc");
    HIPCHK(p->dZ16.ensure((size_t)B * O * 2), "alloc");
    const int hmax = std::max(p->H1, p->H2);
    HIPCHK(p->dA32.ensure((size_t)B * hmax * 4), "alloc");
    HIPCHK(p->dA16.ensure((size_t)B * hmax * 2), "alloc");
    HIPCHK(p->loss.ensure(16), "alloc");
    p->cap_b = B;
    return SSG_OK;
}

// forward over X16 (B rows) -> logits Z
int forward(ssg_mlp* p, int B)
{
    hipStream_t s = p->stream;
    const int O = p->L + p->NB;
    // pre-activations (bias added) into P1 / P2, ReLU'd FP16 activations into A1 / A2
    HIPCHK(gemm(s, B, p->H1, p->D, p->X16.as<_Float16>(), p->D, 1, p->Wh[0].as<_Float16>(), 1, p->D,
                p->P1.as<float>(), p->H1, 0, p->b[0].as<float>(), 1, p->A1.as<_Float16>()), "gemm");
    HIPCHK(gemm(s, B, p->H2, p->H1, p->A1.as<_Float16>(), p->H1, 1, p->Wh[1].as<_Float16>(), 1, p->H1,
                p->P2.as<float>(), p->H2, 0, p->b[1].as<float>(), 1, p->A2.as<_Float16>()), "gemm");
    HIPCHK(gemm(s, B, O, p->H2, p->A2.as<_Float16>(), p->H2, 1, p->Wh[2].as<_Float16>(), 1, p->H2,
                p->Z.as<float>(), O, 0, p->b[2].as<float>()), "gemm");
    return SSG_OK;
}

// backward from dZ16 (B rows); fills gW / gb
int backward(ssg_mlp* p, int B)
{
    hipStream_t s = p->stream;
    const int O = p->L + p->NB;
    // layer 3: gW3[O x H2] = dZ^T . A2 ; dA2 = dZ . W3
    HIPCHK(gemm(s, O, p->H2, B, p->dZ16.as<_Float16>(), 1, O, p->A2.as<_Float16>(), p->H2, 1,
                p->gW[2].as<float>(), p->H2), "gemm");
    HIPCHK(gemm(s, B, p->H2, O, p->dZ16.as<_Float16>(), O, 1, p->Wh[2].as<_Float16>(), p->H2, 1,
                p->dA32.as<float>(), p->H2), "gemm");
    // gb3 = column sums of the FP32 logit gradient, which softmax_xent left in Z
    HIPCHK(colsum(p, s, p->Z.as<float>(), B, O, p->gb[2].as<float>()), "colsum");
    hipLaunchKernelGGL(relu_grad_kernel, dim3(blocks_for((int64_t)B * p->H2)), dim3(256), 0, s,
                       p->dA32.as<float>(), p->P2.as<float>(), (int64_t)B * p->H2,
                       p->C2.as<float>(), p->dA16.as<_Float16>());
    HIPCHK(colsum(p, s, p->C2.as<float>(), B, p->H2, p->gb[1].as<float>()), "colsum");
    // layer 2: gW2[H2 x H1] = dH2^T . A1 ; dA1 = dH2 . W2
    HIPCHK(gemm(s, p->H2, p->H1, B, p->dA16.as<_Float16>(), 1, p->H2, p->A1.as<_Float16>(), p->H1, 1,
                p->gW[1].as<float>(), p->H1), "gemm");
    HIPCHK(gemm(s, B, p->H1, p->H2, p->dA16.as<_Float16>(), p->H2, 1, p->Wh[1].as<_Float16>(), p->H1, 1,
                p->dA32.as<float>(), p->H1), "gemm");
    // reuse dA16 for dH1 after the product above consumed dH2 (stream order)
    hipLaunchKernelGGL(relu_grad_kernel, dim3(blocks_for((int64_t)B * p->H1)), dim3(256), 0, s,
                       p->dA32.as<float>(), p->P1.as<float>(), (int64_t)B * p->H1,
                       p->C1.as<float>(), p->dA16.as<_Float16>());
    HIPCHK(colsum(p, s, p->C1.as<float>(), B, p->H1, p->gb[0].as<float>()), "colsum");
    // layer 1: gW1[H1 x D] = dH1^T . X
    HIPCHK(gemm(s, p->H1, p->D, B, p->dA16.as<_Float16>(), 1, p->H1, p->X16.as<_Float16>(), p->D, 1,
                p->gW[0].as<float>(), p->D), "gemm");
    return SSG_OK;
}

uint64_t splitmix(uint64_t& s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

}  // namespace

extern "C" {

int ssg_mlp_create(ssg_mlp** out, int device, int32_t D, int32_t H1, int32_t H2, int32_t L,
                   int32_t NB, uint64_t seed)
{
    if (!out) return ssg_fail(SSG_E_ARG, "null output");
    *out = nullptr;
    if (D < 1 || H1 < 1 || H2 < 1 || L < 1 || NB < 1) return ssg_fail(SSG_E_ARG, "bad MLP shape");
    HIPCHK(hipSetDevice(device), "hipSetDevice");
    ssg_mlp* p = new ssg_mlp();
    p->device = device;
    p->D = D;
    p->H1 = H1;
    p->H2 = H2;
    p->L = L;
    p->NB = NB;
    hipError_t e = hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete p;
        return ssg_hipfail(e, "stream");
    }
    // He-uniform init from a splitmix stream
    uint64_t st = seed;
    for (int l = 0; l < 3; l++) {
        const int fo = dims_out(p, l), fi = dims_in(p, l);
        std::vector<float> w((size_t)fo * fi), z((size_t)fo * fi, 0.f), bz((size_t)fo, 0.f);
        const float lim = sqrtf(6.f / (float)fi);
        for (auto& x : w) x = lim * (2.f * (float)((splitmix(st) >> 11) * (1.0 / 9007199254740992.0)) - 1.f);
        const size_t nw = w.size() * 4, nb = bz.size() * 4;
        HIPCHK(p->W[l].ensure(nw), "alloc");
        HIPCHK(p->Wh[l].ensure(w.size() * 2), "alloc");
        HIPCHK(p->mW[l].ensure(nw), "alloc");
        HIPCHK(p->vW[l].ensure(nw), "alloc");
        HIPCHK(p->gW[l].ensure(nw), "alloc");
        HIPCHK(p->b[l].ensure(nb), "alloc");
        HIPCHK(p->mb[l].ensure(nb), "alloc");
        HIPCHK(p->vb[l].ensure(nb), "alloc");
        HIPCHK(p->gb[l].ensure(nb), "alloc");
        HIPCHK(hipMemcpy(p->W[l].p, w.data(), nw, hipMemcpyHostToDevice), "copy");
        HIPCHK(hipMemcpy(p->mW[l].p, z.data(), nw, hipMemcpyHostToDevice), "copy");
        HIPCHK(hipMemcpy(p->vW[l].p, z.data(), nw, hipMemcpyHostToDevice), "copy");
        HIPCHK(hipMemcpy(p->b[l].p, bz.data(), nb, hipMemcpyHostToDevice), "copy");
        HIPCHK(hipMemcpy(p->mb[l].p, bz.data(), nb, hipMemcpyHostToDevice), "copy");
        HIPCHK(hipMemcpy(p->vb[l].p, bz.data(), nb, hipMemcpyHostToDevice), "copy");
        hipLaunchKernelGGL(f32_to_f16_kernel, dim3(blocks_for((int64_t)w.size())), dim3(256), 0,
                           p->stream, p->W[l].as<float>(), (int64_t)w.size(), p->Wh[l].as<_Float16>());
    }
    HIPCHK(hipStreamSynchronize(p->stream), "sync");
    *out = p;
    return SSG_OK;
}

int ssg_mlp_destroy(ssg_mlp* p)
{
    if (!p) return SSG_OK;
    (void)hipSetDevice(p->device);
    (void)hipStreamSynchronize(p->stream);
    for (int l = 0; l < 3; l++) {
        DevBuf* bs[] = {&p->W[l], &p->Wh[l], &p->b[l], &p->mW[l], &p->vW[l], &p->mb[l], &p->vb[l],
                        &p->gW[l], &p->gb[l]};
        for (DevBuf* b : bs) b->release();
    }
    DevBuf* ws[] = {&p->X16, &p->C1, &p->P1, &p->A1, &p->C2, &p->P2, &p->A2, &p->Z, &p->dZ16,
                    &p->dA32, &p->dA16, &p->loss, &p->dland, &p->xin, &p->yl, &p->yb, &p->perm};
    for (DevBuf* b : ws) b->release();
    (void)hipStreamDestroy(p->stream);
    delete p;
    return SSG_OK;
}

// Parameters in and out (FP32, row-major W[l] as [out x in], then b[l]); layer = 0..2.
int ssg_mlp_get(ssg_mlp* p, int32_t layer, float* W, float* b)
{
    if (!p || layer < 0 || layer > 2) return ssg_fail(SSG_E_ARG, "bad layer");
    const size_t nw = (size_t)dims_out(p, layer) * dims_in(p, layer);
    if (W) HIPCHK(hipMemcpy(W, p->W[layer].p, nw * 4, hipMemcpyDeviceToHost), "copy");
    if (b) HIPCHK(hipMemcpy(b, p->b[layer].p, (size_t)dims_out(p, layer) * 4, hipMemcpyDeviceToHost), "copy");
    return SSG_OK;
}

int ssg_mlp_set(ssg_mlp* p, int32_t layer, const float* W, const float* b)
{
    if (!p || layer < 0 || layer > 2) return ssg_fail(SSG_E_ARG, "bad layer");
    const size_t nw = (size_t)dims_out(p, layer) * dims_in(p, layer);
    if (W) {
        HIPCHK(hipMemcpy(p->W[layer].p, W, nw * 4, hipMemcpyHostToDevice), "copy");
        hipLaunchKernelGGL(f32_to_f16_kernel, dim3(blocks_for((int64_t)nw)), dim3(256), 0, p->stream,
                           p->W[layer].as<float>(), (int64_t)nw, p->Wh[layer].as<_Float16>());
        HIPCHK(hipStreamSynchronize(p->stream), "sync");
    }
    if (b) HIPCHK(hipMemcpy(p->b[layer].p, b, (size_t)dims_out(p, layer) * 4, hipMemcpyHostToDevice), "copy");
    return SSG_OK;
}

// One forward + backward over a batch (rows of X, FP32 [B x D] host array, labels);
// returns the loss and leaves the gradients readable through ssg_mlp_grad.  `update`
// applies one Adam step (lr, beta1 0.9, beta2 0.999, eps 1e-8).
}  // extern "C"

// forward, loss, backward and (optionally) an Adam step on the batch already in
// X16 / yl / yb; the loss accumulates into p->loss.
static int train_batch(ssg_mlp* p, int B, float bucket_weight, float lr, bool update)
{
    hipStream_t s = p->stream;
    int rc = forward(p, B);
    if (rc) return rc;
    hipLaunchKernelGGL(softmax_xent_kernel, dim3(B), dim3(256), 0, s, p->Z.as<float>(), B, p->L,
                       p->NB, p->yl.as<int32_t>(), p->yb.as<int32_t>(), bucket_weight,
                       p->dZ16.as<_Float16>(), p->Z.as<float>(), p->loss.as<float>());
    rc = backward(p, B);
    if (rc) return rc;
    if (update) {
        p->step++;
        const float c1 = 1.f - powf(0.9f, (float)p->step), c2 = 1.f - powf(0.999f, (float)p->step);
        for (int l = 0; l < 3; l++) {
            const int64_t nw = (int64_t)dims_out(p, l) * dims_in(p, l), nb = dims_out(p, l);
            hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(nw)), dim3(256), 0, s, p->W[l].as<float>(),
                               p->gW[l].as<float>(), p->mW[l].as<float>(), p->vW[l].as<float>(), nw, lr,
                               0.9f, 0.999f, 1e-8f, c1, c2, p->Wh[l].as<_Float16>());
            hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(nb)), dim3(256), 0, s, p->b[l].as<float>(),
                               p->gb[l].as<float>(), p->mb[l].as<float>(), p->vb[l].as<float>(), nb, lr,
                               0.9f, 0.999f, 1e-8f, c1, c2, (_Float16*)nullptr);
        }
    }
    HIPCHK(hipGetLastError(), "train step");
    return SSG_OK;
}

extern "C" {

int ssg_mlp_step(ssg_mlp* p, const float* X, const int32_t* ylandmark, const int32_t* ybucket,
                 int32_t B, float bucket_weight, float lr, int32_t update, float* loss_out)
{
    if (!p || !X || !ylandmark || !ybucket || B < 1) return ssg_fail(SSG_E_ARG, "bad batch");
    HIPCHK(hipSetDevice(p->device), "hipSetDevice");
    int rc = ensure_batch(p, B);
    if (rc) return rc;
    hipStream_t s = p->stream;
    HIPCHK(p->xin.ensure((size_t)B * p->D * 4), "alloc");
    HIPCHK(p->yl.ensure((size_t)B * 4), "alloc");
    HIPCHK(p->yb.ensure((size_t)B * 4), "alloc");
    for (int64_t i = 0; i < B; i++)
        if (ylandmark[i] < 0 || ylandmark[i] >= p->L || ybucket[i] < 0 || ybucket[i] >= p->NB)
            return ssg_fail(SSG_E_ARG, "label out of range");
    HIPCHK(hipMemcpyAsync(p->xin.p, X, (size_t)B * p->D * 4, hipMemcpyHostToDevice, s), "copy");
    HIPCHK(hipMemcpyAsync(p->yl.p, ylandmark, (size_t)B * 4, hipMemcpyHostToDevice, s), "copy");
    HIPCHK(hipMemcpyAsync(p->yb.p, ybucket, (size_t)B * 4, hipMemcpyHostToDevice, s), "copy");
    hipLaunchKernelGGL(gather_rows_f16_kernel, dim3(B), dim3(128), 0, s, p->xin.as<float>(), p->D,
                       (const int32_t*)nullptr, B, p->X16.as<_Float16>());
    HIPCHK(launch_fill_f32(p->loss.as<float>(), 1, 0.f, s), "fill");
    rc = train_batch(p, B, bucket_weight, lr, update != 0);
    if (rc) return rc;
    if (loss_out) HIPCHK(hipMemcpyAsync(loss_out, p->loss.p, 4, hipMemcpyDeviceToHost, s), "copy");
    HIPCHK(hipStreamSynchronize(s), "sync");
    return SSG_OK;
}

// Mini-batch Adam over a device-resident training set: dX [N x D] FP32 rows, dyl / dyb
// labels.  Each epoch visits the rows in a fresh splitmix-shuffled order.
// losses[epochs] (host, nullable) receives the mean batch loss per epoch.
int ssg_mlp_train_dev(ssg_mlp* p, const float* dX, const int32_t* dyl, const int32_t* dyb, int64_t N,
                      int32_t epochs, int32_t batch, float bucket_weight, float lr, uint64_t seed,
                      float* losses)
{
    if (!p || !dX || !dyl || !dyb || N < 1 || epochs < 1 || batch < 1)
        return ssg_fail(SSG_E_ARG, "bad training arguments");
    HIPCHK(hipSetDevice(p->device), "hipSetDevice");
    int rc = ensure_batch(p, batch);
    if (rc) return rc;
    hipStream_t s = p->stream;
    HIPCHK(p->yl.ensure((size_t)batch * 4), "alloc");
    HIPCHK(p->yb.ensure((size_t)batch * 4), "alloc");
    HIPCHK(p->perm.ensure((size_t)N * 4), "alloc");
    std::vector<int32_t> perm((size_t)N);
    uint64_t st = seed;
    for (int ep = 0; ep < epochs; ep++) {
        for (int64_t i = 0; i < N; i++) perm[i] = (int32_t)i;
        for (int64_t i = N - 1; i > 0; i--) std::swap(perm[i], perm[splitmix(st) % (uint64_t)(i + 1)]);
        HIPCHK(hipMemcpyAsync(p->perm.p, perm.data(), (size_t)N * 4, hipMemcpyHostToDevice, s), "copy");
        HIPCHK(launch_fill_f32(p->loss.as<float>(), 1, 0.f, s), "fill");
        int64_t nbatch = 0;
        for (int64_t b0 = 0; b0 + batch <= N; b0 += batch, nbatch++) {
            const int32_t* idx = p->perm.as<int32_t>() + b0;
            hipLaunchKernelGGL(gather_rows_f16_kernel, dim3(batch), dim3(128), 0, s, dX, p->D, idx,
                               batch, p->X16.as<_Float16>());
            hipLaunchKernelGGL(gather_i32_kernel, dim3(blocks_for(batch)), dim3(256), 0, s, dyl, idx,
                               batch, p->yl.as<int32_t>());
            hipLaunchKernelGGL(gather_i32_kernel, dim3(blocks_for(batch)), dim3(256), 0, s, dyb, idx,
                               batch, p->yb.as<int32_t>());
            rc = train_batch(p, batch, bucket_weight, lr, true);
            if (rc) return rc;
        }
        float l = 0.f;
        HIPCHK(hipMemcpyAsync(&l, p->loss.p, 4, hipMemcpyDeviceToHost, s), "copy");
        HIPCHK(hipStreamSynchronize(s), "sync");
        if (losses) losses[ep] = nbatch ? l / (float)nbatch : 0.f;
    }
    return SSG_OK;
}

int ssg_mlp_grad(ssg_mlp* p, int32_t layer, float* gW, float* gb)
{
    if (!p || layer < 0 || layer > 2) return ssg_fail(SSG_E_ARG, "bad layer");
    const size_t nw = (size_t)dims_out(p, layer) * dims_in(p, layer);
    if (gW) HIPCHK(hipMemcpy(gW, p->gW[layer].p, nw * 4, hipMemcpyDeviceToHost), "copy");
    if (gb) HIPCHK(hipMemcpy(gb, p->gb[layer].p, (size_t)dims_out(p, layer) * 4, hipMemcpyDeviceToHost), "copy");
    return SSG_OK;
}

}  // extern "C"

namespace {

// The fused output layer (head_top_kernel) serves hidden widths 64, 128, 192 and 256, up
// to 4 hints and 16 buckets; other shapes store the logits and reduce them afterwards.
bool fused_head_ok(const ssg_mlp* p, int h)
{
    return h <= 4 && p->NB <= 16 && (p->H2 == 64 || p->H2 == 128 || p->H2 == 192 || p->H2 == 256);
}

template <int K>
void launch_head_top(ssg_mlp* p, int B, const uint32_t* dland, int h, float tau, uint32_t* dhints,
                     int32_t* dbucket, float* dconf)
{
    const dim3 grid((unsigned)((B + 64 * HT_NQ - 1) / (64 * HT_NQ)));
    const _Float16* A2 = p->A2i.as<_Float16>();
    const _Float16* W = p->Wh[2].as<_Float16>();
    const float* bias = p->b[2].as<float>();
#define SSG_HT(HH)                                                                                  \
    hipLaunchKernelGGL((head_top_kernel<K, HH>), grid, dim3(256), 0, p->stream, A2, B, W, bias, p->L, \
                       p->NB, tau, dland, dhints, dbucket, dconf)
    switch (h) {
    case 1: SSG_HT(1); break;
    case 2: SSG_HT(2); break;
    case 3: SSG_HT(3); break;
    default: SSG_HT(4); break;
    }
#undef SSG_HT
}

// Inference with the fused output layer: the hidden layers by gemm_f16_kernel (FP16
// activations only), then head_top_kernel, in passes of up to 2^18 rows.
int infer_fused(ssg_mlp* p, const float* dQ, int64_t m, const uint32_t* dland, int h, float tau,
                uint32_t* dhints, int32_t* dbucket, float* dconf)
{
    const int chunk = 1 << 18;
    hipStream_t s = p->stream;
    for (int64_t r0 = 0; r0 < m; r0 += chunk) {
        const int B = (int)std::min<int64_t>(chunk, m - r0);
        if (B > p->cap_i) {
            HIPCHK(p->X16i.ensure((size_t)B * p->D * 2), "alloc");
            HIPCHK(p->A1i.ensure((size_t)B * p->H1 * 2), "alloc");
            HIPCHK(p->A2i.ensure((size_t)B * p->H2 * 2), "alloc");
            p->cap_i = B;
        }
        hipLaunchKernelGGL(gather_rows_f16_kernel, dim3(B), dim3(128), 0, s, dQ + r0 * p->D, p->D,
                           (const int32_t*)nullptr, B, p->X16i.as<_Float16>());
        HIPCHK(gemm(s, B, p->H1, p->D, p->X16i.as<_Float16>(), p->D, 1, p->Wh[0].as<_Float16>(), 1, p->D,
                    nullptr, p->H1, 0, p->b[0].as<float>(), 1, p->A1i.as<_Float16>()), "gemm");
        HIPCHK(gemm(s, B, p->H2, p->H1, p->A1i.as<_Float16>(), p->H1, 1, p->Wh[1].as<_Float16>(), 1, p->H1,
                    nullptr, p->H2, 0, p->b[1].as<float>(), 1, p->A2i.as<_Float16>()), "gemm");
        switch (p->H2) {
        case 64: launch_head_top<64>(p, B, dland, h, tau, dhints + r0 * h, dbucket + r0, dconf + r0); break;
        case 128: launch_head_top<128>(p, B, dland, h, tau, dhints + r0 * h, dbucket + r0, dconf + r0); break;
        case 192: launch_head_top<192>(p, B, dland, h, tau, dhints + r0 * h, dbucket + r0, dconf + r0); break;
        default: launch_head_top<256>(p, B, dland, h, tau, dhints + r0 * h, dbucket + r0, dconf + r0); break;
        }
        HIPCHK(hipGetLastError(), "inference head");
    }
    return SSG_OK;
}

}  // namespace

extern "C" {

// Inference on device query rows dQ [m x D] (FP32): logits, then per query the h
// best landmarks (as 1-based object ids through `landmarks`), bucket and confidence.
int ssg_mlp_infer_dev(ssg_mlp* p, const float* dQ, int64_t m, const uint32_t* dlandmarks,
                      int32_t h, float bucket_quantile, uint32_t* dhints, int32_t* dbucket, float* dconf,
                      void* stream)
{
    if (!p || !dQ || !dlandmarks || !dhints || !dbucket || !dconf || m < 1 || h < 1 || h > p->L)
        return ssg_fail(SSG_E_ARG, "bad inference arguments");
    if (!(bucket_quantile >= 0.f && bucket_quantile <= 1.f))
        return ssg_fail(SSG_E_ARG, "bucket quantile must be in [0, 1]");
    HIPCHK(hipSetDevice(p->device), "hipSetDevice");
    hipStream_t saved = p->stream;
    if (stream) p->stream = (hipStream_t)stream;
    if (fused_head_ok(p, h)) {
        const int rc = infer_fused(p, dQ, m, dlandmarks, h, bucket_quantile, dhints, dbucket, dconf);
        p->stream = saved;
        return rc;
    }
    const int chunk = 32768;  // rows per inference pass (the logits of a pass are stored)
    int rc = SSG_OK;
    for (int64_t r0 = 0; r0 < m && rc == SSG_OK; r0 += chunk) {
        const int B = (int)std::min<int64_t>(chunk, m - r0);
        rc = ensure_batch(p, B);
        if (rc) break;
        hipLaunchKernelGGL(gather_rows_f16_kernel, dim3(B), dim3(128), 0, p->stream, dQ + r0 * p->D,
                           p->D, (const int32_t*)nullptr, B, p->X16.as<_Float16>());
        rc = forward(p, B);
        if (rc) break;
        hipLaunchKernelGGL(topk_heads_kernel, dim3(B), dim3(64), 0, p->stream, p->Z.as<float>(), B,
                           p->L, p->NB, h, bucket_quantile, dlandmarks, dhints + r0 * h, dbucket + r0,
                           dconf + r0);
        if (hipGetLastError() != hipSuccess) rc = ssg_fail(SSG_E_HIP, "inference launch");
    }
    p->stream = saved;
    return rc;
}

}  // extern "C"
