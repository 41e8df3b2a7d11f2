// ssg_consumers.hip -- batch consumers of search results that run on the device.
//
// hsp_queries (src/hsp.jl:77-101 with hsp_should_push :8-33): each query's candidate
// column (ascending by distance, 0 ends it) is scanned in order; a candidate is kept
// unless an already kept candidate is strictly nearer to it than the query is
// (d(candidate, kept) < d(query, candidate)).  Kept candidates enter a sorted queue of
// capacity k (KnnSorted, stable on equal distances), written back as a (k, m) column,
// then 0 / +Inf.
//
// One wave per query.  The candidate rows are staged once in LDS (k rows of dp floats
// per wave), so the kept-vs-candidate distances of the scan read LDS, not HBM: the
// HBM traffic per query is its candidate rows plus the id/distance columns.  Distances
// are the canonical ones (run_partial + the 16-lane tree in FP64, rounded once), the
// same evaluation as rerank! and the beam search; the tree runs on DPP row shifts and
// its total (lane 15 of each 16-lane group) is the butterfly's, bit for bit.
//
// The kept rows are compacted in place: when candidate c is kept as the A-th, its row
// moves to slot A (slots A..c-1 held candidates already decided), so kept row a sits at
// slot a.  Each candidate is tested against every kept row, eight per iteration (two
// independent 4-row steps), with one ballot at the end: the reference stops at the first
// nearer kept row, which decides the same way.
#include <hip/hip_runtime.h>

#include "ssg_device.h"
#include "ssg_kernels.h"

namespace ssg {

// run_partial with the second operand already widened to FP64 (the same operations:
// the widening is exact, so the sum is run_partial's bit for bit).
template <int V>
__device__ __forceinline__ double run_partial_wide(int metric, const double (&p)[V],
                                                   const float (&x)[V])
{
    double s = 0.0;
    if (metric == M_NORMCOS) {
#pragma unroll
        for (int t = 0; t < V; t++) s = __builtin_fma(p[t], (double)x[t], s);
    } else {
#pragma unroll
        for (int t = 0; t < V; t++) {
            const double m = p[t] - (double)x[t];
            s = __builtin_fma(m, m, s);
        }
    }
    return s;
}

template <int V>
__global__ __launch_bounds__(256) void hsp_kernel(const float* __restrict__ X, int metric, int64_t m,
                                                  int k, const uint32_t* __restrict__ ids,
                                                  const float* __restrict__ dists,
                                                  uint32_t* __restrict__ out_ids,
                                                  float* __restrict__ out_dists)
{
    extern __shared__ float hsp_rows[];
    constexpr int dp = 16 * V;
    const int lane = lane_id(), w = threadIdx.x >> 6;
    const int g = lane >> 4, j = lane & 15;
    const int64_t q = (int64_t)blockIdx.x * (blockDim.x >> 6) + w;
    if (q >= m) return;
    float* R = hsp_rows + (size_t)w * k * dp;

    const uint32_t cid = lane < k ? ids[q * k + lane] : 0u;
    const float cd = lane < k ? dists[q * k + lane] : __builtin_inff();
    const uint64_t z = __ballot(lane < k && cid == 0u);
    const int L = z ? (__ffsll((unsigned long long)z) - 1) : k;

    // stage the L candidate rows, 16 lanes per row: up to 32 rows' loads in flight before
    // their LDS stores
    constexpr int SB = V <= 8 ? 8 : 4;  // 4-row steps per batch
    for (int r0 = 0; r0 < L; r0 += 4 * SB) {
        float xv[SB][V];
#pragma unroll
        for (int s = 0; s < SB; s++) {
            const int r = r0 + 4 * s + g;
            const uint32_t id = __shfl(cid, r < L ? r : 0);
            if (r < L) {
                load_run<V>(X + (size_t)(id - 1u) * dp + j * V, xv[s]);
            } else {
#pragma unroll
                for (int t = 0; t < V; t++) xv[s][t] = 0.f;
            }
        }
#pragma unroll
        for (int s = 0; s < SB; s++) {
            const int r = r0 + 4 * s + g;
            if (r < L) {
#pragma unroll
                for (int t = 0; t < V; t++) R[r * dp + j * V + t] = xv[s][t];
            }
        }
    }
    wave_sync();

    RegQueue<1, false> out;
    out.init(k);
    int A = 0;
    for (int c = 0; c < L; c++) {
        float pv[V];
        double pw[V];
#pragma unroll
        for (int t = 0; t < V; t++) {
            pv[t] = R[c * dp + j * V + t];
            pw[t] = (double)pv[t];
        }
        const float pd = rl_f(cd, c);
        bool hit = false;
        for (int a0 = 0; a0 < A; a0 += 8) {
            const int u0 = a0 + g, u1 = a0 + 4 + g;
            const int r0 = u0 < A ? u0 : 0, r1 = u1 < A ? u1 : 0;  // in-range slots only
            float x0[V], x1[V];
#pragma unroll
            for (int t = 0; t < V; t++) {
                x0[t] = R[r0 * dp + j * V + t];
                x1[t] = R[r1 * dp + j * V + t];
            }
            const double s0 = group16_tree(run_partial_wide<V>(metric, pw, x0));
            const double s1 = group16_tree(run_partial_wide<V>(metric, pw, x1));
            if (j == 15) {
                hit |= u0 < A && finish_dist(metric, s0) < pd;
                hit |= u1 < A && finish_dist(metric, s1) < pd;
            }
        }
        if (!__ballot(hit)) {
            if (A != c) {
                if (g == 0) {
#pragma unroll
                    for (int t = 0; t < V; t++) R[A * dp + j * V + t] = pv[t];
                }
                wave_sync();
            }
            A++;
            out.push(pd, rl_u(cid, c), 0, 0);
        }
    }
    if (lane < k) {
        out_ids[q * k + lane] = lane < out.len ? out.id[0] : 0u;
        out_dists[q * k + lane] = lane < out.len ? out.d[0] : __builtin_inff();
    }
}

// Waves per block: as many as keep the block's staged rows within 64 KiB (1 to 4).
static int hsp_waves(int k, int dp)
{
    const int per_wave = k * dp * 4;
    int w = 65536 / (per_wave > 0 ? per_wave : 1);
    return w < 1 ? 1 : (w > 4 ? 4 : w);
}

hipError_t launch_hsp(const float* X, int32_t dp, int32_t metric, int64_t m, int32_t k,
                      const uint32_t* ids, const float* dists, uint32_t* out_ids, float* out_dists,
                      hipStream_t s)
{
    if (k < 1 || k > 64 || m <= 0) return hipErrorInvalidValue;
    const int W = hsp_waves(k, dp);
    const dim3 grid((unsigned)((m + W - 1) / W)), block(64 * W);
    const size_t lds = (size_t)W * k * dp * 4;
    switch (dp / 16) {
#define SSG_HSP(VV) \
    case VV: hipLaunchKernelGGL((hsp_kernel<VV>), grid, block, lds, s, X, metric, m, k, ids, dists, out_ids, out_dists); break;
        SSG_HSP(1) SSG_HSP(2) SSG_HSP(3) SSG_HSP(4) SSG_HSP(5) SSG_HSP(6) SSG_HSP(7) SSG_HSP(8) SSG_HSP(12) SSG_HSP(16)
#undef SSG_HSP
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace ssg
