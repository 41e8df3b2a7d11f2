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

template <int V, int NB>
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

    // candidate c = 64 b + lane in register b (k <= 64 NB)
    uint32_t cid[NB];
    float cd[NB];
    int L = k;
    bool found = false;
#pragma unroll
    for (int b = 0; b < NB; b++) {
        const int c = 64 * b + lane;
        cid[b] = c < k ? ids[q * k + c] : 0u;
        cd[b] = c < k ? dists[q * k + c] : __builtin_inff();
        const uint64_t z = __ballot(c < k && cid[b] == 0u);
        if (!found && z) {
            L = 64 * b + __ffsll((unsigned long long)z) - 1;
            found = true;
        }
    }

    // stage the L candidate rows, 16 lanes per row: up to 32 rows' loads in flight before
    // their LDS stores
    constexpr int SB = V <= 8 ? 8 : 4;  // 4-row steps per batch
    for (int r0 = 0; r0 < L; r0 += 4 * SB) {
        float xv[SB][V];
#pragma unroll
        for (int s = 0; s < SB; s++) {
            const int r = r0 + 4 * s + g;
            // the row's id: from the lanes' registers when the column fits one register
            uint32_t id;
            if constexpr (NB == 1) {
                id = __shfl(cid[0], r < L ? r : 0);
            } else {
                id = r < L ? ids[q * k + r] : 1u;
            }
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

    RegQueue<NB, false> out;
    out.init(k);
    int A = 0;
    for (int c = 0; c < L; c++) {
        float pv[V];
        double pw[V];
        float cdc = cd[0];
        uint32_t cic = cid[0];
#pragma unroll
        for (int b = 1; b < NB; b++)
            if ((c >> 6) == b) {  // wave-uniform selection of the register holding c
                cdc = cd[b];
                cic = cid[b];
            }
#pragma unroll
        for (int t = 0; t < V; t++) {
            pv[t] = R[c * dp + j * V + t];
            pw[t] = (double)pv[t];
        }
        const float pd = rl_f(cdc, c & 63);
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
            out.push(pd, rl_u(cic, c & 63), 0, 0);
        }
    }
#pragma unroll
    for (int b = 0; b < NB; b++) {
        const int c = 64 * b + lane;
        if (c < k) {
            out_ids[q * k + c] = c < out.len ? out.id[b] : 0u;
            out_dists[q * k + c] = c < out.len ? out.d[b] : __builtin_inff();
        }
    }
}

// neardup_block! (src/neardup.jl:186-256), in two launches.
//
// pair_dist_kernel: P[a * mb + b] = canonical d(X[ids[a]], X[ids[b]]) for b < a, one
// wave per row a, four b per step (16 lanes each, DPP tree, lane 15 writes).
template <int V>
__global__ __launch_bounds__(256) void pair_dist_kernel(const float* __restrict__ X, int metric,
                                                        const uint32_t* __restrict__ ids, int mb,
                                                        float* __restrict__ P)
{
    constexpr int dp = 16 * V;
    const int lane = lane_id(), w = threadIdx.x >> 6;
    const int g = lane >> 4, j = lane & 15;
    const int a = blockIdx.x * 4 + w;
    if (a >= mb) return;
    float qv[V];
    load_run<V>(X + (size_t)ids[a] * dp + j * V, qv);
    for (int b0 = 0; b0 < a; b0 += 4) {
        const int b = b0 + g;
        const int bb = b < a ? b : 0;  // in-range row only
        float xv[V];
        load_run<V>(X + (size_t)ids[bb] * dp + j * V, xv);
        const double s = group16_tree(run_partial<V>(metric, qv, xv));
        if (j == 15 && b < a) P[(size_t)a * mb + b] = finish_dist(metric, s);
    }
}

// greedy_net_kernel (one wave): items 0..mb-1 in order; item 0 is a center; item a is a
// center unless the least distance to the centers so far (the first center, i.e. the
// lowest position, on ties) is <= eps, in which case it is covered by that center.
// cover[a] is the block position of a's center (a itself for a center), cdist[a] the
// distance (0 for a center).  Lane l keeps the center flags of items l + 64 i (bit i),
// so mb <= 1024; row a + 1 of P is loaded while row a is reduced.  The minimum is taken
// over one 64-bit key per entry (order-preserving distance bits, then the position):
// the lexicographic minimum, whatever the reduction order.
__device__ __forceinline__ uint64_t dpp_min_u64_shr(uint64_t v, int ctrl_sel)
{
    const uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
    uint32_t olo, ohi;
    // lanes without a source keep their own value (bound_ctrl off, old = own)
    switch (ctrl_sel) {
    case 1:
        olo = (uint32_t)__builtin_amdgcn_update_dpp((int)lo, (int)lo, DPP_ROW_SHR1, 0xF, 0xF, false);
        ohi = (uint32_t)__builtin_amdgcn_update_dpp((int)hi, (int)hi, DPP_ROW_SHR1, 0xF, 0xF, false);
        break;
    case 2:
        olo = (uint32_t)__builtin_amdgcn_update_dpp((int)lo, (int)lo, DPP_ROW_SHR2, 0xF, 0xF, false);
        ohi = (uint32_t)__builtin_amdgcn_update_dpp((int)hi, (int)hi, DPP_ROW_SHR2, 0xF, 0xF, false);
        break;
    case 4:
        olo = (uint32_t)__builtin_amdgcn_update_dpp((int)lo, (int)lo, DPP_ROW_SHR4, 0xF, 0xF, false);
        ohi = (uint32_t)__builtin_amdgcn_update_dpp((int)hi, (int)hi, DPP_ROW_SHR4, 0xF, 0xF, false);
        break;
    default:
        olo = (uint32_t)__builtin_amdgcn_update_dpp((int)lo, (int)lo, DPP_ROW_SHR8, 0xF, 0xF, false);
        ohi = (uint32_t)__builtin_amdgcn_update_dpp((int)hi, (int)hi, DPP_ROW_SHR8, 0xF, 0xF, false);
        break;
    }
    const uint64_t o = ((uint64_t)ohi << 32) | olo;
    return o < v ? o : v;
}

__device__ __forceinline__ uint64_t wave_min_u64(uint64_t v)
{
    v = dpp_min_u64_shr(v, 1);
    v = dpp_min_u64_shr(v, 2);
    v = dpp_min_u64_shr(v, 4);
    v = dpp_min_u64_shr(v, 8);  // lane 15 of each row: the row's minimum
    uint64_t r = ~0ull;
#pragma unroll
    for (int row = 0; row < 4; row++) {
        const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, 16 * row + 15);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), 16 * row + 15);
        const uint64_t x = ((uint64_t)hi << 32) | lo;
        r = x < r ? x : r;
    }
    return r;
}

// The scan, with row a of P's lower triangle read through rowp(a) (b < a valid).
template <class RowP>
__device__ __forceinline__ void greedy_scan(RowP rowp, int mb, float eps, int32_t* __restrict__ cover,
                                            float* __restrict__ cdist)
{
    constexpr int CH = 16;  // mb <= 64 * CH
    const int lane = threadIdx.x;
    uint32_t cflags = lane == 0 ? 1u : 0u;  // item 0 is a center
    if (lane == 0) {
        cover[0] = 0;
        cdist[0] = 0.f;
    }
    float cur[CH];
#pragma unroll
    for (int i = 0; i < CH; i++) cur[i] = 0.f;
    if (mb > 1 && lane == 0) cur[0] = rowp(1)[0];
    for (int a = 1; a < mb; a++) {
        float nxt[CH];
#pragma unroll
        for (int i = 0; i < CH; i++) {
            const int b = lane + 64 * i;
            nxt[i] = (a + 1 < mb && b < a + 1) ? rowp(a + 1)[b] : 0.f;
        }
        uint64_t key = ~0ull;
#pragma unroll
        for (int i = 0; i < CH; i++) {
            const int b = lane + 64 * i;
            const float d = cur[i];
            if (b < a && ((cflags >> i) & 1u) && d == d) {
                const uint32_t u = d == 0.f ? 0u : __float_as_uint(d);  // -0 ties with +0
                const uint32_t kd = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
                const uint64_t k = ((uint64_t)kd << 32) | (uint32_t)b;
                key = k < key ? k : key;
            }
        }
        key = wave_min_u64(key);
        bool center = key == ~0ull;
        float bd = 0.f;
        int bb = 0;
        if (!center) {
            const uint32_t kd = (uint32_t)(key >> 32);
            bd = __uint_as_float((kd & 0x80000000u) ? (kd & 0x7fffffffu) : ~kd);
            bb = (int)(uint32_t)key;
            center = bd > eps;
        }
        if (center) {
            if (lane == (a & 63)) cflags |= 1u << (a >> 6);
            if (lane == 0) {
                cover[a] = a;
                cdist[a] = 0.f;
            }
        } else if (lane == 0) {
            cover[a] = bb;
            cdist[a] = bd;
        }
#pragma unroll
        for (int i = 0; i < CH; i++) cur[i] = nxt[i];
    }
}

__global__ __launch_bounds__(64) void greedy_net_kernel(const float* __restrict__ P, int mb, float eps,
                                                        int32_t* __restrict__ cover,
                                                        float* __restrict__ cdist)
{
    greedy_scan([&](int a) { return P + (size_t)a * mb; }, mb, eps, cover, cdist);
}

// mb <= 256: the whole workgroup first copies P's lower triangle (row a at a(a-1)/2) into
// LDS, then wave 0 scans it there.
constexpr int GREEDY_LDS_MB = 256;
__global__ __launch_bounds__(1024) void greedy_net_lds_kernel(const float* __restrict__ P, int mb,
                                                              float eps, int32_t* __restrict__ cover,
                                                              float* __restrict__ cdist)
{
    __shared__ float tri[GREEDY_LDS_MB * (GREEDY_LDS_MB - 1) / 2];
    const int ntri = mb * (mb - 1) / 2;
    for (int e = threadIdx.x; e < ntri; e += blockDim.x) {
        // row a = the largest a with a(a-1)/2 <= e
        int a = (int)((1.0f + __builtin_sqrtf(1.0f + 8.0f * (float)e)) * 0.5f);
        while (a * (a - 1) / 2 > e) a--;
        while ((a + 1) * a / 2 <= e) a++;
        tri[e] = P[(size_t)a * mb + (e - a * (a - 1) / 2)];
    }
    __syncthreads();
    if (threadIdx.x >= 64) return;
    greedy_scan([&](int a) { return (const float*)(tri + a * (a - 1) / 2); }, mb, eps, cover, cdist);
}

// nearest_center_kernel: for item a, the nearest of the centers F[0..nf) (canonical
// distance; the lowest index on ties) -> pos[a] = 1-based index into F, or 0 when nf = 0,
// and its distance.  One wave per item, four centers per step (16 lanes each).
template <int V>
__global__ __launch_bounds__(256) void nearest_center_kernel(const float* __restrict__ X, int metric,
                                                             const uint32_t* __restrict__ items, int mb,
                                                             const uint32_t* __restrict__ F, int nf,
                                                             int32_t* __restrict__ pos,
                                                             float* __restrict__ pdist)
{
    constexpr int dp = 16 * V;
    const int lane = lane_id(), w = threadIdx.x >> 6;
    const int g = lane >> 4, j = lane & 15;
    const int a = blockIdx.x * 4 + w;
    if (a >= mb) return;
    float qv[V];
    load_run<V>(X + (size_t)items[a] * dp + j * V, qv);
    float bd = __builtin_inff();
    int bi = 0x7fffffff;
    for (int f0 = 0; f0 < nf; f0 += 4) {
        const int f = f0 + g;
        const int ff = f < nf ? f : 0;  // in-range center only
        float xv[V];
        load_run<V>(X + (size_t)F[ff] * dp + j * V, xv);
        const float d = finish_dist(metric, group16_tree(run_partial<V>(metric, qv, xv)));
        if (f < nf && d < bd) {  // each group sees its centers in increasing index
            bd = d;
            bi = f;
        }
    }
    // lane 15 of each group holds its group's best; combine by (distance, index)
#pragma unroll
    for (int o = 16; o <= 32; o <<= 1) {
        const float od = __shfl_xor(bd, o);
        const int oi = __shfl_xor(bi, o);
        if (od < bd || (od == bd && oi < bi)) {
            bd = od;
            bi = oi;
        }
    }
    if (lane == 15) {
        pos[a] = bi == 0x7fffffff ? 0 : bi + 1;
        pdist[a] = bd;
    }
}

hipError_t launch_nearest_center(const float* X, int32_t dp, int32_t metric, const uint32_t* items,
                                 int32_t mb, const uint32_t* F, int32_t nf, int32_t* pos, float* pdist,
                                 hipStream_t s)
{
    if (mb <= 0) return hipSuccess;
    const dim3 grid((unsigned)((mb + 3) / 4)), block(256);
    switch (dp / 16) {
#define SSG_NC(VV) \
    case VV: hipLaunchKernelGGL((nearest_center_kernel<VV>), grid, block, 0, s, X, metric, items, mb, F, nf, pos, pdist); break;
        SSG_NC(1) SSG_NC(2) SSG_NC(3) SSG_NC(4) SSG_NC(5) SSG_NC(6) SSG_NC(7) SSG_NC(8) SSG_NC(12) SSG_NC(16) SSG_NC(24) SSG_NC(32) SSG_NC(48) SSG_NC(64)
#undef SSG_NC
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Rows ids[i] of X (padded, with norms) appended at row n0 + i of Y.
__global__ void gather_padded_rows_kernel(const float* __restrict__ X, const float* __restrict__ xn,
                                          int dp, const uint32_t* __restrict__ ids, int64_t m,
                                          float* __restrict__ Y, float* __restrict__ yn, int64_t n0)
{
    const int64_t r = blockIdx.x;
    if (r >= m) return;
    const size_t src = (size_t)ids[r];
    for (int c = threadIdx.x; c < dp; c += blockDim.x) Y[(size_t)(n0 + r) * dp + c] = X[src * dp + c];
    if (threadIdx.x == 0) yn[n0 + r] = xn[src];
}

hipError_t launch_pair_dist(const float* X, int32_t dp, int32_t metric, const uint32_t* ids, int32_t mb,
                            float* P, hipStream_t s)
{
    if (mb <= 0) return hipSuccess;
    const dim3 grid((unsigned)((mb + 3) / 4)), block(256);
    switch (dp / 16) {
#define SSG_PD(VV) \
    case VV: hipLaunchKernelGGL((pair_dist_kernel<VV>), grid, block, 0, s, X, metric, ids, mb, P); break;
        SSG_PD(1) SSG_PD(2) SSG_PD(3) SSG_PD(4) SSG_PD(5) SSG_PD(6) SSG_PD(7) SSG_PD(8) SSG_PD(12) SSG_PD(16) SSG_PD(24) SSG_PD(32) SSG_PD(48) SSG_PD(64)
#undef SSG_PD
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_greedy_net(const float* P, int32_t mb, float eps, int32_t* cover, float* cdist,
                             hipStream_t s)
{
    if (mb <= 0) return hipSuccess;
    if (mb > 1024) return hipErrorInvalidValue;
    if (mb <= GREEDY_LDS_MB)
        hipLaunchKernelGGL(greedy_net_lds_kernel, dim3(1), dim3(1024), 0, s, P, mb, eps, cover, cdist);
    else
        hipLaunchKernelGGL(greedy_net_kernel, dim3(1), dim3(64), 0, s, P, mb, eps, cover, cdist);
    return hipGetLastError();
}

hipError_t launch_gather_padded_rows(const float* X, const float* xn, int32_t dp, const uint32_t* ids,
                                     int64_t m, float* Y, float* yn, int64_t n0, hipStream_t s)
{
    if (m <= 0) return hipSuccess;
    hipLaunchKernelGGL(gather_padded_rows_kernel, dim3((unsigned)m), dim3(64), 0, s, X, xn, dp, ids, m,
                       Y, yn, n0);
    return hipGetLastError();
}

// Waves per block: as many as keep the block's staged rows within 64 KiB (1 to 4); one
// wave per block beyond that, within the 160 KiB of LDS a gfx950 workgroup may use.
static int hsp_waves(int k, int dp)
{
    const int per_wave = k * dp * 4;
    int w = 65536 / (per_wave > 0 ? per_wave : 1);
    return w < 1 ? 1 : (w > 4 ? 4 : w);
}

template <int NB>
static hipError_t launch_hsp_nb(const float* X, int32_t dp, int32_t metric, int64_t m, int32_t k,
                                const uint32_t* ids, const float* dists, uint32_t* out_ids,
                                float* out_dists, hipStream_t s)
{
    const int W = hsp_waves(k, dp);
    const dim3 grid((unsigned)((m + W - 1) / W)), block(64 * W);
    const size_t lds = (size_t)W * k * dp * 4;
    switch (dp / 16) {
#define SSG_HSP(VV)                                                                                  \
    case VV:                                                                                         \
        if (lds > 65536) {                                                                           \
            hipError_t e = hipFuncSetAttribute((const void*)hsp_kernel<VV, NB>,                      \
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds); \
            if (e != hipSuccess) return e;                                                           \
        }                                                                                            \
        hipLaunchKernelGGL((hsp_kernel<VV, NB>), grid, block, lds, s, X, metric, m, k, ids, dists,   \
                           out_ids, out_dists);                                                      \
        break;
        SSG_HSP(1) SSG_HSP(2) SSG_HSP(3) SSG_HSP(4) SSG_HSP(5) SSG_HSP(6) SSG_HSP(7) SSG_HSP(8) SSG_HSP(12) SSG_HSP(16) SSG_HSP(24) SSG_HSP(32) SSG_HSP(48) SSG_HSP(64)
#undef SSG_HSP
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_hsp(const float* X, int32_t dp, int32_t metric, int64_t m, int32_t k,
                      const uint32_t* ids, const float* dists, uint32_t* out_ids, float* out_dists,
                      hipStream_t s)
{
    if (k < 1 || k > 256 || m <= 0 || (size_t)k * dp * 4 > 160 * 1024) return hipErrorInvalidValue;
    if (k <= 64) return launch_hsp_nb<1>(X, dp, metric, m, k, ids, dists, out_ids, out_dists, s);
    if (k <= 128) return launch_hsp_nb<2>(X, dp, metric, m, k, ids, dists, out_ids, out_dists, s);
    return launch_hsp_nb<4>(X, dp, metric, m, k, ids, dists, out_ids, out_dists, s);
}

}  // namespace ssg
