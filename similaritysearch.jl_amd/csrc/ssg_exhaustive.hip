// ssg_exhaustive.hip -- exhaustive k-NN (ExhaustiveSearch / ParallelExhaustiveSearch)
// and rerank! for gfx950.
//
// Result contract: the k objects that come first in (d, id) order, where d is the
// canonical distance (ssg_device.h) -- exactly what the ascending scan of
// src/exact/sequential-exhaustive.jl:68-80 leaves in a KnnSorted
// (src/pqueue/knnsorted.jl:140-168), and what the batch merge of
// src/exact/parallel-exhaustive.jl:79-98 reproduces.
//
// Stage 1 (MFMA): block = 64 queries x one DB slice, two blocks per CU.  Each wave owns
//   16 queries (the B operand of v_mfma_f32_16x16x4_f32, kept in registers for the
//   whole slice); 64-row DB tiles are staged in LDS (the next one prefetched into
//   registers while the current one is computed) and shared by the 4 waves; row
//   blocks are software-pipelined so the filter below overlaps the next block's MFMAs.  The dot
//   products give approximate distances ||x||^2 + ||q||^2 - 2 x.q (or 1 - x.q)
//   that are filtered against a per-query threshold tau into an LDS candidate
//   buffer; a full buffer is compacted to its KC best (d, id) entries.  The slice
//   emits its KC candidates and tau, a lower bound on the approximate distance of
//   every row of the slice that it dropped.  The K dimension is permuted so each
//   lane's A/B run is contiguous (dot products are order-free sums).
// Stage 2: one wave per query merges the slices' lists, radix-selects the 64 best
//   approximate candidates, evaluates their canonical distances, sorts by (d, id)
//   and writes the top k.  A certificate checks that every dropped object is
//   provably farther than the k-th exact distance, given an error bound on the
//   approximate distances; a query that fails it is flagged.
// Stage 3: flagged queries are rescanned exactly (canonical distance over every
//   row, ascending id, KnnSorted pushes).
#include "ssg_device.h"
#include "ssg_kernels.h"
#include "ssg_select.h"

namespace ssg {

typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int TR = 64;          // DB rows per LDS tile
constexpr int CAP = 64;         // LDS candidate slots per query
constexpr int SEL = 64;         // candidates rescored exactly in stage 2

__device__ __forceinline__ uint32_t sort_key(float f)
{
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float key_float(uint32_t k)
{
    const uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    return __uint_as_float(u);
}

// Keep the KC best (d, id) entries of query slot q; returns the new tau.
template <int KC>
__device__ __forceinline__ float compact(float* cd, uint32_t* ci, int* cnt, int lane)
{
    const int c = min((int)rfl_u((uint32_t)*cnt), CAP);
    const float dv = lane < c ? cd[lane] : __builtin_inff();
    const uint32_t iv = lane < c ? ci[lane] : 0xffffffffu;
    int rank = 0;
    for (int t = 0; t < c; t++) {
        const float dt = rl_f(dv, t);
        const uint32_t it = rl_u(iv, t);
        rank += (dt < dv || (dt == dv && it < iv)) ? 1 : 0;
    }
    __builtin_amdgcn_wave_barrier();
    asm volatile("" ::: "memory");
    if (lane < c && rank < KC) {
        cd[rank] = dv;
        ci[rank] = iv;
    }
    float tau = __builtin_inff();
    if (c >= KC) {
        const uint64_t b = __ballot(lane < c && rank == KC - 1);
        tau = rl_f(dv, __ffsll((unsigned long long)b) - 1);
    }
    if (lane == 0) *cnt = c < KC ? c : KC;
    __builtin_amdgcn_wave_barrier();
    asm volatile("" ::: "memory");
    return tau;
}

template <int DP, int KC>
__global__ __launch_bounds__(256) void exh_stage1_lds(const float* __restrict__ X,
                                                  const float* __restrict__ xnorm, int64_t n,
                                                  int metric, const float* __restrict__ qpad,
                                                  const float* __restrict__ qnorm, int64_t rows_per_slice,
                                                  int nslices, float* cand_d, uint32_t* cand_i,
                                                  float* slice_tau, unsigned int* fail_count)
{
    constexpr int LS = DP + 4;  // LDS row stride (floats)
    constexpr int KS = DP / 4;  // MFMA k-steps
    constexpr int QT = TR;  // 64 queries per block, staged through the row tile
    // the merge that follows on the stream counts its failed certificates from zero
    if (fail_count && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) *fail_count = 0u;
    __shared__ float s_x[TR * LS];
    __shared__ float s_xn[TR];
    __shared__ float s_cd[QT][CAP];
    __shared__ uint32_t s_ci[QT][CAP];
    __shared__ int s_cnt[QT];
    __shared__ float s_qn[QT];

    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int col = lane & 15, gk = lane >> 4;
    const int64_t qt = blockIdx.x;
    const int sl = blockIdx.y;
    const int64_t r0 = (int64_t)sl * rows_per_slice;
    const int64_t r1 = min(n, r0 + rows_per_slice);
    const int qloc = 16 * w + col;

    // The block's 64 query rows are staged through the tile buffer, so the B operands
    // come from LDS reads and no register of the MFMA loop depends on a global load
    // (the tile prefetch below stays in flight across the whole loop).  Query norms go
    // through LDS for the same reason.
    for (int idx = tid; idx < QT * (DP / 4); idx += 256) {
        const int row = idx / (DP / 4), c4 = idx % (DP / 4);
        *reinterpret_cast<float4*>(&s_x[row * LS + 4 * c4]) =
            reinterpret_cast<const float4*>(qpad + (qt * QT + row) * DP)[c4];
    }
    if (tid < QT) {
        s_cnt[tid] = 0;
        s_qn[tid] = qnorm[qt * QT + tid];
    }
    __syncthreads();
    const float qn = s_qn[qloc];
    float bq[KS];
    {
        const float* qp = &s_x[qloc * LS + gk * KS];
#pragma unroll
        for (int s = 0; s < KS; s += 4) {
            const float4 v = *reinterpret_cast<const float4*>(qp + s);
            bq[s] = v.x;
            bq[s + 1] = v.y;
            bq[s + 2] = v.z;
            bq[s + 3] = v.w;
        }
    }
    float tau = __builtin_inff();

    // the next tile is loaded into registers while the current one is computed
    constexpr int PF = (TR * (DP / 4) + 255) / 256;  // float4 per thread per tile
    float4 pf[PF];
    float4 pfn;  // the tile's 64 norms, 4 per thread of the first 16 (tiles start at multiples
                 // of 64 rows and xnorm holds n_pad entries, so the load is always in bounds)
    auto prefetch = [&](int64_t t) {
#pragma unroll
        for (int i = 0; i < PF; i++) {
            const int idx = tid + 256 * i;
            const int row = idx / (DP / 4), c4 = idx % (DP / 4);
            const int64_t gr = t + row;
            pf[i] = (idx < TR * (DP / 4) && gr < r1) ? reinterpret_cast<const float4*>(X + gr * DP)[c4]
                                                     : make_float4(0.f, 0.f, 0.f, 0.f);
        }
        pfn = reinterpret_cast<const float4*>(xnorm + t)[tid & (TR / 4 - 1)];
    };
    if (r0 < r1) prefetch(r0);
    const float* a_base = &s_x[col * LS + gk * KS];  // this lane's A run in row block 0
    for (int64_t t0 = r0; t0 < r1; t0 += TR) {
        __syncthreads();
#pragma unroll
        for (int i = 0; i < PF; i++) {
            const int idx = tid + 256 * i;
            const int row = idx / (DP / 4), c4 = idx % (DP / 4);
            if (idx < TR * (DP / 4)) *reinterpret_cast<float4*>(&s_x[row * LS + 4 * c4]) = pf[i];
        }
        if (tid < TR / 4) {
            const float v[4] = {pfn.x, pfn.y, pfn.z, pfn.w};
#pragma unroll
            for (int c = 0; c < 4; c++) s_xn[4 * tid + c] = (t0 + 4 * tid + c < r1) ? v[c] : 0.f;
        }
        __syncthreads();
        if (t0 + TR < r1) prefetch(t0 + TR);
        // Row blocks are software-pipelined: the MFMA chain of block rb + 1 is issued
        // before block rb's distances are filtered, so the filter's VALU and LDS work
        // runs while the matrix pipe is busy.
        auto chain = [&](int rb) {
            floatx4 acc = {0.f, 0.f, 0.f, 0.f};
            const float* ap = a_base + rb * (16 * LS);
#pragma unroll
            for (int s = 0; s < KS; s++)
                acc = __builtin_amdgcn_mfma_f32_16x16x4f32(ap[s], bq[s], acc, 0, 0, 0);
            return acc;
        };
        floatx4 cur = chain(0);
#pragma unroll
        for (int rb = 0; rb < TR / 16; rb++) {
            floatx4 nxt = {0.f, 0.f, 0.f, 0.f};
            if (rb + 1 < TR / 16) nxt = chain(rb + 1);
            // norms of this lane's 4 output rows of block rb
            const float4 xn4 = *reinterpret_cast<const float4*>(&s_xn[rb * 16 + gk * 4]);
            const float xn[4] = {xn4.x, xn4.y, xn4.z, xn4.w};
            float da[4];
            bool pass[4];
            bool any = false;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int64_t gid = t0 + rb * 16 + gk * 4 + r;
                const float dot = cur[r];
                da[r] = metric == M_NORMCOS ? 1.0f - dot : (xn[r] + qn) - 2.0f * dot;
                pass[r] = gid < r1 && da[r] < tau;
                any = any || pass[r];
            }
            if (__ballot(any)) {
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    if (pass[r]) {
                        const int p = atomicAdd(&s_cnt[qloc], 1);
                        if (p < CAP) {
                            s_cd[qloc][p] = da[r];
                            s_ci[qloc][p] = (uint32_t)(t0 + rb * 16 + gk * 4 + r);
                        }
                    }
                }
                // counts only change where a candidate was added; each query gains at
                // most 16 per row block, so a count at or below CAP - 16 needs no
                // compaction yet
                __builtin_amdgcn_wave_barrier();
                asm volatile("" ::: "memory");
                const int cq = lane < 16 ? s_cnt[16 * w + lane] : 0;
                uint64_t need = __ballot(lane < 16 && cq > CAP - 16);
                while (need) {
                    const int qq = __ffsll((unsigned long long)need) - 1;
                    need &= need - 1;
                    const float nt = compact<KC>(s_cd[16 * w + qq], s_ci[16 * w + qq], &s_cnt[16 * w + qq], lane);
                    if (col == qq) tau = nt;
                }
            }
            cur = nxt;
        }
    }
    __syncthreads();
    // final: every query of this wave down to KC entries, then out
    for (int qq = 0; qq < 16; qq++) {
        const int ql = 16 * w + qq;
        const int c = (int)rfl_u((uint32_t)s_cnt[ql]);
        float nt = rl_f(tau, qq);
        if (c > KC) nt = compact<KC>(s_cd[ql], s_ci[ql], &s_cnt[ql], lane);
        const int cc = min(c, KC);
        const int64_t q = qt * QT + ql;
        const int64_t base = (q * nslices + sl) * KC;
        if (lane < KC) {
            cand_d[base + lane] = lane < cc ? s_cd[ql][lane] : __builtin_inff();
            cand_i[base + lane] = lane < cc ? s_ci[ql][lane] : 0xffffffffu;
        }
        if (lane == 0) slice_tau[q * nslices + sl] = nt;
    }
}

// Stage 1 for dp <= 128: block = 128 queries x one DB slice, 4 waves of 32 queries (two
// independent accumulator chains of v_mfma_f32_16x16x4f32 per wave, the B operands in
// registers for the whole slice); 64-row DB tiles double-buffered in LDS, one barrier per
// tile, the next tile prefetched into registers with unconditional (clamped) loads.  The
// accumulator starts at -||x||^2/2 (SqL2, L2) or 0 (NormCosine), so the MFMA result is a
// score s = x.q - ||x||^2/2 (or x.q): larger is nearer, a row past the slice starts at
// -Inf, and the approximate distance is ||q||^2 - 2 s (or 1 - s).  Each 16-query tile is
// tested against its queries' thresholds; a tile with a passing score appends to the
// queries' candidate lists in global memory (4 sub-lists per list, one per 16-lane group,
// written through a buffer resource at a running offset); a full sub-list triggers a
// compaction to about EXH_KC entries and raises the threshold (ssg_select.h).  At the end
// each list keeps exactly the slice's EXH_KC best rows by (score, id); they are written
// as approximate distances with slice_tau, a lower bound on the approximate distance of
// every row of the slice that was dropped.
constexpr int X_TR = 64;  // DB rows per LDS tile

template <int DP>
__global__ __launch_bounds__(256, 2) void exh_stage1_l(const float* __restrict__ X,
                                                       const float* __restrict__ xnorm, int64_t n,
                                                       int metric, const float* __restrict__ qpad,
                                                       const float* __restrict__ qnorm, int64_t m,
                                                       int64_t rows_per_slice, int nslices, uint2* list,
                                                       float* cand_d, uint32_t* cand_i, float* slice_tau,
                                                       const float* __restrict__ tau0, unsigned int* fail_count)
{
    // the merge that follows on the stream counts its failed certificates from zero
    if (fail_count && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) *fail_count = 0u;
    constexpr int KS = DP / 4;            // MFMA k-steps
    constexpr int LS = DP + 4;            // LDS row stride in floats (staggers banks)
    constexpr int NCT = 2;                // 16-query column tiles per wave
    constexpr int U4 = DP / 4;            // float4 per row
    constexpr int PF = X_TR * U4 / 256;   // float4 per thread per tile
    constexpr int CAPL = 512;             // list entries per (query, slice)
    constexpr int SC = CAPL / 4;          // sub-list entries
    static_assert(X_TR * U4 % 256 == 0 && PF <= 8, "tile must split evenly over the block");
    static_assert(EXH_QT == 4 * 16 * NCT, "queries per block");
    __shared__ __attribute__((aligned(16))) float s_x[2][X_TR * LS];
    __shared__ __attribute__((aligned(16))) float s_h[2][X_TR];
    __shared__ __attribute__((aligned(16))) uint32_t s_hist[4][256];
    __shared__ int s_cnt[4][NCT * 64];
    __shared__ float s_tau[4][NCT * 16];

    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int col = lane & 15, g = lane >> 4;
    const int64_t qb0 = (int64_t)blockIdx.x * EXH_QT;
    const int sl = blockIdx.y;
    const int64_t r0 = (int64_t)sl * rows_per_slice;
    const int64_t r1 = min(n, r0 + rows_per_slice);
    constexpr int kc = EXH_KC;
    const int64_t qstride = (int64_t)nslices * CAPL;
    uint2* lbase = list + (qb0 * nslices + sl) * (int64_t)CAPL;
    auto list_of = [&](int ql) -> uint2* { return lbase + ql * qstride; };

    // B operands: the lane's run of KS elements of query 32w + 16ct + col (query rows past
    // m read row m - 1; their results are never used)
    float bq[NCT][KS];
#pragma unroll
    for (int ct = 0; ct < NCT; ct++) {
        const int64_t qr = min(qb0 + 32 * w + 16 * ct + col, m - 1);
        const float4* qp = reinterpret_cast<const float4*>(qpad + qr * DP + g * KS);
#pragma unroll
        for (int s = 0; s < KS / 4; s++) {
            const float4 v = qp[s];
            bq[ct][4 * s] = v.x;
            bq[ct][4 * s + 1] = v.y;
            bq[ct][4 * s + 2] = v.z;
            bq[ct][4 * s + 3] = v.w;
        }
    }
    const uint32_t smask = (uint32_t)SC * 8u - 1u;
    const uint32_t full_at = (uint32_t)(SC - 5) * 8u;
    const __amdgpu_buffer_rsrc_t lrs = __builtin_amdgcn_make_buffer_rsrc(
        lbase, 0, (int)(((EXH_QT - 1) * qstride + CAPL) * (int64_t)sizeof(uint2)), 0x00020000);
    float tau[NCT];
    uint32_t wp[NCT];
#pragma unroll
    for (int ct = 0; ct < NCT; ct++) {
        // the starting threshold: a seed pass's (rows at or below it cannot be among the k
        // nearest, and slice_tau below reports it as the bound of the dropped rows)
        tau[ct] = tau0 ? tau0[min(qb0 + 32 * w + 16 * ct + col, m - 1)] : -__builtin_inff();
        wp[ct] = (uint32_t)(((32 * w + 16 * ct + col) * qstride + g * SC) * (int64_t)sizeof(uint2));
    }

    if (r0 < r1) {
        float4 pf0, pf1, pf2, pf3, pf4, pf5, pf6, pf7;  // the next tile
        float pfx = 0.f;                                  // the norm of its row tid % 64
        int64_t pft = r0;
        // unconditional loads, consumed only by stage()
        auto prefetch = [&](int64_t t) {
            pft = t;
            pfx = xnorm[min(t + (tid & (X_TR - 1)), r1 - 1)];
#define SSG_PF(I, R)                                                                 \
            if (I < PF) {                                                            \
                const int idx = tid + 256 * I;                                       \
                const int row = idx / U4, c = idx % U4;                              \
                const int64_t gr = min(t + row, r1 - 1);                             \
                R = reinterpret_cast<const float4*>(X + gr * DP)[c];                 \
            }
            SSG_PF(0, pf0) SSG_PF(1, pf1) SSG_PF(2, pf2) SSG_PF(3, pf3)
            SSG_PF(4, pf4) SSG_PF(5, pf5) SSG_PF(6, pf6) SSG_PF(7, pf7)
#undef SSG_PF
        };
        auto stage = [&](int b) {
#define SSG_ST(I, R)                                                                 \
            if (I < PF) {                                                            \
                const int idx = tid + 256 * I;                                       \
                const int row = idx / U4, c = idx % U4;                              \
                *reinterpret_cast<float4*>(&s_x[b][row * LS + 4 * c]) = R;           \
            }
            SSG_ST(0, pf0) SSG_ST(1, pf1) SSG_ST(2, pf2) SSG_ST(3, pf3)
            SSG_ST(4, pf4) SSG_ST(5, pf5) SSG_ST(6, pf6) SSG_ST(7, pf7)
#undef SSG_ST
            if (tid < X_TR)
                s_h[b][tid] = pft + tid < r1 ? (metric == M_NORMCOS ? 0.f : -0.5f * pfx) : -__builtin_inff();
        };
        prefetch(r0);
        stage(0);
        __syncthreads();
        int b = 0;
        for (int64_t t0 = r0; t0 < r1; t0 += X_TR) {
            const bool more = t0 + X_TR < r1;
            if (more) prefetch(t0 + X_TR);
            const float* xb = s_x[b];
            const float* hb = s_h[b];
            auto chain = [&](int rb, floatx4 (&acc)[NCT]) {
                const floatx4 c0 = *reinterpret_cast<const floatx4*>(&hb[rb * 16 + 4 * g]);
                const float* ap = xb + (rb * 16 + col) * LS + g * KS;
                floatx4 c[NCT];
#pragma unroll
                for (int ct = 0; ct < NCT; ct++) c[ct] = c0;
#pragma unroll
                for (int s4 = 0; s4 < KS / 4; s4++) {
                    const float4 av = *reinterpret_cast<const float4*>(ap + 4 * s4);
                    const float a4[4] = {av.x, av.y, av.z, av.w};
#pragma unroll
                    for (int u = 0; u < 4; u++)
#pragma unroll
                        for (int ct = 0; ct < NCT; ct++)
                            c[ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(a4[u], bq[ct][4 * s4 + u], c[ct], 0, 0, 0);
                }
#pragma unroll
                for (int ct = 0; ct < NCT; ct++) acc[ct] = c[ct];
            };
            floatx4 cur[NCT], nxt[NCT];
            chain(0, cur);
#pragma unroll
            for (int rb = 0; rb < X_TR / 16; rb++) {
                if (rb + 1 < X_TR / 16) chain(rb + 1, nxt);
                const uint32_t rowb = (uint32_t)(t0 + rb * 16 + 4 * g);
                uint64_t qfull = 0;  // bit 16 * ct + col: that query's list is compacted now
#pragma unroll
                for (int ct = 0; ct < NCT; ct++) {
                    const float mx = __builtin_elementwise_maximum(
                        __builtin_elementwise_maximum(cur[ct][0], cur[ct][1]),
                        __builtin_elementwise_maximum(cur[ct][2], cur[ct][3]));
                    if (__builtin_expect(__ballot(mx > tau[ct]) != 0, 0)) {
                        const float t = tau[ct];
                        uint32_t o = wp[ct];
#pragma unroll
                        for (int r = 0; r < 4; r++) {
                            if (cur[ct][r] > t) {
                                const u32x2 e = {__float_as_uint(cur[ct][r]), rowb + r};
                                __builtin_amdgcn_raw_buffer_store_b64(e, lrs, (int)o, 0, 0);
                                o += 8u;
                            }
                        }
                        wp[ct] = o;
                        // at most 4 appends per lane, query and row block: a sub-list with
                        // at most SC - 5 entries cannot fill up on the next one
                        const uint64_t f = __ballot((o & smask) > full_at);
                        qfull |= ((f | (f >> 16) | (f >> 32) | (f >> 48)) & 0xffffull) << (16 * ct);
                    }
                }
                if (__builtin_expect(qfull != 0, 0)) {
#pragma unroll
                    for (int ct = 0; ct < NCT; ct++) {
                        s_cnt[w][64 * ct + lane] = (int)((wp[ct] & smask) >> 3);
                        if (g == 0) s_tau[w][16 * ct + col] = tau[ct];
                    }
                    wave_sync();  // this wave's appends are visible to the reads
                    while (qfull) {
                        const int ql = __builtin_ctzll(qfull);
                        qfull &= qfull - 1;
                        int* cq = &s_cnt[w][64 * (ql >> 4) + (ql & 15)];
                        int kept = 0;
                        const float nt = compact_list_j<8>(list_of(32 * w + ql), SC, cq[0], cq[16], cq[32], cq[48],
                                                           kc, false, s_hist[w], lane, &kept);
                        if (lane < 4) cq[16 * lane] = (kept + 3 - lane) >> 2;
                        if (lane == 0) s_tau[w][ql] = nt;
                        lds_order();
                    }
#pragma unroll
                    for (int ct = 0; ct < NCT; ct++) {
                        wp[ct] = (wp[ct] & ~smask) + (uint32_t)s_cnt[w][64 * ct + lane] * 8u;
                        tau[ct] = s_tau[w][16 * ct + col];
                    }
                }
                if (rb + 1 < X_TR / 16) {
#pragma unroll
                    for (int ct = 0; ct < NCT; ct++) cur[ct] = nxt[ct];
                }
            }
            if (more) stage(b ^ 1);
            __syncthreads();
            b ^= 1;
        }
    }
    // final lists: the slice's EXH_KC best rows by (score, id), written as approximate
    // distances; slice_tau bounds every dropped row (its score is at most the larger of
    // the running threshold and the EXH_KC-th best score)
    wave_sync();
#pragma unroll
    for (int ct = 0; ct < NCT; ct++) {
        s_cnt[w][64 * ct + lane] = (int)((wp[ct] & smask) >> 3);
        if (g == 0) s_tau[w][16 * ct + col] = tau[ct];
    }
    lds_order();
    for (int ql = 0; ql < 32; ql++) {
        const int* cq = &s_cnt[w][64 * (ql >> 4) + (ql & 15)];
        static_assert(EXH_KC == 32, "two lists of at most EXH_KC entries fill one wave");
        // Two queries whose lists hold at most kc entries (nothing to select): each half of
        // the wave writes one of them out in list order (sub-list 0 first), as the
        // compaction would, without its round trip through the list.
        if (ql + 1 < 32) {
            const int* cn = &s_cnt[w][64 * ((ql + 1) >> 4) + ((ql + 1) & 15)];
            if (cq[0] + cq[16] + cq[32] + cq[48] <= kc && cn[0] + cn[16] + cn[32] + cn[48] <= kc) {
                const int qq = ql + (lane >> 5), i = lane & 31;
                const int* cc = &s_cnt[w][64 * (qq >> 4) + (qq & 15)];
                const int c0 = cc[0], c1 = cc[16], c2 = cc[32], c3 = cc[48];
                const int tot = c0 + c1 + c2 + c3;
                uint2 v = make_uint2(0u, 0xffffffffu);
                if (i < tot) {
                    int g = 0, pos = i;
                    if (pos >= c0) {
                        pos -= c0;
                        g = 1;
                        if (pos >= c1) {
                            pos -= c1;
                            g = 2;
                            if (pos >= c2) {
                                pos -= c2;
                                g = 3;
                            }
                        }
                    }
                    v = list_of(32 * w + qq)[g * SC + pos];
                }
                const float drop = s_tau[w][qq];
                const int64_t q = qb0 + 32 * w + qq;
                const float qn = q < m ? qnorm[q] : 0.f;
                auto dist_of = [&](float sc) { return metric == M_NORMCOS ? 1.0f - sc : qn - 2.0f * sc; };
                const int64_t base = (q * nslices + sl) * EXH_KC;
                cand_d[base + i] = i < tot ? dist_of(__uint_as_float(v.x)) : __builtin_inff();
                cand_i[base + i] = i < tot ? v.y : 0xffffffffu;
                if (i == 0)
                    slice_tau[q * nslices + sl] = drop == -__builtin_inff() ? __builtin_inff() : dist_of(drop);
                ql++;
                continue;
            }
        }
        int kept = 0;
        const float tk = compact_list_j<8>(list_of(32 * w + ql), SC, cq[0], cq[16], cq[32], cq[48], kc, true,
                                           s_hist[w], lane, &kept);
        const float drop = fmaxf(tk, s_tau[w][ql]);
        const int64_t q = qb0 + 32 * w + ql;
        const float qn = q < m ? qnorm[q] : 0.f;
        auto dist_of = [&](float sc) { return metric == M_NORMCOS ? 1.0f - sc : qn - 2.0f * sc; };
        const int64_t base = (q * nslices + sl) * EXH_KC;
        if (lane < EXH_KC) {
            const uint2 v = lane < kept ? list_of(32 * w + ql)[lane] : make_uint2(0u, 0xffffffffu);
            cand_d[base + lane] = lane < kept ? dist_of(__uint_as_float(v.x)) : __builtin_inff();
            cand_i[base + lane] = lane < kept ? v.y : 0xffffffffu;
        }
        if (lane == 0) slice_tau[q * nslices + sl] = drop == -__builtin_inff() ? __builtin_inff() : dist_of(drop);
    }
}

// Stage 2: merge, exact rescoring of SEL candidates, certificate.
template <int V>
__global__ __launch_bounds__(256) void exh_stage2(const float* __restrict__ X, int64_t n, int metric,
                                                  const float* __restrict__ qpad,
                                                  const float* __restrict__ qnorm, float xnmax,
                                                  int64_t m, int k, int nslices, int kc,
                                                  const float* __restrict__ cand_d,
                                                  const uint32_t* __restrict__ cand_i,
                                                  const float* __restrict__ slice_tau,
                                                  uint32_t* ids, float* dists, int kstride,
                                                  unsigned int* flags)
{
    constexpr int PER = 32;  // entries per lane (nslices * kc <= 2048)
    __shared__ uint32_t s_sel[4][SEL];
    __shared__ __attribute__((aligned(16))) uint32_t s_hist[4][256];
    const int lane = lane_id(), w = threadIdx.x >> 6;
    const int g = lane >> 4, j = lane & 15;
    const int64_t q = (int64_t)blockIdx.x * 4 + w;
    if (q >= m) return;
    const int total = nslices * kc;
    const int dp = 16 * V;

    uint32_t key[PER];
    uint32_t cid[PER];
    // Buffer loads over the query's entries (reads past them return 0), all issued before
    // any is used: plain loads were sunk into the branch of their use, which serialised
    // the wave's 64 round trips.
    const __amdgpu_buffer_rsrc_t ri = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<uint32_t*>(cand_i + q * total), 0, total * 4, 0x00020000);
    const __amdgpu_buffer_rsrc_t rd = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<float*>(cand_d + q * total), 0, total * 4, 0x00020000);
#pragma unroll
    for (int t = 0; t < PER; t++) {
        const int e = lane + 64 * t;
        cid[t] = __builtin_amdgcn_raw_buffer_load_b32(ri, 4 * e, 0, 0);
        key[t] = __builtin_amdgcn_raw_buffer_load_b32(rd, 4 * e, 0, 0);
    }
#pragma unroll
    for (int t = 0; t < PER; t++) {
        const int e = lane + 64 * t;
        const bool ok = e < total && cid[t] != 0xffffffffu;
        key[t] = ok ? sort_key(__uint_as_float(key[t])) : 0xffffffffu;
        cid[t] = ok ? cid[t] : 0xffffffffu;
    }
    float tmin = __builtin_inff();
    for (int s = lane; s < nslices; s += 64) tmin = fminf(tmin, slice_tau[q * nslices + s]);
    for (int o = 32; o >= 1; o >>= 1) tmin = fminf(tmin, __shfl_xor(tmin, o));

    int nvalid = 0;
#pragma unroll
    for (int t = 0; t < PER; t++) nvalid += key[t] != 0xffffffffu ? 1 : 0;
    for (int o = 32; o >= 1; o >>= 1) nvalid += __shfl_xor(nvalid, o);

    // radix select: T = SEL-th smallest key (or the largest valid one), as the SEL-th
    // largest of the complemented keys (8-bit digits, LDS histograms)
    const int want0 = nvalid < SEL ? nvalid : SEL;
    uint32_t T = 0xffffffffu;
    if (want0 > 0) {
        uint32_t ik[PER];
        bool act[PER];
#pragma unroll
        for (int t = 0; t < PER; t++) {
            act[t] = key[t] != 0xffffffffu;
            ik[t] = ~key[t];
        }
        uint32_t kand, diff;
        key_bits<PER>(ik, act, &kand, &diff);
        int above = 0;
        T = ~kth_largest<PER>(ik, act, want0, 4, s_hist[w], lane, &above, kand, diff);
    }
    // gather the selection in index order: keys < T, then keys == T while room
    int nsel = 0;
#pragma unroll
    for (int t = 0; t < PER; t++) {
        const bool lt = key[t] < T;
        const uint64_t bl = __ballot(lt);
        if (lt) s_sel[w][nsel + __popcll(bl & ((1ull << lane) - 1ull))] = cid[t];
        nsel += __popcll(bl);
    }
#pragma unroll
    for (int t = 0; t < PER; t++) {
        const bool eq = key[t] == T && key[t] != 0xffffffffu;
        const uint64_t be = __ballot(eq);
        const int p = nsel + __popcll(be & ((1ull << lane) - 1ull));
        if (eq && p < want0) s_sel[w][p] = cid[t];
        nsel = min(want0, nsel + __popcll(be));
    }
    __builtin_amdgcn_wave_barrier();
    asm volatile("" ::: "memory");
    const bool dropped = nvalid > nsel;

    // exact canonical distances of the selection: 4 per round, 16 rounds
    float qv[V];
    {
        const float* qp = qpad + q * dp + j * V;
#pragma unroll
        for (int t = 0; t < V; t++) qv[t] = qp[t];
    }
    float myd = __builtin_inff();
    double mys = 0.0;
    uint32_t myid = 0xffffffffu;
    if constexpr (V <= 4) {
        // short rows: every selected row's piece is loaded before any is used, so the
        // wave waits for one round of loads instead of SEL / 16
        static_assert(SEL == 64, "16 rows per 16-lane group");
        float xv[16][V];
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int u = 4 * r + g;
            if (u < nsel) {
                load_run<V>(X + (size_t)s_sel[w][u] * dp + j * V, xv[r]);
            } else {
#pragma unroll
                for (int t = 0; t < V; t++) xv[r][t] = 0.f;
            }
        }
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const double sr = group16_sum(run_partial<V>(metric, xv[r], qv));
            const float d = finish_dist(metric, sr);
            // candidate u = 4r + g goes to lane u
            const float dt = __shfl(d, (lane & 3) * 16);
            const double st = __shfl(sr, (lane & 3) * 16);
            if ((lane >> 2) == r && lane < nsel) {
                myd = dt;
                mys = st;
                myid = s_sel[w][lane];
            }
        }
    } else {
        for (int r0 = 0; r0 < nsel; r0 += 16) {
            float xv[4][V];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int u = r0 + 4 * r + g;
                if (u < nsel) {
                    load_run<V>(X + (size_t)s_sel[w][u] * dp + j * V, xv[r]);
                } else {
#pragma unroll
                    for (int t = 0; t < V; t++) xv[r][t] = 0.f;
                }
            }
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const double s = group16_sum(run_partial<V>(metric, xv[r], qv));
                const float d = finish_dist(metric, s);
                // candidate u = r0 + 4r + g goes to lane u
                const float dt = __shfl(d, (lane & 3) * 16);
                const double st = __shfl(s, (lane & 3) * 16);
                if (lane >= r0 && lane < r0 + 16 && ((lane - r0) >> 2) == r && lane < nsel) {
                    myd = dt;
                    mys = st;
                    myid = s_sel[w][lane];
                }
            }
        }
    }
    int rank = 0;
    for (int t = 0; t < nsel; t++) {
        const float dt = rl_f(myd, t);
        const uint32_t it = rl_u(myid, t);
        rank += (dt < myd || (dt == myd && it < myid)) ? 1 : 0;
    }
    if (lane < nsel && rank < k) {
        ids[q * kstride + rank] = myid + 1u;
        dists[q * kstride + rank] = myd;
    }
    if (lane >= nsel && lane < k) {
        ids[q * kstride + lane] = 0u;
        dists[q * kstride + lane] = __builtin_inff();
    }
    // certificate
    float bound = tmin;
    if (dropped) bound = fminf(bound, key_float(T));
    bool ok = true;
    if (bound < __builtin_inff()) {
        if (nsel < k) {
            ok = false;
        } else {
            const uint64_t bk = __ballot(lane < nsel && rank == k - 1);
            const double sk = __shfl(mys, __ffsll((unsigned long long)bk) - 1);
            const double dk = metric == M_NORMCOS ? 1.0 - sk : sk;
            const double err = (double)(dp + 8) * 2.384185791015625e-07 * ((double)xnmax + (double)qnorm[q]);
            ok = ((double)bound - err) > dk + fabs(dk) * 1e-6 + 1e-30;
        }
    }
    if (!ok && lane == 0) {
        const unsigned int p = atomicAdd(&flags[0], 1u);
        flags[1 + p] = (unsigned int)q;
    }
}

// Seed of the starting thresholds from a sample: a block takes 4 queries and scores S <=
// 2048 rows spread over the database (row i * stride) against all four, each thread 8
// rows, in FP32 with the score's form (x.q - |x|^2/2, or x.q).  Lane j of wave w holds
// query w's scores of sampled rows j, j + 64, ...; the wave takes the k-th largest of the
// 128 values formed by each lane's two largest.  At least k sampled scores reach that
// value, so it is at most the k-th largest sampled score (equal to it unless three of the
// k largest share a lane), and lowered by the certificate's error allowance it lies below
// the matrix-core score of each of the k best sampled rows: each of them, and so each of
// the k nearest rows, passes the main pass's threshold.  Fewer than k rows, or k outside
// 1 .. 128: -Inf.
constexpr int SEED_S = 2048;
template <int DP>
__global__ __launch_bounds__(256) void exh_seed_sample_kernel(const float* __restrict__ X,
                                                              const float* __restrict__ xnorm,
                                                              int64_t stride, int S,
                                                              const float* __restrict__ qpad,
                                                              const float* __restrict__ qnorm,
                                                              int64_t m_valid, int64_t m_pad, int k,
                                                              int metric, float xnmax, float* tau0)
{
    __shared__ __attribute__((aligned(16))) float s_q[4][128];
    __shared__ uint32_t s_key[4][SEED_S];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int64_t q0 = (int64_t)blockIdx.x * 4;
    constexpr int U = DP / 4;                             // float4 pieces per row
    constexpr int RU = DP <= 32 ? 4 : DP <= 64 ? 2 : 1;  // rows whose loads are in flight together
    static_assert((SEED_S / 256) % RU == 0, "rows per thread");
    for (int e = tid; e < 4 * DP; e += 256) {
        const int64_t q = min(q0 + e / DP, m_pad - 1);
        s_q[e / DP][e % DP] = qpad[q * DP + e % DP];
    }
    __syncthreads();
    for (int r0 = tid; r0 < SEED_S; r0 += 256 * RU) {
        // every row's loads first (rows past S clamped to the last sampled row)
        float4 xa[RU][U];
        float h[RU];
#pragma unroll
        for (int i = 0; i < RU; i++) {
            const int64_t row = (int64_t)min(r0 + 256 * i, S - 1) * stride;
            const float4* xp = reinterpret_cast<const float4*>(X + row * DP);
#pragma unroll
            for (int u = 0; u < U; u++) xa[i][u] = xp[u];
            h[i] = metric == M_NORMCOS ? 0.f : -0.5f * xnorm[row];
        }
#pragma unroll
        for (int i = 0; i < RU; i++) {
            const int r = r0 + 256 * i;
            float acc[4];
#pragma unroll
            for (int t = 0; t < 4; t++) acc[t] = h[i];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const float4 a = xa[i][u];
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    const float4 b = reinterpret_cast<const float4*>(s_q[t])[u];
                    acc[t] = __builtin_fmaf(a.x, b.x, acc[t]);
                    acc[t] = __builtin_fmaf(a.y, b.y, acc[t]);
                    acc[t] = __builtin_fmaf(a.z, b.z, acc[t]);
                    acc[t] = __builtin_fmaf(a.w, b.w, acc[t]);
                }
            }
#pragma unroll
            for (int t = 0; t < 4; t++) s_key[t][r] = r < S ? score_key(acc[t]) : 0u;  // 0: below every score
        }
    }
    __syncthreads();
    const int64_t q = q0 + w;
    if (q >= m_pad) return;
    if (q >= m_valid || S < k || k < 1 || k > 128) {
        if (lane == 0) tau0[q] = -__builtin_inff();
        return;
    }
    constexpr int PER = SEED_S / 64;
    uint32_t m1 = 0u, m2 = 0u;  // this lane's two largest keys, m1 >= m2
#pragma unroll
    for (int t = 0; t < PER; t++) {
        const uint32_t x = s_key[w][lane + 64 * t];
        m2 = max(m2, min(m1, x));
        m1 = max(m1, x);
    }
    // ranks in the wave's 128 values, ordered by value (larger first), then lane, then
    // the lane's first before its second: a permutation of 0 .. 127
    int r1 = 0, r2 = 0;
#pragma unroll
    for (int j = 0; j < 64; j++) {
        const uint32_t o1 = rl_u(m1, j), o2 = rl_u(m2, j);
        r1 += ((o1 > m1 || (o1 == m1 && j < lane)) ? 1 : 0) + ((o2 > m1 || (o2 == m1 && j < lane)) ? 1 : 0);
        r2 += ((o1 > m2 || (o1 == m2 && j <= lane)) ? 1 : 0) + ((o2 > m2 || (o2 == m2 && j < lane)) ? 1 : 0);
    }
    const uint64_t b1 = __ballot(r1 == k - 1), b2 = __ballot(r2 == k - 1);
    const uint32_t prefix = b1 ? rl_u(m1, __ffsll((unsigned long long)b1) - 1)
                               : rl_u(m2, __ffsll((unsigned long long)b2) - 1);
    if (lane == 0) {
        const float sk = key_score(prefix);
        const float err = (float)(DP + 8) * 2.384185791015625e-07f * (xnmax + qnorm[q]);
        tau0[q] = __builtin_isfinite(sk) ? sk - err : -__builtin_inff();
    }
}

hipError_t exh_launch_seed_sample(const float* X, const float* xnorm, int32_t dp, int64_t n,
                                  const ExhWork& w, int64_t m_valid, int64_t m_pad, int k, int metric,
                                  float xnmax, float* tau0, hipStream_t s)
{
    if (dp > 128) return hipErrorInvalidValue;
    const int S = (int)std::min<int64_t>(n, SEED_S);
    const int64_t stride = n / S;
    const dim3 grid((unsigned)((m_pad + 3) / 4));
    switch (dp) {
#define SSG_SS(D) \
    case D: hipLaunchKernelGGL((exh_seed_sample_kernel<D>), grid, dim3(256), 0, s, X, xnorm, stride, S, w.qpad, \
                               w.qnorm, m_valid, m_pad, k, metric, xnmax, tau0); break;
        SSG_SS(16) SSG_SS(32) SSG_SS(48) SSG_SS(64) SSG_SS(80) SSG_SS(96) SSG_SS(112) SSG_SS(128)
#undef SSG_SS
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Stage 3a: exact rescans.  Block (f, part): part of the rows for flagged query f;
// each wave scans a contiguous range in ascending id order into a KnnSorted.
template <int V>
__global__ __launch_bounds__(256) void exh_rescan(const float* __restrict__ X, int64_t n, int metric,
                                                  const float* __restrict__ qpad,
                                                  const unsigned int* flags, int f0, int k,
                                                  uint32_t* part_i, float* part_d, int nparts)
{
    const int lane = lane_id(), w = threadIdx.x >> 6;
    const int g = lane >> 4, j = lane & 15;
    const int f = f0 + blockIdx.x;
    if (f >= (int)flags[0]) return;
    const int64_t q = flags[1 + f];
    const int dp = 16 * V;
    const int wp = blockIdx.y * 4 + w;
    const int64_t per = (n + (int64_t)nparts * 4 - 1) / ((int64_t)nparts * 4);
    const int64_t a0 = (int64_t)wp * per, a1 = min(n, a0 + per);
    float qv[V];
#pragma unroll
    for (int t = 0; t < V; t++) qv[t] = qpad[q * dp + j * V + t];
    RegQueue<1, false> res;
    res.init(k);
    for (int64_t b = a0; b < a1; b += 4) {
        float xv[V];
        const int64_t row = b + g;
        if (row < a1) {
            load_run<V>(X + row * dp + j * V, xv);
        } else {
#pragma unroll
            for (int t = 0; t < V; t++) xv[t] = 0.f;
        }
        const float d = finish_dist(metric, group16_sum(run_partial<V>(metric, xv, qv)));
        for (int r = 0; r < 4 && b + r < a1; r++) res.push(rl_f(d, 16 * r), (uint32_t)(b + r), 0, 0);
    }
    const int64_t base = ((int64_t)blockIdx.x * nparts * 4 + wp) * k;
    if (lane < k) {
        part_i[base + lane] = lane < res.len ? res.id[0] : 0xffffffffu;
        part_d[base + lane] = lane < res.len ? res.d[0] : __builtin_inff();
    }
}

// Stage 3b: merge the partial lists in ascending row-range order.
__global__ __launch_bounds__(64) void exh_rescan_merge(const unsigned int* flags, int f0, int k,
                                                      const uint32_t* part_i, const float* part_d,
                                                      int nparts, uint32_t* ids, float* dists,
                                                      int kstride)
{
    const int lane = lane_id();
    const int f = f0 + blockIdx.x;
    if (f >= (int)flags[0]) return;
    const int64_t q = flags[1 + f];
    RegQueue<1, false> res;
    res.init(k);
    const int lists = nparts * 4;
    for (int l = 0; l < lists; l++) {
        const int64_t base = ((int64_t)blockIdx.x * lists + l) * k;
        const uint32_t ci = lane < k ? part_i[base + lane] : 0xffffffffu;
        const float cd = lane < k ? part_d[base + lane] : __builtin_inff();
        for (int t = 0; t < k; t++) {
            const uint32_t it = rl_u(ci, t);
            if (it == 0xffffffffu) break;
            res.push(rl_f(cd, t), it, 0, 0);
        }
    }
    if (lane < k) {
        ids[q * kstride + lane] = lane < res.len ? res.id[0] + 1u : 0u;
        dists[q * kstride + lane] = lane < res.len ? res.d[0] : __builtin_inff();
    }
}

// Pad rows to dp floats and compute FP32 squared norms.
// 64 rows per block of 256 threads, in chunks of 128 values: the block reads and writes
// each chunk coalesced (rows past `rows` and values past `dim` are zero) and stages it in
// LDS, where thread t < 64 continues row t's squared norm, summed in element order.
__global__ __launch_bounds__(256) void pad_rows_kernel(const float* __restrict__ src, int64_t rows,
                                                       int64_t rows_pad, int dim, float* __restrict__ dst, int dp,
                                                       float* norms)
{
    constexpr int RB = 64, CH = 128;
    __shared__ float s_v[RB][CH + 1];
    const int tid = threadIdx.x;
    const int64_t r0 = (int64_t)blockIdx.x * RB;
    float s = 0.f;
    for (int c0 = 0; c0 < dp; c0 += CH) {
        const int cw = min(CH, dp - c0);
        for (int e = tid; e < RB * cw; e += 256) {
            const int rl = e / cw, i = c0 + e % cw;
            const int64_t r = r0 + rl;
            float v = 0.f;
            if (r < rows && i < dim) v = src[r * dim + i];
            if (r < rows_pad) dst[r * dp + i] = v;
            s_v[rl][i - c0] = v;
        }
        __syncthreads();
        if (tid < RB)
            for (int i = 0; i < cw; i++) s = __builtin_fmaf(s_v[tid][i], s_v[tid][i], s);
        __syncthreads();
    }
    if (norms && tid < RB && r0 + tid < rows_pad) norms[r0 + tid] = s;
}

hipError_t launch_pad_rows(const float* src, int64_t rows, int32_t dim, float* dst, int32_t dp,
                           float* norms, hipStream_t s, int64_t rows_pad)
{
    if (rows_pad < rows) rows_pad = rows;
    if (rows_pad <= 0) return hipSuccess;
    hipLaunchKernelGGL(pad_rows_kernel, dim3((unsigned)((rows_pad + 63) / 64)), dim3(256), 0, s, src, rows,
                       rows_pad, dim, dst, dp, norms);
    return hipGetLastError();
}

// rerank! batch form: rescore the non-zero-id prefix, stable sort by distance.
template <int V>
__global__ __launch_bounds__(256) void rerank_kernel(const float* __restrict__ X, int metric,
                                                     const float* __restrict__ Q, int qdim,
                                                     int64_t m, int k, uint32_t* ids, float* dists)
{
    const int lane = lane_id(), w = threadIdx.x >> 6;
    const int g = lane >> 4, j = lane & 15;
    const int64_t q = (int64_t)blockIdx.x * 4 + w;
    if (q >= m) return;
    const int dp = 16 * V;
    const uint32_t myid = lane < k ? ids[q * k + lane] : 0u;
    const uint64_t z = __ballot(lane < k && myid == 0u);
    const int len = z ? (__ffsll((unsigned long long)z) - 1) : k;
    float qv[V];
#pragma unroll
    for (int t = 0; t < V; t++) {
        const int i = j * V + t;
        qv[t] = i < qdim ? Q[q * qdim + i] : 0.f;
    }
    float myd = lane < k ? dists[q * k + lane] : 0.f;
    for (int r0 = 0; r0 < len; r0 += 4) {
        const int u = r0 + g;
        float xv[V];
        const uint32_t id = __shfl(myid, u < len ? u : 0);
        if (u < len) {
            load_run<V>(X + (size_t)(id - 1u) * dp + j * V, xv);
        } else {
#pragma unroll
            for (int t = 0; t < V; t++) xv[t] = 0.f;
        }
        const float d = finish_dist(metric, group16_sum(run_partial<V>(metric, xv, qv)));
        const float dt = __shfl(d, ((lane - r0) & 3) * 16);
        if (lane >= r0 && lane < r0 + 4 && lane < len) myd = dt;
    }
    int rank = 0;
    for (int t = 0; t < len; t++) {
        const float dt = rl_f(myd, t);
        rank += (dt < myd || (dt == myd && t < lane)) ? 1 : 0;
    }
    if (lane < len) {
        ids[q * k + rank] = myid;
        dists[q * k + rank] = myd;
    }
}

// rerank! for columns of 65..1024 rows: one wave per column; the prefix's ids and
// canonical distances are staged in LDS (k * 8 bytes per wave), then every lane ranks its
// entries by counting, over LDS broadcast reads, the entries that sort before them
// ((distance, position) order: a stable sort) and writes them at their rank.
template <int V>
__global__ __launch_bounds__(256) void rerank_wide_kernel(const float* __restrict__ X, int metric,
                                                          const float* __restrict__ Q, int qdim,
                                                          int64_t m, int k, uint32_t* ids, float* dists)
{
    extern __shared__ float rr_lds[];
    const int lane = lane_id(), w = threadIdx.x >> 6;
    const int g = lane >> 4, j = lane & 15;
    const int64_t q = (int64_t)blockIdx.x * 4 + w;
    if (q >= m) return;
    const int dp = 16 * V;
    float* sd = rr_lds + (size_t)w * 2 * k;
    uint32_t* si = reinterpret_cast<uint32_t*>(sd + k);
    uint32_t* row = ids + q * k;
    float* drow = dists + q * k;
    int len = k;
    for (int c0 = 0; c0 < k; c0 += 64) {
        const int c = c0 + lane;
        const uint32_t v = c < k ? row[c] : 1u;
        if (c < k) si[c] = v;
        const uint64_t z = __ballot(c < k && v == 0u);
        if (z) {
            len = c0 + __ffsll((unsigned long long)z) - 1;
            break;
        }
    }
    float qv[V];
#pragma unroll
    for (int t = 0; t < V; t++) {
        const int i = j * V + t;
        qv[t] = i < qdim ? Q[q * qdim + i] : 0.f;
    }
    wave_sync();
    for (int r0 = 0; r0 < len; r0 += 4) {
        const int u = r0 + g;
        const uint32_t id = si[u < len ? u : 0];  // in-range entry only
        float xv[V];
        load_run<V>(X + (size_t)(id - 1u) * dp + j * V, xv);
        const double s = group16_tree(run_partial<V>(metric, xv, qv));
        if (j == 15 && u < len) sd[u] = finish_dist(metric, s);
    }
    wave_sync();
    for (int i0 = 0; i0 < len; i0 += 64) {
        const int i = i0 + lane;
        const float di = i < len ? sd[i] : 0.f;
        int rank = 0;
        for (int t = 0; t < len; t++) {
            const float dt = sd[t];
            rank += (dt < di || (dt == di && t < i)) ? 1 : 0;
        }
        if (i < len) {
            row[rank] = si[i];
            drow[rank] = di;
        }
    }
}

// Stage-1 blocks resident per CU: LDS-bound (row tile + candidate lists), at most two
// by registers.
// Stage-1 blocks resident per CU: two for the list kernel (dp <= 128); LDS-bound (row
// tile + candidate lists), at most two, for the dp > 128 kernel.
static int stage1_blocks_per_cu(int dp)
{
    if (dp <= 128) return 2;
    const int lds = TR * (dp + 4) * 4 + TR * 4 + TR * CAP * 8 + TR * 8;
    const int by_lds = (160 * 1024) / lds;
    return by_lds < 1 ? 1 : (by_lds > 2 ? 2 : by_lds);
}

// DB slices per query tile: enough blocks to fill the chip, and among the admissible
// counts the one whose last round of resident blocks is fullest (every block of a launch
// does the same work, so a partly filled last round idles the rest of the chip).
int32_t exhaustive_slices(int64_t n, int64_t m, int num_cu, int dp)
{
    const int64_t qtiles = (m + EXH_QT - 1) / EXH_QT;
    const int64_t cap = (int64_t)num_cu * stage1_blocks_per_cu(dp);
    int64_t smax = (n + 4 * TR - 1) / (4 * TR);  // at least four row tiles per slice
    if (smax > 2048 / EXH_KC) smax = 2048 / EXH_KC;
    if (smax < 1) smax = 1;
    int64_t smin = (cap + qtiles - 1) / qtiles;
    if (smin > smax) smin = smax;
    if (smin < 1) smin = 1;
    int64_t best = smin;
    double best_fill = -1.0;
    for (int64_t s = smin; s <= smax && s <= 4 * smin; s++) {
        const int64_t blocks = qtiles * s;
        const int64_t rounds = (blocks + cap - 1) / cap;
        const double fill = (double)blocks / (double)(rounds * cap);
        if (fill > best_fill + 1e-9) {
            best_fill = fill;
            best = s;
        }
    }
    return (int32_t)best;
}

template <int DP>
static void stage1_dp(dim3 grid, hipStream_t s, const ExhArgs& a, const ExhWork& w, int64_t rps,
                      int ns)
{
    if constexpr (DP <= 128) {
        hipLaunchKernelGGL((exh_stage1_l<DP>), grid, dim3(256), 0, s, a.X, a.xnorm, a.n, a.metric, w.qpad,
                           w.qnorm, w.m_valid, rps, ns, w.list, w.cand_d, w.cand_i, w.slice_tau, w.tau0,
                           w.flags);
    } else {
        // 64-query blocks; query rows up to roundup(m_valid, 64) (the padded rows of X when
        // the queries are the database itself)
        const dim3 g64((unsigned)((w.m_valid + TR - 1) / TR), grid.y);
        hipLaunchKernelGGL((exh_stage1_lds<DP, EXH_KC>), g64, dim3(256), 0, s, a.X, a.xnorm, a.n, a.metric,
                           w.qpad, w.qnorm, rps, ns, w.cand_d, w.cand_i, w.slice_tau, w.flags);
    }
}

hipError_t exh_launch_stage1(const ExhArgs& a, const ExhWork& w, int nslices, hipStream_t s)
{
    const int64_t qtiles = (a.m + EXH_QT - 1) / EXH_QT;
    int64_t rps = (a.n + nslices - 1) / nslices;
    rps = (rps + TR - 1) / TR * TR;
    if (a.dp <= 128 && !w.list) return hipErrorInvalidValue;
    const dim3 grid((unsigned)qtiles, (unsigned)nslices);
    switch (a.dp) {
    case 16: stage1_dp<16>(grid, s, a, w, rps, nslices); break;
    case 32: stage1_dp<32>(grid, s, a, w, rps, nslices); break;
    case 48: stage1_dp<48>(grid, s, a, w, rps, nslices); break;
    case 64: stage1_dp<64>(grid, s, a, w, rps, nslices); break;
    case 80: stage1_dp<80>(grid, s, a, w, rps, nslices); break;
    case 96: stage1_dp<96>(grid, s, a, w, rps, nslices); break;
    case 112: stage1_dp<112>(grid, s, a, w, rps, nslices); break;
    case 128: stage1_dp<128>(grid, s, a, w, rps, nslices); break;
    case 192: stage1_dp<192>(grid, s, a, w, rps, nslices); break;
    case 256: stage1_dp<256>(grid, s, a, w, rps, nslices); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t exh_launch_stage2(const ExhArgs& a, const ExhWork& w, float xnmax, int nslices,
                             uint32_t* ids, float* dists, int kstride, hipStream_t s)
{
    const dim3 grid((unsigned)((a.m + 3) / 4)), block(256);
    switch (a.dp / 16) {
#define SSG_S2(VV) \
    case VV: hipLaunchKernelGGL((exh_stage2<VV>), grid, block, 0, s, a.X, a.n, a.metric, w.qpad, \
                                w.qnorm, xnmax, a.m, a.k, nslices, EXH_KC, w.cand_d, w.cand_i, \
                                w.slice_tau, ids, dists, kstride, w.flags); break;
        SSG_S2(1) SSG_S2(2) SSG_S2(3) SSG_S2(4) SSG_S2(5) SSG_S2(6) SSG_S2(7) SSG_S2(8) SSG_S2(12) SSG_S2(16)
#undef SSG_S2
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t exh_launch_rescan(const ExhArgs& a, const ExhWork& w, int f0, int nf, uint32_t* part_i,
                             float* part_d, int nparts, uint32_t* ids, float* dists, int kstride,
                             hipStream_t s)
{
    const dim3 grid((unsigned)nf, (unsigned)nparts), block(256);
    switch (a.dp / 16) {
#define SSG_S3(VV) \
    case VV: hipLaunchKernelGGL((exh_rescan<VV>), grid, block, 0, s, a.X, a.n, a.metric, w.qpad, \
                                w.flags, f0, a.k, part_i, part_d, nparts); break;
        SSG_S3(1) SSG_S3(2) SSG_S3(3) SSG_S3(4) SSG_S3(5) SSG_S3(6) SSG_S3(7) SSG_S3(8) SSG_S3(12) SSG_S3(16)
#undef SSG_S3
    default: return hipErrorInvalidValue;
    }
    hipLaunchKernelGGL(exh_rescan_merge, dim3((unsigned)nf), dim3(64), 0, s, w.flags, f0, a.k,
                       part_i, part_d, nparts, ids, dists, kstride);
    return hipGetLastError();
}

// Exact exhaustive k-NN for k > EXH_KC (more than stage 1 keeps per slice): one block per
// query scans the rows in ascending id order with the canonical distance (16 lanes per
// row, 16 rows per round, 256 rows per tile).  Keys (d, id) below the current k-th best
// are appended to a pending run in LDS; when the run may overflow, list and run are merged
// by a bitonic sort of the whole buffer and the first k kept.  The result is the first k
// in (d, id) order, which is what the ascending scan with KnnSorted pushes leaves.
// Dynamic LDS: P2 keys (a power of two >= k + 256): [0, k) the sorted list, [k, P2) the
// pending run.
__device__ __forceinline__ void bitonic_sort_lds(uint64_t* B, int P2)
{
    for (int size = 2; size <= P2; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = threadIdx.x; i < (P2 >> 1); i += blockDim.x) {
                const int pos = 2 * i - (i & (stride - 1));
                const int partner = pos + stride;
                const bool up = (pos & size) == 0;
                const uint64_t a = B[pos], b = B[partner];
                if ((a > b) == up) {
                    B[pos] = b;
                    B[partner] = a;
                }
            }
            __syncthreads();
        }
    }
}

template <int V>
__global__ __launch_bounds__(256) void exh_wide_kernel(const float* __restrict__ X, int64_t n,
                                                       int metric, const float* __restrict__ Qp,
                                                       int k, int P2, uint32_t* ids, float* dists,
                                                       int kstride)
{
    extern __shared__ uint64_t B[];
    __shared__ int s_pend, s_len;
    __shared__ uint64_t s_thr;
    const int tid = threadIdx.x, grp = tid >> 4, j = tid & 15;
    const int64_t q = blockIdx.x;
    const int dp = 16 * V;
    float qv[V];
    load_run<V>(Qp + (size_t)q * dp + j * V, qv);
    if (tid == 0) {
        s_pend = 0;
        s_len = 0;
        s_thr = ~0ull;
    }
    __syncthreads();
    const int pend_cap = P2 - k;
    for (int64_t r0 = 0; r0 < n; r0 += 256) {
        const uint64_t thr = s_thr;
        for (int rr = 0; rr < 16; rr++) {
            const int64_t row = r0 + rr * 16 + grp;
            float xv[V];
            load_run<V>(X + (size_t)(row < n ? row : 0) * dp + j * V, xv);
            const float d = finish_dist(metric, group16_sum(run_partial<V>(metric, xv, qv)));
            if (j == 0 && row < n) {
                const uint64_t key = ((uint64_t)sort_key(d) << 32) | (uint64_t)(uint32_t)row;
                if (key < thr) B[k + atomicAdd(&s_pend, 1)] = key;
            }
        }
        __syncthreads();
        const int pend = s_pend;
        if (pend > pend_cap - 256 || (r0 + 256 >= n && pend > 0)) {
            const int len = s_len;
            for (int i = tid; i < P2; i += 256)
                if ((i >= len && i < k) || i >= k + pend) B[i] = ~0ull;
            __syncthreads();
            bitonic_sort_lds(B, P2);
            if (tid == 0) {
                s_len = len + pend < k ? len + pend : k;
                s_pend = 0;
                s_thr = len + pend >= k ? B[k - 1] : ~0ull;
            }
            __syncthreads();
        }
    }
    const int len = s_len;
    for (int i = tid; i < kstride; i += 256) {
        const bool has = i < len;
        ids[(size_t)q * kstride + i] = has ? (uint32_t)(B[i] & 0xffffffffu) + 1u : 0u;
        dists[(size_t)q * kstride + i] = has ? key_float((uint32_t)(B[i] >> 32)) : __builtin_inff();
    }
}

hipError_t launch_exh_wide(const float* X, int64_t n, int32_t dp, int32_t metric, const float* Qp,
                           int64_t m, int32_t k, uint32_t* ids, float* dists, int32_t kstride,
                           hipStream_t s)
{
    if (k < 1 || k > EXH_WIDE_KMAX) return hipErrorInvalidValue;
    int P2 = 512;
    while (P2 < k + 256) P2 <<= 1;
    const size_t lds = (size_t)P2 * sizeof(uint64_t);
    const dim3 grid((unsigned)m), block(256);
    switch (dp / 16) {
#define SSG_EW(VV) \
    case VV: hipLaunchKernelGGL((exh_wide_kernel<VV>), grid, block, lds, s, X, n, metric, Qp, k, P2, ids, dists, kstride); break;
        SSG_EW(1) SSG_EW(2) SSG_EW(3) SSG_EW(4) SSG_EW(5) SSG_EW(6) SSG_EW(7) SSG_EW(8) SSG_EW(12) SSG_EW(16) SSG_EW(24) SSG_EW(32) SSG_EW(48) SSG_EW(64)
#undef SSG_EW
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_rerank(const float* X, int32_t dp, int32_t metric, const float* Q, int32_t qdim,
                         int64_t m, int32_t k, uint32_t* ids, float* dists, hipStream_t s)
{
    const dim3 grid((unsigned)((m + 3) / 4)), block(256);
    if (k > 1024) return hipErrorInvalidValue;
    if (k > 64) {
        const size_t lds = (size_t)4 * 2 * k * sizeof(float);
        switch (dp / 16) {
#define SSG_RW(VV) \
    case VV: hipLaunchKernelGGL((rerank_wide_kernel<VV>), grid, block, lds, s, X, metric, Q, qdim, m, k, ids, dists); break;
            SSG_RW(1) SSG_RW(2) SSG_RW(3) SSG_RW(4) SSG_RW(5) SSG_RW(6) SSG_RW(7) SSG_RW(8) SSG_RW(12) SSG_RW(16) SSG_RW(24) SSG_RW(32) SSG_RW(48) SSG_RW(64)
#undef SSG_RW
        default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    switch (dp / 16) {
#define SSG_RR(VV) \
    case VV: hipLaunchKernelGGL((rerank_kernel<VV>), grid, block, 0, s, X, metric, Q, qdim, m, k, ids, dists); break;
        SSG_RR(1) SSG_RR(2) SSG_RR(3) SSG_RR(4) SSG_RR(5) SSG_RR(6) SSG_RR(7) SSG_RR(8) SSG_RR(12) SSG_RR(16) SSG_RR(24) SSG_RR(32) SSG_RR(48) SSG_RR(64)
#undef SSG_RR
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace ssg
