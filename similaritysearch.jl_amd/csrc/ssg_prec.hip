// ssg_prec.hip -- precision form of exhaustive k-NN and allknn for gfx950: a candidate
// pass with BF16 operands on the matrix cores, then rerank! of the candidates with the
// canonical FP32 distance.
//
// The reference computes every distance in Float32 (src/exact/sequential-exhaustive.jl:68-80,
// src/allknn.jl:66-110); the precision form is the allknn-with-rerank pattern of
// src/rerank.jl:22-98 with the candidate queue filled from low-precision dot products.
//
// Candidate pass (prec_stage1): block = 256 queries x one DB slice, 4 waves of 64
//   queries.  A wave keeps its queries' BF16 rows as the B operands of
//   v_mfma_f32_16x16x32_bf16 in registers for the whole slice; DB rows are staged in
//   LDS in 128-row tiles (double-buffered, one barrier per tile, the next tile
//   prefetched into registers while the current one is multiplied) and shared by the 4
//   waves.  The accumulator starts at -||x||^2/2 (SqL2, L2) or 0 (NormCosine), so the
//   MFMA result is the score s = x.q - ||x||^2/2 (or x.q): larger is nearer, and an
//   invalid row starts at -Inf.  A lane holds 4 rows x 1 query per 16x16 tile; a tile
//   is tested with two max operations and one compare against the query's threshold,
//   and only a wave with a passing score takes the append path.
//   Each (query, slice) has a list of `cap` entries {score, id} in global memory; a list
//   that might overflow on the next row block is compacted to its kc best entries by
//   (score descending, id ascending) with an 8-bit radix select in LDS, and the
//   threshold rises to the kc-th best score.  Rows are visited in ascending id order
//   within a slice, so a later row that ties the threshold loses to every kept entry:
//   at the end each list holds exactly the slice's kc best rows by (score, id).
// Rerank (prec_stage2): one wave per query merges the slices' lists, selects the ncand
//   best by (score, id), evaluates their canonical FP32 distances (ssg_device.h) and
//   writes the k first in (distance, id) order, optionally with every candidate and its
//   approximate distance.
#include "ssg_device.h"
#include "ssg_kernels.h"

namespace ssg {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int PQ_TR = 128;          // DB rows per LDS tile
constexpr int PQ_J = PREC_CAP_MAX / 64;  // list entries per lane in a compaction
constexpr int PQ_JS = 32;           // merged entries per lane in the rerank stage
constexpr int PQ_NS = PREC_NCAND_MAX / 64;  // candidate slots per lane in the rerank stage

// Order-preserving map of floats to unsigned keys (larger float -> larger key).
__device__ __forceinline__ uint32_t score_key(float f)
{
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float key_score(uint32_t k)
{
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

// The want-th largest key among the active entries of the wave (want >= 1 and at most
// the number of active entries), by an 8-bit radix select with a 256-bin histogram in
// LDS (hist: 256 words owned by this wave).  *above = active entries with a larger key.
template <int J>
__device__ uint32_t kth_largest(const uint32_t (&key)[J], const bool (&act)[J], int want,
                                uint32_t* hist, int lane, int* above)
{
    uint32_t prefix = 0;
    int rem = want, abv = 0;
    for (int pass = 0; pass < 4; pass++) {
        const int sh = 24 - 8 * pass;
        reinterpret_cast<uint4*>(hist)[lane] = make_uint4(0u, 0u, 0u, 0u);
        wave_sync();
#pragma unroll
        for (int j = 0; j < J; j++) {
            const bool in = act[j] && (pass == 0 || (key[j] >> (sh + 8)) == prefix);
            if (in) atomicAdd(&hist[(key[j] >> sh) & 255u], 1u);
        }
        wave_sync();
        const uint4 h4 = reinterpret_cast<const uint4*>(hist)[lane];
        const uint32_t tot = h4.x + h4.y + h4.z + h4.w;
        uint32_t suf = tot;  // entries in bins 4*lane .. 255
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t v = __shfl_down(suf, o);
            if (lane + o < 64) suf += v;
        }
        const uint32_t s3 = suf - tot + h4.w, s2 = s3 + h4.z, s1 = s2 + h4.y;
        const uint64_t bl = __ballot(suf >= (uint32_t)rem);
        const int L = 63 - __builtin_clzll(bl);
        const int t3 = (int)rl_u(s3, L), t2 = (int)rl_u(s2, L), t1 = (int)rl_u(s1, L);
        const int t0 = (int)rl_u(suf, L), hw = (int)rl_u(h4.w, L);
        int d, sa;
        if (t3 >= rem) {
            d = 3;
            sa = t3 - hw;
        } else if (t2 >= rem) {
            d = 2;
            sa = t3;
        } else if (t1 >= rem) {
            d = 1;
            sa = t2;
        } else {
            d = 0;
            sa = t1;
        }
        (void)t0;
        rem -= sa;
        abv += sa;
        prefix = (prefix << 8) | (uint32_t)(4 * L + d);
        wave_sync();
    }
    *above = abv;
    return prefix;
}

// Marks the kc best active entries by (key descending, id ascending); returns the kc-th
// best key.  Requires more than kc active entries.
template <int J>
__device__ uint32_t select_best(const uint32_t (&key)[J], const uint32_t (&id)[J], const bool (&act)[J],
                                int kc, uint32_t* hist, int lane, bool (&keep)[J])
{
    int above = 0;
    const uint32_t T = kth_largest<J>(key, act, kc, hist, lane, &above);
    const int need = kc - above;  // entries with key == T to keep, >= 1
    int eq = 0;
    bool tie[J];
    uint32_t nid[J];
#pragma unroll
    for (int j = 0; j < J; j++) {
        tie[j] = act[j] && key[j] == T;
        nid[j] = ~id[j];
        eq += __popcll(__ballot(tie[j]));
    }
    uint32_t TI = 0;
    if (eq > need) {
        int a2 = 0;
        TI = kth_largest<J>(nid, tie, need, hist, lane, &a2);  // ids are distinct
    }
#pragma unroll
    for (int j = 0; j < J; j++) keep[j] = act[j] && (key[j] > T || (tie[j] && nid[j] >= TI));
    return T;
}

// Compaction of one (query, slice) list to its kc best entries; returns the new
// threshold score.  The whole wave calls it with the same list.
__device__ float compact_list(uint2* L, int* cntp, int kc, int cap, uint32_t* hist, int lane)
{
    const int c = min((int)rfl_u((uint32_t)*cntp), cap);
    uint32_t key[PQ_J], id[PQ_J];
    bool act[PQ_J], keep[PQ_J];
#pragma unroll
    for (int j = 0; j < PQ_J; j++) {
        const int e = lane + 64 * j;
        act[j] = e < c;
        const uint2 v = act[j] ? L[e] : make_uint2(0u, 0u);
        key[j] = score_key(__uint_as_float(v.x));
        id[j] = v.y;
    }
    const uint32_t T = select_best<PQ_J>(key, id, act, kc, hist, lane, keep);
    int base = 0;
#pragma unroll
    for (int j = 0; j < PQ_J; j++) {
        const uint64_t b = __ballot(keep[j]);
        if (keep[j]) {
            const int p = base + __popcll(b & ((1ull << lane) - 1ull));
            L[p] = make_uint2(__float_as_uint(key_score(key[j])), id[j]);
        }
        base += __popcll(b);
    }
    if (lane == 0) *cntp = kc;
    wave_sync();
    return key_score(T);
}

template <int KB>
__global__ __launch_bounds__(256, 2) void prec_stage1(PrecArgs a)
{
    constexpr int DPB = 32 * KB;             // BF16 row length
    constexpr int LSB = DPB * 2 + 16;        // LDS row stride in bytes (staggers banks)
    constexpr int NCT = 4;                   // 16-query column tiles per wave
    constexpr int U4 = DPB / 8;              // 16-byte units per row
    constexpr int PF = PQ_TR * U4 / 256;     // 16-byte units per thread per tile
    static_assert(PQ_TR * U4 % 256 == 0, "tile must split evenly over the block");
    __shared__ __attribute__((aligned(16))) uint8_t s_x[2][PQ_TR * LSB];
    __shared__ __attribute__((aligned(16))) float s_h[2][PQ_TR];
    __shared__ int s_cnt[PREC_QB];
    __shared__ __attribute__((aligned(16))) uint32_t s_hist[4][256];

    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int col = lane & 15, g = lane >> 4;
    const int64_t qt = blockIdx.x;
    const int sl = blockIdx.y;
    const int64_t r0 = (int64_t)sl * a.rows_per_slice;
    const int64_t r1 = min(a.n, r0 + a.rows_per_slice);
    const int kc = a.kc, cap = a.cap;
    const int64_t qb0 = qt * PREC_QB;
    auto list_of = [&](int ql) -> uint2* {
        return a.list + ((qb0 + ql) * a.nslices + sl) * (int64_t)cap;
    };

    s_cnt[tid] = 0;
    bf16x8 bq[NCT][KB];
#pragma unroll
    for (int ct = 0; ct < NCT; ct++) {
        const uint16_t* qp = a.Qe + (qb0 + 64 * w + 16 * ct + col) * DPB + g * KB * 8;
#pragma unroll
        for (int s = 0; s < KB; s++) bq[ct][s] = *reinterpret_cast<const bf16x8*>(qp + 8 * s);
    }
    float tau[NCT];
#pragma unroll
    for (int ct = 0; ct < NCT; ct++) tau[ct] = -__builtin_inff();

    if (r0 < r1) {
        uint4 pf0, pf1, pf2, pf3, pf4, pf5, pf6, pf7;  // the next tile (PF <= 8 units per thread)
        float pfh = 0.f;
        static_assert(PF <= 8, "prefetch registers");
        auto prefetch = [&](int64_t t) {
#define SSG_PF(I, R)                                                                        \
            if (I < PF) {                                                                   \
                const int idx = tid + 256 * I;                                              \
                const int row = idx / U4, c = idx % U4;                                     \
                const int64_t gr = min(t + row, r1 - 1); /* rows past the slice: see s_h */ \
                R = reinterpret_cast<const uint4*>(a.Xe + gr * DPB)[c];                     \
            }
            SSG_PF(0, pf0) SSG_PF(1, pf1) SSG_PF(2, pf2) SSG_PF(3, pf3)
            SSG_PF(4, pf4) SSG_PF(5, pf5) SSG_PF(6, pf6) SSG_PF(7, pf7)
#undef SSG_PF
            if (tid < PQ_TR) {
                const int64_t gr = t + tid;
                const float xn = a.xnorm[min(gr, r1 - 1)];
                pfh = gr < r1 ? (a.metric == M_NORMCOS ? 0.f : -0.5f * xn) : -__builtin_inff();
            }
        };
        auto stage = [&](int b) {
#define SSG_ST(I, R)                                                                        \
            if (I < PF) {                                                                   \
                const int idx = tid + 256 * I;                                              \
                const int row = idx / U4, c = idx % U4;                                     \
                *reinterpret_cast<uint4*>(&s_x[b][row * LSB + c * 16]) = R;                 \
            }
            SSG_ST(0, pf0) SSG_ST(1, pf1) SSG_ST(2, pf2) SSG_ST(3, pf3)
            SSG_ST(4, pf4) SSG_ST(5, pf5) SSG_ST(6, pf6) SSG_ST(7, pf7)
#undef SSG_ST
            if (tid < PQ_TR) s_h[b][tid] = pfh;
        };
        prefetch(r0);
        stage(0);
        __syncthreads();
        int b = 0;
        for (int64_t t0 = r0; t0 < r1; t0 += PQ_TR) {
            const bool more = t0 + PQ_TR < r1;
            if (more) prefetch(t0 + PQ_TR);
            const uint8_t* xb = s_x[b];
            const float* hb = s_h[b];
            // MFMA chains of one 16-row block for the wave's 4 query tiles
            auto chain = [&](int rb, floatx4 (&acc)[NCT]) {
                const floatx4 c0 = *reinterpret_cast<const floatx4*>(&hb[rb * 16 + 4 * g]);
                bf16x8 af[KB];
                const uint8_t* ap = xb + (rb * 16 + col) * LSB + g * KB * 16;
#pragma unroll
                for (int s = 0; s < KB; s++) af[s] = *reinterpret_cast<const bf16x8*>(ap + 16 * s);
#pragma unroll
                for (int ct = 0; ct < NCT; ct++) {
                    floatx4 c = c0;
#pragma unroll
                    for (int s = 0; s < KB; s++)
                        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[s], bq[ct][s], c, 0, 0, 0);
                    acc[ct] = c;
                }
            };
            floatx4 cur[NCT], nxt[NCT];
            chain(0, cur);
#pragma unroll
            for (int rb = 0; rb < PQ_TR / 16; rb++) {
                if (rb + 1 < PQ_TR / 16) chain(rb + 1, nxt);
                bool any = false;
#pragma unroll
                for (int ct = 0; ct < NCT; ct++) {
                    const float mx = fmaxf(fmaxf(cur[ct][0], cur[ct][1]), fmaxf(cur[ct][2], cur[ct][3]));
                    any = any || mx > tau[ct];
                }
                if (__ballot(any)) {
                    const uint32_t rowb = (uint32_t)(t0 + rb * 16 + 4 * g);
#pragma unroll
                    for (int ct = 0; ct < NCT; ct++) {
                        const int ql = 64 * w + 16 * ct + col;
#pragma unroll
                        for (int r = 0; r < 4; r++) {
                            if (cur[ct][r] > tau[ct]) {
                                const int p = atomicAdd(&s_cnt[ql], 1);
                                if (p < cap) list_of(ql)[p] = make_uint2(__float_as_uint(cur[ct][r]), rowb + r);
                            }
                        }
                    }
                    // A query gains at most 16 entries per row block, so a list with at
                    // most cap - 16 entries cannot overflow on the next one.
                    wave_sync();
                    const int cq = s_cnt[64 * w + lane];
                    uint64_t need = __ballot(cq > cap - 16);
                    while (need) {
                        const int qq = __builtin_ctzll(need);
                        need &= need - 1;
                        const float nt = compact_list(list_of(64 * w + qq), &s_cnt[64 * w + qq], kc, cap,
                                                      s_hist[w], lane);
#pragma unroll
                        for (int ct = 0; ct < NCT; ct++)
                            if (ct == (qq >> 4) && col == (qq & 15)) tau[ct] = nt;
                    }
                }
                if (rb + 1 < PQ_TR / 16) {
#pragma unroll
                    for (int ct = 0; ct < NCT; ct++) cur[ct] = nxt[ct];
                }
            }
            if (more) stage(b ^ 1);
            __syncthreads();
            b ^= 1;
        }
    }
    // final lists: exactly the slice's kc best (or all of its rows)
    wave_sync();
    {
        const int cq = s_cnt[64 * w + lane];
        uint64_t need = __ballot(cq > kc);
        while (need) {
            const int qq = __builtin_ctzll(need);
            need &= need - 1;
            (void)compact_list(list_of(64 * w + qq), &s_cnt[64 * w + qq], kc, cap, s_hist[w], lane);
        }
        const int cf = s_cnt[64 * w + lane];
        a.list_n[(qb0 + 64 * w + lane) * a.nslices + sl] = cf < kc ? cf : kc;
    }
}

// Rerank stage: merge, select ncand by (score, id), canonical distances, sort, write.
template <int V>
__global__ __launch_bounds__(256) void prec_stage2(PrecArgs a, PrecOut o)
{
    __shared__ __attribute__((aligned(16))) uint32_t s_hist[4][256];
    __shared__ uint32_t s_id[4][PREC_NCAND_MAX];
    __shared__ float s_sc[4][PREC_NCAND_MAX];
    const int lane = lane_id(), w = threadIdx.x >> 6;
    const int g = lane >> 4, j16 = lane & 15;
    const int64_t q = (int64_t)blockIdx.x * 4 + w;
    if (q >= o.m) return;
    const int dp = 16 * V;
    const int ns = a.nslices, kc = a.kc;

    uint32_t key[PQ_JS], cid[PQ_JS];
    bool act[PQ_JS], keep[PQ_JS];
    int total = 0;
#pragma unroll
    for (int j = 0; j < PQ_JS; j++) {
        const int e = lane + 64 * j;
        const int s = e / kc, pos = e - s * kc;
        act[j] = false;
        key[j] = 0u;
        cid[j] = 0u;
        if (s < ns && pos < a.list_n[q * ns + s]) {
            const uint2 v = a.list[(q * ns + s) * (int64_t)a.cap + pos];
            act[j] = true;
            key[j] = score_key(__uint_as_float(v.x));
            cid[j] = v.y;
        }
        total += __popcll(__ballot(act[j]));
    }
    const int nsel = total < o.ncand ? total : o.ncand;
    if (total > o.ncand) {
        (void)select_best<PQ_JS>(key, cid, act, o.ncand, s_hist[w], lane, keep);
    } else {
#pragma unroll
        for (int j = 0; j < PQ_JS; j++) keep[j] = act[j];
    }
    int base = 0;
#pragma unroll
    for (int j = 0; j < PQ_JS; j++) {
        const uint64_t bm = __ballot(keep[j]);
        if (keep[j]) {
            const int p = base + __popcll(bm & ((1ull << lane) - 1ull));
            s_id[w][p] = cid[j];
            s_sc[w][p] = key_score(key[j]);
        }
        base += __popcll(bm);
    }
    wave_sync();

    // canonical distances: candidate u = 64*b + lane lives in slot b of its lane
    float qv[V];
    {
        const float* qp = o.qpad + q * dp + j16 * V;
#pragma unroll
        for (int t = 0; t < V; t++) qv[t] = qp[t];
    }
    float myd[PQ_NS];
    uint32_t myid[PQ_NS];
    float mysc[PQ_NS];
#pragma unroll
    for (int b = 0; b < PQ_NS; b++) {
        myd[b] = __builtin_inff();
        myid[b] = 0xffffffffu;
        mysc[b] = -__builtin_inff();
        if (64 * b < nsel) {
            const int hi = min(nsel, 64 * b + 64);
            for (int u0 = 64 * b; u0 < hi; u0 += 4) {
                const int u = u0 + g;
                float xv[V];
                const uint32_t rid = s_id[w][u < hi ? u : u0];
                load_run<V>(a.X + (size_t)rid * dp + j16 * V, xv);
                const double s = group16_sum(run_partial<V>(a.metric, xv, qv));
                const float d = finish_dist(a.metric, s);
#pragma unroll
                for (int gg = 0; gg < 4; gg++) {
                    const float dt = __shfl(d, gg * 16);
                    if (lane == ((u0 + gg) & 63) && u0 + gg < hi) myd[b] = dt;
                }
            }
            if (64 * b + lane < nsel) {
                myid[b] = s_id[w][64 * b + lane];
                mysc[b] = s_sc[w][64 * b + lane];
            }
        }
    }
    // rank by (distance, id)
    int rank[PQ_NS];
#pragma unroll
    for (int b = 0; b < PQ_NS; b++) rank[b] = 0;
#pragma unroll
    for (int bb = 0; bb < PQ_NS; bb++) {
        if (64 * bb >= nsel) break;
        const int hi = min(64, nsel - 64 * bb);
        for (int t = 0; t < hi; t++) {
            const float dt = rl_f(myd[bb], t);
            const uint32_t it = rl_u(myid[bb], t);
#pragma unroll
            for (int b = 0; b < PQ_NS; b++)
                rank[b] += (dt < myd[b] || (dt == myd[b] && it < myid[b])) ? 1 : 0;
        }
    }
    const float qn = o.qnorm ? o.qnorm[q] : 0.f;
#pragma unroll
    for (int b = 0; b < PQ_NS; b++) {
        const int e = 64 * b + lane;
        if (e < nsel) {
            if (rank[b] < o.k) {
                o.ids[q * o.kstride + rank[b]] = myid[b] + 1u;
                o.dists[q * o.kstride + rank[b]] = myd[b];
            }
            if (o.cand_ids) {
                float ad;
                if (a.metric == M_NORMCOS) ad = 1.0f - mysc[b];
                else {
                    ad = qn - 2.0f * mysc[b];
                    if (a.metric == M_L2) ad = __builtin_sqrtf(fmaxf(ad, 0.f));
                }
                o.cand_ids[q * o.ncand + rank[b]] = myid[b] + 1u;
                o.cand_dists[q * o.ncand + rank[b]] = ad;
            }
        } else {
            if (e < o.k) {
                o.ids[q * o.kstride + e] = 0u;
                o.dists[q * o.kstride + e] = __builtin_inff();
            }
            if (o.cand_ids && e < o.ncand) {
                o.cand_ids[q * o.ncand + e] = 0u;
                o.cand_dists[q * o.ncand + e] = __builtin_inff();
            }
        }
    }
}

// Rows of dp floats -> rows of dpb BF16 (round to nearest even), zero padded.
__global__ void rows_to_bf16_kernel(const float* __restrict__ src, int64_t rows, int32_t dp,
                                    uint16_t* __restrict__ dst, int32_t dpb)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t r = i / dpb, c = i % dpb;
    if (r >= rows) return;
    uint16_t v = 0;
    if (c < dp) {
        const uint32_t u = __float_as_uint(src[r * dp + c]);
        v = (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
    }
    dst[i] = v;
}

hipError_t launch_rows_to_bf16(const float* src, int64_t rows, int32_t dp, uint16_t* dst, int32_t dpb,
                               hipStream_t s)
{
    const int64_t tot = rows * dpb;
    if (tot <= 0) return hipSuccess;
    hipLaunchKernelGGL(rows_to_bf16_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, src,
                       rows, dp, dst, dpb);
    return hipGetLastError();
}

// Slices per query tile: the fewest that fill the chip and leave the last round of
// resident blocks at least 90% full (each slice restarts its threshold, so more slices
// cost more appends).
int32_t prec_slices(int64_t n, int64_t m, int num_cu, int kc)
{
    const int64_t qtiles = (m + PREC_QB - 1) / PREC_QB;
    const int64_t cap = (int64_t)num_cu * 2;
    int64_t smax = (n + 4 * PQ_TR - 1) / (4 * PQ_TR);
    if (smax > (int64_t)PQ_JS * 64 / kc) smax = (int64_t)PQ_JS * 64 / kc;
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
        if (fill >= 0.9) return (int32_t)s;
        if (fill > best_fill + 1e-9) {
            best_fill = fill;
            best = s;
        }
    }
    return (int32_t)best;
}

hipError_t prec_launch_stage1(const PrecArgs& a, hipStream_t s)
{
    const dim3 grid((unsigned)(a.m_pad / PREC_QB), (unsigned)a.nslices), block(256);
    switch (a.dpb / 32) {
    case 1: hipLaunchKernelGGL(prec_stage1<1>, grid, block, 0, s, a); break;
    case 2: hipLaunchKernelGGL(prec_stage1<2>, grid, block, 0, s, a); break;
    case 3: hipLaunchKernelGGL(prec_stage1<3>, grid, block, 0, s, a); break;
    case 4: hipLaunchKernelGGL(prec_stage1<4>, grid, block, 0, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t prec_launch_stage2(const PrecArgs& a, const PrecOut& o, int32_t dp, hipStream_t s)
{
    const dim3 grid((unsigned)((o.m + 3) / 4)), block(256);
    switch (dp / 16) {
#define SSG_P2(VV) \
    case VV: hipLaunchKernelGGL((prec_stage2<VV>), grid, block, 0, s, a, o); break;
        SSG_P2(1) SSG_P2(2) SSG_P2(3) SSG_P2(4) SSG_P2(5) SSG_P2(6) SSG_P2(7) SSG_P2(8)
#undef SSG_P2
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace ssg
