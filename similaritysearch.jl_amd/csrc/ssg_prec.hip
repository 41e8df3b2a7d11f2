// ssg_prec.hip -- precision form of exhaustive k-NN and allknn for gfx950: a candidate
// pass with BF16 or FP8 operands on the matrix cores, then rerank! of the candidates with
// the canonical FP32 distance.
//
// The reference computes every distance in Float32 (src/exact/sequential-exhaustive.jl:68-80,
// src/allknn.jl:66-110); the precision form is the allknn-with-rerank pattern of
// src/rerank.jl:22-98 with the candidate queue filled from low-precision dot products.
//
// Operands: BF16 (round to nearest even) for v_mfma_f32_16x16x32_bf16, or OCP FP8 e4m3fn
//   of the values scaled by a power of two 2^e (e from the largest |value| of the rows,
//   and of the queries, so that it maps to at most 448; round to nearest even) for
//   v_mfma_scale_f32_16x16x128_f8f6f4, whose E8M0 scale operands multiply the product by
//   2^-e of each side exactly.  FP8 rows are padded to 128 values.
//
// Candidate pass (prec_stage1): block = 512 queries x one DB slice, 8 waves of 64
//   queries (one block per CU: each row staged in LDS serves 512 queries).  A wave keeps
//   its queries' encoded rows as the B operands of the matrix-core product in registers
//   for the whole slice; DB rows are staged in LDS in tiles of 512 (FP8) or 256 (BF16)
//   rows (double-buffered, one barrier per tile, the next tile prefetched into
//   registers while the current one is multiplied) and shared by the 8 waves.  Within a
//   tile the products of one 16-row block run while the previous block's scores are
//   tested, and the operand pieces of the block after are read from LDS meanwhile.  The accumulator starts at -||x||^2/2 (SqL2, L2) or 0
//   (NormCosine), so the MFMA result is the score s = x.q - ||x||^2/2 (or x.q): larger
//   is nearer, and an invalid row starts at -Inf.  A lane holds 4 rows x 1 query per 16x16 tile; a tile
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
#include "ssg_select.h"

namespace ssg {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef int intx4 __attribute__((ext_vector_type(4)));
typedef int intx8 __attribute__((ext_vector_type(8)));

constexpr int PQ_TR = 128;          // DB rows per LDS tile
constexpr int PQ_JS = 32;           // merged entries per lane in the rerank stage
constexpr int PQ_NS = PREC_NCAND_MAX / 64;  // candidate slots per lane in the rerank stage

// Operand format of the candidate pass.  A lane's share of one matrix-core product is a
// piece of PB bytes of its row: BF16, 8 values for v_mfma_f32_16x16x32_bf16 (lane l: row
// l % 16, the k group l / 16); FP8 e4m3fn, 32 values for v_mfma_scale_f32_16x16x128_f8f6f4
// (same lane mapping, 32 k per group).  A row of KB products is split into 4 contiguous
// quarters, one per k group, so a lane reads KB pieces that lie next to each other; the
// same split of the query rows keeps the k order of both operands the same.
template <bool F8> struct PrecOp {
    typedef bf16x8 T;
    static constexpr int PB = 16;
    static __device__ inline T ld(const uint8_t* p) { return *reinterpret_cast<const bf16x8*>(p); }
    static __device__ inline floatx4 mma(T x, T q, floatx4 c, int, int)
    {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(x, q, c, 0, 0, 0);
    }
};
template <> struct PrecOp<true> {
    typedef intx8 T;
    static constexpr int PB = 32;
    // two 16-byte reads: rows are 16-byte aligned
    static __device__ inline T ld(const uint8_t* p)
    {
        const intx4 lo = *reinterpret_cast<const intx4*>(p), hi = *reinterpret_cast<const intx4*>(p + 16);
        return __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
    }
    // The E8M0 scale operands (127 - e) undo the power-of-two scaling of each operand
    // exactly, so the product and the -|x|^2/2 start value are in the same units.
    static __device__ inline floatx4 mma(T x, T q, floatx4 c, int sx, int sq)
    {
        return __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(x, q, c, 0, 0, 0, sx, 0, sq);
    }
};

// J = cap / 64 list entries per lane in a compaction (2: lists of 128 entries, 8: of 512,
// 16: of 1024).  Each kernel calls one compaction instance only: the registers that
// instance uses are what the kernel has to keep clear around the call.  F8 selects the
// FP8 form.
template <int KB, int J, bool F8>
__global__ __launch_bounds__(PREC_QB, 8 / PREC_W) void prec_stage1(PrecArgs a)
{
    typedef PrecOp<F8> Op;
    typedef typename Op::T OpT;
    constexpr int PB = Op::PB;
    constexpr int RB = 4 * PB * KB;          // encoded row bytes
    // FP8 rows are 128 bytes, stored unpadded with the 8 16-byte chunks of row r at
    // positions c ^ (r & 7), so that the 8 rows a quarter-wave reads, or the chunks a row
    // is written in, fall in distinct banks.  BF16 rows are padded by 16 bytes.
    constexpr bool SWZ = F8;
    // DB rows per LDS tile, as many as LDS and the prefetch registers allow: FP8 512 (256
    // for 1024-entry lists), BF16 256 (128 for 1024-entry lists of rows of 3-4 pieces)
    constexpr int TR = (F8 && J < 16) ? 4 * PQ_TR : (F8 || J < 16 || KB < 3) ? 2 * PQ_TR : PQ_TR;
    constexpr int LSB = SWZ ? RB : RB + 16;  // LDS row stride in bytes
    constexpr int NCT = 4;                   // 16-query column tiles per wave
    constexpr int U4 = RB / 16;              // 16-byte units per row
    constexpr int PF = TR * U4 / PREC_QB;    // 16-byte units per thread per tile
    static_assert(TR * U4 % PREC_QB == 0, "tile must split evenly over the block");
    static_assert(!SWZ || (U4 == 8 && KB == 1), "the swizzle covers one 128-byte row");
    __shared__ __attribute__((aligned(16))) uint8_t s_x[2][TR * LSB];
    __shared__ __attribute__((aligned(16))) float s_h[2][TR];
    __shared__ __attribute__((aligned(16))) uint32_t s_hist[PREC_W][256];
    __shared__ int s_cnt[PREC_W][NCT * 64];   // sub-list counts of each wave's queries
    __shared__ float s_tau[PREC_W][NCT * 16]; // their thresholds

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

    OpT bq[NCT][KB];
#pragma unroll
    for (int ct = 0; ct < NCT; ct++) {
        const uint8_t* qp = a.Qe + (qb0 + 64 * w + 16 * ct + col) * RB + g * KB * PB;
#pragma unroll
        for (int s = 0; s < KB; s++) bq[ct][s] = Op::ld(qp + PB * s);
    }
    int sx = 127, sq = 127;  // FP8: E8M0 scale operands of the rows and of the queries
    if (F8) {
        sx = 127 - fp8_scale_exp(a.f8max[0]);
        sq = 127 - fp8_scale_exp(a.f8max[1]);
    }
    // Appends go through a buffer resource over the block's lists (stores past it are
    // dropped by the hardware).  Each lane keeps, per query, the byte offset of the next
    // free entry of its sub-list; sub-lists are sc = cap / 4 entries, sc a power of two,
    // so the fill level is the offset's low bits.
    const int sc = cap / 4;
    const uint32_t smask = (uint32_t)sc * 8u - 1u;
    const uint32_t full_at = (uint32_t)(sc - 5) * 8u;  // fill level that needs a compaction
    const int64_t qstride = (int64_t)a.nslices * cap;  // entries between consecutive queries
    const __amdgpu_buffer_rsrc_t lrs = __builtin_amdgcn_make_buffer_rsrc(
        a.list + (qb0 * a.nslices + sl) * (int64_t)cap, 0,
        (int)(((PREC_QB - 1) * qstride + cap) * (int64_t)sizeof(uint2)), 0x00020000);
    float tau[NCT];
    uint32_t wp[NCT];
#pragma unroll
    for (int ct = 0; ct < NCT; ct++) {
        // starting threshold: the main pass's sampled one, or a sample pass's seed
        const float* ts = a.sample ? a.tau_seed : a.tau0;
        tau[ct] = ts ? ts[(qb0 + 64 * w + 16 * ct + col) * a.nslices + sl] : -__builtin_inff();
        wp[ct] = (uint32_t)(((64 * w + 16 * ct + col) * qstride + g * sc) * (int64_t)sizeof(uint2));
    }

    if (r0 < r1) {
        uint4 pf0, pf1, pf2, pf3, pf4, pf5, pf6, pf7;  // the next tile (PF <= 8 units per thread)
        float pfx = 0.f;                                // the norm of its row tid % TR
        int64_t pft = r0;                               // its first row
        static_assert(PF <= 8 && TR <= PREC_QB, "prefetch registers; one row norm per thread");
        // Every load is unconditional (clamped indices) and nothing consumes it before
        // stage(): a conditional load consumed in its branch would make the wave wait for
        // all of its loads in flight, the tile prefetch included.
        auto prefetch = [&](int64_t t) {
            pft = t;
            pfx = a.xnorm[min(t + (tid & (TR - 1)), r1 - 1)];
#define SSG_PF(I, R)                                                                        \
            if (I < PF) {                                                                   \
                const int idx = tid + PREC_QB * I;                                          \
                const int row = idx / U4, c = idx % U4;                                     \
                const int64_t gr = min(t + row, r1 - 1); /* rows past the slice: see s_h */ \
                R = reinterpret_cast<const uint4*>(a.Xe + gr * RB)[c];                      \
            }
            SSG_PF(0, pf0) SSG_PF(1, pf1) SSG_PF(2, pf2) SSG_PF(3, pf3)
            SSG_PF(4, pf4) SSG_PF(5, pf5) SSG_PF(6, pf6) SSG_PF(7, pf7)
#undef SSG_PF
        };
        auto stage = [&](int b) {
#define SSG_ST(I, R)                                                                        \
            if (I < PF) {                                                                   \
                const int idx = tid + PREC_QB * I;                                          \
                const int row = idx / U4, c = idx % U4;                                     \
                const int pc = SWZ ? (c ^ (row & 7)) : c;                                   \
                *reinterpret_cast<uint4*>(&s_x[b][row * LSB + pc * 16]) = R;                \
            }
            SSG_ST(0, pf0) SSG_ST(1, pf1) SSG_ST(2, pf2) SSG_ST(3, pf3)
            SSG_ST(4, pf4) SSG_ST(5, pf5) SSG_ST(6, pf6) SSG_ST(7, pf7)
#undef SSG_ST
            if (tid < TR)
                s_h[b][tid] = pft + tid < r1 ? (a.metric == M_NORMCOS ? 0.f : -0.5f * pfx) : -__builtin_inff();
        };
        prefetch(r0);
        stage(0);
        __syncthreads();
        int b = 0;
        const int64_t step = (int64_t)TR * a.tile_step;
        for (int64_t t0 = r0; t0 < r1; t0 += step) {
            const bool more = t0 + step < r1;
            if (more) prefetch(t0 + step);
            const uint8_t* xb = s_x[b];
            const float* hb = s_h[b];
            // One 16-row block: the lane's operand pieces and start values (LDS reads),
            // then the MFMA chains for the wave's 4 query tiles.  The products of block
            // rb + 1 run while the scores of block rb are tested, and (up to 3 pieces per
            // lane) the pieces of block rb + 2 are read meanwhile, so the LDS latency is
            // behind matrix-core work.
            struct Frag {
                OpT af[KB];
                floatx4 c0;
            };
            auto frag = [&](int rb, Frag& f) {
                f.c0 = *reinterpret_cast<const floatx4*>(&hb[rb * 16 + 4 * g]);
                const int r = rb * 16 + col;
                if constexpr (SWZ) {
                    // the lane's 32 bytes are chunks 2g and 2g + 1 of its row
                    const uint8_t* rp = xb + r * LSB;
                    const intx4 lo = *reinterpret_cast<const intx4*>(rp + ((2 * g) ^ (r & 7)) * 16);
                    const intx4 hi = *reinterpret_cast<const intx4*>(rp + ((2 * g + 1) ^ (r & 7)) * 16);
                    f.af[0] = __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
                } else {
                    const uint8_t* ap = xb + r * LSB + g * KB * PB;
#pragma unroll
                    for (int s = 0; s < KB; s++) f.af[s] = Op::ld(ap + PB * s);
                }
            };
            auto chain = [&](const Frag& f, floatx4 (&acc)[NCT]) {
#pragma unroll
                for (int ct = 0; ct < NCT; ct++) {
                    floatx4 c = f.c0;
#pragma unroll
                    for (int s = 0; s < KB; s++) c = Op::mma(f.af[s], bq[ct][s], c, sx, sq);
                    acc[ct] = c;
                }
            };
            constexpr int NRB = TR / 16;
            constexpr bool LA = KB < 4;  // read ahead (the 4-piece form has no registers for it)
            floatx4 cur[NCT], nxt[NCT];
            Frag fr;
            frag(0, fr);
            chain(fr, cur);
            if (LA) frag(1, fr);
#pragma unroll
            for (int rb = 0; rb < NRB; rb++) {
                if (rb + 1 < NRB) {
                    if (!LA) frag(rb + 1, fr);
                    chain(fr, nxt);
                    if (LA && rb + 2 < NRB) frag(rb + 2, fr);
                }
                // Each 16-query tile is tested on its own (the maximum of its 4 x 16 scores
                // against each query's threshold; 256 scores per decision).  The fast path
                // is one test of all 4 tiles; the append path of a tile runs only when one
                // of its scores passes.
                float mx[NCT];
                bool anyp = false;
#pragma unroll
                for (int ct = 0; ct < NCT; ct++) {
                    mx[ct] = __builtin_elementwise_maximum(
                        __builtin_elementwise_maximum(cur[ct][0], cur[ct][1]),
                        __builtin_elementwise_maximum(cur[ct][2], cur[ct][3]));
                    anyp |= mx[ct] > tau[ct];
                }
                if (__builtin_expect(__ballot(anyp) != 0, 0)) {
                    const uint32_t rowb = (uint32_t)(t0 + rb * 16 + 4 * g);
                    uint64_t qfull = 0;  // bit 16 * ct + col: that query's list is compacted now
#pragma unroll
                    for (int ct = 0; ct < NCT; ct++) {
                        if (__ballot(mx[ct] > tau[ct]) != 0) {
                            const float t = tau[ct];
                            uint32_t o = wp[ct];
                            // No branches: every row's entry is stored, a row that does not
                            // pass at an offset past the buffer's extent (2^31), which the
                            // hardware drops, and only passing rows advance the offset.
#pragma unroll
                            for (int r = 0; r < 4; r++) {
                                const bool p = cur[ct][r] > t;
                                const u32x2 e = {__float_as_uint(cur[ct][r]), rowb + r};
                                __builtin_amdgcn_raw_buffer_store_b64(e, lrs, (int)(p ? o : 0x80000000u), 0, 0);
                                o += p ? 8u : 0u;
                            }
                            wp[ct] = o;
                            // A lane adds at most 4 entries per query and row block, so a
                            // sub-list with at most sc - 5 entries cannot fill up on the next
                            // (and the offset's low bits never wrap to 0).
                            const uint64_t f = __ballot((o & smask) > full_at);
                            qfull |= ((f | (f >> 16) | (f >> 32) | (f >> 48)) & 0xffffull) << (16 * ct);
                        }
                    }
                    if (__builtin_expect(qfull != 0, 0)) {
                        // counts and thresholds go through LDS so that the loop indexes no
                        // register array
#pragma unroll
                        for (int ct = 0; ct < NCT; ct++) {
                            s_cnt[w][64 * ct + lane] = (int)((wp[ct] & smask) >> 3);
                            if (g == 0) s_tau[w][16 * ct + col] = tau[ct];
                        }
                        wave_sync();  // also makes this wave's appends visible to the reads
                        while (qfull) {
                            const int ql = __builtin_ctzll(qfull);
                            qfull &= qfull - 1;
                            int* cq = &s_cnt[w][64 * (ql >> 4) + (ql & 15)];
                            int kept = 0;
                            const float nt = compact_list_j<J>(list_of(64 * w + ql), sc, cq[0], cq[16], cq[32],
                                                               cq[48], kc, false, s_hist[w], lane, &kept);
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
                    // The next block's products are an operand of this path too (at its
                    // end, when they are done), so that the compiler issues them before the
                    // branch instead of sinking them next to their test, which would wait
                    // for each product.
                    if (rb + 1 < NRB) {
#pragma unroll
                        for (int ct = 0; ct < NCT; ct++) asm volatile("" : "+v"(nxt[ct]));
                    }
                }
                if (rb + 1 < NRB) {
#pragma unroll
                    for (int ct = 0; ct < NCT; ct++) cur[ct] = nxt[ct];
                }
            }
            if (more) stage(b ^ 1);
            __syncthreads();
            b ^= 1;
        }
    }
    // final lists: exactly the slice's kc best (or all of its rows), contiguous
    wave_sync();
    // the sub-list counts of the wave's 64 queries go through LDS so that the loop below
    // indexes no register array
#pragma unroll
    for (int ct = 0; ct < NCT; ct++) s_cnt[w][64 * ct + lane] = (int)((wp[ct] & smask) >> 3);
    lds_order();
    for (int ql = 0; ql < 64; ql++) {
        const int ct = ql >> 4, qc = ql & 15;
        const int* cq = &s_cnt[w][64 * ct + qc];
        int kept = 0;
        const float tk = compact_list_j<J>(list_of(64 * w + ql), sc, cq[0], cq[16], cq[32], cq[48], kc, true,
                                      s_hist[w], lane, &kept);
        if (lane == 0) {
            a.list_n[(qb0 + 64 * w + ql) * a.nslices + sl] = kept;
            // sample pass: the kc-th best sampled score (-Inf with at most kc sampled rows)
            if (a.sample) a.tau0[(qb0 + 64 * w + ql) * a.nslices + sl] = tk;
        }
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
    const int64_t qo = o.qmap ? (int64_t)o.qmap[q] : q;  // row of qpad, qnorm and the outputs
    const int dp = 16 * V;
    const int ns = a.nslices, kc = a.kc;

    uint32_t key[PQ_JS], cid[PQ_JS];
    bool act[PQ_JS], keep[PQ_JS];
    int total = 0;
    // Every load is unconditional (clamped slice) and only selected afterwards: a load
    // inside a branch, or one that waited for its slice's count, would serialise the
    // wave's round trips.  Entries past a slice's count are read but not used.
#pragma unroll
    for (int j = 0; j < PQ_JS; j++) {
        const int e = lane + 64 * j;
        const int s = e / kc, pos = e - s * kc;
        const int sc = s < ns ? s : 0;
        const int cnt = a.list_n[q * ns + sc];
        const uint2 v = a.list[(q * ns + sc) * (int64_t)a.cap + pos];
        act[j] = s < ns && pos < cnt;
        key[j] = act[j] ? score_key(__uint_as_float(v.x)) : 0u;
        cid[j] = act[j] ? v.y : 0u;
        total += __popcll(__ballot(act[j]));
    }
    const int nsel = total < o.ncand ? total : o.ncand;
    uint32_t kth = 0u;  // key of the ncand-th best candidate (0 with fewer)
    if (total > o.ncand) {
        uint32_t kand, diff;
        key_bits<PQ_JS>(key, act, &kand, &diff);
        kth = select_best<PQ_JS>(key, cid, act, o.ncand, s_hist[w], lane, keep, kand, diff);
    } else {
        uint32_t mn = 0xffffffffu;
#pragma unroll
        for (int j = 0; j < PQ_JS; j++) {
            keep[j] = act[j];
            if (act[j]) mn = min(mn, key[j]);
        }
        for (int off = 32; off >= 1; off >>= 1) mn = min(mn, (uint32_t)__shfl_xor((int)mn, off));
        if (total == o.ncand) kth = mn;
    }
    // Certificate of a pass that started from sampled thresholds: every row it dropped
    // scored at most tau0 of its slice, so the candidates are exact if the ncand-th best
    // of them scores above every slice's tau0.
    if (a.tau0 && o.flags) {
        float tmax = -__builtin_inff();
        for (int s2 = 0; s2 < ns; s2++) tmax = fmaxf(tmax, a.tau0[q * ns + s2]);
        const bool ok = tmax == -__builtin_inff() || (total >= o.ncand && key_score(kth) > tmax);
        if (!ok && lane == 0) {
            const unsigned int p = atomicAdd(&o.flags[0], 1u);
            o.flags[1 + p] = (unsigned int)q;
        }
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
        const float* qp = o.qpad + qo * dp + j16 * V;
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
            // 4 * RR candidates per round, RR per 16-lane group, all RR rows loaded before
            // any is used (8 per group measured no faster than 4)
            constexpr int RR = 4;
            for (int u0 = 64 * b; u0 < hi; u0 += 4 * RR) {
                float xv[RR][V];
#pragma unroll
                for (int r = 0; r < RR; r++) {
                    const int u = u0 + 4 * r + g;
                    const uint32_t rid = s_id[w][u < hi ? u : u0];
                    load_run<V>(a.X + (size_t)rid * dp + j16 * V, xv[r]);
                }
#pragma unroll
                for (int r = 0; r < RR; r++) {
                    const double s = group16_sum(run_partial<V>(a.metric, xv[r], qv));
                    const float d = finish_dist(a.metric, s);
#pragma unroll
                    for (int gg = 0; gg < 4; gg++) {
                        const float dt = __shfl(d, gg * 16);
                        const int uu = u0 + 4 * r + gg;
                        if (lane == (uu & 63) && uu < hi) myd[b] = dt;
                    }
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
    const float qn = o.qnorm ? o.qnorm[qo] : 0.f;
#pragma unroll
    for (int b = 0; b < PQ_NS; b++) {
        const int e = 64 * b + lane;
        if (e < nsel) {
            if (rank[b] < o.k) {
                o.ids[qo * o.kstride + rank[b]] = myid[b] + 1u;
                o.dists[qo * o.kstride + rank[b]] = myd[b];
            }
            if (o.cand_ids) {
                float ad;
                if (a.metric == M_NORMCOS) ad = 1.0f - mysc[b];
                else {
                    ad = qn - 2.0f * mysc[b];
                    if (a.metric == M_L2) ad = __builtin_sqrtf(fmaxf(ad, 0.f));
                }
                o.cand_ids[qo * o.ncand + rank[b]] = myid[b] + 1u;
                o.cand_dists[qo * o.ncand + rank[b]] = ad;
            }
        } else {
            if (e < o.k) {
                o.ids[qo * o.kstride + e] = 0u;
                o.dists[qo * o.kstride + e] = __builtin_inff();
            }
            if (o.cand_ids && e < o.ncand) {
                o.cand_ids[qo * o.ncand + e] = 0u;
                o.cand_dists[qo * o.ncand + e] = __builtin_inff();
            }
        }
    }
}

// dst row r = src row idx[r] for r < *count (rows of `width` 16-bit words); rows from *count
// to max_rows are zero.
__global__ void gather_rows_u16_kernel(const uint16_t* __restrict__ src, const uint32_t* __restrict__ idx,
                                       const unsigned int* __restrict__ count, int64_t max_rows,
                                       int32_t width, uint16_t* __restrict__ dst)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t r = i / width, c = i % width;
    if (r >= max_rows) return;
    dst[i] = r < (int64_t)*count ? src[(int64_t)idx[r] * width + c] : (uint16_t)0;
}

hipError_t launch_gather_rows_u16(const uint16_t* src, const uint32_t* idx, const unsigned int* count,
                                  int64_t max_rows, int32_t width, uint16_t* dst, hipStream_t s)
{
    const int64_t tot = max_rows * width;
    if (tot <= 0) return hipSuccess;
    hipLaunchKernelGGL(gather_rows_u16_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, src, idx,
                       count, max_rows, width, dst);
    return hipGetLastError();
}

// Redo helpers: out[i] = the largest of src[idx[i] * ns .. + ns), one float step lower
// (i < cnt); out[i * nsf + s]
// = v[i] for i < cnt, -Inf for cnt <= i < rows; out[i] = a[b[i]] (i < cnt).
__global__ void max_slices_kernel(const float* __restrict__ src, int32_t ns, const uint32_t* __restrict__ idx,
                                  int64_t cnt, float* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cnt) return;
    float t = -__builtin_inff();
    for (int s = 0; s < ns; s++) t = fmaxf(t, src[(int64_t)idx[i] * ns + s]);
    // one step down, so that rows scoring exactly the seed (ties, frequent among FP8
    // scores) still pass
    if (t > -__builtin_inff()) {
        const uint32_t u = __float_as_uint(t);
        t = t > 0.f ? __uint_as_float(u - 1u) : t < 0.f ? __uint_as_float(u + 1u) : -__uint_as_float(1u);
    }
    out[i] = t;
}

__global__ void spread_kernel(const float* __restrict__ v, int64_t cnt, int64_t rows, int32_t nsf,
                              float* __restrict__ out)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= rows * nsf) return;
    const int64_t i = e / nsf;
    out[e] = i < cnt ? v[i] : -__builtin_inff();
}

__global__ void compose_kernel(const uint32_t* __restrict__ a, const uint32_t* __restrict__ b, int64_t cnt,
                               uint32_t* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cnt) out[i] = a[b[i]];
}

hipError_t launch_max_slices_f32(const float* src, int32_t ns, const uint32_t* idx, int64_t cnt, float* out,
                                 hipStream_t s)
{
    if (cnt <= 0) return hipSuccess;
    hipLaunchKernelGGL(max_slices_kernel, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s, src, ns, idx, cnt,
                       out);
    return hipGetLastError();
}

hipError_t launch_spread_f32(const float* v, int64_t cnt, int64_t rows, int32_t nsf, float* out, hipStream_t s)
{
    const int64_t tot = rows * nsf;
    if (tot <= 0) return hipSuccess;
    hipLaunchKernelGGL(spread_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, v, cnt, rows, nsf, out);
    return hipGetLastError();
}

hipError_t launch_compose_u32(const uint32_t* a, const uint32_t* b, int64_t cnt, uint32_t* out, hipStream_t s)
{
    if (cnt <= 0) return hipSuccess;
    hipLaunchKernelGGL(compose_kernel, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s, a, b, cnt, out);
    return hipGetLastError();
}

// The rows of every step-th 128-row tile (tiles 0, step, 2 step, ...) of the encoded rows
// (width 16-bit words each) and their norms, packed: sample row i is row
// (i / 128) * 128 * step + i % 128.  Rows at or past n are not copied (the caller passes
// the number of sample rows below n).
__global__ void sample_rows_kernel(const uint16_t* __restrict__ Xe, const float* __restrict__ xnorm,
                                   int64_t ns_rows, int32_t step, int32_t width, uint16_t* __restrict__ Xs,
                                   float* __restrict__ xns)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t r = i / width, c = i % width;
    if (r >= ns_rows) return;
    const int64_t orig = (r / 128) * 128 * step + r % 128;
    Xs[i] = Xe[orig * width + c];
    if (c == 0) xns[r] = xnorm[orig];
}

hipError_t launch_sample_rows(const uint16_t* Xe, const float* xnorm, int64_t ns_rows, int32_t step,
                              int32_t width, uint16_t* Xs, float* xns, hipStream_t s)
{
    const int64_t tot = ns_rows * width;
    if (tot <= 0) return hipSuccess;
    hipLaunchKernelGGL(sample_rows_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, Xe, xnorm,
                       ns_rows, step, width, Xs, xns);
    return hipGetLastError();
}

// max |src| over count values, as float bits (non-negative floats order as their bits),
// folded into amax[0 .. namax) with atomic maxima.
__global__ void absmax_kernel(const float* __restrict__ src, int64_t count, uint32_t* __restrict__ amax,
                              int32_t namax)
{
    uint32_t mx = 0u;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (int64_t)gridDim.x * blockDim.x)
        mx = max(mx, __float_as_uint(src[i]) & 0x7fffffffu);
    for (int off = 32; off >= 1; off >>= 1) mx = max(mx, (uint32_t)__shfl_xor((int)mx, off));
    if ((threadIdx.x & 63) == 0 && mx != 0u)
        for (int t = 0; t < namax; t++) atomicMax(&amax[t], mx);
}

hipError_t launch_absmax(const float* src, int64_t count, uint32_t* amax, int32_t namax, hipStream_t s)
{
    if (count <= 0) return hipSuccess;
    const int64_t blocks = std::min<int64_t>((count + 255) / 256, 4096);
    hipLaunchKernelGGL(absmax_kernel, dim3((unsigned)blocks), dim3(256), 0, s, src, count, amax, namax);
    return hipGetLastError();
}

// Rows of dp floats -> rows of rowb bytes of e4m3fn(v * 2^e), e = fp8_scale_exp(*amax),
// zero padded.  v_cvt_pk_fp8_f32 rounds to nearest even; |v * 2^e| <= 448 by the choice
// of e, so nothing saturates.  One thread per 4 bytes.
__global__ void rows_to_fp8_kernel(const float* __restrict__ src, int64_t rows, int32_t dp,
                                   const uint32_t* __restrict__ amax, uint8_t* __restrict__ dst, int32_t rowb)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int32_t words = rowb / 4;
    const int64_t r = i / words;
    const int32_t c = 4 * (int32_t)(i % words);
    if (r >= rows) return;
    const int e = fp8_scale_exp(*amax);
    float v[4];
#pragma unroll
    for (int t = 0; t < 4; t++) v[t] = c + t < dp ? __builtin_ldexpf(src[r * dp + c + t], e) : 0.f;
    int wv = __builtin_amdgcn_cvt_pk_fp8_f32(v[0], v[1], 0, false);
    wv = __builtin_amdgcn_cvt_pk_fp8_f32(v[2], v[3], wv, true);
    reinterpret_cast<int*>(dst)[i] = wv;
}

hipError_t launch_rows_to_fp8(const float* src, int64_t rows, int32_t dp, const uint32_t* amax, uint8_t* dst,
                              int32_t rowb, hipStream_t s)
{
    const int64_t tot = rows * (rowb / 4);
    if (tot <= 0) return hipSuccess;
    hipLaunchKernelGGL(rows_to_fp8_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, src, rows, dp,
                       amax, dst, rowb);
    return hipGetLastError();
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
    const int64_t cap = (int64_t)num_cu * (8 / PREC_W);  // resident blocks
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
    const dim3 grid((unsigned)(a.m_pad / PREC_QB), (unsigned)a.nslices), block(PREC_QB);
    // list capacities: 128 (kc <= 64; the sample pass), 512 (kc <= 128), 1024
    if (a.cap != 128 && a.cap != 512 && a.cap != 1024) return hipErrorInvalidValue;
    if (a.kc + 64 > a.cap) return hipErrorInvalidValue;
    const int jc = a.cap / 64;
    if (a.fp8) {
        // dim <= 128: one 128-value product per row
        if (a.rowb != 128 || !a.f8max) return hipErrorInvalidValue;
        if (jc == 16) hipLaunchKernelGGL((prec_stage1<1, 16, true>), grid, block, 0, s, a);
        else if (jc == 8) hipLaunchKernelGGL((prec_stage1<1, 8, true>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((prec_stage1<1, 2, true>), grid, block, 0, s, a);
        return hipGetLastError();
    }
    if (a.rowb % 64 != 0) return hipErrorInvalidValue;
    switch (a.rowb / 64) {
#define SSG_P1(KK)                                                                                  \
    case KK:                                                                                        \
        if (jc == 16) hipLaunchKernelGGL((prec_stage1<KK, 16, false>), grid, block, 0, s, a);      \
        else if (jc == 8) hipLaunchKernelGGL((prec_stage1<KK, 8, false>), grid, block, 0, s, a);   \
        else hipLaunchKernelGGL((prec_stage1<KK, 2, false>), grid, block, 0, s, a);                \
        break;
        SSG_P1(1) SSG_P1(2) SSG_P1(3) SSG_P1(4)
#undef SSG_P1
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
