// ssg_kernels.h -- argument blocks and launchers of the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ssg {

// Device layout of one index (one handle):
//   X      n rows of dp floats (dp = roundup(dim,16), zero padded), 64-byte aligned rows
//   xnorm  n floats: squared norms (SqL2/L2) used by the MFMA tiles
//   node   n uint2 {start, deg}: node i's list is adj[start .. start+deg)
//   adj    uint32 pool of 0-based neighbour ids (a frozen CSR, or the growable
//          pool of the GPU graph builder)
//   hints  nhints uint32, 1-based (as SearchGraph.hints)
struct BeamArgs {
    const float* X;
    int64_t n;
    int32_t dp;
    int32_t metric;
    const uint2* node;
    const uint32_t* adj;
    const uint32_t* hints;  // 1-based
    int32_t nhints;
    int32_t log2n;          // ceil(log2(1+n)), beamsearch.jl:30
    const float* Q;         // (m, qdim) rows, unpadded
    int32_t qdim;
    int64_t m;
    int32_t k;
    const int32_t* kq;
    int32_t kstride;
    int32_t bsize;
    const int32_t* bq;
    float delta;
    int64_t maxvisits;
    int32_t beamcap;
    const uint32_t* qvisit;  // 1-based, [m], 0 = none: marked visited before the search
    const uint32_t* qhints;  // 1-based, [m * qhstride]
    const int32_t* qhcount;
    int32_t qhstride;
    uint32_t* ids;
    float* dists;
    int64_t* costd;
    int64_t* costb;
    uint8_t* vis;            // nslots * vis_stride epoch bytes
    uint8_t* epochs;         // nslots
    int64_t vis_stride;
    unsigned int* counter;
    int32_t has_dups;
    int32_t allknn;
    int32_t internal;        // 1: issued by index! / optimize_index!
    int32_t prec;            // 0: canonical FP32 path, 1: BF16 gathers + exact rerank
    const uint16_t* Xh;      // BF16 copy of X (n rows of dp), prec 1
    int32_t krr;             // prec 1: query j reranks max(k_j, krr) result entries exactly
    int32_t rounds;          // gather rounds per batch: 8, or 4 / 10 (beam caps <= 64; 10: BF16 form)
    int32_t expand;          // prec 1: beam entries popped per step (1 = reference step)
    int32_t vnb;             // LDS visited-table buckets per wave (8 tags each; 0 none, < 0 auto)
    int32_t vhb;             // objects < 2^vhb (15 <= vhb <= 24) when vnb != 0
    const uint32_t* order;   // nullable: order[t] = the t-th query to search (0-based)
    // Query layout (nullable): objects stored in another order.  adjn[e] = {start, degree}
    // of the object adj[e] names; perm[stored] = original id (0-based); ipos = inverse.
    const uint2* adjn;
    const uint32_t* perm;
    const uint32_t* ipos;
};

// Launch the persistent beam-search kernel over nslots waves (a multiple of 4).
hipError_t launch_beam(const BeamArgs& a, int nslots, hipStream_t s);
hipError_t launch_zero_u32(unsigned int* p, int64_t n, hipStream_t s);
hipError_t launch_fill_f32(float* p, int64_t n, float v, hipStream_t s);
// FP32 -> BF16 with round-to-nearest-even
hipError_t launch_to_bf16(const float* src, int64_t n, uint16_t* dst, hipStream_t s);
// dst row i = src row perm[i] (rows of dp floats) for i < n; rows n..n_pad are zero
hipError_t launch_permute_rows(const float* src, const uint32_t* perm, int64_t n, int64_t n_pad,
                               int32_t dp, float* dst, hipStream_t s);
// order[0..m) = query indices by descending floor(log2(key)) (key >= 1); counts is
// scratch of 64 words.  Order within a class is unspecified.
hipError_t launch_order_desc(const int32_t* key, int64_t m, uint32_t* order, unsigned int* counts,
                             hipStream_t s);

struct ExhArgs {
    const float* X;
    const float* xnorm;
    int64_t n;
    int32_t dp;
    int32_t metric;
    const float* Q;
    int32_t qdim;
    int64_t m;
    int32_t k;
};

// Exhaustive k-NN: MFMA distance tiles + per-slice candidate selection, then an
// exact rerank/merge with a certificate, then exact rescans of any query whose
// certificate failed.  Work buffers are owned by the caller (the handle).
struct ExhWork {
    float* qpad;        // m_pad * dp
    float* qnorm;       // m_pad
    float* cand_d;      // m_pad * nslices * KC
    uint32_t* cand_i;   // m_pad * nslices * KC
    float* slice_tau;   // m_pad * nslices
    unsigned int* flags;   // [0] = number of flagged queries, then the list
    int64_t m_cap;
    int32_t nslices_cap;
    int64_t m_valid;       // queries actually present (qpad may hold fewer than m_pad rows)
    uint2* list;           // dp <= 128: m_pad * nslices * 512 candidate-list entries
    const float* tau0;     // nullable, dp <= 128: per-query starting threshold (score units)
};
hipError_t exh_launch_stage1(const ExhArgs& a, const ExhWork& w, int nslices, hipStream_t s);
// Starting thresholds: tau0[q] = a lower bound of query q's k-th best score among
// min(n, 2048) rows spread over the database (the k-th largest of each lane's two best),
// scored directly in FP32 and lowered by the certificate's error allowance (xnmax: the
// largest squared row norm); -Inf with fewer than k rows, k outside 1 .. 128, and for
// rows m_valid .. m_pad.
hipError_t exh_launch_seed_sample(const float* X, const float* xnorm, int32_t dp, int64_t n,
                                  const ExhWork& w, int64_t m_valid, int64_t m_pad, int k, int metric,
                                  float xnmax, float* tau0, hipStream_t s);
hipError_t exh_launch_stage2(const ExhArgs& a, const ExhWork& w, float xnmax, int nslices,
                             uint32_t* ids, float* dists, int kstride, hipStream_t s);
hipError_t exh_launch_rescan(const ExhArgs& a, const ExhWork& w, int f0, int nf, uint32_t* part_i,
                             float* part_d, int nparts, uint32_t* ids, float* dists, int kstride,
                             hipStream_t s);
int32_t exhaustive_slices(int64_t n, int64_t m, int num_cu, int dp);
constexpr int EXH_KC = 32;        // candidates kept per (query, slice)
constexpr int EXH_WIDE_KMAX = 3840;  // k of the one-block-per-query path (k > EXH_KC)
constexpr int EXH_QT = 128;       // queries per block of exhaustive stage 1 (query padding)

// Precision form of exhaustive k-NN: BF16 MFMA candidate pass + exact rerank.
constexpr int PREC_W = 8;             // waves (of 64 queries) per candidate-pass block
constexpr int PREC_QB = 64 * PREC_W;  // queries per candidate-pass block
constexpr int PREC_NCAND_MAX = 256;   // candidates per query
constexpr int PREC_CAP_MAX = 1024;    // list capacity per (query, slice): prec_list_cap(ncand)
// 4 sub-lists of a power-of-two number of entries each, room for the kept entries of a
// compaction (at most ncand + 64) plus at least 250 appends
inline int32_t prec_list_cap(int32_t ncand)
{
    return ncand <= 128 ? 512 : 1024;
}
struct PrecArgs {
    const uint8_t* Xe;      // encoded rows of rowb bytes (zero padded): BF16 or scaled FP8
    const float* xnorm;     // FP32 squared norms of the rows
    const float* X;         // n_pad rows of dp FP32 (the rerank reads them)
    int64_t n;
    int32_t rowb;           // bytes per encoded row: 64 * KB (BF16), 128 * KB (FP8)
    int32_t metric;
    const uint8_t* Qe;      // m_pad encoded query rows of rowb bytes
    int32_t fp8;            // 0: BF16 operands; 1: FP8 e4m3fn operands scaled by 2^e
    const uint32_t* f8max;  // fp8: [0] bits of max |x| over the rows, [1] over the queries
                            // (the scale exponents e follow from them: fp8_scale_exp)
    int64_t m_pad;          // multiple of PREC_QB
    int32_t kc;             // candidates kept per (query, slice)
    int32_t cap;            // list capacity per (query, slice)
    int64_t rows_per_slice; // multiple of the 128-row tile
    int32_t nslices;
    uint2* list;            // [m_pad][nslices][cap] {score bits, 0-based id}
    int32_t* list_n;        // [m_pad][nslices]
    int32_t tile_step;      // 1: every 128-row tile of the slice; f: every f-th (sample pass)
    int32_t sample;         // 1: sample pass, writes tau0 (the kc-th best sampled score)
    float* tau0;            // nullable, [m_pad][nslices]: the sample pass writes it; the main
                            // pass starts each query's threshold there (rows at or below it
                            // are dropped) and the rerank stage certifies the result
    const float* tau_seed;  // nullable, sample pass only: its starting thresholds (from a
                            // coarser sample); rows at or below them are dropped
};
struct PrecOut {
    int64_t m;
    int32_t k, kstride, ncand;
    const float* qpad;      // m rows of dp FP32 queries (exact rerank)
    const float* qnorm;     // FP32 squared query norms (approximate distances)
    uint32_t* ids;          // (k, m), 1-based
    float* dists;
    uint32_t* cand_ids;     // nullable: (ncand, m) candidates in (distance, id) order
    float* cand_dists;      // their approximate distances
    const uint32_t* qmap;   // nullable: query j of this launch is row qmap[j] of qpad,
                            // qnorm and the outputs (the redo of flagged queries)
    unsigned int* flags;    // nullable with tau0: [0] = number of queries whose result is
                            // not certified, then their indices
};
// Number of rows of the 1-in-step-tiles sample of n rows, and the packing kernel.
inline int64_t prec_sample_rows(int64_t n, int32_t step)
{
    const int64_t g = (int64_t)128 * step, groups = n / g, rest = n % g;
    return groups * 128 + (rest < 128 ? rest : 128);
}
hipError_t launch_sample_rows(const uint16_t* Xe, const float* xnorm, int64_t ns_rows, int32_t step,
                              int32_t width, uint16_t* Xs, float* xns, hipStream_t s);
// Redo of uncertified queries: out[i] = max over the ns slices of src[idx[i] * ns + s], the
// next float below it;
// out[i * nsf + s] = v[i] (i < cnt) or -Inf (cnt <= i < rows); out[i] = a[b[i]].
hipError_t launch_max_slices_f32(const float* src, int32_t ns, const uint32_t* idx, int64_t cnt, float* out,
                                 hipStream_t s);
hipError_t launch_spread_f32(const float* v, int64_t cnt, int64_t rows, int32_t nsf, float* out, hipStream_t s);
hipError_t launch_compose_u32(const uint32_t* a, const uint32_t* b, int64_t cnt, uint32_t* out, hipStream_t s);
hipError_t launch_gather_rows_u16(const uint16_t* src, const uint32_t* idx, const unsigned int* count,
                                  int64_t max_rows, int32_t width, uint16_t* dst, hipStream_t s);
int32_t prec_slices(int64_t n, int64_t m, int num_cu, int kc);
hipError_t prec_launch_stage1(const PrecArgs& a, hipStream_t s);
hipError_t prec_launch_stage2(const PrecArgs& a, const PrecOut& o, int32_t dp, hipStream_t s);
hipError_t launch_rows_to_bf16(const float* src, int64_t rows, int32_t dp, uint16_t* dst, int32_t dpb,
                               hipStream_t s);
// FP8 operands: amax[0 .. namax) (zeroed by the caller) each receive the bits of max |src|
// over count values; rows_to_fp8 writes rows of rowb bytes, e4m3fn(src * 2^e) rounded to
// nearest even with e = fp8_scale_exp(*amax), zero padded.
hipError_t launch_absmax(const float* src, int64_t count, uint32_t* amax, int32_t namax, hipStream_t s);
hipError_t launch_rows_to_fp8(const float* src, int64_t rows, int32_t dp, const uint32_t* amax,
                              uint8_t* dst, int32_t rowb, hipStream_t s);
// The largest e with max_abs * 2^e <= 448 (the largest finite e4m3fn value), clamped to
// [-120, 120]; 0 when max_abs is 0 or not finite.
__host__ __device__ inline int fp8_scale_exp(uint32_t amax_bits)
{
    const uint32_t ex = (amax_bits >> 23) & 255u, mant = amax_bits & 0x7fffffu;
    if (amax_bits == 0u || ex == 255u) return 0;
    // max_abs = f * 2^E with f in [1, 2): f * 2^(E+e) <= 1.75 * 2^8
    int E;
    uint32_t f;  // the 24-bit significand of max_abs, normalised
    if (ex == 0u) {
        const int lz = __builtin_clz(mant) - 8;  // shift that brings the leading bit to bit 23
        f = mant << lz;
        E = -126 - lz;
    } else {
        f = mant | 0x800000u;
        E = (int)ex - 127;
    }
    int e = (f <= 0xe00000u ? 8 : 7) - E;  // 0xe00000 = 1.75 * 2^23
    return e < -120 ? -120 : (e > 120 ? 120 : e);
}

// Exact exhaustive k-NN for EXH_KC < k <= EXH_WIDE_KMAX: one block per query over the n
// padded rows of X; Qp holds m padded query rows; results (kstride, m), 0 / +Inf past n.
hipError_t launch_exh_wide(const float* X, int64_t n, int32_t dp, int32_t metric, const float* Qp,
                           int64_t m, int32_t k, uint32_t* ids, float* dists, int32_t kstride,
                           hipStream_t s);
hipError_t launch_rerank(const float* X, int32_t dp, int32_t metric, const float* Q, int32_t qdim,
                         int64_t m, int32_t k, uint32_t* ids, float* dists, hipStream_t s);
// hsp_queries filter (ssg_consumers.hip): (k, m) candidate columns -> kept columns.
hipError_t launch_hsp(const float* X, int32_t dp, int32_t metric, int64_t m, int32_t k,
                      const uint32_t* ids, const float* dists, uint32_t* out_ids, float* out_dists,
                      hipStream_t s);
// neardup_block! pieces (ssg_consumers.hip)
hipError_t launch_pair_dist(const float* X, int32_t dp, int32_t metric, const uint32_t* ids, int32_t mb,
                            float* P, hipStream_t s);
hipError_t launch_greedy_net(const float* P, int32_t mb, float eps, int32_t* cover, float* cdist,
                             hipStream_t s);
hipError_t launch_nearest_center(const float* X, int32_t dp, int32_t metric, const uint32_t* items,
                                 int32_t mb, const uint32_t* F, int32_t nf, int32_t* pos, float* pdist,
                                 hipStream_t s);
hipError_t launch_gather_padded_rows(const float* X, const float* xn, int32_t dp, const uint32_t* ids,
                                     int64_t m, float* Y, float* yn, int64_t n0, hipStream_t s);
// dst rows of dp values: rows 0 .. rows - 1 from src (dim values, zero padded) with their
// squared norms, rows .. rows_pad - 1 zero.
hipError_t launch_pad_rows(const float* src, int64_t rows, int32_t dim, float* dst, int32_t dp,
                           float* norms, hipStream_t s, int64_t rows_pad = 0);

}  // namespace ssg
