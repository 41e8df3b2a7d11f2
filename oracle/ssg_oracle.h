/*
 * ssg_oracle.h -- CPU oracle for the query hot path of SimilaritySearch.jl.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only the test suite,
 * the smoke check and the CPU-baseline leg of the benchmark load it.  The
 * product (libssg_hip.so) never links, loads or calls it.
 *
 * Every function restates a reference algorithm serially in plain C; the
 * reference file:line each one follows is cited at its definition in
 * ssg_oracle.c.  Conventions (same as the product C ABI):
 *   - matrices are the reference's column-major (dim, n) Float32 layout,
 *     i.e. object i is the contiguous run X[i*dim .. i*dim+dim-1];
 *   - ids at the boundary are 1-based UInt32, 0 marks an empty slot, and
 *     empty distance slots hold +Inf;
 *   - results are (k, m) column-major: query j owns ids[j*k .. j*k+k-1].
 */
#ifndef SSG_ORACLE_H
#define SSG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_SQL2 = 0, OR_L2 = 1, OR_NORMCOS = 2 };

/* canonical distance (the contract shared with the GPU kernels) */
float or_dist(int metric, const float* a, const float* b, int dim);
/* plain sequential FP64 distance (error reference for reduced precision) */
double or_dist_f64(int metric, const float* a, const float* b, int dim);
void or_dist_batch(int metric, const float* X, int64_t n, int dim, const float* Q, int64_t m,
                   float* out /* (n, m) column-major */);

/* KnnSorted driver: ops[i]==0 pushes (ids_in[i], d_in[i]); ops[i]==1 pops the minimum.
 * accepted[i] receives the push result (1/0) or 1 for a pop. */
int or_knn_run(int32_t cap, int64_t nops, const int8_t* ops, const uint32_t* ids_in,
               const float* d_in, uint8_t* accepted, uint32_t* ids_out, float* d_out,
               int32_t* len_out);

/* ExhaustiveSearch.search over all queries */
void or_exhaustive(int metric, const float* X, int64_t n, int dim, const float* Q, int64_t m,
                   int32_t k, uint32_t* ids, float* dists, int nthreads);
/* ParallelExhaustiveSearch.search: nbatches contiguous DB chunks, merged in chunk order */
void or_parallel_exhaustive(int metric, const float* X, int64_t n, int dim, const float* Q,
                            int64_t m, int32_t k, int32_t nbatches, uint32_t* ids, float* dists);

/* BeamSearch over a CSR graph.  offs has n+1 entries (exclusive prefix, offs[0]=0) and adj
 * holds 1-based neighbour ids (the reference's StaticAdjList end_point).  hints are 1-based.
 * qhints (nullable): per-query hint lists, query j uses qhints[j*qh_stride .. +qh_count[j]-1]
 * instead of the global hints (a negative qh_count[j] keeps the global hints).  kq / bq (nullable) give per-query k and bsize; else k / bsize.
 * The output column stride is kstride (>= every k). */
void or_beamsearch(int metric, const float* X, int64_t n, int dim, const int64_t* offs,
                   const uint32_t* adj, const uint32_t* hints, int64_t nhints,
                   const uint32_t* qhints, const int32_t* qh_count, int32_t qh_stride,
                   const uint32_t* qvisit, const float* Q, int64_t m, int32_t k,
                   const int32_t* kq, int32_t kstride, int32_t bsize, const int32_t* bq,
                   float delta, int64_t maxvisits, int32_t beamcap, uint32_t* ids, float* dists,
                   int64_t* costdists,
                   int64_t* costblocks, int nthreads);
/* The BF16 form of the same search (no reference counterpart): search distance
 * or_dist_bf16, result queue of max(k, krr) entries, then rerank! with or_dist and the
 * first k written. */
void or_beamsearch_bf16(int metric, const float* X, int64_t n, int dim, const int64_t* offs,
                        const uint32_t* adj, const uint32_t* hints, int64_t nhints,
                        const uint32_t* qhints, const int32_t* qh_count, int32_t qh_stride,
                        const uint32_t* qvisit, const float* Q, int64_t m, int32_t k,
                        const int32_t* kq, int32_t kstride, int32_t bsize, const int32_t* bq,
                        float delta, int64_t maxvisits, int32_t beamcap, int32_t krr,
                        uint32_t* ids, float* dists, int64_t* costdists, int64_t* costblocks,
                        int nthreads);
/* The padded dimension of the canonical distance's 16 runs: roundup(dim, 16) up to 128,
 * else 192, 256, 384, 512, 768 or 1024 (the sizes the GPU kernels are built for). */
int or_padded_dim(int dim);
/* BF16 rounding (nearest, ties to even) and the BF16 form's FP32 search distance of row x
 * (rounded to BF16) to the FP32 query q */
float or_bf16_round(float x);
float or_dist_bf16(int metric, const float* x, const float* q, int dim);

/* allknn over a SearchGraph (neighbours of i are the hints, one shared visited set) */
void or_allknn_graph(int metric, const float* X, int64_t n, int dim, const int64_t* offs,
                     const uint32_t* adj, int32_t k, int32_t bsize, float delta,
                     int64_t maxvisits, int32_t beamcap, uint32_t* ids, float* dists,
                     int nthreads);

/* rerank! batch form: rescore the non-zero-id prefix of every column, stable-sort it */
void or_hsp_queries(int metric, const float* X, int dim, int64_t m, int32_t k,
                    const uint32_t* ids, const float* dists, uint32_t* out_ids, float* out_d,
                    int nthreads);
void or_rerank(int metric, const float* X, int dim, const float* Q, int64_t m, int32_t k,
               uint32_t* ids, float* dists);

/* macrorecall over (k, m) id matrices */
double or_macrorecall(const uint32_t* gold, int32_t kg, const uint32_t* res, int32_t kr,
                      int64_t m, int32_t k);

/* fft k-centers; start is 1-based (0 = drawn from seed) */
void or_fft(int metric, const float* X, int64_t n, int dim, int32_t k, int64_t start,
            uint64_t seed, uint32_t* centers, uint32_t* nn, float* nndists, float* eps_list);

/* RandomHints callback restated over a CSR graph.  Returns the number of hints written. */
int64_t or_random_hints(int64_t n, const int64_t* offs, const uint32_t* adj, float logbase,
                        uint64_t seed, uint32_t* hints_out, int64_t cap);

/* Graph construction (index! restated).  Output arrays are malloc'ed by the oracle and
 * released with or_free.  offs_out has n+1 entries, adj_out holds 1-based ids. */
typedef struct {
    float nb_logbase;      /* Neighborhood.logbase (1.3) */
    int32_t nb_minsize;    /* Neighborhood.minsize (2) */
    int32_t parallel_block;/* 1 => sequential insertion loop */
    float hints_logbase;   /* RandomHints.logbase (1.1); <= 0 disables the callback */
    float cb_logbase;      /* logbase_callback (1.5) */
    int32_t cb_starting;   /* starting_callback (256) */
    float opt_minrecall;   /* MinRecall of OptimizeParameters (0.97); <= 0 disables */
    int32_t opt_numqueries;/* 32 */
    int32_t opt_ksearch;   /* 10 */
    int32_t beamcap;       /* rows of beam_ids (32) */
    uint64_t seed;
} or_build_params;

typedef struct {
    int64_t* offs;
    uint32_t* adj;
    uint32_t* hints;
    int64_t nnz, nhints;
    int32_t bsize;
    float delta;
    int64_t maxvisits;
} or_graph;

int or_build_graph(int metric, const float* X, int64_t n, int dim, const or_build_params* p,
                   or_graph* out);
/* The same build in stages: index! over the first ends[0] objects, then append_items! of the
 * objects up to ends[1], ends[2], ... (ends increasing, ends[nends - 1] == n). */
int or_build_graph_staged(int metric, const float* X, int64_t n, int dim, const or_build_params* p,
                          const int64_t* ends, int nends, or_graph* out);
void or_free(void* p);

/* FP8 matrix-core products of the precision form: D[i n + j] is the restated value (before
 * the FP32 accumulation) of the e4m3fn dot product of P row i and R row j (K values each,
 * K a multiple of 8, given as their exact values) started at C[i n + j] (c_mode 0), C[j]
 * (1) or 0 (2), operand scale 2^s; Dmag the magnitude its accumulation bound scales with.
 * or_fma_norms: FP32 fma-chain squared norms of n rows of dim floats. */
void or_fp8_mfma(const float* P, int64_t m, const float* R, int64_t n, int K, const float* C, int c_mode,
                 int s, int variant, double* D, double* Dmag);
void or_fma_norms(const float* X, int64_t n, int dim, float* out);

/* optimize_index!(MinRecall) restated over the BeamSearchSpace grid; queries are given
 * explicitly (1-based ids of database objects).  Returns the chosen (bsize, delta,
 * maxvisits) through the out pointers and the achieved recall. */
double or_optimize_minrecall(int metric, const float* X, int64_t n, int dim, const int64_t* offs,
                             const uint32_t* adj, const uint32_t* hints, int64_t nhints,
                             const float* Q, int64_t m, int32_t ksearch, float minrecall,
                             int64_t cur_maxvisits, int32_t beamcap, int32_t* bsize_out,
                             float* delta_out, int64_t* maxvisits_out, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
