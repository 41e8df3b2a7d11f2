/*
 * ssg.h -- C ABI of libssg_hip.so, the MI355X (gfx950) query hot path of
 * SimilaritySearch.jl: searchbatch/allknn over SearchGraph beam search and over
 * ExhaustiveSearch / ParallelExhaustiveSearch, plus rerank!.
 *
 * The reference has no FFI; its plug-in surface is Julia multiple dispatch.  A
 * GPU index type subtypes AbstractSearchIndex (src/SimilaritySearch.jl:4) and
 * overrides the batch entry points (src/SimilaritySearch.jl:331, :384,
 * src/allknn.jl:66) with one ccall each into the functions below.
 *
 * Conventions (all of them the reference's):
 *   - databases and query sets are column-major (dim, n) Float32 matrices,
 *     i.e. object i is the contiguous run X[i*dim .. i*dim+dim-1];
 *   - results are (kstride, m) column-major: query j owns
 *     ids[j*kstride .. j*kstride+k-1];
 *   - ids are 1-based UInt32, 0 marks an empty slot, empty distances are +Inf;
 *   - each column is sorted by ascending distance, ties keep the lower id
 *     (KnnSorted's stable insertion, src/pqueue/knnsorted.jl:68-101,140-168).
 * Every entry point returns 0 on success and a negative SSG_E* code on error;
 * ssg_last_error() then describes the failure.  Calls on one handle must not
 * overlap; different handles may run concurrently.  Functions without the
 * _dev suffix take host pointers and block until the results are in host
 * memory; _dev functions take device pointers and enqueue on `stream`
 * (a hipStream_t, NULL = the handle's own stream) without synchronising.
 *
 * Size limits (SSG_E_LIMIT beyond them; the reference has none): dim <= 1024 (rows pad
 * to roundup(dim, 16) up to 128, else to 192, 256, 384, 512, 768 or 1024; above 256 the
 * exact exhaustive search takes the one-block-per-query scan, and the BF16 and precision
 * forms are refused); n < 2^32; beam search k <= kstride <= 256 (<= 64 in the BF16 form);
 * exact exhaustive k <= 3840 (k > 32 takes a one-block-per-query scan); precision form
 * k <= ncand <= 256 with dim <= 128; rerank! columns <= 1024 rows; hsp_queries
 * columns <= 256 rows; neardup blocks <= 1024 objects when filtering.
 */
#ifndef SSG_H
#define SSG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ssg_handle ssg_handle;

enum {
    SSG_OK = 0,
    SSG_E_ARG = -1,     /* invalid argument (the reference throws ArgumentError) */
    SSG_E_STATE = -2,   /* missing upload (no database / no graph) */
    SSG_E_HIP = -3,     /* HIP runtime failure */
    SSG_E_NOMEM = -4,   /* device allocation failed */
    SSG_E_LIMIT = -5    /* size beyond a documented kernel limit */
};

/* Distance functions (src/dist/vecs.jl:52-102, src/dist/cos.jl:39,76-85). */
enum { SSG_SQL2 = 0, SSG_L2 = 1, SSG_NORMCOS = 2 };

/* Storage precision of the device copy of the database.  SSG_FP8 (OCP e4m3fn) is taken by
 * the precision form of exhaustive search only. */
enum { SSG_F32 = 0, SSG_BF16 = 1, SSG_FP8 = 2 };

/* Device memory, streams and events (thin HIP wrappers, no reference counterpart) for hosts
 * that have no GPU runtime of their own; the _dev entry points take these pointers.
 * ssg_memcpy kind: 0 host->device, 1 device->host, 2 device->device (synchronous). */
int ssg_dev_malloc(void** p, size_t bytes);
int ssg_dev_free(void* p);
int ssg_memcpy(void* dst, const void* src, size_t bytes, int32_t kind);
int ssg_device_synchronize(void);
int ssg_stream_create(void** s);
int ssg_stream_destroy(void* s);
int ssg_stream_synchronize(void* s);
int ssg_event_create(void** e);
int ssg_event_destroy(void* e);
int ssg_event_record(void* e, void* stream);
int ssg_event_elapsed_ms(void* e0, void* e1, float* ms);

/* Lifecycle.  Replaces constructing the Julia index object; device = HIP ordinal. */
int ssg_create(ssg_handle** h, int device);
int ssg_destroy(ssg_handle* h);
/* Thread-local message of the last failed call (h may be NULL). */
const char* ssg_last_error(const ssg_handle* h);
/* Library version string. */
const char* ssg_version(void);

/* Upload the database (MatrixDatabase, src/db/matrixdatabase.jl:27-59) and fix the
 * index's distance (the `dist` field of ExhaustiveSearch / SearchGraph).
 * X: column-major (dim, n) Float32 host matrix. */
int ssg_upload_db(ssg_handle* h, const float* X, int32_t dim, int64_t n, int32_t metric);

/* Upload a frozen graph in StaticAdjList form (src/adj/adjstatic.jl:27-30,102-120):
 * offset[i] is the inclusive end of node i+1's list in end_point (offset[0] of the
 * reference is implicit 0), end_point holds 1-based ids.  hints are SearchGraph.hints
 * (src/searchgraph/SearchGraph.jl:147-158), 1-based. */
/* append_items!(index, ctx, X) for an exhaustive index (src/exact/sequential-exhaustive.jl:33-43
 * appends to its database): add m objects (m x dim row-major FP32, the reference's
 * column-major (dim, m)) after the uploaded ones; they get ids n+1..n+m.  Device capacity
 * doubles as needed, so repeated appends move each object once.  Rejected on a handle
 * that holds a graph. */
int ssg_append_db(ssg_handle* h, const float* X, int64_t m);
/* neardup(dist, X, eps; blocksize, filterblocks) with an exact index (src/neardup.jl:69-256,
 * recall = 1) over the handle's database, run on the device: blocks of blocksize objects
 * search the net built so far (nearest within eps: a near duplicate of it); the rest of a
 * block is filtered in order against the block's own new centers (the first center at the
 * least distance), or with filterblocks = 0 every such object becomes a center; the new
 * centers join the net.  Outputs: map (n entries of room; *nmap are filled: net position
 * -> 1-based object id, the centers in construction order), nn (n: each object's
 * covering object, itself for a center), dist (n: the distance to it), *costdists (the
 * distance evaluations: each block searched against a non-empty net charges the net's
 * size per object, as the exhaustive search does, plus the block filters' evaluations).
 * An object with no finite nearest center (NaN distances) becomes a center.  blocksize
 * <= 1024 when filtering. */
int ssg_neardup(ssg_handle* h, float eps, int32_t blocksize, int32_t filterblocks, uint32_t* map,
                int64_t* nmap, int32_t* nn, float* dist, int64_t* costdists);
int ssg_upload_graph(ssg_handle* h, const int64_t* offset, int64_t n, const uint32_t* end_point,
                     int64_t nnz, const uint32_t* hints, int64_t nhints);

/* Beam-search parameters (BeamSearch, src/searchgraph/SearchGraph.jl:94-100) plus the
 * rows of the context's beam buffers (getbeam clamps bsize to them,
 * src/searchgraph/context.jl:119-126,220-225). */
typedef struct {
    int32_t bsize;
    float delta;
    int64_t maxvisits;
    int32_t beamcap;   /* 32 in the default SearchGraphContext; at most 256 */
    int32_t precision; /* SSG_F32: canonical distances.  SSG_BF16 (no reference
                          counterpart): BF16 vector gathers with FP32 arithmetic during the
                          search, then rerank! of the result list with the canonical distance */
    int32_t rerank;    /* SSG_BF16: query j keeps and reranks max(k_j, rerank) result entries
                          (max(kstride, rerank) <= 64) */
    int32_t expand;    /* SSG_BF16 only (no reference counterpart): beam entries popped and
                          expanded together per step, 0..4; 0 and 1 are the reference step */
} ssg_beam_params;

/* Context knobs of index! (src/searchgraph/insertions.jl:94-113) with the defaults of
 * SearchGraphContext (src/searchgraph/context.jl:105-138), Neighborhood
 * (src/searchgraph/SearchGraph.jl:51-56), RandomHints (hints.jl:87-89) and
 * OptimizeParameters (optbs.jl:62-127): 1.3, 2, 8*nthreads, 1.1, 1.5, 256, 0.97, 32, 10, 32. */
typedef struct {
    float nb_logbase;       /* Neighborhood.logbase */
    int32_t nb_minsize;     /* Neighborhood.minsize */
    int32_t parallel_block; /* insertion block; >= 2 (the parallel insertion loop) */
    float hints_logbase;    /* RandomHints.logbase; <= 0 disables the callback */
    float cb_logbase;       /* logbase_callback */
    int32_t cb_starting;    /* starting_callback */
    float opt_minrecall;    /* MinRecall of the OptimizeParameters callback; <= 0 disables it */
    int32_t opt_numqueries;
    int32_t opt_ksearch;
    int32_t beamcap;        /* rows of the context's beam buffers */
    uint64_t seed;          /* draws of RandomHints and of the optimiser's query sample */
    int32_t verbose;        /* > 0: progress lines on stderr every `verbose` percent */
} ssg_build_params;

/* index!(G::SearchGraph, ctx) over the uploaded database: builds the graph on the device
 * and installs it in the handle, together with its hints and the tuned BeamSearch. */
int ssg_build_graph(ssg_handle* h, const ssg_build_params* p);
/* append_items!(G::SearchGraph, ctx, db) (src/searchgraph/insertions.jl:27-34): appends the
 * m objects of X (column-major (dim, m)) to the database of a graph that ssg_build_graph
 * built on this handle and continues index! from length(index) + 1 with the same build
 * parameters and the same generator of draws (blocks of parallel_block objects start at
 * the first new object, callbacks fire as their positions say). */
int ssg_append_items_graph(ssg_handle* h, const float* X, int64_t m);
/* neardup(dist, X, eps; recall < 1, k, blocksize, filterblocks) (src/neardup.jl:71-76): the
 * same loop with the net held in a SearchGraph grown by append_items! (index! with the build
 * parameters bp, whose opt_minrecall carries MinRecall(recall); continued over each block's
 * new centers).  Each block searches the net as it stands with KnnSorted(k) queues (the
 * graph's current BeamSearch; the nearest result decides, 0 < k <= 64).  *costdists counts
 * the block searches' evaluations (costdists of the beam searches) and the block filters';
 * index!'s own searches and its optimiser's are not counted.  net: a handle holding an empty
 * database of h's dimension and distance (ssg_upload_db with n = 0); it returns holding the
 * net's centers (rows in map order) and their graph (the reference's idx).  Outputs as
 * ssg_neardup. */
int ssg_neardup_graph(ssg_handle* h, ssg_handle* net, float eps, int32_t blocksize, int32_t filterblocks,
                      int32_t k, const ssg_build_params* bp, uint32_t* map, int64_t* nmap, int32_t* nn,
                      float* dist, int64_t* costdists);
/* Size of the installed graph and its current BeamSearch parameters. */
int ssg_graph_info(ssg_handle* h, int64_t* nnz, int64_t* nhints, ssg_beam_params* algo);
/* Copy the installed graph out in StaticAdjList form (offset[n] inclusive ends, 1-based
 * end_point[nnz]) plus its 1-based hints (nullable). */
int ssg_download_graph(ssg_handle* h, int64_t* offset, uint32_t* end_point, uint32_t* hints);
/* optimize_index!(G, ctx, MinRecall(minrecall); queries=Q, ksearch) (src/opt.jl:220-317):
 * algo_inout gives beamcap and the current maxvisits and receives the chosen
 * (bsize, delta, maxvisits); *recall gets the sample recall.  Q: column-major (dim, m). */
int ssg_optimize_minrecall(ssg_handle* h, const float* Q, int64_t m, int32_t ksearch,
                           float minrecall, ssg_beam_params* algo_inout, double* recall);

/* searchbatch!(::ExhaustiveSearch / ::ParallelExhaustiveSearch, ctx, Q, ids, dists)
 * (src/SimilaritySearch.jl:331-348 over src/exact/sequential-exhaustive.jl:68-80 and
 * src/exact/parallel-exhaustive.jl:64-102).  Q: column-major (dim, m). */
int ssg_exhaustive_knn(ssg_handle* h, const float* Q, int64_t m, int32_t k, uint32_t* ids,
                       float* dists);
int ssg_exhaustive_knn_dev(ssg_handle* h, const float* dQ, int64_t m, int32_t k, uint32_t* dids,
                           float* ddists, void* stream);

/* Per-query optional inputs of the queue-vector form searchbatch!(index, ctx, Q, knns)
 * (src/SimilaritySearch.jl:384-398): per-query k and bsize, and per-query hint lists
 * that replace SearchGraph.hints (the approx_by_hints! injection point,
 * src/searchgraph/hints.jl:8-14).  Every pointer may be NULL. */
typedef struct {
    const int32_t* k;        /* [m], each 1 <= k[j] <= kstride */
    const int32_t* bsize;    /* [m] */
    const uint32_t* hints;   /* [m * hint_stride], 1-based */
    const int32_t* nhints;   /* [m] */
    int32_t hint_stride;
    const uint32_t* order;   /* [m] nullable, no reference counterpart: the order in which
                                queries are taken (a permutation of 0..m-1; results do not
                                depend on it) */
    const uint32_t* visited; /* [m] nullable: a 1-based object marked visited before query j's
                                search starts (0 = none), as bichromatic_search! marks the
                                query's own vertex (src/bichromatic/bichromaticclosestpair.jl:
                                107-113); the n < 64 brute-force branch ignores it, as the
                                reference's does */
} ssg_query_opts;

/* searchbatch!(::SearchGraph, ctx, Q, ids, dists) with search(::BeamSearch, ...)
 * (src/searchgraph/beamsearch.jl:88-110, src/searchgraph/SearchGraph.jl:210-213).
 * costdists / costblocks (nullable, [m]) receive the per-query counters that the
 * reference accumulates in ctx.costdists / ctx.costblocks
 * (src/SimilaritySearch.jl:224-225). */
int ssg_beamsearch(ssg_handle* h, const float* Q, int64_t m, int32_t k, int32_t kstride,
                   const ssg_beam_params* bp, const ssg_query_opts* qo, uint32_t* ids,
                   float* dists, int64_t* costdists, int64_t* costblocks);
/* Device-pointer form; qo's arrays are device pointers too. */
int ssg_beamsearch_dev(ssg_handle* h, const float* dQ, int64_t m, int32_t k, int32_t kstride,
                       const ssg_beam_params* bp, const ssg_query_opts* qo, uint32_t* dids,
                       float* ddists, int64_t* dcostdists, int64_t* dcostblocks, void* stream);

/* The form of ssg_beamsearch_dev that allocates nothing.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  Same arguments, results and errors as
 * ssg_beamsearch_dev; its visited slots, slot epochs and work counter live in a static
 * 256 MiB device array that each GPU the library runs on holds until the process ends, so
 * it serves databases of at most 67,076,096 objects (SSG_E_LIMIT beyond; fewer wave slots
 * for larger databases: 255 at 1M objects against 32 per CU in ssg_beamsearch_dev).  It
 * searches the query layout only if an earlier ssg_beamsearch_dev call built it (results
 * do not depend on the layout), and the BF16 form needs the BF16 rows an earlier BF16
 * ssg_beamsearch_dev call on the handle made (SSG_E_STATE otherwise).  The array keeps the
 * last call's contents and each call clears what it uses first; calls whose kernels can
 * run at the same time on one GPU (other streams, other handles) race on it. */
int ssg_beamsearch_noalloc_dev(ssg_handle* h, const float* dQ, int64_t m, int32_t k,
                               int32_t kstride, const ssg_beam_params* bp,
                               const ssg_query_opts* qo, uint32_t* dids, float* ddists,
                               int64_t* dcostdists, int64_t* dcostblocks, void* stream);

/* HIP graphs (no reference counterpart): launches issued by ssg_*_dev calls on `stream`
 * between ssg_capture_begin and ssg_capture_end are recorded into a graph that
 * ssg_graph_launch replays on any stream.  Every buffer those calls use must already be
 * sized, so make each call once before capturing it; the calls record only kernels (no
 * memsets, copies or synchronisation).  A graph replays with the arguments it was
 * recorded with, so it is valid while those buffers and the handle's graph live. */
typedef struct ssg_graph ssg_graph;
int ssg_capture_begin(void* stream);
int ssg_capture_end(void* stream, ssg_graph** graph);
int ssg_graph_launch(ssg_graph* graph, void* stream);
int ssg_graph_destroy(ssg_graph* graph);

/* Launch shape of the handle's beam searches (no reference counterpart): persistent wave
 * slots per CU (each wave owns a visited slot of n bytes; 4..64, default 32), gather
 * rounds per batch (each round gathers one row per 16-lane group: 8, the default; 4, fewer
 * registers; 10, in the BF16 form, 40 children per batch, for beam caps <= 64, others run
 * 8) and the LDS visited table of each wave (vbuckets buckets
 * of eight 16-bit tags, 16 bytes each: a visited mark goes to the table while its bucket
 * has room, else to the wave's HBM slot; -1, the default: the largest table that leaves
 * the kernel's occupancy unchanged; 0: every mark in HBM; 1..2048: that many buckets,
 * the LDS of a block, 4 * 16 * vbuckets bytes, then bounding how many waves share a CU.
 * A table is used only when n <= 2^24 and vbuckets >= n / 2^15).  Results do not depend
 * on it. */
typedef struct {
    int32_t slots_per_cu;
    int32_t rounds;
    int32_t vbuckets;
} ssg_launch_params;
int ssg_set_launch(ssg_handle* h, const ssg_launch_params* lp);
int ssg_get_launch(ssg_handle* h, ssg_launch_params* lp);

/* Launch autotuner (no reference counterpart; it extends optimize_index!'s recall-targeted
 * search, src/opt.jl:220-317, from BeamSearch's (bsize, Δ) to the launch).  Over m
 * device-resident queries with per-query options qo (device arrays, nullable), it runs
 * the search of *bp for each recall-changing knob -- in the BF16 form the rerank length
 * (r, r + 2, r + 4, 1.5 r and 2 r from r = bp->rerank, or k when it is 0; max(kstride,
 * rerank) <= 64) and expand 1 or 2 -- and
 * measures recall against gold (host, m x kg 1-based ids in any order; query j is scored
 * on its first min(k_j, kg) results against its first min(k_j, kg) gold ids).  Among the
 * knobs reaching minrecall (else the best recall) it times every launch shape
 * (slots_per_cu in {8, 16, 24, 32, 48, 64} x rounds in {4, 8, 10 (BF16)} x vbuckets in
 * {-1, 0}) with HIP events and keeps
 * the fastest.  It installs that shape, writes the chosen rerank/expand into *bp, and
 * returns the shape, its recall and its time (ms).  Uses the handle's stream. */
int ssg_optimize_launch(ssg_handle* h, const float* dQ, int64_t m, int32_t k, int32_t kstride,
                        ssg_beam_params* bp, const ssg_query_opts* qo, const uint32_t* gold,
                        int32_t kg, float minrecall, ssg_launch_params* lp_out, double* recall_out,
                        float* ms_out);

/* Handle options (no reference counterpart).
 * SSG_OPT_QUERY_LAYOUT: 1 (default) lets searchbatch! run on a copy of the graph and the
 * vectors renumbered breadth-first from the hints, each list entry carrying its target's
 * {start, degree}; results, ids and cost counters are unchanged, memory traffic is
 * lower.  0 searches the installed layout directly. */
/* SSG_OPT_PREC_SLICES: number of DB slices of the precision form's candidate pass (0, the
 * default, chooses it from the device size); results do not depend on it.
 * SSG_OPT_PREC_SAMPLE: the precision form's sampled thresholds (a pass over every 16th
 * 128-row tile sets each query's starting threshold at its T-th best sampled score; queries
 * the result cannot certify are searched again from a lower threshold, or from none).
 * 0 (default): used on slices of at least 131072 rows with T = max(16, ncand / 8); -1:
 * never; T in 1..256: always, with that T.  Results do not depend on it. */
enum { SSG_OPT_QUERY_LAYOUT = 1, SSG_OPT_PREC_SLICES = 2, SSG_OPT_PREC_SAMPLE = 3 };
int ssg_set_option(ssg_handle* h, int32_t option, int64_t value);

/* Longest-first order for a heterogeneous batch (no reference counterpart): dorder[0..m)
 * receives the query indices grouped by floor(log2(bsize[j])), largest first, for
 * ssg_query_opts.order.  dbsize and dorder are device arrays. */
int ssg_order_by_bsize_dev(ssg_handle* h, const int32_t* dbsize, int64_t m, uint32_t* dorder,
                           void* stream);

/* allknn!(::SearchGraph, ctx, ids, dists) (src/allknn.jl:66-106): every database object
 * is a query; the neighbours of object i seed one search each, sharing one visited set. */
int ssg_allknn_graph(ssg_handle* h, int32_t k, const ssg_beam_params* bp, uint32_t* ids,
                     float* dists);
/* allknn!(::ExhaustiveSearch, ...) (src/allknn.jl:66-91,108-110): exact, self included. */
int ssg_allknn_exhaustive(ssg_handle* h, int32_t k, uint32_t* ids, float* dists);

/* Precision form of searchbatch!/allknn! over ExhaustiveSearch (no reference counterpart;
 * the reference computes Float32 distances only).  It is allknn/searchbatch with a
 * candidate queue of ncand entries followed by rerank! (src/rerank.jl:22-98):
 *   1. a candidate pass on the matrix cores (FP32 accumulation) with precision SSG_BF16
 *      operands (round to nearest even) or SSG_FP8 operands (OCP e4m3fn of the values
 *      times 2^e, round to nearest even; e is the largest integer with
 *      max |value| * 2^e <= 448, one e for the database and one for the queries, clamped
 *      to [-120, 120]) keeps, per query, the ncand objects with the best approximate
 *      distance (ties to the lower id);
 *   2. those are rescored with the canonical FP32 distance and the k first in
 *      (distance, id) order are returned, (k, m) column-major, 1-based ids.
 * The returned distances are exact; an object outside the ncand best approximate
 * candidates can be missed (recall < 1).  k <= ncand <= 256, dim <= 128.  cand_ids /
 * cand_dists (nullable together, (ncand, m)) receive every candidate in (distance, id)
 * order with its approximate distance.  The allknn forms take every database object as
 * a query, self included (src/allknn.jl:66-91). */
int ssg_exhaustive_knn_prec(ssg_handle* h, const float* Q, int64_t m, int32_t k, int32_t precision,
                            int32_t ncand, uint32_t* ids, float* dists, uint32_t* cand_ids,
                            float* cand_dists);
int ssg_exhaustive_knn_prec_dev(ssg_handle* h, const float* dQ, int64_t m, int32_t k,
                                int32_t precision, int32_t ncand, uint32_t* dids, float* ddists,
                                uint32_t* dcand_ids, float* dcand_dists, void* stream);
int ssg_allknn_exhaustive_prec(ssg_handle* h, int32_t k, int32_t precision, int32_t ncand,
                               uint32_t* ids, float* dists, uint32_t* cand_ids, float* cand_dists);
int ssg_allknn_exhaustive_prec_dev(ssg_handle* h, int32_t k, int32_t precision, int32_t ncand,
                                   uint32_t* dids, float* ddists, uint32_t* dcand_ids,
                                   float* dcand_dists, void* stream);

/* fft(dist, X, k; start) over the uploaded database (src/kcenters/fft.jl:53-87): farthest-first
 * k-centers.  centers[k] (1-based), nn[n] (1-based nearest center, nullable), nndists[n]
 * (nullable), eps[k] (the ε recorded before each center was added, nullable). */
int ssg_fft(ssg_handle* h, int32_t k, int64_t start, uint32_t* centers, uint32_t* nn,
            float* nndists, float* eps);

/* ---------------------------------------------------------------------------------
 * Learned entry points (no reference counterpart).  An MLP D -> H1 -> H2 -> (L + NB),
 * ReLU, FP16 operands on MFMA with FP32 accumulation and FP32 Adam master weights,
 * scores L landmarks (entry points, from ssg_fft) and NB beam-size buckets.  It is
 * trained on traces of the plain search and feeds per-query hints and bsize into
 * ssg_beamsearch_dev through ssg_learned_plan_dev; low-confidence queries fall back
 * to the plain search (SearchGraph.hints and the plain bsize).
 * --------------------------------------------------------------------------------- */
typedef struct ssg_mlp ssg_mlp;
int ssg_mlp_create(ssg_mlp** p, int device, int32_t D, int32_t H1, int32_t H2, int32_t L,
                   int32_t NB, uint64_t seed);
int ssg_mlp_destroy(ssg_mlp* p);
/* FP32 parameters of layer 0..2: W row-major [out x in], b [out] (nullable). */
int ssg_mlp_get(ssg_mlp* p, int32_t layer, float* W, float* b);
int ssg_mlp_set(ssg_mlp* p, int32_t layer, const float* W, const float* b);
/* Gradients of the last step (same layout as the parameters). */
int ssg_mlp_grad(ssg_mlp* p, int32_t layer, float* gW, float* gb);
/* One forward/backward on a host batch X [B x D]; update != 0 applies one Adam step.
 * Loss = mean CE(landmark head) + bucket_weight * mean CE(bucket head). */
int ssg_mlp_step(ssg_mlp* p, const float* X, const int32_t* ylandmark, const int32_t* ybucket,
                 int32_t B, float bucket_weight, float lr, int32_t update, float* loss);
/* Mini-batch Adam epochs over a device-resident training set dX [N x D]. */
int ssg_mlp_train_dev(ssg_mlp* p, const float* dX, const int32_t* dyl, const int32_t* dyb,
                      int64_t N, int32_t epochs, int32_t batch, float bucket_weight, float lr,
                      uint64_t seed, float* losses);
/* Inference on device rows dQ [m x D]: the h best landmarks (mapped through dlandmarks to
 * 1-based object ids) into dhints [m x h], the bucket, and the top-landmark softmax
 * probability.  The bucket is the most probable one when bucket_quantile is 0, else the
 * smallest bucket whose cumulative predicted probability reaches bucket_quantile. */
int ssg_mlp_infer_dev(ssg_mlp* p, const float* dQ, int64_t m, const uint32_t* dlandmarks,
                      int32_t h, float bucket_quantile, uint32_t* dhints, int32_t* dbucket,
                      float* dconf, void* stream);
/* Per-query search inputs from the inference outputs: nhints[j] = h when conf[j] >=
 * min_conf else -1 (the plain hints), bsize[j] = bucket_bsize[bucket[j]] or
 * fallback_bsize.  bucket_bsize is a host array of NB entries. */
int ssg_learned_plan_dev(ssg_handle* h, const int32_t* dbucket, const float* dconf, int64_t m,
                         int32_t nh, float min_conf, const int32_t* bucket_bsize, int32_t nb,
                         int32_t fallback_bsize, int32_t* dnhints, int32_t* dbsize, void* stream);

/* The learned form end to end on `stream`: ssg_mlp_infer_dev (nh landmark hints per query),
 * ssg_learned_plan_dev, ssg_order_by_bsize_dev (largest beams first), then
 * ssg_beamsearch_dev with those per-query hints, beam sizes and order (bp->bsize is the
 * fallback beam size; bp->precision selects the BF16 form).  Scratch is owned by the
 * handle, so repeated calls allocate nothing once sized and the sequence can be captured
 * into a graph (ssg_capture_begin). */
int ssg_learned_search_dev(ssg_handle* h, ssg_mlp* mlp, const uint32_t* dlandmarks,
                           const float* dQ, int64_t m, int32_t k, int32_t kstride,
                           const ssg_beam_params* bp, int32_t nh, float min_conf,
                           float bucket_quantile, const int32_t* bucket_bsize, int32_t nb,
                           uint32_t* dids, float* ddists, int64_t* dcd, int64_t* dcb, void* stream);

/* rerank!(dist, db, queries, ids, dists) batch form (src/rerank.jl:77-86): recompute the
 * non-zero-id prefix of each column with the handle's distance and stable-sort it
 * (columns of up to 1024 rows).
 * ids / dists are updated in place. */
int ssg_rerank(ssg_handle* h, const float* Q, int64_t m, int32_t k, uint32_t* ids, float* dists);
int ssg_rerank_dev(ssg_handle* h, const float* dQ, int64_t m, int32_t k, uint32_t* dids,
                   float* ddists, void* stream);

/* hsp_queries(dist, X, Q, knns_ids, knns_dists) (src/hsp.jl:77-101, hsp_should_push
 * :8-33) over the handle's database and distance: each of the m (k, m) candidate columns
 * (ascending, 0 ends a column; 1 <= k <= 256 with k * padded dim * 4 <= 160 KiB) is scanned in order and a candidate is kept
 * unless an already kept one is strictly nearer to it than its given distance.  The kept
 * candidates fill the top of the hsp_ids / hsp_dists columns as a sorted queue (stable on
 * equal distances), then 0 / +Inf.  The queries themselves are not needed: the given
 * distances stand for them. */
int ssg_hsp_queries(ssg_handle* h, int64_t m, int32_t k, const uint32_t* ids, const float* dists,
                    uint32_t* hsp_ids, float* hsp_dists);
int ssg_hsp_queries_dev(ssg_handle* h, int64_t m, int32_t k, const uint32_t* dids,
                        const float* ddists, uint32_t* dhsp_ids, float* dhsp_dists, void* stream);

#ifdef __cplusplus
}
#endif
#endif
