// ssg_handle.h -- the handle behind the C ABI and the helpers shared by the ABI
// entry points (ssg_abi.hip) and the graph builder (ssg_build.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ssg.h"
#include "ssg_kernels.h"

// Padded row size of dimension dim: roundup(dim, 16) up to 128, else the next of 192, 256,
// 384, 512, 768 and 1024 (the row sizes the kernels are built for; 0 above 1024).
inline int32_t ssg_padded_dim(int64_t dim)
{
    const int32_t dp = (int32_t)((dim + 15) / 16 * 16);
    if (dp <= 128) return dp;
    for (int32_t s : {192, 256, 384, 512, 768, 1024})
        if (dp <= s) return s;
    return 0;
}

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    hipError_t ensure(size_t bytes)
    {
        if (bytes <= cap) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        const size_t b = std::max<size_t>(bytes, 256);
        hipError_t e = hipMalloc(&p, b);
        if (e == hipSuccess) cap = b;
        return e;
    }
    // grow keeping the first `keep` bytes
    hipError_t grow(size_t bytes, size_t keep, hipStream_t s)
    {
        if (bytes <= cap) return hipSuccess;
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, bytes);
        if (e != hipSuccess) return e;
        if (p && keep) e = hipMemcpyAsync(q, p, keep, hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (p) (void)hipFree(p);
        p = q;
        cap = bytes;
        return e;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

struct ssg_handle {
    int device = 0;
    hipStream_t stream = nullptr;
    int num_cu = 0;
    std::mutex mu;
    // database
    int32_t dim = 0, dp = 0, metric = 0;
    int64_t n = 0, n_pad = 0;
    DevBuf X, xnorm;
    DevBuf Xh;  // BF16 copy of X, made on first use of the BF16 form
    float xnmax = 0.f;
    // graph: node[i] = {start, deg} into the adj pool (0-based ids), hints 1-based
    bool has_graph = false;
    DevBuf node, adj, hints;
    int64_t nnz = 0, nhints = 0, pool_used = 0;
    int has_dups = 0;
    std::vector<std::vector<uint32_t>> lists;  // host copy of the lists (0-based)
    std::vector<uint32_t> hints_host;          // 1-based
    ssg_beam_params algo{4, 1.0f, 1000000, 32};
    // index! state kept between calls, so append_items! continues the insertion loop
    // (src/searchgraph/insertions.jl:27-34, index! from length(index) + 1): the build
    // parameters, objects inserted so far, the draws' generator state and each list's
    // segment of the device pool (start, capacity, length on the device)
    bool b_valid = false;
    ssg_build_params b_params{};
    int64_t b_len = 0;
    uint64_t b_rng = 0;
    std::vector<uint32_t> b_seg_start, b_seg_cap, b_dev_len;
    // beam scratch
    DevBuf vis, epochs, counter;
    int64_t vis_stride = 0;
    int nslots = 0;
    // launch shape of the beam searches (ssg_set_launch / ssg_optimize_launch): persistent
    // wave slots (each with its visited slot) per CU, and gather rounds per batch
    int32_t slots_per_cu = 32;
    int32_t rounds = 8;
    // LDS visited-table buckets per wave (8 tags each; -1 = as many as leave the
    // occupancy unchanged, 0 = visited bytes in HBM only)
    int32_t vbuckets = -1;
    // exhaustive work
    DevBuf qpad, qnorm, cand_d, cand_i, slice_tau, flags, part_i, part_d, xList, seed_tau;
    // host-call staging
    DevBuf sQ, sIds, sD, sCd, sCb, sK, sB, sH, sHc, sO;
    DevBuf cIds, cD;  // hsp_queries outputs (host entry point)
    DevBuf sV;        // per-query visited ids (host entry point)
    // builder scratch
    DevBuf bIds, bD, bOut, bCnt, bPairs, bNodeUpd;
    // learned-search scratch: hints, bucket, confidence, hint counts, beam sizes
    DevBuf lH, lB, lC, lN, lS, lO, lCnt;
    // query layout: the installed graph and the vectors renumbered in breadth-first
    // order from the hints (neighbours get nearby numbers, so their visited bytes,
    // list entries and rows share cache lines and DRAM pages), with each list entry
    // carrying its target's {start, degree}.  Built on the first query search after the
    // graph or database changes; searchbatch! uses it, index!/allknn! do not.
    // precision form of exhaustive k-NN: encoded rows (made on first use in the format
    // xe_fmt: BF16 rows of 32 * ceil(dim / 32) elements, or FP8 rows of 128 bytes), encoded
    // queries, per-(query, slice) candidate lists and their lengths; prec_slices > 0 forces
    // the number of DB slices
    DevBuf Xe, pQe, pList, pListN, pTau0, pFlags, pQf;
    DevBuf pFlags2, pTauQ, pTauF, pIdx3;  // the seeded redo: its flags, thresholds, last queries
    DevBuf Xs, xns;  // rows of every 16th tile of Xe, packed, and their norms (sample pass)
    DevBuf Xs2, xns2, pTauSeed;  // every 16th tile of Xs (seed pass) and its thresholds
    DevBuf xAmax;    // FP8 max |value| bits: [0] and [1] the rows, [2] the last queries
    int32_t xe_fmt = 0;  // SSG_BF16 or SSG_FP8 while Xe holds rows of that format
    int32_t prec_slices = 0;
    int32_t prec_sample = 0;  // 0: sampled thresholds when slices are long, -1: never, T > 0: always, rank T
    bool reorder = true;
    bool q_ready = false, q_bf16 = false;
    DevBuf qnode, qadj, qadjn, qperm, qipos, qhints, Xq, Xqh;
};

// Invalidate the query layout (after any change of graph or database).
inline void ssg_query_layout_stale(ssg_handle* h)
{
    h->q_ready = false;
    h->q_bf16 = false;
}

int ssg_fail(int code, const std::string& msg);
int ssg_hipfail(hipError_t e, const char* what);

#define HIPCHK(call, what)                                      \
    do {                                                        \
        hipError_t _e = (call);                                 \
        if (_e != hipSuccess) return ssg_hipfail(_e, what);     \
    } while (0)

// Visited slots, their epochs and the work counter of a beam-search launch held in
// memory the caller provides (the no-allocation entry point): nslots slots of
// vis_stride bytes at vis, nslots epoch bytes and a 4-byte counter, all cleared by the
// launch itself before the search.
struct BeamScratch {
    uint8_t* vis;
    uint8_t* epochs;
    unsigned int* counter;
    int64_t vis_stride;
    int nslots;
};

// Beam search over the handle's graph with the first n_eff objects as the index
// (n_eff = h->n outside graph construction).  Queries are rows of dQ with qdim
// floats each.  With `ext` the launch uses that scratch and allocates nothing: it
// searches the query layout only if it is already built (the results do not depend on
// the layout) and the BF16 form needs BF16 rows made by an earlier BF16 search.
int ssg_beam_run(ssg_handle* h, const float* dQ, int32_t qdim, int64_t m, int32_t k,
                 int32_t kstride, const ssg_beam_params* bp, const ssg_query_opts* qo,
                 uint32_t* dids, float* ddists, int64_t* dcd, int64_t* dcb, hipStream_t s,
                 bool allknn, int64_t n_eff, bool internal = false,
                 const BeamScratch* ext = nullptr);

// Resolves the device address of the no-allocation scratch on `device` (current); called
// once per device by ssg_create.
int ssg_noalloc_init(int device);

// Exhaustive k-NN over the first n_eff objects.  Either dQ (m rows of h->dim
// floats) or a prepared (qpad, qnorm) pair with m_pad = roundup(m, 64) rows of dp
// floats is given.
int ssg_exh_run(ssg_handle* h, const float* dQ, const float* qpad, const float* qnorm, int64_t m,
                int32_t k, int32_t kstride, uint32_t* dids, float* ddists, hipStream_t s,
                int64_t n_eff);

// Install a frozen CSR graph (0-based lists) on the device.
int ssg_install_lists(ssg_handle* h);
extern "C" {
// Append m rows to the database: host rows of dim floats (column-major (dim, m)), or
// device rows already padded to dp (dev = true).  Encoded copies are dropped.
int ssg_append_rows(ssg_handle* h, const float* X, int64_t m, bool dev);
// index! over the objects after h->b_len (the whole database on the first call).
int ssg_build_continue(ssg_handle* h);
// Start an index! state with parameters p over the handle's (possibly empty) database.
int ssg_graph_begin(ssg_handle* h, const ssg_build_params* p);
}
