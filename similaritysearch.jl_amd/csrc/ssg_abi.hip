// ssg_abi.hip -- the C ABI of libssg_hip.so (declared in include/ssg.h).
//
// A handle is one GPU index: the device copy of a MatrixDatabase, optionally a
// frozen SearchGraph (CSR), and the scratch the kernels need (visited epoch
// bytes per wave slot, exhaustive work buffers, staging for the host-pointer
// entry points).  Host-pointer calls copy in, run the _dev path on the handle's
// stream and copy out; _dev calls only enqueue (plus, for the exhaustive path,
// one small read of the certificate flag count).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ssg.h"
#include "ssg_kernels.h"

using namespace ssg;

namespace {

thread_local std::string g_err;

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
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
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

}  // namespace

struct ssg_handle {
    int device = 0;
    hipStream_t stream = nullptr;
    int num_cu = 0;
    std::mutex mu;
    // database
    int32_t dim = 0, dp = 0, metric = 0;
    int64_t n = 0, n_pad = 0;
    DevBuf X, xnorm;
    float xnmax = 0.f;
    // graph
    bool has_graph = false;
    DevBuf offs, adj, hints;
    int64_t nnz = 0, nhints = 0;
    int has_dups = 0;
    // beam scratch
    DevBuf vis, epochs, counter;
    int64_t vis_stride = 0;
    int nslots = 0;
    // exhaustive work
    DevBuf qpad, qnorm, cand_d, cand_i, slice_tau, flags, part_i, part_d;
    // host-call staging
    DevBuf sQ, sIds, sD, sCd, sCb, sK, sB, sH, sHc;
};

static int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

static int hipfail(hipError_t e, const char* what)
{
    return fail(e == hipErrorOutOfMemory ? SSG_E_NOMEM : SSG_E_HIP,
                std::string(what) + ": " + hipGetErrorString(e));
}

#define HIPCHK(call, what)                          \
    do {                                            \
        hipError_t _e = (call);                     \
        if (_e != hipSuccess) return hipfail(_e, what); \
    } while (0)

extern "C" {

const char* ssg_version(void) { return "ssg_hip 0.1 gfx950"; }

const char* ssg_last_error(const ssg_handle*) { return g_err.c_str(); }

int ssg_create(ssg_handle** out, int device)
{
    if (!out) return fail(SSG_E_ARG, "ssg_create: null output pointer");
    *out = nullptr;
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev), "hipGetDeviceCount");
    if (device < 0 || device >= ndev) return fail(SSG_E_ARG, "ssg_create: no such device");
    HIPCHK(hipSetDevice(device), "hipSetDevice");
    ssg_handle* h = new ssg_handle();
    h->device = device;
    hipDeviceProp_t prop;
    hipError_t e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) {
        delete h;
        return hipfail(e, "hipGetDeviceProperties");
    }
    h->num_cu = prop.multiProcessorCount;
    e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete h;
        return hipfail(e, "hipStreamCreate");
    }
    *out = h;
    return SSG_OK;
}

int ssg_destroy(ssg_handle* h)
{
    if (!h) return SSG_OK;
    (void)hipSetDevice(h->device);
    (void)hipStreamSynchronize(h->stream);
    DevBuf* bufs[] = {&h->X, &h->xnorm, &h->offs, &h->adj, &h->hints, &h->vis, &h->epochs,
                      &h->counter, &h->qpad, &h->qnorm, &h->cand_d, &h->cand_i, &h->slice_tau,
                      &h->flags, &h->part_i, &h->part_d, &h->sQ, &h->sIds, &h->sD, &h->sCd,
                      &h->sCb, &h->sK, &h->sB, &h->sH, &h->sHc};
    for (DevBuf* b : bufs) b->release();
    (void)hipStreamDestroy(h->stream);
    delete h;
    return SSG_OK;
}

int ssg_upload_db(ssg_handle* h, const float* X, int32_t dim, int64_t n, int32_t metric)
{
    if (!h) return fail(SSG_E_ARG, "null handle");
    std::lock_guard<std::mutex> lk(h->mu);
    if (!X && n > 0) return fail(SSG_E_ARG, "ssg_upload_db: null matrix");
    if (dim <= 0 || n < 0) return fail(SSG_E_ARG, "ssg_upload_db: bad shape");
    if (metric < SSG_SQL2 || metric > SSG_NORMCOS) return fail(SSG_E_ARG, "ssg_upload_db: unknown distance");
    const int32_t dp = (dim + 15) / 16 * 16;
    const int v = dp / 16;
    if (!(v <= 8 || v == 12 || v == 16))
        return fail(SSG_E_LIMIT, "ssg_upload_db: dim must pad to 16..128, 192 or 256");
    if (n >= (int64_t)0xffffffffLL) return fail(SSG_E_LIMIT, "ssg_upload_db: n must fit UInt32 ids");
    HIPCHK(hipSetDevice(h->device), "hipSetDevice");
    h->dim = dim;
    h->dp = dp;
    h->metric = metric;
    h->n = n;
    h->n_pad = (n + 63) / 64 * 64;
    h->has_graph = false;
    const int64_t np = std::max<int64_t>(h->n_pad, 64);
    HIPCHK(h->X.ensure((size_t)np * dp * sizeof(float)), "alloc X");
    HIPCHK(h->xnorm.ensure((size_t)np * sizeof(float)), "alloc xnorm");
    HIPCHK(launch_fill_f32(h->X.as<float>(), np * dp, 0.f, h->stream), "fill X");
    HIPCHK(launch_fill_f32(h->xnorm.as<float>(), np, 0.f, h->stream), "fill xnorm");
    if (n > 0) {
        DevBuf tmp;
        const int64_t chunk = std::max<int64_t>(1, (int64_t)(256 << 20) / ((int64_t)dim * 4));
        HIPCHK(tmp.ensure((size_t)std::min(chunk, n) * dim * sizeof(float)), "alloc staging");
        for (int64_t r = 0; r < n; r += chunk) {
            const int64_t rows = std::min(chunk, n - r);
            hipError_t e = hipMemcpyAsync(tmp.p, X + r * dim, (size_t)rows * dim * sizeof(float),
                                          hipMemcpyHostToDevice, h->stream);
            if (e == hipSuccess)
                e = launch_pad_rows(tmp.as<float>(), rows, dim, h->X.as<float>() + r * dp, dp,
                                    h->xnorm.as<float>() + r, h->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
            if (e != hipSuccess) {
                tmp.release();
                return hipfail(e, "upload X");
            }
        }
        tmp.release();
        std::vector<float> nr((size_t)n);
        HIPCHK(hipMemcpy(nr.data(), h->xnorm.p, (size_t)n * sizeof(float), hipMemcpyDeviceToHost),
               "read norms");
        float mx = 0.f;
        for (float f : nr) mx = std::max(mx, f);
        h->xnmax = mx;
    }
    HIPCHK(hipStreamSynchronize(h->stream), "sync");
    return SSG_OK;
}

int ssg_upload_graph(ssg_handle* h, const int64_t* offset, int64_t n, const uint32_t* end_point,
                     int64_t nnz, const uint32_t* hints, int64_t nhints)
{
    if (!h) return fail(SSG_E_ARG, "null handle");
    std::lock_guard<std::mutex> lk(h->mu);
    if (n != h->n) return fail(SSG_E_ARG, "ssg_upload_graph: graph size differs from the database");
    if (nnz < 0 || (nnz > 0 && !end_point) || (n > 0 && !offset) || nhints < 0 || (nhints > 0 && !hints))
        return fail(SSG_E_ARG, "ssg_upload_graph: bad arguments");
    if (nnz >= (int64_t)0xffffffffLL) return fail(SSG_E_LIMIT, "ssg_upload_graph: nnz must be < 2^32");
    std::vector<int64_t> offs((size_t)n + 1);
    offs[0] = 0;
    for (int64_t i = 0; i < n; i++) {
        if (offset[i] < offs[i]) return fail(SSG_E_ARG, "ssg_upload_graph: offsets must be non-decreasing");
        offs[i + 1] = offset[i];
    }
    if (offs[n] != nnz) return fail(SSG_E_ARG, "ssg_upload_graph: offset[n] must equal nnz");
    std::vector<uint32_t> adj((size_t)std::max<int64_t>(nnz, 1));
    int dups = 0;
    std::vector<uint32_t> tmp;
    for (int64_t i = 0; i < n; i++) {
        const int64_t s = offs[i], e = offs[i + 1];
        for (int64_t t = s; t < e; t++) {
            const uint32_t v = end_point[t];
            if (v < 1 || (int64_t)v > n) return fail(SSG_E_ARG, "ssg_upload_graph: neighbour id out of range");
            adj[t] = v - 1;
        }
        if (!dups && e - s > 1) {
            tmp.assign(adj.begin() + s, adj.begin() + e);
            std::sort(tmp.begin(), tmp.end());
            if (std::adjacent_find(tmp.begin(), tmp.end()) != tmp.end()) dups = 1;
        }
    }
    for (int64_t i = 0; i < nhints; i++)
        if (hints[i] < 1 || (int64_t)hints[i] > n) return fail(SSG_E_ARG, "ssg_upload_graph: hint out of range");
    HIPCHK(hipSetDevice(h->device), "hipSetDevice");
    HIPCHK(h->offs.ensure((size_t)(n + 1) * sizeof(int64_t)), "alloc offs");
    HIPCHK(h->adj.ensure((size_t)std::max<int64_t>(nnz, 1) * sizeof(uint32_t)), "alloc adj");
    HIPCHK(h->hints.ensure((size_t)std::max<int64_t>(nhints, 1) * sizeof(uint32_t)), "alloc hints");
    HIPCHK(hipMemcpy(h->offs.p, offs.data(), (size_t)(n + 1) * sizeof(int64_t), hipMemcpyHostToDevice), "copy offs");
    if (nnz > 0) HIPCHK(hipMemcpy(h->adj.p, adj.data(), (size_t)nnz * sizeof(uint32_t), hipMemcpyHostToDevice), "copy adj");
    if (nhints > 0) HIPCHK(hipMemcpy(h->hints.p, hints, (size_t)nhints * sizeof(uint32_t), hipMemcpyHostToDevice), "copy hints");
    h->nnz = nnz;
    h->nhints = nhints;
    h->has_dups = dups;
    h->has_graph = true;
    return SSG_OK;
}

}  // extern "C"

// Visited-byte slots: one per resident wave, n bytes each (rounded to 1 KiB).
static int ensure_vis(ssg_handle* h)
{
    const int64_t stride = (std::max<int64_t>(h->n, 1) + 1023) / 1024 * 1024;
    int want = h->num_cu * 32;
    const int64_t budget = (int64_t)8 << 30;
    if ((int64_t)want * stride > budget) want = (int)std::max<int64_t>(4, budget / stride);
    want = want / 4 * 4;
    if (h->vis.p && h->vis_stride == stride && h->nslots == want) return SSG_OK;
    h->vis.release();
    h->epochs.release();
    HIPCHK(h->vis.ensure((size_t)want * stride), "alloc visited slots");
    HIPCHK(h->epochs.ensure((size_t)want), "alloc epochs");
    HIPCHK(h->counter.ensure(sizeof(unsigned int) * 4), "alloc counter");
    HIPCHK(hipMemset(h->vis.p, 0, (size_t)want * stride), "clear visited slots");
    HIPCHK(hipMemset(h->epochs.p, 0, (size_t)want), "clear epochs");
    h->vis_stride = stride;
    h->nslots = want;
    return SSG_OK;
}

static int beam_common(ssg_handle* h, const float* dQ, int64_t m, int32_t k, int32_t kstride,
                       const ssg_beam_params* bp, const ssg_query_opts* qo, uint32_t* dids,
                       float* ddists, int64_t* dcd, int64_t* dcb, hipStream_t s, bool allknn)
{
    if (!h->X.p) return fail(SSG_E_STATE, "no database uploaded");
    if (!h->has_graph) return fail(SSG_E_STATE, "no graph uploaded");
    if (!bp) return fail(SSG_E_ARG, "null beam parameters");
    if (m <= 0) return fail(SSG_E_ARG, "empty set of queries");
    if (k < 1 || kstride < k || kstride > 64) return fail(SSG_E_LIMIT, "k must satisfy 1 <= k <= kstride <= 64");
    if (bp->bsize < 1 || bp->beamcap < 1 || bp->beamcap > 256)
        return fail(SSG_E_ARG, "bsize >= 1 and 1 <= beamcap <= 256 required");
    if (!(bp->delta >= 0.f) || !std::isfinite(bp->delta)) return fail(SSG_E_ARG, "delta must be finite and >= 0");
    if (bp->maxvisits < 0) return fail(SSG_E_ARG, "maxvisits must be >= 0");
    if (qo && qo->hints && (!qo->nhints || qo->hint_stride < 0)) return fail(SSG_E_ARG, "per-query hints need counts");
    int rc = ensure_vis(h);
    if (rc) return rc;
    BeamArgs a{};
    a.X = h->X.as<float>();
    a.n = h->n;
    a.dp = h->dp;
    a.metric = h->metric;
    a.offs = h->offs.as<int64_t>();
    a.adj = h->adj.as<uint32_t>();
    a.hints = h->hints.as<uint32_t>();
    a.nhints = (int32_t)h->nhints;
    a.log2n = (int32_t)std::ceil(std::log((double)(1 + h->n)) / std::log(2.0));
    a.Q = dQ;
    a.qdim = h->dim;
    a.m = m;
    a.k = k;
    a.kq = qo ? qo->k : nullptr;
    a.kstride = kstride;
    a.bsize = bp->bsize;
    a.bq = qo ? qo->bsize : nullptr;
    a.delta = bp->delta;
    a.maxvisits = bp->maxvisits;
    a.beamcap = bp->beamcap;
    a.qhints = qo ? qo->hints : nullptr;
    a.qhcount = qo ? qo->nhints : nullptr;
    a.qhstride = qo ? qo->hint_stride : 0;
    a.ids = dids;
    a.dists = ddists;
    a.costd = dcd;
    a.costb = dcb;
    a.vis = h->vis.as<uint8_t>();
    a.epochs = h->epochs.as<uint8_t>();
    a.vis_stride = h->vis_stride;
    a.counter = h->counter.as<unsigned int>();
    a.has_dups = h->has_dups;
    a.allknn = allknn ? 1 : 0;
    int slots = (int)std::min<int64_t>(h->nslots, (m + 3) / 4 * 4);
    HIPCHK(launch_zero_u32(a.counter, 1, s), "zero counter");
    HIPCHK(launch_beam(a, slots, s), "beam kernel launch");
    return SSG_OK;
}

static hipStream_t pick(ssg_handle* h, void* stream)
{
    return stream ? (hipStream_t)stream : h->stream;
}

static int exh_common(ssg_handle* h, const float* dQ, bool q_is_db, int64_t m, int32_t k,
                      int32_t kstride, uint32_t* dids, float* ddists, hipStream_t s)
{
    if (!h->X.p) return fail(SSG_E_STATE, "no database uploaded");
    if (m <= 0) return fail(SSG_E_ARG, "empty set of queries");
    if (k < 1 || kstride < k) return fail(SSG_E_ARG, "k must satisfy 1 <= k <= kstride");
    if (k > EXH_KC) return fail(SSG_E_LIMIT, "exhaustive k is limited to 32");
    if (h->n == 0) {
        // every slot stays empty
        HIPCHK(launch_fill_f32(ddists, m * kstride, __builtin_inff(), s), "fill");
        HIPCHK(launch_zero_u32(dids, m * kstride, s), "fill");
        return SSG_OK;
    }
    const int64_t m_pad = (m + EXH_QT - 1) / EXH_QT * EXH_QT;
    const int ns = exhaustive_slices(h->n, m, h->num_cu);
    ExhWork w{};
    if (q_is_db) {
        w.qpad = h->X.as<float>();
        w.qnorm = h->xnorm.as<float>();
    } else {
        HIPCHK(h->qpad.ensure((size_t)m_pad * h->dp * sizeof(float)), "alloc qpad");
        HIPCHK(h->qnorm.ensure((size_t)m_pad * sizeof(float)), "alloc qnorm");
        w.qpad = h->qpad.as<float>();
        w.qnorm = h->qnorm.as<float>();
        HIPCHK(launch_fill_f32(w.qpad + m * h->dp, (m_pad - m) * h->dp, 0.f, s), "fill qpad");
        HIPCHK(launch_fill_f32(w.qnorm + m, m_pad - m, 0.f, s), "fill qnorm");
        HIPCHK(launch_pad_rows(dQ, m, h->dim, w.qpad, h->dp, w.qnorm, s), "pad queries");
    }
    HIPCHK(h->cand_d.ensure((size_t)m_pad * ns * EXH_KC * sizeof(float)), "alloc cand");
    HIPCHK(h->cand_i.ensure((size_t)m_pad * ns * EXH_KC * sizeof(uint32_t)), "alloc cand");
    HIPCHK(h->slice_tau.ensure((size_t)m_pad * ns * sizeof(float)), "alloc tau");
    HIPCHK(h->flags.ensure((size_t)(m + 1) * sizeof(unsigned int)), "alloc flags");
    w.cand_d = h->cand_d.as<float>();
    w.cand_i = h->cand_i.as<uint32_t>();
    w.slice_tau = h->slice_tau.as<float>();
    w.flags = h->flags.as<unsigned int>();
    ExhArgs a{};
    a.X = h->X.as<float>();
    a.xnorm = h->xnorm.as<float>();
    a.n = h->n;
    a.dp = h->dp;
    a.metric = h->metric;
    a.Q = dQ;
    a.qdim = h->dim;
    a.m = m;
    a.k = k;
    HIPCHK(launch_zero_u32(w.flags, 1, s), "zero flags");
    // unused columns beyond k stay empty
    if (kstride > k) {
        HIPCHK(launch_fill_f32(ddists, m * kstride, __builtin_inff(), s), "fill");
        HIPCHK(launch_zero_u32(dids, m * kstride, s), "fill");
    }
    ExhArgs a1 = a;
    a1.m = m_pad;
    HIPCHK(exh_launch_stage1(a1, w, ns, s), "exhaustive stage 1");
    HIPCHK(exh_launch_stage2(a, w, h->xnmax, ns, dids, ddists, kstride, s), "exhaustive stage 2");
    unsigned int nf = 0;
    HIPCHK(hipMemcpyAsync(&nf, w.flags, sizeof(unsigned int), hipMemcpyDeviceToHost, s), "read flags");
    HIPCHK(hipStreamSynchronize(s), "sync");
    if (nf > 0) {
        const int nparts = 16, per = 64;
        HIPCHK(h->part_i.ensure((size_t)per * nparts * 4 * k * sizeof(uint32_t)), "alloc rescan");
        HIPCHK(h->part_d.ensure((size_t)per * nparts * 4 * k * sizeof(float)), "alloc rescan");
        for (unsigned int f0 = 0; f0 < nf; f0 += per) {
            const int cnt = (int)std::min<unsigned int>(per, nf - f0);
            HIPCHK(exh_launch_rescan(a, w, (int)f0, cnt, h->part_i.as<uint32_t>(), h->part_d.as<float>(),
                                     nparts, dids, ddists, kstride, s),
                   "exhaustive rescan");
        }
    }
    return SSG_OK;
}

extern "C" {

int ssg_exhaustive_knn_dev(ssg_handle* h, const float* dQ, int64_t m, int32_t k, uint32_t* dids,
                           float* ddists, void* stream)
{
    if (!h) return fail(SSG_E_ARG, "null handle");
    std::lock_guard<std::mutex> lk(h->mu);
    HIPCHK(hipSetDevice(h->device), "hipSetDevice");
    if (!dQ || !dids || !ddists) return fail(SSG_E_ARG, "null buffer");
    return exh_common(h, dQ, false, m, k, k, dids, ddists, pick(h, stream));
}

int ssg_exhaustive_knn(ssg_handle* h, const float* Q, int64_t m, int32_t k, uint32_t* ids,
                       float* dists)
{
    if (!h) return fail(SSG_E_ARG, "null handle");
    std::lock_guard<std::mutex> lk(h->mu);
    HIPCHK(hipSetDevice(h->device), "hipSetDevice");
    if (!Q || !ids || !dists) return fail(SSG_E_ARG, "null buffer");
    if (m <= 0) return fail(SSG_E_ARG, "empty set of queries");
    if (k < 1) return fail(SSG_E_ARG, "k must be >= 1");
    hipStream_t s = h->stream;
    HIPCHK(h->sQ.ensure((size_t)m * h->dim * sizeof(float)), "alloc");
    HIPCHK(h->sIds.ensure((size_t)m * k * sizeof(uint32_t)), "alloc");
    HIPCHK(h->sD.ensure((size_t)m * k * sizeof(float)), "alloc");
    HIPCHK(hipMemcpyAsync(h->sQ.p, Q, (size_t)m * h->dim * sizeof(float), hipMemcpyHostToDevice, s), "copy Q");
    int rc = exh_common(h, h->sQ.as<float>(), false, m, k, k, h->sIds.as<uint32_t>(), h->sD.as<float>(), s);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(ids, h->sIds.p, (size_t)m * k * sizeof(uint32_t), hipMemcpyDeviceToHost, s), "copy ids");
    HIPCHK(hipMemcpyAsync(dists, h->sD.p, (size_t)m * k * sizeof(float), hipMemcpyDeviceToHost, s), "copy dists");
    HIPCHK(hipStreamSynchronize(s), "sync");
    return SSG_OK;
}

int ssg_allknn_exhaustive(ssg_handle* h, int32_t k, uint32_t* ids, float* dists)
{
    if (!h) return fail(SSG_E_ARG, "null handle");
    std::lock_guard<std::mutex> lk(h->mu);
    HIPCHK(hipSetDevice(h->device), "hipSetDevice");
    if (!ids || !dists) return fail(SSG_E_ARG, "null buffer");
    const int64_t n = h->n;
    if (n <= 0) return fail(SSG_E_ARG, "allknn needs a non-empty database");
    if (k < 1 || k > n) return fail(SSG_E_ARG, "allknn needs 0 < k <= n");
    hipStream_t s = h->stream;
    HIPCHK(h->sIds.ensure((size_t)n * k * sizeof(uint32_t)), "alloc");
    HIPCHK(h->sD.ensure((size_t)n * k * sizeof(float)), "alloc");
    int rc = exh_common(h, nullptr, true, n, k, k, h->sIds.as<uint32_t>(), h->sD.as<float>(), s);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(ids, h->sIds.p, (size_t)n * k * sizeof(uint32_t), hipMemcpyDeviceToHost, s), "copy ids");
    HIPCHK(hipMemcpyAsync(dists, h->sD.p, (size_t)n * k * sizeof(float), hipMemcpyDeviceToHost, s), "copy dists");
    HIPCHK(hipStreamSynchronize(s), "sync");
    return SSG_OK;
}

int ssg_beamsearch_dev(ssg_handle* h, const float* dQ, int64_t m, int32_t k, int32_t kstride,
                       const ssg_beam_params* bp, const ssg_query_opts* qo, uint32_t* dids,
                       float* ddists, int64_t* dcd, int64_t* dcb, void* stream)
{
    if (!h) return fail(SSG_E_ARG, "null handle");
    std::lock_guard<std::mutex> lk(h->mu);
    HIPCHK(hipSetDevice(h->device), "hipSetDevice");
    if (!dQ || !dids || !ddists) return fail(SSG_E_ARG, "null buffer");
    return beam_common(h, dQ, m, k, kstride, bp, qo, dids, ddists, dcd, dcb, pick(h, stream), false);
}

int ssg_beamsearch(ssg_handle* h, const float* Q, int64_t m, int32_t k, int32_t kstride,
                   const ssg_beam_params* bp, const ssg_query_opts* qo, uint32_t* ids,
                   float* dists, int64_t* costdists, int64_t* costblocks)
{
    if (!h) return fail(SSG_E_ARG, "null handle");
    std::lock_guard<std::mutex> lk(h->mu);
    HIPCHK(hipSetDevice(h->device), "hipSetDevice");
    if (!Q || !ids || !dists) return fail(SSG_E_ARG, "null buffer");
    if (m <= 0) return fail(SSG_E_ARG, "empty set of queries");
    if (k < 1 || kstride < k || kstride > 64) return fail(SSG_E_LIMIT, "k must satisfy 1 <= k <= kstride <= 64");
    hipStream_t s = h->stream;
    ssg_query_opts dq{};
    if (qo) {
        if (qo->k) {
            for (int64_t j = 0; j < m; j++)
                if (qo->k[j] < 1 || qo->k[j] > kstride) return fail(SSG_E_ARG, "per-query k out of range");
            HIPCHK(h->sK.ensure((size_t)m * 4), "alloc");
            HIPCHK(hipMemcpyAsync(h->sK.p, qo->k, (size_t)m * 4, hipMemcpyHostToDevice, s), "copy k");
            dq.k = h->sK.as<int32_t>();
        }
        if (qo->bsize) {
            for (int64_t j = 0; j < m; j++)
                if (qo->bsize[j] < 1) return fail(SSG_E_ARG, "per-query bsize must be >= 1");
            HIPCHK(h->sB.ensure((size_t)m * 4), "alloc");
            HIPCHK(hipMemcpyAsync(h->sB.p, qo->bsize, (size_t)m * 4, hipMemcpyHostToDevice, s), "copy bsize");
            dq.bsize = h->sB.as<int32_t>();
        }
        if (qo->hints) {
            if (!qo->nhints || qo->hint_stride < 1) return fail(SSG_E_ARG, "per-query hints need counts and a stride");
            for (int64_t j = 0; j < m; j++) {
                if (qo->nhints[j] < 0 || qo->nhints[j] > qo->hint_stride) return fail(SSG_E_ARG, "per-query hint count out of range");
                for (int32_t t = 0; t < qo->nhints[j]; t++) {
                    const uint32_t v = qo->hints[j * qo->hint_stride + t];
                    if (v < 1 || (int64_t)v > h->n) return fail(SSG_E_ARG, "per-query hint out of range");
                }
            }
            HIPCHK(h->sH.ensure((size_t)m * qo->hint_stride * 4), "alloc");
            HIPCHK(h->sHc.ensure((size_t)m * 4), "alloc");
            HIPCHK(hipMemcpyAsync(h->sH.p, qo->hints, (size_t)m * qo->hint_stride * 4, hipMemcpyHostToDevice, s), "copy hints");
            HIPCHK(hipMemcpyAsync(h->sHc.p, qo->nhints, (size_t)m * 4, hipMemcpyHostToDevice, s), "copy hint counts");
            dq.hints = h->sH.as<uint32_t>();
            dq.nhints = h->sHc.as<int32_t>();
            dq.hint_stride = qo->hint_stride;
        }
    }
    HIPCHK(h->sQ.ensure((size_t)m * h->dim * sizeof(float)), "alloc");
    HIPCHK(h->sIds.ensure((size_t)m * kstride * 4), "alloc");
    HIPCHK(h->sD.ensure((size_t)m * kstride * 4), "alloc");
    HIPCHK(h->sCd.ensure((size_t)m * 8), "alloc");
    HIPCHK(h->sCb.ensure((size_t)m * 8), "alloc");
    HIPCHK(hipMemcpyAsync(h->sQ.p, Q, (size_t)m * h->dim * sizeof(float), hipMemcpyHostToDevice, s), "copy Q");
    int rc = beam_common(h, h->sQ.as<float>(), m, k, kstride, bp, qo ? &dq : nullptr, h->sIds.as<uint32_t>(),
                         h->sD.as<float>(), h->sCd.as<int64_t>(), h->sCb.as<int64_t>(), s, false);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(ids, h->sIds.p, (size_t)m * kstride * 4, hipMemcpyDeviceToHost, s), "copy ids");
    HIPCHK(hipMemcpyAsync(dists, h->sD.p, (size_t)m * kstride * 4, hipMemcpyDeviceToHost, s), "copy dists");
    if (costdists) HIPCHK(hipMemcpyAsync(costdists, h->sCd.p, (size_t)m * 8, hipMemcpyDeviceToHost, s), "copy costs");
    if (costblocks) HIPCHK(hipMemcpyAsync(costblocks, h->sCb.p, (size_t)m * 8, hipMemcpyDeviceToHost, s), "copy costs");
    HIPCHK(hipStreamSynchronize(s), "sync");
    return SSG_OK;
}

int ssg_allknn_graph(ssg_handle* h, int32_t k, const ssg_beam_params* bp, uint32_t* ids, float* dists)
{
    if (!h) return fail(SSG_E_ARG, "null handle");
    std::lock_guard<std::mutex> lk(h->mu);
    HIPCHK(hipSetDevice(h->device), "hipSetDevice");
    if (!ids || !dists) return fail(SSG_E_ARG, "null buffer");
    const int64_t n = h->n;
    if (n <= 0) return fail(SSG_E_ARG, "allknn needs a non-empty database");
    if (k < 1 || k > n) return fail(SSG_E_ARG, "allknn needs 0 < k <= n");
    hipStream_t s = h->stream;
    HIPCHK(h->sIds.ensure((size_t)n * k * 4), "alloc");
    HIPCHK(h->sD.ensure((size_t)n * k * 4), "alloc");
    int rc = beam_common(h, nullptr, n, k, k, bp, nullptr, h->sIds.as<uint32_t>(), h->sD.as<float>(),
                         nullptr, nullptr, s, true);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(ids, h->sIds.p, (size_t)n * k * 4, hipMemcpyDeviceToHost, s), "copy ids");
    HIPCHK(hipMemcpyAsync(dists, h->sD.p, (size_t)n * k * 4, hipMemcpyDeviceToHost, s), "copy dists");
    HIPCHK(hipStreamSynchronize(s), "sync");
    return SSG_OK;
}

int ssg_rerank_dev(ssg_handle* h, const float* dQ, int64_t m, int32_t k, uint32_t* dids,
                   float* ddists, void* stream)
{
    if (!h) return fail(SSG_E_ARG, "null handle");
    std::lock_guard<std::mutex> lk(h->mu);
    HIPCHK(hipSetDevice(h->device), "hipSetDevice");
    if (!h->X.p) return fail(SSG_E_STATE, "no database uploaded");
    if (m <= 0 || k < 1) return fail(SSG_E_ARG, "empty rerank");
    if (k > 64) return fail(SSG_E_LIMIT, "rerank columns are limited to 64 rows");
    HIPCHK(launch_rerank(h->X.as<float>(), h->dp, h->metric, dQ, h->dim, m, k, dids, ddists, pick(h, stream)), "rerank");
    return SSG_OK;
}

int ssg_rerank(ssg_handle* h, const float* Q, int64_t m, int32_t k, uint32_t* ids, float* dists)
{
    if (!h) return fail(SSG_E_ARG, "null handle");
    std::lock_guard<std::mutex> lk(h->mu);
    HIPCHK(hipSetDevice(h->device), "hipSetDevice");
    if (!h->X.p) return fail(SSG_E_STATE, "no database uploaded");
    if (!Q || !ids || !dists) return fail(SSG_E_ARG, "null buffer");
    if (m <= 0 || k < 1) return fail(SSG_E_ARG, "empty rerank");
    if (k > 64) return fail(SSG_E_LIMIT, "rerank columns are limited to 64 rows");
    for (int64_t j = 0; j < m; j++)
        for (int32_t t = 0; t < k; t++) {
            const uint32_t v = ids[j * k + t];
            if (v == 0) break;
            if ((int64_t)v > h->n) return fail(SSG_E_ARG, "rerank id out of range");
        }
    hipStream_t s = h->stream;
    HIPCHK(h->sQ.ensure((size_t)m * h->dim * 4), "alloc");
    HIPCHK(h->sIds.ensure((size_t)m * k * 4), "alloc");
    HIPCHK(h->sD.ensure((size_t)m * k * 4), "alloc");
    HIPCHK(hipMemcpyAsync(h->sQ.p, Q, (size_t)m * h->dim * 4, hipMemcpyHostToDevice, s), "copy");
    HIPCHK(hipMemcpyAsync(h->sIds.p, ids, (size_t)m * k * 4, hipMemcpyHostToDevice, s), "copy");
    HIPCHK(hipMemcpyAsync(h->sD.p, dists, (size_t)m * k * 4, hipMemcpyHostToDevice, s), "copy");
    HIPCHK(launch_rerank(h->X.as<float>(), h->dp, h->metric, h->sQ.as<float>(), h->dim, m, k,
                         h->sIds.as<uint32_t>(), h->sD.as<float>(), s), "rerank");
    HIPCHK(hipMemcpyAsync(ids, h->sIds.p, (size_t)m * k * 4, hipMemcpyDeviceToHost, s), "copy");
    HIPCHK(hipMemcpyAsync(dists, h->sD.p, (size_t)m * k * 4, hipMemcpyDeviceToHost, s), "copy");
    HIPCHK(hipStreamSynchronize(s), "sync");
    return SSG_OK;
}

}  // extern "C"
