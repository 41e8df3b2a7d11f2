// ssg_abi.hip -- the C ABI of libssg_hip.so (declared in include/ssg.h).
//
// A handle is one GPU index: the device copy of a MatrixDatabase, optionally a
// SearchGraph (node table + adjacency pool, plus a host copy of the lists), and
// the scratch the kernels need (visited epoch bytes per wave slot, exhaustive
// work buffers, staging for the host-pointer entry points).  Host-pointer calls
// copy in, run on the handle's stream and copy out; _dev calls only enqueue
// (plus, for the exhaustive path, one small read of the certificate flag count).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "ssg_handle.h"

using namespace ssg;

static thread_local std::string g_err;

int ssg_fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

int ssg_hipfail(hipError_t e, const char* what)
{
    return ssg_fail(e == hipErrorOutOfMemory ? SSG_E_NOMEM : SSG_E_HIP,
                    std::string(what) + ": " + hipGetErrorString(e));
}

// Visited-byte slots: one per resident wave, n bytes each (rounded to 1 KiB).  They
// are cleared on stream s, the stream of the launch that follows (a memset on the
// null stream would not be ordered before work on a non-blocking stream).
static int ensure_vis(ssg_handle* h, hipStream_t s)
{
    const int64_t stride = (std::max<int64_t>(h->n, 1) + 1023) / 1024 * 1024;
    int want = h->num_cu * h->slots_per_cu;
    const int64_t budget = (int64_t)8 << 30;
    if ((int64_t)want * stride > budget) want = (int)std::max<int64_t>(4, budget / stride);
    want = want / 4 * 4;
    if (h->vis.p && h->vis_stride == stride && h->nslots == want) return SSG_OK;
    h->vis.release();
    h->epochs.release();
    HIPCHK(h->vis.ensure((size_t)want * stride), "alloc visited slots");
    HIPCHK(h->epochs.ensure((size_t)want), "alloc epochs");
    HIPCHK(h->counter.ensure(sizeof(unsigned int) * 4), "alloc counter");
    HIPCHK(launch_zero_u32(h->vis.as<unsigned int>(), (int64_t)want * stride / 4, s),
           "clear visited slots");
    HIPCHK(launch_zero_u32(h->epochs.as<unsigned int>(), ((int64_t)want + 3) / 4, s), "clear epochs");
    h->vis_stride = stride;
    h->nslots = want;
    return SSG_OK;
}

// Renumbering of the installed graph into cells of about 256 objects: a multi-source
// breadth-first search from n/256 evenly spaced seeds assigns each object to the first
// seed that reaches it (objects reached by none start cells of their own), and objects
// are numbered cell by cell in discovery order.  A search visits a spatially compact
// set of objects, so its visited bytes, list entries and rows fall into a few cells.
// Then the renumbered CSR with per-entry target {start, degree} and the vectors in the
// new order (FP32, plus BF16 when asked).
static void cell_order(const std::vector<std::vector<uint32_t>>& lists, int64_t n,
                       std::vector<uint32_t>& order, std::vector<uint32_t>& ipos)
{
    const int64_t S = std::max<int64_t>(1, n / 256);
    std::vector<uint32_t> cell((size_t)n, 0xffffffffu), disc;
    disc.reserve((size_t)n);
    uint32_t ncell = 0;
    auto run = [&](size_t head) {
        while (head < disc.size()) {
            const uint32_t u = disc[head++];
            for (uint32_t v : lists[u])
                if (cell[v] == 0xffffffffu) {
                    cell[v] = cell[u];
                    disc.push_back(v);
                }
        }
    };
    for (int64_t i = 0; i < S; i++) {
        const uint32_t sd = (uint32_t)(i * n / S);
        if (cell[sd] != 0xffffffffu) continue;
        cell[sd] = ncell++;
        disc.push_back(sd);
    }
    run(0);
    for (int64_t i = 0; i < n; i++)
        if (cell[i] == 0xffffffffu) {
            const size_t head = disc.size();
            cell[i] = ncell++;
            disc.push_back((uint32_t)i);
            run(head);
        }
    std::vector<uint32_t> start((size_t)ncell + 1, 0);
    for (int64_t i = 0; i < n; i++) start[cell[i] + 1]++;
    for (uint32_t c = 0; c < ncell; c++) start[c + 1] += start[c];
    order.assign((size_t)n, 0);
    ipos.assign((size_t)n, 0);
    for (uint32_t u : disc) {
        const uint32_t p = start[cell[u]]++;
        order[p] = u;
        ipos[u] = p;
    }
}

static int ensure_query_layout(ssg_handle* h, hipStream_t s, bool bf16)
{
    if (!h->q_ready) {
        HIPCHK(hipDeviceSynchronize(), "sync");  // no launch may still read the old layout
        const int64_t n = h->n;
        const auto& lists = h->lists;
        std::vector<uint32_t> order, ipos;
        cell_order(lists, n, order, ipos);
        std::vector<uint2> node((size_t)n);
        std::vector<uint32_t> adj((size_t)std::max<int64_t>(h->nnz, 1));
        std::vector<uint2> adjn((size_t)std::max<int64_t>(h->nnz, 1));
        int64_t pos = 0;
        for (int64_t i = 0; i < n; i++) {
            const auto& L = lists[order[i]];
            node[i] = make_uint2((uint32_t)pos, (uint32_t)L.size());
            for (uint32_t v : L) adj[(size_t)pos++] = ipos[v];
        }
        if (pos != h->nnz) return ssg_fail(SSG_E_STATE, "query layout: lists and nnz disagree");
        for (int64_t e = 0; e < pos; e++) adjn[(size_t)e] = node[adj[(size_t)e]];
        std::vector<uint32_t> hints(h->hints_host.size());
        for (size_t i = 0; i < hints.size(); i++) hints[i] = ipos[h->hints_host[i] - 1u] + 1u;
        HIPCHK(h->qnode.ensure(node.size() * sizeof(uint2)), "alloc");
        HIPCHK(h->qadj.ensure(adj.size() * sizeof(uint32_t)), "alloc");
        HIPCHK(h->qadjn.ensure(adjn.size() * sizeof(uint2)), "alloc");
        HIPCHK(h->qperm.ensure((size_t)n * sizeof(uint32_t)), "alloc");
        HIPCHK(h->qipos.ensure((size_t)n * sizeof(uint32_t)), "alloc");
        HIPCHK(h->qhints.ensure(std::max<size_t>(hints.size(), 1) * sizeof(uint32_t)), "alloc");
        HIPCHK(hipMemcpy(h->qnode.p, node.data(), node.size() * sizeof(uint2), hipMemcpyHostToDevice), "copy");
        HIPCHK(hipMemcpy(h->qadj.p, adj.data(), adj.size() * sizeof(uint32_t), hipMemcpyHostToDevice), "copy");
        HIPCHK(hipMemcpy(h->qadjn.p, adjn.data(), adjn.size() * sizeof(uint2), hipMemcpyHostToDevice), "copy");
        HIPCHK(hipMemcpy(h->qperm.p, order.data(), (size_t)n * sizeof(uint32_t), hipMemcpyHostToDevice), "copy");
        HIPCHK(hipMemcpy(h->qipos.p, ipos.data(), (size_t)n * sizeof(uint32_t), hipMemcpyHostToDevice), "copy");
        if (!hints.empty())
            HIPCHK(hipMemcpy(h->qhints.p, hints.data(), hints.size() * sizeof(uint32_t), hipMemcpyHostToDevice), "copy");
        const int64_t np = std::max<int64_t>(h->n_pad, 64);
        HIPCHK(h->Xq.ensure((size_t)np * h->dp * sizeof(float)), "alloc permuted rows");
        HIPCHK(launch_permute_rows(h->X.as<float>(), h->qperm.as<uint32_t>(), n, np, h->dp, h->Xq.as<float>(), s),
               "permute rows");
        h->q_ready = true;
        h->q_bf16 = false;
    }
    if (bf16 && !h->q_bf16) {
        const int64_t np = std::max<int64_t>(h->n_pad, 64);
        HIPCHK(h->Xqh.ensure((size_t)np * h->dp * 2), "alloc BF16 rows");
        HIPCHK(launch_to_bf16(h->Xq.as<float>(), np * h->dp, h->Xqh.as<uint16_t>(), s), "BF16 rows");
        h->q_bf16 = true;
    }
    return SSG_OK;
}

int ssg_beam_run(ssg_handle* h, const float* dQ, int32_t qdim, int64_t m, int32_t k,
                 int32_t kstride, const ssg_beam_params* bp, const ssg_query_opts* qo,
                 uint32_t* dids, float* ddists, int64_t* dcd, int64_t* dcb, hipStream_t s,
                 bool allknn, int64_t n_eff, bool internal, const BeamScratch* ext)
{
    if (!h->X.p) return ssg_fail(SSG_E_STATE, "no database uploaded");
    if (!h->has_graph) return ssg_fail(SSG_E_STATE, "no graph uploaded");
    if (!bp) return ssg_fail(SSG_E_ARG, "null beam parameters");
    if (m <= 0) return ssg_fail(SSG_E_ARG, "empty set of queries");
    if (k < 1 || kstride < k || kstride > 256)
        return ssg_fail(SSG_E_LIMIT, "k must satisfy 1 <= k <= kstride <= 256");
    if (kstride > 64 && (bp->precision == SSG_BF16 || internal))
        return ssg_fail(SSG_E_LIMIT, "k is limited to 64 in the BF16 form and in index!");
    if (bp->bsize < 1 || bp->beamcap < 1 || bp->beamcap > 256)
        return ssg_fail(SSG_E_ARG, "bsize >= 1 and 1 <= beamcap <= 256 required");
    if (!(bp->delta >= 0.f) || !std::isfinite(bp->delta))
        return ssg_fail(SSG_E_ARG, "delta must be finite and >= 0");
    if (bp->maxvisits < 0) return ssg_fail(SSG_E_ARG, "maxvisits must be >= 0");
    if (bp->precision != SSG_F32 && bp->precision != SSG_BF16)
        return ssg_fail(SSG_E_ARG, "precision must be SSG_F32 or SSG_BF16");
    if (bp->precision == SSG_BF16 && (allknn || internal))
        return ssg_fail(SSG_E_ARG, "the BF16 form serves searchbatch! only");
    if (bp->precision == SSG_BF16 && h->dp > 256)
        return ssg_fail(SSG_E_LIMIT, "the BF16 form is limited to dim <= 256");
    if (bp->expand < 0 || bp->expand > 4) return ssg_fail(SSG_E_ARG, "expand must be in 0..4");
    if (bp->expand > 1 && bp->precision != SSG_BF16)
        return ssg_fail(SSG_E_ARG, "wide steps (expand > 1) belong to the BF16 form");
    // each query reranks max(its own k, rerank) entries; kstride bounds every k
    const int32_t krr = std::max(kstride, bp->rerank);
    if (bp->precision == SSG_BF16 && krr > 64)
        return ssg_fail(SSG_E_LIMIT, "BF16 rerank list is limited to 64");
    int rc = ext ? SSG_OK : ensure_vis(h, s);
    if (rc) return rc;
    bool use_q = h->reorder && !allknn && !internal && n_eff == h->n && h->n >= 64;
    if (ext) {
        // nothing may be allocated: the query layout only if it is already there
        use_q = use_q && h->q_ready && (bp->precision != SSG_BF16 || h->q_bf16);
        if (bp->precision == SSG_BF16 && !use_q && !h->Xh.p)
            return ssg_fail(SSG_E_STATE, "the no-allocation form needs the BF16 rows that an "
                                         "earlier BF16 ssg_beamsearch_dev call on this handle makes");
    }
    if (bp->precision == SSG_BF16 && !use_q && !h->Xh.p) {
        HIPCHK(h->Xh.ensure((size_t)std::max<int64_t>(h->n_pad, 64) * h->dp * 2), "alloc BF16 copy");
        HIPCHK(launch_to_bf16(h->X.as<float>(), std::max<int64_t>(h->n_pad, 64) * h->dp,
                              h->Xh.as<uint16_t>(), s), "BF16 copy");
    }
    BeamArgs a{};
    a.X = h->X.as<float>();
    a.n = n_eff;
    a.dp = h->dp;
    a.metric = h->metric;
    a.node = h->node.as<uint2>();
    a.adj = h->adj.as<uint32_t>();
    a.hints = h->hints.as<uint32_t>();
    a.nhints = (int32_t)h->nhints;
    a.log2n = (int32_t)std::ceil(std::log((double)(1 + n_eff)) / std::log(2.0));
    a.Q = dQ;
    a.qdim = qdim;
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
    a.qvisit = qo ? qo->visited : nullptr;
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
    a.internal = internal ? 1 : 0;
    a.prec = bp->precision == SSG_BF16 ? 1 : 0;
    a.Xh = h->Xh.as<uint16_t>();
    a.krr = std::max(bp->rerank, 0);
    a.rounds = internal ? 8 : h->rounds;
    // LDS visited tables: tags name at most 2^24 objects
    a.vhb = 15;
    while (a.vhb < 24 && ((int64_t)1 << a.vhb) < n_eff) a.vhb++;
    a.vnb = n_eff <= ((int64_t)1 << 24) ? h->vbuckets : 0;
    a.expand = bp->expand > 1 ? bp->expand : 1;
    a.order = qo ? qo->order : nullptr;
    if (use_q) {
        rc = ensure_query_layout(h, s, bp->precision == SSG_BF16);
        if (rc) return rc;
        a.X = h->Xq.as<float>();
        a.Xh = h->Xqh.as<uint16_t>();
        a.node = h->qnode.as<uint2>();
        a.adj = h->qadj.as<uint32_t>();
        a.adjn = h->qadjn.as<uint2>();
        a.hints = h->qhints.as<uint32_t>();
        a.perm = h->qperm.as<uint32_t>();
        a.ipos = h->qipos.as<uint32_t>();
    }
    int slots = (int)std::min<int64_t>(h->nslots, (m + 3) / 4 * 4);
    if (ext) {
        // the scratch keeps whatever the last call on this GPU left: clear the slots used
        slots = (int)std::min<int64_t>(ext->nslots, (m + 3) / 4 * 4);
        a.vis = ext->vis;
        a.epochs = ext->epochs;
        a.counter = ext->counter;
        a.vis_stride = ext->vis_stride;
        HIPCHK(launch_zero_u32(reinterpret_cast<unsigned int*>(ext->vis),
                               (int64_t)slots * ext->vis_stride / 4, s), "clear visited slots");
        HIPCHK(launch_zero_u32(reinterpret_cast<unsigned int*>(ext->epochs), ((int64_t)slots + 3) / 4, s),
               "clear epochs");
    }
    HIPCHK(launch_zero_u32(a.counter, 1, s), "zero counter");
    HIPCHK(launch_beam(a, slots, s), "beam kernel launch");
    return SSG_OK;
}

int ssg_exh_run(ssg_handle* h, const float* dQ, const float* qpad_in, const float* qnorm_in,
                int64_t m, int32_t k, int32_t kstride, uint32_t* dids, float* ddists,
                hipStream_t s, int64_t n_eff)
{
    if (!h->X.p) return ssg_fail(SSG_E_STATE, "no database uploaded");
    if (m <= 0) return ssg_fail(SSG_E_ARG, "empty set of queries");
    if (k < 1 || kstride < k) return ssg_fail(SSG_E_ARG, "k must satisfy 1 <= k <= kstride");
    if (k > EXH_WIDE_KMAX) return ssg_fail(SSG_E_LIMIT, "exhaustive k is limited to 3840");
    if (kstride > k || n_eff == 0) {
        HIPCHK(launch_fill_f32(ddists, m * kstride, __builtin_inff(), s), "fill");
        HIPCHK(launch_zero_u32(dids, m * kstride, s), "fill");
    }
    if (n_eff == 0) return SSG_OK;
    const int64_t m_pad = (m + EXH_QT - 1) / EXH_QT * EXH_QT;
    const int ns = exhaustive_slices(n_eff, m, h->num_cu, h->dp);
    ExhWork w{};
    if (qpad_in) {
        w.qpad = const_cast<float*>(qpad_in);
        w.qnorm = const_cast<float*>(qnorm_in);
    } else {
        HIPCHK(h->qpad.ensure((size_t)m_pad * h->dp * sizeof(float)), "alloc qpad");
        HIPCHK(h->qnorm.ensure((size_t)m_pad * sizeof(float)), "alloc qnorm");
        w.qpad = h->qpad.as<float>();
        w.qnorm = h->qnorm.as<float>();
        HIPCHK(launch_pad_rows(dQ, m, h->dim, w.qpad, h->dp, w.qnorm, s, m_pad), "pad queries");
    }
    if (k > EXH_KC || h->dp > 256) {
        // more than stage 1 keeps per slice, or rows above 256 dims: the one-block-per-query
        // scan
        HIPCHK(launch_exh_wide(h->X.as<float>(), n_eff, h->dp, h->metric, w.qpad, m, k, dids, ddists,
                               kstride, s), "exhaustive k > 32");
        return SSG_OK;
    }
    if (h->dp <= 128) {
        HIPCHK(h->xList.ensure((size_t)m_pad * ns * 512 * sizeof(uint2)), "alloc candidate lists");
        w.list = h->xList.as<uint2>();
    }
    w.m_valid = m;
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
    a.n = n_eff;
    a.dp = h->dp;
    a.metric = h->metric;
    a.Q = dQ;
    a.qdim = h->dim;
    a.m = m;
    a.k = k;
    ExhArgs a1 = a;
    a1.m = m_pad;
    // Short slices spend most of stage 1 filling and compacting candidate lists from a
    // threshold of -Inf.  Then each query's threshold starts at a lower bound of its k-th
    // best score over 2048 rows spread over the database (and so of the k-th best over all
    // rows), less the rounding allowance, and the main pass keeps only rows above it.
    const int64_t rps = ((n_eff + ns - 1) / ns + 63) / 64 * 64;
    if (h->dp <= 128 && n_eff >= 32768 && rps < 32768) {
        HIPCHK(h->seed_tau.ensure((size_t)m_pad * sizeof(float)), "alloc seed");
        HIPCHK(exh_launch_seed_sample(h->X.as<float>(), h->xnorm.as<float>(), h->dp, n_eff, w, m, m_pad, k,
                                      h->metric, h->xnmax, h->seed_tau.as<float>(), s), "seed");
        w.tau0 = h->seed_tau.as<float>();
    }
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
            HIPCHK(exh_launch_rescan(a, w, (int)f0, cnt, h->part_i.as<uint32_t>(),
                                     h->part_d.as<float>(), nparts, dids, ddists, kstride, s),
                   "exhaustive rescan");
        }
    }
    return SSG_OK;
}

// Precision form of exhaustive k-NN over the whole database: queries are dQ (m rows of
// h->dim floats), or the database itself (allknn, dQ == nullptr).
static int ssg_prec_run(ssg_handle* h, const float* dQ, int64_t m, int32_t k, int32_t kstride,
                        int32_t precision, int32_t ncand, uint32_t* dids, float* ddists,
                        uint32_t* dcand_i, float* dcand_d, hipStream_t s)
{
    if (!h->X.p) return ssg_fail(SSG_E_STATE, "no database uploaded");
    if (m <= 0) return ssg_fail(SSG_E_ARG, "empty set of queries");
    if (precision != SSG_BF16 && precision != SSG_FP8)
        return ssg_fail(SSG_E_ARG, "the precision form takes SSG_BF16 or SSG_FP8");
    if (k < 1 || kstride < k) return ssg_fail(SSG_E_ARG, "k must satisfy 1 <= k <= kstride");
    if (ncand < k || ncand > PREC_NCAND_MAX)
        return ssg_fail(SSG_E_LIMIT, "ncand must satisfy k <= ncand <= 256");
    if (h->dp > 128) return ssg_fail(SSG_E_LIMIT, "the precision form needs dim <= 128");
    if ((dcand_i == nullptr) != (dcand_d == nullptr))
        return ssg_fail(SSG_E_ARG, "candidate ids and distances go together");
    const int64_t n = h->n;
    if (kstride > k || n < k) {
        HIPCHK(launch_fill_f32(ddists, m * kstride, __builtin_inff(), s), "fill");
        HIPCHK(launch_zero_u32(dids, m * kstride, s), "fill");
    }
    if (n == 0) {
        if (dcand_i) {
            HIPCHK(launch_fill_f32(dcand_d, m * ncand, __builtin_inff(), s), "fill");
            HIPCHK(launch_zero_u32(dcand_i, m * ncand, s), "fill");
        }
        return SSG_OK;
    }
    // encoded row bytes: BF16 rows of 32 * ceil(dp / 32) values, FP8 rows of 128 values
    const bool f8 = precision == SSG_FP8;
    const int32_t rowb = f8 ? 128 : 2 * ((h->dp + 31) / 32 * 32);
    const int64_t m_pad = (m + PREC_QB - 1) / PREC_QB * PREC_QB;
    const int64_t xe_rows = std::max<int64_t>(h->n_pad, (n + PREC_QB - 1) / PREC_QB * PREC_QB);
    if (h->xe_fmt != precision) {
        h->Xe.release();
        h->Xs.release();
        h->Xs2.release();
        h->xe_fmt = 0;
    }
    if (f8) HIPCHK(h->xAmax.ensure(4 * sizeof(uint32_t)), "alloc scales");
    if (!h->Xe.p) {
        // rows past n_pad are zero (the FP32 matrix holds n_pad rows)
        HIPCHK(h->Xe.ensure((size_t)xe_rows * rowb), "alloc encoded rows");
        HIPCHK(launch_zero_u32(h->Xe.as<unsigned int>(), xe_rows * rowb / 4, s), "clear encoded rows");
        if (f8) {
            uint32_t* am = h->xAmax.as<uint32_t>();
            HIPCHK(launch_zero_u32(am, 2, s), "clear scale");
            HIPCHK(launch_absmax(h->X.as<float>(), n * h->dp, am, 2, s), "row scale");
            HIPCHK(launch_rows_to_fp8(h->X.as<float>(), n, h->dp, am, h->Xe.as<uint8_t>(), rowb, s), "FP8 rows");
        } else {
            HIPCHK(launch_rows_to_bf16(h->X.as<float>(), n, h->dp, h->Xe.as<uint16_t>(), rowb / 2, s),
                   "BF16 rows");
        }
        h->xe_fmt = precision;
    }
    PrecOut o{};
    PrecArgs a{};
    a.fp8 = f8 ? 1 : 0;
    if (dQ) {
        HIPCHK(h->qpad.ensure((size_t)m_pad * h->dp * sizeof(float)), "alloc qpad");
        HIPCHK(h->qnorm.ensure((size_t)m_pad * sizeof(float)), "alloc qnorm");
        HIPCHK(h->pQe.ensure((size_t)m_pad * rowb), "alloc encoded queries");
        float* qpad = h->qpad.as<float>();
        HIPCHK(launch_fill_f32(qpad + m * h->dp, (m_pad - m) * h->dp, 0.f, s), "fill qpad");
        HIPCHK(launch_fill_f32(h->qnorm.as<float>() + m, m_pad - m, 0.f, s), "fill qnorm");
        HIPCHK(launch_pad_rows(dQ, m, h->dim, qpad, h->dp, h->qnorm.as<float>(), s), "pad queries");
        if (f8) {
            uint32_t* am = h->xAmax.as<uint32_t>();
            HIPCHK(launch_zero_u32(am + 2, 1, s), "clear scale");
            HIPCHK(launch_absmax(qpad, m * h->dp, am + 2, 1, s), "query scale");
            HIPCHK(launch_rows_to_fp8(qpad, m_pad, h->dp, am + 2, h->pQe.as<uint8_t>(), rowb, s), "FP8 queries");
            a.f8max = am + 1;  // {rows, queries}
        } else {
            HIPCHK(launch_rows_to_bf16(qpad, m_pad, h->dp, h->pQe.as<uint16_t>(), rowb / 2, s), "BF16 queries");
        }
        a.Qe = h->pQe.as<uint8_t>();
        o.qpad = qpad;
        o.qnorm = h->qnorm.as<float>();
    } else {
        a.Qe = h->Xe.as<uint8_t>();  // xe_rows >= m_pad
        if (f8) a.f8max = h->xAmax.as<uint32_t>();  // {rows, rows}
        o.qpad = h->X.as<float>();
        o.qnorm = h->xnorm.as<float>();
    }
    const int32_t kc = ncand, cap = prec_list_cap(ncand);
    int32_t ns = h->prec_slices > 0 ? h->prec_slices : prec_slices(n, m, h->num_cu, kc);
    ns = std::max(1, std::min<int32_t>(ns, 2048 / kc));
    // Sampled thresholds (below) on long slices; their slices hold whole groups of STEP tiles
    // so that each slice's sample is a contiguous range of the packed sample rows.
    constexpr int32_t STEP = 16;
    const int64_t rps0 = (n + ns - 1) / ns;
    const bool sample = h->prec_sample > 0 || (h->prec_sample == 0 && rps0 >= (int64_t)STEP * 128 * 64);
    // On slices of at least 256 Ki rows, a seed pass over every 16th tile of the sample
    // gives the sample pass its starting thresholds (slices then hold whole groups of
    // STEP^2 tiles, so each slice's seed rows are contiguous too).
    const bool seed = sample && rps0 >= (int64_t)STEP * STEP * 128 * 8;
    const int64_t unit = seed ? (int64_t)STEP * STEP * 128 : (sample ? (int64_t)STEP * 128 : 128);
    const int64_t rps = (rps0 + unit - 1) / unit * unit;
    ns = (int32_t)((n + rps - 1) / rps);
    HIPCHK(h->pList.ensure((size_t)m_pad * ns * cap * sizeof(uint2)), "alloc candidate lists");
    HIPCHK(h->pListN.ensure((size_t)m_pad * ns * sizeof(int32_t)), "alloc candidate counts");
    a.Xe = h->Xe.as<uint8_t>();
    a.xnorm = h->xnorm.as<float>();
    a.X = h->X.as<float>();
    a.n = n;
    a.rowb = rowb;
    a.metric = h->metric;
    a.m_pad = m_pad;
    a.kc = kc;
    a.cap = cap;
    a.rows_per_slice = rps;
    a.nslices = ns;
    a.list = h->pList.as<uint2>();
    a.list_n = h->pListN.as<int32_t>();
    a.tile_step = 1;
    o.m = m;
    o.k = k;
    o.kstride = kstride;
    o.ncand = ncand;
    o.ids = dids;
    o.dists = ddists;
    o.cand_ids = dcand_i;
    o.cand_dists = dcand_d;
    // Sampled thresholds: a pass over every 16th tile of each slice finds, per query, the
    // T-th best sampled score (about rank 16 T of the slice); the main pass drops every row
    // at or below it, which removes most appends, and the rerank stage certifies each
    // query (the ncand-th best candidate must score above the threshold).  Queries it
    // cannot certify are searched again from no threshold.
    // T = kc / 8 puts the threshold near rank 2 kc: a query's kc-th best lies below it
    // unless 16 sampled rows of 128 fall in its kc best (binomial tail 0.6% at kc = 128).
    const int32_t T = h->prec_sample > 0 ? h->prec_sample : std::max(16, kc / 8);
    if (!sample) {
        HIPCHK(prec_launch_stage1(a, s), "precision candidate pass");
        HIPCHK(prec_launch_stage2(a, o, h->dp, s), "precision rerank");
        return SSG_OK;
    }
    HIPCHK(h->pTau0.ensure((size_t)m_pad * ns * sizeof(float)), "alloc thresholds");
    HIPCHK(h->pFlags.ensure((size_t)(m + 1) * sizeof(unsigned int)), "alloc flags");
    unsigned int* flags = h->pFlags.as<unsigned int>();
    const int64_t nsr = prec_sample_rows(n, STEP);
    if (!h->Xs.p) {
        HIPCHK(h->Xs.ensure((size_t)std::max<int64_t>(nsr, 1) * rowb), "alloc sample rows");
        HIPCHK(h->xns.ensure((size_t)std::max<int64_t>(nsr, 1) * sizeof(float)), "alloc sample norms");
        HIPCHK(launch_sample_rows(h->Xe.as<uint16_t>(), h->xnorm.as<float>(), nsr, STEP, rowb / 2,
                                  h->Xs.as<uint16_t>(), h->xns.as<float>(), s), "sample rows");
    }
    PrecArgs as = a;
    as.Xe = h->Xs.as<uint8_t>();
    as.xnorm = h->xns.as<float>();
    as.n = nsr;
    as.rows_per_slice = rps / STEP;
    as.kc = std::min(T, PREC_NCAND_MAX);
    // short lists: the sample pass starts with no threshold, and a list that is compacted
    // after about a hundred appends raises it sooner
    as.cap = as.kc <= 32 ? 128 : prec_list_cap(as.kc);
    as.tile_step = 1;
    as.sample = 1;
    as.tau0 = h->pTau0.as<float>();
    if (seed) {
        // Seed pass: the T2-th best score of every 16th tile of the sample (about rank
        // 16 T2 = 80 of the sample for T = 16) starts the sample pass's threshold, so it
        // appends about 16 T2 rows per query instead of filling its lists from nothing.
        // A query with fewer than T sample rows above its seed gets no threshold from the
        // sample pass (its main pass then starts from none): slower, still exact.
        const int64_t nsr2 = prec_sample_rows(nsr, STEP);
        if (!h->Xs2.p) {
            HIPCHK(h->Xs2.ensure((size_t)nsr2 * rowb), "alloc seed rows");
            HIPCHK(h->xns2.ensure((size_t)nsr2 * sizeof(float)), "alloc seed norms");
            HIPCHK(launch_sample_rows(h->Xs.as<uint16_t>(), h->xns.as<float>(), nsr2, STEP, rowb / 2,
                                      h->Xs2.as<uint16_t>(), h->xns2.as<float>(), s), "seed rows");
        }
        HIPCHK(h->pTauSeed.ensure((size_t)m_pad * ns * sizeof(float)), "alloc seed thresholds");
        PrecArgs a2 = as;
        a2.Xe = h->Xs2.as<uint8_t>();
        a2.xnorm = h->xns2.as<float>();
        a2.n = nsr2;
        a2.rows_per_slice = rps / (STEP * STEP);
        a2.kc = std::max(4, as.kc / 3);
        a2.cap = a2.kc <= 32 ? 128 : prec_list_cap(a2.kc);
        a2.tau0 = h->pTauSeed.as<float>();
        a2.tau_seed = nullptr;
        HIPCHK(prec_launch_stage1(a2, s), "precision seed pass");
        as.tau_seed = a2.tau0;
    }
    HIPCHK(prec_launch_stage1(as, s), "precision sample pass");
    a.tau0 = as.tau0;
    o.flags = flags;
    HIPCHK(launch_zero_u32(flags, 1, s), "zero flags");
    HIPCHK(prec_launch_stage1(a, s), "precision candidate pass");
    HIPCHK(prec_launch_stage2(a, o, h->dp, s), "precision rerank");
    unsigned int nf = 0;
    HIPCHK(hipMemcpyAsync(&nf, flags, sizeof(unsigned int), hipMemcpyDeviceToHost, s), "read flags");
    HIPCHK(hipStreamSynchronize(s), "sync");
    if (nf == 0) return SSG_OK;
    // Redo of the queries in idx[0 .. cnt) (count: the same number on the device), gathered,
    // over enough slices to fill the device.  tau_q (nullable): one starting threshold per
    // gathered query for all of its slices; then the redo certifies its queries into
    // fl_out as the main pass does.
    auto redo = [&](const uint32_t* idx, const unsigned int* count, unsigned int cnt, const float* tau_q,
                    unsigned int* fl_out) -> int {
        const int64_t mf_pad = ((int64_t)cnt + PREC_QB - 1) / PREC_QB * PREC_QB;
        int32_t nsf = (int32_t)std::max<int64_t>(
            1, ((int64_t)h->num_cu * (8 / PREC_W) + mf_pad / PREC_QB - 1) / (mf_pad / PREC_QB));
        nsf = std::max(1, std::min<int32_t>(nsf, 2048 / kc));
        int64_t rpsf = (n + nsf - 1) / nsf;
        rpsf = (rpsf + 127) / 128 * 128;
        nsf = (int32_t)((n + rpsf - 1) / rpsf);
        HIPCHK(h->pQf.ensure((size_t)mf_pad * rowb), "alloc redo queries");
        HIPCHK(h->pList.grow((size_t)mf_pad * nsf * cap * sizeof(uint2), 0, s), "alloc candidate lists");
        HIPCHK(h->pListN.grow((size_t)mf_pad * nsf * sizeof(int32_t), 0, s), "alloc candidate counts");
        HIPCHK(launch_gather_rows_u16(reinterpret_cast<const uint16_t*>(a.Qe), idx, count, mf_pad, rowb / 2,
                                      h->pQf.as<uint16_t>(), s), "gather");
        PrecArgs af = a;
        af.Qe = h->pQf.as<uint8_t>();
        af.m_pad = mf_pad;
        af.rows_per_slice = rpsf;
        af.nslices = nsf;
        af.list = h->pList.as<uint2>();
        af.list_n = h->pListN.as<int32_t>();
        af.tau0 = nullptr;
        if (tau_q) {
            HIPCHK(h->pTauF.ensure((size_t)mf_pad * nsf * sizeof(float)), "alloc redo thresholds");
            HIPCHK(launch_spread_f32(tau_q, cnt, mf_pad, nsf, h->pTauF.as<float>(), s), "redo thresholds");
            af.tau0 = h->pTauF.as<float>();
        }
        PrecOut of = o;
        of.m = cnt;
        of.qmap = idx;
        of.flags = fl_out;
        HIPCHK(prec_launch_stage1(af, s), "precision candidate pass (redo)");
        HIPCHK(prec_launch_stage2(af, of, h->dp, s), "precision rerank (redo)");
        return SSG_OK;
    };
    if (!seed) return redo(flags + 1, flags, nf, nullptr, nullptr);
    // With seed thresholds (about rank 16 * 16 * T2 of the slice, far below the 128th best
    // of all but a few queries in a million), the redo of each query starts at the largest
    // of its slices' seeds instead of nothing: a fraction of the appends.  The queries that
    // redo cannot certify are searched once more from no threshold.
    HIPCHK(h->pFlags2.ensure((size_t)(nf + 1) * sizeof(unsigned int)), "alloc redo flags");
    HIPCHK(h->pTauQ.ensure((size_t)nf * sizeof(float)), "alloc redo thresholds");
    unsigned int* flags2 = h->pFlags2.as<unsigned int>();
    HIPCHK(launch_max_slices_f32(h->pTauSeed.as<float>(), ns, flags + 1, nf, h->pTauQ.as<float>(), s), "seeds");
    HIPCHK(launch_zero_u32(flags2, 1, s), "zero flags");
    if (int rc = redo(flags + 1, flags, nf, h->pTauQ.as<float>(), flags2)) return rc;
    unsigned int nf2 = 0;
    HIPCHK(hipMemcpyAsync(&nf2, flags2, sizeof(unsigned int), hipMemcpyDeviceToHost, s), "read flags");
    HIPCHK(hipStreamSynchronize(s), "sync");
    if (nf2 == 0) return SSG_OK;
    // flags2 holds positions among the redone queries: map them to query rows
    HIPCHK(h->pIdx3.ensure((size_t)nf2 * sizeof(uint32_t)), "alloc last queries");
    HIPCHK(launch_compose_u32(flags + 1, flags2 + 1, nf2, h->pIdx3.as<uint32_t>(), s), "compose");
    return redo(h->pIdx3.as<uint32_t>(), flags2, nf2, nullptr, nullptr);
}

// Frozen lists -> node table + packed pool on the device.
int ssg_install_lists(ssg_handle* h)
{
    HIPCHK(hipStreamSynchronize(h->stream), "sync");  // no launch may still read the old graph
    const int64_t n = (int64_t)h->lists.size();
    std::vector<uint2> node((size_t)std::max<int64_t>(n, 1));
    int64_t nnz = 0;
    for (int64_t i = 0; i < n; i++) nnz += (int64_t)h->lists[i].size();
    if (nnz >= (int64_t)0xffffffffLL) return ssg_fail(SSG_E_LIMIT, "graph needs nnz < 2^32");
    std::vector<uint32_t> pool((size_t)std::max<int64_t>(nnz, 1));
    int64_t pos = 0;
    int dups = 0;
    std::vector<uint32_t> tmp;
    for (int64_t i = 0; i < n; i++) {
        const auto& L = h->lists[i];
        node[i] = make_uint2((uint32_t)pos, (uint32_t)L.size());
        std::copy(L.begin(), L.end(), pool.begin() + pos);
        pos += (int64_t)L.size();
        if (!dups && L.size() > 1) {
            tmp.assign(L.begin(), L.end());
            std::sort(tmp.begin(), tmp.end());
            if (std::adjacent_find(tmp.begin(), tmp.end()) != tmp.end()) dups = 1;
        }
    }
    HIPCHK(h->node.ensure(node.size() * sizeof(uint2)), "alloc node table");
    HIPCHK(h->adj.ensure(pool.size() * sizeof(uint32_t)), "alloc adjacency");
    HIPCHK(h->hints.ensure(std::max<size_t>(h->hints_host.size(), 1) * sizeof(uint32_t)), "alloc hints");
    HIPCHK(hipMemcpy(h->node.p, node.data(), node.size() * sizeof(uint2), hipMemcpyHostToDevice), "copy node table");
    HIPCHK(hipMemcpy(h->adj.p, pool.data(), pool.size() * sizeof(uint32_t), hipMemcpyHostToDevice), "copy adjacency");
    if (!h->hints_host.empty())
        HIPCHK(hipMemcpy(h->hints.p, h->hints_host.data(), h->hints_host.size() * sizeof(uint32_t),
                         hipMemcpyHostToDevice), "copy hints");
    h->nnz = nnz;
    h->pool_used = nnz;
    h->nhints = (int64_t)h->hints_host.size();
    h->has_dups = dups;
    h->has_graph = true;
    ssg_query_layout_stale(h);
    return SSG_OK;
}

static hipStream_t pick(ssg_handle* h, void* stream)
{
    return stream ? (hipStream_t)stream : h->stream;
}

extern "C" {

// device memory, streams and events for hosts without their own GPU runtime
int ssg_dev_malloc(void** p, size_t bytes)
{
    if (!p) return ssg_fail(SSG_E_ARG, "null output");
    *p = nullptr;
    HIPCHK(hipMalloc(p, std::max<size_t>(bytes, 16)), "hipMalloc");
    return SSG_OK;
}

int ssg_dev_free(void* p)
{
    if (p) HIPCHK(hipFree(p), "hipFree");
    return SSG_OK;
}

int ssg_memcpy(void* dst, const void* src, size_t bytes, int32_t kind)
{
    const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice
                          : kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    if (bytes) HIPCHK(hipMemcpy(dst, src, bytes, k), "hipMemcpy");
    return SSG_OK;
}

int ssg_device_synchronize(void)
{
    HIPCHK(hipDeviceSynchronize(), "hipDeviceSynchronize");
    return SSG_OK;
}

int ssg_stream_create(void** s)
{
    if (!s) return ssg_fail(SSG_E_ARG, "null output");
    hipStream_t st = nullptr;
    HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
    *s = st;
    return SSG_OK;
}

int ssg_stream_destroy(void* s)
{
    if (s) HIPCHK(hipStreamDestroy((hipStream_t)s), "hipStreamDestroy");
    return SSG_OK;
}

int ssg_stream_synchronize(void* s)
{
    HIPCHK(hipStreamSynchronize((hipStream_t)s), "hipStreamSynchronize");
    return SSG_OK;
}

int ssg_event_create(void** e)
{
    if (!e) return ssg_fail(SSG_E_ARG, "null output");
    hipEvent_t ev = nullptr;
    HIPCHK(hipEventCreate(&ev), "hipEventCreate");
    *e = ev;
    return SSG_OK;
}

int ssg_event_destroy(void* e)
{
    if (e) HIPCHK(hipEventDestroy((hipEvent_t)e), "hipEventDestroy");
    return SSG_OK;
}

int ssg_event_record(void* e, void* s)
{
    HIPCHK(hipEventRecord((hipEvent_t)e, (hipStream_t)s), "hipEventRecord");
    return SSG_OK;
}

int ssg_event_elapsed_ms(void* e0, void* e1, float* ms)
{
    if (!ms) return ssg_fail(SSG_E_ARG, "null output");
    HIPCHK(hipEventSynchronize((hipEvent_t)e1), "hipEventSynchronize");
    HIPCHK(hipEventElapsedTime(ms, (hipEvent_t)e0, (hipEvent_t)e1), "hipEventElapsedTime");
    return SSG_OK;
}

const char* ssg_version(void) { return "ssg_hip 0.1 gfx950"; }

const char* ssg_last_error(const ssg_handle*) { return g_err.c_str(); }

int ssg_create(ssg_handle** out, int device)
{
    if (!out) return ssg_fail(SSG_E_ARG, "ssg_create: null output pointer");
    *out = nullptr;
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev), "hipGetDeviceCount");
    if (device < 0 || device >= ndev) return ssg_fail(SSG_E_ARG, "ssg_create: no such device");
    HIPCHK(hipSetDevice(device), "hipSetDevice");
    ssg_handle* h = new ssg_handle();
    h->device = device;
    hipDeviceProp_t prop;
    hipError_t e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) {
        delete h;
        return ssg_hipfail(e, "hipGetDeviceProperties");
    }
    h->num_cu = prop.multiProcessorCount;
    const int rc = ssg_noalloc_init(device);
    if (rc) {
        delete h;
        return rc;
    }
    e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete h;
        return ssg_hipfail(e, "hipStreamCreate");
    }
    *out = h;
    return SSG_OK;
}

int ssg_destroy(ssg_handle* h)
{
    if (!h) return SSG_OK;
    (void)hipSetDevice(h->device);
    (void)hipStreamSynchronize(h->stream);
    DevBuf* bufs[] = {&h->X, &h->xnorm, &h->Xh, &h->node, &h->adj, &h->hints, &h->vis, &h->epochs,
                      &h->counter, &h->qpad, &h->qnorm, &h->cand_d, &h->cand_i, &h->slice_tau,
                      &h->flags, &h->part_i, &h->part_d, &h->sQ, &h->sIds, &h->sD, &h->sCd,
                      &h->sCb, &h->sK, &h->sB, &h->sH, &h->sHc, &h->bIds, &h->bD, &h->bOut,
                      &h->bCnt, &h->bPairs, &h->bNodeUpd, &h->lH, &h->lB, &h->lC, &h->lN,
                      &h->lS};
    for (DevBuf* b : bufs) b->release();
    (void)hipStreamDestroy(h->stream);
    delete h;
    return SSG_OK;
}

int ssg_upload_db(ssg_handle* h, const float* X, int32_t dim, int64_t n, int32_t metric)
{
    if (!h) return ssg_fail(SSG_E_ARG, "null handle");
    std::lock_guard<std::mutex> lk(h->mu);
    if (!X && n > 0) return ssg_fail(SSG_E_ARG, "ssg_upload_db: null matrix");
    if (dim <= 0 || n < 0) return ssg_fail(SSG_E_ARG, "ssg_upload_db: bad shape");
    if (metric < SSG_SQL2 || metric > SSG_NORMCOS)
        return ssg_fail(SSG_E_ARG, "ssg_upload_db: unknown distance");
    // rows are padded to the next size the kernels are built for (roundup(dim, 16) up to
    // 128, else 192, 256, 384, 512, 768 or 1024); the canonical distance's 16 runs cover
    // the padded row
    if (dim > 1024) return ssg_fail(SSG_E_LIMIT, "ssg_upload_db: dim is limited to 1024");
    const int32_t dp = ssg_padded_dim(dim);
    if (n >= (int64_t)0xffffffffLL) return ssg_fail(SSG_E_LIMIT, "ssg_upload_db: n must fit UInt32 ids");
    HIPCHK(hipSetDevice(h->device), "hipSetDevice");
    HIPCHK(hipStreamSynchronize(h->stream), "sync");  // no launch may still read the old database
    h->dim = dim;
    h->dp = dp;
    h->metric = metric;
    h->n = n;
    h->n_pad = (n + 63) / 64 * 64;
    h->b_valid = false;
    h->has_graph = false;
    h->lists.clear();
    h->hints_host.clear();
    h->Xh.release();
    h->Xe.release();
    h->Xs.release();
    h->Xs2.release();
    h->xe_fmt = 0;
    ssg_query_layout_stale(h);
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
                return ssg_hipfail(e, "upload X");
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

// Room for n2 padded rows (and norms) in h, keeping the first rows; capacity doubles, and
// rows past n stay zero (the padding the kernels read).
static hipError_t grow_rows(ssg_handle* h, int64_t n2, hipStream_t s)
{
    const int64_t np2 = std::max<int64_t>((n2 + 63) / 64 * 64, 64);
    const size_t rowb = (size_t)h->dp * sizeof(float);
    if ((size_t)np2 * rowb <= h->X.cap && (size_t)np2 * sizeof(float) <= h->xnorm.cap) return hipSuccess;
    const int64_t have = (int64_t)(h->X.cap / rowb);
    const int64_t rows = std::max<int64_t>(np2, (2 * have + 63) / 64 * 64);
    const size_t xold = h->X.cap, nold = h->xnorm.cap;
    hipError_t e = h->X.grow((size_t)rows * rowb, xold, s);
    if (e == hipSuccess) e = h->xnorm.grow((size_t)rows * sizeof(float), nold, s);
    if (e == hipSuccess && h->X.cap > xold)
        e = launch_fill_f32(h->X.as<float>() + xold / 4, (int64_t)((h->X.cap - xold) / 4), 0.f, s);
    if (e == hipSuccess && h->xnorm.cap > nold)
        e = launch_fill_f32(h->xnorm.as<float>() + nold / 4, (int64_t)((h->xnorm.cap - nold) / 4), 0.f, s);
    return e;
}

int ssg_append_rows(ssg_handle* h, const float* X, int64_t m, bool dev)
{
    if (h->n + m >= (int64_t)0xffffffffLL) return ssg_fail(SSG_E_LIMIT, "append: n must fit UInt32 ids");
    if (m == 0) return SSG_OK;
    HIPCHK(hipSetDevice(h->device), "hipSetDevice");
    HIPCHK(hipStreamSynchronize(h->stream), "sync");  // no launch may still read the old rows
    const int64_t n = h->n, n2 = n + m, dp = h->dp, dim = h->dim;
    HIPCHK(grow_rows(h, n2, h->stream), "grow X");
    if (dev) {
        // padded rows: a copy, with the norms (zero padding adds exactly nothing to them)
        HIPCHK(launch_pad_rows(X, m, (int32_t)dp, h->X.as<float>() + n * dp, (int32_t)dp,
                               h->xnorm.as<float>() + n, h->stream), "append X");
        HIPCHK(hipStreamSynchronize(h->stream), "sync");
    } else {
        DevBuf tmp;
        const int64_t chunk = std::max<int64_t>(1, (int64_t)(256 << 20) / (dim * 4));
        HIPCHK(tmp.ensure((size_t)std::min(chunk, m) * dim * sizeof(float)), "alloc staging");
        for (int64_t r = 0; r < m; r += chunk) {
            const int64_t rows = std::min(chunk, m - r);
            hipError_t e = hipMemcpyAsync(tmp.p, X + r * dim, (size_t)rows * dim * sizeof(float),
                                          hipMemcpyHostToDevice, h->stream);
            if (e == hipSuccess)
                e = launch_pad_rows(tmp.as<float>(), rows, (int32_t)dim, h->X.as<float>() + (n + r) * dp,
                                    (int32_t)dp, h->xnorm.as<float>() + n + r, h->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
            if (e != hipSuccess) {
                tmp.release();
                return ssg_hipfail(e, "append X");
            }
        }
    }
    std::vector<float> nr((size_t)m);
    HIPCHK(hipMemcpy(nr.data(), h->xnorm.as<float>() + n, (size_t)m * sizeof(float),
                     hipMemcpyDeviceToHost), "read norms");
    for (float f : nr) h->xnmax = std::max(h->xnmax, f);
    h->n = n2;
    h->n_pad = (n2 + 63) / 64 * 64;
    // encoded copies and layouts of the old rows are rebuilt on next use
    h->Xh.release();
    h->Xe.release();
    h->Xs.release();
    h->Xs2.release();
    h->xe_fmt = 0;
    ssg_query_layout_stale(h);
    return SSG_OK;
}

int ssg_append_db(ssg_handle* h, const float* X, int64_t m)
{
    if (!h) return ssg_fail(SSG_E_ARG, "null handle");
    std::lock_guard<std::mutex> lk(h->mu);
    if (!h->X.p) return ssg_fail(SSG_E_STATE, "ssg_append_db: no database uploaded");
    if (m < 0 || (m > 0 && !X)) return ssg_fail(SSG_E_ARG, "ssg_append_db: bad arguments");
    if (h->has_graph)
        return ssg_fail(SSG_E_STATE, "ssg_append_db: the handle holds a graph (inserting into a "
                                     "SearchGraph is ssg_append_items_graph)");
    return ssg_append_rows(h, X, m, false);
}

int ssg_neardup(ssg_handle* h, float eps, int32_t blocksize, int32_t filterblocks, uint32_t* map,
                int64_t* nmap, int32_t* nn, float* dist, int64_t* costdists)
{
    if (!h) return ssg_fail(SSG_E_ARG, "null handle");
    std::lock_guard<std::mutex> lk(h->mu);
    if (!h->X.p) return ssg_fail(SSG_E_STATE, "ssg_neardup: no database uploaded");
    if (!map || !nmap || !nn || !dist || !costdists) return ssg_fail(SSG_E_ARG, "null buffer");
    if (blocksize < 1) return ssg_fail(SSG_E_ARG, "ssg_neardup: blocksize must be >= 1");
    const int64_t n = h->n;
    *nmap = 0;
    *costdists = 0;
    if (n == 0) return SSG_OK;
    const int32_t bs = (int32_t)std::min<int64_t>(blocksize, n);
    if (filterblocks && bs > 1024) return ssg_fail(SSG_E_LIMIT, "ssg_neardup: blocksize is limited to 1024");
    HIPCHK(hipSetDevice(h->device), "hipSetDevice");
    hipStream_t s = h->stream;
    HIPCHK(hipStreamSynchronize(s), "sync");
    // the exact index over the net: its own handle, rows gathered from h on the device
    ssg_handle* g = nullptr;
    int rc = ssg_create(&g, h->device);
    if (rc != SSG_OK) return rc;
    struct Free {
        ssg_handle* g;
        ~Free() { ssg_destroy(g); }
    } free_g{g};
    rc = ssg_upload_db(g, nullptr, h->dim, 0, h->metric);
    if (rc != SSG_OK) return rc;
    g->xnmax = h->xnmax;  // a bound on the net's norms (the exact stage's error bound)
    const int64_t dp = h->dp;
    // Blocks are searched in batches of about 4096 objects against the net as it stood at
    // the batch start (one exhaustive call); each block then checks the centers added
    // since (nearest_center, ties to the lower net position), which together is the
    // nearest over the whole net, as searchbatch! over the current net returns it.
    const int32_t nbatch = std::max<int32_t>(1, 4096 / bs);
    const int64_t qmax = (int64_t)nbatch * bs;
    const int64_t qt = (qmax + EXH_QT - 1) / EXH_QT * EXH_QT;
    DevBuf qpad, qnorm, rid, rd, items, P, cover, cdist, fresh_d, fpos, fd;
    HIPCHK(qpad.ensure((size_t)qt * dp * 4), "alloc");
    HIPCHK(qnorm.ensure((size_t)qt * 4), "alloc");
    HIPCHK(rid.ensure((size_t)qmax * 4), "alloc");
    HIPCHK(rd.ensure((size_t)qmax * 4), "alloc");
    HIPCHK(items.ensure((size_t)bs * 4), "alloc");
    if (filterblocks) HIPCHK(P.ensure((size_t)bs * bs * 4), "alloc");  // bs <= 1024 here
    HIPCHK(cover.ensure((size_t)bs * 4), "alloc");
    HIPCHK(cdist.ensure((size_t)bs * 4), "alloc");
    HIPCHK(fresh_d.ensure((size_t)qmax * 4), "alloc");
    HIPCHK(fpos.ensure((size_t)bs * 4), "alloc");
    HIPCHK(fd.ensure((size_t)bs * 4), "alloc");
    std::vector<uint32_t> M;       // net position -> 1-based object id (map)
    std::vector<uint32_t> sid((size_t)qmax), imap, fresh, blk(bs);
    std::vector<float> sd((size_t)qmax), hcd(bs), hfd(bs);
    std::vector<int32_t> hcov(bs), hfp(bs);
    int64_t cost = 0;
    for (int64_t b0 = 0; b0 < n; b0 += qmax) {
        const int64_t mq = std::min<int64_t>(qmax, n - b0);
        const int64_t net0 = g->n;  // net size at the batch start
        if (net0 > 0) {
            const int64_t m_pad = (mq + EXH_QT - 1) / EXH_QT * EXH_QT;
            HIPCHK(hipMemcpyAsync(qpad.p, h->X.as<float>() + b0 * dp, (size_t)mq * dp * 4,
                                  hipMemcpyDeviceToDevice, s), "copy");
            HIPCHK(hipMemcpyAsync(qnorm.p, h->xnorm.as<float>() + b0, (size_t)mq * 4,
                                  hipMemcpyDeviceToDevice, s), "copy");
            if (m_pad > mq) {
                HIPCHK(launch_fill_f32(qpad.as<float>() + mq * dp, (m_pad - mq) * dp, 0.f, s), "fill");
                HIPCHK(launch_fill_f32(qnorm.as<float>() + mq, m_pad - mq, 0.f, s), "fill");
            }
            rc = ssg_exh_run(g, nullptr, qpad.as<float>(), qnorm.as<float>(), mq, 1, 1,
                             rid.as<uint32_t>(), rd.as<float>(), s, net0);
            if (rc != SSG_OK) return rc;
            HIPCHK(hipMemcpyAsync(sid.data(), rid.p, (size_t)mq * 4, hipMemcpyDeviceToHost, s), "copy");
            HIPCHK(hipMemcpyAsync(sd.data(), rd.p, (size_t)mq * 4, hipMemcpyDeviceToHost, s), "copy");
            HIPCHK(hipStreamSynchronize(s), "sync");
        }
        std::vector<uint32_t> F;  // centers added during this batch (object ids, net order)
        for (int64_t r0 = b0; r0 < b0 + mq; r0 += bs) {
            const int32_t mb = (int32_t)std::min<int64_t>(bs, b0 + mq - r0);
            imap.clear();
            if (g->n == 0) {
                for (int32_t t = 0; t < mb; t++) imap.push_back((uint32_t)(r0 + t));
            } else {
                // searchbatch! of the block against the net (neardup.jl:121-127): each
                // exhaustive search charges the net's length (sequential-exhaustive.jl:78)
                cost += (int64_t)mb * g->n;
                const int32_t nf = (int32_t)F.size();
                if (nf > 0) {
                    for (int32_t t = 0; t < mb; t++) blk[t] = (uint32_t)(r0 + t);
                    HIPCHK(hipMemcpyAsync(items.p, blk.data(), (size_t)mb * 4, hipMemcpyHostToDevice, s),
                           "copy");
                    HIPCHK(launch_nearest_center(h->X.as<float>(), (int32_t)dp, h->metric,
                                                 items.as<uint32_t>(), mb, fresh_d.as<uint32_t>(), nf,
                                                 fpos.as<int32_t>(), fd.as<float>(), s), "nearest center");
                    HIPCHK(hipMemcpyAsync(hfp.data(), fpos.p, (size_t)mb * 4, hipMemcpyDeviceToHost, s),
                           "copy");
                    HIPCHK(hipMemcpyAsync(hfd.data(), fd.p, (size_t)mb * 4, hipMemcpyDeviceToHost, s), "copy");
                    HIPCHK(hipStreamSynchronize(s), "sync");
                }
                for (int32_t t = 0; t < mb; t++) {
                    const int64_t j = r0 + t, q = j - b0;
                    int64_t pos = net0 > 0 ? (int64_t)sid[q] : 0;  // 1-based net position
                    float d = net0 > 0 ? sd[q] : __builtin_inff();
                    if (nf > 0 && hfp[t] > 0 && hfd[t] < d) {  // later positions lose ties
                        pos = net0 + hfp[t];
                        d = hfd[t];
                    }
                    if (pos < 1 || d > eps) {  // no nearest center (NaN or +Inf distances): a center
                        imap.push_back((uint32_t)j);
                    } else {
                        dist[j] = d;
                        nn[j] = (int32_t)M[pos - 1];
                    }
                }
            }
            if (imap.empty()) continue;
            const int32_t mi = (int32_t)imap.size();
            fresh.clear();
            if (!filterblocks) {
                for (uint32_t j : imap) {
                    M.push_back(j + 1);
                    nn[j] = (int32_t)(j + 1);
                    dist[j] = 0.f;
                    fresh.push_back(j);
                }
            } else {
                HIPCHK(hipMemcpyAsync(items.p, imap.data(), (size_t)mi * 4, hipMemcpyHostToDevice, s),
                       "copy");
                HIPCHK(launch_pair_dist(h->X.as<float>(), (int32_t)dp, h->metric, items.as<uint32_t>(), mi,
                                        P.as<float>(), s), "pair distances");
                HIPCHK(launch_greedy_net(P.as<float>(), mi, eps, cover.as<int32_t>(), cdist.as<float>(), s),
                       "greedy net");
                HIPCHK(hipMemcpyAsync(hcov.data(), cover.p, (size_t)mi * 4, hipMemcpyDeviceToHost, s),
                       "copy");
                HIPCHK(hipMemcpyAsync(hcd.data(), cdist.p, (size_t)mi * 4, hipMemcpyDeviceToHost, s), "copy");
                HIPCHK(hipStreamSynchronize(s), "sync");
                int64_t T = 0;  // centers of this block so far (add_distance_evaluations!, :240)
                for (int32_t a = 0; a < mi; a++) {
                    const uint32_t j = imap[a];
                    if (a > 0) cost += T;
                    if (hcov[a] == a) {
                        T++;
                        M.push_back(j + 1);
                        nn[j] = (int32_t)(j + 1);
                        dist[j] = 0.f;
                        fresh.push_back(j);
                    } else {
                        nn[j] = (int32_t)(imap[hcov[a]] + 1);
                        dist[j] = hcd[a];
                    }
                }
            }
            if (fresh.empty()) continue;
            // append_items!(idx, ctx, X[tmp]) (:255): the new centers join the net (and the
            // batch's list of centers added since its search)
            HIPCHK(hipMemcpyAsync(fresh_d.as<uint32_t>() + F.size(), fresh.data(), fresh.size() * 4,
                                  hipMemcpyHostToDevice, s), "copy");
            HIPCHK(grow_rows(g, g->n + (int64_t)fresh.size(), s), "grow net");
            HIPCHK(launch_gather_padded_rows(h->X.as<float>(), h->xnorm.as<float>(), (int32_t)dp,
                                             fresh_d.as<uint32_t>() + F.size(), (int64_t)fresh.size(),
                                             g->X.as<float>(), g->xnorm.as<float>(), g->n, s),
                   "gather net rows");
            F.insert(F.end(), fresh.begin(), fresh.end());
            g->n += (int64_t)fresh.size();
            g->n_pad = (g->n + 63) / 64 * 64;
            // no synchronisation here: fresh (the copy's source) is next modified after the
            // next block's fix-up or batch-search synchronisation, which orders this copy
        }
    }
    HIPCHK(hipStreamSynchronize(s), "sync");
    std::copy(M.begin(), M.end(), map);
    *nmap = (int64_t)M.size();
    *costdists = cost;
    return SSG_OK;
}

// neardup with recall < 1 (src/neardup.jl:71-76): the net is a SearchGraph grown by
// append_items! (index! continued over each block's new centers, with OptimizeParameters
// (MinRecall(recall)) among its callbacks); every block searches the net as it stands
// (searchbatch! with KnnSorted(k) queues, the nearest result decides), then its remaining
// objects are filtered in order against the block's new centers as in the exact form.
int ssg_neardup_graph(ssg_handle* h, ssg_handle* g, float eps, int32_t blocksize, int32_t filterblocks,
                      int32_t k, const ssg_build_params* bp, uint32_t* map, int64_t* nmap, int32_t* nn,
                      float* dist, int64_t* costdists)
{
    if (!h || !g || !bp) return ssg_fail(SSG_E_ARG, "null argument");
    if (g == h) return ssg_fail(SSG_E_ARG, "ssg_neardup_graph: the net needs its own handle");
    std::lock_guard<std::mutex> lk(h->mu);
    std::lock_guard<std::mutex> lkg(g->mu);
    if (g->n != 0 || g->dim != h->dim || g->metric != h->metric || g->device != h->device)
        return ssg_fail(SSG_E_ARG, "ssg_neardup_graph: the net handle needs an empty database of "
                                   "the same dimension and distance on the same device");
    if (!h->X.p) return ssg_fail(SSG_E_STATE, "ssg_neardup_graph: no database uploaded");
    if (!map || !nmap || !nn || !dist || !costdists) return ssg_fail(SSG_E_ARG, "null buffer");
    if (blocksize < 1) return ssg_fail(SSG_E_ARG, "ssg_neardup_graph: blocksize must be >= 1");
    if (k < 1 || k > 64) return ssg_fail(SSG_E_ARG, "ssg_neardup_graph: k must be in 1..64");
    if (h->dp > 256) return ssg_fail(SSG_E_LIMIT, "ssg_neardup_graph: dim is limited to 256");
    const int64_t n = h->n;
    *nmap = 0;
    *costdists = 0;
    if (n == 0) return SSG_OK;
    const int32_t bs = (int32_t)std::min<int64_t>(blocksize, n);
    if (filterblocks && bs > 1024) return ssg_fail(SSG_E_LIMIT, "ssg_neardup_graph: blocksize is limited to 1024");
    HIPCHK(hipSetDevice(h->device), "hipSetDevice");
    hipStream_t s = h->stream;
    HIPCHK(hipStreamSynchronize(s), "sync");
    int rc = ssg_graph_begin(g, bp);
    if (rc != SSG_OK) return rc;
    const int64_t dp = h->dp;
    DevBuf rid, rd, rcd, items, P, cover, cdist, fresh_d, rows;
    HIPCHK(rid.ensure((size_t)bs * k * 4), "alloc");
    HIPCHK(rd.ensure((size_t)bs * k * 4), "alloc");
    HIPCHK(rcd.ensure((size_t)bs * 8), "alloc");
    HIPCHK(items.ensure((size_t)bs * 4), "alloc");
    if (filterblocks) HIPCHK(P.ensure((size_t)bs * bs * 4), "alloc");
    HIPCHK(cover.ensure((size_t)bs * 4), "alloc");
    HIPCHK(cdist.ensure((size_t)bs * 4), "alloc");
    HIPCHK(fresh_d.ensure((size_t)bs * 4), "alloc");
    HIPCHK(rows.ensure((size_t)bs * dp * 4 + (size_t)bs * 4), "alloc");
    std::vector<uint32_t> M, imap, fresh, sid((size_t)bs * k);
    std::vector<float> sd((size_t)bs * k), hcd(bs);
    std::vector<int64_t> scd(bs);
    std::vector<int32_t> hcov(bs);
    int64_t cost = 0;
    for (int64_t r0 = 0; r0 < n; r0 += bs) {
        const int32_t mb = (int32_t)std::min<int64_t>(bs, n - r0);
        imap.clear();
        if (g->n == 0) {
            for (int32_t t = 0; t < mb; t++) imap.push_back((uint32_t)(r0 + t));
        } else {
            // searchbatch!(idx, ctx, X[range], knns) over the net graph (neardup.jl:121-127)
            rc = ssg_beam_run(g, h->X.as<float>() + r0 * dp, (int32_t)dp, mb, k, k, &g->algo, nullptr,
                              rid.as<uint32_t>(), rd.as<float>(), rcd.as<int64_t>(), nullptr, s, false,
                              g->n, true);
            if (rc != SSG_OK) return rc;
            HIPCHK(hipMemcpyAsync(sid.data(), rid.p, (size_t)mb * k * 4, hipMemcpyDeviceToHost, s), "copy");
            HIPCHK(hipMemcpyAsync(sd.data(), rd.p, (size_t)mb * k * 4, hipMemcpyDeviceToHost, s), "copy");
            HIPCHK(hipMemcpyAsync(scd.data(), rcd.p, (size_t)mb * 8, hipMemcpyDeviceToHost, s), "copy");
            HIPCHK(hipStreamSynchronize(s), "sync");
            for (int32_t t = 0; t < mb; t++) {
                cost += scd[t];
                const int64_t j = r0 + t;
                const uint32_t pid = sid[(size_t)t * k];  // 1-based net position, 0 = none
                const float d = sd[(size_t)t * k];
                if (pid == 0 || d > eps) {
                    imap.push_back((uint32_t)j);
                } else {
                    dist[j] = d;
                    nn[j] = (int32_t)M[pid - 1];
                }
            }
        }
        if (imap.empty()) continue;
        const int32_t mi = (int32_t)imap.size();
        fresh.clear();
        if (!filterblocks) {
            for (uint32_t j : imap) {
                M.push_back(j + 1);
                nn[j] = (int32_t)(j + 1);
                dist[j] = 0.f;
                fresh.push_back(j);
            }
        } else {
            HIPCHK(hipMemcpyAsync(items.p, imap.data(), (size_t)mi * 4, hipMemcpyHostToDevice, s), "copy");
            HIPCHK(launch_pair_dist(h->X.as<float>(), (int32_t)dp, h->metric, items.as<uint32_t>(), mi,
                                    P.as<float>(), s), "pair distances");
            HIPCHK(launch_greedy_net(P.as<float>(), mi, eps, cover.as<int32_t>(), cdist.as<float>(), s),
                   "greedy net");
            HIPCHK(hipMemcpyAsync(hcov.data(), cover.p, (size_t)mi * 4, hipMemcpyDeviceToHost, s), "copy");
            HIPCHK(hipMemcpyAsync(hcd.data(), cdist.p, (size_t)mi * 4, hipMemcpyDeviceToHost, s), "copy");
            HIPCHK(hipStreamSynchronize(s), "sync");
            int64_t T = 0;  // centers of this block so far (add_distance_evaluations!, :240)
            for (int32_t a = 0; a < mi; a++) {
                const uint32_t j = imap[a];
                if (a > 0) cost += T;
                if (hcov[a] == a) {
                    T++;
                    M.push_back(j + 1);
                    nn[j] = (int32_t)(j + 1);
                    dist[j] = 0.f;
                    fresh.push_back(j);
                } else {
                    nn[j] = (int32_t)(imap[hcov[a]] + 1);
                    dist[j] = hcd[a];
                }
            }
        }
        // append_items!(idx, ctx, X[tmp]) (:255): the new centers' rows join the net and
        // index! continues over them
        const int64_t nf = (int64_t)fresh.size();
        HIPCHK(hipMemcpyAsync(fresh_d.p, fresh.data(), (size_t)nf * 4, hipMemcpyHostToDevice, s), "copy");
        HIPCHK(launch_gather_padded_rows(h->X.as<float>(), h->xnorm.as<float>(), (int32_t)dp,
                                         fresh_d.as<uint32_t>(), nf, rows.as<float>(),
                                         rows.as<float>() + (size_t)bs * dp, 0, s), "gather net rows");
        HIPCHK(hipStreamSynchronize(s), "sync");
        rc = ssg_append_rows(g, rows.as<float>(), nf, true);
        if (rc == SSG_OK) rc = ssg_build_continue(g);
        if (rc != SSG_OK) return rc;
    }
    HIPCHK(hipStreamSynchronize(s), "sync");
    std::copy(M.begin(), M.end(), map);
    *nmap = (int64_t)M.size();
    *costdists = cost;
    return SSG_OK;
}

int ssg_upload_graph(ssg_handle* h, const int64_t* offset, int64_t n, const uint32_t* end_point,
                     int64_t nnz, const uint32_t* hints, int64_t nhints)
{
    if (!h) return ssg_fail(SSG_E_ARG, "null handle");
    std::lock_guard<std::mutex> lk(h->mu);
    if (n != h->n) return ssg_fail(SSG_E_ARG, "ssg_upload_graph: graph size differs from the database");
    if (nnz < 0 || (nnz > 0 && !end_point) || (n > 0 && !offset) || nhints < 0 ||
        (nhints > 0 && !hints))
        return ssg_fail(SSG_E_ARG, "ssg_upload_graph: bad arguments");
    if (nnz >= (int64_t)0xffffffffLL) return ssg_fail(SSG_E_LIMIT, "ssg_upload_graph: nnz must be < 2^32");
    std::vector<std::vector<uint32_t>> lists((size_t)n);
    int64_t prev = 0;
    for (int64_t i = 0; i < n; i++) {
        const int64_t e = offset[i];
        if (e < prev || e > nnz) return ssg_fail(SSG_E_ARG, "ssg_upload_graph: offsets must be non-decreasing");
        lists[i].resize((size_t)(e - prev));
        for (int64_t t = prev; t < e; t++) {
            const uint32_t v = end_point[t];
            if (v < 1 || (int64_t)v > n) return ssg_fail(SSG_E_ARG, "ssg_upload_graph: neighbour id out of range");
            lists[i][(size_t)(t - prev)] = v - 1;
        }
        prev = e;
    }
    if (prev != nnz) return ssg_fail(SSG_E_ARG, "ssg_upload_graph: offset[n] must equal nnz");
    for (int64_t i = 0; i < nhints; i++)
        if (hints[i] < 1 || (int64_t)hints[i] > n) return ssg_fail(SSG_E_ARG, "ssg_upload_graph: hint out of range");
    HIPCHK(hipSetDevice(h->device), "hipSetDevice");
    h->b_valid = false;  // not built by index! here: append_items! cannot continue it
    h->lists.swap(lists);
    h->hints_host.assign(hints, hints + nhints);
    return ssg_install_lists(h);
}

int ssg_graph_info(ssg_handle* h, int64_t* nnz, int64_t* nhints, ssg_beam_params* algo)
{
    if (!h) return ssg_fail(SSG_E_ARG, "null handle");
    std::lock_guard<std::mutex> lk(h->mu);
    if (!h->has_graph) return ssg_fail(SSG_E_STATE, "no graph");
    if (nnz) *nnz = h->nnz;
    if (nhints) *nhints = h->nhints;
    if (algo) *algo = h->algo;
    return SSG_OK;
}

int ssg_download_graph(ssg_handle* h, int64_t* offset, uint32_t* end_point, uint32_t* hints)
{
    if (!h) return ssg_fail(SSG_E_ARG, "null handle");
    std::lock_guard<std::mutex> lk(h->mu);
    if (!h->has_graph) return ssg_fail(SSG_E_STATE, "no graph");
    int64_t pos = 0;
    for (size_t i = 0; i < h->lists.size(); i++) {
        for (uint32_t v : h->lists[i]) end_point[pos++] = v + 1;
        offset[i] = pos;
    }
    if (hints) std::copy(h->hints_host.begin(), h->hints_host.end(), hints);
    return SSG_OK;
}

int ssg_exhaustive_knn_dev(ssg_handle* h, const float* dQ, int64_t m, int32_t k, uint32_t* dids,
                           float* ddists, void* stream)
{
    if (!h) return ssg_fail(SSG_E_ARG, "null handle");
    std::lock_guard<std::mutex> lk(h->mu);
    HIPCHK(hipSetDevice(h->device), "hipSetDevice");
    if (!dQ || !dids || !ddists) return ssg_fail(SSG_E_ARG, "null buffer");
    return ssg_exh_run(h, dQ, nullptr, nullptr, m, k, k, dids, ddists, pick(h, stream), h->n);
}

int ssg_exhaustive_knn(ssg_handle* h, const float* Q, int64_t m, int32_t k, uint32_t* ids,
                       float* dists)
{
    if (!h) return ssg_fail(SSG_E_ARG, "null handle");
    std::lock_guard<std::mutex> lk(h->mu);
    HIPCHK(hipSetDevice(h->device), "hipSetDevice");
    if (!Q || !ids || !dists) return ssg_fail(SSG_E_ARG, "null buffer");
    if (m <= 0) return ssg_fail(SSG_E_ARG, "empty set of queries");
    if (k < 1) return ssg_fail(SSG_E_ARG, "k must be >= 1");
    hipStream_t s = h->stream;
    HIPCHK(h->sQ.ensure((size_t)m * h->dim * sizeof(float)), "alloc");
    HIPCHK(h->sIds.ensure((size_t)m * k * sizeof(uint32_t)), "alloc");
    HIPCHK(h->sD.ensure((size_t)m * k * sizeof(float)), "alloc");
    HIPCHK(hipMemcpyAsync(h->sQ.p, Q, (size_t)m * h->dim * sizeof(float), hipMemcpyHostToDevice, s), "copy Q");
    int rc = ssg_exh_run(h, h->sQ.as<float>(), nullptr, nullptr, m, k, k, h->sIds.as<uint32_t>(),
                         h->sD.as<float>(), s, h->n);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(ids, h->sIds.p, (size_t)m * k * sizeof(uint32_t), hipMemcpyDeviceToHost, s), "copy ids");
    HIPCHK(hipMemcpyAsync(dists, h->sD.p, (size_t)m * k * sizeof(float), hipMemcpyDeviceToHost, s), "copy dists");
    HIPCHK(hipStreamSynchronize(s), "sync");
    return SSG_OK;
}

int ssg_allknn_exhaustive(ssg_handle* h, int32_t k, uint32_t* ids, float* dists)
{
    if (!h) return ssg_fail(SSG_E_ARG, "null handle");
    std::lock_guard<std::mutex> lk(h->mu);
    HIPCHK(hipSetDevice(h->device), "hipSetDevice");
    if (!ids || !dists) return ssg_fail(SSG_E_ARG, "null buffer");
    const int64_t n = h->n;
    if (n <= 0) return ssg_fail(SSG_E_ARG, "allknn needs a non-empty database");
    if (k < 1 || k > n) return ssg_fail(SSG_E_ARG, "allknn needs 0 < k <= n");
    hipStream_t s = h->stream;
    HIPCHK(h->sIds.ensure((size_t)n * k * sizeof(uint32_t)), "alloc");
    HIPCHK(h->sD.ensure((size_t)n * k * sizeof(float)), "alloc");
    int rc = ssg_exh_run(h, nullptr, h->X.as<float>(), h->xnorm.as<float>(), n, k, k,
                         h->sIds.as<uint32_t>(), h->sD.as<float>(), s, n);
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
    if (!h) return ssg_fail(SSG_E_ARG, "null handle");
    std::lock_guard<std::mutex> lk(h->mu);
    HIPCHK(hipSetDevice(h->device), "hipSetDevice");
    if (!dQ || !dids || !ddists) return ssg_fail(SSG_E_ARG, "null buffer");
    if (qo && qo->hints && (!qo->nhints || qo->hint_stride < 1))
        return ssg_fail(SSG_E_ARG, "per-query hints need counts and a stride");
    return ssg_beam_run(h, dQ, h->dim, m, k, kstride, bp, qo, dids, ddists, dcd, dcb,
                        pick(h, stream), false, h->n);
}

// HIP graphs of device-pointer call sequences
}  // extern "C"

struct ssg_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

extern "C" {

int ssg_capture_begin(void* stream)
{
    if (!stream) return ssg_fail(SSG_E_ARG, "capture needs a created stream (not the null stream)");
    HIPCHK(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeRelaxed), "begin capture");
    return SSG_OK;
}

int ssg_capture_end(void* stream, ssg_graph** g)
{
    if (!stream || !g) return ssg_fail(SSG_E_ARG, "null argument");
    *g = nullptr;
    ssg_graph* r = new ssg_graph();
    hipError_t e = hipStreamEndCapture((hipStream_t)stream, &r->graph);
    if (e == hipSuccess) e = hipGraphInstantiate(&r->exec, r->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        if (r->graph) (void)hipGraphDestroy(r->graph);
        delete r;
        return ssg_hipfail(e, "end capture");
    }
    *g = r;
    return SSG_OK;
}

int ssg_graph_launch(ssg_graph* g, void* stream)
{
    if (!g || !stream) return ssg_fail(SSG_E_ARG, "null argument");
    HIPCHK(hipGraphLaunch(g->exec, (hipStream_t)stream), "graph launch");
    return SSG_OK;
}

int ssg_graph_destroy(ssg_graph* g)
{
    if (!g) return SSG_OK;
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
    return SSG_OK;
}

int ssg_set_option(ssg_handle* h, int32_t option, int64_t value)
{
    if (!h) return ssg_fail(SSG_E_ARG, "null handle");
    std::lock_guard<std::mutex> lk(h->mu);
    if (option == SSG_OPT_QUERY_LAYOUT) {
        if (value != 0 && value != 1) return ssg_fail(SSG_E_ARG, "query layout option is 0 or 1");
        HIPCHK(hipSetDevice(h->device), "hipSetDevice");
        HIPCHK(hipDeviceSynchronize(), "sync");
        h->reorder = value == 1;
        ssg_query_layout_stale(h);
        return SSG_OK;
    }
    if (option == SSG_OPT_PREC_SAMPLE) {
        if (value < -1 || value > PREC_NCAND_MAX) return ssg_fail(SSG_E_ARG, "sample option must be in -1..256");
        h->prec_sample = (int32_t)value;
        return SSG_OK;
    }
    if (option == SSG_OPT_PREC_SLICES) {
        if (value < 0 || value > 2048) return ssg_fail(SSG_E_ARG, "slice count must be in 0..2048");
        h->prec_slices = (int32_t)value;
        return SSG_OK;
    }
    return ssg_fail(SSG_E_ARG, "unknown option");
}

int ssg_order_by_bsize_dev(ssg_handle* h, const int32_t* dbsize, int64_t m, uint32_t* dorder, void* stream)
{
    if (!h || !dbsize || !dorder) return ssg_fail(SSG_E_ARG, "null argument");
    if (m < 1) return ssg_fail(SSG_E_ARG, "empty set of queries");
    std::lock_guard<std::mutex> lk(h->mu);
    HIPCHK(hipSetDevice(h->device), "hipSetDevice");
    HIPCHK(h->lCnt.ensure(64 * 4), "alloc");
    HIPCHK(launch_order_desc(dbsize, m, dorder, h->lCnt.as<unsigned int>(), pick(h, stream)), "order");
    return SSG_OK;
}

int ssg_beamsearch(ssg_handle* h, const float* Q, int64_t m, int32_t k, int32_t kstride,
                   const ssg_beam_params* bp, const ssg_query_opts* qo, uint32_t* ids,
                   float* dists, int64_t* costdists, int64_t* costblocks)
{
    if (!h) return ssg_fail(SSG_E_ARG, "null handle");
    std::lock_guard<std::mutex> lk(h->mu);
    HIPCHK(hipSetDevice(h->device), "hipSetDevice");
    if (!Q || !ids || !dists) return ssg_fail(SSG_E_ARG, "null buffer");
    if (m <= 0) return ssg_fail(SSG_E_ARG, "empty set of queries");
    if (k < 1 || kstride < k || kstride > 256)
        return ssg_fail(SSG_E_LIMIT, "k must satisfy 1 <= k <= kstride <= 256");
    hipStream_t s = h->stream;
    ssg_query_opts dq{};
    if (qo) {
        if (qo->k) {
            for (int64_t j = 0; j < m; j++)
                if (qo->k[j] < 1 || qo->k[j] > kstride) return ssg_fail(SSG_E_ARG, "per-query k out of range");
            HIPCHK(h->sK.ensure((size_t)m * 4), "alloc");
            HIPCHK(hipMemcpyAsync(h->sK.p, qo->k, (size_t)m * 4, hipMemcpyHostToDevice, s), "copy k");
            dq.k = h->sK.as<int32_t>();
        }
        if (qo->bsize) {
            for (int64_t j = 0; j < m; j++)
                if (qo->bsize[j] < 1) return ssg_fail(SSG_E_ARG, "per-query bsize must be >= 1");
            HIPCHK(h->sB.ensure((size_t)m * 4), "alloc");
            HIPCHK(hipMemcpyAsync(h->sB.p, qo->bsize, (size_t)m * 4, hipMemcpyHostToDevice, s), "copy bsize");
            dq.bsize = h->sB.as<int32_t>();
        }
        if (qo->hints) {
            if (!qo->nhints || qo->hint_stride < 1)
                return ssg_fail(SSG_E_ARG, "per-query hints need counts and a stride");
            for (int64_t j = 0; j < m; j++) {
                if (qo->nhints[j] < 0 || qo->nhints[j] > qo->hint_stride)
                    return ssg_fail(SSG_E_ARG, "per-query hint count out of range");
                for (int32_t t = 0; t < qo->nhints[j]; t++) {
                    const uint32_t v = qo->hints[j * qo->hint_stride + t];
                    if (v < 1 || (int64_t)v > h->n) return ssg_fail(SSG_E_ARG, "per-query hint out of range");
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
        if (qo->visited) {
            for (int64_t j = 0; j < m; j++)
                if ((int64_t)qo->visited[j] > h->n) return ssg_fail(SSG_E_ARG, "visited id out of range");
            HIPCHK(h->sV.ensure((size_t)m * 4), "alloc");
            HIPCHK(hipMemcpyAsync(h->sV.p, qo->visited, (size_t)m * 4, hipMemcpyHostToDevice, s),
                   "copy visited");
            dq.visited = h->sV.as<uint32_t>();
        }
        if (qo->order) {
            std::vector<uint8_t> seen((size_t)m, 0);
            for (int64_t j = 0; j < m; j++) {
                const uint32_t v = qo->order[j];
                if ((int64_t)v >= m || seen[v]) return ssg_fail(SSG_E_ARG, "order must be a permutation of 0..m-1");
                seen[v] = 1;
            }
            HIPCHK(h->sO.ensure((size_t)m * 4), "alloc");
            HIPCHK(hipMemcpyAsync(h->sO.p, qo->order, (size_t)m * 4, hipMemcpyHostToDevice, s), "copy order");
            dq.order = h->sO.as<uint32_t>();
        }
    }
    HIPCHK(h->sQ.ensure((size_t)m * h->dim * sizeof(float)), "alloc");
    HIPCHK(h->sIds.ensure((size_t)m * kstride * 4), "alloc");
    HIPCHK(h->sD.ensure((size_t)m * kstride * 4), "alloc");
    HIPCHK(h->sCd.ensure((size_t)m * 8), "alloc");
    HIPCHK(h->sCb.ensure((size_t)m * 8), "alloc");
    HIPCHK(hipMemcpyAsync(h->sQ.p, Q, (size_t)m * h->dim * sizeof(float), hipMemcpyHostToDevice, s), "copy Q");
    int rc = ssg_beam_run(h, h->sQ.as<float>(), h->dim, m, k, kstride, bp, qo ? &dq : nullptr,
                          h->sIds.as<uint32_t>(), h->sD.as<float>(), h->sCd.as<int64_t>(),
                          h->sCb.as<int64_t>(), s, false, h->n);
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
    if (!h) return ssg_fail(SSG_E_ARG, "null handle");
    std::lock_guard<std::mutex> lk(h->mu);
    HIPCHK(hipSetDevice(h->device), "hipSetDevice");
    if (!ids || !dists) return ssg_fail(SSG_E_ARG, "null buffer");
    const int64_t n = h->n;
    if (n <= 0) return ssg_fail(SSG_E_ARG, "allknn needs a non-empty database");
    if (k < 1 || k > n) return ssg_fail(SSG_E_ARG, "allknn needs 0 < k <= n");
    hipStream_t s = h->stream;
    HIPCHK(h->sIds.ensure((size_t)n * k * 4), "alloc");
    HIPCHK(h->sD.ensure((size_t)n * k * 4), "alloc");
    int rc = ssg_beam_run(h, nullptr, h->dim, n, k, k, bp, nullptr, h->sIds.as<uint32_t>(),
                          h->sD.as<float>(), nullptr, nullptr, s, true, n);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(ids, h->sIds.p, (size_t)n * k * 4, hipMemcpyDeviceToHost, s), "copy ids");
    HIPCHK(hipMemcpyAsync(dists, h->sD.p, (size_t)n * k * 4, hipMemcpyDeviceToHost, s), "copy dists");
    HIPCHK(hipStreamSynchronize(s), "sync");
    return SSG_OK;
}

int ssg_exhaustive_knn_prec_dev(ssg_handle* h, const float* dQ, int64_t m, int32_t k,
                                int32_t precision, int32_t ncand, uint32_t* dids, float* ddists,
                                uint32_t* dcand_ids, float* dcand_dists, void* stream)
{
    if (!h) return ssg_fail(SSG_E_ARG, "null handle");
    std::lock_guard<std::mutex> lk(h->mu);
    HIPCHK(hipSetDevice(h->device), "hipSetDevice");
    if (!dQ || !dids || !ddists) return ssg_fail(SSG_E_ARG, "null buffer");
    return ssg_prec_run(h, dQ, m, k, k, precision, ncand, dids, ddists, dcand_ids, dcand_dists,
                        pick(h, stream));
}

int ssg_allknn_exhaustive_prec_dev(ssg_handle* h, int32_t k, int32_t precision, int32_t ncand,
                                   uint32_t* dids, float* ddists, uint32_t* dcand_ids,
                                   float* dcand_dists, void* stream)
{
    if (!h) return ssg_fail(SSG_E_ARG, "null handle");
    std::lock_guard<std::mutex> lk(h->mu);
    HIPCHK(hipSetDevice(h->device), "hipSetDevice");
    if (!dids || !ddists) return ssg_fail(SSG_E_ARG, "null buffer");
    if (h->n <= 0) return ssg_fail(SSG_E_ARG, "allknn needs a non-empty database");
    if (k < 1 || k > h->n) return ssg_fail(SSG_E_ARG, "allknn needs 0 < k <= n");
    return ssg_prec_run(h, nullptr, h->n, k, k, precision, ncand, dids, ddists, dcand_ids,
                        dcand_dists, pick(h, stream));
}

// Host-buffer forms: Q (m rows) or the database itself when Q is null.
static int prec_host(ssg_handle* h, const float* Q, int64_t m, int32_t k, int32_t precision,
                     int32_t ncand, uint32_t* ids, float* dists, uint32_t* cand_ids,
                     float* cand_dists)
{
    if (!ids || !dists) return ssg_fail(SSG_E_ARG, "null buffer");
    if ((cand_ids == nullptr) != (cand_dists == nullptr))
        return ssg_fail(SSG_E_ARG, "candidate ids and distances go together");
    if (m <= 0) return ssg_fail(SSG_E_ARG, "empty set of queries");
    if (k < 1) return ssg_fail(SSG_E_ARG, "k must be >= 1");
    if (ncand < k || ncand > PREC_NCAND_MAX)
        return ssg_fail(SSG_E_LIMIT, "ncand must satisfy k <= ncand <= 256");
    hipStream_t s = h->stream;
    const float* dQ = nullptr;
    if (Q) {
        HIPCHK(h->sQ.ensure((size_t)m * h->dim * sizeof(float)), "alloc");
        HIPCHK(hipMemcpyAsync(h->sQ.p, Q, (size_t)m * h->dim * sizeof(float), hipMemcpyHostToDevice, s),
               "copy Q");
        dQ = h->sQ.as<float>();
    }
    HIPCHK(h->sIds.ensure((size_t)m * k * sizeof(uint32_t)), "alloc");
    HIPCHK(h->sD.ensure((size_t)m * k * sizeof(float)), "alloc");
    if (cand_ids) {
        HIPCHK(h->sCd.ensure((size_t)m * ncand * sizeof(uint32_t)), "alloc");
        HIPCHK(h->sCb.ensure((size_t)m * ncand * sizeof(float)), "alloc");
    }
    int rc = ssg_prec_run(h, dQ, m, k, k, precision, ncand, h->sIds.as<uint32_t>(), h->sD.as<float>(),
                          cand_ids ? h->sCd.as<uint32_t>() : nullptr,
                          cand_ids ? h->sCb.as<float>() : nullptr, s);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(ids, h->sIds.p, (size_t)m * k * sizeof(uint32_t), hipMemcpyDeviceToHost, s), "copy ids");
    HIPCHK(hipMemcpyAsync(dists, h->sD.p, (size_t)m * k * sizeof(float), hipMemcpyDeviceToHost, s), "copy dists");
    if (cand_ids) {
        HIPCHK(hipMemcpyAsync(cand_ids, h->sCd.p, (size_t)m * ncand * sizeof(uint32_t), hipMemcpyDeviceToHost, s),
               "copy candidates");
        HIPCHK(hipMemcpyAsync(cand_dists, h->sCb.p, (size_t)m * ncand * sizeof(float), hipMemcpyDeviceToHost, s),
               "copy candidates");
    }
    HIPCHK(hipStreamSynchronize(s), "sync");
    return SSG_OK;
}

int ssg_exhaustive_knn_prec(ssg_handle* h, const float* Q, int64_t m, int32_t k, int32_t precision,
                            int32_t ncand, uint32_t* ids, float* dists, uint32_t* cand_ids,
                            float* cand_dists)
{
    if (!h) return ssg_fail(SSG_E_ARG, "null handle");
    std::lock_guard<std::mutex> lk(h->mu);
    HIPCHK(hipSetDevice(h->device), "hipSetDevice");
    if (!Q) return ssg_fail(SSG_E_ARG, "null buffer");
    return prec_host(h, Q, m, k, precision, ncand, ids, dists, cand_ids, cand_dists);
}

int ssg_allknn_exhaustive_prec(ssg_handle* h, int32_t k, int32_t precision, int32_t ncand,
                               uint32_t* ids, float* dists, uint32_t* cand_ids, float* cand_dists)
{
    if (!h) return ssg_fail(SSG_E_ARG, "null handle");
    std::lock_guard<std::mutex> lk(h->mu);
    HIPCHK(hipSetDevice(h->device), "hipSetDevice");
    if (h->n <= 0) return ssg_fail(SSG_E_ARG, "allknn needs a non-empty database");
    if (k < 1 || k > h->n) return ssg_fail(SSG_E_ARG, "allknn needs 0 < k <= n");
    return prec_host(h, nullptr, h->n, k, precision, ncand, ids, dists, cand_ids, cand_dists);
}

int ssg_rerank_dev(ssg_handle* h, const float* dQ, int64_t m, int32_t k, uint32_t* dids,
                   float* ddists, void* stream)
{
    if (!h) return ssg_fail(SSG_E_ARG, "null handle");
    std::lock_guard<std::mutex> lk(h->mu);
    HIPCHK(hipSetDevice(h->device), "hipSetDevice");
    if (!h->X.p) return ssg_fail(SSG_E_STATE, "no database uploaded");
    if (m <= 0 || k < 1) return ssg_fail(SSG_E_ARG, "empty rerank");
    if (k > 1024) return ssg_fail(SSG_E_LIMIT, "rerank columns are limited to 1024 rows");
    HIPCHK(launch_rerank(h->X.as<float>(), h->dp, h->metric, dQ, h->dim, m, k, dids, ddists,
                         pick(h, stream)), "rerank");
    return SSG_OK;
}

int ssg_hsp_queries_dev(ssg_handle* h, int64_t m, int32_t k, const uint32_t* dids,
                        const float* ddists, uint32_t* dhsp_ids, float* dhsp_dists, void* stream)
{
    if (!h) return ssg_fail(SSG_E_ARG, "null handle");
    std::lock_guard<std::mutex> lk(h->mu);
    HIPCHK(hipSetDevice(h->device), "hipSetDevice");
    if (!h->X.p) return ssg_fail(SSG_E_STATE, "no database uploaded");
    if (!dids || !ddists || !dhsp_ids || !dhsp_dists) return ssg_fail(SSG_E_ARG, "null buffer");
    if (m <= 0 || k < 1) return ssg_fail(SSG_E_ARG, "hsp_queries: empty candidate set");
    if (k > 256 || (int64_t)k * h->dp * 4 > 160 * 1024)
        return ssg_fail(SSG_E_LIMIT, "hsp_queries: columns are limited to 256 rows and k * padded dim * 4 to 160 KiB");
    HIPCHK(launch_hsp(h->X.as<float>(), h->dp, h->metric, m, k, dids, ddists, dhsp_ids, dhsp_dists,
                      pick(h, stream)), "hsp_queries");
    return SSG_OK;
}

int ssg_hsp_queries(ssg_handle* h, int64_t m, int32_t k, const uint32_t* ids, const float* dists,
                    uint32_t* hsp_ids, float* hsp_dists)
{
    if (!h) return ssg_fail(SSG_E_ARG, "null handle");
    std::lock_guard<std::mutex> lk(h->mu);
    HIPCHK(hipSetDevice(h->device), "hipSetDevice");
    if (!h->X.p) return ssg_fail(SSG_E_STATE, "no database uploaded");
    if (!ids || !dists || !hsp_ids || !hsp_dists) return ssg_fail(SSG_E_ARG, "null buffer");
    if (m <= 0 || k < 1) return ssg_fail(SSG_E_ARG, "hsp_queries: empty candidate set");
    if (k > 256 || (int64_t)k * h->dp * 4 > 160 * 1024)
        return ssg_fail(SSG_E_LIMIT, "hsp_queries: columns are limited to 256 rows and k * padded dim * 4 to 160 KiB");
    for (int64_t j = 0; j < m; j++)
        for (int32_t t = 0; t < k; t++) {
            const uint32_t v = ids[j * k + t];
            if (v == 0) break;
            if ((int64_t)v > h->n) return ssg_fail(SSG_E_ARG, "hsp_queries: id out of range");
        }
    hipStream_t s = h->stream;
    const size_t bytes = (size_t)m * k * 4;
    HIPCHK(h->sIds.ensure(bytes), "alloc");
    HIPCHK(h->sD.ensure(bytes), "alloc");
    HIPCHK(h->cIds.ensure(bytes), "alloc");
    HIPCHK(h->cD.ensure(bytes), "alloc");
    HIPCHK(hipMemcpyAsync(h->sIds.p, ids, bytes, hipMemcpyHostToDevice, s), "copy");
    HIPCHK(hipMemcpyAsync(h->sD.p, dists, bytes, hipMemcpyHostToDevice, s), "copy");
    HIPCHK(launch_hsp(h->X.as<float>(), h->dp, h->metric, m, k, h->sIds.as<uint32_t>(),
                      h->sD.as<float>(), h->cIds.as<uint32_t>(), h->cD.as<float>(), s),
           "hsp_queries");
    HIPCHK(hipMemcpyAsync(hsp_ids, h->cIds.p, bytes, hipMemcpyDeviceToHost, s), "copy");
    HIPCHK(hipMemcpyAsync(hsp_dists, h->cD.p, bytes, hipMemcpyDeviceToHost, s), "copy");
    HIPCHK(hipStreamSynchronize(s), "sync");
    return SSG_OK;
}

int ssg_rerank(ssg_handle* h, const float* Q, int64_t m, int32_t k, uint32_t* ids, float* dists)
{
    if (!h) return ssg_fail(SSG_E_ARG, "null handle");
    std::lock_guard<std::mutex> lk(h->mu);
    HIPCHK(hipSetDevice(h->device), "hipSetDevice");
    if (!h->X.p) return ssg_fail(SSG_E_STATE, "no database uploaded");
    if (!Q || !ids || !dists) return ssg_fail(SSG_E_ARG, "null buffer");
    if (m <= 0 || k < 1) return ssg_fail(SSG_E_ARG, "empty rerank");
    if (k > 1024) return ssg_fail(SSG_E_LIMIT, "rerank columns are limited to 1024 rows");
    for (int64_t j = 0; j < m; j++)
        for (int32_t t = 0; t < k; t++) {
            const uint32_t v = ids[j * k + t];
            if (v == 0) break;
            if ((int64_t)v > h->n) return ssg_fail(SSG_E_ARG, "rerank id out of range");
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
