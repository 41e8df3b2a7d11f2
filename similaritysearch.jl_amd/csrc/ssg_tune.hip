// ssg_tune.hip -- the launch shape of the beam searches and its autotuner (no reference
// counterpart).  ssg_optimize_launch extends optimize_index!'s recall-targeted search
// (src/opt.jl:220-317: MinRecall keeps configurations that reach the recall target and
// prefers the cheapest) from BeamSearch's (bsize, Δ) to the launch: the recall-changing
// knobs of the BF16 form (rerank length, wide steps) are scored on recall against gold ids,
// and among those that reach the target the launch shape (wave slots per CU, gather rounds
// per batch, LDS visited table or none), which changes no result, is chosen by measured time.
#include <algorithm>
#include <unordered_set>
#include <vector>

#include "ssg_handle.h"

namespace {

int check_shape(const ssg_launch_params* lp)
{
    if (!lp) return ssg_fail(SSG_E_ARG, "null launch parameters");
    if (lp->slots_per_cu < 4 || lp->slots_per_cu > 64)
        return ssg_fail(SSG_E_ARG, "slots_per_cu must be in 4..64");
    if (lp->rounds != 4 && lp->rounds != 8 && lp->rounds != 10)
        return ssg_fail(SSG_E_ARG, "rounds must be 4, 8 or 10");
    if (lp->vbuckets < -1 || lp->vbuckets > 2048)
        return ssg_fail(SSG_E_ARG, "vbuckets must be -1 (automatic), 0 (none) or at most 2048");
    return SSG_OK;
}

// Mean over queries of |first kk results ∩ first kk gold ids| / kk, kk = min(k_j, kg).
double score(const std::vector<uint32_t>& ids, int32_t kstride, const std::vector<int32_t>& kq,
             const uint32_t* gold, int32_t kg, int64_t m)
{
    double acc = 0.0;
    std::unordered_set<uint32_t> g;
    for (int64_t j = 0; j < m; j++) {
        const int32_t kk = std::min(kq[(size_t)j], kg);
        g.clear();
        for (int32_t t = 0; t < kk; t++) g.insert(gold[(size_t)j * kg + t]);
        int32_t hit = 0;
        for (int32_t t = 0; t < kk; t++) hit += g.count(ids[(size_t)j * kstride + t]) ? 1 : 0;
        acc += kk > 0 ? (double)hit / kk : 1.0;
    }
    return m > 0 ? acc / (double)m : 0.0;
}

}  // namespace

extern "C" {

int ssg_set_launch(ssg_handle* h, const ssg_launch_params* lp)
{
    if (!h) return ssg_fail(SSG_E_ARG, "null handle");
    const int rc = check_shape(lp);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(h->mu);
    h->slots_per_cu = lp->slots_per_cu;
    h->rounds = lp->rounds;
    h->vbuckets = lp->vbuckets;
    return SSG_OK;
}

int ssg_get_launch(ssg_handle* h, ssg_launch_params* lp)
{
    if (!h || !lp) return ssg_fail(SSG_E_ARG, "null argument");
    std::lock_guard<std::mutex> lk(h->mu);
    lp->slots_per_cu = h->slots_per_cu;
    lp->rounds = h->rounds;
    lp->vbuckets = h->vbuckets;
    return SSG_OK;
}

int ssg_optimize_launch(ssg_handle* h, const float* dQ, int64_t m, int32_t k, int32_t kstride,
                        ssg_beam_params* bp, const ssg_query_opts* qo, const uint32_t* gold,
                        int32_t kg, float minrecall, ssg_launch_params* lp_out, double* recall_out,
                        float* ms_out)
{
    if (!h || !dQ || !bp || !gold || !lp_out) return ssg_fail(SSG_E_ARG, "null argument");
    if (m < 1 || kg < 1 || k < 1 || kstride < k || kstride > 64)
        return ssg_fail(SSG_E_ARG, "bad sizes (m >= 1, kg >= 1, 1 <= k <= kstride <= 64)");
    std::lock_guard<std::mutex> lk(h->mu);
    HIPCHK(hipSetDevice(h->device), "hipSetDevice");
    hipStream_t s = h->stream;
    DevBuf dI, dD;
    HIPCHK(dI.ensure((size_t)m * kstride * 4), "alloc");
    HIPCHK(dD.ensure((size_t)m * kstride * 4), "alloc");
    std::vector<int32_t> kq((size_t)m, k);
    if (qo && qo->k)
        HIPCHK(hipMemcpy(kq.data(), qo->k, (size_t)m * 4, hipMemcpyDeviceToHost), "copy k");
    std::vector<uint32_t> ids((size_t)m * kstride);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIPCHK(hipEventCreate(&e0), "event");
    hipError_t ee = hipEventCreate(&e1);
    if (ee != hipSuccess) {
        (void)hipEventDestroy(e0);
        return ssg_hipfail(ee, "event");
    }
    struct Guard {
        hipEvent_t a, b;
        ~Guard()
        {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    } guard{e0, e1};

    const int32_t sp0 = h->slots_per_cu, r0 = h->rounds, vb0 = h->vbuckets;
    auto run = [&](const ssg_beam_params& p, float* ms) -> int {
        HIPCHK(hipEventRecord(e0, s), "event");
        int rc = ssg_beam_run(h, dQ, h->dim, m, k, kstride, &p, qo, dI.as<uint32_t>(), dD.as<float>(),
                              nullptr, nullptr, s, false, h->n);
        if (rc) return rc;
        HIPCHK(hipEventRecord(e1, s), "event");
        HIPCHK(hipEventSynchronize(e1), "sync");
        if (ms) HIPCHK(hipEventElapsedTime(ms, e0, e1), "elapsed");
        return SSG_OK;
    };

    // 1. recall-changing knobs at the current launch shape
    std::vector<ssg_beam_params> knobs;
    if (bp->precision == SSG_BF16) {
        // query j reranks max(k_j, rerank): the grid grows the caller's rerank (or k)
        const int32_t r_lo = bp->rerank > 0 ? bp->rerank : k;
        std::vector<int32_t> rr = {r_lo, r_lo + 2, r_lo + 4, (r_lo * 3 + 1) / 2, 2 * r_lo};
        std::sort(rr.begin(), rr.end());
        rr.erase(std::unique(rr.begin(), rr.end()), rr.end());
        for (int32_t r : rr) {
            if (std::max(r, kstride) > 64) continue;
            for (int32_t ex : {1, 2}) {
                ssg_beam_params p = *bp;
                p.rerank = r;
                p.expand = ex;
                knobs.push_back(p);
            }
        }
    } else {
        knobs.push_back(*bp);
    }
    std::vector<double> rec(knobs.size(), 0.0);
    std::vector<float> kms(knobs.size(), 0.f);
    for (size_t i = 0; i < knobs.size(); i++) {
        int rc = run(knobs[i], nullptr);  // sizes buffers and warms up
        if (!rc) rc = run(knobs[i], &kms[i]);
        if (rc) return rc;
        HIPCHK(hipMemcpy(ids.data(), dI.p, ids.size() * 4, hipMemcpyDeviceToHost), "copy ids");
        rec[i] = score(ids, kstride, kq, gold, kg, m);
    }
    // the knobs that reach the target, fastest first (else the best recall alone)
    std::vector<size_t> keep;
    for (size_t i = 0; i < knobs.size(); i++)
        if (rec[i] >= (double)minrecall) keep.push_back(i);
    if (keep.empty()) {
        keep.push_back((size_t)(std::max_element(rec.begin(), rec.end()) - rec.begin()));
    } else {
        std::sort(keep.begin(), keep.end(), [&](size_t a, size_t b) { return kms[a] < kms[b]; });
        if (keep.size() > 2) keep.resize(2);
    }

    // 2. launch shapes for those knobs, by measured time
    float best_ms = 0.f;
    size_t best_knob = keep[0];
    ssg_launch_params best{sp0, r0, vb0};
    bool have = false;
    for (size_t kn : keep) {
        for (int32_t sp : {8, 16, 24, 32, 48, 64}) {
            for (int32_t rd : {4, 8, 10}) {
                if (rd == 10 && knobs[kn].precision != SSG_BF16) continue;  // BF16 kernels only
                if (rd == 10 && knobs[kn].beamcap > 64) continue;  // one-register beams only
                for (int32_t vb : {-1, 0}) {
                    h->slots_per_cu = sp;
                    h->rounds = rd;
                    h->vbuckets = vb;
                    float ms = 0.f;
                    int rc = run(knobs[kn], nullptr);
                    if (!rc) rc = run(knobs[kn], &ms);
                    if (rc) {
                        h->slots_per_cu = sp0;
                        h->rounds = r0;
                        h->vbuckets = vb0;
                        return rc;
                    }
                    if (!have || ms < best_ms) {
                        have = true;
                        best_ms = ms;
                        best = {sp, rd, vb};
                        best_knob = kn;
                    }
                }
            }
        }
    }
    h->slots_per_cu = best.slots_per_cu;
    h->rounds = best.rounds;
    h->vbuckets = best.vbuckets;
    bp->rerank = knobs[best_knob].rerank;
    bp->expand = knobs[best_knob].expand;
    *lp_out = best;
    if (recall_out) *recall_out = rec[best_knob];
    if (ms_out) *ms_out = best_ms;
    return SSG_OK;
}

}  // extern "C"
