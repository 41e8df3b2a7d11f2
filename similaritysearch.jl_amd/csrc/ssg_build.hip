// ssg_build.hip -- SearchGraph construction (index!) with the searches on the GPU.
//
// Restates _parallel_append_items_loop! (src/searchgraph/insertions.jl:46-79):
// objects are inserted in blocks of parallel_block; every object of a block
//   1. searches the graph built so far (beam search, hints, the current
//      BeamSearch) for its ksearch = neighborhoodsize(ep) nearest objects
//      (find_neighborhood!, src/searchgraph/neighborhood.jl:25-45),
//   2. pushes every earlier object of its own block (the interblock loop :33-38),
//   3. keeps the SAT/HSP-filtered neighbourhood (hsp_proximal_neighborhood_filter!,
//      src/hsp.jl:103-119; neardup is typemin(Float32) so its random branch never
//      fires) as its list;
// then reverse links are connected (connect_reverse_links!, neighborhood.jl:73-89,
// appended in ascending node order) and the callbacks fire when the size crosses
// a logbase_callback power (callbacks.jl:20-25): RandomHints (hints.jl:96-120)
// and OptimizeParameters(MinRecall) (optbs.jl:62-127 -> opt.jl:220-317).
//
// Steps 1-3 run on the GPU (beam_kernel, then block_filter_kernel, one wave per
// object); the lists live on the host and are mirrored on the device in a pool
// with slack (node[i] = {start, deg}), updated per block by a scatter kernel.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "ssg_device.h"
#include "ssg_handle.h"

namespace ssg {

template <int V>
__global__ __launch_bounds__(256) void block_filter_kernel(const float* __restrict__ X, int metric,
                                                           int64_t sp0, int nb, int ks,
                                                           const uint32_t* __restrict__ tmp_ids,
                                                           const float* __restrict__ tmp_d,
                                                           int has_tmp, uint32_t* out_ids,
                                                           int32_t* out_cnt)
{
    const int lane = lane_id(), w = threadIdx.x >> 6;
    const int g = lane >> 4, j = lane & 15;
    const int t = blockIdx.x * 4 + w;
    if (t >= nb) return;
    constexpr int dp = 16 * V;
    const int64_t obj = sp0 + t;
    float qv[V];
    load_run<V>(X + obj * dp + j * V, qv);

    RegQueue<1, false> tmp;
    tmp.init(ks);
    if (has_tmp) {
        const uint32_t id = lane < ks ? tmp_ids[(int64_t)t * ks + lane] : 0u;
        const float dd = lane < ks ? tmp_d[(int64_t)t * ks + lane] : __builtin_inff();
        const uint64_t z = __ballot(lane >= ks || id == 0u);
        const int L = z ? __ffsll((unsigned long long)z) - 1 : 64;
        tmp.d[0] = lane < L ? dd : __builtin_inff();
        tmp.id[0] = lane < L ? id - 1u : 0xffffffffu;
        tmp.len = L;
    }
    // interblock neighbours: earlier objects of the block, ascending
    for (int64_t b = sp0; b < obj; b += 4) {
        const int64_t row = b + g;
        float xv[V];
        if (row < obj) {
            load_run<V>(X + row * dp + j * V, xv);
        } else {
#pragma unroll
            for (int u = 0; u < V; u++) xv[u] = 0.f;
        }
        const float d = finish_dist(metric, group16_sum(run_partial<V>(metric, xv, qv)));
        for (int r = 0; r < 4 && b + r < obj; r++) {
            const float dr = rl_f(d, 16 * r);
            if (tmp.len < tmp.cap || dr < tmp.maximum()) tmp.push(dr, (uint32_t)(b + r), 0, 0);
        }
    }
    // SAT / HSP filter
    RegQueue<1, false> out;
    out.init(ks);
    const int L = tmp.len;
    if (L > 0) {
        out.push(tmp.dist_at(0), tmp.id_at(0), 0, 0);
        for (int c = 1; c < L; c++) {
            const uint32_t pid = tmp.id_at(c);
            const float pd = tmp.dist_at(c);
            float pv[V];
            load_run<V>(X + (int64_t)pid * dp + j * V, pv);
            bool ok = true;
            const int A = out.len;
            for (int a0 = 0; a0 < A && ok; a0 += 4) {
                const int u = a0 + g;
                const uint32_t hid = __shfl(out.id[0], u < A ? u : 0);
                float xh[V];
                if (u < A) {
                    load_run<V>(X + (int64_t)hid * dp + j * V, xh);
                } else {
#pragma unroll
                    for (int q = 0; q < V; q++) xh[q] = 0.f;
                }
                const float dph = finish_dist(metric, group16_sum(run_partial<V>(metric, xh, pv)));
                if (__ballot(u < A && dph < pd)) ok = false;
            }
            if (ok) out.push(pd, pid, 0, 0);
        }
    }
    if (lane < ks) out_ids[(int64_t)t * ks + lane] = lane < out.len ? out.id[0] : 0xffffffffu;
    if (lane == 0) out_cnt[t] = out.len;
}

__global__ void scatter_kernel(uint32_t* pool, const uint32_t* pos, const uint32_t* val, int64_t np,
                               uint2* node, const uint32_t* nidx, const uint2* nval, int64_t nn)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < np) pool[pos[i]] = val[i];
    if (i < nn) node[nidx[i]] = nval[i];
}

__global__ void gather_rows_kernel(const float* __restrict__ X, const float* __restrict__ xnorm,
                                   int dp, const uint32_t* ids, int m, int m_pad, float* qpad,
                                   float* qnorm)
{
    const int r = blockIdx.x;
    for (int c = threadIdx.x; c < dp; c += blockDim.x)
        qpad[(int64_t)r * dp + c] = r < m ? X[(int64_t)ids[r] * dp + c] : 0.f;
    if (threadIdx.x == 0) qnorm[r] = r < m ? xnorm[ids[r]] : 0.f;
    (void)m_pad;
}

static hipError_t launch_block_filter(const float* X, int dp, int metric, int64_t sp0, int nb,
                                      int ks, const uint32_t* tids, const float* td, int has_tmp,
                                      uint32_t* out, int32_t* cnt, hipStream_t s)
{
    const dim3 grid((unsigned)((nb + 3) / 4)), block(256);
    switch (dp / 16) {
#define SSG_BF(VV) \
    case VV: hipLaunchKernelGGL((block_filter_kernel<VV>), grid, block, 0, s, X, metric, sp0, nb, ks, tids, td, has_tmp, out, cnt); break;
        SSG_BF(1) SSG_BF(2) SSG_BF(3) SSG_BF(4) SSG_BF(5) SSG_BF(6) SSG_BF(7) SSG_BF(8) SSG_BF(12) SSG_BF(16) SSG_BF(24) SSG_BF(32) SSG_BF(48) SSG_BF(64)
#undef SSG_BF
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace ssg

using namespace ssg;

namespace {

struct Rng {  // splitmix64; rand(1:n) -> 1 + next % n
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    int64_t rand_1n(int64_t n) { return 1 + (int64_t)(next() % (uint64_t)n); }
};

int32_t nb_size(const ssg_build_params* p, int64_t n)
{
    if (n == 0) return p->nb_minsize;
    return (int32_t)ceilf((float)p->nb_minsize + logf((float)n) / logf(p->nb_logbase));
}

bool cb_fires(const ssg_build_params* p, int64_t n, int64_t m)
{
    if (n < p->cb_starting) return false;
    const float lb = logf(p->cb_logbase);
    return (int64_t)ceilf(logf((float)n) / lb) != (int64_t)ceilf(logf((float)m) / lb);
}

// RandomHints (src/searchgraph/hints.jl:96-120); V gains {objID} ∪ N(objID).
void random_hints(const std::vector<std::vector<uint32_t>>& lists, int64_t n, float logbase,
                  Rng& rng, std::vector<uint32_t>& out)
{
    const int64_t m = (int64_t)ceilf(logf((float)n) / logf(logbase));
    std::vector<uint8_t> V((size_t)n + 2, 0);
    out.clear();
    for (int64_t it = 0; it < n; it++) {
        const int64_t obj = rng.rand_1n(n);
        if (V[obj]) continue;
        const auto& N = lists[obj - 1];
        if (N.size() <= 2) continue;
        V[obj] = 1;
        for (uint32_t c : N) V[c + 1] = 1;
        out.push_back((uint32_t)obj);
        if ((int64_t)out.size() >= m) break;
    }
}

// MinRecall cost of opt.jl:281 over the BeamSearchSpace grid (bsize 2:2:16,
// Δ 0.9:0.025:1.1) instead of SearchModels' evolutionary search; per-query
// costs stand for the per-batch counters.  Gold comes from the exhaustive path
// over the first n_eff objects.  qp/qn: m_pad padded query rows and norms.
// On success algo gets (bsize, Δ, maxvisits = ceil(2 max visited)) (optbs.jl:141-145).
int optimize_grid(ssg_handle* h, const float* qp, const float* qn, int m, int32_t ks,
                  float minrecall, ssg_beam_params& algo, int64_t n_eff, hipStream_t s,
                  double* recall_out)
{
    DevBuf gi, gd, ri, rd, cd;
    HIPCHK(gi.ensure((size_t)m * ks * 4), "alloc");
    HIPCHK(gd.ensure((size_t)m * ks * 4), "alloc");
    HIPCHK(ri.ensure((size_t)m * ks * 4), "alloc");
    HIPCHK(rd.ensure((size_t)m * ks * 4), "alloc");
    HIPCHK(cd.ensure((size_t)m * 8), "alloc");
    int rc = ssg_exh_run(h, nullptr, qp, qn, m, ks, ks, gi.as<uint32_t>(), gd.as<float>(), s, n_eff);
    if (rc) return rc;
    std::vector<uint32_t> gold((size_t)m * ks), res((size_t)m * ks);
    std::vector<int64_t> cost((size_t)m);
    HIPCHK(hipMemcpyAsync(gold.data(), gi.p, gold.size() * 4, hipMemcpyDeviceToHost, s), "copy");
    HIPCHK(hipStreamSynchronize(s), "sync");
    std::vector<int> glen((size_t)m);
    for (int q = 0; q < m; q++) {
        int l = 0;
        while (l < ks && gold[(size_t)q * ks + l] != 0) l++;
        glen[q] = l;
    }
    const int NB = 8, ND = 9;
    struct Perf { double recall, vmean; int64_t vmax; bool valid; };
    std::vector<Perf> P(NB * ND);
    for (int bi = 0; bi < NB; bi++)
        for (int di = 0; di < ND; di++) {
            ssg_beam_params bp{2 + 2 * bi, (float)((900.0 + 25.0 * di) / 1000.0),
                               2 * algo.maxvisits, algo.beamcap};
            rc = ssg_beam_run(h, qp, h->dp, m, ks, ks, &bp, nullptr, ri.as<uint32_t>(),
                              rd.as<float>(), cd.as<int64_t>(), nullptr, s, false, n_eff, true);
            if (rc) return rc;
            HIPCHK(hipMemcpyAsync(res.data(), ri.p, res.size() * 4, hipMemcpyDeviceToHost, s), "copy");
            HIPCHK(hipMemcpyAsync(cost.data(), cd.p, cost.size() * 8, hipMemcpyDeviceToHost, s), "copy");
            HIPCHK(hipStreamSynchronize(s), "sync");
            Perf& pf = P[bi * ND + di];
            int64_t full = 0, vmax = 0, act = 0;
            double rs = 0.0, vs = 0.0;
            for (int q = 0; q < m; q++) {
                int l = 0;
                while (l < ks && res[(size_t)q * ks + l] != 0) l++;
                if (l == ks) full++;
                int c = 0;
                for (int a = 0; a < glen[q]; a++)
                    for (int b = 0; b < l; b++)
                        if (gold[(size_t)q * ks + a] == res[(size_t)q * ks + b]) {
                            c++;
                            break;
                        }
                rs += glen[q] > 0 ? (double)c / glen[q] : 0.0;
                if (cost[q] != 0) {
                    vs += (double)cost[q];
                    act++;
                    vmax = std::max(vmax, cost[q]);
                }
            }
            pf.valid = full > 3;
            pf.recall = rs / m;
            pf.vmean = act ? vs / act : 0.0;
            pf.vmax = vmax;
        }
    double M = 0.0;
    for (const Perf& pf : P)
        if (pf.valid) M = std::max(M, (double)pf.vmax);
    int best = -1;
    double bestc = INFINITY;
    for (int i = 0; i < NB * ND; i++) {
        if (!P[i].valid) continue;
        const double c = P[i].recall < (double)minrecall ? 1.0 + ((double)minrecall - P[i].recall)
                                                         : (M > 0 ? P[i].vmean / M : 0.0);
        if (c < bestc) {
            bestc = c;
            best = i;
        }
    }
    if (best < 0) return ssg_fail(SSG_E_STATE, "optimisation found no usable configuration");
    algo.bsize = 2 + 2 * (best / ND);
    algo.delta = (float)((900.0 + 25.0 * (best % ND)) / 1000.0);
    algo.maxvisits = std::max<int64_t>(1, (int64_t)ceil(2.0 * (double)P[best].vmax));
    if (recall_out) *recall_out = P[best].recall;
    return SSG_OK;
}

struct Builder {
    ssg_handle* h;
    const ssg_build_params* p;
    hipStream_t s;
    std::vector<uint32_t>& seg_start;
    std::vector<uint32_t>& seg_cap;
    std::vector<uint32_t>& dev_len;
    std::vector<uint32_t> pos, val, nidx;
    std::vector<uint2> nval;

    int upload_hints()
    {
        const auto& H = h->hints_host;
        HIPCHK(h->hints.ensure(std::max<size_t>(H.size(), 1) * 4), "alloc hints");
        if (!H.empty())
            HIPCHK(hipMemcpyAsync(h->hints.p, H.data(), H.size() * 4, hipMemcpyHostToDevice, s), "copy hints");
        HIPCHK(hipStreamSynchronize(s), "sync");
        h->nhints = (int64_t)H.size();
        return SSG_OK;
    }

    // Bring the device copy of the changed lists up to date.
    int flush(const std::vector<uint32_t>& changed)
    {
        pos.clear();
        val.clear();
        nidx.clear();
        nval.clear();
        for (uint32_t c : changed) {
            const auto& L = h->lists[c];
            const uint32_t newlen = (uint32_t)L.size();
            uint32_t from = dev_len[c];
            if (newlen > seg_cap[c]) {
                seg_cap[c] = std::max<uint32_t>(16u, 2u * newlen);
                seg_start[c] = (uint32_t)h->pool_used;
                h->pool_used += seg_cap[c];
                from = 0;
            }
            for (uint32_t t = from; t < newlen; t++) {
                pos.push_back(seg_start[c] + t);
                val.push_back(L[t]);
            }
            nidx.push_back(c);
            nval.push_back(make_uint2(seg_start[c], newlen));
            dev_len[c] = newlen;
        }
        if ((size_t)h->pool_used * 4 > h->adj.cap) {
            const size_t want = std::max<size_t>((size_t)h->pool_used * 4 * 3 / 2, h->adj.cap * 2);
            HIPCHK(h->adj.grow(want, h->adj.cap, s), "grow adjacency pool");
        }
        const size_t np = pos.size(), nn = nidx.size();
        const size_t bytes = np * 8 + nn * 12 + 64;
        HIPCHK(h->bPairs.ensure(bytes), "alloc update buffer");
        uint8_t* base = h->bPairs.as<uint8_t>();
        uint32_t* dpos = (uint32_t*)base;
        uint32_t* dval = dpos + np;
        uint2* dnval = (uint2*)(((uintptr_t)(dval + np) + 15) & ~(uintptr_t)15);
        uint32_t* dnidx = (uint32_t*)(dnval + nn);
        if (np) {
            HIPCHK(hipMemcpyAsync(dpos, pos.data(), np * 4, hipMemcpyHostToDevice, s), "copy");
            HIPCHK(hipMemcpyAsync(dval, val.data(), np * 4, hipMemcpyHostToDevice, s), "copy");
        }
        if (nn) {
            HIPCHK(hipMemcpyAsync(dnval, nval.data(), nn * 8, hipMemcpyHostToDevice, s), "copy");
            HIPCHK(hipMemcpyAsync(dnidx, nidx.data(), nn * 4, hipMemcpyHostToDevice, s), "copy");
        }
        const size_t mx = std::max(np, nn);
        if (mx) {
            hipLaunchKernelGGL(scatter_kernel, dim3((unsigned)((mx + 255) / 256)), dim3(256), 0, s,
                               h->adj.as<uint32_t>(), dpos, dval, (int64_t)np, h->node.as<uint2>(),
                               dnidx, dnval, (int64_t)nn);
            HIPCHK(hipGetLastError(), "scatter");
        }
        return SSG_OK;
    }

    // optimize_index!(index, ctx, MinRecall(r)) on a sample of database objects.
    int optimize(int64_t len, const std::vector<uint32_t>& qids, int32_t ks, float minrecall,
                 ssg_beam_params& algo)
    {
        const int m = (int)qids.size();
        const int m_pad = (m + EXH_QT - 1) / EXH_QT * EXH_QT;
        DevBuf qp, qn, qi;
        HIPCHK(qp.ensure((size_t)m_pad * h->dp * 4), "alloc");
        HIPCHK(qn.ensure((size_t)m_pad * 4), "alloc");
        HIPCHK(qi.ensure((size_t)m * 4), "alloc");
        HIPCHK(hipMemcpyAsync(qi.p, qids.data(), (size_t)m * 4, hipMemcpyHostToDevice, s), "copy");
        hipLaunchKernelGGL(gather_rows_kernel, dim3((unsigned)m_pad), dim3(64), 0, s,
                           h->X.as<float>(), h->xnorm.as<float>(), h->dp, qi.as<uint32_t>(), m,
                           m_pad, qp.as<float>(), qn.as<float>());
        HIPCHK(hipGetLastError(), "gather rows");
        ssg_beam_params cand = algo;
        double rec = -1.0;
        int rc = optimize_grid(h, qp.as<float>(), qn.as<float>(), m, ks, minrecall, cand, len, s, &rec);
        if (rc == SSG_OK) algo = cand;
        return rc == SSG_E_STATE ? SSG_OK : rc;  // no usable configuration: keep the current one
    }

    int callbacks(int64_t len, Rng& rng)
    {
        if (p->hints_logbase > 0) {
            random_hints(h->lists, len, p->hints_logbase, rng, h->hints_host);
            int rc = upload_hints();
            if (rc) return rc;
        }
        if (p->opt_minrecall > 0) {
            std::vector<uint32_t> smp;
            for (int i = 0; i < p->opt_numqueries; i++) {
                const uint32_t x = (uint32_t)rng.rand_1n(len);
                if (std::find(smp.begin(), smp.end(), x) == smp.end()) smp.push_back(x);
            }
            for (auto& x : smp) x -= 1;
            int rc = optimize(len, smp, p->opt_ksearch, p->opt_minrecall, h->algo);
            if (rc) return rc;
        }
        return SSG_OK;
    }
};

}  // namespace

extern "C" {

static int check_build_params(const ssg_build_params* p, int64_t n)
{
    if (p->parallel_block < 2) return ssg_fail(SSG_E_ARG, "the GPU builder needs parallel_block >= 2");
    if (p->beamcap < 1 || p->beamcap > 256) return ssg_fail(SSG_E_ARG, "beamcap must be in 1..256");
    if (!(p->nb_logbase > 1.f) || p->nb_minsize < 0) return ssg_fail(SSG_E_ARG, "bad Neighborhood");
    if (nb_size(p, n) > 64) return ssg_fail(SSG_E_LIMIT, "neighborhood size above 64");
    if (p->opt_minrecall > 0 && (p->opt_ksearch < 1 || p->opt_ksearch > EXH_KC || p->opt_numqueries < 1))
        return ssg_fail(SSG_E_ARG, "bad OptimizeParameters");
    return SSG_OK;
}

// The insertion loop of index! (src/searchgraph/insertions.jl:46-79, 96-113) over the objects
// h->b_len + 1 .. n with the parameters and generator state the handle keeps: blocks of
// parallel_block objects from the first uninserted one, so a build in several calls has the
// block boundaries of the reference's index! after each append_items!.
static int build_continue(ssg_handle* h)
{
    const int64_t n = h->n;
    const ssg_build_params* p = &h->b_params;
    if (!h->X.p || n <= 0) return ssg_fail(SSG_E_STATE, "index! needs a non-empty database");
    int rc = check_build_params(p, n);
    if (rc) return rc;
    hipStream_t s = h->stream;
    const int64_t n0 = h->b_len;
    Builder B{h, p, s, h->b_seg_start, h->b_seg_cap, h->b_dev_len, {}, {}, {}, {}};
    if (n0 == 0) {
        h->lists.assign((size_t)n, {});
        h->hints_host.clear();
        h->algo = ssg_beam_params{4, 1.0f, 1000000, p->beamcap};
        B.seg_start.assign((size_t)n, 0);
        B.seg_cap.assign((size_t)n, 0);
        B.dev_len.assign((size_t)n, 0);
        HIPCHK(h->node.ensure((size_t)n * sizeof(uint2)), "alloc node table");
        HIPCHK(launch_zero_u32(h->node.as<unsigned int>(), 2 * n, s), "zero node table");
        h->adj.release();
        HIPCHK(h->adj.ensure((size_t)n * 16 * 4), "alloc adjacency pool");
        h->pool_used = 0;
        h->nhints = 0;
        rc = B.upload_hints();
        if (rc) return rc;
    } else {
        h->lists.resize((size_t)n);
        B.seg_start.resize((size_t)n, 0);
        B.seg_cap.resize((size_t)n, 0);
        B.dev_len.resize((size_t)n, 0);
        HIPCHK(h->node.grow((size_t)n * sizeof(uint2), (size_t)n0 * sizeof(uint2), s), "grow node table");
        HIPCHK(launch_zero_u32(h->node.as<unsigned int>() + 2 * n0, 2 * (n - n0), s), "zero node table");
    }
    h->has_dups = 0;
    h->has_graph = true;
    ssg_query_layout_stale(h);
    const int PB = p->parallel_block;
    const int ksmax = nb_size(p, n);
    HIPCHK(h->bIds.ensure((size_t)PB * ksmax * 4), "alloc");
    HIPCHK(h->bD.ensure((size_t)PB * ksmax * 4), "alloc");
    HIPCHK(h->bOut.ensure((size_t)PB * ksmax * 4), "alloc");
    HIPCHK(h->bCnt.ensure((size_t)PB * 4), "alloc");
    std::vector<uint32_t> outs((size_t)PB * ksmax);
    std::vector<int32_t> cnts((size_t)PB);
    std::vector<uint8_t> mark((size_t)n, 0);
    std::vector<uint32_t> changed, Lc;
    Rng rng{h->b_rng};
    int64_t len = n0;
    for (int64_t sp = n0 + 1; sp <= n;) {
        const int64_t ep = std::min<int64_t>(n, sp + PB - 1);
        const int nb = (int)(ep - sp + 1);
        const int ks = nb_size(p, ep);
        if (len > 0) {
            rc = ssg_beam_run(h, h->X.as<float>() + (sp - 1) * h->dp, h->dp, nb, ks, ks, &h->algo,
                              nullptr, h->bIds.as<uint32_t>(), h->bD.as<float>(), nullptr, nullptr,
                              s, false, len, true);
            if (rc) return rc;
        }
        HIPCHK(launch_block_filter(h->X.as<float>(), h->dp, h->metric, sp - 1, nb, ks,
                                   h->bIds.as<uint32_t>(), h->bD.as<float>(), len > 0 ? 1 : 0,
                                   h->bOut.as<uint32_t>(), h->bCnt.as<int32_t>(), s),
               "block filter");
        HIPCHK(hipMemcpyAsync(outs.data(), h->bOut.p, (size_t)nb * ks * 4, hipMemcpyDeviceToHost, s), "copy");
        HIPCHK(hipMemcpyAsync(cnts.data(), h->bCnt.p, (size_t)nb * 4, hipMemcpyDeviceToHost, s), "copy");
        HIPCHK(hipStreamSynchronize(s), "sync");
        changed.clear();
        for (int t = 0; t < nb; t++) {
            auto& L = h->lists[(size_t)(sp - 1 + t)];
            L.insert(L.end(), outs.begin() + (size_t)t * ks, outs.begin() + (size_t)t * ks + cnts[t]);
        }
        // connect_reverse_links!(adj, sp, ep)
        for (int64_t node = sp; node <= ep; node++)
            for (uint32_t rel : h->lists[(size_t)node - 1])
                if ((int64_t)rel + 1 < sp) {
                    h->lists[rel].push_back((uint32_t)(node - 1));
                    if (!mark[rel]) {
                        mark[rel] = 1;
                        changed.push_back(rel);
                    }
                }
        Lc.resize((size_t)nb);
        for (int t = 0; t < nb; t++) Lc[t] = (uint32_t)h->lists[(size_t)(sp - 1 + t)].size();
        for (int t = 0; t < nb; t++) {
            const int64_t node = sp + t;
            for (uint32_t q = 0; q < Lc[t]; q++) {
                const uint32_t rel = h->lists[(size_t)node - 1][q];
                if ((int64_t)rel + 1 >= sp && (int64_t)rel + 1 != node)
                    h->lists[rel].push_back((uint32_t)(node - 1));
            }
        }
        for (int64_t node = sp; node <= ep; node++) {
            if (!mark[node - 1]) {
                mark[node - 1] = 1;
                changed.push_back((uint32_t)(node - 1));
            }
        }
        rc = B.flush(changed);
        if (rc) return rc;
        for (uint32_t c : changed) mark[c] = 0;
        if (p->verbose > 0) {
            const int64_t step = std::max<int64_t>(1, n * p->verbose / 100);
            if (ep / step != len / step || ep == n) {
                fprintf(stderr, "index!: %lld / %lld inserted, adjacency pool %lld slots (lists and their "
                        "free room), algo bsize=%d delta=%.3f maxvisits=%lld\n",
                        (long long)ep, (long long)n, (long long)h->pool_used, h->algo.bsize,
                        h->algo.delta, (long long)h->algo.maxvisits);
                fflush(stderr);
            }
        }
        len = ep;
        h->b_len = len;
        if (cb_fires(p, sp, ep)) {
            HIPCHK(hipStreamSynchronize(s), "sync");
            rc = B.callbacks(len, rng);
            if (rc) return rc;
        }
        h->b_rng = rng.s;
        sp = ep + 1;
    }
    HIPCHK(hipStreamSynchronize(s), "sync");
    int64_t nnz = 0;
    int dups = 0;
    for (const auto& L : h->lists) nnz += (int64_t)L.size();
    std::vector<uint32_t> tmp;
    for (const auto& L : h->lists) {
        if (L.size() < 2) continue;
        tmp.assign(L.begin(), L.end());
        std::sort(tmp.begin(), tmp.end());
        if (std::adjacent_find(tmp.begin(), tmp.end()) != tmp.end()) {
            dups = 1;
            break;
        }
    }
    h->nnz = nnz;
    h->has_dups = dups;
    ssg_query_layout_stale(h);
    return SSG_OK;
}

int ssg_graph_begin(ssg_handle* h, const ssg_build_params* p)
{
    const int rc = check_build_params(p, std::max<int64_t>(h->n, 1));
    if (rc) return rc;
    h->b_params = *p;
    h->b_len = 0;
    h->b_rng = p->seed;
    h->b_valid = true;
    return SSG_OK;
}

int ssg_build_continue(ssg_handle* h)
{
    if (!h->b_valid) return ssg_fail(SSG_E_STATE, "the graph was not built by index! on this handle");
    const int rc = build_continue(h);
    if (rc) h->b_valid = false;  // an interrupted insertion loop cannot be continued
    return rc;
}

int ssg_build_graph(ssg_handle* h, const ssg_build_params* p)
{
    if (!h || !p) return ssg_fail(SSG_E_ARG, "null argument");
    std::lock_guard<std::mutex> lk(h->mu);
    HIPCHK(hipSetDevice(h->device), "hipSetDevice");
    if (!h->X.p || h->n <= 0) return ssg_fail(SSG_E_STATE, "index! needs a non-empty database");
    const int rc = check_build_params(p, h->n);
    if (rc) return rc;
    ssg_graph_begin(h, p);
    return ssg_build_continue(h);
}

int ssg_append_items_graph(ssg_handle* h, const float* X, int64_t m)
{
    if (!h) return ssg_fail(SSG_E_ARG, "null handle");
    std::lock_guard<std::mutex> lk(h->mu);
    HIPCHK(hipSetDevice(h->device), "hipSetDevice");
    if (m < 0 || (m > 0 && !X)) return ssg_fail(SSG_E_ARG, "append_items!: bad arguments");
    if (!h->has_graph || !h->b_valid || h->b_len != h->n)
        return ssg_fail(SSG_E_STATE, "append_items!: the handle's graph was not built by index! here");
    if (check_build_params(&h->b_params, h->n + m))
        return ssg_fail(SSG_E_LIMIT, "append_items!: neighborhood size above 64 at the new size");
    int rc = ssg_append_rows(h, X, m, false);
    if (rc) return rc;
    return ssg_build_continue(h);
}

int ssg_optimize_minrecall(ssg_handle* h, const float* Q, int64_t m, int32_t ksearch,
                           float minrecall, ssg_beam_params* algo_inout, double* recall)
{
    if (!h || !Q || !algo_inout) return ssg_fail(SSG_E_ARG, "null argument");
    std::lock_guard<std::mutex> lk(h->mu);
    HIPCHK(hipSetDevice(h->device), "hipSetDevice");
    if (!h->has_graph) return ssg_fail(SSG_E_STATE, "no graph");
    if (m < 1 || ksearch < 1 || ksearch > EXH_KC) return ssg_fail(SSG_E_ARG, "bad optimisation queries");
    hipStream_t s = h->stream;
    // queries are appended as rows of a temporary padded matrix via the exhaustive path
    const int64_t m_pad = (m + EXH_QT - 1) / EXH_QT * EXH_QT;
    DevBuf dq, qp, qn;
    HIPCHK(dq.ensure((size_t)m * h->dim * 4), "alloc");
    HIPCHK(qp.ensure((size_t)m_pad * h->dp * 4), "alloc");
    HIPCHK(qn.ensure((size_t)m_pad * 4), "alloc");
    HIPCHK(hipMemcpyAsync(dq.p, Q, (size_t)m * h->dim * 4, hipMemcpyHostToDevice, s), "copy");
    HIPCHK(launch_fill_f32(qp.as<float>(), m_pad * h->dp, 0.f, s), "fill");
    HIPCHK(launch_fill_f32(qn.as<float>(), m_pad, 0.f, s), "fill");
    HIPCHK(launch_pad_rows(dq.as<float>(), m, h->dim, qp.as<float>(), h->dp, qn.as<float>(), s), "pad");
    double rec = -1.0;
    int rc = optimize_grid(h, qp.as<float>(), qn.as<float>(), (int)m, ksearch, minrecall, *algo_inout,
                           h->n, s, &rec);
    if (rc) return rc;
    if (recall) *recall = rec;
    h->algo = *algo_inout;
    return SSG_OK;
}

}  // extern "C"
