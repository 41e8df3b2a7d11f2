// ssg_beam.h -- wavefront-per-query graph beam search for gfx950 (kernel templates;
// the instantiations are spread over ssg_beam_v*.hip so they compile in parallel).
//
// Restates search(bs::BeamSearch, ...) of src/searchgraph/beamsearch.jl:88-110:
//   n == 0 -> nothing; n < 64 -> brute force over 1..n (no visited marks);
//   otherwise beamsearch_init (:26-34, every hint through enqueue_item! :12-18,
//   then ids 1:ceil(log2(1+n)) if the result is still empty) and
//   beamsearch_inner_beam (:44-73): a KnnSorted beam of min(bsize, beamcap)
//   slots seeded with nearest(res); pop_min, expand the CSR list, test-and-set
//   visited, evaluate, push into res, stop once costdists > maxvisits, and admit
//   the child into the beam when deg(child) > 1 and d <= Δ * maximum(res).
//
// GPU shape: one wave64 per query, persistent over an atomic work counter.
//   * res (k <= 64, or k <= 256 in four registers) and the beam (<= 256 slots) live
//     in registers (RegQueue).
//   * Neighbour lists are expanded 64 ids at a time: one coalesced id load, one
//     visited test per id (below), then the unvisited ids are compacted in LDS and
//     their vectors gathered in batches (one child per 16-lane group per round,
//     8 rounds = 32 children), with
//     the child degrees loaded in the same round trip.  Every row index of a
//     batch is clamped into the batch, so all its loads are unconditional and in
//     flight together before the first one is consumed.
//   * Child distances are transposed through LDS to one lane per child; the
//     order-dependent part (pushes into res and beam, the maxvisits cut) is then
//     replayed serially over the children in list order with readlane, skipping
//     children that provably change neither queue.
//   * Visited set: a per-wave table in LDS (vnb buckets of eight 16-bit tags; each
//     object has one bucket and one tag) takes a query's marks while their buckets
//     have room; the others go to the wave slot's epoch bytes in HBM (one byte per
//     (wave slot, object); a query takes the next epoch of its slot, the slot is
//     cleared once every 255 queries, a byte equal to the current epoch means
//     visited), which are read only for ids whose bucket is full.  vnb = 0 keeps
//     every mark in HBM.
#pragma once
#include <algorithm>

#include "ssg_device.h"
#include "ssg_kernels.h"

namespace ssg {

constexpr int WAVES = 4;            // waves per block

enum { MODE_HINT = 0, MODE_EXPAND = 1, MODE_BRUTE = 2 };

__device__ __forceinline__ void lds_order()
{
    __builtin_amdgcn_wave_barrier();
    asm volatile("" ::: "memory");
}

// Wait for every outstanding vector memory operation of this wave (gfx9 vmcnt
// covers stores), so visited marks are visible to the next visited loads.
__device__ __forceinline__ void vm_drain()
{
    __builtin_amdgcn_s_waitcnt(0x0F70);
    asm volatile("" ::: "memory");
}

// Visited bytes: all reads, marks, unmarks and the slot clear use the same
// (workgroup-scope, relaxed) atomic access, so they share one cache policy and a
// clear is never hidden behind a copy held at another cache level.
__device__ __forceinline__ uint8_t vis_get(const uint8_t* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void vis_put(uint8_t* p, uint8_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// PREC 0: canonical distance on the FP32 vectors.  PREC 1: BF16 vectors with FP32
// arithmetic during the search, then an exact (canonical) rerank of the result list.
// R: gather rounds per batch (the launch autotuner picks 4, 8 or, in the BF16 form, 10;
// fewer rounds hold fewer rows in registers, so more waves may fit a SIMD, each with fewer
// loads in flight; more rounds take an expansion's children in fewer round trips).
// NR: registers of the result queue (k <= 64 * NR; 4 only in the FP32 form).
template <int V, int NB, int PREC, int R = 8, int NR = 1>
struct Searcher {
    static constexpr int ROUNDS = R;
    static constexpr int BATCH = 4 * ROUNDS;           // children per batch
    const BeamArgs& a;
    float qv[V];
    RegQueue<NR, false> res;
    RegQueue<NB, true> beam;
    int64_t cd_init;   // distance evaluations charged by enqueue_item!
    int64_t cd_inner;  // costdists of beamsearch_inner_beam
    int64_t cb;        // costblocks
    uint32_t epoch;
    uint8_t* vis;  // wave-private slot; every access is a workgroup-scope relaxed atomic
    uint4* vt;         // wave-private LDS visited table: vnb buckets of 8 tags
    uint32_t vnb;      // 0: every mark goes to vis
    uint32_t vhb;      // objects are named by vhb bits
    uint32_t* cidbuf;  // 64 ids, wave-private LDS
    float* dbuf;       // 64 distances, wave-private LDS
    uint32_t* stbuf;   // 64 list starts of the compacted children (adjn layout), LDS
    uint32_t* dgbuf;   // 64 degrees of the compacted children (adjn layout), LDS
    int lane, g, j;

    __device__ __forceinline__ Searcher(const BeamArgs& args) : a(args) {}

    // ---- LDS visited table ----
    // Object c < 2^hb (15 <= hb <= 24) maps bijectively to h = c * 0x9E3779B1 mod 2^hb;
    // its bucket is h * vnb >> hb (vnb >= 2^(hb-15), so a bucket covers at most 2^15
    // consecutive h) and its tag (h mod 2^15) + 1, which with the bucket names c.  A
    // bucket holds 8 tags in 16 bytes (0 = free, VT_ERASED = an erased tag, occupied for
    // good), filled left to right and never emptied.  A mark goes to the bucket while it has a free slot and to
    // the HBM epoch byte otherwise, so c is visited iff its tag is in its bucket, or its
    // bucket is full and its byte holds the epoch.
    static constexpr uint32_t VT_ERASED = 0xFFFFu;
    __device__ __forceinline__ void vt_key(uint32_t c, uint32_t& b, uint32_t& tag) const
    {
        const uint32_t h = (c * 0x9E3779B1u) & ((1u << vhb) - 1u);
        b = (uint32_t)(((uint64_t)h * vnb) >> vhb);
        tag = (h & 0x7FFFu) + 1u;
    }
    __device__ __forceinline__ static bool vt_has(uint32_t w, uint32_t tag)
    {
        return (w & 0xFFFFu) == tag || (w >> 16) == tag;
    }
    __device__ __forceinline__ static bool vt_fullw(uint32_t w)
    {
        return (w & 0xFFFFu) != 0u && (w >> 16) != 0u;
    }
    // 1: visited, 0: not visited, 2: not in the table and the bucket is full (ask HBM)
    __device__ __forceinline__ int vt_test(uint32_t c) const
    {
        uint32_t b, tag;
        vt_key(c, b, tag);
        const uint4 e = vt[b];
        if (vt_has(e.x, tag) || vt_has(e.y, tag) || vt_has(e.z, tag) || vt_has(e.w, tag)) return 1;
        return vt_fullw(e.x) && vt_fullw(e.y) && vt_fullw(e.z) && vt_fullw(e.w) ? 2 : 0;
    }
    // Insert c (absent): claim the first free slot by compare-and-swap of its 32-bit word
    // (lanes of one call may race for a bucket; a loser looks again).  False: bucket full.
    __device__ __forceinline__ bool vt_insert(uint32_t c)
    {
        uint32_t b, tag;
        vt_key(c, b, tag);
        uint32_t* wp = reinterpret_cast<uint32_t*>(vt + b);
        for (;;) {
            const uint4 e = vt[b];
            const uint32_t ws[4] = {e.x, e.y, e.z, e.w};
            int q = -1;
            uint32_t ow = 0u, nw = 0u;
#pragma unroll
            for (int t = 3; t >= 0; t--) {
                if ((ws[t] & 0xFFFFu) == 0u) {
                    q = t;
                    ow = ws[t];
                    nw = ws[t] | tag;
                } else if ((ws[t] >> 16) == 0u) {
                    q = t;
                    ow = ws[t];
                    nw = ws[t] | (tag << 16);
                }
            }
            if (q < 0) return false;
            if (atomicCAS(wp + q, ow, nw) == ow) return true;
        }
    }
    // Erase c: its tag becomes VT_ERASED (still occupied).  False: c is not in the table.
    __device__ __forceinline__ bool vt_erase(uint32_t c)
    {
        uint32_t b, tag;
        vt_key(c, b, tag);
        uint16_t* tp = reinterpret_cast<uint16_t*>(vt + b);
#pragma unroll
        for (int t = 0; t < 8; t++)
            if (tp[t] == tag) {
                tp[t] = (uint16_t)VT_ERASED;
                return true;
            }
        return false;
    }
    // Start of a query: an empty table.
    __device__ __forceinline__ void vt_reset()
    {
        if (!vnb) return;
        for (uint32_t b = (uint32_t)lane; b < vnb; b += 64) vt[b] = make_uint4(0u, 0u, 0u, 0u);
        lds_order();
    }
    // Has c been visited by this query (the same answer in every lane)?
    __device__ __forceinline__ bool visited_uniform(uint32_t c)
    {
        int st = vnb ? vt_test(c) : 2;
        if (st == 2) {
            vm_drain();
            st = vis_get(vis + c) == (uint8_t)epoch ? 1 : 0;
        }
        return rfl_u((uint32_t)st) == 1u;
    }
    // Mark one id (the same c in every lane) visited.
    __device__ __forceinline__ void mark_uniform(uint32_t c)
    {
        bool in_t = false;
        if (vnb) {
            if (lane == 0) in_t = vt_insert(c);
            lds_order();
        }
        if (!in_t && lane == 0) vis_put(vis + c, (uint8_t)epoch);
        vm_drain();
    }

    // Evaluate cidbuf[ub .. ub+cnt) (1 <= cnt <= BATCH).  Child u = 4r + q is gathered
    // by row q in round r; rows past the batch repeat its last child, so every load is
    // unconditional.  Returns the distance of child `lane` (meaningful for lane < cnt);
    // lane u also receives the child's id, list start and degree.
    __device__ __forceinline__ float eval_batch(int ub, int cnt, uint32_t& cid, uint32_t& st,
                                                uint32_t& dg, bool need_deg)
    {
        const int dp = 16 * V;
        cid = cidbuf[ub + (lane < cnt ? lane : cnt - 1)];
        uint32_t cr[ROUNDS];
#pragma unroll
        for (int r = 0; r < ROUNDS; r++) {
            const int u = 4 * r + g;
            cr[r] = cidbuf[ub + (u < cnt ? u : cnt - 1)];
        }
        uint2 nd = make_uint2(0u, 0u);
        if (need_deg) {
            if (a.adjn) {
                const int u = ub + (lane < cnt ? lane : cnt - 1);
                nd = make_uint2(stbuf[u], dgbuf[u]);
            } else {
                nd = a.node[cid];
            }
        }
        RunWords<V, PREC> xw[ROUNDS];
#pragma unroll
        for (int r = 0; r < ROUNDS; r++) {
            if (4 * r < cnt) {
                if constexpr (PREC == 1)
                    load_words<V>(a.Xh + (size_t)cr[r] * dp + j * V, xw[r]);
                else
                    load_words<V>(a.X + (size_t)cr[r] * dp + j * V, xw[r]);
            }
        }
#pragma unroll
        for (int r = 0; r < ROUNDS; r++) {
            if (4 * r < cnt) {
                float xv[V];
                unpack_words<V>(xw[r], xv);
                float d;
                if constexpr (PREC == 1)
                    d = finish_f32(a.metric, group16_tree_f32(run_partial_f32<V>(a.metric, xv, qv)));
                else
                    d = finish_dist(a.metric, group16_tree(run_partial<V>(a.metric, xv, qv)));
                if (j == 15) dbuf[4 * r + g] = d;
            }
        }
        lds_order();
        const float dl = dbuf[lane];
        st = nd.x;
        dg = nd.y;
        return dl;
    }

    // One chunk of up to 64 candidate ids (lane-indexed, `valid` lanes).  Returns
    // true when the maxvisits cut ends the search.
    __device__ __forceinline__ bool chunk(int mode, uint32_t cid, bool valid, bool dedupe,
                                          uint2 cnd = make_uint2(0u, 0u))
    {
        valid = valid && (int64_t)cid < a.n;
        bool unv = valid;
        if (mode != MODE_BRUTE) {
            if (vnb) {
                int st = valid ? vt_test(cid) : 1;
                if (__ballot(st == 2)) {
                    vm_drain();
                    if (st == 2) st = vis_get(vis + cid) == (uint8_t)epoch ? 1 : 0;
                }
                unv = st == 0;
            } else {
                vm_drain();
                if (valid) unv = vis_get(vis + cid) != (uint8_t)epoch;
            }
            if (dedupe) {
                // a later copy of an id in the chunk is visited by its first copy; only
                // unvisited lanes can be first copies (a visited id is visited in every copy)
                uint64_t um = __ballot(unv);
                while (um) {
                    const int t = __ffsll((unsigned long long)um) - 1;
                    um &= um - 1;
                    const uint32_t c = rl_u(cid, t);
                    if (lane > t && cid == c) unv = false;
                }
            }
            if (vnb) {
                if (unv && !vt_insert(cid)) vis_put(vis + cid, (uint8_t)epoch);
                lds_order();
            } else if (unv) {
                vis_put(vis + cid, (uint8_t)epoch);
            }
        }
        const uint64_t mask = __ballot(unv);
        const int U = __popcll(mask);
        const int pos = __popcll(mask & ((1ull << lane) - 1ull));
        if (unv) {
            cidbuf[pos] = cid;
            if (a.adjn && mode == MODE_EXPAND) {
                stbuf[pos] = cnd.x;
                dgbuf[pos] = cnd.y;
            }
        }
        lds_order();
        bool stopped = false;
        for (int ub = 0; ub < U && !stopped; ub += BATCH) {
            const int cnt = U - ub < BATCH ? U - ub : BATCH;
            uint32_t c, st, dg;
            const float dl = eval_batch(ub, cnt, c, st, dg, mode == MODE_EXPAND);
            int lim = cnt, stop = -1;
            if (mode == MODE_EXPAND && cd_inner + cnt > a.maxvisits) {
                stop = (int)(a.maxvisits - cd_inner);
                lim = stop + 1;
            }
            const bool full = res.len == res.cap;
            const float mx0 = full ? res.maximum() : __builtin_inff();
            const float th0 = a.delta * mx0;
            // candidates: children that may enter res or the beam (maximum(res) only
            // decreases while res is full, so a child failing now fails later too)
            bool cand = lane < lim;
            if (full) cand = cand && (dl < mx0 || (mode == MODE_EXPAND && dl <= th0));
            uint64_t cm = __ballot(cand);
            while (cm) {
                const int u = __ffsll((unsigned long long)cm) - 1;
                cm &= cm - 1;
                const float du = rl_f(dl, u);
                const uint32_t cu = rl_u(c, u);
                res.push(du, cu, 0, 0);
                if (mode != MODE_EXPAND) continue;
                if (u == stop) break;
                const uint32_t gu = rl_u(dg, u);
                const float thr = a.delta * res.maximum();
                if (gu > 1 && du <= thr) beam.push(du, cu, rl_u(st, u), gu);
            }
            if (mode == MODE_EXPAND) {
                cd_inner += lim;
                if (stop >= 0) {
                    // children after the cut were marked ahead of time: unmark them
                    if (lane > ub + stop && lane < U) {
                        const uint32_t cl = cidbuf[lane];
                        if (!vnb || !vt_erase(cl)) vis_put(vis + cl, 0);
                    }
                    stopped = true;
                }
            } else {
                cd_init += cnt;
            }
            lds_order();
        }
        lds_order();
        return stopped;
    }

    // `map`: the list names objects by their original ids, the graph is stored in
    // another order (a.ipos maps original -> stored)
    __device__ __forceinline__ void hint_list(const uint32_t* h, int nh, bool one_based, bool map)
    {
        for (int hb = 0; hb < nh; hb += 64) {
            const bool valid = hb + lane < nh;
            uint32_t c = valid ? h[hb + lane] : 1u;
            if (one_based) c -= 1u;
            if (map && valid && (int64_t)c < a.n) c = a.ipos[c];
            chunk(MODE_HINT, c, valid, true);
        }
    }

    // search(bs::BeamSearch, ...) for the current query with the given hints.
    __device__ __forceinline__ void search(const uint32_t* h, int nh, bool one_based, int bsize,
                                           bool map_hints = false)
    {
        const int64_t n = a.n;
        if (n == 0) return;
        if (n < 64) {
            chunk(MODE_BRUTE, (uint32_t)lane, lane < n, false);
            return;
        }
        hint_list(h, nh, one_based, map_hints);
        if (res.len == 0) {
            // ids 1:ceil(log2(1+n)) in the original numbering
            uint32_t c = (uint32_t)lane;
            if (a.ipos && lane < a.log2n) c = a.ipos[lane];
            chunk(MODE_HINT, c, lane < a.log2n, false);
        }
        const int cap = bsize < a.beamcap ? bsize : a.beamcap;
        beam.init(cap);
        {
            if (res.len == 0) return;
            const uint32_t nid = res.id_at(0);
            const float ndist = res.dist_at(0);
            const uint2 ni = a.node[nid];
            beam.push(ndist, nid, ni.x, ni.y);
        }
        if (PREC == 1 && a.expand > 1) {
            expand_wide();
            return;
        }
        while (beam.len > 0) {
            float pd;
            uint32_t pid, pst, pdeg;
            beam.pop_min(pd, pid, pst, pdeg);
            cb++;
            bool stop = false;
            for (uint32_t base = 0; base < pdeg && !stop; base += 64) {
                const uint32_t rem = pdeg - base;
                const bool valid = (uint32_t)lane < rem;
                const size_t e = (size_t)pst + base + lane;
                const uint32_t c = valid ? a.adj[e] : 0u;
                const uint2 cn = valid && a.adjn ? a.adjn[e] : make_uint2(0u, 0u);
                stop = chunk(MODE_EXPAND, c, valid, a.has_dups != 0, cn);
            }
            if (stop) break;
        }
    }

    // Wide step (BF16 form only, no reference counterpart): pop the min(expand, len)
    // nearest beam entries at once and expand their lists as one concatenated list, in
    // pop order, 64 ids per chunk.  Ids shared by several lists are visited once, by
    // their first copy.  expand = 1 is the reference step.
    __device__ __forceinline__ void expand_wide()
    {
        constexpr int PMAX = 4;
        const int P = a.expand < PMAX ? a.expand : PMAX;
        while (beam.len > 0) {
            uint32_t st[PMAX], off[PMAX + 1];
            off[0] = 0;
            int np = 0;
#pragma unroll
            for (int i = 0; i < PMAX; i++) {
                st[i] = 0;
                off[i + 1] = off[i];
                if (i < P && beam.len > 0) {
                    float pd;
                    uint32_t pid, pdeg;
                    beam.pop_min(pd, pid, st[i], pdeg);
                    off[i + 1] = off[i] + pdeg;
                    np++;
                }
            }
            cb += np;
            const uint32_t total = off[PMAX];
            bool stop = false;
            for (uint32_t base = 0; base < total && !stop; base += 64) {
                const uint32_t t = base + (uint32_t)lane;
                const bool valid = t < total;
                uint32_t s0 = st[0], o0 = 0;
#pragma unroll
                for (int i = 1; i < PMAX; i++)
                    if (t >= off[i]) {
                        s0 = st[i];
                        o0 = off[i];
                    }
                const size_t e = (size_t)s0 + (t - o0);
                const uint32_t c = valid ? a.adj[e] : 0u;
                const uint2 cn = valid && a.adjn ? a.adjn[e] : make_uint2(0u, 0u);
                stop = chunk(MODE_EXPAND, c, valid, true, cn);
            }
            if (stop) break;
        }
    }

    // 1-based original id of stored object c
    __device__ __forceinline__ uint32_t out_id(uint32_t c) const { return (a.perm ? a.perm[c] : c) + 1u; }

    // Exact rerank of the result list (rerank!, src/rerank.jl:22-41): canonical
    // distances from the FP32 vectors, stable sort, top kk written to row qi.
    __device__ __forceinline__ void rerank_out(int64_t qi, int kk)
    {
        const int dp = 16 * V;
        const int len = res.len;
        float dex = __builtin_inff();
        for (int r0 = 0; r0 < len; r0 += 16) {
            // rows past the list repeat its last entry: unconditional loads, in flight together
            RunWords<V, 0> xw[4];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int u = r0 + 4 * r + g;
                const uint32_t c = __shfl(res.id[0], u < len ? u : len - 1);
                load_words<V>(a.X + (size_t)c * dp + j * V, xw[r]);
            }
#pragma unroll
            for (int r = 0; r < 4; r++) {
                float xv[V];
                unpack_words<V>(xw[r], xv);
                const float d = finish_dist(a.metric, group16_sum(run_partial<V>(a.metric, xv, qv)));
                const float dt = __shfl(d, (lane & 3) * 16);
                if (lane >= r0 && lane < r0 + 16 && ((lane - r0) >> 2) == r && lane < len) dex = dt;
            }
        }
        int rank = 0;
        for (int t = 0; t < len; t++) {
            const float dt = rl_f(dex, t);
            rank += (dt < dex || (dt == dex && t < lane)) ? 1 : 0;
        }
        if (lane < len && rank < kk) {
            a.ids[(size_t)qi * a.kstride + rank] = out_id(res.id[0]);
            a.dists[(size_t)qi * a.kstride + rank] = dex;
        }
        const int filled = len < kk ? len : kk;
        if (lane >= filled && lane < a.kstride) {
            a.ids[(size_t)qi * a.kstride + lane] = 0u;
            a.dists[(size_t)qi * a.kstride + lane] = __builtin_inff();
        }
    }
};

template <int V, int NB, int PREC, int R, int NR = 1>
__device__ __forceinline__ void beam_body(const BeamArgs& a)
{
    __shared__ uint32_t s_cid[WAVES][64];
    __shared__ float s_dist[WAVES][64];
    __shared__ uint32_t s_st[WAVES][64];
    __shared__ uint32_t s_dg[WAVES][64];
    extern __shared__ uint4 s_vt[];  // WAVES * a.vnb buckets (dynamic)
    const int w = threadIdx.x >> 6;
    const int64_t slot = (int64_t)blockIdx.x * WAVES + w;
    Searcher<V, NB, PREC, R, NR> S(a);
    S.lane = lane_id();
    S.g = S.lane >> 4;
    S.j = S.lane & 15;
    S.cidbuf = s_cid[w];
    S.dbuf = s_dist[w];
    S.stbuf = s_st[w];
    S.dgbuf = s_dg[w];
    S.vis = a.vis + slot * a.vis_stride;
    S.vnb = (uint32_t)a.vnb;
    S.vhb = (uint32_t)a.vhb;
    S.vt = s_vt + rfl_u((uint32_t)w * (uint32_t)a.vnb);
    uint32_t epoch = rfl_u(a.epochs[slot]);
    const int lane = S.lane;
    const int dp = 16 * V;

    for (;;) {
        uint32_t qi = 0;
        if (lane == 0) qi = atomicAdd(a.counter, 1u);
        qi = rl_u(qi, 0);
        if ((int64_t)qi >= a.m) break;
        if (a.order) {
            qi = rfl_u(a.order[qi]);
            if ((int64_t)qi >= a.m) continue;  // not a permutation: skip, never index past m
        }
        epoch++;
        if (epoch > 255) {
            uint64_t* vp = reinterpret_cast<uint64_t*>(a.vis + slot * a.vis_stride);
            for (int64_t w8 = lane; w8 < a.vis_stride / 8; w8 += 64)
                __hip_atomic_store(vp + w8, (uint64_t)0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            vm_drain();
            epoch = 1;
        }
        S.epoch = epoch;
        S.vt_reset();
        if (a.qvisit) {  // the query's own vertex, marked before seeding (bichromatic_search!)
            const uint32_t x = rfl_u(a.qvisit[qi]);
            if (x != 0u && (int64_t)x <= a.n) {
                uint32_t c = x - 1u;
                if (a.ipos) c = rfl_u(a.ipos[c]);
                S.mark_uniform(c);
            }
        }
        if (a.allknn) {
            load_run<V>(a.X + (size_t)qi * dp + S.j * V, S.qv);
        } else {
#pragma unroll
            for (int t = 0; t < V; t++) {
                const int i = S.j * V + t;
                S.qv[t] = i < a.qdim ? a.Q[(size_t)qi * a.qdim + i] : 0.f;
            }
        }
        const int kk = a.kq ? a.kq[qi] : a.k;
        const int bs = a.bq ? a.bq[qi] : a.bsize;
        S.res.init(PREC == 1 ? (kk > a.krr ? kk : a.krr) : kk);
        S.cd_init = 0;
        S.cd_inner = 0;
        S.cb = 0;
        if (a.allknn) {
            const uint2 ni = a.node[qi];
            for (int64_t t = ni.x; t < (int64_t)ni.x + ni.y; t++) {
                const uint32_t h = a.adj[t];
                if (S.visited_uniform(h)) continue;
                const uint32_t h1 = h + 1u;
                S.search(&h1, 1, true, bs);
                S.cd_inner = 0;
            }
        } else if (a.qhints && a.qhcount[qi] >= 0) {
            S.search(a.qhints + (size_t)qi * a.qhstride, a.qhcount[qi], true, bs, a.ipos != nullptr);
        } else {
            S.search(a.hints, a.nhints, true, bs);
        }
        if constexpr (PREC == 1) {
            S.rerank_out(qi, kk);
        } else {
#pragma unroll
            for (int b = 0; b < NR; b++) {
                const int e = 64 * b + lane;
                if (e < a.kstride) {
                    const bool has = e < S.res.len;
                    a.ids[(size_t)qi * a.kstride + e] = has ? S.out_id(S.res.id[b]) : 0u;
                    a.dists[(size_t)qi * a.kstride + e] = has ? S.res.d[b] : __builtin_inff();
                }
            }
        }
        if (lane == 0) {
            if (a.costd) a.costd[qi] = S.cd_init + S.cd_inner;
            if (a.costb) a.costb[qi] = S.cb;
        }
    }
    if (lane == 0) a.epochs[slot] = (uint8_t)epoch;
}

// searchbatch! / allknn! launches (NR = 4: k up to 256)
template <int V, int NB, int R = 8, int NR = 1>
__global__ __launch_bounds__(256) void beam_kernel(BeamArgs a)
{
    beam_body<V, NB, 0, R, NR>(a);
}

// precision form: BF16 gathers, FP32 arithmetic, exact rerank of the result list.  Four
// waves per SIMD: below four the launch has too few loads in flight (3 waves: 12% slower
// at 1M x 128), and the few registers spilled to hold four cost less.
template <int V, int NB, int R = 8>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void beam_bf16_kernel(BeamArgs a)
{
    beam_body<V, NB, 1, R>(a);
}

// Gather rounds per batch for rows above 256 dims (V > 16), where each round's rows take
// V registers per lane: 2 rounds up to 512 dims, 1 beyond.
template <int V>
constexpr int wide_rounds()
{
    return V <= 32 ? 2 : 1;
}

// the same search issued by index! and optimize_index! (a separate symbol keeps
// their launches apart from query traffic in kernel traces)
template <int V, int NB>
__global__ __launch_bounds__(256) void beam_build_kernel(BeamArgs a)
{
    beam_body<V, NB, 0, (V > 16 ? wide_rounds<V>() : 8)>(a);
}

// Launch with the visited tables as dynamic LDS (WAVES * vnb * 16 bytes per block).  vnb < 0
// picks the largest table that keeps the kernel's occupancy: the LDS a CU has, shared by the
// blocks its registers let run, less their static LDS.  A table smaller than 2^(vhb-15)
// buckets (or than 256, when chosen automatically) is not used.
template <typename K>
hipError_t launch_lds(K kern, dim3 grid, dim3 block, const BeamArgs& args, hipStream_t s)
{
    BeamArgs a = args;
    if (a.vnb < 0) {
        int blocks = 0, dev = 0;
        size_t lds_cu = 0;
        hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, kern, 256, 0);
        if (e != hipSuccess) return e;
        if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
        hipDeviceProp_t prop;
        if ((e = hipGetDeviceProperties(&prop, dev)) != hipSuccess) return e;
        lds_cu = prop.maxSharedMemoryPerMultiProcessor;
        hipFuncAttributes fa;
        if ((e = hipFuncGetAttributes(&fa, (const void*)kern)) != hipSuccess) return e;
        int64_t v = blocks > 0 ? ((int64_t)(lds_cu / (size_t)blocks) - (int64_t)fa.sharedSizeBytes) /
                                     (int64_t)(WAVES * sizeof(uint4))
                               : 0;
        v = std::min<int64_t>(v, 2048);
        a.vnb = v >= 256 ? (int32_t)v : 0;
    }
    if (a.vnb > 0 && ((int64_t)a.vnb << 15) < ((int64_t)1 << a.vhb)) a.vnb = 0;
    const size_t lds = (size_t)WAVES * (size_t)a.vnb * sizeof(uint4);
    if (lds > 65536) {
        const hipError_t e =
            hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kern, grid, block, lds, s, a);
    return hipGetLastError();
}

template <int V>
hipError_t launch_v(const BeamArgs& a, int nslots, hipStream_t s)
{
    const dim3 grid((unsigned)(nslots / WAVES)), block(256);
    // 4 rounds for every beam cap (fewer registers: the 4-register FP32 kernel fits 3 waves
    // per SIMD with them, 2 with 8 rounds); 10 rounds (40 children per batch) for one beam
    // register in the BF16 form, where its rows still leave 4 waves per SIMD
    if constexpr (V > 16) {
        // rows above 256 dims: the FP32 form only, with fewer gather rounds
        constexpr int RW = wide_rounds<V>();
        if (a.prec == 1) return hipErrorInvalidValue;
        if (a.internal) {
            if (a.beamcap <= 64) return launch_lds(beam_build_kernel<V, 1>, grid, block, a, s);
            return launch_lds(beam_build_kernel<V, 4>, grid, block, a, s);
        }
        if (a.kstride > 64) {
            if (a.beamcap <= 64) return launch_lds(beam_kernel<V, 1, RW, 4>, grid, block, a, s);
            return launch_lds(beam_kernel<V, 4, RW, 4>, grid, block, a, s);
        }
        if (a.beamcap <= 64) return launch_lds(beam_kernel<V, 1, RW>, grid, block, a, s);
        return launch_lds(beam_kernel<V, 4, RW>, grid, block, a, s);
    } else {
        const bool r4 = a.rounds == 4;
        const bool r10 = a.rounds == 10 && a.beamcap <= 64;
        if (a.internal) {
            if (a.beamcap <= 64) return launch_lds(beam_build_kernel<V, 1>, grid, block, a, s);
            return launch_lds(beam_build_kernel<V, 4>, grid, block, a, s);
        }
        if (a.prec == 1) {
            if (r10) return launch_lds(beam_bf16_kernel<V, 1, 10>, grid, block, a, s);
            if (a.beamcap <= 64)
                return r4 ? launch_lds(beam_bf16_kernel<V, 1, 4>, grid, block, a, s)
                          : launch_lds(beam_bf16_kernel<V, 1>, grid, block, a, s);
            if (a.beamcap <= 128)
                return r4 ? launch_lds(beam_bf16_kernel<V, 2, 4>, grid, block, a, s)
                          : launch_lds(beam_bf16_kernel<V, 2>, grid, block, a, s);
            return r4 ? launch_lds(beam_bf16_kernel<V, 4, 4>, grid, block, a, s)
                      : launch_lds(beam_bf16_kernel<V, 4>, grid, block, a, s);
        }
        if (a.kstride > 64) {
            if (a.beamcap <= 64) return launch_lds(beam_kernel<V, 1, 8, 4>, grid, block, a, s);
            return launch_lds(beam_kernel<V, 4, 8, 4>, grid, block, a, s);
        }
        if (a.beamcap <= 64)
            return r4 ? launch_lds(beam_kernel<V, 1, 4>, grid, block, a, s)
                      : launch_lds(beam_kernel<V, 1>, grid, block, a, s);
        return r4 ? launch_lds(beam_kernel<V, 4, 4>, grid, block, a, s)
                  : launch_lds(beam_kernel<V, 4>, grid, block, a, s);
    }
}

}  // namespace ssg
