// ssg_select.h -- wave-level selection helpers shared by the candidate passes of the
// exhaustive searches (ssg_exhaustive.hip, ssg_prec.hip): order-preserving float keys, an
// 8-bit radix select with LDS histograms, exact selection by (score, id) and the
// compaction of per-(query, slice) candidate lists split into 4 sub-lists.
#pragma once
#include "ssg_device.h"
#include "ssg_kernels.h"

namespace ssg {

typedef uint32_t u32x2 __attribute__((__vector_size__(8)));

// Order-preserving map of floats to unsigned keys (larger float -> larger key).
__device__ __forceinline__ uint32_t score_key(float f)
{
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float key_score(uint32_t k)
{
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

// Orders a wave's LDS accesses for the compiler (a wave's LDS operations complete in
// issue order, so no wait is needed between them).
__device__ __forceinline__ void lds_order()
{
    asm volatile("" ::: "memory");
    __builtin_amdgcn_wave_barrier();
}

// Radix select over the active entries of the wave: the top 8 * npass bits of the
// want-th largest key (want >= 1, at most the number of active entries), one 8-bit digit
// per pass with a 256-bin histogram in LDS (hist: 256 words owned by this wave).
// *above = active entries whose top 8 * npass bits are larger.
// kand / diff (wave-uniform): the AND of the active keys and the bits in which they
// differ; a byte in which no active key differs needs no histogram (all of its increments
// would go to one bin).  diff = ~0 skips nothing.
template <int J>
__device__ uint32_t kth_largest(const uint32_t (&key)[J], const bool (&act)[J], int want, int npass,
                                uint32_t* hist, int lane, int* above, uint32_t kand, uint32_t diff)
{
    uint32_t prefix = 0;
    int rem = want, abv = 0;
    for (int pass = 0; pass < npass; pass++) {
        const int sh = 24 - 8 * pass;
        if (((diff >> sh) & 255u) == 0u) {
            prefix = (prefix << 8) | ((kand >> sh) & 255u);
            continue;
        }
        reinterpret_cast<uint4*>(hist)[lane] = make_uint4(0u, 0u, 0u, 0u);
        lds_order();
#pragma unroll
        for (int j = 0; j < J; j++) {
            const bool in = act[j] && (pass == 0 || (key[j] >> (sh + 8)) == prefix);
            if (in) atomicAdd(&hist[(key[j] >> sh) & 255u], 1u);
        }
        lds_order();
        const uint4 h4 = reinterpret_cast<const uint4*>(hist)[lane];
        const uint32_t tot = h4.x + h4.y + h4.z + h4.w;
        uint32_t suf = tot;  // entries in bins 4*lane .. 255
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t v = __shfl_down(suf, o);
            if (lane + o < 64) suf += v;
        }
        const uint32_t s3 = suf - tot + h4.w, s2 = s3 + h4.z, s1 = s2 + h4.y;
        const uint64_t bl = __ballot(suf >= (uint32_t)rem);
        const int L = 63 - __builtin_clzll(bl);
        const int t3 = (int)rl_u(s3, L), t2 = (int)rl_u(s2, L), t1 = (int)rl_u(s1, L);
        const int hw = (int)rl_u(h4.w, L);
        int d, sa;
        if (t3 >= rem) {
            d = 3;
            sa = t3 - hw;
        } else if (t2 >= rem) {
            d = 2;
            sa = t3;
        } else if (t1 >= rem) {
            d = 1;
            sa = t2;
        } else {
            d = 0;
            sa = t1;
        }
        rem -= sa;
        abv += sa;
        prefix = (prefix << 8) | (uint32_t)(4 * L + d);
        lds_order();
    }
    *above = abv;
    return prefix;
}

// Wave-uniform AND of the active keys and the bits in which they differ.
template <int J>
__device__ __forceinline__ void key_bits(const uint32_t (&key)[J], const bool (&act)[J], uint32_t* kand,
                                         uint32_t* diff)
{
    uint32_t a = 0xffffffffu, o = 0u;
#pragma unroll
    for (int j = 0; j < J; j++) {
        a &= act[j] ? key[j] : 0xffffffffu;
        o |= act[j] ? key[j] : 0u;
    }
    for (int off = 32; off >= 1; off >>= 1) {
        a &= (uint32_t)__shfl_xor((int)a, off);
        o |= (uint32_t)__shfl_xor((int)o, off);
    }
    *kand = rfl_u(a);
    *diff = rfl_u(a ^ o);
}

// Marks the kc best active entries by (key descending, id ascending); returns the kc-th
// best key.  Requires more than kc active entries.
template <int J>
__device__ uint32_t select_best(const uint32_t (&key)[J], const uint32_t (&id)[J], const bool (&act)[J],
                                int kc, uint32_t* hist, int lane, bool (&keep)[J], uint32_t kand, uint32_t diff)
{
    int above = 0;
    const uint32_t T = kth_largest<J>(key, act, kc, 4, hist, lane, &above, kand, diff);
    const int need = kc - above;  // entries with key == T to keep, >= 1
    int eq = 0;
    bool tie[J];
    uint32_t nid[J];
#pragma unroll
    for (int j = 0; j < J; j++) {
        tie[j] = act[j] && key[j] == T;
        nid[j] = ~id[j];
        eq += __popcll(__ballot(tie[j]));
    }
    uint32_t TI = 0;
    if (eq > need) {
        int a2 = 0;
        TI = kth_largest<J>(nid, tie, need, 4, hist, lane, &a2, 0u, 0xffffffffu);  // ids are distinct
    }
#pragma unroll
    for (int j = 0; j < J; j++) keep[j] = act[j] && (key[j] > T || (tie[j] && nid[j] >= TI));
    return T;
}

// A (query, slice) list of cap entries is split into 4 sub-lists of sc = cap / 4
// entries; sub-list g is appended to by the lanes 16g .. 16g + 15 of the owning wave
// (the lanes that hold rows 4g .. 4g + 3 of each 16-row block), each with its own count
// in a register, so appends need no atomics.
//
// Compaction of one list with sub-list counts c0..c3.  final == false (during the
// scan): keeps every entry whose key reaches the 16-bit prefix of the kc-th best key
// (at least kc entries, at most kc + 64, else the exact selection below), spread
// round-robin over the sub-lists; returns the new threshold (a row passes if its score
// is larger) and *kept.  final == true: keeps exactly the kc best by (score descending,
// id ascending), or every entry if there are at most kc, written contiguously at the
// start of the list.
// J = cap / 64 entries per lane (8 for lists of up to 512 entries, 16 for larger ones).
template <int J>
__device__ __attribute__((noinline)) float compact_list_j(uint2* L, int sc, int c0, int c1, int c2, int c3, int kc,
                                                          bool final, uint32_t* hist, int lane, int* kept)
{
    uint32_t key[J], id[J];
    bool act[J], keep[J];
    int total = 0;
#pragma unroll
    for (int j = 0; j < J; j++) {
        const int e = lane + 64 * j;
        const int g = e / sc, pos = e - g * sc;
        const int c = g == 0 ? c0 : g == 1 ? c1 : g == 2 ? c2 : g == 3 ? c3 : 0;
        act[j] = pos < c;
        const uint2 v = act[j] ? L[e] : make_uint2(0u, 0u);
        key[j] = score_key(__uint_as_float(v.x));
        id[j] = v.y;
        total += __popcll(__ballot(act[j]));
    }
    // fewer than kc entries: no threshold yet (and none was set before, since a
    // compaction leaves at least kc); the entries are only spread over the sub-lists
    float tau = -__builtin_inff();
    bool exact = final;
    if (!final && total < kc) {
#pragma unroll
        for (int j = 0; j < J; j++) keep[j] = act[j];
    } else if (!final) {
        int above = 0;
        uint32_t kand, diff;
        key_bits<J>(key, act, &kand, &diff);
        const uint32_t K = kth_largest<J>(key, act, kc, 2, hist, lane, &above, kand, diff) << 16;
        int nk = 0;
#pragma unroll
        for (int j = 0; j < J; j++) {
            keep[j] = act[j] && key[j] >= K;
            nk += __popcll(__ballot(keep[j]));
        }
        if (nk <= kc + 64) tau = key_score(K - 1u);
        else exact = true;
    }
    if (exact) {
        if (total > kc) {
            uint32_t kand, diff;
            key_bits<J>(key, act, &kand, &diff);
            tau = key_score(select_best<J>(key, id, act, kc, hist, lane, keep, kand, diff));
        } else {
#pragma unroll
            for (int j = 0; j < J; j++) keep[j] = act[j];
        }
    }
    // at most kc + 64 entries stay (ids are distinct, so the selections above keep at most
    // that many; the bound also holds the list in its space for any input)
    const int lim = kc + 64;
    int base = 0;
#pragma unroll
    for (int j = 0; j < J; j++) {
        const uint64_t b = __ballot(keep[j]);
        if (keep[j]) {
            const int i = base + __popcll(b & ((1ull << lane) - 1ull));
            const int p = final ? i : (i & 3) * sc + (i >> 2);
            if (i < lim) L[p] = make_uint2(__float_as_uint(key_score(key[j])), id[j]);
        }
        base += __popcll(b);
    }
    *kept = base < lim ? base : lim;
    wave_sync();  // the rewritten entries are visible to later appends and compactions
    return tau;
}


}  // namespace ssg
