// ssg_device.h -- device-side building blocks shared by the gfx950 kernels.
//
// * Canonical distance: a 16-lane group evaluates one (object, query) pair.
//   Lane j of the group owns the contiguous run [j*V, (j+1)*V) of the padded
//   dimension Dp = 16*V, accumulates it in FP64 with fma in increasing index,
//   and the group combines the 16 partial sums by the xor-butterfly 8,4,2,1.
//   The FP64 result is rounded once to Float32.  The CPU oracle evaluates the
//   identical expression tree, so both sides produce the same bits.
// * RegQueue: KnnSorted (src/pqueue/knnsorted.jl:31-233) held in registers,
//   slot e = 64*b + lane.  push keeps equal distances in insertion order and a
//   full queue rejects d >= maximum (knnsorted.jl:140-168); pop_min removes the
//   first slot (knnsorted.jl:191-196).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ssg {

enum { M_SQL2 = 0, M_L2 = 1, M_NORMCOS = 2 };

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63); }

__device__ __forceinline__ float rl_f(float v, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ uint32_t rl_u(uint32_t v, int l)
{
    return (uint32_t)__builtin_amdgcn_readlane((int)v, l);
}
__device__ __forceinline__ uint32_t rfl_u(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}

// Orders a wave's LDS and global accesses before the ones that follow.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

template <int V>
__device__ __forceinline__ void load_run(const float* __restrict__ p, float (&x)[V])
{
    if constexpr (V % 4 == 0) {
#pragma unroll
        for (int t = 0; t < V / 4; t++) {
            const float4 w = reinterpret_cast<const float4*>(p)[t];
            x[4 * t + 0] = w.x;
            x[4 * t + 1] = w.y;
            x[4 * t + 2] = w.z;
            x[4 * t + 3] = w.w;
        }
    } else if constexpr (V % 2 == 0) {
#pragma unroll
        for (int t = 0; t < V / 2; t++) {
            const float2 w = reinterpret_cast<const float2*>(p)[t];
            x[2 * t + 0] = w.x;
            x[2 * t + 1] = w.y;
        }
    } else {
#pragma unroll
        for (int t = 0; t < V; t++) x[t] = p[t];
    }
}

// BF16 run (upper halves of FP32 words) widened to FP32.
template <int V>
__device__ __forceinline__ void load_run_bf16(const uint16_t* __restrict__ p, float (&x)[V])
{
    if constexpr (V % 8 == 0) {
#pragma unroll
        for (int t = 0; t < V / 8; t++) {
            const uint4 w = reinterpret_cast<const uint4*>(p)[t];
            const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
            for (int u = 0; u < 4; u++) {
                x[8 * t + 2 * u] = __uint_as_float(ws[u] << 16);
                x[8 * t + 2 * u + 1] = __uint_as_float(ws[u] & 0xffff0000u);
            }
        }
    } else if constexpr (V % 2 == 0) {
#pragma unroll
        for (int t = 0; t < V / 2; t++) {
            const uint32_t w = reinterpret_cast<const uint32_t*>(p)[t];
            x[2 * t] = __uint_as_float(w << 16);
            x[2 * t + 1] = __uint_as_float(w & 0xffff0000u);
        }
    } else {
#pragma unroll
        for (int t = 0; t < V; t++) x[t] = __uint_as_float((uint32_t)p[t] << 16);
    }
}

// Raw words of one lane's run, loaded without being consumed so that many rows can
// be in flight at once: FP32 runs as V words; BF16 runs as V/2 packed words (even V)
// or V zero-extended halves (odd V).
template <int V, int PREC>
struct RunWords {
    static constexpr int N = PREC == 1 ? (V % 2 == 0 ? V / 2 : V) : V;
    uint32_t w[N];
};

template <int V>
__device__ __forceinline__ void load_words(const float* __restrict__ p, RunWords<V, 0>& r)
{
    float x[V];
    load_run<V>(p, x);
#pragma unroll
    for (int t = 0; t < V; t++) r.w[t] = __float_as_uint(x[t]);
}

template <int V>
__device__ __forceinline__ void load_words(const uint16_t* __restrict__ p, RunWords<V, 1>& r)
{
    if constexpr (V % 8 == 0) {
#pragma unroll
        for (int t = 0; t < V / 8; t++) {
            const uint4 w = reinterpret_cast<const uint4*>(p)[t];
            r.w[4 * t + 0] = w.x;
            r.w[4 * t + 1] = w.y;
            r.w[4 * t + 2] = w.z;
            r.w[4 * t + 3] = w.w;
        }
    } else if constexpr (V % 2 == 0) {
#pragma unroll
        for (int t = 0; t < V / 2; t++) r.w[t] = reinterpret_cast<const uint32_t*>(p)[t];
    } else {
#pragma unroll
        for (int t = 0; t < V; t++) r.w[t] = p[t];
    }
}

template <int V>
__device__ __forceinline__ void unpack_words(const RunWords<V, 0>& r, float (&x)[V])
{
#pragma unroll
    for (int t = 0; t < V; t++) x[t] = __uint_as_float(r.w[t]);
}

template <int V>
__device__ __forceinline__ void unpack_words(const RunWords<V, 1>& r, float (&x)[V])
{
    if constexpr (V % 2 == 0) {
#pragma unroll
        for (int t = 0; t < V / 2; t++) {
            x[2 * t] = __uint_as_float(r.w[t] << 16);
            x[2 * t + 1] = __uint_as_float(r.w[t] & 0xffff0000u);
        }
    } else {
#pragma unroll
        for (int t = 0; t < V; t++) x[t] = __uint_as_float(r.w[t] << 16);
    }
}

// FP32 partial sum (reduced-precision form; its result is never final).
template <int V>
__device__ __forceinline__ float run_partial_f32(int metric, const float (&x)[V], const float (&q)[V])
{
    float s = 0.f;
    if (metric == M_NORMCOS) {
#pragma unroll
        for (int t = 0; t < V; t++) s = __builtin_fmaf(x[t], q[t], s);
    } else {
#pragma unroll
        for (int t = 0; t < V; t++) {
            const float m = x[t] - q[t];
            s = __builtin_fmaf(m, m, s);
        }
    }
    return s;
}

__device__ __forceinline__ float group16_sum_f32(float s)
{
    s = s + __shfl_xor(s, 8);
    s = s + __shfl_xor(s, 4);
    s = s + __shfl_xor(s, 2);
    s = s + __shfl_xor(s, 1);
    return s;
}

__device__ __forceinline__ float finish_f32(int metric, float s)
{
    if (metric == M_L2) return __builtin_sqrtf(s);
    if (metric == M_NORMCOS) return 1.0f - s;
    return s;
}

// Partial sum of one lane's run.
template <int V>
__device__ __forceinline__ double run_partial(int metric, const float (&x)[V], const float (&q)[V])
{
    double s = 0.0;
    if (metric == M_NORMCOS) {
#pragma unroll
        for (int t = 0; t < V; t++) s = __builtin_fma((double)x[t], (double)q[t], s);
    } else {
#pragma unroll
        for (int t = 0; t < V; t++) {
            const double m = (double)x[t] - (double)q[t];
            s = __builtin_fma(m, m, s);
        }
    }
    return s;
}

// ---------------------------------------------------------------------------
// DPP data movement (VALU, no LDS round trip).  Controls of the GFX9 family:
// row_shr:n = 0x110 + n (within rows of 16 lanes), wave_shr:1 = 0x138,
// wave_shl:1 = 0x130.  Lanes without a source read 0 (bound_ctrl).
// ---------------------------------------------------------------------------
enum { DPP_ROW_SHR1 = 0x111, DPP_ROW_SHR2 = 0x112, DPP_ROW_SHR4 = 0x114, DPP_ROW_SHR8 = 0x118,
       DPP_WAVE_SHL1 = 0x130, DPP_WAVE_SHR1 = 0x138 };

template <int CTRL>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, true);
}
template <int CTRL>
__device__ __forceinline__ float dpp_f32(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, true));
}
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v)
{
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = dpp_u32<CTRL>((uint32_t)u), hi = dpp_u32<CTRL>((uint32_t)(u >> 32));
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// The same 16-lane tree as group16_sum -- pairs (j, j+8), then (j, j+4), (j, j+2),
// (j, j+1) of the partial results -- evaluated with row shifts; only lane 15 of each
// 16-lane row holds the total (bit-identical to group16_sum's).
__device__ __forceinline__ double group16_tree(double s)
{
    s = s + dpp_f64<DPP_ROW_SHR8>(s);
    s = s + dpp_f64<DPP_ROW_SHR4>(s);
    s = s + dpp_f64<DPP_ROW_SHR2>(s);
    s = s + dpp_f64<DPP_ROW_SHR1>(s);
    return s;
}
__device__ __forceinline__ float group16_tree_f32(float s)
{
    s = s + dpp_f32<DPP_ROW_SHR8>(s);
    s = s + dpp_f32<DPP_ROW_SHR4>(s);
    s = s + dpp_f32<DPP_ROW_SHR2>(s);
    s = s + dpp_f32<DPP_ROW_SHR1>(s);
    return s;
}

// xor-butterfly over the 16 lanes of a group; every lane ends with the total.
__device__ __forceinline__ double group16_sum(double s)
{
    s = s + __shfl_xor(s, 8);
    s = s + __shfl_xor(s, 4);
    s = s + __shfl_xor(s, 2);
    s = s + __shfl_xor(s, 1);
    return s;
}

__device__ __forceinline__ float finish_dist(int metric, double s)
{
    if (metric == M_L2) return (float)__builtin_sqrt(s);
    if (metric == M_NORMCOS) return (float)(1.0 - s);
    return (float)s;
}

// ---------------------------------------------------------------------------
// KnnSorted in registers.  NB registers per field, capacity cap <= 64*NB.
// PAY adds two uint32 payload fields carried with each entry.
// ---------------------------------------------------------------------------
template <int NB, bool PAY>
struct RegQueue {
    float d[NB];
    uint32_t id[NB];
    uint32_t pa[PAY ? NB : 1];
    uint32_t pb[PAY ? NB : 1];
    int len;
    int cap;

    __device__ __forceinline__ void init(int capacity)
    {
#pragma unroll
        for (int b = 0; b < NB; b++) {
            d[b] = __builtin_inff();
            id[b] = 0xffffffffu;
            if constexpr (PAY) {
                pa[b] = 0;
                pb[b] = 0;
            }
        }
        len = 0;
        cap = capacity;
    }

    // distance of slot e (e wave-uniform)
    __device__ __forceinline__ float dist_at(int e) const
    {
        float r = 0.f;
#pragma unroll
        for (int b = 0; b < NB; b++)
            if ((e >> 6) == b) r = rl_f(d[b], e & 63);
        return r;
    }
    __device__ __forceinline__ uint32_t id_at(int e) const
    {
        uint32_t r = 0;
#pragma unroll
        for (int b = 0; b < NB; b++)
            if ((e >> 6) == b) r = rl_u(id[b], e & 63);
        return r;
    }
    __device__ __forceinline__ float maximum() const { return dist_at(len - 1); }

    // push_item! with wave-uniform arguments; returns whether the item entered.
    __device__ __forceinline__ bool push(float x, uint32_t xid, uint32_t xa, uint32_t xb)
    {
        const int lane = lane_id();
        if (len == cap) {
            if (x >= maximum()) return false;
            len--;
        }
        int pos = 0;
#pragma unroll
        for (int b = 0; b < NB; b++) {
            const int e = 64 * b + lane;
            if (64 * b < len) pos += __popcll(__ballot(e < len && d[b] <= x));
        }
        // registers past the last occupied slot hold nothing to move
#pragma unroll
        for (int b = NB - 1; b >= 0; b--) {
            if (64 * b > len) continue;
            const int e = 64 * b + lane;
            float ud = dpp_f32<DPP_WAVE_SHR1>(d[b]);
            uint32_t ui = dpp_u32<DPP_WAVE_SHR1>(id[b]);
            uint32_t ua = 0, ub = 0;
            if constexpr (PAY) {
                ua = dpp_u32<DPP_WAVE_SHR1>(pa[b]);
                ub = dpp_u32<DPP_WAVE_SHR1>(pb[b]);
            }
            if (b > 0) {
                const float cd = rl_f(d[b > 0 ? b - 1 : 0], 63);
                const uint32_t ci = rl_u(id[b > 0 ? b - 1 : 0], 63);
                uint32_t ca = 0, cb = 0;
                if constexpr (PAY) {
                    ca = rl_u(pa[b > 0 ? b - 1 : 0], 63);
                    cb = rl_u(pb[b > 0 ? b - 1 : 0], 63);
                }
                if (lane == 0) {
                    ud = cd;
                    ui = ci;
                    ua = ca;
                    ub = cb;
                }
            }
            if (e > pos && e <= len) {
                d[b] = ud;
                id[b] = ui;
                if constexpr (PAY) {
                    pa[b] = ua;
                    pb[b] = ub;
                }
            } else if (e == pos) {
                d[b] = x;
                id[b] = xid;
                if constexpr (PAY) {
                    pa[b] = xa;
                    pb[b] = xb;
                }
            }
        }
        len++;
        return true;
    }

    // pop_min!: returns slot 0 and shifts the rest down.
    __device__ __forceinline__ void pop_min(float& x, uint32_t& xid, uint32_t& xa, uint32_t& xb)
    {
        const int lane = lane_id();
        x = rl_f(d[0], 0);
        xid = rl_u(id[0], 0);
        if constexpr (PAY) {
            xa = rl_u(pa[0], 0);
            xb = rl_u(pb[0], 0);
        } else {
            xa = 0;
            xb = 0;
        }
#pragma unroll
        for (int b = 0; b < NB; b++) {
            if (64 * b >= len) continue;
            float nd = dpp_f32<DPP_WAVE_SHL1>(d[b]);
            uint32_t ni = dpp_u32<DPP_WAVE_SHL1>(id[b]);
            uint32_t na = 0, nb = 0;
            if constexpr (PAY) {
                na = dpp_u32<DPP_WAVE_SHL1>(pa[b]);
                nb = dpp_u32<DPP_WAVE_SHL1>(pb[b]);
            }
            if (b + 1 < NB) {
                const int bn = b + 1 < NB ? b + 1 : b;
                const float cd = rl_f(d[bn], 0);
                const uint32_t ci = rl_u(id[bn], 0);
                uint32_t ca = 0, cb = 0;
                if constexpr (PAY) {
                    ca = rl_u(pa[bn], 0);
                    cb = rl_u(pb[bn], 0);
                }
                if (lane == 63) {
                    nd = cd;
                    ni = ci;
                    na = ca;
                    nb = cb;
                }
            }
            d[b] = nd;
            id[b] = ni;
            if constexpr (PAY) {
                pa[b] = na;
                pb[b] = nb;
            }
        }
        len--;
    }
};

}  // namespace ssg
