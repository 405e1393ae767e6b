// em_device.hpp -- device-side building blocks of the gfx950 field-line tracer.
//
// Layout and work split (CDNA4, wave64):
//   * one line is owned by S adjacent lanes (S = 1..32); the S lanes split the
//     charge sum (lane k takes charges k, k+S, k+2S, ... of every tile) and
//     combine it with an xor butterfly, so every lane of the group ends the
//     step holding the same bits;
//   * charges are staged through LDS in tiles of float4 {x, y, z, q}; inside a
//     tile all lanes of a wave read one of S consecutive records -- a
//     broadcast from LDS with no bank conflict for S <= 8;
//   * the Coulomb term is FP32 with v_rsq_f32 (12 VALU + 1 transcendental per
//     charge-point interaction); accumulation is blocked: plain FP32 partial
//     sums over <= kBlk charges, folded into a (sum, compensation) pair with
//     Knuth's TwoSum;
//   * the step rule and the coalesced row store run in the same kernel.
//
// This file is compiled with -ffp-contract=off: every fused multiply-add in
// the hot loop is written out as __builtin_fmaf, and the step arithmetic
// below is evaluated exactly in source order, which is the order of the
// reference's vec3 helpers (include/Vector.h:158-197).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace em {

constexpr int kUnroll = 8;    // charges per lane per unrolled group
constexpr int kGroup = 2 * kUnroll;  // padding granule per lane (two pipelined groups)
constexpr int kBlk = 128;     // charges per lane per compensated block (multiple of kGroup)
constexpr float kFar = 1.0e15f;  // padding charge position (q = 0)

enum : int { kRulePlain = 0, kRuleRef = 1, kRuleAngle = 2 };

// Internal per-configuration descriptor (the public em_batch_cfg plus the
// offsets of the adjoint's buffers).
struct Cfg {
    uint64_t charge_off;  // first charge record
    uint64_t line_off;    // element offset of (row 0, line 0) in x/y/z
    uint64_t pitch;       // row pitch (elements)
    uint64_t ckpt_off;    // first float of this configuration's checkpoints
    uint64_t line_index;  // global index of this configuration's first line (step counts)
    uint32_t p, n, len, grad_line_off;
};

struct Item {
    uint32_t cfg;
    uint32_t line0;
};

struct StepP {
    int rule;
    float res, h0, hmin, hmax, turn;
    float scale;  // factor applied to the Coulomb sum (k_e for plain/ref)
    int stop;     // angle rule: end a line when its direction reverses
};

__host__ __device__ constexpr uint32_t round_up(uint32_t a, uint32_t b) { return (a + b - 1) / b * b; }

// r^2 starts from kSoft2 instead of 0 (same instruction count: three FMAs).
// A point that lands exactly on a charge -- float rounding snaps a line that
// oscillates about a sink onto the charge's coordinates -- then gets a zero
// contribution from it instead of 0 * inf = NaN; inv^3 <= 1e30 keeps q * inv^3
// finite for |q| < 3e8.  Relative effect at r >= 1e-5: below 1e-10.
constexpr float kSoft2 = 1.0e-20f;

// One charge-point interaction: E += q (p - c) / |p - c|^3.
__device__ __forceinline__ void coulomb(const float4 c, float px, float py, float pz,
                                        float& ex, float& ey, float& ez)
{
    const float dx = px - c.x;
    const float dy = py - c.y;
    const float dz = pz - c.z;
    const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, __builtin_fmaf(dx, dx, kSoft2)));
    const float inv = __builtin_amdgcn_rsqf(r2);
    const float s = (c.w * inv) * (inv * inv);
    ex = __builtin_fmaf(s, dx, ex);
    ey = __builtin_fmaf(s, dy, ey);
    ez = __builtin_fmaf(s, dz, ez);
}

// Knuth TwoSum: (sum, comp) += part, error-free.
__device__ __forceinline__ void fold(float& sum, float& comp, float part)
{
    const float t = sum + part;
    const float bp = t - sum;
    const float err = (sum - (t - bp)) + (part - bp);
    sum = t;
    comp += err;
}

struct Acc {
    float sx, sy, sz, cx, cy, cz;
    __device__ void zero() { sx = sy = sz = cx = cy = cz = 0.f; }
};

// Software-pipelined walk over one LDS tile of records (R float4 each).
// Lane k of a line group visits records k, k+S, k+2S, ...; `count` records
// (a multiple of S*kGroup; the tile is padded with records that contribute
// exactly zero).  While one group of kUnroll records is consumed, the LDS reads
// of the next group are in flight (two register sets, no copies).
// Body::term(const float4* rec) consumes one record into plain FP32 partial
// sums; Body::fold() moves the partials into compensated totals every kBlk
// records per lane and at the end of the tile.  U records per group (kUnroll
// by default; a body that needs many registers of its own may take fewer):
// every record is still consumed in order, so the sums do not depend on U.
template <int S, int R, int U = kUnroll, class Body>
__device__ __forceinline__ void pipelined_tile(const float4* __restrict__ tile, uint32_t count,
                                               int k, Body& body)
{
    static_assert(kGroup % (2 * U) == 0 && kBlk % U == 0, "U must divide the padding granule");
    const uint32_t ngroups = count / (S * U);  // even: count is a multiple of S*kGroup
    const float4* t = tile + k * R;
    constexpr uint32_t kStride = S * U * R;
    constexpr uint32_t kFoldEvery = kBlk / U;  // groups per compensated block (even)
    float4 a[U][R], b[U][R];
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
        for (int r = 0; r < R; ++r) a[u][r] = t[u * S * R + r];
    for (uint32_t g = 0; g < ngroups; g += 2) {
        // sched_barrier(0) keeps each group's LDS reads issued ahead of the
        // other group's arithmetic (the scheduler would otherwise sink them
        // next to their uses and expose the LDS latency every iteration).
        const float4* tb = t + (g + 1) * kStride;
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int r = 0; r < R; ++r) b[u][r] = tb[u * S * R + r];
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < U; ++u) body.term(a[u]);
        __builtin_amdgcn_sched_barrier(0);
        const uint32_t gn = min(g + 2, ngroups - 2);  // last pair re-reads itself
        const float4* ta = t + gn * kStride;
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int r = 0; r < R; ++r) a[u][r] = ta[u * S * R + r];
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < U; ++u) body.term(b[u]);
        __builtin_amdgcn_sched_barrier(0);
        if ((g + 2) % kFoldEvery == 0 || g + 2 >= ngroups) body.fold();
    }
}

// Forward body: E = sum q (p - c)/|p - c|^3 at one point.
struct FieldBody {
    float px, py, pz;
    float ex, ey, ez;  // partials of the current block
    Acc acc;           // compensated totals
    __device__ void init(float x, float y, float z)
    {
        px = x; py = y; pz = z;
        ex = ey = ez = 0.f;
        acc.zero();
    }
    __device__ __forceinline__ void term(const float4* c) { coulomb(c[0], px, py, pz, ex, ey, ez); }
    __device__ __forceinline__ void fold()
    {
        em::fold(acc.sx, acc.cx, ex);
        em::fold(acc.sy, acc.cy, ey);
        em::fold(acc.sz, acc.cz, ez);
        ex = ey = ez = 0.f;
    }
};

template <int S, int U = kUnroll>
__device__ __forceinline__ void sum_tile(const float4* __restrict__ tile, uint32_t count, int k,
                                         FieldBody& body)
{
    pipelined_tile<S, 1, U>(tile, count, k, body);
}

// Cooperative global -> LDS copy of `cnt` charge records, padded to `cntp`.
__device__ __forceinline__ void stage(float4* __restrict__ tile, const float4* __restrict__ src,
                                      uint32_t cnt, uint32_t cntp)
{
    for (uint32_t i = threadIdx.x; i < cntp; i += blockDim.x)
        tile[i] = i < cnt ? src[i] : make_float4(kFar, kFar, kFar, 0.f);
}

// v + (v of the partner lane under a DPP pattern); partners pair up, and IEEE
// addition commutes, so both lanes of a pair get the same bits.
template <int CTRL>
__device__ __forceinline__ float dpp_pair_add(float v)
{
    return v + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

// Sum over the S lanes of a line group (S consecutive lanes, S a power of
// two <= 64), in registers: lane pairs by DPP within a row of 16 (quad xor 1,
// quad xor 2, half-row mirror, row mirror), then row pairs and wave halves by
// the gfx950 lane swaps (v_permlane16_swap / v_permlane32_swap give every
// lane the even and the odd row of its pair, added in that order).  Every
// lane of a group ends with the same bits.
template <int S>
__device__ __forceinline__ float group_sum(float v)
{
    static_assert(S >= 1 && S <= 64 && (S & (S - 1)) == 0, "group of 1..64 lanes");
    if constexpr (S >= 2) v = dpp_pair_add<0xB1>(v);   // quad_perm [1, 0, 3, 2]
    if constexpr (S >= 4) v = dpp_pair_add<0x4E>(v);   // quad_perm [2, 3, 0, 1]
    if constexpr (S >= 8) v = dpp_pair_add<0x141>(v);  // row_half_mirror
    if constexpr (S >= 16) v = dpp_pair_add<0x140>(v); // row_mirror
    if constexpr (S >= 32) {
        const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v), false, false);
        v = __uint_as_float(r[0]) + __uint_as_float(r[1]);
    }
    if constexpr (S >= 64) {
        const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v), false, false);
        v = __uint_as_float(r[0]) + __uint_as_float(r[1]);
    }
    return v;
}

// Full Coulomb sum at (px, py, pz) over p charges, using the LDS tile.
// If p fits the tile (pp <= tile_cap) the tile is assumed already staged.
// Every thread of the block must call this (it contains barriers when the
// charges do not fit).  A wave none of whose lines needs the sum passes
// sum = false (a wave-uniform value): it still stages and meets the barriers
// but leaves the arithmetic, and its SIMD's issue slots, to the other waves;
// its result is then zero.
template <int S, int U = kUnroll>
__device__ __forceinline__ void field_at(float4* __restrict__ tile, uint32_t tile_cap,
                                         const float4* __restrict__ charges, uint32_t p,
                                         int k, float px, float py, float pz,
                                         float& ex, float& ey, float& ez, bool sum = true)
{
    FieldBody body;
    body.init(px, py, pz);
    constexpr uint32_t G = S * kGroup;
    const uint32_t pp = round_up(p, G);
    if (pp <= tile_cap) {
        if (sum) sum_tile<S, U>(tile, pp, k, body);
    } else {
        for (uint32_t t0 = 0; t0 < p; t0 += tile_cap) {
            const uint32_t cnt = min(tile_cap, p - t0);
            const uint32_t cntp = round_up(cnt, G);
            __syncthreads();
            stage(tile, charges + t0, cnt, cntp);
            __syncthreads();
            if (sum) sum_tile<S, U>(tile, cntp, k, body);
        }
    }
    ex = group_sum<S>(body.acc.sx + body.acc.cx);
    ey = group_sum<S>(body.acc.sy + body.acc.cy);
    ez = group_sum<S>(body.acc.sz + body.acc.cz);
}

// Per-line state carried across steps.
struct LineState {
    float px, py, pz;  // current point x_{s-1}
    float ux, uy, uz;  // previous unit direction (angle rule)
    float h;           // previous step length (angle rule)
    float Px, Py, Pz;  // seed point (ref rule: the SSE path's prevAccum)
    bool alive;        // false once the line has ended (stop rule)
    uint32_t steps;    // steps taken (rows that moved the line)
};

// Record of one step, kept by the adjoint's recompute.
struct StepRec {
    float L;       // |E|
    float c;       // chord |u - u_prev|
    float h;       // step taken
    int branch;    // 0: turn ratio, 1: clamp hmax, 2: clamp hmin, 3: constant,
                   // 4: the line ended here or earlier (no move)
    float ux, uy, uz;
};

// Advance one step.  E must already carry sp.scale.  Operation order is that
// of oracle/restate.c (and, for plain/ref, of the reference's vec3 helpers).
__device__ __forceinline__ void step_update(const StepP& sp, uint32_t s, float ex, float ey,
                                            float ez, LineState& st, StepRec* rec = nullptr)
{
    if (sp.rule == kRulePlain) {
        // CPU_Implement.cpp:89-90 with Vector.h:158-164
        const float len = sqrtf(ex * ex + ey * ey + ez * ez);
        const float sc = sp.res * len;
        st.px = st.px + ex / sc;
        st.py = st.py + ey / sc;
        st.pz = st.pz + ez / sc;
        if (rec) {
            rec->L = len; rec->c = 0.f; rec->h = 1.f / sp.res; rec->branch = 3;
            rec->ux = ex / len; rec->uy = ey / len; rec->uz = ez / len;
        }
    } else if (sp.rule == kRuleRef) {
        // CPU_Implement.cpp:327-331 (P = seed point, never updated)
        float k = ex * ex + ey * ey + ez * ez;
        const float ix = ex - st.Px, iy = ey - st.Py, iz = ez - st.Pz;
        const float cx = iy * st.Pz - iz * st.Py;
        const float cy = iz * st.Px - ix * st.Pz;
        const float cz = ix * st.Py - iy * st.Px;
        k = sqrtf(cx * cx + cy * cy + cz * cz) / (k * sqrtf(k));
        float sc = (k + 1.f) * sp.res;
        sc = sc * sqrtf(ex * ex + ey * ey + ez * ez);
        st.px = st.px + ex / sc;
        st.py = st.py + ey / sc;
        st.pz = st.pz + ez / sc;
    } else {
        const float len = sqrtf(ex * ex + ey * ey + ez * ez);
        const float ux = ex / len, uy = ey / len, uz = ez / len;
        if (sp.stop && s > 1 && ux * st.ux + uy * st.uy + uz * st.uz < 0.f) {
            // the direction reversed: the line went through a sink; it ends at x_{s-1}
            st.alive = false;
            if (rec) rec->branch = 4;
            return;
        }
        float h = sp.h0, c = 0.f;
        int branch = 3;
        if (s > 1) {
            const float dx = ux - st.ux, dy = uy - st.uy, dz = uz - st.uz;
            c = sqrtf(dx * dx + dy * dy + dz * dz);
            const float th = sp.turn * st.h;
            if (c * sp.hmax > th) {
                h = th / c;
                branch = 0;
            } else {
                h = sp.hmax;
                branch = 1;
            }
            if (!(h > sp.hmin)) {
                h = sp.hmin;
                branch = 2;
            }
        }
        st.px = st.px + h * ux;
        st.py = st.py + h * uy;
        st.pz = st.pz + h * uz;
        st.ux = ux; st.uy = uy; st.uz = uz; st.h = h;
        if (rec) {
            rec->L = len; rec->c = c; rec->h = h; rec->branch = branch;
            rec->ux = ux; rec->uy = uy; rec->uz = uz;
        }
    }
}

}  // namespace em
