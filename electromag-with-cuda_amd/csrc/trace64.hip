// trace64.hip -- FP64 field-line tracer (the CalcField_CPU<double> counterpart).
//
// The reference's double path (ElectroMag/src/CPU_Implement.cpp:366-498 SSE2
// curvature form, 38-102 generic plain form, dispatch 515-527) evaluated with
// IEEE binary64 operations in the reference's source order: for each line the
// charges are summed in index order 0..p-1, each PartField term is
//   curvature / angle:  r*k*q / (|r|^2 sqrt(|r|^2))        (Electrostatics.h:38-51)
//   plain:              r * ((k_e*q) / (|r|^2 sqrt(|r|^2)))   (Electrostatics.h:53-60)
// with |r|^2 = (rx*rx + ry*ry) + rz*rz, and the step updates are those of the
// FP32 kernels (em_device.hpp step_update) in double.  sqrt and division are
// correctly rounded and the file is compiled without FMA contraction, so every
// row equals the FP64 restatement of the reference algorithm bit for bit.
//
// Layout: one lane per line (a lane owns its line's whole charge sum, which
// fixes the summation order), or, for problems with too few lines to give
// every SIMD a wave, a group of 2 or 4 lanes per line that splits the terms
// and adds them in the same order; 256 lanes per workgroup; charges staged through
// LDS as 32-byte records, all of them once when they fit a tile, else tile by
// tile each step.  The charge loop is unrolled so the independent per-charge
// terms (a square root and one or three divisions each) overlap while the
// additions stay in order.
//
// Square root and division: the compiler expands IEEE sqrt and division into
// v_rsq_f64 / v_rcp_f64 plus FMA refinement, wrapped in range scaling and
// special-case steps (the square root's ldexp scaling and class test;
// v_div_scale, the scaling in v_div_fmas and v_div_fixup).  For operands in the
// fast domain below those wrapper steps change no value, so the refinement
// alone (sqrt_core, recip_core + div_core) gives the same correctly rounded
// bits, and one refined reciprocal of |r|^2 sqrt(|r|^2) serves a term's three
// divisions.  Tiles and points outside the domain take the compiler's
// expansion (EMTRACE_FP64_EXACT=1 forces it everywhere).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>

#include "em_internal.hpp"

namespace em {

struct StepP64 {
    int rule;
    double res, h0, hmin, hmax, turn;
    int stop;
    int exact;  // 1: the compiler's sqrt and division expansions for every term
};

namespace {

constexpr int kBlock64 = 256;
constexpr uint32_t kTile64 = 2048;  // 64 KiB of LDS: two workgroups per CU
constexpr double kCoulomb64 = 8.987551787E9;  // include/Electrostatics.h:27

// The fast domain.  Every coordinate of the charges and of the current point is
// 0 or has a magnitude in [2^-100, 2^190], and every charge magnitude |q| is in
// [2^-100, 2^100].  Then each nonzero component of r = p - c is a nonzero
// multiple of 2^-152 (both operands are multiples of it), so
//   2^-304 <= |r|^2 <= 2^384 when r != 0,   |r|^2 sqrt(|r|^2) in [2^-456, 2^576],
// and every numerator (k_e q, or r_i k q with k = k_e or 1; k_e < 2^34) is 0
// or has a magnitude in [2^-252, 2^324], its quotient in [2^-828, 2^590]: no
// operand or result is denormal, near overflow, or far enough apart in
// exponent (768) for the wrapper steps to scale.  A zero numerator gives a
// zero term whose sign may differ from the IEEE quotient's; added to a sum
// that starts at +0 it changes no bit.  r = 0 (a point exactly on a charge)
// gives NaN terms in both forms.
constexpr double kDomLo = 0x1p-100, kDomHiX = 0x1p190, kDomHiQ = 0x1p100;

__device__ __forceinline__ bool coord_in_domain(double v)
{
    const double a = fabs(v);
    return v == 0.0 || (a >= kDomLo && a <= kDomHiX);
}

__device__ __forceinline__ bool charge_in_domain(const double4 c)
{
    const double a = fabs(c.w);
    return coord_in_domain(c.x) && coord_in_domain(c.y) && coord_in_domain(c.z) &&
           a >= kDomLo && a <= kDomHiQ;
}

// The charge record as staged in LDS: the plain rule's numerator k_e * q is the
// same product for every line, formed once here.
template <int RULE>
__device__ __forceinline__ double4 staged(const double4 c)
{
    return RULE == kRulePlain ? make_double4(c.x, c.y, c.z, kCoulomb64 * c.w) : c;
}

// sqrt(x) for x in [2^-767, inf): the compiler's refinement of v_rsq_f64
// without its scaling and class test (no-ops there).
__device__ __forceinline__ double sqrt_core(double x)
{
    const double y0 = __builtin_amdgcn_rsq(x);
    double g = x * y0;
    double h = y0 * 0.5;
    const double r = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    double d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    d = __builtin_fma(-g, g, x);
    return __builtin_fma(d, h, g);
}

// The reciprocal refinement of the compiler's division (two Newton steps from
// v_rcp_f64) and its final correction: div_core(n, d, recip_core(d)) == n / d
// whenever v_div_scale and v_div_fixup would leave the operands and the
// quotient unchanged.
__device__ __forceinline__ double recip_core(double d)
{
    double r = __builtin_amdgcn_rcp(d);
    double e = __builtin_fma(-d, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-d, r, 1.0);
    return __builtin_fma(r, e, r);
}

__device__ __forceinline__ double div_core(double n, double d, double r)
{
    const double y = n * r;
    const double e = __builtin_fma(-d, y, n);
    return __builtin_fma(e, r, y);
}

// One PartField term; c.w as staged (k_e q for the plain rule, q otherwise).
template <int RULE, bool FAST>
__device__ __forceinline__ void term64(const double4 c, double px, double py, double pz,
                                       double& ex, double& ey, double& ez)
{
    const double rx = px - c.x, ry = py - c.y, rz = pz - c.z;
    const double r2 = rx * rx + ry * ry + rz * rz;
    const double k = RULE == kRuleRef ? kCoulomb64 : 1.0;
    if constexpr (FAST) {
        const double den = r2 * sqrt_core(r2);
        const double rd = recip_core(den);
        if (RULE == kRulePlain) {
            const double s = div_core(c.w, den, rd);
            ex = rx * s;
            ey = ry * s;
            ez = rz * s;
        } else {
            ex = div_core(rx * k * c.w, den, rd);
            ey = div_core(ry * k * c.w, den, rd);
            ez = div_core(rz * k * c.w, den, rd);
        }
    } else {
        const double den = r2 * sqrt(r2);
        if (RULE == kRulePlain) {
            const double s = c.w / den;
            ex = rx * s;
            ey = ry * s;
            ez = rz * s;
        } else {
            ex = rx * k * c.w / den;
            ey = ry * k * c.w / den;
            ez = rz * k * c.w / den;
        }
    }
}

// Work threaded through the stages of terms64: after stage I, side.at<I>().
struct NoSide {
    template <int I>
    __device__ __forceinline__ void at() {}
};

#define EM_STAGE(I, body)                            \
    _Pragma("unroll") for (int u = 0; u < U; ++u) { \
        body;                                        \
    }                                                \
    side.template at<I>();                           \
    __builtin_amdgcn_sched_barrier(0)

// the division of component J: e_J = div_core(r_J k q, den, rd), stages I..I+3
#define EM_DIV_STAGES(I, J)                                               \
    EM_STAGE(I, r2[u] = e[u][J] * k * c[u].w);                            \
    EM_STAGE(I + 1, g[u] = r2[u] * h[u]);                                 \
    EM_STAGE(I + 2, t[u] = __builtin_fma(-d[u], g[u], r2[u]));            \
    EM_STAGE(I + 3, e[u][J] = __builtin_fma(t[u], h[u], g[u]))

// U terms at once.  The fast form is written stage by stage across the U
// terms (each a dependent chain of ~30 operations) so that their operations
// interleave: with one wave per SIMD nothing else hides a chain's latency.
// sched_barrier(0) after each stage keeps the compiler's scheduler from
// regrouping the stages chain by chain.  Every term's own operations are those
// of term64, in the same order.
template <int RULE, bool FAST, int U, class Side = NoSide>
__device__ __forceinline__ void terms64(const double4 (&c)[U], double px, double py, double pz,
                                        double (&e)[U][3], Side side = Side())
{
    if constexpr (!FAST) {
#pragma unroll
        for (int u = 0; u < U; ++u) term64<RULE, false>(c[u], px, py, pz, e[u][0], e[u][1], e[u][2]);
    } else {
        const double k = RULE == kRuleRef ? kCoulomb64 : 1.0;
        double r2[U], g[U], h[U], t[U], d[U];
        EM_STAGE(0, e[u][0] = px - c[u].x; e[u][1] = py - c[u].y; e[u][2] = pz - c[u].z);
        EM_STAGE(1, r2[u] = e[u][0] * e[u][0] + e[u][1] * e[u][1] + e[u][2] * e[u][2]);
        // sqrt_core
        EM_STAGE(2, h[u] = __builtin_amdgcn_rsq(r2[u]));
        EM_STAGE(3, g[u] = r2[u] * h[u]; h[u] = h[u] * 0.5);
        EM_STAGE(4, t[u] = __builtin_fma(-h[u], g[u], 0.5));
        EM_STAGE(5, g[u] = __builtin_fma(g[u], t[u], g[u]); h[u] = __builtin_fma(h[u], t[u], h[u]));
        EM_STAGE(6, t[u] = __builtin_fma(-g[u], g[u], r2[u]));
        EM_STAGE(7, g[u] = __builtin_fma(t[u], h[u], g[u]));
        EM_STAGE(8, t[u] = __builtin_fma(-g[u], g[u], r2[u]));
        EM_STAGE(9, g[u] = __builtin_fma(t[u], h[u], g[u]));
        // den = |r|^2 sqrt(|r|^2) and recip_core(den)
        EM_STAGE(10, d[u] = r2[u] * g[u]);
        EM_STAGE(11, h[u] = __builtin_amdgcn_rcp(d[u]));
        EM_STAGE(12, t[u] = __builtin_fma(-d[u], h[u], 1.0));
        EM_STAGE(13, h[u] = __builtin_fma(h[u], t[u], h[u]));
        EM_STAGE(14, t[u] = __builtin_fma(-d[u], h[u], 1.0));
        EM_STAGE(15, h[u] = __builtin_fma(h[u], t[u], h[u]));
        if constexpr (RULE == kRulePlain) {
            // s = div_core(k_e q, den, rd), then r * s
            EM_STAGE(16, g[u] = c[u].w * h[u]);
            EM_STAGE(17, t[u] = __builtin_fma(-d[u], g[u], c[u].w));
            EM_STAGE(18, g[u] = __builtin_fma(t[u], h[u], g[u]));
            EM_STAGE(19, e[u][0] = e[u][0] * g[u]; e[u][1] = e[u][1] * g[u]; e[u][2] = e[u][2] * g[u]);
        } else {
            EM_DIV_STAGES(16, 0);
            EM_DIV_STAGES(20, 1);
            EM_DIV_STAGES(24, 2);
        }
    }
}
#undef EM_DIV_STAGES
#undef EM_STAGE

// Charges whose terms are in flight together.  The compiler's expansions
// (kUnroll64): with one wave per SIMD (65,536 lines) 4 is fastest, the plain
// rule's 250-row trace of 16,384 charges taking 440 ms, against 489 ms at 2,
// 497 ms at 8 and 564 ms without unrolling.  The fast form (kUnroll64Fast,
// stages interleaved): 321 ms at 8, against 359 ms at 4, 351 ms at 12 and
// 324 ms at 16 (156 VGPRs at 8).
constexpr int kUnroll64 = 4;
#ifndef EMTRACE_FP64_UNROLL_FAST
#define EMTRACE_FP64_UNROLL_FAST 8
#endif
constexpr int kUnroll64Fast = EMTRACE_FP64_UNROLL_FAST;

template <int RULE, bool FAST>
__device__ __forceinline__ void sum_range(const double4* __restrict__ tile, uint32_t count,
                                          double px, double py, double pz, double& sx, double& sy,
                                          double& sz)
{
    constexpr int U = FAST ? kUnroll64Fast : kUnroll64;
    uint32_t i = 0;
    for (; i + U <= count; i += U) {
        double4 c[U];
#pragma unroll
        for (int u = 0; u < U; ++u) c[u] = tile[i + u];
        double e[U][3];
        terms64<RULE, FAST, U>(c, px, py, pz, e);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            sx += e[u][0];
            sy += e[u][1];
            sz += e[u][2];
        }
    }
    for (; i < count; ++i) {
        double ex, ey, ez;
        term64<RULE, FAST>(tile[i], px, py, pz, ex, ey, ez);
        sx += ex;
        sy += ey;
        sz += ez;
    }
}

// Lines with fewer than one lane each per SIMD's wave slot leave SIMDs idle:
// such problems give each line an S-lane group (S = 2 or 4, within one quad of
// lanes).  Lane k computes the terms of charges i+k, i+S+k, ... and every lane
// of the group adds all of the group's terms in charge order, lane j's term
// read by a quad permutation; the group then holds identical sums, added in
// the order of the one-lane loop.
template <int S, int J>
__device__ __forceinline__ double from_lane(double v)
{
    static_assert(S == 2 || S == 4 || S == 8, "groups of 2, 4 or 8 lanes");
    if constexpr (S == 8) {
        return __shfl(v, J, 8);  // eight-lane groups span two quads: ds_bpermute
    } else {
        // quad_perm selects: lane J of {0,1} and of {2,3}, or lane J of {0..3}
        constexpr int ctrl = S == 4 ? J * 0x55 : (J | (J << 2) | ((2 + J) << 4) | ((2 + J) << 6));
        const long long b = __double_as_longlong(v);
        const int lo = __builtin_amdgcn_update_dpp(0, (int)b, ctrl, 0xF, 0xF, false);
        const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), ctrl, 0xF, 0xF, false);
        return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
    }
}

template <int S, int J = 0>
__device__ __forceinline__ void add_lanes_from(double& s, double v, uint32_t m)
{
    if constexpr (J < S) {
        if (m > (uint32_t)J) {
            s += from_lane<S, J>(v);
            add_lanes_from<S, J + 1>(s, v, m);
        }
    }
}

// s += the terms of lanes 0 .. m-1 of the group, in lane order (m <= S)
template <int S>
__device__ __forceinline__ void add_in_order(double& s, double v, uint32_t m)
{
    if constexpr (S == 1)
        s += v;
    else
        add_lanes_from<S>(s, v, m);
}

// Four- and eight-lane groups in the fast form: lane k of a group adds only
// component k (lanes 3 .. S-1 repeat component 2), so each lane adds S U terms
// per pass instead of reading all 3 S U term components from the other lanes.
// The group's terms pass through a per-wave LDS buffer: written at the end of
// a pass, read at the start of the next, before that pass overwrites them (LDS
// operations of one wave complete in program order, so a lane reads what the
// other lanes of its group wrote one pass earlier).  The additions run in
// charge order, as in the one-lane loop, threaded through the next pass's term
// stages; the three totals are broadcast back to the group before the last
// charges of the tile.
//
// Buffer layout, in 16-byte chunks of two doubles: chunk (q, g, c) at
// q * split_q(S) + 3 g + c holds elements 2q, 2q+1 of component c of group g,
// where element j U + u is lane j's term u.  A reader (g, c) takes chunks
// q = 0 .. S U / 2 - 1 in turn; a quarter-wave's distinct chunks (12 for S = 4,
// 6 for S = 8) are consecutive, so at most two share a bank.  The extra chunk
// in split_q (49, not 48, for S = 4; 25 for S = 8) staggers the writers'
// chunks (lane k writes q = 2k, 2k+1 for U = 4) across banks as well.
constexpr int kSplitU = 4;  // terms per lane per pass
__host__ __device__ constexpr uint32_t split_q(int S) { return 64 / S * 3 + 1; }
__host__ __device__ constexpr uint32_t split_buf(int S) { return S * kSplitU / 2 * split_q(S) * 2; }  // doubles per wave
__host__ __device__ constexpr uint32_t split_lds_bytes(int S)
{
    return (kBlock64 / 64) * split_buf(S) * sizeof(double);  // 24.5 KiB (S = 4), 25 KiB (S = 8)
}
constexpr uint32_t kTileSplit = 1024;  // records staged at once: 32 KiB + ~25 KiB, two per CU

// The previous pass's additions, threaded through this pass's term stages
// (from stage 5 on, by when the LDS reads have landed): element e of the pass
// is lane e % S's term u = e / S, added in e order.
template <int S>
struct SplitAdds {
    double& acc;
    const double (&v)[S * kSplitU];
    template <int I>
    __device__ __forceinline__ void at()
    {
        constexpr int kFirst = 5, kN = S * kSplitU, kPer = (kN + 14) / 15;  // within stages 5 .. 19
        if constexpr (I >= kFirst) {
#pragma unroll
            for (int e = (I - kFirst) * kPer; e < (I - kFirst + 1) * kPer && e < kN; ++e)
                acc += v[(e % S) * kSplitU + e / S];
        }
    }
};

template <int RULE, int S>
__device__ __forceinline__ uint32_t sum_passes_split(const double4* __restrict__ tile,
                                                     uint32_t count, uint32_t k, double px,
                                                     double py, double pz, double& sx, double& sy,
                                                     double& sz, double* __restrict__ xw)
{
    static_assert(S == 4 || S == 8, "four- or eight-lane groups");
    constexpr int U = kSplitU;
    constexpr uint32_t G = S * U, Q = split_q(S);
    const uint32_t g = (threadIdx.x % 64) / S;
    const uint32_t comp = k < 3 ? k : 2;
    double acc = k == 0 ? sx : (k == 1 ? sy : sz);
    double2* const xc = reinterpret_cast<double2*>(xw);
    const uint32_t passes = count / G;
    auto write = [&](const double (&e)[U][3]) {
        // lane k's term u is element k U + u: chunk (k U + u) / 2
#pragma unroll
        for (int cc = 0; cc < 3; ++cc)
#pragma unroll
            for (int u = 0; u < U; u += 2)
                xc[((k * U + u) / 2) * Q + 3 * g + cc] = make_double2(e[u][cc], e[u + 1][cc]);
        __builtin_amdgcn_wave_barrier();
    };
    auto read = [&](double (&v)[S * U]) {
#pragma unroll
        for (int q = 0; q < S * U / 2; ++q) {
            const double2 t = xc[q * Q + 3 * g + comp];
            v[2 * q] = t.x;
            v[2 * q + 1] = t.y;
        }
    };
    // each pass's charge records are read one pass ahead
    double4 c[U];
    if (passes > 0) {
#pragma unroll
        for (int u = 0; u < U; ++u) c[u] = tile[u * S + k];
        const uint32_t nx = passes > 1 ? G : 0;
        double4 cn[U];
#pragma unroll
        for (int u = 0; u < U; ++u) cn[u] = tile[nx + u * S + k];
        __builtin_amdgcn_sched_barrier(0);
        double e[U][3];
        terms64<RULE, true, U>(c, px, py, pz, e);
        write(e);
#pragma unroll
        for (int u = 0; u < U; ++u) c[u] = cn[u];
    }
    for (uint32_t pass = 1; pass < passes; ++pass) {
        double v[S * U];  // the previous pass's terms of component comp, [lane][u]
        read(v);
        const uint32_t nx = (pass + 1 < passes ? pass + 1 : pass) * G;
        double4 cn[U];
#pragma unroll
        for (int u = 0; u < U; ++u) cn[u] = tile[nx + u * S + k];
        __builtin_amdgcn_sched_barrier(0);
        double e[U][3];
        terms64<RULE, true, U>(c, px, py, pz, e, SplitAdds<S>{acc, v});
        write(e);
#pragma unroll
        for (int u = 0; u < U; ++u) c[u] = cn[u];
    }
    if (passes > 0) {
        double v[S * U];
        read(v);
#pragma unroll
        for (int e = 0; e < S * U; ++e) acc += v[(e % S) * U + e / S];
    }
    sx = from_lane<S, 0>(acc);
    sy = from_lane<S, 1>(acc);
    sz = from_lane<S, 2>(acc);
    return passes * G;
}

template <int RULE, int S, bool FAST>
__device__ __forceinline__ void sum_range_group(const double4* __restrict__ tile, uint32_t count,
                                                uint32_t k, double px, double py, double pz,
                                                double& sx, double& sy, double& sz,
                                                double* __restrict__ xw)
{
    if constexpr (S == 1) {
        sum_range<RULE, FAST>(tile, count, px, py, pz, sx, sy, sz);
    } else {
        constexpr int U = FAST ? kUnroll64Fast : kUnroll64;
        constexpr uint32_t G = S * U;  // charges per pass
        uint32_t i = 0;
        if constexpr ((S == 4 || S == 8) && FAST)
            i = sum_passes_split<RULE, S>(tile, count, k, px, py, pz, sx, sy, sz, xw);
        for (; i + G <= count; i += G) {
            double4 c[U];
#pragma unroll
            for (int u = 0; u < U; ++u) c[u] = tile[i + u * S + k];
            double e[U][3];
            terms64<RULE, FAST, U>(c, px, py, pz, e);
#pragma unroll
            for (int u = 0; u < U; ++u) {
                add_in_order<S>(sx, e[u][0], S);
                add_in_order<S>(sy, e[u][1], S);
                add_in_order<S>(sz, e[u][2], S);
            }
        }
        for (; i < count; i += S) {
            // the group's last charges: m = fewer than S terms may remain
            // (uniform over the workgroup: every lane sees the same count)
            const uint32_t m = min((uint32_t)S, count - i);
            double ex = 0.0, ey = 0.0, ez = 0.0;
            if (k < m) term64<RULE, FAST>(tile[i + k], px, py, pz, ex, ey, ez);
            add_in_order<S>(sx, ex, m);
            add_in_order<S>(sy, ey, m);
            add_in_order<S>(sz, ez, m);
        }
    }
}

struct Line64 {
    double px, py, pz;  // current point
    double Px, Py, Pz;  // seed (the reference's prevAccum)
    double ux, uy, uz;  // previous direction (angle rule)
    double h;           // previous step length (angle rule)
    bool alive;
    uint32_t steps;
};

template <int RULE>
__device__ __forceinline__ void step64(const StepP64& sp, uint32_t s, double ex, double ey,
                                       double ez, Line64& st)
{
    if (RULE == kRulePlain) {
        const double len = sqrt(ex * ex + ey * ey + ez * ez);
        const double sc = sp.res * len;
        st.px = st.px + ex / sc;
        st.py = st.py + ey / sc;
        st.pz = st.pz + ez / sc;
    } else if (RULE == kRuleRef) {
        double k = ex * ex + ey * ey + ez * ez;
        const double ix = ex - st.Px, iy = ey - st.Py, iz = ez - st.Pz;
        const double cx = iy * st.Pz - iz * st.Py;
        const double cy = iz * st.Px - ix * st.Pz;
        const double cz = ix * st.Py - iy * st.Px;
        k = sqrt(cx * cx + cy * cy + cz * cz) / (k * sqrt(k));
        double sc = (k + 1.0) * sp.res;
        sc = sc * sqrt(ex * ex + ey * ey + ez * ez);
        st.px = st.px + ex / sc;
        st.py = st.py + ey / sc;
        st.pz = st.pz + ez / sc;
    } else {
        const double len = sqrt(ex * ex + ey * ey + ez * ez);
        const double ux = ex / len, uy = ey / len, uz = ez / len;
        if (sp.stop && s > 1 && ux * st.ux + uy * st.uy + uz * st.uz < 0.0) {
            st.alive = false;  // went through a sink: the line ends at x_{s-1}
            return;
        }
        double h = sp.h0;
        if (s > 1) {
            const double dx = ux - st.ux, dy = uy - st.uy, dz = uz - st.uz;
            const double c = sqrt(dx * dx + dy * dy + dz * dz);
            const double th = sp.turn * st.h;
            h = (c * sp.hmax > th) ? th / c : sp.hmax;
            h = h > sp.hmin ? h : sp.hmin;
        }
        st.px = st.px + h * ux;
        st.py = st.py + h * uy;
        st.pz = st.pz + h * uz;
        st.ux = ux;
        st.uy = uy;
        st.uz = uz;
        st.h = h;
    }
    st.steps = s;
}

template <int RULE, int S>
__global__ __launch_bounds__(kBlock64) void k_trace64(const double4* __restrict__ charges,
                                                      uint32_t p, double* __restrict__ x,
                                                      double* __restrict__ y,
                                                      double* __restrict__ z, uint32_t n,
                                                      uint32_t len, uint64_t pitch, StepP64 sp,
                                                      uint32_t tile, uint32_t* steps_out,
                                                      uint32_t row0)
{
    extern __shared__ double4 lds64[];
    const uint32_t k = threadIdx.x % S;
    const uint32_t line = blockIdx.x * (kBlock64 / S) + threadIdx.x / S;
    const bool valid = line < n;
    const bool writer = valid && k == 0;
    Line64 st{};
    if (valid) {
        // rows row0 .. len-1; row0 > 1 (plain and ref rules) continues from
        // row row0-1, the seed staying row 0
        const uint64_t from = (uint64_t)(row0 - 1) * pitch + line;
        st.px = x[from];
        st.py = y[from];
        st.pz = z[from];
        st.Px = x[line];
        st.Py = y[line];
        st.Pz = z[line];
    }
    st.h = sp.h0;
    st.steps = row0 - 1;
    st.alive = valid;
    const bool single = p <= tile;
    // this wave's exchange buffer (four- and eight-lane groups), after the tile
    double* const xw = S >= 4 ? reinterpret_cast<double*>(lds64 + tile) + (threadIdx.x / 64) * split_buf(S)
                              : nullptr;
    // the tile holds a charge outside the fast domain (staged tiles: every step)
    bool tile_out = sp.exact != 0;
    if (single) {
        bool out = false;
        for (uint32_t i = threadIdx.x; i < p; i += kBlock64) {
            const double4 c = charges[i];
            out |= !charge_in_domain(c);
            lds64[i] = staged<RULE>(c);
        }
        tile_out = __syncthreads_or(out) || tile_out;
    }
    for (uint32_t s = row0; s < len; ++s) {
        // the workgroup skips the field once none of its lines is alive
        if (__syncthreads_or(st.alive)) {
            double ex = 0.0, ey = 0.0, ez = 0.0;
            // wave-uniform: every moving line of the wave has its point in the domain
            const bool pts = __all(!st.alive || (coord_in_domain(st.px) && coord_in_domain(st.py) &&
                                                  coord_in_domain(st.pz)));
            if (single) {
                if (st.alive) {
                    if (pts && !tile_out)
                        sum_range_group<RULE, S, true>(lds64, p, k, st.px, st.py, st.pz, ex, ey, ez, xw);
                    else
                        sum_range_group<RULE, S, false>(lds64, p, k, st.px, st.py, st.pz, ex, ey, ez, xw);
                }
            } else {
                for (uint32_t t0 = 0; t0 < p; t0 += tile) {
                    const uint32_t cnt = min(tile, p - t0);
                    __syncthreads();
                    bool out = false;
                    for (uint32_t i = threadIdx.x; i < cnt; i += kBlock64) {
                        const double4 c = charges[t0 + i];
                        out |= !charge_in_domain(c);
                        lds64[i] = staged<RULE>(c);
                    }
                    const bool fast = !__syncthreads_or(out) && pts && sp.exact == 0;
                    if (st.alive) {
                        if (fast)
                            sum_range_group<RULE, S, true>(lds64, cnt, k, st.px, st.py, st.pz, ex,
                                                           ey, ez, xw);
                        else
                            sum_range_group<RULE, S, false>(lds64, cnt, k, st.px, st.py, st.pz, ex,
                                                            ey, ez, xw);
                    }
                }
            }
            if (st.alive) step64<RULE>(sp, s, ex, ey, ez, st);
        }
        if (writer) {
            const uint64_t at = (uint64_t)s * pitch + line;
            x[at] = st.px;
            y[at] = st.py;
            z[at] = st.pz;
        }
    }
    if (writer && steps_out) steps_out[line] = st.steps;
}

int validate64(size_t p, size_t n, size_t len, const StepP64& sp)
{
    // the reference's order: n, resolution, rows (CPU_Implement.cpp:43-52, 213-230)
    if (!n) EM_FAIL(EM_ERR_NO_LINES, "n == 0");
    if (sp.rule == kRulePlain || sp.rule == kRuleRef) {
        if (sp.res == 0.0) EM_FAIL(EM_ERR_RESOLUTION, "resolution == 0");
        if (!std::isfinite(sp.res)) EM_FAIL(EM_ERR_INVALID, "resolution not finite");
    } else if (sp.rule == kRuleAngle) {
        if (!(sp.h0 > 0.0 && sp.hmin > 0.0 && sp.hmax >= sp.hmin && sp.turn > 0.0) ||
            !std::isfinite(sp.hmax) || !std::isfinite(sp.h0))
            EM_FAIL(EM_ERR_INVALID, "angle rule needs h0 > 0, 0 < hmin <= hmax, turn > 0");
    } else {
        EM_FAIL(EM_ERR_INVALID, "unknown step rule %d", sp.rule);
    }
    if (sp.stop && sp.rule != kRuleAngle)
        EM_FAIL(EM_ERR_INVALID, "stop is defined for the angle rule only");
    if (len < 2) EM_FAIL(EM_ERR_LENGTH, "len %zu < 2", len);
    if (p > 0xFFFFFFFFull || n > 0xFFFFFFFFull || len > 0xFFFFFFFFull)
        EM_FAIL(EM_ERR_INVALID, "p, n and len must each fit 32 bits");
    return EM_OK;
}

StepP64 from_step(const em_step& st)
{
    StepP64 sp;
    sp.rule = st.rule;
    sp.res = st.resolution;
    sp.h0 = st.h0;
    sp.hmin = st.hmin;
    sp.hmax = st.hmax;
    sp.turn = st.turn;
    sp.stop = st.stop ? 1 : 0;
    sp.exact = 0;
    return sp;
}

// Lanes per line.  A wave's time is set by its lanes' share of the terms, and
// a SIMD runs its waves one round after another: with one lane per line every
// wave sums all p terms, with four lanes a quarter (11.2 against 28.6 ms per
// round for 1,024 charges and 300 rows, ref rule).  Eight lanes per line while
// they fill at most one wave per SIMD (n <= 8,192 on 256 CUs: 6.9 against
// 10.7 ms with four lanes at 8,192 lines; at 16,384, two waves per SIMD, 11.8
// against 10.6 ms).  Four lanes per line up to half a lane per SIMD slot
// (n <= 32,768: at most two rounds, two workgroups per CU); beyond, one lane per
// line (one round up to 65,536 lines) is faster (at 40,000 / 60,000 lines:
// 28.7 / 28.6 ms, against 29.7 / 38.2 ms with four lanes and 39.4 / 38.9 ms
// with two).
uint32_t lanes64(const DeviceInfo& dev, uint32_t n)
{
    // EMTRACE_FP64_LANES = 1, 2 or 4 overrides the choice (the rows do not
    // depend on it)
    if (const char* env = std::getenv("EMTRACE_FP64_LANES")) {
        const int v = std::atoi(env);
        if (v == 1 || v == 2 || v == 4 || v == 8) return (uint32_t)v;
    }
    const uint64_t simd_lanes = (uint64_t)dev.cus * 4 * 64;
    if ((uint64_t)n * 8 <= simd_lanes) return 8;
    return (uint64_t)n * 2 <= simd_lanes ? 4 : 1;
}

template <int RULE, int S>
void launch64_s(dim3 grid, size_t lds, hipStream_t stream, const double4* q, uint32_t p, double* x,
                double* y, double* z, uint32_t n, uint32_t len, uint64_t pitch, const StepP64& sp,
                uint32_t tile, uint32_t* steps, uint32_t row0)
{
    hipLaunchKernelGGL((k_trace64<RULE, S>), grid, dim3(kBlock64), lds, stream, q, p, x, y, z, n,
                       len, pitch, sp, tile, steps, row0);
}

template <int RULE>
void launch64_rule(uint32_t S, size_t lds, hipStream_t stream, const double4* q, uint32_t p,
                   double* x, double* y, double* z, uint32_t n, uint32_t len, uint64_t pitch,
                   const StepP64& sp, uint32_t tile, uint32_t* steps, uint32_t row0)
{
    const uint32_t per = kBlock64 / S;
    const dim3 grid((n + per - 1) / per);
    if (S == 8)
        launch64_s<RULE, 8>(grid, lds, stream, q, p, x, y, z, n, len, pitch, sp, tile, steps, row0);
    else if (S == 4)
        launch64_s<RULE, 4>(grid, lds, stream, q, p, x, y, z, n, len, pitch, sp, tile, steps, row0);
    else if (S == 2)
        launch64_s<RULE, 2>(grid, lds, stream, q, p, x, y, z, n, len, pitch, sp, tile, steps, row0);
    else
        launch64_s<RULE, 1>(grid, lds, stream, q, p, x, y, z, n, len, pitch, sp, tile, steps, row0);
}

// Charge records staged in LDS at once: all of them up to kTile64 (64 KiB, two
// workgroups per CU), else tiles of half that; four- and eight-lane groups also
// hold their exchange buffers (~25 KiB), so their tiles stop at kTileSplit.
uint32_t tile64(uint32_t p, uint32_t S)
{
    const uint32_t cap = S >= 4 ? kTileSplit : kTile64;
    return p <= cap ? std::max<uint32_t>(p, 1) : kTile64 / 2;
}

// Rows row0 .. len-1 (row0 > 1: plain and ref rules only).
int launch64(const DeviceInfo& dev, const double* charges, uint32_t p, double* x, double* y,
             double* z, uint32_t n, uint32_t len, uint64_t pitch, const StepP64& sp,
             uint32_t* steps, hipStream_t stream, uint32_t row0 = 1)
{
    if (row0 < 1 || (row0 > 1 && sp.rule == kRuleAngle))
        EM_FAIL(EM_ERR_INVALID, "row0 %u: the angle rule traces from row 1", row0);
    const uint32_t S = lanes64(dev, n);
    const uint32_t tile = tile64(p, S);
    const size_t lds = (size_t)tile * sizeof(double4) +
                       (S == 8 ? split_lds_bytes(8) : S == 4 ? split_lds_bytes(4) : 0);
    const double4* q = reinterpret_cast<const double4*>(charges);
    // EMTRACE_FP64_EXACT=1: every term through the compiler's sqrt and division
    // expansions (the same bits; for tests and timing of the fast domain)
    const char* env = std::getenv("EMTRACE_FP64_EXACT");
    StepP64 spk = sp;
    spk.exact = env && *env && *env != '0' ? 1 : 0;
    switch (sp.rule) {
    case kRulePlain:
        launch64_rule<kRulePlain>(S, lds, stream, q, p, x, y, z, n, len, pitch, spk, tile, steps, row0);
        break;
    case kRuleRef:
        launch64_rule<kRuleRef>(S, lds, stream, q, p, x, y, z, n, len, pitch, spk, tile, steps, row0);
        break;
    default:
        launch64_rule<kRuleAngle>(S, lds, stream, q, p, x, y, z, n, len, pitch, spk, tile, steps, row0);
        break;
    }
    EM_HIP(hipGetLastError());
    return EM_OK;
}

double now64()
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// Host arrays in, host arrays out, through the per-device buffers, streams and
// events the FP32 host-array forms keep (host_path_acquire).  The plain and ref
// rules trace in row chunks whose copies back overlap the next chunk's kernel.
int host_trace64(const double* charges, size_t p, double* x, double* y, double* z, size_t n,
                 size_t len, const StepP64& sp, em_perf* perf, uint32_t* steps)
{
    const double t0 = now64();
    if (perf) {
        set_progress(perf, 0.0);
        perf->time_s = perf->kernel_s = perf->h2d_s = perf->d2h_s = 0.0;
        perf->h2d_bytes = perf->d2h_bytes = perf->gflops = 0.0;
    }
    EM_TRY(validate64(p, n, len, sp));
    if ((p && !charges) || !x || !y || !z) EM_FAIL(EM_ERR_INVALID, "NULL array");
    DeviceInfo dev;
    EM_TRY(current_device_info(dev));

    const size_t row = n * sizeof(double);
    std::unique_lock<std::mutex> lk;
    HostPathBuffers* hp = nullptr;
    EM_TRY(host_path_acquire(dev.device, p * sizeof(double4), 3 * len * row,
                             steps ? n * sizeof(uint32_t) : 0, lk, &hp));
    HostPathBuffers& b = *hp;
    double* dx = static_cast<double*>(b.lines);
    double* dy = dx + len * n;
    double* dz = dy + len * n;
    uint32_t* dsteps = steps ? static_cast<uint32_t*>(b.steps) : nullptr;

    EM_HIP(hipEventRecord(b.ev[0], b.stream));
    const CopySeg in[4] = {{dx, x, row}, {dy, y, row}, {dz, z, row},
                           {b.charges, charges, p * sizeof(double4)}};
    EM_TRY(h2d_staged(dev.device, in, p ? 4 : 3, b.stream));
    EM_HIP(hipEventRecord(b.ev[1], b.stream));
    const size_t out_bytes = (len - 1) * row;
    const uint32_t chunks =
        host_chunks(sp.rule == kRulePlain || sp.rule == kRuleRef, len, 3 * out_bytes);
    // chunk c covers rows bound(c) .. bound(c+1)-1
    auto bound = [&](uint32_t c) { return (uint32_t)(1 + (uint64_t)(len - 1) * c / chunks); };
    for (uint32_t c = 0; c < chunks; ++c) {
        EM_TRY(launch64(dev, static_cast<const double*>(b.charges), (uint32_t)p, dx, dy, dz,
                        (uint32_t)n, bound(c + 1), n, sp, dsteps, b.stream, bound(c)));
        EM_HIP(hipEventRecord(b.chunk_ev[c], b.stream));
    }
    EM_HIP(hipEventRecord(b.ev[2], b.stream));
    double d2h = 0.0;
    for (uint32_t c = 0; c < chunks; ++c) {
        const size_t r0 = bound(c), bytes = (size_t)(bound(c + 1) - r0) * row;
        const size_t o = r0 * n;
        const CopySeg segs[4] = {{x + o, dx + o, bytes}, {y + o, dy + o, bytes},
                                 {z + o, dz + o, bytes}, {steps, dsteps, n * sizeof(uint32_t)}};
        prefault_pages(segs, 3);  // while the GPU traces
        EM_HIP(hipEventSynchronize(b.chunk_ev[c]));  // d2h_s: transfer time alone
        EM_HIP(hipStreamWaitEvent(b.copy, b.chunk_ev[c], 0));
        const double td = now64();
        EM_TRY(d2h_staged(dev.device, segs, (steps && c + 1 == chunks) ? 4 : 3, b.copy));
        d2h += now64() - td;
        set_progress(perf, (double)bound(c + 1) / (double)len);
    }
    EM_HIP(hipEventSynchronize(b.ev[2]));

    float ms[2] = {0, 0};
    EM_HIP(hipEventElapsedTime(&ms[0], b.ev[0], b.ev[1]));
    EM_HIP(hipEventElapsedTime(&ms[1], b.ev[1], b.ev[2]));  // the kernels (chunks overlap copies)
    if (perf) {
        const double ks = ms[1] * 1e-3, dn = (double)n, dl = (double)len, dp = (double)p;
        perf->time_s = now64() - t0;
        perf->kernel_s = ks;
        perf->h2d_s = ms[0] * 1e-3;
        perf->d2h_s = d2h;
        perf->h2d_bytes = (double)(p * sizeof(double4) + 3 * row);
        perf->d2h_bytes = (double)(3 * out_bytes);
        perf->interactions = dp * dn * (dl - 1.0);
        perf->interactions_per_s = ks > 0 ? perf->interactions / ks : 0.0;
        perf->line_steps_per_s = ks > 0 ? dn * (dl - 1.0) / ks : 0.0;
        // CPU_Implement.cpp:347-349
        perf->gflops = (dn * (dl - 1.0) * (dp * 18.0 + 13.0)) / perf->time_s / 1e9;
        perf->lanes_per_line = (int32_t)lanes64(dev, (uint32_t)n);
        perf->tile = (int32_t)tile64((uint32_t)p, lanes64(dev, (uint32_t)n));
        perf->block = kBlock64;
        perf->unroll = 0;  // the FP64 kernel's loop has no pipelined groups
        set_progress(perf, 1.0);
    }
    return EM_OK;
}

}  // namespace
}  // namespace em

using namespace em;

extern "C" {

int em_calc_lines_f64(const double* charges, size_t p, double* x, double* y, double* z, size_t n,
                      size_t len, double resolution, int use_curvature, em_perf* perf)
{
    clear_error();
    StepP64 sp{};
    sp.rule = use_curvature ? kRuleRef : kRulePlain;
    sp.res = resolution;
    return host_trace64(charges, p, x, y, z, n, len, sp, perf, nullptr);
}

int em_trace_host_f64(const double* charges, size_t p, double* x, double* y, double* z, size_t n,
                      size_t len, const em_step* step, em_perf* perf, uint32_t* steps)
{
    clear_error();
    if (!step) EM_FAIL(EM_ERR_INVALID, "step is NULL");
    return host_trace64(charges, p, x, y, z, n, len, from_step(*step), perf, steps);
}

int em_trace_dev_f64(const double* d_charges, size_t p, double* d_x, double* d_y, double* d_z,
                     size_t n, size_t len, size_t pitch, const em_step* step, uint32_t* d_steps,
                     void* stream)
{
    clear_error();
    if (!step) EM_FAIL(EM_ERR_INVALID, "step is NULL");
    const StepP64 sp = from_step(*step);
    EM_TRY(validate64(p, n, len, sp));
    if ((p && !d_charges) || !d_x || !d_y || !d_z) EM_FAIL(EM_ERR_INVALID, "NULL device pointer");
    if (pitch < n) EM_FAIL(EM_ERR_INVALID, "pitch %zu < n %zu", pitch, n);
    DeviceInfo dev;
    EM_TRY(current_device_info(dev));
    return launch64(dev, d_charges, (uint32_t)p, d_x, d_y, d_z, (uint32_t)n, (uint32_t)len, pitch, sp,
                    d_steps, static_cast<hipStream_t>(stream));
}

}  // extern "C"
