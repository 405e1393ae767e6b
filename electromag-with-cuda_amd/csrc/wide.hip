// wide.hip -- the angle rule with line ends for 4,096 to 16,384 charges:
// one 1,024-lane workgroup per CU holds every charge record in registers and
// traces up to kWideSlots lines at a time.
//
// Lane t keeps records t, t + 1024, ..., t + 1024 (R - 1) (zero-charge padding
// beyond p), so a step needs no memory traffic for charges, and the 16 waves
// (four per SIMD) keep the SIMDs issuing.  A workgroup's lines form two
// groups that alternate: while every wave sums one group's lines (each lane
// in plain FP32 over its records, in record order), wave 0 first takes the
// step of the other group, whose sums the previous half-step produced.  The
// step's dependent chain (square root, divisions, the stop and the clamps)
// then runs beside the other waves' sums instead of between them, and a
// half-step needs one barrier.  Within a wave, the lines' sums are combined
// by a reduce-scatter over the 64 lanes (lane-pair swaps across rows, DPP
// within a row), so each lane ends with one (line, component) total; the 16
// waves' totals meet in LDS and wave 0 adds them in wave order.  A line's
// arithmetic does not depend on which workgroup traces it or when, so the
// rows do not depend on the schedule.
//
// Line lengths are very uneven (in the 16,384-charge, 65,536-line problem the
// median line ends after 52 steps while 1.8 % run all 2,499), and a line's
// steps are a serial dependence: the lines that never end must not be left to
// start late.  Three passes (three launches of k_wide):
//   FRESH  -- lines from a counter, in index order, into free slots.  The
//             workgroup that finds the counter exhausted raises a flag; each
//             workgroup, at its next half-step, writes its lines' states to
//             list 1 and leaves (no workgroup drains alone).
//   SETTLE -- list 1, ordered by steps taken (most first) and dealt to the
//             workgroups in snake order, each line for up to kWideSettle more
//             steps; lines still moving then go to list 2.  After these steps
//             the lines still moving are, but for a few, the ones that never
//             end, and each one's work left is known.
//   FINISH -- list 2, ordered by steps taken (fewest first: most work left)
//             and dealt in snake order, to the end.  Its idle waves also
//             write the rows after the end of every line that ended before
//             it; k_wide_fill_list then writes those of list 2's lines.
// k_wide_order sorts a list's entries by their step (one workgroup, bitonic
// in LDS).
#include <cstdlib>

#include "em_internal.hpp"
#include "wide.hpp"

namespace em {

namespace {

constexpr int kWideLanes = 1024;
constexpr int kWideWaves = kWideLanes / 64;
constexpr int kGroupMax = kWideMaxSlots / 2;  // lines of one group at most (16)
constexpr uint32_t kWideSortMax = 8192;       // entries k_wide_order sorts (64 KiB of keys)
static_assert(kWideLanes * 16 >= (int)kWideMaxCharges, "capacity");
static_assert(kGroupMax == 16, "slot lanes: group g holds lanes 16 g .. 16 g + 15 of wave 0");

enum : int { kFresh = 0, kSettle = 1, kFinish = 2 };

// LDS shared by the waves of a workgroup.
struct WideShared {
    float part[2][kGroupMax][4][kWideWaves];  // each wave's totals: group, line, component
    float4 pts[2][kGroupMax];                  // the points the next sums of a group use
    float4 slot[2 * kGroupMax][3];             // wave 0's line states (see slot_load)
    uint32_t cnt[2];                           // lines of each group
    uint32_t exit[2];                          // written in half-step h at [h & 1]
    uint32_t fill_k;                           // FINISH: this workgroup's next fill tile
};

// ---- rows after a line's end (FINISH fills them for the lines that ended before it) ----
// A tile is 64 consecutive lines x kWideFillRows rows; workgroup b of G takes
// tiles b, b + G, ...  Before FINISH the lines of list 2 (still moving) get
// kUnended in steps[], and the tiles skip them.  A list-2 line that ends
// during FINISH has its step written then, and a tile read after that may
// fill its rows from an end point not yet visible across the chip's L2s;
// k_wide_fill_list rewrites every list-2 line's rows after FINISH, so the
// rows end the same either way.
constexpr uint32_t kWideFillRows = 16;
constexpr uint32_t kUnended = 0xFFFFFFFFu;

__device__ __forceinline__ uint64_t fill_tiles(const Cfg& cfg)
{
    return (uint64_t)((cfg.n + 63u) / 64u) * ((cfg.len + kWideFillRows - 1u) / kWideFillRows);
}

// Rows of tile t after each (ended) line's end get the line's end point.
__device__ __forceinline__ void fill_tile(const Cfg& cfg, float* __restrict__ X, float* __restrict__ Y,
                                          float* __restrict__ Z, const uint32_t* __restrict__ steps,
                                          uint64_t t, int lane)
{
    const uint32_t nb = (cfg.n + 63u) / 64u;
    const uint32_t line = (uint32_t)(t % nb) * 64u + (uint32_t)lane;
    const uint32_t rc = (uint32_t)(t / nb);
    if (line >= cfg.n) return;
    const uint32_t e = steps[cfg.line_index + line];
    if (e == kUnended) return;
    const uint32_t r0 = max(e + 1u, rc * kWideFillRows);
    const uint32_t r1 = min(cfg.len, (rc + 1u) * kWideFillRows);
    if (r0 >= r1) return;
    const uint64_t base = cfg.line_off + line;
    const uint64_t src = base + (uint64_t)e * cfg.pitch;
    const float x = X[src], y = Y[src], z = Z[src];
    for (uint32_t r = r0; r < r1; ++r) {
        const uint64_t o = base + (uint64_t)r * cfg.pitch;
        X[o] = x;
        Y[o] = y;
        Z[o] = z;
    }
}

// One line's state as wave 0 keeps it in LDS: {x, y, z, h}, {ux, uy, uz, line},
// {s (next step), steps taken, steps taken in this pass, 0}.
struct WideSlot {
    LineState st;
    uint32_t line;
    uint32_t s;
    uint32_t taken;
};

__device__ __forceinline__ void slot_to_rec(const WideSlot& sl, float4* __restrict__ e)
{
    e[0] = make_float4(sl.st.px, sl.st.py, sl.st.pz, sl.st.h);
    e[1] = make_float4(sl.st.ux, sl.st.uy, sl.st.uz, __uint_as_float(sl.line));
    e[2] = make_float4(__uint_as_float(sl.s), __uint_as_float(sl.st.steps), __uint_as_float(sl.taken), 0.f);
}

__device__ __forceinline__ void rec_to_slot(const float4* __restrict__ e, WideSlot& sl)
{
    const float4 a = e[0], b = e[1], c = e[2];
    sl.st.px = a.x; sl.st.py = a.y; sl.st.pz = a.z; sl.st.h = a.w;
    sl.st.ux = b.x; sl.st.uy = b.y; sl.st.uz = b.z;
    sl.st.Px = sl.st.Py = sl.st.Pz = 0.f;
    sl.st.alive = true;
    sl.line = __float_as_uint(b.w);
    sl.s = __float_as_uint(c.x);
    sl.st.steps = __float_as_uint(c.y);
    sl.taken = __float_as_uint(c.z);
}

// ---- reduce-scatter of M values over the 64 lanes of a wave ----------------
// After the call lane L holds in a[0] the wave's total of value index
// L >> (6 - log2 M).  Levels go from lane bit 5 down: at each, the lanes with
// the bit clear keep the lower half of the values and add their partner's
// copy; the other lanes keep the upper half.  Bits 5 and 4 pair rows through
// v_permlane32_swap / v_permlane16_swap (the sum r[0] + r[1] is lower + upper
// in both lanes of a pair); bits 3..0 pair lanes of a row through DPP (row
// rotate by 8, half-row mirror, quad permutes).  Once one value is left the
// remaining levels add in both lanes of each pair (IEEE addition commutes, so
// both hold the same bits).
template <int CTRL>
__device__ __forceinline__ float dpp_partner(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

template <int CTRL, int BIT, int C>
__device__ __forceinline__ void rs_dpp_level(float (&a)[C], int lane)
{
    const bool hi = (lane >> BIT) & 1;
    if constexpr (C == 1) {
        a[0] = a[0] + dpp_partner<CTRL>(a[0]);
    } else {
#pragma unroll
        for (int i = 0; i < C / 2; ++i) {
            const float x = a[i] + dpp_partner<CTRL>(a[i]);
            const float y = a[i + C / 2] + dpp_partner<CTRL>(a[i + C / 2]);
            a[i] = hi ? y : x;
        }
    }
}

template <int M>
__device__ __forceinline__ float reduce_scatter(float (&v)[M], int lane)
{
    static_assert(M == 4 || M == 8 || M == 16, "4, 8 or 16 values");
    float a[M];
#pragma unroll
    for (int i = 0; i < M; ++i) a[i] = v[i];
#pragma unroll
    for (int i = 0; i < M / 2; ++i) {
        const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(a[i]), __float_as_uint(a[i + M / 2]),
                                                        false, false);
        a[i] = __uint_as_float(r[0]) + __uint_as_float(r[1]);
    }
#pragma unroll
    for (int i = 0; i < M / 4; ++i) {
        const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(a[i]), __float_as_uint(a[i + M / 4]),
                                                        false, false);
        a[i] = __uint_as_float(r[0]) + __uint_as_float(r[1]);
    }
    // M / 4 values left per lane
    float b[M / 4];
#pragma unroll
    for (int i = 0; i < M / 4; ++i) b[i] = a[i];
    rs_dpp_level<0x128, 3, M / 4>(b, lane);                      // row_ror:8
    constexpr int C2 = M / 8 > 0 ? M / 8 : 1;
    float c[C2];
#pragma unroll
    for (int i = 0; i < C2; ++i) c[i] = b[i];
    rs_dpp_level<0x141, 2, C2>(c, lane);                         // row_half_mirror
    constexpr int C3 = M / 16 > 0 ? M / 16 : 1;
    float d[C3];
#pragma unroll
    for (int i = 0; i < C3; ++i) d[i] = c[i];
    rs_dpp_level<0x4E, 1, C3>(d, lane);                          // quad_perm [2, 3, 0, 1]
    float e[1] = {d[0]};
    rs_dpp_level<0xB1, 0, 1>(e, lane);                           // quad_perm [1, 0, 3, 2]
    return e[0];
}

// Sums of NL lines (points pts[0 .. NL-1], broadcast from LDS) over this
// lane's R records; the wave's totals go to part[l0 + i][c][wave].
template <int R, int NL>
__device__ __forceinline__ void wide_sums(const float4 (&rec)[R], const float4* __restrict__ pts,
                                          float (*part)[4][kWideWaves], uint32_t l0, int lane, int wave,
                                          uint32_t n)
{
    float px[NL], py[NL], pz[NL], ex[NL], ey[NL], ez[NL];
#pragma unroll
    for (int i = 0; i < NL; ++i) {
        const float4 q = pts[l0 + i];
        px[i] = q.x; py[i] = q.y; pz[i] = q.z;
        ex[i] = ey[i] = ez[i] = 0.f;
    }
#pragma unroll
    for (int j = 0; j < R; ++j) {
#pragma unroll
        for (int i = 0; i < NL; ++i) coulomb(rec[j], px[i], py[i], pz[i], ex[i], ey[i], ez[i]);
    }
    constexpr int M = 4 * NL;
    float v[M];
#pragma unroll
    for (int i = 0; i < NL; ++i) {
        v[4 * i] = ex[i];
        v[4 * i + 1] = ey[i];
        v[4 * i + 2] = ez[i];
        v[4 * i + 3] = 0.f;
    }
    const float t = reduce_scatter<M>(v, lane);
    constexpr int SH = NL == 4 ? 2 : NL == 2 ? 3 : 4;  // 6 - log2(M)
    const int vi = lane >> SH;
    if ((lane & ((1 << SH) - 1)) == 0 && (vi & 3) != 3) part[l0 + (vi >> 2)][vi & 3][wave] = t;
}

// The sums of the n lines of a group: four at a time, then two, then one.
template <int R>
__device__ __forceinline__ void wide_sums_all(const float4 (&rec)[R], const float4* __restrict__ pts,
                                              float (*part)[4][kWideWaves], uint32_t n, int lane, int wave)
{
    uint32_t l = 0;
    for (; l + 4 <= n; l += 4) wide_sums<R, 4>(rec, pts, part, l, lane, wave, n);
    if (l + 2 <= n) {
        wide_sums<R, 2>(rec, pts, part, l, lane, wave, n);
        l += 2;
    }
    if (l < n) wide_sums<R, 1>(rec, pts, part, l, lane, wave, n);
}

// Entry position of round r of workgroup b in a list dealt to G workgroups in
// snake order (round 0: b, round 1: 2G - 1 - b, ...).
__device__ __forceinline__ uint32_t deal_pos(uint32_t r, uint32_t b, uint32_t G)
{
    return r * G + ((r & 1u) ? G - 1u - b : b);
}

// Wave 0: write the n lines of group g (slot lanes 16 g ..) to list_out.
__device__ __forceinline__ void flush_group(WideShared& sh, uint32_t g, uint32_t n, int lane,
                                            float4* __restrict__ list_out, uint32_t* __restrict__ count_out,
                                            uint32_t out_cap)
{
    if (n == 0) return;
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(count_out, n);
    base = __builtin_amdgcn_readfirstlane(__shfl(base, 0, 64));
    if ((uint32_t)lane < n && base + (uint32_t)lane < out_cap) {
        const float4* s = sh.slot[kGroupMax * g + lane];
        float4* e = list_out + 3 * (uint64_t)(base + (uint32_t)lane);
        e[0] = s[0];
        e[1] = s[1];
        e[2] = s[2];
    }
}

// One pass of the wide trace (see the file comment).  `list_in` / `order_in`
// / `count_in` / `in_cap`: the entries a list pass takes (unused in FRESH);
// `list_out` / `count_out` / `out_cap`: where lines leave this pass (FRESH:
// at the flag; SETTLE: after `cap` steps); `gslots`: lines per group.
// `iter_max` bounds the half-steps of a workgroup (a guard: every pass ends
// well before it).
template <int R, int MODE>
__global__ void __launch_bounds__(kWideLanes) k_wide(Cfg cfg, const float4* __restrict__ charges,
                                                     float* __restrict__ X, float* __restrict__ Y,
                                                     float* __restrict__ Z, StepP sp,
                                                     uint32_t* __restrict__ steps,
                                                     WideCtrl* __restrict__ ctrl, uint32_t gslots,
                                                     const float4* __restrict__ list_in,
                                                     const uint32_t* __restrict__ order_in,
                                                     const uint32_t* __restrict__ count_in,
                                                     uint32_t in_cap, float4* __restrict__ list_out,
                                                     uint32_t* __restrict__ count_out,
                                                     uint32_t out_cap, uint32_t cap,
                                                     uint64_t iter_max, uint32_t fill)
{
    __shared__ WideShared sh;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const uint32_t p = cfg.p;

    // wave 0 takes the steps; the other waves hold the records and sum
    const bool ctl = __builtin_amdgcn_readfirstlane(wave) == 0;

    // ---- wave 0's bookkeeping (meaningful in wave 0 only) ----
    const uint32_t G = gridDim.x;
    const uint32_t b = blockIdx.x;
    const uint32_t m = MODE == kFresh ? 0u : min(__builtin_amdgcn_readfirstlane(*count_in), in_cap);
    uint32_t r_next = 0;       // list passes: next dealing round of this workgroup
    uint32_t gc0 = 0, gc1 = 0;  // lines of group 0 / 1
    bool exhausted = false;    // FRESH: the counter had no more lines
    uint32_t flag_v = 0;       // FRESH: the flag as last read

    // Take up to `want` new lines into group g's slots from `first` on; returns
    // how many (wave 0).
    auto take = [&](uint32_t g, uint32_t first, uint32_t want) -> uint32_t {
        uint32_t got = 0;
        if (MODE == kFresh) {
            if (exhausted || want == 0) return 0u;
            uint32_t g0 = 0;
            if (lane == 0) {
                g0 = atomicAdd(&ctrl->next_line, want);
                if (g0 + want >= cfg.n)
                    __hip_atomic_store(&ctrl->flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            g0 = __builtin_amdgcn_readfirstlane(__shfl(g0, 0, 64));
            got = g0 >= cfg.n ? 0u : min(want, cfg.n - g0);
            exhausted = g0 + want >= cfg.n;
            if ((uint32_t)lane < got) {
                WideSlot sl;
                sl.line = g0 + (uint32_t)lane;
                const uint64_t o = cfg.line_off + sl.line;
                sl.st.px = X[o];
                sl.st.py = Y[o];
                sl.st.pz = Z[o];
                sl.st.ux = sl.st.uy = sl.st.uz = 0.f;
                sl.st.h = sp.h0;
                sl.st.steps = 0;
                sl.s = 1;
                sl.taken = 0;
                slot_to_rec(sl, sh.slot[kGroupMax * g + first + lane]);
            }
        } else {
            while (got < want && deal_pos(r_next + got, b, G) < m) ++got;
            if ((uint32_t)lane < got) {
                const uint32_t e = order_in[deal_pos(r_next + (uint32_t)lane, b, G)];
                const float4* src = list_in + 3 * (uint64_t)e;
                float4* dst = sh.slot[kGroupMax * g + first + lane];
                dst[0] = src[0];
                dst[1] = src[1];
                dst[2] = make_float4(src[2].x, src[2].y, __uint_as_float(0u), 0.f);  // taken = 0
            }
            r_next += got;
        }
        return got;
    };
    auto publish = [&](uint32_t g, uint32_t cnt) {
        if ((uint32_t)lane < cnt) {
            const float4 a = sh.slot[kGroupMax * g + lane][0];
            sh.pts[g][lane] = make_float4(a.x, a.y, a.z, 0.f);
        }
        if (lane == 0) sh.cnt[g] = cnt;
    };

    // Records: lane t of the workgroup holds charges t, t + 1024, ..., t + 1024 (R - 1).
    float4 rec[R];
#pragma unroll
    for (int j = 0; j < R; ++j) {
        const uint32_t i = (uint32_t)tid + (uint32_t)j * (uint32_t)kWideLanes;
        rec[j] = i < p ? charges[cfg.charge_off + i] : make_float4(kFar, kFar, kFar, 0.f);
    }

    if (ctl) {
        // the first lines (slot records 0 .. got-1), alternately into the two groups
        const uint32_t got = take(0u, 0u, 2u * gslots);
        float4 s0, s1, s2;
        const bool mine = (uint32_t)lane < got;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (mine) {
            s0 = sh.slot[lane][0];
            s1 = sh.slot[lane][1];
            s2 = sh.slot[lane][2];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (mine) {
            const uint32_t dst = (lane & 1) ? kGroupMax + (uint32_t)lane / 2u : (uint32_t)lane / 2u;
            sh.slot[dst][0] = s0;
            sh.slot[dst][1] = s1;
            sh.slot[dst][2] = s2;
        }
        gc1 = got / 2u;
        gc0 = got - gc1;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        publish(0u, gc0);
        publish(1u, gc1);
        if (lane == 0) {
            sh.exit[0] = 0u;
            sh.exit[1] = got == 0u ? 1u : 0u;  // read at half-step 0
            sh.fill_k = 0u;
        }
        if (MODE == kFresh && lane == 0)
            flag_v = __hip_atomic_load(&ctrl->flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();

    bool pd0 = false, pd1 = false;  // wave 0: a group's sums of the last half-step await their step
    uint32_t grp = 0;
    for (uint64_t h = 0;; ++h) {
        if (__builtin_amdgcn_readfirstlane(sh.exit[(h + 1u) & 1u]) != 0u || h > iter_max) break;
        const uint32_t n = __builtin_amdgcn_readfirstlane(sh.cnt[grp]);
        const uint32_t og = grp ^ 1u;
        if (ctl) {
                // ---- the step of group og (its sums are complete), then its free slots ----
                const bool flagged = MODE == kFresh && __builtin_amdgcn_readfirstlane(flag_v) != 0u;
                uint32_t nkeep = 0;
                const uint32_t cn = og ? gc1 : gc0;
                const uint32_t cg = og ? gc0 : gc1;  // lines of group grp (n)
                if (og ? pd1 : pd0) {
                    bool keep = false, leave = false;
                    WideSlot sl;
                    if ((uint32_t)lane < cn) {
                        rec_to_slot(sh.slot[kGroupMax * og + lane], sl);
                        // the 16 waves' totals, added as a fixed tree
                        float e[3];
#pragma unroll
                        for (int c = 0; c < 3; ++c) {
                            const float4* q = reinterpret_cast<const float4*>(sh.part[og][lane][c]);
                            const float4 u0 = q[0], u1 = q[1], u2 = q[2], u3 = q[3];
                            const float a0 = (u0.x + u0.y) + (u0.z + u0.w);
                            const float a1 = (u1.x + u1.y) + (u1.z + u1.w);
                            const float a2 = (u2.x + u2.y) + (u2.z + u2.w);
                            const float a3 = (u3.x + u3.y) + (u3.z + u3.w);
                            e[c] = (a0 + a1) + (a2 + a3);
                        }
                        step_update(sp, sl.s, e[0] * sp.scale, e[1] * sp.scale, e[2] * sp.scale, sl.st);
                        if (sl.st.alive) {
                            sl.st.steps = sl.s;
                            const uint64_t o = cfg.line_off + (uint64_t)sl.s * cfg.pitch + sl.line;
                            X[o] = sl.st.px;
                            Y[o] = sl.st.py;
                            Z[o] = sl.st.pz;
                        }
                        const bool done = !sl.st.alive || sl.s + 1 >= cfg.len;
                        if (done) {
                            steps[cfg.line_index + sl.line] = sl.st.steps;
                        } else {
                            ++sl.s;
                            ++sl.taken;
                            leave = MODE == kSettle && sl.taken >= cap;
                            keep = !leave;
                        }
                    }
                    if (MODE == kSettle) {
                        const uint64_t leave_mask = __ballot(leave);
                        if (leave_mask) {
                            const uint32_t nleave = (uint32_t)__popcll(leave_mask);
                            uint32_t base = 0;
                            if (lane == 0) base = atomicAdd(count_out, nleave);
                            base = __builtin_amdgcn_readfirstlane(__shfl(base, 0, 64));
                            const uint32_t rank = __builtin_amdgcn_mbcnt_hi(
                                (uint32_t)(leave_mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)leave_mask, 0u));
                            if (leave && base + rank < out_cap) slot_to_rec(sl, list_out + 3 * (uint64_t)(base + rank));
                        }
                    }
                    // the lines still held go to the front of the group's slots
                    const uint64_t keep_mask = __ballot(keep);
                    nkeep = (uint32_t)__popcll(keep_mask);
                    if (nkeep == cn) {
                        // every line stays where it is
                        if (keep) slot_to_rec(sl, sh.slot[kGroupMax * og + lane]);
                    } else {
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    if (keep) {
                        const uint32_t pos = __builtin_amdgcn_mbcnt_hi(
                            (uint32_t)(keep_mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)keep_mask, 0u));
                        slot_to_rec(sl, sh.slot[kGroupMax * og + pos]);
                    }
                    }
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                } else {
                    nkeep = cn;
                }
                uint32_t cnew = nkeep;
                bool flush = MODE == kFresh && (flagged || exhausted);
                if (!flush) {
                    cnew += take(og, nkeep, gslots - nkeep);
                    flush = MODE == kFresh && exhausted;
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                if (flush) {
                    // every line of the workgroup goes to list 1 as it stands (the
                    // sums of group grp now running are discarded)
                    flush_group(sh, og, cnew, lane, list_out, count_out, out_cap);
                    flush_group(sh, grp, cg, lane, list_out, count_out, out_cap);
                    if (lane == 0) sh.exit[h & 1u] = 1u;
                } else {
                    publish(og, cnew);
                    // done once neither group has lines left to sum or take
                    if (lane == 0) sh.exit[h & 1u] = (n == 0u && cnew == 0u) ? 1u : 0u;
                }
                if (og) gc1 = cnew; else gc0 = cnew;
                if (MODE == kFresh && lane == 0)
                    flag_v = __hip_atomic_load(&ctrl->flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }

        // ---- every wave: the sums of group grp ----
        wide_sums_all<R>(rec, sh.pts[grp], sh.part[grp], n, lane, wave);
        if (grp) { pd1 = n > 0u; pd0 = false; } else { pd0 = n > 0u; pd1 = false; }
        // FINISH: wave 1 (the first of its SIMD to end its sums) fills one tile
        if (MODE == kFinish && fill && __builtin_amdgcn_readfirstlane(wave) == 1) {
            const uint32_t k = __builtin_amdgcn_readfirstlane(sh.fill_k);
            const uint64_t t = b + (uint64_t)k * G;
            if (t < fill_tiles(cfg)) {
                fill_tile(cfg, X, Y, Z, steps, t, lane);
                if (lane == 0) sh.fill_k = k + 1u;
            }
        }
        __syncthreads();
        grp = og;
    }
    // FINISH: every wave fills this workgroup's tiles left
    if (MODE == kFinish && fill) {
        __syncthreads();
        const uint64_t k0 = __builtin_amdgcn_readfirstlane(sh.fill_k);
        const uint64_t nt = fill_tiles(cfg);
        for (uint64_t t = b + (k0 + (uint32_t)wave) * G; t < nt; t += (uint64_t)kWideWaves * G)
            fill_tile(cfg, X, Y, Z, steps, t, lane);
    }
}

// Rows after the end of each line of `list` (the lines FINISH traced): lanes
// are (entry, row chunk) pairs.
__global__ void k_wide_fill_list(Cfg cfg, float* __restrict__ X, float* __restrict__ Y,
                                 float* __restrict__ Z, const uint32_t* __restrict__ steps,
                                 const float4* __restrict__ list, const uint32_t* __restrict__ count,
                                 uint32_t cap)
{
    const uint32_t m = min(__builtin_amdgcn_readfirstlane(*count), cap);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const uint32_t line = __float_as_uint(list[3 * (uint64_t)i + 1].w);
    const uint32_t e = steps[cfg.line_index + line];
    const uint32_t r0 = max(e + 1u, blockIdx.y * 64u);
    const uint32_t r1 = min(cfg.len, (blockIdx.y + 1u) * 64u);
    if (e == kUnended || r0 >= r1) return;
    const uint64_t base = cfg.line_off + line;
    const uint64_t src = base + (uint64_t)e * cfg.pitch;
    const float x = X[src], y = Y[src], z = Z[src];
    for (uint32_t r = r0; r < r1; ++r) {
        const uint64_t o = base + (uint64_t)r * cfg.pitch;
        X[o] = x;
        Y[o] = y;
        Z[o] = z;
    }
}

// order[i] = the list position of the entry of rank i, entries ranked by
// their next step (ascending, or descending when `desc`), ties by position.
// One workgroup; lists longer than kWideSortMax keep their order.
__global__ void __launch_bounds__(1024) k_wide_order(const float4* __restrict__ list,
                                                     const uint32_t* __restrict__ count,
                                                     uint32_t* __restrict__ order, uint32_t cap,
                                                     int desc, uint32_t* __restrict__ mark)
{
    __shared__ uint64_t key[kWideSortMax];
    const uint32_t m = min(__builtin_amdgcn_readfirstlane(*count), cap);
    const uint32_t tid = threadIdx.x;
    // `mark` (steps + line_index) non-null: the list's lines have not ended
    if (mark)
        for (uint32_t i = tid; i < m; i += blockDim.x) mark[__float_as_uint(list[3 * (uint64_t)i + 1].w)] = kUnended;
    if (m > kWideSortMax) {
        for (uint32_t i = tid; i < m; i += blockDim.x) order[i] = i;
        return;
    }
    uint32_t P = 1;
    while (P < m) P <<= 1;
    for (uint32_t i = tid; i < P; i += blockDim.x) {
        if (i < m) {
            const uint32_t s = __float_as_uint(list[3 * (uint64_t)i + 2].x);
            key[i] = ((uint64_t)(desc ? ~s : s) << 32) | i;
        } else {
            key[i] = ~0ull;
        }
    }
    __syncthreads();
    for (uint32_t k = 2; k <= P; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = tid; i < P; i += blockDim.x) {
                const uint32_t o = i ^ j;
                if (o > i) {
                    const uint64_t a = key[i], c = key[o];
                    const bool up = (i & k) == 0;
                    if (up ? a > c : a < c) {
                        key[i] = c;
                        key[o] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (uint32_t i = tid; i < m; i += blockDim.x) order[i] = (uint32_t)key[i];
}

__global__ void k_wide_prepare(WideCtrl* __restrict__ ctrl)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        ctrl->next_line = 0u;
        ctrl->flag = 0u;
        ctrl->count1 = 0u;
        ctrl->count2 = 0u;
    }
}

template <int R>
int wide_launch_r(const DeviceInfo& dev, const Cfg& cfg, const float4* charges, float* x, float* y,
                  float* z, const StepP& sp, uint32_t* steps, hipStream_t stream, char* buf,
                  const WideLayout& L)
{
    WideCtrl* ctrl = reinterpret_cast<WideCtrl*>(buf);
    float4* l1 = reinterpret_cast<float4*>(buf + L.off_l1);
    float4* l2 = reinterpret_cast<float4*>(buf + L.off_l2);
    uint32_t* o1 = reinterpret_cast<uint32_t*>(buf + L.off_o1);
    uint32_t* o2 = reinterpret_cast<uint32_t*>(buf + L.off_o2);
    int per_cu = 0;
    EM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_wide<R, kFresh>, kWideLanes, 0));
    if (per_cu < 1) EM_FAIL(EM_ERR_INVALID, "wide line-ends workgroup does not fit a CU");
    const uint32_t G = L.grid;
    const uint32_t gs = L.slots / 2u;
    // a half-step steps at least one line of a workgroup's (two per line step)
    const uint64_t iter_max = 2u * (uint64_t)cfg.len * ((uint64_t)cfg.n + 1u) + 16u;
    hipLaunchKernelGGL(k_wide_prepare, dim3(1), dim3(64), 0, stream, ctrl);
    EM_HIP(hipGetLastError());
    hipLaunchKernelGGL((k_wide<R, kFresh>), dim3(G), dim3(kWideLanes), 0, stream, cfg, charges, x, y,
                       z, sp, steps, ctrl, gs, nullptr, nullptr, nullptr, 0u, l1, &ctrl->count1,
                       L.list_cap, 0u, iter_max, 0u);
    EM_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_wide_order, dim3(1), dim3(1024), 0, stream, l1, &ctrl->count1, o1,
                       L.list_cap, 1, (uint32_t*)nullptr);
    EM_HIP(hipGetLastError());
    hipLaunchKernelGGL((k_wide<R, kSettle>), dim3(G), dim3(kWideLanes), 0, stream, cfg, charges, x,
                       y, z, sp, steps, ctrl, gs, l1, o1, &ctrl->count1, L.list_cap, l2,
                       &ctrl->count2, L.list_cap, kWideSettle, iter_max, 0u);
    EM_HIP(hipGetLastError());
    // list 2's lines are still moving: FINISH's fill leaves them to k_wide_fill_list
    hipLaunchKernelGGL(k_wide_order, dim3(1), dim3(1024), 0, stream, l2, &ctrl->count2, o2,
                       L.list_cap, 0, steps + cfg.line_index);
    EM_HIP(hipGetLastError());
    hipLaunchKernelGGL((k_wide<R, kFinish>), dim3(G), dim3(kWideLanes), 0, stream, cfg, charges, x,
                       y, z, sp, steps, ctrl, gs, l2, o2, &ctrl->count2, L.list_cap, nullptr,
                       nullptr, 0u, 0u, iter_max, 1u);
    EM_HIP(hipGetLastError());
    const dim3 fill_grid((L.list_cap + 255) / 256, (cfg.len + 63) / 64);
    hipLaunchKernelGGL(k_wide_fill_list, fill_grid, dim3(256), 0, stream, cfg, x, y, z, steps, l2,
                       &ctrl->count2, L.list_cap);
    EM_HIP(hipGetLastError());
    return EM_OK;
}

}  // namespace

bool wide_applies(const Cfg& cfg)
{
    if (const char* env = std::getenv("EMTRACE_ENDS_WIDE"))
        if (env[0] == '0') return false;
    return cfg.p >= kEndsMinCharges && cfg.p <= kWideMaxCharges;
}

WideLayout wide_layout(const DeviceInfo& dev, size_t off)
{
    WideLayout L;
    L.slots = kWideSlots;
    if (const char* env = std::getenv("EMTRACE_WIDE_SLOTS")) {
        const int v = std::atoi(env);
        if (v >= 2 && v <= kWideMaxSlots) L.slots = (uint32_t)v & ~1u;
    }
    L.grid = (uint32_t)dev.cus;
    if (const char* env = std::getenv("EMTRACE_WIDE_GRID")) {
        const int v = std::atoi(env);
        if (v >= 1 && v <= dev.cus) L.grid = (uint32_t)v;
    }
    L.list_cap = L.grid * L.slots;
    auto a256 = [](size_t v) { return (v + 255) / 256 * 256; };
    L.off_l1 = a256(off);
    L.off_l2 = L.off_l1 + a256((size_t)L.list_cap * 3 * sizeof(float4));
    L.off_o1 = L.off_l2 + a256((size_t)L.list_cap * 3 * sizeof(float4));
    L.off_o2 = L.off_o1 + a256((size_t)L.list_cap * sizeof(uint32_t));
    L.end = L.off_o2 + a256((size_t)L.list_cap * sizeof(uint32_t));
    return L;
}

int wide_launch(const DeviceInfo& dev, const Cfg& cfg, const float4* charges, float* x, float* y,
                float* z, const StepP& sp, uint32_t* steps, hipStream_t stream, char* buf,
                const WideLayout& L)
{
    const uint32_t R = (cfg.p + kWideLanes - 1) / kWideLanes;
    if (R <= 4) return wide_launch_r<4>(dev, cfg, charges, x, y, z, sp, steps, stream, buf, L);
    if (R <= 8) return wide_launch_r<8>(dev, cfg, charges, x, y, z, sp, steps, stream, buf, L);
    if (R <= 12) return wide_launch_r<12>(dev, cfg, charges, x, y, z, sp, steps, stream, buf, L);
    if (R <= 16) return wide_launch_r<16>(dev, cfg, charges, x, y, z, sp, steps, stream, buf, L);
    EM_FAIL(EM_ERR_INVALID, "wide line ends: %u charges above %u", cfg.p, kWideMaxCharges);
}

}  // namespace em
