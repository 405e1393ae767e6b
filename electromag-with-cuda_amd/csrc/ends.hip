// ends.hip -- the angle rule with line ends: the entry points, their work
// memory, and the lane-group trace for problems above 16,384 charges (4,096
// to 16,384 charges take the register-resident trace of wide.hip; fewer, the
// one-kernel path of trace.hip).
//
// With line ends most lines of a problem are short (a line that runs into a
// sink ends there) and a few never end (they escape and take every row).  In
// the 16,384-charge, 65,536-line problem the median line ends after 52 steps
// while 1.8 % of the lines run all 2,499 and hold 46 % of the steps.  One
// launch shape cannot serve both: groups of S lanes per line keep a wave busy
// until its slowest line ends, and a line that never ends is then traced at
// the rate of S lanes.  The lane-group trace, in two phases:
//   1. k_trace_ends<S, U>: a persistent grid of S-lane groups.  A group whose
//      line ends takes the next line from a counter; a line still moving
//      after kEndsCap steps is set aside (its state goes to a list) and the
//      group moves on.  The last lines handed out (one per lane group) are
//      the ones the phase drains on once the counter is exhausted: they are
//      set aside after kEndsDrainCap steps already (the lines that end do so
//      early: in the problem above 99 % of them within 110 steps).  Charges
//      go through LDS tiles as in the other kernels.
//   2. k_trace_tail: two 256-lane workgroups (4 waves, one per SIMD) per CU;
//      each owns a share of the lines set aside and traces up to 8 of them at
//      once.  The 256 lanes split every charge sum; each charge record a lane
//      loads (straight from a padded copy in global memory, which stays in
//      L2) serves all the workgroup's lines; wave 0 takes the steps.
//   3. k_ends_fill: rows after a line's end get its end point (the reference
//      convention of the one-kernel path: an ended line keeps its last point),
//      written row-major and coalesced.
// The second phase runs for problems of at least 4,096 charges, where a
// line's step is a long sum; below that the first phase takes every line to
// its end.  Every line takes the same steps with the same arithmetic as in
// the one-kernel path; only the split of each charge sum over lanes differs
// (and not at all below the cap or the 4,096-charge threshold), so results
// agree with it to FP32 summation rounding.
#include <mutex>
#include <vector>

#include "em_internal.hpp"
#include "wide.hpp"

namespace em {

namespace {

constexpr uint32_t kEndsCap = 192;     // steps a line may take in the first phase
constexpr uint32_t kEndsDrainCap = 128;  // ... for the last lines handed out, one per lane group
constexpr int kSurvRec = 3;            // float4 records per set-aside line
constexpr int kTailLanes = 256;        // lanes of a second-phase workgroup
constexpr int kTailWaves = kTailLanes / 64;
constexpr int kSlots = 8;              // lines a second-phase workgroup traces at once
// Records per lane in each of the second phase's two pipelined register sets.
// With 2 or 3 lines per workgroup each record feeds few interactions, and
// eight sets of the unrolled sum (one per active-line count) share the
// instruction cache: on the 16,384-charge, 65,536-line problem 4 records take
// 20.6 ms, against 21.9 ms with 8, 23.6 ms with 16 and 21.1 ms with 2.
constexpr int kTailU = 4;
constexpr uint32_t kTailMinCharges = kEndsMinCharges;  // 4,096
static_assert(kTailMinCharges >= 2u * kTailLanes * kTailU, "the second phase's bulk needs 2 groups");
constexpr size_t kTailReserve = 64 * 1024;  // LDS that keeps two second-phase workgroups per CU
constexpr uint32_t kFillRows = 64;     // rows per fill thread

struct EndsCtrl {
    uint32_t next_line;  // first phase: next line to hand out
    uint32_t nsurv;      // lines set aside at the cap
    uint32_t next_resume;  // continuation: next set-aside line to hand out
    uint32_t unused;
};

// Zero the counters (a kernel, not a memset: the sequence may be captured
// into a graph) and write the padded charge copy of the second phase.
__global__ void k_ends_prepare(EndsCtrl* __restrict__ ctrl, const float4* __restrict__ charges,
                               uint32_t p, float4* __restrict__ padded, uint32_t pp)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        ctrl->next_line = 0u;
        ctrl->nsurv = 0u;
        ctrl->next_resume = 0u;
        ctrl->unused = 0u;
    }
    if (i < pp) padded[i] = i < p ? charges[i] : make_float4(kFar, kFar, kFar, 0.f);
}

// First phase (RESUME = false): lines 0..n-1 from their seeds; a line still
// moving after `cap` steps -- `drain` steps for the last lines handed out, one
// per lane group of the grid -- is set aside in `surv`, kSurvRec float4 per line:
// {x, y, z, h}, {ux, uy, uz, line index as bits}, {step as bits, 0, 0, 0}.
// Continuation (RESUME = true): when more lines were set aside than the second
// phase holds at once (more than `tail_max`), this kernel takes them instead,
// each from the step after the one it was set aside at, to the end, with the
// same refill; otherwise it exits at once (the second phase has them).
template <int S, int U, bool RESUME>
__global__ void __launch_bounds__(512) k_trace_ends(Cfg cfg, const float4* __restrict__ charges,
                                                    float* __restrict__ X, float* __restrict__ Y,
                                                    float* __restrict__ Z, StepP sp,
                                                    uint32_t tile_cap, uint32_t cap, uint32_t drain,
                                                    uint32_t tail_max, uint32_t* __restrict__ steps,
                                                    EndsCtrl* __restrict__ ctrl,
                                                    float4* __restrict__ surv)
{
    extern __shared__ float4 tile[];
    const uint32_t nsurv = RESUME ? __builtin_amdgcn_readfirstlane(ctrl->nsurv) : 0u;
    if (RESUME && nsurv <= tail_max) return;  // uniform: before any barrier
    const int k = threadIdx.x % S;
    const int leader = (int)(threadIdx.x % 64) - k;  // wave lane of the group's lane 0
    const uint32_t groups = gridDim.x * (blockDim.x / S);
    const uint32_t late = cfg.n > groups ? cfg.n - groups : 0u;
    const uint32_t p = cfg.p;
    const uint32_t pp = round_up(p, S * kGroup);
    const bool single = pp <= tile_cap;
    if (single) {
        stage(tile, charges, p, pp);
        __syncthreads();
    }
    LineState st;
    st.px = st.py = st.pz = 0.f;
    st.ux = st.uy = st.uz = 0.f;
    st.Px = st.Py = st.Pz = 0.f;
    st.h = sp.h0;
    st.alive = false;
    st.steps = 0;
    uint32_t line = 0, s = 0;
    bool active = false, drained = false;
    for (;;) {
        if (!active && !drained) {
            // the whole group is here (the branch is uniform over it)
            uint32_t got = 0;
            if (k == 0) got = atomicAdd(RESUME ? &ctrl->next_resume : &ctrl->next_line, 1u);
            got = __shfl(got, leader, 64);
            if (RESUME && got < nsurv) {
                const float4 a = surv[kSurvRec * (uint64_t)got];
                const float4 b = surv[kSurvRec * (uint64_t)got + 1];
                const float4 c = surv[kSurvRec * (uint64_t)got + 2];
                line = __float_as_uint(b.w);
                st.px = a.x; st.py = a.y; st.pz = a.z; st.h = a.w;
                st.ux = b.x; st.uy = b.y; st.uz = b.z;
                st.Px = st.Py = st.Pz = 0.f;
                st.alive = true;
                st.steps = __float_as_uint(c.x);
                s = st.steps + 1;
                active = true;
            } else if (!RESUME && got < cfg.n) {
                line = got;
                const uint64_t b = cfg.line_off + line;
                st.px = X[b];
                st.py = Y[b];
                st.pz = Z[b];
                st.Px = st.px; st.Py = st.py; st.Pz = st.pz;
                st.ux = st.uy = st.uz = 0.f;
                st.h = sp.h0;
                st.alive = true;
                st.steps = 0;
                s = 1;
                active = true;
            } else {
                drained = true;
            }
        }
        // a wave (the block, when the charges are re-staged with barriers)
        // leaves once none of its groups holds a line
        const int any = single ? __any(active) : __syncthreads_or(active);
        if (__builtin_amdgcn_readfirstlane(any) == 0) break;  // uniform: a scalar branch
        float ex, ey, ez;
        field_at<S, U>(tile, tile_cap, charges, p, k, st.px, st.py, st.pz, ex, ey, ez,
                       __any(active) != 0);
        if (active) {
            step_update(sp, s, ex * sp.scale, ey * sp.scale, ez * sp.scale, st);
            if (st.alive) {
                st.steps = s;
                if (k == 0) {
                    const uint64_t o = cfg.line_off + (uint64_t)s * cfg.pitch + line;
                    X[o] = st.px;
                    Y[o] = st.py;
                    Z[o] = st.pz;
                }
            }
            // the last `groups` lines handed out are the ones the phase
            // drains on: they are set aside at `drain` steps, the others at
            // `cap` (a rule on the line index, so the rows do not depend on
            // which group took a line when)
            const bool aside = !RESUME && st.alive && s + 1 < cfg.len && s == (line >= late ? drain : cap);
            if (!st.alive || s + 1 >= cfg.len) {
                if (k == 0) steps[cfg.line_index + line] = st.steps;
                active = false;
            } else if (aside) {
                if (k == 0) {
                    const uint64_t i = atomicAdd(&ctrl->nsurv, 1u);
                    surv[kSurvRec * i] = make_float4(st.px, st.py, st.pz, st.h);
                    surv[kSurvRec * i + 1] = make_float4(st.ux, st.uy, st.uz, __uint_as_float(line));
                    surv[kSurvRec * i + 2] = make_float4(__uint_as_float(s), 0.f, 0.f, 0.f);
                }
                active = false;
            } else {
                ++s;
            }
        }
    }
}

// One float4 record through a buffer load: a single lane-offset VGPR serves
// every load of the walk (the record offsets go in scalar registers); a load
// beyond the buffer returns zeros instead of faulting.
__device__ __forceinline__ float4 buffer_record(__amdgpu_buffer_rsrc_t rsrc, uint32_t lane_off,
                                                uint32_t rec)
{
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(rsrc, lane_off, rec * 16u, 0);
    return make_float4(__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]),
                       __uint_as_float(v[3]));
}

// Coulomb sums at the NACT points of a workgroup's active slots over pp padded
// records (a multiple of kTailLanes, at least kTailMinCharges) split over the
// kTailLanes lanes: every record a lane loads serves all NACT points.  The
// bulk runs the software pipeline of pipelined_tile (two register sets of
// kTailU records, one set's loads in flight while the other is consumed),
// reading global memory; the last pp mod (2 kTailU kTailLanes) records
// follow one per lane at a time.  Each lane
// sums pp / kTailLanes records per point in plain FP32 (the blocked
// compensation of the other kernels folds every 128 records, so it would
// change nothing here); the lanes' sums are then combined by a butterfly and
// across the waves in LDS.  Lane 0 of every wave writes its wave's sums to
// part[wave][l].
template <int NACT>
__device__ __forceinline__ void tail_sums(__amdgpu_buffer_rsrc_t rsrc, uint32_t pp, int tid,
                                          const float4* __restrict__ pts, float4* __restrict__ part)
{
    constexpr uint32_t kStride = kTailLanes * kTailU;  // records per group
    const uint32_t ngroups = (pp / kStride) & ~1u;  // even, >= 2 since pp >= kTailMinCharges
    const uint32_t lane_off = (uint32_t)tid * 16u;
    float px[NACT], py[NACT], pz[NACT], ex[NACT], ey[NACT], ez[NACT];
#pragma unroll
    for (int l = 0; l < NACT; ++l) {
        const float4 q = pts[l];
        px[l] = q.x; py[l] = q.y; pz[l] = q.z;
        ex[l] = ey[l] = ez[l] = 0.f;
    }
    float4 a[kTailU], b[kTailU];
#pragma unroll
    for (int u = 0; u < kTailU; ++u) a[u] = buffer_record(rsrc, lane_off, u * kTailLanes);
    for (uint32_t g = 0; g < ngroups; g += 2) {
        const uint32_t rb = (g + 1) * kStride;
#pragma unroll
        for (int u = 0; u < kTailU; ++u) b[u] = buffer_record(rsrc, lane_off, rb + u * kTailLanes);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < kTailU; ++u)
#pragma unroll
            for (int l = 0; l < NACT; ++l) coulomb(a[u], px[l], py[l], pz[l], ex[l], ey[l], ez[l]);
        __builtin_amdgcn_sched_barrier(0);
        const uint32_t ra = min(g + 2, ngroups - 2) * kStride;  // the last pair re-reads itself
#pragma unroll
        for (int u = 0; u < kTailU; ++u) a[u] = buffer_record(rsrc, lane_off, ra + u * kTailLanes);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < kTailU; ++u)
#pragma unroll
            for (int l = 0; l < NACT; ++l) coulomb(b[u], px[l], py[l], pz[l], ex[l], ey[l], ez[l]);
        __builtin_amdgcn_sched_barrier(0);
    }
    for (uint32_t r = ngroups * kStride; r < pp; r += kTailLanes) {
        const float4 c = buffer_record(rsrc, lane_off, r);
#pragma unroll
        for (int l = 0; l < NACT; ++l) coulomb(c, px[l], py[l], pz[l], ex[l], ey[l], ez[l]);
    }
    const int wave = tid / 64;
#pragma unroll
    for (int l = 0; l < NACT; ++l) {
        const float sx = group_sum<64>(ex[l]);
        const float sy = group_sum<64>(ey[l]);
        const float sz = group_sum<64>(ez[l]);
        if ((tid & 63) == 0) part[wave * kSlots + l] = make_float4(sx, sy, sz, 0.f);
    }
}

// State of one slot of the second phase, held by lane `slot` of wave 0.
struct Slot {
    LineState st;
    uint32_t line;
    uint32_t s;  // the step this slot's line takes next
};

// Wave 0 only.  Slots [0, nact) hold lines; `done` marks those whose line
// has just finished.  Finished and empty slots take the next lines of this
// workgroup's list (set-aside lines blockIdx.x + k * gridDim.x, k = next_k,
// next_k + 1, ...); the slots still holding lines are then packed to the
// front through LDS (`xfer`), their points published in `pts` and their count
// returned.
__device__ __forceinline__ uint32_t refill_slots(Slot& sl, bool done, uint32_t nact,
                                                 uint32_t& next_k, uint32_t nsurv,
                                                 const float4* __restrict__ surv,
                                                 float4* __restrict__ xfer, float4* __restrict__ pts)
{
    const uint32_t lane = threadIdx.x;  // wave 0: 0..63
    const bool holds = lane < nact && !done;
    const bool need = lane < (uint32_t)kSlots && !holds;
    const uint64_t need_mask = __ballot(need);
    const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(need_mask >> 32),
                                                    __builtin_amdgcn_mbcnt_lo((uint32_t)need_mask, 0u));
    const uint32_t j = blockIdx.x + (next_k + rank) * gridDim.x;
    const bool got = need && j < nsurv;
    // the lines handed out are the list's next ones in rank order, so the
    // list position advances by their count (and stops at the list's end)
    next_k += (uint32_t)__popcll(__ballot(got));
    if (got) {
        const float4 a = surv[kSurvRec * (uint64_t)j];
        const float4 b = surv[kSurvRec * (uint64_t)j + 1];
        const float4 c = surv[kSurvRec * (uint64_t)j + 2];
        sl.st.px = a.x; sl.st.py = a.y; sl.st.pz = a.z; sl.st.h = a.w;
        sl.st.ux = b.x; sl.st.uy = b.y; sl.st.uz = b.z;
        sl.st.Px = sl.st.Py = sl.st.Pz = 0.f;
        sl.st.alive = true;
        sl.st.steps = __float_as_uint(c.x);  // the step it was set aside at
        sl.line = __float_as_uint(b.w);
        sl.s = sl.st.steps + 1;
    }
    const bool keep = holds || got;
    const uint64_t keep_mask = __ballot(keep);
    const uint32_t pos = __builtin_amdgcn_mbcnt_hi((uint32_t)(keep_mask >> 32),
                                                   __builtin_amdgcn_mbcnt_lo((uint32_t)keep_mask, 0u));
    const uint32_t nkeep = (uint32_t)__popcll(keep_mask);
    if (keep) {
        xfer[3 * pos] = make_float4(sl.st.px, sl.st.py, sl.st.pz, sl.st.h);
        xfer[3 * pos + 1] = make_float4(sl.st.ux, sl.st.uy, sl.st.uz, __uint_as_float(sl.line));
        xfer[3 * pos + 2] = make_float4(__uint_as_float(sl.s), __uint_as_float(sl.st.steps), 0.f, 0.f);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (lane < nkeep) {
        const float4 a = xfer[3 * lane];
        const float4 b = xfer[3 * lane + 1];
        const float4 c = xfer[3 * lane + 2];
        sl.st.px = a.x; sl.st.py = a.y; sl.st.pz = a.z; sl.st.h = a.w;
        sl.st.ux = b.x; sl.st.uy = b.y; sl.st.uz = b.z;
        sl.st.alive = true;
        sl.line = __float_as_uint(b.w);
        sl.s = __float_as_uint(c.x);
        sl.st.steps = __float_as_uint(c.y);
        pts[lane] = make_float4(sl.st.px, sl.st.py, sl.st.pz, 0.f);
    }
    return nkeep;
}

// Second phase: the lines set aside at step `cap`.  Two workgroups per CU
// (the launch reserves enough LDS that no third one fits): one wave per SIMD
// each, so while one takes its steps (a serial stretch in wave 0 between two
// barriers) the other sums.  Workgroup b owns set-aside lines b, b + grid,
// b + 2 grid, ...; with workgroups b and b + grid/2 on one CU (the dispatch
// order of a 1-D grid over the CUs), a CU holds every line congruent to b
// modulo grid/2.  A workgroup traces up to kSlots of its lines at once, so
// each charge record it loads from L2 serves every one of them.
// The assignment is static on purpose: a loop whose head is a branch on the
// thread index (one thread fetching the next item from a counter) and whose
// body holds barriers can be threaded by the compiler into a loop that the
// other threads re-enter without that thread, which then never publishes the
// next item.  Loop exits on LDS values go through readfirstlane, so they are
// scalar branches.
// `padded` holds pp records (a multiple of kTailLanes, >= kTailMinCharges),
// charges then zero-charge padding; lane t sums records t, t + 256, t + 512, ...
__global__ void __launch_bounds__(kTailLanes) k_trace_tail(Cfg cfg, const float4* __restrict__ padded,
                                                           uint32_t pp, float* __restrict__ X,
                                                           float* __restrict__ Y, float* __restrict__ Z,
                                                           StepP sp, uint32_t tail_max,
                                                           uint32_t* __restrict__ steps,
                                                           const EndsCtrl* __restrict__ ctrl,
                                                           const float4* __restrict__ surv)
{
    __shared__ float4 part[kTailWaves * kSlots];
    __shared__ float4 pts[kSlots];
    __shared__ float4 xfer[3 * kSlots];
    __shared__ uint32_t s_nact;
    const int tid = threadIdx.x;
    const int wave = tid / 64;
    // final: the first phase has completed; beyond tail_max lines the
    // continuation of the first phase takes them instead
    const uint32_t nsurv = __builtin_amdgcn_readfirstlane(ctrl->nsurv);
    if (nsurv > tail_max) return;  // uniform: before any barrier
    const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<float4*>(padded), (short)0, (int)(pp * sizeof(float4)), 0x00020000);
    Slot sl;  // meaningful in lanes 0..kSlots-1 of wave 0
    sl.st.px = sl.st.py = sl.st.pz = 0.f;
    sl.st.ux = sl.st.uy = sl.st.uz = 0.f;
    sl.st.Px = sl.st.Py = sl.st.Pz = 0.f;
    sl.st.h = sp.h0;
    sl.st.alive = false;
    sl.st.steps = 0;
    sl.line = 0;
    sl.s = 0;
    uint32_t next_k = 0, nact = 0;
    if (wave == 0) {
        nact = refill_slots(sl, false, 0u, next_k, nsurv, surv, xfer, pts);
        if (tid == 0) s_nact = nact;
    }
    for (;;) {
        __syncthreads();  // pts and s_nact published
        const uint32_t n_now = __builtin_amdgcn_readfirstlane(s_nact);
        if (n_now == 0u) break;
        // A CU's two workgroups hold unequal line counts (3 and 2 where it has
        // 5 lines): the waves of the one with more lines issue first, and the
        // other fills their stalls, so both end closer together (19.7 ms
        // against 20.6 ms at equal priority on the problem above).
        if (n_now >= 3u)
            __builtin_amdgcn_s_setprio(3);
        else if (n_now == 2u)
            __builtin_amdgcn_s_setprio(2);
        else
            __builtin_amdgcn_s_setprio(1);
        switch (n_now) {
            case 1: tail_sums<1>(rsrc, pp, tid, pts, part); break;
            case 2: tail_sums<2>(rsrc, pp, tid, pts, part); break;
            case 3: tail_sums<3>(rsrc, pp, tid, pts, part); break;
            case 4: tail_sums<4>(rsrc, pp, tid, pts, part); break;
            case 5: tail_sums<5>(rsrc, pp, tid, pts, part); break;
            case 6: tail_sums<6>(rsrc, pp, tid, pts, part); break;
            case 7: tail_sums<7>(rsrc, pp, tid, pts, part); break;
            default: tail_sums<8>(rsrc, pp, tid, pts, part); break;
        }
        __syncthreads();  // part complete
        if (wave == 0) {
            bool done = false;
            if ((uint32_t)tid < n_now) {
                float4 e = part[tid];
#pragma unroll
                for (int w = 1; w < kTailWaves; ++w) {
                    const float4 q = part[w * kSlots + tid];
                    e.x += q.x;
                    e.y += q.y;
                    e.z += q.z;
                }
                step_update(sp, sl.s, e.x * sp.scale, e.y * sp.scale, e.z * sp.scale, sl.st);
                if (sl.st.alive) {
                    sl.st.steps = sl.s;
                    const uint64_t o = cfg.line_off + (uint64_t)sl.s * cfg.pitch + sl.line;
                    X[o] = sl.st.px;
                    Y[o] = sl.st.py;
                    Z[o] = sl.st.pz;
                }
                done = !sl.st.alive || sl.s + 1 >= cfg.len;
                if (done) steps[cfg.line_index + sl.line] = sl.st.steps;
                else ++sl.s;
            }
            // slots change only when a line finished or an empty slot can
            // take a line from the list; otherwise just publish the points
            const bool change = __ballot(done) != 0ull ||
                                (n_now < (uint32_t)kSlots && blockIdx.x + next_k * gridDim.x < nsurv);
            if (change) {
                nact = refill_slots(sl, done, n_now, next_k, nsurv, surv, xfer, pts);
            } else {
                if ((uint32_t)tid < n_now) pts[tid] = make_float4(sl.st.px, sl.st.py, sl.st.pz, 0.f);
                nact = n_now;
            }
            if (tid == 0) s_nact = nact;
        }
    }
}

// Rows steps+1 .. len-1 of every line get row `steps` (the end point).
__global__ void k_ends_fill(Cfg cfg, float* __restrict__ X, float* __restrict__ Y,
                            float* __restrict__ Z, const uint32_t* __restrict__ steps)
{
    const uint32_t line = blockIdx.x * blockDim.x + threadIdx.x;
    if (line >= cfg.n) return;
    const uint32_t e = steps[cfg.line_index + line];
    const uint32_t r0 = max(e + 1u, blockIdx.y * kFillRows);
    const uint32_t r1 = min(cfg.len, (blockIdx.y + 1u) * kFillRows);
    if (r0 >= r1) return;
    const uint64_t b = cfg.line_off + line;
    const uint64_t src = b + (uint64_t)e * cfg.pitch;
    const float x = X[src], y = Y[src], z = Z[src];
    for (uint32_t r = r0; r < r1; ++r) {
        const uint64_t o = b + (uint64_t)r * cfg.pitch;
        X[o] = x;
        Y[o] = y;
        Z[o] = z;
    }
}

template <int S, int U, bool RESUME>
int launch_phase1(const Shape& sh, const DeviceInfo& dev, const Cfg& cfg, const float4* charges,
                  float* x, float* y, float* z, const StepP& sp, uint32_t cap, uint32_t drain,
                  uint32_t tail_max, uint32_t* steps, EndsCtrl* ctrl, float4* surv,
                  hipStream_t stream)
{
    const size_t lds = (size_t)sh.tile * sizeof(float4);
    EM_TRY(allow_lds(k_trace_ends<S, U, RESUME>, lds));
    int per_cu = 0;
    EM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_trace_ends<S, U, RESUME>,
                                                         sh.block, lds));
    if (per_cu < 1)
        EM_FAIL(EM_ERR_INVALID, "launch shape does not fit a CU (block %d, tile %u)", sh.block, sh.tile);
    const uint64_t groups_per_block = (uint64_t)sh.block / S;
    const uint64_t want = (cfg.n + groups_per_block - 1) / groups_per_block;
    const uint32_t grid = (uint32_t)std::min<uint64_t>((uint64_t)per_cu * dev.cus, want);
    hipLaunchKernelGGL((k_trace_ends<S, U, RESUME>), dim3(grid), dim3(sh.block), lds, stream, cfg,
                       charges, x, y, z, sp, sh.tile, cap, drain, tail_max, steps, ctrl, surv);
    EM_HIP(hipGetLastError());
    return EM_OK;
}

template <bool RESUME>
int dispatch_phase1(const Shape& sh, const DeviceInfo& dev, const Cfg& cfg, const float4* charges,
                    float* x, float* y, float* z, const StepP& sp, uint32_t cap, uint32_t drain,
                    uint32_t tail_max, uint32_t* steps, EndsCtrl* ctrl, float4* surv,
                    hipStream_t stream)
{
#define EM_PHASE1_U(SS, UU)                                                                   \
    case UU:                                                                                  \
        return launch_phase1<SS, UU, RESUME>(sh, dev, cfg, charges, x, y, z, sp, cap, drain,   \
                                             tail_max, steps, ctrl, surv, stream)
#define EM_PHASE1(SS)                                                                         \
    case SS:                                                                                  \
        switch (sh.unroll) {                                                                  \
            EM_PHASE1_U(SS, 2);                                                               \
            EM_PHASE1_U(SS, 4);                                                               \
            EM_PHASE1_U(SS, 8);                                                               \
            default: EM_FAIL(EM_ERR_INVALID, "unroll %d not in {2, 4, 8}", sh.unroll);        \
        }
    switch (sh.S) {
        EM_PHASE1(1);
        EM_PHASE1(2);
        EM_PHASE1(4);
        EM_PHASE1(8);
        EM_PHASE1(16);
        EM_PHASE1(32);
        default: EM_FAIL(EM_ERR_INVALID, "lanes_per_line %d not in {1,2,4,8,16,32}", sh.S);
    }
#undef EM_PHASE1
#undef EM_PHASE1_U
}

// Per-device work memory of the line-ends trace: counters, step counts (when
// the caller keeps none), and the register-resident trace's two lists and
// their orders, or the lane-group trace's set-aside list and padded charge
// copy.
// Eager calls on one device use it in turn: each call's work waits for the
// previous eager call's (an event).  A captured graph uses the buffer current
// at capture; its replays are ordered only by the caller's streams, so a
// replay that overlaps another call on the same device races with it.
// A buffer that is outgrown is retired, not freed: a graph captured earlier
// still holds its address.  Retired buffers are freed by em_shutdown.
struct EndsScratch {
    std::mutex mu;
    int device = -1;
    char* buf = nullptr;
    size_t bytes = 0;
    hipEvent_t done = nullptr;
    bool recorded = false;
    std::vector<char*> retired;
};
EndsScratch g_ends[64];

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// Where the pieces of the work memory lie for one problem, and how much it
// needs.  Problems of 4,096..16,384 charges take the register-resident trace
// of wide.hip (`wide`); larger ones the two phases here.
struct EndsLayout {
    bool wide;        // the wide trace (wide.hip)
    WideLayout w;
    bool tail;        // the second phase runs
    uint32_t cap, drain, pp;
    size_t off_steps, off_surv, off_pad, need;
};

EndsLayout ends_layout(const DeviceInfo& dev, const Cfg& cfg, bool own_steps)
{
    EndsLayout L;
    L.wide = wide_applies(cfg);
    if (L.wide) {
        L.tail = false;
        L.cap = L.drain = cfg.len;
        L.pp = 0;
        L.off_steps = align256(sizeof(EndsCtrl) > sizeof(WideCtrl) ? sizeof(EndsCtrl) : sizeof(WideCtrl));
        const size_t steps_bytes = own_steps ? align256((size_t)cfg.n * sizeof(uint32_t)) : 0;
        L.w = wide_layout(dev, L.off_steps + steps_bytes);
        L.off_surv = L.off_pad = L.w.end;
        L.need = L.w.end;
        return L;
    }
    // The second phase pays off when a line's step is a long sum.  Below
    // kTailMinCharges charges the first phase's wide groups take every line
    // to its end (a step is then at most 128 records per lane), and no
    // padded copy is made.
    L.tail = cfg.p >= kTailMinCharges && cfg.len > kEndsCap + 1;
    L.cap = L.tail ? kEndsCap : cfg.len;
    L.drain = L.tail ? kEndsDrainCap : cfg.len;
    L.pp = L.tail ? round_up(cfg.p, kTailLanes) : 0u;
    L.off_steps = align256(sizeof(EndsCtrl));
    const size_t steps_bytes = own_steps ? align256((size_t)cfg.n * sizeof(uint32_t)) : 0;
    L.off_surv = L.off_steps + steps_bytes;
    L.off_pad = L.off_surv + align256((size_t)cfg.n * kSurvRec * sizeof(float4));
    L.need = L.off_pad + (size_t)L.pp * sizeof(float4);
    return L;
}

int ends_check(const DeviceInfo& dev, const Cfg& cfg)
{
    if (dev.device < 0 || dev.device >= 64) EM_FAIL(EM_ERR_INVALID, "device %d", dev.device);
    // the line counter passes n by up to one fetch per lane group, and the
    // second phase's list positions by up to kSlots per workgroup: 32 bits hold both
    if (cfg.n > 0xFFF00000u) EM_FAIL(EM_ERR_INVALID, "line ends: n %u above 2^32 - 2^20", cfg.n);
    // the second phase addresses the padded charges through a buffer
    // descriptor whose size field counts bytes in 32 bits
    if (cfg.p > (1u << 28) - kTailLanes) EM_FAIL(EM_ERR_INVALID, "line ends: p %u above 2^28 - 256", cfg.p);
    if ((cfg.len + kFillRows - 1) / kFillRows > 65535u)
        EM_FAIL(EM_ERR_INVALID, "len %u exceeds the fill grid", cfg.len);
    return EM_OK;
}

// The launches of one line-ends trace over work memory `buf` (laid out by
// `L`).  Allocates nothing; the counters are zeroed by a kernel, so the
// sequence may be captured into a graph.
int ends_launch(const Shape& sh, const DeviceInfo& dev, const Cfg& cfg, const float4* charges,
                float* x, float* y, float* z, const StepP& sp, uint32_t* steps, hipStream_t stream,
                char* buf, const EndsLayout& L)
{
    EndsCtrl* ctrl = reinterpret_cast<EndsCtrl*>(buf);
    uint32_t* st = steps ? steps : reinterpret_cast<uint32_t*>(buf + L.off_steps);
    float4* surv = reinterpret_cast<float4*>(buf + L.off_surv);
    float4* padded = reinterpret_cast<float4*>(buf + L.off_pad);
    const dim3 fill_grid((cfg.n + 255) / 256, (cfg.len + kFillRows - 1) / kFillRows);
    if (L.wide) return wide_launch(dev, cfg, charges, x, y, z, sp, st, stream, buf, L.w);

    hipLaunchKernelGGL(k_ends_prepare, dim3(std::max<uint32_t>(1u, (L.pp + 255) / 256)), dim3(256), 0,
                       stream, ctrl, charges, cfg.p, padded, L.pp);
    EM_HIP(hipGetLastError());
    // the second phase holds 8 lines in each of its 2-per-CU workgroups at once
    const uint32_t tail_max = 2u * (uint32_t)dev.cus * (uint32_t)kSlots;
    EM_TRY(dispatch_phase1<false>(sh, dev, cfg, charges, x, y, z, sp, L.cap, L.drain, tail_max, st,
                                  ctrl, surv, stream));
    if (L.tail) {
        // two workgroups per CU: each loads every charge record once per step
        // for all the lines it holds
        EM_TRY(allow_lds(k_trace_tail, kTailReserve));
        int per_cu = 0;
        EM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_trace_tail, kTailLanes,
                                                             kTailReserve));
        if (per_cu != 2) EM_FAIL(EM_ERR_INVALID, "second-phase workgroups per CU: %d, not 2", per_cu);
        hipLaunchKernelGGL(k_trace_tail, dim3(2 * dev.cus), dim3(kTailLanes), kTailReserve, stream,
                           cfg, padded, L.pp, x, y, z, sp, tail_max, st, ctrl, surv);
        EM_HIP(hipGetLastError());
        // more set-aside lines than that: the first phase's groups continue them
        EM_TRY(dispatch_phase1<true>(sh, dev, cfg, charges, x, y, z, sp, cfg.len, cfg.len,
                                     tail_max, st, ctrl, surv, stream));
    }
    hipLaunchKernelGGL(k_ends_fill, fill_grid, dim3(256), 0, stream, cfg, x, y, z, st);
    EM_HIP(hipGetLastError());
    return EM_OK;
}

// Work memory of em_trace_dev_f32_noalloc: one fixed array in the device code,
// so a call allocates nothing.  Host code never passes the array's name to a
// kernel (on the host it is only a stand-in for the symbol): each device's
// address comes from hipGetSymbolAddress, once per device, and is kept in
// g_scratch_ptr.  Every call on a device shares its copy; it keeps the last
// call's contents (each call zeroes its counters with a kernel first).
static __device__ __align__(256) char scratch[kNoAllocScratchBytes];

std::mutex g_scratch_mu;
char* g_scratch_ptr[64] = {};

}  // namespace

int noalloc_scratch(int device, char** out)
{
    if (device < 0 || device >= 64) EM_FAIL(EM_ERR_INVALID, "device %d", device);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        int cur = -1;
        EM_HIP(hipGetDevice(&cur));
        if (cur != device) EM_FAIL(EM_ERR_INVALID, "scratch of device %d asked with device %d current", device, cur);
        void* ptr = nullptr;
        EM_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) EM_FAIL(EM_ERR_HIP, "hipGetSymbolAddress(scratch) gave NULL on device %d", device);
        g_scratch_ptr[device] = static_cast<char*>(ptr);
    }
    if (out) *out = g_scratch_ptr[device];
    return EM_OK;
}

int launch_trace_ends_noalloc(const Shape& sh, const DeviceInfo& dev, const Cfg& cfg,
                              const float4* charges, float* x, float* y, float* z, const StepP& sp,
                              uint32_t* steps, hipStream_t stream)
{
    EM_TRY(ends_check(dev, cfg));
    const EndsLayout L = ends_layout(dev, cfg, steps == nullptr);
    if (L.need > kNoAllocScratchBytes)
        EM_FAIL(EM_ERR_INVALID,
                "line ends without allocation: %u lines and %u charges need %zu bytes of work "
                "memory, above the fixed %zu", cfg.n, cfg.p, L.need, kNoAllocScratchBytes);
    char* buf = nullptr;
    EM_TRY(noalloc_scratch(dev.device, &buf));
    return ends_launch(sh, dev, cfg, charges, x, y, z, sp, steps, stream, buf, L);
}

int launch_trace_ends(const Shape& sh, const DeviceInfo& dev, const Cfg& cfg, const float4* charges,
                      float* x, float* y, float* z, const StepP& sp, uint32_t* steps,
                      hipStream_t stream)
{
    EM_TRY(ends_check(dev, cfg));
    const EndsLayout L = ends_layout(dev, cfg, steps == nullptr);
    const size_t need = L.need;

    hipStreamCaptureStatus capture = hipStreamCaptureStatusNone;
    EM_HIP(hipStreamIsCapturing(stream, &capture));
    const bool capturing = capture != hipStreamCaptureStatusNone;

    EndsScratch& w = g_ends[dev.device];
    std::lock_guard<std::mutex> lk(w.mu);
    w.device = dev.device;
    if (!w.done) EM_HIP(hipEventCreateWithFlags(&w.done, hipEventDisableTiming));
    if (w.bytes < need) {
        if (capturing)
            EM_FAIL(EM_ERR_INVALID,
                    "line-ends work memory must grow to %zu bytes: make a first call of this "
                    "size outside stream capture", need);
        // at least doubling: the retired buffers total less than the live one
        const size_t size = std::max(need, 2 * w.bytes);
        char* grown = nullptr;
        EM_HIP(hipMalloc(reinterpret_cast<void**>(&grown), size));
        if (w.buf) w.retired.push_back(w.buf);
        w.buf = grown;
        w.bytes = size;
    }
    if (w.recorded && !capturing) EM_HIP(hipStreamWaitEvent(stream, w.done, 0));
    EM_TRY(ends_launch(sh, dev, cfg, charges, x, y, z, sp, steps, stream, w.buf, L));
    if (!capturing) {
        EM_HIP(hipEventRecord(w.done, stream));
        w.recorded = true;
    }
    return EM_OK;
}

int ends_release_all()
{
    int rc = EM_OK;
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) return EM_ERR_HIP;
    for (auto& w : g_ends) {
        std::lock_guard<std::mutex> lk(w.mu);
        if (!w.buf && !w.done) continue;
        if (hipSetDevice(w.device) != hipSuccess) rc = EM_ERR_HIP;
        if (w.recorded && hipEventSynchronize(w.done) != hipSuccess) rc = EM_ERR_HIP;
        if (w.buf && hipFree(w.buf) != hipSuccess) rc = EM_ERR_HIP;
        for (char* r : w.retired)
            if (hipFree(r) != hipSuccess) rc = EM_ERR_HIP;
        w.retired.clear();
        if (w.done && hipEventDestroy(w.done) != hipSuccess) rc = EM_ERR_HIP;
        w.buf = nullptr;
        w.bytes = 0;
        w.done = nullptr;
        w.recorded = false;
    }
    if (hipSetDevice(cur) != hipSuccess) rc = EM_ERR_HIP;
    return rc;
}

}  // namespace em
