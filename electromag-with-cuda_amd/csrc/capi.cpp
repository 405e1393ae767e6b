// capi.cpp -- the extern "C" boundary of libemtrace.so (declared in
// include/emtrace.h): argument checks with the reference's status codes,
// launch-shape choice, the host-array drop-in with timed transfers, and
// batch plans (sorted work items + persistent grid + optional hipGraph).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "em_internal.hpp"
#include "emtrace/legacy.hpp"

namespace em {

// ---------------------------------------------------------------------------
// error state
static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

void clear_error() { g_err[0] = 0; }

// ---------------------------------------------------------------------------
// device info
static std::mutex g_mu;
static DeviceInfo g_dev[64];

int current_device_info(DeviceInfo& out)
{
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
        EM_FAIL(EM_ERR_NO_DEVICE, "no HIP device visible");
    int dev = 0;
    EM_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) EM_FAIL(EM_ERR_NO_DEVICE, "device index %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_mu);
    DeviceInfo& d = g_dev[dev];
    if (d.device != dev) {
        hipDeviceProp_t prop;
        EM_HIP(hipGetDeviceProperties(&prop, dev));
        d.cus = prop.multiProcessorCount;
        d.lds_per_block = prop.sharedMemPerBlock;
        std::snprintf(d.arch, sizeof(d.arch), "%s", prop.gcnArchName);
        if (std::strncmp(d.arch, "gfx950", 6) != 0)
            EM_FAIL(EM_ERR_NO_DEVICE, "device %d is %s; this library is built for gfx950", dev,
                    d.arch);
        d.device = dev;
    }
    out = d;
    return EM_OK;
}

// ---------------------------------------------------------------------------
// launch shape
static bool pow2_in(int v, int lo, int hi) { return v >= lo && v <= hi && (v & (v - 1)) == 0; }

int validate_shape(const DeviceInfo& dev, const Shape& s)
{
    if (!pow2_in(s.S, 1, 32)) EM_FAIL(EM_ERR_INVALID, "lanes_per_line %d not in {1..32}", s.S);
    if (s.unroll != 2 && s.unroll != 4 && s.unroll != 8)
        EM_FAIL(EM_ERR_INVALID, "unroll %d not in {2, 4, 8}", s.unroll);
    if (!pow2_in(s.block, 64, 512)) EM_FAIL(EM_ERR_INVALID, "block %d not in {64..512}", s.block);
    if (s.tile == 0 || s.tile % (uint32_t)(s.S * kGroup) || s.tile % 256)
        EM_FAIL(EM_ERR_INVALID, "tile %u not a positive multiple of 256 and of %d", s.tile, s.S * kGroup);
    if ((uint64_t)s.tile > tile_budget(dev, 1))
        EM_FAIL(EM_ERR_INVALID, "tile %u x 16 B + %zu B of static LDS exceeds the %zu B of LDS per "
                "workgroup", s.tile, kLdsStatic, dev.lds_per_block);
    return EM_OK;
}

static uint32_t tile_for(const DeviceInfo& dev, uint64_t p, int S, int resident)
{
    const uint64_t G = (uint64_t)S * kGroup;
    const uint64_t budget = tile_budget(dev, resident);
    const uint64_t Q = std::max<uint64_t>(256, G);  // tile granule
    uint64_t tmax = std::max<uint64_t>(Q, budget / Q * Q);
    const uint64_t need = (std::max<uint64_t>(p, 1) + Q - 1) / Q * Q;
    if (need <= tmax) return (uint32_t)need;
    const uint64_t tiles = (p + tmax - 1) / tmax;
    const uint64_t t = ((p + tiles - 1) / tiles + Q - 1) / Q * Q;
    return (uint32_t)std::min(t, tmax);
}

// Charge records per pipelined group (the trace kernel takes 100 VGPRs at 8,
// 68 at 4, 52 at 2; the rows do not depend on it).  Eight keep the most LDS
// reads in flight per wave, which pays with few waves per SIMD and with line
// ends.  With four or more resident waves per SIMD other waves hide the LDS
// latency, and fewer records per group issue better: 2 when 512-thread blocks
// re-stage the charges tile by tile, else 4.  Measured on MI355X (ms):
//   16,384 charges x 65,536 lines, 4 lanes, 512 threads (re-staged): U=2 831, 4 847, 8 917
//   the same, 256 threads, 2,304-charge tiles, 4 resident: U=4 best, 2 +4 %, 8 +2 %
//   4,000 x 65,536, 512 threads (one tile): U=4 25.9, 2 27.3, 8 27.2
//   512-configuration batch, 256 threads, 2,048-charge tiles: U=4 432, 2 448, 8 454;
//     4,096-charge tiles (2 waves per SIMD): U=8 467, 4 482, 2 519
//   64-configuration checkpointed forward with line ends: U=8 13.1, 2 14.5
int default_unroll(const DeviceInfo& dev, const Shape& s, uint64_t p, uint64_t n, bool stop)
{
    if (stop) return 8;
    const uint64_t blocks = (n * (uint64_t)s.S + s.block - 1) / s.block;
    const uint64_t by_lds = dev.lds_per_block / ((uint64_t)s.tile * sizeof(float4) + kLdsStatic);
    const uint64_t per_cu = std::min<uint64_t>((blocks + dev.cus - 1) / dev.cus, std::max<uint64_t>(1, by_lds));
    const uint64_t waves_per_simd = per_cu * (uint64_t)s.block / 256;
    if (waves_per_simd < 4) return 8;
    return (s.block >= 512 && p > s.tile) ? 2 : 4;
}

int choose_shape(const DeviceInfo& dev, uint64_t p, uint64_t n, const em_launch* req, Shape& out,
                 bool stop, bool refill)
{
    Shape s;
    if (refill) {
        // first phase of the lane-group line-ends trace: a group takes a new
        // line as soon as its line ends, so blocks stay full and may be
        // large; wide groups keep the slowest line of the phase short
        int S = 1;
        while (S < 32 && p / (2ull * S) >= 64) S *= 2;
        s.S = (req && req->lanes_per_line) ? req->lanes_per_line : S;
        // blocks of up to 512 threads, small enough that the grid reaches every CU
        int block = 512;
        while (block > 64 && n * (uint64_t)s.S / (uint64_t)block < (uint64_t)dev.cus) block /= 2;
        s.block = (req && req->block) ? req->block : block;
        s.tile = (req && req->tile) ? (uint32_t)req->tile : tile_for(dev, p, s.S, 2);
        s.unroll = (req && req->unroll) ? req->unroll : default_unroll(dev, s, p, n, true);
        EM_TRY(validate_shape(dev, s));
        out = s;
        return EM_OK;
    }
    const bool restaged = (uint64_t)p > tile_budget(dev, 1);
    if (req && req->lanes_per_line) {
        s.S = req->lanes_per_line;
    } else if (stop) {
        // With line ends a wave runs until the longest of its 64/S lines has
        // ended, so wide groups pay even at few charges per lane: S doubles
        // while each lane keeps >= 16 charges.  Measured (checkpointed
        // forward, 500 rows, ms): p = 128: S=8 0.74 (S=2 1.37); p = 256 x 2,048
        // lines: S=16 0.75 (S=4 1.37); p = 1,024: S=32 1.03 (S=16 1.45);
        // p = 2,000 x 65,536 lines: S=32 5.14 (S=16 6.60).
        int S = 1;
        while (S < 32 && p / (2ull * S) >= 16) S *= 2;
        s.S = S;
    } else {
        // Split lines over more lanes while the grid is short of waves, as long
        // as each lane keeps enough charges to amortize the per-step reduction
        // and step rule: >= 64 charges per lane up to 4 waves per SIMD, >= 32 up
        // to 2, >= 16 up to 1.  Measured on MI355X (1,000 rows, default block):
        // this picks the fastest S, or one within 4 %, for p x n = 300 x 8,192,
        // 1,000 x 8,192, 1,000 x 2,048, 2,000 x 16,384, 600 x 30,000,
        // 150 x 65,536, 5,000 x 1,024 and 16,384 x 65,536.
        const uint64_t simds = 4ull * (uint64_t)dev.cus;
        auto split = [&](int S) {
            const uint64_t waves = (n * (uint64_t)S + 63) / 64;
            const uint64_t per_lane = p / (2ull * (uint64_t)S);  // after doubling S
            return (waves < 4 * simds && per_lane >= 64) || (waves < 2 * simds && per_lane >= 32) ||
                   (waves < simds && per_lane >= 16);
        };
        int S = 1;
        while (S < 32 && split(S)) S *= 2;
        s.S = S;
    }
    // 512-thread blocks once the grid holds two of them per CU (four waves per
    // SIMD): 12 % faster than 256 threads on 16,384 x 65,536 and 8,000 x
    // 32,768, within 2 % on one-tile problems.  With line ends a whole block
    // keeps evaluating until its last line has ended when the charges are
    // re-staged tile by tile (barriers): one-wave blocks over small tiles then.
    if (req && req->block)
        s.block = req->block;
    else if (stop)
        s.block = restaged ? 64 : 256;
    else
        s.block = n * (uint64_t)s.S >= 2ull * 512ull * (uint64_t)dev.cus ? 512 : 256;
    if (req && req->tile) {
        s.tile = (uint32_t)req->tile;
    } else if (stop && restaged) {
        s.tile = std::max<uint32_t>(512u, (uint32_t)(s.S * kGroup));  // many one-wave blocks per CU
    } else {
        const uint64_t blocks = (n * (uint64_t)s.S + s.block - 1) / s.block;
        const int resident = (int)std::min<uint64_t>(std::max<uint64_t>(1, (blocks + dev.cus - 1) / dev.cus),
                                                      std::max(1, 1024 / s.block));
        s.tile = tile_for(dev, p, s.S, resident);
    }
    s.unroll = (req && req->unroll) ? req->unroll : default_unroll(dev, s, p, n, stop);
    EM_TRY(validate_shape(dev, s));
    out = s;
    return EM_OK;
}

int validate_step(const em_step* st)
{
    if (!st) EM_FAIL(EM_ERR_INVALID, "step is NULL");
    if (st->rule == kRulePlain || st->rule == kRuleRef) {
        if (st->resolution == 0.f) EM_FAIL(EM_ERR_RESOLUTION, "resolution == 0");
        if (!std::isfinite(st->resolution)) EM_FAIL(EM_ERR_INVALID, "resolution not finite");
    } else if (st->rule == kRuleAngle) {
        if (!(st->h0 > 0.f && st->hmin > 0.f && st->hmax >= st->hmin && st->turn > 0.f) ||
            !std::isfinite(st->hmax) || !std::isfinite(st->h0))
            EM_FAIL(EM_ERR_INVALID, "angle rule needs h0 > 0, 0 < hmin <= hmax, turn > 0");
    } else {
        EM_FAIL(EM_ERR_INVALID, "unknown step rule %d", st->rule);
    }
    if (st->stop && st->rule != kRuleAngle)
        EM_FAIL(EM_ERR_INVALID, "stop is defined for the angle rule only");
    return EM_OK;
}

StepP to_stepp(const em_step& st)
{
    StepP sp;
    sp.rule = st.rule;
    sp.res = st.resolution;
    sp.h0 = st.h0;
    sp.hmin = st.hmin;
    sp.hmax = st.hmax;
    sp.turn = st.turn;
    sp.scale = (st.rule == kRuleAngle) ? 1.0f : (float)8.987551787E9;  // Electrostatics.h:27
    sp.stop = st.stop ? 1 : 0;
    return sp;
}

// common size checks, reference order: n, resolution, len
static int check_sizes(size_t p, size_t n, size_t len, const em_step* st)
{
    if (!n) EM_FAIL(EM_ERR_NO_LINES, "n == 0");
    EM_TRY(validate_step(st));
    if (len < 2) EM_FAIL(EM_ERR_LENGTH, "len %zu < 2", len);
    if (p > 0xFFFFFFFFull || n > 0xFFFFFFFFull || len > 0xFFFFFFFFull)
        EM_FAIL(EM_ERR_INVALID, "p, n and len must each fit 32 bits");
    return EM_OK;
}

// ---------------------------------------------------------------------------
// device buffer cache for the host-array forms (one set per device)
constexpr size_t kChunkBytes = 48ull << 20;  // output bytes per row chunk

static std::mutex g_hp_mu;  // held by a host-array call from set-up to the last copy
static HostPathBuffers g_hp[64];

// Only the last chunk's copy is exposed: on the reference's default problem
// (491 MB of rows) eight chunks take 20.1-20.4 ms per call against 21.7 ms
// with four (sixteen: within 2 %).
uint32_t host_chunks(bool chunkable, size_t len, size_t out_bytes)
{
    if (!chunkable || len < 16) return 1;
    return (uint32_t)std::min<size_t>(kMaxChunks, std::max<size_t>(1, out_bytes / kChunkBytes));
}

static int ensure(void** ptr, size_t* have, size_t need)
{
    if (*have >= need && *ptr) return EM_OK;
    if (*ptr) EM_HIP(hipFree(*ptr));
    *ptr = nullptr;
    *have = 0;
    EM_HIP(hipMalloc(ptr, need));
    *have = need;
    return EM_OK;
}

int host_path_acquire(int device, size_t charge_bytes, size_t line_bytes, size_t step_bytes,
                      std::unique_lock<std::mutex>& lk, HostPathBuffers** out)
{
    if (device < 0 || device >= 64) EM_FAIL(EM_ERR_NO_DEVICE, "device index %d out of range", device);
    lk = std::unique_lock<std::mutex>(g_hp_mu);
    HostPathBuffers& b = g_hp[device];
    if (!b.stream) {
        EM_HIP(hipStreamCreateWithFlags(&b.stream, hipStreamNonBlocking));
        EM_HIP(hipStreamCreateWithFlags(&b.copy, hipStreamNonBlocking));
        for (auto& e : b.ev) EM_HIP(hipEventCreate(&e));
        for (auto& e : b.chunk_ev) EM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    EM_TRY(ensure(&b.charges, &b.charges_bytes, std::max<size_t>(charge_bytes, 16)));
    EM_TRY(ensure(&b.lines, &b.lines_bytes, line_bytes));
    if (step_bytes) EM_TRY(ensure(&b.steps, &b.steps_bytes, step_bytes));
    *out = &b;
    return EM_OK;
}

int host_path_release_all()
{
    std::lock_guard<std::mutex> lk(g_hp_mu);
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) return EM_OK;
    for (int i = 0; i < 64; ++i) {
        HostPathBuffers& b = g_hp[i];
        if (!b.stream && !b.charges && !b.lines) continue;
        EM_HIP(hipSetDevice(i));
        if (b.charges) EM_HIP(hipFree(b.charges));
        if (b.lines) EM_HIP(hipFree(b.lines));
        if (b.steps) EM_HIP(hipFree(b.steps));
        for (auto& e : b.ev)
            if (e) EM_HIP(hipEventDestroy(e));
        for (auto& e : b.chunk_ev)
            if (e) EM_HIP(hipEventDestroy(e));
        if (b.stream) EM_HIP(hipStreamDestroy(b.stream));
        if (b.copy) EM_HIP(hipStreamDestroy(b.copy));
        b = HostPathBuffers();
    }
    EM_HIP(hipSetDevice(cur));
    return EM_OK;
}

static double now_s()
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

static void fill_perf(em_perf* perf, double p, double n, double len, double kernel_s)
{
    if (!perf) return;
    perf->kernel_s = kernel_s;
    perf->interactions = p * n * (len - 1.0);
    perf->interactions_per_s = kernel_s > 0 ? perf->interactions / kernel_s : 0.0;
    perf->line_steps_per_s = kernel_s > 0 ? n * (len - 1.0) / kernel_s : 0.0;
}

static int host_trace(const float* charges, size_t p, float* x, float* y, float* z, size_t n,
                      size_t len, const em_step* st, const em_launch* launch, em_perf* perf,
                      uint32_t* steps)
{
    const double t0 = now_s();
    if (perf) {
        set_progress(perf, 0.0);
        perf->time_s = perf->kernel_s = perf->h2d_s = perf->d2h_s = 0.0;
        perf->h2d_bytes = perf->d2h_bytes = 0.0;
        perf->gflops = 0.0;
    }
    EM_TRY(check_sizes(p, n, len, st));
    if ((p && !charges) || !x || !y || !z) EM_FAIL(EM_ERR_INVALID, "NULL array");
    DeviceInfo dev;
    EM_TRY(current_device_info(dev));
    Shape sh;
    // line ends (no checkpoints here) over many charges: launch_trace_ends
    const bool ends = st->stop != 0 && p >= kEndsMinCharges;
    EM_TRY(choose_shape(dev, p, n, launch, sh, st->stop != 0, ends));

    const size_t row = n * sizeof(float);
    std::unique_lock<std::mutex> lk;
    HostPathBuffers* hp = nullptr;
    EM_TRY(host_path_acquire(dev.device, p * sizeof(float4), 3 * len * row,
                             steps ? n * sizeof(uint32_t) : 0, lk, &hp));
    HostPathBuffers& b = *hp;
    float* dx = static_cast<float*>(b.lines);
    float* dy = dx + len * n;
    float* dz = dy + len * n;

    EM_HIP(hipEventRecord(b.ev[0], b.stream));
    const CopySeg in[4] = {{dx, x, row}, {dy, y, row}, {dz, z, row},
                           {b.charges, charges, p * sizeof(float4)}};
    EM_TRY(h2d_staged(dev.device, in, p ? 4 : 3, b.stream));
    EM_HIP(hipEventRecord(b.ev[1], b.stream));

    Cfg cfg{};
    cfg.charge_off = 0;
    cfg.line_off = 0;
    cfg.pitch = n;
    cfg.p = (uint32_t)p;
    cfg.n = (uint32_t)n;
    cfg.len = (uint32_t)len;
    const StepP sp = to_stepp(*st);
    uint32_t* dsteps = steps ? static_cast<uint32_t*>(b.steps) : nullptr;
    const size_t out_bytes = (len - 1) * row;
    const uint32_t chunks =
        host_chunks(!ends && (sp.rule == kRulePlain || sp.rule == kRuleRef), len, 3 * out_bytes);
    // chunk c covers rows bound(c) .. bound(c+1)-1
    auto bound = [&](uint32_t c) { return (uint32_t)(1 + (uint64_t)(len - 1) * c / chunks); };
    if (ends) {
        EM_TRY(launch_trace_ends(sh, dev, cfg, static_cast<const float4*>(b.charges), dx, dy, dz, sp,
                                 dsteps, b.stream));
        EM_HIP(hipEventRecord(b.chunk_ev[0], b.stream));
    } else {
        for (uint32_t c = 0; c < chunks; ++c) {
            Cfg part = cfg;
            part.len = bound(c + 1);
            EM_TRY(launch_trace_direct(sh, part, static_cast<const float4*>(b.charges), dx, dy, dz,
                                       sp, nullptr, 1, dsteps, b.stream, bound(c)));
            EM_HIP(hipEventRecord(b.chunk_ev[c], b.stream));
        }
    }
    EM_HIP(hipEventRecord(b.ev[2], b.stream));
    double d2h = 0.0;
    for (uint32_t c = 0; c < chunks; ++c) {
        const size_t r0 = bound(c), bytes = (size_t)(bound(c + 1) - r0) * row;
        const size_t o = r0 * n;
        const CopySeg segs[4] = {{x + o, dx + o, bytes}, {y + o, dy + o, bytes},
                                 {z + o, dz + o, bytes}, {steps, dsteps, n * sizeof(uint32_t)}};
        // while the GPU traces: fault in the caller's output rows of this chunk
        prefault_pages(segs, 3);
        EM_HIP(hipEventSynchronize(b.chunk_ev[c]));  // d2h_s: transfer time alone
        EM_HIP(hipStreamWaitEvent(b.copy, b.chunk_ev[c], 0));
        const double td = now_s();
        EM_TRY(d2h_staged(dev.device, segs, (steps && c + 1 == chunks) ? 4 : 3, b.copy));
        d2h += now_s() - td;
        set_progress(perf, (double)bound(c + 1) / (double)len);
    }
    EM_HIP(hipEventSynchronize(b.ev[2]));

    float ms[2] = {0, 0};
    EM_HIP(hipEventElapsedTime(&ms[0], b.ev[0], b.ev[1]));
    EM_HIP(hipEventElapsedTime(&ms[1], b.ev[1], b.ev[2]));  // the kernels (chunks overlap copies)
    if (perf) {
        perf->time_s = now_s() - t0;
        perf->h2d_s = ms[0] * 1e-3;
        perf->d2h_s = d2h;
        perf->h2d_bytes = (double)(p * sizeof(float4) + 3 * row);
        perf->d2h_bytes = (double)(3 * out_bytes);
        fill_perf(perf, (double)p, (double)n, (double)len, ms[1] * 1e-3);
        // CPU_Implement.cpp:347-349
        perf->gflops = ((double)n * ((double)len - 1) * ((double)p * 18.0 + 13.0)) / perf->time_s / 1e9;
        perf->lanes_per_line = sh.S;
        perf->tile = (int32_t)sh.tile;
        perf->block = sh.block;
        perf->unroll = sh.unroll;
        set_progress(perf, 1.0);
    }
    return EM_OK;
}

}  // namespace em

using namespace em;

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char* em_last_error(void) { return g_err; }

const char* em_version(void) { return "emtrace 0.1 (gfx950)"; }

int em_device_count(void)
{
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) return 0;
    return c;
}

int em_init(int device)
{
    clear_error();
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
        EM_FAIL(EM_ERR_NO_DEVICE, "no HIP device visible");
    if (device < 0 || device >= count) EM_FAIL(EM_ERR_INVALID, "device %d of %d", device, count);
    EM_HIP(hipSetDevice(device));
    DeviceInfo d;
    EM_TRY(current_device_info(d));
    // the device address of em_trace_dev_f32_noalloc's fixed work memory
    // (looked up here once per device, else on that entry point's first call)
    return noalloc_scratch(device, nullptr);
}

int em_shutdown(void)
{
    clear_error();
    EM_TRY(host_path_release_all());
    EM_TRY(staging_release_all());
    EM_TRY(ends_release_all());
    return adj_release_all();
}

int em_calc_lines_f32(const float* charges, size_t p, float* x, float* y, float* z, size_t n,
                      size_t len, float resolution, int use_curvature, em_perf* perf)
{
    clear_error();
    em_step st{};
    st.rule = use_curvature ? EM_RULE_REF : EM_RULE_PLAIN;
    st.resolution = resolution;
    return host_trace(charges, p, x, y, z, n, len, &st, nullptr, perf, nullptr);
}

int em_trace_host_f32(const float* charges, size_t p, float* x, float* y, float* z, size_t n,
                      size_t len, const em_step* step, const em_launch* launch, em_perf* perf,
                      uint32_t* steps)
{
    clear_error();
    return host_trace(charges, p, x, y, z, n, len, step, launch, perf, steps);
}

size_t em_ckpt_floats(size_t n, size_t len, uint32_t ckpt_every)
{
    if (!ckpt_every || len < 1) return 0;
    return 8ull * n * ((len - 1) / ckpt_every + 1);
}

static int trace_dev(const float* d_charges, size_t p, float* d_x, float* d_y, float* d_z,
                     size_t n, size_t len, size_t pitch, const em_step* step,
                     const em_launch* launch, float* d_ckpt, uint32_t ckpt_every,
                     uint32_t* d_steps, void* stream, bool noalloc)
{
    EM_TRY(check_sizes(p, n, len, step));
    if ((p && !d_charges) || !d_x || !d_y || !d_z) EM_FAIL(EM_ERR_INVALID, "NULL device pointer");
    if (pitch < n) EM_FAIL(EM_ERR_INVALID, "pitch %zu < n %zu", pitch, n);
    if (d_ckpt && !ckpt_every) EM_FAIL(EM_ERR_INVALID, "checkpoint buffer given with ckpt_every = 0");
    DeviceInfo dev;
    EM_TRY(current_device_info(dev));
    Shape sh;
    // line ends without checkpoints over many charges: launch_trace_ends;
    // with checkpoints the one-kernel path, whose steps the adjoint's
    // recompute repeats
    const bool ends = step->stop && !d_ckpt && p >= kEndsMinCharges;
    EM_TRY(choose_shape(dev, p, n, launch, sh, step->stop != 0, ends));
    Cfg cfg{};
    cfg.pitch = pitch;
    cfg.p = (uint32_t)p;
    cfg.n = (uint32_t)n;
    cfg.len = (uint32_t)len;
    if (ends && noalloc)
        return launch_trace_ends_noalloc(sh, dev, cfg, reinterpret_cast<const float4*>(d_charges),
                                         d_x, d_y, d_z, to_stepp(*step), d_steps,
                                         static_cast<hipStream_t>(stream));
    if (ends)
        return launch_trace_ends(sh, dev, cfg, reinterpret_cast<const float4*>(d_charges), d_x, d_y,
                                 d_z, to_stepp(*step), d_steps, static_cast<hipStream_t>(stream));
    return launch_trace_direct(sh, cfg, reinterpret_cast<const float4*>(d_charges), d_x, d_y, d_z,
                               to_stepp(*step), d_ckpt, ckpt_every ? ckpt_every : 1, d_steps,
                               static_cast<hipStream_t>(stream));
}

int em_trace_dev_f32(const float* d_charges, size_t p, float* d_x, float* d_y, float* d_z,
                     size_t n, size_t len, size_t pitch, const em_step* step,
                     const em_launch* launch, float* d_ckpt, uint32_t ckpt_every,
                     uint32_t* d_steps, void* stream)
{
    clear_error();
    return trace_dev(d_charges, p, d_x, d_y, d_z, n, len, pitch, step, launch, d_ckpt, ckpt_every,
                     d_steps, stream, false);
}

int em_trace_dev_f32_noalloc(const float* d_charges, size_t p, float* d_x, float* d_y,
                             float* d_z, size_t n, size_t len, size_t pitch,
                             const em_step* step, const em_launch* launch, float* d_ckpt,
                             uint32_t ckpt_every, uint32_t* d_steps, void* stream)
{
    clear_error();
    return trace_dev(d_charges, p, d_x, d_y, d_z, n, len, pitch, step, launch, d_ckpt, ckpt_every,
                     d_steps, stream, true);
}

int em_field_dev_f32(const float* d_charges, size_t p, const float* d_px, const float* d_py,
                     const float* d_pz, size_t m, float scale, float* d_ex, float* d_ey,
                     float* d_ez, const em_launch* launch, void* stream)
{
    clear_error();
    if (!m) return EM_OK;
    if ((p && !d_charges) || !d_px || !d_py || !d_pz || !d_ex || !d_ey || !d_ez)
        EM_FAIL(EM_ERR_INVALID, "NULL device pointer");
    if (p > 0xFFFFFFFFull) EM_FAIL(EM_ERR_INVALID, "p must fit 32 bits");
    DeviceInfo dev;
    EM_TRY(current_device_info(dev));
    Shape sh;
    EM_TRY(choose_shape(dev, p, m, launch, sh));
    return launch_field(sh, reinterpret_cast<const float4*>(d_charges), (uint32_t)p, d_px, d_py,
                        d_pz, m, scale, d_ex, d_ey, d_ez, static_cast<hipStream_t>(stream));
}

int em_legacy_rand(uint32_t seed, int32_t* out, size_t count)
{
    clear_error();
    if (count && !out) EM_FAIL(EM_ERR_INVALID, "NULL output");
    emtrace::legacy::LibcRand g(seed);
    for (size_t i = 0; i < count; ++i) out[i] = g.next();
    return EM_OK;
}

int em_legacy_problem_f32(float* charges, size_t p, uint32_t charge_seed, float* x, float* y,
                          float* z, size_t width, size_t height, size_t depth, int random_lines,
                          uint32_t line_seed)
{
    clear_error();
    const size_t n = width * height * depth;
    if (p && !charges) EM_FAIL(EM_ERR_INVALID, "NULL charges");
    if (n && (!x || !y || !z)) EM_FAIL(EM_ERR_INVALID, "NULL line arrays");
    emtrace::legacy::LibcRand g(charge_seed);
    emtrace::legacy::init_charges(charges, p, g);
    if (random_lines) {
        g.reseed(line_seed);
        emtrace::legacy::init_random_lines(x, y, z, n, g);
    } else {
        emtrace::legacy::init_grid(x, y, z, width, height, depth);
    }
    return EM_OK;
}

}  // extern "C"
