// em_internal.hpp -- host-side internals shared by the translation units of
// libemtrace.so (error state, HIP checks, device properties, launch choice).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdarg>
#include <cstdio>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/emtrace.h"
#include "em_device.hpp"

namespace em {

// ---- error state (thread-local message, C-ABI status codes) ----------------
void set_error(const char* fmt, ...);
void clear_error();

#define EM_FAIL(code, ...)            \
    do {                              \
        ::em::set_error(__VA_ARGS__); \
        return (code);                \
    } while (0)

#define EM_HIP(call)                                                                     \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess)                                                            \
            EM_FAIL(e_ == hipErrorOutOfMemory ? EM_ERR_NOMEM : EM_ERR_HIP, "%s: %s (%s:%d)", \
                    #call, hipGetErrorString(e_), __FILE__, __LINE__);                   \
    } while (0)

#define EM_TRY(call)            \
    do {                        \
        int rc_ = (call);       \
        if (rc_) return rc_;    \
    } while (0)

// ---- device ----------------------------------------------------------------
struct DeviceInfo {
    int device = -1;
    int cus = 0;
    size_t lds_per_block = 0;
    char arch[64] = {0};
};
int current_device_info(DeviceInfo& out);  // inits lazily; EM_ERR_NO_DEVICE if none

// LDS a workgroup needs beside its charge tile: the tracer's kernels declare
// 256-272 B of static LDS, and static plus dynamic LDS must fit the
// workgroup's (and, for `resident` workgroups per CU, the CU's) budget.
constexpr size_t kLdsStatic = 512;

// Charge records that fit one tile when `resident` workgroups share a CU.
inline uint64_t tile_budget(const DeviceInfo& dev, int resident)
{
    const size_t share = dev.lds_per_block / (size_t)(resident < 1 ? 1 : resident);
    return share > kLdsStatic ? (share - kLdsStatic) / sizeof(float4) : 0;
}

// ---- launch shape ----------------------------------------------------------
struct Shape {
    int S = 1;         // lanes per line
    int block = 256;   // threads per workgroup
    uint32_t tile = 0; // charges per LDS tile
    int unroll = kUnroll;  // charge records per pipelined group: 2, 4 or 8
};

// Heuristic launch shape for one configuration (p charges, n lines); fields
// set in `req` (non-zero) override it.  Fails with EM_ERR_INVALID on an
// unsupported request.
// `stop`: the angle rule with line ends (favours wide line groups and small
// blocks); the forward and the adjoint of one problem must choose alike.
// `refill`: the first phase of the lane-group line-ends trace (launch_trace_ends,
// above 16,384 charges; the register-resident trace takes no shape).
int choose_shape(const DeviceInfo& dev, uint64_t p, uint64_t n, const em_launch* req, Shape& out,
                 bool stop = false, bool refill = false);
int validate_shape(const DeviceInfo& dev, const Shape& s);
// The unroll a shape gets when the request leaves it 0 (from the resident
// waves per SIMD, whether the charges are re-staged, and line ends).
int default_unroll(const DeviceInfo& dev, const Shape& s, uint64_t p, uint64_t n, bool stop);

// perf->progress = v, mirrored into *perf->progress_out when the caller gave one
inline void set_progress(em_perf* perf, double v)
{
    if (!perf) return;
    perf->progress = v;
    if (perf->progress_out) *perf->progress_out = v;
}

StepP to_stepp(const em_step& st);
int validate_step(const em_step* st);

// Dynamic LDS above 64 KiB (up to the 160 KiB a gfx950 workgroup may hold)
// is requested explicitly per kernel.
template <typename K>
int allow_lds(K kernel, size_t bytes)
{
    if (bytes > 65536)
        EM_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    return EM_OK;
}

// ---- kernels (defined in trace.hip / ends.hip / adjoint.hip) ---------------
// Rows row0 .. cfg.len-1 (row0 > 1: plain and ref rules only, no
// checkpoints; the trace continues from row row0-1).
int launch_trace_direct(const Shape& sh, const Cfg& cfg, const float4* charges, float* x, float* y,
                        float* z, const StepP& sp, float* ckpt, uint32_t ckpt_every,
                        uint32_t* steps, hipStream_t stream, uint32_t row0 = 1);
int launch_trace_persistent(const Shape& sh, int grid, const Cfg* cfgs, const Item* items,
                            uint32_t n_items, uint32_t* counter, const float4* charges, float* x,
                            float* y, float* z, const StepP& sp, float* ckpt, uint32_t ckpt_every,
                            uint32_t* steps, hipStream_t stream);
int persistent_grid(const Shape& sh, const DeviceInfo& dev, int* grid);
// Problems with at least this many charges trace the angle rule with line
// ends (and no checkpoints) through launch_trace_ends; smaller ones, whose
// steps are short sums, through the one-kernel path.
constexpr uint32_t kEndsMinCharges = 4096;

// The angle rule with line ends and no checkpoints (ends.hip): up to 16,384
// charges the register-resident trace of wide.hip (1,024-lane workgroups
// holding every charge record, three passes); above, lane groups that take a
// new line whenever theirs ends, up to a step cap, then 256-lane workgroups
// for the lines still moving.  Both end with a coalesced fill of the rows
// after each line's end.  `steps` may be NULL.
int launch_trace_ends(const Shape& sh, const DeviceInfo& dev, const Cfg& cfg, const float4* charges,
                      float* x, float* y, float* z, const StepP& sp, uint32_t* steps,
                      hipStream_t stream);
int ends_release_all();
// The same trace over a fixed device array of kNoAllocScratchBytes per device
// (em_trace_dev_f32_noalloc): no allocation; problems whose work memory
// exceeds the array fail with EM_ERR_INVALID before any launch.
constexpr size_t kNoAllocScratchBytes = EM_NOALLOC_SCRATCH_BYTES;
int launch_trace_ends_noalloc(const Shape& sh, const DeviceInfo& dev, const Cfg& cfg,
                              const float4* charges, float* x, float* y, float* z, const StepP& sp,
                              uint32_t* steps, hipStream_t stream);
// The device address of that array on `device` (which must be current),
// looked up once per device with hipGetSymbolAddress.
int noalloc_scratch(int device, char** out);
int launch_field(const Shape& sh, const float4* charges, uint32_t p, const float* px,
                 const float* py, const float* pz, uint64_t m, float scale, float* ex, float* ey,
                 float* ez, hipStream_t stream);

// Device buffers, two streams and events kept per device for the host-array
// entry points (the FP32 and FP64 forms share one set): charges, the lines'
// rows and step counts, grown on demand and kept until em_shutdown.
constexpr size_t kMaxChunks = 8;  // row chunks of a host-array trace
struct HostPathBuffers {
    void* charges = nullptr;
    size_t charges_bytes = 0;
    void* lines = nullptr;  // 3 * len * n elements
    size_t lines_bytes = 0;
    void* steps = nullptr;  // n step counts
    size_t steps_bytes = 0;
    hipStream_t stream = nullptr;
    hipStream_t copy = nullptr;  // copies back of finished row chunks, beside the kernels
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    hipEvent_t chunk_ev[kMaxChunks] = {};
};
// Takes `lk` on the set of `device` (current) for the caller's whole call,
// creates its streams and events on first use and grows its buffers to at
// least the given sizes.
int host_path_acquire(int device, size_t charge_bytes, size_t line_bytes, size_t step_bytes,
                      std::unique_lock<std::mutex>& lk, HostPathBuffers** out);
int host_path_release_all();
// Row chunks of a host-array trace whose rule carries no state besides the
// point and the seed (chunk k's copy back overlaps chunk k+1's kernel): one
// per 48 MiB of output, at most kMaxChunks; 1 when `chunkable` is false.
uint32_t host_chunks(bool chunkable, size_t len, size_t out_bytes);

// device -> pageable host copies through pinned staging buffers (staging.cpp);
// issued on `stream` after the work that produces `src`, returns when every
// byte is in `dst`
struct CopySeg {
    void* dst;
    const void* src;
    size_t bytes;
};
int d2h_staged(int device, const CopySeg* segs, int nseg, hipStream_t stream);
// host (pageable) -> device, through the same buffers; returns when the
// copies have completed
int h2d_staged(int device, const CopySeg* segs, int nseg, hipStream_t stream);
int staging_release_all();
// Write one byte of every 4 KiB page of each segment's dst (output memory the
// copy back will overwrite), with a small host thread team; a no-op below 16 MiB.
void prefault_pages(const CopySeg* segs, int nseg);

// adjoint (adjoint.hip)
struct AdjPlan;
AdjPlan* adj_new();
void adj_delete(AdjPlan* a);
int adj_prepare(AdjPlan* a, const std::vector<Cfg>& cfgs, const Shape& sh, uint32_t every,
                const DeviceInfo& dev);
int adj_launch(AdjPlan* a, const Cfg* d_cfgs, const float4* charges, const float* x,
               const float* y, const float* z, const float* ckpt, const float* gx,
               const float* gy, const float* gz, const StepP& sp, float4* grad, float* gseed,
               hipStream_t stream);
int adj_release_all();

}  // namespace em
