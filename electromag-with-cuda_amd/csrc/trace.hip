// trace.hip -- forward tracing kernels for gfx950.
//
// What the reference does on the CPU (ElectroMag/src/CPU_Implement.cpp:207-351,
// one OpenMP thread per 16-line SSE block, every step a sequential sum over all
// charges) and in its dormant OpenCL design (GPGPU_Segment/src/
// Electrostatics.cl.c:282-367, one work-item per line, 64-128-charge __local
// tiles) is done here as:
//   * k_trace_direct<S, U>  -- one workgroup per block of lines of a single
//                              configuration; all len-1 steps in one launch;
//   * k_trace_persistent<S, U> -- a persistent grid that pulls (configuration,
//                              line-block) items, largest first, from an atomic
//                              counter (the batched form);
//   * k_field<S>            -- the Coulomb sum alone at arbitrary points.
// Charges live in LDS for the whole launch when the padded set fits the tile
// (up to 160 KiB per workgroup on gfx950); otherwise they are re-staged tile by
// tile every step.  Row stores are coalesced: the lanes of a wave own
// consecutive lines, so each row component is one contiguous run.
#include "em_internal.hpp"

namespace em {

namespace {

__device__ __forceinline__ void store_ckpt(float* ckpt, uint64_t off, uint32_t m, uint32_t n,
                                           uint32_t line, const LineState& st)
{
    float4* c = reinterpret_cast<float4*>(ckpt + off) + 2ull * ((uint64_t)m * n + line);
    c[0] = make_float4(st.px, st.py, st.pz, st.h);
    c[1] = make_float4(st.ux, st.uy, st.uz, st.alive ? 1.f : 0.f);
}

// Trace lines [line0, line0 + blockDim/S) of one configuration, rows
// row0 .. cfg.len-1 (row0 > 1 continues from row row0-1: the plain and ref
// rules carry no state besides the point and the seed).
template <int S, int U>
__device__ __forceinline__ void trace_block(const Cfg& cfg, uint32_t line0, uint32_t row0,
                                            const float4* __restrict__ charges, float* __restrict__ X,
                                            float* __restrict__ Y, float* __restrict__ Z,
                                            const StepP& sp, float4* tile, uint32_t tile_cap,
                                            float* __restrict__ ckpt, uint32_t ckpt_every,
                                            uint32_t* __restrict__ steps_out)
{
    const int tid = threadIdx.x;
    const int k = tid % S;
    const uint32_t line = line0 + tid / S;
    const bool valid = line < cfg.n;
    const uint32_t li = valid ? line : cfg.n - 1;
    const float4* ch = charges + cfg.charge_off;
    const uint32_t p = cfg.p;
    const uint32_t pp = round_up(p, S * kGroup);
    const bool single = pp <= tile_cap;
    if (single) {
        stage(tile, ch, p, pp);
        __syncthreads();
    }
    const uint64_t base = cfg.line_off + li;
    const uint64_t from = base + (uint64_t)(row0 - 1) * cfg.pitch;
    LineState st;
    st.px = X[from];
    st.py = Y[from];
    st.pz = Z[from];
    st.Px = X[base]; st.Py = Y[base]; st.Pz = Z[base];
    st.ux = st.uy = st.uz = 0.f;
    st.h = sp.h0;
    st.alive = valid;
    st.steps = row0 - 1;
    const bool writer = valid && k == 0;
    if (ckpt && writer) store_ckpt(ckpt, cfg.ckpt_off, 0, cfg.n, line, st);
    for (uint32_t s = row0; s < cfg.len; ++s) {
        // Lines that ended keep writing their end point; the charge loop runs
        // while any line of the wave (of the block, when charges are re-staged
        // with barriers) is still moving.
        const bool need = single ? __any(st.alive) : __syncthreads_or(st.alive);
        if (need) {
            float ex, ey, ez;
            field_at<S, U>(tile, tile_cap, ch, p, k, st.px, st.py, st.pz, ex, ey, ez,
                           single || __any(st.alive) != 0);
            if (st.alive) {
                ex *= sp.scale;
                ey *= sp.scale;
                ez *= sp.scale;
                step_update(sp, s, ex, ey, ez, st);
                if (st.alive) st.steps = s;
            }
        }
        if (writer) {
            // ordinary (L2 write-back) stores: with S lanes per line a wave
            // writes 64/S consecutive lines, part of a 128-byte line per row;
            // the neighbouring waves' parts meet in L2 before they reach HBM
            const uint64_t o = cfg.line_off + (uint64_t)s * cfg.pitch + line;
            X[o] = st.px;
            Y[o] = st.py;
            Z[o] = st.pz;
            if (ckpt && s % ckpt_every == 0) store_ckpt(ckpt, cfg.ckpt_off, s / ckpt_every, cfg.n, line, st);
        }
    }
    if (writer && steps_out) steps_out[cfg.line_index + line] = st.steps;
}

template <int S, int U>
__global__ void __launch_bounds__(512) k_trace_direct(Cfg cfg, const float4* __restrict__ charges,
                                                      float* __restrict__ X, float* __restrict__ Y,
                                                      float* __restrict__ Z, StepP sp,
                                                      uint32_t tile_cap, float* __restrict__ ckpt,
                                                      uint32_t ckpt_every, uint32_t* __restrict__ steps,
                                                      uint32_t row0)
{
    extern __shared__ float4 tile[];
    const uint32_t line0 = blockIdx.x * (blockDim.x / S);
    trace_block<S, U>(cfg, line0, row0, charges, X, Y, Z, sp, tile, tile_cap, ckpt, ckpt_every, steps);
}

template <int S, int U>
__global__ void __launch_bounds__(512) k_trace_persistent(const Cfg* __restrict__ cfgs,
                                                          const Item* __restrict__ items,
                                                          uint32_t n_items, uint32_t* counter,
                                                          const float4* __restrict__ charges,
                                                          float* __restrict__ X, float* __restrict__ Y,
                                                          float* __restrict__ Z, StepP sp,
                                                          uint32_t tile_cap, float* __restrict__ ckpt,
                                                          uint32_t ckpt_every, uint32_t* __restrict__ steps)
{
    extern __shared__ float4 tile[];
    __shared__ uint32_t s_item;
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(counter, 1u);
        __syncthreads();
        const uint32_t it = s_item;
        __syncthreads();
        if (it >= n_items) break;  // uniform across the block: every wave exits
        const Item item = items[it];
        const Cfg cfg = cfgs[item.cfg];
        trace_block<S, U>(cfg, item.line0, 1u, charges, X, Y, Z, sp, tile, tile_cap, ckpt, ckpt_every,
                          steps);
        __syncthreads();  // the next item re-stages the tile
    }
}

__global__ void k_zero_counter(uint32_t* counter) { *counter = 0u; }

template <int S>
__global__ void __launch_bounds__(512) k_field(const float4* __restrict__ charges, uint32_t p,
                                               const float* __restrict__ PX, const float* __restrict__ PY,
                                               const float* __restrict__ PZ, uint64_t m, float scale,
                                               float* __restrict__ EX, float* __restrict__ EY,
                                               float* __restrict__ EZ, uint32_t tile_cap)
{
    extern __shared__ float4 tile[];
    const int k = threadIdx.x % S;
    const uint64_t i = (uint64_t)blockIdx.x * (blockDim.x / S) + threadIdx.x / S;
    const bool valid = i < m;
    const uint64_t ii = valid ? i : m - 1;
    const uint32_t pp = round_up(p, S * kGroup);
    if (pp <= tile_cap) {
        stage(tile, charges, p, pp);
        __syncthreads();
    }
    float ex, ey, ez;
    field_at<S>(tile, tile_cap, charges, p, k, PX[ii], PY[ii], PZ[ii], ex, ey, ez);
    if (valid && k == 0) {
        EX[i] = ex * scale;
        EY[i] = ey * scale;
        EZ[i] = ez * scale;
    }
}

// Lanes per line S and records per pipelined group U (2, 4 or 8: the sums
// consume the records in the same order whatever U, so U changes the speed
// only, not the bits).
template <template <int, int> class F, int S, typename... A>
int dispatch_U(int U, A&&... a)
{
    switch (U) {
        case 2: return F<S, 2>::run(a...);
        case 4: return F<S, 4>::run(a...);
        case 8: return F<S, 8>::run(a...);
        default: EM_FAIL(EM_ERR_INVALID, "unroll %d not in {2, 4, 8}", U);
    }
}

template <template <int, int> class F, typename... A>
int dispatch_SU(int S, int U, A&&... a)
{
    switch (S) {
        case 1: return dispatch_U<F, 1>(U, a...);
        case 2: return dispatch_U<F, 2>(U, a...);
        case 4: return dispatch_U<F, 4>(U, a...);
        case 8: return dispatch_U<F, 8>(U, a...);
        case 16: return dispatch_U<F, 16>(U, a...);
        case 32: return dispatch_U<F, 32>(U, a...);
        default: EM_FAIL(EM_ERR_INVALID, "lanes_per_line %d not in {1,2,4,8,16,32}", S);
    }
}

template <template <int> class F, typename... A>
int dispatch_S(int S, A&&... a)
{
    switch (S) {
        case 1: return F<1>::run(a...);
        case 2: return F<2>::run(a...);
        case 4: return F<4>::run(a...);
        case 8: return F<8>::run(a...);
        case 16: return F<16>::run(a...);
        case 32: return F<32>::run(a...);
        default: EM_FAIL(EM_ERR_INVALID, "lanes_per_line %d not in {1,2,4,8,16,32}", S);
    }
}

template <int S, int U>
struct DirectL {
    static int run(const Shape& sh, const Cfg& cfg, const float4* charges, float* x, float* y,
                   float* z, const StepP& sp, float* ckpt, uint32_t every, uint32_t* steps,
                   uint32_t row0, hipStream_t stream)
    {
        const uint32_t lpb = sh.block / S;
        const uint32_t grid = (cfg.n + lpb - 1) / lpb;
        const size_t lds = (size_t)sh.tile * sizeof(float4);
        EM_TRY(allow_lds(k_trace_direct<S, U>, lds));
        hipLaunchKernelGGL((k_trace_direct<S, U>), dim3(grid), dim3(sh.block), lds, stream, cfg,
                           charges, x, y, z, sp, sh.tile, ckpt, every, steps, row0);
        EM_HIP(hipGetLastError());
        return EM_OK;
    }
};

template <int S, int U>
struct PersistentL {
    static int run(const Shape& sh, int grid, const Cfg* cfgs, const Item* items, uint32_t n_items,
                   uint32_t* counter, const float4* charges, float* x, float* y, float* z,
                   const StepP& sp, float* ckpt, uint32_t every, uint32_t* steps,
                   hipStream_t stream)
    {
        const size_t lds = (size_t)sh.tile * sizeof(float4);
        EM_TRY(allow_lds(k_trace_persistent<S, U>, lds));
        hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(1), 0, stream, counter);
        EM_HIP(hipGetLastError());
        hipLaunchKernelGGL((k_trace_persistent<S, U>), dim3(grid), dim3(sh.block), lds, stream,
                           cfgs, items, n_items, counter, charges, x, y, z, sp, sh.tile, ckpt,
                           every, steps);
        EM_HIP(hipGetLastError());
        return EM_OK;
    }
};

template <int S, int U>
struct GridL {
    static int run(const Shape& sh, const DeviceInfo& dev, int* grid)
    {
        int per_cu = 0;
        const size_t lds = (size_t)sh.tile * sizeof(float4);
        EM_TRY(allow_lds(k_trace_persistent<S, U>, lds));
        EM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_trace_persistent<S, U>,
                                                             sh.block, lds));
        if (per_cu < 1) EM_FAIL(EM_ERR_INVALID, "launch shape does not fit a CU (block %d, tile %u)",
                                sh.block, sh.tile);
        *grid = per_cu * dev.cus;
        return EM_OK;
    }
};

template <int S>
struct FieldL {
    static int run(const Shape& sh, const float4* charges, uint32_t p, const float* px,
                   const float* py, const float* pz, uint64_t m, float scale, float* ex, float* ey,
                   float* ez, hipStream_t stream)
    {
        const uint32_t lpb = sh.block / S;
        const uint64_t grid = (m + lpb - 1) / lpb;
        const size_t lds = (size_t)sh.tile * sizeof(float4);
        EM_TRY(allow_lds(k_field<S>, lds));
        hipLaunchKernelGGL(k_field<S>, dim3((uint32_t)grid), dim3(sh.block), lds, stream, charges,
                           p, px, py, pz, m, scale, ex, ey, ez, sh.tile);
        EM_HIP(hipGetLastError());
        return EM_OK;
    }
};

}  // namespace

int launch_trace_direct(const Shape& sh, const Cfg& cfg, const float4* charges, float* x, float* y,
                        float* z, const StepP& sp, float* ckpt, uint32_t ckpt_every,
                        uint32_t* steps, hipStream_t stream, uint32_t row0)
{
    if (row0 < 1 || (row0 > 1 && (sp.rule == kRuleAngle || ckpt)))
        EM_FAIL(EM_ERR_INVALID, "a trace may start past row 1 only for the plain and ref rules, "
                                "without checkpoints");
    return dispatch_SU<DirectL>(sh.S, sh.unroll, sh, cfg, charges, x, y, z, sp, ckpt, ckpt_every,
                                steps, row0, stream);
}

int launch_trace_persistent(const Shape& sh, int grid, const Cfg* cfgs, const Item* items,
                            uint32_t n_items, uint32_t* counter, const float4* charges, float* x,
                            float* y, float* z, const StepP& sp, float* ckpt, uint32_t ckpt_every,
                            uint32_t* steps, hipStream_t stream)
{
    return dispatch_SU<PersistentL>(sh.S, sh.unroll, sh, grid, cfgs, items, n_items, counter,
                                    charges, x, y, z, sp, ckpt, ckpt_every, steps, stream);
}

int persistent_grid(const Shape& sh, const DeviceInfo& dev, int* grid)
{
    return dispatch_SU<GridL>(sh.S, sh.unroll, sh, dev, grid);
}

int launch_field(const Shape& sh, const float4* charges, uint32_t p, const float* px,
                 const float* py, const float* pz, uint64_t m, float scale, float* ex, float* ey,
                 float* ez, hipStream_t stream)
{
    return dispatch_S<FieldL>(sh.S, sh, charges, p, px, py, pz, m, scale, ex, ey, ez, stream);
}

}  // namespace em
