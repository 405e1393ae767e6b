// batch.cpp -- batch plans: many independent configurations traced by one
// persistent launch (no reference counterpart: the reference traces a single
// configuration per CalcField_CPU call, CPU_Implement.cpp:501-513).
//
// A plan splits every configuration into blocks of block/S lines, sorts the
// (configuration, line-block) items by cost p * lines * (len - 1), largest
// first, and uploads descriptors and items once.  A run resets the work
// counter with a one-thread kernel (no memset node, so the run can be
// captured) and launches the persistent grid, sized to fill every CU.  With
// use_graph the first run captures both launches into a hipGraph and later
// runs with the same buffers replay it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <numeric>
#include <vector>

#include "em_internal.hpp"

struct em_batch {
    int device = -1;
    em::Shape sh;
    int grid = 0;
    em::StepP sp{};
    em_step step{};
    std::vector<em::Cfg> cfgs;
    std::vector<em::Item> items;
    em::Cfg* d_cfgs = nullptr;
    em::Item* d_items = nullptr;
    uint32_t* d_counter = nullptr;
    double interactions = 0.0;
    uint64_t total_lines = 0;
    uint32_t ckpt_every = 0;  // offsets currently in d_cfgs
    bool use_graph = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const void* graph_key[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    uint32_t graph_every = 0;
    hipStream_t cap_stream = nullptr;  // graphs are recorded here (the caller's stream may be
                                       // the null stream, which cannot be captured)
    em::AdjPlan* adj = nullptr;  // adjoint work lists, built on the first adjoint call
    uint32_t adj_every = 0;
};

namespace em {

static int free_plan(em_batch* b)
{
    if (b->adj) adj_delete(b->adj);
    b->adj = nullptr;
    if (b->exec) EM_HIP(hipGraphExecDestroy(b->exec));
    if (b->graph) EM_HIP(hipGraphDestroy(b->graph));
    if (b->d_cfgs) EM_HIP(hipFree(b->d_cfgs));
    if (b->d_items) EM_HIP(hipFree(b->d_items));
    if (b->d_counter) EM_HIP(hipFree(b->d_counter));
    if (b->cap_stream) EM_HIP(hipStreamDestroy(b->cap_stream));
    b->cap_stream = nullptr;
    b->exec = nullptr;
    b->graph = nullptr;
    b->d_cfgs = nullptr;
    b->d_items = nullptr;
    b->d_counter = nullptr;
    return EM_OK;
}

// Set the checkpoint offsets of every configuration for `every` and upload.
static int set_ckpt_offsets(em_batch* b, uint32_t every, hipStream_t stream)
{
    if (every == b->ckpt_every) return EM_OK;
    EM_HIP(hipStreamSynchronize(stream));  // earlier runs may still read d_cfgs
    uint64_t off = 0;
    for (auto& c : b->cfgs) {
        c.ckpt_off = off;
        if (every) off += 8ull * c.n * ((c.len - 1) / every + 1);
    }
    EM_HIP(hipMemcpy(b->d_cfgs, b->cfgs.data(), sizeof(Cfg) * b->cfgs.size(), hipMemcpyHostToDevice));
    b->ckpt_every = every;
    return EM_OK;
}

static int run_plan(em_batch* b, const float* d_charges, float* d_x, float* d_y, float* d_z,
                    float* d_ckpt, uint32_t every, uint32_t* d_steps, hipStream_t stream)
{
    int dev = 0;
    EM_HIP(hipGetDevice(&dev));
    if (dev != b->device) EM_FAIL(EM_ERR_INVALID, "plan was created on device %d, current is %d", b->device, dev);
    // the descriptors' checkpoint offsets are read only when checkpoints are
    // written: a plain run leaves them as the last checkpointed run set them
    // (a graph captured with checkpoints keeps seeing its own offsets)
    if (d_ckpt) EM_TRY(set_ckpt_offsets(b, every, stream));
    const float4* ch = reinterpret_cast<const float4*>(d_charges);
    const uint32_t ev = every ? every : 1;
    if (!b->use_graph)
        return launch_trace_persistent(b->sh, b->grid, b->d_cfgs, b->d_items, (uint32_t)b->items.size(),
                                       b->d_counter, ch, d_x, d_y, d_z, b->sp, d_ckpt, ev, d_steps,
                                       stream);
    const void* key[6] = {d_charges, d_x, d_y, d_z, d_ckpt, d_steps};
    if (!b->exec || std::memcmp(key, b->graph_key, sizeof(key)) != 0 || b->graph_every != ev) {
        if (b->exec) EM_HIP(hipGraphExecDestroy(b->exec));
        if (b->graph) EM_HIP(hipGraphDestroy(b->graph));
        b->exec = nullptr;
        b->graph = nullptr;
        if (!b->cap_stream) EM_HIP(hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking));
        EM_HIP(hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal));
        int rc = launch_trace_persistent(b->sh, b->grid, b->d_cfgs, b->d_items,
                                         (uint32_t)b->items.size(), b->d_counter, ch, d_x, d_y, d_z,
                                         b->sp, d_ckpt, ev, d_steps, b->cap_stream);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(b->cap_stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        EM_HIP(e);
        b->graph = g;
        EM_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        std::memcpy(b->graph_key, key, sizeof(key));
        b->graph_every = ev;
    }
    EM_HIP(hipGraphLaunch(b->exec, stream));
    return EM_OK;
}

}  // namespace em

using namespace em;

extern "C" {

int em_batch_create(const em_batch_cfg* cfgs, size_t count, const em_step* step,
                    const em_launch* launch, int use_graph, em_batch** out)
{
    clear_error();
    if (!out) EM_FAIL(EM_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (!cfgs || !count) EM_FAIL(EM_ERR_NO_LINES, "no configurations");
    if (count > 0xFFFFFFFFull) EM_FAIL(EM_ERR_INVALID, "too many configurations");
    EM_TRY(validate_step(step));
    uint64_t max_p = 0, total_lines = 0;
    for (size_t i = 0; i < count; ++i) {
        const em_batch_cfg& c = cfgs[i];
        if (!c.n) EM_FAIL(EM_ERR_NO_LINES, "configuration %zu has n == 0", i);
        if (c.len < 2) EM_FAIL(EM_ERR_LENGTH, "configuration %zu has len %u < 2", i, c.len);
        if (c.pitch < c.n) EM_FAIL(EM_ERR_INVALID, "configuration %zu: pitch < n", i);
        max_p = std::max<uint64_t>(max_p, c.p);
        total_lines += c.n;
    }
    DeviceInfo dev;
    EM_TRY(current_device_info(dev));
    em_batch* b = new (std::nothrow) em_batch();
    if (!b) EM_FAIL(EM_ERR_NOMEM, "host allocation failed");
    b->device = dev.device;
    b->step = *step;
    b->sp = to_stepp(*step);
    b->use_graph = use_graph != 0;
    b->total_lines = total_lines;

    // Launch shape: lanes per line from the mean line count, tile from the
    // largest configuration and the persistent residency.
    em_launch req{};
    if (launch) req = *launch;
    if (!req.lanes_per_line) {
        int S = 4;
        if (step->stop) {
            // line ends: a wave runs until its longest line has ended, so
            // wide groups (up to 16 lanes, >= 16 charges each); the 512
            // configurations with line ends: S=16 113 ms, S=32 119, S=4 199;
            // 64 configurations of <= 1,024 charges, checkpointed forward +
            // adjoint: S=16 25.0 ms, S=4 39.8
            S = 1;
            while (S < 16 && max_p / (2ull * S) >= 16) S *= 2;
        } else {
            // 4 lanes per line: 64-line work items even out the persistent
            // grid's finishing times (measured on 512 uneven configurations:
            // S = 4 ran 1.3x faster than S = 1), as long as every lane keeps
            // >= 64 charges
            while (S > 1 && max_p / (uint64_t)S < 64) S /= 2;
        }
        req.lanes_per_line = S;
    }
    // 256-thread blocks: with small tiles four workgroups share a CU, so while
    // one re-stages a tile or finishes a short item the others keep issuing
    // (512 configurations, up to 8,192 charges: 432 ms against 452-467 ms with
    // 512 threads and tiles for two workgroups per CU).  With line ends a
    // block runs until its longest line has ended: 128 threads (the same 512
    // configurations: 103 ms against 112 ms with 256 threads; 64
    // configurations of <= 1,024 charges, checkpointed forward + adjoint in
    // the inverse fit: 21.6 against 22.2 ms per iteration)
    if (!req.block) req.block = step->stop ? 128 : 256;
    Shape sh;
    int rc = choose_shape(dev, max_p, total_lines, &req, sh);
    if (rc) {
        delete b;
        return rc;
    }
    if (!req.tile) {
        // residency of the persistent grid: as many blocks per CU as fit 4 waves/SIMD
        const int resident = std::max(1, 1024 / sh.block);
        const uint64_t Q = std::max<uint64_t>(256, (uint64_t)sh.S * kGroup);
        uint64_t tmax = tile_budget(dev, resident) / Q * Q;
        tmax = std::max<uint64_t>(tmax, Q);
        const uint64_t need = (std::max<uint64_t>(max_p, 1) + Q - 1) / Q * Q;
        sh.tile = (uint32_t)std::min(need, tmax);
    }
    if (!req.unroll) sh.unroll = default_unroll(dev, sh, max_p, total_lines, step->stop != 0);
    b->sh = sh;
    rc = persistent_grid(sh, dev, &b->grid);
    if (rc) {
        delete b;
        return rc;
    }

    const uint32_t lpb = (uint32_t)(sh.block / sh.S);
    uint64_t line_index = 0;
    b->cfgs.resize(count);
    for (size_t i = 0; i < count; ++i) {
        const em_batch_cfg& c = cfgs[i];
        Cfg& d = b->cfgs[i];
        d = Cfg{};
        d.charge_off = c.charge_offset;
        d.line_off = c.line_offset;
        d.pitch = c.pitch;
        d.p = c.p;
        d.n = c.n;
        d.len = c.len;
        d.line_index = line_index;
        line_index += c.n;
        b->interactions += (double)c.p * (double)c.n * (double)(c.len - 1);
        for (uint32_t l0 = 0; l0 < c.n; l0 += lpb) b->items.push_back(Item{(uint32_t)i, l0});
    }
    std::vector<double> cost(b->items.size());
    for (size_t j = 0; j < b->items.size(); ++j) {
        const Item& it = b->items[j];
        const Cfg& c = b->cfgs[it.cfg];
        const uint32_t lines = std::min(lpb, c.n - it.line0);
        // a block costs its wave count (not its line count) times the charge sum
        const double waves = std::ceil((double)lines * sh.S / 64.0);
        cost[j] = (double)std::max(c.p, 1u) * waves * (double)(c.len - 1);
    }
    std::vector<size_t> order(b->items.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t c) { return cost[a] > cost[c]; });
    std::vector<Item> sorted(order.size());
    for (size_t j = 0; j < order.size(); ++j) sorted[j] = b->items[order[j]];
    b->items.swap(sorted);

    hipError_t e = hipMalloc(reinterpret_cast<void**>(&b->d_cfgs), sizeof(Cfg) * count);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&b->d_items), sizeof(Item) * b->items.size());
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&b->d_counter), sizeof(uint32_t) * 4);
    if (e == hipSuccess)
        e = hipMemcpy(b->d_cfgs, b->cfgs.data(), sizeof(Cfg) * count, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(b->d_items, b->items.data(), sizeof(Item) * b->items.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        free_plan(b);
        delete b;
        EM_FAIL(e == hipErrorOutOfMemory ? EM_ERR_NOMEM : EM_ERR_HIP, "batch plan upload: %s",
                hipGetErrorString(e));
    }
    *out = b;
    return EM_OK;
}

int em_batch_run(em_batch* plan, const float* d_charges, float* d_x, float* d_y, float* d_z,
                 uint32_t* d_steps, void* stream)
{
    clear_error();
    if (!plan || !d_charges || !d_x || !d_y || !d_z) EM_FAIL(EM_ERR_INVALID, "NULL argument");
    return run_plan(plan, d_charges, d_x, d_y, d_z, nullptr, 0, d_steps,
                    static_cast<hipStream_t>(stream));
}

int em_batch_info(const em_batch* plan, double* interactions, int32_t* items, int32_t* grid,
                  em_launch* used)
{
    clear_error();
    if (!plan) EM_FAIL(EM_ERR_INVALID, "NULL plan");
    if (interactions) *interactions = plan->interactions;
    if (items) *items = (int32_t)plan->items.size();
    if (grid) *grid = plan->grid;
    if (used) {
        used->lanes_per_line = plan->sh.S;
        used->tile = (int32_t)plan->sh.tile;
        used->unroll = plan->sh.unroll;
        used->block = plan->sh.block;
    }
    return EM_OK;
}

int em_batch_destroy(em_batch* plan)
{
    clear_error();
    if (!plan) return EM_OK;
    int rc = free_plan(plan);
    delete plan;
    return rc;
}

size_t em_batch_ckpt_floats(const em_batch* plan, uint32_t ckpt_every)
{
    if (!plan || !ckpt_every) return 0;
    size_t total = 0;
    for (const auto& c : plan->cfgs) total += 8ull * c.n * ((c.len - 1) / ckpt_every + 1);
    return total;
}

int em_batch_adjoint(em_batch* plan, const float* d_charges, const float* d_x, const float* d_y,
                     const float* d_z, const float* d_ckpt, uint32_t ckpt_every, const float* d_gx,
                     const float* d_gy, const float* d_gz, float* d_grad_charges, void* stream)
{
    clear_error();
    if (!plan || !d_charges || !d_x || !d_y || !d_z || !d_ckpt || !d_gx || !d_gy || !d_gz ||
        !d_grad_charges)
        EM_FAIL(EM_ERR_INVALID, "NULL argument");
    if (!ckpt_every) EM_FAIL(EM_ERR_INVALID, "ckpt_every == 0");
    if (plan->step.rule == kRuleRef)
        EM_FAIL(EM_ERR_UNSUPPORTED, "the adjoint supports the plain and angle rules");
    for (const auto& c : plan->cfgs)
        if ((uint64_t)c.n * (c.len - 1) > 0xFFFFFFFFull)
            EM_FAIL(EM_ERR_INVALID, "a configuration has more than 2^32 line-steps");
    hipStream_t st = static_cast<hipStream_t>(stream);
    int dev = 0;
    EM_HIP(hipGetDevice(&dev));
    if (dev != plan->device) EM_FAIL(EM_ERR_INVALID, "plan belongs to device %d", plan->device);
    EM_TRY(set_ckpt_offsets(plan, ckpt_every, st));
    if (!plan->adj || plan->adj_every != ckpt_every) {
        EM_HIP(hipStreamSynchronize(st));
        if (!plan->adj) plan->adj = adj_new();
        if (!plan->adj) EM_FAIL(EM_ERR_NOMEM, "host allocation failed");
        DeviceInfo di;
        EM_TRY(current_device_info(di));
        EM_TRY(adj_prepare(plan->adj, plan->cfgs, plan->sh, ckpt_every, di));
        plan->adj_every = ckpt_every;
    }
    return adj_launch(plan->adj, plan->d_cfgs, reinterpret_cast<const float4*>(d_charges), d_x, d_y,
                      d_z, d_ckpt, d_gx, d_gy, d_gz, plan->sp,
                      reinterpret_cast<float4*>(d_grad_charges), nullptr, st);
}

int em_batch_run_ckpt(em_batch* plan, const float* d_charges, float* d_x, float* d_y, float* d_z,
                      float* d_ckpt, uint32_t ckpt_every, uint32_t* d_steps, void* stream)
{
    clear_error();
    if (!plan || !d_charges || !d_x || !d_y || !d_z || !d_ckpt)
        EM_FAIL(EM_ERR_INVALID, "NULL argument");
    if (!ckpt_every) EM_FAIL(EM_ERR_INVALID, "ckpt_every == 0");
    return run_plan(plan, d_charges, d_x, d_y, d_z, d_ckpt, ckpt_every, d_steps,
                    static_cast<hipStream_t>(stream));
}

}  // extern "C"
