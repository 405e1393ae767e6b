// wide.hpp -- the register-resident line-ends trace (wide.hip), as used by
// the line-ends entry points in ends.hip.
#pragma once

#include "em_internal.hpp"

namespace em {

constexpr uint32_t kWideSlots = 16;     // lines a workgroup traces at once (default)
constexpr int kWideMaxSlots = 32;       // ... at most (EMTRACE_WIDE_SLOTS)
constexpr uint32_t kWideMaxCharges = 16384;  // charges a workgroup holds in registers
#ifndef EMTRACE_WIDE_SETTLE
#define EMTRACE_WIDE_SETTLE 128
#endif
constexpr uint32_t kWideSettle = EMTRACE_WIDE_SETTLE;  // steps of the settling pass

struct WideCtrl {
    uint32_t next_line;  // FRESH: next line to hand out
    uint32_t flag;       // FRESH: set once the counter is exhausted
    uint32_t count1;     // entries of list 1 (lines that left FRESH)
    uint32_t count2;     // entries of list 2 (lines that left SETTLE)
};

// Where the wide trace's pieces lie in the line-ends work memory (after the
// control block at offset 0 and the step counts), and the launch it uses.
struct WideLayout {
    uint32_t grid = 0;      // workgroups (one per CU)
    uint32_t slots = 0;     // lines per workgroup at once
    uint32_t list_cap = 0;  // entries of each list: grid * slots
    size_t off_l1 = 0, off_l2 = 0, off_o1 = 0, off_o2 = 0, end = 0;
};

// The problem takes the wide trace: 4,096..16,384 charges (unless
// EMTRACE_ENDS_WIDE=0 selects the lane-group trace of ends.hip).
bool wide_applies(const Cfg& cfg);
// Layout from byte offset `off` on; grid and slots from the device (and the
// EMTRACE_WIDE_GRID / EMTRACE_WIDE_SLOTS overrides).
WideLayout wide_layout(const DeviceInfo& dev, size_t off);
// The launches of one wide trace over work memory `buf` (the WideCtrl at
// offset 0); allocates nothing, so the sequence may be captured.  The rows
// after each line's end are left to the caller's fill.
int wide_launch(const DeviceInfo& dev, const Cfg& cfg, const float4* charges, float* x, float* y,
                float* z, const StepP& sp, uint32_t* steps, hipStream_t stream, char* buf,
                const WideLayout& L);

}  // namespace em
