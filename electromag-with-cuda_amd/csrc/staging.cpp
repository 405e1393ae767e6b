// staging.cpp -- host <-> device copies for the host-array entry points.
//
// Output arrays handed in by a caller are often fresh allocations whose pages
// the kernel has not touched yet: a copy into them takes a page fault per
// 4 KiB.  prefault_pages() takes those faults with a thread team while the
// GPU computes, so the copy back afterwards runs into resident pages.
//
// The host-array forms (the drop-in calc-lines calls, the FP64 and particle
// host forms) hand rows back into the caller's own arrays, which are ordinary
// pageable memory.  A plain hipMemcpy into pageable memory runs at a few GB/s.
// Here the rows go through two pinned staging buffers per device (hipHostMalloc,
// kept between calls, grown up to 64 MiB each): while the DMA engine fills one
// buffer with chunk c+1, a small team of host threads copies chunk c out of
// the other into the caller's arrays.  Host-to-device copies use the same
// buffers the other way round.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "em_internal.hpp"

namespace em {
namespace {

constexpr size_t kStageChunk = 64ull << 20;  // bytes per staging buffer

struct Staging {
    std::mutex mu;
    void* pin[2] = {nullptr, nullptr};
    size_t cap = 0;  // bytes of each buffer
    hipEvent_t ev[2] = {nullptr, nullptr};
};
Staging g_stage[64];

// Buffers of at least min(largest segment, kStageChunk) bytes (pinned memory
// is slow to allocate: small transfers keep small buffers).
int ensure_staging(Staging& st, const CopySeg* segs, int nseg)
{
    size_t need = 4096;
    for (int i = 0; i < nseg; ++i) need = std::max(need, std::min(segs[i].bytes, kStageChunk));
    need = (need + 4095) / 4096 * 4096;
    if (st.cap < need) {
        for (int i = 0; i < 2; ++i) {
            if (st.pin[i]) EM_HIP(hipHostFree(st.pin[i]));
            st.pin[i] = nullptr;
        }
        st.cap = 0;
        for (int i = 0; i < 2; ++i) EM_HIP(hipHostMalloc(&st.pin[i], need, hipHostMallocDefault));
        st.cap = need;
    }
    for (int i = 0; i < 2; ++i)
        if (!st.ev[i]) EM_HIP(hipEventCreateWithFlags(&st.ev[i], hipEventDisableTiming));
    return EM_OK;
}

struct Chunk {
    char* dst;
    const char* src;
    size_t bytes;
};

void copy_part(const Chunk& c, const void* pin, unsigned t, unsigned T)
{
    const size_t per = (c.bytes + T - 1) / T;
    const size_t lo = std::min(c.bytes, per * t), hi = std::min(c.bytes, lo + per);
    if (hi > lo) std::memcpy(c.dst + lo, static_cast<const char*>(pin) + lo, hi - lo);
}

}  // namespace

int d2h_staged(int device, const CopySeg* segs, int nseg, hipStream_t stream)
{
    if (device < 0 || device >= 64) EM_FAIL(EM_ERR_INVALID, "device %d", device);
    Staging& st = g_stage[device];
    std::lock_guard<std::mutex> lk(st.mu);
    EM_TRY(ensure_staging(st, segs, nseg));
    std::vector<Chunk> ch;
    for (int i = 0; i < nseg; ++i)
        for (size_t off = 0; off < segs[i].bytes; off += st.cap)
            ch.push_back(Chunk{static_cast<char*>(segs[i].dst) + off,
                               static_cast<const char*>(segs[i].src) + off,
                               std::min(st.cap, segs[i].bytes - off)});
    if (ch.empty()) return EM_OK;
    auto issue = [&](size_t c) -> int {
        EM_HIP(hipMemcpyAsync(st.pin[c & 1], ch[c].src, ch[c].bytes, hipMemcpyDeviceToHost, stream));
        EM_HIP(hipEventRecord(st.ev[c & 1], stream));
        return EM_OK;
    };
    EM_TRY(issue(0));

    // host team: thread 0 is this thread; the others copy their share of each
    // chunk once `ready` names it, and count themselves in `done`
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const unsigned T = ch.size() > 1 ? std::min(8u, std::max(1u, hw / 2)) : 1u;
    std::atomic<long> ready(-1);
    std::atomic<unsigned> done(0);
    std::atomic<bool> stop(false);
    std::vector<std::thread> team;
    for (unsigned t = 1; t < T; ++t)
        team.emplace_back([&, t] {
            long seen = -1;
            for (;;) {
                long c = ready.load(std::memory_order_acquire);
                if (c == seen) {
                    if (stop.load(std::memory_order_acquire)) return;
                    std::this_thread::yield();
                    continue;
                }
                seen = c;
                copy_part(ch[(size_t)c], st.pin[c & 1], t, T);
                done.fetch_add(1, std::memory_order_acq_rel);
            }
        });
    int rc = EM_OK;
    for (size_t c = 0; c < ch.size(); ++c) {
        const hipError_t e = hipEventSynchronize(st.ev[c & 1]);
        if (e != hipSuccess) {
            set_error("hipEventSynchronize: %s", hipGetErrorString(e));
            rc = EM_ERR_HIP;
            break;
        }
        // the other buffer's chunk (c-1) has been copied out: refill it
        if (c + 1 < ch.size() && (rc = issue(c + 1)) != EM_OK) break;
        done.store(0, std::memory_order_release);
        ready.store((long)c, std::memory_order_release);
        copy_part(ch[c], st.pin[c & 1], 0, T);
        while (done.load(std::memory_order_acquire) < T - 1) std::this_thread::yield();
    }
    stop.store(true, std::memory_order_release);
    for (auto& th : team) th.join();
    if (rc != EM_OK) (void)hipStreamSynchronize(stream);  // no DMA may outlive the call
    return rc;
}

int h2d_staged(int device, const CopySeg* segs, int nseg, hipStream_t stream)
{
    if (device < 0 || device >= 64) EM_FAIL(EM_ERR_INVALID, "device %d", device);
    Staging& st = g_stage[device];
    std::lock_guard<std::mutex> lk(st.mu);
    EM_TRY(ensure_staging(st, segs, nseg));
    // chunk c goes through buffer c & 1 once that buffer's previous DMA is done
    bool used[2] = {false, false};
    size_t c = 0;
    for (int i = 0; i < nseg; ++i)
        for (size_t off = 0; off < segs[i].bytes; off += st.cap, ++c) {
            const size_t b = std::min(st.cap, segs[i].bytes - off);
            if (used[c & 1]) EM_HIP(hipEventSynchronize(st.ev[c & 1]));
            std::memcpy(st.pin[c & 1], static_cast<const char*>(segs[i].src) + off, b);
            EM_HIP(hipMemcpyAsync(static_cast<char*>(segs[i].dst) + off, st.pin[c & 1], b,
                                  hipMemcpyHostToDevice, stream));
            EM_HIP(hipEventRecord(st.ev[c & 1], stream));
            used[c & 1] = true;
        }
    // the buffers are reused by the next call: wait for their last DMA
    for (int i = 0; i < 2; ++i)
        if (used[i]) EM_HIP(hipEventSynchronize(st.ev[i]));
    return EM_OK;
}

void prefault_pages(const CopySeg* segs, int nseg)
{
    constexpr size_t kPage = 4096;
    size_t total = 0;
    for (int i = 0; i < nseg; ++i) total += segs[i].bytes;
    if (total < (16ull << 20)) return;  // a few thousand faults: not worth a team
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const unsigned T = std::min(16u, std::max(1u, hw / 2));
    auto touch = [&](unsigned t) {
        for (int i = 0; i < nseg; ++i) {
            char* base = static_cast<char*>(segs[i].dst);
            const size_t per = (segs[i].bytes + T - 1) / T;
            const size_t lo = std::min(segs[i].bytes, per * t), hi = std::min(segs[i].bytes, lo + per);
            // one write per page; the copy back overwrites every byte later
            for (size_t off = lo; off < hi; off += kPage) *reinterpret_cast<volatile char*>(base + off) = 0;
        }
    };
    std::vector<std::thread> team;
    for (unsigned t = 1; t < T; ++t) team.emplace_back(touch, t);
    touch(0);
    for (auto& th : team) th.join();
}

int staging_release_all()
{
    int rc = EM_OK;
    for (auto& st : g_stage) {
        std::lock_guard<std::mutex> lk(st.mu);
        for (int i = 0; i < 2; ++i) {
            if (st.pin[i] && hipHostFree(st.pin[i]) != hipSuccess) rc = EM_ERR_HIP;
            if (st.ev[i] && hipEventDestroy(st.ev[i]) != hipSuccess) rc = EM_ERR_HIP;
            st.pin[i] = nullptr;
            st.ev[i] = nullptr;
        }
        st.cap = 0;
    }
    return rc;
}

}  // namespace em
