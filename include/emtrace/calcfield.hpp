// emtrace/calcfield.hpp -- C++ drop-in for the reference's calc-lines call.
//
// Replaces   CalcField_CPU<float>(fieldLines, pointCharges, n, resolution,
//                                 perfData, useCurvature)   (and <double>)
//   (ElectroMag/src/CPU Implement.h:25-30, called at ElectroMag.cpp:305-306)
// and the GPU stub CalcField(...) (GPGPU_Segment/src/Electrostatics.cpp:28-42)
// with emtrace::CalcField_GPU, which has the same arguments and return codes
// and runs the gfx950 kernels through the C ABI in emtrace.h; emtrace::CalcField
// takes the stub's array-of-structures lines.
//
// CalcField_GPU is generic over the containers: it only uses the public
// interface the reference's types expose --
//   fieldLines.GetSize(), fieldLines.GetDataPointers() -> {x, y, z} float*
//     (Vector3<Array<float>>, include/SOA_utils.hpp:70-101),
//   pointCharges.GetSize(), pointCharges.GetDataPointer() -> 16-byte records
//     {x, y, z, q} (Array<pointCharge<float>>, include/Data Structures.h:71-81,
//      include/Electrostatics.h:30-35),
//   perfData.time, .performance, .progress, .add(TimingInfo)
//     (perfPacket, include/Data Structures.h:221-236)
// -- so it accepts the reference's own objects unchanged.  For callers without
// those headers, emtrace::compat below provides types with the same members.
//
// Status codes: 0 ok, 1 n == 0, 2 resolution == 0, 3 fewer than two rows
// (the reference's), 16+ device errors (see emtrace.h; message via
// em_last_error()).  Unlike the SSE path, n need not be a multiple of 64.
#pragma once

#include <cstddef>
#include <cstdlib>
#include <string>
#include <type_traits>
#include <vector>

#include "../emtrace.h"

namespace emtrace {

namespace detail {
inline int calc_lines(float* x, float* y, float* z, const void* q, size_t p, size_t n, size_t len,
                      double resolution, bool curvature, em_perf* perf)
{
    return em_calc_lines_f32(static_cast<const float*>(q), p, x, y, z, n, len, (float)resolution,
                             curvature ? 1 : 0, perf);
}
// FP64 lines: the IEEE double kernel (the reference's GPU stub for doubles is
// not implemented, GPGPU_Segment/src/Electrostatics.cpp:36-42).
inline int calc_lines(double* x, double* y, double* z, const void* q, size_t p, size_t n,
                      size_t len, double resolution, bool curvature, em_perf* perf)
{
    return em_calc_lines_f64(static_cast<const double*>(q), p, x, y, z, n, len, resolution,
                             curvature ? 1 : 0, perf);
}
}  // namespace detail

namespace detail {
template <class Perf>
void fill_perf(Perf& perfData, const em_perf& p)
{
    typedef typename decltype(perfData.stepTimes)::value_type Timing;
    perfData.time = p.time_s;
    perfData.performance = p.gflops;
    perfData.add(Timing("Host to device transfer", p.h2d_s, (size_t)p.h2d_bytes));
    perfData.add(Timing("Kernel execution time", p.kernel_s));
    perfData.add(Timing("Device to host transfer", p.d2h_s, (size_t)p.d2h_bytes));
    perfData.progress = 1.0;
}
}  // namespace detail

template <class LinesSoA, class ChargeArray, class Perf>
int CalcField_GPU(LinesSoA& fieldLines, ChargeArray& pointCharges, const size_t n,
                  double resolution, Perf& perfData, bool useCurvature = false)
{
    auto lines = fieldLines.GetDataPointers();
    static_assert(sizeof(*pointCharges.GetDataPointer()) == 4 * sizeof(*lines.x),
                  "point charges must be {x, y, z, q} records of the line element type");
    const size_t len = n ? fieldLines.GetSize() / n : 0;
    em_perf p = {};
    p.progress_out = &perfData.progress;  // the caller's monitor sees each row chunk
    const int rc = detail::calc_lines(lines.x, lines.y, lines.z, pointCharges.GetDataPointer(),
                                      pointCharges.GetSize(), n, len, resolution, useCurvature, &p);
    if (rc != EM_OK) return rc;
    detail::fill_perf(perfData, p);
    return rc;
}

// The reference's GPU entry point with array-of-structures lines,
//   int CalcField(Array<Vector3<float>>& fieldLines, Array<pointCharge<float>>&,
//                 size_t n, float resolution, perfPacket&, bool useCurvature)
// (GPGPU_Segment/src/Electrostatics.cpp:28-42, a stub returning -1 there; the
// double overload likewise).  Element (line, step) is fieldLines[step*n + line]
// with x, y, z members; row 0 holds the seeds.  The rows go through planar
// buffers to the same kernels as CalcField_GPU.
template <class LinesAoS, class ChargeArray, class Perf>
int CalcField(LinesAoS& fieldLines, ChargeArray& pointCharges, const size_t n, double resolution,
              Perf& perfData, bool useCurvature = false)
{
    typedef typename std::remove_reference<decltype(fieldLines[0].x)>::type T;
    static_assert(sizeof(*pointCharges.GetDataPointer()) == 4 * sizeof(T),
                  "point charges must be {x, y, z, q} records of the line element type");
    const size_t total = fieldLines.GetSize();
    const size_t len = n ? total / n : 0;
    std::vector<T> x(n * len), y(n * len), z(n * len);
    for (size_t i = 0; i < n && i < total; ++i) {
        x[i] = fieldLines[i].x;
        y[i] = fieldLines[i].y;
        z[i] = fieldLines[i].z;
    }
    em_perf p = {};
    p.progress_out = &perfData.progress;
    const int rc = detail::calc_lines(x.data(), y.data(), z.data(), pointCharges.GetDataPointer(),
                                      pointCharges.GetSize(), n, len, resolution, useCurvature, &p);
    if (rc != EM_OK) return rc;
    for (size_t i = n; i < n * len; ++i) {
        fieldLines[i].x = x[i];
        fieldLines[i].y = y[i];
        fieldLines[i].z = z[i];
    }
    detail::fill_perf(perfData, p);
    return rc;
}

namespace compat {

// Minimal types with the reference's member names, for callers that do not
// include the reference's headers.
template <class T>
struct Vector3 {
    T x, y, z;
};

template <class T>
class Array {
public:
    Array() = default;
    Array(const Array&) = delete;
    Array& operator=(const Array&) = delete;
    ~Array() { Free(); }
    int AlignAlloc(size_t elements, size_t alignment = 256)
    {
        if (size_) return 1;
        raw_ = std::malloc(elements * sizeof(T) + alignment - 1);
        if (!raw_) return 1;
        data_ = reinterpret_cast<T*>((reinterpret_cast<size_t>(raw_) + alignment - 1) / alignment *
                                     alignment);
        size_ = elements;
        return 0;
    }
    void Free()
    {
        std::free(raw_);
        raw_ = nullptr;
        data_ = nullptr;
        size_ = 0;
    }
    T& operator[](size_t i) { return data_[i]; }
    size_t GetSize() const { return size_; }
    size_t GetSizeBytes() const { return size_ * sizeof(T); }
    T* GetDataPointer() { return data_; }

private:
    void* raw_ = nullptr;
    T* data_ = nullptr;
    size_t size_ = 0;
};

template <class T>
struct LinesSoA {  // the shape of Vector3<Array<T>>
    Array<T> x, y, z;
    int AlignAlloc(size_t elements, size_t alignment = 256)
    {
        int e = x.AlignAlloc(elements, alignment) | y.AlignAlloc(elements, alignment) |
                z.AlignAlloc(elements, alignment);
        if (e) Free();
        return e;
    }
    void Free()
    {
        x.Free();
        y.Free();
        z.Free();
    }
    size_t GetSize() const { return x.GetSize(); }
    Vector3<T*> GetDataPointers() { return {x.GetDataPointer(), y.GetDataPointer(), z.GetDataPointer()}; }
    Vector3<T> operator[](size_t i) { return {x[i], y[i], z[i]}; }
    void write(Vector3<T> v, size_t i)
    {
        x[i] = v.x;
        y[i] = v.y;
        z[i] = v.z;
    }
};

template <class T>
struct pointCharge {
    Vector3<T> position;
    T magnitude;
};

struct TimingInfo {
    double time;
    std::string message;
    double bandwidth;  // MB/s when a transfer is involved
    TimingInfo(const char* msg, double t) : time(t), message(msg), bandwidth(0) {}
    TimingInfo(const char* msg, double t, size_t bytes)
        : time(t), message(msg), bandwidth(t > 0 ? (double)bytes / t / (1024.0 * 1024.0) : 0)
    {
    }
};

struct perfPacket {
    double performance = 0, time = 0;  // GFLOP/s (reference formula), seconds
    std::vector<TimingInfo> stepTimes;
    volatile double progress = 0;
    void add(const TimingInfo& t) { stepTimes.push_back(t); }
};

}  // namespace compat
}  // namespace emtrace
