// oracle/ref_driver.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A flat extern "C" driver around the reference's own CPU tracer,
// CalcField_CPU<float/double> (ElectroMag/src/CPU Implement.h:25-30,
// ElectroMag/src/CPU_Implement.cpp:501-527).  oracle/Makefile compiles this
// file together with the reference's CPU_Implement.cpp, where it lies under
// the reference tree, into oracle/_ref/libemref.so.  Tests use it to write the
// golden vectors and to pin the C restatement; bench.py may time it as the CPU
// baseline.  It is never part of the product library.
//
// Marshalling only: plain arrays are copied into the reference's own
// Array<pointCharge<T>> and Vector3<Array<T>> containers (256-B aligned, as
// ElectroMag.cpp:237-240 allocates them), the reference function runs, and
// the rows are copied back.
#include <cstring>

#include "CPU Implement.h"

template <class T>
static int run(const T* charges, size_t p, T* x, T* y, T* z, size_t n,
               size_t len, T resolution, int use_curvature, double* time_s,
               double* gflops)
{
    Array<electro::pointCharge<T> > q;
    if (q.AlignAlloc(p, 256)) return -100;
    for (size_t i = 0; i < p; ++i) {
        q[i].position.x = charges[4 * i + 0];
        q[i].position.y = charges[4 * i + 1];
        q[i].position.z = charges[4 * i + 2];
        q[i].magnitude = charges[4 * i + 3];
    }
    Vector3<Array<T> > lines;
    if (lines.AlignAlloc(n * len, 256)) return -101;
    std::memcpy(lines.x.GetDataPointer(), x, sizeof(T) * n * len);
    std::memcpy(lines.y.GetDataPointer(), y, sizeof(T) * n * len);
    std::memcpy(lines.z.GetDataPointer(), z, sizeof(T) * n * len);
    perfPacket perf;
    perf.time = 0;
    perf.performance = 0;
    int rc = CalcField_CPU<T>(lines, q, n, resolution, perf, use_curvature != 0);
    std::memcpy(x, lines.x.GetDataPointer(), sizeof(T) * n * len);
    std::memcpy(y, lines.y.GetDataPointer(), sizeof(T) * n * len);
    std::memcpy(z, lines.z.GetDataPointer(), sizeof(T) * n * len);
    if (time_s) *time_s = perf.time;
    if (gflops) *gflops = perf.performance;
    return rc;
}

extern "C" int ref_calc_lines_f32(const float* charges, size_t p, float* x,
                                  float* y, float* z, size_t n, size_t len,
                                  float resolution, int use_curvature,
                                  double* time_s, double* gflops)
{
    return run<float>(charges, p, x, y, z, n, len, resolution, use_curvature,
                      time_s, gflops);
}

extern "C" int ref_calc_lines_f64(const double* charges, size_t p, double* x,
                                  double* y, double* z, size_t n, size_t len,
                                  double resolution, int use_curvature,
                                  double* time_s, double* gflops)
{
    return run<double>(charges, p, x, y, z, n, len, resolution, use_curvature,
                       time_s, gflops);
}
