// cpu_tracer.hpp -- the CPU field-line tracer of the electromag command line.
//
// The reference application compares its GPU rows against its own CPU tracer
// when run with --cpu, and runs only the CPU tracer with --gpudisable
// (ElectroMag/src/ElectroMag.cpp:101-107, 300-322).  This is that CPU side,
// selected by those flags only: the GPU side always runs the HIP library and
// never substitutes this code.
//
// Arithmetic per line, in the order of ElectroMag/src/CPU_Implement.cpp:
//   field   E = sum over charges in index order of PartField
//           curvature: r*k_e*q / (|r|^2 sqrt(|r|^2))          (Electrostatics.h:38-51)
//           plain:     r * ((k_e*q) / (|r|^2 sqrt(|r|^2)))      (Electrostatics.h:53-60)
//   step    plain:     x += E / (res*|E|)                       (CPU_Implement.cpp:89-92)
//           curvature: k = |(E-P) x P| / |E|^3, P = the seed point,
//                      x += E / ((k+1)*res*|E|)                 (CPU_Implement.cpp:322-331)
// with IEEE single/double operations (the caller compiles this file without
// -ffast-math and without FMA contraction).  Lines are processed in blocks of
// kBlock so the compiler vectorises across lines, which keeps every line's
// operation sequence unchanged.  Status codes follow the reference: 1 n == 0,
// 2 resolution == 0, 5 curvature with n % 64 != 0 (the SSE path's
// alignment rule), 3 fewer than two rows (CPU_Implement.cpp:43-52, 213-230).
#pragma once

#include <atomic>
#include <cmath>
#include <cstddef>

#include <omp.h>

namespace emcli {

constexpr double kCoulomb = 8.987551787E9;  // include/Electrostatics.h:27
constexpr size_t kBlock = 16;

template <class T>
int cpu_calc_lines(const T* q, size_t p, T* x, T* y, T* z, size_t n, size_t len, T resolution,
                   bool curvature, volatile double* progress)
{
    if (!n) return 1;
    if (resolution == T(0)) return 2;
    if (curvature && (n & 63)) return 5;
    if (len < 2) return 3;
    const T ke = static_cast<T>(kCoulomb);
    const long long blocks = static_cast<long long>((n + kBlock - 1) / kBlock);
    std::atomic<long long> done(0);
    if (progress) *progress = 0.0;

#pragma omp parallel for schedule(dynamic, 4)
    for (long long b = 0; b < blocks; ++b) {
        const size_t l0 = static_cast<size_t>(b) * kBlock;
        const size_t m = n - l0 < kBlock ? n - l0 : kBlock;
        T px[kBlock], py[kBlock], pz[kBlock], sx[kBlock], sy[kBlock], sz[kBlock];
        T Px[kBlock], Py[kBlock], Pz[kBlock];
        for (size_t l = 0; l < m; ++l) {
            Px[l] = px[l] = x[l0 + l];
            Py[l] = py[l] = y[l0 + l];
            Pz[l] = pz[l] = z[l0 + l];
        }
        for (size_t step = 1; step < len; ++step) {
            for (size_t l = 0; l < m; ++l) sx[l] = sy[l] = sz[l] = T(0);
            for (size_t i = 0; i < p; ++i) {
                const T cx = q[4 * i], cy = q[4 * i + 1], cz = q[4 * i + 2], cq = q[4 * i + 3];
                if (curvature) {
                    for (size_t l = 0; l < m; ++l) {
                        const T rx = px[l] - cx, ry = py[l] - cy, rz = pz[l] - cz;
                        const T r2 = rx * rx + ry * ry + rz * rz;
                        const T den = r2 * std::sqrt(r2);
                        sx[l] += rx * ke * cq / den;
                        sy[l] += ry * ke * cq / den;
                        sz[l] += rz * ke * cq / den;
                    }
                } else {
                    const T kq = ke * cq;
                    for (size_t l = 0; l < m; ++l) {
                        const T rx = px[l] - cx, ry = py[l] - cy, rz = pz[l] - cz;
                        const T r2 = rx * rx + ry * ry + rz * rz;
                        const T s = kq / (r2 * std::sqrt(r2));
                        sx[l] += rx * s;
                        sy[l] += ry * s;
                        sz[l] += rz * s;
                    }
                }
            }
            for (size_t l = 0; l < m; ++l) {
                const T ex = sx[l], ey = sy[l], ez = sz[l];
                const T e2 = ex * ex + ey * ey + ez * ez;
                T scale;
                if (curvature) {
                    const T ix = ex - Px[l], iy = ey - Py[l], iz = ez - Pz[l];
                    const T kx = iy * Pz[l] - iz * Py[l];
                    const T ky = iz * Px[l] - ix * Pz[l];
                    const T kz = ix * Py[l] - iy * Px[l];
                    const T k = std::sqrt(kx * kx + ky * ky + kz * kz) / (e2 * std::sqrt(e2));
                    scale = (k + T(1)) * resolution;
                    scale = scale * std::sqrt(e2);
                } else {
                    scale = resolution * std::sqrt(e2);
                }
                px[l] = px[l] + ex / scale;
                py[l] = py[l] + ey / scale;
                pz[l] = pz[l] + ez / scale;
                const size_t at = step * n + l0 + l;
                x[at] = px[l];
                y[at] = py[l];
                z[at] = pz[l];
            }
        }
        const long long d = ++done;
        if (progress && omp_get_thread_num() == 0) *progress = static_cast<double>(d) / blocks;
    }
    if (progress) *progress = 1.0;
    return 0;
}

}  // namespace emcli
