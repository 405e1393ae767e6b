/*
 * oracle/restate.c -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * A plain-C restatement of the reference's CPU field-line tracer, used by
 * tests/ and by bench.py's cpu_baseline leg as the parity checker.  Nothing in
 * the product library links, calls or loads this file.
 *
 * What it restates (reference = t0b1as/electromag-with-cuda):
 *   - electro::PartField, 3-argument form r*k*q/(|r|^2*sqrt(|r|^2))
 *       include/Electrostatics.h:38-51
 *   - electro::PartField, 2-argument form r*((k*q)/(|r|^2*sqrt(|r|^2)))
 *       include/Electrostatics.h:53-60
 *   - vec3SetInvLen / vec3LenSq / vec3Len / vec3Cross
 *       include/Vector.h:158-164, 166-175, 189-197
 *   - CalcField_CPU_T (plain step, generic template, 2-arg PartField)
 *       ElectroMag/src/CPU_Implement.cpp:38-102
 *   - CalcField_CPU_T_Curvature<float>/<double> SSE specialisations
 *     (3-arg PartField, curvature term with prevAccum = the seed point,
 *      n % 64 alignment error 5)
 *       ElectroMag/src/CPU_Implement.cpp:207-351, 366-498
 *   - the error-code order of those functions (1: n==0, 2: resolution==0,
 *     5: n not a multiple of 64 [SSE paths only], 3: fewer than 2 rows)
 *
 * The SSE code computes each lane with IEEE single/double operations; this
 * restatement performs the same per-lane operations in the same source order
 * and is compiled without -ffast-math and without FMA contraction, so it is
 * the literal-order reading of the reference source.  (The reference itself is
 * built with -ffast-math, which lets GCC reassociate; oracle/_ref measures
 * how far that moves the bits.)
 *
 * Project-defined addition (no reference counterpart): the "angle" step rule,
 *   u_s = E_s/|E_s|,  c_s = |u_s - u_{s-1}|,
 *   h_1 = h0,  h_s = max(hmin, c_s*hmax > turn*h_{s-1} ? turn*h_{s-1}/c_s : hmax)
 *   x_s = x_{s-1} + h_s*u_s
 * with the field summed by the 3-argument PartField in charge order and k = 1.
 * With stop = 1 a line ends at the first step s > 1 whose direction reverses
 * (u_s . u_{s-1} < 0, the line went through a sink): it does not move at that
 * step or later, and its remaining rows repeat the end point.  Its op order is
 * shared bit-for-bit with the HIP step code.
 *
 * Parity pins for this file: tests/test_oracle.py (golden vectors written by
 * the compiled reference, closed forms for a single charge, scale and
 * translation invariance, REF == PLAIN when k+1 rounds to 1).
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <string.h>

#define ORC_KE 8.987551787E9 /* include/Electrostatics.h:27 */

enum { ORC_PLAIN = 0, ORC_REF = 1, ORC_ANGLE = 2 };

typedef struct orc_step_params {
    int rule;          /* ORC_PLAIN / ORC_REF / ORC_ANGLE */
    double resolution; /* plain & ref: step = 1/resolution */
    double h0, hmin, hmax, turn; /* angle rule */
    int stop;                    /* angle rule: end lines whose direction reverses */
} orc_step_params;

/* ---------------------------------------------------------------------------
 * Generated for T = float and T = double.
 * ------------------------------------------------------------------------- */
#define ORC_DEFINE(T, SUF, SQRT)                                               \
                                                                               \
/* Electrostatics.h:38-51: r*k*q / (lenSq*sqrt(lenSq)) */                      \
static inline void part_field3_##SUF(const T* c, T px, T py, T pz, T k,      \
                                     T* ex, T* ey, T* ez)                      \
{                                                                              \
    T rx = px - c[0], ry = py - c[1], rz = pz - c[2];                          \
    T lenSq = rx * rx + ry * ry + rz * rz;                                     \
    T den = lenSq * SQRT(lenSq);                                               \
    *ex = rx * k * c[3] / den;                                                 \
    *ey = ry * k * c[3] / den;                                                 \
    *ez = rz * k * c[3] / den;                                                 \
}                                                                              \
                                                                               \
/* Electrostatics.h:53-60: r * ((k*q) / (lenSq*sqrt(lenSq))) */                \
static inline void part_field2_##SUF(const T* c, T px, T py, T pz,           \
                                     T* ex, T* ey, T* ez)                      \
{                                                                              \
    T rx = px - c[0], ry = py - c[1], rz = pz - c[2];                          \
    T lenSq = rx * rx + ry * ry + rz * rz;                                     \
    T s = (T)ORC_KE * c[3] / (lenSq * SQRT(lenSq));                            \
    *ex = rx * s;                                                              \
    *ey = ry * s;                                                              \
    *ez = rz * s;                                                              \
}                                                                              \
                                                                               \
static void field_sum_##SUF(const T* ch, size_t p, int two_arg, T k,          \
                            T px, T py, T pz, T* ax, T* ay, T* az)             \
{                                                                              \
    T sx = 0, sy = 0, sz = 0;                                                  \
    for (size_t i = 0; i < p; ++i) {                                           \
        T ex, ey, ez;                                                          \
        if (two_arg)                                                           \
            part_field2_##SUF(ch + 4 * i, px, py, pz, &ex, &ey, &ez);          \
        else                                                                   \
            part_field3_##SUF(ch + 4 * i, px, py, pz, k, &ex, &ey, &ez);       \
        sx += ex; sy += ey; sz += ez; /* Vector.h:50-56 operator+= */          \
    }                                                                          \
    *ax = sx; *ay = sy; *az = sz;                                              \
}                                                                              \
                                                                               \
int orc_calc_lines_##SUF(const T* charges, size_t p, T* x, T* y, T* z,        \
                         size_t n, size_t len, const orc_step_params* sp,      \
                         uint32_t* steps)                                      \
{                                                                              \
    const T res = (T)sp->resolution;                                           \
    if (!n) return 1;                                                          \
    if (res == 0 && sp->rule != ORC_ANGLE) return 2;                           \
    if (sp->rule == ORC_REF && (n & 63)) return 5;                             \
    if (len < 2) return 3;                                                     \
    long long nl = (long long)n;                                               \
    _Pragma("omp parallel for schedule(static)")                               \
    for (long long line = 0; line < nl; ++line) {                              \
        T px = x[line], py = y[line], pz = z[line];                            \
        /* SSE path: prevAccum = seed point, never updated (207-351) */        \
        const T Px = px, Py = py, Pz = pz;                                     \
        T upx = 0, upy = 0, upz = 0, hprev = (T)sp->h0;                        \
        int alive = 1;                                                         \
        uint32_t taken = 0;                                                    \
        for (size_t step = 1; step < len; ++step) {                            \
            T ex, ey, ez;                                                      \
            if (!alive) goto store_##SUF;                                      \
            if (sp->rule == ORC_PLAIN) {                                       \
                /* CPU_Implement.cpp:79-92 */                                  \
                field_sum_##SUF(charges, p, 1, 0, px, py, pz, &ex, &ey, &ez);  \
                T len_e = SQRT(ex * ex + ey * ey + ez * ez);                   \
                T sc = res * len_e;                                            \
                px = px + ex / sc; py = py + ey / sc; pz = pz + ez / sc;       \
            } else if (sp->rule == ORC_REF) {                                  \
                /* CPU_Implement.cpp:271-337 */                                \
                field_sum_##SUF(charges, p, 0, (T)ORC_KE, px, py, pz,          \
                                &ex, &ey, &ez);                                \
                T k = ex * ex + ey * ey + ez * ez;                             \
                T ix = ex - Px, iy = ey - Py, iz = ez - Pz;                    \
                T cx = iy * Pz - iz * Py;                                      \
                T cy = iz * Px - ix * Pz;                                      \
                T cz = ix * Py - iy * Px;                                      \
                k = SQRT(cx * cx + cy * cy + cz * cz) / (k * SQRT(k));         \
                T sc = (k + (T)1) * res;                                       \
                sc = sc * SQRT(ex * ex + ey * ey + ez * ez);                   \
                px = px + ex / sc; py = py + ey / sc; pz = pz + ez / sc;       \
            } else {                                                           \
                field_sum_##SUF(charges, p, 0, (T)1, px, py, pz,               \
                                &ex, &ey, &ez);                                \
                T len_e = SQRT(ex * ex + ey * ey + ez * ez);                   \
                T ux = ex / len_e, uy = ey / len_e, uz = ez / len_e;           \
                if (sp->stop && step > 1 &&                                    \
                    ux * upx + uy * upy + uz * upz < (T)0) {                   \
                    alive = 0;                                                 \
                    goto store_##SUF;                                          \
                }                                                              \
                T h = (T)sp->h0;                                               \
                if (step > 1) {                                                \
                    T dx = ux - upx, dy = uy - upy, dz = uz - upz;             \
                    T c = SQRT(dx * dx + dy * dy + dz * dz);                   \
                    T th = (T)sp->turn * hprev;                                \
                    h = (c * (T)sp->hmax > th) ? th / c : (T)sp->hmax;         \
                    h = h > (T)sp->hmin ? h : (T)sp->hmin;                     \
                }                                                              \
                px = px + h * ux; py = py + h * uy; pz = pz + h * uz;          \
                upx = ux; upy = uy; upz = uz; hprev = h;                       \
            }                                                                  \
            taken = (uint32_t)step;                                            \
        store_##SUF:                                                           \
            x[step * n + line] = px;                                           \
            y[step * n + line] = py;                                           \
            z[step * n + line] = pz;                                           \
        }                                                                      \
        if (steps) steps[line] = taken;                                        \
    }                                                                          \
    return 0;                                                                  \
}                                                                              \
                                                                               \
/* E at m points, reference arithmetic (3-arg PartField, scale k). */         \
void orc_field_##SUF(const T* charges, size_t p, const T* pts, size_t m,      \
                     T k, T* E)                                                \
{                                                                              \
    long long ml = (long long)m;                                               \
    _Pragma("omp parallel for schedule(static)")                               \
    for (long long i = 0; i < ml; ++i)                                         \
        field_sum_##SUF(charges, p, 0, k, pts[3 * i], pts[3 * i + 1],          \
                        pts[3 * i + 2], &E[3 * i], &E[3 * i + 1],              \
                        &E[3 * i + 2]);                                        \
}

ORC_DEFINE(float, f32, sqrtf)
ORC_DEFINE(double, f64, sqrt)
