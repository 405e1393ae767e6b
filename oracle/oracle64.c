/*
 * oracle/oracle64.c -- TEST INFRASTRUCTURE ONLY (the FP64 oracle).
 *
 * Double-precision statements of the physics the tracer computes, used by
 * tests/ as the checker for the HIP kernels.  Nothing in the product library
 * links, calls or loads this file.
 *
 *   o64_field      E(x) = s * sum_i q_i (x - c_i)/|x - c_i|^3, Neumaier-
 *                  compensated double sum (the Coulomb sum of
 *                  include/Electrostatics.h:38-60 without the FP32 rounding).
 *   o64_field_jvp  (dE/dx) w = s * sum_i q_i (w/r^3 - 3 d (d.w)/r^5), the
 *                  field-gradient tensor applied to w (closed form of the
 *                  derivative of the expression above).
 *   o64_trace      the plain (CPU_Implement.cpp:38-102) and angle step rules
 *                  in double with the physics-form field.
 *   o64_trace_jitter the same with float32-sized relative perturbations of
 *                  the field and the new point at every step (deterministic
 *                  signs): the parity tests' measure of how far a line
 *                  amplifies float32 rounding.  Pinned by: no perturbation
 *                  -> o64_trace bit for bit; a single charge's straight lines
 *                  stay within the injected size; the spread is linear in it.
 *   o64_min_dist_f32 / o64_path_dist_f32  the parity metrics (nearest
 *                  charge; distance to a reference polyline), pinned by
 *                  brute force and known geometry.
 *   o64_adjoint    reverse-mode derivative of sum_s G_s . x_s through
 *                  o64_trace, w.r.t. every charge (x, y, z, q) and every seed.
 *                  No reference counterpart (the reference has no adjoint); it
 *                  is pinned by central finite differences of o64_trace in
 *                  tests/test_oracle.py.
 *   o64_verlet_step the particle step of Particle_System.cpp:108-190 (new
 *                  accelerations, optionally the written-back state), pinned
 *                  by closed forms (Coulomb and Biot-Savart pairs, box
 *                  reflection, Newton's third law) in tests/test_oracle.py.
 *
 * Parity pins: single-charge closed form (|E| = |q|/r^2, radial), linearity
 * in q, the divergence of J being zero away from charges (trace of the
 * gradient tensor), and finite differences for the JVP and the adjoint.
 */
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

enum { O64_PLAIN = 0, O64_REF = 1, O64_ANGLE = 2 };

typedef struct o64_step_params {
    int rule;
    double resolution, h0, hmin, hmax, turn;
    int stop; /* angle rule: a line ends at the first step whose direction reverses */
} o64_step_params;

static int reversed(const o64_step_params* sp, int step, const double* u, const double* up)
{
    return sp->stop && step > 1 && u[0] * up[0] + u[1] * up[1] + u[2] * up[2] < 0;
}

static inline void neu_add(double* s, double* c, double v)
{
    double t = *s + v;
    if (fabs(*s) >= fabs(v))
        *c += (*s - t) + v;
    else
        *c += (v - t) + *s;
    *s = t;
}

static void field64(const double* ch, size_t p, double scale, const double* x,
                    double* E)
{
    double s[3] = {0, 0, 0}, c[3] = {0, 0, 0};
    for (size_t i = 0; i < p; ++i) {
        const double* q = ch + 4 * i;
        double d[3] = {x[0] - q[0], x[1] - q[1], x[2] - q[2]};
        double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        double i3 = q[3] / (r2 * sqrt(r2));
        for (int k = 0; k < 3; ++k) neu_add(&s[k], &c[k], i3 * d[k]);
    }
    for (int k = 0; k < 3; ++k) E[k] = scale * (s[k] + c[k]);
}

static void jvp64(const double* ch, size_t p, double scale, const double* x,
                  const double* w, double* out)
{
    double s[3] = {0, 0, 0}, c[3] = {0, 0, 0};
    for (size_t i = 0; i < p; ++i) {
        const double* q = ch + 4 * i;
        double d[3] = {x[0] - q[0], x[1] - q[1], x[2] - q[2]};
        double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        double r = sqrt(r2);
        double i3 = 1.0 / (r2 * r), i5 = i3 / r2;
        double a = d[0] * w[0] + d[1] * w[1] + d[2] * w[2];
        for (int k = 0; k < 3; ++k)
            neu_add(&s[k], &c[k], q[3] * (w[k] * i3 - 3.0 * d[k] * a * i5));
    }
    for (int k = 0; k < 3; ++k) out[k] = scale * (s[k] + c[k]);
}

void o64_field(const double* ch, size_t p, const double* pts, size_t m,
               double scale, double* E)
{
    long long ml = (long long)m;
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < ml; ++i) field64(ch, p, scale, pts + 3 * i, E + 3 * i);
}

void o64_field_jvp(const double* ch, size_t p, const double* pts, size_t m,
                   double scale, const double* w, double* out)
{
    long long ml = (long long)m;
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < ml; ++i)
        jvp64(ch, p, scale, pts + 3 * i, w + 3 * i, out + 3 * i);
}

/* One step of the rule.  Returns h; writes u and the rule's branch flag
 * (0: h from the turn ratio, 1: clamped to hmax, 2: clamped to hmin,
 *  3: constant step). */
static double step_rule(const o64_step_params* sp, int step, const double* E,
                        const double* up, double hprev, double* u,
                        double* L, double* c, int* branch)
{
    double l = sqrt(E[0] * E[0] + E[1] * E[1] + E[2] * E[2]);
    *L = l;
    for (int k = 0; k < 3; ++k) u[k] = E[k] / l;
    *c = 0;
    if (sp->rule != O64_ANGLE) {
        *branch = 3;
        return 1.0 / sp->resolution;
    }
    if (step == 1) {
        *branch = 3;
        return sp->h0;
    }
    double d[3] = {u[0] - up[0], u[1] - up[1], u[2] - up[2]};
    double cc = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    *c = cc;
    double th = sp->turn * hprev, h;
    if (cc * sp->hmax > th) {
        h = th / cc;
        *branch = 0;
    } else {
        h = sp->hmax;
        *branch = 1;
    }
    if (!(h > sp->hmin)) {
        h = sp->hmin;
        *branch = 2;
    }
    return h;
}

/* Relative perturbation +-rel, its sign from a SplitMix64 hash of (key, line,
 * step, slot): the same draws on every run and thread count. */
static double jitter(unsigned long long key, long long line, size_t step, int slot, double rel)
{
    unsigned long long z = key * 0x9E3779B97F4A7C15ull ^ ((unsigned long long)line << 20) ^
                           ((unsigned long long)step << 3) ^ (unsigned long long)slot;
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return 1.0 + ((z >> 63) ? rel : -rel);
}

/* rel > 0: every field component and every new coordinate is scaled by
 * 1 +- rel at every step (signs from `key`) -- rounding of the size a
 * float32 evaluation makes, injected into the FP64 path. */
static int trace_impl(const double* ch, size_t p, const double* seeds, size_t n, size_t len,
                      const o64_step_params* sp, double scale, double* rows, unsigned* steps,
                      double rel, unsigned long long key)
{
    if (!n) return 1;
    if (sp->rule == O64_REF) return 4; /* the REF rule is restated in restate.c */
    if (sp->rule == O64_PLAIN && sp->resolution == 0) return 2;
    if (len < 2) return 3;
    long long nl = (long long)n;
#pragma omp parallel for schedule(static)
    for (long long line = 0; line < nl; ++line) {
        double x[3] = {seeds[3 * line], seeds[3 * line + 1], seeds[3 * line + 2]};
        double up[3] = {0, 0, 0}, hprev = sp->h0;
        int alive = 1;
        unsigned taken = 0;
        for (int k = 0; k < 3; ++k) rows[3 * line + k] = x[k];
        for (size_t s = 1; s < len; ++s) {
            if (alive) {
                double E[3], u[3], L, c;
                int br;
                field64(ch, p, scale, x, E);
                if (rel > 0)
                    for (int k = 0; k < 3; ++k) E[k] *= jitter(key, line, s, k, rel);
                double h = step_rule(sp, (int)s, E, up, hprev, u, &L, &c, &br);
                if (reversed(sp, (int)s, u, up)) {
                    alive = 0;
                } else {
                    for (int k = 0; k < 3; ++k) {
                        x[k] = x[k] + h * u[k];
                        if (rel > 0) x[k] *= jitter(key, line, s, 3 + k, rel);
                        up[k] = u[k];
                    }
                    hprev = h;
                    taken = (unsigned)s;
                }
            }
            for (int k = 0; k < 3; ++k) rows[(s * n + line) * 3 + k] = x[k];
        }
        if (steps) steps[line] = taken;
    }
    return 0;
}

/* rows: len*n*3, step-major, rows[(s*n + line)*3 + k]; row 0 = seeds. */
int o64_trace(const double* ch, size_t p, const double* seeds, size_t n,
              size_t len, const o64_step_params* sp, double scale, double* rows,
              unsigned* steps)
{
    return trace_impl(ch, p, seeds, n, len, sp, scale, rows, steps, 0.0, 0ull);
}

/* The same FP64 trace with float32-sized rounding injected at every step (see
 * trace_impl): the spread of such paths around o64_trace's measures how far
 * a line amplifies the rounding any float32 implementation makes. */
int o64_trace_jitter(const double* ch, size_t p, const double* seeds, size_t n,
                     size_t len, const o64_step_params* sp, double scale, double* rows,
                     unsigned* steps, double rel, unsigned long long key)
{
    return trace_impl(ch, p, seeds, n, len, sp, scale, rows, steps, rel, key);
}

/* Reverse-mode derivative of  L = sum_{s=0}^{len-1} G_s . x_s  (G: len*n*3,
 * same layout as rows) w.r.t. charges (gch: p*4 = d/dx, d/dy, d/dz, d/dq) and
 * seeds (gseed: n*3). */
static int adjoint_impl(const double* ch, size_t p, const double* seeds, size_t n,
                        size_t len, const o64_step_params* sp, double scale,
                        const double* G, double* gch, double* gseed,
                        const float* path, const unsigned* path_steps)
{
    if (!n) return 1;
    if (sp->rule == O64_REF) return 4;
    if (sp->rule == O64_PLAIN && sp->resolution == 0) return 2;
    if (len < 2) return 3;
    memset(gch, 0, sizeof(double) * 4 * p);
    long long nl = (long long)n;
#pragma omp parallel
    {
        double* acc = (double*)calloc(4 * p, sizeof(double));
        /* per-step forward record */
        double* X = (double*)malloc(sizeof(double) * 3 * len); /* x_{s-1} */
        double* U = (double*)malloc(sizeof(double) * 3 * len);
        double* Lr = (double*)malloc(sizeof(double) * len);
        double* Hs = (double*)malloc(sizeof(double) * len);
        double* Cs = (double*)malloc(sizeof(double) * len);
        int* Br = (int*)malloc(sizeof(int) * len);
#pragma omp for schedule(static)
        for (long long line = 0; line < nl; ++line) {
            double x[3] = {seeds[3 * line], seeds[3 * line + 1], seeds[3 * line + 2]};
            double up[3] = {0, 0, 0}, hprev = sp->h0;
            int alive = 1;
            for (size_t s = 1; s < len; ++s) {
                double E[3], u[3];
                if (path) /* linearise along the given trajectory */
                    for (int k = 0; k < 3; ++k) x[k] = path[((s - 1) * n + line) * 3 + k];
                for (int k = 0; k < 3; ++k) X[3 * s + k] = x[k];
                Br[s] = 4; /* ended: x_s = x_{s-1} */
                if (path && s > path_steps[line]) alive = 0;
                if (!alive) continue;
                field64(ch, p, scale, x, E);
                double h = step_rule(sp, (int)s, E, up, hprev, u, &Lr[s], &Cs[s], &Br[s]);
                if (reversed(sp, (int)s, u, up) && !path) {
                    alive = 0;
                    Br[s] = 4;
                    continue;
                }
                Hs[s] = h;
                for (int k = 0; k < 3; ++k) {
                    U[3 * s + k] = u[k];
                    x[k] = x[k] + h * u[k];
                    up[k] = u[k];
                }
                hprev = h;
            }
            double Xb[3] = {0, 0, 0}, Ub[3] = {0, 0, 0}, Hb = 0;
            for (size_t s = len - 1; s >= 1; --s) {
                const double* u = U + 3 * s;
                for (int k = 0; k < 3; ++k) Xb[k] += G[(s * n + line) * 3 + k];
                if (Br[s] == 4) { /* the line had ended: identity, no field dependence */
                    Ub[0] = Ub[1] = Ub[2] = 0;
                    Hb = 0;
                    continue;
                }
                Hb += Xb[0] * u[0] + Xb[1] * u[1] + Xb[2] * u[2];
                for (int k = 0; k < 3; ++k) Ub[k] += Hs[s] * Xb[k];
                double Upb[3] = {0, 0, 0}, hpb = 0;
                if (sp->rule == O64_ANGLE && s >= 2 && Br[s] == 0) {
                    const double* u0 = U + 3 * (s - 1);
                    double cb = -Hb * Hs[s] / Cs[s];
                    hpb = Hb * sp->turn / Cs[s];
                    for (int k = 0; k < 3; ++k) {
                        double dv = (u[k] - u0[k]) / Cs[s];
                        Ub[k] += cb * dv;
                        Upb[k] = -cb * dv;
                    }
                }
                double ud = u[0] * Ub[0] + u[1] * Ub[1] + u[2] * Ub[2];
                double Eb[3];
                for (int k = 0; k < 3; ++k) Eb[k] = (Ub[k] - u[k] * ud) / Lr[s];
                const double* xp = X + 3 * s;
                double jv[3];
                jvp64(ch, p, scale, xp, Eb, jv);
                for (int k = 0; k < 3; ++k) Xb[k] += jv[k];
                /* charge gradients: dE/dq_i = s d/r^3,
                 * dE/dc_i = -s q_i (I/r^3 - 3 d d^T / r^5) */
                for (size_t i = 0; i < p; ++i) {
                    const double* q = ch + 4 * i;
                    double d[3] = {xp[0] - q[0], xp[1] - q[1], xp[2] - q[2]};
                    double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                    double r = sqrt(r2);
                    double i3 = 1.0 / (r2 * r), i5 = i3 / r2;
                    double a = d[0] * Eb[0] + d[1] * Eb[1] + d[2] * Eb[2];
                    acc[4 * i + 3] += scale * a * i3;
                    for (int k = 0; k < 3; ++k)
                        acc[4 * i + k] -= scale * q[3] * (Eb[k] * i3 - 3.0 * d[k] * a * i5);
                }
                for (int k = 0; k < 3; ++k) Ub[k] = Upb[k];
                Hb = hpb;
            }
            for (int k = 0; k < 3; ++k) gseed[3 * line + k] = Xb[k] + G[3 * line + k];
        }
#pragma omp critical
        for (size_t i = 0; i < 4 * p; ++i) gch[i] += acc[i];
        free(acc); free(X); free(U); free(Lr); free(Hs); free(Cs); free(Br);
    }
    return 0;
}

/* Distance from each of m points (float32 xyz, stride 3) to the nearest of
 * p charges: the "stays more than r from any charge" filter of the parity
 * tests.  Plain brute force in double. */
void o64_min_dist_f32(const float* ch, size_t p, const float* pts, size_t m, float* out)
{
    long long ml = (long long)m;
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < ml; ++i) {
        double best = INFINITY;
        for (size_t j = 0; j < p; ++j) {
            double dx = (double)pts[3 * i] - ch[4 * j];
            double dy = (double)pts[3 * i + 1] - ch[4 * j + 1];
            double dz = (double)pts[3 * i + 2] - ch[4 * j + 2];
            double d = dx * dx + dy * dy + dz * dz;
            if (d < best) best = d;
        }
        out[i] = (float)sqrt(best);
    }
}

/* Path distance: for every point j < first[l] of line l in `test` (SoA rows
 * [3][len][n], float32), the distance to the polyline through the points of
 * the same line in `ref` (same layout) up to and including point first[l]
 * (when it exists and is finite: the reference path continues there, so a
 * test point a little further along the line than the last compared
 * reference point still finds its segment), searching segments k in
 * [j-W, j+W].  Measures how far a traced point lies from the reference path,
 * independent of where along the path it falls.  out: [len][n] (0 beyond
 * first[l]). */
void o64_path_dist_f32(const float* test, const float* ref, size_t len, size_t n,
                       const long long* first, long long window, float* out)
{
    long long nl = (long long)n;
    const size_t plane = len * n;
#pragma omp parallel for schedule(dynamic, 16)
    for (long long l = 0; l < nl; ++l) {
        long long fl = first[l] < (long long)len ? first[l] : (long long)len;
        /* last reference point of the polyline */
        long long last = fl - 1;
        if (fl < (long long)len && fabsf(ref[fl * n + l]) < 1e29f &&
            fabsf(ref[plane + fl * n + l]) < 1e29f && fabsf(ref[2 * plane + fl * n + l]) < 1e29f)
            last = fl;
        for (long long j = 0; j < (long long)len; ++j) {
            if (j >= fl) { out[j * n + l] = 0.f; continue; }
            double px = test[j * n + l], py = test[plane + j * n + l], pz = test[2 * plane + j * n + l];
            long long k0 = j - window < 0 ? 0 : j - window;
            long long k1 = j + window < last ? j + window : last;
            double best = INFINITY;
            if (k1 <= k0) {
                double dx = px - ref[j * n + l], dy = py - ref[plane + j * n + l],
                       dz = pz - ref[2 * plane + j * n + l];
                best = dx * dx + dy * dy + dz * dz;
            }
            for (long long k = k0; k < k1; ++k) {
                double ax = ref[k * n + l], ay = ref[plane + k * n + l], az = ref[2 * plane + k * n + l];
                double bx = ref[(k + 1) * n + l] - ax, by = ref[plane + (k + 1) * n + l] - ay,
                       bz = ref[2 * plane + (k + 1) * n + l] - az;
                double wx = px - ax, wy = py - ay, wz = pz - az;
                double bb = bx * bx + by * by + bz * bz;
                double t = bb > 0 ? (wx * bx + wy * by + wz * bz) / bb : 0.0;
                t = t < 0 ? 0 : (t > 1 ? 1 : t);
                double dx = wx - t * bx, dy = wy - t * by, dz = wz - t * bz;
                double d = dx * dx + dy * dy + dz * dz;
                if (d < best) best = d;
            }
            out[j * n + l] = (float)sqrt(best);
        }
    }
}

int o64_adjoint(const double* ch, size_t p, const double* seeds, size_t n, size_t len,
                const o64_step_params* sp, double scale, const double* G, double* gch,
                double* gseed)
{
    return adjoint_impl(ch, p, seeds, n, len, sp, scale, G, gch, gseed, NULL, NULL);
}

/* The same reverse sweep, linearised along a given float trajectory (points
 * [len][n][3], e.g. the GPU's rows) with its per-line step counts: the FP64
 * derivative of the discrete path that trajectory took.  Isolates the
 * arithmetic error of a float adjoint from differences between float and
 * double trajectories. */
int o64_adjoint_path(const double* ch, size_t p, const double* seeds, size_t n, size_t len,
                     const o64_step_params* sp, double scale, const double* G, double* gch,
                     double* gseed, const float* path, const unsigned* path_steps)
{
    return adjoint_impl(ch, p, seeds, n, len, sp, scale, G, gch, gseed, path, path_steps);
}

/* ---------------------------------------------------------------------------
 * Particle step, FP64 -- ElectroMag/src/Particle_System.cpp:108-190 read as its
 * evident algorithm (the reference never instantiates it, and as written it
 * does not compile: magneto_k is "1E-7;" and there is no scalar * vector or
 * vector /= scalar operator).  Per particle i, the others at their old state:
 *   v' = v + a dt/2;  x' = x + v' dt                         (134-140)
 *   E  = sum_static r k_e q/(|r|^2 sqrt|r|^2)                 (143-148, Electrostatics.h:53-60)
 *      + sum_{j != i} (k_e q_j) r/(|r|^2 sqrt|r|^2)            (PartFieldOp, Electrostatics.h:66-75)
 *   B  = sum_{j != i} (q_j 1e-7) v_j x r/(|r|^2 sqrt|r|^2)     (Magnetics.h PartFieldOp)
 *   a  = (q E + q (v' x B)) / m                              (173-178, Magnetics.h Force)
 *   accel[i] = a;  with `out`: v' + a dt/2 and x', reflected into the box
 *   (BoundToBox, 19-56), are written to out[i] (the reference discards them).
 * Operation order is that of the reference's vec3 helpers.
 * ------------------------------------------------------------------------- */
void o64_verlet_step(const double* parts, size_t n, const double* charges, size_t p,
                     double* accel, const double* bmin, const double* bmax, double dt,
                     double* out)
{
    const double ke = 8.987551787E9, mk = 1E-7, dt2 = dt / 2;
    long long nl = (long long)n;
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < nl; ++i) {
        const double* P = parts + 8 * i;
        double vx = P[4], vy = P[5], vz = P[6];
        const double q = P[3], m = P[7];
        vx = vx + accel[3 * i] * dt2;
        vy = vy + accel[3 * i + 1] * dt2;
        vz = vz + accel[3 * i + 2] * dt2;
        double x = P[0] + vx * dt, y = P[1] + vy * dt, z = P[2] + vz * dt;
        double ex = 0, ey = 0, ez = 0, bx = 0, by = 0, bz = 0;
        for (size_t j = 0; j < p; ++j) {
            const double* c = charges + 4 * j;
            double rx = x - c[0], ry = y - c[1], rz = z - c[2];
            double l2 = rx * rx + ry * ry + rz * rz;
            double s = ke * c[3] / (l2 * sqrt(l2));
            ex += rx * s; ey += ry * s; ez += rz * s;
        }
        for (size_t j = 0; j < n; ++j) {
            if ((long long)j == i) continue;
            const double* S = parts + 8 * j;
            double rx = x - S[0], ry = y - S[1], rz = z - S[2];
            double l2 = rx * rx + ry * ry + rz * rz;
            double d = l2 * sqrt(l2);
            double ux = rx / d, uy = ry / d, uz = rz / d;     /* vec3InverseSquare */
            double kq = ke * S[3];
            ex += kq * ux; ey += kq * uy; ez += kq * uz;
            double qm = S[3] * mk;                          /* q * magneto_k * cross(v, u) */
            bx += qm * (S[5] * uz - S[6] * uy);
            by += qm * (S[6] * ux - S[4] * uz);
            bz += qm * (S[4] * uy - S[5] * ux);
        }
        double ax = q * ex + q * (vy * bz - vz * by);
        double ay = q * ey + q * (vz * bx - vx * bz);
        double az = q * ez + q * (vx * by - vy * bx);
        ax = ax / m; ay = ay / m; az = az / m;
        accel[3 * i] = ax; accel[3 * i + 1] = ay; accel[3 * i + 2] = az;
        if (out) {
            double wx = vx + ax * dt2, wy = vy + ay * dt2, wz = vz + az * dt2;
            if (x > bmax[0]) { x = 2 * bmax[0] - x; wx = -wx; }
            if (y > bmax[1]) { y = 2 * bmax[1] - y; wy = -wy; }
            if (z > bmax[2]) { z = 2 * bmax[2] - z; wz = -wz; }
            if (x < bmin[0]) { x = 2 * bmin[0] - x; wx = -wx; }
            if (y < bmin[1]) { y = 2 * bmin[1] - y; wy = -wy; }
            if (z < bmin[2]) { z = 2 * bmin[2] - z; wz = -wz; }
            double* O = out + 8 * i;
            O[0] = x; O[1] = y; O[2] = z; O[3] = q; O[4] = wx; O[5] = wy; O[6] = wz; O[7] = m;
        }
    }
}
