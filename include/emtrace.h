/*
 * emtrace.h -- C ABI of the MI355X (gfx950) field-line tracer.
 *
 * Plain C: plain pointers and sizes, no C++ or torch types, no exceptions
 * across the boundary.  Every entry point returns an int status (0 = success,
 * codes below) and leaves a thread-local message for em_last_error().
 *
 * Data model (same as the reference's calc-lines interface):
 *   charges : p records {x, y, z, q} of float32, 16 bytes each -- the layout
 *             of electro::pointCharge<float> (include/Electrostatics.h:30-35).
 *   lines   : three float32 arrays x, y, z, step-major: element (line, step)
 *             at [step * pitch + line]; row 0 holds the seeds (input), rows
 *             1..len-1 are written (ElectroMag/src/CPU_Implement.cpp:80,92,
 *             334-337; include/SOA_utils.hpp:31-112).  pitch == n in the
 *             reference; the device forms accept pitch >= n.
 */
#ifndef EMTRACE_H
#define EMTRACE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
/* 1..3 keep the reference's meaning (CPU_Implement.cpp:43-52, 213-230). */
#define EM_OK 0
#define EM_ERR_NO_LINES 1      /* n == 0                                     */
#define EM_ERR_RESOLUTION 2    /* resolution == 0 (plain / ref rules)        */
#define EM_ERR_LENGTH 3        /* fewer than 2 rows per line                 */
/* 16 and up are new: the reference reports device failures only as text.   */
#define EM_ERR_INVALID 16      /* null pointer, bad size or parameter        */
#define EM_ERR_HIP 17          /* HIP runtime / launch failure               */
#define EM_ERR_NOMEM 18        /* device allocation failed                   */
#define EM_ERR_NO_DEVICE 19    /* no usable gfx950 device                    */
#define EM_ERR_UNSUPPORTED 20  /* combination not implemented (see message)  */

/* ---- step rules -------------------------------------------------------- */
#define EM_RULE_PLAIN 0 /* x += E/(|E| res)               CPU_Implement.cpp:89-90   */
#define EM_RULE_REF 1   /* x += E/(|E| (k+1) res), k = |(E-P)xP|/|E|^3, P = seed,
                           E with k_e                      CPU_Implement.cpp:322-331 */
#define EM_RULE_ANGLE 2 /* h = clamp(turn*h_prev/|u - u_prev|, hmin, hmax), h_1 = h0
                           (project-defined; no reference counterpart)              */

typedef struct em_step {
    int32_t rule;
    float resolution; /* plain / ref: step length 1/resolution          */
    float h0;         /* angle: first step                              */
    float hmin, hmax; /* angle: clamp                                   */
    float turn;       /* angle: target chord between successive directions */
    int32_t stop;     /* angle only: 1 = a line ends at the first step whose direction
                         reverses (u_s . u_{s-1} < 0: it passed through a sink); its
                         remaining rows repeat the end point.  0 = always len-1 steps. */
} em_step;

/* Launch shape.  Any field left 0 is chosen by the library from the problem
 * size and the device's CU count (the Python autotuner passes all four). */
typedef struct em_launch {
    int32_t lanes_per_line; /* 1, 2, 4, 8, 16 or 32 lanes split one line's charge sum */
    int32_t tile;           /* charges per LDS tile: a multiple of 256 and of
                               16 x lanes_per_line; tile x 16 B plus 512 B of static
                               LDS must fit a workgroup (at most 9,984 on gfx950)   */
    int32_t block;          /* threads per workgroup: 64, 128, 256 or 512             */
    int32_t unroll;         /* charge records per lane per pipelined group: 2, 4 or 8;
                               any value gives the same bits (speed only)             */
} em_launch;

/* Timing/performance record -- the fields of the reference's perfPacket
 * (include/Data Structures.h:221-236: performance [GFLOP/s], time [s],
 * progress) plus the per-phase times of CL_Electrostatics.cpp:472-511. */
typedef struct em_perf {
    double time_s;             /* whole call, wall clock                       */
    double kernel_s;           /* HIP-event time of the tracing kernel         */
    double h2d_s, d2h_s;       /* host<->device copies (host-array forms)      */
    double h2d_bytes, d2h_bytes;
    double interactions;       /* p * n * (len-1) charge-point interactions    */
    double interactions_per_s; /* interactions / kernel_s                      */
    double line_steps_per_s;   /* n * (len-1) / kernel_s                       */
    double gflops;             /* n*(len-1)*(p*18+13)/time_s/1e9, the reference
                                  formula (CPU_Implement.cpp:347-349)          */
    volatile double progress;  /* 0 at entry, 1.0 when the rows are on the host */
    int32_t lanes_per_line, tile, block, unroll; /* launch shape actually used */
    volatile double* progress_out; /* optional (may be NULL): every update of
                                  progress is also stored here while the call
                                  runs -- e.g. &perfPacket.progress, which the
                                  reference's monitor thread polls
                                  (Electromag utils.h:133-156)            */
} em_perf;

/* ---- library lifetime ----------------------------------------------------*/
int em_init(int device);           /* optional; entry points init lazily  */
int em_shutdown(void);             /* frees cached device buffers          */
const char* em_last_error(void);   /* thread-local, never NULL             */
int em_device_count(void);         /* HIP devices visible (0 without GPU)  */
const char* em_version(void);

/* ---- drop-in for CalcField_CPU<float> ------------------------------------
 * Replaces: template<> int CalcField_CPU<float>(Vector3<Array<float>>&,
 *   Array<pointCharge<float>>&, size_t n, float resolution, perfPacket&,
 *   bool useCurvature)   (ElectroMag/src/CPU Implement.h:25-30,
 *   CPU_Implement.cpp:501-513).  Host arrays, synchronous; x/y/z hold len*n
 * floats with row 0 = seeds.  use_curvature selects EM_RULE_REF, else
 * EM_RULE_PLAIN.  Unlike the SSE path it does not require n % 64 == 0. */
int em_calc_lines_f32(const float* charges, size_t p, float* x, float* y,
                      float* z, size_t n, size_t len, float resolution,
                      int use_curvature, em_perf* perf);

/* Host arrays, any step rule and launch shape (launch may be NULL).
 * steps (may be NULL): n counts of the steps each line took (len-1 unless it
 * stopped). */
int em_trace_host_f32(const float* charges, size_t p, float* x, float* y,
                      float* z, size_t n, size_t len, const em_step* step,
                      const em_launch* launch, em_perf* perf, uint32_t* steps);

/* ---- FP64 ------------------------------------------------------------------
 * Replaces: template<> int CalcField_CPU<double>(...) (CPU Implement.h:25-30,
 *   CPU_Implement.cpp:366-498, 515-527).  charges: p records {x, y, z, q} of
 * float64 (pointCharge<double>); lines: float64 step-major as above.  Every
 * operation is IEEE binary64 in the reference's source order (charges summed
 * in index order, correctly rounded sqrt and division, no fused
 * multiply-add), so the rows equal the FP64 restatement of the reference
 * algorithm bit for bit.  Same status codes as the FP32 forms. */
int em_calc_lines_f64(const double* charges, size_t p, double* x, double* y,
                      double* z, size_t n, size_t len, double resolution,
                      int use_curvature, em_perf* perf);
/* Any step rule (parameters are the em_step floats, widened exactly). */
int em_trace_host_f64(const double* charges, size_t p, double* x, double* y,
                      double* z, size_t n, size_t len, const em_step* step,
                      em_perf* perf, uint32_t* steps);
/* Device pointers, asynchronous on `stream`; pitch >= n. */
int em_trace_dev_f64(const double* d_charges, size_t p, double* d_x,
                     double* d_y, double* d_z, size_t n, size_t len,
                     size_t pitch, const em_step* step, uint32_t* d_steps,
                     void* stream);

/* ---- device-pointer forms (asynchronous on `stream`, a hipStream_t) -------
 * Replaces the OpenCL functor's device path (GPGPU_Segment/src/
 * CL_Electrostatics.cpp:396-513 MainFunctor; Electrostatics.cl.c:282-367).
 * d_ckpt (may be NULL): when non-NULL, the forward also stores the state
 * needed by em_adjoint_dev_f32 every ckpt_every steps; size it with
 * em_ckpt_floats().  d_steps (may be NULL): n step counts.
 * The angle rule with stop, no checkpoints and p >= 4096 is traced in two
 * phases (lane groups that take a new line when theirs ends, then workgroups
 * that trace the lines still moving after 192 steps -- 128 for the last lines
 * handed out, one per lane group -- up to 8 at a time, or the lane groups
 * again when more than that many lines are still moving);
 * `launch` then sets the first phase's shape.
 * That path keeps per-device work memory (about 52 bytes per line plus a
 * padded copy of the charges), allocated on first use and grown by doubling;
 * outgrown buffers stay allocated, since a graph captured earlier may use
 * them, until em_shutdown.  Under stream capture it must already be large
 * enough (a call of the same or a larger size made outside capture), else
 * EM_ERR_INVALID.  Eager calls on one device are ordered among themselves; a
 * graph replay that overlaps another call on the same device races with it. */
int em_trace_dev_f32(const float* d_charges, size_t p, float* d_x, float* d_y,
                     float* d_z, size_t n, size_t len, size_t pitch,
                     const em_step* step, const em_launch* launch,
                     float* d_ckpt, uint32_t ckpt_every, uint32_t* d_steps,
                     void* stream);

size_t em_ckpt_floats(size_t n, size_t len, uint32_t ckpt_every);

/* ---- allocation-free form of em_trace_dev_f32 -------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory: code under stream/graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  Same arguments, same rows and step counts as
 * em_trace_dev_f32 (the same kernels), but the work memory of the line-ends
 * trace is a fixed array in the library's device code,
 * EM_NOALLOC_SCRATCH_BYTES per device, held from the first call on a device
 * (or em_init) until the process ends.  The call allocates nothing, creates
 * no stream or event, and never waits on the host.
 * Work memory needed by that path (angle rule with stop, no checkpoints,
 * p >= 4096), where a() rounds up to 256 and C is the device's CU count:
 *   4096 <= p <= 16384 (register-resident trace): 256 + 104 * 16 C bytes
 *     (425,984 + 256 on 256 CUs, whatever n), plus a(4 n) when d_steps is NULL;
 *   p > 16384 (lane-group trace): 256 + a(48 n) + 16 round_up(p, 256) bytes,
 *     plus a(4 n) when d_steps is NULL.
 * More than EM_NOALLOC_SCRATCH_BYTES -> EM_ERR_INVALID before any launch
 * (on 256 CUs: about 1,990,000 lines without a steps buffer up to 16,384
 * charges; about 170,000 lines with one above).  The other paths use no work
 * memory.
 * Every call on one device uses that device's one array: calls that overlap
 * on a device (on different streams, or a graph replay beside another call)
 * race with each other; calls on one stream are ordered. */
#define EM_NOALLOC_SCRATCH_BYTES (8u << 20)
int em_trace_dev_f32_noalloc(const float* d_charges, size_t p, float* d_x, float* d_y,
                             float* d_z, size_t n, size_t len, size_t pitch,
                             const em_step* step, const em_launch* launch,
                             float* d_ckpt, uint32_t ckpt_every, uint32_t* d_steps,
                             void* stream);

/* E(x) = scale * sum_i q_i (x - c_i)/|x - c_i|^3 at m points (SoA px/py/pz),
 * written to ex/ey/ez.  The same device code as the tracer's charge loop
 * (include/Electrostatics.h:38-60 summed over all charges). */
int em_field_dev_f32(const float* d_charges, size_t p, const float* d_px,
                     const float* d_py, const float* d_pz, size_t m,
                     float scale, float* d_ex, float* d_ey, float* d_ez,
                     const em_launch* launch, void* stream);

/* ---- batched form: many independent configurations, one persistent launch
 * No reference counterpart (the reference traces one configuration per call).
 * A plan sorts (configuration, line-block) work items by cost and sizes a
 * persistent grid; em_batch_run launches it (the first run of a captured plan
 * records a hipGraph, later runs replay it). */
typedef struct em_batch_cfg {
    uint64_t charge_offset; /* first charge record of this configuration     */
    uint64_t line_offset;   /* element offset of (row 0, line 0) in x/y/z    */
    uint64_t pitch;         /* row pitch in elements (>= n)                  */
    uint32_t p;             /* charges                                       */
    uint32_t n;             /* lines                                         */
    uint32_t len;           /* rows (>= 2)                                   */
    uint32_t reserved;
} em_batch_cfg;

typedef struct em_batch em_batch; /* opaque */

int em_batch_create(const em_batch_cfg* cfgs, size_t count, const em_step* step,
                    const em_launch* launch, int use_graph, em_batch** out);
int em_batch_run(em_batch* plan, const float* d_charges, float* d_x, float* d_y,
                 float* d_z, uint32_t* d_steps, void* stream);
int em_batch_info(const em_batch* plan, double* interactions, int32_t* items,
                  int32_t* grid, em_launch* used);
int em_batch_destroy(em_batch* plan);

/* ---- adjoint --------------------------------------------------------------
 * No reference counterpart.  Given the forward's rows and checkpoints
 * (d_ckpt from em_trace_dev_f32 with the same ckpt_every) and the loss
 * gradient G_s = dL/dx_s for every row (gx/gy/gz, same layout as x/y/z),
 * recomputes each checkpoint segment, sweeps it backward, and writes
 *   d_grad_charges : p x 4 floats, dL/d{x, y, z, q} of every charge,
 *   d_grad_seeds   : 3 x n floats (SoA), dL/d(seed point)  (may be NULL).
 * Rules: EM_RULE_PLAIN and EM_RULE_ANGLE (EM_RULE_REF -> EM_ERR_UNSUPPORTED).
 * Work lists and buffers are kept per device between calls: a call with the
 * same sizes, shape and ckpt_every reuses them asynchronously (waiting on
 * the device for the previous call when the stream differs); other calls
 * rebuild them and block the host until the previous call has finished.
 * Not for graph capture (use the batch form, em_batch_adjoint). */
int em_adjoint_dev_f32(const float* d_charges, size_t p, const float* d_x,
                       const float* d_y, const float* d_z, size_t n, size_t len,
                       size_t pitch, const em_step* step, const em_launch* launch,
                       const float* d_ckpt, uint32_t ckpt_every,
                       const float* d_gx, const float* d_gy, const float* d_gz,
                       float* d_grad_charges, float* d_grad_seeds, void* stream);

/* Batched adjoint over the configurations of a plan (charges, rows and
 * gradients packed as for em_batch_run; checkpoints from
 * em_batch_run_ckpt; grad arrays indexed like charges / lines).
 * The plan keeps device descriptors and work lists for one ckpt_every at a
 * time: a call with another ckpt_every rebuilds them, so a graph that
 * captured em_batch_run_ckpt / em_batch_adjoint of this plan must be captured
 * again after such a call.  Plain em_batch_run calls change nothing. */
size_t em_batch_ckpt_floats(const em_batch* plan, uint32_t ckpt_every);
int em_batch_run_ckpt(em_batch* plan, const float* d_charges, float* d_x,
                      float* d_y, float* d_z, float* d_ckpt,
                      uint32_t ckpt_every, uint32_t* d_steps, void* stream);
int em_batch_adjoint(em_batch* plan, const float* d_charges, const float* d_x,
                     const float* d_y, const float* d_z, const float* d_ckpt,
                     uint32_t ckpt_every, const float* d_gx, const float* d_gy,
                     const float* d_gz, float* d_grad_charges, void* stream);

/* ---- particle step (the N-body client of the Coulomb sum) ------------------
 * Replaces: emgVerletStep<float>(dynCharges, pointCharges, verletAccel,
 *   minBound, maxBound, timeStep, n, perfData)  (ElectroMag/src/
 *   Particle_System.cpp:108-190, never instantiated by the reference).
 * particles : n records of 8 floats {x, y, z, q, vx, vy, vz, m}
 *             (electro::dynamicPointCharge<float>, include/Electrodynamics.h)
 * charges   : p static records {x, y, z, q}
 * accel     : n x 3 floats, a(t) in, a(t + dt) out
 * For each particle: v' = v + a dt/2, x' = x + v' dt, then with the OTHER
 * particles at their old state
 *   a = (q E(x') + q v' x B(x')) / m,
 *   E = k_e [sum_static q_c r/|r|^3 + sum_j q_j r/|r|^3],
 *   B = 1e-7 sum_j q_j v_j x r/|r|^3,   r = x' - source.
 * The reference stores only a (it never writes the particle back).  With
 * write_back (host form) or a non-NULL d_particles_out (device form, must not
 * alias the input) the step also writes x' and v' + a dt/2, each reflected
 * into [box_min, box_max] as BoundToBox does (Particle_System.cpp:19-56).
 * box_min / box_max are host float[3].  Status: 1 when n == 0 (the
 * reference's only check), 16+ as elsewhere.
 * Each particle's sum over the sources is split over 1-8 lanes (as many as
 * bring the grid to four waves per SIMD); the environment variable
 * EMTRACE_VERLET_LANES = 1, 2, 4 or 8 overrides that count (any other value:
 * EM_ERR_INVALID).  The split changes the order of the FP32 sums, so the
 * last bits of a may differ between counts. */
int em_verlet_step_f32(float* particles, size_t n, const float* charges,
                       size_t p, float* accel, const float* box_min,
                       const float* box_max, float dt, int write_back,
                       em_perf* perf);
int em_verlet_step_dev_f32(const float* d_particles, size_t n,
                           const float* d_charges, size_t p, float* d_accel,
                           const float* box_min, const float* box_max,
                           float dt, float* d_particles_out, void* stream);

/* ---- the reference application's problem generator -----------------------
 * Host-only.  Restates ElectroMag/src/Electromag utils.h:29-111 (libc rand()
 * under srand) with a built-in copy of glibc's rand(), so the reference's
 * problems are reproduced bit for bit on any host.
 * em_legacy_rand: out[i] = the i-th rand() after srand(seed).
 * em_legacy_problem_f32: InitializePointChargeArray (p records {x,y,z,q},
 *   srand(charge_seed); the reference uses 1) followed by
 *   InitializeFieldLineArray into x/y/z (n = width*height*depth seeds): the
 *   grid 10*((-w/2 + 1e-5) + i), ... when random_lines == 0, else three
 *   rand() draws per line after srand(line_seed). */
int em_legacy_rand(uint32_t seed, int32_t* out, size_t count);
int em_legacy_problem_f32(float* charges, size_t p, uint32_t charge_seed,
                          float* x, float* y, float* z, size_t width,
                          size_t height, size_t depth, int random_lines,
                          uint32_t line_seed);

#ifdef __cplusplus
}
#endif

#endif /* EMTRACE_H */
