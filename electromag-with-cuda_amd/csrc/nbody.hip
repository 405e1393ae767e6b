// nbody.hip -- the particle step of the reference (emgVerletStep), on the
// tracer's LDS-tiled Coulomb sum.
//
// ElectroMag/src/Particle_System.cpp:108-190, per dynamic particle i, all
// particles reading the *old* state of the others (the reference copies
// particle i, never writes it back, and writes only the new acceleration):
//   v' = v_i + a_i dt/2                     (half kick)
//   x' = x_i + v' dt                        (drift)
//   E  = sum_static  k_e q_c (x' - c)/|x' - c|^3            (Electrostatics.h:53-60)
//      + sum_{j != i} k_e q_j (x' - x_j)/|x' - x_j|^3        (PartFieldOp, 66-75)
//   B  = sum_{j != i} 1e-7 q_j v_j x (x' - x_j)/|x' - x_j|^3 (Magnetics.h PartFieldOp)
//   a  = (q_i E + q_i v' x B) / m_i          (Force, Magnetics.h; 173-178)
// The reference writes a into verletAccel[i] and discards x', v'.  With an
// output array the step also writes the full velocity-Verlet state
// v'' = v' + a dt/2 and x', reflected into the box (BoundToBox, 19-56) -- the
// evident intent of the reference's TODO comments.
//
// Device layout: particles are 8-float records {x, y, z, q, vx, vy, vz, m}
// (electro::dynamicPointCharge<float>: pointCharge, velocity, mass); static
// charges are {x, y, z, q}; accelerations are n x 3 floats (Array<Vector3>).
// S lanes per particle (S = 1..8: until the grid holds four waves per SIMD;
// 65,536 particles among 16,384 charges: 3.03-3.05 ms at S = 4 with 2-record
// pipelined groups (4 records: 3.14-3.20 ms, 8: 3.41-3.46 ms), against
// 3.81-3.89 ms at S = 1, where one wave per SIMD leaves every latency
// exposed).  Static charges go through the tracer's pipelined tile walk
// (FieldBody).  Dynamic sources are staged as two float4 records,
// {x, y, z, q} and {1e-7 q v, index}; the index zeroes the self term.
// 21 VALU + 2 for the self mask + 1 transcendental per particle pair.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>

#include "em_internal.hpp"

namespace em {
namespace {

constexpr int kNbBlock = 256;
constexpr uint32_t kNbTile = 4096;          // float4 records per LDS tile (64 KiB)
constexpr uint32_t kNbTileSplit = 2048;     // with S > 1: 32 KiB, four workgroups per CU
constexpr int kNbMaxLanes = 8;
constexpr float kMagneto = 1e-7f;           // mu_0 / (4 pi) (Magnetics.h magneto_k)
constexpr float kCoulombF = 8.987551787E9f; // include/Electrostatics.h:27

struct VerletP {
    float dt, dt2;
    float bmin[3], bmax[3];
};

// One dynamic source: E += q u, B += w x u, u = (x - x_j)/|x - x_j|^3, the
// self term (index == self) masked to zero.
struct DynBody {
    float px, py, pz;
    int self;
    float ex, ey, ez, bx, by, bz;  // partials of the current block
    Acc e, b;                      // compensated totals
    __device__ void init(float x, float y, float z, int i)
    {
        px = x; py = y; pz = z; self = i;
        ex = ey = ez = bx = by = bz = 0.f;
        e.zero();
        b.zero();
    }
    __device__ __forceinline__ void term(const float4* r)
    {
        const float dx = px - r[0].x, dy = py - r[0].y, dz = pz - r[0].z;
        const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, __builtin_fmaf(dx, dx, kSoft2)));
        const float inv = __builtin_amdgcn_rsqf(r2);
        float inv3 = inv * (inv * inv);
        inv3 = (__float_as_int(r[1].w) == self) ? 0.f : inv3;
        const float ux = dx * inv3, uy = dy * inv3, uz = dz * inv3;
        ex = __builtin_fmaf(r[0].w, ux, ex);
        ey = __builtin_fmaf(r[0].w, uy, ey);
        ez = __builtin_fmaf(r[0].w, uz, ez);
        // w x u with w = 1e-7 q v (Vector.h:190-197 cross order: index x middle)
        bx = __builtin_fmaf(r[1].y, uz, __builtin_fmaf(-r[1].z, uy, bx));
        by = __builtin_fmaf(r[1].z, ux, __builtin_fmaf(-r[1].x, uz, by));
        bz = __builtin_fmaf(r[1].x, uy, __builtin_fmaf(-r[1].y, ux, bz));
    }
    __device__ __forceinline__ void fold()
    {
        em::fold(e.sx, e.cx, ex);
        em::fold(e.sy, e.cy, ey);
        em::fold(e.sz, e.cz, ez);
        em::fold(b.sx, b.cx, bx);
        em::fold(b.sy, b.cy, by);
        em::fold(b.sz, b.cz, bz);
        ex = ey = ez = bx = by = bz = 0.f;
    }
};

__device__ __forceinline__ void stage_dyn(float4* __restrict__ tile,
                                          const float4* __restrict__ parts, uint32_t j0,
                                          uint32_t cnt, uint32_t cntp)
{
    for (uint32_t s = threadIdx.x; s < cntp; s += blockDim.x) {
        float4 r0 = make_float4(kFar, kFar, kFar, 0.f), r1 = make_float4(0.f, 0.f, 0.f, 0.f);
        if (s < cnt) {
            r0 = parts[2 * (j0 + s)];
            const float4 v = parts[2 * (j0 + s) + 1];
            const float qk = r0.w * kMagneto;
            r1 = make_float4(qk * v.x, qk * v.y, qk * v.z, __int_as_float((int)(j0 + s)));
        } else {
            r1.w = __int_as_float(-1);
        }
        tile[2 * s] = r0;
        tile[2 * s + 1] = r1;
    }
}

// S lanes per particle (lane k sums records k, k + S, ... of every tile, an
// xor butterfly combines the S partial sums); `tile` float4 records of LDS.
template <int S, int U>
__global__ __launch_bounds__(kNbBlock) void k_verlet_step(const float4* __restrict__ parts,
                                                          uint32_t n,
                                                          const float4* __restrict__ charges,
                                                          uint32_t p, float* __restrict__ accel,
                                                          VerletP vp, float4* __restrict__ out,
                                                          uint32_t tile)
{
    extern __shared__ float4 nb_tile[];
    const uint32_t i = (blockIdx.x * kNbBlock + threadIdx.x) / S;
    const int k = threadIdx.x % S;
    const bool valid = i < n;
    float4 c0 = make_float4(0.f, 0.f, 0.f, 0.f), c1 = make_float4(0.f, 0.f, 0.f, 1.f);
    float ax = 0.f, ay = 0.f, az = 0.f;
    if (valid) {
        c0 = parts[2 * i];
        c1 = parts[2 * i + 1];
        ax = accel[3 * i];
        ay = accel[3 * i + 1];
        az = accel[3 * i + 2];
    }
    // v(t + dt/2) = v(t) + a(t) dt/2; r(t + dt) = r(t) + v(t + dt/2) dt  (Particle_System.cpp:134-140)
    const float vx = c1.x + ax * vp.dt2, vy = c1.y + ay * vp.dt2, vz = c1.z + az * vp.dt2;
    float x = c0.x + vx * vp.dt, y = c0.y + vy * vp.dt, z = c0.z + vz * vp.dt;

    // static charges
    constexpr uint32_t G = S * kGroup;
    FieldBody fs;
    fs.init(x, y, z);
    for (uint32_t t0 = 0; t0 < p; t0 += tile) {
        const uint32_t cnt = min(tile, p - t0), cntp = round_up(cnt, G);
        __syncthreads();
        stage(nb_tile, charges + t0, cnt, cntp);
        __syncthreads();
        pipelined_tile<S, 1, U>(nb_tile, cntp, k, fs);
    }
    // other particles (old positions and velocities)
    DynBody fd;
    fd.init(x, y, z, valid ? (int)i : -2);
    const uint32_t dyn_tile = tile / 2;
    for (uint32_t t0 = 0; t0 < n; t0 += dyn_tile) {
        const uint32_t cnt = min(dyn_tile, n - t0), cntp = round_up(cnt, G);
        __syncthreads();
        stage_dyn(nb_tile, parts, t0, cnt, cntp);
        __syncthreads();
        pipelined_tile<S, 2, U>(nb_tile, cntp, k, fd);
    }
    // every lane of the group ends with the same bits; lane 0 writes
    const float Esx = group_sum<S>(fs.acc.sx + fs.acc.cx), Edx = group_sum<S>(fd.e.sx + fd.e.cx);
    const float Esy = group_sum<S>(fs.acc.sy + fs.acc.cy), Edy = group_sum<S>(fd.e.sy + fd.e.cy);
    const float Esz = group_sum<S>(fs.acc.sz + fs.acc.cz), Edz = group_sum<S>(fd.e.sz + fd.e.cz);
    const float Bx = group_sum<S>(fd.b.sx + fd.b.cx);
    const float By = group_sum<S>(fd.b.sy + fd.b.cy);
    const float Bz = group_sum<S>(fd.b.sz + fd.b.cz);
    if (!valid || k != 0) return;
    const float Ex = kCoulombF * (Esx + Edx);
    const float Ey = kCoulombF * (Esy + Edy);
    const float Ez = kCoulombF * (Esz + Edz);
    // a = q E + q (v x B), then / m  (Particle_System.cpp:173-178)
    const float q = c0.w;
    ax = q * Ex + q * (vy * Bz - vz * By);
    ay = q * Ey + q * (vz * Bx - vx * Bz);
    az = q * Ez + q * (vx * By - vy * Bx);
    ax = ax / c1.w;
    ay = ay / c1.w;
    az = az / c1.w;
    accel[3 * i] = ax;
    accel[3 * i + 1] = ay;
    accel[3 * i + 2] = az;
    if (out) {
        // v(t + dt) = v(t + dt/2) + a(t + dt) dt/2, then BoundToBox (19-56)
        float wx = vx + ax * vp.dt2, wy = vy + ay * vp.dt2, wz = vz + az * vp.dt2;
        if (x > vp.bmax[0]) { x = 2.f * vp.bmax[0] - x; wx = -wx; }
        if (y > vp.bmax[1]) { y = 2.f * vp.bmax[1] - y; wy = -wy; }
        if (z > vp.bmax[2]) { z = 2.f * vp.bmax[2] - z; wz = -wz; }
        if (x < vp.bmin[0]) { x = 2.f * vp.bmin[0] - x; wx = -wx; }
        if (y < vp.bmin[1]) { y = 2.f * vp.bmin[1] - y; wy = -wy; }
        if (z < vp.bmin[2]) { z = 2.f * vp.bmin[2] - z; wz = -wz; }
        out[2 * i] = make_float4(x, y, z, q);
        out[2 * i + 1] = make_float4(wx, wy, wz, c1.w);
    }
}

template <int S, int U>
int launch_verlet_s(const float* parts, uint32_t n, const float* charges, uint32_t p, float* accel,
                    const VerletP& vp, float* out, uint32_t tile, hipStream_t stream)
{
    const uint32_t ppb = kNbBlock / S;  // particles per workgroup
    const dim3 grid((n + ppb - 1) / ppb), block(kNbBlock);
    hipLaunchKernelGGL((k_verlet_step<S, U>), grid, block, tile * sizeof(float4), stream,
                       reinterpret_cast<const float4*>(parts), n,
                       reinterpret_cast<const float4*>(charges), p, accel, vp,
                       reinterpret_cast<float4*>(out), tile);
    EM_HIP(hipGetLastError());
    return EM_OK;
}

int launch_verlet(const float* parts, uint32_t n, const float* charges, uint32_t p, float* accel,
                  const VerletP& vp, float* out, hipStream_t stream)
{
    DeviceInfo dev;
    EM_TRY(current_device_info(dev));
    // lanes per particle: until the grid holds four waves per SIMD, unless
    // EMTRACE_VERLET_LANES (1, 2, 4 or 8) names a number
    int S = 1;
    while (S < kNbMaxLanes && (uint64_t)n * S < 4ull * 4ull * 64ull * (uint64_t)dev.cus) S *= 2;
    if (const char* env = std::getenv("EMTRACE_VERLET_LANES")) {
        const int want = std::atoi(env);
        if (want != 1 && want != 2 && want != 4 && want != 8)
            EM_FAIL(EM_ERR_INVALID, "EMTRACE_VERLET_LANES=%s not in {1, 2, 4, 8}", env);
        S = want;
    }
    const uint32_t tile = S == 1 ? kNbTile : kNbTileSplit;
    switch (S) {
        case 1: return launch_verlet_s<1, 8>(parts, n, charges, p, accel, vp, out, tile, stream);
        case 2: return launch_verlet_s<2, 4>(parts, n, charges, p, accel, vp, out, tile, stream);
        case 4: return launch_verlet_s<4, 2>(parts, n, charges, p, accel, vp, out, tile, stream);
        default: return launch_verlet_s<8, 4>(parts, n, charges, p, accel, vp, out, tile, stream);
    }
}

int check_verlet(size_t n, size_t p, const float* box_min, const float* box_max, float dt)
{
    if (!n) EM_FAIL(EM_ERR_NO_LINES, "n == 0");  // the reference's only check (Particle_System.cpp:121-122)
    if (n > 0x7FFFFFFFull || p > 0xFFFFFFFFull) EM_FAIL(EM_ERR_INVALID, "n must fit 31 bits, p 32");
    if (!box_min || !box_max) EM_FAIL(EM_ERR_INVALID, "box bounds are NULL");
    if (!std::isfinite(dt)) EM_FAIL(EM_ERR_INVALID, "dt not finite");
    return EM_OK;
}

VerletP make_vp(const float* box_min, const float* box_max, float dt)
{
    VerletP vp;
    vp.dt = dt;
    vp.dt2 = dt / 2;
    for (int a = 0; a < 3; ++a) {
        vp.bmin[a] = box_min[a];
        vp.bmax[a] = box_max[a];
    }
    return vp;
}

}  // namespace
}  // namespace em

using namespace em;

extern "C" {

int em_verlet_step_dev_f32(const float* d_particles, size_t n, const float* d_charges, size_t p,
                           float* d_accel, const float* box_min, const float* box_max, float dt,
                           float* d_particles_out, void* stream)
{
    clear_error();
    EM_TRY(check_verlet(n, p, box_min, box_max, dt));
    if (!d_particles || !d_accel || (p && !d_charges)) EM_FAIL(EM_ERR_INVALID, "NULL device pointer");
    if (d_particles_out == d_particles)
        EM_FAIL(EM_ERR_INVALID, "the output particles must not alias the input (others read the old state)");
    DeviceInfo dev;
    EM_TRY(current_device_info(dev));
    return launch_verlet(d_particles, (uint32_t)n, d_charges, (uint32_t)p, d_accel,
                         make_vp(box_min, box_max, dt), d_particles_out,
                         static_cast<hipStream_t>(stream));
}

int em_verlet_step_f32(float* particles, size_t n, const float* charges, size_t p, float* accel,
                       const float* box_min, const float* box_max, float dt, int write_back,
                       em_perf* perf)
{
    clear_error();
    const auto t0 = std::chrono::steady_clock::now();
    if (perf) {
        set_progress(perf, 0.0);
        perf->time_s = perf->kernel_s = perf->h2d_s = perf->d2h_s = 0.0;
    }
    EM_TRY(check_verlet(n, p, box_min, box_max, dt));
    if (!particles || !accel || (p && !charges)) EM_FAIL(EM_ERR_INVALID, "NULL array");
    DeviceInfo dev;
    EM_TRY(current_device_info(dev));
    struct Res {
        void* mem = nullptr;
        hipStream_t s = nullptr;
        hipEvent_t ev[2] = {nullptr, nullptr};
        ~Res()
        {
            if (s) (void)hipStreamSynchronize(s);
            if (mem) (void)hipFree(mem);
            for (auto& e : ev)
                if (e) (void)hipEventDestroy(e);
            if (s) (void)hipStreamDestroy(s);
        }
    } r;
    const size_t pb = n * 8 * sizeof(float), qb = std::max<size_t>(p, 1) * 4 * sizeof(float),
                 ab = n * 3 * sizeof(float);
    EM_HIP(hipStreamCreateWithFlags(&r.s, hipStreamNonBlocking));
    for (auto& e : r.ev) EM_HIP(hipEventCreate(&e));
    EM_HIP(hipMalloc(&r.mem, 2 * pb + qb + ab));
    float* dp = static_cast<float*>(r.mem);
    float* dout = dp + n * 8;
    float* dq = dout + n * 8;
    float* da = dq + std::max<size_t>(p, 1) * 4;
    const CopySeg in[3] = {{dp, particles, pb}, {da, accel, ab}, {dq, charges, p * 4 * sizeof(float)}};
    EM_TRY(h2d_staged(dev.device, in, p ? 3 : 2, r.s));
    EM_HIP(hipEventRecord(r.ev[0], r.s));
    EM_TRY(launch_verlet(dp, (uint32_t)n, dq, (uint32_t)p, da, make_vp(box_min, box_max, dt),
                         write_back ? dout : nullptr, r.s));
    EM_HIP(hipEventRecord(r.ev[1], r.s));
    const CopySeg segs[2] = {{accel, da, ab}, {particles, dout, pb}};
    EM_TRY(d2h_staged(dev.device, segs, write_back ? 2 : 1, r.s));
    EM_HIP(hipStreamSynchronize(r.s));
    float ms = 0.f;
    EM_HIP(hipEventElapsedTime(&ms, r.ev[0], r.ev[1]));
    if (perf) {
        perf->kernel_s = ms * 1e-3;
        perf->time_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        perf->interactions = (double)n * ((double)p + (double)n - 1.0);
        perf->interactions_per_s = perf->kernel_s > 0 ? perf->interactions / perf->kernel_s : 0.0;
        set_progress(perf, 1.0);
    }
    return EM_OK;
}

}  // extern "C"
