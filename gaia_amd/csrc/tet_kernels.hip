// HIP kernels of the tet VBD hot path (CDNA4 / gfx950, wave64).
//
// Reference behaviour (paths relative to Simulator/Modules/):
//   per-colour vertex block solve   VBD/VBDPhysics.cpp:2402-2483 (CPU, the parity target),
//                                   VBD/VBDPhysicsCompute.cu:803-968 (CUDA launch structure)
//   neo-Hookean force / Hessian     VBD/VBD_NeoHookean.cpp:379-581
//   inertia                         VBD/VBD_BaseMaterial.h:359-366
//   plane boundary + friction       VBD/VBDPhysics.cpp:2821-2954, 2563-2592
//   initial step                    VBD/VBD_BaseMaterial.h:201-259, VBD/VBD_BaseMaterial.cpp:7-50
//   velocity update + damping       VBD/VBDPhysics.cpp:2486-2561
//
// Design (not a translation of the CUDA kernel): LPV lanes of one wave64 share
// one vertex and stride over its colour-major incidence records; each record
// is 64 B streamed with 4 float4 loads plus 3 float4 neighbour gathers. The
// per-tet 9x9 assembly of the reference is replaced by its exact contraction
//   H_tet = A (mu |m|^2 I + lambda g g^T),  g = cof(F) m,
//   f_tet = -A (mu F m + lambda (det F - a) g),
// because the second derivative of det F contracted with m (x) m vanishes
// (antisymmetric in the two shape indices). Partial sums are combined with
// butterfly lane exchanges; positions are updated in place (vertices of one
// colour share no tet), so one launch per colour suffices.
#include <hip/hip_runtime.h>

#include "device_types.h"
#include "kernels.h"
#include "vec_math.h"

namespace gvbd {

namespace {

template <int LPV>
__device__ __forceinline__ float lane_sum(float v) {
#pragma unroll
    for (int o = LPV / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Plane-ground boundary penalty and IPC-mollified friction for one vertex
// (VBDPhysics.cpp:2877-2952 with computeVertexFriction :2563-2592).
__device__ __forceinline__ void boundary_force(const SolveParams& sp, f3 x, f3 xp, bool applyFriction, f3& f,
                                               Sym3& h) {
    const float k = sp.boundaryStiffness;
    const float xs[3] = {x.x, x.y, x.z};
    const float dxs[3] = {x.x - xp.x, x.y - xp.y, x.z - xp.z};
    float fs[3] = {f.x, f.y, f.z};
    float hd[3] = {h.xx, h.yy, h.zz};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        float pen;
        if (xs[d] < sp.lower[d]) {
            pen = sp.lower[d] - xs[d];
            fs[d] += pen * k;
        } else if (xs[d] > sp.upper[d]) {
            pen = xs[d] - sp.upper[d];
            fs[d] -= pen * k;
        } else {
            continue;
        }
        hd[d] += k;
        if (applyFriction) {
            const int d1 = (d + 1) % 3, d2 = (d + 2) % 3;
            const float u0 = dxs[d1], u1 = dxs[d2];
            const float uNorm = sqrtf(u0 * u0 + u1 * u1);
            if (uNorm > 0.f) {
                const float epsU = sp.frictionEpsU;
                const float f1 = uNorm > epsU ? 1.0f / uNorm : (-uNorm / epsU + 2.0f) / epsU;
                const float lambda = pen * k;
                const float c = sp.frictionMu * lambda * f1;
                fs[d1] -= c * u0;
                fs[d2] -= c * u1;
                hd[d1] += c;
                hd[d2] += c;
            }
        }
    }
    f = mk(fs[0], fs[1], fs[2]);
    h.xx = hd[0];
    h.yy = hd[1];
    h.zz = hd[2];
}

}  // namespace

// One launch = one parallel group (colour) of every scene referenced by `items`.
template <int LPV, int BLOCK>
__global__ __launch_bounds__(BLOCK) void tet_group_solve_kernel(const SceneView* __restrict__ scenes,
                                                               const WorkItem* __restrict__ items,
                                                               int iteration) {
    const WorkItem it = items[blockIdx.x];
    const SceneView& S = scenes[it.scene];
    const int local = threadIdx.x / LPV;
    const int sub = threadIdx.x % LPV;
    const bool active = local < it.count;

    int4 slot = make_int4(0, 0, 0, 0);
    if (active) slot = S.slots[it.slotBegin + local];
    const int v = slot.x;

    f3 x = mk(0.f, 0.f, 0.f);
    if (active) x = xyz(S.pos[v]);

    float f0 = 0.f, f1 = 0.f, f2 = 0.f;
    float gxx = 0.f, gxy = 0.f, gxz = 0.f, gyy = 0.f, gyz = 0.f, gzz = 0.f;
    float cmu = 0.f;

    const float4* __restrict__ pos = S.pos;
    const TetInc* __restrict__ tinc = S.tinc;
    for (int k = slot.y + sub; k < slot.z; k += LPV) {
        const float4* r = reinterpret_cast<const float4*>(tinc + k);
        const float4 r0 = r[0];
        const float4 r1 = r[1];
        const float4 r2 = r[2];
        const float4 r3 = r[3];
        const int n0 = __float_as_int(r0.x), n1 = __float_as_int(r0.y), n2 = __float_as_int(r0.z);
        const float Amu = r0.w;
        const f3 m0 = mk(r1.x, r1.y, r1.z);
        const f3 m1 = mk(r1.w, r2.x, r2.y);
        const f3 m2 = mk(r2.z, r2.w, r3.x);
        const float Alm = r3.y;
        const float a = r3.z;
        const f3 d0 = xyz(pos[n0]) - x;
        const f3 d1 = xyz(pos[n1]) - x;
        const f3 d2 = xyz(pos[n2]) - x;
        // F columns: F e_c = sum_k d_k m_k[c]
        const f3 c0 = m0.x * d0 + m1.x * d1 + m2.x * d2;
        const f3 c1 = m0.y * d0 + m1.y * d1 + m2.y * d2;
        const f3 c2 = m0.z * d0 + m1.z * d1 + m2.z * d2;
        const f3 mv = mk(-(m0.x + m1.x + m2.x), -(m0.y + m1.y + m2.y), -(m0.z + m1.z + m2.z));
        // cof(F) columns are c1 x c2, c2 x c0, c0 x c1
        const f3 k0 = cross(c1, c2);
        const f3 k1 = cross(c2, c0);
        const f3 k2 = cross(c0, c1);
        const float J = dot(c0, k0);
        const f3 g = mv.x * k0 + mv.y * k1 + mv.z * k2;
        const f3 Fm = mv.x * c0 + mv.y * c1 + mv.z * c2;
        const float s = Alm * (J - a);
        f0 -= Amu * Fm.x + s * g.x;
        f1 -= Amu * Fm.y + s * g.y;
        f2 -= Amu * Fm.z + s * g.z;
        const f3 ag = Alm * g;
        gxx += ag.x * g.x;
        gxy += ag.x * g.y;
        gxz += ag.x * g.z;
        gyy += ag.y * g.y;
        gyz += ag.y * g.z;
        gzz += ag.z * g.z;
        cmu += Amu * dot(mv, mv);
    }
    if (LPV > 1) {
        f0 = lane_sum<LPV>(f0);
        f1 = lane_sum<LPV>(f1);
        f2 = lane_sum<LPV>(f2);
        gxx = lane_sum<LPV>(gxx);
        gxy = lane_sum<LPV>(gxy);
        gxz = lane_sum<LPV>(gxz);
        gyy = lane_sum<LPV>(gyy);
        gyz = lane_sum<LPV>(gyz);
        gzz = lane_sum<LPV>(gzz);
        cmu = lane_sum<LPV>(cmu);
    }
    if (!active || sub != 0) return;

    const SolveParams& sp = S.sp;
    const float4 pv = S.prev[v];
    const f3 xp = xyz(pv);
    f3 f = mk(f0, f1, f2);
    Sym3 h{gxx + cmu, gxy, gxz, gyy + cmu, gyz, gzz + cmu};
    if (slot.w & SLOT_DAMPING) {
        // sum over tets of VBD_NeoHookean.cpp:562-574: (gamma_h * lambda part + gamma_d * mu part) / dt
        const float4 vp = S.vparam[v];
        const float ch = vp.z / sp.dt, cd = vp.w * cmu / sp.dt;
        const Sym3 dh{ch * gxx + cd, ch * gxy, ch * gxz, ch * gyy + cd, ch * gyz, ch * gzz + cd};
        const f3 disp = x - xp;
        f.x -= dh.xx * disp.x + dh.xy * disp.y + dh.xz * disp.z;
        f.y -= dh.xy * disp.x + dh.yy * disp.y + dh.yz * disp.z;
        f.z -= dh.xz * disp.x + dh.yz * disp.y + dh.zz * disp.z;
        h.xx += dh.xx; h.xy += dh.xy; h.xz += dh.xz;
        h.yy += dh.yy; h.yz += dh.yz; h.zz += dh.zz;
    }
    // inertia
    const float4 yv = S.inertia[v];
    const float mass = yv.w;
    f.x += mass * (yv.x - x.x) * sp.dtSqrRecip;
    f.y += mass * (yv.y - x.y) * sp.dtSqrRecip;
    f.z += mass * (yv.z - x.z) * sp.dtSqrRecip;
    const float hi = mass * sp.dtSqrRecip;
    h.xx += hi;
    h.yy += hi;
    h.zz += hi;
    if (sp.usePlaneGround) boundary_force(sp, x, xp, iteration >= sp.frictionStartIter, f, h);

    if (dot(f, f) > 1e-10f) {
        f3 dx;
        float step = sp.stepSize;
        if (!solve_sym3(h, f, dx)) {
            step = sp.stepSizeGD;
            dx = f;
        }
        x = x + step * dx;
        S.pos[v] = make_float4(x.x, x.y, x.z, 0.f);
    }
}

// VBD_BaseMaterial.h:244-259 applyInitialStep (computeInertia + forwardStep).
__global__ __launch_bounds__(256) void tet_initial_step_kernel(const SceneView* __restrict__ scenes,
                                                               const WorkItem* __restrict__ items) {
    const WorkItem it = items[blockIdx.x];
    const int i = threadIdx.x;
    if (i >= it.count) return;
    const SceneView& S = scenes[it.scene];
    const SolveParams& sp = S.sp;
    const int v = it.slotBegin + i;
    const float4 p = S.pos[v];
    const float4 vel = S.vel[v];
    const uint8_t flag = S.vflag[v];
    const float dt = sp.dt;
    S.prev[v] = p;
    f3 vn = xyz(vel);
    if (sp.associateGravity) {
        const f3 dg = mk(dt * sp.gravity[0], dt * sp.gravity[1], dt * sp.gravity[2]);
        vn = vn + dg;
    }
    if (flag & 1) vn = mk(0.f, 0.f, 0.f);
    const f3 x = xyz(p);
    const f3 y = x + dt * vn;
    const float mass = S.inertia[v].w;
    S.inertia[v] = make_float4(y.x, y.y, y.z, mass);
    f3 xn;
    switch ((flag >> 1) & 7) {
    case 0: xn = y; break;
    case 2: xn = x + dt * xyz(vel); break;
    case 3: xn = 0.5f * x + 0.5f * y; break;
    case 4: {
        const float c = dt * dt * sp.initRatioG;
        xn = (x + dt * xyz(vel)) + mk(c * sp.gravity[0], c * sp.gravity[1], c * sp.gravity[2]);
        break;
    }
    default: xn = x; break;
    }
    S.pos[v] = make_float4(xn.x, xn.y, xn.z, 0.f);
    S.vel[v] = make_float4(vn.x, vn.y, vn.z, 0.f);
}

// VBDPhysics.cpp:2486-2561 updateVelocities + dampVelocities.
__global__ __launch_bounds__(256) void tet_update_velocity_kernel(const SceneView* __restrict__ scenes,
                                                                  const WorkItem* __restrict__ items) {
    const WorkItem it = items[blockIdx.x];
    const int i = threadIdx.x;
    if (i >= it.count) return;
    const SceneView& S = scenes[it.scene];
    const int v = it.slotBegin + i;
    const float dt = S.sp.dt;
    const float4 p = S.pos[v];
    const float4 q = S.prev[v];
    f3 vel = mk((p.x - q.x) / dt, (p.y - q.y) / dt, (p.z - q.z) / dt);
    const float vMag = sqrtf(dot(vel, vel));
    if (vMag > 1e-6f) {
        const float4 vp = S.vparam[v];
        float vNew = vMag * vp.x - vp.y;
        vNew = vNew > 1e-6f ? vNew : 0.f;
        vel = (vNew / vMag) * vel;
    } else {
        vel = 0.f * vel;
    }
    if (S.vflag[v] & 1) vel = mk(0.f, 0.f, 0.f);
    S.vel[v] = make_float4(vel.x, vel.y, vel.z, 0.f);
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
template <int LPV, int BLOCK>
static hipError_t launch_solve(const SceneView* scenes, const WorkItem* items, int nitems, int iteration,
                               hipStream_t stream) {
    hipLaunchKernelGGL((tet_group_solve_kernel<LPV, BLOCK>), dim3(nitems), dim3(BLOCK), 0, stream, scenes, items,
                       iteration);
    return hipGetLastError();
}

hipError_t launch_tet_group_solve(const SceneView* scenes, const WorkItem* items, int nitems, int iteration,
                                  int lpv, int block, hipStream_t stream) {
    if (nitems <= 0) return hipSuccess;
#define GVBD_CASE(L, B) \
    if (lpv == L && block == B) return launch_solve<L, B>(scenes, items, nitems, iteration, stream);
    GVBD_CASE(1, 64) GVBD_CASE(1, 128) GVBD_CASE(1, 256)
    GVBD_CASE(2, 64) GVBD_CASE(2, 128) GVBD_CASE(2, 256)
    GVBD_CASE(4, 64) GVBD_CASE(4, 128) GVBD_CASE(4, 256)
    GVBD_CASE(8, 64) GVBD_CASE(8, 128) GVBD_CASE(8, 256)
    GVBD_CASE(16, 64) GVBD_CASE(16, 128) GVBD_CASE(16, 256)
#undef GVBD_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_tet_initial_step(const SceneView* scenes, const WorkItem* items, int nitems, hipStream_t stream) {
    if (nitems <= 0) return hipSuccess;
    hipLaunchKernelGGL(tet_initial_step_kernel, dim3(nitems), dim3(256), 0, stream, scenes, items);
    return hipGetLastError();
}

hipError_t launch_tet_update_velocity(const SceneView* scenes, const WorkItem* items, int nitems,
                                      hipStream_t stream) {
    if (nitems <= 0) return hipSuccess;
    hipLaunchKernelGGL(tet_update_velocity_kernel, dim3(nitems), dim3(256), 0, stream, scenes, items);
    return hipGetLastError();
}

// Plain streaming copy (float4 per lane, grid-stride): the reference kernel
// whose HBM counters calibrate the profiler and whose rate is the measured
// copy bandwidth of the device.
__global__ __launch_bounds__(256) void stream_copy_kernel(const float4* __restrict__ src, float4* __restrict__ dst,
                                                          int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

hipError_t launch_stream_copy(const float4* src, float4* dst, int64_t n, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 256 * 64);
    hipLaunchKernelGGL(stream_copy_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, src, dst, n);
    return hipGetLastError();
}

bool valid_schedule(int lpv, int block) {
    const bool l = lpv == 1 || lpv == 2 || lpv == 4 || lpv == 8 || lpv == 16;
    const bool b = block == 64 || block == 128 || block == 256;
    return l && b;
}

}  // namespace gvbd
