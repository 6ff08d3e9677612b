// Convergence evaluation of the tet solve (VBDPhysics::evaluateConvergenceGPU, VBD/VBDPhysics.cpp:
// 3510-3556, with evaluateMeritEnergyGPU, :3850-3870): the merit energy (inertia + stable
// neo-Hookean elastic energy, evaluateNeoHookeanEnergy, VBDPhysicsCompute.cu:1775-1807;
// computeInertiaEnergy, VBD_BaseMaterial.h:223-240) and the average norm over the tet vertices of
// the material (with damping) + inertia force. Per-block partial sums in double.
#include <hip/hip_runtime.h>

#include <mutex>

#include "device_types.h"
#include "kernels.h"
#include "vec_math.h"

namespace gvbd {
namespace {

constexpr int kBlock = 256;

__device__ double block_sum(double v, double* sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    return sh[0];
}

// per free tet vertex (its slot): |material + damping + inertia force|
__global__ __launch_bounds__(kBlock) void residual_kernel(SceneView S, int nslots, double* forcePartial) {
    __shared__ double sh[kBlock];
    const int t = blockIdx.x * kBlock + threadIdx.x;
    double fnorm = 0.0;
    if (t < nslots) {
        const int4 slot = S.slots[t];
        const int v = slot.x;
        const f3 x = xyz(S.pos[v]);
        f3 f = mk(0.f, 0.f, 0.f);
        Sym3 h{0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        float cmu = 0.f;
        const float4 mat = S.tmat[slot.w >> SLOT_MAT_SHIFT];
        for (int r = slot.y; r < slot.z; ++r) {
            const TetInc rec = S.tinc[r];
            const float A = S.tvol[r];
            const float Amu = A * mat.x, Alm = A * mat.y;
            const f3 m0 = mk(rec.m[0], rec.m[1], rec.m[2]);
            const f3 m1 = mk(rec.m[3], rec.m[4], rec.m[5]);
            const f3 m2 = mk(rec.m[6], rec.m[7], rec.m[8]);
            const f3 d0 = xyz(S.pos[rec.nb[0]]) - x;
            const f3 d1 = xyz(S.pos[rec.nb[1]]) - x;
            const f3 d2 = xyz(S.pos[rec.nb[2]]) - x;
            const f3 c0 = m0.x * d0 + m1.x * d1 + m2.x * d2;
            const f3 c1 = m0.y * d0 + m1.y * d1 + m2.y * d2;
            const f3 c2 = m0.z * d0 + m1.z * d1 + m2.z * d2;
            const f3 mv = mk(-(m0.x + m1.x + m2.x), -(m0.y + m1.y + m2.y), -(m0.z + m1.z + m2.z));
            const f3 k0 = cross(c1, c2), k1 = cross(c2, c0), k2 = cross(c0, c1);
            const float J = dot(c0, k0);
            const f3 g = mv.x * k0 + mv.y * k1 + mv.z * k2;
            const f3 Fm = mv.x * c0 + mv.y * c1 + mv.z * c2;
            const float s = Alm * (J - mat.z);
            f = f - (Amu * Fm + s * g);
            const f3 ag = Alm * g;
            h.xx += ag.x * g.x;
            h.xy += ag.x * g.y;
            h.xz += ag.x * g.z;
            h.yy += ag.y * g.y;
            h.yz += ag.y * g.z;
            h.zz += ag.z * g.z;
            cmu += Amu * dot(mv, mv);
        }
        if (slot.w & SLOT_DAMPING) {
            // the solve's damping term (VBD_NeoHookean.cpp:562-574)
            const float4 vp = S.vparam[v];
            const float ch = vp.z / S.sp.dt, cd = vp.w * cmu / S.sp.dt;
            const Sym3 dh{ch * h.xx + cd, ch * h.xy, ch * h.xz, ch * h.yy + cd, ch * h.yz, ch * h.zz + cd};
            f = f - sym_mul(dh, x - xyz(S.prev[v]));
        }
        const float4 yv = S.inertia[v];
        const f3 dy = mk(yv.x - x.x, yv.y - x.y, yv.z - x.z);
        f = f + (yv.w * S.sp.dtSqrRecip) * dy;
        fnorm = sqrt((double)dot(f, f));
    }
    const double a = block_sum(fnorm, sh);
    if (threadIdx.x == 0) forcePartial[blockIdx.x] = a;
}

// per tet vertex, fixed ones included: m |y - x|^2
__global__ __launch_bounds__(kBlock) void inertia_kernel(SceneView S, const int* __restrict__ verts, int n,
                                                         double* partial) {
    __shared__ double sh[kBlock];
    const int t = blockIdx.x * kBlock + threadIdx.x;
    double e = 0.0;
    if (t < n) {
        const int v = verts[t];
        const float4 x = S.pos[v], y = S.inertia[v];
        const f3 dy = mk(y.x - x.x, y.y - x.y, y.z - x.z);
        e = (double)y.w * (double)dot(dy, dy);
    }
    const double a = block_sum(e, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = a;
}

// per tet: A (mu (|F|^2 - 3) / 2 + lambda (det F - a)^2 / 2) - A lambda (1 - a)^2 / 2
__global__ __launch_bounds__(kBlock) void tet_energy_kernel(const TetEnergyRec* __restrict__ tets, int nt,
                                                            const float4* __restrict__ pos, double* partial) {
    __shared__ double sh[kBlock];
    const int t = blockIdx.x * kBlock + threadIdx.x;
    double e = 0.0;
    if (t < nt) {
        const TetEnergyRec r = tets[t];
        const f3 x0 = xyz(pos[r.v.x]);
        const f3 e1 = xyz(pos[r.v.y]) - x0, e2 = xyz(pos[r.v.z]) - x0, e3 = xyz(pos[r.v.w]) - x0;
        // F = Ds DmInv, DmInv column-major: column c of F = sum_k e_k DmInv(k, c)
        const f3 F0 = r.dminv[0] * e1 + r.dminv[1] * e2 + r.dminv[2] * e3;
        const f3 F1 = r.dminv[3] * e1 + r.dminv[4] * e2 + r.dminv[5] * e3;
        const f3 F2 = r.dminv[6] * e1 + r.dminv[7] * e2 + r.dminv[8] * e3;
        const float a = 1.f + r.mu / r.lmbd;
        const float phiD = 0.5f * (dot(F0, F0) + dot(F1, F1) + dot(F2, F2) - 3.f);
        const float detF = dot(F0, cross(F1, F2));
        const float phiH = 0.5f * (detF - a) * (detF - a);
        const float E = r.A * (r.mu * phiD + r.lmbd * phiH);
        const float rest = 0.5f * r.A * r.lmbd * (1.f - a) * (1.f - a);
        e = (double)(E - rest);
    }
    const double s = block_sum(e, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// out[k] = partial[off[k]] + partial[off[k] + 1] + ... over n[k] entries, k = 0..2, summed in
// ascending order by one thread (the order of the host-side sum of the allocating entry point,
// so both give the same doubles)
__global__ void final_sum_kernel(const double* __restrict__ partial, int3 off, int3 n, double* __restrict__ out) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const int o[3] = {off.x, off.y, off.z}, c[3] = {n.x, n.y, n.z};
    for (int k = 0; k < 3; ++k) {
        double s = 0.0;
        for (int b = 0; b < c[k]; ++b) s += partial[o[k] + b];
        out[k] = s;
    }
}

// Scratch of the allocation-free convergence evaluation: per-block partial sums, then the three
// totals. One copy per GPU, held until the process ends; every call on a GPU shares it.
static __device__ __align__(256) double scratch[kConvergenceScratchDoubles];

std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

}  // namespace

int convergence_blocks(int n) { return (n + kBlock - 1) / kBlock; }

hipError_t convergence_scratch_init(int device, double** out) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur != device) return hipErrorInvalidDevice;  // the caller makes `device` current
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = p;
    }
    *out = (double*)g_scratch_ptr[device];
    return hipSuccess;
}

hipError_t launch_convergence_final_sum(const double* partial, int off0, int n0, int off1, int n1, int off2, int n2,
                                        double* out, hipStream_t stream) {
    hipLaunchKernelGGL(final_sum_kernel, dim3(1), dim3(64), 0, stream, partial, make_int3(off0, off1, off2),
                       make_int3(n0, n1, n2), out);
    return hipGetLastError();
}

hipError_t launch_convergence(const SceneView& view, int nslots, const TetEnergyRec* tets, int nt,
                              const int* tetVerts, int ntv, double* forcePartial, double* inertiaPartial,
                              double* energyPartial, hipStream_t stream) {
    if (nslots > 0)
        hipLaunchKernelGGL(residual_kernel, dim3(convergence_blocks(nslots)), dim3(kBlock), 0, stream, view, nslots,
                           forcePartial);
    if (ntv > 0)
        hipLaunchKernelGGL(inertia_kernel, dim3(convergence_blocks(ntv)), dim3(kBlock), 0, stream, view, tetVerts, ntv,
                           inertiaPartial);
    if (nt > 0)
        hipLaunchKernelGGL(tet_energy_kernel, dim3(convergence_blocks(nt)), dim3(kBlock), 0, stream, tets, nt,
                           view.pos, energyPartial);
    return hipGetLastError();
}

}  // namespace gvbd
