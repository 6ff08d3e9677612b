// Small float3 helpers and the 3x3 block solve used by every per-vertex kernel.
#pragma once
#include <hip/hip_runtime.h>

namespace gvbd {

struct f3 {
    float x, y, z;
};

__device__ __forceinline__ f3 mk(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ f3 xyz(float4 a) { return {a.x, a.y, a.z}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 operator*(float s, f3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross(f3 a, f3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// Symmetric 3x3 block: (xx, xy, xz, yy, yz, zz).
struct Sym3 {
    float xx, xy, xz, yy, yz, zz;
};

__device__ __forceinline__ f3 sym_mul(const Sym3& h, f3 v) {
    return mk(h.xx * v.x + h.xy * v.y + h.xz * v.z, h.xy * v.x + h.yy * v.y + h.yz * v.z,
              h.xz * v.x + h.yz * v.y + h.zz * v.z);
}

// Solves H dx = b for a symmetric 3x3 block by adjugate / determinant.
// Stands in for CuMatrix::solve3x3_psd_stable (VBDPhysics.cpp:2449,
// VBDPhysicsCompute.cu:896), whose source is not available: failure iff the
// determinant is not finite or |det| <= 1e-12 |h_xx h_yy h_zz|. Returns false on
// failure, and then the caller takes a gradient step (VBDPhysics.cpp:2452-2459).
__device__ __forceinline__ bool solve_sym3(const Sym3& h, f3 b, f3& x) {
    const float C00 = h.yy * h.zz - h.yz * h.yz;
    const float C01 = h.yz * h.xz - h.xy * h.zz;
    const float C02 = h.xy * h.yz - h.yy * h.xz;
    const float C11 = h.xx * h.zz - h.xz * h.xz;
    const float C12 = h.xy * h.xz - h.xx * h.yz;
    const float C22 = h.xx * h.yy - h.xy * h.xy;
    const float det = h.xx * C00 + h.xy * C01 + h.xz * C02;
    const float scale = fabsf(h.xx * h.yy * h.zz);
    if (!isfinite(det) || !(fabsf(det) > 1e-12f * scale)) return false;
    const float inv = 1.0f / det;
    x.x = (C00 * b.x + C01 * b.y + C02 * b.z) * inv;
    x.y = (C01 * b.x + C11 * b.y + C12 * b.z) * inv;
    x.z = (C02 * b.x + C12 * b.y + C22 * b.z) * inv;
    return isfinite(x.x) && isfinite(x.y) && isfinite(x.z);
}

}  // namespace gvbd
