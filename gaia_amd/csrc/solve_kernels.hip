// HIP kernels of the VBD hot path (CDNA4 / gfx950, wave64): tet and cloth vertex block solves.
//
// Reference behaviour (paths relative to Simulator/Modules/):
//   tet per-colour vertex block solve  VBD/VBDPhysics.cpp:2402-2483 (CPU, the parity target),
//                                      VBD/VBDPhysicsCompute.cu:803-968 (CUDA launch structure)
//   neo-Hookean force / Hessian        VBD/VBD_NeoHookean.cpp:379-581
//   inertia                            VBD/VBD_BaseMaterial.h:359-366, VBDCloth/VBDBaseTriMesh.cpp:311-318
//   plane boundary + friction          VBD/VBDPhysics.cpp:2821-2954, 2563-2592;
//                                      VBDCloth/VBDClothPhysics.cpp:1401-1463, 1500-1531
//   cloth per-vertex solve             VBDCloth/VBDClothPhysics.cpp:1026-1154
//   StVK membrane                      VBDCloth/VBDTriMeshStVK.cpp:721-911
//   quadratic bending                  VBDCloth/VBDTriMeshStVK.cpp:913-982
//   initial step                       VBD/VBD_BaseMaterial.h:201-259, VBD/VBD_BaseMaterial.cpp:7-50,
//                                      VBDCloth/VBDBaseTriMesh.cpp:58-178
//   velocity update + damping          VBD/VBDPhysics.cpp:2486-2561, VBDCloth/VBDClothPhysics.cpp:2318-2360
//
// Design (not a translation of the CUDA kernel): LPV lanes of one wave64 share one vertex
// and stride over its colour-major incidence records (64 B per tet, 32 B per face or
// bending stencil, streamed with float4 loads) plus float4 neighbour gathers; partial sums
// are combined with butterfly lane exchanges. Element Hessians are evaluated in contracted
// closed form instead of the reference's 9x9 / 6x6 assemblies:
//   tet:  H = A (mu |m|^2 I + lambda g g^T), g = cof(F) m;  f = -A (mu F m + lambda (det F - a) g)
//         (the second derivative of det F is antisymmetric in the shape indices, so it
//         vanishes against m (x) m);
//   StVK: H = A [(m^T S m) I + mu |m|^2 F F^T + (mu + lambda)(F m)(F m)^T];  f = -A F S m.
// One launch covers one parallel group of every scene of a plan; each block's work item is
// either tet slots (updated in place: vertices of one colour share no tet, as in the
// reference CPU path) or cloth slots (written to pnext and copied back after the group, the
// reference's positionsNext scheme).
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
// (tet: VBDPhysics.cpp:2877-2952; cloth: VBDClothPhysics.cpp:1401-1463; friction :2563-2592 / :1500-1531).
__device__ __forceinline__ void boundary_force(const SolveParams& sp, float k, float frictionEpsU, f3 x, f3 xp,
                                               bool applyFriction, f3& f, Sym3& h) {
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
                const float f1 = uNorm > frictionEpsU ? 1.0f / uNorm : (-uNorm / frictionEpsU + 2.0f) / frictionEpsU;
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

// ---------------------------------------------------------------------------
// Incidence records: read straight from HBM by the lanes that use them (STAGE = 0), or staged
// through LDS in chunks of whole-block coalesced loads (STAGE records of Q x 16 B per chunk). A
// lane handles the same records in the same order either way, so the results are identical.
// ---------------------------------------------------------------------------

// Calls body(r, k) for the records k = first, first + LPV, ... < end of one lane, where r points
// at record k's `Q` float4s (in HBM, or in the LDS copy of the chunk that holds it). The block's
// records are [blockBegin, blockEnd); every thread of the block must call this (barriers).
template <int LPV, int STAGE, int Q, class Body>
__device__ __forceinline__ void for_lane_records(const float4* __restrict__ recs, int blockBegin, int blockEnd,
                                                 int first, int end, float4* lds, Body&& body) {
    if constexpr (STAGE == 0) {
        // the next record's loads are issued before this record's gathers and arithmetic, so a
        // lane has two records in flight
        if (first >= end) return;
        float4 r[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) r[q] = recs[(int64_t)Q * first + q];
        for (int k = first; k < end; k += LPV) {
            float4 nx[Q];
            const int kn = k + LPV < end ? k + LPV : k;
#pragma unroll
            for (int q = 0; q < Q; ++q) nx[q] = recs[(int64_t)Q * kn + q];
            body(r, k);
#pragma unroll
            for (int q = 0; q < Q; ++q) r[q] = nx[q];
        }
    } else {
        for (int c0 = blockBegin; c0 < blockEnd; c0 += STAGE) {
            const int n = min(STAGE, blockEnd - c0);
            const float4* src = recs + (int64_t)Q * c0;
            for (int i = threadIdx.x; i < Q * n; i += blockDim.x) lds[i] = src[i];
            __syncthreads();
            int k = first;
            if (k < c0) k += ((c0 - k + LPV - 1) / LPV) * LPV;
            const int hi = min(end, c0 + n);
            for (; k < hi; k += LPV) body(lds + Q * (k - c0), k);
            __syncthreads();
        }
    }
}

template <int LPV>
__device__ void contact_forces(const SceneView& S, int v, int sub, const int* faceRef, int fb, int fe,
                               const int* edgeRef, int eb, int ee, f3& f, Sym3& h);

// Sums a vertex's contact force and Hessian over its LPV lanes (every lane gets the sums).
template <int LPV>
__device__ __forceinline__ void lane_sum_contact(f3& f, Sym3& h) {
    if (LPV > 1) {
        f.x = lane_sum<LPV>(f.x);
        f.y = lane_sum<LPV>(f.y);
        f.z = lane_sum<LPV>(f.z);
        h.xx = lane_sum<LPV>(h.xx);
        h.xy = lane_sum<LPV>(h.xy);
        h.xz = lane_sum<LPV>(h.xz);
        h.yy = lane_sum<LPV>(h.yy);
        h.yz = lane_sum<LPV>(h.yz);
        h.zz = lane_sum<LPV>(h.zz);
    }
}

// ---------------------------------------------------------------------------
// Tet vertex block (VBDPhysics.cpp:2402-2483)
// ---------------------------------------------------------------------------
template <int LPV, int STAGE, bool CONTACT>
__device__ __forceinline__ void tet_block(const SceneView& S, const WorkItem& it, int iteration, float4* lds) {
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
    int blockBegin = 0, blockEnd = 0;
    if (STAGE > 0) {
        blockBegin = S.slots[it.slotBegin].y;
        blockEnd = S.slots[it.slotBegin + it.count - 1].z;
    }
    // the mesh's material (mu, lambda, a); A mu and A lambda are formed per record from its tet's
    // rest volume, the same float products as precomputed ones
    float4 mat = make_float4(0.f, 0.f, 0.f, 0.f);
    if (active) mat = S.tmat[slot.w >> SLOT_MAT_SHIFT];
    const float* __restrict__ tvol = S.tvol;
    for_lane_records<LPV, STAGE, 3>(reinterpret_cast<const float4*>(S.tinc), blockBegin, blockEnd, slot.y + sub,
                                    slot.z, lds, [&](const float4* r, int k) {
        const float4 r0 = r[0];
        const float4 r1 = r[1];
        const float4 r2 = r[2];
        const int n0 = __float_as_int(r0.x), n1 = __float_as_int(r0.y), n2 = __float_as_int(r0.z);
        const f3 m0 = mk(r0.w, r1.x, r1.y);
        const f3 m1 = mk(r1.z, r1.w, r2.x);
        const f3 m2 = mk(r2.y, r2.z, r2.w);
        const float A = tvol[k];
        const float Amu = A * mat.x;
        const float Alm = A * mat.y;
        const float a = mat.z;
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
    });
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
    // collision handling: cloth-model contact on the mesh's boundary surface, each of the
    // vertex's lanes taking a share of its contact pairs (interior vertices have empty ranges)
    f3 cf = mk(0.f, 0.f, 0.f);
    Sym3 ch{0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if constexpr (CONTACT) {
        if (active) {
            const int4 tc = S.tcontact[it.slotBegin + local];
            contact_forces<LPV>(S, v, sub, S.cv.tFaceRef, tc.x, tc.y, S.cv.tEdgeRef, tc.z, tc.w, cf, ch);
        }
        lane_sum_contact<LPV>(cf, ch);
    }
    if (!active || sub != 0) return;

    const SolveParams& sp = S.sp;
    const f3 xp = xyz(S.prev[v]);
    f3 f = mk(f0, f1, f2);
    Sym3 h{gxx + cmu, gxy, gxz, gyy + cmu, gyz, gzz + cmu};
    if (slot.w & SLOT_DAMPING) {
        // sum over tets of VBD_NeoHookean.cpp:562-574: (gamma_h * lambda part + gamma_d * mu part) / dt
        const float4 vp = S.vparam[v];
        const float ch = vp.z / sp.dt, cd = vp.w * cmu / sp.dt;
        const Sym3 dh{ch * gxx + cd, ch * gxy, ch * gxz, ch * gyy + cd, ch * gyz, ch * gzz + cd};
        f = f - sym_mul(dh, x - xp);
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
    if (sp.usePlaneGround)
        boundary_force(sp, sp.boundaryStiffness, sp.frictionEpsU, x, xp, iteration >= sp.frictionStartIter, f, h);
    if constexpr (CONTACT) {
        f = f + cf;
        h.xx += ch.xx; h.xy += ch.xy; h.xz += ch.xz;
        h.yy += ch.yy; h.yz += ch.yz; h.zz += ch.zz;
    }

    f3 xn = x;
    const bool moved = dot(f, f) > 1e-10f;
    if (moved) {
        f3 dx;
        float step = sp.stepSize;
        if (!solve_sym3(h, f, dx)) {
            // the 3x3 block is singular: gradient step (VBDPhysics.cpp:2449-2466), counted
            step = sp.stepSizeGD;
            dx = f;
            atomicAdd(S.solveFail, 1);
        }
        xn = x + step * dx;
    }
    if (CONTACT && S.cv.model == 0) {
        // applyConservativeBoundTruncation (VBDClothPhysics.cpp:1223-1246). The tet initial
        // guess is not bounded, so every vertex is truncated, moved or not (interior vertices
        // have an unbounded step).
        const f3 x0 = xyz(S.cv.xAtCD[v]);
        const f3 acc = xn - x0;
        const float size = sqrtf(dot(acc, acc));
        const float b = S.cv.bound[v];
        if (size > b) {
            xn = x0 + (b / size) * acc;
            atomicOr(S.cv.flag, 1);
        }
    }
    if (CONTACT || moved) S.pos[v] = make_float4(xn.x, xn.y, xn.z, 0.f);
}

// ---------------------------------------------------------------------------
// Cloth contact (VBDClothPhysics.cpp:1248-1399, 1500-1573): every evaluation recomputes the
// closest points from the current positions (the reference's updateContactInfo = true).
// ---------------------------------------------------------------------------
__device__ __forceinline__ f3 normalized(f3 v) {
    const float n = sqrtf(dot(v, v));
    return n > 0.f ? (1.f / n) * v : v;
}

// CollisionGeometry.h:90-157 with barycentrics; true when the closest point is interior to the face
__device__ bool closest_point_bary(f3 p, f3 a, f3 b, f3 c, f3& cp, f3& bary) {
    const f3 ab = b - a, ac = c - a, ap = p - a;
    const float d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.f && d2 <= 0.f) {
        cp = a;
        bary = mk(1.f, 0.f, 0.f);
        return false;
    }
    const f3 bp = p - b;
    const float d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0.f && d4 <= d3) {
        cp = b;
        bary = mk(0.f, 1.f, 0.f);
        return false;
    }
    const f3 cpv = p - c;
    const float d5 = dot(ab, cpv), d6 = dot(ac, cpv);
    if (d6 >= 0.f && d5 <= d6) {
        cp = c;
        bary = mk(0.f, 0.f, 1.f);
        return false;
    }
    const float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.f && d1 >= 0.f && d3 <= 0.f) {
        const float v = d1 / (d1 - d3);
        cp = a + v * ab;
        bary = mk(1.f - v, v, 0.f);
        return false;
    }
    const float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.f && d2 >= 0.f && d6 <= 0.f) {
        const float v = d2 / (d2 - d6);
        cp = a + v * ac;
        bary = mk(1.f - v, 0.f, v);
        return false;
    }
    const float va = d3 * d6 - d5 * d4;
    if (va <= 0.f && (d4 - d3) >= 0.f && (d5 - d6) >= 0.f) {
        const float v = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        cp = b + v * (c - b);
        bary = mk(0.f, 1.f - v, v);
        return false;
    }
    const float denom = 1.f / (va + vb + vc);
    const float v = vb * denom, w = vc * denom;
    cp = (a + v * ab) + w * ac;
    bary = mk(1.f - v - w, v, w);
    return true;
}

// CollisionGeometry.h:160-267 (absolute CMP_EPSILON on the determinant, as the reference)
__device__ void segment_closest(f3 p0, f3 p1, f3 q0, f3 q1, f3& c1, f3& c2, float& s, float& t) {
    const f3 p = p1 - p0, q = q1 - q0, r = p0 - q0;
    const float a = dot(p, p), b = dot(p, q), c = dot(q, q), d = dot(p, r), e = dot(q, r);
    s = 0.f;
    t = 0.f;
    const float det = a * c - b * b;
    if (det > 0.00001f) {
        const float bte = b * e, ctd = c * d;
        if (bte <= ctd) {
            if (e <= 0.f) {
                s = (-d >= a ? 1.f : (-d > 0.f ? -d / a : 0.f));
            } else if (e < c) {
                t = e / c;
            } else {
                s = (b - d >= a ? 1.f : (b - d > 0.f ? (b - d) / a : 0.f));
                t = 1.f;
            }
        } else {
            s = bte - ctd;
            if (s >= det) {
                if (b + e <= 0.f) {
                    s = (-d <= 0.f ? 0.f : (-d < a ? -d / a : 1.f));
                    t = 0.f;
                } else if (b + e < c) {
                    s = 1.f;
                    t = (b + e) / c;
                } else {
                    s = (b - d <= 0.f ? 0.f : (b - d < a ? (b - d) / a : 1.f));
                    t = 1.f;
                }
            } else {
                const float ate = a * e, btd = b * d;
                if (ate <= btd) {
                    s = (-d <= 0.f ? 0.f : (-d >= a ? 1.f : -d / a));
                    t = 0.f;
                } else {
                    t = ate - btd;
                    if (t >= det) {
                        s = (b - d <= 0.f ? 0.f : (b - d >= a ? 1.f : (b - d) / a));
                        t = 1.f;
                    } else {
                        s /= det;
                        t /= det;
                    }
                }
            }
        }
    } else {
        if (e <= 0.f) {
            s = (-d <= 0.f ? 0.f : (-d >= a ? 1.f : -d / a));
        } else if (e >= c) {
            s = (b - d <= 0.f ? 0.f : (b - d >= a ? 1.f : (b - d) / a));
            t = 1.f;
        } else {
            t = e / c;
        }
    }
    c1 = (1.f - s) * p0 + s * p1;
    c2 = (1.f - t) * q0 + t * q1;
}

// VBDClothPhysics.cpp:1533-1573 (0 quadratic, 1 quadratic + log below radius / 2, 2 IPC-like)
__device__ void repulsive(const SolveParams& sp, float dis, float& dEdD, float& d2E) {
    const float adj = dis - sp.thickness;
    const float pen = sp.contactRadius - adj;
    const float k = sp.contactStiffness;
    if (sp.contactEnergyType == 1) {
        const float tau = sp.contactRadius * 0.5f;
        if (adj < tau && adj > 0.00001f) {
            const float k2 = 0.5f * tau * tau * k;
            dEdD = -k2 / adj;
            d2E = k2 / (adj * adj);
        } else {
            dEdD = -k * pen;
            d2E = k;
        }
    } else if (sp.contactEnergyType == 2) {
        const float dr = adj - sp.contactRadius;
        const float lg = logf(adj / sp.contactRadius);
        dEdD = -k * dr * dr / adj - 2.f * k * lg * dr;
        d2E = k * (dr * dr / (adj * adj) - 4.f * dr / adj - 2.f * lg);
    } else {
        dEdD = -k * pen;
        d2E = k;
    }
}

// repulsion b lambda n, Hessian d2E b^2 n n^T, and IPC-mollified friction in the tangent frame
// (t0 normalised, (t0 x n) normalised), VBDClothPhysics.cpp:1500-1531
__device__ void contact_apply(float dt, float b, float lambda, float d2E, f3 n, f3 dx, f3 t0, float mu, float epsV,
                              f3& f, Sym3& h) {
    f = f + (b * lambda) * n;
    const float cn = d2E * b * b;
    h.xx += cn * n.x * n.x; h.xy += cn * n.x * n.y; h.xz += cn * n.x * n.z;
    h.yy += cn * n.y * n.y; h.yz += cn * n.y * n.z; h.zz += cn * n.z * n.z;
    const f3 e0 = normalized(t0);
    const f3 e1 = normalized(cross(e0, n));
    const float u0 = dot(e0, dx), u1 = dot(e1, dx);
    const float uNorm = sqrtf(u0 * u0 + u1 * u1);
    if (!(uNorm > 0.f)) return;
    const float epsU = epsV * dt;
    const float f1 = uNorm > epsU ? 1.f / uNorm : (-uNorm / epsU + 2.f) / epsU;
    const float c = mu * lambda * f1;
    f = f - (b * c) * (u0 * e0 + u1 * e1);
    const float ch = b * b * c;
    h.xx += ch * (e0.x * e0.x + e1.x * e1.x); h.xy += ch * (e0.x * e0.y + e1.x * e1.y);
    h.xz += ch * (e0.x * e0.z + e1.x * e1.z); h.yy += ch * (e0.y * e0.y + e1.y * e1.y);
    h.yz += ch * (e0.y * e0.z + e1.y * e1.z); h.zz += ch * (e0.z * e0.z + e1.z * e1.z);
}

// Position of contact partner q for a vertex of parallel group g (see ContactView::snap).
__device__ __forceinline__ f3 cpos(const SceneView& S, int q, int g) {
    return xyz(S.cv.vgroup[q] == g ? S.cv.snap[q] : S.pos[q]);
}

// VF contact of vertex u with face `face`, on the vertex of order `order` (3: u itself,
// 0..2: face vertex; weights 1 and -bary_order), VBDClothPhysics.cpp:1248-1321; g is the
// solving vertex's parallel group
__device__ void vf_contact(const SceneView& S, int u, int face, int order, int g, f3& f, Sym3& h) {
    const ContactView& C = S.cv;
    const int4 fv = C.faces[face];
    const f3 x = cpos(S, u, g);
    const f3 a = cpos(S, fv.x, g), b = cpos(S, fv.y, g), c = cpos(S, fv.z, g);
    f3 cp, bary;
    closest_point_bary(x, a, b, c, cp, bary);
    const f3 diff = x - cp;
    const f3 n = normalized(diff);
    const float dis = dot(diff, n);
    if (!(dis < S.sp.contactRadius + S.sp.thickness)) return;
    const float w = order == 3 ? 1.f : -(order == 0 ? bary.x : (order == 1 ? bary.y : bary.z));
    float dEdD, d2E;
    repulsive(S.sp, dis, dEdD, d2E);
    const f3 cpPrev = bary.x * xyz(S.prev[fv.x]) + bary.y * xyz(S.prev[fv.y]) + bary.z * xyz(S.prev[fv.z]);
    const f3 dx = (x - xyz(S.prev[u])) - (cp - cpPrev);
    const float2 fa = C.friction[fv.x], fb = C.friction[u];
    contact_apply(S.sp.dt, w, -dEdD, d2E, n, dx, b - a, (fa.x + fb.x) * 0.5f, (fa.y + fb.y) * 0.5f, f, h);
}

// EE contact of edge e1 with e2, on the vertex of order `order` of e1 (weights 1 - mu1, mu1),
// VBDClothPhysics.cpp:1323-1399
__device__ void ee_contact(const SceneView& S, int e1, int e2, int order, int g, f3& f, Sym3& h) {
    const ContactView& C = S.cv;
    const int4 A = C.edges[e1], B = C.edges[e2];
    const f3 p1 = cpos(S, A.x, g), p2 = cpos(S, A.y, g), q1 = cpos(S, B.x, g), q2 = cpos(S, B.y, g);
    f3 c1, c2;
    float mu1, mu2;
    segment_closest(p1, p2, q1, q2, c1, c2, mu1, mu2);
    const f3 diff = c1 - c2;
    const float dis = sqrtf(dot(diff, diff));
    const f3 v1 = p2 - p1, v2 = q2 - q1;
    const f3 cr = cross(v1, v2);
    const float parallelEps = dot(cr, cr) / (sqrtf(dot(v1, v1)) * sqrtf(dot(v2, v2)));
    if (!(dis < S.sp.contactRadius + S.sp.thickness && parallelEps > 0.00001f)) return;
    const f3 n = (1.f / dis) * diff;
    const float w = order == 0 ? 1.f - mu1 : mu1;
    float dEdD, d2E;
    repulsive(S.sp, dis, dEdD, d2E);
    const f3 c1Prev = (1.f - mu1) * xyz(S.prev[A.x]) + mu1 * xyz(S.prev[A.y]);
    const f3 c2Prev = (1.f - mu2) * xyz(S.prev[B.x]) + mu2 * xyz(S.prev[B.y]);
    const f3 dx = (c1 - c1Prev) - (c2 - c2Prev);
    const float2 fa = C.friction[A.x], fb = C.friction[B.x];
    contact_apply(S.sp.dt, w, -dEdD, d2E, n, dx, v1, (fa.x + fb.x) * 0.5f, (fa.y + fb.y) * 0.5f, f, h);
}

// One penetration contact of the tet solver (VBDPhysicsCompute.cu:1831-1923): vertex u pushed out
// of face `face`, evaluated on the vertex of order `order` (3: u, weight 1; 0..2: a face vertex,
// weight -bary). The closest point is recomputed from the current positions; the normal is the
// face normal at an interior closest point, else the direction to the closest point turned to
// the face normal's side; with depth d = (c - x) . n > 0: force k d n and Hessian k n n^T plus
// friction (the contact_apply frame, t0 = b - a), times the weight (its square for the Hessian).
__device__ void pen_contact(const SceneView& S, int u, int face, int order, int g, f3& f, Sym3& h) {
    const ContactView& C = S.cv;
    const int4 fv = C.faces[face];
    const f3 x = cpos(S, u, g);
    const f3 a = cpos(S, fv.x, g), b = cpos(S, fv.y, g), c = cpos(S, fv.z, g);
    f3 cp, bary;
    const bool interior = closest_point_bary(x, a, b, c, cp, bary);
    const f3 fn = normalized(cross(b - a, c - a));
    const f3 diff = cp - x;
    f3 n = fn;
    if (!interior) {
        n = (1.f / sqrtf(dot(diff, diff))) * diff;
        if (dot(n, fn) < 0.f) n = -1.f * n;
    }
    const float depth = dot(diff, n);
    if (!(depth > 0.f)) return;
    const float w = order == 3 ? 1.f : -(order == 0 ? bary.x : (order == 1 ? bary.y : bary.z));
    const float k = C.penStiffness;
    // relative displacement over the step as the reference forms it (VBDPhysicsCompute.cu:1870-1887):
    // from each vertex's own displacement, so that vertices at rest in a direction give exactly zero
    // there (the friction's Hessian switches on at |u| > 0)
    const f3 dx = (x - xyz(S.prev[u])) - ((bary.x * (a - xyz(S.prev[fv.x])) + bary.y * (b - xyz(S.prev[fv.y]))) +
                                          bary.z * (c - xyz(S.prev[fv.z])));
    const float2 fa = C.friction[fv.x], fb = C.friction[u];
    contact_apply(S.sp.dt, w, depth * k, k, n, dx, b - a, (fa.x + fb.x) * 0.5f, (fa.y + fb.y) * 0.5f, f, h);
}

// VBDClothPhysics.cpp:1033-1094: the vertex's own VF contacts, the FV contacts of its incident
// faces (faceRef[fb, fe)) and the EE contacts of its incident edges (edgeRef[eb, ee))
// A vertex's contact pairs (VBDClothPhysics.cpp:1033-1094): its own VF list, the vertex lists of
// its incident faces (FV) and the edge lists of its incident edges (EE). Lane `sub` of the
// vertex's LPV lanes takes every LPV-th pair of each of the three sequences (the FV and EE
// sequences run over all incident primitives back to back), so the pairs spread evenly however
// they fall on the primitives; LPV = 1 is the reference's serial order.
template <int LPV>
__device__ void contact_forces(const SceneView& S, int v, int sub, const int* faceRef, int fb, int fe,
                               const int* edgeRef, int eb, int ee, f3& f, Sym3& h) {
    const ContactView& C = S.cv;
    const int g = C.vgroup[v];
    const int nvf = C.vfCount[v];
    if (C.model == 1) {
        // penetration contact: the vertex's own hit, then the hits on its incident faces
        for (int k = sub; k < nvf; k += LPV) pen_contact(S, v, C.vfRec[(int64_t)v * C.vfCap + k].x, 3, g, f, h);
        int t = sub, base = 0;
        for (int r = fb; r < fe; ++r) {
            const int ref = faceRef[r];
            const int face = ref & 0x3FFFFFFF, order = (ref >> 30) & 3;
            const int2 range = C.fvRange[face];
            const int end = base + (range.y - range.x);
            for (; t < end; t += LPV) pen_contact(S, C.fvVerts[range.x + (t - base)], face, order, g, f, h);
            base = end;
        }
        return;
    }
    for (int k = sub; k < nvf; k += LPV) vf_contact(S, v, C.vfRec[(int64_t)v * C.vfCap + k].x, 3, g, f, h);
    int t = sub, base = 0;
    for (int r = fb; r < fe; ++r) {
        const int ref = faceRef[r];
        const int face = ref & 0x3FFFFFFF, order = (ref >> 30) & 3;
        const int2 range = C.fvRange[face];
        const int end = base + (range.y - range.x);
        for (; t < end; t += LPV) vf_contact(S, C.fvVerts[range.x + (t - base)], face, order, g, f, h);
        base = end;
    }
    t = sub;
    base = 0;
    for (int r = eb; r < ee; ++r) {
        const int ref = edgeRef[r];
        const int e = ref & 0x3FFFFFFF, order = (ref >> 30) & 3;
        const int end = base + min(C.eeCount[e], C.eeCap);
        for (; t < end; t += LPV)
            ee_contact(S, e, __float_as_int(C.eeRec[(int64_t)e * C.eeCap + (t - base)].x), order, g, f, h);
        base = end;
    }
}

// ---------------------------------------------------------------------------
// Cloth vertex block (VBDClothPhysics.cpp:1026-1154)
// ---------------------------------------------------------------------------
template <int LPV, int STAGE, bool CONTACT>
__device__ __forceinline__ void cloth_block(const SceneView& S, const WorkItem& it, float4* lds) {
    const int local = threadIdx.x / LPV;
    const int sub = threadIdx.x % LPV;
    const bool active = local < it.count;

    int4 sa = make_int4(0, 0, 0, 0), sb = make_int4(0, 0, 0, 0);
    if (active) {
        const ClothSlot cs = S.cslots[it.slotBegin + local];
        sa = cs.a;
        sb = cs.b;
    }
    const int v = sa.x;
    f3 x = mk(0.f, 0.f, 0.f);
    if (active) x = xyz(S.pos[v]);

    const float4* __restrict__ pos = S.pos;
    int4 first = make_int4(0, 0, 0, 0), last = make_int4(0, 0, 0, 0);
    if (STAGE > 0) {
        const ClothSlot c0 = S.cslots[it.slotBegin], c1 = S.cslots[it.slotBegin + it.count - 1];
        first = make_int4(c0.a.y, c0.b.x, 0, 0);
        last = make_int4(c1.a.z, c1.b.y, 0, 0);
    }
    // membrane force and Hessian (kept apart for the StVK damping term)
    float fm0 = 0.f, fm1 = 0.f, fm2 = 0.f;
    float hxx = 0.f, hxy = 0.f, hxz = 0.f, hyy = 0.f, hyz = 0.f, hzz = 0.f;
    for_lane_records<LPV, 2 * STAGE, 2>(reinterpret_cast<const float4*>(S.cface), first.x, last.x, sa.y + sub, sa.z,
                                        lds, [&](const float4* r, int) {
        const float4 r0 = r[0];
        const float4 r1 = r[1];
        const int n0 = __float_as_int(r0.x), n1 = __float_as_int(r0.y);
        const float Amu = r0.z, Alm = r0.w;
        const f3 d0 = xyz(pos[n0]) - x;
        const f3 d1 = xyz(pos[n1]) - x;
        const f3 c0 = r1.x * d0 + r1.z * d1;  // F column 0
        const f3 c1 = r1.y * d0 + r1.w * d1;  // F column 1
        const float mv0 = -(r1.x + r1.z), mv1 = -(r1.y + r1.w);
        const float G00 = 0.5f * (dot(c0, c0) - 1.f), G11 = 0.5f * (dot(c1, c1) - 1.f), G01 = 0.5f * dot(c0, c1);
        const float trG = G00 + G11;
        // A S
        const float S00 = 2.f * Amu * G00 + Alm * trG, S11 = 2.f * Amu * G11 + Alm * trG, S01 = 2.f * Amu * G01;
        const float Sm0 = S00 * mv0 + S01 * mv1, Sm1 = S01 * mv0 + S11 * mv1;
        const f3 ff = Sm0 * c0 + Sm1 * c1;  // A F S m
        fm0 -= ff.x;
        fm1 -= ff.y;
        fm2 -= ff.z;
        const float mSm = mv0 * Sm0 + mv1 * Sm1;
        const float cm = Amu * (mv0 * mv0 + mv1 * mv1);
        const f3 Fm = mv0 * c0 + mv1 * c1;
        const float cf = Amu + Alm;
        hxx += mSm + cm * (c0.x * c0.x + c1.x * c1.x) + cf * Fm.x * Fm.x;
        hxy += cm * (c0.x * c0.y + c1.x * c1.y) + cf * Fm.x * Fm.y;
        hxz += cm * (c0.x * c0.z + c1.x * c1.z) + cf * Fm.x * Fm.z;
        hyy += mSm + cm * (c0.y * c0.y + c1.y * c1.y) + cf * Fm.y * Fm.y;
        hyz += cm * (c0.y * c0.z + c1.y * c1.z) + cf * Fm.y * Fm.z;
        hzz += mSm + cm * (c0.z * c0.z + c1.z * c1.z) + cf * Fm.z * Fm.z;
    });
    // bending: f -= sum_j kQ(o,j) X_j, H += kQ(o,o) I, skipping degenerate stencils
    float fb0 = 0.f, fb1 = 0.f, fb2 = 0.f, hb = 0.f;
    const float thr2 = S.sp.degenerateThresSq;
    for_lane_records<LPV, 2 * STAGE, 2>(reinterpret_cast<const float4*>(S.cbend), first.y, last.y, sb.x + sub, sb.y,
                                        lds, [&](const float4* r, int) {
        const float4 r0 = r[0];
        const float4 r1 = r[1];
        const int o = __float_as_int(r0.w);
        const f3 y0 = xyz(pos[__float_as_int(r0.x)]);
        const f3 y1 = xyz(pos[__float_as_int(r0.y)]);
        const f3 y2 = xyz(pos[__float_as_int(r0.z)]);
        // stencil in role order (eV1, eV2, eV12Next, eV21Next) with v at role o
        const f3 X0 = o == 0 ? x : y0;
        const f3 X1 = o == 0 ? y0 : (o == 1 ? x : y1);
        const f3 X2 = o <= 1 ? y1 : (o == 2 ? x : y2);
        const f3 X3 = o == 3 ? x : y2;
        const f3 n1 = cross(X1 - X0, X2 - X0);
        const f3 n2 = cross(X2 - X3, X1 - X3);
        if (dot(n1, n1) < thr2 || dot(n2, n2) < thr2) return;
        const f3 q = r1.x * y0 + r1.y * y1 + r1.z * y2 + r1.w * x;
        fb0 -= q.x;
        fb1 -= q.y;
        fb2 -= q.z;
        hb += r1.w;
    });
    if (LPV > 1) {
        fm0 = lane_sum<LPV>(fm0);
        fm1 = lane_sum<LPV>(fm1);
        fm2 = lane_sum<LPV>(fm2);
        hxx = lane_sum<LPV>(hxx);
        hxy = lane_sum<LPV>(hxy);
        hxz = lane_sum<LPV>(hxz);
        hyy = lane_sum<LPV>(hyy);
        hyz = lane_sum<LPV>(hyz);
        hzz = lane_sum<LPV>(hzz);
        fb0 = lane_sum<LPV>(fb0);
        fb1 = lane_sum<LPV>(fb1);
        fb2 = lane_sum<LPV>(fb2);
        hb = lane_sum<LPV>(hb);
    }
    // contact pairs, shared over the vertex's lanes
    f3 cf = mk(0.f, 0.f, 0.f);
    Sym3 ch{0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if constexpr (CONTACT) {
        if (active) contact_forces<LPV>(S, v, sub, S.cv.faceRef, sa.y, sa.z, S.cv.edgeRef, sb.z, sb.w, cf, ch);
        lane_sum_contact<LPV>(cf, ch);
    }
    if (!active || sub != 0) return;

    const SolveParams& sp = S.sp;
    const f3 xp = xyz(S.prev[v]);
    const f3 disp = x - xp;
    const float4 vp = S.vparam[v];  // (exp damping, const damping, StVK damping, bending damping)
    // membrane + its damping (VBDTriMeshStVK.cpp:904-909): dampingH = H_membrane * gamma / dt
    const float cs = vp.z / sp.dt;
    Sym3 hm{hxx, hxy, hxz, hyy, hyz, hzz};
    f3 f = mk(fm0, fm1, fm2) - cs * sym_mul(hm, disp);
    Sym3 h{hxx + cs * hxx, hxy + cs * hxy, hxz + cs * hxz, hyy + cs * hyy, hyz + cs * hyz, hzz + cs * hzz};
    // bending with damping (VBDTriMeshStVK.cpp:963-966)
    const float cb = vp.w / sp.dt;
    f = f + mk(fb0, fb1, fb2) - (hb * cb) * disp;
    const float hbd = hb * (1.f + cb);
    h.xx += hbd;
    h.yy += hbd;
    h.zz += hbd;
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
    if (sp.usePlaneGround) boundary_force(sp, sp.clothBoundaryStiffness, sp.clothFrictionEpsU, x, xp, true, f, h);
    if constexpr (CONTACT) {
        f = f + cf;
        h.xx += ch.xx; h.xy += ch.xy; h.xz += ch.xz;
        h.yy += ch.yy; h.yz += ch.yz; h.zz += ch.zz;
    }
    if (sa.w & SLOT_ATTACH) {
        // vertex attachments (VBDClothPhysics.cpp:1465-1498)
        for (int r = S.attBegin[v]; r < S.attBegin[v + 1]; ++r) {
            const AttachRec A = S.att[r];
            const f3 xr = xyz(S.pos[A.ids.w]);
            f3 tgt = mk(A.p.x, A.p.y, A.p.z);
            if (A.ids.x >= 0)
                tgt = A.p.x * xyz(S.pos[A.ids.x]) + A.p.y * xyz(S.pos[A.ids.y]) + A.p.z * xyz(S.pos[A.ids.z]);
            f = f - A.p.w * (xr - tgt);
            const float hk = A.p.w * A.b;
            h.xx += hk;
            h.yy += hk;
            h.zz += hk;
        }
    }
    f3 dx;
    f3 xn = x;
    const bool solved = solve_sym3(h, f, dx);
    if (!solved) atomicAdd(S.solveFail, 1);  // singular block: the vertex keeps its position, counted
    if (solved) {
        xn = x + sp.clothStepSize * dx;
        if constexpr (CONTACT) {
            // applyConservativeBoundTruncation (VBDClothPhysics.cpp:1223-1246); any truncation
            // requests detection before the next iteration
            const f3 x0 = xyz(S.cv.xAtCD[v]);
            const f3 acc = xn - x0;
            const float size = sqrtf(dot(acc, acc));
            const float b = S.cv.bound[v];
            if (size > b) {
                xn = x0 + (b / size) * acc;
                atomicOr(S.cv.flag, 1);
            }
        }
    }
    float4* dst = (sa.w & SLOT_INPLACE) ? S.pos : S.pnext;
    dst[v] = make_float4(xn.x, xn.y, xn.z, 0.f);
}

}  // namespace

// One launch = one parallel group of every scene referenced by `items`; the block's
// work kind is uniform, so the branch does not diverge.
// STAGE > 0 stages the incidence records through LDS, STAGE 64-byte records (or 2 * STAGE cloth
// records of 32 bytes) per chunk.
// CONTACT selects the variant for scenes with collision handling (contact forces, bound
// truncation); the others carry none of that code, so their register use is unchanged by it.
template <int LPV, int BLOCK, int STAGE, bool CONTACT>
__global__ __launch_bounds__(BLOCK) void group_solve_kernel(const SceneView* __restrict__ scenes,
                                                           const WorkItem* __restrict__ items, int iteration,
                                                           int xcdBlocks) {
    __shared__ float4 lds[STAGE > 0 ? 4 * STAGE : 1];
    const WorkItem it =
        items[xcdBlocks == 0 ? (int)blockIdx.x : (xcdBlocks == 1 ? xcd_block_index() : xcd_chunk_index(xcdBlocks))];
    const SceneView& S = scenes[it.scene];
    if (it.kind == WORK_TET)
        tet_block<LPV, STAGE, CONTACT>(S, it, iteration, lds);
    else
        cloth_block<LPV, STAGE, CONTACT>(S, it, lds);
}

// After a group's cloth solves: x = positionsNext for the group's cloth vertices
// (VBDClothPhysics.cpp:483-493).
__global__ __launch_bounds__(256) void cloth_commit_kernel(const SceneView* __restrict__ scenes,
                                                           const WorkItem* __restrict__ items) {
    const WorkItem it = items[blockIdx.x];
    const SceneView& S = scenes[it.scene];
    for (int i = threadIdx.x; i < it.count; i += blockDim.x) {
        const int v = S.cslots[it.slotBegin + i].a.x;
        S.pos[v] = S.pnext[v];
    }
}

// One whole VBD iteration of every scene of a plan in one persistent launch.
//
// The queue holds (scene, colour level, slot chunk) entries ordered by (level, scene); an
// entry may run once `need + epoch * perIteration` entries of its scene have completed
// (per-scene counters accumulate over the launches (epochs) since counter_reset_kernel last
// cleared them; `iteration` is the VBD iteration index, used for the friction start). Blocks
// take entries in queue order from the epoch's ticket counter, so every entry a block waits for
// was taken earlier by a block that is
// already running: the lowest unfinished entry never waits, and the grid drains without
// co-residency assumptions. A wait longer than `spinLimit` polls sets *status and proceeds, so
// no wave can spin forever; the host reports the flag as an error.
//
// Within a scene this keeps the reference's order exactly (colour g of iteration i sees colours
// < g of iteration i and colours > g of iteration i - 1), so results are bitwise those of one
// launch per parallel group; across scenes there is no barrier, so the tail of one colour
// overlaps the head of the next.
template <int LPV, int BLOCK>
__global__ __launch_bounds__(BLOCK) void queue_solve_kernel(const SceneView* __restrict__ scenes,
                                                           const QueueEntry* __restrict__ queue, int nq,
                                                           int* __restrict__ ticket, int* __restrict__ progress,
                                                           int iteration, int epoch, int* __restrict__ status,
                                                           int spinLimit) {
    constexpr int VPB = BLOCK / LPV;
    __shared__ int sTicket;
    // The scheduling work (count the previous entry, take a ticket, wait for its dependencies)
    // is done by the whole first wave, with lane 0 alone issuing the counter atomics: every
    // loop and branch around a barrier is then uniform within each wave, so no lane can be
    // parked in a loop that other lanes of its wave keep executing.
    const bool wave0 = threadIdx.x < 64;
    const int lane = threadIdx.x & 63;
    int prevScene = -1;
    for (;;) {
        if (wave0) {
            if (prevScene >= 0 && lane == 0)
                __hip_atomic_fetch_add(progress + prevScene, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            int t = 0;
            if (lane == 0) t = atomicAdd(ticket, 1);
            t = __builtin_amdgcn_readfirstlane(__shfl(t, 0, 64));
            if (t < nq) {
                const int scene = queue[t].scene;
                const int target = queue[t].need + epoch * queue[t].perIteration;
                for (int spins = 0;; ++spins) {
                    const int done = __builtin_amdgcn_readfirstlane(
                        __hip_atomic_load(progress + scene, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
                    if (done >= target) break;
                    if (spins >= spinLimit) {
                        if (lane == 0) atomicOr(status, 1);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            }
            if (lane == 0) sTicket = t;
        }
        // wave 0's agent-scope acquire happens before every thread's loads below through this
        // barrier (workgroup-scope synchronisation)
        __syncthreads();
        const int t = sTicket;
        if (t >= nq) break;
        const QueueEntry e = queue[t];
        const SceneView& S = scenes[e.scene];
        if (e.kind == WORK_COMMIT) {
            for (int i = threadIdx.x; i < e.count; i += BLOCK) {
                const int v = S.cslots[e.slotBegin + i].a.x;
                S.pos[v] = S.pnext[v];
            }
        } else {
            for (int p = 0; p < e.count; p += VPB) {
                const WorkItem it{e.scene, e.slotBegin + p, min(VPB, e.count - p), e.kind};
                if (e.kind == WORK_TET)
                    tet_block<LPV, 0, false>(S, it, iteration, nullptr);
                else
                    cloth_block<LPV, 0, false>(S, it, nullptr);
            }
        }
        // every thread's stores happen before lane 0's agent-scope release (the completion count
        // at the top of the next round) through this barrier: each wave first waits until its own
        // stores are acknowledged by L2 (s_waitcnt vmcnt(0); expcnt and lgkmcnt left free), so the
        // L2 write-back of that release covers them. The barrier also keeps sTicket until every
        // thread has read it.
        __builtin_amdgcn_s_waitcnt(0x0F70);
        __syncthreads();
        prevScene = e.scene;
    }
}

// Clears the work-queue counters (a kernel rather than a memset so that it can sit in a
// captured graph).
__global__ __launch_bounds__(256) void counter_reset_kernel(int* __restrict__ counters, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) counters[i] = 0;
}

// VBD_BaseMaterial.h:244-259 applyInitialStep (computeInertia + forwardStep) and
// VBDBaseTriMesh.cpp:163-178; the per-vertex initialisation code is
// 0: x = y, 1: keep x, 2: x + dt v_prev, 3: (x + y)/2, 4: x + dt v_prev + dt^2 r g,
// 5: x + (y - x) shortened to the vertex's conservative bound (cloth type 3).
__global__ __launch_bounds__(256) void initial_step_kernel(const SceneView* __restrict__ scenes,
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
    // curTime += dt at the start of the substep (one thread per scene)
    if (v == 0) S.clock[0] = S.clock[0] + dt;
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
    case 5: {
        // cloth inertia within the conservative bounds (VBDBaseTriMesh.cpp:137-152)
        f3 step = y - x;
        const float size = sqrtf(dot(step, step));
        if (S.initBound) {
            const float b = S.initBound[v];
            if (size > b && size != 0.f) step = (b / size) * step;
        }
        xn = x + step;
        break;
    }
    default: xn = x; break;
    }
    S.pos[v] = make_float4(xn.x, xn.y, xn.z, 0.f);
    S.vel[v] = make_float4(vn.x, vn.y, vn.z, 0.f);
}

// Deformers (VBD/VBDDeformer.h:45-178, VBDCloth/VBDClothDeformer.h:44-196). Items cover one
// phase's deformed vertices: cloth = 1 for cloth meshes' (launched before the substep's initial
// step, so the substep's time is clock + dt), 0 for tet meshes' (after it: clock). Rotator:
// p = root + R(theta) RootP, theta = min(t, end) omega (end > 0), R's entries in double rounded
// to float; tet Translator: p = speed min(t, end) + p0; cloth Translator: p += speed dt while
// start <= t and (t < end or end < 0). Products and sums unfused, as the reference's float code.
__global__ __launch_bounds__(256) void deform_kernel(const SceneView* __restrict__ scenes,
                                                     const WorkItem* __restrict__ items, int cloth) {
    const WorkItem it = items[blockIdx.x];
    const int i = threadIdx.x;
    if (i >= it.count) return;
    const SceneView& S = scenes[it.scene];
    const DeformVert r = S.dverts[it.slotBegin + i];
    const DeformParams P = S.dparams[r.d];
    const float dt = S.sp.dt;
    const float t = cloth ? __fadd_rn(S.clock[0], dt) : S.clock[0];
    float p[3];
    if (P.kind == 0) {
        double tc = t;
        if (tc >= P.endTime && P.endTime > 0) tc = cloth ? (double)(float)P.endTime : P.endTime;
        const double theta = tc * P.omega;
        const double c = cos(theta), s = sin(theta);
        const double ux = P.axis[0], uy = P.axis[1], uz = P.axis[2];
        const float R[9] = {(float)(c + ux * ux * (1 - c)), (float)(ux * uy * (1 - c) - uz * s),
                            (float)(ux * uz * (1 - c) + uy * s), (float)(uy * ux * (1 - c) + uz * s),
                            (float)(c + uy * uy * (1 - c)),      (float)(uy * uz * (1 - c) - ux * s),
                            (float)(uz * ux * (1 - c) - uy * s), (float)(uz * uy * (1 - c) + ux * s),
                            (float)(c + uz * uz * (1 - c))};
        for (int k = 0; k < 3; ++k) {
            const float q = __fadd_rn(__fadd_rn(__fmul_rn(R[3 * k], r.b[0]), __fmul_rn(R[3 * k + 1], r.b[1])),
                                      __fmul_rn(R[3 * k + 2], r.b[2]));
            p[k] = __fadd_rn(r.a[k], q);
        }
    } else if (!cloth) {
        float tc = t;
        const float end = (float)P.endTime;
        if (tc >= end && end > 0.f) tc = end;
        for (int k = 0; k < 3; ++k) p[k] = __fadd_rn(__fmul_rn(P.speed[k], tc), r.a[k]);
    } else {
        const float end = (float)P.endTime;
        if (!(t >= P.startTime && (t < end || end < 0.f))) return;
        const float4 x = S.pos[r.v];
        p[0] = __fadd_rn(x.x, __fmul_rn(P.speed[0], dt));
        p[1] = __fadd_rn(x.y, __fmul_rn(P.speed[1], dt));
        p[2] = __fadd_rn(x.z, __fmul_rn(P.speed[2], dt));
    }
    S.pos[r.v] = make_float4(p[0], p[1], p[2], 0.f);
}

// Chebyshev acceleration (runStepGPU, VBDPhysics.cpp:872-906): after the initial step both
// history buffers take the current positions (recordInitialPositionForAccelerator, :3577-3592).
__global__ __launch_bounds__(256) void accel_init_kernel(const SceneView* __restrict__ scenes,
                                                         const WorkItem* __restrict__ items) {
    const WorkItem it = items[blockIdx.x];
    const int i = threadIdx.x;
    if (i >= it.count) return;
    const SceneView& S = scenes[it.scene];
    if (!S.accOmega) return;
    const int v = it.slotBegin + i;
    if (!(S.vflag[v] & VFLAG_ACCEL)) return;
    const float4 p = S.pos[v];
    S.accPrevPrev[v] = p;
    S.accPrevIter[v] = p;
}

// After iteration k's sweep (VBDSolveParallelGroup_applyAcceleration_kernel,
// VBDPhysicsCompute.cu:910-944): x <- y_{k-1} + omega_k (x - y_{k-1}); y_{k-1} <- y_k (the positions
// before this sweep); and the accelerated positions become the next sweep's y_k
// (recordPrevIterPositionsAccelerator, VBDPhysics.cpp:3594-3606). Fixed vertices are left alone.
__global__ __launch_bounds__(256) void accel_apply_kernel(const SceneView* __restrict__ scenes,
                                                          const WorkItem* __restrict__ items, int iteration) {
    const WorkItem it = items[blockIdx.x];
    const int i = threadIdx.x;
    if (i >= it.count) return;
    const SceneView& S = scenes[it.scene];
    if (!S.accOmega) return;
    const int v = it.slotBegin + i;
    const uint8_t flag = S.vflag[v];
    if (!(flag & VFLAG_ACCEL) || (flag & VFLAG_FIXED)) return;
    const float omega = S.accOmega[iteration];
    const f3 x = xyz(S.pos[v]);
    const f3 pp = xyz(S.accPrevPrev[v]);
    const f3 xn = pp + omega * (x - pp);
    S.pos[v] = make_float4(xn.x, xn.y, xn.z, 0.f);
    S.accPrevPrev[v] = S.accPrevIter[v];
    S.accPrevIter[v] = make_float4(xn.x, xn.y, xn.z, 0.f);
}

// VBDPhysics.cpp:2486-2561 updateVelocities + dampVelocities (cloth: VBDClothPhysics.cpp:2318-2360).
__global__ __launch_bounds__(256) void update_velocity_kernel(const SceneView* __restrict__ scenes,
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
    const uint8_t flags = S.vflag[v];
    // no-gravity zone (VBDPhysics.cpp:2534-2542 on the position's y, VBDClothPhysics.cpp:2330-2338
    // on the velocity's y): the speed is capped at maxVelocityMagnitude, no damping
    bool zone = false;
    float2 z = make_float2(0.f, 0.f);
    if (flags & (VFLAG_NOGRAV_POS | VFLAG_NOGRAV_VEL)) {
        z = S.vzone[v];
        zone = ((flags & VFLAG_NOGRAV_POS) ? p.y : vel.y) > z.x;
    }
    if (zone) {
        if (vMag > z.y) vel = (z.y / vMag) * vel;
    } else if (vMag > 1e-6f) {
        const float4 vp = S.vparam[v];
        float vNew = vMag * vp.x - vp.y;
        vNew = vNew > 1e-6f ? vNew : 0.f;
        vel = (vNew / vMag) * vel;
    } else {
        vel = 0.f * vel;
    }
    if (flags & VFLAG_FIXED) vel = mk(0.f, 0.f, 0.f);
    S.vel[v] = make_float4(vel.x, vel.y, vel.z, 0.f);
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
template <int LPV, int BLOCK, int STAGE, bool CONTACT>
static hipError_t launch_solve(const SceneView* scenes, const WorkItem* items, int nitems, int iteration,
                               int xcd, hipStream_t stream) {
    hipLaunchKernelGGL((group_solve_kernel<LPV, BLOCK, STAGE, CONTACT>), dim3(nitems), dim3(BLOCK), 0, stream, scenes,
                       items, iteration, xcd);
    return hipGetLastError();
}

// LDS-staged variants: the lane/block shapes the autotuner finds competitive, two chunk sizes.
#define GVBD_STAGED(X) X(4, 128) X(4, 256) X(8, 128) X(8, 256) X(16, 128) X(16, 256)

hipError_t launch_group_solve(const SceneView* scenes, const WorkItem* items, int nitems, int iteration, int lpv,
                              int block, int stage, bool contact, int xcd, hipStream_t stream) {
    if (nitems <= 0) return hipSuccess;
#define GVBD_CASE(L, B)                                                                                        \
    if (stage == 0 && lpv == L && block == B)                                                                  \
        return contact ? launch_solve<L, B, 0, true>(scenes, items, nitems, iteration, xcd, stream)            \
                       : launch_solve<L, B, 0, false>(scenes, items, nitems, iteration, xcd, stream);
    GVBD_CASE(1, 64) GVBD_CASE(1, 128) GVBD_CASE(1, 256)
    GVBD_CASE(2, 64) GVBD_CASE(2, 128) GVBD_CASE(2, 256)
    GVBD_CASE(4, 64) GVBD_CASE(4, 128) GVBD_CASE(4, 256)
    GVBD_CASE(8, 64) GVBD_CASE(8, 128) GVBD_CASE(8, 256)
    GVBD_CASE(16, 64) GVBD_CASE(16, 128) GVBD_CASE(16, 256)
#undef GVBD_CASE
    if (contact) return hipErrorInvalidValue;  // staged variants carry no contact code
#define GVBD_CASE(L, B)                                                                                  \
    if (lpv == L && block == B && stage == 32)                                                           \
        return launch_solve<L, B, 32, false>(scenes, items, nitems, iteration, xcd, stream);             \
    if (lpv == L && block == B && stage == 64)                                                           \
        return launch_solve<L, B, 64, false>(scenes, items, nitems, iteration, xcd, stream);
    GVBD_STAGED(GVBD_CASE)
#undef GVBD_CASE
    return hipErrorInvalidValue;
}

bool valid_stage(int lpv, int block, int stage) {
    if (stage == 0) return true;
    if (stage != 32 && stage != 64) return false;
    bool ok = false;
#define GVBD_CASE(L, B) ok = ok || (lpv == L && block == B);
    GVBD_STAGED(GVBD_CASE)
#undef GVBD_CASE
    return ok;
}

template <int LPV, int BLOCK>
static hipError_t launch_queue(const QueueArgs& a, int iteration, int epoch, int grid, hipStream_t stream) {
    hipLaunchKernelGGL((queue_solve_kernel<LPV, BLOCK>), dim3(grid), dim3(BLOCK), 0, stream, a.scenes, a.queue,
                       a.nq, a.tickets + epoch, a.progress, iteration, epoch, a.status, a.spinLimit);
    return hipGetLastError();
}

template <int LPV, int BLOCK>
static int queue_occupancy() {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, queue_solve_kernel<LPV, BLOCK>, BLOCK, 0) != hipSuccess) {
        (void)hipGetLastError();
        return 1;
    }
    return n > 0 ? n : 1;
}

#define GVBD_SCHEDULES(X)                                        \
    X(1, 64) X(1, 128) X(1, 256) X(2, 64) X(2, 128) X(2, 256)    \
    X(4, 64) X(4, 128) X(4, 256) X(8, 64) X(8, 128) X(8, 256)    \
    X(16, 64) X(16, 128) X(16, 256)

hipError_t launch_queue_solve(const QueueArgs& a, int iteration, int epoch, int grid, int lpv, int block,
                              hipStream_t stream) {
    if (a.nq <= 0) return hipSuccess;
    if (grid <= 0) return hipErrorInvalidValue;
#define GVBD_CASE(L, B) \
    if (lpv == L && block == B) return launch_queue<L, B>(a, iteration, epoch, grid, stream);
    GVBD_SCHEDULES(GVBD_CASE)
#undef GVBD_CASE
    return hipErrorInvalidValue;
}

int queue_blocks_per_cu(int lpv, int block) {
#define GVBD_CASE(L, B) \
    if (lpv == L && block == B) return queue_occupancy<L, B>();
    GVBD_SCHEDULES(GVBD_CASE)
#undef GVBD_CASE
    return 1;
}

hipError_t launch_counter_reset(int* counters, int n, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const int blocks = std::min((n + 255) / 256, 1024);
    hipLaunchKernelGGL(counter_reset_kernel, dim3(blocks), dim3(256), 0, stream, counters, n);
    return hipGetLastError();
}

hipError_t launch_cloth_commit(const SceneView* scenes, const WorkItem* items, int nitems, hipStream_t stream) {
    if (nitems <= 0) return hipSuccess;
    hipLaunchKernelGGL(cloth_commit_kernel, dim3(nitems), dim3(256), 0, stream, scenes, items);
    return hipGetLastError();
}

hipError_t launch_accel_init(const SceneView* scenes, const WorkItem* items, int nitems, hipStream_t stream) {
    if (nitems <= 0) return hipSuccess;
    hipLaunchKernelGGL(accel_init_kernel, dim3(nitems), dim3(256), 0, stream, scenes, items);
    return hipGetLastError();
}

hipError_t launch_accel_apply(const SceneView* scenes, const WorkItem* items, int nitems, int iteration,
                              hipStream_t stream) {
    if (nitems <= 0) return hipSuccess;
    hipLaunchKernelGGL(accel_apply_kernel, dim3(nitems), dim3(256), 0, stream, scenes, items, iteration);
    return hipGetLastError();
}

hipError_t launch_deform(const SceneView* scenes, const WorkItem* items, int nitems, bool cloth, hipStream_t stream) {
    if (nitems > 0)
        hipLaunchKernelGGL(deform_kernel, dim3(nitems), dim3(256), 0, stream, scenes, items, cloth ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_initial_step(const SceneView* scenes, const WorkItem* items, int nitems, hipStream_t stream) {
    if (nitems <= 0) return hipSuccess;
    hipLaunchKernelGGL(initial_step_kernel, dim3(nitems), dim3(256), 0, stream, scenes, items);
    return hipGetLastError();
}

hipError_t launch_update_velocity(const SceneView* scenes, const WorkItem* items, int nitems, hipStream_t stream) {
    if (nitems <= 0) return hipSuccess;
    hipLaunchKernelGGL(update_velocity_kernel, dim3(nitems), dim3(256), 0, stream, scenes, items);
    return hipGetLastError();
}

// Plain streaming copy (float4 per lane, grid-stride): the reference kernel
// whose HBM counters calibrate the profiler and whose rate is the measured
// copy bandwidth of the device.
// Grid-stride copy, one 16-byte non-temporal load and store per lane and step (the fastest of
// the streaming variants measured on MI355X: 1-4 loads in flight per lane, 1k-64k blocks,
// temporal or non-temporal accesses).
__global__ __launch_bounds__(256) void stream_copy_kernel(const float4* __restrict__ src, float4* __restrict__ dst,
                                                          int64_t n) {
    typedef float f4v __attribute__((ext_vector_type(4)));
    const f4v* s = reinterpret_cast<const f4v*>(src);
    f4v* d = reinterpret_cast<f4v*>(dst);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        __builtin_nontemporal_store(__builtin_nontemporal_load(s + i), d + i);
}

hipError_t launch_stream_copy(const float4* src, float4* dst, int64_t n, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(stream_copy_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, src, dst, n);
    return hipGetLastError();
}

bool valid_schedule(int lpv, int block) {
    const bool l = lpv == 1 || lpv == 2 || lpv == 4 || lpv == 8 || lpv == 16;
    const bool b = block == 64 || block == 128 || block == 256;
    return l && b;
}

}  // namespace gvbd
