// CPU oracle: restatement of the reference's CPU tet VBD path.
// TEST INFRASTRUCTURE ONLY (see oracle.h). Never linked by the product.
//
// Reference files restated here (paths relative to Simulator/Modules/):
//   TetMesh/TetMeshFEM.cpp:332-443        rest-pose precompute (DmInv, volume, lumped mass)
//   TetMesh/TetMeshFEM.cpp:275-325        vertex -> tet incidence (order here: ascending tet id;
//                                         MeshFrame's V_TV order is unavailable: parity unpinned)
//   TetMesh/TetMeshFEM.cpp:122-133        colour categories sorted by size, descending
//   VBD/VBDPhysics.cpp:99-122,235-274     parallel-group merge of per-mesh colours
//   VBD/VBDPhysics.cpp:1580-1632          runStep_serialCollisionHandling (no DCD/CCD)
//   VBD/VBD_BaseMaterial.h:201-259        applyInitialStep / computeInertia
//   VBD/VBD_BaseMaterial.cpp:7-50         forwardStep (initialisation types 0-4)
//   VBD/VBDPhysics.cpp:2402-2483          VBDStepWithCollision
//   VBD/VBD_NeoHookean.cpp:126-152,379-581 neo-Hookean force/Hessian with full 9x9 assembly
//   VBD/VBD_BaseMaterial.h:359-366        inertia force/Hessian
//   VBD/VBDPhysics.cpp:2821-2954,2563-2592 plane-ground boundary penalty + IPC friction
//   VBD/VBDPhysics.cpp:2486-2561,794-807  velocity update + damping (step-invariant scaling)
//   tet_collision.h                       penetration contact (DCD / CCD) of runStepGPU
#include "oracle.h"
#include "deform.h"
#include "pool.h"
#include "tet_collision.h"
#include "tet_core.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

namespace {

using namespace tetcore;

template <class R> struct V3 {
    R x, y, z;
};

template <class R> inline bool solve3(const R* h, const R* b, R* x) {
    // a_ij = h[i + 3j]; C_ij = cofactor of a_ij
    const R a00 = h[0], a10 = h[1], a20 = h[2];
    const R a01 = h[3], a11 = h[4], a21 = h[5];
    const R a02 = h[6], a12 = h[7], a22 = h[8];
    const R C00 = a11 * a22 - a12 * a21;
    const R C01 = a12 * a20 - a10 * a22;
    const R C02 = a10 * a21 - a11 * a20;
    const R C10 = a02 * a21 - a01 * a22;
    const R C11 = a00 * a22 - a02 * a20;
    const R C12 = a01 * a20 - a00 * a21;
    const R C20 = a01 * a12 - a02 * a11;
    const R C21 = a02 * a10 - a00 * a12;
    const R C22 = a00 * a11 - a01 * a10;
    const R det = a00 * C00 + a01 * C01 + a02 * C02;
    const R scale = std::fabs(a00 * a11 * a22);
    if (!std::isfinite(det) || !(std::fabs(det) > R(1e-12) * scale)) return false;
    // (A^-1)_ij = C_ji / det
    const R inv = R(1) / det;
    x[0] = (C00 * b[0] + C10 * b[1] + C20 * b[2]) * inv;
    x[1] = (C01 * b[0] + C11 * b[1] + C21 * b[2]) * inv;
    x[2] = (C02 * b[0] + C12 * b[1] + C22 * b[2]) * inv;
    return std::isfinite(x[0]) && std::isfinite(x[1]) && std::isfinite(x[2]);
}
// VBDPhysics.cpp:2563-2592 computeVertexFriction (IPC mollifier; the u u^T
// term is dropped in the reference's Hessian, and here).
// T is 3x2 column-major, u = T^T dx.
template <class R>
inline void vertex_friction(R mu, R lambda, const R* T, const R* u, R epsU, R* f, R* H) {
    R uNorm = std::sqrt(u[0] * u[0] + u[1] * u[1]);
    if (uNorm > 0) {
        R f1_SF_over_x;
        if (uNorm > epsU) {
            f1_SF_over_x = R(1) / uNorm;
        } else {
            f1_SF_over_x = (-uNorm / epsU + R(2)) / epsU;
        }
        for (int r = 0; r < 3; ++r) f[r] = -mu * lambda * (T[r] * u[0] + T[r + 3] * u[1]) * f1_SF_over_x;
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 3; ++r)
                H[r + 3 * c] = mu * lambda * f1_SF_over_x * (T[r] * T[c] + T[r + 3] * T[c + 3]);
    } else {
        for (int i = 0; i < 3; ++i) f[i] = 0;
        for (int i = 0; i < 9; ++i) H[i] = 0;
    }
}

template <class R> struct Mesh {
    int nv = 0, nt = 0;
    std::vector<R> x, xprev, v, vprev, y, dminv, vol, mass;
    std::vector<R> accPrevPrev, accPrevIter;  // Chebyshev acceleration history
    std::vector<int32_t> tets;
    std::vector<uint8_t> fixed;
    std::vector<int32_t> fixedList;
    std::vector<int32_t> incBegin, incTet, incOrder;
    std::vector<std::vector<int32_t>> categories;
    R mu, lmbd, dampH, dampD, expVelDamp, constVelDamp, density;
    int initType = 0;
    bool noGravZone = false;        // hasNoGravZone (with maxVelocityMagnitude > 0)
    R noGravThreshold = 0, maxVel = -1;
    R frictionMu = 0, frictionEpsV = 0;  // contact friction (frictionDynamic, frictionEpsV)
};

template <class R> struct Scene {
    double p[OP_COUNT];
    R dt, dtSqrRecip;
    std::vector<Mesh<R>> meshes;
    std::vector<std::vector<int32_t>> groups;  // interleaved (mesh, vertex)
    oracle::ThreadPool* pool = nullptr;
    std::vector<oracle::Deformer> deformers;
    float clock = 0.f;  // curTime (float in the reference)
    std::atomic<long long> solveFailures{0};  // vertex solves whose 3x3 block had no solution
    // penetration contact (tet_collision.h): per mesh its surface topology, per vertex the last
    // detection (hit mesh and face, whether a tet contained it, whether CCD found it) and the
    // relations (colliding mesh, colliding vertex, order) of the vertices of hit faces
    bool collide = false;
    bool boxPrefilter = true;  // skip meshes / faces by box in the brute-force queries
    tetcol::Params col;
    std::vector<tetcol::Topo> topo;
    std::vector<std::vector<tetcol::Hit>> hits;
    std::vector<std::vector<uint8_t>> penetrated, fromCCD;
    std::vector<std::vector<std::vector<std::array<int, 3>>>> relations;
    std::vector<std::vector<R>> snap;  // positions at the start of the current parallel group
};

// VBD_BaseMaterial.h:244-259, 201-221; VBD_BaseMaterial.cpp:7-50.
template <class R> void apply_initial_step(Scene<R>& s, Mesh<R>& m) {
    const R dt = s.dt;
    const R g[3] = {(R)s.p[OP_GRAVITY_X], (R)s.p[OP_GRAVITY_Y], (R)s.p[OP_GRAVITY_Z]};
    const int initType = m.initType;
    m.xprev = m.x;
    m.vprev = m.v;
    // external force is zero (evaluateExternalForce sets it to 0)
    if (s.p[OP_ASSOCIATE_GRAVITY] != 0) {
        R dg[3] = {dt * g[0], dt * g[1], dt * g[2]};
        for (int i = 0; i < m.nv; ++i)
            for (int d = 0; d < 3; ++d) m.v[3 * i + d] += dg[d];
    }
    for (int f : m.fixedList)
        for (int d = 0; d < 3; ++d) m.v[3 * f + d] = 0;
    for (int i = 0; i < 3 * m.nv; ++i) m.y[i] = m.x[i] + dt * m.v[i];
    switch (initType) {
    case 0: m.x = m.y; break;
    case 1: break;
    case 2:
        for (int i = 0; i < 3 * m.nv; ++i) m.x[i] = m.xprev[i] + dt * m.vprev[i];
        break;
    case 3:
        for (int i = 0; i < 3 * m.nv; ++i) m.x[i] = R(0.5) * m.x[i] + R(0.5) * m.y[i];
        break;
    case 4: {
        const R r = (R)s.p[OP_INIT_RATIO_G];
        for (int i = 0; i < m.nv; ++i)
            for (int d = 0; d < 3; ++d) m.x[3 * i + d] = m.x[3 * i + d] + dt * m.vprev[3 * i + d] + dt * dt * r * g[d];
        break;
    }
    default: break;
    }
}

template <class R> void collision_terms(Scene<R>& s, int mi, int vid, R* f, R* h);

// VBDPhysics.cpp:2402-2483 VBDStepWithCollision (no line search); with penetration contact its
// collision terms follow the boundary terms
template <class R> void vbd_vertex(Scene<R>& s, Mesh<R>& m, int vid, bool applyFriction, int mi = -1) {
    R h[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    R f[3] = {0, 0, 0};
    R* xv = &m.x[3 * vid];
    const R disp[3] = {xv[0] - m.xprev[3 * vid], xv[1] - m.xprev[3 * vid + 1], xv[2] - m.xprev[3 * vid + 2]};
    for (int k = m.incBegin[vid]; k < m.incBegin[vid + 1]; ++k) {
        const int t = m.incTet[k];
        R x4[12];
        for (int j = 0; j < 4; ++j)
            for (int d = 0; d < 3; ++d) x4[3 * j + d] = m.x[3 * m.tets[4 * t + j] + d];
        neohookean_vertex<R>(x4, &m.dminv[9 * t], m.vol[t], m.mu, m.lmbd, m.incOrder[k], m.dampH, m.dampD, s.dt,
                             disp, f, h);
    }
    // inertia (VBD_BaseMaterial.h:359-366)
    const R mass = m.mass[vid];
    for (int d = 0; d < 3; ++d) f[d] += mass * (m.y[3 * vid + d] - xv[d]) * s.dtSqrRecip;
    R hi = mass * s.dtSqrRecip;
    h[0] += hi; h[4] += hi; h[8] += hi;
    // plane-ground boundary (VBDPhysics.cpp:2877-2952)
    if (s.p[OP_USE_PLANE_GROUND] != 0) {
        const R k = (R)s.p[OP_BOUNDARY_STIFFNESS];
        for (int dim = 0; dim < 3; ++dim) {
            const R lower = (R)s.p[OP_LOWER_X + dim];
            const R upper = (R)s.p[OP_UPPER_X + dim];
            R pen;
            if (xv[dim] < lower) {
                pen = lower - xv[dim];
                f[dim] += pen * k;
            } else if (xv[dim] > upper) {
                pen = xv[dim] - upper;
                f[dim] -= pen * k;
            } else {
                continue;
            }
            h[dim + 3 * dim] += k;
            if (applyFriction) {
                R dx[3] = {xv[0] - m.xprev[3 * vid], xv[1] - m.xprev[3 * vid + 1], xv[2] - m.xprev[3 * vid + 2]};
                R T[6] = {0, 0, 0, 0, 0, 0};
                T[(dim + 1) % 3] = 1;
                T[3 + (dim + 2) % 3] = 1;
                R u[2] = {T[0] * dx[0] + T[1] * dx[1] + T[2] * dx[2], T[3] * dx[0] + T[4] * dx[1] + T[5] * dx[2]};
                R lambda = pen * k;
                R mu = (R)s.p[OP_BOUNDARY_FRICTION_MU];
                R epsU = (R)s.p[OP_BOUNDARY_FRICTION_EPS_V] * s.dt;
                R ff[3], fh[9];
                vertex_friction<R>(mu, lambda, T, u, epsU, ff, fh);
                for (int d = 0; d < 3; ++d) f[d] += ff[d];
                for (int i = 0; i < 9; ++i) h[i] += fh[i];
            }
        }
    }
    if (s.collide && mi >= 0) collision_terms(s, mi, vid, f, h);
    if (f[0] * f[0] + f[1] * f[1] + f[2] * f[2] > R(1e-10)) {
        R dx[3];
        R step = (R)s.p[OP_STEP_SIZE];
        if (!solve3<R>(h, f, dx)) {
            step = (R)s.p[OP_STEP_SIZE_GD];
            dx[0] = f[0]; dx[1] = f[1]; dx[2] = f[2];
            ++s.solveFailures;
        }
        for (int d = 0; d < 3; ++d) xv[d] += step * dx[d];
    }
}

// VBDPhysics.cpp:2486-2561 updateVelocities + dampVelocities.
template <class R> void update_velocities(Scene<R>& s, Mesh<R>& m) {
    for (int i = 0; i < m.nv; ++i) {
        R* v = &m.v[3 * i];
        for (int d = 0; d < 3; ++d) v[d] = (m.x[3 * i + d] - m.xprev[3 * i + d]) / s.dt;
        tetcore::damp_velocity<R>(v, m.noGravZone, m.x[3 * i + 1], m.noGravThreshold, m.maxVel, m.expVelDamp,
                                  m.constVelDamp);
    }
    for (int f : m.fixedList)
        for (int d = 0; d < 3; ++d) m.v[3 * f + d] = 0;
}

// VBDPhysics::getAcceleratorOmega (VBDPhysics.cpp:3558-3575), order from 1
template <class R> R accelerator_omega(int order, R rho, R prevOmega) {
    if (order == 1) return R(1);
    if (order == 2) return R(2) / (R(2) - rho * rho);
    return R(4) / (R(4) - rho * rho * prevOmega);
}

// The collision terms of vertex vid of mesh mi (VBDPhysicsCompute.cu:1731-1770): its own
// collision (weight 1) and the collisions whose closest face it belongs to (weight -bary), every
// position taken at the start of the parallel group
template <class R> void collision_terms(Scene<R>& s, int mi, int vid, R* f, R* h) {
    using tetcol::ld;
    auto contact = [&](int qm, int qv, int order) {
        const tetcol::Hit& hit = s.hits[qm][qv];
        const tetcol::Topo& T = s.topo[hit.mesh];
        const int32_t* fv = &T.faces[3 * hit.face];
        const R* Xq = s.snap[qm].data();
        const R* Xf = s.snap[hit.mesh].data();
        const Mesh<R>& mq = s.meshes[qm];
        const Mesh<R>& mf = s.meshes[hit.mesh];
        // contact friction: the average of the two meshes' coefficients
        const R mu = (mq.frictionMu + mf.frictionMu) * R(0.5), epsV = (mq.frictionEpsV + mf.frictionEpsV) * R(0.5);
        tetcol::penetration_force<R>(ld(Xq, qv), ld(Xf, fv[0]), ld(Xf, fv[1]), ld(Xf, fv[2]), ld(mq.xprev.data(), qv),
                                     ld(mf.xprev.data(), fv[0]), ld(mf.xprev.data(), fv[1]),
                                     ld(mf.xprev.data(), fv[2]), order, (R)s.col.stiffness, mu, epsV, s.dt, f, h);
    };
    if (s.hits[mi][vid].face >= 0) contact(mi, vid, 3);
    for (const auto& r : s.relations[mi][vid]) contact(r[0], r[1], r[2]);
}

// dcd() / ccd() of runStepGPU (VBDPhysics.cpp:1958-2014, 2066-2102) and prepareCollisionData
// (:1708-1760): detection at the start of the substep (DCD) and after the initial step (CCD, for
// surface vertices no tet contained); relations in ascending (mesh, vertex) order
template <class R> void detect_dcd(Scene<R>& s) {
    const int nm = (int)s.meshes.size();
    std::vector<const R*> X(nm);
    std::vector<tetcol::Box3<R>> boxes(nm);
    for (int m = 0; m < nm; ++m) {
        X[m] = s.meshes[m].x.data();
        for (int v = 0; v < s.meshes[m].nv; ++v) boxes[m].add(tetcol::ld(X[m], v));
        boxes[m].pad(R(1e-3) * boxes[m].extent() + R(1e-6));
    }
    for (int m = 0; m < nm; ++m) {
        auto& H = s.hits[m];
        std::fill(H.begin(), H.end(), tetcol::Hit{});
        std::fill(s.penetrated[m].begin(), s.penetrated[m].end(), 0);
        std::fill(s.fromCCD[m].begin(), s.fromCCD[m].end(), 0);
        if (!s.col.dcd) continue;
        const tetcol::Topo& T = s.topo[m];
        auto body = [&](int v) {
            if (!T.surf[v] || !(s.meshes[m].x[3 * v + 1] < (R)s.col.offHeight)) return;
            bool pen = false;
            H[v] = tetcol::dcd_vertex<R>(s.topo, X, s.col, m, v, pen, s.boxPrefilter ? boxes.data() : nullptr);
            s.penetrated[m][v] = pen;
        };
        if (s.pool) s.pool->parallel_for(0, T.nv, body);
        else for (int v = 0; v < T.nv; ++v) body(v);
    }
}

template <class R> void detect_ccd(Scene<R>& s) {
    const int nm = (int)s.meshes.size();
    if (s.col.ccd) {
        std::vector<const R*> X(nm), Xp(nm);
        std::vector<tetcol::Box3<R>> swept(nm);
        std::vector<R> move(nm, R(0));
        for (int m = 0; m < nm; ++m) {
            X[m] = s.meshes[m].x.data();
            Xp[m] = s.meshes[m].xprev.data();
            for (int v = 0; v < s.meshes[m].nv; ++v) {
                const tetcol::P3<R> a = tetcol::ld(Xp[m], v), b = tetcol::ld(X[m], v);
                swept[m].add(a);
                swept[m].add(b);
                move[m] = std::max(move[m], tetcol::norm(b - a));
            }
        }
        for (int m = 0; m < nm; ++m) {
            const tetcol::Topo& T = s.topo[m];
            auto body = [&](int v) {
                if (!T.surf[v] || s.penetrated[m][v] || !(s.meshes[m].x[3 * v + 1] < (R)s.col.offHeight)) return;
                s.hits[m][v] = s.boxPrefilter ? tetcol::ccd_vertex<R>(s.topo, X, Xp, m, v, swept.data(), move.data())
                                              : tetcol::ccd_vertex<R>(s.topo, X, Xp, m, v);
                s.fromCCD[m][v] = s.hits[m][v].face >= 0;
            };
            if (s.pool) s.pool->parallel_for(0, T.nv, body);
            else for (int v = 0; v < T.nv; ++v) body(v);
        }
    }
    for (int m = 0; m < nm; ++m)
        for (auto& r : s.relations[m]) r.clear();
    for (int m = 0; m < nm; ++m)
        for (int v = 0; v < s.topo[m].nv; ++v) {
            const tetcol::Hit& hit = s.hits[m][v];
            if (hit.face < 0) continue;
            const int32_t* fv = &s.topo[hit.mesh].faces[3 * hit.face];
            for (int k = 0; k < 3; ++k) s.relations[hit.mesh][fv[k]].push_back({m, v, k});
        }
}

template <class R> void scene_step(Scene<R>& s, int nthreads) {
    const int nsub = (int)s.p[OP_NUM_SUBSTEPS];
    const int iters = (int)s.p[OP_ITERATIONS];
    const int fsi = (int)s.p[OP_FRICTION_START_ITER];
    const bool accel = s.p[OP_USE_ACCELERATOR] != 0;
    const R rho = (R)s.p[OP_ACCELERATOR_RHO];
    for (int sub = 0; sub < nsub; ++sub) {
        s.clock += (float)s.dt;
        if (s.collide) detect_dcd(s);
        for (auto& m : s.meshes) apply_initial_step(s, m);
        // runStepGPU (VBDPhysics.cpp:839-862): deformers move their (fixed) vertices after the
        // initial step
        for (const auto& D : s.deformers)
            for (size_t k = 0; k < D.meshes.size(); ++k)
                oracle::deformer_apply<R>(D, (int)k, false, s.clock, (float)s.dt, s.meshes[D.meshes[k]].x.data());
        if (s.collide) detect_ccd(s);
        // Chebyshev acceleration (runStepGPU, VBDPhysics.cpp:872-906): both histories start at
        // the initial guess
        R omega = R(1);
        if (accel)
            for (auto& m : s.meshes) m.accPrevPrev = m.accPrevIter = m.x;
        for (int it = 0; it < iters; ++it) {
            if (accel) omega = accelerator_omega<R>(it + 1, rho, omega);
            const bool fr = it >= fsi;
            for (auto& grp : s.groups) {
                const int n = (int)grp.size() / 2;
                if (s.collide)
                    for (size_t m = 0; m < s.meshes.size(); ++m) s.snap[m] = s.meshes[m].x;
                auto body = [&](int i) {
                    Mesh<R>& m = s.meshes[grp[2 * i]];
                    int vid = grp[2 * i + 1];
                    if (!m.fixed[vid]) vbd_vertex(s, m, vid, fr, grp[2 * i]);
                };
                if (s.pool && nthreads > 1)
                    s.pool->parallel_for(0, n, body);
                else
                    for (int i = 0; i < n; ++i) body(i);
            }
            // VBDSolveParallelGroup_applyAcceleration_kernel (VBDPhysicsCompute.cu:910-944):
            // x <- y_{k-1} + omega (x - y_{k-1}), y_{k-1} <- y_k; the next sweep's y_k is the
            // accelerated x (recordPrevIterPositionsAccelerator, VBDPhysics.cpp:3594-3606)
            if (accel)
                for (auto& m : s.meshes)
                    for (int v = 0; v < m.nv; ++v) {
                        if (m.fixed[v]) continue;
                        for (int d = 0; d < 3; ++d) {
                            const size_t k = 3 * (size_t)v + d;
                            const R pp = m.accPrevPrev[k];
                            m.x[k] = pp + omega * (m.x[k] - pp);
                            m.accPrevPrev[k] = m.accPrevIter[k];
                            m.accPrevIter[k] = m.x[k];
                        }
                    }
        }
        for (auto& m : s.meshes) update_velocities(s, m);
    }
}

// VBDPhysics::evaluateConvergence (VBDPhysics.cpp:3459-3508) with evaluateMeritEnergy
// (:3784-3798), computeInertiaEnergy (VBD_BaseMaterial.h:223-240) and computeElasticEnergy
// (VBD_NeoHookean.h:64-84): out = {energy, inertia energy, elastic energy, average force norm}.
template <class R> void evaluate_convergence(Scene<R>& s, double* out) {
    double eIn = 0.0, eEl = 0.0, fsum = 0.0;
    int64_t nAll = 0;
    for (auto& m : s.meshes) {
        nAll += m.nv;
        const bool started = m.y.size() == m.x.size();
        for (int t = 0; t < m.nt; ++t) {
            R x4[12];
            for (int j = 0; j < 4; ++j)
                for (int d = 0; d < 3; ++d) x4[3 * j + d] = m.x[3 * m.tets[4 * t + j] + d];
            const R* DmInv = &m.dminv[9 * t];
            R F[9];
            for (int c = 0; c < 3; ++c)
                for (int r = 0; r < 3; ++r) {
                    R acc = 0;
                    for (int k = 0; k < 3; ++k) acc += (x4[3 * (k + 1) + r] - x4[r]) * DmInv[k + 3 * c];
                    F[r + 3 * c] = acc;
                }
            const R a = R(1) + m.mu / m.lmbd;
            const R A = m.vol[t];
            R fn = 0;
            for (int i = 0; i < 9; ++i) fn += F[i] * F[i];
            const R k = det3(F) - a;
            R energy = R(0.5) * A * (m.lmbd * k * k + m.mu * (fn - R(3)));
            energy -= R(0.5) * A * m.lmbd * (R(1) - a) * (R(1) - a);
            eEl += (double)energy;
        }
        for (int v = 0; v < m.nv; ++v) {
            const R* xv = &m.x[3 * v];
            const R* yv = started ? &m.y[3 * v] : xv;
            R dy[3] = {yv[0] - xv[0], yv[1] - xv[1], yv[2] - xv[2]};
            eIn += (double)(m.mass[v] * (dy[0] * dy[0] + dy[1] * dy[1] + dy[2] * dy[2]));
            if (m.fixed[v]) continue;
            R f[3] = {0, 0, 0}, h[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            R disp[3] = {0, 0, 0};
            if (started)
                for (int d = 0; d < 3; ++d) disp[d] = xv[d] - m.xprev[3 * v + d];
            for (int k = m.incBegin[v]; k < m.incBegin[v + 1]; ++k) {
                const int t = m.incTet[k];
                R x4[12];
                for (int j = 0; j < 4; ++j)
                    for (int d = 0; d < 3; ++d) x4[3 * j + d] = m.x[3 * m.tets[4 * t + j] + d];
                neohookean_vertex<R>(x4, &m.dminv[9 * t], m.vol[t], m.mu, m.lmbd, m.incOrder[k], m.dampH, m.dampD,
                                     s.dt, disp, f, h);
            }
            for (int d = 0; d < 3; ++d) f[d] += m.mass[v] * dy[d] * s.dtSqrRecip;
            fsum += std::sqrt((double)(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]));
        }
    }
    eIn *= 0.5 * (double)s.dtSqrRecip;
    out[0] = eIn + eEl;
    out[1] = eIn;
    out[2] = eEl;
    out[3] = nAll ? fsum / (double)nAll : 0.0;
}

}  // namespace

// ---------------------------------------------------------------------------
// Scene wrapper: one struct holds either precision.
// ---------------------------------------------------------------------------
struct OracleTetScene {
    int precision;
    Scene<float> s32;
    Scene<double> s64;
    std::unique_ptr<oracle::ThreadPool> pool;
};

template <class R> static int add_mesh(Scene<R>& s, int nv, const double* pos, const double* vel, int nt,
                                       const int32_t* tets, const double* mat, int nfixed, const int32_t* fixed,
                                       const int32_t* colors) {
    Mesh<R> m;
    m.nv = nv;
    m.nt = nt;
    m.x.resize(3 * nv);
    m.v.assign(3 * nv, R(0));
    for (int i = 0; i < 3 * nv; ++i) m.x[i] = (R)pos[i];
    if (vel)
        for (int i = 0; i < 3 * nv; ++i) m.v[i] = (R)vel[i];
    m.xprev = m.x;
    m.vprev = m.v;
    m.y = m.x;
    m.tets.assign(tets, tets + 4 * nt);
    m.density = (R)mat[OM_DENSITY];
    m.mu = (R)mat[OM_MU];
    m.lmbd = (R)mat[OM_LAMBDA];
    m.dampH = (R)mat[OM_DAMP_HYDRO];
    m.dampD = (R)mat[OM_DAMP_DEV];
    m.initType = (int)mat[OM_INIT_TYPE];
    // VBDPhysics.cpp:794-807 timeStepEqualize
    m.expVelDamp = (R)std::pow((R)mat[OM_EXP_VEL_DAMPING], s.dt);
    m.constVelDamp = (R)mat[OM_CONST_VEL_DAMPING] * s.dt;
    m.noGravZone = mat[OM_NO_GRAV_ZONE] != 0;
    m.noGravThreshold = (R)mat[OM_NO_GRAV_THRESHOLD];
    m.maxVel = (R)mat[OM_MAX_VELOCITY];
    m.frictionMu = (R)mat[OM_FRICTION_MU];
    m.frictionEpsV = (R)mat[OM_FRICTION_EPS_V];
    m.dminv.resize(9 * nt);
    m.vol.resize(nt);
    m.mass.resize(nv);
    if (tet_prepare<R>(nv, m.x.data(), nt, tets, m.density, m.dminv.data(), m.vol.data(), m.mass.data())) return -1;
    m.fixed.assign(nv, 0);
    for (int i = 0; i < nfixed; ++i) {
        m.fixed[fixed[i]] = 1;
        m.fixedList.push_back(fixed[i]);
    }
    build_incidence(m);
    // categories: vertices per colour, ascending id (ColoringAlgorithms.cpp:69-80), then
    // sorted by size descending (TetMeshFEM.cpp:122-133; stable here, std::sort there).
    int nc = 0;
    for (int i = 0; i < nv; ++i) nc = std::max(nc, colors[i] + 1);
    m.categories.assign(nc, {});
    for (int i = 0; i < nv; ++i) m.categories[colors[i]].push_back(i);
    std::stable_sort(m.categories.begin(), m.categories.end(),
                     [](const std::vector<int32_t>& a, const std::vector<int32_t>& b) { return a.size() > b.size(); });
    s.meshes.push_back(std::move(m));
    return (int)s.meshes.size() - 1;
}

// VBDPhysics.cpp:99-122 findSmallestParallelGroup and :235-274 group merge.
template <class R> static void build_groups(Scene<R>& s) {
    size_t ng = 0;
    for (auto& m : s.meshes) ng = std::max(ng, m.categories.size());
    s.groups.assign(ng, {});
    for (size_t im = 0; im < s.meshes.size(); ++im) {
        auto& m = s.meshes[im];
        std::vector<int> avail(m.categories.size());
        for (size_t c = 0; c < avail.size(); ++c) avail[c] = (int)c;
        for (size_t c = 0; c < m.categories.size(); ++c) {
            size_t best = s.groups[avail[0]].size();
            int bestId = avail[0];
            size_t bestI = 0;
            for (size_t i = 1; i < avail.size(); ++i)
                if (s.groups[avail[i]].size() < best) {
                    best = s.groups[avail[i]].size();
                    bestId = avail[i];
                    bestI = i;
                }
            avail.erase(avail.begin() + bestI);
            for (int v : m.categories[c]) {
                s.groups[bestId].push_back((int32_t)im);
                s.groups[bestId].push_back(v);
            }
        }
    }
}

template <class R> static int get_state(Scene<R>& s, int mesh, double* pos, double* vel) {
    if (mesh < 0 || mesh >= (int)s.meshes.size()) return -1;
    auto& m = s.meshes[mesh];
    for (int i = 0; i < 3 * m.nv; ++i) {
        if (pos) pos[i] = (double)m.x[i];
        if (vel) vel[i] = (double)m.v[i];
    }
    return 0;
}
template <class R> static int set_state(Scene<R>& s, int mesh, const double* pos, const double* vel) {
    if (mesh < 0 || mesh >= (int)s.meshes.size()) return -1;
    auto& m = s.meshes[mesh];
    for (int i = 0; i < 3 * m.nv; ++i) {
        if (pos) m.x[i] = (R)pos[i];
        if (vel) m.v[i] = (R)vel[i];
    }
    return 0;
}

template <class R> static void finalize_scene(Scene<R>& s) {
    build_groups(s);
    if (!s.collide) return;
    const size_t nm = s.meshes.size();
    s.topo.assign(nm, {});
    s.hits.assign(nm, {});
    s.penetrated.assign(nm, {});
    s.fromCCD.assign(nm, {});
    s.relations.assign(nm, {});
    s.snap.assign(nm, {});
    for (size_t m = 0; m < nm; ++m) {
        Mesh<R>& M = s.meshes[m];
        tetcol::build_topo<R>(s.topo[m], M.nv, M.nt, M.tets.data(), M.x.data());
        s.hits[m].assign(M.nv, tetcol::Hit{});
        s.penetrated[m].assign(M.nv, 0);
        s.fromCCD[m].assign(M.nv, 0);
        s.relations[m].assign(M.nv, {});
    }
}

extern "C" {

int oracle_tet_prepare_f64(int nv, const double* pos, int nt, const int32_t* tets, double density, double* dminv,
                           double* vol, double* mass) {
    return tet_prepare<double>(nv, pos, nt, tets, density, dminv, vol, mass);
}
int oracle_tet_prepare_f32(int nv, const float* pos, int nt, const int32_t* tets, float density, float* dminv,
                           float* vol, float* mass) {
    return tet_prepare<float>(nv, pos, nt, tets, density, dminv, vol, mass);
}

void oracle_neohookean_vertex_f64(const double* x4, const double* dminv, double vol, double mu, double lmbd,
                                  int order, double* force3, double* hess9) {
    const double disp[3] = {0, 0, 0};
    for (int i = 0; i < 3; ++i) force3[i] = 0;
    for (int i = 0; i < 9; ++i) hess9[i] = 0;
    neohookean_vertex<double>(x4, dminv, vol, mu, lmbd, order, 0.0, 0.0, 1.0, disp, force3, hess9);
}

int oracle_solve3x3_f64(const double* h9, const double* b3, double* x3) { return solve3<double>(h9, b3, x3) ? 1 : 0; }
int oracle_solve3x3_f32(const float* h9, const float* b3, float* x3) { return solve3<float>(h9, b3, x3) ? 1 : 0; }

OracleTetScene* oracle_tet_scene_create(int precision, const double* physics) {
    auto* sc = new OracleTetScene();
    sc->precision = precision == 64 ? 64 : 32;
    std::memcpy(sc->s32.p, physics, sizeof(double) * OP_COUNT);
    std::memcpy(sc->s64.p, physics, sizeof(double) * OP_COUNT);
    // BasePhysicsFramework.cpp:245-249: dt in FloatingType, dtSqrReciprocal = 1/(dt*dt)
    sc->s32.dt = (float)physics[OP_DT];
    sc->s32.dtSqrRecip = 1.f / (sc->s32.dt * sc->s32.dt);
    sc->s64.dt = physics[OP_DT];
    sc->s64.dtSqrRecip = 1.0 / (sc->s64.dt * sc->s64.dt);
    return sc;
}

void oracle_tet_scene_destroy(OracleTetScene* sc) { delete sc; }

int oracle_tet_scene_add_mesh(OracleTetScene* sc, int nv, const double* pos, const double* vel, int nt,
                              const int32_t* tets, const double* material, int nfixed, const int32_t* fixed,
                              const int32_t* colors) {
    if (sc->precision == 64) return add_mesh<double>(sc->s64, nv, pos, vel, nt, tets, material, nfixed, fixed, colors);
    return add_mesh<float>(sc->s32, nv, pos, vel, nt, tets, material, nfixed, fixed, colors);
}

int oracle_tet_scene_finalize(OracleTetScene* sc) {
    if (sc->precision == 64) finalize_scene(sc->s64);
    else finalize_scene(sc->s32);
    return 0;
}

// Box prefilter of the brute-force DCD / CCD on (default) or off; a speed-up only, the results are
// the same either way
void oracle_tet_scene_set_box_prefilter(OracleTetScene* sc, int on) {
    sc->s32.boxPrefilter = on != 0;
    sc->s64.boxPrefilter = on != 0;
}

int oracle_tet_scene_set_collision(OracleTetScene* sc, const double* c) {
    auto set = [&](auto& s) {
        s.collide = true;
        tetcol::Params& P = s.col;
        P.dcd = c[OC_ALLOW_DCD] != 0;
        P.ccd = c[OC_ALLOW_CCD] != 0;
        P.selfCollision = c[OC_SELF_COLLISION] != 0;
        P.feasibleRegion = c[OC_FEASIBLE_REGION] != 0;
        P.tetTraverse = c[OC_TET_TRAVERSE] != 0;
        P.traverseNonSelf = c[OC_TRAVERSE_NON_SELF] != 0;
        P.shiftToCenter = c[OC_SHIFT_TO_CENTER] != 0;
        P.stopAfterPassing = c[OC_STOP_AFTER_PASSING] != 0;
        P.stiffness = c[OC_STIFFNESS];
        P.offHeight = c[OC_OFF_HEIGHT];
        P.feasibleEps = c[OC_FEASIBLE_EPS];
        P.rayEps = c[OC_RAY_EPS];
        P.maxSearchMult = c[OC_MAX_SEARCH_MULT];
        P.centerShift = c[OC_CENTER_SHIFT];
    };
    set(sc->s32);
    set(sc->s64);
    return 0;
}

// the last detection: per surface vertex of every mesh (mesh, vertex, hit mesh, hit face,
// penetrated, from CCD); returns the number of rows (out may be NULL to count)
int oracle_tet_scene_collisions(OracleTetScene* sc, int32_t* out) {
    auto get = [&](auto& s) {
        int n = 0;
        for (size_t m = 0; m < s.topo.size(); ++m)
            for (int v = 0; v < s.topo[m].nv; ++v) {
                if (!s.topo[m].surf[v]) continue;
                if (out) {
                    int32_t* r = out + 6 * n;
                    r[0] = (int32_t)m;
                    r[1] = v;
                    r[2] = s.hits[m][v].mesh;
                    r[3] = s.hits[m][v].face;
                    r[4] = s.penetrated[m][v];
                    r[5] = s.fromCCD[m][v];
                }
                ++n;
            }
        return n;
    };
    return sc->precision == 64 ? get(sc->s64) : get(sc->s32);
}

// the boundary surface of mesh `mesh` as the collision model sees it: 3 ids per face (out may be
// NULL), followed in owner (if not NULL) by the owning tet and local face of each face
int oracle_tet_scene_surface(OracleTetScene* sc, int mesh, int32_t* faces, int32_t* owner) {
    auto get = [&](auto& s) {
        if (mesh < 0 || mesh >= (int)s.topo.size()) return -1;
        const tetcol::Topo& T = s.topo[mesh];
        const int nf = (int)T.faceTet.size();
        if (faces) std::memcpy(faces, T.faces.data(), sizeof(int32_t) * 3 * nf);
        if (owner)
            for (int f = 0; f < nf; ++f) {
                owner[2 * f] = T.faceTet[f];
                owner[2 * f + 1] = T.faceLocal[f];
            }
        return nf;
    };
    return sc->precision == 64 ? get(sc->s64) : get(sc->s32);
}

int oracle_tet_scene_num_groups(OracleTetScene* sc) {
    return sc->precision == 64 ? (int)sc->s64.groups.size() : (int)sc->s32.groups.size();
}

int oracle_tet_scene_group(OracleTetScene* sc, int g, int32_t* out) {
    auto& grp = sc->precision == 64 ? sc->s64.groups : sc->s32.groups;
    if (g < 0 || g >= (int)grp.size()) return -1;
    if (out) std::memcpy(out, grp[g].data(), grp[g].size() * sizeof(int32_t));
    return (int)grp[g].size() / 2;
}

int oracle_tet_scene_step(OracleTetScene* sc, int nsteps, int nthreads) {
    if (nthreads > 1 && (!sc->pool || sc->pool->size() != nthreads)) sc->pool.reset(new oracle::ThreadPool(nthreads));
    sc->s32.pool = sc->s64.pool = nthreads > 1 ? sc->pool.get() : nullptr;
    for (int i = 0; i < nsteps; ++i) {
        if (sc->precision == 64) scene_step(sc->s64, nthreads);
        else scene_step(sc->s32, nthreads);
    }
    return 0;
}

int oracle_tet_scene_add_deformer(OracleTetScene* sc, int kind, const double* params, int nmeshes,
                                  const int32_t* meshes, const int32_t* counts, const int32_t* verts) {
    if (sc->precision == 64)
        return oracle::deformer_add(sc->s64.deformers, sc->s64.meshes, kind, params, nmeshes, meshes, counts, verts);
    return oracle::deformer_add(sc->s32.deformers, sc->s32.meshes, kind, params, nmeshes, meshes, counts, verts);
}
long long oracle_tet_scene_solve_failures(OracleTetScene* sc) {
    return sc->precision == 64 ? sc->s64.solveFailures.load() : sc->s32.solveFailures.load();
}
double oracle_tet_scene_time(OracleTetScene* sc) { return sc->precision == 64 ? sc->s64.clock : sc->s32.clock; }
void oracle_tet_scene_set_time(OracleTetScene* sc, double t) { sc->s64.clock = sc->s32.clock = (float)t; }

int oracle_tet_scene_get_state(OracleTetScene* sc, int mesh, double* pos, double* vel) {
    return sc->precision == 64 ? get_state(sc->s64, mesh, pos, vel) : get_state(sc->s32, mesh, pos, vel);
}
int oracle_tet_scene_evaluate_convergence(OracleTetScene* sc, double* out) {
    if (sc->precision == 64)
        evaluate_convergence(sc->s64, out);
    else
        evaluate_convergence(sc->s32, out);
    return 0;
}
int oracle_point_in_tet_f32(const float* p, const float* a, const float* b, const float* c, const float* d) {
    using tetcol::ld;
    return tetcol::point_in_tet<float>(ld(p, 0), ld(a, 0), ld(b, 0), ld(c, 0), ld(d, 0)) ? 1 : 0;
}
int oracle_ccd_point_triangle_f64(const double* f0, const double* f1, const double* p0, const double* p1, double* t,
                                  double* bary) {
    using tetcol::ld;
    const tetcol::P3<double> a0[3] = {ld(f0, 0), ld(f0, 1), ld(f0, 2)};
    const tetcol::P3<double> a1[3] = {ld(f1, 0), ld(f1, 1), ld(f1, 2)};
    return tetcol::ccd_point_triangle<double>(a0, a1, ld(p0, 0), ld(p1, 0), *t, bary) ? 1 : 0;
}
int oracle_tet_walk_f64(int nv, const double* x, int nt, const int32_t* tets, const double* origin, const double* dir,
                        double maxDis, int startTet, int startLocal, int target) {
    tetcol::Topo T;
    tetcol::build_topo<double>(T, nv, nt, tets, x);
    using tetcol::ld;
    return tetcol::walk<double>(T, x, ld(origin, 0), ld(dir, 0), maxDis, startTet, startLocal, target, 1e-10) ? 1 : 0;
}
void oracle_penetration_force_f64(const double* x4, const double* xp4, int order, double k, double mu, double epsV,
                                  double dt, double* f, double* H) {
    using tetcol::ld;
    for (int i = 0; i < 3; ++i) f[i] = 0;
    for (int i = 0; i < 9; ++i) H[i] = 0;
    tetcol::penetration_force<double>(ld(x4, 0), ld(x4, 1), ld(x4, 2), ld(x4, 3), ld(xp4, 0), ld(xp4, 1), ld(xp4, 2),
                                      ld(xp4, 3), order, k, mu, epsV, dt, f, H);
}
int oracle_cubic_roots01_f64(const double* coef, double* roots) {
    return tetcol::cubic_roots01<double>(coef, roots, 6e-7);
}
int oracle_tet_scene_set_state(OracleTetScene* sc, int mesh, const double* pos, const double* vel) {
    return sc->precision == 64 ? set_state(sc->s64, mesh, pos, vel) : set_state(sc->s32, mesh, pos, vel);
}

}  // extern "C"
