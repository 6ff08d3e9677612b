// CPU oracle: deformers and vertex attachments. TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// Restated from (paths relative to Simulator/Modules/):
//   VBD/VBDDeformer.h:20-43               setFixed: selected vertices join the mesh's fixed points
//   VBD/VBDDeformer.h:45-131              tet Rotator: root = ((P - o) . axis) axis, RootP = P - root
//                                         at construction; p = root + R(theta) RootP with
//                                         theta = min(t, end) * angularVelocity (end > 0), R the
//                                         axis-angle matrix of the (unnormalised) axis, built in
//                                         double and rounded to float
//   VBD/VBDDeformer.h:134-178             tet Translator: p = speed * min(t, end) + P0 (end > 0)
//   VBDCloth/VBDClothDeformer.h:44-128    cloth Rotator (curTime is float: the clamped end time is
//                                         rounded to float)
//   VBDCloth/VBDClothDeformer.h:130-196   cloth Translator: p += speed * dt while
//                                         start <= t and (t < end or end < 0)
//   VBDCloth/VBDClothPhysics.cpp:148-224  loadConstraints: vertex-face attachments
//   VBDCloth/VBDClothPhysics.cpp:1465-1498 accumulateConstraintForceAndHessian
// The scene time is float (BasePhysicsFramework.h:100, FloatingType curTime), advanced by dt at
// the start of every substep.
#pragma once
#include <cmath>
#include <cstdint>
#include <vector>

namespace oracle {

struct Deformer {
    int kind = 0;  // 0 rotator, 1 translator
    float o[3] = {0, 0, 0}, axis[3] = {0, 0, 0}, speed[3] = {0, 0, 0};
    double omega = 0, endTime = -1;
    float startTime = 0;
    std::vector<int> meshes;
    std::vector<std::vector<int>> verts;
    std::vector<std::vector<float>> a, b;  // per mesh, 3 per vertex: root / P0, and RootP
};

// The per-vertex constants taken at construction from the current (float) positions.
inline void deformer_capture(Deformer& D, int k, const float* P) {
    float OP[3], root[3], rootP[3];
    for (int d = 0; d < 3; ++d) OP[d] = P[d] - D.o[d];
    const float dot = OP[0] * D.axis[0] + OP[1] * D.axis[1] + OP[2] * D.axis[2];
    for (int d = 0; d < 3; ++d) root[d] = dot * D.axis[d];
    for (int d = 0; d < 3; ++d) rootP[d] = P[d] - root[d];
    for (int d = 0; d < 3; ++d) {
        D.a[k].push_back(D.kind == 0 ? root[d] : P[d]);
        D.b[k].push_back(D.kind == 0 ? rootP[d] : 0.f);
    }
}

// Rotation matrix of angle theta about `axis` (row-major), each entry evaluated in double.
inline void rotation(const float* axis, double theta, double* Rm) {
    const double ux = axis[0], uy = axis[1], uz = axis[2];
    const double c = std::cos(theta), s = std::sin(theta);
    Rm[0] = c + ux * ux * (1 - c);
    Rm[1] = ux * uy * (1 - c) - uz * s;
    Rm[2] = ux * uz * (1 - c) + uy * s;
    Rm[3] = uy * ux * (1 - c) + uz * s;
    Rm[4] = c + uy * uy * (1 - c);
    Rm[5] = uy * uz * (1 - c) - ux * s;
    Rm[6] = uz * ux * (1 - c) - uy * s;
    Rm[7] = uz * uy * (1 - c) + ux * s;
    Rm[8] = c + uz * uz * (1 - c);
}

// Moves the deformer's vertices of mesh slot k at scene time t (float). `cloth` selects the cloth
// framework's semantics. Positions x (3 per vertex, precision R); R = float follows the
// reference's float arithmetic, R = double evaluates the same formulas in double.
template <class R> void deformer_apply(const Deformer& D, int k, bool cloth, float t, float dt, R* x) {
    const std::vector<int>& vs = D.verts[k];
    if (D.kind == 0) {
        double tc = t;
        if (tc >= D.endTime && D.endTime > 0) tc = cloth ? (double)(float)D.endTime : D.endTime;
        double Rd[9];
        rotation(D.axis, tc * D.omega, Rd);
        R Rm[9];
        for (int i = 0; i < 9; ++i) Rm[i] = (R)(float)Rd[i];
        for (size_t i = 0; i < vs.size(); ++i) {
            const float* a = &D.a[k][3 * i];
            const float* b = &D.b[k][3 * i];
            R* p = &x[3 * (size_t)vs[i]];
            for (int r = 0; r < 3; ++r) p[r] = (R)a[r] + (Rm[3 * r] * b[0] + Rm[3 * r + 1] * b[1] + Rm[3 * r + 2] * b[2]);
        }
    } else if (!cloth) {
        float tc = t;
        if (tc >= (float)D.endTime && (float)D.endTime > 0) tc = (float)D.endTime;
        for (size_t i = 0; i < vs.size(); ++i) {
            R* p = &x[3 * (size_t)vs[i]];
            for (int d = 0; d < 3; ++d) p[d] = (R)D.speed[d] * (R)tc + (R)D.a[k][3 * i + d];
        }
    } else {
        const float end = (float)D.endTime;
        if (!(t >= D.startTime && (t < end || end < 0))) return;
        for (size_t i = 0; i < vs.size(); ++i) {
            R* p = &x[3 * (size_t)vs[i]];
            for (int d = 0; d < 3; ++d) p[d] = p[d] + (R)D.speed[d] * (R)dt;
        }
    }
}

// One attachment record of a vertex: force -= k b (x_read - target), H += k b^2 I, where target
// is the barycentric point of face f on mesh `tgtMesh` or, with tgtMesh < 0, the point p.
// `readMesh`/`readVert` is the vertex-side mesh and the solved vertex's own index in its mesh
// (VBDClothPhysics.cpp:1471 reads the vertex-side mesh at the solved vertex's index, also for
// the face-side copies).
struct Attachment {
    int readMesh, readVert, tgtMesh;
    int f[3];
    double p[3];
    float k, b;
};

}  // namespace oracle

namespace oracle {

// setFixed + construction of a deformer over meshes of a scene (any mesh type with nv, x, fixed,
// fixedList). params: o[3], axis[3], angularVelocity, endTime, speed[3], startTime.
template <class MeshVec>
int deformer_add(std::vector<Deformer>& list, MeshVec& meshes, int kind, const double* prm, int nm,
                 const int32_t* meshIds, const int32_t* counts, const int32_t* verts) {
    if (kind != 0 && kind != 1) return -1;
    Deformer D;
    D.kind = kind;
    for (int d = 0; d < 3; ++d) {
        D.o[d] = (float)prm[d];
        D.axis[d] = (float)prm[3 + d];
        D.speed[d] = (float)prm[8 + d];
    }
    D.omega = prm[6];
    D.endTime = prm[7];
    D.startTime = (float)prm[11];
    size_t at = 0;
    for (int k = 0; k < nm; ++k) {
        const int mi = meshIds[k];
        if (mi < 0 || mi >= (int)meshes.size()) return -1;
        auto& m = meshes[mi];
        D.meshes.push_back(mi);
        D.verts.emplace_back();
        D.a.emplace_back();
        D.b.emplace_back();
        for (int i = 0; i < counts[k]; ++i) {
            const int v = verts[at++];
            if (v < 0 || v >= m.nv) return -1;
            D.verts.back().push_back(v);
            const float P[3] = {(float)m.x[3 * (size_t)v], (float)m.x[3 * (size_t)v + 1], (float)m.x[3 * (size_t)v + 2]};
            deformer_capture(D, k, P);
            if (!m.fixed[v]) {
                m.fixed[v] = 1;
                m.fixedList.push_back(v);
            }
        }
    }
    list.push_back(std::move(D));
    return (int)list.size() - 1;
}

}  // namespace oracle
