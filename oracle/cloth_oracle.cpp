// CPU oracle: restatement of the reference's CPU cloth VBD path (no self-contact).
// TEST INFRASTRUCTURE ONLY (see oracle.h). Never linked by the product.
//
// Reference files restated here (paths relative to Simulator/Modules/):
//   TriMesh/TriMesh.cpp:213-244            rest pose from 3D: 2D frame, Dm, area = det(Dm)/2, mass += rho A / 3
//   TriMesh/TriMesh.cpp:481-753            edge infos, vertex->face, vertex->relevant-bending lists
//   VBDCloth/VBDTriMeshStVK.cpp:50-88      quadratic bending forms Q = 3 K0 K0^T / (A0 + A1)
//   VBDCloth/VBDClothPhysics.cpp:72-111    parallel groups (remove = false)
//   VBDCloth/VBDClothPhysics.cpp:389-518   runStep_CDEveryIter, collision handling off
//   VBDCloth/VBDBaseTriMesh.cpp:58-178     applyInitialStep / forwardStep_intertia (types 0, 2, 4)
//   VBDCloth/VBDClothPhysics.cpp:1026-1154 VBDStepWithExistingCollisions (no contacts)
//   VBDCloth/VBDTriMeshStVK.cpp:721-911    StVK force, closed-form 6x6 d2W/dF2, membrane damping
//   VBDCloth/VBDBaseTriMesh.h:149-160      assembleMembraneHessian
//   VBDCloth/VBDTriMeshStVK.cpp:913-982    bending force/Hessian with degenerate-triangle skip
//   VBDCloth/VBDClothPhysics.cpp:1401-1463,1500-1531 boundary penalty + friction (always on)
//   VBDCloth/VBDClothPhysics.cpp:2318-2360 updateVelocity (damping NOT made step-invariant)
//
// Orders MeshFrame decides are defined here (parity unpinned: summation order only):
// edges are numbered in order of first appearance scanning faces (v0v1, v1v2, v2v0); eV1->eV2
// follows the first face (fId1); vertex->faces ascending; relevant bendings = incident interior
// edges ascending, then the opposite interior edge of each incident face, faces ascending.
// A failed 3x3 solve leaves the vertex where it is (the reference copies a stale
// positionsNext, VBDClothPhysics.cpp:1141-1153 and :483-493).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <utility>
#include <vector>

#include "oracle.h"
#include "pool.h"

namespace {

template <class R> struct ClothMesh {
    int nv = 0, nf = 0, ne = 0;
    std::vector<R> x, xprev, v, y, xnext;
    std::vector<int32_t> faces;
    std::vector<R> dminv, area, mass;                 // 4 * nf (col-major 2x2), nf, nv
    std::vector<int32_t> edges;                       // 6 * ne: eV1 eV2 eV12Next eV21Next fId1 fId2
    std::vector<R> Q;                                 // 16 * ne (row-major 4x4); zero for boundary edges
    std::vector<int32_t> vfBegin, vfFace, vfOrder;    // vertex -> faces
    std::vector<int32_t> vbBegin, vbEdge, vbOrder;    // vertex -> relevant bendings
    std::vector<uint8_t> fixed;
    std::vector<int32_t> fixedList;
    std::vector<std::vector<int32_t>> categories;
    R mu, lmbd, kb, dampS, dampB, expVel, constVel, density;
    int initType = 2;
};

template <class R> struct ClothScene {
    double p[OP_COUNT];
    R dt, dtSqrRecip;
    std::vector<ClothMesh<R>> meshes;
    std::vector<std::vector<int32_t>> groups;
};

template <class R> inline void cross3(const R* a, const R* b, R* c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}
template <class R> inline R dot3(const R* a, const R* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// TriMesh.cpp:213-244 with a right-handed orthonormal frame (a1 along v0->v1, a2 = n x a1):
// the reference's CuMatrix::buildOrthonormalBasis is unavailable; the StVK energy does not
// depend on the in-plane rotation of the rest frame.
template <class R> int tri_prepare(ClothMesh<R>& m) {
    m.dminv.assign(4 * (size_t)m.nf, 0);
    m.area.assign(m.nf, 0);
    m.mass.assign(m.nv, 0);
    for (int f = 0; f < m.nf; ++f) {
        const int32_t* fv = &m.faces[3 * f];
        R e1[3], e2[3], n[3];
        for (int d = 0; d < 3; ++d) {
            e1[d] = m.x[3 * fv[1] + d] - m.x[3 * fv[0] + d];
            e2[d] = m.x[3 * fv[2] + d] - m.x[3 * fv[0] + d];
        }
        cross3(e1, e2, n);
        R nn = std::sqrt(dot3(n, n));
        R l1 = std::sqrt(dot3(e1, e1));
        if (!(nn > 0) || !(l1 > 0)) return -1;
        R a1[3] = {e1[0] / l1, e1[1] / l1, e1[2] / l1};
        R un[3] = {n[0] / nn, n[1] / nn, n[2] / nn};
        R a2[3];
        cross3(un, a1, a2);
        // Dm columns: 2D coordinates of v1 - v0 and v2 - v0
        R Dm[4] = {dot3(a1, e1), dot3(a2, e1), dot3(a1, e2), dot3(a2, e2)};
        R det = Dm[0] * Dm[3] - Dm[2] * Dm[1];
        m.area[f] = det * R(0.5);
        R inv = R(1) / det;
        R* Di = &m.dminv[4 * f];
        Di[0] = Dm[3] * inv;
        Di[1] = -Dm[1] * inv;
        Di[2] = -Dm[2] * inv;
        Di[3] = Dm[0] * inv;
        for (int k = 0; k < 3; ++k) m.mass[fv[k]] += m.density * m.area[f] / R(3);
    }
    return 0;
}

// TriMesh.cpp:481-753 (orders defined in the header comment).
template <class R> int tri_topology(ClothMesh<R>& m) {
    std::map<std::pair<int, int>, int> edgeOf;
    m.edges.clear();
    for (int f = 0; f < m.nf; ++f) {
        for (int k = 0; k < 3; ++k) {
            int a = m.faces[3 * f + k], b = m.faces[3 * f + (k + 1) % 3], c = m.faces[3 * f + (k + 2) % 3];
            auto key = std::make_pair(std::min(a, b), std::max(a, b));
            auto it = edgeOf.find(key);
            if (it == edgeOf.end()) {
                edgeOf[key] = (int)(m.edges.size() / 6);
                int e[6] = {a, b, c, -1, f, -1};
                m.edges.insert(m.edges.end(), e, e + 6);
            } else {
                int32_t* e = &m.edges[6 * it->second];
                if (e[5] != -1 || e[0] != b || e[1] != a) return -1;  // non-manifold or inconsistent orientation
                e[3] = c;
                e[5] = f;
            }
        }
    }
    m.ne = (int)(m.edges.size() / 6);
    // vertex -> faces
    m.vfBegin.assign(m.nv + 1, 0);
    for (int f = 0; f < m.nf; ++f)
        for (int k = 0; k < 3; ++k) m.vfBegin[m.faces[3 * f + k] + 1]++;
    for (int i = 0; i < m.nv; ++i) m.vfBegin[i + 1] += m.vfBegin[i];
    m.vfFace.resize(m.vfBegin[m.nv]);
    m.vfOrder.resize(m.vfBegin[m.nv]);
    std::vector<int32_t> fill(m.vfBegin.begin(), m.vfBegin.end() - 1);
    for (int f = 0; f < m.nf; ++f)
        for (int k = 0; k < 3; ++k) {
            int v = m.faces[3 * f + k];
            m.vfFace[fill[v]] = f;
            m.vfOrder[fill[v]] = k;
            fill[v]++;
        }
    // vertex -> incident edges (ascending)
    std::vector<std::vector<int>> vEdges(m.nv);
    for (int e = 0; e < m.ne; ++e) {
        vEdges[m.edges[6 * e]].push_back(e);
        vEdges[m.edges[6 * e + 1]].push_back(e);
    }
    // face -> edge opposite each local vertex
    auto edge_id = [&](int a, int b) { return edgeOf[std::make_pair(std::min(a, b), std::max(a, b))]; };
    m.vbBegin.assign(m.nv + 1, 0);
    m.vbEdge.clear();
    m.vbOrder.clear();
    for (int v = 0; v < m.nv; ++v) {
        for (int e : vEdges[v]) {
            if (m.edges[6 * e + 5] == -1) continue;
            m.vbEdge.push_back(e);
            m.vbOrder.push_back(m.edges[6 * e] == v ? 0 : 1);
        }
        for (int k = m.vfBegin[v]; k < m.vfBegin[v + 1]; ++k) {
            int f = m.vfFace[k], o = m.vfOrder[k];
            int a = m.faces[3 * f + (o + 1) % 3], b = m.faces[3 * f + (o + 2) % 3];
            int e = edge_id(a, b);
            if (m.edges[6 * e + 5] == -1) continue;
            m.vbEdge.push_back(e);
            m.vbOrder.push_back(m.edges[6 * e + 4] == f ? 2 : 3);
        }
        m.vbBegin[v + 1] = (int)m.vbEdge.size();
    }
    return 0;
}

template <class R> inline R cotangent(const R* a, const R* b) {
    R c[3];
    cross3(a, b, c);
    return dot3(a, b) / std::sqrt(dot3(c, c));
}

// VBDTriMeshStVK.cpp:50-88
template <class R> void bending_forms(ClothMesh<R>& m) {
    m.Q.assign(16 * (size_t)m.ne, 0);
    for (int e = 0; e < m.ne; ++e) {
        const int32_t* E = &m.edges[6 * e];
        if (E[5] == -1) continue;
        const R* X0 = &m.x[3 * E[0]];
        const R* X1 = &m.x[3 * E[1]];
        const R* X2 = &m.x[3 * E[2]];
        const R* X3 = &m.x[3 * E[3]];
        R e01[3], e03[3], e02[3], e12[3], e13[3], me01[3];
        for (int d = 0; d < 3; ++d) {
            e01[d] = X1[d] - X0[d];
            e03[d] = X3[d] - X0[d];
            e02[d] = X2[d] - X0[d];
            e12[d] = X2[d] - X1[d];
            e13[d] = X3[d] - X1[d];
            me01[d] = -e01[d];
        }
        const R c01 = cotangent(e02, e01), c02 = cotangent(e01, e03);
        const R c03 = cotangent(me01, e12), c04 = cotangent(me01, e13);
        const R K0[4] = {c03 + c04, c01 + c02, -c01 - c03, -c02 - c04};
        auto normal_len = [&](int f) {
            const int32_t* fv = &m.faces[3 * f];
            R a[3], b[3], n[3];
            for (int d = 0; d < 3; ++d) {
                a[d] = m.x[3 * fv[1] + d] - m.x[3 * fv[0] + d];
                b[d] = m.x[3 * fv[2] + d] - m.x[3 * fv[0] + d];
            }
            cross3(a, b, n);
            return std::sqrt(dot3(n, n));
        };
        const R A0 = normal_len(E[4]), A1 = normal_len(E[5]);
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) m.Q[16 * e + 4 * i + j] = R(3) * K0[i] * K0[j] / (A0 + A1);
    }
}

template <class R> inline bool solve3(const R* h, const R* b, R* x) {
    const R a00 = h[0], a10 = h[1], a20 = h[2], a01 = h[3], a11 = h[4], a21 = h[5], a02 = h[6], a12 = h[7],
            a22 = h[8];
    const R C00 = a11 * a22 - a12 * a21, C01 = a12 * a20 - a10 * a22, C02 = a10 * a21 - a11 * a20;
    const R C10 = a02 * a21 - a01 * a22, C11 = a00 * a22 - a02 * a20, C12 = a01 * a20 - a00 * a21;
    const R C20 = a01 * a12 - a02 * a11, C21 = a02 * a10 - a00 * a12, C22 = a00 * a11 - a01 * a10;
    const R det = a00 * C00 + a01 * C01 + a02 * C02;
    const R scale = std::fabs(a00 * a11 * a22);
    if (!std::isfinite(det) || !(std::fabs(det) > R(1e-12) * scale)) return false;
    const R inv = R(1) / det;
    x[0] = (C00 * b[0] + C10 * b[1] + C20 * b[2]) * inv;
    x[1] = (C01 * b[0] + C11 * b[1] + C21 * b[2]) * inv;
    x[2] = (C02 * b[0] + C12 * b[1] + C22 * b[2]) * inv;
    return std::isfinite(x[0]) && std::isfinite(x[1]) && std::isfinite(x[2]);
}

// VBDTriMeshStVK.cpp:721-911: StVK membrane of every incident face, then damping.
template <class R> void stvk_vertex(const ClothMesh<R>& m, int vid, R dt, R* force, R* hess) {
    R Hmem[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const R lmbd = m.lmbd, miu = m.mu;
    for (int k = m.vfBegin[vid]; k < m.vfBegin[vid + 1]; ++k) {
        const int f = m.vfFace[k], order = m.vfOrder[k];
        const int32_t* fv = &m.faces[3 * f];
        const R* D = &m.dminv[4 * f];  // col-major: D[0]=(0,0) D[1]=(1,0) D[2]=(0,1) D[3]=(1,1)
        R Ds[6];
        for (int d = 0; d < 3; ++d) {
            Ds[d] = m.x[3 * fv[1] + d] - m.x[3 * fv[0] + d];
            Ds[3 + d] = m.x[3 * fv[2] + d] - m.x[3 * fv[0] + d];
        }
        R F[6];  // 3x2 col-major
        for (int c = 0; c < 2; ++c)
            for (int r = 0; r < 3; ++r) F[r + 3 * c] = Ds[r] * D[0 + 2 * c] + Ds[3 + r] * D[1 + 2 * c];
        R FtF[4];
        FtF[0] = F[0] * F[0] + F[1] * F[1] + F[2] * F[2];
        FtF[1] = F[3] * F[0] + F[4] * F[1] + F[5] * F[2];
        FtF[2] = FtF[1];
        FtF[3] = F[3] * F[3] + F[4] * F[4] + F[5] * F[5];
        const R G[4] = {R(0.5) * (FtF[0] - 1), R(0.5) * FtF[1], R(0.5) * FtF[2], R(0.5) * (FtF[3] - 1)};
        const R trG = G[0] + G[3];
        const R S[4] = {2 * miu * G[0] + lmbd * trG, 2 * miu * G[1], 2 * miu * G[2], 2 * miu * G[3] + lmbd * trG};
        const R A = m.area[f];
        // F12 = -A F S DmInv^T (3x2)
        R FS[6];
        for (int c = 0; c < 2; ++c)
            for (int r = 0; r < 3; ++r) FS[r + 3 * c] = F[r] * S[0 + 2 * c] + F[r + 3] * S[1 + 2 * c];
        R F12[6];
        for (int c = 0; c < 2; ++c)
            for (int r = 0; r < 3; ++r) F12[r + 3 * c] = -A * (FS[r] * D[c + 0] + FS[r + 3] * D[c + 2]);
        const R Dm_inv1_1 = D[0], Dm_inv2_1 = D[1], Dm_inv1_2 = D[2], Dm_inv2_2 = D[3];
        const R F1_1 = F[0], F2_1 = F[1], F3_1 = F[2], F1_2 = F[3], F2_2 = F[4], F3_2 = F[5];
        const R F1_1_sqr = F1_1 * F1_1, F2_1_sqr = F2_1 * F2_1, F3_1_sqr = F3_1 * F3_1;
        const R F1_2_sqr = F1_2 * F1_2, F2_2_sqr = F2_2 * F2_2, F3_2_sqr = F3_2 * F3_2;
        const R e_uu = G[0], e_vv = G[3], e_uv = G[2], e_uuvvSum = e_uu + e_vv;
        R W[36];
        auto w = [&](int i, int j, R val) {
            W[i + 6 * j] = val;
            W[j + 6 * i] = val;
        };
        w(0, 0, F1_1 * (F1_1 * lmbd + 2 * F1_1 * miu) + 2 * miu * e_uu + lmbd * (e_uuvvSum) + F1_2_sqr * miu);
        w(0, 1, F1_1 * (F2_1 * lmbd + 2 * F2_1 * miu) + F1_2 * F2_2 * miu);
        w(0, 2, F1_1 * (F3_1 * lmbd + 2 * F3_1 * miu) + F1_2 * F3_2 * miu);
        w(0, 3, 2 * miu * e_uv + F1_1 * F1_2 * lmbd + F1_1 * F1_2 * miu);
        w(0, 4, F1_1 * F2_2 * lmbd + F1_2 * F2_1 * miu);
        w(0, 5, F1_1 * F3_2 * lmbd + F1_2 * F3_1 * miu);
        w(1, 1, F2_1 * (F2_1 * lmbd + 2 * F2_1 * miu) + 2 * miu * e_uu + lmbd * (e_uuvvSum) + F2_2_sqr * miu);
        w(1, 2, F2_1 * (F3_1 * lmbd + 2 * F3_1 * miu) + F2_2 * F3_2 * miu);
        w(1, 3, F1_2 * F2_1 * lmbd + F1_1 * F2_2 * miu);
        w(1, 4, 2 * miu * e_uv + F2_1 * F2_2 * lmbd + F2_1 * F2_2 * miu);
        w(1, 5, F2_1 * F3_2 * lmbd + F2_2 * F3_1 * miu);
        w(2, 2, F3_1 * (F3_1 * lmbd + 2 * F3_1 * miu) + 2 * miu * e_uu + lmbd * (e_uuvvSum) + F3_2_sqr * miu);
        w(2, 3, F1_2 * F3_1 * lmbd + F1_1 * F3_2 * miu);
        w(2, 4, F2_2 * F3_1 * lmbd + F2_1 * F3_2 * miu);
        w(2, 5, 2 * miu * e_uv + F3_1 * F3_2 * lmbd + F3_1 * F3_2 * miu);
        w(3, 3, F1_2 * (F1_2 * lmbd + 2 * F1_2 * miu) + 2 * miu * e_vv + lmbd * (e_uuvvSum) + F1_1_sqr * miu);
        w(3, 4, F1_2 * (F2_2 * lmbd + 2 * F2_2 * miu) + F1_1 * F2_1 * miu);
        w(3, 5, F1_2 * (F3_2 * lmbd + 2 * F3_2 * miu) + F1_1 * F3_1 * miu);
        w(4, 4, F2_2 * (F2_2 * lmbd + 2 * F2_2 * miu) + 2 * miu * e_vv + lmbd * (e_uuvvSum) + F2_1_sqr * miu);
        w(4, 5, F2_2 * (F3_2 * lmbd + 2 * F3_2 * miu) + F2_1 * F3_1 * miu);
        w(5, 5, F3_2 * (F3_2 * lmbd + 2 * F3_2 * miu) + 2 * miu * e_vv + lmbd * (e_uuvvSum) + F3_1_sqr * miu);
        for (int i = 0; i < 36; ++i) W[i] *= A;
        R m1, m2;
        switch (order) {
        case 0:
            m1 = -Dm_inv1_1 - Dm_inv2_1;
            m2 = -Dm_inv1_2 - Dm_inv2_2;
            for (int d = 0; d < 3; ++d) force[d] += -F12[d] - F12[3 + d];
            break;
        case 1:
            m1 = Dm_inv1_1;
            m2 = Dm_inv1_2;
            for (int d = 0; d < 3; ++d) force[d] += F12[d];
            break;
        default:
            m1 = Dm_inv2_1;
            m2 = Dm_inv2_2;
            for (int d = 0; d < 3; ++d) force[d] += F12[3 + d];
            break;
        }
        // VBDBaseTriMesh.h:149-160 assembleMembraneHessian
        R HL[3][6];
        for (int c = 0; c < 6; ++c)
            for (int r = 0; r < 3; ++r) HL[r][c] = W[r + 6 * c] * m1 + W[(r + 3) + 6 * c] * m2;
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 3; ++r) Hmem[r + 3 * c] += HL[r][c] * m1 + HL[r][c + 3] * m2;
    }
    const R c = m.dampS / dt;
    R disp[3];
    for (int d = 0; d < 3; ++d) disp[d] = m.x[3 * vid + d] - m.xprev[3 * vid + d];
    for (int r = 0; r < 3; ++r) {
        R df = 0;
        for (int cc = 0; cc < 3; ++cc) df += Hmem[r + 3 * cc] * c * disp[cc];
        force[r] -= df;
    }
    for (int i = 0; i < 9; ++i) hess[i] += Hmem[i] + Hmem[i] * c;
}

// VBDTriMeshStVK.cpp:913-982
template <class R> void bending_vertex(const ClothMesh<R>& m, int vid, R dt, R degThres, R* force, R* hess) {
    const R kb = m.kb, gb = m.dampB;
    for (int k = m.vbBegin[vid]; k < m.vbBegin[vid + 1]; ++k) {
        const int e = m.vbEdge[k], o = m.vbOrder[k];
        const int32_t* E = &m.edges[6 * e];
        if (E[5] == -1) continue;
        const R* X[4] = {&m.x[3 * E[0]], &m.x[3 * E[1]], &m.x[3 * E[2]], &m.x[3 * E[3]]};
        R a[3], b[3], n1[3], n2[3];
        for (int d = 0; d < 3; ++d) {
            a[d] = X[1][d] - X[0][d];
            b[d] = X[2][d] - X[0][d];
        }
        cross3(a, b, n1);
        for (int d = 0; d < 3; ++d) {
            a[d] = X[2][d] - X[3][d];
            b[d] = X[1][d] - X[3][d];
        }
        cross3(a, b, n2);
        if (dot3(n1, n1) < degThres * degThres || dot3(n2, n2) < degThres * degThres) continue;
        const R* Q = &m.Q[16 * e];
        R dE[3] = {0, 0, 0};
        for (int j = 0; j < 4; ++j)
            for (int d = 0; d < 3; ++d) dE[d] += kb * Q[4 * o + j] * X[j][d];
        const R hq = kb * Q[4 * o + o];
        for (int d = 0; d < 3; ++d)
            force[d] -= dE[d] + hq * (m.x[3 * vid + d] - m.xprev[3 * vid + d]) * (gb / dt);
        const R hd = hq * (R(1) + gb / dt);
        hess[0] += hd;
        hess[4] += hd;
        hess[8] += hd;
    }
}

template <class R> void friction(R mu, R lambda, const R* T, const R* u, R epsU, R* f, R* H) {
    R uNorm = std::sqrt(u[0] * u[0] + u[1] * u[1]);
    if (uNorm > 0) {
        R f1 = uNorm > epsU ? R(1) / uNorm : (-uNorm / epsU + R(2)) / epsU;
        for (int r = 0; r < 3; ++r) f[r] = -mu * lambda * (T[r] * u[0] + T[r + 3] * u[1]) * f1;
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 3; ++r) H[r + 3 * c] = mu * lambda * f1 * (T[r] * T[c] + T[r + 3] * T[c + 3]);
    } else {
        for (int i = 0; i < 3; ++i) f[i] = 0;
        for (int i = 0; i < 9; ++i) H[i] = 0;
    }
}

template <class R> void cloth_vertex(ClothScene<R>& s, ClothMesh<R>& m, int vid) {
    R h[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    R f[3] = {0, 0, 0};
    const R* xv = &m.x[3 * vid];
    // inertia (VBDBaseTriMesh.cpp:311-318)
    for (int d = 0; d < 3; ++d) f[d] += m.mass[vid] * (m.y[3 * vid + d] - xv[d]) * s.dtSqrRecip;
    const R hi = m.mass[vid] * s.dtSqrRecip;
    h[0] += hi; h[4] += hi; h[8] += hi;
    stvk_vertex(m, vid, s.dt, f, h);
    bending_vertex(m, vid, s.dt, (R)s.p[OP_DEGENERATE_THRES], f, h);
    // boundary (VBDClothPhysics.cpp:1401-1463)
    if (s.p[OP_USE_PLANE_GROUND] != 0) {
        const R k = (R)s.p[OP_CLOTH_BOUNDARY_STIFFNESS];
        for (int dim = 0; dim < 3; ++dim) {
            const R lower = (R)s.p[OP_LOWER_X + dim], upper = (R)s.p[OP_UPPER_X + dim];
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
            R dx[3] = {xv[0] - m.xprev[3 * vid], xv[1] - m.xprev[3 * vid + 1], xv[2] - m.xprev[3 * vid + 2]};
            R T[6] = {0, 0, 0, 0, 0, 0};
            T[(dim + 1) % 3] = 1;
            T[3 + (dim + 2) % 3] = 1;
            R u[2] = {T[0] * dx[0] + T[1] * dx[1] + T[2] * dx[2], T[3] * dx[0] + T[4] * dx[1] + T[5] * dx[2]};
            R ff[3], fh[9];
            friction<R>((R)s.p[OP_BOUNDARY_FRICTION_MU], pen * k, T, u,
                        (R)s.p[OP_CLOTH_BOUNDARY_FRICTION_EPS] * s.dt, ff, fh);
            for (int d = 0; d < 3; ++d) f[d] += ff[d];
            for (int i = 0; i < 9; ++i) h[i] += fh[i];
        }
    }
    R dx[3];
    R* xn = &m.xnext[3 * vid];
    if (solve3<R>(h, f, dx)) {
        const R step = (R)s.p[OP_STEP_SIZE];
        for (int d = 0; d < 3; ++d) xn[d] = xv[d] + step * dx[d];
    } else {
        for (int d = 0; d < 3; ++d) xn[d] = xv[d];
    }
}

template <class R> void cloth_initial_step(ClothScene<R>& s, ClothMesh<R>& m) {
    const R dt = s.dt;
    const R g[3] = {(R)s.p[OP_GRAVITY_X], (R)s.p[OP_GRAVITY_Y], (R)s.p[OP_GRAVITY_Z]};
    m.xprev = m.x;
    if (s.p[OP_ASSOCIATE_GRAVITY] != 0) {
        const R dg[3] = {dt * g[0], dt * g[1], dt * g[2]};
        for (int i = 0; i < m.nv; ++i)
            for (int d = 0; d < 3; ++d) m.v[3 * i + d] += dg[d];
    }
    for (int f : m.fixedList)
        for (int d = 0; d < 3; ++d) m.v[3 * f + d] = 0;
    for (int i = 0; i < 3 * m.nv; ++i) m.y[i] = m.x[i] + dt * m.v[i];
    if (m.initType == 0 || m.initType == 3) m.x = m.y;  // 3: inertia within (infinite) conservative bounds
    // 2: previous position (x unchanged), 4: nothing
}

template <class R> void cloth_velocity(ClothScene<R>& s, ClothMesh<R>& m) {
    for (int i = 0; i < m.nv; ++i) {
        R* v = &m.v[3 * i];
        for (int d = 0; d < 3; ++d) v[d] = (m.x[3 * i + d] - m.xprev[3 * i + d]) / s.dt;
        R vMag = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        if (vMag > R(1e-6)) {
            R vNew = vMag * m.expVel - m.constVel;
            vNew = vNew > R(1e-6) ? vNew : R(0);
            for (int d = 0; d < 3; ++d) v[d] *= vNew / vMag;
        } else {
            for (int d = 0; d < 3; ++d) v[d] *= R(0);
        }
    }
    for (int f : m.fixedList)
        for (int d = 0; d < 3; ++d) m.v[3 * f + d] = 0;
}

template <class R> void cloth_step(ClothScene<R>& s, oracle::ThreadPool* pool) {
    const int nsub = (int)s.p[OP_NUM_SUBSTEPS], iters = (int)s.p[OP_ITERATIONS];
    for (int sub = 0; sub < nsub; ++sub) {
        for (int it = 0; it < iters; ++it) {
            if (it == 0)
                for (auto& m : s.meshes) cloth_initial_step(s, m);
            for (auto& grp : s.groups) {
                const int n = (int)grp.size() / 2;
                auto solve = [&](int i) {
                    ClothMesh<R>& m = s.meshes[grp[2 * i]];
                    int vid = grp[2 * i + 1];
                    if (!m.fixed[vid]) cloth_vertex(s, m, vid);
                };
                auto copy = [&](int i) {
                    ClothMesh<R>& m = s.meshes[grp[2 * i]];
                    int vid = grp[2 * i + 1];
                    if (!m.fixed[vid])
                        for (int d = 0; d < 3; ++d) m.x[3 * vid + d] = m.xnext[3 * vid + d];
                };
                if (pool) {
                    pool->parallel_for(0, n, solve);
                    pool->parallel_for(0, n, copy);
                } else {
                    for (int i = 0; i < n; ++i) solve(i);
                    for (int i = 0; i < n; ++i) copy(i);
                }
            }
        }
        if (iters == 0)
            for (auto& m : s.meshes) cloth_initial_step(s, m);
        for (auto& m : s.meshes) cloth_velocity(s, m);
    }
}

}  // namespace

struct OracleClothScene {
    int precision;
    ClothScene<float> s32;
    ClothScene<double> s64;
    std::unique_ptr<oracle::ThreadPool> pool;
};

template <class R>
static int cloth_add(ClothScene<R>& s, int nv, const double* pos, const double* vel, int nf, const int32_t* faces,
                     const double* mat, int nfixed, const int32_t* fixed, const int32_t* colors) {
    ClothMesh<R> m;
    m.nv = nv;
    m.nf = nf;
    m.x.resize(3 * nv);
    m.v.assign(3 * nv, 0);
    for (int i = 0; i < 3 * nv; ++i) m.x[i] = (R)pos[i];
    if (vel)
        for (int i = 0; i < 3 * nv; ++i) m.v[i] = (R)vel[i];
    m.xprev = m.x;
    m.y = m.x;
    m.xnext = m.x;
    m.faces.assign(faces, faces + 3 * nf);
    m.density = (R)mat[OMC_DENSITY];
    m.mu = (R)mat[OMC_MU];
    m.lmbd = (R)mat[OMC_LAMBDA];
    m.kb = (R)mat[OMC_BENDING];
    m.dampS = (R)mat[OMC_DAMP_STVK];
    m.dampB = (R)mat[OMC_DAMP_BENDING];
    m.expVel = (R)mat[OMC_EXP_VEL_DAMPING];
    m.constVel = (R)mat[OMC_CONST_VEL_DAMPING];
    m.initType = (int)mat[OMC_INIT_TYPE];
    if (tri_prepare(m) || tri_topology(m)) return -1;
    bending_forms(m);
    m.fixed.assign(nv, 0);
    for (int i = 0; i < nfixed; ++i) {
        m.fixed[fixed[i]] = 1;
        m.fixedList.push_back(fixed[i]);
    }
    int nc = 0;
    for (int i = 0; i < nv; ++i) nc = std::max(nc, colors[i] + 1);
    m.categories.assign(nc, {});
    for (int i = 0; i < nv; ++i) m.categories[colors[i]].push_back(i);
    std::stable_sort(m.categories.begin(), m.categories.end(),
                     [](const std::vector<int32_t>& a, const std::vector<int32_t>& b) { return a.size() > b.size(); });
    s.meshes.push_back(std::move(m));
    return (int)s.meshes.size() - 1;
}

// VBDClothPhysics.cpp:72-111 (a mesh's colours may share a group: remove = false)
template <class R> static void cloth_groups(ClothScene<R>& s) {
    size_t ng = 0;
    for (auto& m : s.meshes) ng = std::max(ng, m.categories.size());
    s.groups.assign(ng, {});
    for (size_t im = 0; im < s.meshes.size(); ++im) {
        auto& m = s.meshes[im];
        const size_t nc = m.categories.size();
        for (size_t c = 0; c < nc; ++c) {
            size_t best = s.groups[0].size();
            int bestId = 0;
            for (size_t i = 1; i < nc; ++i)
                if (s.groups[i].size() < best) {
                    best = s.groups[i].size();
                    bestId = (int)i;
                }
            for (int v : m.categories[c]) {
                s.groups[bestId].push_back((int32_t)im);
                s.groups[bestId].push_back(v);
            }
        }
    }
}

template <class R> static int cloth_state(ClothScene<R>& s, int mesh, double* pos, double* vel, bool get) {
    if (mesh < 0 || mesh >= (int)s.meshes.size()) return -1;
    auto& m = s.meshes[mesh];
    for (int i = 0; i < 3 * m.nv; ++i) {
        if (get) {
            if (pos) pos[i] = (double)m.x[i];
            if (vel) vel[i] = (double)m.v[i];
        } else {
            if (pos) m.x[i] = (R)pos[i];
            if (vel) m.v[i] = (R)vel[i];
        }
    }
    return 0;
}

extern "C" {

int oracle_cloth_prepare(int precision, int nv, const double* pos, int nf, const int32_t* faces, double density,
                         double* dminv, double* area, double* mass, int32_t* num_edges, int32_t* edges, double* Q) {
    ClothMesh<double> md;
    ClothMesh<float> mf;
    auto run = [&](auto& m) -> int {
        using R = typename std::decay_t<decltype(m.x)>::value_type;
        m.nv = nv;
        m.nf = nf;
        m.x.resize(3 * nv);
        for (int i = 0; i < 3 * nv; ++i) m.x[i] = (R)pos[i];
        m.faces.assign(faces, faces + 3 * nf);
        m.density = (R)density;
        if (tri_prepare(m) || tri_topology(m)) return -1;
        bending_forms(m);
        for (int i = 0; i < 4 * nf; ++i) if (dminv) dminv[i] = m.dminv[i];
        for (int i = 0; i < nf; ++i) if (area) area[i] = m.area[i];
        for (int i = 0; i < nv; ++i) if (mass) mass[i] = m.mass[i];
        if (num_edges) *num_edges = m.ne;
        if (edges) std::memcpy(edges, m.edges.data(), m.edges.size() * sizeof(int32_t));
        if (Q) for (size_t i = 0; i < m.Q.size(); ++i) Q[i] = m.Q[i];
        return 0;
    };
    return precision == 64 ? run(md) : run(mf);
}

void oracle_stvk_vertex_f64(const double* x3, const double* dminv4, double area, double mu, double lmbd, int order,
                            double* force3, double* hess9) {
    ClothMesh<double> m;
    m.nv = 3;
    m.nf = 1;
    m.x.assign(x3, x3 + 9);
    m.xprev = m.x;
    m.faces = {0, 1, 2};
    m.dminv.assign(dminv4, dminv4 + 4);
    m.area = {area};
    m.mu = mu;
    m.lmbd = lmbd;
    m.dampS = 0;
    m.vfBegin = {0, 1, 2, 3};
    m.vfFace = {0, 0, 0};
    m.vfOrder = {0, 1, 2};
    for (int i = 0; i < 3; ++i) force3[i] = 0;
    for (int i = 0; i < 9; ++i) hess9[i] = 0;
    stvk_vertex(m, order, 1.0, force3, hess9);
}

OracleClothScene* oracle_cloth_scene_create(int precision, const double* physics) {
    auto* sc = new OracleClothScene();
    sc->precision = precision == 64 ? 64 : 32;
    std::memcpy(sc->s32.p, physics, sizeof(double) * OP_COUNT);
    std::memcpy(sc->s64.p, physics, sizeof(double) * OP_COUNT);
    sc->s32.dt = (float)physics[OP_DT];
    sc->s32.dtSqrRecip = 1.f / (sc->s32.dt * sc->s32.dt);
    sc->s64.dt = physics[OP_DT];
    sc->s64.dtSqrRecip = 1.0 / (sc->s64.dt * sc->s64.dt);
    return sc;
}

void oracle_cloth_scene_destroy(OracleClothScene* sc) { delete sc; }

int oracle_cloth_scene_add_mesh(OracleClothScene* sc, int nv, const double* pos, const double* vel, int nf,
                                const int32_t* faces, const double* material, int nfixed, const int32_t* fixed,
                                const int32_t* colors) {
    if (sc->precision == 64) return cloth_add(sc->s64, nv, pos, vel, nf, faces, material, nfixed, fixed, colors);
    return cloth_add(sc->s32, nv, pos, vel, nf, faces, material, nfixed, fixed, colors);
}

int oracle_cloth_scene_finalize(OracleClothScene* sc) {
    if (sc->precision == 64) cloth_groups(sc->s64);
    else cloth_groups(sc->s32);
    return 0;
}

int oracle_cloth_scene_num_groups(OracleClothScene* sc) {
    return sc->precision == 64 ? (int)sc->s64.groups.size() : (int)sc->s32.groups.size();
}

int oracle_cloth_scene_group(OracleClothScene* sc, int g, int32_t* out) {
    auto& grp = sc->precision == 64 ? sc->s64.groups : sc->s32.groups;
    if (g < 0 || g >= (int)grp.size()) return -1;
    if (out) std::memcpy(out, grp[g].data(), grp[g].size() * sizeof(int32_t));
    return (int)grp[g].size() / 2;
}

int oracle_cloth_scene_step(OracleClothScene* sc, int nsteps, int nthreads) {
    if (nthreads > 1 && (!sc->pool || sc->pool->size() != nthreads)) sc->pool.reset(new oracle::ThreadPool(nthreads));
    oracle::ThreadPool* pool = nthreads > 1 ? sc->pool.get() : nullptr;
    for (int i = 0; i < nsteps; ++i) {
        if (sc->precision == 64) cloth_step(sc->s64, pool);
        else cloth_step(sc->s32, pool);
    }
    return 0;
}

int oracle_cloth_scene_get_state(OracleClothScene* sc, int mesh, double* pos, double* vel) {
    return sc->precision == 64 ? cloth_state(sc->s64, mesh, pos, vel, true) : cloth_state(sc->s32, mesh, pos, vel, true);
}
int oracle_cloth_scene_set_state(OracleClothScene* sc, int mesh, const double* pos, const double* vel) {
    return sc->precision == 64 ? cloth_state(sc->s64, mesh, (double*)pos, (double*)vel, false)
                               : cloth_state(sc->s32, mesh, (double*)pos, (double*)vel, false);
}

}  // extern "C"
