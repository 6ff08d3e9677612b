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
#include <array>
#include <map>
#include <atomic>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <utility>
#include <vector>

#include "oracle.h"
#include "tet_core.h"
#include "deform.h"
#include "pool.h"

// contact detector oracle (detector_oracle.cpp)
extern "C" {
void* oracle_cd_create(int nmesh, const int* mesh_nv, const int* mesh_nf, const int* faces);
void oracle_cd_destroy(void* h);
void oracle_cd_detect(void* h, const float* pos, float radius, int mode, int threads);
void oracle_cd_counts(void* h, int64_t* counts);
void oracle_cd_vf_contacts(void* h, int* rec, float* d);
void oracle_cd_ee_contacts(void* h, int* ee, float* mu, float* d);
void oracle_cd_bounds(void* h, float thickness, float relaxation, float* bounds);
void oracle_cd_set_query_mask(void* h, const uint8_t* mask);
}

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
    R frictionMu = R(0.1), frictionEpsV = R(0.1);  // contact friction (TriMesh.h:41-42 defaults)
    int initType = 2;
    bool noGravZone = false;                          // hasNoGravZone (Materials.h:34)
    R noGravThreshold = 0, maxVel = -1;               // noGravZoneThreshold, maxVelocityMagnitude
    std::vector<int32_t> veBegin, veEdge;             // vertex -> incident edges (ascending)
    // tet meshes in a contact scene: neo-Hookean elements; faces/edges above are the boundary
    // surface, which takes part in contact
    bool tet = false;
    int nt = 0;
    std::vector<int32_t> tets, incBegin, incTet, incOrder;
    std::vector<R> tdminv, tvol, vprev;
    std::vector<uint8_t> surface;                     // vertex lies on the boundary surface
    R dampH = 0, dampD = 0;
    std::vector<std::vector<oracle::Attachment>> att; // per vertex attachment records (cloth)
};

// Contact state of a scene with handle_collision (global ids: meshes concatenated in order).
struct ContactState {
    void* cd = nullptr;                      // detector oracle (detector_oracle.cpp)
    std::vector<int> vOff, fOff, eOff;       // per mesh, + total
    std::vector<int> vMesh;                  // global vertex -> mesh
    std::vector<std::vector<int>> vfFaces;   // global vertex -> its contact faces (lowest 8 face ids)
    std::vector<std::vector<int>> fvVerts;   // global face -> vertices in contact with it (lowest 32 ids)
    std::vector<std::vector<int>> eeOther;   // global edge -> other edges in contact (lowest 8 ids)
    std::vector<float> bound;                // conservative bound per global vertex
    std::vector<double> xAtCD;               // positions at the last detection (3 per global vertex)
    std::atomic<int> required{0};            // a vertex left its bound: detect before the next iteration
    ~ContactState() {
        if (cd) oracle_cd_destroy(cd);
    }
};

template <class R> struct ClothScene {
    double p[OP_COUNT];
    R dt, dtSqrRecip;
    std::vector<ClothMesh<R>> meshes;
    std::vector<std::vector<int32_t>> groups;
    std::unique_ptr<ContactState> contact;   // set by finalize when OP_HANDLE_COLLISION != 0
    int threads = 1;
    std::vector<oracle::Deformer> deformers;
    float clock = 0.f;                       // curTime (float in the reference)
    std::atomic<long long> solveFailures{0}; // vertex solves whose 3x3 block had no solution
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
    m.veBegin.assign(m.nv + 1, 0);
    m.veEdge.clear();
    for (int v = 0; v < m.nv; ++v) {
        m.veEdge.insert(m.veEdge.end(), vEdges[v].begin(), vEdges[v].end());
        m.veBegin[v + 1] = (int)m.veEdge.size();
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

// ---------------------------------------------------------------------------
// Cloth contact (VBDClothPhysics.cpp:1248-1399, 1500-1573), evaluated from the current positions
// as the reference does (updateContactInfo = true recomputes closest points every evaluation).
// ---------------------------------------------------------------------------

// CollisionGeometry.h:90-157 in precision R; bary receives the closest point's barycentrics
template <class R> void closest_point_tri(const R* p, const R* a, const R* b, const R* c, R* cp, R* bary) {
    R ab[3], ac[3], ap[3], bp[3], q[3];
    for (int d = 0; d < 3; ++d) {
        ab[d] = b[d] - a[d];
        ac[d] = c[d] - a[d];
        ap[d] = p[d] - a[d];
    }
    const R d1 = dot3(ab, ap), d2 = dot3(ac, ap);
    auto set = [&](const R* x, R b0, R b1, R b2) {
        for (int d = 0; d < 3; ++d) cp[d] = x[d];
        bary[0] = b0;
        bary[1] = b1;
        bary[2] = b2;
    };
    if (d1 <= 0 && d2 <= 0) return set(a, 1, 0, 0);
    for (int d = 0; d < 3; ++d) bp[d] = p[d] - b[d];
    const R d3 = dot3(ab, bp), d4 = dot3(ac, bp);
    if (d3 >= 0 && d4 <= d3) return set(b, 0, 1, 0);
    R cpv[3];
    for (int d = 0; d < 3; ++d) cpv[d] = p[d] - c[d];
    const R d5 = dot3(ab, cpv), d6 = dot3(ac, cpv);
    if (d6 >= 0 && d5 <= d6) return set(c, 0, 0, 1);
    const R vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) {
        const R v = d1 / (d1 - d3);
        for (int d = 0; d < 3; ++d) q[d] = a[d] + v * ab[d];
        return set(q, R(1) - v, v, 0);
    }
    const R vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) {
        const R v = d2 / (d2 - d6);
        for (int d = 0; d < 3; ++d) q[d] = a[d] + v * ac[d];
        return set(q, R(1) - v, 0, v);
    }
    const R va = d3 * d6 - d5 * d4;
    if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
        const R v = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        for (int d = 0; d < 3; ++d) q[d] = b[d] + v * (c[d] - b[d]);
        return set(q, 0, R(1) - v, v);
    }
    const R denom = R(1) / (va + vb + vc);
    const R v = vb * denom, w = vc * denom;
    for (int d = 0; d < 3; ++d) q[d] = (a[d] + v * ab[d]) + w * ac[d];
    return set(q, R(1) - v - w, v, w);
}

// CollisionGeometry.h:160-267 in precision R (absolute CMP_EPSILON as in the reference)
template <class R> void segment_closest(const R* p0, const R* p1, const R* q0, const R* q1, R* c1, R* c2, R& s, R& t) {
    R p[3], q[3], r[3];
    for (int d = 0; d < 3; ++d) {
        p[d] = p1[d] - p0[d];
        q[d] = q1[d] - q0[d];
        r[d] = p0[d] - q0[d];
    }
    const R a = dot3(p, p), b = dot3(p, q), c = dot3(q, q), d = dot3(p, r), e = dot3(q, r);
    s = 0;
    t = 0;
    const R det = a * c - b * b;
    if (det > R(0.00001)) {
        const R bte = b * e, ctd = c * d;
        if (bte <= ctd) {
            if (e <= 0) {
                s = (-d >= a ? R(1) : (-d > 0 ? -d / a : R(0)));
                t = 0;
            } else if (e < c) {
                s = 0;
                t = e / c;
            } else {
                s = (b - d >= a ? R(1) : (b - d > 0 ? (b - d) / a : R(0)));
                t = 1;
            }
        } else {
            s = bte - ctd;
            if (s >= det) {
                if (b + e <= 0) {
                    s = (-d <= 0 ? R(0) : (-d < a ? -d / a : R(1)));
                    t = 0;
                } else if (b + e < c) {
                    s = 1;
                    t = (b + e) / c;
                } else {
                    s = (b - d <= 0 ? R(0) : (b - d < a ? (b - d) / a : R(1)));
                    t = 1;
                }
            } else {
                const R ate = a * e, btd = b * d;
                if (ate <= btd) {
                    s = (-d <= 0 ? R(0) : (-d >= a ? R(1) : -d / a));
                    t = 0;
                } else {
                    t = ate - btd;
                    if (t >= det) {
                        s = (b - d <= 0 ? R(0) : (b - d >= a ? R(1) : (b - d) / a));
                        t = 1;
                    } else {
                        s /= det;
                        t /= det;
                    }
                }
            }
        }
    } else {
        if (e <= 0) {
            s = (-d <= 0 ? R(0) : (-d >= a ? R(1) : -d / a));
            t = 0;
        } else if (e >= c) {
            s = (b - d <= 0 ? R(0) : (b - d >= a ? R(1) : (b - d) / a));
            t = 1;
        } else {
            s = 0;
            t = e / c;
        }
    }
    for (int k = 0; k < 3; ++k) {
        c1[k] = (R(1) - s) * p0[k] + s * p1[k];
        c2[k] = (R(1) - t) * q0[k] + t * q1[k];
    }
}

// VBDClothPhysics.cpp:1533-1573 (energy types 0 quadratic, 1 quadratic-log two-stage, 2 IPC-like)
template <class R> void repulsive(const ClothScene<R>& s, R dis, R& dEdD, R& d2E) {
    const R adj = dis - (R)s.p[OP_THICKNESS];
    const R radius = (R)s.p[OP_CONTACT_RADIUS];
    const R pen = radius - adj;
    const R k = (R)s.p[OP_CONTACT_STIFFNESS];
    switch ((int)s.p[OP_CONTACT_ENERGY_TYPE]) {
    case 0:
        dEdD = -k * pen;
        d2E = k;
        break;
    case 1: {
        const R tau = radius * R(0.5);
        if (adj < tau && adj > R(0.00001)) {
            const R k2 = R(0.5) * tau * tau * k;
            dEdD = -k2 / adj;
            d2E = k2 / (adj * adj);
        } else {
            dEdD = -k * pen;
            d2E = k;
        }
        break;
    }
    case 2:
        dEdD = -k * (adj - radius) * (adj - radius) / adj - R(2) * k * std::log(adj / radius) * (adj - radius);
        d2E = k * ((adj - radius) * (adj - radius) / (adj * adj) - (R(4) * adj - R(4) * radius) / adj -
                   R(2) * std::log(adj / radius));
        break;
    default:
        dEdD = 0;
        d2E = 0;
    }
}

template <class R> inline const R* gx(ClothScene<R>& s, int gv) {
    const ContactState& C = *s.contact;
    const int m = C.vMesh[gv];
    return &s.meshes[m].x[3 * (size_t)(gv - C.vOff[m])];
}
template <class R> inline const R* gxp(ClothScene<R>& s, int gv) {
    const ContactState& C = *s.contact;
    const int m = C.vMesh[gv];
    return &s.meshes[m].xprev[3 * (size_t)(gv - C.vOff[m])];
}
inline int mesh_of(const std::vector<int>& off, int g) {
    return (int)(std::upper_bound(off.begin(), off.end(), g) - off.begin()) - 1;
}
template <class R> inline void normalized(const R* v, R* out) {
    const R n = std::sqrt(dot3(v, v));
    for (int d = 0; d < 3; ++d) out[d] = n > 0 ? v[d] / n : v[d];
}

// contact force / Hessian on the vertex of weight b (1 for the vertex side, -bary_o for face vertex o)
template <class R> void contact_apply(ClothScene<R>& s, R b, R lambda, R d2E, const R* n, const R* dx, const R* t0,
                                      R mu, R epsV, R* f, R* h) {
    for (int r = 0; r < 3; ++r) {
        f[r] += b * lambda * n[r];
        for (int c = 0; c < 3; ++c) h[r + 3 * c] += d2E * b * b * n[r] * n[c];
    }
    R e0[3], e1[3], cr[3];
    normalized(t0, e0);
    cross3(e0, n, cr);
    normalized(cr, e1);
    const R T[6] = {e0[0], e0[1], e0[2], e1[0], e1[1], e1[2]};
    const R u[2] = {dot3(e0, dx), dot3(e1, dx)};
    R ff[3], fh[9];
    friction<R>(mu, lambda, T, u, epsV * s.dt, ff, fh);
    for (int r = 0; r < 3; ++r) f[r] += b * ff[r];
    for (int i = 0; i < 9; ++i) h[i] += b * b * fh[i];
}

// VF contact of vertex gu with face gf, seen from the vertex of order `order` (3 = the vertex side)
template <class R> void vf_contact(ClothScene<R>& s, int gu, int gf, int order, R* f, R* h) {
    const ContactState& C = *s.contact;
    const int mf = mesh_of(C.fOff, gf), mu_ = C.vMesh[gu];
    const ClothMesh<R>& F = s.meshes[mf];
    const int32_t* fv = &F.faces[3 * (size_t)(gf - C.fOff[mf])];
    const int g0 = C.vOff[mf] + fv[0], g1 = C.vOff[mf] + fv[1], g2 = C.vOff[mf] + fv[2];
    const R* x = gx(s, gu);
    R cp[3], bary[3];
    closest_point_tri(x, gx(s, g0), gx(s, g1), gx(s, g2), cp, bary);
    R diff[3] = {x[0] - cp[0], x[1] - cp[1], x[2] - cp[2]}, n[3];
    normalized(diff, n);
    const R dis = dot3(diff, n);
    if (!(dis < (R)(s.p[OP_CONTACT_RADIUS] + s.p[OP_THICKNESS]))) return;
    const R bs[4] = {-bary[0], -bary[1], -bary[2], R(1)};
    R dEdD, d2E;
    repulsive(s, dis, dEdD, d2E);
    const R *p0 = gxp(s, g0), *p1 = gxp(s, g1), *p2 = gxp(s, g2), *xp = gxp(s, gu);
    R dx[3], t0[3];
    for (int d = 0; d < 3; ++d) {
        const R cpPrev = bary[0] * p0[d] + bary[1] * p1[d] + bary[2] * p2[d];
        dx[d] = (x[d] - xp[d]) - (cp[d] - cpPrev);
        t0[d] = gx(s, g1)[d] - gx(s, g0)[d];
    }
    const R mu = (F.frictionMu + s.meshes[mu_].frictionMu) * R(0.5);
    const R epsV = (F.frictionEpsV + s.meshes[mu_].frictionEpsV) * R(0.5);
    contact_apply(s, bs[order], -dEdD, d2E, n, dx, t0, mu, epsV, f, h);
}

// EE contact of edge ge1 with ge2, seen from the vertex of order `order` (0 eV1, 1 eV2 of ge1)
template <class R> void ee_contact(ClothScene<R>& s, int ge1, int ge2, int order, R* f, R* h) {
    const ContactState& C = *s.contact;
    const int m1 = mesh_of(C.eOff, ge1), m2 = mesh_of(C.eOff, ge2);
    const int32_t* E1 = &s.meshes[m1].edges[6 * (size_t)(ge1 - C.eOff[m1])];
    const int32_t* E2 = &s.meshes[m2].edges[6 * (size_t)(ge2 - C.eOff[m2])];
    const int a0 = C.vOff[m1] + E1[0], a1 = C.vOff[m1] + E1[1], b0 = C.vOff[m2] + E2[0], b1 = C.vOff[m2] + E2[1];
    R c1[3], c2[3], mu1, mu2;
    segment_closest(gx(s, a0), gx(s, a1), gx(s, b0), gx(s, b1), c1, c2, mu1, mu2);
    R diff[3] = {c1[0] - c2[0], c1[1] - c2[1], c1[2] - c2[2]};
    const R dis = std::sqrt(dot3(diff, diff));
    R v1[3], v2[3], cr[3];
    for (int d = 0; d < 3; ++d) {
        v1[d] = gx(s, a1)[d] - gx(s, a0)[d];
        v2[d] = gx(s, b1)[d] - gx(s, b0)[d];
    }
    cross3(v1, v2, cr);
    const R parallelEps = dot3(cr, cr) / (std::sqrt(dot3(v1, v1)) * std::sqrt(dot3(v2, v2)));
    if (!(dis < (R)(s.p[OP_CONTACT_RADIUS] + s.p[OP_THICKNESS]) && parallelEps > R(0.00001))) return;
    R n[3] = {diff[0] / dis, diff[1] / dis, diff[2] / dis};
    const R bs[4] = {R(1) - mu1, mu1, R(-1) + mu2, -mu2};
    R dEdD, d2E;
    repulsive(s, dis, dEdD, d2E);
    R dx[3];
    for (int d = 0; d < 3; ++d) {
        const R c1Prev = (R(1) - mu1) * gxp(s, a0)[d] + mu1 * gxp(s, a1)[d];
        const R c2Prev = (R(1) - mu2) * gxp(s, b0)[d] + mu2 * gxp(s, b1)[d];
        dx[d] = (c1[d] - c1Prev) - (c2[d] - c2Prev);
    }
    const R mu = (s.meshes[m1].frictionMu + s.meshes[m2].frictionMu) * R(0.5);
    const R epsV = (s.meshes[m1].frictionEpsV + s.meshes[m2].frictionEpsV) * R(0.5);
    contact_apply(s, bs[order], -dEdD, d2E, n, dx, v1, mu, epsV, f, h);
}

// applyCollisionDetection (VBDClothPhysics.cpp:735-819): detection at the current positions,
// per-primitive contact lists (lowest ids kept: 8 VF per vertex, 32 FV per face, 8 EE per edge),
// conservative bounds and the positions they are measured from.
template <class R> void cloth_detect(ClothScene<R>& s) {
    ContactState& C = *s.contact;
    const int nv = C.vOff.back(), nf = C.fOff.back(), ne = C.eOff.back();
    std::vector<float> pos(3 * (size_t)nv);
    for (int gv = 0; gv < nv; ++gv)
        for (int d = 0; d < 3; ++d) pos[3 * (size_t)gv + d] = (float)gx(s, gv)[d];
    oracle_cd_detect(C.cd, pos.data(), (float)s.p[OP_MAX_QUERY_DIS], 1, s.threads);
    int64_t cnt[3];
    oracle_cd_counts(C.cd, cnt);
    std::vector<int> vf(4 * (size_t)cnt[1]), ee(2 * (size_t)cnt[2]);
    oracle_cd_vf_contacts(C.cd, vf.data(), nullptr);
    oracle_cd_ee_contacts(C.cd, ee.data(), nullptr, nullptr);
    C.vfFaces.assign(nv, {});
    C.fvVerts.assign(nf, {});
    C.eeOther.assign(ne, {});
    for (int64_t i = 0; i < cnt[1]; ++i) C.vfFaces[vf[4 * i]].push_back(vf[4 * i + 3]);
    for (int gv = 0; gv < nv; ++gv) {
        auto& L = C.vfFaces[gv];
        std::sort(L.begin(), L.end());
        if (L.size() > 8) L.resize(8);
        for (int f : L) C.fvVerts[f].push_back(gv);
    }
    for (auto& L : C.fvVerts)
        if (L.size() > 32) L.resize(32);  // vertices were appended in ascending order
    for (int64_t i = 0; i < cnt[2]; ++i) C.eeOther[ee[2 * i]].push_back(ee[2 * i + 1]);
    for (auto& L : C.eeOther) {
        std::sort(L.begin(), L.end());
        if (L.size() > 8) L.resize(8);
    }
    C.bound.assign(nv, 0.f);
    oracle_cd_bounds(C.cd, (float)s.p[OP_THICKNESS], (float)s.p[OP_CONSERVATIVE_RELAXATION], C.bound.data());
    C.xAtCD.resize(3 * (size_t)nv);
    for (int gv = 0; gv < nv; ++gv)
        for (int d = 0; d < 3; ++d) C.xAtCD[3 * (size_t)gv + d] = (double)gx(s, gv)[d];
    C.required = 0;
}

template <class R> void cloth_vertex(ClothScene<R>& s, ClothMesh<R>& m, int vid) {
    R h[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    R f[3] = {0, 0, 0};
    const R* xv = &m.x[3 * vid];
    const int mi = (int)(&m - s.meshes.data());
    if (s.contact) {
        // VBDClothPhysics.cpp:1033-1094: own VF contacts, FV contacts of incident faces, EE
        // contacts of incident edges
        const ContactState& C = *s.contact;
        const int gv = C.vOff[mi] + vid;
        for (int gf : C.vfFaces[gv]) vf_contact(s, gv, gf, 3, f, h);
        for (int k = m.vfBegin[vid]; k < m.vfBegin[vid + 1]; ++k) {
            const int gf = C.fOff[mi] + m.vfFace[k];
            for (int gu : C.fvVerts[gf]) vf_contact(s, gu, gf, m.vfOrder[k], f, h);
        }
        for (int k = m.veBegin[vid]; k < m.veBegin[vid + 1]; ++k) {
            const int e = m.veEdge[k];
            const int ge = C.eOff[mi] + e;
            const int order = m.edges[6 * e] == vid ? 0 : 1;
            for (int ge2 : C.eeOther[ge]) ee_contact(s, ge, ge2, order, f, h);
        }
    }
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
    // vertex attachments (VBDClothPhysics.cpp:1465-1498)
    if (!m.att.empty())
        for (const oracle::Attachment& A : m.att[vid]) {
            const R* xr = &s.meshes[A.readMesh].x[3 * (size_t)A.readVert];
            R tgt[3];
            if (A.tgtMesh >= 0) {
                const std::vector<R>& tx = s.meshes[A.tgtMesh].x;
                for (int d = 0; d < 3; ++d)
                    tgt[d] = (R)A.p[0] * tx[3 * (size_t)A.f[0] + d] + (R)A.p[1] * tx[3 * (size_t)A.f[1] + d] +
                             (R)A.p[2] * tx[3 * (size_t)A.f[2] + d];
            } else {
                for (int d = 0; d < 3; ++d) tgt[d] = (R)A.p[d];
            }
            const R kb = (R)A.k * (R)A.b;
            for (int d = 0; d < 3; ++d) f[d] -= kb * (xr[d] - tgt[d]);
            const R hk = kb * (R)A.b;
            h[0] += hk;
            h[4] += hk;
            h[8] += hk;
        }
    R dx[3];
    R* xn = &m.xnext[3 * vid];
    if (solve3<R>(h, f, dx)) {
        const R step = (R)s.p[OP_STEP_SIZE];
        for (int d = 0; d < 3; ++d) xn[d] = xv[d] + step * dx[d];
        if (s.contact) {
            // applyConservativeBoundTruncation (VBDClothPhysics.cpp:1223-1246). The reference
            // assigns the shared re-detection flag from every vertex (a race); any truncation
            // requests detection here.
            ContactState& C = *s.contact;
            const int gv = C.vOff[mi] + vid;
            R acc[3];
            for (int d = 0; d < 3; ++d) acc[d] = xn[d] - (R)C.xAtCD[3 * (size_t)gv + d];
            const R size = std::sqrt(dot3(acc, acc));
            const R b = (R)C.bound[gv];
            if (size > b) {
                for (int d = 0; d < 3; ++d) xn[d] = (R)C.xAtCD[3 * (size_t)gv + d] + acc[d] * (b / size);
                C.required = 1;
            }
        }
    } else {
        for (int d = 0; d < 3; ++d) xn[d] = xv[d];
        ++s.solveFailures;
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
    if (m.initType == 0) m.x = m.y;
    if (m.initType == 3) {
        // inertia within the conservative bounds (VBDBaseTriMesh.cpp:137-152; bounds infinite
        // without collision handling)
        const int mi = (int)(&m - s.meshes.data());
        for (int i = 0; i < m.nv; ++i) {
            R step[3];
            for (int d = 0; d < 3; ++d) step[d] = m.y[3 * i + d] - m.xprev[3 * i + d];
            const R size = std::sqrt(dot3(step, step));
            if (s.contact) {
                const R b = (R)s.contact->bound[s.contact->vOff[mi] + i];
                if (size > b && size != 0)
                    for (int d = 0; d < 3; ++d) step[d] = step[d] * (b / size);
            }
            for (int d = 0; d < 3; ++d) m.x[3 * i + d] = m.xprev[3 * i + d] + step[d];
        }
    }
    // 2: previous position (x unchanged), 4: nothing
}

template <class R> void cloth_velocity(ClothScene<R>& s, ClothMesh<R>& m) {
    for (int i = 0; i < m.nv; ++i) {
        R* v = &m.v[3 * i];
        for (int d = 0; d < 3; ++d) v[d] = (m.x[3 * i + d] - m.xprev[3 * i + d]) / s.dt;
        tetcore::damp_velocity<R>(v, m.noGravZone, v[1], m.noGravThreshold, m.maxVel, m.expVel, m.constVel);
    }
    for (int f : m.fixedList)
        for (int d = 0; d < 3; ++d) m.v[3 * f + d] = 0;
}

// ---------------------------------------------------------------------------
// Tet meshes with cloth-model contact on their boundary surface. The reference's tet solver
// uses its own penetration (DCD/CCD) contact; this composition keeps each tet vertex's solve as
// the reference's (VBDPhysics.cpp:2402-2483) and adds, for boundary vertices, the cloth VF/EE
// contact energies, friction and conservative-bound truncation (VBDClothPhysics.cpp:1033-1094,
// 1223-1246). Vertices of one colour are solved Jacobi-style (positionsNext) as cloth is, since
// contact pairs may share a colour.
// ---------------------------------------------------------------------------

// Boundary faces: faces of exactly one tet, visited in (tet, face) order, each oriented with
// its normal away from the tet's fourth vertex.
template <class R> void tet_surface(int nt, const int32_t* tets, const R* x, std::vector<int32_t>& faces) {
    static const int opp[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
    std::map<std::array<int32_t, 3>, int> count;
    for (int t = 0; t < nt; ++t)
        for (int k = 0; k < 4; ++k) {
            std::array<int32_t, 3> key = {tets[4 * t + opp[k][0]], tets[4 * t + opp[k][1]], tets[4 * t + opp[k][2]]};
            std::sort(key.begin(), key.end());
            count[key]++;
        }
    faces.clear();
    for (int t = 0; t < nt; ++t)
        for (int k = 0; k < 4; ++k) {
            int32_t a = tets[4 * t + opp[k][0]], b = tets[4 * t + opp[k][1]], c = tets[4 * t + opp[k][2]];
            std::array<int32_t, 3> key = {a, b, c};
            std::sort(key.begin(), key.end());
            if (count[key] != 1) continue;
            const int32_t o = tets[4 * t + k];
            R e1[3], e2[3], n[3], w[3];
            for (int d = 0; d < 3; ++d) {
                e1[d] = x[3 * b + d] - x[3 * a + d];
                e2[d] = x[3 * c + d] - x[3 * a + d];
                w[d] = x[3 * o + d] - x[3 * a + d];
            }
            cross3(e1, e2, n);
            if (dot3(n, w) > 0) std::swap(b, c);
            faces.push_back(a);
            faces.push_back(b);
            faces.push_back(c);
        }
}

// the contacts of global vertex gv of mesh mi (own VF, FV of incident faces, EE of incident edges)
template <class R> void vertex_contacts(ClothScene<R>& s, ClothMesh<R>& m, int mi, int vid, R* f, R* h) {
    const ContactState& C = *s.contact;
    const int gv = C.vOff[mi] + vid;
    for (int gf : C.vfFaces[gv]) vf_contact(s, gv, gf, 3, f, h);
    for (int k = m.vfBegin[vid]; k < m.vfBegin[vid + 1]; ++k) {
        const int gf = C.fOff[mi] + m.vfFace[k];
        for (int gu : C.fvVerts[gf]) vf_contact(s, gu, gf, m.vfOrder[k], f, h);
    }
    for (int k = m.veBegin[vid]; k < m.veBegin[vid + 1]; ++k) {
        const int e = m.veEdge[k];
        const int ge = C.eOff[mi] + e;
        const int order = m.edges[6 * e] == vid ? 0 : 1;
        for (int ge2 : C.eeOther[ge]) ee_contact(s, ge, ge2, order, f, h);
    }
}

// VBDPhysics.cpp:2402-2483 for one tet vertex, plus surface contact; writes positionsNext
template <class R> void tet_vertex(ClothScene<R>& s, ClothMesh<R>& m, int vid, bool applyFriction) {
    R h[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    R f[3] = {0, 0, 0};
    const R* xv = &m.x[3 * vid];
    const int mi = (int)(&m - s.meshes.data());
    const R disp[3] = {xv[0] - m.xprev[3 * vid], xv[1] - m.xprev[3 * vid + 1], xv[2] - m.xprev[3 * vid + 2]};
    for (int k = m.incBegin[vid]; k < m.incBegin[vid + 1]; ++k) {
        const int t = m.incTet[k];
        R x4[12];
        for (int j = 0; j < 4; ++j)
            for (int d = 0; d < 3; ++d) x4[3 * j + d] = m.x[3 * m.tets[4 * t + j] + d];
        tetcore::neohookean_vertex<R>(x4, &m.tdminv[9 * t], m.tvol[t], m.mu, m.lmbd, m.incOrder[k], m.dampH, m.dampD,
                                      s.dt, disp, f, h);
    }
    const R mass = m.mass[vid];
    for (int d = 0; d < 3; ++d) f[d] += mass * (m.y[3 * vid + d] - xv[d]) * s.dtSqrRecip;
    const R hi = mass * s.dtSqrRecip;
    h[0] += hi; h[4] += hi; h[8] += hi;
    if (s.p[OP_USE_PLANE_GROUND] != 0) {
        const R k = (R)s.p[OP_BOUNDARY_STIFFNESS];
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
            if (applyFriction) {
                R T[6] = {0, 0, 0, 0, 0, 0};
                T[(dim + 1) % 3] = 1;
                T[3 + (dim + 2) % 3] = 1;
                R u[2] = {T[0] * disp[0] + T[1] * disp[1] + T[2] * disp[2],
                          T[3] * disp[0] + T[4] * disp[1] + T[5] * disp[2]};
                R ff[3], fh[9];
                friction<R>((R)s.p[OP_BOUNDARY_FRICTION_MU], pen * k, T, u,
                            (R)s.p[OP_BOUNDARY_FRICTION_EPS_V] * s.dt, ff, fh);
                for (int d = 0; d < 3; ++d) f[d] += ff[d];
                for (int i = 0; i < 9; ++i) h[i] += fh[i];
            }
        }
    }
    if (s.contact && m.surface[vid]) vertex_contacts(s, m, mi, vid, f, h);
    R* xn = &m.xnext[3 * vid];
    for (int d = 0; d < 3; ++d) xn[d] = xv[d];
    if (f[0] * f[0] + f[1] * f[1] + f[2] * f[2] > R(1e-10)) {
        R dx[3];
        R step = (R)s.p[OP_STEP_SIZE];
        if (!solve3<R>(h, f, dx)) {
            step = (R)s.p[OP_STEP_SIZE_GD];
            dx[0] = f[0]; dx[1] = f[1]; dx[2] = f[2];
            ++s.solveFailures;
        }
        for (int d = 0; d < 3; ++d) xn[d] = xv[d] + step * dx[d];
    }
    // The tet initial guess is not bounded, so the truncation applies to every boundary vertex,
    // moved or not: a vertex whose force is under the update threshold still returns within its
    // conservative bound (otherwise it could stay past it, and whether it is truncated would hang
    // on rounding of a near-zero force).
    if (s.contact && m.surface[vid]) {
        ContactState& C = *s.contact;
        const int gv = C.vOff[mi] + vid;
        R acc[3];
        for (int d = 0; d < 3; ++d) acc[d] = xn[d] - (R)C.xAtCD[3 * (size_t)gv + d];
        const R size = std::sqrt(dot3(acc, acc));
        const R b = (R)C.bound[gv];
        if (size > b) {
            for (int d = 0; d < 3; ++d) xn[d] = (R)C.xAtCD[3 * (size_t)gv + d] + acc[d] * (b / size);
            C.required = 1;
        }
    }
}

// VBD_BaseMaterial.h:244-259, VBD_BaseMaterial.cpp:7-50 (initialisation types 0-4)
template <class R> void tet_initial_step(ClothScene<R>& s, ClothMesh<R>& m) {
    const R dt = s.dt;
    const R g[3] = {(R)s.p[OP_GRAVITY_X], (R)s.p[OP_GRAVITY_Y], (R)s.p[OP_GRAVITY_Z]};
    m.xprev = m.x;
    m.vprev = m.v;
    if (s.p[OP_ASSOCIATE_GRAVITY] != 0) {
        const R dg[3] = {dt * g[0], dt * g[1], dt * g[2]};
        for (int i = 0; i < m.nv; ++i)
            for (int d = 0; d < 3; ++d) m.v[3 * i + d] += dg[d];
    }
    for (int f : m.fixedList)
        for (int d = 0; d < 3; ++d) m.v[3 * f + d] = 0;
    for (int i = 0; i < 3 * m.nv; ++i) m.y[i] = m.x[i] + dt * m.v[i];
    switch (m.initType) {
    case 0: m.x = m.y; break;
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

// VBDPhysics.cpp:2486-2561 (damping pre-scaled by timeStepEqualize at load)
template <class R> void tet_velocity(ClothScene<R>& s, ClothMesh<R>& m) {
    for (int i = 0; i < m.nv; ++i) {
        R* v = &m.v[3 * i];
        for (int d = 0; d < 3; ++d) v[d] = (m.x[3 * i + d] - m.xprev[3 * i + d]) / s.dt;
        tetcore::damp_velocity<R>(v, m.noGravZone, m.x[3 * i + 1], m.noGravThreshold, m.maxVel, m.expVel,
                                  m.constVel);
    }
    for (int f : m.fixedList)
        for (int d = 0; d < 3; ++d) m.v[3 * f + d] = 0;
}

template <class R> void cloth_step(ClothScene<R>& s, oracle::ThreadPool* pool) {
    const int nsub = (int)s.p[OP_NUM_SUBSTEPS], iters = (int)s.p[OP_ITERATIONS];
    // deformers (VBDClothPhysics.cpp:408-412): cloth meshes' deformed vertices move at iteration 0
    // before detection and the initial guess; tet meshes' after the initial step (as the tet
    // framework does, VBDPhysics.cpp:849-862)
    auto deform = [&](bool clothPhase) {
        for (const auto& D : s.deformers)
            for (size_t k = 0; k < D.meshes.size(); ++k) {
                ClothMesh<R>& m = s.meshes[D.meshes[k]];
                if (m.tet == clothPhase) continue;
                oracle::deformer_apply<R>(D, (int)k, clothPhase, s.clock, (float)s.dt, m.x.data());
            }
    };
    auto initial = [&]() {
        for (auto& m : s.meshes) m.tet ? tet_initial_step(s, m) : cloth_initial_step(s, m);
        deform(false);
    };
    for (int sub = 0; sub < nsub; ++sub) {
        s.clock += (float)s.dt;
        deform(true);
        for (int it = 0; it < iters; ++it) {
            // detection before the iteration's sweep: always at iteration 0 (before the initial
            // guess), later when a vertex left its bound (VBDClothPhysics.cpp:418-434)
            if (s.contact && (it == 0 || s.contact->required)) cloth_detect(s);
            if (it == 0) initial();
            const bool friction = it >= (int)s.p[OP_FRICTION_START_ITER];
            for (auto& grp : s.groups) {
                const int n = (int)grp.size() / 2;
                auto solve = [&](int i) {
                    ClothMesh<R>& m = s.meshes[grp[2 * i]];
                    int vid = grp[2 * i + 1];
                    if (m.fixed[vid]) return;
                    if (m.tet)
                        tet_vertex(s, m, vid, friction);
                    else
                        cloth_vertex(s, m, vid);
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
        if (iters == 0) initial();
        for (auto& m : s.meshes) m.tet ? tet_velocity(s, m) : cloth_velocity(s, m);
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
    m.frictionMu = (R)mat[OMC_FRICTION_MU];
    m.frictionEpsV = (R)mat[OMC_FRICTION_EPS_V];
    m.noGravZone = mat[OMC_NO_GRAV_ZONE] != 0;
    m.noGravThreshold = (R)mat[OMC_NO_GRAV_THRESHOLD];
    m.maxVel = (R)mat[OMC_MAX_VELOCITY];
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

// A tet mesh (VBD/VBDPhysics.cpp material setup, TetMeshFEM.cpp:332-443) in a scene with
// collision handling: its boundary surface takes part in contact.
template <class R>
static int tet_add(ClothScene<R>& s, int nv, const double* pos, const double* vel, int nt, const int32_t* tets,
                   const double* mat, int nfixed, const int32_t* fixed, const int32_t* colors) {
    ClothMesh<R> m;
    m.tet = true;
    m.nv = nv;
    m.nt = nt;
    m.x.resize(3 * nv);
    m.v.assign(3 * nv, 0);
    for (int i = 0; i < 3 * nv; ++i) m.x[i] = (R)pos[i];
    if (vel)
        for (int i = 0; i < 3 * nv; ++i) m.v[i] = (R)vel[i];
    m.xprev = m.x;
    m.vprev = m.v;
    m.y = m.x;
    m.xnext = m.x;
    m.tets.assign(tets, tets + 4 * nt);
    m.density = (R)mat[OM_DENSITY];
    m.mu = (R)mat[OM_MU];
    m.lmbd = (R)mat[OM_LAMBDA];
    m.dampH = (R)mat[OM_DAMP_HYDRO];
    m.dampD = (R)mat[OM_DAMP_DEV];
    m.initType = (int)mat[OM_INIT_TYPE];
    m.expVel = (R)std::pow((R)mat[OM_EXP_VEL_DAMPING], s.dt);  // timeStepEqualize (VBDPhysics.cpp:794-807)
    m.constVel = (R)mat[OM_CONST_VEL_DAMPING] * s.dt;
    m.frictionMu = (R)mat[OM_FRICTION_MU];
    m.frictionEpsV = (R)mat[OM_FRICTION_EPS_V];
    m.noGravZone = mat[OM_NO_GRAV_ZONE] != 0;
    m.noGravThreshold = (R)mat[OM_NO_GRAV_THRESHOLD];
    m.maxVel = (R)mat[OM_MAX_VELOCITY];
    m.tdminv.resize(9 * (size_t)nt);
    m.tvol.resize(nt);
    m.mass.resize(nv);
    if (tetcore::tet_prepare<R>(nv, m.x.data(), nt, tets, m.density, m.tdminv.data(), m.tvol.data(), m.mass.data()))
        return -1;
    tetcore::build_incidence(m);
    tet_surface(nt, tets, m.x.data(), m.faces);
    m.nf = (int)m.faces.size() / 3;
    if (tri_topology(m)) return -1;
    m.surface.assign(nv, 0);
    for (int32_t v : m.faces) m.surface[v] = 1;
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
        // tet meshes take each group at most once (VBDPhysics.cpp:235-274), cloth meshes may reuse
        std::vector<int> avail(nc);
        for (size_t c = 0; c < nc; ++c) avail[c] = (int)c;
        for (size_t c = 0; c < nc; ++c) {
            size_t best = s.groups[avail[0]].size();
            int bestId = avail[0];
            size_t bestI = 0;
            for (size_t i = 1; i < avail.size(); ++i)
                if (s.groups[avail[i]].size() < best) {
                    best = s.groups[avail[i]].size();
                    bestId = avail[i];
                    bestI = i;
                }
            if (m.tet) avail.erase(avail.begin() + bestI);
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

template <class R> static int attachments_add(ClothScene<R>& s, int n, const int32_t* ints, const double* vals) {
    const int nm = (int)s.meshes.size();
    auto cloth = [&](int mi) { return mi >= 0 && mi < nm && !s.meshes[mi].tet; };
    for (int i = 0; i < n; ++i) {
        const int32_t* q = ints + 6 * i;
        const double* w = vals + 4 * i;
        const int mv = q[0], v = q[1], tg = q[2];
        if (!cloth(mv) || v < 0 || v >= s.meshes[mv].nv) return -1;
        if (tg >= 0) {
            if (!cloth(tg)) return -1;
            for (int k = 0; k < 3; ++k)
                if (q[3 + k] < 0 || q[3 + k] >= s.meshes[tg].nv) return -1;
        } else if (tg != -1) {
            return -1;
        }
        oracle::Attachment A;
        A.readMesh = mv;
        A.readVert = v;
        A.tgtMesh = tg;
        for (int k = 0; k < 3; ++k) {
            A.f[k] = q[3 + k];
            A.p[k] = (double)(float)w[k];
        }
        A.k = (float)w[3];
        A.b = 1.f;
        auto push = [&](int mi, int vi, const oracle::Attachment& r) {
            ClothMesh<R>& m = s.meshes[mi];
            if (m.att.empty()) m.att.resize(m.nv);
            m.att[vi].push_back(r);
        };
        push(mv, v, A);
        if (tg >= 0)
            for (int k = 0; k < 3; ++k) {
                if ((float)A.p[k] == 0.f) continue;
                oracle::Attachment F = A;
                F.readVert = A.f[k];  // the vertex-side mesh at the face vertex's index
                if (F.readVert >= s.meshes[mv].nv) return -1;
                F.b = -(float)A.p[k];
                push(tg, A.f[k], F);
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

int oracle_cloth_scene_add_tet(OracleClothScene* sc, int nv, const double* pos, const double* vel, int nt,
                               const int32_t* tets, const double* material, int nfixed, const int32_t* fixed,
                               const int32_t* colors) {
    if (sc->precision == 64) return tet_add(sc->s64, nv, pos, vel, nt, tets, material, nfixed, fixed, colors);
    return tet_add(sc->s32, nv, pos, vel, nt, tets, material, nfixed, fixed, colors);
}

int oracle_cloth_scene_surface(OracleClothScene* sc, int mesh, int32_t* faces) {
    auto get = [&](auto& s) -> int {
        if (mesh < 0 || mesh >= (int)s.meshes.size()) return -1;
        if (faces) std::memcpy(faces, s.meshes[mesh].faces.data(), s.meshes[mesh].faces.size() * sizeof(int32_t));
        return s.meshes[mesh].nf;
    };
    return sc->precision == 64 ? get(sc->s64) : get(sc->s32);
}

int oracle_cloth_scene_finalize(OracleClothScene* sc) {
    auto fin = [&](auto& s) -> int {
        cloth_groups(s);
        if (s.p[OP_HANDLE_COLLISION] == 0) return 0;
        auto C = std::make_unique<ContactState>();
        std::vector<int> nv, nf, faces;
        C->vOff.push_back(0);
        C->fOff.push_back(0);
        C->eOff.push_back(0);
        for (size_t im = 0; im < s.meshes.size(); ++im) {
            const auto& m = s.meshes[im];
            nv.push_back(m.nv);
            nf.push_back(m.nf);
            faces.insert(faces.end(), m.faces.begin(), m.faces.end());
            C->vOff.push_back(C->vOff.back() + m.nv);
            C->fOff.push_back(C->fOff.back() + m.nf);
            C->eOff.push_back(C->eOff.back() + m.ne);
            for (int i = 0; i < m.nv; ++i) C->vMesh.push_back((int)im);
        }
        C->cd = oracle_cd_create((int)s.meshes.size(), nv.data(), nf.data(), faces.data());
        if (!C->cd) return -1;
        bool anyTet = false;
        std::vector<uint8_t> mask;
        for (const auto& m : s.meshes) {
            anyTet = anyTet || m.tet;
            for (int i = 0; i < m.nv; ++i) mask.push_back(m.tet ? m.surface[i] : 1);
        }
        if (anyTet) oracle_cd_set_query_mask(C->cd, mask.data());
        s.contact = std::move(C);
        return 0;
    };
    return sc->precision == 64 ? fin(sc->s64) : fin(sc->s32);
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
    sc->s64.threads = sc->s32.threads = std::max(1, nthreads);
    for (int i = 0; i < nsteps; ++i) {
        if (sc->precision == 64) cloth_step(sc->s64, pool);
        else cloth_step(sc->s32, pool);
    }
    return 0;
}

int oracle_cloth_scene_add_deformer(OracleClothScene* sc, int kind, const double* params, int nmeshes,
                                    const int32_t* meshes, const int32_t* counts, const int32_t* verts) {
    if (sc->precision == 64)
        return oracle::deformer_add(sc->s64.deformers, sc->s64.meshes, kind, params, nmeshes, meshes, counts, verts);
    return oracle::deformer_add(sc->s32.deformers, sc->s32.meshes, kind, params, nmeshes, meshes, counts, verts);
}
long long oracle_cloth_scene_solve_failures(OracleClothScene* sc) {
    return sc->precision == 64 ? sc->s64.solveFailures.load() : sc->s32.solveFailures.load();
}
double oracle_cloth_scene_time(OracleClothScene* sc) { return sc->precision == 64 ? sc->s64.clock : sc->s32.clock; }
void oracle_cloth_scene_set_time(OracleClothScene* sc, double t) { sc->s64.clock = sc->s32.clock = (float)t; }

// n attachments: ints (mesh, vertex, target mesh, face v0, v1, v2) and values (bary or point,
// stiffness) per record (loadConstraints, VBDClothPhysics.cpp:148-224).
int oracle_cloth_scene_add_attachments(OracleClothScene* sc, int n, const int32_t* ints, const double* vals) {
    if (sc->precision == 64) return attachments_add(sc->s64, n, ints, vals);
    return attachments_add(sc->s32, n, ints, vals);
}

int oracle_cloth_scene_get_state(OracleClothScene* sc, int mesh, double* pos, double* vel) {
    return sc->precision == 64 ? cloth_state(sc->s64, mesh, pos, vel, true) : cloth_state(sc->s32, mesh, pos, vel, true);
}
int oracle_cloth_scene_set_state(OracleClothScene* sc, int mesh, const double* pos, const double* vel) {
    return sc->precision == 64 ? cloth_state(sc->s64, mesh, (double*)pos, (double*)vel, false)
                               : cloth_state(sc->s32, mesh, (double*)pos, (double*)vel, false);
}

}  // extern "C"
