// CPU oracle: penetration contact of the reference's tet solver (discrete and continuous
// collision detection of surface vertices against tet meshes, and the per-collision force in the
// vertex solve). TEST INFRASTRUCTURE ONLY (see oracle.h). Never linked by the product.
//
// Reference files restated here (paths relative to Simulator/Modules/):
//   VBD/VBDPhysics.cpp:839-862,1958-2014,2066-2102   runStepGPU: dcd() at the start of a substep,
//                                                    initial step, ccd(), collision data
//   VBD/VBDPhysics.cpp:1708-1760                     prepareCollisionData: vertex side and the
//                                                    relations of the closest face's vertices
//   VBD/VBDPhysicsCompute.cu:1731-1923               collision force and Hessian evaluated in the
//                                                    vertex solve from the current positions
//   CollisionDetector/CollisionGeometry.h:56-88      pointInTet (strict, |6V| >= CMP_EPSILON)
//   CollisionDetector/DiscreteCollisionDetector.cpp:13-95,97-300,514-700
//                                                    vertex-in-tet query, closest surface point
//                                                    with feasible regions and tet traversal
//   TetMesh/TetMeshFEM.cpp:503-735, TetMesh/TetMeshFEM.h:423-640
//                                                    ray walk through tets (2D projection, exit
//                                                    faces with an epsilon, candidate stack,
//                                                    recent-tet loop check, pass-through check)
//   CollisionDetector/ContinuousCollisionDetector.cpp:67-190,349-382, CollisionDetector/CCDSolver.h
//                                                    moving point vs moving triangle, earliest hit
//   3rdParty/CyCodeBase/cyPolynomial.h:383-460,800-880 cubic roots in [0, 1]: the interval split at
//                                                    the derivative's roots and a Newton iteration
//                                                    safeguarded by bisection (C. Yuksel, "High-
//                                                    Performance Polynomial Root Finding for
//                                                    Graphics", HPG 2022), tolerance 6e-4
//
// Choices where the reference's result depends on code that is absent or on Embree's traversal
// order (parity unpinned): among several tets containing a vertex the lowest (mesh, tet) is
// taken (maxNumberOfCollisions = 1); equal closest-face distances keep the lower face id; equal
// CCD depths keep the lower (mesh, face); the closest-point query always completes (the
// reference stops after maxNumberOfBVHQuery Embree callbacks, a count that depends on Embree's
// tree); the orthonormal basis of the ray walk is Duff et al. 2017 ("Building an Orthonormal
// Basis, Revisited"; CuMatrix::buildOrthonormalBasis is absent); the candidate stack and the
// recent-tet list hold 64 and 32 entries, a walk that overflows the stack fails (the reference
// then repeats it with an unbounded stack).
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

namespace tetcol {

// CollisionDetertionParameters.h:43-80 defaults; collisionStiffness and collisionOffHeight from
// VBDPhysicsParameters.h:36,70
struct Params {
    bool dcd = true, ccd = true;
    bool selfCollision = true;         // handleSelfCollision
    bool feasibleRegion = true;        // checkFeasibleRegion
    bool tetTraverse = true;           // checkTetTraverse
    bool traverseNonSelf = true;       // tetrahedralTraverseForNonSelfIntersection
    bool shiftToCenter = true;         // shiftQueryPointToCenter
    bool stopAfterPassing = true;      // stopTraversingAfterPassingQueryPoint
    double stiffness = 1e5;            // collisionStiffness
    double offHeight = 1e7;            // collisionOffHeight (vertices at or above skip detection)
    double feasibleEps = 1e-2;         // feasibleRegionEpsilon
    double rayEps = 1e-10;             // rayTriIntersectionEPSILON
    double maxSearchMult = 1.8;        // maxSearchDistanceMultiplier
    double centerShift = 0.01;         // centerShiftLevel
};

constexpr int kStack = 64;   // candidate exit faces of a walk
constexpr int kRecent = 32;  // recently entered tets checked for loops
constexpr float kCmpEpsilon = 0.00001f;

// local face k of a tet is opposite its vertex k; its vertices in this order (TetMeshFEM.cpp:22)
constexpr int kTetFace[4][3] = {{1, 2, 3}, {2, 0, 3}, {0, 1, 3}, {1, 0, 2}};

// Boundary surface of one tet mesh: faces seen by one tet, in ascending (tet, local face) order,
// each turned to face away from its tet's fourth vertex at the given positions (MeshFrame's
// surface face order and vertex rotation are unavailable: parity unpinned for both); the owning tet
// and local face; the neighbouring surface face across edges (v0 v1), (v1 v2), (v2 v0); the
// neighbour tet across every local face; surface vertices and their neighbours over surface edges.
struct Topo {
    int nv = 0, nt = 0;
    std::vector<int32_t> tets;
    std::vector<int32_t> faces, faceTet, faceLocal, faceNbr, tetNbr;
    std::vector<uint8_t> surf;
    std::vector<std::vector<int32_t>> vNbr;
};

template <class R> void build_topo(Topo& T, int nv, int nt, const int32_t* tets, const R* x) {
    T.nv = nv;
    T.nt = nt;
    T.tets.assign(tets, tets + 4 * (size_t)nt);
    struct Key {
        int32_t a, b, c;
        int64_t at;
    };
    std::vector<Key> keys;
    keys.reserve(4 * (size_t)nt);
    for (int t = 0; t < nt; ++t)
        for (int k = 0; k < 4; ++k) {
            int32_t v[3] = {tets[4 * t + kTetFace[k][0]], tets[4 * t + kTetFace[k][1]], tets[4 * t + kTetFace[k][2]]};
            std::sort(v, v + 3);
            keys.push_back({v[0], v[1], v[2], 4 * (int64_t)t + k});
        }
    std::sort(keys.begin(), keys.end(), [](const Key& p, const Key& q) {
        if (p.a != q.a) return p.a < q.a;
        if (p.b != q.b) return p.b < q.b;
        if (p.c != q.c) return p.c < q.c;
        return p.at < q.at;
    });
    T.tetNbr.assign(4 * (size_t)nt, -1);
    std::vector<int64_t> boundary;
    for (size_t i = 0; i < keys.size();) {
        size_t j = i + 1;
        while (j < keys.size() && keys[j].a == keys[i].a && keys[j].b == keys[i].b && keys[j].c == keys[i].c) ++j;
        if (j - i == 1) boundary.push_back(keys[i].at);
        if (j - i == 2) {
            T.tetNbr[keys[i].at] = (int32_t)(keys[i + 1].at / 4);
            T.tetNbr[keys[i + 1].at] = (int32_t)(keys[i].at / 4);
        }
        i = j;
    }
    std::sort(boundary.begin(), boundary.end());
    T.faces.clear();
    T.faceTet.clear();
    T.faceLocal.clear();
    // a surface face lists its tet's other three vertices in ascending local order, then swaps
    // the last two if needed to face outwards (the boundary surface of the product's contact scenes)
    static const int kAsc[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
    for (int64_t at : boundary) {
        const int t = (int)(at / 4), k = (int)(at % 4);
        int32_t a = tets[4 * t + kAsc[k][0]], b = tets[4 * t + kAsc[k][1]], c = tets[4 * t + kAsc[k][2]];
        const int32_t o = tets[4 * t + k];
        R e1[3], e2[3], w[3];
        for (int d = 0; d < 3; ++d) {
            e1[d] = x[3 * b + d] - x[3 * a + d];
            e2[d] = x[3 * c + d] - x[3 * a + d];
            w[d] = x[3 * o + d] - x[3 * a + d];
        }
        const R n0 = e1[1] * e2[2] - e1[2] * e2[1], n1 = e1[2] * e2[0] - e1[0] * e2[2],
                n2 = e1[0] * e2[1] - e1[1] * e2[0];
        if (n0 * w[0] + n1 * w[1] + n2 * w[2] > 0) std::swap(b, c);
        T.faces.insert(T.faces.end(), {a, b, c});
        T.faceTet.push_back(t);
        T.faceLocal.push_back(k);
    }
    const int nf = (int)T.faceTet.size();
    // neighbour faces over the directed edges: (u, v) of one face meets (v, u) of another
    std::vector<std::pair<int64_t, int>> dir;
    dir.reserve(3 * (size_t)nf);
    for (int f = 0; f < nf; ++f)
        for (int e = 0; e < 3; ++e) {
            const int64_t u = T.faces[3 * f + e], v = T.faces[3 * f + (e + 1) % 3];
            dir.push_back({u * nv + v, 3 * f + e});
        }
    std::sort(dir.begin(), dir.end());
    T.faceNbr.assign(3 * (size_t)nf, -1);
    for (int f = 0; f < nf; ++f)
        for (int e = 0; e < 3; ++e) {
            const int64_t u = T.faces[3 * f + e], v = T.faces[3 * f + (e + 1) % 3];
            auto it = std::lower_bound(dir.begin(), dir.end(), std::make_pair(v * nv + u, -1));
            if (it != dir.end() && it->first == v * nv + u) T.faceNbr[3 * f + e] = it->second / 3;
        }
    T.surf.assign(nv, 0);
    T.vNbr.assign(nv, {});
    for (int f = 0; f < nf; ++f)
        for (int e = 0; e < 3; ++e) {
            const int u = T.faces[3 * f + e], v = T.faces[3 * f + (e + 1) % 3];
            T.surf[u] = 1;
            T.vNbr[u].push_back(v);
            T.vNbr[v].push_back(u);
        }
    for (auto& l : T.vNbr) {
        std::sort(l.begin(), l.end());
        l.erase(std::unique(l.begin(), l.end()), l.end());
    }
}

template <class R> struct P3 {
    R x, y, z;
};
template <class R> inline P3<R> ld(const R* X, int i) { return {X[3 * i], X[3 * i + 1], X[3 * i + 2]}; }
template <class R> inline P3<R> operator-(P3<R> a, P3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> inline P3<R> operator+(P3<R> a, P3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> inline P3<R> operator*(R s, P3<R> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class R> inline R dot(P3<R> a, P3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> inline P3<R> cross(P3<R> a, P3<R> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class R> inline R norm(P3<R> a) { return std::sqrt(dot(a, a)); }
template <class R> inline P3<R> normalize(P3<R> a) { return (R(1) / norm(a)) * a; }

// CollisionGeometry.h:56-88: strictly inside, and |AB . (AC x AD)| >= CMP_EPSILON
template <class R> bool point_in_tet(P3<R> p, P3<R> a, P3<R> b, P3<R> c, P3<R> d) {
    const R vol = dot(b - a, cross(c - a, d - a));
    if (std::fabs(vol) < (R)kCmpEpsilon) return false;
    const P3<R> vs[4] = {a, b, c, d};
    for (int i = 0; i < 4; ++i) {
        const P3<R> v1 = vs[kTetFace[i][0]], v2 = vs[kTetFace[i][1]], v3 = vs[kTetFace[i][2]];
        if (dot(p - v1, cross(v2 - v1, v3 - v2)) * vol >= 0) return false;
    }
    return true;
}

// Ericson's closest point on a triangle (CollisionGeometry.h:90-157) with its region:
// 0 vertex A, 1 B, 2 C, 3 edge AB, 4 edge AC, 5 edge BC, 6 interior
template <class R> int closest_point(P3<R> p, P3<R> a, P3<R> b, P3<R> c, P3<R>& cp, R bary[3]) {
    const P3<R> ab = b - a, ac = c - a, ap = p - a;
    const R d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0 && d2 <= 0) {
        cp = a, bary[0] = 1, bary[1] = 0, bary[2] = 0;
        return 0;
    }
    const P3<R> bp = p - b;
    const R d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0 && d4 <= d3) {
        cp = b, bary[0] = 0, bary[1] = 1, bary[2] = 0;
        return 1;
    }
    const P3<R> cpv = p - c;
    const R d5 = dot(ab, cpv), d6 = dot(ac, cpv);
    if (d6 >= 0 && d5 <= d6) {
        cp = c, bary[0] = 0, bary[1] = 0, bary[2] = 1;
        return 2;
    }
    const R vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) {
        const R v = d1 / (d1 - d3);
        cp = a + v * ab, bary[0] = 1 - v, bary[1] = v, bary[2] = 0;
        return 3;
    }
    const R vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) {
        const R v = d2 / (d2 - d6);
        cp = a + v * ac, bary[0] = 1 - v, bary[1] = 0, bary[2] = v;
        return 4;
    }
    const R va = d3 * d6 - d5 * d4;
    if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
        const R v = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        cp = b + v * (c - b), bary[0] = 0, bary[1] = 1 - v, bary[2] = v;
        return 5;
    }
    const R denom = R(1) / (va + vb + vc);
    const R v = vb * denom, w = vc * denom;
    cp = (a + v * ab) + w * ac, bary[0] = 1 - v - w, bary[1] = v, bary[2] = w;
    return 6;
}

// DiscreteCollisionDetector.cpp:697-780 (edge 0 = v0 v1, 1 = v1 v2, 2 = v2 v0) and :782-815
template <class R> bool feasible(const Topo& T, const R* X, P3<R> p, int face, int region, R eps) {
    const int32_t* fv = &T.faces[3 * face];
    auto fnormal = [&](int f) {
        const P3<R> a = ld(X, T.faces[3 * f]), b = ld(X, T.faces[3 * f + 1]), c = ld(X, T.faces[3 * f + 2]);
        return normalize(cross(b - a, c - a));
    };
    if (region == 6) return true;
    if (region >= 3) {
        const int e = region == 3 ? 0 : (region == 5 ? 1 : 2);
        const int i1 = fv[e], i2 = fv[(e + 1) % 3];
        const int nb = T.faceNbr[3 * face + e];
        if (nb < 0) return true;  // open boundary (the reference asserts a closed surface)
        const P3<R> v1 = ld(X, i1), v2 = ld(X, i2);
        const P3<R> n1 = (R)-1 * fnormal(face), n2 = (R)-1 * fnormal(nb);
        const P3<R> AP = p - v1, BP = p - v2, AB = v2 - v1, BA = (R)-1 * AB;
        const R relaxed = -dot(AB, AB) * eps;
        if (dot(AP, AB) < relaxed) return false;
        if (dot(BP, BA) < relaxed) return false;
        if (dot(AP, cross(n1, AB)) < relaxed) return false;
        if (dot(AP, cross(n2, BA)) < relaxed) return false;
        return true;
    }
    const int vid = fv[region];
    const P3<R> A = ld(X, vid), AP = p - A;
    for (int nb : T.vNbr[vid]) {
        const P3<R> BA = A - ld(X, nb);
        if (dot(AP, BA) < -dot(BA, BA) * eps) return false;
    }
    return true;
}

// Duff et al. 2017: a right-handed orthonormal basis (b1, b2) of the plane normal to unit n
template <class R> void ortho_basis(P3<R> n, P3<R>& b1, P3<R>& b2) {
    const R s = std::copysign(R(1), n.z);
    const R a = R(-1) / (s + n.z);
    const R b = n.x * n.y * a;
    b1 = {R(1) + s * n.x * n.x * a, s * b, -s * n.x};
    b2 = {b, s + n.y * n.y * a, -n.y};
}

template <class R> inline R cross2(R ax, R ay, R bx, R by) { return ax * by - ay * bx; }

// The ray walk of TetMeshFEM::tetrahedralTraverseTo (TetMeshFEM.cpp:503-735): from the surface
// face (startTet, startLocal) along the ray o + s dir, through face-adjacent tets whose exit faces
// the ray may cross, until the target tet is entered (true), the surface is left, the crossing
// lies beyond maxDis (checked every 5 walked tets, when maxDis > 0) or no candidate is left.
template <class R> bool walk(const Topo& T, const R* X, P3<R> o, P3<R> dir, R maxDis, int startTet, int startLocal,
                             int target, R eps) {
    if (startTet == target) return true;
    P3<R> b1, b2;
    ortho_basis(dir, b1, b2);
    struct Pt {
        R x, y;
    };
    // projected tet vertices: incoming face's three in kTetFace order, then the opposite one
    auto project = [&](int t, int in, Pt q[4], int ids[4]) {
        for (int i = 0; i < 3; ++i) ids[i] = T.tets[4 * t + kTetFace[in][i]];
        ids[3] = T.tets[4 * t + in];
        for (int i = 0; i < 4; ++i) {
            const P3<R> d = ld(X, ids[i]) - o;
            q[i] = {dot(b1, d), dot(b2, d)};
        }
    };
    // TetMeshFEM.h:423-474: exit face i (opposite incoming vertex i) may be crossed
    auto exits = [&](const Pt q[4], int out[3]) {
        const R area = cross2(q[1].x - q[0].x, q[1].y - q[0].y, q[2].x - q[0].x, q[2].y - q[0].y);
        const R s = std::copysign(R(1), area);
        const R d0 = cross2(q[3].x, q[3].y, q[0].x, q[0].y) * s;
        const R d1 = cross2(q[3].x, q[3].y, q[1].x, q[1].y) * s;
        const R d2 = cross2(q[3].x, q[3].y, q[2].x, q[2].y) * s;
        out[0] = d1 >= -eps && d2 <= eps;
        out[1] = d2 >= -eps && d0 <= eps;
        out[2] = d0 >= -eps && d1 <= eps;
    };
    int stackTet[kStack], stackFace[kStack], ns = 0;
    int recent[kRecent], nrec = 0, head = 0;
    auto remember = [&](int t) {
        recent[head] = t;
        head = (head + 1) % kRecent;
        nrec = std::min(nrec + 1, kRecent);
    };
    auto seen = [&](int t) {
        for (int i = 0; i < nrec; ++i)
            if (recent[i] == t) return true;
        return false;
    };
    Pt q[4];
    int ids[4], ex[3];
    project(startTet, startLocal, q, ids);
    exits(q, ex);
    remember(startTet);
    for (int i = 0; i < 3; ++i)
        if (ex[i]) {
            stackTet[ns] = startTet;
            stackFace[ns++] = kTetFace[startLocal][i];
        }
    int steps = 0;
    while (ns > 0) {
        --ns;
        const int prev = stackTet[ns], face = stackFace[ns];
        const int cur = T.tetNbr[4 * prev + face];
        if (cur == target) return true;
        if (cur < 0) return false;  // reached the surface
        if (seen(cur)) continue;
        remember(cur);
        // the incoming face of cur is the one shared with prev: opposite cur's vertex not on it
        int fv[3];
        for (int i = 0; i < 3; ++i) fv[i] = T.tets[4 * prev + kTetFace[face][i]];
        int in = 3;
        for (int i = 0; i < 4; ++i) {
            const int v = T.tets[4 * cur + i];
            if (v != fv[0] && v != fv[1] && v != fv[2]) {
                in = i;
                break;
            }
        }
        project(cur, in, q, ids);
        exits(q, ex);
        for (int i = 0; i < 3; ++i) {
            if (!ex[i]) continue;
            const int exitLocal = kTetFace[in][i];
            if (T.tetNbr[4 * cur + exitLocal] == target) return true;
            if (ns == kStack) return false;  // candidate stack overflow
            stackTet[ns] = cur;
            stackFace[ns++] = exitLocal;
            if (steps % 5 == 0 && maxDis > 0) {
                // crossing of the ray with exit face i: the origin's 2D barycentrics in its projection
                static const int kExit[3][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}};
                const Pt a = q[kExit[i][0]], b = q[kExit[i][1]], c = q[kExit[i][2]];
                const R v0x = b.x - a.x, v0y = b.y - a.y, v1x = c.x - a.x, v1y = c.y - a.y, v2x = -a.x, v2y = -a.y;
                const R den = v0x * v1y - v1x * v0y;
                const R l1 = (v2x * v1y - v1x * v2y) / den, l2 = (v0x * v2y - v2x * v0y) / den, l0 = 1 - l1 - l2;
                const P3<R> hit = l0 * ld(X, ids[kExit[i][0]]) + l1 * ld(X, ids[kExit[i][1]]) +
                                  l2 * ld(X, ids[kExit[i][2]]);
                const P3<R> d = hit - o;
                if (dot(d, d) > maxDis * maxDis) return false;
            }
        }
        ++steps;
    }
    return false;
}

// The result of one surface vertex's detection: the mesh and surface face it is pushed out of
// (face -1: none; the reference's shortestPathFound)
struct Hit {
    int mesh = -1, face = -1;
};

// Axis-aligned box used only to skip meshes and faces that cannot hold a hit (a speed-up of
// the brute-force queries that leaves their results unchanged: the padding covers rounding in
// the containment test and the root tolerance of the CCD).
template <class R> struct Box3 {
    R lo[3] = {R(1e30), R(1e30), R(1e30)}, hi[3] = {R(-1e30), R(-1e30), R(-1e30)};
    void add(P3<R> p) {
        const R c[3] = {p.x, p.y, p.z};
        for (int k = 0; k < 3; ++k) lo[k] = std::min(lo[k], c[k]), hi[k] = std::max(hi[k], c[k]);
    }
    void pad(R d) {
        for (int k = 0; k < 3; ++k) lo[k] -= d, hi[k] += d;
    }
    bool overlaps(const Box3& o) const {
        for (int k = 0; k < 3; ++k)
            if (o.hi[k] < lo[k] || hi[k] < o.lo[k]) return false;
        return true;
    }
    R extent() const { return std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2])); }
};

// DCD of surface vertex v of mesh qm (DiscreteCollisionDetector.cpp:514-560 with the callbacks
// :13-95 and :97-300): the lowest containing tet of any mesh (of its own mesh only with
// selfCollision, and no tet of its own), then the closest face of that mesh's surface that passes
// the self-vertex, feasible-region and walk checks. `penetrated` = a containing tet was found.
// `meshBoxes` (optional): per mesh the box of its positions padded by 1e-3 of its extent; a mesh
// whose box does not hold the vertex holds no tet that contains it.
template <class R> Hit dcd_vertex(const std::vector<Topo>& topo, const std::vector<const R*>& X, const Params& P,
                                  int qm, int v, bool& penetrated, const Box3<R>* meshBoxes = nullptr) {
    Hit h;
    penetrated = false;
    const P3<R> p = ld(X[qm], v);
    Box3<R> pb;
    pb.add(p);
    int hm = -1, ht = -1;
    for (int m = 0; m < (int)topo.size() && hm < 0; ++m) {
        if (m == qm && !P.selfCollision) continue;
        if (meshBoxes && !meshBoxes[m].overlaps(pb)) continue;
        const Topo& T = topo[m];
        for (int t = 0; t < T.nt; ++t) {
            const int32_t* tv = &T.tets[4 * t];
            if (m == qm && (tv[0] == v || tv[1] == v || tv[2] == v || tv[3] == v)) continue;
            if (point_in_tet(p, ld(X[m], tv[0]), ld(X[m], tv[1]), ld(X[m], tv[2]), ld(X[m], tv[3]))) {
                hm = m;
                ht = t;
                break;
            }
        }
    }
    if (hm < 0) return h;
    penetrated = true;
    const Topo& T = topo[hm];
    const R* Y = X[hm];
    const int nf = (int)T.faceTet.size();
    struct Cand {
        R d;
        int f, region;
        P3<R> cp;
    };
    std::vector<Cand> cands;
    cands.reserve(nf);
    for (int f = 0; f < nf; ++f) {
        const int32_t* fv = &T.faces[3 * f];
        P3<R> cp;
        R bary[3];
        const int region = closest_point(p, ld(Y, fv[0]), ld(Y, fv[1]), ld(Y, fv[2]), cp, bary);
        if (hm == qm && region <= 2 && fv[region] == v) continue;  // its own vertex is the closest point
        cands.push_back({norm(p - cp), f, region, cp});
    }
    std::sort(cands.begin(), cands.end(), [](const Cand& a, const Cand& b) { return a.d < b.d || (a.d == b.d && a.f < b.f); });
    const R c = (R)P.centerShift;
    for (const Cand& k : cands) {
        if (P.feasibleRegion && !feasible(T, Y, p, k.f, k.region, (R)P.feasibleEps)) continue;
        if (P.tetTraverse && (hm == qm || P.traverseNonSelf)) {
            const int32_t* fv = &T.faces[3 * k.f];
            const P3<R> a = ld(Y, fv[0]), b = ld(Y, fv[1]), cc = ld(Y, fv[2]);
            const P3<R> from = k.region == 6 ? k.cp : (R(1) - c) * k.cp + (c / R(3)) * ((a + b) + cc);
            P3<R> to = p;
            if (P.shiftToCenter) {
                const int32_t* tv = &T.tets[4 * ht];
                const P3<R> cen = R(0.25) * (((ld(Y, tv[0]) + ld(Y, tv[1])) + ld(Y, tv[2])) + ld(Y, tv[3]));
                to = (R(1) - c) * p + c * cen;
            }
            P3<R> dir = to - from;
            const R len = norm(dir);
            dir = (R(1) / len) * dir;
            const R maxDis = P.stopAfterPassing ? (R)P.maxSearchMult * len : R(-1);
            if (!walk(T, Y, from, dir, maxDis, T.faceTet[k.f], T.faceLocal[k.f], ht, (R)P.rayEps)) continue;
        }
        h.mesh = hm;
        h.face = k.f;
        return h;
    }
    return h;
}

// ---- continuous detection --------------------------------------------------------------------

template <class R> inline R poly3(const R c[4], R x) { return ((c[3] * x + c[2]) * x + c[1]) * x + c[0]; }
template <class R> inline bool diff_sign(R a, R b) { return (a < 0) != (b < 0); }

// root of the cubic c in [x0, x1] where its values at the ends differ in sign (y0 at x0), to
// within err (cyPolynomial.h:383-460): Newton steps from the midpoint clamped to the interval;
// if 16 of them do not settle, Newton steps kept inside a shrinking bracket, bisecting whenever
// a step would leave it
template <class R> R bracketed_root(const R c[4], R x0, R x1, R y0, R err) {
    auto dpoly = [&](R x) { return (R(3) * c[3] * x + R(2) * c[2]) * x + c[1]; };
    const R ep2 = 2 * err;
    R xr = (x0 + x1) / 2;
    if (x1 - x0 <= ep2) return xr;
    const R xr0 = xr;
    for (int it = 0; it < 16; ++it) {
        R xn = xr - poly3(c, xr) / dpoly(xr);
        xn = std::min(std::max(xn, x0), x1);
        if (std::fabs(xr - xn) <= err) return xn;
        xr = xn;
    }
    if (!std::isfinite(xr)) xr = xr0;
    R yr = poly3(c, xr);
    R xb0 = x0, xb1 = x1;
    for (int guard = 0; guard < 200; ++guard) {
        if (diff_sign(y0, yr)) xb1 = xr;
        else xb0 = xr;
        const R xn = xr - yr / dpoly(xr);
        if (xn > xb0 && xn < xb1) {
            const R step = std::fabs(xr - xn);
            xr = xn;
            if (step > err) yr = poly3(c, xr);
            else break;
        } else {
            xr = (xb0 + xb1) / 2;
            if (xr == xb0 || xr == xb1 || xb1 - xb0 <= ep2) break;
            yr = poly3(c, xr);
        }
    }
    return xr;
}

// real roots of c in [0, 1] where the cubic changes sign, ascending (cyPolynomial.h:800-880)
template <class R> int cubic_roots01(const R c[4], R roots[3], R err) {
    R pts[4];
    int np = 0;
    pts[np++] = 0;
    const R a = 3 * c[3], b2 = c[2], cc = c[1];
    const R disc = b2 * b2 - a * cc;
    if (a != 0 && disc > 0) {
        const R sq = std::sqrt(disc);
        const R q = -(b2 + std::copysign(sq, b2));
        R r0 = q / a, r1 = cc / q;
        if (r0 > r1) std::swap(r0, r1);
        if (r0 > 0 && r0 < 1) pts[np++] = r0;
        if (r1 > 0 && r1 < 1 && r1 != r0) pts[np++] = r1;
    } else if (a == 0 && b2 != 0) {
        const R r = -cc / (2 * b2);
        if (r > 0 && r < 1) pts[np++] = r;
    }
    pts[np++] = 1;
    int n = 0;
    R yprev = poly3(c, pts[0]);
    for (int i = 1; i < np; ++i) {
        const R y = poly3(c, pts[i]);
        if (diff_sign(yprev, y)) roots[n++] = bracketed_root(c, pts[i - 1], pts[i], yprev, err);
        yprev = y;
    }
    return n;
}

// CCDSolver.h:104-202: the earliest t in [0, 1] at which the moving point lies in the moving
// triangle's plane and inside it (dominant-axis 2D sign test); barycentrics at that t
template <class R> bool ccd_point_triangle(const P3<R> f0[3], const P3<R> f1[3], P3<R> p0, P3<R> p1, R& tt,
                                           R bary[3]) {
    const P3<R> x0[4] = {f0[0], f0[1], f0[2], p0};
    const P3<R> x1[4] = {f1[0], f1[1], f1[2], p1};
    const P3<R> dx0 = x1[0] - x0[0], dx1 = x1[1] - x0[1], dx2 = x1[2] - x0[2], dx3 = x1[3] - x0[3];
    const P3<R> e1 = x0[1] - x0[0], e2 = x0[2] - x0[0], e3 = x0[3] - x0[0];
    const P3<R> de0 = dx3 - dx0, de1 = dx1 - dx0, de2 = dx2 - dx0;
    const P3<R> e1xe2 = cross(e1, e2), de1xde2 = cross(de1, de2), mixed = cross(e1, de2) + cross(de1, e2);
    const R c[4] = {dot(e3, e1xe2), dot(de0, e1xe2) + dot(e3, mixed), dot(e3, de1xde2) + dot(de0, mixed),
                    dot(de0, de1xde2)};
    R ts[3];
    const int n = cubic_roots01(c, ts, (R)6e-4);
    for (int i = 0; i < n; ++i) {
        const R t = ts[i];
        P3<R> y3[4];
        for (int k = 0; k < 4; ++k) y3[k] = x0[k] + t * (x1[k] - x0[k]);
        const P3<R> nrm = cross(y3[1] - y3[0], y3[2] - y3[0]);
        const R ax = std::fabs(nrm.x), ay = std::fabs(nrm.y), az = std::fabs(nrm.z);
        const int drop = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
        R u[4], w[4];
        for (int k = 0; k < 4; ++k) {
            u[k] = drop == 0 ? y3[k].y : y3[k].x;
            w[k] = drop == 2 ? y3[k].y : y3[k].z;
        }
        for (int k = 0; k < 3; ++k) u[k] -= u[3], w[k] -= w[3];
        const bool s01 = std::signbit(cross2(u[0], w[0], u[1], w[1]));
        const bool s12 = std::signbit(cross2(u[1], w[1], u[2], w[2]));
        const bool s20 = std::signbit(cross2(u[2], w[2], u[0], w[0]));
        if (s01 == s12 && s01 == s20) {
            tt = t;
            const P3<R> q = p0 + t * (p1 - p0);
            const P3<R> a = f0[0] + t * (f1[0] - f0[0]), b = f0[1] + t * (f1[1] - f0[1]),
                        cc = f0[2] + t * (f1[2] - f0[2]);
            const P3<R> v0 = b - a, v1 = cc - a, v2 = q - a;
            const R d00 = dot(v0, v0), d01 = dot(v0, v1), d11 = dot(v1, v1), d20 = dot(v2, v0), d21 = dot(v2, v1);
            const R den = d00 * d11 - d01 * d01;
            bary[1] = (d11 * d20 - d01 * d21) / den;
            bary[2] = (d00 * d21 - d01 * d20) / den;
            bary[0] = R(1) - bary[1] - bary[2];
            return true;
        }
    }
    return false;
}

// CCD of surface vertex v of mesh qm (ContinuousCollisionDetector.cpp:67-190, 349-382): the face
// (of any mesh; of its own mesh only faces without v) that the vertex's path from Xp to X meets
// first, i.e. with the smallest t |p1 - p0|
// `sweptBoxes` (optional): per mesh the box of its start and end positions; `maxMove`: per mesh the
// largest vertex displacement. A face whose swept box, padded by the point's and the face's
// largest displacement (far above the error of a root within 6e-4 in t), misses the point's path
// cannot be hit.
template <class R> Hit ccd_vertex(const std::vector<Topo>& topo, const std::vector<const R*>& X,
                                  const std::vector<const R*>& Xp, int qm, int v,
                                  const Box3<R>* sweptBoxes = nullptr, const R* maxMove = nullptr) {
    Hit h;
    const P3<R> p0 = ld(Xp[qm], v), p1 = ld(X[qm], v);
    const R len = norm(p1 - p0);
    Box3<R> path;
    path.add(p0);
    path.add(p1);
    R best = 0;
    for (int m = 0; m < (int)topo.size(); ++m) {
        const Topo& T = topo[m];
        const R margin = sweptBoxes ? (len + maxMove[m]) * R(1.01) + R(1e-6) * (R(1) + sweptBoxes[m].extent()) : R(0);
        if (sweptBoxes) {
            Box3<R> mb = sweptBoxes[m];
            mb.pad(margin);
            if (!mb.overlaps(path)) continue;
        }
        for (int f = 0; f < (int)T.faceTet.size(); ++f) {
            const int32_t* fv = &T.faces[3 * f];
            if (m == qm && (fv[0] == v || fv[1] == v || fv[2] == v)) continue;
            const P3<R> a0[3] = {ld(Xp[m], fv[0]), ld(Xp[m], fv[1]), ld(Xp[m], fv[2])};
            const P3<R> a1[3] = {ld(X[m], fv[0]), ld(X[m], fv[1]), ld(X[m], fv[2])};
            if (sweptBoxes) {
                Box3<R> fb;
                for (int k = 0; k < 3; ++k) fb.add(a0[k]), fb.add(a1[k]);
                fb.pad(margin);
                if (!fb.overlaps(path)) continue;
            }
            R tt, bary[3];
            if (!ccd_point_triangle(a0, a1, p0, p1, tt, bary)) continue;
            const R depth = tt * len;
            if (h.face < 0 || depth < best) {
                best = depth;
                h.mesh = m;
                h.face = f;
            }
        }
    }
    return h;
}

// One penetration contact evaluated in the vertex solve (VBDPhysicsCompute.cu:1831-1923): p is the
// colliding vertex, (a, b, c) its face, with their positions at the start of the step (pp, ap, bp,
// cp); `order` 3 for the vertex side (weight 1), 0..2 for a face vertex (weight -bary[order]).
// Adds b f and b^2 H of the force k d n + friction and the Hessian k n n^T + friction.
template <class R> void penetration_force(P3<R> p, P3<R> a, P3<R> b, P3<R> c, P3<R> pp, P3<R> ap, P3<R> bp, P3<R> cpp,
                                          int order, R k, R mu, R epsV, R dt, R* f, R* H) {
    P3<R> cp;
    R bary[3];
    const int region = closest_point(p, a, b, c, cp, bary);
    const P3<R> fn = normalize(cross(b - a, c - a));
    const P3<R> diff = cp - p;
    P3<R> n = fn;
    if (region != 6) {
        n = (R(1) / norm(diff)) * diff;
        if (dot(n, fn) < 0) n = (R)-1 * n;
    }
    const R depth = dot(diff, n);
    if (!(depth > 0)) return;
    const R w = order == 3 ? R(1) : -bary[order];
    const R lambda = depth * k;
    R cf[3] = {lambda * n.x, lambda * n.y, lambda * n.z};
    const R nn[3] = {n.x, n.y, n.z};
    R ch[9];
    for (int cI = 0; cI < 3; ++cI)
        for (int r = 0; r < 3; ++r) ch[r + 3 * cI] = k * nn[r] * nn[cI];
    const P3<R> dx = (p - pp) - ((bary[0] * (a - ap) + bary[1] * (b - bp)) + bary[2] * (c - cpp));
    const P3<R> e0 = normalize(b - a);
    const P3<R> e1 = normalize(cross(e0, n));
    const R u0 = dot(e0, dx), u1 = dot(e1, dx);
    const R uNorm = std::sqrt(u0 * u0 + u1 * u1);
    if (uNorm > 0) {
        const R epsU = epsV * dt;
        const R f1 = uNorm > epsU ? R(1) / uNorm : (-uNorm / epsU + R(2)) / epsU;
        const R T[6] = {e0.x, e0.y, e0.z, e1.x, e1.y, e1.z};
        for (int r = 0; r < 3; ++r) cf[r] += -mu * lambda * (T[r] * u0 + T[r + 3] * u1) * f1;
        for (int cI = 0; cI < 3; ++cI)
            for (int r = 0; r < 3; ++r) ch[r + 3 * cI] += mu * lambda * f1 * (T[r] * T[cI] + T[r + 3] * T[cI + 3]);
    }
    for (int r = 0; r < 3; ++r) f[r] += cf[r] * w;
    for (int i = 0; i < 9; ++i) H[i] += ch[i] * w * w;
}

}  // namespace tetcol
