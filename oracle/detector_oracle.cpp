// Oracle (test infrastructure only): CPU restatement of the cloth contact detector's
// discrete proximity queries. Paths relative to Simulator/Modules/:
//   VF radius query + topology filter + face min distance   CollisionDetector/ClothContactDetector.cpp:144-288
//   EE query (segment closest points, feasibility)           CollisionDetector/ClothContactDetector.cpp:731-813
//   query setup (radius, centre, min-distance init)          CollisionDetector/ClothContactDetector.cpp:1052-1135
//   closest point on triangle / between segments            CollisionDetector/CollisionGeometry.h:90-267
//   feasible regions                                         CollisionDetector/TriMeshCollisionGeometry.cpp:7-186
//   closest-point primitive type                             CollisionDetector/CollisionDetertionParameters.h:224-243
//   conservative bounds                                      VBDCloth/VBDClothPhysics.cpp:775-819
//
// Every quantity the reference derives from its Embree traversal is restated in a
// traversal-order-free form:
//   * VF candidates = every (vertex, face) with the face not incident to the vertex (same mesh)
//     and d(x_v, closest point) < r; the per-vertex and per-face minima are taken over them;
//   * VF contacts = one per (closest-primitive type, primitive id) key whose feasible-region test
//     passes (the test depends on the key only, so the key set does not depend on the order);
//     the representative face is the lowest candidate face id with that key;
//   * EE: the per-edge minimum over every non-adjacent edge (a candidate edge closer than r
//     always overlaps the query sphere, so Embree visits it) and every feasible pair with d < r.
// Arithmetic follows Embree's Vec3fa (dot = (x + y) + z, no fused multiply-add; the library is
// built with -ffp-contract=off); sqrt and division are IEEE.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <vector>

#include "oracle.h"
#include "pool.h"

namespace {

struct V3 {
    float x, y, z;
};
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator*(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline float distance(V3 a, V3 b) { return std::sqrt(dot(a - b, a - b)); }

constexpr float kCmpEpsilon = 0.00001f;                 // Types.h:20
constexpr float kCmpEpsilon2 = kCmpEpsilon * kCmpEpsilon;
constexpr float kAbsoluteRelaxation = 1e-6f;            // TriMeshCollisionGeometry.cpp:4
constexpr float kFeasibleEpsilon = 1e-3f;               // ClothContactDetector.cpp:226

enum PointType { AT_A = 0, AT_B, AT_C, AT_AB, AT_BC, AT_AC, AT_INTERIOR };
enum PrimType { PRIM_VERTEX = 0, PRIM_EDGE = 1, PRIM_FACE = 2 };

V3 closest_point_triangle(V3 p, V3 a, V3 b, V3 c, V3& bary, int& type) {
    const V3 ab = b - a, ac = c - a, ap = p - a;
    const float d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.f && d2 <= 0.f) {
        type = AT_A;
        bary = {1.f, 0.f, 0.f};
        return a;
    }
    const V3 bp = p - b;
    const float d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0.f && d4 <= d3) {
        type = AT_B;
        bary = {0.f, 1.f, 0.f};
        return b;
    }
    const V3 cp = p - c;
    const float d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0.f && d5 <= d6) {
        type = AT_C;
        bary = {0.f, 0.f, 1.f};
        return c;
    }
    const float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.f && d1 >= 0.f && d3 <= 0.f) {
        const float v = d1 / (d1 - d3);
        type = AT_AB;
        bary = {1.0f - v, v, 0.f};
        return a + v * ab;
    }
    const float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.f && d2 >= 0.f && d6 <= 0.f) {
        const float v = d2 / (d2 - d6);
        type = AT_AC;
        bary = {1.0f - v, 0.f, v};
        return a + v * ac;
    }
    const float va = d3 * d6 - d5 * d4;
    if (va <= 0.f && (d4 - d3) >= 0.f && (d5 - d6) >= 0.f) {
        type = AT_BC;
        const float v = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        bary = {0.f, 1.f - v, v};
        return b + v * (c - b);
    }
    const float denom = 1.f / (va + vb + vc);
    const float v = vb * denom, w = vc * denom;
    type = AT_INTERIOR;
    bary = {1.f - v - w, v, w};
    return (a + v * ab) + w * ac;
}

void segment_closest_points(V3 p0, V3 p1, V3 q0, V3 q1, V3& cp, V3& cq, float& s, float& t) {
    const V3 p = p1 - p0, q = q1 - q0, r = p0 - q0;
    const float a = dot(p, p), b = dot(p, q), c = dot(q, q), d = dot(p, r), e = dot(q, r);
    s = 0.0f;
    t = 0.0f;
    const float det = a * c - b * b;
    if (det > kCmpEpsilon) {
        const float bte = b * e, ctd = c * d;
        if (bte <= ctd) {
            if (e <= 0.0f) {
                s = (-d >= a ? 1 : (-d > 0.0f ? -d / a : 0.0f));
                t = 0.0f;
            } else if (e < c) {
                s = 0.0f;
                t = e / c;
            } else {
                s = (b - d >= a ? 1 : (b - d > 0.0f ? (b - d) / a : 0.0f));
                t = 1;
            }
        } else {
            s = bte - ctd;
            if (s >= det) {
                if (b + e <= 0.0f) {
                    s = (-d <= 0.0f ? 0.0f : (-d < a ? -d / a : 1));
                    t = 0.0f;
                } else if (b + e < c) {
                    s = 1;
                    t = (b + e) / c;
                } else {
                    s = (b - d <= 0.0f ? 0.0f : (b - d < a ? (b - d) / a : 1));
                    t = 1;
                }
            } else {
                const float ate = a * e, btd = b * d;
                if (ate <= btd) {
                    s = (-d <= 0.0f ? 0.0f : (-d >= a ? 1 : -d / a));
                    t = 0.0f;
                } else {
                    t = ate - btd;
                    if (t >= det) {
                        s = (b - d <= 0.0f ? 0.0f : (b - d >= a ? 1 : (b - d) / a));
                        t = 1;
                    } else {
                        s /= det;
                        t /= det;
                    }
                }
            }
        }
    } else {
        if (e <= 0.0f) {
            s = (-d <= 0.0f ? 0.0f : (-d >= a ? 1 : -d / a));
            t = 0.0f;
        } else if (e >= c) {
            s = (b - d <= 0.0f ? 0.0f : (b - d >= a ? 1 : (b - d) / a));
            t = 1;
        } else {
            s = 0.0f;
            t = e / c;
        }
    }
    cp = (1 - s) * p0 + s * p1;
    cq = (1 - t) * q0 + t * q1;
}

struct VFPair {
    int v, f;
    float d;
};
struct VFContact {
    int v, type, prim, face;
    float d;
};
struct EEContact {
    int e1, e2;
    float mu1, mu2, d;
};

struct Detector {
    int nmesh = 0, nv = 0, nf = 0, ne = 0;
    std::vector<int> vMesh, fMesh, eMesh;
    std::vector<int> faces;      // 3 * nf global ids
    std::vector<int> faceEdges;  // 3 * nf: edges AB, BC, CA
    std::vector<int> edges;      // 4 * ne: eV1 eV2 eV12Next eV21Next (global ids, -1 none)
    std::vector<int> vfBegin, vfFace, veBegin, veEdge, vvBegin, vvNei;
    std::vector<uint8_t> query;  // per vertex: take part in VF queries (empty: all vertices)
    // results
    std::vector<V3> pos;
    float r = 0.f;
    std::vector<float> vfMin, faceMin, eeMin;
    std::vector<VFPair> vfPairs;
    std::vector<VFContact> vfContacts;
    std::vector<EEContact> eeContacts;
};

// Topology per mesh, numbered as the product host code numbers it (TriMesh.cpp:481-753):
// edges by first appearance over faces in order, eV1 -> eV2 the direction in the first face,
// eV12Next the third vertex of that face, eV21Next the third vertex of the other face.
int build_topology(Detector& D, int nmesh, const int* meshNv, const int* meshNf, const int* facesLocal) {
    D.nmesh = nmesh;
    int vOff = 0, fOff = 0;
    for (int m = 0; m < nmesh; ++m) {
        std::map<std::pair<int, int>, int> edgeOf;
        const int eOff = (int)(D.edges.size() / 4);
        for (int f = 0; f < meshNf[m]; ++f) {
            const int* fl = facesLocal + 3 * (size_t)(fOff + f);
            int fe[3];
            for (int k = 0; k < 3; ++k) {
                const int a = fl[k], b = fl[(k + 1) % 3], c = fl[(k + 2) % 3];
                if (a < 0 || a >= meshNv[m]) return -1;
                const auto key = std::make_pair(std::min(a, b), std::max(a, b));
                auto it = edgeOf.find(key);
                if (it == edgeOf.end()) {
                    const int id = (int)(D.edges.size() / 4) - eOff;
                    edgeOf[key] = id;
                    const int e[4] = {vOff + a, vOff + b, vOff + c, -1};
                    D.edges.insert(D.edges.end(), e, e + 4);
                    D.eMesh.push_back(m);
                    fe[k] = eOff + id;
                } else {
                    int* e = &D.edges[4 * (size_t)(eOff + it->second)];
                    if (e[3] != -1 || e[0] != vOff + b || e[1] != vOff + a) return -1;
                    e[3] = vOff + c;
                    fe[k] = eOff + it->second;
                }
                D.faces.push_back(vOff + a);
            }
            D.faceEdges.insert(D.faceEdges.end(), fe, fe + 3);
            D.fMesh.push_back(m);
        }
        for (int i = 0; i < meshNv[m]; ++i) D.vMesh.push_back(m);
        vOff += meshNv[m];
        fOff += meshNf[m];
    }
    D.nv = vOff;
    D.nf = fOff;
    D.ne = (int)(D.edges.size() / 4);
    auto csr = [&](int n, int count, auto&& items, std::vector<int>& begin, std::vector<int>& out) {
        begin.assign(n + 1, 0);
        items([&](int v, int) { begin[v + 1]++; });
        for (int i = 0; i < n; ++i) begin[i + 1] += begin[i];
        out.assign(begin[n], 0);
        std::vector<int> fill(begin.begin(), begin.end() - 1);
        items([&](int v, int x) { out[fill[v]++] = x; });
        (void)count;
    };
    csr(D.nv, 0, [&](auto&& emit) {
        for (int f = 0; f < D.nf; ++f)
            for (int k = 0; k < 3; ++k) emit(D.faces[3 * f + k], f);
    }, D.vfBegin, D.vfFace);
    csr(D.nv, 0, [&](auto&& emit) {
        for (int e = 0; e < D.ne; ++e) {
            emit(D.edges[4 * e], e);
            emit(D.edges[4 * e + 1], e);
        }
    }, D.veBegin, D.veEdge);
    csr(D.nv, 0, [&](auto&& emit) {
        for (int e = 0; e < D.ne; ++e) {
            emit(D.edges[4 * e], D.edges[4 * e + 1]);
            emit(D.edges[4 * e + 1], D.edges[4 * e]);
        }
    }, D.vvBegin, D.vvNei);
    return 0;
}

bool vertex_feasible(const Detector& D, V3 p, int v) {
    const V3 A = D.pos[v];
    const V3 AP = p - A;
    const float relaxed = -kFeasibleEpsilon - kAbsoluteRelaxation;
    for (int k = D.vvBegin[v]; k < D.vvBegin[v + 1]; ++k) {
        const V3 BA = A - D.pos[D.vvNei[k]];
        if (dot(AP, BA) < relaxed) return false;
    }
    return true;
}

bool edge_feasible(const Detector& D, V3 p, int e) {
    const int* ei = &D.edges[4 * (size_t)e];
    const V3 v1 = D.pos[ei[0]], v2 = D.pos[ei[1]];
    const V3 AP = p - v1, BP = p - v2, AB = v2 - v1;
    const float relaxed = kFeasibleEpsilon + kAbsoluteRelaxation;
    if (dot(AP, AB) < -relaxed) return false;
    if (dot(BP, AB) > relaxed) return false;
    const float ab2 = dot(AB, AB);
    if (ab2 < kCmpEpsilon2) return false;
    for (int k = 2; k < 4; ++k) {
        if (ei[k] == -1) continue;
        const V3 vn = D.pos[ei[k]];
        const float t = dot(AB, vn - v1) / ab2;
        const V3 foot = v1 + t * AB;
        if (dot(vn - foot, p - foot) > relaxed) return false;
    }
    return true;
}

// primitive key of a closest-point type on face f (ClothContactDetector.cpp:196-224)
void primitive_of(const Detector& D, int f, int type, int& prim, int& primType) {
    switch (type) {
    case AT_A: case AT_B: case AT_C:
        prim = D.faces[3 * f + (type - AT_A)];
        primType = PRIM_VERTEX;
        break;
    case AT_AB: prim = D.faceEdges[3 * f + 0]; primType = PRIM_EDGE; break;
    case AT_BC: prim = D.faceEdges[3 * f + 1]; primType = PRIM_EDGE; break;
    case AT_AC: prim = D.faceEdges[3 * f + 2]; primType = PRIM_EDGE; break;
    default: prim = f; primType = PRIM_FACE; break;
    }
}

bool feasible(const Detector& D, V3 p, int f, int type) {
    switch (type) {
    case AT_INTERIOR: return true;
    case AT_AB: return edge_feasible(D, p, D.faceEdges[3 * f + 0]);
    case AT_BC: return edge_feasible(D, p, D.faceEdges[3 * f + 1]);
    case AT_AC: return edge_feasible(D, p, D.faceEdges[3 * f + 2]);
    default: return vertex_feasible(D, p, D.faces[3 * f + (type - AT_A)]);
    }
}

// Uniform grid over primitive boxes grown by the query reach: cell -> primitive ids, sorted.
struct Grid {
    float h = 1.f;
    V3 lo{0, 0, 0};
    std::vector<std::pair<uint64_t, int>> cells;
    static uint64_t key(int64_t i, int64_t j, int64_t k) {
        return ((uint64_t)(i & 0x1FFFFF) << 42) | ((uint64_t)(j & 0x1FFFFF) << 21) | (uint64_t)(k & 0x1FFFFF);
    }
    int64_t cell(float x, float o) const { return (int64_t)std::floor((x - o) / h); }
    void insert(int id, V3 bmin, V3 bmax) {
        for (int64_t i = cell(bmin.x, lo.x); i <= cell(bmax.x, lo.x); ++i)
            for (int64_t j = cell(bmin.y, lo.y); j <= cell(bmax.y, lo.y); ++j)
                for (int64_t k = cell(bmin.z, lo.z); k <= cell(bmax.z, lo.z); ++k) cells.push_back({key(i, j, k), id});
    }
    void finish() { std::sort(cells.begin(), cells.end()); }
    template <class F> void visit(V3 bmin, V3 bmax, F&& f) const {
        for (int64_t i = cell(bmin.x, lo.x); i <= cell(bmax.x, lo.x); ++i)
            for (int64_t j = cell(bmin.y, lo.y); j <= cell(bmax.y, lo.y); ++j)
                for (int64_t k = cell(bmin.z, lo.z); k <= cell(bmax.z, lo.z); ++k) {
                    const uint64_t kk = key(i, j, k);
                    auto it = std::lower_bound(cells.begin(), cells.end(), std::make_pair(kk, INT32_MIN));
                    for (; it != cells.end() && it->first == kk; ++it) f(it->second);
                }
    }
};

V3 vmin3(V3 a, V3 b) { return {std::min(a.x, b.x), std::min(a.y, b.y), std::min(a.z, b.z)}; }
V3 vmax3(V3 a, V3 b) { return {std::max(a.x, b.x), std::max(a.y, b.y), std::max(a.z, b.z)}; }

void detect(Detector& D, const float* pos, float r, int mode, int threads) {
    D.pos.resize(D.nv);
    for (int i = 0; i < D.nv; ++i) D.pos[i] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
    D.r = r;
    const int T = std::max(1, threads);
    // candidate generation
    Grid fg, eg;
    if (mode == 1) {
        float ext = 0.f;
        V3 lo{1e30f, 1e30f, 1e30f};
        for (int i = 0; i < D.nv; ++i) lo = vmin3(lo, D.pos[i]);
        for (int f = 0; f < D.nf; ++f) {
            V3 a = D.pos[D.faces[3 * f]], b = D.pos[D.faces[3 * f + 1]], c = D.pos[D.faces[3 * f + 2]];
            V3 bmin = vmin3(vmin3(a, b), c), bmax = vmax3(vmax3(a, b), c);
            ext = std::max({ext, bmax.x - bmin.x, bmax.y - bmin.y, bmax.z - bmin.z});
        }
        const float pad = 2.f * r;
        fg.h = eg.h = std::max(ext + 2.f * pad, 1e-6f);
        fg.lo = eg.lo = {lo.x - 4 * fg.h, lo.y - 4 * fg.h, lo.z - 4 * fg.h};
        for (int f = 0; f < D.nf; ++f) {
            V3 a = D.pos[D.faces[3 * f]], b = D.pos[D.faces[3 * f + 1]], c = D.pos[D.faces[3 * f + 2]];
            V3 bmin = vmin3(vmin3(a, b), c), bmax = vmax3(vmax3(a, b), c);
            fg.insert(f, {bmin.x - pad, bmin.y - pad, bmin.z - pad}, {bmax.x + pad, bmax.y + pad, bmax.z + pad});
        }
        for (int e = 0; e < D.ne; ++e) {
            V3 a = D.pos[D.edges[4 * e]], b = D.pos[D.edges[4 * e + 1]];
            V3 bmin = vmin3(a, b), bmax = vmax3(a, b);
            eg.insert(e, {bmin.x - pad, bmin.y - pad, bmin.z - pad}, {bmax.x + pad, bmax.y + pad, bmax.z + pad});
        }
        fg.finish();
        eg.finish();
    }
    oracle::ThreadPool pool(T);
    // ---- VF (per vertex, candidate faces ascending)
    std::vector<std::vector<VFPair>> pairsV(D.nv);
    std::vector<std::vector<VFContact>> contactsV(D.nv);
    D.vfMin.assign(D.nv, r);
    pool.parallel_for(0, D.nv, [&](int v) {
        if (!D.query.empty() && !D.query[v]) return;  // not a surface vertex
        const V3 p = D.pos[v];
        std::vector<int> cand;
        if (mode == 1) {
            fg.visit(p, p, [&](int f) { cand.push_back(f); });
            std::sort(cand.begin(), cand.end());
            cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
        }
        std::vector<std::pair<int, int>> keys;  // (primType, prim)
        auto test = [&](int f) {
            if (D.fMesh[f] == D.vMesh[v] &&
                (D.faces[3 * f] == v || D.faces[3 * f + 1] == v || D.faces[3 * f + 2] == v))
                return;
            V3 bary;
            int type;
            const V3 cp = closest_point_triangle(p, D.pos[D.faces[3 * f]], D.pos[D.faces[3 * f + 1]],
                                                 D.pos[D.faces[3 * f + 2]], bary, type);
            const float d = distance(p, cp);
            if (!(d < r)) return;
            D.vfMin[v] = std::min(d, D.vfMin[v]);
            pairsV[v].push_back({v, f, d});
            int prim, primType;
            primitive_of(D, f, type, prim, primType);
            for (auto& k : keys)
                if (k.first == primType && k.second == prim) return;
            if (!feasible(D, p, f, type)) return;
            keys.push_back({primType, prim});
            contactsV[v].push_back({v, primType, prim, f, d});
        };
        if (mode == 1)
            for (int f : cand) test(f);
        else
            for (int f = 0; f < D.nf; ++f) test(f);
    });
    D.vfPairs.clear();
    D.vfContacts.clear();
    for (int v = 0; v < D.nv; ++v) {
        D.vfPairs.insert(D.vfPairs.end(), pairsV[v].begin(), pairsV[v].end());
        D.vfContacts.insert(D.vfContacts.end(), contactsV[v].begin(), contactsV[v].end());
    }
    std::sort(D.vfPairs.begin(), D.vfPairs.end(),
              [](const VFPair& a, const VFPair& b) { return a.v != b.v ? a.v < b.v : a.f < b.f; });
    std::sort(D.vfContacts.begin(), D.vfContacts.end(), [](const VFContact& a, const VFContact& b) {
        if (a.v != b.v) return a.v < b.v;
        if (a.type != b.type) return a.type < b.type;
        return a.prim < b.prim;
    });
    D.faceMin.assign(D.nf, r);
    for (const VFPair& q : D.vfPairs) D.faceMin[q.f] = std::min(D.faceMin[q.f], q.d);
    // ---- EE (per edge, candidate edges ascending)
    std::vector<std::vector<EEContact>> eeV(D.ne);
    D.eeMin.assign(D.ne, r);
    pool.parallel_for(0, D.ne, [&](int e1) {
        const int* a = &D.edges[4 * (size_t)e1];
        const V3 p1 = D.pos[a[0]], p2 = D.pos[a[1]];
        std::vector<int> cand;
        if (mode == 1) {
            V3 bmin = vmin3(p1, p2), bmax = vmax3(p1, p2);
            eg.visit(bmin, bmax, [&](int e) { cand.push_back(e); });
            std::sort(cand.begin(), cand.end());
            cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
        }
        auto test = [&](int e2) {
            const int* b = &D.edges[4 * (size_t)e2];
            if (D.eMesh[e2] == D.eMesh[e1]) {
                if (e2 == e1) return;
                if (a[0] == b[0] || a[0] == b[1] || a[1] == b[0] || a[1] == b[1]) return;
            }
            const V3 q1 = D.pos[b[0]], q2 = D.pos[b[1]];
            V3 c1, c2;
            float mua, mub;
            segment_closest_points(p1, p2, q1, q2, c1, c2, mua, mub);
            const float d = distance(c1, c2);
            D.eeMin[e1] = std::min(d, D.eeMin[e1]);
            if (!(d < r)) return;
            const V3 ev1 = p2 - p1, ev2 = q2 - q1;
            bool ok = false;
            if (dot(c2 - p1, ev1) >= 0 && dot(p2 - c2, ev1) >= 0)
                ok = true;
            else if (dot(c1 - q1, ev2) >= 0 && dot(q2 - c1, ev2) >= 0)
                ok = true;
            if (ok) eeV[e1].push_back({e1, e2, mua, mub, d});
        };
        if (mode == 1)
            for (int e : cand) test(e);
        else
            for (int e = 0; e < D.ne; ++e) test(e);
    });
    D.eeContacts.clear();
    for (int e = 0; e < D.ne; ++e) D.eeContacts.insert(D.eeContacts.end(), eeV[e].begin(), eeV[e].end());
    std::sort(D.eeContacts.begin(), D.eeContacts.end(),
              [](const EEContact& x, const EEContact& y) { return x.e1 != y.e1 ? x.e1 < y.e1 : x.e2 < y.e2; });
}

}  // namespace

extern "C" {

void* oracle_cd_create(int nmesh, const int* mesh_nv, const int* mesh_nf, const int* faces) {
    auto D = std::make_unique<Detector>();
    if (build_topology(*D, nmesh, mesh_nv, mesh_nf, faces)) return nullptr;
    return D.release();
}

void oracle_cd_destroy(void* h) { delete (Detector*)h; }

// Vertices with mask 0 take no VF queries and get an unbounded step (tet interior vertices).
void oracle_cd_set_query_mask(void* h, const uint8_t* mask) {
    Detector& D = *(Detector*)h;
    D.query.assign(mask, mask + D.nv);
}

void oracle_cd_sizes(void* h, int* nv, int* nf, int* ne) {
    const Detector& D = *(Detector*)h;
    *nv = D.nv;
    *nf = D.nf;
    *ne = D.ne;
}

void oracle_cd_topology(void* h, int* edges, int* face_edges) {
    const Detector& D = *(Detector*)h;
    if (edges) std::memcpy(edges, D.edges.data(), D.edges.size() * sizeof(int));
    if (face_edges) std::memcpy(face_edges, D.faceEdges.data(), D.faceEdges.size() * sizeof(int));
}

void oracle_cd_detect(void* h, const float* pos, float radius, int mode, int threads) {
    detect(*(Detector*)h, pos, radius, mode, threads);
}

void oracle_cd_counts(void* h, int64_t* counts) {
    const Detector& D = *(Detector*)h;
    counts[0] = (int64_t)D.vfPairs.size();
    counts[1] = (int64_t)D.vfContacts.size();
    counts[2] = (int64_t)D.eeContacts.size();
}

void oracle_cd_vf_pairs(void* h, int* vf, float* d) {
    const Detector& D = *(Detector*)h;
    for (size_t i = 0; i < D.vfPairs.size(); ++i) {
        vf[2 * i] = D.vfPairs[i].v;
        vf[2 * i + 1] = D.vfPairs[i].f;
        if (d) d[i] = D.vfPairs[i].d;
    }
}

void oracle_cd_vf_contacts(void* h, int* rec, float* d) {
    const Detector& D = *(Detector*)h;
    for (size_t i = 0; i < D.vfContacts.size(); ++i) {
        const VFContact& c = D.vfContacts[i];
        rec[4 * i] = c.v;
        rec[4 * i + 1] = c.type;
        rec[4 * i + 2] = c.prim;
        rec[4 * i + 3] = c.face;
        if (d) d[i] = c.d;
    }
}

void oracle_cd_ee_contacts(void* h, int* ee, float* mu, float* d) {
    const Detector& D = *(Detector*)h;
    for (size_t i = 0; i < D.eeContacts.size(); ++i) {
        const EEContact& c = D.eeContacts[i];
        ee[2 * i] = c.e1;
        ee[2 * i + 1] = c.e2;
        if (mu) {
            mu[2 * i] = c.mu1;
            mu[2 * i + 1] = c.mu2;
        }
        if (d) d[i] = c.d;
    }
}

void oracle_cd_min_dis(void* h, float* vf, float* face, float* edge) {
    const Detector& D = *(Detector*)h;
    if (vf) std::memcpy(vf, D.vfMin.data(), D.vfMin.size() * sizeof(float));
    if (face) std::memcpy(face, D.faceMin.data(), D.faceMin.size() * sizeof(float));
    if (edge) std::memcpy(edge, D.eeMin.data(), D.eeMin.size() * sizeof(float));
}

// VBDClothPhysics.cpp:775-812: b_v = relax * max(0, min(r, vf_v, faces of v, edges of v) - thickness)
void oracle_cd_bounds(void* h, float thickness, float relaxation, float* bounds) {
    const Detector& D = *(Detector*)h;
    for (int v = 0; v < D.nv; ++v) {
        if (!D.query.empty() && !D.query[v]) {
            bounds[v] = FLT_MAX;
            continue;
        }
        float dMin = D.r;
        dMin = std::min(dMin, D.vfMin[v]);
        for (int k = D.vfBegin[v]; k < D.vfBegin[v + 1]; ++k) dMin = std::min(dMin, D.faceMin[D.vfFace[k]]);
        for (int k = D.veBegin[v]; k < D.veBegin[v + 1]; ++k) dMin = std::min(dMin, D.eeMin[D.veEdge[k]]);
        dMin = std::max(0.f, dMin - thickness);
        bounds[v] = dMin * relaxation;
    }
}

int oracle_closest_point_triangle(const float* p, const float* a, const float* b, const float* c, float* closest,
                                  float* bary) {
    V3 br;
    int type;
    const V3 r = closest_point_triangle({p[0], p[1], p[2]}, {a[0], a[1], a[2]}, {b[0], b[1], b[2]}, {c[0], c[1], c[2]},
                                        br, type);
    closest[0] = r.x;
    closest[1] = r.y;
    closest[2] = r.z;
    bary[0] = br.x;
    bary[1] = br.y;
    bary[2] = br.z;
    return type;
}

void oracle_segment_closest_points(const float* p0, const float* p1, const float* q0, const float* q1, float* cp,
                                   float* cq, float* st) {
    V3 a, b;
    segment_closest_points({p0[0], p0[1], p0[2]}, {p1[0], p1[1], p1[2]}, {q0[0], q0[1], q0[2]}, {q1[0], q1[1], q1[2]},
                           a, b, st[0], st[1]);
    cp[0] = a.x;
    cp[1] = a.y;
    cp[2] = a.z;
    cq[0] = b.x;
    cq[1] = b.y;
    cq[2] = b.z;
}

}  // extern "C"
