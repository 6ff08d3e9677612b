// GPU contact detector (CDNA4 / gfx950): linear BVHs over faces and edges replace the
// reference's Embree scenes (CollisionDetector/ClothContactDetector.cpp:944-1181).
//
//   build   scene box of primitive centroids -> 30-bit Morton codes in the box (the 10-bit-per-
//           axis interleave of BVH/MortonCode.h:9-31) with the primitive id in the low 32 bits ->
//           device radix sort -> binary radix tree from the sorted keys (Karras 2012, HPG) ->
//           node depths -> internal nodes sorted by depth (the refit schedule)
//   refit   leaf boxes from the current positions, then one launch per depth level, deepest
//           first (no atomics, no inter-block waits)
//   query   one thread per vertex (VF, query box of half-size r around the vertex) or per edge
//           (EE, half-size r + |e|/2 around the midpoint), stack traversal, exact tests with the
//           reference's predicates (ClothContactDetector.cpp:144-288, 731-813; CollisionGeometry.h:
//           90-267; TriMeshCollisionGeometry.cpp:7-186)
//
// The exact tests use the same float operations in the same order as the CPU restatement
// (dot = (x + y) + z; no contraction into fused multiply-adds; IEEE division and square root),
// so candidate sets match it bit for bit. Outputs that the reference takes from its Embree
// traversal order are made order-free: the representative face of a VF contact is the lowest
// face id with that closest primitive; per-primitive minima are exact minima.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <string>

#include "detector.h"
#include "errors.h"
#include "mesh_prep.h"

#pragma clang fp contract(off)

namespace gvbd {
namespace {

constexpr int kMaxCap = 16;       // largest per-query contact capacity
// Traversal is stackless: every node's hi.w holds its skip link (the next node in depth-first
// order after its subtree, -1 at the end), so a query visits left children first and follows
// skip links; no per-thread stack lives in scratch memory.
constexpr int kMaxDepth = 128;    // depth histogram size
constexpr int kBlockLevel = 1024;  // refit levels of at most this many nodes share one-block launches
constexpr float kCmpEpsilon = 0.00001f;
constexpr float kCmpEpsilon2 = kCmpEpsilon * kCmpEpsilon;
constexpr float kAbsoluteRelaxation = 1e-6f;
constexpr float kFeasibleEpsilon = 1e-3f;

struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 operator*(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ float dist(V3 a, V3 b) {
    const V3 d = a - b;
    return sqrtf(dot(d, d));
}
__device__ __forceinline__ V3 ld(const float4* p, int i) {
    const float4 v = p[i];
    return {v.x, v.y, v.z};
}

enum PointType { AT_A = 0, AT_B, AT_C, AT_AB, AT_BC, AT_AC, AT_INTERIOR };

__device__ V3 closest_point_triangle(V3 p, V3 a, V3 b, V3 c, int& type) {
    const V3 ab = b - a, ac = c - a, ap = p - a;
    const float d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.f && d2 <= 0.f) {
        type = AT_A;
        return a;
    }
    const V3 bp = p - b;
    const float d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0.f && d4 <= d3) {
        type = AT_B;
        return b;
    }
    const V3 cp = p - c;
    const float d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0.f && d5 <= d6) {
        type = AT_C;
        return c;
    }
    const float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.f && d1 >= 0.f && d3 <= 0.f) {
        const float v = d1 / (d1 - d3);
        type = AT_AB;
        return a + v * ab;
    }
    const float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.f && d2 >= 0.f && d6 <= 0.f) {
        const float v = d2 / (d2 - d6);
        type = AT_AC;
        return a + v * ac;
    }
    const float va = d3 * d6 - d5 * d4;
    if (va <= 0.f && (d4 - d3) >= 0.f && (d5 - d6) >= 0.f) {
        type = AT_BC;
        const float v = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        return b + v * (c - b);
    }
    const float denom = 1.f / (va + vb + vc);
    const float v = vb * denom, w = vc * denom;
    type = AT_INTERIOR;
    return (a + v * ab) + w * ac;
}

__device__ void segment_closest_points(V3 p0, V3 p1, V3 q0, V3 q1, V3& cp, V3& cq, float& s, float& t) {
    const V3 p = p1 - p0, q = q1 - q0, r = p0 - q0;
    const float a = dot(p, p), b = dot(p, q), c = dot(q, q), d = dot(p, r), e = dot(q, r);
    s = 0.0f;
    t = 0.0f;
    const float det = a * c - b * b;
    if (det > kCmpEpsilon) {
        const float bte = b * e, ctd = c * d;
        if (bte <= ctd) {
            if (e <= 0.0f) {
                s = (-d >= a ? 1.f : (-d > 0.0f ? -d / a : 0.0f));
                t = 0.0f;
            } else if (e < c) {
                s = 0.0f;
                t = e / c;
            } else {
                s = (b - d >= a ? 1.f : (b - d > 0.0f ? (b - d) / a : 0.0f));
                t = 1.f;
            }
        } else {
            s = bte - ctd;
            if (s >= det) {
                if (b + e <= 0.0f) {
                    s = (-d <= 0.0f ? 0.0f : (-d < a ? -d / a : 1.f));
                    t = 0.0f;
                } else if (b + e < c) {
                    s = 1.f;
                    t = (b + e) / c;
                } else {
                    s = (b - d <= 0.0f ? 0.0f : (b - d < a ? (b - d) / a : 1.f));
                    t = 1.f;
                }
            } else {
                const float ate = a * e, btd = b * d;
                if (ate <= btd) {
                    s = (-d <= 0.0f ? 0.0f : (-d >= a ? 1.f : -d / a));
                    t = 0.0f;
                } else {
                    t = ate - btd;
                    if (t >= det) {
                        s = (b - d <= 0.0f ? 0.0f : (b - d >= a ? 1.f : (b - d) / a));
                        t = 1.f;
                    } else {
                        s /= det;
                        t /= det;
                    }
                }
            }
        }
    } else {
        if (e <= 0.0f) {
            s = (-d <= 0.0f ? 0.0f : (-d >= a ? 1.f : -d / a));
            t = 0.0f;
        } else if (e >= c) {
            s = (b - d <= 0.0f ? 0.0f : (b - d >= a ? 1.f : (b - d) / a));
            t = 1.f;
        } else {
            s = 0.0f;
            t = e / c;
        }
    }
    cp = (1.f - s) * p0 + s * p1;
    cq = (1.f - t) * q0 + t * q1;
}

// Device view of a detector's topology and outputs.
struct CDView {
    const float4* pos;
    int nv, nf, ne;
    const int* vMesh;
    const int4* faces;      // v0, v1, v2, mesh
    const int4* faceEdges;  // edges (v0 v1), (v1 v2), (v2 v0)
    const int4* edges;      // eV1, eV2, eV12Next, eV21Next
    const int* eMesh;
    const int* vvBegin;
    const int* vvNei;
    const BvhNode* fNodes;
    int fN;
    const BvhNode* eNodes;
    int eN;
    float r;
    int vfCap, eeCap;
    float* vfMin;
    int* faceMin;
    float* eeMin;
    int* vfPairCount;
    int* vfStored;
    int* vfOverflow;
    int* eeCount;
    int* eeStored;
    int* eeOverflow;
    VFContactRec* vfRec;
    EEContactRec* eeRec;
    // record mode
    const int64_t* vfOffset;
    const int64_t* eeOffset;
    int2* recVF;
    float* recVFd;
    int4* recKeys;
    float* recKeysD;
    int2* recEE;
    float2* recEEmu;
    float* recEEd;
    const uint64_t* eOrder;  // edge tree's sorted Morton keys: query edges in leaf order (or NULL)
    const uint64_t* vOrder;  // vertices' sorted Morton keys: query vertices in that order (or NULL)
    const uint8_t* vQuery;   // per vertex: takes VF queries (NULL: every vertex)
    const int* gate;         // device flag: 0 skips the pass's kernels (NULL: always run)
    int* error;
    const float4* boxes;     // kind 3 primitives: lo, hi per primitive
};

__device__ __forceinline__ bool box_overlap(V3 qlo, V3 qhi, const BvhNode& n) {
    return qlo.x <= n.hi.x && qhi.x >= n.lo.x && qlo.y <= n.hi.y && qhi.y >= n.lo.y && qlo.z <= n.hi.z &&
           qhi.z >= n.lo.z;
}

__device__ bool vertex_feasible(const CDView& S, V3 p, int v) {
    const V3 A = ld(S.pos, v);
    const V3 AP = p - A;
    const float relaxed = -kFeasibleEpsilon - kAbsoluteRelaxation;
    for (int k = S.vvBegin[v]; k < S.vvBegin[v + 1]; ++k) {
        const V3 BA = A - ld(S.pos, S.vvNei[k]);
        if (dot(AP, BA) < relaxed) return false;
    }
    return true;
}

__device__ bool edge_feasible(const CDView& S, V3 p, int e) {
    const int4 ei = S.edges[e];
    const V3 v1 = ld(S.pos, ei.x), v2 = ld(S.pos, ei.y);
    const V3 AP = p - v1, BP = p - v2, AB = v2 - v1;
    const float relaxed = kFeasibleEpsilon + kAbsoluteRelaxation;
    if (dot(AP, AB) < -relaxed) return false;
    if (dot(BP, AB) > relaxed) return false;
    const float ab2 = dot(AB, AB);
    if (ab2 < kCmpEpsilon2) return false;
    const int nx[2] = {ei.z, ei.w};
    for (int k = 0; k < 2; ++k) {
        if (nx[k] == -1) continue;
        const V3 vn = ld(S.pos, nx[k]);
        const float t = dot(AB, vn - v1) / ab2;
        const V3 foot = v1 + t * AB;
        if (dot(vn - foot, p - foot) > relaxed) return false;
    }
    return true;
}

// VF query of one vertex (ClothContactDetector.cpp:144-288 per candidate face)
// CAP bounds the per-thread candidate list (the runtime vfCap is at most CAP); only the face
// id and the packed (type, primitive) key are kept per candidate, the distance is recomputed
// for the stored ones at the end, so the list costs 2 CAP registers.
template <bool RECORD, int CAP>
__global__ __launch_bounds__(256) void vf_query_kernel(CDView S) {
    // vertices in Morton order: a wavefront's queries are spatial neighbours, and so are the
    // blocks of one XCD
    if (S.gate && !*S.gate) return;
    const int t = xcd_block_index() * blockDim.x + threadIdx.x;
    if (t >= S.nv) return;
    const int v = S.vOrder ? (int)(S.vOrder[t] & 0xFFFFFFFFu) : t;
    if (S.vQuery && !S.vQuery[v]) {  // not a surface vertex: no VF contacts
        if (!RECORD) {
            S.vfMin[v] = S.r;
            S.vfPairCount[v] = 0;
            S.vfStored[v] = 0;
            S.vfOverflow[v] = 0;
        }
        return;
    }
    const V3 p = ld(S.pos, v);
    const int mesh = S.vMesh[v];
    const float r = S.r;
    const float rq = r * 1.0001f + 1e-7f;
    const V3 qlo = {p.x - rq, p.y - rq, p.z - rq}, qhi = {p.x + rq, p.y + rq, p.z + rq};
    float vmin = r;
    int pairs = 0, stored = 0, overflow = 0;
    int kKey[CAP], kFace[CAP];  // key: primitive | type << 30
    int64_t pairBase = 0, keyBase = 0;
    int recKeys = 0, pairCap = 0;
    if (RECORD) {
        pairCap = S.vfPairCount[v];
        pairBase = S.vfOffset[v];
        keyBase = pairBase;  // the key segment has the pair segment's size (an upper bound)
    }
    const int firstLeaf = S.fN - 1;
    int node = S.fN > 0 ? 0 : -1;
    for (int visits = 0; node >= 0; ++visits) {
        if (visits > 2 * S.fN) {  // a depth-first walk visits every node at most once
            atomicOr(S.error, 1);
            break;
        }
        const BvhNode nd = S.fNodes[node];
        const int skip = __float_as_int(nd.hi.w);
        if (!box_overlap(qlo, qhi, nd)) {
            node = skip;
            continue;
        }
        if (node < firstLeaf) {
            node = __float_as_int(nd.lo.w);
            continue;
        }
        node = skip;
        const int f = __float_as_int(nd.lo.w);
        const int4 fv = S.faces[f];
        if (fv.w == mesh && (fv.x == v || fv.y == v || fv.z == v)) continue;
        int type;
        const V3 cp = closest_point_triangle(p, ld(S.pos, fv.x), ld(S.pos, fv.y), ld(S.pos, fv.z), type);
        const float d = dist(p, cp);
        if (!(d < r)) continue;
        vmin = fminf(d, vmin);
        atomicMin(S.faceMin + f, __float_as_int(d));
        if (RECORD) {
            // the record pass repeats the counting pass's traversal; never write past its count
            if (pairs >= pairCap) {
                atomicOr(S.error, 4);
                break;
            }
            S.recVF[pairBase + pairs] = make_int2(v, f);
            S.recVFd[pairBase + pairs] = d;
        }
        ++pairs;
        int prim, ptype;
        if (type <= AT_C) {
            prim = type == AT_A ? fv.x : (type == AT_B ? fv.y : fv.z);
            ptype = 0;
        } else if (type == AT_INTERIOR) {
            prim = f;
            ptype = 2;
        } else {
            const int4 fe = S.faceEdges[f];
            prim = type == AT_AB ? fe.x : (type == AT_BC ? fe.y : fe.z);
            ptype = 1;
        }
        // dedup on (type, primitive): keep the lowest face id as the representative
        bool dup = false;
        if (!RECORD) {
            const int key = prim | (ptype << 30);
#pragma unroll
            for (int k = 0; k < CAP; ++k)
                if (k < stored && !dup && kKey[k] == key) {
                    if (f < kFace[k]) kFace[k] = f;
                    dup = true;
                }
        } else {
            for (int k = 0; k < recKeys; ++k) {
                int4& rk = S.recKeys[keyBase + k];
                if (rk.y == ptype && rk.z == prim) {
                    if (f < rk.w) {
                        rk.w = f;
                        S.recKeysD[keyBase + k] = d;
                    }
                    dup = true;
                    break;
                }
            }
        }
        if (dup) continue;
        bool ok = true;
        if (ptype == 0)
            ok = vertex_feasible(S, p, prim);
        else if (ptype == 1)
            ok = edge_feasible(S, p, prim);
        if (!ok) continue;
        if (RECORD) {
            S.recKeys[keyBase + recKeys] = make_int4(v, ptype, prim, f);
            S.recKeysD[keyBase + recKeys] = d;
            ++recKeys;
        } else {
            // keep the vfCap keys with the lowest representative faces (order-free selection)
            int slot = stored;
            if (stored == S.vfCap) {
                overflow = 1;
                slot = 0;
                int worst = kFace[0];
#pragma unroll
                for (int k = 1; k < CAP; ++k)
                    if (k < stored && kFace[k] > worst) {
                        worst = kFace[k];
                        slot = k;
                    }
                if (f > worst) slot = -1;
            } else {
                ++stored;
            }
#pragma unroll
            for (int k = 0; k < CAP; ++k)
                if (k == slot) {
                    kKey[k] = prim | (ptype << 30);
                    kFace[k] = f;
                }
        }
    }
    if (RECORD) {
        for (int k = recKeys; k < pairs; ++k) S.recKeys[keyBase + k] = make_int4(v, -1, -1, -1);
        return;
    }
    S.vfMin[v] = vmin;
    S.vfPairCount[v] = pairs;
    S.vfStored[v] = stored;
    S.vfOverflow[v] = overflow;
    // stored in ascending face id, whatever the traversal order (so the list, and the solver's sum
    // over it, does not depend on the tree's shape); representative faces are distinct
    int last = -1;
    for (int o = 0; o < stored; ++o) {
        int bf = INT_MAX, bk = 0;
#pragma unroll
        for (int k = 0; k < CAP; ++k)
            if (k < stored && kFace[k] > last && kFace[k] < bf) {
                bf = kFace[k];
                bk = kKey[k];
            }
        // the same closest-point computation as the traversal's, so the same distance
        const int4 fv = S.faces[bf];
        int type;
        const float d = dist(p, closest_point_triangle(p, ld(S.pos, fv.x), ld(S.pos, fv.y), ld(S.pos, fv.z), type));
        S.vfRec[(int64_t)v * S.vfCap + o] = VFContactRec{bf, (bk >> 30) & 3, bk & 0x3FFFFFFF, d};
        last = bf;
    }
}

// EE query of one edge (ClothContactDetector.cpp:731-813 per candidate edge, :1106-1135 setup)
// CAP as for the VF query: only the other edge's id is kept per candidate; the closest points and
// distance of the stored ones are recomputed at the end.
template <bool RECORD, int CAP>
__global__ __launch_bounds__(256) void ee_query_kernel(CDView S) {
    // neighbouring threads take edges that are neighbours in the tree's Morton order, so a
    // wavefront's traversals share their paths (and cache lines), and an XCD's blocks a range
    if (S.gate && !*S.gate) return;
    const int t = xcd_block_index() * blockDim.x + threadIdx.x;
    if (t >= S.ne) return;
    const int e1 = S.eOrder ? (int)(S.eOrder[t] & 0xFFFFFFFFu) : t;
    const int4 a = S.edges[e1];
    const int m1 = S.eMesh[e1];
    const V3 p1 = ld(S.pos, a.x), p2 = ld(S.pos, a.y);
    const float r = S.r;
    // Every edge within r of this one overlaps the segment's box grown by r: a tighter region
    // than the box of the reference's sphere (midpoint, r + len / 2), with the same contacts and
    // minimum distances (both are decided by the exact tests below).
    const float rq = r * 1.0001f + 1e-7f;
    const V3 qlo = {fminf(p1.x, p2.x) - rq, fminf(p1.y, p2.y) - rq, fminf(p1.z, p2.z) - rq};
    const V3 qhi = {fmaxf(p1.x, p2.x) + rq, fmaxf(p1.y, p2.y) + rq, fmaxf(p1.z, p2.z) + rq};
    float emin = r;
    int count = 0, stored = 0, overflow = 0;
    int cE[CAP];
    const int64_t base = RECORD ? S.eeOffset[e1] : 0;
    const int countCap = RECORD ? S.eeCount[e1] : 0;
    const int firstLeaf = S.eN - 1;
    const V3 ev1 = p2 - p1;
    int node = S.eN > 0 ? 0 : -1;
    for (int visits = 0; node >= 0; ++visits) {
        if (visits > 2 * S.eN) {
            atomicOr(S.error, 1);
            break;
        }
        const BvhNode nd = S.eNodes[node];
        const int skip = __float_as_int(nd.hi.w);
        if (!box_overlap(qlo, qhi, nd)) {
            node = skip;
            continue;
        }
        if (node < firstLeaf) {
            node = __float_as_int(nd.lo.w);
            continue;
        }
        node = skip;
        // the other edge's endpoints from its leaf (the positions of the last refit, which are
        // the query's): no edge or position loads for the many leaves beyond r
        const int tag = __float_as_int(nd.lo.w);
        const int e2 = tag & kLeafPrimMask;
        const V3 q1 = {(tag >> kLeafPrimBits) & 1 ? nd.hi.x : nd.lo.x, (tag >> (kLeafPrimBits + 1)) & 1 ? nd.hi.y : nd.lo.y,
                       (tag >> (kLeafPrimBits + 2)) & 1 ? nd.hi.z : nd.lo.z};
        const V3 q2 = {(tag >> kLeafPrimBits) & 1 ? nd.lo.x : nd.hi.x, (tag >> (kLeafPrimBits + 1)) & 1 ? nd.lo.y : nd.hi.y,
                       (tag >> (kLeafPrimBits + 2)) & 1 ? nd.lo.z : nd.hi.z};
        V3 c1, c2;
        float mua, mub;
        segment_closest_points(p1, p2, q1, q2, c1, c2, mua, mub);
        const float d = dist(c1, c2);
        // emin starts at r, so an edge at d >= r (or NaN) leaves it unchanged whether or not the
        // topology filter below would have skipped it
        if (!(d < r)) continue;
        if (S.eMesh[e2] == m1) {
            if (e2 == e1) continue;
            const int4 b = S.edges[e2];
            if (a.x == b.x || a.x == b.y || a.y == b.x || a.y == b.y) continue;
        }
        emin = fminf(d, emin);
        const V3 ev2 = q2 - q1;
        bool ok = false;
        if (dot(c2 - p1, ev1) >= 0.f && dot(p2 - c2, ev1) >= 0.f)
            ok = true;
        else if (dot(c1 - q1, ev2) >= 0.f && dot(q2 - c1, ev2) >= 0.f)
            ok = true;
        if (!ok) continue;
        if (RECORD) {
            if (count >= countCap) {
                atomicOr(S.error, 4);
                break;
            }
            S.recEE[base + count] = make_int2(e1, e2);
            S.recEEmu[base + count] = make_float2(mua, mub);
            S.recEEd[base + count] = d;
        } else {
            // keep the eeCap contacts with the lowest other-edge ids (order-free selection)
            int slot = stored;
            if (stored == S.eeCap) {
                overflow = 1;
                slot = 0;
                int worst = cE[0];
#pragma unroll
                for (int k = 1; k < CAP; ++k)
                    if (k < stored && cE[k] > worst) {
                        worst = cE[k];
                        slot = k;
                    }
                if (e2 > worst) slot = -1;
            } else {
                ++stored;
            }
#pragma unroll
            for (int k = 0; k < CAP; ++k)
                if (k == slot) cE[k] = e2;
        }
        ++count;
    }
    if (RECORD) return;
    S.eeMin[e1] = emin;
    S.eeCount[e1] = count;
    S.eeStored[e1] = stored;
    S.eeOverflow[e1] = overflow;
    // stored in ascending other-edge id (independent of the traversal order)
    int last = -1;
    for (int o = 0; o < stored; ++o) {
        int be = INT_MAX;
#pragma unroll
        for (int k = 0; k < CAP; ++k)
            if (k < stored && cE[k] > last && cE[k] < be) be = cE[k];
        const int4 b = S.edges[be];
        V3 c1, c2;
        float mua, mub;
        segment_closest_points(p1, p2, ld(S.pos, b.x), ld(S.pos, b.y), c1, c2, mua, mub);
        S.eeRec[(int64_t)e1 * S.eeCap + o] = EEContactRec{be, mua, mub, dist(c1, c2)};
        last = be;
    }
}

// VBDClothPhysics.cpp:775-812 per vertex
__global__ __launch_bounds__(256) void bounds_kernel(const float* vfMin, const int* faceMin, const float* eeMin,
                                                     const int* vfBegin, const int* vfFace, const int* veBegin,
                                                     const int* veEdge, const uint8_t* vQuery, int nv, float r,
                                                     float thickness, float relaxation, float* out,
                                                     const int* gate) {
    if (gate && !*gate) return;
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    if (vQuery && !vQuery[v]) {  // interior vertices of tet meshes: unbounded
        out[v] = FLT_MAX;
        return;
    }
    float dMin = r;
    dMin = fminf(dMin, vfMin[v]);
    for (int k = vfBegin[v]; k < vfBegin[v + 1]; ++k) dMin = fminf(dMin, __int_as_float(faceMin[vfFace[k]]));
    for (int k = veBegin[v]; k < veBegin[v + 1]; ++k) dMin = fminf(dMin, eeMin[veEdge[k]]);
    dMin = fmaxf(0.f, dMin - thickness);
    out[v] = dMin * relaxation;
}

// ---- build / refit -------------------------------------------------------------------------

__device__ __forceinline__ unsigned int order_bits(float f) {
    const unsigned int u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float unorder_bits(unsigned int u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}

__device__ __forceinline__ V3 prim_centroid(const CDView& S, int kind, int i) {
    if (kind == 0) {
        const int4 f = S.faces[i];
        const V3 a = ld(S.pos, f.x), b = ld(S.pos, f.y), c = ld(S.pos, f.z);
        return (1.f / 3.f) * ((a + b) + c);
    }
    if (kind == 1) {
        const int4 e = S.edges[i];
        return 0.5f * (ld(S.pos, e.x) + ld(S.pos, e.y));
    }
    if (kind == 3) {
        const float4 lo = S.boxes[2 * i], hi = S.boxes[2 * i + 1];
        return 0.5f * (V3{lo.x, lo.y, lo.z} + V3{hi.x, hi.y, hi.z});
    }
    return ld(S.pos, i);  // kind 2: vertices
}

// Centroid bounds in two passes without atomics: each 256-thread block writes its 6 ordered
// bounds to partial[6 * block], then one block reduces the partials into box.
__global__ __launch_bounds__(256) void centroid_box_kernel(CDView S, int kind, int n, unsigned int* partial) {
    __shared__ unsigned int red[6][4];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned int v[6] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u, 0u};
    if (i < n) {
        const V3 c = prim_centroid(S, kind, i);
        const float cc[3] = {c.x, c.y, c.z};
        for (int d = 0; d < 3; ++d) v[d] = v[3 + d] = order_bits(cc[d]);
    }
    for (int d = 0; d < 6; ++d) {
        unsigned int a = v[d];
        for (int o = 32; o > 0; o >>= 1) {
            const unsigned int b = (unsigned int)__shfl_xor((int)a, o, 64);
            a = d < 3 ? min(a, b) : max(a, b);
        }
        if ((threadIdx.x & 63) == 0) red[d][threadIdx.x >> 6] = a;
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int d = threadIdx.x;
        unsigned int a = red[d][0];
        for (int w = 1; w < 4; ++w) a = d < 3 ? min(a, red[d][w]) : max(a, red[d][w]);
        partial[6 * (int64_t)blockIdx.x + d] = a;
    }
}

__global__ __launch_bounds__(256) void box_final_kernel(const unsigned int* partial, int blocks, unsigned int* box) {
    __shared__ unsigned int red[6][256];
    unsigned int v[6] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u, 0u};
    for (int b = threadIdx.x; b < blocks; b += blockDim.x)
        for (int d = 0; d < 6; ++d) {
            const unsigned int x = partial[6 * (int64_t)b + d];
            v[d] = d < 3 ? min(v[d], x) : max(v[d], x);
        }
    for (int d = 0; d < 6; ++d) red[d][threadIdx.x] = v[d];
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int d = 0; d < 6; ++d)
                red[d][threadIdx.x] = d < 3 ? min(red[d][threadIdx.x], red[d][threadIdx.x + s])
                                            : max(red[d][threadIdx.x], red[d][threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x < 6) box[threadIdx.x] = red[threadIdx.x][0];
}

__device__ __forceinline__ unsigned int expand_bits10(unsigned int v) {
    // spread the low 10 bits of v to every third bit
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

__global__ __launch_bounds__(256) void morton_kernel(CDView S, int kind, int n, const unsigned int* box,
                                                     uint64_t* keys) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const V3 c = prim_centroid(S, kind, i);
    const float cc[3] = {c.x, c.y, c.z};
    unsigned int q[3];
    for (int d = 0; d < 3; ++d) {
        const float lo = unorder_bits(box[d]), hi = unorder_bits(box[3 + d]);
        const float ext = hi - lo;
        float u = ext > 0.f ? (cc[d] - lo) / ext : 0.f;
        u = fminf(fmaxf(u * 1024.f, 0.f), 1023.f);
        q[d] = (unsigned int)u;
    }
    const unsigned int code = expand_bits10(q[0]) * 4u + expand_bits10(q[1]) * 2u + expand_bits10(q[2]);
    keys[i] = ((uint64_t)code << 32) | (uint32_t)i;
}

__device__ __forceinline__ int common_prefix(const uint64_t* keys, int n, int i, int64_t j) {
    if (j < 0 || j >= n) return -1;
    return __clzll(keys[i] ^ keys[j]);
}

// Binary radix tree over sorted unique keys (Karras 2012): internal node i covers a key range
// starting or ending at i; its split is the highest differing bit in the range.
__global__ __launch_bounds__(256) void build_tree_kernel(const uint64_t* keys, int n, BvhNode* nodes, int* parent,
                                                         int* rightChild) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {  // leaves
        BvhNode& L = nodes[n - 1 + i];
        L.lo.w = __int_as_float((int)(keys[i] & 0xFFFFFFFFu));
    }
    if (i == 0) parent[0] = -1;
    if (i >= n - 1) return;
    const int d = (common_prefix(keys, n, i, i + 1) - common_prefix(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = common_prefix(keys, n, i, (int64_t)i - d);
    int64_t lmax = 2;
    while (common_prefix(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
    int64_t l = 0;
    for (int64_t t = lmax / 2; t >= 1; t /= 2)
        if (common_prefix(keys, n, i, i + (l + t) * d) > dmin) l += t;
    const int64_t j = i + l * d;
    const int dnode = common_prefix(keys, n, i, j);
    int64_t s = 0;
    for (int64_t t = (l + 1) / 2;; t = (t + 1) / 2) {
        if (common_prefix(keys, n, i, i + (s + t) * d) > dnode) s += t;
        if (t == 1) break;
    }
    const int gamma = (int)(i + s * d + min(d, 0));
    const int lo = (int)min((int64_t)i, j), hi = (int)max((int64_t)i, j);
    const int left = lo == gamma ? n - 1 + gamma : gamma;
    const int right = hi == gamma + 1 ? n - 1 + gamma + 1 : gamma + 1;
    nodes[i].lo.w = __int_as_float(left);
    rightChild[i] = right;
    parent[left] = i;
    parent[right] = i;
}

// Skip link of every node (written to hi.w): the right sibling of the nearest node on the path
// to the root (itself included) that is a left child; -1 past the last subtree.
__global__ __launch_bounds__(256) void skip_kernel(BvhNode* nodes, const int* parent, const int* rightChild,
                                                   int total, int* error) {
    const int c0 = blockIdx.x * blockDim.x + threadIdx.x;
    if (c0 >= total) return;
    int c = c0, skip = -1;
    for (int steps = 0;; ++steps) {
        const int p = parent[c];
        if (p < 0) break;
        if (__float_as_int(nodes[p].lo.w) == c) {
            skip = rightChild[p];
            break;
        }
        if (steps >= 4 * kMaxDepth) {
            atomicOr(error, 2);
            break;
        }
        c = p;
    }
    nodes[c0].hi.w = __int_as_float(skip);
}

// Depths of the internal nodes by pointer jumping: after round k every node knows its distance
// to its 2^k-th ancestor (or to the root), so log2(kMaxDepth) rounds give every depth.
__global__ __launch_bounds__(256) void depth_init_kernel(const int* parent, int nInternal, int* d, int* anc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nInternal) return;
    const int p = parent[i];
    d[i] = p >= 0 ? 1 : 0;
    anc[i] = p;
}

__global__ __launch_bounds__(256) void depth_jump_kernel(const int* dIn, const int* ancIn, int* dOut, int* ancOut,
                                                         int nInternal) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nInternal) return;
    const int a = ancIn[i];
    dOut[i] = a >= 0 ? dIn[i] + dIn[a] : dIn[i];
    ancOut[i] = a >= 0 ? ancIn[a] : -1;
}

// sort keys, deepest first: ascending (kMaxDepth - depth, node)
__global__ __launch_bounds__(256) void depth_keys_kernel(const int* d, const int* anc, int nInternal, int* depth,
                                                         uint64_t* keys, int* error) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nInternal) return;
    int dpt = d[i];
    if (anc[i] >= 0 || dpt >= kMaxDepth) {
        atomicOr(error, 2);
        dpt = kMaxDepth - 1;
    }
    depth[i] = dpt;
    keys[i] = ((uint64_t)(kMaxDepth - dpt) << 32) | (uint32_t)i;
}

// node ids in refit order, and the first index of each depth's run (starts[depth], -1 if none)
__global__ __launch_bounds__(256) void level_nodes_kernel(const uint64_t* keys, int n, int* out, int* starts) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = (int)(keys[i] & 0xFFFFFFFFu);
    const int dk = kMaxDepth - (int)(keys[i] >> 32);
    if (i == 0 || (int)(keys[i - 1] >> 32) != (int)(keys[i] >> 32)) starts[dk] = i;
}

__global__ __launch_bounds__(256) void leaf_box_kernel(CDView S, int kind, BvhNode* nodes, int n) {
    if (S.gate && !*S.gate) return;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    BvhNode& L = nodes[n - 1 + k];
    const int prim = __float_as_int(L.lo.w) & kLeafPrimMask;
    int tag = prim;
    V3 lo, hi;
    if (kind == 3) {
        const float4 a = S.boxes[2 * prim], b = S.boxes[2 * prim + 1];
        lo = {a.x, a.y, a.z};
        hi = {b.x, b.y, b.z};
    } else if (kind == 0) {
        const int4 f = S.faces[prim];
        const V3 a = ld(S.pos, f.x), b = ld(S.pos, f.y), c = ld(S.pos, f.z);
        lo = {fminf(fminf(a.x, b.x), c.x), fminf(fminf(a.y, b.y), c.y), fminf(fminf(a.z, b.z), c.z)};
        hi = {fmaxf(fmaxf(a.x, b.x), c.x), fmaxf(fmaxf(a.y, b.y), c.y), fmaxf(fmaxf(a.z, b.z), c.z)};
    } else {
        // the same box as min / max of the endpoints, each bound copied from one endpoint, which
        // the tag bits name
        const int4 e = S.edges[prim];
        const V3 a = ld(S.pos, e.x), b = ld(S.pos, e.y);
        const bool sx = a.x > b.x, sy = a.y > b.y, sz = a.z > b.z;
        lo = {sx ? b.x : a.x, sy ? b.y : a.y, sz ? b.z : a.z};
        hi = {sx ? a.x : b.x, sy ? a.y : b.y, sz ? a.z : b.z};
        tag = prim | ((int)sx | (int)sy << 1 | (int)sz << 2) << kLeafPrimBits;
    }
    L.lo = make_float4(lo.x, lo.y, lo.z, __int_as_float(tag));
    L.hi = make_float4(hi.x, hi.y, hi.z, L.hi.w);
}

__global__ __launch_bounds__(256) void internal_box_kernel(const int* levelNodes, int begin, int end,
                                                           BvhNode* nodes, const int* rightChild, const int* gate) {
    if (gate && !*gate) return;
    const int k = begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= end) return;
    const int i = levelNodes[k];
    BvhNode& N = nodes[i];
    const BvhNode A = nodes[__float_as_int(N.lo.w)];
    const BvhNode B = nodes[rightChild[i]];
    N.lo = make_float4(fminf(A.lo.x, B.lo.x), fminf(A.lo.y, B.lo.y), fminf(A.lo.z, B.lo.z), N.lo.w);
    N.hi = make_float4(fmaxf(A.hi.x, B.hi.x), fmaxf(A.hi.y, B.hi.y), fmaxf(A.hi.z, B.hi.z), N.hi.w);
}

// Several consecutive refit levels of few nodes each in one block: the levels run in order with a
// block barrier between them, so a level reads children finished by the previous ones (levels
// before the segment were finished by earlier launches).
__global__ __launch_bounds__(1024) void internal_box_levels_kernel(const int* levelNodes, const int* levelBegin,
                                                                   int l0, int l1, BvhNode* nodes,
                                                                   const int* rightChild, const int* gate) {
    if (gate && !*gate) return;  // uniform over the block: no barrier is left waiting
    for (int l = l0; l < l1; ++l) {
        const int b = levelBegin[l], e = levelBegin[l + 1];
        for (int k = b + (int)threadIdx.x; k < e; k += (int)blockDim.x) {
            const int i = levelNodes[k];
            BvhNode& N = nodes[i];
            const BvhNode A = nodes[__float_as_int(N.lo.w)];
            const BvhNode B = nodes[rightChild[i]];
            N.lo = make_float4(fminf(A.lo.x, B.lo.x), fminf(A.lo.y, B.lo.y), fminf(A.lo.z, B.lo.z), N.lo.w);
            N.hi = make_float4(fmaxf(A.hi.x, B.hi.x), fmaxf(A.hi.y, B.hi.y), fmaxf(A.hi.z, B.hi.z), N.hi.w);
        }
        __syncthreads();
    }
}

// Mesh of every node of a face tree (the faces' mesh ids, w of int4 faces), -1 for nodes whose
// faces belong to several meshes: leaves, then the internal levels deepest first as the refit.
__global__ __launch_bounds__(256) void node_mesh_leaf_kernel(const int4* faces, const BvhNode* nodes, int n,
                                                             int* mesh) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    mesh[n - 1 + k] = faces[__float_as_int(nodes[n - 1 + k].lo.w)].w;
}

__device__ __forceinline__ int merge_mesh(int a, int b) { return a == b ? a : -1; }

__global__ __launch_bounds__(256) void node_mesh_internal_kernel(const int* levelNodes, int begin, int end,
                                                                 const BvhNode* nodes, const int* rightChild,
                                                                 int* mesh) {
    const int k = begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= end) return;
    const int i = levelNodes[k];
    mesh[i] = merge_mesh(mesh[__float_as_int(nodes[i].lo.w)], mesh[rightChild[i]]);
}

__global__ __launch_bounds__(1024) void node_mesh_levels_kernel(const int* levelNodes, const int* levelBegin, int l0,
                                                                int l1, const BvhNode* nodes, const int* rightChild,
                                                                int* mesh) {
    for (int l = l0; l < l1; ++l) {
        const int b = levelBegin[l], e = levelBegin[l + 1];
        for (int k = b + (int)threadIdx.x; k < e; k += (int)blockDim.x) {
            const int i = levelNodes[k];
            mesh[i] = merge_mesh(mesh[__float_as_int(nodes[i].lo.w)], mesh[rightChild[i]]);
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void gated_fill_int_kernel(int* p, int64_t n, int value, const int* gate) {
    if (gate && !*gate) return;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = value;
}

// Whether this iteration's detection pass runs: always at the first iteration of a substep,
// afterwards when the previous iteration truncated a vertex to its bound (the re-detection flag,
// VBDClothPhysics.cpp:430-446); counts the passes that run.
__global__ void detection_gate_kernel(const int* flag, int* gate, int* passes, int first) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const int g = (first || *flag) ? 1 : 0;
        *gate = g;
        if (g) ++*passes;
    }
}

// positionsAtPrevCollisionDetection = x and the re-detection flag cleared (a gated pass only
// when its gate is set)
__global__ __launch_bounds__(256) void snapshot_kernel(const float4* pos, float4* atCD, int n, int* flag,
                                                       const int* gate) {
    if (gate && !*gate) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) *flag = 0;
    for (int k = i; k < n; k += gridDim.x * blockDim.x) atCD[k] = pos[k];
}

__global__ __launch_bounds__(256) void fill_int_kernel(int* p, int64_t n, int value) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = value;
}

// ---- workload generator: x = rest + step * drift + sigma * N(0, 1) --------------------------

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ float gaussian(uint64_t seed, int step, int64_t idx) {
    const uint64_t z = mix64(seed ^ mix64((uint64_t)step * 0x9E3779B97F4A7C15ull + (uint64_t)idx));
    const float u1 = ((float)(z >> 40) + 0.5f) * (1.0f / 16777216.0f);
    const float u2 = ((float)((z >> 16) & 0xFFFFFFull) + 0.5f) * (1.0f / 16777216.0f);
    return sqrtf(-2.f * logf(u1)) * cosf(6.28318530718f * u2);
}

__global__ __launch_bounds__(256) void perturb_kernel(const float4* rest, float4* pos, const int* vMesh,
                                                      const float* drift, int nv, uint64_t seed, int step,
                                                      float sigma) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    const float4 r = rest[v];
    const int m = vMesh[v];
    const float fs = (float)step;
    pos[v] = make_float4(r.x + fs * drift[3 * m] + sigma * gaussian(seed, step, 3 * (int64_t)v),
                         r.y + fs * drift[3 * m + 1] + sigma * gaussian(seed, step, 3 * (int64_t)v + 1),
                         r.z + fs * drift[3 * m + 2] + sigma * gaussian(seed, step, 3 * (int64_t)v + 2), 0.f);
}

// Face back-links (the FV lists): for every face, the vertices whose stored VF list holds it,
// ascending, at most `cap` of them (the lowest ids). Built from the stored records only:
// count per face, exclusive scan, scatter, then an in-place selection of each face's lowest ids.
__global__ __launch_bounds__(256) void fv_count_kernel(const int* stored, const VFContactRec* rec, int nv, int cap,
                                                       int* counts, const int* gate) {
    if (gate && !*gate) return;
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= (int64_t)nv * cap) return;
    const int v = (int)(i / cap), k = (int)(i % cap);
    if (k < stored[v]) atomicAdd(&counts[rec[i].face], 1);
}

__global__ __launch_bounds__(256) void fv_scatter_kernel(const int* stored, const VFContactRec* rec, int nv, int cap,
                                                         const int* offsets, int* fill, int* verts, const int* gate) {
    if (gate && !*gate) return;
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= (int64_t)nv * cap) return;
    const int v = (int)(i / cap), k = (int)(i % cap);
    if (k >= stored[v]) return;
    const int f = rec[i].face;
    verts[offsets[f] + atomicAdd(&fill[f], 1)] = v;
}

// selection of the `cap` lowest vertex ids of each face into its first slots, ascending (the
// scatter order is arbitrary; this makes the lists, and every sum over them, deterministic)
__global__ __launch_bounds__(256) void fv_select_kernel(const int* offsets, int nf, int cap, int* verts,
                                                        int2* range, const int* gate) {
    if (gate && !*gate) return;
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    const int b = offsets[f], e = offsets[f + 1];
    const int m = min(e - b, cap);
    for (int i = b; i < b + m; ++i) {
        int best = i;
        for (int j = i + 1; j < e; ++j)
            if (verts[j] < verts[best]) best = j;
        const int t = verts[i];
        verts[i] = verts[best];
        verts[best] = t;
    }
    range[f] = make_int2(b, b + m);
}

// ---- penetration contact of tet meshes (the reference tet solver's DCD and CCD) -------------
//
// The same float operations in the same order as the CPU restatement (oracle tet_collision.h),
// so detections match it bit for bit: DCD finds the lowest tet containing a surface vertex
// (strictly, |6V| >= CMP_EPSILON; CollisionGeometry.h:56-88), then the closest face of that
// mesh's surface passing the self-vertex, feasible-region and tet-walk checks
// (DiscreteCollisionDetector.cpp:97-300,697-815; TetMeshFEM.cpp:503-735); CCD finds the face a
// surface vertex's path meets first (ContinuousCollisionDetector.cpp:67-190; CCDSolver.h).

constexpr int kPenCounts = 8;   // penetration counters (5 used)
constexpr int kNearestRounds = 8;  // nearest-face rounds of the DCD before its single-traversal search
constexpr int kPenStack = 64;   // candidate exit faces of a walk
constexpr int kPenRecent = 32;  // recently entered tets checked for loops
// local face k of a tet is opposite its vertex k, vertices in this order (TetMeshFEM.cpp:22)
__constant__ int kTetFace[4][3] = {{1, 2, 3}, {2, 0, 3}, {0, 1, 3}, {1, 0, 2}};

struct PenView {
    const float4* pos;
    const float4* prev;      // positions at the start of the step (CCD)
    const int4* faces;       // v0, v1, v2, mesh (detector faces = the meshes' boundary surfaces)
    const int4* tets;        // scene vertex ids
    const int* tetMesh;
    const int4* tetNbr;      // per local face k, -1 on the surface
    const int2* faceOwner;   // (tet, local face)
    const int4* faceNbr;     // neighbour face across (v0 v1), (v1 v2), (v2 v0)
    const int* vNbrBegin;    // surface-edge neighbours (CSR over scene vertices)
    const int* vNbr;
    const int* surfVerts;
    int nSurf;
    const int* vMesh;
    const BvhNode* tNodes;   // tet tree
    int tN;
    const BvhNode* fNodes;   // face tree (current positions)
    const int* fMesh;        // per face-tree node: the mesh of its faces, -1 for several
    int fN;
    const BvhNode* sNodes;   // swept-face tree (prev and current positions)
    int sN;
    PenetrationParams P;
    int* stored;             // per vertex: 1 with a hit face (the VF record slot's count)
    VFContactRec* rec;       // per vertex, cap records: rec[v * cap].face = hit face
    int cap;
    uint8_t* flags;          // per vertex: 1 a tet contained it, 2 the hit came from CCD
    int* counts;             // [0] DCD hits, [1] penetrated without a face, [2] CCD hits, [3] walk overflows,
                             // [4] contained vertices listed for the closest-face pass, [5] closest-face
                             // searches past the nearest rounds
    int* inTet;              // per contained vertex: the lowest tet containing it
    int* list;               // contained vertices (order of discovery), counts[4] of them
    int* error;
};

__device__ __forceinline__ V3 vcross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float vnorm(V3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ V3 vnormalize(V3 a) { return (1.f / vnorm(a)) * a; }

__device__ bool point_in_tet(V3 p, V3 a, V3 b, V3 c, V3 d) {
    const float vol = dot(b - a, vcross(c - a, d - a));
    if (fabsf(vol) < kCmpEpsilon) return false;
    const V3 vs[4] = {a, b, c, d};
    for (int i = 0; i < 4; ++i) {
        const V3 v1 = vs[kTetFace[i][0]], v2 = vs[kTetFace[i][1]], v3 = vs[kTetFace[i][2]];
        if (dot(p - v1, vcross(v2 - v1, v3 - v2)) * vol >= 0.f) return false;
    }
    return true;
}

__device__ V3 face_normal(const PenView& S, int f) {
    const int4 fv = S.faces[f];
    const V3 a = ld(S.pos, fv.x), b = ld(S.pos, fv.y), c = ld(S.pos, fv.z);
    return vnormalize(vcross(b - a, c - a));
}

// DiscreteCollisionDetector.cpp:697-815 (edge 0 = v0 v1, 1 = v1 v2, 2 = v2 v0; no absolute relaxation)
__device__ bool pen_feasible(const PenView& S, V3 p, int face, int type) {
    if (type == AT_INTERIOR) return true;
    const float eps = S.P.feasibleEps;
    const int4 fv = S.faces[face];
    if (type >= AT_AB) {
        const int e = type == AT_AB ? 0 : (type == AT_BC ? 1 : 2);
        const int i1 = e == 0 ? fv.x : (e == 1 ? fv.y : fv.z);
        const int i2 = e == 0 ? fv.y : (e == 1 ? fv.z : fv.x);
        const int4 nbs = S.faceNbr[face];
        const int nb = e == 0 ? nbs.x : (e == 1 ? nbs.y : nbs.z);
        if (nb < 0) return true;
        const V3 v1 = ld(S.pos, i1), v2 = ld(S.pos, i2);
        const V3 n1 = -1.f * face_normal(S, face), n2 = -1.f * face_normal(S, nb);
        const V3 AP = p - v1, BP = p - v2, AB = v2 - v1, BA = -1.f * AB;
        const float relaxed = -dot(AB, AB) * eps;
        if (dot(AP, AB) < relaxed) return false;
        if (dot(BP, BA) < relaxed) return false;
        if (dot(AP, vcross(n1, AB)) < relaxed) return false;
        if (dot(AP, vcross(n2, BA)) < relaxed) return false;
        return true;
    }
    const int vid = type == AT_A ? fv.x : (type == AT_B ? fv.y : fv.z);
    const V3 A = ld(S.pos, vid), AP = p - A;
    for (int k = S.vNbrBegin[vid]; k < S.vNbrBegin[vid + 1]; ++k) {
        const V3 BA = A - ld(S.pos, S.vNbr[k]);
        if (dot(AP, BA) < -dot(BA, BA) * eps) return false;
    }
    return true;
}

__device__ __forceinline__ float cross2(float ax, float ay, float bx, float by) { return ax * by - ay * bx; }

// The ray walk of the closest-point check (TetMeshFEM.cpp:503-735): from the surface face
// (startTet, startLocal) along o + s dir through face-adjacent tets whose exit faces the ray may
// cross (2D projection on a Duff et al. 2017 basis, epsilon rayEps), until the target tet is
// entered (true), the surface is left, the crossing checked every 5 tets lies beyond maxDis, the
// candidates run out or the candidate stack overflows (counted).
__device__ bool pen_walk(const PenView& S, V3 o, V3 dir, float maxDis, int startTet, int startLocal, int target) {
    if (startTet == target) return true;
    const float sg = copysignf(1.f, dir.z);
    const float ca = -1.f / (sg + dir.z);
    const float cb = dir.x * dir.y * ca;
    const V3 b1 = {1.f + sg * dir.x * dir.x * ca, sg * cb, -sg * dir.x};
    const V3 b2 = {cb, sg + dir.y * dir.y * ca, -dir.y};
    const float eps = S.P.rayEps;
    float qx[4], qy[4];
    int ids[4];
    auto project = [&](int t, int in) {
        const int4 tv = S.tets[t];
        const int tvs[4] = {tv.x, tv.y, tv.z, tv.w};
        for (int i = 0; i < 3; ++i) ids[i] = tvs[kTetFace[in][i]];
        ids[3] = tvs[in];
        for (int i = 0; i < 4; ++i) {
            const V3 d = ld(S.pos, ids[i]) - o;
            qx[i] = dot(b1, d);
            qy[i] = dot(b2, d);
        }
    };
    int ex[3];
    auto exits = [&]() {
        const float area = cross2(qx[1] - qx[0], qy[1] - qy[0], qx[2] - qx[0], qy[2] - qy[0]);
        const float s = copysignf(1.f, area);
        const float d0 = cross2(qx[3], qy[3], qx[0], qy[0]) * s;
        const float d1 = cross2(qx[3], qy[3], qx[1], qy[1]) * s;
        const float d2 = cross2(qx[3], qy[3], qx[2], qy[2]) * s;
        ex[0] = d1 >= -eps && d2 <= eps;
        ex[1] = d2 >= -eps && d0 <= eps;
        ex[2] = d0 >= -eps && d1 <= eps;
    };
    int stackTet[kPenStack], stackFace[kPenStack], ns = 0;
    int recent[kPenRecent], nrec = 0, head = 0;
    project(startTet, startLocal);
    exits();
    recent[head] = startTet;
    head = 1;
    nrec = 1;
    for (int i = 0; i < 3; ++i)
        if (ex[i]) {
            stackTet[ns] = startTet;
            stackFace[ns++] = kTetFace[startLocal][i];
        }
    int steps = 0;
    while (ns > 0) {
        --ns;
        const int prevT = stackTet[ns], face = stackFace[ns];
        const int4 pn = S.tetNbr[prevT];
        const int cur = face == 0 ? pn.x : (face == 1 ? pn.y : (face == 2 ? pn.z : pn.w));
        if (cur == target) return true;
        if (cur < 0) return false;
        bool seen = false;
        for (int i = 0; i < nrec; ++i) seen = seen || recent[i] == cur;
        if (seen) continue;
        recent[head] = cur;
        head = (head + 1) % kPenRecent;
        nrec = min(nrec + 1, kPenRecent);
        const int4 ptv = S.tets[prevT];
        const int ptvs[4] = {ptv.x, ptv.y, ptv.z, ptv.w};
        const int fa = ptvs[kTetFace[face][0]], fb = ptvs[kTetFace[face][1]], fc = ptvs[kTetFace[face][2]];
        const int4 ctv = S.tets[cur];
        const int ctvs[4] = {ctv.x, ctv.y, ctv.z, ctv.w};
        int in = 3;
        for (int i = 0; i < 4; ++i)
            if (ctvs[i] != fa && ctvs[i] != fb && ctvs[i] != fc) {
                in = i;
                break;
            }
        project(cur, in);
        exits();
        const int4 cn = S.tetNbr[cur];
        const int cns[4] = {cn.x, cn.y, cn.z, cn.w};
        for (int i = 0; i < 3; ++i) {
            if (!ex[i]) continue;
            const int exitLocal = kTetFace[in][i];
            if (cns[exitLocal] == target) return true;
            if (ns == kPenStack) {
                atomicAdd(S.counts + 3, 1);
                return false;
            }
            stackTet[ns] = cur;
            stackFace[ns++] = exitLocal;
            if (steps % 5 == 0 && maxDis > 0.f) {
                // crossing with exit face i: the origin's barycentrics in its projection
                const int e0 = i == 0 ? 1 : 0, e1 = i == 2 ? 1 : 2;
                const float v0x = qx[e1] - qx[e0], v0y = qy[e1] - qy[e0], v1x = qx[3] - qx[e0], v1y = qy[3] - qy[e0];
                const float v2x = -qx[e0], v2y = -qy[e0];
                const float den = v0x * v1y - v1x * v0y;
                const float l1 = (v2x * v1y - v1x * v2y) / den, l2 = (v0x * v2y - v2x * v0y) / den;
                const float l0 = 1.f - l1 - l2;
                const V3 hit = (l0 * ld(S.pos, ids[e0]) + l1 * ld(S.pos, ids[e1])) + l2 * ld(S.pos, ids[3]);
                const V3 d = hit - o;
                if (dot(d, d) > maxDis * maxDis) return false;
            }
        }
        ++steps;
    }
    return false;
}

__device__ __forceinline__ float box_dist2(V3 p, const BvhNode& n) {
    const float dx = fmaxf(fmaxf(n.lo.x - p.x, 0.f), p.x - n.hi.x);
    const float dy = fmaxf(fmaxf(n.lo.y - p.y, 0.f), p.y - n.hi.y);
    const float dz = fmaxf(fmaxf(n.lo.z - p.z, 0.f), p.z - n.hi.z);
    return dx * dx + dy * dy + dz * dz;
}

// DCD pass 1, one thread per surface vertex: the lowest tet containing the vertex (of any mesh;
// of its own mesh only with self collision, and no tet of its own). Contained vertices are
// listed for pass 2, so that its waves hold only vertices with a closest-face search to do.
__global__ __launch_bounds__(64) void pen_contain_kernel(PenView S) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S.nSurf) return;
    const int v = S.surfVerts[i];
    S.stored[v] = 0;
    S.flags[v] = 0;
    const V3 p = ld(S.pos, v);
    if (!S.P.dcd || !(p.y < S.P.offHeight)) return;
    const int qm = S.vMesh[v];
    int ht = -1;
    const int firstLeaf = S.tN - 1;
    int node = S.tN > 0 ? 0 : -1;
    for (int visits = 0; node >= 0; ++visits) {
        if (visits > 2 * S.tN) {
            atomicOr(S.error, 1);
            break;
        }
        const BvhNode nd = S.tNodes[node];
        const int skip = __float_as_int(nd.hi.w);
        if (!(p.x >= nd.lo.x && p.x <= nd.hi.x && p.y >= nd.lo.y && p.y <= nd.hi.y && p.z >= nd.lo.z &&
              p.z <= nd.hi.z)) {
            node = skip;
            continue;
        }
        if (node < firstLeaf) {
            node = __float_as_int(nd.lo.w);
            continue;
        }
        node = skip;
        const int t = __float_as_int(nd.lo.w);
        if (ht >= 0 && t > ht) continue;
        const int m = S.tetMesh[t];
        if (m == qm && !S.P.self) continue;
        const int4 tv = S.tets[t];
        if (m == qm && (tv.x == v || tv.y == v || tv.z == v || tv.w == v)) continue;
        if (point_in_tet(p, ld(S.pos, tv.x), ld(S.pos, tv.y), ld(S.pos, tv.z), ld(S.pos, tv.w))) ht = t;
    }
    if (ht < 0) return;
    S.flags[v] = 1;
    S.inTet[v] = ht;
    S.list[atomicAdd(S.counts + 4, 1)] = v;
}

// Whether surface face f (closest point cp of region `type`) is a valid hit of vertex v (at p,
// inside tet ht of mesh hm): the feasible-region check and the ray walk from the face to the
// query point (DiscreteCollisionDetector.cpp:97-300).
__device__ bool pen_face_valid(const PenView& S, int qm, int hm, int ht, V3 p, int f, int type, V3 cp) {
    if (S.P.feasible && !pen_feasible(S, p, f, type)) return false;
    if (S.P.traverse && (hm == qm || S.P.traverseNonSelf)) {
        const float c = S.P.centerShift;
        const int4 fv = S.faces[f];
        const V3 a = ld(S.pos, fv.x), b = ld(S.pos, fv.y), cc = ld(S.pos, fv.z);
        const V3 from = type == AT_INTERIOR ? cp : (1.f - c) * cp + (c / 3.f) * ((a + b) + cc);
        V3 to = p;
        if (S.P.shiftToCenter) {
            const int4 tv = S.tets[ht];
            const V3 cen = 0.25f * (((ld(S.pos, tv.x) + ld(S.pos, tv.y)) + ld(S.pos, tv.z)) + ld(S.pos, tv.w));
            to = (1.f - c) * p + c * cen;
        }
        V3 dir = to - from;
        const float len = vnorm(dir);
        dir = (1.f / len) * dir;
        const float maxDis = S.P.stopAfterPassing ? S.P.maxSearchMult * len : -1.f;
        const int2 own = S.faceOwner[f];
        if (!pen_walk(S, from, dir, maxDis, own.x, own.y, ht)) return false;
    }
    return true;
}

// Squared distance from p to the farthest point of a node's box.
__device__ __forceinline__ float box_far2(V3 p, const BvhNode& n) {
    const float dx = fmaxf(p.x - n.lo.x, n.hi.x - p.x);
    const float dy = fmaxf(p.y - n.lo.y, n.hi.y - p.y);
    const float dz = fmaxf(p.z - n.lo.z, n.hi.z - p.z);
    return dx * dx + dy * dy + dz * dz;
}

// DCD pass 2, one thread per contained vertex: the closest face of the containing tet's mesh
// that passes the checks, lower face id on equal distance (DiscreteCollisionDetector.cpp:
// 514-560). The faces are first taken in ascending (distance, id), the first one passing the
// checks being the answer: each round is a nearest-face query on the face tree over the faces
// after the last rejected one, pruning subtrees of other meshes, boxes surely farther than the
// round's best and boxes surely closer than the rejected face. The nearest face passes for
// nearly every vertex, so the costly checks run about once per vertex. After kNearestRounds
// rejected faces the rest is searched in one traversal that checks every face closer than the
// best valid one so far (each round repeats a traversal, which costs more than the extra checks
// once many faces are rejected). Both orders give the same answer.
__global__ __launch_bounds__(64) void pen_closest_kernel(PenView S) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S.counts[4]) return;
    const int v = S.list[i];
    const int ht = S.inTet[v];
    const V3 p = ld(S.pos, v);
    const int qm = S.vMesh[v];
    const int hm = S.tetMesh[ht];
    float lastD = -1.f;
    int lastF = -1;
    int hit = -1;
    bool exhausted = false;
    const int firstLeaf = S.fN - 1;
    for (int round = 0; round < kNearestRounds && hit < 0 && !exhausted; ++round) {
        float best = FLT_MAX;
        int bestF = -1, bestType = 0;
        V3 bestCp = p;
        const float near2 = lastD * lastD;
        int node = S.fN > 0 ? 0 : -1;
        for (int visits = 0; node >= 0; ++visits) {
            if (visits > 2 * S.fN) {
                atomicOr(S.error, 1);
                break;
            }
            const BvhNode nd = S.fNodes[node];
            const int skip = __float_as_int(nd.hi.w);
            // prune subtrees of other meshes' faces, boxes surely farther than the round's best and
            // boxes surely closer than the last rejected face (margins for rounding)
            const int nm = S.fMesh[node];
            if ((nm >= 0 && nm != hm) || (bestF >= 0 && box_dist2(p, nd) > best * best * 1.0001f + 1e-12f) ||
                (lastF >= 0 && box_far2(p, nd) * 1.0001f + 1e-12f < near2)) {
                node = skip;
                continue;
            }
            if (node < firstLeaf) {
                node = __float_as_int(nd.lo.w);
                continue;
            }
            node = skip;
            const int f = __float_as_int(nd.lo.w);
            const int4 fv = S.faces[f];
            if (fv.w != hm) continue;
            int type;
            const V3 cp = closest_point_triangle(p, ld(S.pos, fv.x), ld(S.pos, fv.y), ld(S.pos, fv.z), type);
            if (hm == qm && type <= AT_C && (type == AT_A ? fv.x : (type == AT_B ? fv.y : fv.z)) == v) continue;
            const float d = vnorm(p - cp);
            if (!(d > lastD || (d == lastD && f > lastF))) continue;  // taken in an earlier round
            if (bestF >= 0 && !(d < best || (d == best && f < bestF))) continue;
            best = d;
            bestF = f;
            bestType = type;
            bestCp = cp;
        }
        if (bestF < 0) {
            exhausted = true;  // no face left
            break;
        }
        if (pen_face_valid(S, qm, hm, ht, p, bestF, bestType, bestCp)) hit = bestF;
        lastD = best;
        lastF = bestF;
    }
    if (hit < 0 && !exhausted) {
        // many near faces rejected (a vertex deep in the other body): one traversal over the faces
        // after the last rejected one, checking each face closer than the best valid face so far
        atomicAdd(S.counts + 5, 1);
        float best = FLT_MAX;
        const float near2 = lastD * lastD;
        int node = S.fN > 0 ? 0 : -1;
        for (int visits = 0; node >= 0; ++visits) {
            if (visits > 2 * S.fN) {
                atomicOr(S.error, 1);
                break;
            }
            const BvhNode nd = S.fNodes[node];
            const int skip = __float_as_int(nd.hi.w);
            const int nm = S.fMesh[node];
            if ((nm >= 0 && nm != hm) || (hit >= 0 && box_dist2(p, nd) > best * best * 1.0001f + 1e-12f) ||
                box_far2(p, nd) * 1.0001f + 1e-12f < near2) {
                node = skip;
                continue;
            }
            if (node < firstLeaf) {
                node = __float_as_int(nd.lo.w);
                continue;
            }
            node = skip;
            const int f = __float_as_int(nd.lo.w);
            const int4 fv = S.faces[f];
            if (fv.w != hm) continue;
            int type;
            const V3 cp = closest_point_triangle(p, ld(S.pos, fv.x), ld(S.pos, fv.y), ld(S.pos, fv.z), type);
            if (hm == qm && type <= AT_C && (type == AT_A ? fv.x : (type == AT_B ? fv.y : fv.z)) == v) continue;
            const float d = vnorm(p - cp);
            if (!(d > lastD || (d == lastD && f > lastF))) continue;
            if (hit >= 0 && !(d < best || (d == best && f < hit))) continue;
            if (!pen_face_valid(S, qm, hm, ht, p, f, type, cp)) continue;
            best = d;
            hit = f;
        }
    }
    if (hit < 0) {
        atomicAdd(S.counts + 1, 1);
        return;
    }
    S.stored[v] = 1;
    S.rec[(int64_t)v * S.cap].face = hit;
    atomicAdd(S.counts, 1);
}

__device__ __forceinline__ float poly3(const float c[4], float x) { return ((c[3] * x + c[2]) * x + c[1]) * x + c[0]; }
__device__ __forceinline__ bool diff_sign(float a, float b) { return (a < 0.f) != (b < 0.f); }

// cyPolynomial.h:383-460 (see the oracle): Newton from the midpoint, then a safeguarded bracket
__device__ float bracketed_root(const float c[4], float x0, float x1, float y0, float err) {
    auto dpoly = [&](float x) { return (3.f * c[3] * x + 2.f * c[2]) * x + c[1]; };
    const float ep2 = 2.f * err;
    float xr = (x0 + x1) / 2.f;
    if (x1 - x0 <= ep2) return xr;
    const float xr0 = xr;
    for (int it = 0; it < 16; ++it) {
        float xn = xr - poly3(c, xr) / dpoly(xr);
        xn = fminf(fmaxf(xn, x0), x1);
        if (fabsf(xr - xn) <= err) return xn;
        xr = xn;
    }
    if (!isfinite(xr)) xr = xr0;
    float yr = poly3(c, xr);
    float xb0 = x0, xb1 = x1;
    for (int guard = 0; guard < 200; ++guard) {
        if (diff_sign(y0, yr)) xb1 = xr;
        else xb0 = xr;
        const float xn = xr - yr / dpoly(xr);
        if (xn > xb0 && xn < xb1) {
            const float step = fabsf(xr - xn);
            xr = xn;
            if (step > err) yr = poly3(c, xr);
            else break;
        } else {
            xr = (xb0 + xb1) / 2.f;
            if (xr == xb0 || xr == xb1 || xb1 - xb0 <= ep2) break;
            yr = poly3(c, xr);
        }
    }
    return xr;
}

__device__ int cubic_roots01(const float c[4], float roots[3], float err) {
    float pts[4];
    int np = 0;
    pts[np++] = 0.f;
    const float a = 3.f * c[3], b2 = c[2], cc = c[1];
    const float disc = b2 * b2 - a * cc;
    if (a != 0.f && disc > 0.f) {
        const float sq = sqrtf(disc);
        const float q = -(b2 + copysignf(sq, b2));
        float r0 = q / a, r1 = cc / q;
        if (r0 > r1) {
            const float t = r0;
            r0 = r1;
            r1 = t;
        }
        if (r0 > 0.f && r0 < 1.f) pts[np++] = r0;
        if (r1 > 0.f && r1 < 1.f && r1 != r0) pts[np++] = r1;
    } else if (a == 0.f && b2 != 0.f) {
        const float r = -cc / (2.f * b2);
        if (r > 0.f && r < 1.f) pts[np++] = r;
    }
    pts[np++] = 1.f;
    int n = 0;
    float yprev = poly3(c, pts[0]);
    for (int i = 1; i < np; ++i) {
        const float y = poly3(c, pts[i]);
        if (diff_sign(yprev, y)) roots[n++] = bracketed_root(c, pts[i - 1], pts[i], yprev, err);
        yprev = y;
    }
    return n;
}

// CCDSolver.h:104-202: the earliest t in [0, 1] at which the moving point lies in the moving
// triangle's plane and inside it
__device__ bool ccd_point_triangle(const V3 f0[3], const V3 f1[3], V3 p0, V3 p1, float& tt) {
    const V3 x0[4] = {f0[0], f0[1], f0[2], p0};
    const V3 x1[4] = {f1[0], f1[1], f1[2], p1};
    const V3 dx0 = x1[0] - x0[0], dx1 = x1[1] - x0[1], dx2 = x1[2] - x0[2], dx3 = x1[3] - x0[3];
    const V3 e1 = x0[1] - x0[0], e2 = x0[2] - x0[0], e3 = x0[3] - x0[0];
    const V3 de0 = dx3 - dx0, de1 = dx1 - dx0, de2 = dx2 - dx0;
    const V3 e1xe2 = vcross(e1, e2), de1xde2 = vcross(de1, de2), mixed = vcross(e1, de2) + vcross(de1, e2);
    const float c[4] = {dot(e3, e1xe2), dot(de0, e1xe2) + dot(e3, mixed), dot(e3, de1xde2) + dot(de0, mixed),
                        dot(de0, de1xde2)};
    float ts[3];
    const int n = cubic_roots01(c, ts, 6e-4f);
    for (int i = 0; i < n; ++i) {
        const float t = ts[i];
        V3 y3[4];
        for (int k = 0; k < 4; ++k) y3[k] = x0[k] + t * (x1[k] - x0[k]);
        const V3 nrm = vcross(y3[1] - y3[0], y3[2] - y3[0]);
        const float ax = fabsf(nrm.x), ay = fabsf(nrm.y), az = fabsf(nrm.z);
        const int drop = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
        float u[4], w[4];
        for (int k = 0; k < 4; ++k) {
            u[k] = drop == 0 ? y3[k].y : y3[k].x;
            w[k] = drop == 2 ? y3[k].y : y3[k].z;
        }
        for (int k = 0; k < 3; ++k) {
            u[k] -= u[3];
            w[k] -= w[3];
        }
        const bool s01 = signbit(cross2(u[0], w[0], u[1], w[1]));
        const bool s12 = signbit(cross2(u[1], w[1], u[2], w[2]));
        const bool s20 = signbit(cross2(u[2], w[2], u[0], w[0]));
        if (s01 == s12 && s01 == s20) {
            tt = t;
            return true;
        }
    }
    return false;
}

// CCD of every surface vertex that no tet contained, on its path from prev to pos
__global__ __launch_bounds__(256) void pen_ccd_kernel(PenView S) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S.nSurf) return;
    const int v = S.surfVerts[i];
    if (S.flags[v] & 1) return;
    const V3 p1 = ld(S.pos, v);
    if (!(p1.y < S.P.offHeight)) return;
    const V3 p0 = ld(S.prev, v);
    const int qm = S.vMesh[v];
    const float len = vnorm(p1 - p0);
    const V3 mid = 0.5f * (p0 + p1);
    // boxes within the half path length of the midpoint (with a margin for the root's tolerance)
    const float r = 0.5f * len + 1e-3f * len + 1e-6f;
    const float r2 = r * r;
    float best = 0.f;
    int bestF = -1;
    const int firstLeaf = S.sN - 1;
    int node = S.sN > 0 ? 0 : -1;
    for (int visits = 0; node >= 0; ++visits) {
        if (visits > 2 * S.sN) {
            atomicOr(S.error, 1);
            break;
        }
        const BvhNode nd = S.sNodes[node];
        const int skip = __float_as_int(nd.hi.w);
        if (box_dist2(mid, nd) > r2) {
            node = skip;
            continue;
        }
        if (node < firstLeaf) {
            node = __float_as_int(nd.lo.w);
            continue;
        }
        node = skip;
        const int f = __float_as_int(nd.lo.w);
        const int4 fv = S.faces[f];
        if (fv.w == qm && (fv.x == v || fv.y == v || fv.z == v)) continue;
        const V3 a0[3] = {ld(S.prev, fv.x), ld(S.prev, fv.y), ld(S.prev, fv.z)};
        const V3 a1[3] = {ld(S.pos, fv.x), ld(S.pos, fv.y), ld(S.pos, fv.z)};
        float tt;
        if (!ccd_point_triangle(a0, a1, p0, p1, tt)) continue;
        const float depth = tt * len;
        if (bestF < 0 || depth < best || (depth == best && f < bestF)) {
            best = depth;
            bestF = f;
        }
    }
    if (bestF < 0) return;
    S.stored[v] = 1;
    S.rec[(int64_t)v * S.cap].face = bestF;
    S.flags[v] = 2;
    atomicAdd(S.counts + 2, 1);
}

// boxes of tets (kind 3 primitives): lo, hi per tet
__global__ __launch_bounds__(256) void tet_box_kernel(const float4* pos, const int4* tets, int n, float4* boxes) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int4 tv = tets[t];
    const V3 a = ld(pos, tv.x), b = ld(pos, tv.y), c = ld(pos, tv.z), d = ld(pos, tv.w);
    boxes[2 * t] = make_float4(fminf(fminf(a.x, b.x), fminf(c.x, d.x)), fminf(fminf(a.y, b.y), fminf(c.y, d.y)),
                               fminf(fminf(a.z, b.z), fminf(c.z, d.z)), 0.f);
    boxes[2 * t + 1] = make_float4(fmaxf(fmaxf(a.x, b.x), fmaxf(c.x, d.x)), fmaxf(fmaxf(a.y, b.y), fmaxf(c.y, d.y)),
                                   fmaxf(fmaxf(a.z, b.z), fmaxf(c.z, d.z)), 0.f);
}

// boxes of faces over their path from prev to pos (movingFaceBoundsFunc, ContinuousCollisionDetector.cpp:19-65)
__global__ __launch_bounds__(256) void swept_box_kernel(const float4* pos, const float4* prev, const int4* faces, int n,
                                                        float4* boxes) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n) return;
    const int4 fv = faces[f];
    const V3 q[6] = {ld(pos, fv.x), ld(pos, fv.y), ld(pos, fv.z), ld(prev, fv.x), ld(prev, fv.y), ld(prev, fv.z)};
    V3 lo = q[0], hi = q[0];
    for (int k = 1; k < 6; ++k) {
        lo = {fminf(lo.x, q[k].x), fminf(lo.y, q[k].y), fminf(lo.z, q[k].z)};
        hi = {fmaxf(hi.x, q[k].x), fmaxf(hi.y, q[k].y), fmaxf(hi.z, q[k].z)};
    }
    boxes[2 * f] = make_float4(lo.x, lo.y, lo.z, 0.f);
    boxes[2 * f + 1] = make_float4(hi.x, hi.y, hi.z, 0.f);
}

__global__ __launch_bounds__(256) void fill_u8_kernel(uint8_t* p, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

inline int blocks_for(int64_t n) { return (int)std::max<int64_t>(1, (n + 255) / 256); }

}  // namespace

void ContactDetector::useExternalPositions(float4* pos) {
    if (ready_) fail(GVBD_ERR_STATE, "detector already finalized");
    external_ = pos;
}

void ContactDetector::rebindPositions(float4* pos) {
    if (!external_) fail(GVBD_ERR_STATE, "the detector owns its positions");
    external_ = pos;
    pos_ = pos;
}

void ContactDetector::buildFaceLinks(int cap, const int* gate) {
    if (!ready_) fail(GVBD_ERR_STATE, "detector not finalized");
    const int64_t n = (int64_t)nv_ * vfCap_;
    int* counts = fvCounts_;         // nf + 1 (the last stays 0: the scan's total lands at offsets[nf])
    int* fill = fvCounts_ + nf_ + 1;  // nf
    // a gated-off pass leaves the counts as they are, so the scan (not gated) rewrites the same
    // offsets
    hipLaunchKernelGGL(gated_fill_int_kernel, dim3(blocks_for(2 * (int64_t)nf_ + 1)), dim3(256), 0, stream_,
                       fvCounts_, 2 * (int64_t)nf_ + 1, 0, gate);
    hipLaunchKernelGGL(fv_count_kernel, dim3(blocks_for(n)), dim3(256), 0, stream_, vfStored_, vfRec_, nv_, vfCap_,
                       counts, gate);
    size_t tb = fvScanTempBytes_;
    hip_check(rocprim::exclusive_scan(fvScanTemp_, tb, counts, fvOffsets_, 0, (size_t)nf_ + 1, rocprim::plus<int>(),
                                      stream_),
              "face link scan");
    hipLaunchKernelGGL(fv_scatter_kernel, dim3(blocks_for(n)), dim3(256), 0, stream_, vfStored_, vfRec_, nv_, vfCap_,
                       fvOffsets_, fill, fvVerts_, gate);
    hipLaunchKernelGGL(fv_select_kernel, dim3(blocks_for(nf_)), dim3(256), 0, stream_, fvOffsets_, nf_, cap, fvVerts_,
                       fvRange_, gate);
    hip_check(hipGetLastError(), "face link launch");
}

void ContactDetector::snapshotPositions(const int* gate) {
    if (!ready_) fail(GVBD_ERR_STATE, "detector not finalized");
    hipLaunchKernelGGL(snapshot_kernel, dim3(std::min(blocks_for(nv_), 4096)), dim3(256), 0, stream_, pos_, xAtCD_, nv_,
                       flag_, gate);
    hip_check(hipGetLastError(), "snapshot launch");
}

void ContactDetector::detectGated(bool first, bool rebuild, float thickness, float relaxation) {
    if (!ready_) fail(GVBD_ERR_STATE, "detector not finalized");
    if (rebuild && !first) fail(GVBD_ERR_STATE, "tree rebuilds belong to a substep's first detection");
    hipLaunchKernelGGL(detection_gate_kernel, dim3(1), dim3(64), 0, stream_, flag_, gate_, gatePasses_, first ? 1 : 0);
    if (rebuild) {
        buildBvh(fb_, 0);
        buildBvh(eb_, 1);
    } else {
        fork();
        refitBvh(fb_, 0, stream_, nullptr, gate_);
        refitBvh(eb_, 1, side_, nullptr, gate_);
        join();
    }
    query(false, gate_);
    bounds(thickness, relaxation, gate_);
    buildFaceLinks(32, gate_);
    snapshotPositions(gate_);
}

int64_t ContactDetector::gatedPasses() {
    if (!gatePasses_) return 0;
    int h = 0;
    hip_check(hipMemcpyAsync(&h, gatePasses_, sizeof(int), hipMemcpyDeviceToHost, stream_), "passes");
    hip_check(hipStreamSynchronize(stream_), "passes sync");
    return h;
}

DetectorDeviceView ContactDetector::deviceView() const {
    return DetectorDeviceView{faces_, edges_,  vfStored_, vfRec_, vfCap_, eeStored_, eeRec_,
                              eeCap_, fvRange_, fvVerts_, bounds_, xAtCD_, flag_};
}

// ---------------------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------------------

template <class T> T* ContactDetector::dalloc(size_t n) {
    void* p = nullptr;
    const size_t bytes = std::max<size_t>(n * sizeof(T), 16);
    if (hipMalloc(&p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        fail(GVBD_ERR_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes for the contact detector failed");
    }
    allocs_.push_back(p);
    bytes_ += bytes;
    return (T*)p;
}

ContactDetector::ContactDetector(int device, hipStream_t stream, float maxQueryDis, int vfCap, int eeCap)
    : device_(device), stream_(stream), r_(maxQueryDis), vfCap_(vfCap), eeCap_(eeCap) {
    require(maxQueryDis > 0.f && std::isfinite(maxQueryDis), "max_query_distance must be positive");
    require(vfCap >= 1 && vfCap <= kMaxCap && eeCap >= 1 && eeCap <= kMaxCap,
            "contact capacities must be in [1, 16]");
    meshV_.push_back(0);
    meshF_.push_back(0);
    meshE_.push_back(0);
}

ContactDetector::~ContactDetector() {
    if (side_) {
        (void)hipStreamSynchronize(side_);
        (void)hipStreamDestroy(side_);
    }
    if (forkEv_) (void)hipEventDestroy(forkEv_);
    if (joinEv_) (void)hipEventDestroy(joinEv_);
    freeAll();
}

void ContactDetector::fork() {
    hip_check(hipEventRecord(forkEv_, stream_), "fork record");
    hip_check(hipStreamWaitEvent(side_, forkEv_, 0), "fork wait");
}

void ContactDetector::join() {
    hip_check(hipEventRecord(joinEv_, side_), "join record");
    hip_check(hipStreamWaitEvent(stream_, joinEv_, 0), "join wait");
}

void ContactDetector::freeAll() {
    for (void* p : allocs_) (void)hipFree(p);
    allocs_.clear();
    bytes_ = 0;
}

int ContactDetector::addMesh(int nv, const float* pos, int nf, const int* faces, const uint8_t* queryMask) {
    if (ready_) fail(GVBD_ERR_STATE, "cannot add meshes after finalize");
    require(nv > 0 && nf >= 0 && pos != nullptr && (nf == 0 || faces != nullptr), "bad mesh arguments");
    std::vector<TriEdge> edges;
    std::vector<int32_t> faceEdges;
    std::string err;
    if (!build_tri_edges(nv, nf, faces, edges, &faceEdges, err)) fail(GVBD_ERR_INVALID, err);
    const int m = (int)meshV_.size() - 1;
    const int vo = nv_, eo = ne_;
    hostPos_.insert(hostPos_.end(), pos, pos + 3 * (size_t)nv);
    for (int i = 0; i < nv; ++i) hVMesh_.push_back(m);
    if (queryMask) masked_ = true;
    for (int i = 0; i < nv; ++i) hQuery_.push_back(queryMask ? (uint8_t)(queryMask[i] != 0) : (uint8_t)1);
    for (int f = 0; f < nf; ++f) {
        for (int k = 0; k < 3; ++k) hFaces_.push_back(vo + faces[3 * (size_t)f + k]);
        hFaces_.push_back(m);
        for (int k = 0; k < 3; ++k) hFaceEdges_.push_back(eo + faceEdges[3 * (size_t)f + k]);
        hFaceEdges_.push_back(0);
    }
    for (const TriEdge& e : edges) {
        hEdges_.push_back(vo + e.v1);
        hEdges_.push_back(vo + e.v2);
        hEdges_.push_back(e.v12Next < 0 ? -1 : vo + e.v12Next);
        hEdges_.push_back(e.v21Next < 0 ? -1 : vo + e.v21Next);
        hEMesh_.push_back(m);
    }
    nv_ += nv;
    nf_ += nf;
    ne_ += (int)edges.size();
    meshV_.push_back(nv_);
    meshF_.push_back(nf_);
    meshE_.push_back(ne_);
    return m;
}

void ContactDetector::meshOffsets(int mesh, int* v, int* f, int* e) const {
    require(mesh >= 0 && mesh < numMeshes(), "mesh index out of range");
    if (v) *v = meshV_[mesh];
    if (f) *f = meshF_[mesh];
    if (e) *e = meshE_[mesh];
}

void ContactDetector::topology(int32_t* edges, int32_t* faceEdges) const {
    if (edges) std::memcpy(edges, hEdges_.data(), hEdges_.size() * sizeof(int32_t));
    if (faceEdges)
        for (int f = 0; f < nf_; ++f)
            for (int k = 0; k < 3; ++k) faceEdges[3 * f + k] = hFaceEdges_[4 * (size_t)f + k];
}

void ContactDetector::finalize() {
    if (ready_) fail(GVBD_ERR_STATE, "detector already finalized");
    if (nv_ == 0) fail(GVBD_ERR_STATE, "detector has no meshes");
    hip_check(hipSetDevice(device_), "hipSetDevice");
    hip_check(hipStreamCreateWithFlags(&side_, hipStreamNonBlocking), "side stream");
    hip_check(hipEventCreateWithFlags(&forkEv_, hipEventDisableTiming), "event");
    hip_check(hipEventCreateWithFlags(&joinEv_, hipEventDisableTiming), "event");
    // CSR adjacency: vertex -> neighbour vertices, incident faces, incident edges
    auto csr = [&](const std::vector<std::pair<int, int>>& items, std::vector<int>& begin, std::vector<int>& out) {
        begin.assign(nv_ + 1, 0);
        for (auto& it : items) begin[it.first + 1]++;
        for (int i = 0; i < nv_; ++i) begin[i + 1] += begin[i];
        out.assign(begin[nv_], 0);
        std::vector<int> fill(begin.begin(), begin.end() - 1);
        for (auto& it : items) out[fill[it.first]++] = it.second;
    };
    std::vector<std::pair<int, int>> vv, vf, ve;
    for (int e = 0; e < ne_; ++e) {
        const int a = hEdges_[4 * (size_t)e], b = hEdges_[4 * (size_t)e + 1];
        vv.push_back({a, b});
        vv.push_back({b, a});
        ve.push_back({a, e});
        ve.push_back({b, e});
    }
    for (int f = 0; f < nf_; ++f)
        for (int k = 0; k < 3; ++k) vf.push_back({hFaces_[4 * (size_t)f + k], f});
    std::vector<int> vvB, vvN, vfB, vfF, veB, veE;
    csr(vv, vvB, vvN);
    csr(vf, vfB, vfF);
    csr(ve, veB, veE);
    auto up = [&](void* dst, const void* src, size_t bytes) {
        if (bytes) hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream_), "detector upload");
    };
    std::vector<float4> p4(nv_);
    for (int i = 0; i < nv_; ++i) p4[i] = make_float4(hostPos_[3 * i], hostPos_[3 * i + 1], hostPos_[3 * i + 2], 0.f);
    if (external_) {
        pos_ = external_;  // the owner keeps it current
    } else {
        pos_ = dalloc<float4>(nv_);
        rest_ = dalloc<float4>(nv_);
        up(pos_, p4.data(), sizeof(float4) * nv_);
        up(rest_, p4.data(), sizeof(float4) * nv_);
    }
    // face back-links and solver-side state
    fvVerts_ = dalloc<int>(std::max<size_t>((size_t)nv_ * vfCap_, 1));
    fvCounts_ = dalloc<int>(2 * (size_t)nf_ + 1);
    fvOffsets_ = dalloc<int>((size_t)nf_ + 1);
    fvRange_ = dalloc<int2>(std::max(nf_, 1));
    xAtCD_ = dalloc<float4>(nv_);
    flag_ = dalloc<int>(1);
    gate_ = dalloc<int>(1);
    gatePasses_ = dalloc<int>(1);
    hip_check(rocprim::exclusive_scan(nullptr, fvScanTempBytes_, fvCounts_, fvOffsets_, 0, (size_t)nf_ + 1,
                                      rocprim::plus<int>(), stream_),
              "scan size");
    fvScanTemp_ = dalloc<char>(std::max<size_t>(fvScanTempBytes_, 1));
    hip_check(hipMemsetAsync(flag_, 0, sizeof(int), stream_), "flag clear");
    hip_check(hipMemsetAsync(gate_, 0, sizeof(int), stream_), "gate clear");
    hip_check(hipMemsetAsync(gatePasses_, 0, sizeof(int), stream_), "pass count clear");
    vMesh_ = dalloc<int>(nv_);
    up(vMesh_, hVMesh_.data(), sizeof(int) * nv_);
    if (masked_) {
        vQuery_ = dalloc<uint8_t>(nv_);
        up(vQuery_, hQuery_.data(), nv_);
    }
    faces_ = dalloc<int4>(nf_);
    up(faces_, hFaces_.data(), sizeof(int4) * nf_);
    faceEdges_ = dalloc<int4>(nf_);
    up(faceEdges_, hFaceEdges_.data(), sizeof(int4) * nf_);
    edges_ = dalloc<int4>(ne_);
    up(edges_, hEdges_.data(), sizeof(int4) * ne_);
    eMesh_ = dalloc<int>(ne_);
    up(eMesh_, hEMesh_.data(), sizeof(int) * ne_);
    vvBegin_ = dalloc<int>(nv_ + 1);
    up(vvBegin_, vvB.data(), sizeof(int) * (nv_ + 1));
    vvNei_ = dalloc<int>(vvN.size());
    up(vvNei_, vvN.data(), sizeof(int) * vvN.size());
    vfBegin_ = dalloc<int>(nv_ + 1);
    up(vfBegin_, vfB.data(), sizeof(int) * (nv_ + 1));
    vfFace_ = dalloc<int>(vfF.size());
    up(vfFace_, vfF.data(), sizeof(int) * vfF.size());
    veBegin_ = dalloc<int>(nv_ + 1);
    up(veBegin_, veB.data(), sizeof(int) * (nv_ + 1));
    veEdge_ = dalloc<int>(veE.size());
    up(veEdge_, veE.data(), sizeof(int) * veE.size());
    const int M = numMeshes();
    drift_ = dalloc<float>(3 * (size_t)M);
    sceneBox_ = dalloc<float>(8);
    vfMin_ = dalloc<float>(nv_);
    eeMin_ = dalloc<float>(ne_);
    bounds_ = dalloc<float>(nv_);
    faceMin_ = dalloc<int>(nf_);
    vfPairCount_ = dalloc<int>(nv_ + 1);  // one extra zero: the scans' totals
    vfStored_ = dalloc<int>(nv_);
    vfOverflow_ = dalloc<int>(nv_);
    eeCount_ = dalloc<int>(ne_ + 1);
    eeStored_ = dalloc<int>(ne_);
    eeOverflow_ = dalloc<int>(ne_);
    vfRec_ = dalloc<VFContactRec>((size_t)nv_ * vfCap_);
    eeRec_ = dalloc<EEContactRec>((size_t)ne_ * eeCap_);
    vfOffset_ = dalloc<int64_t>(nv_ + 1);
    eeOffset_ = dalloc<int64_t>(ne_ + 1);
    error_ = dalloc<int>(64);
    depthHist_ = dalloc<int>(kMaxDepth);
    // per-block partial boxes of a tree build: the largest primitive count of any tree (faces,
    // edges, and with penetration contact the tets)
    boxPartialBlocks_ = blocks_for(std::max<int64_t>(std::max(nf_, ne_), pen_ ? (int64_t)penHost_.tetMesh.size() : 0));
    boxPartial_ = dalloc<unsigned int>(6 * (size_t)boxPartialBlocks_);
    hip_check(hipMemsetAsync(error_, 0, 64 * sizeof(int), stream_), "detector clear");
    hip_check(hipMemsetAsync(vfPairCount_, 0, sizeof(int) * (nv_ + 1), stream_), "detector clear");
    hip_check(hipMemsetAsync(eeCount_, 0, sizeof(int) * (ne_ + 1), stream_), "detector clear");
    // scan temp storage (sized for the larger of the two scans)
    size_t a = 0, b = 0;
    hip_check(rocprim::exclusive_scan(nullptr, a, vfPairCount_, vfOffset_, (int64_t)0, (size_t)nv_ + 1,
                                      rocprim::plus<int64_t>(), stream_),
              "scan size");
    hip_check(rocprim::exclusive_scan(nullptr, b, eeCount_, eeOffset_, (int64_t)0, (size_t)ne_ + 1,
                                      rocprim::plus<int64_t>(), stream_),
              "scan size");
    scanTempBytes_ = std::max(a, b);
    scanTemp_ = dalloc<char>(scanTempBytes_);
    // vertex query order (sorted Morton keys, rebuilt with the face tree)
    if (nv_ > 0) {
        vKeys_ = dalloc<uint64_t>(nv_);
        vKeysAlt_ = dalloc<uint64_t>(nv_);
        hip_check(rocprim::radix_sort_keys(nullptr, vSortTempBytes_, vKeys_, vKeysAlt_, (size_t)nv_, 0, 62, stream_),
                  "sort size");
        vSortTemp_ = dalloc<char>(vSortTempBytes_);
    }
    // BVH storage
    allocBvh(fb_, nf_);
    allocBvh(eb_, ne_);
    if (pen_ && nf_ > 0) fb_.mesh = dalloc<int>(2 * (size_t)nf_ - 1);
    if (pen_) setupPenetrationDevice();
    hip_check(hipStreamSynchronize(stream_), "detector upload sync");
    ready_ = true;
    updateBvh(true);
    if (pen_) {
        buildBvh(tb_, 3, tetBoxes_);
        // the swept-face tree over zero-length paths, so that every later pass refits (or
        // rebuilds on schedule) without a first-use build
        hipLaunchKernelGGL(swept_box_kernel, dim3(blocks_for(nf_)), dim3(256), 0, stream_, pos_, pos_, faces_, nf_,
                           sweptBoxes_);
        buildBvh(sb_, 3, sweptBoxes_);
        sweptBuilt_ = true;
    }
}

void ContactDetector::allocBvh(DeviceBvh& B, int n) {
    if (n > kLeafPrimMask) fail(GVBD_ERR_INVALID, "a tree of " + std::to_string(n) + " primitives exceeds 2^28 - 1");
    B.n = n;
    if (B.n == 0) return;
    B.nodes = dalloc<BvhNode>(2 * (size_t)B.n - 1);
    B.parent = dalloc<int>(2 * (size_t)B.n - 1);
    B.right = dalloc<int>(std::max(B.n - 1, 1));
    B.jump = dalloc<int>(4 * (size_t)std::max(B.n - 1, 1));
    B.keys = dalloc<uint64_t>(B.n);
    B.keysAlt = dalloc<uint64_t>(B.n);
    B.depthKeys = dalloc<uint64_t>(std::max(B.n - 1, 1));
    B.depthKeysAlt = dalloc<uint64_t>(std::max(B.n - 1, 1));
    B.depth = dalloc<int>(std::max(B.n - 1, 1));
    B.levelNodes = dalloc<int>(std::max(B.n - 1, 1));
    B.dLevelBegin = dalloc<int>(kMaxDepth + 1);
    size_t s1 = 0, s2 = 0;
    hip_check(rocprim::radix_sort_keys(nullptr, s1, B.keys, B.keysAlt, (size_t)B.n, 0, 62, stream_), "sort size");
    hip_check(rocprim::radix_sort_keys(nullptr, s2, B.depthKeys, B.depthKeysAlt, (size_t)std::max(B.n - 1, 1), 0, 40,
                                       stream_),
              "sort size");
    B.sortTempBytes = std::max(s1, s2);
    B.sortTemp = dalloc<char>(B.sortTempBytes);
}

namespace {
CDView make_view(float4* pos, int nv, int nf, int ne) {
    CDView v{};
    v.pos = pos;
    v.nv = nv;
    v.nf = nf;
    v.ne = ne;
    return v;
}
}  // namespace

void ContactDetector::buildBvh(DeviceBvh& B, int kind, const float4* boxes) {
    if (B.n == 0) return;
    CDView S = make_view(pos_, nv_, nf_, ne_);
    S.faces = faces_;
    S.edges = edges_;
    S.boxes = boxes;
    const int n = B.n;
    if (blocks_for(n) > boxPartialBlocks_)
        fail(GVBD_ERR_STATE, "tree of " + std::to_string(n) + " primitives exceeds the partial-box buffer");
    unsigned int* box = (unsigned int*)sceneBox_;
    hipLaunchKernelGGL(centroid_box_kernel, dim3(blocks_for(n)), dim3(256), 0, stream_, S, kind, n, boxPartial_);
    hipLaunchKernelGGL(box_final_kernel, dim3(1), dim3(256), 0, stream_, boxPartial_, blocks_for(n), box);
    hipLaunchKernelGGL(morton_kernel, dim3(blocks_for(n)), dim3(256), 0, stream_, S, kind, n, box, B.keys);
    size_t tb = B.sortTempBytes;
    hip_check(rocprim::radix_sort_keys(B.sortTemp, tb, B.keys, B.keysAlt, (size_t)n, 0, 62, stream_), "morton sort");
    if (kind == 0 && nv_ > 0) {
        // VF query order: vertices sorted by Morton code in the face tree's box
        hipLaunchKernelGGL(morton_kernel, dim3(blocks_for(nv_)), dim3(256), 0, stream_, S, 2, nv_, box, vKeys_);
        size_t tv = vSortTempBytes_;
        hip_check(rocprim::radix_sort_keys(vSortTemp_, tv, vKeys_, vKeysAlt_, (size_t)nv_, 0, 62, stream_),
                  "vertex order sort");
    }
    hipLaunchKernelGGL(build_tree_kernel, dim3(blocks_for(n)), dim3(256), 0, stream_, B.keysAlt, n, B.nodes, B.parent,
                       B.right);
    hipLaunchKernelGGL(skip_kernel, dim3(blocks_for(2 * (int64_t)n - 1)), dim3(256), 0, stream_, B.nodes, B.parent,
                       B.right, 2 * n - 1, error_);
    hip_check(hipGetLastError(), "bvh build launch");
    // refit schedule: internal nodes by depth, deepest first
    B.levelBegin.clear();
    const int ni = n - 1;
    if (ni > 0) {
        int* d0 = B.jump;
        int* d1 = B.jump + ni;
        int* a0 = B.jump + 2 * (size_t)ni;
        int* a1 = B.jump + 3 * (size_t)ni;
        hipLaunchKernelGGL(depth_init_kernel, dim3(blocks_for(ni)), dim3(256), 0, stream_, B.parent, ni, d0, a0);
        for (int round = 0; (1 << round) < kMaxDepth; ++round) {
            hipLaunchKernelGGL(depth_jump_kernel, dim3(blocks_for(ni)), dim3(256), 0, stream_, d0, a0, d1, a1, ni);
            std::swap(d0, d1);
            std::swap(a0, a1);
        }
        hipLaunchKernelGGL(depth_keys_kernel, dim3(blocks_for(ni)), dim3(256), 0, stream_, d0, a0, ni, B.depth,
                           B.depthKeys, error_);
        size_t tb2 = B.sortTempBytes;
        hip_check(rocprim::radix_sort_keys(B.sortTemp, tb2, B.depthKeys, B.depthKeysAlt, (size_t)ni, 0, 40, stream_),
                  "depth sort");
        hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(256), 0, stream_, depthHist_, (int64_t)kMaxDepth, -1);
        hipLaunchKernelGGL(level_nodes_kernel, dim3(blocks_for(ni)), dim3(256), 0, stream_, B.depthKeysAlt, ni,
                           B.levelNodes, depthHist_);
        std::vector<int> starts(kMaxDepth, -1);
        hip_check(hipMemcpyAsync(starts.data(), depthHist_, kMaxDepth * sizeof(int), hipMemcpyDeviceToHost, stream_),
                  "level read");
        hip_check(hipStreamSynchronize(stream_), "bvh build sync");
        // runs are stored deepest first
        for (int d = kMaxDepth - 1; d >= 0; --d)
            if (starts[d] >= 0) B.levelBegin.push_back(starts[d]);
        B.levelBegin.push_back(ni);
        // refit launches: every level of more than kBlockLevel nodes alone, runs of smaller levels
        // together in one block
        B.segments.clear();
        const int nl = (int)B.levelBegin.size() - 1;
        for (int l = 0; l < nl;) {
            if (B.levelBegin[l + 1] - B.levelBegin[l] > kBlockLevel) {
                B.segments.push_back({l, l + 1});
                ++l;
                continue;
            }
            int e = l;
            while (e < nl && B.levelBegin[e + 1] - B.levelBegin[e] <= kBlockLevel) ++e;
            B.segments.push_back({l, e});
            l = e;
        }
        hip_check(hipMemcpyAsync(B.dLevelBegin, B.levelBegin.data(), sizeof(int) * B.levelBegin.size(),
                                 hipMemcpyHostToDevice, stream_),
                  "level upload");
    }
    refitBvh(B, kind, stream_, boxes);
    if (B.mesh && kind == 0) {
        // mesh membership changes only with the tree's topology: set at every build
        hipLaunchKernelGGL(node_mesh_leaf_kernel, dim3(blocks_for(n)), dim3(256), 0, stream_, faces_, B.nodes, n,
                           B.mesh);
        for (const auto& seg : B.segments) {
            const int b = B.levelBegin[seg.first], e = B.levelBegin[seg.first + 1];
            if (seg.second == seg.first + 1 && e - b > kBlockLevel)
                hipLaunchKernelGGL(node_mesh_internal_kernel, dim3(blocks_for(e - b)), dim3(256), 0, stream_,
                                   B.levelNodes, b, e, B.nodes, B.right, B.mesh);
            else
                hipLaunchKernelGGL(node_mesh_levels_kernel, dim3(1), dim3(kBlockLevel), 0, stream_, B.levelNodes,
                                   B.dLevelBegin, seg.first, seg.second, B.nodes, B.right, B.mesh);
        }
        hip_check(hipGetLastError(), "node mesh launch");
    }
}

void ContactDetector::refitBvh(DeviceBvh& B, int kind, hipStream_t stream, const float4* boxes, const int* gate) {
    if (B.n == 0) return;
    CDView S = make_view(pos_, nv_, nf_, ne_);
    S.faces = faces_;
    S.edges = edges_;
    S.boxes = boxes;
    S.gate = gate;
    hipLaunchKernelGGL(leaf_box_kernel, dim3(blocks_for(B.n)), dim3(256), 0, stream, S, kind, B.nodes, B.n);
    for (const auto& seg : B.segments) {
        const int b = B.levelBegin[seg.first], e = B.levelBegin[seg.first + 1];
        if (seg.second == seg.first + 1 && e - b > kBlockLevel)
            hipLaunchKernelGGL(internal_box_kernel, dim3(blocks_for(e - b)), dim3(256), 0, stream, B.levelNodes, b,
                               e, B.nodes, B.right, gate);
        else
            hipLaunchKernelGGL(internal_box_levels_kernel, dim3(1), dim3(kBlockLevel), 0, stream, B.levelNodes,
                               B.dLevelBegin, seg.first, seg.second, B.nodes, B.right, gate);
    }
    hip_check(hipGetLastError(), "bvh refit launch");
}

void ContactDetector::setPenetration(PenetrationSetup&& setup) {
    if (ready_) fail(GVBD_ERR_STATE, "detector already finalized");
    require(setup.tets.size() % 4 == 0 && setup.tetNbr.size() == setup.tets.size() &&
                setup.tetMesh.size() * 4 == setup.tets.size(),
            "penetration setup: tet arrays do not match");
    penHost_ = std::move(setup);
    pen_ = true;
}

void ContactDetector::setupPenetrationDevice() {
    const PenetrationSetup& H = penHost_;
    nt_ = (int)H.tetMesh.size();
    nSurf_ = (int)H.surfVerts.size();
    require((int)H.faceOwner.size() == 2 * nf_ && (int)H.faceNbr.size() == 3 * nf_ &&
                (int)H.vNbrBegin.size() == nv_ + 1,
            "penetration setup: surface arrays do not match the detector's faces and vertices");
    auto up = [&](void* dst, const void* src, size_t bytes) {
        if (bytes) hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream_), "penetration upload");
    };
    penTets_ = dalloc<int4>(std::max(nt_, 1));
    up(penTets_, H.tets.data(), sizeof(int32_t) * H.tets.size());
    penTetNbr_ = dalloc<int4>(std::max(nt_, 1));
    up(penTetNbr_, H.tetNbr.data(), sizeof(int32_t) * H.tetNbr.size());
    penTetMesh_ = dalloc<int>(std::max(nt_, 1));
    up(penTetMesh_, H.tetMesh.data(), sizeof(int32_t) * H.tetMesh.size());
    penFaceOwner_ = dalloc<int2>(std::max(nf_, 1));
    up(penFaceOwner_, H.faceOwner.data(), sizeof(int32_t) * H.faceOwner.size());
    std::vector<int4> nb(nf_);
    for (int f = 0; f < nf_; ++f) nb[f] = make_int4(H.faceNbr[3 * f], H.faceNbr[3 * f + 1], H.faceNbr[3 * f + 2], 0);
    penFaceNbr_ = dalloc<int4>(std::max(nf_, 1));
    up(penFaceNbr_, nb.data(), sizeof(int4) * nb.size());
    penVNbrBegin_ = dalloc<int>(nv_ + 1);
    up(penVNbrBegin_, H.vNbrBegin.data(), sizeof(int32_t) * H.vNbrBegin.size());
    penVNbr_ = dalloc<int>(std::max<size_t>(H.vNbr.size(), 1));
    up(penVNbr_, H.vNbr.data(), sizeof(int32_t) * H.vNbr.size());
    penSurf_ = dalloc<int>(std::max(nSurf_, 1));
    up(penSurf_, H.surfVerts.data(), sizeof(int32_t) * H.surfVerts.size());
    tetBoxes_ = dalloc<float4>(2 * (size_t)std::max(nt_, 1));
    sweptBoxes_ = dalloc<float4>(2 * (size_t)std::max(nf_, 1));
    penFlags_ = dalloc<uint8_t>(std::max(nv_, 1));
    penCounts_ = dalloc<int>(kPenCounts);
    penInTet_ = dalloc<int>(std::max(nv_, 1));
    penList_ = dalloc<int>(std::max(nSurf_, 1));
    hipLaunchKernelGGL(fill_u8_kernel, dim3(std::min(blocks_for(nv_), 4096)), dim3(256), 0, stream_, penFlags_,
                       (int64_t)nv_);
    hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(256), 0, stream_, penCounts_, (int64_t)kPenCounts, 0);
    allocBvh(tb_, nt_);
    allocBvh(sb_, nf_);
    // tet boxes for the first tree build (at the end of finalize)
    hipLaunchKernelGGL(tet_box_kernel, dim3(blocks_for(nt_)), dim3(256), 0, stream_, pos_, penTets_, nt_, tetBoxes_);
    hip_check(hipGetLastError(), "penetration setup launch");
}

namespace {
PenView pen_view(float4* pos, const float4* prev, const int4* faces) {
    PenView v{};
    v.pos = pos;
    v.prev = prev;
    v.faces = faces;
    return v;
}
}  // namespace

void ContactDetector::penetrationDcd(bool rebuild) {
    if (!ready_ || !pen_) fail(GVBD_ERR_STATE, "detector not finalized for penetration contact");
    hipLaunchKernelGGL(fill_int_kernel, dim3(std::min(blocks_for(nv_), 4096)), dim3(256), 0, stream_, vfStored_,
                       (int64_t)nv_, 0);
    hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(256), 0, stream_, penCounts_, (int64_t)kPenCounts, 0);
    hipLaunchKernelGGL(tet_box_kernel, dim3(blocks_for(nt_)), dim3(256), 0, stream_, pos_, penTets_, nt_, tetBoxes_);
    if (rebuild) {
        buildBvh(tb_, 3, tetBoxes_);
        buildBvh(fb_, 0);
    } else {
        refitBvh(tb_, 3, stream_, tetBoxes_);
        refitBvh(fb_, 0, stream_);
    }
    PenView S = pen_view(pos_, nullptr, faces_);
    S.tets = penTets_;
    S.tetMesh = penTetMesh_;
    S.tetNbr = penTetNbr_;
    S.faceOwner = penFaceOwner_;
    S.faceNbr = penFaceNbr_;
    S.vNbrBegin = penVNbrBegin_;
    S.vNbr = penVNbr_;
    S.surfVerts = penSurf_;
    S.nSurf = nSurf_;
    S.vMesh = vMesh_;
    S.tNodes = tb_.nodes;
    S.tN = tb_.n;
    S.fNodes = fb_.nodes;
    S.fMesh = fb_.mesh;
    S.fN = fb_.n;
    S.P = penHost_.params;
    S.stored = vfStored_;
    S.rec = vfRec_;
    S.cap = vfCap_;
    S.flags = penFlags_;
    S.counts = penCounts_;
    S.inTet = penInTet_;
    S.list = penList_;
    S.error = error_;
    if (nSurf_ > 0) {
        hipLaunchKernelGGL(pen_contain_kernel, dim3((nSurf_ + 63) / 64), dim3(64), 0, stream_, S);
        hipLaunchKernelGGL(pen_closest_kernel, dim3((nSurf_ + 63) / 64), dim3(64), 0, stream_, S);
    }
    hip_check(hipGetLastError(), "penetration DCD launch");
}

void ContactDetector::penetrationCcd(const float4* prev, bool rebuild) {
    if (!ready_ || !pen_) fail(GVBD_ERR_STATE, "detector not finalized for penetration contact");
    if (!penHost_.params.ccd || nf_ == 0) return;
    hipLaunchKernelGGL(swept_box_kernel, dim3(blocks_for(nf_)), dim3(256), 0, stream_, pos_, prev, faces_, nf_,
                       sweptBoxes_);
    if (rebuild || !sweptBuilt_) {
        buildBvh(sb_, 3, sweptBoxes_);
        sweptBuilt_ = true;
    } else {
        refitBvh(sb_, 3, stream_, sweptBoxes_);
    }
    PenView S = pen_view(pos_, prev, faces_);
    S.surfVerts = penSurf_;
    S.nSurf = nSurf_;
    S.vMesh = vMesh_;
    S.sNodes = sb_.nodes;
    S.sN = sb_.n;
    S.P = penHost_.params;
    S.stored = vfStored_;
    S.rec = vfRec_;
    S.cap = vfCap_;
    S.flags = penFlags_;
    S.counts = penCounts_;
    S.error = error_;
    if (nSurf_ > 0) hipLaunchKernelGGL(pen_ccd_kernel, dim3(blocks_for(nSurf_)), dim3(256), 0, stream_, S);
    hip_check(hipGetLastError(), "penetration CCD launch");
}

PenetrationCounts ContactDetector::penetrationCounts() {
    PenetrationCounts c{};
    if (!pen_) return c;
    int h[kPenCounts];
    hip_check(hipMemcpyAsync(h, penCounts_, sizeof(h), hipMemcpyDeviceToHost, stream_), "counts");
    hip_check(hipStreamSynchronize(stream_), "counts sync");
    c.dcdHits = h[0];
    c.penetratedNoFace = h[1];
    c.ccdHits = h[2];
    c.walkOverflows = h[3];
    c.fullSearches = h[5];
    return c;
}

void ContactDetector::penetrationResults(int32_t* face, uint8_t* flags) {
    if (!pen_) fail(GVBD_ERR_STATE, "detector has no penetration contact");
    std::vector<int> stored(nv_);
    std::vector<VFContactRec> rec((size_t)nv_ * vfCap_);
    hip_check(hipMemcpyAsync(stored.data(), vfStored_, sizeof(int) * nv_, hipMemcpyDeviceToHost, stream_), "download");
    hip_check(hipMemcpyAsync(rec.data(), vfRec_, sizeof(VFContactRec) * rec.size(), hipMemcpyDeviceToHost, stream_),
              "download");
    if (flags) hip_check(hipMemcpyAsync(flags, penFlags_, nv_, hipMemcpyDeviceToHost, stream_), "download");
    hip_check(hipStreamSynchronize(stream_), "download sync");
    for (int v = 0; v < nv_; ++v) face[v] = stored[v] ? rec[(size_t)v * vfCap_].face : -1;
}

void ContactDetector::updateBvh(bool rebuild) {
    if (!ready_) fail(GVBD_ERR_STATE, "detector not finalized");
    if (rebuild) {
        buildBvh(fb_, 0);
        buildBvh(eb_, 1);
    } else {
        fork();
        refitBvh(fb_, 0, stream_);
        refitBvh(eb_, 1, side_);
        join();
    }
    treeCurrent_ = true;
}

void ContactDetector::setPositions(int mesh, const float* pos) {
    if (!ready_) fail(GVBD_ERR_STATE, "detector not finalized");
    require(mesh >= 0 && mesh < numMeshes(), "mesh index out of range");
    const int off = meshV_[mesh], n = meshV_[mesh + 1] - off;
    std::vector<float4> p4(n);
    for (int i = 0; i < n; ++i) p4[i] = make_float4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], 0.f);
    hip_check(hipMemcpyAsync(pos_ + off, p4.data(), sizeof(float4) * n, hipMemcpyHostToDevice, stream_), "positions");
    hip_check(hipStreamSynchronize(stream_), "positions sync");
    treeCurrent_ = false;
}

void ContactDetector::getPositions(int mesh, float* pos) const {
    if (!ready_) fail(GVBD_ERR_STATE, "detector not finalized");
    require(mesh >= 0 && mesh < numMeshes(), "mesh index out of range");
    const int off = meshV_[mesh], n = meshV_[mesh + 1] - off;
    std::vector<float4> p4(n);
    hip_check(hipMemcpyAsync(p4.data(), pos_ + off, sizeof(float4) * n, hipMemcpyDeviceToHost, stream_), "positions");
    hip_check(hipStreamSynchronize(stream_), "positions sync");
    for (int i = 0; i < n; ++i) {
        pos[3 * i] = p4[i].x;
        pos[3 * i + 1] = p4[i].y;
        pos[3 * i + 2] = p4[i].z;
    }
}

void ContactDetector::perturb(uint64_t seed, int step, float sigma, const float* drift) {
    if (!ready_) fail(GVBD_ERR_STATE, "detector not finalized");
    if (!rest_) fail(GVBD_ERR_STATE, "perturb needs a detector that owns its positions");
    const int M = numMeshes();
    std::vector<float> d(3 * (size_t)M, 0.f);
    if (drift) std::memcpy(d.data(), drift, d.size() * sizeof(float));
    hip_check(hipMemcpyAsync(drift_, d.data(), d.size() * sizeof(float), hipMemcpyHostToDevice, stream_), "drift");
    hipLaunchKernelGGL(perturb_kernel, dim3(blocks_for(nv_)), dim3(256), 0, stream_, rest_, pos_, vMesh_, drift_, nv_,
                       seed, step, sigma);
    hip_check(hipGetLastError(), "perturb launch");
    hip_check(hipStreamSynchronize(stream_), "perturb sync");
    treeCurrent_ = false;
}

void ContactDetector::query(bool record, const int* gate) {
    if (!ready_) fail(GVBD_ERR_STATE, "detector not finalized");
    // the EE query reads the other edge's endpoints from the tree's leaves
    if (!treeCurrent_) fail(GVBD_ERR_STATE, "positions changed since the last tree update (update the BVH before querying)");
    if (record && gate) fail(GVBD_ERR_STATE, "the record pass is not gated");
    CDView S = make_view(pos_, nv_, nf_, ne_);
    S.gate = gate;
    S.vMesh = vMesh_;
    S.vQuery = vQuery_;
    S.faces = faces_;
    S.faceEdges = faceEdges_;
    S.edges = edges_;
    S.eMesh = eMesh_;
    S.vvBegin = vvBegin_;
    S.vvNei = vvNei_;
    S.fNodes = fb_.nodes;
    S.fN = fb_.n;
    S.eNodes = eb_.nodes;
    S.eN = eb_.n;
    S.eOrder = eb_.n > 0 ? eb_.keysAlt : nullptr;
    S.vOrder = fb_.n > 0 ? vKeysAlt_ : nullptr;  // computed with the face tree
    S.r = r_;
    S.vfCap = vfCap_;
    S.eeCap = eeCap_;
    S.vfMin = vfMin_;
    S.faceMin = faceMin_;
    S.eeMin = eeMin_;
    S.vfPairCount = vfPairCount_;
    S.vfStored = vfStored_;
    S.vfOverflow = vfOverflow_;
    S.eeCount = eeCount_;
    S.eeStored = eeStored_;
    S.eeOverflow = eeOverflow_;
    S.vfRec = vfRec_;
    S.eeRec = eeRec_;
    S.error = error_;
    float rr = r_;
    int rbits;
    std::memcpy(&rbits, &rr, sizeof(int));
    hipLaunchKernelGGL(gated_fill_int_kernel, dim3(std::min(blocks_for(nf_), 4096)), dim3(256), 0, stream_, faceMin_,
                       (int64_t)nf_, rbits, gate);
    fork();
    if (vfCap_ <= 8)
        hipLaunchKernelGGL((vf_query_kernel<false, 8>), dim3(blocks_for(nv_)), dim3(256), 0, stream_, S);
    else
        hipLaunchKernelGGL((vf_query_kernel<false, kMaxCap>), dim3(blocks_for(nv_)), dim3(256), 0, stream_, S);
    if (eeCap_ <= 8)
        hipLaunchKernelGGL((ee_query_kernel<false, 8>), dim3(blocks_for(ne_)), dim3(256), 0, side_, S);
    else
        hipLaunchKernelGGL((ee_query_kernel<false, kMaxCap>), dim3(blocks_for(ne_)), dim3(256), 0, side_, S);
    join();
    hip_check(hipGetLastError(), "query launch");
    recorded_ = false;
    if (!record) return;
    // record pass: offsets from the counts, then the uncapped lists
    size_t tb = scanTempBytes_;
    hip_check(rocprim::exclusive_scan(scanTemp_, tb, vfPairCount_, vfOffset_, (int64_t)0, (size_t)nv_ + 1,
                                      rocprim::plus<int64_t>(), stream_),
              "pair scan");
    tb = scanTempBytes_;
    hip_check(rocprim::exclusive_scan(scanTemp_, tb, eeCount_, eeOffset_, (int64_t)0, (size_t)ne_ + 1,
                                      rocprim::plus<int64_t>(), stream_),
              "ee scan");
    hip_check(hipMemcpyAsync(&nRecPairs_, vfOffset_ + nv_, sizeof(int64_t), hipMemcpyDeviceToHost, stream_), "total");
    hip_check(hipMemcpyAsync(&nRecEE_, eeOffset_ + ne_, sizeof(int64_t), hipMemcpyDeviceToHost, stream_), "total");
    hip_check(hipStreamSynchronize(stream_), "record sync");
    if (nRecPairs_ > capPairs_) {
        capPairs_ = std::max<int64_t>(nRecPairs_, 2 * capPairs_);
        recVF_ = dalloc<int2>(capPairs_);
        recVFd_ = dalloc<float>(capPairs_);
        recKeys_ = dalloc<int4>(capPairs_);
        recKeysD_ = dalloc<float>(capPairs_);
    }
    if (nRecEE_ > capEE_) {
        capEE_ = std::max<int64_t>(nRecEE_, 2 * capEE_);
        recEE_ = dalloc<int2>(capEE_);
        recEEmu_ = dalloc<float2>(capEE_);
        recEEd_ = dalloc<float>(capEE_);
    }
    S.vfOffset = vfOffset_;
    S.eeOffset = eeOffset_;
    S.recVF = recVF_;
    S.recVFd = recVFd_;
    S.recKeys = recKeys_;
    S.recKeysD = recKeysD_;
    S.recEE = recEE_;
    S.recEEmu = recEEmu_;
    S.recEEd = recEEd_;
    hipLaunchKernelGGL((vf_query_kernel<true, 1>), dim3(blocks_for(nv_)), dim3(256), 0, stream_, S);
    hipLaunchKernelGGL((ee_query_kernel<true, 1>), dim3(blocks_for(ne_)), dim3(256), 0, stream_, S);
    hip_check(hipGetLastError(), "record query launch");
    hip_check(hipStreamSynchronize(stream_), "record query sync");
    recorded_ = true;
}

void ContactDetector::bounds(float thickness, float relaxation, const int* gate) {
    if (!ready_) fail(GVBD_ERR_STATE, "detector not finalized");
    hipLaunchKernelGGL(bounds_kernel, dim3(blocks_for(nv_)), dim3(256), 0, stream_, vfMin_, faceMin_, eeMin_,
                       vfBegin_, vfFace_, veBegin_, veEdge_, vQuery_, nv_, r_, thickness, relaxation, bounds_, gate);
    hip_check(hipGetLastError(), "bounds launch");
}

namespace {
template <class T> std::vector<T> download(const T* d, size_t n, hipStream_t s) {
    std::vector<T> h(n);
    if (n) hip_check(hipMemcpyAsync(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost, s), "download");
    hip_check(hipStreamSynchronize(s), "download sync");
    return h;
}
}  // namespace

DetectorResultCounts ContactDetector::counts() {
    if (!ready_) fail(GVBD_ERR_STATE, "detector not finalized");
    const auto err = download(error_, 1, stream_);
    if (err[0] & 1) fail(GVBD_ERR_INTERNAL, "BVH traversal visited more nodes than the tree holds");
    if (err[0] & 2) fail(GVBD_ERR_INTERNAL, "BVH deeper than the refit schedule supports");
    if (err[0] & 4) fail(GVBD_ERR_INTERNAL, "record pass found more contacts than the counting pass");
    const auto pc = download(vfPairCount_, nv_, stream_);
    const auto st = download(vfStored_, nv_, stream_);
    const auto vo = download(vfOverflow_, nv_, stream_);
    const auto ec = download(eeCount_, ne_, stream_);
    const auto eo = download(eeOverflow_, ne_, stream_);
    DetectorResultCounts c{};
    for (int i = 0; i < nv_; ++i) {
        c.vfPairs += pc[i];
        c.vfContacts += st[i];
        c.vfOverflow += vo[i];
    }
    for (int i = 0; i < ne_; ++i) {
        c.eeContacts += ec[i];
        c.eeOverflow += eo[i];
    }
    if (recorded_) {
        // exact distinct-key count from the record pass
        const auto keys = download(recKeys_, nRecPairs_, stream_);
        c.vfContacts = 0;
        for (const int4& k : keys) c.vfContacts += k.y >= 0;
    }
    return c;
}

void ContactDetector::vfPairs(int32_t* vf, float* d) {
    if (!recorded_) fail(GVBD_ERR_STATE, "no recorded query (call query with record = 1)");
    const auto p = download(recVF_, nRecPairs_, stream_);
    const auto dd = download(recVFd_, nRecPairs_, stream_);
    for (int64_t i = 0; i < nRecPairs_; ++i) {
        if (vf) {
            vf[2 * i] = p[i].x;
            vf[2 * i + 1] = p[i].y;
        }
        if (d) d[i] = dd[i];
    }
}

void ContactDetector::vfContactsAll(int32_t* rec, float* d) {
    if (!recorded_) fail(GVBD_ERR_STATE, "no recorded query (call query with record = 1)");
    const auto k = download(recKeys_, nRecPairs_, stream_);
    const auto dd = download(recKeysD_, nRecPairs_, stream_);
    int64_t o = 0;
    for (int64_t i = 0; i < nRecPairs_; ++i) {
        if (k[i].y < 0) continue;
        if (rec) {
            rec[4 * o] = k[i].x;
            rec[4 * o + 1] = k[i].y;
            rec[4 * o + 2] = k[i].z;
            rec[4 * o + 3] = k[i].w;
        }
        if (d) d[o] = dd[i];
        ++o;
    }
}

void ContactDetector::eeContactsAll(int32_t* ee, float* mu, float* d) {
    if (!recorded_) fail(GVBD_ERR_STATE, "no recorded query (call query with record = 1)");
    const auto p = download(recEE_, nRecEE_, stream_);
    const auto m = download(recEEmu_, nRecEE_, stream_);
    const auto dd = download(recEEd_, nRecEE_, stream_);
    for (int64_t i = 0; i < nRecEE_; ++i) {
        if (ee) {
            ee[2 * i] = p[i].x;
            ee[2 * i + 1] = p[i].y;
        }
        if (mu) {
            mu[2 * i] = m[i].x;
            mu[2 * i + 1] = m[i].y;
        }
        if (d) d[i] = dd[i];
    }
}

void ContactDetector::vfContactsCapped(int32_t* count, VFContactRec* rec) {
    if (count) {
        const auto c = download(vfStored_, nv_, stream_);
        std::memcpy(count, c.data(), sizeof(int) * nv_);
    }
    if (rec) {
        const auto r = download(vfRec_, (size_t)nv_ * vfCap_, stream_);
        std::memcpy(rec, r.data(), sizeof(VFContactRec) * r.size());
    }
}

void ContactDetector::eeContactsCapped(int32_t* count, EEContactRec* rec) {
    if (count) {
        const auto c = download(eeStored_, ne_, stream_);
        std::memcpy(count, c.data(), sizeof(int) * ne_);
    }
    if (rec) {
        const auto r = download(eeRec_, (size_t)ne_ * eeCap_, stream_);
        std::memcpy(rec, r.data(), sizeof(EEContactRec) * r.size());
    }
}

void ContactDetector::minDistances(float* vf, float* face, float* edge) {
    if (vf) {
        const auto a = download(vfMin_, nv_, stream_);
        std::memcpy(vf, a.data(), sizeof(float) * nv_);
    }
    if (face) {
        const auto a = download(faceMin_, nf_, stream_);
        std::memcpy(face, a.data(), sizeof(float) * nf_);
    }
    if (edge) {
        const auto a = download(eeMin_, ne_, stream_);
        std::memcpy(edge, a.data(), sizeof(float) * ne_);
    }
}

void ContactDetector::boundsOut(float* b) {
    const auto a = download(bounds_, nv_, stream_);
    std::memcpy(b, a.data(), sizeof(float) * nv_);
}

}  // namespace gvbd
