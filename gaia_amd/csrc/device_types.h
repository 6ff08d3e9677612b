// Device-side data layout of the VBD hot path (shared by host code and HIP kernels).
//
// HBM layout of one scene (all meshes of the scene share one vertex index space):
//   pos[nv]      float4  current positions x (w unused)           -- gathered by neighbours
//   prev[nv]     float4  positions at the start of the step       -- streamed
//   inertia[nv]  float4  inertia target y in xyz, lumped mass in w -- streamed
//   vel[nv]      float4  velocities v (w unused)                   -- streamed
//   vparam[nv]   float4  (exp vel damping, const vel damping, hydrostatic damping, deviatoric damping)
//   vflag[nv]    uint8   bit0 fixed, bits1-3 initialisation type
//   slots[ns]    int4    colour-major work slots: (vertex, incidence begin, incidence end, flags)
//   tinc[ni]     TetInc  colour-major vertex->tet incidence records (48 B each, streamed)
//   tvol[ni]     float   rest volume of each incidence record's tet (streamed beside tinc)
//   tmat[nm]     float4  per tet mesh (mu, lambda, a = 1 + mu / lambda, 0), indexed by slot flags >> 8
// Slots of parallel group g are slots[groupOffset[g] .. groupOffset[g+1]); the incidence
// records of consecutive slots are consecutive, so a block's records form one range.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gvbd {

// One incident tet of one vertex v, expressed relative to v:
//   F = sum_k (x_nb[k] - x_v) m_k^T,  m_v = -(m_0 + m_1 + m_2)
// where m_k is the shape-matrix row of the k-th other vertex (rows of DmInv, or
// minus their sum for tet vertex 0). The tet's rest volume A is the same record's entry of
// the tvol stream and the material (mu, lambda, a) is its mesh's entry of tmat, so the solve
// forms A * mu and A * lambda itself (the same float products the host would store).
struct alignas(16) TetInc {
    int32_t nb[3];
    float m[9];  // m_0 (3), m_1 (3), m_2 (3)
};
static_assert(sizeof(TetInc) == 48, "TetInc must be 48 bytes");

// One tet for energy evaluation: scene vertex ids, DmInv (column-major), rest volume, moduli.
struct alignas(16) TetEnergyRec {
    int4 v;
    float dminv[9];
    float A, mu, lmbd;
};
static_assert(sizeof(TetEnergyRec) == 64, "TetEnergyRec must be 64 bytes");

enum SlotFlags : int32_t {
    SLOT_DAMPING = 1,  // the vertex's mesh has non-zero material damping
    SLOT_INPLACE = 2,  // cloth: write x directly instead of positionsNext (no commit pass)
    SLOT_ATTACH = 8,   // cloth: the vertex has attachment records (SceneView::attBegin)
};
constexpr int SLOT_MAT_SHIFT = 8;  // tet slots: flags >> 8 = the mesh's entry of SceneView::tmat

// One deformed vertex (VBD/VBDDeformer.h, VBDCloth/VBDClothDeformer.h): scene vertex v moved by
// deformer d of its scene; a = root (Rotator) or initial position (tet Translator), b = RootP.
struct alignas(16) DeformVert {
    int32_t v;
    int32_t d;
    float a[3];
    float b[3];
};
static_assert(sizeof(DeformVert) == 32, "DeformVert must be 32 bytes");

// A deformer's parameters: kind 0 Rotator (axis, angular velocity omega, end time), 1 Translator
// (speed, start and end time).
struct alignas(16) DeformParams {
    int32_t kind;
    float axis[3];
    float speed[3];
    float startTime;
    double omega;
    double endTime;
};
static_assert(sizeof(DeformParams) == 48, "DeformParams must be 48 bytes");

// One vertex-face attachment record of a cloth vertex (VBDClothPhysics.cpp:1465-1498):
// force -= k b (x[read] - target), H += k b^2 I, target = p.x x[f0] + p.y x[f1] + p.z x[f2]
// (f0 >= 0) or the point p.xyz (f0 < 0). ids = (f0, f1, f2, read) in scene vertex ids; p.w = k b.
struct alignas(16) AttachRec {
    int4 ids;
    float4 p;
    float b;
    float pad[3];
};
static_assert(sizeof(AttachRec) == 48, "AttachRec must be 48 bytes");

// One incident triangle of a cloth vertex v, relative to v:
//   F (3x2) = (x_nb0 - x_v) m0^T + (x_nb1 - x_v) m1^T,  m_v = -(m0 + m1)
// with m0, m1 the rows of the face's 2x2 DmInv (or minus their sum for face vertex 0).
struct alignas(16) FaceInc {
    int32_t nb[2];
    float Amu;    // rest area * mu
    float Alm;    // rest area * lambda
    float m[4];   // m0 (2), m1 (2)
};
static_assert(sizeof(FaceInc) == 32, "FaceInc must be 32 bytes");

// One relevant bending stencil of a cloth vertex v (VBDTriMeshStVK.cpp:913-982): v plays role
// `order` in (eV1, eV2, eV12Next, eV21Next); nb are the other three in role order; q holds
// bendingStiffness * Q(order, role of nb[k]) and qself bendingStiffness * Q(order, order).
struct alignas(16) BendInc {
    int32_t nb[3];
    int32_t order;
    float q[3];
    float qself;
};
static_assert(sizeof(BendInc) == 32, "BendInc must be 32 bytes");

// Cloth work slot: (vertex, face begin, face end, flags), (bending begin, bending end, edge begin,
// edge end); the edge range indexes ContactView::edgeRef (incident edges, contact scenes only).
struct alignas(16) ClothSlot {
    int4 a;
    int4 b;
};

enum WorkKind : int32_t {
    WORK_TET = 0,
    WORK_CLOTH = 1,
};

// Scene-wide constants of the per-vertex solve (VBDPhysicsParameters / BasePhysicsParams).
struct SolveParams {
    float dt;
    float dtSqrRecip;
    float gravity[3];
    float lower[3];
    float upper[3];
    float boundaryStiffness;
    float frictionMu;
    float frictionEpsU;  // boundaryFrictionEpsV * dt
    float stepSize;
    float stepSizeGD;
    float initRatioG;
    int32_t usePlaneGround;
    int32_t associateGravity;
    int32_t frictionStartIter;
    int32_t pad;
    // cloth (VBDClothPhysicsParameters)
    float clothBoundaryStiffness;  // boundaryContactStiffness
    float clothFrictionEpsU;       // boundaryFrictionEpsValue * dt
    float degenerateThresSq;       // degenerateTriangleThres^2
    float clothStepSize;
    // cloth contact (VBDClothPhysicsParameters.h:60-71)
    float contactStiffness;
    float contactRadius;
    float thickness;
    int32_t contactEnergyType;
};

// Contact data of a cloth scene with collision handling: the detector's per-primitive lists
// (scene vertex / face / edge ids), face back-links and conservative bounds.
struct ContactView {
    const int4* faces;        // v0, v1, v2, mesh
    const int4* edges;        // eV1, eV2, eV12Next, eV21Next
    const float2* friction;   // per vertex: (frictionDynamic, frictionEpsV) of its mesh
    const int* vfCount;       // per vertex: stored VF contacts
    const int4* vfRec;        // per vertex vfCap records (face, type, prim, d bits)
    int vfCap;
    int eeCap;
    const int* eeCount;       // per edge
    const float4* eeRec;      // per edge eeCap records (edge bits, mu1, mu2, d)
    const int2* fvRange;      // per face: [begin, end) into fvVerts (at most 32)
    const int* fvVerts;       // per face, ascending vertex ids of its VF contacts
    const int* faceRef;       // per FaceInc record: face id | order << 30
    const int* edgeRef;       // per cloth-slot edge record: edge id | order << 30
    const int* tFaceRef;      // tet slots: incident boundary faces, face id | order << 30
    const int* tEdgeRef;      // tet slots: incident boundary edges, edge id | order << 30
    const float* bound;       // per vertex conservative bound
    const float4* xAtCD;      // positions at the last detection
    int* flag;                // set when a vertex is truncated: detect before the next iteration
    // Tet vertices of contact scenes are updated in place; a contact pair may join two vertices
    // of one parallel group, so contact partners of the solving vertex's own group are read from
    // the iteration-start snapshot (their value at the start of the group, as the reference's
    // positionsNext + copy gives); vgroup: per vertex parallel group (-1: not solved)
    const float4* snap;
    const int* vgroup;
    // 0: cloth contact model (VF / FV / EE repulsion, bounds and truncation); 1: penetration
    // contact of tet meshes (the vertex's own hit face in its first VF record, the hits on its
    // incident faces in the FV lists; force k d n + friction, no EE, no bounds)
    int model;
    float penStiffness;  // collisionStiffness of the penetration model
};

// Device view of one scene, referenced by work items of a (batched) launch.
struct SceneView {
    float4* pos;
    float4* prev;
    float4* inertia;
    float4* vel;
    const float4* vparam;
    const uint8_t* vflag;
    const int4* slots;
    const TetInc* tinc;
    const float* tvol;   // per incidence record: its tet's rest volume
    const float4* tmat;  // per tet mesh: (mu, lambda, a, 0)
    // cloth
    float4* pnext;  // Jacobi target of cloth solves (positionsNext)
    const ClothSlot* cslots;
    const FaceInc* cface;
    const BendInc* cbend;
    int32_t nv;
    int32_t hasContact;
    SolveParams sp;
    ContactView cv;
    const float* initBound;  // per vertex bound of initialisation type 3 (nullptr: unbounded)
    const int4* tcontact;    // contact scenes, per tet slot: [tFaceRef range), [tEdgeRef range)
    // Chebyshev acceleration (scenes with use_accelerator; nullptr otherwise): positions two and
    // one iterations back, and omega per iteration
    float4* accPrevPrev;
    float4* accPrevIter;
    const float* accOmega;
    // scene time (curTime, float as in the reference), advanced by the initial step
    float* clock;
    // deformers: deformed vertices (cloth meshes' first, then tet meshes'), per-deformer parameters
    const DeformVert* dverts;
    const DeformParams* dparams;
    // cloth attachments: records of scene vertex v are att[attBegin[v] .. attBegin[v + 1])
    const int* attBegin;
    const AttachRec* att;
    // no-gravity zone: per vertex (noGravZoneThreshold, maxVelocityMagnitude) of its mesh, read for
    // vertices flagged VFLAG_NOGRAV_POS / VFLAG_NOGRAV_VEL (nullptr when no mesh has a zone)
    const float2* vzone;
    // count of vertex solves whose 3x3 block could not be solved (gradient step or no move)
    int* solveFail;
};

// vflag bits: 1 fixed, (code << 1) initialisation type, 16 accelerated (tet vertex of a scene
// with use_accelerator), 32 / 64 no-gravity zone on the position's / the velocity's y (tet / cloth
// meshes with hasNoGravZone and maxVelocityMagnitude > 0; SceneView::vzone holds the constants)
constexpr uint8_t VFLAG_FIXED = 1;
constexpr uint8_t VFLAG_ACCEL = 16;
constexpr uint8_t VFLAG_NOGRAV_POS = 32;
constexpr uint8_t VFLAG_NOGRAV_VEL = 64;

// Workgroups are dealt to the 8 XCDs round-robin (block b runs on XCD b % 8), each XCD with its
// own L2. This index maps the blocks of one XCD to one contiguous range of [0, gridDim.x), so the
// work an XCD has in flight at a time is a run of neighbouring items. A bijection.
__device__ __forceinline__ int xcd_block_index() {
    constexpr int kXcds = 8;
    const int nb = (int)gridDim.x, b = (int)blockIdx.x;
    const int xcd = b % kXcds, j = b / kXcds, q = nb / kXcds, rem = nb % kXcds;
    return xcd < rem ? xcd * (q + 1) + j : rem * (q + 1) + (xcd - rem) * q + j;
}

// The same with the range cut into chunks of `chunk` consecutive items dealt to the XCDs in turn
// (chunk 16: items 0-15 on XCD 0, 16-31 on XCD 1, ...), so neighbouring items share an L2 while
// the XCDs' shares of a long range stay balanced; the tail after the last full round of 8 chunks
// keeps block order. A bijection.
__device__ __forceinline__ int xcd_chunk_index(int chunk) {
    constexpr int kXcds = 8;
    const int b = (int)blockIdx.x;
    const int round = kXcds * chunk;
    if (b >= ((int)gridDim.x / round) * round) return b;
    const int xcd = b % kXcds, j = b / kXcds;
    return ((j / chunk) * kXcds + xcd) * chunk + j % chunk;
}

// One block's share of a launch: slots [slotBegin, slotBegin + count) of scene `scene`
// (tet slots or cloth slots by `kind`; vertex ranges for the per-vertex kernels).
struct WorkItem {
    int32_t scene;
    int32_t slotBegin;
    int32_t count;
    int32_t kind;
};

// Entry of the flattened (scene, colour, vertex-chunk) work queue of one iteration.
// kind: WORK_TET / WORK_CLOTH solve chunk, or WORK_COMMIT (pnext -> pos for the cloth
// slots of one scene colour). The entry may start once `need` earlier entries of its
// scene have completed in this iteration; entries are ordered by (colour level, scene),
// so every dependency lies earlier in the queue.
struct QueueEntry {
    int32_t scene;
    int32_t slotBegin;
    int32_t count;
    int32_t kind;
    int32_t need;
    int32_t perIteration;  // entries of this scene per iteration (progress counters accumulate)
    int32_t pad[2];
};

enum : int32_t { WORK_COMMIT = 2 };

}  // namespace gvbd
