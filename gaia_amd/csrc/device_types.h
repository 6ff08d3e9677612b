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
//   tinc[ni]     TetInc  colour-major vertex->tet incidence records (64 B each, streamed)
// Slots of parallel group g are slots[groupOffset[g] .. groupOffset[g+1]); the incidence
// records of consecutive slots are consecutive, so a block's records form one range.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gvbd {

// One incident tet of one vertex v, expressed relative to v:
//   F = sum_k (x_nb[k] - x_v) m_k^T,  m_v = -(m_0 + m_1 + m_2)
// where m_k is the shape-matrix row of the k-th other vertex (rows of DmInv, or
// minus their sum for tet vertex 0). Material constants are folded in:
//   Amu = restVolume * mu, Alm = restVolume * lambda, a = 1 + mu / lambda.
struct alignas(16) TetInc {
    int32_t nb[3];
    float Amu;
    float m[9];  // m_0 (3), m_1 (3), m_2 (3)
    float Alm;
    float a;
    float pad;
};
static_assert(sizeof(TetInc) == 64, "TetInc must be 64 bytes");

enum SlotFlags : int32_t {
    SLOT_DAMPING = 1,  // the vertex's mesh has non-zero material damping
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
    int32_t nv;
    int32_t pad;
    SolveParams sp;
};

// One block's share of a parallel-group launch: slots [slotBegin, slotBegin + count) of scene `scene`.
struct WorkItem {
    int32_t scene;
    int32_t slotBegin;
    int32_t count;
    int32_t pad;
};

}  // namespace gvbd
