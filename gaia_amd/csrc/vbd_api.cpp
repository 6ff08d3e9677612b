// C-ABI implementation: contexts, scenes, batched plans and HIP graph capture.
// Declarations and reference citations: include/gaia_vbd.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstddef>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <functional>
#include <limits>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/gaia_vbd.h"
#include "detector.h"
#include "device_types.h"
#include "errors.h"
#include "kernels.h"
#include "mesh_prep.h"
#include "scene_io.h"
#include "scene_loader.h"

namespace gvbd {

namespace {

thread_local std::string g_lastError;

constexpr int kVertexChunk = 256;
constexpr int kSolveFailSlot = 16;  // int index of the solve-failure counter in the clock block

// All device arrays of one scene live in one allocation.
struct DeviceScene {
    void* arena = nullptr;
    size_t bytes = 0;
    float4 *pos = nullptr, *prev = nullptr, *inertia = nullptr, *vel = nullptr, *vparam = nullptr, *pnext = nullptr;
    uint8_t* vflag = nullptr;
    int4* slots = nullptr;
    TetInc* tinc = nullptr;
    float* tvol = nullptr;     // per incidence record: rest volume of its tet
    float4* tmat = nullptr;    // per tet mesh: (mu, lambda, a, 0)
    ClothSlot* cslots = nullptr;
    FaceInc* cface = nullptr;
    BendInc* cbend = nullptr;
    int* faceRef = nullptr;     // contact scenes: per face record, face id | order << 30
    int* edgeRef = nullptr;     // contact scenes: per slot incident edge, edge id | order << 30
    float2* friction = nullptr; // contact scenes: per vertex (frictionDynamic, frictionEpsV)
    int* tFaceRef = nullptr;    // contact scenes: per tet slot incident boundary face, face id | order << 30
    int* tEdgeRef = nullptr;    // contact scenes: per tet slot incident boundary edge, edge id | order << 30
    int4* tcontact = nullptr;   // contact scenes: per tet slot, its tFaceRef and tEdgeRef ranges
    float4* snap = nullptr;     // contact scenes: positions at the start of the iteration
    int* vgroup = nullptr;      // contact scenes: per vertex parallel group (-1: not solved)
    float4 *accPrevPrev = nullptr, *accPrevIter = nullptr;  // use_accelerator: iteration history
    float* accOmega = nullptr;                              // use_accelerator: omega per iteration
    float* clock = nullptr;                                 // scene time
    float2* vzone = nullptr;                                // no-gravity zone constants per vertex
    DeformVert* dverts = nullptr;                           // deformed vertices (deformers)
    DeformParams* dparams = nullptr;
    int* attBegin = nullptr;                                // cloth attachments: per vertex CSR
    AttachRec* att = nullptr;
};

// A deformer as added (gvbd_scene_add_deformer): parameters and, per selected mesh, its vertices
// with the constants taken at construction (root / initial position, RootP).
struct DeformerHost {
    DeformParams prm{};
    std::vector<int32_t> meshes;
    std::vector<std::vector<int32_t>> verts;
    std::vector<std::vector<float>> a, b;  // 3 per vertex
};

// Boundary surface of a tet mesh in a scene with collision handling: its contact faces, their
// edges (numbered as the detector numbers them) and, per vertex, the incident ones with the
// vertex's order in each.
struct TetSurface {
    std::vector<int32_t> faces;  // 3 per face, outward
    std::vector<int32_t> owner, tetNbr;  // penetration contact: owning (tet, local face); tet across each local face
    std::vector<TriEdge> edges;
    std::vector<int32_t> vfBegin, vfRef;  // vertex -> (face | order << 30)
    std::vector<int32_t> veBegin, veRef;  // vertex -> (edge | order << 30), ascending edge id
    std::vector<uint8_t> surface;         // vertex on the boundary
};

enum MeshKind { MESH_TET = 0, MESH_CLOTH = 1 };

// Page-locked host memory (hipHostMalloc) that a served scene's arena is copied out to: copies
// between it and HBM run at the link's rate without the runtime's staging through a pageable
// bounce buffer. The allocation is kept when the contents are copied back (size() drops to 0),
// so a scene that is copied out again reuses it.
class PinnedBuffer {
  public:
    PinnedBuffer() = default;
    PinnedBuffer(const PinnedBuffer&) = delete;
    PinnedBuffer& operator=(const PinnedBuffer&) = delete;
    ~PinnedBuffer() {
        if (p_) (void)hipHostFree(p_);
    }
    void resize(size_t n) {
        if (n > cap_) {
            if (p_) (void)hipHostFree(p_);
            p_ = nullptr;
            cap_ = 0;
            if (hipHostMalloc((void**)&p_, n, hipHostMallocDefault) != hipSuccess) {
                (void)hipGetLastError();
                p_ = nullptr;
                fail(GVBD_ERR_NOMEM, "page-locked host buffer of " + std::to_string(n) + " bytes");
            }
            cap_ = n;
        }
        n_ = n;
    }
    void clear() { n_ = 0; }
    bool empty() const { return n_ == 0; }
    size_t size() const { return n_; }
    char* data() { return p_; }
    const char* data() const { return p_; }

  private:
    char* p_ = nullptr;
    size_t n_ = 0, cap_ = 0;
};

struct SceneObj {
    gvbd_physics_params params{};
    SolveParams sp{};
    std::vector<TetMeshHost> tets;
    std::vector<ClothMeshHost> cloths;
    std::vector<std::pair<int, int>> meshes;  // (kind, index into tets / cloths), in addition order
    std::vector<int32_t> vertexOffset;        // per mesh, into the scene's vertex space
    int32_t nv = 0;
    bool initialized = false;
    std::vector<std::vector<int32_t>> groups;  // interleaved (mesh, vertex)
    std::vector<int32_t> tetGroupOffset;       // G + 1, into tet slots
    std::vector<int32_t> clothGroupOffset;     // G + 1, into cloth slots
    std::vector<int64_t> slotWorkBytes;        // algorithmic bytes per tet slot then per cloth slot
    std::vector<uint8_t> clothSlotInplace;     // per cloth slot: solved in place (no commit pass)
    std::vector<TetSurface> tetSurfaces;       // per tet mesh, scenes with collision handling
    DeviceScene dev;
    SceneView view{};
    int32_t plan = -1;  // the scene's own single-scene plan
    // cloth collision handling (handle_collision): detector over the scene's positions
    std::unique_ptr<ContactDetector> cd;
    int64_t frame = 0;
    gvbd_contact_stats contactStats{};       // gvbd_scene_contact_stats
    void* convBuf = nullptr;                 // convergence evaluation: tet records + partial sums
    size_t convBytes = 0;
    hipEvent_t cdEvents[4] = {nullptr, nullptr, nullptr, nullptr};  // around a frame's detection passes
    bool detectPending = false;  // penetration frames: cdEvents hold a detection not yet timed
    // serving loop residency: the server that owns it (-1: none) and, while copied out to host
    // memory, the arena's bytes and the device address they were copied from
    int32_t server = -1;
    PinnedBuffer hostCopy;
    uintptr_t hostCopyBase = 0;
    std::vector<DeformerHost> deformers;            // gvbd_scene_add_deformer
    std::vector<gvbd_attachment_desc> attachments;  // gvbd_scene_add_attachments
    float time0 = 0.f;                              // scene time at initialisation
    int32_t nDeformPre = 0, nDeformPost = 0;        // deformed vertices of cloth / tet meshes

    int32_t meshVertices(int32_t m) const {
        return meshes[m].first == MESH_TET ? tets[meshes[m].second].nv : cloths[meshes[m].second].nv;
    }
    size_t numTetSlots() const { return tetGroupOffset.empty() ? 0 : (size_t)tetGroupOffset.back(); }
};

constexpr int kCommitChunk = 1024;
constexpr int kDefaultSpinLimit = 1 << 20;

struct PlanObj {
    std::vector<int32_t> scenes;
    int lpv = 4, block = 256;
    int stage = 0;  // incidence records per LDS staging chunk (0: read straight from HBM)
    int xcd = 0;    // group launches: each XCD's blocks take one contiguous range of the work items
    int mode = GVBD_MODE_GROUP_LAUNCHES;
    int vertsPerEntry = 0;  // work-queue entry size (0: 4 passes of block / lpv vertices)
    int blocksPerCu = 0;    // persistent grid per CU (0: resident maximum)
    int spinLimit = kDefaultSpinLimit;
    int numGroups = 0, iterations = 0, substeps = 0;
    SceneView* dViews = nullptr;
    WorkItem* dItems = nullptr;   // per group: solve items (tet and cloth)
    WorkItem* dCommit = nullptr;  // per group: cloth commit items
    WorkItem* dVItems = nullptr;  // per-vertex items
    std::vector<int32_t> groupItemBegin;    // G + 1
    std::vector<int32_t> groupContactBegin;  // G: items of scenes with collision handling start here
    std::vector<int32_t> groupAt;            // scene x launch: the scene's parallel group solved there, or -1
    std::vector<int32_t> groupCommitBegin;  // G + 1
    int32_t nVItems = 0;
    WorkItem* dDeformPre = nullptr;   // deformed vertices of cloth meshes (before the initial step)
    WorkItem* dDeformPost = nullptr;  // deformed vertices of tet meshes (after it)
    int32_t nDeformPre = 0, nDeformPost = 0;
    // work queue (mode GVBD_MODE_WORK_QUEUE)
    QueueEntry* dQueue = nullptr;
    int* dCounters = nullptr;  // tickets[iterations] then progress[scenes]
    int32_t nQueue = 0, nCounters = 0, queueGrid = 0;
    int32_t ticketSlots = 0;  // queue launches allowed between counter resets
    int32_t epoch = 0;        // queue launches since the last reset (host-issued iterations)
    size_t bytes = 0;  // device bytes held (the capacity of dBuf)
    // every device array above lives in one buffer, uploaded in one copy from a pinned host
    // staging buffer of the same capacity; both grow on demand and are kept across rebuilds, so a
    // rebuild (a serving tick with a new scene set) allocates and synchronises nothing
    char* dBuf = nullptr;
    char* hBuf = nullptr;
    size_t cap = 0;
    hipEvent_t uploadDone = nullptr;  // the last copy out of hBuf
    bool wantGraph = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool dirty = true;
    bool capturing = false;  // a frame is being captured into the plan's graph (no events, no host bookkeeping)
    bool contact = false;      // some scene has collision handling (contact kernel variant for its items)
    bool clothContact = false;  // a cloth-model contact scene: re-detection decided per iteration (on the
                                // host for a scene stepped alone without a graph, on the device otherwise)
    bool accel = false;    // some scene uses the Chebyshev accelerator
};

// Serving loop: registered scenes, their queued step requests, and the plan of one tick.
struct ServerObj {
    using Clock = std::chrono::steady_clock;
    struct Request {
        int64_t tag;
        Clock::time_point at;
    };
    struct Entry {
        std::deque<Request> pending;  // the first `launched` have a frame in flight
        int32_t launched = 0;
        int64_t lastUse = -1;  // tick of the scene's last frame: idle scenes leave HBM in this order
        int32_t frames = 0;
        int64_t needBytes = 0;  // device bytes, once known
    };
    // A launched frame: its scenes, the plan slot whose buffers its kernels read, launch time.
    struct Frame {
        std::vector<int32_t> scenes;
        int slot;
        Clock::time_point launched;
    };
    int64_t budget = 0;
    int iterations = -1, substeps = -1;
    // Two plans used alternately, so that one tick's frame can be built and queued while the
    // previous one still runs; each slot has the completion event of its last frame.
    PlanObj plans[2];
    int64_t planLayout[2] = {-1, -1};  // `layout` when each plan's device views were built
    hipEvent_t frameDone[2] = {nullptr, nullptr};
    int nextSlot = 0;
    std::deque<Frame> inflight;
    hipEvent_t ref = nullptr;  // completed at creation; frame completion times are measured from it
    Clock::time_point refTime;
    int64_t layout = 0;  // bumped whenever a scene enters or leaves HBM (device addresses move)
    std::map<int32_t, Entry> scenes;
    std::deque<gvbd_served_step> done;
    gvbd_server_stats stats{};
};

}  // namespace
}  // namespace gvbd

struct gvbd_context {
    int device = 0;
    int numCUs = 0;
    hipStream_t stream = nullptr;
    int* dStatus = nullptr;  // set by work-queue kernels when a dependency wait exceeds its limit
    double* convScratch = nullptr;  // this device's scratch of the allocation-free convergence evaluation
    hipEvent_t events[64] = {};
    std::string lastError;
    std::map<int32_t, std::unique_ptr<gvbd::SceneObj>> scenes;
    std::map<int32_t, std::unique_ptr<gvbd::PlanObj>> plans;
    std::map<int32_t, int32_t> batches;  // batch id -> plan id
    std::map<int32_t, std::unique_ptr<gvbd::ContactDetector>> detectors;
    std::map<int32_t, std::unique_ptr<gvbd::ServerObj>> servers;
    int32_t nextScene = 0, nextPlan = 0, nextBatch = 0, nextDetector = 0, nextServer = 0;
};

namespace gvbd {
namespace {

SceneObj& scene_of(gvbd_context* ctx, int32_t id) {
    auto it = ctx->scenes.find(id);
    if (it == ctx->scenes.end()) fail(GVBD_ERR_INVALID, "unknown scene id " + std::to_string(id));
    return *it->second;
}

PlanObj& plan_of(gvbd_context* ctx, int32_t id) {
    auto it = ctx->plans.find(id);
    if (it == ctx->plans.end()) fail(GVBD_ERR_INVALID, "unknown plan id " + std::to_string(id));
    return *it->second;
}

PlanObj& batch_plan(gvbd_context* ctx, int32_t batch) {
    auto it = ctx->batches.find(batch);
    if (it == ctx->batches.end()) fail(GVBD_ERR_INVALID, "unknown batch id " + std::to_string(batch));
    return plan_of(ctx, it->second);
}

SolveParams make_solve_params(const gvbd_physics_params& p, const SceneObj& s) {
    SolveParams sp{};
    if (p.num_substeps <= 0) fail(GVBD_ERR_INVALID, "num_substeps must be positive");
    sp.dt = p.time_step / (float)p.num_substeps;
    sp.dtSqrRecip = 1.f / (sp.dt * sp.dt);
    for (int d = 0; d < 3; ++d) {
        sp.gravity[d] = p.gravity[d];
        sp.lower[d] = p.world_lower[d];
        sp.upper[d] = p.world_upper[d];
    }
    sp.boundaryStiffness = p.boundary_collision_stiffness;
    sp.frictionMu = p.boundary_friction_dynamic;
    sp.frictionEpsU = p.boundary_friction_eps_v * sp.dt;
    sp.stepSize = p.step_size;
    sp.stepSizeGD = p.step_size_gd;
    sp.usePlaneGround = p.use_plane_ground;
    sp.associateGravity = p.associate_gravity_with_inertia;
    sp.frictionStartIter = p.friction_start_iter;
    sp.initRatioG = s.tets.empty() ? 0.5f : s.tets[0].mat.initRatioG;
    sp.clothBoundaryStiffness = p.cloth_boundary_contact_stiffness;
    sp.clothFrictionEpsU = p.cloth_boundary_friction_eps * sp.dt;
    sp.degenerateThresSq = p.degenerate_triangle_thres * p.degenerate_triangle_thres;
    sp.clothStepSize = p.cloth_step_size;
    sp.contactStiffness = p.contact_stiffness;
    sp.contactRadius = p.contact_radius;
    sp.thickness = p.thickness;
    sp.contactEnergyType = p.contact_energy_type;
    return sp;
}

void free_plan_graph(PlanObj& p) {
    if (p.exec) (void)hipGraphExecDestroy(p.exec);
    if (p.graph) (void)hipGraphDestroy(p.graph);
    p.exec = nullptr;
    p.graph = nullptr;
}

// Marks a plan for rebuilding (its scene set, schedule or scene addresses changed); its buffers
// stay allocated for the rebuild.
void free_plan(PlanObj& p) {
    free_plan_graph(p);
    p.dDeformPre = p.dDeformPost = nullptr;
    p.nDeformPre = p.nDeformPost = 0;
    p.dViews = nullptr;
    p.dItems = nullptr;
    p.dCommit = nullptr;
    p.dVItems = nullptr;
    p.dQueue = nullptr;
    p.dCounters = nullptr;
    p.nQueue = p.nCounters = p.queueGrid = p.ticketSlots = p.epoch = 0;
    p.dirty = true;
}

// Frees a plan's buffers (the plan is being destroyed; its frames must be complete).
void release_plan(PlanObj& p) {
    free_plan(p);
    if (p.dBuf) (void)hipFree(p.dBuf);
    if (p.hBuf) (void)hipHostFree(p.hBuf);
    if (p.uploadDone) (void)hipEventDestroy(p.uploadDone);
    p.dBuf = p.hBuf = nullptr;
    p.uploadDone = nullptr;
    p.cap = p.bytes = 0;
}

// Work-queue status: reports (and clears) a dependency wait that hit its spin limit.
void check_queue_status(gvbd_context* ctx) {
    int st = 0;
    hip_check(hipMemcpy(&st, ctx->dStatus, sizeof(int), hipMemcpyDeviceToHost), "status read");
    if (st != 0) {
        hip_check(hipMemset(ctx->dStatus, 0, sizeof(int)), "status clear");
        fail(GVBD_ERR_HIP, "work-queue dependency wait exceeded its spin limit; the affected steps are invalid");
    }
}

// Colour-major tet slot and incidence records of one vertex.
// Contact references of a tet vertex (scenes with collision handling): its incident boundary
// faces and edges as scene-global ids with the vertex's order, and the slot's ranges of them.
struct TetContactRefs {
    int32_t faceOff = 0, edgeOff = 0;
    const TetSurface* surf = nullptr;
    std::vector<int32_t>* faceRef = nullptr;
    std::vector<int32_t>* edgeRef = nullptr;
    std::vector<int4>* ranges = nullptr;
};

void emit_tet_slot(const TetMeshHost& m, int32_t off, int32_t lv, int32_t mat, std::vector<int4>& slots,
                   std::vector<TetInc>& tinc, std::vector<float>& tvol, std::vector<int64_t>& work,
                   const TetContactRefs* refs) {
    int4 slot;
    slot.x = off + lv;
    slot.y = (int32_t)tinc.size();
    for (int32_t k = m.incBegin[lv]; k < m.incBegin[lv + 1]; ++k) {
        const int32_t t = m.incTet[k];
        const int32_t ord = m.incOrder[k];
        const float* D = &m.dminv[9 * (size_t)t];
        // shape-matrix row j of tet vertex j: j = 0 -> minus column sums, j >= 1 -> row j-1 of DmInv
        float B[4][3];
        for (int c = 0; c < 3; ++c) {
            B[0][c] = -D[0 + 3 * c] - D[1 + 3 * c] - D[2 + 3 * c];
            for (int j = 1; j < 4; ++j) B[j][c] = D[(j - 1) + 3 * c];
        }
        TetInc rec{};
        int n = 0;
        for (int j = 0; j < 4; ++j) {
            if (j == ord) continue;
            rec.nb[n] = off + m.tets[4 * (size_t)t + j];
            for (int c = 0; c < 3; ++c) rec.m[3 * n + c] = B[j][c];
            ++n;
        }
        tinc.push_back(rec);
        tvol.push_back(m.vol[t]);
    }
    slot.z = (int32_t)tinc.size();
    const bool damping = m.mat.dampingHydrostatic != 0.f || m.mat.dampingDeviatoric != 0.f;
    slot.w = (damping ? SLOT_DAMPING : 0) | (mat << SLOT_MAT_SHIFT);
    if (refs) {
        const TetSurface& S = *refs->surf;
        int4 r;
        r.x = (int32_t)refs->faceRef->size();
        for (int32_t k = S.vfBegin[lv]; k < S.vfBegin[lv + 1]; ++k) {
            const int32_t ref = S.vfRef[k];
            refs->faceRef->push_back((refs->faceOff + (ref & 0x3FFFFFFF)) | (ref & ~0x3FFFFFFF));
        }
        r.y = (int32_t)refs->faceRef->size();
        r.z = (int32_t)refs->edgeRef->size();
        for (int32_t k = S.veBegin[lv]; k < S.veBegin[lv + 1]; ++k) {
            const int32_t ref = S.veRef[k];
            refs->edgeRef->push_back((refs->edgeOff + (ref & 0x3FFFFFFF)) | (ref & ~0x3FFFFFFF));
        }
        r.w = (int32_t)refs->edgeRef->size();
        refs->ranges->push_back(r);
    }
    slots.push_back(slot);
    // gather model: 72 B per vertex (slot, material, own position, previous position, inertia
    // target, result) + 88 B per incident tet (48 B record, 4 B rest volume, 3 x 12 B neighbours)
    work.push_back(72 + 88 * (int64_t)(slot.z - slot.y));
}

// Colour-major cloth slot, face and bending records of one vertex.
// Contact references of a cloth vertex (scene-global face / edge ids with the vertex's order
// in the primitive); only built for scenes with collision handling.
struct ClothContactRefs {
    int32_t faceOff = 0, edgeOff = 0;
    const std::vector<std::vector<int32_t>>* vertexEdges = nullptr;  // per local vertex, ascending
    std::vector<int32_t>* faceRef = nullptr;
    std::vector<int32_t>* edgeRef = nullptr;
};

void emit_cloth_slot(const ClothMeshHost& m, int32_t off, int32_t lv, bool inplace, std::vector<ClothSlot>& slots,
                     std::vector<FaceInc>& faces, std::vector<BendInc>& bends, std::vector<int64_t>& work,
                     const ClothContactRefs* refs) {
    ClothSlot cs{};
    cs.a.x = off + lv;
    cs.a.y = (int32_t)faces.size();
    for (int32_t k = m.vfBegin[lv]; k < m.vfBegin[lv + 1]; ++k) {
        const int32_t f = m.vfFace[k], o = m.vfOrder[k];
        if (refs) refs->faceRef->push_back((refs->faceOff + f) | (o << 30));
        const float* D = &m.dminv[4 * (size_t)f];  // (0,0) (1,0) (0,1) (1,1)
        // shape row of face vertex j: j = 0 -> minus the sum of DmInv rows, j >= 1 -> row j-1
        float B[3][2];
        for (int c = 0; c < 2; ++c) {
            B[0][c] = -D[0 + 2 * c] - D[1 + 2 * c];
            B[1][c] = D[0 + 2 * c];
            B[2][c] = D[1 + 2 * c];
        }
        FaceInc rec{};
        int n = 0;
        for (int j = 0; j < 3; ++j) {
            if (j == o) continue;
            rec.nb[n] = off + m.faces[3 * (size_t)f + j];
            rec.m[2 * n] = B[j][0];
            rec.m[2 * n + 1] = B[j][1];
            ++n;
        }
        rec.Amu = m.area[f] * m.mat.mu;
        rec.Alm = m.area[f] * m.mat.lambda;
        faces.push_back(rec);
    }
    cs.a.z = (int32_t)faces.size();
    cs.a.w = inplace ? SLOT_INPLACE : 0;
    cs.b.x = (int32_t)bends.size();
    const float kb = m.mat.bendingStiffness;
    for (int32_t k = m.vbBegin[lv]; k < m.vbBegin[lv + 1]; ++k) {
        const int32_t id = m.vbEdge[k], o = m.vbOrder[k];
        const TriEdge& e = m.edges[id];
        const int32_t roles[4] = {e.v1, e.v2, e.v12Next, e.v21Next};
        const float* Q = &m.Q[16 * (size_t)id];
        BendInc rec{};
        int n = 0;
        for (int j = 0; j < 4; ++j) {
            if (j == o) continue;
            rec.nb[n] = off + roles[j];
            rec.q[n] = kb * Q[4 * o + j];
            ++n;
        }
        rec.order = o;
        rec.qself = kb * Q[4 * o + o];
        bends.push_back(rec);
    }
    cs.b.y = (int32_t)bends.size();
    if (refs) {
        cs.b.z = (int32_t)refs->edgeRef->size();
        for (int32_t e : (*refs->vertexEdges)[lv])
            refs->edgeRef->push_back((refs->edgeOff + e) | ((m.edges[e].v1 == lv ? 0 : 1) << 30));
        cs.b.w = (int32_t)refs->edgeRef->size();
    }
    slots.push_back(cs);
    // gather model: 60 B per vertex + 56 B per face + 116 B per bending stencil
    work.push_back(60 + 56 * (int64_t)(cs.a.z - cs.a.y) + 116 * (int64_t)(cs.b.y - cs.b.x));
}

// Builds the device arrays of an initialised scene. `reserve` (when set) sees the arena's size
// before it is allocated and may refuse it by throwing.
void upload_scene(gvbd_context* ctx, SceneObj& s, const std::function<void(size_t)>& reserve) {
    const int32_t G = (int32_t)s.groups.size();
    std::vector<int4> slots;
    std::vector<TetInc> tinc;
    std::vector<float> tvol;
    std::vector<ClothSlot> cslots;
    std::vector<FaceInc> cface;
    std::vector<BendInc> cbend;
    std::vector<int64_t> tetWork, clothWork;
    s.tetGroupOffset.assign(G + 1, 0);
    s.clothGroupOffset.assign(G + 1, 0);
    // fixed vertices are not solved (VBDPhysics.cpp:1617, VBDClothPhysics.cpp:470)
    std::vector<std::vector<uint8_t>> fixedMask(s.meshes.size());
    for (size_t im = 0; im < s.meshes.size(); ++im) {
        fixedMask[im].assign(s.meshVertices((int32_t)im), 0);
        const auto& fx = s.meshes[im].first == MESH_TET ? s.tets[s.meshes[im].second].fixed
                                                        : s.cloths[s.meshes[im].second].fixed;
        for (int32_t f : fx) fixedMask[im][f] = 1;
    }
    // contact references (collision handling): scene-global face and edge ids follow the mesh
    // order, as the detector numbers them
    const bool contact = s.params.handle_collision != 0;
    const bool accel = s.params.use_accelerator != 0 && !s.tets.empty();
    std::vector<int32_t> faceRef, edgeRef, tFaceRef, tEdgeRef;
    std::vector<int4> tcontact;
    std::vector<ClothContactRefs> refs(s.meshes.size());
    std::vector<TetContactRefs> trefs(s.meshes.size());
    std::vector<std::vector<std::vector<int32_t>>> vertexEdges(s.meshes.size());
    if (contact) {
        int32_t fo = 0, eo = 0;
        for (size_t im = 0; im < s.meshes.size(); ++im) {
            if (s.meshes[im].first == MESH_TET) {
                const TetSurface& S = s.tetSurfaces[s.meshes[im].second];
                trefs[im] = TetContactRefs{fo, eo, &S, &tFaceRef, &tEdgeRef, &tcontact};
                fo += (int32_t)S.faces.size() / 3;
                eo += (int32_t)S.edges.size();
                continue;
            }
            const ClothMeshHost& m = s.cloths[s.meshes[im].second];
            vertexEdges[im].assign(m.nv, {});
            for (int32_t e = 0; e < (int32_t)m.edges.size(); ++e) {
                vertexEdges[im][m.edges[e].v1].push_back(e);
                vertexEdges[im][m.edges[e].v2].push_back(e);
            }
            refs[im] = ClothContactRefs{fo, eo, &vertexEdges[im], &faceRef, &edgeRef};
            fo += m.nf;
            eo += (int32_t)m.edges.size();
        }
    }
    s.clothSlotInplace.clear();
    // vertex-face attachments (loadConstraints, VBDClothPhysics.cpp:148-224): the vertex side gets
    // b = 1, each face vertex with a non-zero barycentric b = -bary (its position read from the
    // vertex-side mesh at the face vertex's own index, VBDClothPhysics.cpp:1471); records of one
    // vertex keep the constraint order
    std::vector<std::vector<AttachRec>> attOf;
    if (!s.attachments.empty()) {
        attOf.assign(s.nv, {});
        for (const gvbd_attachment_desc& A : s.attachments) {
            const int32_t offV = s.vertexOffset[A.mesh];
            AttachRec r{};
            r.ids = make_int4(-1, -1, -1, offV + A.vertex);
            if (A.target_mesh >= 0) {
                const int32_t offT = s.vertexOffset[A.target_mesh];
                r.ids = make_int4(offT + A.face[0], offT + A.face[1], offT + A.face[2], offV + A.vertex);
            }
            r.p = make_float4(A.bary[0], A.bary[1], A.bary[2], A.stiffness * 1.f);
            r.b = 1.f;
            attOf[offV + A.vertex].push_back(r);
            if (A.target_mesh < 0) continue;
            for (int j = 0; j < 3; ++j) {
                if (A.bary[j] == 0.f) continue;
                AttachRec q = r;
                q.ids.w = offV + A.face[j];
                q.b = -A.bary[j];
                q.p.w = A.stiffness * q.b;
                attOf[s.vertexOffset[A.target_mesh] + A.face[j]].push_back(q);
            }
        }
    }
    std::vector<int32_t> attBegin;
    std::vector<AttachRec> att;
    if (!attOf.empty()) {
        attBegin.push_back(0);
        for (const auto& l : attOf) {
            att.insert(att.end(), l.begin(), l.end());
            attBegin.push_back((int32_t)att.size());
        }
    }
    for (int32_t g = 0; g < G; ++g) {
        const auto& grp = s.groups[g];
        // A cloth mesh with a single colour in this group can be updated in place: its solved
        // vertices share no element, so the reference's positionsNext + copy (Jacobi over the
        // group) gives the same result. Without contact, meshes do not interact.
        std::vector<int32_t> firstColor(s.meshes.size(), -1);
        std::vector<uint8_t> multiColor(s.meshes.size(), 0);
        for (size_t i = 0; i + 1 < grp.size(); i += 2) {
            const int32_t im = grp[i], lv = grp[i + 1];
            if (s.meshes[im].first != MESH_CLOTH) continue;
            const int32_t c = s.cloths[s.meshes[im].second].colors[lv];
            if (firstColor[im] < 0) firstColor[im] = c;
            else if (firstColor[im] != c) multiColor[im] = 1;
        }
        for (size_t i = 0; i + 1 < grp.size(); i += 2) {
            const int32_t im = grp[i], lv = grp[i + 1];
            if (fixedMask[im][lv]) continue;
            const int32_t off = s.vertexOffset[im];
            if (s.meshes[im].first == MESH_TET) {
                emit_tet_slot(s.tets[s.meshes[im].second], off, lv, s.meshes[im].second, slots, tinc, tvol,
                              tetWork, contact ? &trefs[im] : nullptr);
            } else {
                // attachments read other vertices of the group: positionsNext keeps them Jacobi
                const bool inplace = !multiColor[im] && !s.params.handle_collision && attOf.empty();
                emit_cloth_slot(s.cloths[s.meshes[im].second], off, lv, inplace, cslots, cface, cbend, clothWork,
                                contact ? &refs[im] : nullptr);
                if (!attOf.empty() && !attOf[off + lv].empty()) cslots.back().a.w |= SLOT_ATTACH;
                s.clothSlotInplace.push_back(inplace ? 1 : 0);
            }
        }
        s.tetGroupOffset[g + 1] = (int32_t)slots.size();
        s.clothGroupOffset[g + 1] = (int32_t)cslots.size();
    }
    s.slotWorkBytes = tetWork;
    s.slotWorkBytes.insert(s.slotWorkBytes.end(), clothWork.begin(), clothWork.end());
    // per-vertex arrays
    const int32_t nv = s.nv;
    std::vector<float4> pos(nv), inertia(nv), vel(nv), vparam(nv);
    std::vector<uint8_t> vflag(nv, 0);
    std::vector<float2> friction(contact ? nv : 0);
    std::vector<float2> vzone;  // (noGravZoneThreshold, maxVelocityMagnitude), when a mesh has a zone
    for (size_t im = 0; im < s.meshes.size(); ++im) {
        const int32_t off = s.vertexOffset[im];
        const std::vector<float>* mp;
        const std::vector<float>* mv;
        const std::vector<float>* mm;
        float4 vp;
        uint8_t code, zoneFlag;
        float2 zone;
        if (s.meshes[im].first == MESH_TET) {
            const TetMeshHost& m = s.tets[s.meshes[im].second];
            mp = &m.pos;
            mv = &m.vel;
            mm = &m.mass;
            // step-invariant damping (VBDPhysics.cpp:794-807)
            vp = make_float4(std::pow(m.mat.exponentialVelDamping, s.sp.dt), m.mat.constantVelDamping * s.sp.dt,
                             m.mat.dampingHydrostatic, m.mat.dampingDeviatoric);
            code = (uint8_t)(m.mat.initializationType & 7);
            zoneFlag = (m.mat.hasNoGravZone && m.mat.maxVelocityMagnitude > 0.f) ? VFLAG_NOGRAV_POS : 0;
            zone = make_float2(m.mat.noGravZoneThreshold, m.mat.maxVelocityMagnitude);
            if (contact)
                for (int32_t i = 0; i < m.nv; ++i)
                    friction[s.vertexOffset[im] + i] = make_float2(m.mat.frictionDynamic, m.mat.frictionEpsV);
        } else {
            const ClothMeshHost& m = s.cloths[s.meshes[im].second];
            mp = &m.pos;
            mv = &m.vel;
            mm = &m.mass;
            vp = make_float4(m.mat.exponentialVelDamping, m.mat.constantVelDamping, m.mat.dampingStVK,
                             m.mat.dampingBending);
            // cloth initialisation types (VBDBaseTriMesh.cpp:72-153) as per-vertex codes:
            // 0 inertia -> x = y; 2 previous position, 4 none -> keep x;
            // 3 inertia within the conservative bounds -> x + (y - x) truncated (code 5)
            const int t = m.mat.initializationType;
            code = (uint8_t)(t == 0 ? 0 : (t == 3 ? 5 : 1));
            zoneFlag = (m.mat.hasNoGravZone && m.mat.maxVelocityMagnitude > 0.f) ? VFLAG_NOGRAV_VEL : 0;
            zone = make_float2(m.mat.noGravZoneThreshold, m.mat.maxVelocityMagnitude);
            if (contact)
                for (int32_t i = 0; i < m.nv; ++i)
                    friction[s.vertexOffset[im] + i] = make_float2(m.mat.frictionDynamic, m.mat.frictionEpsV);
        }
        const int32_t n = s.meshVertices((int32_t)im);
        for (int32_t i = 0; i < n; ++i) {
            const float4 p = make_float4((*mp)[3 * i], (*mp)[3 * i + 1], (*mp)[3 * i + 2], 0.f);
            pos[off + i] = p;
            inertia[off + i] = make_float4(p.x, p.y, p.z, (*mm)[i]);
            vel[off + i] = make_float4((*mv)[3 * i], (*mv)[3 * i + 1], (*mv)[3 * i + 2], 0.f);
            vparam[off + i] = vp;
            vflag[off + i] = (uint8_t)(code << 1);
        }
        for (int32_t i = 0; i < n; ++i)
            if (fixedMask[im][i]) vflag[off + i] |= VFLAG_FIXED;
        if (zoneFlag) {
            vzone.resize(nv, make_float2(0.f, 0.f));
            for (int32_t i = 0; i < n; ++i) {
                vflag[off + i] |= zoneFlag;
                vzone[off + i] = zone;
            }
        }
        if (accel && s.meshes[im].first == MESH_TET)
            for (int32_t i = 0; i < n; ++i) vflag[off + i] |= VFLAG_ACCEL;
    }
    // omega per iteration (getAcceleratorOmega, VBDPhysics.cpp:3558-3575), in float as there
    std::vector<float> omega;
    if (accel) {
        const float rho2 = s.params.accelerator_rho * s.params.accelerator_rho;
        float w = 1.f;
        for (int k = 1; k <= std::max(s.params.iterations, 1); ++k) {
            w = k == 1 ? 1.f : (k == 2 ? 2.f / (2.f - rho2) : 4.f / (4.f - rho2 * w));
            omega.push_back(w);
        }
    }
    // deformed vertices: cloth meshes' first, then tet meshes' (the two launch phases)
    std::vector<DeformVert> dverts;
    std::vector<DeformParams> dparams;
    for (int phase = 0; phase < 2; ++phase)
        for (size_t di = 0; di < s.deformers.size(); ++di) {
            const DeformerHost& D = s.deformers[di];
            for (size_t k = 0; k < D.meshes.size(); ++k) {
                const bool cloth = s.meshes[D.meshes[k]].first == MESH_CLOTH;
                if (cloth != (phase == 0)) continue;
                for (size_t i = 0; i < D.verts[k].size(); ++i) {
                    DeformVert r{};
                    r.v = s.vertexOffset[D.meshes[k]] + D.verts[k][i];
                    r.d = (int32_t)di;
                    for (int c = 0; c < 3; ++c) {
                        r.a[c] = D.a[k][3 * i + c];
                        r.b[c] = D.b[k][3 * i + c];
                    }
                    dverts.push_back(r);
                }
            }
            if (phase == 0) s.nDeformPre = (int32_t)dverts.size();
        }
    if (s.deformers.empty()) s.nDeformPre = 0;
    s.nDeformPost = (int32_t)dverts.size() - s.nDeformPre;
    for (const DeformerHost& D : s.deformers) dparams.push_back(D.prm);
    // one arena
    auto align = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t bV = align(sizeof(float4) * (size_t)std::max(nv, 1));
    const size_t bF = align((size_t)std::max(nv, 1));
    const size_t bS = align(sizeof(int4) * std::max<size_t>(slots.size(), 1));
    const size_t bI = align(sizeof(TetInc) * std::max<size_t>(tinc.size(), 1));
    const size_t bIV = align(sizeof(float) * std::max<size_t>(tvol.size(), 1));
    std::vector<float4> tmat(std::max<size_t>(s.tets.size(), 1), make_float4(0.f, 0.f, 0.f, 0.f));
    for (size_t t = 0; t < s.tets.size(); ++t) {
        const auto& mt = s.tets[t].mat;
        tmat[t] = make_float4(mt.mu, mt.lambda, 1.f + mt.mu / mt.lambda, 0.f);
    }
    const size_t bTM = align(sizeof(float4) * tmat.size());
    const size_t bCS = align(sizeof(ClothSlot) * std::max<size_t>(cslots.size(), 1));
    const size_t bCF = align(sizeof(FaceInc) * std::max<size_t>(cface.size(), 1));
    const size_t bCB = align(sizeof(BendInc) * std::max<size_t>(cbend.size(), 1));
    const size_t bFR = contact ? align(sizeof(int32_t) * std::max<size_t>(faceRef.size(), 1)) : 0;
    const size_t bER = contact ? align(sizeof(int32_t) * std::max<size_t>(edgeRef.size(), 1)) : 0;
    const size_t bFr = contact ? align(sizeof(float2) * (size_t)std::max(nv, 1)) : 0;
    const size_t bTF = contact ? align(sizeof(int32_t) * std::max<size_t>(tFaceRef.size(), 1)) : 0;
    const size_t bTE = contact ? align(sizeof(int32_t) * std::max<size_t>(tEdgeRef.size(), 1)) : 0;
    const size_t bTC = contact ? align(sizeof(int4) * std::max<size_t>(tcontact.size(), 1)) : 0;
    const size_t bSn = contact ? bV : 0, bVG = contact ? align(sizeof(int32_t) * (size_t)std::max(nv, 1)) : 0;
    std::vector<int32_t> vgroup;
    if (contact) {
        vgroup.assign(nv, -1);
        for (int32_t g = 0; g < G; ++g)
            for (size_t i = 0; i + 1 < s.groups[g].size(); i += 2) {
                const int32_t im = s.groups[g][i], lv = s.groups[g][i + 1];
                if (!fixedMask[im][lv]) vgroup[s.vertexOffset[im] + lv] = g;
            }
    }
    const size_t bA = accel ? bV : 0, bW = accel ? align(sizeof(float) * omega.size()) : 0;
    const size_t bClk = 256;
    const size_t bZ = vzone.empty() ? 0 : align(sizeof(float2) * vzone.size());
    const size_t bDV = dverts.empty() ? 0 : align(sizeof(DeformVert) * dverts.size());
    const size_t bDP = dparams.empty() ? 0 : align(sizeof(DeformParams) * dparams.size());
    const size_t bAB = attBegin.empty() ? 0 : align(sizeof(int32_t) * attBegin.size());
    const size_t bAR = att.empty() ? 0 : align(sizeof(AttachRec) * att.size());
    const size_t total = 6 * bV + bF + bS + bI + bIV + bTM + bCS + bCF + bCB + bFR + bER + bFr + bTF + bTE + bTC + 2 * bA + bW +
                         bClk + bZ + bDV + bDP + bAB + bAR + bSn + bVG;
    DeviceScene& d = s.dev;
    if (reserve) reserve(total);
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    if (hipMalloc(&d.arena, total) != hipSuccess) {
        (void)hipGetLastError();
        d.arena = nullptr;
        fail(GVBD_ERR_NOMEM, "hipMalloc of " + std::to_string(total) + " bytes for scene failed");
    }
    d.bytes = total;
    char* base = (char*)d.arena;
    size_t at = 0;
    auto take = [&](size_t n) {
        char* p = base + at;
        at += n;
        return p;
    };
    d.pos = (float4*)take(bV);
    d.prev = (float4*)take(bV);
    d.inertia = (float4*)take(bV);
    d.vel = (float4*)take(bV);
    d.vparam = (float4*)take(bV);
    d.pnext = (float4*)take(bV);
    d.vflag = (uint8_t*)take(bF);
    d.slots = (int4*)take(bS);
    d.tinc = (TetInc*)take(bI);
    d.tvol = (float*)take(bIV);
    d.tmat = (float4*)take(bTM);
    d.cslots = (ClothSlot*)take(bCS);
    d.cface = (FaceInc*)take(bCF);
    d.cbend = (BendInc*)take(bCB);
    if (contact) {
        d.faceRef = (int*)take(bFR);
        d.edgeRef = (int*)take(bER);
        d.friction = (float2*)take(bFr);
        d.tFaceRef = (int*)take(bTF);
        d.tEdgeRef = (int*)take(bTE);
        d.tcontact = (int4*)take(bTC);
        d.snap = (float4*)take(bSn);
        d.vgroup = (int*)take(bVG);
    }
    if (accel) {
        d.accPrevPrev = (float4*)take(bA);
        d.accPrevIter = (float4*)take(bA);
        d.accOmega = (float*)take(bW);
    }
    d.clock = (float*)take(bClk);
    if (bZ) d.vzone = (float2*)take(bZ);
    if (bDV) {
        d.dverts = (DeformVert*)take(bDV);
        d.dparams = (DeformParams*)take(bDP);
    }
    if (bAB) {
        d.attBegin = (int*)take(bAB);
        d.att = (AttachRec*)take(bAR);
    }
    auto up = [&](void* dst, const void* src, size_t n) {
        if (n) hip_check(hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, ctx->stream), "upload");
    };
    up(d.pos, pos.data(), sizeof(float4) * nv);
    up(d.prev, pos.data(), sizeof(float4) * nv);
    up(d.pnext, pos.data(), sizeof(float4) * nv);
    up(d.inertia, inertia.data(), sizeof(float4) * nv);
    up(d.vel, vel.data(), sizeof(float4) * nv);
    up(d.vparam, vparam.data(), sizeof(float4) * nv);
    up(d.vflag, vflag.data(), nv);
    up(d.slots, slots.data(), sizeof(int4) * slots.size());
    up(d.tinc, tinc.data(), sizeof(TetInc) * tinc.size());
    up(d.tvol, tvol.data(), sizeof(float) * tvol.size());
    up(d.tmat, tmat.data(), sizeof(float4) * tmat.size());
    up(d.cslots, cslots.data(), sizeof(ClothSlot) * cslots.size());
    up(d.cface, cface.data(), sizeof(FaceInc) * cface.size());
    up(d.cbend, cbend.data(), sizeof(BendInc) * cbend.size());
    if (contact) {
        up(d.faceRef, faceRef.data(), sizeof(int32_t) * faceRef.size());
        up(d.edgeRef, edgeRef.data(), sizeof(int32_t) * edgeRef.size());
        up(d.friction, friction.data(), sizeof(float2) * friction.size());
        up(d.tFaceRef, tFaceRef.data(), sizeof(int32_t) * tFaceRef.size());
        up(d.tEdgeRef, tEdgeRef.data(), sizeof(int32_t) * tEdgeRef.size());
        up(d.tcontact, tcontact.data(), sizeof(int4) * tcontact.size());
        up(d.snap, pos.data(), sizeof(float4) * nv);
        up(d.vgroup, vgroup.data(), sizeof(int32_t) * nv);
    }
    if (accel) {
        up(d.accPrevPrev, pos.data(), sizeof(float4) * nv);
        up(d.accPrevIter, pos.data(), sizeof(float4) * nv);
        up(d.accOmega, omega.data(), sizeof(float) * omega.size());
    }
    // scene clock block: the time at [0], the solve-failure counter at int [16]
    float clockBlock[64] = {};
    clockBlock[0] = s.time0;
    up(d.clock, clockBlock, sizeof(clockBlock));
    if (bZ) up(d.vzone, vzone.data(), sizeof(float2) * vzone.size());
    if (bDV) {
        up(d.dverts, dverts.data(), sizeof(DeformVert) * dverts.size());
        up(d.dparams, dparams.data(), sizeof(DeformParams) * dparams.size());
    }
    if (bAB) {
        up(d.attBegin, attBegin.data(), sizeof(int32_t) * attBegin.size());
        up(d.att, att.data(), sizeof(AttachRec) * att.size());
    }
    hip_check(hipStreamSynchronize(ctx->stream), "upload sync");
    SceneView& v = s.view;
    v.pos = d.pos;
    v.prev = d.prev;
    v.inertia = d.inertia;
    v.vel = d.vel;
    v.vparam = d.vparam;
    v.vflag = d.vflag;
    v.slots = d.slots;
    v.tinc = d.tinc;
    v.tvol = d.tvol;
    v.tmat = d.tmat;
    v.pnext = d.pnext;
    v.cslots = d.cslots;
    v.cface = d.cface;
    v.cbend = d.cbend;
    v.tcontact = d.tcontact;
    v.accPrevPrev = d.accPrevPrev;
    v.accPrevIter = d.accPrevIter;
    v.accOmega = d.accOmega;
    v.clock = d.clock;
    v.dverts = d.dverts;
    v.dparams = d.dparams;
    v.attBegin = d.attBegin;
    v.att = d.att;
    v.vzone = d.vzone;
    v.solveFail = reinterpret_cast<int*>(d.clock) + kSolveFailSlot;
    v.nv = nv;
    v.sp = s.sp;
}

// (Re)builds the work items of a plan for its current schedule.
// Launch slot of every parallel group of every scene of a plan (group launches): a scene's
// groups keep their order (group g in a later launch than group g - 1), which is all a scene's
// results depend on, so any such assignment gives the same bits. A scene with as many groups as
// the plan has launches takes them one to one; the others, largest first, take the increasing
// slots that keep the largest launch smallest (a dynamic program over groups and slots), which
// moves the groups of scenes with few colours into the launches the many-colour scenes leave
// small.
void assign_group_launches(gvbd_context* ctx, PlanObj& p) {
    const int L = p.numGroups;
    const size_t nS = p.scenes.size();
    p.groupAt.assign(nS * (size_t)L, -1);
    std::vector<std::vector<int64_t>> w(nS);
    std::vector<int64_t> total(nS, 0);
    for (size_t si = 0; si < nS; ++si) {
        const SceneObj& s = scene_of(ctx, p.scenes[si]);
        const size_t nt = s.numTetSlots();
        w[si].assign(s.groups.size(), 0);
        for (size_t g = 0; g < s.groups.size(); ++g) {
            for (int32_t k = s.tetGroupOffset[g]; k < s.tetGroupOffset[g + 1]; ++k) w[si][g] += s.slotWorkBytes[k];
            for (int32_t k = s.clothGroupOffset[g]; k < s.clothGroupOffset[g + 1]; ++k)
                w[si][g] += s.slotWorkBytes[nt + k];
            total[si] += w[si][g];
        }
    }
    std::vector<size_t> order(nS);
    for (size_t i = 0; i < nS; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
        const bool fa = (int)w[a].size() == L, fb = (int)w[b].size() == L;
        return fa != fb ? fa : total[a] > total[b];
    });
    std::vector<int64_t> load(L, 0);
    const int64_t INF = std::numeric_limits<int64_t>::max();
    for (size_t si : order) {
        const int G = (int)w[si].size();
        if (G == 0) continue;
        std::vector<int> slot(G);
        if (G == L || p.mode != GVBD_MODE_GROUP_LAUNCHES) {
            for (int g = 0; g < G; ++g) slot[g] = g;
        } else {
            // best[g][t]: the smallest largest load over groups 0..g with group g in slot t
            std::vector<int64_t> best((size_t)G * L, INF);
            std::vector<int> from((size_t)G * L, -1);
            for (int t = 0; t <= L - G; ++t) best[t] = load[t] + w[si][0];
            for (int g = 1; g < G; ++g) {
                int64_t run = INF;
                int arg = -1;
                for (int t = g; t <= L - G + g; ++t) {
                    const int64_t prev = best[(size_t)(g - 1) * L + t - 1];
                    if (prev < run) {
                        run = prev;
                        arg = t - 1;
                    }
                    if (arg < 0) continue;
                    best[(size_t)g * L + t] = std::max(load[t] + w[si][g], run);
                    from[(size_t)g * L + t] = arg;
                }
            }
            int t = -1;
            for (int u = G - 1; u <= L - 1; ++u)
                if (best[(size_t)(G - 1) * L + u] != INF && (t < 0 || best[(size_t)(G - 1) * L + u] <
                                                                           best[(size_t)(G - 1) * L + t]))
                    t = u;
            for (int g = G - 1; g >= 0; --g) {
                slot[g] = t;
                t = from[(size_t)g * L + t];
            }
        }
        for (int g = 0; g < G; ++g) {
            p.groupAt[si * L + slot[g]] = g;
            load[slot[g]] += w[si][g];
        }
    }
}

void build_plan(gvbd_context* ctx, PlanObj& p) {
    if (!p.dirty) return;
    free_plan(p);
    const int vpb = p.block / p.lpv;
    std::vector<SceneView> views;
    p.numGroups = 0;
    p.accel = false;
    p.contact = p.clothContact = false;
    for (int32_t sid : p.scenes) {
        SceneObj& s = scene_of(ctx, sid);
        if (s.cd) {
            p.contact = true;
            if (!s.cd->penetration()) p.clothContact = true;
        }
        views.push_back(s.view);
        p.numGroups = std::max(p.numGroups, (int)s.groups.size());
        if (s.view.accOmega) p.accel = true;
    }
    std::vector<WorkItem> items, commits, vitems;
    p.groupItemBegin.assign(p.numGroups + 1, 0);
    p.groupCommitBegin.assign(p.numGroups + 1, 0);
    // runs of consecutive cloth slots of group g that need the positionsNext copy
    auto commit_runs = [](const SceneObj& s, int g, int chunk, auto&& emit) {
        const int32_t e = s.clothGroupOffset[g + 1];
        for (int32_t k = s.clothGroupOffset[g]; k < e;) {
            if (s.clothSlotInplace[k]) {
                ++k;
                continue;
            }
            int32_t r = k;
            while (r < e && !s.clothSlotInplace[r] && r - k < chunk) ++r;
            emit(k, r - k);
            k = r;
        }
    };
    p.groupContactBegin.assign(p.numGroups, 0);
    assign_group_launches(ctx, p);
    for (int t = 0; t < p.numGroups; ++t) {
        // items of scenes without collision handling first (the plain solve variant), then those
        // of scenes with it (the contact variant): each scene's work is the same either way
        for (int pass = 0; pass < 2; ++pass) {
            if (pass == 1) p.groupContactBegin[t] = (int32_t)items.size();
            for (size_t si = 0; si < p.scenes.size(); ++si) {
                SceneObj& s = scene_of(ctx, p.scenes[si]);
                const int g = p.groupAt[si * p.numGroups + t];
                if (g < 0 || (s.cd != nullptr) != (pass == 1)) continue;
                for (int32_t k = s.tetGroupOffset[g], e = s.tetGroupOffset[g + 1]; k < e; k += vpb)
                    items.push_back({(int32_t)si, k, std::min(vpb, e - k), WORK_TET});
                for (int32_t k = s.clothGroupOffset[g], e = s.clothGroupOffset[g + 1]; k < e; k += vpb)
                    items.push_back({(int32_t)si, k, std::min(vpb, e - k), WORK_CLOTH});
                commit_runs(s, g, kVertexChunk,
                            [&](int32_t k, int32_t n) { commits.push_back({(int32_t)si, k, n, WORK_CLOTH}); });
            }
        }
        p.groupItemBegin[t + 1] = (int32_t)items.size();
        p.groupCommitBegin[t + 1] = (int32_t)commits.size();
    }
    // Work queue of one iteration, ordered by (level, scene): solve entries of level g of every
    // scene, then the level's commit entries. Solve entries of (s, g) need every entry of scene s
    // in levels < g; commits of (s, g) also need the level's solves of s.
    std::vector<QueueEntry> queue;
    if (p.mode == GVBD_MODE_WORK_QUEUE) {
        const int per = p.vertsPerEntry > 0 ? p.vertsPerEntry : 4 * vpb;
        const size_t nS = p.scenes.size();
        std::vector<int32_t> done(nS, 0), levelSolves(nS, 0);
        for (int g = 0; g < p.numGroups; ++g) {
            for (size_t si = 0; si < nS; ++si) {
                SceneObj& s = scene_of(ctx, p.scenes[si]);
                levelSolves[si] = 0;
                if (g >= (int)s.groups.size()) continue;
                auto solves = [&](int32_t b, int32_t e, int32_t kind) {
                    for (int32_t k = b; k < e; k += per) {
                        queue.push_back({(int32_t)si, k, std::min(per, e - k), kind, done[si], 0, {0, 0}});
                        ++levelSolves[si];
                    }
                };
                solves(s.tetGroupOffset[g], s.tetGroupOffset[g + 1], WORK_TET);
                solves(s.clothGroupOffset[g], s.clothGroupOffset[g + 1], WORK_CLOTH);
            }
            std::vector<int32_t> levelCommits(nS, 0);
            for (size_t si = 0; si < nS; ++si) {
                SceneObj& s = scene_of(ctx, p.scenes[si]);
                if (g >= (int)s.groups.size()) continue;
                commit_runs(s, g, kCommitChunk, [&](int32_t k, int32_t n) {
                    queue.push_back({(int32_t)si, k, n, WORK_COMMIT, done[si] + levelSolves[si], 0, {0, 0}});
                    ++levelCommits[si];
                });
            }
            for (size_t si = 0; si < nS; ++si) done[si] += levelSolves[si] + levelCommits[si];
        }
        for (QueueEntry& q : queue) q.perIteration = done[q.scene];
        p.nQueue = (int32_t)queue.size();
        p.ticketSlots = std::max(p.iterations, 64);
        p.nCounters = p.ticketSlots + (int32_t)nS;
        const int perCu = p.blocksPerCu > 0 ? p.blocksPerCu : queue_blocks_per_cu(p.lpv, p.block);
        p.queueGrid = std::max(1, std::min(p.nQueue, perCu * std::max(ctx->numCUs, 1)));
    }
    for (size_t si = 0; si < p.scenes.size(); ++si) {
        SceneObj& s = scene_of(ctx, p.scenes[si]);
        for (int32_t k = 0; k < s.nv; k += kVertexChunk)
            vitems.push_back({(int32_t)si, k, std::min(kVertexChunk, s.nv - k), 0});
    }
    p.nVItems = (int32_t)vitems.size();
    std::vector<WorkItem> dpre, dpost;
    for (size_t si = 0; si < p.scenes.size(); ++si) {
        SceneObj& s = scene_of(ctx, p.scenes[si]);
        for (int32_t k = 0; k < s.nDeformPre; k += 256)
            dpre.push_back({(int32_t)si, k, std::min(256, s.nDeformPre - k), 0});
        const int32_t e = s.nDeformPre + s.nDeformPost;
        for (int32_t k = s.nDeformPre; k < e; k += 256) dpost.push_back({(int32_t)si, k, std::min(256, e - k), 0});
    }
    p.nDeformPre = (int32_t)dpre.size();
    p.nDeformPost = (int32_t)dpost.size();
    // one buffer: each array at a 256-byte aligned offset
    struct Part {
        void** dst;
        const void* src;
        size_t n;
        size_t off;
    };
    std::vector<int> zeros;
    std::vector<Part> parts = {{(void**)&p.dViews, views.data(), sizeof(SceneView) * views.size(), 0},
                               {(void**)&p.dItems, items.data(), sizeof(WorkItem) * items.size(), 0},
                               {(void**)&p.dCommit, commits.data(), sizeof(WorkItem) * commits.size(), 0},
                               {(void**)&p.dVItems, vitems.data(), sizeof(WorkItem) * vitems.size(), 0}};
    if (p.nDeformPre) parts.push_back({(void**)&p.dDeformPre, dpre.data(), sizeof(WorkItem) * dpre.size(), 0});
    if (p.nDeformPost) parts.push_back({(void**)&p.dDeformPost, dpost.data(), sizeof(WorkItem) * dpost.size(), 0});
    if (p.mode == GVBD_MODE_WORK_QUEUE) {
        zeros.assign(p.nCounters, 0);
        parts.push_back({(void**)&p.dQueue, queue.data(), sizeof(QueueEntry) * queue.size(), 0});
        parts.push_back({(void**)&p.dCounters, zeros.data(), sizeof(int) * zeros.size(), 0});
    }
    size_t total = 0;
    for (Part& q : parts) {
        q.off = total;
        total += (std::max<size_t>(q.n, 16) + 255) & ~(size_t)255;
    }
    if (total > p.cap) {
        // grow: the old buffers may still be read by frames in flight and by the last upload
        hip_check(hipStreamSynchronize(ctx->stream), "plan buffer sync");
        if (p.dBuf) (void)hipFree(p.dBuf);
        if (p.hBuf) (void)hipHostFree(p.hBuf);
        p.dBuf = p.hBuf = nullptr;
        p.cap = p.bytes = 0;
        const size_t cap = std::max(total, total + total / 2);
        if (hipMalloc((void**)&p.dBuf, cap) != hipSuccess) {
            (void)hipGetLastError();
            p.dBuf = nullptr;
            fail(GVBD_ERR_NOMEM, "hipMalloc for plan failed");
        }
        if (hipHostMalloc((void**)&p.hBuf, cap, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            (void)hipFree(p.dBuf);
            p.dBuf = p.hBuf = nullptr;
            fail(GVBD_ERR_NOMEM, "pinned host allocation for plan staging failed");
        }
        p.cap = p.bytes = cap;
        if (!p.uploadDone) hip_check(hipEventCreateWithFlags(&p.uploadDone, hipEventDisableTiming), "event create");
    } else {
        // the staging buffer is rewritten only after its last copy has left it
        hip_check(hipEventSynchronize(p.uploadDone), "plan staging wait");
    }
    for (Part& q : parts) {
        if (q.n) std::memcpy(p.hBuf + q.off, q.src, q.n);
        *q.dst = p.dBuf + q.off;
    }
    hip_check(hipMemcpyAsync(p.dBuf, p.hBuf, total, hipMemcpyHostToDevice, ctx->stream), "plan upload");
    hip_check(hipEventRecord(p.uploadDone, ctx->stream), "event record");
    p.dirty = false;
}

void plan_reset_counters(gvbd_context* ctx, PlanObj& p);

// Deformers of cloth meshes (VBDClothPhysics.cpp:408-412: at iteration 0, before detection and
// the initial guess).
void plan_deform_cloth(gvbd_context* ctx, PlanObj& p) {
    if (p.nDeformPre)
        hip_check(launch_deform(p.dViews, p.dDeformPre, p.nDeformPre, true, ctx->stream), "deformer launch");
}

// Initial step of a substep, preceded by the cloth meshes' deformers unless the caller already
// ran them and followed by the tet meshes' (VBDPhysics.cpp:849-862: after the initial step); in
// work-queue mode it also starts a new counter epoch.
void plan_initial_step(gvbd_context* ctx, PlanObj& p, bool clothDeformed = false) {
    if (!clothDeformed) plan_deform_cloth(ctx, p);
    hip_check(launch_initial_step(p.dViews, p.dVItems, p.nVItems, ctx->stream), "initial step launch");
    if (p.nDeformPost)
        hip_check(launch_deform(p.dViews, p.dDeformPost, p.nDeformPost, false, ctx->stream), "deformer launch");
    if (p.accel) hip_check(launch_accel_init(p.dViews, p.dVItems, p.nVItems, ctx->stream), "accelerator launch");
    if (p.mode == GVBD_MODE_WORK_QUEUE) plan_reset_counters(ctx, p);
}

void plan_group(gvbd_context* ctx, PlanObj& p, int g, int iter) {
    const int b = p.groupItemBegin[g], c = p.groupContactBegin[g], e = p.groupItemBegin[g + 1];
    if (c > b)
        hip_check(launch_group_solve(p.dViews, p.dItems + b, c - b, iter, p.lpv, p.block, p.stage, false, p.xcd,
                                     ctx->stream),
                  "group solve launch");
    if (e > c)
        hip_check(launch_group_solve(p.dViews, p.dItems + c, e - c, iter, p.lpv, p.block, 0, true, p.xcd, ctx->stream),
                  "group solve launch");
    const int cb = p.groupCommitBegin[g], ce = p.groupCommitBegin[g + 1];
    if (ce > cb) hip_check(launch_cloth_commit(p.dViews, p.dCommit + cb, ce - cb, ctx->stream), "commit launch");
}

void plan_velocity(gvbd_context* ctx, PlanObj& p) {
    hip_check(launch_update_velocity(p.dViews, p.dVItems, p.nVItems, ctx->stream), "velocity launch");
}

QueueArgs queue_args(gvbd_context* ctx, PlanObj& p) {
    return QueueArgs{p.dViews, p.dQueue, p.nQueue, p.dCounters, p.dCounters + p.ticketSlots, ctx->dStatus,
                     p.spinLimit};
}

void plan_reset_counters(gvbd_context* ctx, PlanObj& p) {
    hip_check(launch_counter_reset(p.dCounters, p.nCounters, ctx->stream), "counter reset launch");
    p.epoch = 0;
}

// One iteration of every scene: one persistent launch (work queue, the `epoch`-th launch since
// the counters were reset) or one launch per parallel group.
// Contact scenes: the iteration-start snapshot read for same-group contact partners.
void plan_contact_snapshot(gvbd_context* ctx, PlanObj& p) {
    for (int32_t sid : p.scenes) {
        SceneObj& s = scene_of(ctx, sid);
        if (s.cd) hip_check(launch_stream_copy(s.view.pos, s.dev.snap, s.nv, ctx->stream), "snapshot launch");
    }
}

void plan_iteration(gvbd_context* ctx, PlanObj& p, int it, int epoch) {
    if (p.contact) plan_contact_snapshot(ctx, p);
    if (p.mode == GVBD_MODE_WORK_QUEUE) {
        hip_check(launch_queue_solve(queue_args(ctx, p), it, epoch, p.queueGrid, p.lpv, p.block, ctx->stream),
                  "work-queue solve launch");
    } else {
        for (int g = 0; g < p.numGroups; ++g) plan_group(ctx, p, g, it);
    }
    // Chebyshev acceleration of the sweep (omega of iteration it; the per-iteration hook may be
    // called past the configured count, which keeps the last omega)
    if (p.accel)
        hip_check(launch_accel_apply(p.dViews, p.dVItems, p.nVItems, std::max(0, std::min(it, p.iterations - 1)),
                                     ctx->stream),
                  "accelerator launch");
}

bool scene_collision_detection(SceneObj& s, bool firstIteration);

// One frame of a cloth scene with collision handling (VBDClothPhysics.cpp:389-518): detection
// before iteration 0 (ahead of the initial guess) and before every iteration after one in which
// a vertex was truncated to its conservative bound; the host reads that flag per iteration.
void plan_contact_frame(gvbd_context* ctx, PlanObj& p) {
    SceneObj& s = scene_of(ctx, p.scenes[0]);
    const DetectorDeviceView dv = s.cd->deviceView();
    for (int sub = 0; sub < p.substeps; ++sub) {
        bool required = true;
        for (int it = 0; it < p.iterations; ++it) {
            const bool detect = it == 0 || required;
            if (it == 0) plan_deform_cloth(ctx, p);
            if (detect) {
                if (!s.cdEvents[0]) {
                    hip_check(hipEventCreate(&s.cdEvents[0]), "event create");
                    hip_check(hipEventCreate(&s.cdEvents[1]), "event create");
                }
                hip_check(hipEventRecord(s.cdEvents[0], ctx->stream), "event record");
                if (scene_collision_detection(s, it == 0)) ++s.contactStats.rebuilds;
                hip_check(hipEventRecord(s.cdEvents[1], ctx->stream), "event record");
                ++s.contactStats.detections;
            }
            if (it == 0) plan_initial_step(ctx, p, true);
            plan_iteration(ctx, p, it, it);
            int flag = 0;
            hip_check(hipMemcpyAsync(&flag, dv.flag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream), "flag read");
            hip_check(hipStreamSynchronize(ctx->stream), "flag sync");
            required = flag != 0;
            if (detect) {
                float ms = 0.f;
                hip_check(hipEventElapsedTime(&ms, s.cdEvents[0], s.cdEvents[1]), "event time");
                s.contactStats.detect_ms += ms;
            }
        }
        if (p.iterations == 0) plan_initial_step(ctx, p);
        plan_velocity(ctx, p);
        ++s.contactStats.frames;
        s.contactStats.iterations += p.iterations;
    }
    ++s.frame;
}

// Penetration frames time their detection with events read one substep later (or when the
// stats are read), so the frame itself has no host wait.
void collect_detection_time(SceneObj& s) {
    if (!s.detectPending) return;
    float a = 0.f, b = 0.f;
    hip_check(hipEventSynchronize(s.cdEvents[3]), "event sync");
    hip_check(hipEventElapsedTime(&a, s.cdEvents[0], s.cdEvents[1]), "event time");
    hip_check(hipEventElapsedTime(&b, s.cdEvents[2], s.cdEvents[3]), "event time");
    s.contactStats.detect_ms += a + b;
    s.detectPending = false;
}

// Whether scene s rebuilds its detection trees in the frame it is about to step
// (dcdTetMeshSceneBVHRebuildSteps, VBDPhysics.cpp:1962-1963: at substep 0 of every such frame
// but the first).
bool penetration_rebuild_due(const SceneObj& s) {
    const int steps = s.params.dcd_bvh_rebuild_steps;
    return steps > 0 && s.frame > 0 && s.frame % steps == 0;
}

bool cloth_rebuild_due(const SceneObj& s);

// One frame of a plan with contact scenes, batched with its other scenes, without host reads
// (runStepGPU, VBDPhysics.cpp:839-969, and runStep_CDEveryIter, VBDClothPhysics.cpp:389-518).
// Per substep: DCD of each penetration-contact scene at its start (the trees rebuilt at substep 0
// of a rebuild frame, refit otherwise); the cloth deformers; the first detection pass of each
// cloth-model scene; the batched initial step; CCD of each penetration-contact scene over its
// vertices' paths from their start positions and its face links; then per iteration a detection
// pass of each cloth-model scene gated on the device by its re-detection flag (iterations after
// the first) and the batched iteration; the velocities. With `timing` the penetration passes are
// timed with events read a substep later.
void plan_device_contact_frame(gvbd_context* ctx, PlanObj& p, bool timing) {
    std::vector<SceneObj*> pen, cloth;
    for (int32_t sid : p.scenes) {
        SceneObj& s = scene_of(ctx, sid);
        if (s.cd) (s.cd->penetration() ? pen : cloth).push_back(&s);
    }
    auto gated = [&](bool first) {
        for (SceneObj* s : cloth)
            s->cd->detectGated(first, first && cloth_rebuild_due(*s), s->params.thickness,
                               s->params.conservative_step_relaxation);
    };
    for (int sub = 0; sub < p.substeps; ++sub) {
        for (SceneObj* s : pen) {
            if (timing) {
                collect_detection_time(*s);
                for (hipEvent_t& e : s->cdEvents)
                    if (!e) hip_check(hipEventCreate(&e), "event create");
                hip_check(hipEventRecord(s->cdEvents[0], ctx->stream), "event record");
            }
            s->cd->penetrationDcd(sub == 0 && penetration_rebuild_due(*s));
            if (timing) hip_check(hipEventRecord(s->cdEvents[1], ctx->stream), "event record");
        }
        if (p.iterations > 0 && !cloth.empty()) {
            plan_deform_cloth(ctx, p);
            gated(true);
            plan_initial_step(ctx, p, true);
        } else {
            plan_initial_step(ctx, p);
        }
        for (SceneObj* s : pen) {
            if (timing) hip_check(hipEventRecord(s->cdEvents[2], ctx->stream), "event record");
            s->cd->penetrationCcd(s->view.prev, sub == 0 && penetration_rebuild_due(*s));
            s->cd->buildFaceLinks(32);
            if (timing) {
                hip_check(hipEventRecord(s->cdEvents[3], ctx->stream), "event record");
                s->detectPending = true;
            }
        }
        for (int it = 0; it < p.iterations; ++it) {
            if (it > 0) gated(false);
            plan_iteration(ctx, p, it, it);
        }
        plan_velocity(ctx, p);
    }
}

// Whether a cloth-model contact scene rebuilds its trees in the frame it is about to step
// (contactBVHReconstructionSteps, VBDClothPhysics.cpp:727-735: at the first detection of every
// substep of such a frame but the first).
bool cloth_rebuild_due(const SceneObj& s) {
    const int steps = s.params.contact_bvh_rebuild_steps;
    return steps > 0 && s.frame > 0 && s.frame % steps == 0;
}

bool contact_rebuild_due(const SceneObj& s) {
    return s.cd->penetration() ? penetration_rebuild_due(s) : cloth_rebuild_due(s);
}

// Host bookkeeping of one device-driven frame of every contact scene of a plan (counters, frame
// index); the gated detection passes of cloth-model scenes are counted on the device.
void contact_frame_done(PlanObj& p, gvbd_context* ctx) {
    for (int32_t sid : p.scenes) {
        SceneObj& s = scene_of(ctx, sid);
        if (!s.cd) continue;
        if (s.cd->penetration()) {
            if (penetration_rebuild_due(s)) ++s.contactStats.rebuilds;
            s.contactStats.detections += p.substeps;
        } else if (cloth_rebuild_due(s) && p.iterations > 0) {
            s.contactStats.rebuilds += p.substeps;
        }
        s.contactStats.frames += p.substeps;
        s.contactStats.iterations += (int64_t)p.substeps * p.iterations;
        ++s.frame;
    }
}

// One frame: substeps x (initial step, iterations x groups, velocity update)
// (VBDPhysics.cpp:839-969 / VBDClothPhysics.cpp:389-518 structure, all on the device).
void plan_frame_launches(gvbd_context* ctx, PlanObj& p) {
    if (p.clothContact && !p.capturing && p.scenes.size() == 1) {
        // a cloth-model scene stepped alone: the host reads the re-detection flag per iteration
        plan_contact_frame(ctx, p);
        p.epoch = p.iterations;
        return;
    }
    if (p.contact) {
        // detection is timed for a scene stepped alone (gvbd_contact_stats.detect_ms); batched and
        // served frames are not timed, so that no event wait slows their launch
        plan_device_contact_frame(ctx, p, !p.capturing && p.scenes.size() == 1);
        if (!p.capturing) contact_frame_done(p, ctx);
        p.epoch = p.iterations;
        return;
    }
    for (int sub = 0; sub < p.substeps; ++sub) {
        plan_initial_step(ctx, p);
        for (int it = 0; it < p.iterations; ++it) plan_iteration(ctx, p, it, it);
        plan_velocity(ctx, p);
    }
    p.epoch = p.iterations;  // device counters after the last substep
}

// Whether a penetration-contact scene of the plan rebuilds its trees in the next frame (such a
// frame is launched directly: the rebuild reads level sizes back to the host, and the captured
// refits of the old trees no longer apply after it).
bool plan_rebuild_due(gvbd_context* ctx, const PlanObj& p) {
    if (!p.contact) return false;
    for (int32_t sid : p.scenes) {
        const SceneObj& s = scene_of(ctx, sid);
        if (s.cd && contact_rebuild_due(s)) return true;
    }
    return false;
}

void plan_capture(gvbd_context* ctx, PlanObj& p) {
    free_plan_graph(p);
    if (plan_rebuild_due(ctx, p)) return;  // captured after the rebuild frame (see plan_run)
    p.capturing = true;
    hip_check(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal), "begin capture");
    try {
        plan_frame_launches(ctx, p);
    } catch (...) {
        p.capturing = false;
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(ctx->stream, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    p.capturing = false;
    hip_check(hipStreamEndCapture(ctx->stream, &p.graph), "end capture");
    hip_check(hipGraphInstantiate(&p.exec, p.graph, nullptr, nullptr, 0), "graph instantiate");
}

void plan_run(gvbd_context* ctx, PlanObj& p, int n) {
    build_plan(ctx, p);
    for (int i = 0; i < n; ++i) {
        if (p.wantGraph && plan_rebuild_due(ctx, p)) {
            // a tree rebuild: launched directly, the graph captured again for the next frames
            free_plan_graph(p);
            plan_frame_launches(ctx, p);
            continue;
        }
        if (p.wantGraph && !p.exec) plan_capture(ctx, p);
        if (p.exec) {
            hip_check(hipGraphLaunch(p.exec, ctx->stream), "graph launch");
            if (p.contact) contact_frame_done(p, ctx);
        } else {
            plan_frame_launches(ctx, p);
        }
    }
    p.epoch = p.iterations;
}

int32_t new_plan(gvbd_context* ctx, const std::vector<int32_t>& scenes) {
    auto p = std::make_unique<PlanObj>();
    p->scenes = scenes;
    SceneObj& s0 = scene_of(ctx, scenes[0]);
    p->iterations = s0.params.iterations;
    p->substeps = s0.params.num_substeps;
    for (int32_t sid : scenes) {
        SceneObj& s = scene_of(ctx, sid);
        if (!s.initialized) fail(GVBD_ERR_STATE, "scene " + std::to_string(sid) + " is not initialised");
        if (s.server >= 0 && scenes.size() > 1)
            fail(GVBD_ERR_STATE, "scene " + std::to_string(sid) + " is registered with a server");
        if (s.params.iterations != p->iterations || s.params.num_substeps != p->substeps)
            fail(GVBD_ERR_INVALID, "scenes of one batch must share iterations and substeps");
        if (s.cd) {
            if (!s.cd->penetration()) p->clothContact = true;
            p->contact = true;
        }
    }
    const int32_t id = ctx->nextPlan++;
    ctx->plans[id] = std::move(p);
    return id;
}

// The tets and boundary-surface topology of a penetration-contact scene in scene ids (meshes in
// order; the detector's faces are the meshes' boundary faces in the same order).
PenetrationSetup penetration_setup(const SceneObj& s) {
    PenetrationSetup P;
    int32_t tetOff = 0, faceOff = 0;
    std::vector<std::vector<int32_t>> nbr(s.nv);
    for (size_t im = 0; im < s.meshes.size(); ++im) {
        const TetMeshHost& m = s.tets[s.meshes[im].second];
        const TetSurface& S = s.tetSurfaces[s.meshes[im].second];
        const int32_t vOff = s.vertexOffset[im];
        for (int32_t t = 0; t < m.nt; ++t) {
            for (int k = 0; k < 4; ++k) {
                P.tets.push_back(vOff + m.tets[4 * (size_t)t + k]);
                const int32_t n = S.tetNbr[4 * (size_t)t + k];
                P.tetNbr.push_back(n < 0 ? -1 : tetOff + n);
            }
            P.tetMesh.push_back((int32_t)im);
        }
        const int32_t nf = (int32_t)S.faces.size() / 3;
        // neighbour faces over directed edges: (u, v) of one face meets (v, u) of another
        std::map<std::pair<int32_t, int32_t>, int32_t> dir;
        for (int32_t f = 0; f < nf; ++f)
            for (int e = 0; e < 3; ++e) dir[{S.faces[3 * (size_t)f + e], S.faces[3 * (size_t)f + (e + 1) % 3]}] = f;
        for (int32_t f = 0; f < nf; ++f) {
            P.faceOwner.push_back(tetOff + S.owner[2 * (size_t)f]);
            P.faceOwner.push_back(S.owner[2 * (size_t)f + 1]);
            for (int e = 0; e < 3; ++e) {
                const int32_t u = S.faces[3 * (size_t)f + e], v = S.faces[3 * (size_t)f + (e + 1) % 3];
                auto it = dir.find({v, u});
                P.faceNbr.push_back(it == dir.end() ? -1 : faceOff + it->second);
                nbr[vOff + u].push_back(vOff + v);
                nbr[vOff + v].push_back(vOff + u);
            }
        }
        for (int32_t v = 0; v < m.nv; ++v)
            if (S.surface[v]) P.surfVerts.push_back(vOff + v);
        tetOff += m.nt;
        faceOff += nf;
    }
    P.vNbrBegin.assign(s.nv + 1, 0);
    for (int32_t v = 0; v < s.nv; ++v) {
        std::sort(nbr[v].begin(), nbr[v].end());
        nbr[v].erase(std::unique(nbr[v].begin(), nbr[v].end()), nbr[v].end());
        P.vNbrBegin[v + 1] = P.vNbrBegin[v] + (int32_t)nbr[v].size();
        P.vNbr.insert(P.vNbr.end(), nbr[v].begin(), nbr[v].end());
    }
    const gvbd_physics_params& q = s.params;
    PenetrationParams& c = P.params;
    c.dcd = q.allow_dcd != 0;
    c.ccd = q.allow_ccd != 0;
    c.self = q.handle_self_collision != 0;
    c.feasible = q.check_feasible_region != 0;
    c.traverse = q.check_tet_traverse != 0;
    c.traverseNonSelf = q.tet_traverse_non_self != 0;
    c.shiftToCenter = q.shift_query_point_to_center != 0;
    c.stopAfterPassing = q.stop_traversing_after_passing != 0;
    c.offHeight = q.collision_off_height;
    c.feasibleEps = q.feasible_region_epsilon;
    c.rayEps = q.ray_tri_intersection_epsilon;
    c.maxSearchMult = q.max_search_distance_multiplier;
    c.centerShift = q.center_shift_level;
    return P;
}

void initialize_scene(gvbd_context* ctx, SceneObj& s, const std::function<void(size_t)>& reserve = {}) {
    if (s.initialized) fail(GVBD_ERR_STATE, "scene already initialised");
    if (s.meshes.empty()) fail(GVBD_ERR_STATE, "scene has no meshes");
    const bool contact = s.params.handle_collision != 0;
    if (contact && !(s.params.contact_query_radius > 0.f))
        fail(GVBD_ERR_INVALID, "contact_query_radius must be positive");
    if (contact)
        for (const DeformerHost& D : s.deformers)
            for (int32_t m : D.meshes)
                if (s.meshes[m].first == MESH_TET)
                    fail(GVBD_ERR_INVALID,
                         "deformers of tet meshes are not supported with collision handling: they move vertices "
                         "after the frame's first detection, so contact bounds would not see the motion");
    if (s.params.use_accelerator) {
        if (contact)
            fail(GVBD_ERR_INVALID, "use_accelerator applies to tet scenes without collision handling (the "
                                   "reference's accelerator is part of its tet GPU path)");
        if (!(s.params.accelerator_rho >= 0.f && s.params.accelerator_rho < 1.f))
            fail(GVBD_ERR_INVALID, "accelerator_rho must be in [0, 1)");
    }
    const bool penetration = contact && s.params.tet_contact_model == 1;
    if (contact && s.params.tet_contact_model != 0 && s.params.tet_contact_model != 1)
        fail(GVBD_ERR_INVALID, "tet_contact_model must be 0 (cloth-model contact) or 1 (penetration contact)");
    if (penetration) {
        if (!s.cloths.empty())
            fail(GVBD_ERR_INVALID, "the penetration contact model (tet_contact_model 1) is the tet solver's: its scenes "
                                   "hold tet meshes only");
        if (!(s.params.collision_stiffness >= 0.f))
            fail(GVBD_ERR_INVALID, "collision_stiffness must be non-negative");
    }
    s.sp = make_solve_params(s.params, s);
    // tet meshes take part in contact through their boundary surface
    s.tetSurfaces.clear();
    if (contact) {
        for (const TetMeshHost& m : s.tets) {
            TetSurface S;
            std::string err;
            if (!build_tet_surface(m.nv, m.nt, m.tets.data(), m.pos.data(), S.faces, err,
                                   penetration ? &S.owner : nullptr, penetration ? &S.tetNbr : nullptr) ||
                !build_tri_edges(m.nv, (int32_t)S.faces.size() / 3, S.faces.data(), S.edges, nullptr, err))
                fail(GVBD_ERR_INVALID, "tet mesh boundary surface: " + err);
            const int32_t nf = (int32_t)S.faces.size() / 3;
            S.surface.assign(m.nv, 0);
            for (int32_t v : S.faces) S.surface[v] = 1;
            S.vfBegin.assign(m.nv + 1, 0);
            for (int32_t v : S.faces) S.vfBegin[v + 1]++;
            for (int32_t i = 0; i < m.nv; ++i) S.vfBegin[i + 1] += S.vfBegin[i];
            S.vfRef.assign(S.vfBegin[m.nv], 0);
            std::vector<int32_t> fill(S.vfBegin.begin(), S.vfBegin.end() - 1);
            for (int32_t f = 0; f < nf; ++f)
                for (int k = 0; k < 3; ++k) S.vfRef[fill[S.faces[3 * (size_t)f + k]]++] = f | (k << 30);
            S.veBegin.assign(m.nv + 1, 0);
            for (const TriEdge& e : S.edges) {
                S.veBegin[e.v1 + 1]++;
                S.veBegin[e.v2 + 1]++;
            }
            for (int32_t i = 0; i < m.nv; ++i) S.veBegin[i + 1] += S.veBegin[i];
            S.veRef.assign(S.veBegin[m.nv], 0);
            fill.assign(S.veBegin.begin(), S.veBegin.end() - 1);
            for (int32_t e = 0; e < (int32_t)S.edges.size(); ++e) {
                S.veRef[fill[S.edges[e].v1]++] = e;
                S.veRef[fill[S.edges[e].v2]++] = e | (1 << 30);
            }
            s.tetSurfaces.push_back(std::move(S));
        }
    }
    std::vector<const std::vector<std::vector<int32_t>>*> cats;
    std::vector<bool> removeUsed;
    for (auto& mk : s.meshes) {
        if (mk.first == MESH_TET) {
            cats.push_back(&s.tets[mk.second].categories);
            removeUsed.push_back(true);
        } else {
            cats.push_back(&s.cloths[mk.second].categories);
            removeUsed.push_back(false);
        }
    }
    s.groups = merge_parallel_groups(cats, removeUsed);
    upload_scene(ctx, s, reserve);
    if (contact) {
        // ClothContactDetector over the scene's own positions (VBDClothPhysics initialize ->
        // ClothContactDetector::initialize, ClothContactDetector.cpp:944-1050)
        s.cd = std::make_unique<ContactDetector>(ctx->device, ctx->stream, s.params.contact_query_radius, 8, 8);
        s.cd->useExternalPositions(s.dev.pos);
        for (auto& mk : s.meshes) {
            if (mk.first == MESH_TET) {
                const TetMeshHost& m = s.tets[mk.second];
                const TetSurface& S = s.tetSurfaces[mk.second];
                s.cd->addMesh(m.nv, m.pos.data(), (int)S.faces.size() / 3, S.faces.data(), S.surface.data());
            } else {
                const ClothMeshHost& m = s.cloths[mk.second];
                s.cd->addMesh(m.nv, m.pos.data(), m.nf, m.faces.data());
            }
        }
        if (penetration) s.cd->setPenetration(penetration_setup(s));
        s.cd->finalize();
        const DetectorDeviceView dv = s.cd->deviceView();
        ContactView& cv = s.view.cv;
        cv.faces = dv.faces;
        cv.edges = dv.edges;
        cv.friction = s.dev.friction;
        cv.vfCount = dv.vfStored;
        cv.vfRec = (const int4*)dv.vfRec;
        cv.vfCap = dv.vfCap;
        cv.eeCap = dv.eeCap;
        cv.eeCount = dv.eeStored;
        cv.eeRec = (const float4*)dv.eeRec;
        cv.fvRange = dv.fvRange;
        cv.fvVerts = dv.fvVerts;
        cv.faceRef = s.dev.faceRef;
        cv.edgeRef = s.dev.edgeRef;
        cv.tFaceRef = s.dev.tFaceRef;
        cv.tEdgeRef = s.dev.tEdgeRef;
        cv.bound = dv.bounds;
        cv.xAtCD = dv.xAtDetection;
        cv.flag = dv.flag;
        cv.snap = s.dev.snap;
        cv.vgroup = s.dev.vgroup;
        cv.model = penetration ? 1 : 0;
        cv.penStiffness = s.params.collision_stiffness;
        s.view.hasContact = 1;
        s.view.initBound = dv.bounds;
    }
    s.initialized = true;
}

// applyCollisionDetection (VBDClothPhysics.cpp:719-819): BVH refit (full rebuild every
// contact_bvh_rebuild_steps frames at the first iteration), VF/EE queries, conservative bounds,
// face back-links (FV_CONTACT_PREALLOCATE = 32 per face), positions at detection.
// Returns whether the BVH was rebuilt (a refit otherwise).
bool scene_collision_detection(SceneObj& s, bool firstIteration) {
    if (s.cd->penetration()) {
        // the tet solver's DCD at the current positions and the face links (a path for CCD only
        // exists within a frame)
        s.cd->penetrationDcd(false);
        s.cd->buildFaceLinks(32);
        return false;
    }
    const int steps = s.params.contact_bvh_rebuild_steps;
    const bool rebuild = firstIteration && steps > 0 && s.frame > 0 && s.frame % steps == 0;
    s.cd->updateBvh(rebuild);
    s.cd->query(false);
    s.cd->bounds(s.params.thickness, s.params.conservative_step_relaxation);
    s.cd->buildFaceLinks(32);
    s.cd->snapshotPositions();
    return rebuild;
}

template <class F> int guard(gvbd_context* ctx, F&& f) {
    try {
        f();
        return GVBD_OK;
    } catch (const Failure& e) {
        g_lastError = e.msg;
        if (ctx) ctx->lastError = e.msg;
        return e.code;
    } catch (const std::bad_alloc&) {
        g_lastError = "host allocation failed";
        if (ctx) ctx->lastError = g_lastError;
        return GVBD_ERR_NOMEM;
    } catch (const std::exception& e) {
        g_lastError = e.what();
        if (ctx) ctx->lastError = g_lastError;
        return GVBD_ERR_INTERNAL;
    }
}

SceneObj& uninitialized_scene(gvbd_context* ctx, int32_t scene) {
    SceneObj& s = scene_of(ctx, scene);
    if (s.initialized) fail(GVBD_ERR_STATE, "cannot add meshes after initialisation");
    return s;
}

}  // namespace

int32_t add_tet_mesh_host(gvbd_context* ctx, int32_t scene, TetMeshHost&& m) {
    SceneObj& s = uninitialized_scene(ctx, scene);
    std::string err;
    if (!prepare_tet_mesh(m, err)) fail(GVBD_ERR_INVALID, err);
    if ((int32_t)m.colors.size() != m.nv) fail(GVBD_ERR_INTERNAL, "mesh colours missing");
    m.categories = sorted_categories(m.colors);
    s.vertexOffset.push_back(s.nv);
    s.nv += m.nv;
    s.tets.push_back(std::move(m));
    s.meshes.push_back({MESH_TET, (int)s.tets.size() - 1});
    return (int32_t)s.meshes.size() - 1;
}

int32_t add_cloth_mesh_host(gvbd_context* ctx, int32_t scene, ClothMeshHost&& m) {
    SceneObj& s = uninitialized_scene(ctx, scene);
    std::string err;
    if (!prepare_cloth_mesh(m, err)) fail(GVBD_ERR_INVALID, err);
    if ((int32_t)m.colors.size() != m.nv) fail(GVBD_ERR_INTERNAL, "mesh colours missing");
    m.categories = sorted_categories(m.colors);
    s.vertexOffset.push_back(s.nv);
    s.nv += m.nv;
    s.cloths.push_back(std::move(m));
    s.meshes.push_back({MESH_CLOTH, (int)s.cloths.size() - 1});
    return (int32_t)s.meshes.size() - 1;
}

}  // namespace gvbd

using namespace gvbd;

// The gvbd_* entry points below keep the C linkage of their declarations in gaia_vbd.h; the
// helpers between them are internal C++.

int gvbd_abi_struct_sizes(int32_t* physics_params_bytes, int32_t* tet_mesh_desc_bytes,
                          int32_t* cloth_mesh_desc_bytes, int32_t* schedule_bytes) {
    if (physics_params_bytes) *physics_params_bytes = (int32_t)sizeof(gvbd_physics_params);
    if (tet_mesh_desc_bytes) *tet_mesh_desc_bytes = (int32_t)sizeof(gvbd_tet_mesh_desc);
    if (cloth_mesh_desc_bytes) *cloth_mesh_desc_bytes = (int32_t)sizeof(gvbd_cloth_mesh_desc);
    if (schedule_bytes) *schedule_bytes = (int32_t)sizeof(gvbd_schedule);
    return GVBD_OK;
}

int gvbd_context_create(int32_t device, gvbd_context** out) {
    return guard(nullptr, [&] {
        require(out != nullptr, "out is NULL");
        int n = 0;
        hip_check(hipGetDeviceCount(&n), "hipGetDeviceCount");
        if (device < 0 || device >= n) fail(GVBD_ERR_INVALID, "device index out of range");
        auto ctx = std::make_unique<gvbd_context>();
        ctx->device = device;
        hip_check(hipSetDevice(device), "hipSetDevice");
        hip_check(hipDeviceGetAttribute(&ctx->numCUs, hipDeviceAttributeMultiprocessorCount, device),
                  "hipDeviceGetAttribute");
        hip_check(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking), "hipStreamCreate");
        for (auto& e : ctx->events) hip_check(hipEventCreate(&e), "hipEventCreate");
        hip_check(hipMalloc((void**)&ctx->dStatus, 256), "hipMalloc status");
        hip_check(hipMemset(ctx->dStatus, 0, 256), "status clear");
        hip_check(gvbd::convergence_scratch_init(device, &ctx->convScratch), "convergence scratch address");
        *out = ctx.release();
    });
}

int gvbd_context_destroy(gvbd_context* ctx) {
    if (!ctx) return GVBD_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& kv : ctx->plans) release_plan(*kv.second);
    for (auto& kv : ctx->servers) {
        for (gvbd::PlanObj& P : kv.second->plans) release_plan(P);
        for (hipEvent_t e : kv.second->frameDone)
            if (e) (void)hipEventDestroy(e);
        if (kv.second->ref) (void)hipEventDestroy(kv.second->ref);
    }
    ctx->detectors.clear();
    for (auto& kv : ctx->scenes)
        if (kv.second->dev.arena) (void)hipFree(kv.second->dev.arena);
    for (auto& e : ctx->events)
        if (e) (void)hipEventDestroy(e);
    if (ctx->dStatus) (void)hipFree(ctx->dStatus);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return GVBD_OK;
}

const char* gvbd_last_error(const gvbd_context* ctx) { return ctx ? ctx->lastError.c_str() : g_lastError.c_str(); }

int gvbd_synchronize(gvbd_context* ctx) {
    return guard(ctx, [&] {
        require(ctx != nullptr, "ctx is NULL");
        hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        check_queue_status(ctx);
    });
}

int gvbd_context_memory(gvbd_context* ctx, int64_t* bytes) {
    return guard(ctx, [&] {
        require(ctx && bytes, "NULL argument");
        int64_t b = 0;
        for (auto& kv : ctx->scenes) {
            if (kv.second->dev.arena) b += (int64_t)kv.second->dev.bytes;
            if (kv.second->cd) b += (int64_t)kv.second->cd->deviceBytes();
        }
        for (auto& kv : ctx->plans) b += (int64_t)kv.second->bytes;
        for (auto& kv : ctx->servers)
            for (const PlanObj& P : kv.second->plans) b += (int64_t)P.bytes;
        for (auto& kv : ctx->detectors) b += (int64_t)kv.second->deviceBytes();
        *bytes = b;
    });
}

int gvbd_scene_device_bytes(gvbd_context* ctx, int32_t scene, int64_t* arena, int64_t* detector) {
    return guard(ctx, [&] {
        require(ctx && arena && detector, "NULL argument");
        const SceneObj& s = scene_of(ctx, scene);
        *arena = s.dev.arena ? (int64_t)s.dev.bytes : 0;
        *detector = s.cd ? (int64_t)s.cd->deviceBytes() : 0;
    });
}

int gvbd_scene_create(gvbd_context* ctx, const gvbd_physics_params* params, int32_t* scene_out) {
    return guard(ctx, [&] {
        require(ctx && params && scene_out, "NULL argument");
        require(params->num_substeps > 0 && params->iterations >= 0, "bad substeps/iterations");
        require(params->time_step > 0.f, "time_step must be positive");
        auto s = std::make_unique<SceneObj>();
        s->params = *params;
        const int32_t id = ctx->nextScene++;
        ctx->scenes[id] = std::move(s);
        *scene_out = id;
    });
}

int gvbd_scene_clone(gvbd_context* ctx, int32_t scene, int32_t* scene_out) {
    return guard(ctx, [&] {
        require(ctx && scene_out, "NULL argument");
        const SceneObj& src = scene_of(ctx, scene);
        auto s = std::make_unique<SceneObj>();
        s->params = src.params;
        s->tets = src.tets;
        s->cloths = src.cloths;
        s->meshes = src.meshes;
        s->vertexOffset = src.vertexOffset;
        s->nv = src.nv;
        s->deformers = src.deformers;
        s->attachments = src.attachments;
        s->time0 = src.time0;
        const int32_t id = ctx->nextScene++;
        ctx->scenes[id] = std::move(s);
        *scene_out = id;
    });
}

int gvbd_scene_destroy(gvbd_context* ctx, int32_t scene) {
    return guard(ctx, [&] {
        require(ctx != nullptr, "ctx is NULL");
        SceneObj& s = scene_of(ctx, scene);
        if (s.server >= 0) fail(GVBD_ERR_STATE, "scene is registered with a server; remove it from the server first");
        for (auto& kv : ctx->plans)
            for (int32_t sid : kv.second->scenes)
                if (sid == scene && kv.first != s.plan)
                    fail(GVBD_ERR_STATE, "scene is part of a live batch; destroy the batch first");
        hip_check(hipStreamSynchronize(ctx->stream), "sync");
        if (s.plan >= 0) {
            release_plan(plan_of(ctx, s.plan));
            ctx->plans.erase(s.plan);
        }
        s.cd.reset();  // reads the arena's positions: released first
        if (s.convBuf) (void)hipFree(s.convBuf);
        s.convBuf = nullptr;
        for (hipEvent_t& e : s.cdEvents)
            if (e) {
                (void)hipEventDestroy(e);
                e = nullptr;
            }
        if (s.dev.arena) hip_check(hipFree(s.dev.arena), "hipFree");
        ctx->scenes.erase(scene);
    });
}

int gvbd_scene_add_tet_mesh(gvbd_context* ctx, int32_t scene, const gvbd_tet_mesh_desc* desc, int32_t* mesh_out) {
    return guard(ctx, [&] {
        require(ctx && desc, "NULL argument");
        require(desc->num_vertices > 0 && desc->num_tets > 0, "empty tet mesh");
        require(desc->positions && desc->tets, "positions/tets are NULL");
        require(desc->num_fixed == 0 || desc->fixed_points, "fixed_points is NULL");
        require(desc->initialization_type >= 0 && desc->initialization_type <= 4, "initialization_type must be 0..4");
        TetMeshHost m;
        m.nv = desc->num_vertices;
        m.nt = desc->num_tets;
        m.pos.assign(desc->positions, desc->positions + 3 * (size_t)m.nv);
        if (desc->velocities) {
            m.vel.assign(desc->velocities, desc->velocities + 3 * (size_t)m.nv);
        } else {
            m.vel.resize(3 * (size_t)m.nv);
            for (int32_t i = 0; i < m.nv; ++i)
                for (int d = 0; d < 3; ++d) m.vel[3 * i + d] = desc->initial_velocity[d];
        }
        m.tets.assign(desc->tets, desc->tets + 4 * (size_t)m.nt);
        m.mat.density = desc->density;
        m.mat.mu = desc->mu;
        m.mat.lambda = desc->lambda;
        m.mat.dampingHydrostatic = desc->damping_hydrostatic;
        m.mat.dampingDeviatoric = desc->damping_deviatoric;
        m.mat.exponentialVelDamping = desc->exponential_vel_damping;
        m.mat.constantVelDamping = desc->constant_vel_damping;
        m.mat.initializationType = desc->initialization_type;
        m.mat.initRatioG = desc->init_ratio_g;
        m.mat.frictionDynamic = desc->friction_dynamic;
        m.mat.frictionEpsV = desc->friction_eps_v;
        m.mat.hasNoGravZone = desc->has_no_grav_zone;
        m.mat.noGravZoneThreshold = desc->no_grav_zone_threshold;
        m.mat.maxVelocityMagnitude = desc->max_velocity_magnitude;
        m.fixed.assign(desc->fixed_points, desc->fixed_points + desc->num_fixed);
        for (int32_t t : m.tets) require(t >= 0 && t < m.nv, "tet vertex index out of range");
        Graph g = tet_vertex_graph(m.nv, m.nt, m.tets.data());
        if (desc->vertex_colors) {
            m.colors.assign(desc->vertex_colors, desc->vertex_colors + m.nv);
            for (int32_t c : m.colors) require(c >= 0, "negative vertex colour");
            require(coloring_valid(g, m.colors), "vertex colouring is not valid for this mesh");
        } else {
            m.colors = desc->coloring_algorithm == GVBD_COLOR_GREEDY ? color_greedy(g) : color_mcs(g);
        }
        const int32_t id = add_tet_mesh_host(ctx, scene, std::move(m));
        if (mesh_out) *mesh_out = id;
    });
}

int gvbd_scene_add_cloth_mesh(gvbd_context* ctx, int32_t scene, const gvbd_cloth_mesh_desc* desc,
                              int32_t* mesh_out) {
    return guard(ctx, [&] {
        require(ctx && desc, "NULL argument");
        require(desc->num_vertices > 0 && desc->num_faces > 0, "empty cloth mesh");
        require(desc->positions && desc->faces, "positions/faces are NULL");
        require(desc->num_fixed == 0 || desc->fixed_points, "fixed_points is NULL");
        const int t = desc->initialization_type;
        require(t == 0 || t == 2 || t == 3 || t == 4, "cloth initialization_type must be 0, 2, 3 or 4");
        ClothMeshHost m;
        m.nv = desc->num_vertices;
        m.nf = desc->num_faces;
        m.pos.assign(desc->positions, desc->positions + 3 * (size_t)m.nv);
        if (desc->velocities) {
            m.vel.assign(desc->velocities, desc->velocities + 3 * (size_t)m.nv);
        } else {
            m.vel.resize(3 * (size_t)m.nv);
            for (int32_t i = 0; i < m.nv; ++i)
                for (int d = 0; d < 3; ++d) m.vel[3 * i + d] = desc->initial_velocity[d];
        }
        m.faces.assign(desc->faces, desc->faces + 3 * (size_t)m.nf);
        m.mat.density = desc->density;
        m.mat.mu = desc->mu;
        m.mat.lambda = desc->lambda;
        m.mat.bendingStiffness = desc->bending_stiffness;
        m.mat.dampingStVK = desc->damping_stvk;
        m.mat.dampingBending = desc->damping_bending;
        m.mat.exponentialVelDamping = desc->exponential_vel_damping;
        m.mat.constantVelDamping = desc->constant_vel_damping;
        m.mat.initializationType = t;
        m.mat.frictionDynamic = desc->friction_dynamic;
        m.mat.frictionEpsV = desc->friction_eps_v;
        m.mat.hasNoGravZone = desc->has_no_grav_zone;
        m.mat.noGravZoneThreshold = desc->no_grav_zone_threshold;
        m.mat.maxVelocityMagnitude = desc->max_velocity_magnitude;
        m.fixed.assign(desc->fixed_points, desc->fixed_points + desc->num_fixed);
        for (int32_t v : m.faces) require(v >= 0 && v < m.nv, "face vertex index out of range");
        if (desc->vertex_colors) {
            m.colors.assign(desc->vertex_colors, desc->vertex_colors + m.nv);
            for (int32_t c : m.colors) require(c >= 0, "negative vertex colour");
        } else {
            Graph g = tri_vertex_graph(m.nv, m.nf, m.faces.data(), true);
            m.colors = desc->coloring_algorithm == GVBD_COLOR_GREEDY ? color_greedy(g) : color_mcs(g);
        }
        const int32_t id = add_cloth_mesh_host(ctx, scene, std::move(m));
        if (mesh_out) *mesh_out = id;
    });
}

namespace gvbd {
namespace {

std::vector<int32_t>& mesh_fixed(SceneObj& s, int32_t m) {
    return s.meshes[m].first == MESH_TET ? s.tets[s.meshes[m].second].fixed : s.cloths[s.meshes[m].second].fixed;
}
const std::vector<float>& mesh_positions(const SceneObj& s, int32_t m) {
    return s.meshes[m].first == MESH_TET ? s.tets[s.meshes[m].second].pos : s.cloths[s.meshes[m].second].pos;
}

// The scene time: the device copy once initialised (resident, or in the served scene's host copy).
float* scene_clock_host(SceneObj& s) {
    if (!s.hostCopy.empty() && s.dev.clock)
        return reinterpret_cast<float*>(s.hostCopy.data() + (reinterpret_cast<uintptr_t>(s.dev.clock) - s.hostCopyBase));
    return nullptr;
}

void set_scene_time(gvbd_context* ctx, SceneObj& s, float t) {
    if (!s.initialized) {
        s.time0 = t;
        return;
    }
    if (float* h = scene_clock_host(s)) {
        *h = t;
        return;
    }
    hip_check(hipStreamSynchronize(ctx->stream), "sync");
    hip_check(hipMemcpy(s.dev.clock, &t, sizeof(float), hipMemcpyHostToDevice), "time upload");
}

}  // namespace
}  // namespace gvbd

int gvbd_scene_add_deformer(gvbd_context* ctx, int32_t scene, const gvbd_deformer_desc* desc) {
    return guard(ctx, [&] {
        require(ctx && desc, "NULL argument");
        SceneObj& s = scene_of(ctx, scene);
        if (s.initialized) fail(GVBD_ERR_STATE, "deformers are added before the scene is initialised");
        require(desc->kind == GVBD_DEFORMER_ROTATOR || desc->kind == GVBD_DEFORMER_TRANSLATOR,
                "unrecognised deformer kind");
        require(desc->num_meshes >= 0 && (desc->num_meshes == 0 || (desc->meshes && desc->num_vertices)),
                "deformer meshes are NULL");
        DeformerHost D;
        D.prm.kind = desc->kind;
        for (int c = 0; c < 3; ++c) {
            D.prm.axis[c] = desc->axis[c];
            D.prm.speed[c] = desc->speed[c];
        }
        D.prm.startTime = desc->start_time;
        D.prm.omega = desc->angular_velocity;
        D.prm.endTime = desc->end_time;
        size_t at = 0;
        for (int32_t k = 0; k < desc->num_meshes; ++k) {
            const int32_t m = desc->meshes[k];
            require(m >= 0 && m < (int32_t)s.meshes.size(), "deformer mesh index out of range");
            require(desc->num_vertices[k] >= 0, "negative vertex count");
            require(desc->num_vertices[k] == 0 || desc->vertices, "deformer vertices are NULL");
            const std::vector<float>& P = mesh_positions(s, m);
            D.meshes.push_back(m);
            D.verts.emplace_back();
            D.a.emplace_back();
            D.b.emplace_back();
            for (int32_t i = 0; i < desc->num_vertices[k]; ++i) {
                const int32_t v = desc->vertices[at++];
                require(v >= 0 && v < s.meshVertices(m), "deformer vertex index out of range");
                // the device moves every deformed vertex of a phase in one launch: a vertex selected
                // twice would be written by two threads (the reference applies deformers in list
                // order), so each vertex may belong to one deformer only
                const bool taken = std::any_of(s.deformers.begin(), s.deformers.end(), [&](const DeformerHost& o) {
                    for (size_t j = 0; j < o.meshes.size(); ++j)
                        if (o.meshes[j] == m && std::find(o.verts[j].begin(), o.verts[j].end(), v) != o.verts[j].end())
                            return true;
                    return false;
                });
                if (taken || std::find(D.verts.back().begin(), D.verts.back().end(), v) != D.verts.back().end())
                    fail(GVBD_ERR_INVALID, "vertex " + std::to_string(v) + " of mesh " + std::to_string(m) +
                                               " is already selected by a deformer");
                D.verts.back().push_back(v);
                // Rotator: root = ((P - o) . axis) axis, RootP = P - root; Translator: P
                // (VBDDeformer.h:52-67, 139-148), in float
                const float p[3] = {P[3 * (size_t)v], P[3 * (size_t)v + 1], P[3 * (size_t)v + 2]};
                float op[3], root[3];
                for (int c = 0; c < 3; ++c) op[c] = p[c] - desc->center[c];
                const float dot = op[0] * desc->axis[0] + op[1] * desc->axis[1] + op[2] * desc->axis[2];
                for (int c = 0; c < 3; ++c) root[c] = dot * desc->axis[c];
                for (int c = 0; c < 3; ++c) {
                    D.a.back().push_back(desc->kind == GVBD_DEFORMER_ROTATOR ? root[c] : p[c]);
                    D.b.back().push_back(desc->kind == GVBD_DEFORMER_ROTATOR ? p[c] - root[c] : 0.f);
                }
            }
        }
        // setFixed / setFixedPointCloth (VBDDeformer.h:20-43, VBDClothDeformer.h:5-26)
        for (size_t k = 0; k < D.meshes.size(); ++k) {
            std::vector<int32_t>& fx = mesh_fixed(s, D.meshes[k]);
            for (int32_t v : D.verts[k])
                if (std::find(fx.begin(), fx.end(), v) == fx.end()) fx.push_back(v);
        }
        s.deformers.push_back(std::move(D));
    });
}

int gvbd_scene_add_attachments(gvbd_context* ctx, int32_t scene, const gvbd_attachment_desc* list, int32_t n) {
    return guard(ctx, [&] {
        require(ctx && (n == 0 || list) && n >= 0, "NULL argument");
        SceneObj& s = scene_of(ctx, scene);
        if (s.initialized) fail(GVBD_ERR_STATE, "attachments are added before the scene is initialised");
        auto cloth = [&](int32_t m) { return m >= 0 && m < (int32_t)s.meshes.size() && s.meshes[m].first == MESH_CLOTH; };
        for (int32_t i = 0; i < n; ++i) {
            const gvbd_attachment_desc& A = list[i];
            require(cloth(A.mesh), "attachment " + std::to_string(i) + ": mesh must be a cloth mesh of the scene");
            require(A.vertex >= 0 && A.vertex < s.meshVertices(A.mesh),
                    "attachment " + std::to_string(i) + ": vertex index out of range");
            if (A.target_mesh == -1) continue;
            require(cloth(A.target_mesh),
                    "attachment " + std::to_string(i) + ": target mesh must be -1 or a cloth mesh of the scene");
            for (int j = 0; j < 3; ++j) {
                require(A.face[j] >= 0 && A.face[j] < s.meshVertices(A.target_mesh),
                        "attachment " + std::to_string(i) + ": face vertex index out of range");
                // the face side reads the vertex-side mesh at the face vertex's index
                require(A.bary[j] == 0.f || A.face[j] < s.meshVertices(A.mesh),
                        "attachment " + std::to_string(i) +
                            ": face vertex index exceeds the vertex-side mesh (its face-side force reads it)");
            }
        }
        s.attachments.insert(s.attachments.end(), list, list + n);
    });
}

int gvbd_scene_get_time(gvbd_context* ctx, int32_t scene, float* cur_time) {
    return guard(ctx, [&] {
        require(ctx && cur_time, "NULL argument");
        SceneObj& s = scene_of(ctx, scene);
        if (!s.initialized) {
            *cur_time = s.time0;
        } else if (float* h = scene_clock_host(s)) {
            *cur_time = *h;
        } else {
            hip_check(hipStreamSynchronize(ctx->stream), "sync");
            hip_check(hipMemcpy(cur_time, s.dev.clock, sizeof(float), hipMemcpyDeviceToHost), "time read");
        }
    });
}

int gvbd_scene_set_time(gvbd_context* ctx, int32_t scene, float cur_time) {
    return guard(ctx, [&] {
        require(ctx != nullptr, "ctx is NULL");
        set_scene_time(ctx, scene_of(ctx, scene), cur_time);
    });
}

int gvbd_scene_initialize(gvbd_context* ctx, int32_t scene) {
    return guard(ctx, [&] {
        require(ctx != nullptr, "ctx is NULL");
        SceneObj& s = scene_of(ctx, scene);
        initialize_scene(ctx, s);
        s.plan = new_plan(ctx, {scene});
    });
}

int gvbd_scene_load_json(gvbd_context* ctx, const char* models_json, const char* params_json, const char* repo_root,
                         int32_t* scene_out) {
    return guard(ctx, [&] {
        require(ctx && models_json && params_json && scene_out, "NULL argument");
        LoadedScene ls;
        std::string err;
        if (!load_scene_json(models_json, params_json, repo_root ? repo_root : "", ls, err)) fail(GVBD_ERR_IO, err);
        int32_t sid = -1;
        int rc = gvbd_scene_create(ctx, &ls.params, &sid);
        if (rc != GVBD_OK) fail(rc, ctx->lastError);
        try {
            for (auto& mk : ls.order) {
                if (mk.first == MESH_TET)
                    add_tet_mesh_host(ctx, sid, std::move(ls.tets[mk.second]));
                else
                    add_cloth_mesh_host(ctx, sid, std::move(ls.cloths[mk.second]));
            }
            for (const LoadedDeformer& d : ls.deformers) {
                gvbd_deformer_desc dd = d.desc;
                std::vector<int32_t> counts;
                std::vector<int32_t> verts;
                for (const auto& l : d.vertices) {
                    counts.push_back((int32_t)l.size());
                    verts.insert(verts.end(), l.begin(), l.end());
                }
                dd.num_meshes = (int32_t)d.meshes.size();
                dd.meshes = d.meshes.data();
                dd.num_vertices = counts.data();
                dd.vertices = verts.data();
                rc = gvbd_scene_add_deformer(ctx, sid, &dd);
                if (rc != GVBD_OK) fail(rc, ctx->lastError);
            }
            if (!ls.attachments.empty()) {
                rc = gvbd_scene_add_attachments(ctx, sid, ls.attachments.data(), (int32_t)ls.attachments.size());
                if (rc != GVBD_OK) fail(rc, ctx->lastError);
            }
            SceneObj& s = scene_of(ctx, sid);
            initialize_scene(ctx, s);
            s.plan = new_plan(ctx, {sid});
        } catch (...) {
            (void)gvbd_scene_destroy(ctx, sid);
            throw;
        }
        *scene_out = sid;
    });
}

int gvbd_scene_num_meshes(gvbd_context* ctx, int32_t scene, int32_t* n) {
    return guard(ctx, [&] {
        require(ctx && n, "NULL argument");
        *n = (int32_t)scene_of(ctx, scene).meshes.size();
    });
}

int gvbd_scene_mesh_info(gvbd_context* ctx, int32_t scene, int32_t mesh, int32_t* num_vertices, int32_t* num_elements) {
    return guard(ctx, [&] {
        require(ctx != nullptr, "ctx is NULL");
        SceneObj& s = scene_of(ctx, scene);
        require(mesh >= 0 && mesh < (int32_t)s.meshes.size(), "mesh index out of range");
        if (num_vertices) *num_vertices = s.meshVertices(mesh);
        if (num_elements)
            *num_elements = s.meshes[mesh].first == MESH_TET ? s.tets[s.meshes[mesh].second].nt
                                                            : s.cloths[s.meshes[mesh].second].nf;
    });
}

int gvbd_scene_num_parallel_groups(gvbd_context* ctx, int32_t scene, int32_t* n) {
    return guard(ctx, [&] {
        require(ctx && n, "NULL argument");
        SceneObj& s = scene_of(ctx, scene);
        if (!s.initialized) fail(GVBD_ERR_STATE, "scene not initialised");
        *n = (int32_t)s.groups.size();
    });
}

int gvbd_scene_parallel_group(gvbd_context* ctx, int32_t scene, int32_t group, int32_t* pairs, int32_t* count) {
    return guard(ctx, [&] {
        require(ctx && count, "NULL argument");
        SceneObj& s = scene_of(ctx, scene);
        if (!s.initialized) fail(GVBD_ERR_STATE, "scene not initialised");
        require(group >= 0 && group < (int32_t)s.groups.size(), "group index out of range");
        const auto& g = s.groups[group];
        *count = (int32_t)g.size() / 2;
        if (pairs) std::memcpy(pairs, g.data(), g.size() * sizeof(int32_t));
    });
}

int gvbd_scene_mesh_constants(gvbd_context* ctx, int32_t scene, int32_t mesh, float* dminv, float* volume,
                              float* mass) {
    return guard(ctx, [&] {
        require(ctx != nullptr, "ctx is NULL");
        SceneObj& s = scene_of(ctx, scene);
        require(mesh >= 0 && mesh < (int32_t)s.meshes.size(), "mesh index out of range");
        const std::vector<float>*D, *V, *M;
        if (s.meshes[mesh].first == MESH_TET) {
            const TetMeshHost& m = s.tets[s.meshes[mesh].second];
            D = &m.dminv;
            V = &m.vol;
            M = &m.mass;
        } else {
            const ClothMeshHost& m = s.cloths[s.meshes[mesh].second];
            D = &m.dminv;
            V = &m.area;
            M = &m.mass;
        }
        if (dminv) std::memcpy(dminv, D->data(), D->size() * sizeof(float));
        if (volume) std::memcpy(volume, V->data(), V->size() * sizeof(float));
        if (mass) std::memcpy(mass, M->data(), M->size() * sizeof(float));
    });
}

static PlanObj& scene_plan(gvbd_context* ctx, int32_t scene) {
    require(ctx != nullptr, "ctx is NULL");
    SceneObj& s = scene_of(ctx, scene);
    if (s.plan < 0) fail(GVBD_ERR_STATE, "scene not initialised");
    if (!s.dev.arena) fail(GVBD_ERR_STATE, "scene is held in host memory by its server (not resident)");
    return plan_of(ctx, s.plan);
}

static void set_schedule(PlanObj& p, const gvbd_schedule& sc) {
    require(valid_schedule(sc.lanes_per_vertex, sc.block),
            "schedule must be lanes_per_vertex in {1,2,4,8,16}, block in {64,128,256}");
    require(sc.mode == GVBD_MODE_GROUP_LAUNCHES || sc.mode == GVBD_MODE_WORK_QUEUE, "unknown schedule mode");
    const int vpb = sc.block / sc.lanes_per_vertex;
    require(sc.vertices_per_entry >= 0 && sc.vertices_per_entry <= 65536 && sc.vertices_per_entry % vpb == 0,
            "vertices_per_entry must be a multiple of block / lanes_per_vertex (at most 65536)");
    require(sc.blocks_per_cu >= 0 && sc.blocks_per_cu <= 32, "blocks_per_cu must be in [0, 32]");
    require(sc.spin_limit >= 0, "spin_limit must be non-negative");
    require(valid_stage(sc.lanes_per_vertex, sc.block, sc.stage_records),
            "stage_records must be 0, or 32 or 64 with lanes_per_vertex in {4,8,16} and block in {128,256}");
    require(sc.stage_records == 0 || sc.mode == GVBD_MODE_GROUP_LAUNCHES,
            "LDS staging applies to group launches only");
    require(!(p.contact && sc.mode == GVBD_MODE_WORK_QUEUE),
            "scenes with collision handling run with group launches (detection runs between the launches)");
    require(!(p.contact && sc.stage_records != 0), "LDS staging is not available for scenes with collision handling");
    require(sc.xcd_blocks == 0 || sc.xcd_blocks == 1 || sc.xcd_blocks == 4 || sc.xcd_blocks == 16 ||
                sc.xcd_blocks == 64,
            "xcd_blocks must be 0, 1, 4, 16 or 64");
    p.lpv = sc.lanes_per_vertex;
    p.xcd = sc.xcd_blocks;
    p.stage = sc.stage_records;
    p.block = sc.block;
    p.mode = sc.mode;
    p.vertsPerEntry = sc.vertices_per_entry;
    p.blocksPerCu = sc.blocks_per_cu;
    p.spinLimit = sc.spin_limit > 0 ? sc.spin_limit : kDefaultSpinLimit;
    free_plan(p);
}

static void get_schedule(const PlanObj& p, gvbd_schedule* sc) {
    require(sc != nullptr, "schedule is NULL");
    *sc = gvbd_schedule{};
    sc->lanes_per_vertex = p.lpv;
    sc->block = p.block;
    sc->stage_records = p.stage;
    sc->xcd_blocks = p.xcd;
    sc->mode = p.mode;
    sc->vertices_per_entry = p.vertsPerEntry;
    sc->blocks_per_cu = p.blocksPerCu;
    sc->spin_limit = p.spinLimit;
}

// One VBD iteration issued by the host: in work-queue mode it is the next epoch of the
// counters (a fresh epoch range is started when the ticket slots run out), so it is exactly
// one queue launch.
static void run_single_iteration(gvbd_context* ctx, PlanObj& p, int32_t iteration) {
    require(iteration >= 0, "iteration must be non-negative");
    build_plan(ctx, p);
    if (p.mode == GVBD_MODE_WORK_QUEUE && p.epoch >= p.ticketSlots) plan_reset_counters(ctx, p);
    const int epoch = p.epoch;
    plan_iteration(ctx, p, iteration, epoch);
    if (p.mode == GVBD_MODE_WORK_QUEUE) p.epoch = epoch + 1;
}

int gvbd_scene_set_schedule(gvbd_context* ctx, int32_t scene, int32_t lanes_per_vertex, int32_t block) {
    return guard(ctx, [&] {
        PlanObj& p = scene_plan(ctx, scene);
        gvbd_schedule sc;
        get_schedule(p, &sc);
        sc.lanes_per_vertex = lanes_per_vertex;
        sc.block = block;
        sc.vertices_per_entry = 0;
        set_schedule(p, sc);
    });
}

int gvbd_scene_set_schedule_ex(gvbd_context* ctx, int32_t scene, const gvbd_schedule* schedule) {
    return guard(ctx, [&] {
        require(schedule != nullptr, "schedule is NULL");
        set_schedule(scene_plan(ctx, scene), *schedule);
    });
}

int gvbd_scene_get_schedule(gvbd_context* ctx, int32_t scene, gvbd_schedule* schedule) {
    return guard(ctx, [&] { get_schedule(scene_plan(ctx, scene), schedule); });
}

int gvbd_scene_run_iteration(gvbd_context* ctx, int32_t scene, int32_t iteration) {
    return guard(ctx, [&] { run_single_iteration(ctx, scene_plan(ctx, scene), iteration); });
}

int gvbd_scene_run_step(gvbd_context* ctx, int32_t scene) { return gvbd_scene_run_steps(ctx, scene, 1); }

int gvbd_scene_run_steps(gvbd_context* ctx, int32_t scene, int32_t n) {
    return guard(ctx, [&] {
        require(n >= 0, "bad argument");
        plan_run(ctx, scene_plan(ctx, scene), n);
    });
}

int gvbd_scene_apply_initial_step(gvbd_context* ctx, int32_t scene) {
    return guard(ctx, [&] {
        PlanObj& p = scene_plan(ctx, scene);
        build_plan(ctx, p);
        plan_initial_step(ctx, p);
    });
}

int gvbd_scene_solve_parallel_group(gvbd_context* ctx, int32_t scene, int32_t group, int32_t iteration) {
    return guard(ctx, [&] {
        PlanObj& p = scene_plan(ctx, scene);
        build_plan(ctx, p);
        require(group >= 0 && group < p.numGroups, "group index out of range");
        if (p.contact) plan_contact_snapshot(ctx, p);
        plan_group(ctx, p, group, iteration);
    });
}

int gvbd_scene_update_velocities(gvbd_context* ctx, int32_t scene) {
    return guard(ctx, [&] {
        PlanObj& p = scene_plan(ctx, scene);
        build_plan(ctx, p);
        plan_velocity(ctx, p);
    });
}

int gvbd_scene_collision_detection(gvbd_context* ctx, int32_t scene) {
    return guard(ctx, [&] {
        require(ctx != nullptr, "ctx is NULL");
        SceneObj& s = scene_of(ctx, scene);
        if (!s.cd) fail(GVBD_ERR_STATE, "scene has no collision handling");
        scene_collision_detection(s, true);
    });
}

int gvbd_scene_contact_stats(gvbd_context* ctx, int32_t scene, gvbd_contact_stats* out) {
    return guard(ctx, [&] {
        require(ctx && out, "NULL argument");
        SceneObj& s = scene_of(ctx, scene);
        if (!s.cd) fail(GVBD_ERR_STATE, "scene has no collision handling");
        collect_detection_time(s);
        *out = s.contactStats;
        if (!s.cd->penetration()) out->detections += s.cd->gatedPasses();  // device-decided passes
        out->vf_pairs = out->ee_pairs = 0;
        if (!s.cd->penetration() && out->detections > 0) {
            hip_check(hipStreamSynchronize(ctx->stream), "sync");
            const DetectorResultCounts c = s.cd->counts();
            out->vf_pairs = c.vfPairs;
            out->ee_pairs = c.eeContacts;
            return;
        }
        if (s.contactStats.detections > 0) {
            hip_check(hipStreamSynchronize(ctx->stream), "sync");
            if (s.cd->penetration()) {
                // penetration contact: vertices with a hit face (DCD + CCD) in vf_pairs
                const PenetrationCounts c = s.cd->penetrationCounts();
                out->vf_pairs = c.dcdHits + c.ccdHits;
            } else {
                const DetectorResultCounts c = s.cd->counts();
                out->vf_pairs = c.vfPairs;
                out->ee_pairs = c.eeContacts;
            }
        }
    });
}

int gvbd_scene_penetration_results(gvbd_context* ctx, int32_t scene, int32_t* hit_mesh, int32_t* hit_face,
                                   uint8_t* flags, int32_t* counts) {
    return guard(ctx, [&] {
        require(ctx != nullptr, "ctx is NULL");
        SceneObj& s = scene_of(ctx, scene);
        if (!s.cd || !s.cd->penetration()) fail(GVBD_ERR_STATE, "scene has no penetration contact");
        hip_check(hipStreamSynchronize(ctx->stream), "sync");
        std::vector<int32_t> face(s.nv);
        std::vector<uint8_t> fl(s.nv);
        s.cd->penetrationResults(face.data(), fl.data());
        // scene face ids -> (mesh, boundary face of that mesh)
        std::vector<int32_t> faceOff(s.meshes.size() + 1, 0);
        for (size_t im = 0; im < s.meshes.size(); ++im)
            faceOff[im + 1] = faceOff[im] + (int32_t)s.tetSurfaces[s.meshes[im].second].faces.size() / 3;
        for (int32_t v = 0; v < s.nv; ++v) {
            int32_t m = -1, f = -1;
            if (face[v] >= 0) {
                m = (int32_t)(std::upper_bound(faceOff.begin(), faceOff.end(), face[v]) - faceOff.begin()) - 1;
                f = face[v] - faceOff[m];
            }
            if (hit_mesh) hit_mesh[v] = m;
            if (hit_face) hit_face[v] = f;
            if (flags) flags[v] = fl[v];
        }
        if (counts) {
            const PenetrationCounts c = s.cd->penetrationCounts();
            counts[0] = c.dcdHits;
            counts[1] = c.penetratedNoFace;
            counts[2] = c.ccdHits;
            counts[3] = c.walkOverflows;
            counts[4] = c.fullSearches;
        }
    });
}

namespace gvbd {
namespace {

// Sizes of one scene's convergence evaluation: tets, tet vertices, tet slots and the partial-sum
// blocks of each kernel.
struct ConvergenceSizes {
    int64_t ntet = 0, ntetVerts = 0;
    int nslots = 0, bs = 0, bv = 0, bt = 0;
    size_t recBytes = 0, vertBytes = 0;
};

ConvergenceSizes convergence_sizes(const SceneObj& s) {
    ConvergenceSizes z;
    for (const TetMeshHost& m : s.tets) {
        z.ntet += m.nt;
        z.ntetVerts += m.nv;
    }
    z.nslots = (int)s.numTetSlots();
    z.bs = convergence_blocks(std::max(z.nslots, 1));
    z.bt = convergence_blocks((int)std::max<int64_t>(z.ntet, 1));
    z.bv = convergence_blocks((int)std::max<int64_t>(z.ntetVerts, 1));
    z.recBytes = sizeof(TetEnergyRec) * (size_t)std::max<int64_t>(z.ntet, 1);
    z.vertBytes = (sizeof(int32_t) * (size_t)std::max<int64_t>(z.ntetVerts, 1) + 15) & ~(size_t)15;
    return z;
}

void check_convergence_scene(const SceneObj& s) {
    if (!s.initialized) fail(GVBD_ERR_STATE, "scene not initialised");
    if (!s.dev.arena) fail(GVBD_ERR_STATE, "scene is held in host memory by its server (not resident)");
}

// merit energy parts and the average force norm from the three totals (force, inertia, elastic)
void convergence_outputs(const SceneObj& s, const ConvergenceSizes& z, double fsum, double isum, double esum,
                         double* energy, double* inertia_energy, double* elastic_energy, double* avg_force_norm) {
    const double eIn = 0.5 * (double)s.sp.dtSqrRecip * isum;
    if (inertia_energy) *inertia_energy = eIn;
    if (elastic_energy) *elastic_energy = esum;
    if (energy) *energy = eIn + esum;
    if (avg_force_norm) *avg_force_norm = z.ntetVerts ? fsum / (double)z.ntetVerts : 0.0;
}

}  // namespace
}  // namespace gvbd

int gvbd_scene_evaluate_convergence(gvbd_context* ctx, int32_t scene, double* energy, double* inertia_energy,
                                    double* elastic_energy, double* avg_force_norm) {
    return guard(ctx, [&] {
        require(ctx != nullptr, "ctx is NULL");
        SceneObj& s = scene_of(ctx, scene);
        check_convergence_scene(s);
        const ConvergenceSizes z = convergence_sizes(s);
        const int bs = z.bs, bv = z.bv, bt = z.bt;
        const size_t need = z.recBytes + z.vertBytes + sizeof(double) * (size_t)(bs + bv + bt);
        if (!s.convBuf) {
            // the tets of every tet mesh with scene vertex ids (built once)
            std::vector<TetEnergyRec> recs;
            std::vector<int32_t> verts;
            recs.reserve((size_t)z.ntet);
            for (size_t im = 0; im < s.meshes.size(); ++im) {
                if (s.meshes[im].first != MESH_TET) continue;
                const TetMeshHost& m = s.tets[s.meshes[im].second];
                const int32_t off = s.vertexOffset[im];
                for (int32_t i = 0; i < m.nv; ++i) verts.push_back(off + i);
                for (int32_t t = 0; t < m.nt; ++t) {
                    TetEnergyRec r{};
                    r.v = make_int4(off + m.tets[4 * (size_t)t], off + m.tets[4 * (size_t)t + 1],
                                    off + m.tets[4 * (size_t)t + 2], off + m.tets[4 * (size_t)t + 3]);
                    for (int k = 0; k < 9; ++k) r.dminv[k] = m.dminv[9 * (size_t)t + k];
                    r.A = m.vol[t];
                    r.mu = m.mat.mu;
                    r.lmbd = m.mat.lambda;
                    recs.push_back(r);
                }
            }
            hip_check(hipSetDevice(ctx->device), "hipSetDevice");
            if (hipMalloc(&s.convBuf, need) != hipSuccess) {
                (void)hipGetLastError();
                s.convBuf = nullptr;
                fail(GVBD_ERR_NOMEM, "hipMalloc for convergence evaluation failed");
            }
            s.convBytes = need;
            if (!recs.empty())
                hip_check(hipMemcpyAsync(s.convBuf, recs.data(), sizeof(TetEnergyRec) * recs.size(),
                                         hipMemcpyHostToDevice, ctx->stream),
                          "upload");
            if (!verts.empty())
                hip_check(hipMemcpyAsync((char*)s.convBuf + z.recBytes, verts.data(), sizeof(int32_t) * verts.size(),
                                         hipMemcpyHostToDevice, ctx->stream),
                          "upload");
            hip_check(hipStreamSynchronize(ctx->stream), "upload sync");
        }
        auto* recs = (const TetEnergyRec*)s.convBuf;
        auto* tverts = (const int*)((char*)s.convBuf + z.recBytes);
        auto* part = (double*)((char*)s.convBuf + z.recBytes + z.vertBytes);
        hip_check(launch_convergence(s.view, z.nslots, recs, (int)z.ntet, tverts, (int)z.ntetVerts, part, part + bs,
                                     part + bs + bv, ctx->stream),
                  "convergence launch");
        std::vector<double> h((size_t)(bs + bv + bt), 0.0);
        hip_check(hipMemcpyAsync(h.data(), part, sizeof(double) * h.size(), hipMemcpyDeviceToHost, ctx->stream),
                  "download");
        hip_check(hipStreamSynchronize(ctx->stream), "sync");
        double fsum = 0.0, isum = 0.0, esum = 0.0;
        if (z.nslots > 0)
            for (int b = 0; b < bs; ++b) fsum += h[b];
        if (z.ntetVerts > 0)
            for (int b = 0; b < bv; ++b) isum += h[bs + b];
        if (z.ntet > 0)
            for (int b = 0; b < bt; ++b) esum += h[bs + bv + b];
        convergence_outputs(s, z, fsum, isum, esum, energy, inertia_energy, elastic_energy, avg_force_norm);
    });
}

int gvbd_scene_evaluate_convergence_noalloc(gvbd_context* ctx, int32_t scene, double* energy, double* inertia_energy,
                                            double* elastic_energy, double* avg_force_norm) {
    return guard(ctx, [&] {
        require(ctx != nullptr, "ctx is NULL");
        SceneObj& s = scene_of(ctx, scene);
        check_convergence_scene(s);
        if (!s.convBuf)
            fail(GVBD_ERR_STATE, "the scene's tet energy records are not resident: call "
                                 "gvbd_scene_evaluate_convergence once at set-up (it allocates them)");
        if (!ctx->convScratch) fail(GVBD_ERR_STATE, "convergence scratch address not set up for this device");
        const ConvergenceSizes z = convergence_sizes(s);
        // partials of the three kernels, then the three totals, within the fixed scratch
        const int64_t needDoubles = (int64_t)z.bs + z.bv + z.bt + 3;
        if (needDoubles > kConvergenceScratchDoubles)
            fail(GVBD_ERR_INVALID, "scene too large for the allocation-free convergence evaluation: " +
                                       std::to_string(needDoubles) + " scratch doubles needed, at most " +
                                       std::to_string(kConvergenceScratchDoubles) + " (about " +
                                       std::to_string((int64_t)(kConvergenceScratchDoubles - 3) * 256) +
                                       " tets + tet vertices + free tet vertices)");
        double* part = ctx->convScratch;
        double* totals = ctx->convScratch + kConvergenceScratchDoubles - 3;
        const auto* recs = (const TetEnergyRec*)s.convBuf;
        const auto* tverts = (const int*)((const char*)s.convBuf + z.recBytes);
        hip_check(launch_convergence(s.view, z.nslots, recs, (int)z.ntet, tverts, (int)z.ntetVerts, part,
                                     part + z.bs, part + z.bs + z.bv, ctx->stream),
                  "convergence launch");
        // kernels that did not run (no slots, tets or tet vertices) contribute zero partials
        hip_check(launch_convergence_final_sum(part, 0, z.nslots > 0 ? z.bs : 0, z.bs, z.ntetVerts > 0 ? z.bv : 0,
                                               z.bs + z.bv, z.ntet > 0 ? z.bt : 0, totals, ctx->stream),
                  "final sum launch");
        double h[3] = {0.0, 0.0, 0.0};
        hip_check(hipMemcpyAsync(h, totals, sizeof(h), hipMemcpyDeviceToHost, ctx->stream), "download");
        hip_check(hipStreamSynchronize(ctx->stream), "sync");
        convergence_outputs(s, z, h[0], h[1], h[2], energy, inertia_energy, elastic_energy, avg_force_norm);
    });
}

int gvbd_scene_record_graph(gvbd_context* ctx, int32_t scene, int32_t enable) {
    return guard(ctx, [&] {
        PlanObj& p = scene_plan(ctx, scene);
        p.wantGraph = enable != 0;
        free_plan_graph(p);
        if (p.wantGraph) {
            build_plan(ctx, p);
            plan_capture(ctx, p);
        }
    });
}

static void mesh_state_io(gvbd_context* ctx, int32_t scene, int32_t mesh, float* pos, float* vel, bool get) {
    require(ctx != nullptr, "ctx is NULL");
    SceneObj& s = scene_of(ctx, scene);
    if (!s.initialized) fail(GVBD_ERR_STATE, "scene not initialised");
    require(mesh >= 0 && mesh < (int32_t)s.meshes.size(), "mesh index out of range");
    const int32_t off = s.vertexOffset[mesh], n = s.meshVertices(mesh);
    std::vector<float4> tmp(n);
    hip_check(hipStreamSynchronize(ctx->stream), "sync");
    check_queue_status(ctx);
    float4* dev[2] = {s.dev.pos + off, s.dev.vel + off};
    float* host[2] = {pos, vel};
    // a scene its server holds in host memory is read and written there
    const bool resident = s.dev.arena != nullptr;
    auto held = [&](float4* d) { return s.hostCopy.data() + ((uintptr_t)d - s.hostCopyBase); };
    for (int a = 0; a < 2; ++a) {
        if (!host[a]) continue;
        if (get) {
            if (resident)
                hip_check(hipMemcpy(tmp.data(), dev[a], sizeof(float4) * n, hipMemcpyDeviceToHost), "download");
            else
                std::memcpy(tmp.data(), held(dev[a]), sizeof(float4) * n);
            for (int32_t i = 0; i < n; ++i) {
                host[a][3 * i] = tmp[i].x;
                host[a][3 * i + 1] = tmp[i].y;
                host[a][3 * i + 2] = tmp[i].z;
            }
        } else {
            for (int32_t i = 0; i < n; ++i) tmp[i] = make_float4(host[a][3 * i], host[a][3 * i + 1], host[a][3 * i + 2], 0.f);
            if (resident)
                hip_check(hipMemcpy(dev[a], tmp.data(), sizeof(float4) * n, hipMemcpyHostToDevice), "upload");
            else
                std::memcpy(held(dev[a]), tmp.data(), sizeof(float4) * n);
        }
    }
}

int gvbd_scene_solve_failures(gvbd_context* ctx, int32_t scene, int32_t reset, int64_t* count) {
    return guard(ctx, [&] {
        require(ctx && count, "NULL argument");
        SceneObj& s = scene_of(ctx, scene);
        if (!s.initialized) fail(GVBD_ERR_STATE, "scene not initialised");
        hip_check(hipStreamSynchronize(ctx->stream), "sync");
        int32_t n = 0;
        int* dev = s.view.solveFail;
        if (s.dev.arena) {
            hip_check(hipMemcpy(&n, dev, sizeof(n), hipMemcpyDeviceToHost), "download");
            if (reset) hip_check(hipMemset(dev, 0, sizeof(n)), "counter reset");
        } else {  // held in host memory by its server
            char* held = s.hostCopy.data() + ((uintptr_t)dev - s.hostCopyBase);
            std::memcpy(&n, held, sizeof(n));
            if (reset) std::memset(held, 0, sizeof(n));
        }
        *count = n;
    });
}

int gvbd_scene_get_mesh_state(gvbd_context* ctx, int32_t scene, int32_t mesh, float* positions, float* velocities) {
    return guard(ctx, [&] { mesh_state_io(ctx, scene, mesh, positions, velocities, true); });
}

int gvbd_scene_set_mesh_state(gvbd_context* ctx, int32_t scene, int32_t mesh, const float* positions,
                              const float* velocities) {
    return guard(ctx, [&] {
        mesh_state_io(ctx, scene, mesh, const_cast<float*>(positions), const_cast<float*>(velocities), false);
    });
}

// ---- scene files ----

static std::vector<io::MeshState> scene_states(gvbd_context* ctx, int32_t scene) {
    SceneObj& s = scene_of(ctx, scene);
    std::vector<io::MeshState> out(s.meshes.size());
    for (size_t m = 0; m < s.meshes.size(); ++m) {
        const size_t n = 3 * (size_t)s.meshVertices((int32_t)m);
        out[m].position.resize(n);
        out[m].velocities.resize(n);
        mesh_state_io(ctx, scene, (int32_t)m, out[m].position.data(), out[m].velocities.data(), true);
    }
    return out;
}

// The surface of one mesh at its current positions: a tet mesh's boundary vertices (ascending
// tet-vertex id) and boundary faces renumbered to them; a cloth mesh whole.
static void mesh_surface(SceneObj& s, int32_t m, const std::vector<float>& pos, std::vector<float>& verts,
                         std::vector<int32_t>& faces) {
    if (s.meshes[m].first == MESH_CLOTH) {
        verts = pos;
        faces = s.cloths[s.meshes[m].second].faces;
        return;
    }
    const TetMeshHost& t = s.tets[s.meshes[m].second];
    std::vector<int32_t> tf;
    std::string err;
    if (!build_tet_surface(t.nv, t.nt, t.tets.data(), t.pos.data(), tf, err)) fail(GVBD_ERR_INVALID, err);
    std::vector<int32_t> remap(t.nv, -1);
    for (int32_t v : tf) remap[v] = 0;
    verts.clear();
    int32_t k = 0;
    for (int32_t v = 0; v < t.nv; ++v)
        if (remap[v] == 0) {
            remap[v] = k++;
            verts.insert(verts.end(), {pos[3 * v], pos[3 * v + 1], pos[3 * v + 2]});
        }
    faces.resize(tf.size());
    for (size_t i = 0; i < tf.size(); ++i) faces[i] = remap[tf[i]];
}

int gvbd_scene_save_state(gvbd_context* ctx, int32_t scene, const char* json_path, int32_t frame_id, double cur_time,
                          int32_t binary) {
    return guard(ctx, [&] {
        require(ctx && json_path, "NULL argument");
        io::write_state(json_path, frame_id, cur_time, binary != 0, scene_states(ctx, scene));
    });
}

int gvbd_scene_load_state(gvbd_context* ctx, int32_t scene, const char* json_path, int32_t* frame_id,
                          double* cur_time) {
    return guard(ctx, [&] {
        require(ctx && json_path, "NULL argument");
        SceneObj& s = scene_of(ctx, scene);
        if (!s.initialized) fail(GVBD_ERR_STATE, "scene not initialised");
        int32_t fid = 0;
        double t = 0.0;
        std::vector<io::MeshState> st;
        io::read_state(json_path, fid, t, st);
        const size_t nm = std::min(st.size(), s.meshes.size());
        for (size_t m = 0; m < nm; ++m)
            if (st[m].position.size() != 3 * (size_t)s.meshVertices((int32_t)m))
                fail(GVBD_ERR_INVALID, "recovery state: mesh " + std::to_string(m) + " holds " +
                                           std::to_string(st[m].position.size() / 3) + " vertices, the scene's " +
                                           std::to_string(s.meshVertices((int32_t)m)));
        for (size_t m = 0; m < nm; ++m)
            mesh_state_io(ctx, scene, (int32_t)m, st[m].position.data(),
                          st[m].velocities.empty() ? nullptr : st[m].velocities.data(), false);
        set_scene_time(ctx, s, (float)t);  // PhysicsState::toPhysics sets curTime (BasePhysicsFramework.h:178)
        if (frame_id) *frame_id = fid;
        if (cur_time) *cur_time = t;
    });
}

int gvbd_scene_write_ply(gvbd_context* ctx, int32_t scene, const char* path, int32_t all_in_one) {
    return guard(ctx, [&] {
        require(ctx && path, "NULL argument");
        SceneObj& s = scene_of(ctx, scene);
        const std::vector<io::MeshState> st = scene_states(ctx, scene);
        std::vector<float> allV, v;
        std::vector<int32_t> allF, f;
        const std::string p(path);
        const size_t slash = p.find_last_of("/\\"), dot = p.find_last_of('.');
        const bool hasExt = dot != std::string::npos && (slash == std::string::npos || dot > slash);
        const std::string base = hasExt ? p.substr(0, dot) : p, ext = hasExt ? p.substr(dot) : std::string();
        for (size_t m = 0; m < s.meshes.size(); ++m) {
            mesh_surface(s, (int32_t)m, st[m].position, v, f);
            if (all_in_one) {
                const int32_t off = (int32_t)(allV.size() / 3);
                allV.insert(allV.end(), v.begin(), v.end());
                for (int32_t i : f) allF.push_back(i + off);
            } else {
                char num[16];
                std::snprintf(num, sizeof(num), "%06d", (int)m);
                io::write_ply(base + "_Model_" + num + ext, v, f);
            }
        }
        if (all_in_one) io::write_ply(p, allV, allF);
    });
}

int gvbd_scene_write_binary_frame(gvbd_context* ctx, int32_t scene, const char* path) {
    return guard(ctx, [&] {
        require(ctx && path, "NULL argument");
        SceneObj& s = scene_of(ctx, scene);
        const std::vector<io::MeshState> st = scene_states(ctx, scene);
        std::vector<float> all, v;
        std::vector<int32_t> f;
        for (size_t m = 0; m < s.meshes.size(); ++m) {
            mesh_surface(s, (int32_t)m, st[m].position, v, f);
            all.insert(all.end(), v.begin(), v.end());
        }
        io::write_binary(path, all);
    });
}

int gvbd_state_write(const char* json_path, int32_t frame_id, double cur_time, int32_t binary, int32_t num_meshes,
                     const int32_t* num_vertices, const float* positions, const float* velocities) {
    return guard(nullptr, [&] {
        require(json_path && num_meshes >= 0 && (num_meshes == 0 || (num_vertices && positions)), "bad argument");
        std::vector<io::MeshState> st(num_meshes);
        size_t off = 0;
        for (int32_t m = 0; m < num_meshes; ++m) {
            require(num_vertices[m] >= 0, "negative vertex count");
            const size_t n = 3 * (size_t)num_vertices[m];
            st[m].position.assign(positions + off, positions + off + n);
            if (velocities) st[m].velocities.assign(velocities + off, velocities + off + n);
            off += n;
        }
        require(velocities != nullptr || !binary, "a binary recovery state needs velocities");
        io::write_state(json_path, frame_id, cur_time, binary != 0, st);
    });
}

int gvbd_state_read(const char* json_path, int32_t max_meshes, int32_t* num_meshes, int32_t* num_vertices,
                    int64_t max_floats, float* positions, float* velocities, int32_t* frame_id, double* cur_time) {
    return guard(nullptr, [&] {
        require(json_path && num_meshes, "bad argument");
        int32_t fid = 0;
        double t = 0.0;
        std::vector<io::MeshState> st;
        io::read_state(json_path, fid, t, st);
        *num_meshes = (int32_t)st.size();
        if (frame_id) *frame_id = fid;
        if (cur_time) *cur_time = t;
        size_t total = 0;
        for (size_t m = 0; m < st.size(); ++m) {
            if (num_vertices && (int32_t)m < max_meshes) num_vertices[m] = (int32_t)(st[m].position.size() / 3);
            total += st[m].position.size();
        }
        if (!positions) return;
        require((int64_t)st.size() <= max_meshes && (int64_t)total <= max_floats, "buffers too small");
        size_t off = 0;
        for (const auto& m : st) {
            std::memcpy(positions + off, m.position.data(), sizeof(float) * m.position.size());
            if (velocities) {
                if (m.velocities.empty())
                    std::fill(velocities + off, velocities + off + m.position.size(), 0.f);
                else
                    std::memcpy(velocities + off, m.velocities.data(), sizeof(float) * m.velocities.size());
            }
            off += m.position.size();
        }
    });
}

int gvbd_write_ply(const char* path, int32_t num_vertices, const float* positions, int32_t num_faces,
                   const int32_t* faces) {
    return guard(nullptr, [&] {
        require(path && num_vertices >= 0 && num_faces >= 0 && (num_vertices == 0 || positions) &&
                    (num_faces == 0 || faces),
                "bad argument");
        for (int32_t i = 0; i < 3 * num_faces; ++i) require(faces[i] >= 0 && faces[i] < num_vertices, "face index out of range");
        io::write_ply(path, std::vector<float>(positions, positions + 3 * (size_t)num_vertices),
                      std::vector<int32_t>(faces, faces + 3 * (size_t)num_faces));
    });
}

int gvbd_batch_create(gvbd_context* ctx, const int32_t* scenes, int32_t n, int32_t* batch_out) {
    return guard(ctx, [&] {
        require(ctx && scenes && batch_out && n > 0, "bad argument");
        std::vector<int32_t> ids(scenes, scenes + n);
        std::vector<int32_t> sorted = ids;
        std::sort(sorted.begin(), sorted.end());
        require(std::adjacent_find(sorted.begin(), sorted.end()) == sorted.end(), "duplicate scene in batch");
        const int32_t pid = new_plan(ctx, ids);
        const int32_t bid = ctx->nextBatch++;
        ctx->batches[bid] = pid;
        *batch_out = bid;
    });
}

int gvbd_batch_destroy(gvbd_context* ctx, int32_t batch) {
    return guard(ctx, [&] {
        require(ctx != nullptr, "ctx is NULL");
        PlanObj& p = batch_plan(ctx, batch);
        hip_check(hipStreamSynchronize(ctx->stream), "sync");
        release_plan(p);
        ctx->plans.erase(ctx->batches[batch]);
        ctx->batches.erase(batch);
    });
}

int gvbd_batch_set_schedule(gvbd_context* ctx, int32_t batch, int32_t lanes_per_vertex, int32_t block) {
    return guard(ctx, [&] {
        require(ctx != nullptr, "ctx is NULL");
        PlanObj& p = batch_plan(ctx, batch);
        gvbd_schedule sc;
        get_schedule(p, &sc);
        sc.lanes_per_vertex = lanes_per_vertex;
        sc.block = block;
        sc.vertices_per_entry = 0;
        set_schedule(p, sc);
    });
}

int gvbd_batch_set_schedule_ex(gvbd_context* ctx, int32_t batch, const gvbd_schedule* schedule) {
    return guard(ctx, [&] {
        require(ctx != nullptr && schedule != nullptr, "NULL argument");
        set_schedule(batch_plan(ctx, batch), *schedule);
    });
}

int gvbd_batch_get_schedule(gvbd_context* ctx, int32_t batch, gvbd_schedule* schedule) {
    return guard(ctx, [&] {
        require(ctx != nullptr, "ctx is NULL");
        get_schedule(batch_plan(ctx, batch), schedule);
    });
}

int gvbd_batch_queue_info(gvbd_context* ctx, int32_t batch, int32_t* entries, int32_t* grid) {
    return guard(ctx, [&] {
        require(ctx != nullptr, "ctx is NULL");
        PlanObj& p = batch_plan(ctx, batch);
        build_plan(ctx, p);
        if (entries) *entries = p.nQueue;
        if (grid) *grid = p.queueGrid;
    });
}

int gvbd_batch_apply_initial_step(gvbd_context* ctx, int32_t batch) {
    return guard(ctx, [&] {
        require(ctx != nullptr, "ctx is NULL");
        PlanObj& p = batch_plan(ctx, batch);
        build_plan(ctx, p);
        plan_initial_step(ctx, p);
    });
}

int gvbd_batch_run_iteration(gvbd_context* ctx, int32_t batch, int32_t iteration) {
    return guard(ctx, [&] {
        require(ctx != nullptr, "ctx is NULL");
        run_single_iteration(ctx, batch_plan(ctx, batch), iteration);
    });
}

int gvbd_batch_update_velocities(gvbd_context* ctx, int32_t batch) {
    return guard(ctx, [&] {
        require(ctx != nullptr, "ctx is NULL");
        PlanObj& p = batch_plan(ctx, batch);
        build_plan(ctx, p);
        plan_velocity(ctx, p);
    });
}

int gvbd_batch_record_graph(gvbd_context* ctx, int32_t batch, int32_t enable) {
    return guard(ctx, [&] {
        require(ctx != nullptr, "ctx is NULL");
        PlanObj& p = batch_plan(ctx, batch);
        p.wantGraph = enable != 0;
        free_plan_graph(p);
        if (p.wantGraph) {
            build_plan(ctx, p);
            plan_capture(ctx, p);
        }
    });
}

int gvbd_batch_run_steps(gvbd_context* ctx, int32_t batch, int32_t n) {
    return guard(ctx, [&] {
        require(ctx != nullptr && n >= 0, "bad argument");
        plan_run(ctx, batch_plan(ctx, batch), n);
    });
}

int gvbd_batch_solve_parallel_group(gvbd_context* ctx, int32_t batch, int32_t group, int32_t iteration) {
    return guard(ctx, [&] {
        require(ctx != nullptr, "ctx is NULL");
        PlanObj& p = batch_plan(ctx, batch);
        build_plan(ctx, p);
        require(group >= 0 && group < p.numGroups, "group index out of range");
        // as the single-scene unit: contact scenes read same-group partners from the snapshot, and
        // cloth slots solved through positionsNext are committed after the group
        if (p.contact) plan_contact_snapshot(ctx, p);
        plan_group(ctx, p, group, iteration);
    });
}

int gvbd_batch_num_parallel_groups(gvbd_context* ctx, int32_t batch, int32_t* n) {
    return guard(ctx, [&] {
        require(ctx && n, "NULL argument");
        PlanObj& p = batch_plan(ctx, batch);
        build_plan(ctx, p);
        *n = p.numGroups;
    });
}

int gvbd_batch_group_work(gvbd_context* ctx, int32_t batch, int32_t group, int64_t* bytes, int64_t* vertices) {
    return guard(ctx, [&] {
        require(ctx != nullptr, "ctx is NULL");
        PlanObj& p = batch_plan(ctx, batch);
        build_plan(ctx, p);
        require(group >= 0 && group < p.numGroups, "group index out of range");
        int64_t b = 0, nvx = 0;
        for (size_t si = 0; si < p.scenes.size(); ++si) {
            SceneObj& s = scene_of(ctx, p.scenes[si]);
            const int g = p.groupAt[si * p.numGroups + group];  // the scene's group in this launch
            if (g < 0) continue;
            const size_t nt = s.numTetSlots();
            for (int32_t k = s.tetGroupOffset[g]; k < s.tetGroupOffset[g + 1]; ++k, ++nvx) b += s.slotWorkBytes[k];
            for (int32_t k = s.clothGroupOffset[g]; k < s.clothGroupOffset[g + 1]; ++k, ++nvx)
                b += s.slotWorkBytes[nt + k];
        }
        if (bytes) *bytes = b;
        if (vertices) *vertices = nvx;
    });
}

// ---- serving loop ----

// Frees a server's device resources (its frames must be complete).
static void release_server(ServerObj& S) {
    for (PlanObj& P : S.plans) release_plan(P);
    for (hipEvent_t& e : S.frameDone)
        if (e) (void)hipEventDestroy(e), e = nullptr;
    if (S.ref) (void)hipEventDestroy(S.ref), S.ref = nullptr;
}

static ServerObj& server_of(gvbd_context* ctx, int32_t id) {
    require(ctx != nullptr, "ctx is NULL");
    auto it = ctx->servers.find(id);
    if (it == ctx->servers.end()) fail(GVBD_ERR_INVALID, "unknown server id " + std::to_string(id));
    return *it->second;
}

// Moves a scene's device pointers from one arena base to another. The arena's arrays hold
// indices, never addresses, so a copied arena is valid at any base.
static void rebase_scene(SceneObj& s, uintptr_t from, uintptr_t to) {
    auto mv = [&](auto*& p) {
        using P = std::remove_reference_t<decltype(p)>;
        if (p) p = reinterpret_cast<P>(to + (reinterpret_cast<uintptr_t>(p) - from));
    };
    DeviceScene& d = s.dev;
    mv(d.pos), mv(d.prev), mv(d.inertia), mv(d.vel), mv(d.vparam), mv(d.pnext), mv(d.vflag), mv(d.slots);
    mv(d.tinc), mv(d.tvol), mv(d.tmat), mv(d.cslots), mv(d.cface), mv(d.cbend), mv(d.faceRef), mv(d.edgeRef);
    mv(d.friction);
    mv(d.tFaceRef), mv(d.tEdgeRef), mv(d.tcontact), mv(d.accPrevPrev), mv(d.accPrevIter), mv(d.accOmega);
    mv(d.snap), mv(d.vgroup);
    mv(d.clock), mv(d.vzone), mv(d.dverts), mv(d.dparams), mv(d.attBegin), mv(d.att);
    SceneView& v = s.view;
    mv(v.pos), mv(v.prev), mv(v.inertia), mv(v.vel), mv(v.vparam), mv(v.vflag), mv(v.slots), mv(v.tinc);
    mv(v.tvol), mv(v.tmat);
    mv(v.pnext), mv(v.cslots), mv(v.cface), mv(v.cbend), mv(v.tcontact), mv(v.accPrevPrev), mv(v.accPrevIter);
    mv(v.accOmega), mv(v.clock), mv(v.vzone), mv(v.dverts), mv(v.dparams), mv(v.attBegin), mv(v.att);
    mv(v.solveFail);
    if (s.cd) {
        // contact arrays in the arena; the detector's own arrays stay where they are
        ContactView& c = v.cv;
        mv(c.friction), mv(c.faceRef), mv(c.edgeRef), mv(c.tFaceRef), mv(c.tEdgeRef), mv(c.snap), mv(c.vgroup);
        s.cd->rebindPositions(d.pos);
    }
}

// Copies an idle scene's arena out to host memory and frees it.
static void server_evict(gvbd_context* ctx, SceneObj& s) {
    hip_check(hipStreamSynchronize(ctx->stream), "sync");
    s.hostCopy.resize(s.dev.bytes);
    hip_check(hipMemcpy(s.hostCopy.data(), s.dev.arena, s.dev.bytes, hipMemcpyDeviceToHost), "copy out");
    s.hostCopyBase = reinterpret_cast<uintptr_t>(s.dev.arena);
    hip_check(hipFree(s.dev.arena), "hipFree");
    s.dev.arena = nullptr;
    if (s.plan >= 0) free_plan(plan_of(ctx, s.plan));  // its device views hold the old addresses
}

// Copies a scene held in host memory back into a new arena.
static void server_restore(gvbd_context* ctx, SceneObj& s) {
    void* arena = nullptr;
    if (hipMalloc(&arena, s.dev.bytes) != hipSuccess) {
        (void)hipGetLastError();
        fail(GVBD_ERR_NOMEM, "hipMalloc of " + std::to_string(s.dev.bytes) + " bytes for a served scene failed");
    }
    // synchronous on the null stream: the frame in flight on the library's (non-blocking) stream
    // keeps running meanwhile
    if (hipMemcpy(arena, s.hostCopy.data(), s.dev.bytes, hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipFree(arena);
        fail(GVBD_ERR_HIP, "copy back of a served scene failed");
    }
    rebase_scene(s, s.hostCopyBase, reinterpret_cast<uintptr_t>(arena));
    s.dev.arena = arena;
    s.hostCopy.clear();  // the page-locked buffer stays for the next copy-out
    s.hostCopyBase = 0;
    if (s.plan >= 0) free_plan(plan_of(ctx, s.plan));
}

static int64_t server_resident_bytes(gvbd_context* ctx, const ServerObj& S, int32_t* count) {
    int64_t b = 0;
    int32_t n = 0;
    for (auto& kv : S.scenes) {
        const SceneObj& s = scene_of(ctx, kv.first);
        if (s.dev.arena) b += (int64_t)s.dev.bytes, ++n;
        if (s.cd) b += (int64_t)s.cd->deviceBytes();  // detection trees and lists stay resident
    }
    if (count) *count = n;
    return b;
}

// Makes room for `need` more device bytes by copying idle resident scenes (no pending request)
// out, least recently stepped first. False, with nothing copied out, when even every idle
// scene would not make room.
static bool server_make_room(gvbd_context* ctx, ServerObj& S, int64_t need, int32_t self) {
    int64_t resident = server_resident_bytes(ctx, S, nullptr);
    if (resident + need <= S.budget) return true;
    std::vector<std::pair<int64_t, int32_t>> idle;
    int64_t freeable = 0;
    for (auto& kv : S.scenes) {
        const SceneObj& s = scene_of(ctx, kv.first);
        if (kv.first == self || !s.dev.arena || !kv.second.pending.empty()) continue;
        idle.push_back({kv.second.lastUse, kv.first});
        freeable += (int64_t)s.dev.bytes;
    }
    if (resident - freeable + need > S.budget) return false;
    std::sort(idle.begin(), idle.end());
    for (auto& c : idle) {
        if (resident + need <= S.budget) break;
        SceneObj& s = scene_of(ctx, c.second);
        resident -= (int64_t)s.dev.bytes;
        server_evict(ctx, s);
        ++S.stats.evictions;
        ++S.layout;
    }
    return true;
}

struct ServeBudgetWait {};

// Makes a registered scene resident (first upload, or copy back). False when it does not fit
// yet; a scene larger than the whole budget fails with GVBD_ERR_NOMEM.
static bool server_admit(gvbd_context* ctx, ServerObj& S, int32_t sid, ServerObj::Entry& e) {
    SceneObj& s = scene_of(ctx, sid);
    auto too_large = [&](int64_t need) {
        if (need > S.budget)
            fail(GVBD_ERR_NOMEM, "scene " + std::to_string(sid) + " needs " + std::to_string(need) +
                                     " device bytes, more than the server's budget");
    };
    if (s.initialized) {
        if (!server_make_room(ctx, S, (int64_t)s.dev.bytes, sid)) return false;
        server_restore(ctx, s);
    } else {
        if (e.needBytes > 0 && !server_make_room(ctx, S, e.needBytes, sid)) return false;
        try {
            initialize_scene(ctx, s, [&](size_t total) {
                e.needBytes = (int64_t)total;
                too_large(e.needBytes);
                if (!server_make_room(ctx, S, (int64_t)total, sid)) throw ServeBudgetWait{};
            });
        } catch (const ServeBudgetWait&) {
            return false;
        }
        s.plan = new_plan(ctx, {sid});
    }
    ++S.stats.admissions;
    ++S.layout;
    return true;
}

static int64_t free_vertices(const SceneObj& s) {
    return (int64_t)s.tetGroupOffset.back() + (int64_t)s.clothGroupOffset.back();
}

// Completes the oldest frame in flight: one request of each of its scenes.
static void server_complete_oldest(gvbd_context* ctx, ServerObj& S) {
    using Clock = ServerObj::Clock;
    ServerObj::Frame f = std::move(S.inflight.front());
    S.inflight.pop_front();
    hip_check(hipEventSynchronize(S.frameDone[f.slot]), "frame wait");
    check_queue_status(ctx);
    float ms = 0.f;
    hip_check(hipEventElapsedTime(&ms, S.ref, S.frameDone[f.slot]), "frame time");
    const Clock::time_point finished =
        S.refTime + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double, std::milli>(ms));
    for (int32_t sid : f.scenes) {
        ServerObj::Entry& e = S.scenes[sid];
        const ServerObj::Request r = e.pending.front();
        e.pending.pop_front();
        --e.launched;
        e.lastUse = S.stats.ticks;
        ++e.frames;
        gvbd_served_step d{};
        d.tag = r.tag;
        d.scene = sid;
        d.frame = e.frames;
        d.latency_ms = std::chrono::duration<double, std::milli>(finished - r.at).count();
        d.queue_ms = std::chrono::duration<double, std::milli>(f.launched - r.at).count();
        S.done.push_back(d);
        ++S.stats.served;
        S.stats.vertex_solves += free_vertices(scene_of(ctx, sid)) * S.iterations * S.substeps;
    }
}

static void server_drain(gvbd_context* ctx, ServerObj& S) {
    while (!S.inflight.empty()) server_complete_oldest(ctx, S);
}

static void server_tick(gvbd_context* ctx, ServerObj& S) {
    using Clock = ServerObj::Clock;
    // one frame of every resident scene with a request not launched yet; a scene whose previous
    // frame is still in flight is stepped again behind it (one stream orders the two)
    std::vector<int32_t> ready;
    for (auto& kv : S.scenes)
        if (kv.second.pending.size() > (size_t)kv.second.launched && scene_of(ctx, kv.first).dev.arena)
            ready.push_back(kv.first);
    if (!ready.empty()) {
        const int slot = S.nextSlot;
        while (!S.inflight.empty() && S.inflight.front().slot == slot) server_complete_oldest(ctx, S);
        PlanObj& P = S.plans[slot];
        if (ready != P.scenes || S.layout != S.planLayout[slot]) {
            free_plan(P);
            P.scenes = ready;
            S.planLayout[slot] = S.layout;
        }
        build_plan(ctx, P);
        const Clock::time_point launched = Clock::now();
        plan_frame_launches(ctx, P);
        hip_check(hipEventRecord(S.frameDone[slot], ctx->stream), "frame event");
        for (int32_t sid : ready) ++S.scenes[sid].launched;
        S.inflight.push_back({ready, slot, launched});
        S.nextSlot = slot ^ 1;
    }
    // admissions while that frame runs, oldest request first and without overtaking: a scene
    // that does not fit holds back the ones behind it. A failed admission is reported after the
    // bookkeeping below; a scene that can never fit has its requests completed as rejected
    // (frame -1, latency -1).
    std::vector<std::pair<Clock::time_point, int32_t>> waiting;
    for (auto& kv : S.scenes)
        if (!kv.second.pending.empty() && !scene_of(ctx, kv.first).dev.arena)
            waiting.push_back({kv.second.pending.front().at, kv.first});
    std::sort(waiting.begin(), waiting.end());
    Failure admitError{GVBD_OK, ""};
    for (auto& w : waiting) {
        ServerObj::Entry& e = S.scenes[w.second];
        try {
            if (!server_admit(ctx, S, w.second, e)) break;
        } catch (const Failure& f) {
            admitError = f;
            if (f.code == GVBD_ERR_NOMEM && e.needBytes > S.budget) {
                for (const ServerObj::Request& r : e.pending) S.done.push_back({r.tag, w.second, -1, -1.0, -1.0});
                e.pending.clear();
                continue;
            }
            break;
        }
    }
    // keep at most the frame just launched in flight; with nothing launched, finish everything
    if (!ready.empty()) {
        while (S.inflight.size() > 1) server_complete_oldest(ctx, S);
    } else {
        server_drain(ctx, S);
    }
    S.stats.last_stepped = (int32_t)ready.size();
    ++S.stats.ticks;
    if (admitError.code != GVBD_OK) throw admitError;
}

static void server_fill_stats(gvbd_context* ctx, const ServerObj& S, gvbd_server_stats* out) {
    gvbd_server_stats st = S.stats;
    st.budget_bytes = S.budget;
    st.resident_bytes = server_resident_bytes(ctx, S, &st.resident);
    st.pending = 0;
    st.waiting = 0;
    for (auto& kv : S.scenes) {
        st.pending += (int32_t)kv.second.pending.size();
        if (!kv.second.pending.empty() && !scene_of(ctx, kv.first).dev.arena) ++st.waiting;
    }
    *out = st;
}

int gvbd_server_create(gvbd_context* ctx, int64_t budget_bytes, int32_t* server_out) {
    return guard(ctx, [&] {
        require(ctx && server_out, "NULL argument");
        require(budget_bytes > 0, "budget_bytes must be positive");
        size_t freeBytes = 0, totalBytes = 0;
        hip_check(hipSetDevice(ctx->device), "hipSetDevice");
        hip_check(hipMemGetInfo(&freeBytes, &totalBytes), "hipMemGetInfo");
        auto S = std::make_unique<ServerObj>();
        S->budget = std::min<int64_t>(budget_bytes, (int64_t)totalBytes);
        for (hipEvent_t& e : S->frameDone) hip_check(hipEventCreate(&e), "hipEventCreate");
        hip_check(hipEventCreate(&S->ref), "hipEventCreate");
        hip_check(hipEventRecord(S->ref, ctx->stream), "hipEventRecord");
        hip_check(hipEventSynchronize(S->ref), "hipEventSynchronize");
        S->refTime = ServerObj::Clock::now();
        const int32_t id = ctx->nextServer++;
        ctx->servers[id] = std::move(S);
        *server_out = id;
    });
}

int gvbd_server_destroy(gvbd_context* ctx, int32_t server) {
    return guard(ctx, [&] {
        ServerObj& S = server_of(ctx, server);
        server_drain(ctx, S);
        for (auto& kv : S.scenes) {
            SceneObj& s = scene_of(ctx, kv.first);
            if (s.initialized && !s.dev.arena) server_restore(ctx, s);
        }
        for (auto& kv : S.scenes) scene_of(ctx, kv.first).server = -1;
        release_server(S);
        ctx->servers.erase(server);
    });
}

int gvbd_server_set_schedule(gvbd_context* ctx, int32_t server, const gvbd_schedule* schedule) {
    return guard(ctx, [&] {
        require(schedule != nullptr, "schedule is NULL");
        ServerObj& S = server_of(ctx, server);
        server_drain(ctx, S);
        for (PlanObj& P : S.plans) {
            free_plan(P);
            set_schedule(P, *schedule);
        }
    });
}

int gvbd_server_add_scene(gvbd_context* ctx, int32_t server, int32_t scene) {
    return guard(ctx, [&] {
        ServerObj& S = server_of(ctx, server);
        SceneObj& s = scene_of(ctx, scene);
        if (s.server >= 0) fail(GVBD_ERR_STATE, "scene is already registered with a server");
        require(!s.meshes.empty(), "scene has no meshes");
        for (auto& kv : ctx->plans)
            if (kv.first != s.plan)
                for (int32_t sid : kv.second->scenes)
                    if (sid == scene) fail(GVBD_ERR_STATE, "scene is part of a live batch; destroy the batch first");
        if (S.iterations < 0) {
            S.iterations = s.params.iterations;
            S.substeps = s.params.num_substeps;
            for (PlanObj& P : S.plans) {
                P.iterations = S.iterations;
                P.substeps = S.substeps;
            }
        } else if (s.params.iterations != S.iterations || s.params.num_substeps != S.substeps) {
            fail(GVBD_ERR_INVALID, "scenes of one server must share iterations and substeps");
        }
        ServerObj::Entry e;
        if (s.initialized) {
            e.needBytes = (int64_t)s.dev.bytes;
            if (!server_make_room(ctx, S, e.needBytes, scene))
                fail(GVBD_ERR_NOMEM, "the resident scene does not fit the server's budget");
        }
        S.scenes[scene] = e;
        s.server = server;
        ++S.layout;
    });
}

int gvbd_server_remove_scene(gvbd_context* ctx, int32_t server, int32_t scene) {
    return guard(ctx, [&] {
        ServerObj& S = server_of(ctx, server);
        auto it = S.scenes.find(scene);
        if (it == S.scenes.end()) fail(GVBD_ERR_INVALID, "scene is not registered with this server");
        if (!it->second.pending.empty()) fail(GVBD_ERR_STATE, "scene has pending step requests");
        SceneObj& s = scene_of(ctx, scene);
        if (s.initialized && !s.dev.arena) server_restore(ctx, s);
        s.server = -1;
        S.scenes.erase(it);
        for (PlanObj& P : S.plans) {
            free_plan(P);
            P.scenes.clear();
        }
        ++S.layout;
    });
}

int gvbd_server_request_step(gvbd_context* ctx, int32_t server, int32_t scene, int64_t tag) {
    return guard(ctx, [&] {
        ServerObj& S = server_of(ctx, server);
        auto it = S.scenes.find(scene);
        if (it == S.scenes.end()) fail(GVBD_ERR_INVALID, "scene is not registered with this server");
        if (it->second.needBytes > S.budget)
            fail(GVBD_ERR_NOMEM, "scene needs more device bytes than the server's budget");
        it->second.pending.push_back({tag, ServerObj::Clock::now()});
    });
}

int gvbd_server_tick(gvbd_context* ctx, int32_t server, gvbd_server_stats* stats) {
    return guard(ctx, [&] {
        ServerObj& S = server_of(ctx, server);
        server_tick(ctx, S);
        if (stats) server_fill_stats(ctx, S, stats);
    });
}

int gvbd_server_poll(gvbd_context* ctx, int32_t server, gvbd_served_step* out, int32_t capacity, int32_t* count) {
    return guard(ctx, [&] {
        require(count != nullptr && capacity >= 0 && (out || capacity == 0), "bad argument");
        ServerObj& S = server_of(ctx, server);
        int32_t n = 0;
        while (n < capacity && !S.done.empty()) {
            out[n++] = S.done.front();
            S.done.pop_front();
        }
        *count = n;
    });
}

int gvbd_server_get_stats(gvbd_context* ctx, int32_t server, gvbd_server_stats* stats) {
    return guard(ctx, [&] {
        require(stats != nullptr, "stats is NULL");
        server_fill_stats(ctx, server_of(ctx, server), stats);
    });
}

int gvbd_server_scene_resident(gvbd_context* ctx, int32_t server, int32_t scene, int32_t* resident) {
    return guard(ctx, [&] {
        require(resident != nullptr, "resident is NULL");
        ServerObj& S = server_of(ctx, server);
        if (!S.scenes.count(scene)) fail(GVBD_ERR_INVALID, "scene is not registered with this server");
        *resident = scene_of(ctx, scene).dev.arena ? 1 : 0;
    });
}

// ---- contact detector ----

static ContactDetector& detector_of(gvbd_context* ctx, int32_t cd) {
    require(ctx != nullptr, "ctx is NULL");
    auto it = ctx->detectors.find(cd);
    if (it == ctx->detectors.end()) fail(GVBD_ERR_INVALID, "unknown detector id " + std::to_string(cd));
    return *it->second;
}

int gvbd_cd_create(gvbd_context* ctx, float max_query_distance, int32_t vf_capacity, int32_t ee_capacity,
                   int32_t* cd_out) {
    return guard(ctx, [&] {
        require(ctx != nullptr && cd_out != nullptr, "NULL argument");
        hip_check(hipSetDevice(ctx->device), "hipSetDevice");
        auto d = std::make_unique<ContactDetector>(ctx->device, ctx->stream, max_query_distance, vf_capacity,
                                                   ee_capacity);
        const int32_t id = ctx->nextDetector++;
        ctx->detectors[id] = std::move(d);
        *cd_out = id;
    });
}

int gvbd_cd_destroy(gvbd_context* ctx, int32_t cd) {
    return guard(ctx, [&] {
        detector_of(ctx, cd);
        hip_check(hipStreamSynchronize(ctx->stream), "sync");
        ctx->detectors.erase(cd);
    });
}

int gvbd_cd_add_mesh(gvbd_context* ctx, int32_t cd, int32_t num_vertices, const float* positions, int32_t num_faces,
                     const int32_t* faces, int32_t* mesh_out) {
    return guard(ctx, [&] {
        const int32_t m = detector_of(ctx, cd).addMesh(num_vertices, positions, num_faces, faces);
        if (mesh_out) *mesh_out = m;
    });
}

int gvbd_cd_finalize(gvbd_context* ctx, int32_t cd) {
    return guard(ctx, [&] { detector_of(ctx, cd).finalize(); });
}

int gvbd_cd_sizes(gvbd_context* ctx, int32_t cd, int32_t* num_vertices, int32_t* num_faces, int32_t* num_edges) {
    return guard(ctx, [&] {
        ContactDetector& d = detector_of(ctx, cd);
        if (num_vertices) *num_vertices = d.numVertices();
        if (num_faces) *num_faces = d.numFaces();
        if (num_edges) *num_edges = d.numEdges();
    });
}

int gvbd_cd_mesh_offsets(gvbd_context* ctx, int32_t cd, int32_t mesh, int32_t* vertex, int32_t* face, int32_t* edge) {
    return guard(ctx, [&] { detector_of(ctx, cd).meshOffsets(mesh, vertex, face, edge); });
}

int gvbd_cd_topology(gvbd_context* ctx, int32_t cd, int32_t* edges, int32_t* face_edges) {
    return guard(ctx, [&] { detector_of(ctx, cd).topology(edges, face_edges); });
}

int gvbd_cd_set_positions(gvbd_context* ctx, int32_t cd, int32_t mesh, const float* positions) {
    return guard(ctx, [&] {
        require(positions != nullptr, "positions is NULL");
        detector_of(ctx, cd).setPositions(mesh, positions);
    });
}

int gvbd_cd_get_positions(gvbd_context* ctx, int32_t cd, int32_t mesh, float* positions) {
    return guard(ctx, [&] {
        require(positions != nullptr, "positions is NULL");
        detector_of(ctx, cd).getPositions(mesh, positions);
    });
}

int gvbd_cd_perturb(gvbd_context* ctx, int32_t cd, uint64_t seed, int32_t step, float sigma, const float* drift) {
    return guard(ctx, [&] { detector_of(ctx, cd).perturb(seed, step, sigma, drift); });
}

int gvbd_cd_update_bvh(gvbd_context* ctx, int32_t cd, int32_t rebuild) {
    return guard(ctx, [&] { detector_of(ctx, cd).updateBvh(rebuild != 0); });
}

int gvbd_cd_query(gvbd_context* ctx, int32_t cd, int32_t record) {
    return guard(ctx, [&] { detector_of(ctx, cd).query(record != 0); });
}

int gvbd_cd_bounds(gvbd_context* ctx, int32_t cd, float thickness, float relaxation, float* bounds) {
    return guard(ctx, [&] {
        ContactDetector& d = detector_of(ctx, cd);
        d.bounds(thickness, relaxation);
        if (bounds) d.boundsOut(bounds);
    });
}

int gvbd_cd_counts(gvbd_context* ctx, int32_t cd, int64_t* counts) {
    return guard(ctx, [&] {
        require(counts != nullptr, "counts is NULL");
        const DetectorResultCounts c = detector_of(ctx, cd).counts();
        counts[0] = c.vfPairs;
        counts[1] = c.vfContacts;
        counts[2] = c.eeContacts;
        counts[3] = c.vfOverflow;
        counts[4] = c.eeOverflow;
    });
}

int gvbd_cd_vf_pairs(gvbd_context* ctx, int32_t cd, int32_t* pairs, float* d) {
    return guard(ctx, [&] { detector_of(ctx, cd).vfPairs(pairs, d); });
}

int gvbd_cd_vf_contacts(gvbd_context* ctx, int32_t cd, int32_t* records, float* d) {
    return guard(ctx, [&] { detector_of(ctx, cd).vfContactsAll(records, d); });
}

int gvbd_cd_ee_contacts(gvbd_context* ctx, int32_t cd, int32_t* pairs, float* mu, float* d) {
    return guard(ctx, [&] { detector_of(ctx, cd).eeContactsAll(pairs, mu, d); });
}

static_assert(sizeof(gvbd_vf_contact) == sizeof(VFContactRec), "gvbd_vf_contact layout");
static_assert(sizeof(gvbd_server_stats) == 72 && sizeof(gvbd_served_step) == 32, "server record layout");
static_assert(sizeof(gvbd_deformer_desc) == 96 && offsetof(gvbd_deformer_desc, angular_velocity) == 48 &&
                  offsetof(gvbd_deformer_desc, meshes) == 72,
              "gvbd_deformer_desc layout");
static_assert(sizeof(gvbd_attachment_desc) == 40 && offsetof(gvbd_attachment_desc, bary) == 24,
              "gvbd_attachment_desc layout");
static_assert(sizeof(gvbd_ee_contact) == sizeof(EEContactRec), "gvbd_ee_contact layout");

int gvbd_cd_vf_results(gvbd_context* ctx, int32_t cd, int32_t* count, gvbd_vf_contact* records) {
    return guard(ctx, [&] { detector_of(ctx, cd).vfContactsCapped(count, (VFContactRec*)records); });
}

int gvbd_cd_ee_results(gvbd_context* ctx, int32_t cd, int32_t* count, gvbd_ee_contact* records) {
    return guard(ctx, [&] { detector_of(ctx, cd).eeContactsCapped(count, (EEContactRec*)records); });
}

int gvbd_cd_min_distances(gvbd_context* ctx, int32_t cd, float* vertex, float* face, float* edge) {
    return guard(ctx, [&] { detector_of(ctx, cd).minDistances(vertex, face, edge); });
}

int gvbd_stream_copy(gvbd_context* ctx, int64_t bytes, int32_t reps, float* avg_ms) {
    return gvbd_copy_bandwidth(ctx, bytes, reps, 0, avg_ms);
}

int gvbd_copy_bandwidth(gvbd_context* ctx, int64_t bytes, int32_t reps, int32_t method, float* avg_ms) {
    return guard(ctx, [&] {
        require(ctx && avg_ms && bytes > 0 && bytes % 16 == 0 && reps > 0 && (method == 0 || method == 1),
                "bad argument");
        void *a = nullptr, *b = nullptr;
        if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
            (void)hipGetLastError();
            if (a) (void)hipFree(a);
            fail(GVBD_ERR_NOMEM, "hipMalloc for stream copy failed");
        }
        try {
            hip_check(hipMemsetAsync(a, 0, bytes, ctx->stream), "memset");
            const int64_t n = bytes / 16;
            auto copy = [&] {
                if (method == 0)
                    hip_check(launch_stream_copy((const float4*)a, (float4*)b, n, ctx->stream), "copy launch");
                else
                    hip_check(hipMemcpyAsync(b, a, bytes, hipMemcpyDeviceToDevice, ctx->stream), "copy");
            };
            copy();
            hip_check(hipEventRecord(ctx->events[62], ctx->stream), "event");
            for (int32_t r = 0; r < reps; ++r) copy();
            hip_check(hipEventRecord(ctx->events[63], ctx->stream), "event");
            hip_check(hipEventSynchronize(ctx->events[63]), "event sync");
            float ms = 0.f;
            hip_check(hipEventElapsedTime(&ms, ctx->events[62], ctx->events[63]), "elapsed");
            *avg_ms = ms / (float)reps;
        } catch (...) {
            (void)hipFree(a);
            (void)hipFree(b);
            throw;
        }
        hip_check(hipFree(a), "hipFree");
        hip_check(hipFree(b), "hipFree");
    });
}

int gvbd_event_record(gvbd_context* ctx, int32_t slot) {
    return guard(ctx, [&] {
        require(ctx != nullptr && slot >= 0 && slot < 64, "bad argument");
        hip_check(hipEventRecord(ctx->events[slot], ctx->stream), "hipEventRecord");
    });
}

int gvbd_event_elapsed_ms(gvbd_context* ctx, int32_t slot_a, int32_t slot_b, float* ms) {
    return guard(ctx, [&] {
        require(ctx && ms && slot_a >= 0 && slot_a < 64 && slot_b >= 0 && slot_b < 64, "bad argument");
        hip_check(hipEventSynchronize(ctx->events[slot_b]), "hipEventSynchronize");
        hip_check(hipEventElapsedTime(ms, ctx->events[slot_a], ctx->events[slot_b]), "hipEventElapsedTime");
    });
}

int gvbd_tet_mesh_constants(int32_t num_vertices, const float* positions, int32_t num_tets, const int32_t* tets,
                            float density, float* dminv, float* volume, float* mass) {
    return guard(nullptr, [&] {
        require(num_vertices > 0 && num_tets > 0 && positions && tets, "bad argument");
        TetMeshHost m;
        m.nv = num_vertices;
        m.nt = num_tets;
        m.pos.assign(positions, positions + 3 * (size_t)num_vertices);
        m.tets.assign(tets, tets + 4 * (size_t)num_tets);
        m.mat.density = density;
        std::string err;
        if (!prepare_tet_mesh(m, err)) fail(GVBD_ERR_INVALID, err);
        if (dminv) std::memcpy(dminv, m.dminv.data(), m.dminv.size() * sizeof(float));
        if (volume) std::memcpy(volume, m.vol.data(), m.vol.size() * sizeof(float));
        if (mass) std::memcpy(mass, m.mass.data(), m.mass.size() * sizeof(float));
    });
}

int gvbd_cloth_mesh_constants(int32_t num_vertices, const float* positions, int32_t num_faces, const int32_t* faces,
                              float density, float* dminv, float* area, float* mass, int32_t* num_edges,
                              int32_t* edges, float* Q) {
    return guard(nullptr, [&] {
        require(num_vertices > 0 && num_faces > 0 && positions && faces, "bad argument");
        ClothMeshHost m;
        m.nv = num_vertices;
        m.nf = num_faces;
        m.pos.assign(positions, positions + 3 * (size_t)num_vertices);
        m.faces.assign(faces, faces + 3 * (size_t)num_faces);
        m.mat.density = density;
        std::string err;
        if (!prepare_cloth_mesh(m, err)) fail(GVBD_ERR_INVALID, err);
        if (dminv) std::memcpy(dminv, m.dminv.data(), m.dminv.size() * sizeof(float));
        if (area) std::memcpy(area, m.area.data(), m.area.size() * sizeof(float));
        if (mass) std::memcpy(mass, m.mass.data(), m.mass.size() * sizeof(float));
        if (num_edges) *num_edges = (int32_t)m.edges.size();
        if (edges)
            for (size_t i = 0; i < m.edges.size(); ++i) {
                const TriEdge& e = m.edges[i];
                const int32_t row[6] = {e.v1, e.v2, e.v12Next, e.v21Next, e.f1, e.f2};
                std::memcpy(edges + 6 * i, row, sizeof(row));
            }
        if (Q) std::memcpy(Q, m.Q.data(), m.Q.size() * sizeof(float));
    });
}

int gvbd_color_graph(int32_t num_vertices, int32_t num_edges, const int32_t* edges, int32_t algorithm,
                     int32_t* colors, int32_t* num_colors_out) {
    return guard(nullptr, [&] {
        require(num_vertices >= 0 && num_edges >= 0 && colors, "bad argument");
        require(num_edges == 0 || edges, "edges is NULL");
        for (int32_t i = 0; i < 2 * num_edges; ++i)
            require(edges[i] >= 0 && edges[i] < num_vertices, "edge index out of range");
        Graph g = graph_from_edges(num_vertices, edges, num_edges);
        std::vector<int32_t> c = algorithm == GVBD_COLOR_GREEDY ? color_greedy(g) : color_mcs(g);
        std::memcpy(colors, c.data(), c.size() * sizeof(int32_t));
        if (num_colors_out) *num_colors_out = num_colors(c);
    });
}

int gvbd_color_tet_mesh(int32_t num_vertices, int32_t num_tets, const int32_t* tets, int32_t algorithm,
                        int32_t* colors, int32_t* num_colors_out) {
    return guard(nullptr, [&] {
        require(num_vertices >= 0 && num_tets >= 0 && colors && (num_tets == 0 || tets), "bad argument");
        for (int32_t i = 0; i < 4 * num_tets; ++i)
            require(tets[i] >= 0 && tets[i] < num_vertices, "tet index out of range");
        Graph g = tet_vertex_graph(num_vertices, num_tets, tets);
        std::vector<int32_t> c = algorithm == GVBD_COLOR_GREEDY ? color_greedy(g) : color_mcs(g);
        std::memcpy(colors, c.data(), c.size() * sizeof(int32_t));
        if (num_colors_out) *num_colors_out = num_colors(c);
    });
}

int gvbd_color_tri_mesh(int32_t num_vertices, int32_t num_faces, const int32_t* faces, int32_t with_bending,
                        int32_t algorithm, int32_t* colors, int32_t* num_colors_out) {
    return guard(nullptr, [&] {
        require(num_vertices >= 0 && num_faces >= 0 && colors && (num_faces == 0 || faces), "bad argument");
        for (int32_t i = 0; i < 3 * num_faces; ++i)
            require(faces[i] >= 0 && faces[i] < num_vertices, "face index out of range");
        Graph g = tri_vertex_graph(num_vertices, num_faces, faces, with_bending != 0);
        std::vector<int32_t> c = algorithm == GVBD_COLOR_GREEDY ? color_greedy(g) : color_mcs(g);
        std::memcpy(colors, c.data(), c.size() * sizeof(int32_t));
        if (num_colors_out) *num_colors_out = num_colors(c);
    });
}

