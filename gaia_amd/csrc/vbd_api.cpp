// C-ABI implementation: contexts, scenes, batched plans and HIP graph capture.
// Declarations and reference citations: include/gaia_vbd.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/gaia_vbd.h"
#include "device_types.h"
#include "kernels.h"
#include "mesh_prep.h"
#include "scene_loader.h"

namespace gvbd {

namespace {

thread_local std::string g_lastError;

struct Failure {
    int code;
    std::string msg;
};

[[noreturn]] void fail(int code, const std::string& msg) { throw Failure{code, msg}; }

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) fail(GVBD_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

constexpr int kVertexChunk = 256;

// All device arrays of one scene live in one allocation.
struct DeviceScene {
    void* arena = nullptr;
    size_t bytes = 0;
    float4 *pos = nullptr, *prev = nullptr, *inertia = nullptr, *vel = nullptr, *vparam = nullptr;
    uint8_t* vflag = nullptr;
    int4* slots = nullptr;
    TetInc* tinc = nullptr;
};

struct SceneObj {
    gvbd_physics_params params{};
    SolveParams sp{};
    std::vector<TetMeshHost> meshes;
    std::vector<int32_t> vertexOffset;  // per mesh, into the scene's vertex space
    int32_t nv = 0;
    bool initialized = false;
    std::vector<std::vector<int32_t>> groups;  // interleaved (mesh, vertex)
    std::vector<int32_t> groupSlotOffset;      // G + 1
    std::vector<int32_t> slotIncCount;         // incidence count per slot (work model)
    DeviceScene dev;
    SceneView view{};
    int32_t plan = -1;  // the scene's own single-scene plan
};

struct PlanObj {
    std::vector<int32_t> scenes;
    int lpv = 4, block = 256;
    int numGroups = 0, iterations = 0, substeps = 0;
    SceneView* dViews = nullptr;
    WorkItem* dItems = nullptr;
    WorkItem* dVItems = nullptr;
    std::vector<int32_t> groupItemBegin;  // G + 1
    int32_t nVItems = 0;
    size_t bytes = 0;
    bool wantGraph = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool dirty = true;
};

}  // namespace
}  // namespace gvbd

struct gvbd_context {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t events[64] = {};
    std::string lastError;
    std::map<int32_t, std::unique_ptr<gvbd::SceneObj>> scenes;
    std::map<int32_t, std::unique_ptr<gvbd::PlanObj>> plans;
    std::map<int32_t, int32_t> batches;  // batch id -> plan id
    int32_t nextScene = 0, nextPlan = 0, nextBatch = 0;
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

SolveParams make_solve_params(const gvbd_physics_params& p, const std::vector<TetMeshHost>& meshes) {
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
    sp.initRatioG = meshes.empty() ? 0.5f : meshes[0].mat.initRatioG;
    return sp;
}

void free_plan_graph(PlanObj& p) {
    if (p.exec) hipGraphExecDestroy(p.exec);
    if (p.graph) hipGraphDestroy(p.graph);
    p.exec = nullptr;
    p.graph = nullptr;
}

void free_plan(PlanObj& p) {
    free_plan_graph(p);
    if (p.dViews) hipFree(p.dViews);
    if (p.dItems) hipFree(p.dItems);
    if (p.dVItems) hipFree(p.dVItems);
    p.dViews = nullptr;
    p.dItems = nullptr;
    p.dVItems = nullptr;
    p.bytes = 0;
}

// Builds the device arrays of an initialised scene (colour-major slots and incidence records).
void upload_scene(gvbd_context* ctx, SceneObj& s) {
    const int32_t G = (int32_t)s.groups.size();
    // slots and incidence in colour-major order; fixed vertices are not solved (VBDPhysics.cpp:1617)
    std::vector<int4> slots;
    std::vector<TetInc> tinc;
    s.groupSlotOffset.assign(G + 1, 0);
    s.slotIncCount.clear();
    for (int32_t g = 0; g < G; ++g) {
        const auto& grp = s.groups[g];
        for (size_t i = 0; i + 1 < grp.size(); i += 2) {
            const TetMeshHost& m = s.meshes[grp[i]];
            const int32_t lv = grp[i + 1];
            bool fixed = false;
            for (int32_t f : m.fixed) fixed = fixed || f == lv;
            if (fixed) continue;
            const int32_t off = s.vertexOffset[grp[i]];
            int4 slot;
            slot.x = off + lv;
            slot.y = (int32_t)tinc.size();
            const bool damping = m.mat.dampingHydrostatic != 0.f || m.mat.dampingDeviatoric != 0.f;
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
                rec.Amu = m.vol[t] * m.mat.mu;
                rec.Alm = m.vol[t] * m.mat.lambda;
                rec.a = 1.f + m.mat.mu / m.mat.lambda;
                rec.pad = 0.f;
                tinc.push_back(rec);
            }
            slot.z = (int32_t)tinc.size();
            slot.w = damping ? SLOT_DAMPING : 0;
            slots.push_back(slot);
            s.slotIncCount.push_back(slot.z - slot.y);
        }
        s.groupSlotOffset[g + 1] = (int32_t)slots.size();
    }
    // per-vertex arrays
    const int32_t nv = s.nv;
    std::vector<float4> pos(nv), prev(nv), inertia(nv), vel(nv), vparam(nv);
    std::vector<uint8_t> vflag(nv, 0);
    for (size_t im = 0; im < s.meshes.size(); ++im) {
        const TetMeshHost& m = s.meshes[im];
        const int32_t off = s.vertexOffset[im];
        const float expD = std::pow(m.mat.exponentialVelDamping, s.sp.dt);
        const float constD = m.mat.constantVelDamping * s.sp.dt;
        for (int32_t i = 0; i < m.nv; ++i) {
            const float4 p = make_float4(m.pos[3 * i], m.pos[3 * i + 1], m.pos[3 * i + 2], 0.f);
            pos[off + i] = p;
            prev[off + i] = p;
            inertia[off + i] = make_float4(p.x, p.y, p.z, m.mass[i]);
            vel[off + i] = make_float4(m.vel[3 * i], m.vel[3 * i + 1], m.vel[3 * i + 2], 0.f);
            vparam[off + i] = make_float4(expD, constD, m.mat.dampingHydrostatic, m.mat.dampingDeviatoric);
            vflag[off + i] = (uint8_t)((m.mat.initializationType & 7) << 1);
        }
        for (int32_t f : m.fixed) vflag[off + f] |= 1;
    }
    // one arena
    auto align = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t bV = align(sizeof(float4) * (size_t)std::max(nv, 1));
    const size_t bF = align((size_t)std::max(nv, 1));
    const size_t bS = align(sizeof(int4) * std::max<size_t>(slots.size(), 1));
    const size_t bI = align(sizeof(TetInc) * std::max<size_t>(tinc.size(), 1));
    const size_t total = 5 * bV + bF + bS + bI;
    DeviceScene& d = s.dev;
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    if (hipMalloc(&d.arena, total) != hipSuccess) {
        (void)hipGetLastError();
        d.arena = nullptr;
        fail(GVBD_ERR_NOMEM, "hipMalloc of " + std::to_string(total) + " bytes for scene failed");
    }
    d.bytes = total;
    char* base = (char*)d.arena;
    d.pos = (float4*)base;
    d.prev = (float4*)(base + bV);
    d.inertia = (float4*)(base + 2 * bV);
    d.vel = (float4*)(base + 3 * bV);
    d.vparam = (float4*)(base + 4 * bV);
    d.vflag = (uint8_t*)(base + 5 * bV);
    d.slots = (int4*)(base + 5 * bV + bF);
    d.tinc = (TetInc*)(base + 5 * bV + bF + bS);
    auto up = [&](void* dst, const void* src, size_t n) {
        if (n) hip_check(hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, ctx->stream), "upload");
    };
    up(d.pos, pos.data(), sizeof(float4) * nv);
    up(d.prev, prev.data(), sizeof(float4) * nv);
    up(d.inertia, inertia.data(), sizeof(float4) * nv);
    up(d.vel, vel.data(), sizeof(float4) * nv);
    up(d.vparam, vparam.data(), sizeof(float4) * nv);
    up(d.vflag, vflag.data(), nv);
    up(d.slots, slots.data(), sizeof(int4) * slots.size());
    up(d.tinc, tinc.data(), sizeof(TetInc) * tinc.size());
    hip_check(hipStreamSynchronize(ctx->stream), "upload sync");
    s.view.pos = d.pos;
    s.view.prev = d.prev;
    s.view.inertia = d.inertia;
    s.view.vel = d.vel;
    s.view.vparam = d.vparam;
    s.view.vflag = d.vflag;
    s.view.slots = d.slots;
    s.view.tinc = d.tinc;
    s.view.nv = nv;
    s.view.sp = s.sp;
}

// (Re)builds the work items of a plan for its current schedule.
void build_plan(gvbd_context* ctx, PlanObj& p) {
    if (!p.dirty) return;
    free_plan(p);
    const int vpb = p.block / p.lpv;
    std::vector<SceneView> views;
    p.numGroups = 0;
    for (int32_t sid : p.scenes) {
        SceneObj& s = scene_of(ctx, sid);
        views.push_back(s.view);
        p.numGroups = std::max(p.numGroups, (int)s.groups.size());
    }
    std::vector<WorkItem> items, vitems;
    p.groupItemBegin.assign(p.numGroups + 1, 0);
    for (int g = 0; g < p.numGroups; ++g) {
        for (size_t si = 0; si < p.scenes.size(); ++si) {
            SceneObj& s = scene_of(ctx, p.scenes[si]);
            if (g >= (int)s.groups.size()) continue;
            const int32_t b = s.groupSlotOffset[g], e = s.groupSlotOffset[g + 1];
            for (int32_t k = b; k < e; k += vpb) items.push_back({(int32_t)si, k, std::min(vpb, e - k), 0});
        }
        p.groupItemBegin[g + 1] = (int32_t)items.size();
    }
    for (size_t si = 0; si < p.scenes.size(); ++si) {
        SceneObj& s = scene_of(ctx, p.scenes[si]);
        for (int32_t k = 0; k < s.nv; k += kVertexChunk)
            vitems.push_back({(int32_t)si, k, std::min(kVertexChunk, s.nv - k), 0});
    }
    p.nVItems = (int32_t)vitems.size();
    auto alloc_up = [&](void** dst, const void* src, size_t n) {
        if (hipMalloc(dst, std::max<size_t>(n, 16)) != hipSuccess) {
            (void)hipGetLastError();
            *dst = nullptr;
            fail(GVBD_ERR_NOMEM, "hipMalloc for plan failed");
        }
        p.bytes += std::max<size_t>(n, 16);
        if (n) hip_check(hipMemcpyAsync(*dst, src, n, hipMemcpyHostToDevice, ctx->stream), "plan upload");
    };
    alloc_up((void**)&p.dViews, views.data(), sizeof(SceneView) * views.size());
    alloc_up((void**)&p.dItems, items.data(), sizeof(WorkItem) * items.size());
    alloc_up((void**)&p.dVItems, vitems.data(), sizeof(WorkItem) * vitems.size());
    hip_check(hipStreamSynchronize(ctx->stream), "plan upload sync");
    p.dirty = false;
}

void plan_initial_step(gvbd_context* ctx, PlanObj& p) {
    hip_check(launch_tet_initial_step(p.dViews, p.dVItems, p.nVItems, ctx->stream), "initial step launch");
}

void plan_group(gvbd_context* ctx, PlanObj& p, int g, int iter) {
    const int b = p.groupItemBegin[g], e = p.groupItemBegin[g + 1];
    hip_check(launch_tet_group_solve(p.dViews, p.dItems + b, e - b, iter, p.lpv, p.block, ctx->stream),
              "group solve launch");
}

void plan_velocity(gvbd_context* ctx, PlanObj& p) {
    hip_check(launch_tet_update_velocity(p.dViews, p.dVItems, p.nVItems, ctx->stream), "velocity launch");
}

// One frame: substeps x (initial step, iterations x groups, velocity update)
// (VBDPhysics.cpp:839-969 structure, all on the device).
void plan_frame_launches(gvbd_context* ctx, PlanObj& p) {
    for (int sub = 0; sub < p.substeps; ++sub) {
        plan_initial_step(ctx, p);
        for (int it = 0; it < p.iterations; ++it)
            for (int g = 0; g < p.numGroups; ++g) plan_group(ctx, p, g, it);
        plan_velocity(ctx, p);
    }
}

void plan_capture(gvbd_context* ctx, PlanObj& p) {
    free_plan_graph(p);
    hip_check(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal), "begin capture");
    try {
        plan_frame_launches(ctx, p);
    } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(ctx->stream, &g);
        if (g) hipGraphDestroy(g);
        throw;
    }
    hip_check(hipStreamEndCapture(ctx->stream, &p.graph), "end capture");
    hip_check(hipGraphInstantiate(&p.exec, p.graph, nullptr, nullptr, 0), "graph instantiate");
}

void plan_run(gvbd_context* ctx, PlanObj& p, int n) {
    build_plan(ctx, p);
    if (p.wantGraph && !p.exec) plan_capture(ctx, p);
    for (int i = 0; i < n; ++i) {
        if (p.exec)
            hip_check(hipGraphLaunch(p.exec, ctx->stream), "graph launch");
        else
            plan_frame_launches(ctx, p);
    }
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
        if (s.params.iterations != p->iterations || s.params.num_substeps != p->substeps)
            fail(GVBD_ERR_INVALID, "scenes of one batch must share iterations and substeps");
    }
    const int32_t id = ctx->nextPlan++;
    ctx->plans[id] = std::move(p);
    return id;
}

void initialize_scene(gvbd_context* ctx, SceneObj& s) {
    if (s.initialized) fail(GVBD_ERR_STATE, "scene already initialised");
    if (s.meshes.empty()) fail(GVBD_ERR_STATE, "scene has no meshes");
    s.sp = make_solve_params(s.params, s.meshes);
    std::vector<const std::vector<std::vector<int32_t>>*> cats;
    for (auto& m : s.meshes) cats.push_back(&m.categories);
    s.groups = merge_parallel_groups(cats, true);
    upload_scene(ctx, s);
    s.initialized = true;
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

void require(bool cond, const char* msg) {
    if (!cond) fail(GVBD_ERR_INVALID, msg);
}

}  // namespace

// Used by the JSON scene loader.
int32_t add_tet_mesh_host(gvbd_context* ctx, int32_t scene, TetMeshHost&& m) {
    SceneObj& s = scene_of(ctx, scene);
    if (s.initialized) fail(GVBD_ERR_STATE, "cannot add meshes after initialisation");
    std::string err;
    if (!prepare_tet_mesh(m, err)) fail(GVBD_ERR_INVALID, err);
    if (m.colors.empty()) {
        fail(GVBD_ERR_INTERNAL, "mesh colours missing");
    }
    m.categories = sorted_categories(m.colors);
    s.vertexOffset.push_back(s.nv);
    s.nv += m.nv;
    s.meshes.push_back(std::move(m));
    return (int32_t)s.meshes.size() - 1;
}

}  // namespace gvbd

using namespace gvbd;

extern "C" {

int gvbd_abi_struct_sizes(int32_t* physics_params_bytes, int32_t* tet_mesh_desc_bytes) {
    if (physics_params_bytes) *physics_params_bytes = (int32_t)sizeof(gvbd_physics_params);
    if (tet_mesh_desc_bytes) *tet_mesh_desc_bytes = (int32_t)sizeof(gvbd_tet_mesh_desc);
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
        hip_check(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking), "hipStreamCreate");
        for (auto& e : ctx->events) hip_check(hipEventCreate(&e), "hipEventCreate");
        *out = ctx.release();
    });
}

int gvbd_context_destroy(gvbd_context* ctx) {
    if (!ctx) return GVBD_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& kv : ctx->plans) free_plan(*kv.second);
    for (auto& kv : ctx->scenes)
        if (kv.second->dev.arena) (void)hipFree(kv.second->dev.arena);
    for (auto& e : ctx->events)
        if (e) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return GVBD_OK;
}

const char* gvbd_last_error(const gvbd_context* ctx) {
    return ctx ? ctx->lastError.c_str() : g_lastError.c_str();
}

int gvbd_synchronize(gvbd_context* ctx) {
    return guard(ctx, [&] {
        require(ctx != nullptr, "ctx is NULL");
        hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    });
}

int gvbd_context_memory(gvbd_context* ctx, int64_t* bytes) {
    return guard(ctx, [&] {
        require(ctx && bytes, "NULL argument");
        int64_t b = 0;
        for (auto& kv : ctx->scenes) b += (int64_t)kv.second->dev.bytes;
        for (auto& kv : ctx->plans) b += (int64_t)kv.second->bytes;
        *bytes = b;
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

int gvbd_scene_destroy(gvbd_context* ctx, int32_t scene) {
    return guard(ctx, [&] {
        require(ctx != nullptr, "ctx is NULL");
        SceneObj& s = scene_of(ctx, scene);
        for (auto& kv : ctx->plans)
            for (int32_t sid : kv.second->scenes)
                if (sid == scene && kv.first != s.plan)
                    fail(GVBD_ERR_STATE, "scene is part of a live batch; destroy the batch first");
        hip_check(hipStreamSynchronize(ctx->stream), "sync");
        if (s.plan >= 0) {
            free_plan(plan_of(ctx, s.plan));
            ctx->plans.erase(s.plan);
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
        m.fixed.assign(desc->fixed_points, desc->fixed_points + desc->num_fixed);
        for (int32_t t : m.tets) require(t >= 0 && t < m.nv, "tet vertex index out of range");
        if (desc->vertex_colors) {
            m.colors.assign(desc->vertex_colors, desc->vertex_colors + m.nv);
            for (int32_t c : m.colors) require(c >= 0, "negative vertex colour");
            Graph g = tet_vertex_graph(m.nv, m.nt, m.tets.data());
            require(coloring_valid(g, m.colors), "vertex colouring is not valid for this mesh");
        } else {
            Graph g = tet_vertex_graph(m.nv, m.nt, m.tets.data());
            m.colors = desc->coloring_algorithm == GVBD_COLOR_GREEDY ? color_greedy(g) : color_mcs(g);
        }
        const int32_t id = add_tet_mesh_host(ctx, scene, std::move(m));
        if (mesh_out) *mesh_out = id;
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
            for (auto& m : ls.meshes) add_tet_mesh_host(ctx, sid, std::move(m));
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

int gvbd_scene_mesh_info(gvbd_context* ctx, int32_t scene, int32_t mesh, int32_t* num_vertices, int32_t* num_tets) {
    return guard(ctx, [&] {
        require(ctx != nullptr, "ctx is NULL");
        SceneObj& s = scene_of(ctx, scene);
        require(mesh >= 0 && mesh < (int32_t)s.meshes.size(), "mesh index out of range");
        if (num_vertices) *num_vertices = s.meshes[mesh].nv;
        if (num_tets) *num_tets = s.meshes[mesh].nt;
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
        const TetMeshHost& m = s.meshes[mesh];
        if (dminv) std::memcpy(dminv, m.dminv.data(), m.dminv.size() * sizeof(float));
        if (volume) std::memcpy(volume, m.vol.data(), m.vol.size() * sizeof(float));
        if (mass) std::memcpy(mass, m.mass.data(), m.mass.size() * sizeof(float));
    });
}

int gvbd_scene_set_schedule(gvbd_context* ctx, int32_t scene, int32_t lanes_per_vertex, int32_t block) {
    return guard(ctx, [&] {
        require(ctx != nullptr, "ctx is NULL");
        require(valid_schedule(lanes_per_vertex, block), "schedule must be lpv in {1,2,4,8,16}, block in {64,128,256}");
        SceneObj& s = scene_of(ctx, scene);
        if (s.plan < 0) fail(GVBD_ERR_STATE, "scene not initialised");
        PlanObj& p = plan_of(ctx, s.plan);
        p.lpv = lanes_per_vertex;
        p.block = block;
        p.dirty = true;
        free_plan_graph(p);
    });
}

int gvbd_scene_run_step(gvbd_context* ctx, int32_t scene) { return gvbd_scene_run_steps(ctx, scene, 1); }

int gvbd_scene_run_steps(gvbd_context* ctx, int32_t scene, int32_t n) {
    return guard(ctx, [&] {
        require(ctx != nullptr && n >= 0, "bad argument");
        SceneObj& s = scene_of(ctx, scene);
        if (s.plan < 0) fail(GVBD_ERR_STATE, "scene not initialised");
        plan_run(ctx, plan_of(ctx, s.plan), n);
    });
}

int gvbd_scene_apply_initial_step(gvbd_context* ctx, int32_t scene) {
    return guard(ctx, [&] {
        require(ctx != nullptr, "ctx is NULL");
        SceneObj& s = scene_of(ctx, scene);
        if (s.plan < 0) fail(GVBD_ERR_STATE, "scene not initialised");
        PlanObj& p = plan_of(ctx, s.plan);
        build_plan(ctx, p);
        plan_initial_step(ctx, p);
    });
}

int gvbd_scene_solve_parallel_group(gvbd_context* ctx, int32_t scene, int32_t group, int32_t iteration) {
    return guard(ctx, [&] {
        require(ctx != nullptr, "ctx is NULL");
        SceneObj& s = scene_of(ctx, scene);
        if (s.plan < 0) fail(GVBD_ERR_STATE, "scene not initialised");
        PlanObj& p = plan_of(ctx, s.plan);
        build_plan(ctx, p);
        require(group >= 0 && group < p.numGroups, "group index out of range");
        plan_group(ctx, p, group, iteration);
    });
}

int gvbd_scene_update_velocities(gvbd_context* ctx, int32_t scene) {
    return guard(ctx, [&] {
        require(ctx != nullptr, "ctx is NULL");
        SceneObj& s = scene_of(ctx, scene);
        if (s.plan < 0) fail(GVBD_ERR_STATE, "scene not initialised");
        PlanObj& p = plan_of(ctx, s.plan);
        build_plan(ctx, p);
        plan_velocity(ctx, p);
    });
}

int gvbd_scene_record_graph(gvbd_context* ctx, int32_t scene, int32_t enable) {
    return guard(ctx, [&] {
        require(ctx != nullptr, "ctx is NULL");
        SceneObj& s = scene_of(ctx, scene);
        if (s.plan < 0) fail(GVBD_ERR_STATE, "scene not initialised");
        PlanObj& p = plan_of(ctx, s.plan);
        p.wantGraph = enable != 0;
        free_plan_graph(p);
        if (p.wantGraph) {
            build_plan(ctx, p);
            plan_capture(ctx, p);
        }
    });
}

int gvbd_scene_get_mesh_state(gvbd_context* ctx, int32_t scene, int32_t mesh, float* positions, float* velocities) {
    return guard(ctx, [&] {
        require(ctx != nullptr, "ctx is NULL");
        SceneObj& s = scene_of(ctx, scene);
        if (!s.initialized) fail(GVBD_ERR_STATE, "scene not initialised");
        require(mesh >= 0 && mesh < (int32_t)s.meshes.size(), "mesh index out of range");
        const int32_t off = s.vertexOffset[mesh], n = s.meshes[mesh].nv;
        std::vector<float4> tmp(n);
        hip_check(hipStreamSynchronize(ctx->stream), "sync");
        if (positions) {
            hip_check(hipMemcpy(tmp.data(), s.dev.pos + off, sizeof(float4) * n, hipMemcpyDeviceToHost), "download");
            for (int32_t i = 0; i < n; ++i) {
                positions[3 * i] = tmp[i].x;
                positions[3 * i + 1] = tmp[i].y;
                positions[3 * i + 2] = tmp[i].z;
            }
        }
        if (velocities) {
            hip_check(hipMemcpy(tmp.data(), s.dev.vel + off, sizeof(float4) * n, hipMemcpyDeviceToHost), "download");
            for (int32_t i = 0; i < n; ++i) {
                velocities[3 * i] = tmp[i].x;
                velocities[3 * i + 1] = tmp[i].y;
                velocities[3 * i + 2] = tmp[i].z;
            }
        }
    });
}

int gvbd_scene_set_mesh_state(gvbd_context* ctx, int32_t scene, int32_t mesh, const float* positions,
                              const float* velocities) {
    return guard(ctx, [&] {
        require(ctx != nullptr, "ctx is NULL");
        SceneObj& s = scene_of(ctx, scene);
        if (!s.initialized) fail(GVBD_ERR_STATE, "scene not initialised");
        require(mesh >= 0 && mesh < (int32_t)s.meshes.size(), "mesh index out of range");
        const int32_t off = s.vertexOffset[mesh], n = s.meshes[mesh].nv;
        std::vector<float4> tmp(n);
        hip_check(hipStreamSynchronize(ctx->stream), "sync");
        if (positions) {
            for (int32_t i = 0; i < n; ++i)
                tmp[i] = make_float4(positions[3 * i], positions[3 * i + 1], positions[3 * i + 2], 0.f);
            hip_check(hipMemcpy(s.dev.pos + off, tmp.data(), sizeof(float4) * n, hipMemcpyHostToDevice), "upload");
        }
        if (velocities) {
            for (int32_t i = 0; i < n; ++i)
                tmp[i] = make_float4(velocities[3 * i], velocities[3 * i + 1], velocities[3 * i + 2], 0.f);
            hip_check(hipMemcpy(s.dev.vel + off, tmp.data(), sizeof(float4) * n, hipMemcpyHostToDevice), "upload");
        }
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
        free_plan(p);
        ctx->plans.erase(ctx->batches[batch]);
        ctx->batches.erase(batch);
    });
}

int gvbd_batch_set_schedule(gvbd_context* ctx, int32_t batch, int32_t lanes_per_vertex, int32_t block) {
    return guard(ctx, [&] {
        require(ctx != nullptr, "ctx is NULL");
        require(valid_schedule(lanes_per_vertex, block), "schedule must be lpv in {1,2,4,8,16}, block in {64,128,256}");
        PlanObj& p = batch_plan(ctx, batch);
        p.lpv = lanes_per_vertex;
        p.block = block;
        p.dirty = true;
        free_plan_graph(p);
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
        for (int32_t sid : p.scenes) {
            SceneObj& s = scene_of(ctx, sid);
            if (group >= (int32_t)s.groups.size()) continue;
            for (int32_t k = s.groupSlotOffset[group]; k < s.groupSlotOffset[group + 1]; ++k) {
                // tet vertex-solve gather model: 60 B per vertex + 96 B per incident tet
                b += 60 + 96 * (int64_t)s.slotIncCount[k];
                ++nvx;
            }
        }
        if (bytes) *bytes = b;
        if (vertices) *vertices = nvx;
    });
}

int gvbd_stream_copy(gvbd_context* ctx, int64_t bytes, int32_t reps, float* avg_ms) {
    return guard(ctx, [&] {
        require(ctx && avg_ms && bytes > 0 && bytes % 16 == 0 && reps > 0, "bad argument");
        void *a = nullptr, *b = nullptr;
        if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
            (void)hipGetLastError();
            if (a) (void)hipFree(a);
            fail(GVBD_ERR_NOMEM, "hipMalloc for stream copy failed");
        }
        try {
            hip_check(hipMemsetAsync(a, 0, bytes, ctx->stream), "memset");
            const int64_t n = bytes / 16;
            hip_check(launch_stream_copy((const float4*)a, (float4*)b, n, ctx->stream), "copy launch");
            hip_check(hipEventRecord(ctx->events[62], ctx->stream), "event");
            for (int32_t r = 0; r < reps; ++r)
                hip_check(launch_stream_copy((const float4*)a, (float4*)b, n, ctx->stream), "copy launch");
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

int gvbd_color_graph(int32_t num_vertices, int32_t num_edges, const int32_t* edges, int32_t algorithm,
                     int32_t* colors, int32_t* num_colors_out) {
    return guard(nullptr, [&] {
        require(num_vertices >= 0 && num_edges >= 0 && colors, "bad argument");
        require(num_edges == 0 || edges, "edges is NULL");
        for (int32_t i = 0; i < 2 * num_edges; ++i) require(edges[i] >= 0 && edges[i] < num_vertices, "edge index out of range");
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
        for (int32_t i = 0; i < 4 * num_tets; ++i) require(tets[i] >= 0 && tets[i] < num_vertices, "tet index out of range");
        Graph g = tet_vertex_graph(num_vertices, num_tets, tets);
        std::vector<int32_t> c = algorithm == GVBD_COLOR_GREEDY ? color_greedy(g) : color_mcs(g);
        std::memcpy(colors, c.data(), c.size() * sizeof(int32_t));
        if (num_colors_out) *num_colors_out = num_colors(c);
    });
}

}  // extern "C"
