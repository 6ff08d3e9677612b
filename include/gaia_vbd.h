/*
 * gaia_vbd.h -- C ABI of the MI355X-native Vertex Block Descent hot path.
 *
 * This is the drop-in boundary between a VBD host framework (scene loading,
 * frame loop, output) and the device hot path. Every entry point replaces a
 * reference interface, cited as Simulator/Modules/<file>:<line>:
 *
 *   gvbd_scene_load_json        VBDPhysics::loadRunningparameters + initialize   (VBD/VBDPhysics.h:50-52)
 *   gvbd_scene_add_tet_mesh     BasePhysicFramework::initialize mesh load +
 *                               TetMeshFEM::initialize                           (Framework/BasePhysicsFramework.cpp:101-202,
 *                                                                                 TetMesh/TetMeshFEM.cpp:332-443)
 *   gvbd_scene_initialize       VBDPhysics::initialize (parallel groups, GPU upload) (VBD/VBDPhysics.cpp:208-454,563-651)
 *   gvbd_scene_run_step         VBDPhysics::runStep / runStepGPU                 (VBD/VBDPhysics.cpp:809-969)
 *   gvbd_scene_apply_initial_step VBDBaseTetMesh::applyInitialStep              (VBD/VBD_BaseMaterial.h:244-259)
 *   gvbd_scene_solve_parallel_group VBDSolveParallelGroup_vertexSweep_2hierarchiesGPU +
 *                               ..._updateVertexPositionGPU                      (VBD/VBDPhysicsCompute.h:10-32)
 *   gvbd_scene_update_velocities VBDUpdateVelocityGPU / updateVelocities         (VBD/VBDPhysicsCompute.h:30,
 *                                                                                 VBD/VBDPhysics.cpp:2486-2561)
 *   gvbd_scene_record_graph     VBDPhysics::recordVBDSolveGraph                  (VBD/VBDPhysics.cpp:1546-1570)
 *   gvbd_scene_get/set_mesh_state PhysicsState positions/velocities (checkpoint) (Framework/BasePhysicsFramework.h:145-319)
 *   gvbd_color_*                GraphColoring Mcs / OrderedGreedy                (GraphColoring/mcs.cpp:14-168,
 *                                                                                 GraphColoring/gready.cpp:15-117)
 *   gvbd_scene_add_cloth_mesh   TriMeshFEM / VBDBaseTriMesh initialize           (TriMesh/TriMesh.cpp:15-111,481-753,
 *                                                                                 VBDCloth/VBDTriMeshStVK.cpp:50-88)
 *   gvbd_scene_add_deformer     loadDeformers + DeformerRotator / DeformerTranslater (VBD/VBDDeformer.h:6-222,
 *                               ClothDeformer*, VBDCloth/VBDClothDeformer.h:6-246)
 *   gvbd_scene_add_attachments  VBDClothSimulationFramework::loadConstraints       (VBDCloth/VBDClothPhysics.cpp:148-224)
 *   gvbd_scene_get/set_time     BasePhysicFramework::curTime                     (Framework/BasePhysicsFramework.h:100)
 *   gvbd_scene_collision_detection VBDClothSimulationFramework::applyCollisionDetection
 *                                                                                (VBDCloth/VBDClothPhysics.cpp:719-819)
 *   gvbd_scene_penetration_results VBDPhysics::collisionResultsAll after dcd() / ccd() (tet solver
 *                               penetration contact)                             (VBD/VBDPhysics.h:154,
 *                                                                                 VBD/VBDPhysics.cpp:1958-2116)
 *   gvbd_cd_*                   ClothContactDetector (initialize, updateBVH, contactQueryVF/EE)
 *                                                                                (CollisionDetector/ClothContactDetector.h:133-182)
 *   gvbd_batch_*                (no reference counterpart: many scenes per launch; the reference's
 *                                nearest form is the multi-mesh parallel-group merge, VBD/VBDPhysics.cpp:235-274)
 *   gvbd_server_*, gvbd_scene_clone  (no reference counterpart: serving loop over many scenes)
 *   gvbd_scene_evaluate_convergence VBDPhysics::evaluateConvergenceGPU           (VBD/VBDPhysics.cpp:3510-3556)
 *   gvbd_scene_evaluate_convergence_noalloc (no reference counterpart: allocation-free variant of
 *                                gvbd_scene_evaluate_convergence for callers that cannot allocate)
 *
 * Conventions: every function returns GVBD_OK (0) or a negative status; the
 * message is available from gvbd_last_error(ctx) (or gvbd_last_error(NULL) for
 * calls without a context). No exceptions cross the ABI. The caller owns host
 * arrays; the library owns device memory. One context owns one HIP stream;
 * calls on one context must be externally synchronised. Stepping calls are
 * asynchronous on the context's stream; gvbd_synchronize waits for them.
 * Positions/velocities cross the ABI as float xyz triples (3 * n floats).
 */
#ifndef GAIA_VBD_H
#define GAIA_VBD_H

#include <stdint.h>

/* The library is built with hidden symbol visibility; only the entry points below are exported. */
#if defined(__GNUC__) || defined(__clang__)
#define GVBD_API __attribute__((visibility("default")))
#else
#define GVBD_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

#define GVBD_OK 0
#define GVBD_ERR_INVALID (-1)
#define GVBD_ERR_HIP (-2)
#define GVBD_ERR_IO (-3)
#define GVBD_ERR_STATE (-4)
#define GVBD_ERR_NOMEM (-5)
#define GVBD_ERR_INTERNAL (-6)

#define GVBD_COLOR_MCS 0
#define GVBD_COLOR_GREEDY 1

typedef struct gvbd_context gvbd_context;

/* VBDPhysicsParameters + BasePhysicsParams fields on the hot path
 * (VBD/VBDPhysicsParameters.h:8-81, Parameters/PhysicsParameters.h:18-120).
 * dt = time_step / num_substeps (Framework/BasePhysicsFramework.cpp:245-249). */
typedef struct gvbd_physics_params {
    float gravity[3];
    float time_step;
    int32_t num_substeps;
    int32_t iterations;
    float world_lower[3]; /* worldBounds[0] */
    float world_upper[3]; /* worldBounds[1] */
    int32_t use_plane_ground;
    float boundary_collision_stiffness;
    float boundary_friction_dynamic;
    float boundary_friction_eps_v;
    float step_size;
    float step_size_gd;
    int32_t friction_start_iter;
    int32_t associate_gravity_with_inertia;
    /* cloth meshes (VBDClothPhysicsParameters, VBDCloth/VBDClothPhysicsParameters.h:15-81) */
    float cloth_boundary_contact_stiffness; /* boundaryContactStiffness (1e4; the reference never reads it from JSON) */
    float cloth_boundary_friction_eps;      /* boundaryFrictionEpsValue (1e-2) */
    float degenerate_triangle_thres;        /* degenerateTriangleThres (1e-6) */
    float cloth_step_size;                  /* stepSize (1) */
    int32_t handle_collision;               /* VF/EE contact between cloth meshes (cloth-only scenes) */
    float contact_stiffness;                /* contactStiffness (1e4) */
    float contact_radius;                   /* contactRadius (1) */
    float thickness;                        /* thickness (0) */
    float conservative_step_relaxation;     /* conservativeStepRelaxation (0.45) */
    int32_t contact_energy_type;            /* contactEnergyTypes (1) */
    float contact_query_radius;             /* ContactDetectorParams.maxQueryDis (0.4) */
    int32_t contact_bvh_rebuild_steps;      /* contactBVHReconstructionSteps (32): full BVH rebuild period in frames */
    /* Chebyshev acceleration of the tet iterations (useAccelerator, acceleratorRho 0.93,
     * VBD/VBDPhysicsParameters.h:64-65; runStepGPU, VBD/VBDPhysics.cpp:872-906) */
    int32_t use_accelerator;
    float accelerator_rho;                  /* in [0, 1) */
    /* contact model of tet meshes with handle_collision: 0 = their boundary surfaces take part in
     * the cloth contact model (VF/FV/EE repulsion, friction, conservative bounds); 1 = the reference
     * tet solver's penetration contact: DCD of surface vertices at the start of a substep and CCD of
     * their paths after the initial step (VBDPhysics::dcd / ccd, VBD/VBDPhysics.cpp:1958-2116), the
     * force k d n + friction of each collision in the vertex solve (VBDPhysicsCompute.cu:1731-1923).
     * Model 1 scenes hold tet meshes only; their frames need no host decisions. */
    int32_t tet_contact_model;
    int32_t allow_dcd;                      /* allowDCD (1) */
    int32_t allow_ccd;                      /* allowCCD (1) */
    int32_t handle_self_collision;          /* handleSelfCollision (1) */
    float collision_stiffness;              /* collisionStiffness (1e5) */
    float collision_off_height;             /* collisionOffHeight (1e7): vertices at or above it are not detected */
    int32_t check_feasible_region;          /* checkFeasibleRegion (1) */
    int32_t check_tet_traverse;             /* checkTetTraverse (1) */
    int32_t tet_traverse_non_self;          /* tetrahedralTraverseForNonSelfIntersection (1) */
    int32_t shift_query_point_to_center;    /* shiftQueryPointToCenter (1) */
    int32_t stop_traversing_after_passing;  /* stopTraversingAfterPassingQueryPoint (1) */
    float feasible_region_epsilon;          /* feasibleRegionEpsilon (1e-2) */
    float ray_tri_intersection_epsilon;     /* rayTriIntersectionEPSILON (1e-10) */
    float max_search_distance_multiplier;   /* maxSearchDistanceMultiplier (1.8) */
    float center_shift_level;               /* centerShiftLevel (0.01) */
    int32_t dcd_bvh_rebuild_steps;          /* dcdTetMeshSceneBVHRebuildSteps (16): full rebuild period in frames */
} gvbd_physics_params;

/* One tet mesh (ObjectParametersVBDNeoHookean, VBD/VBD_NeoHookean.cpp:1613-1625,
 * VBD/VBD_BaseMaterial.h:21-46, Materials/Materials.h:12-75). Positions are
 * world-space (rotation/scale/translation already applied). */
typedef struct gvbd_tet_mesh_desc {
    int32_t num_vertices;
    const float* positions;   /* 3 * num_vertices */
    const float* velocities;  /* 3 * num_vertices, or NULL for initial_velocity */
    float initial_velocity[3];
    int32_t num_tets;
    const int32_t* tets;      /* 4 * num_tets */
    float density;
    float mu;                 /* "miu" */
    float lambda;             /* "lmbd" */
    float damping_hydrostatic;
    float damping_deviatoric;
    float exponential_vel_damping; /* raw value; pow(., dt) is applied (VBD/VBDPhysics.cpp:794-807) */
    float constant_vel_damping;    /* raw value; multiplied by dt */
    int32_t initialization_type;   /* 0..4 (VBD/VBD_BaseMaterial.cpp:7-50) */
    float init_ratio_g;
    int32_t num_fixed;
    const int32_t* fixed_points;
    const int32_t* vertex_colors;  /* colour per vertex, or NULL to colour here */
    int32_t coloring_algorithm;    /* GVBD_COLOR_MCS / GVBD_COLOR_GREEDY when vertex_colors is NULL */
    float friction_dynamic;        /* contact friction of the boundary surface (collision handling; 0.1) */
    float friction_eps_v;          /* (0.1) */
    /* no-gravity zone (ObjectParams::hasNoGravZone / noGravZoneThreshold / maxVelocityMagnitude,
     * Materials/Materials.h:34-35; dampVelocities, VBD/VBDPhysics.cpp:2529-2555): a vertex whose
     * position y exceeds the threshold has its speed capped at max_velocity_magnitude and no
     * damping; the zone needs max_velocity_magnitude > 0 */
    int32_t has_no_grav_zone;
    float no_grav_zone_threshold;
    float max_velocity_magnitude;  /* <= 0: none (-1) */
} gvbd_tet_mesh_desc;

/* One cloth (triangle) mesh with StVK membrane + quadratic bending (VBDObjectParamsTriMeshStVk,
 * VBDCloth/VBDTriMeshStVK.cpp:1171-1180, VBDCloth/VBDBaseTriMesh.h:8-18, TriMesh/TriMesh.h:21-40).
 * The rest pose is the given positions (use3DRestpose). */
typedef struct gvbd_cloth_mesh_desc {
    int32_t num_vertices;
    const float* positions;   /* 3 * num_vertices */
    const float* velocities;  /* 3 * num_vertices, or NULL for initial_velocity */
    float initial_velocity[3];
    int32_t num_faces;
    const int32_t* faces;     /* 3 * num_faces, consistently oriented */
    float density;
    float mu;                 /* "miu" */
    float lambda;
    float bending_stiffness;
    float damping_stvk;
    float damping_bending;
    float exponential_vel_damping; /* applied per substep as given (cloth does not rescale by dt) */
    float constant_vel_damping;
    int32_t initialization_type;   /* 0 inertia, 2 previous position (default), 3 inertia within the conservative
                                      bounds, 4 none */
    int32_t num_fixed;
    const int32_t* fixed_points;
    const int32_t* vertex_colors;  /* colour per vertex, or NULL to colour the vertex graph with bending edges */
    int32_t coloring_algorithm;
    float friction_dynamic;        /* frictionDynamic (0.1): contact friction, averaged over the two meshes */
    float friction_eps_v;          /* frictionEpsV (0.1) */
    /* no-gravity zone (Materials/Materials.h:34-35; updateVelocity, VBDCloth/VBDClothPhysics.cpp:2325-2351):
     * as for tet meshes, except that the reference compares the vertex's velocity y (not its
     * position) with the threshold */
    int32_t has_no_grav_zone;
    float no_grav_zone_threshold;
    float max_velocity_magnitude;  /* <= 0: none (-1) */
} gvbd_cloth_mesh_desc;

#define GVBD_MODE_GROUP_LAUNCHES 0 /* one kernel launch per parallel group (the reference's structure) */
#define GVBD_MODE_WORK_QUEUE 1     /* one persistent launch per iteration over a (scene, colour, vertex) queue */

/* Launch schedule of the vertex solve (no reference counterpart: the reference fixes one block
 * per vertex, VBD/VBDPhysicsCompute.cu:803-968). */
typedef struct gvbd_schedule {
    int32_t lanes_per_vertex;   /* 1, 2, 4, 8, 16: lanes of one wave64 sharing a vertex's incident elements */
    int32_t block;              /* 64, 128, 256 threads */
    int32_t mode;               /* GVBD_MODE_* */
    int32_t vertices_per_entry; /* work-queue entry size, a multiple of block / lanes_per_vertex (0: 4x that) */
    int32_t blocks_per_cu;      /* persistent grid per compute unit (0: the resident maximum) */
    int32_t spin_limit;         /* polls before a dependency wait is reported as an error (0: default) */
    int32_t stage_records;      /* group launches: incidence records staged through LDS per chunk by the whole
                                   block (32 or 64, with lanes_per_vertex 4/8/16 and block 128/256), or 0 to
                                   read them straight from HBM; results are identical either way */
    int32_t xcd_blocks;         /* group launches: 1 maps the blocks one XCD runs (the hardware deals blocks to the
                                   8 XCDs round-robin) to one contiguous range of the work items, so the vertices
                                   an XCD solves at a time are neighbours whose gathers share its L2; 4, 16 or 64
                                   deal chunks of that many consecutive items to the XCDs in turn; 0 takes work
                                   item = block index. Results are identical either way */
} gvbd_schedule;

/* sizes of the ABI structs as compiled into the library (binding self-check) */
GVBD_API int gvbd_abi_struct_sizes(int32_t* physics_params_bytes, int32_t* tet_mesh_desc_bytes,
                          int32_t* cloth_mesh_desc_bytes, int32_t* schedule_bytes);

/* ---- context ---- */
GVBD_API int gvbd_context_create(int32_t device, gvbd_context** out);
GVBD_API int gvbd_context_destroy(gvbd_context* ctx);
GVBD_API const char* gvbd_last_error(const gvbd_context* ctx);
GVBD_API int gvbd_synchronize(gvbd_context* ctx);
/* device bytes currently held by scenes and plans of this context */
GVBD_API int gvbd_context_memory(gvbd_context* ctx, int64_t* bytes);
/* device bytes of one scene as a server's memory budget counts them (no reference counterpart):
 * its arena (0 while not uploaded or copied out) and its contact detector (trees and detection
 * lists, resident while the scene exists); its launch plans are not included */
GVBD_API int gvbd_scene_device_bytes(gvbd_context* ctx, int32_t scene, int64_t* arena, int64_t* detector);

/* ---- scenes ---- */
GVBD_API int gvbd_scene_create(gvbd_context* ctx, const gvbd_physics_params* params, int32_t* scene_out);
GVBD_API int gvbd_scene_destroy(gvbd_context* ctx, int32_t scene);
/* a new, not yet initialised scene with the parameters and meshes of `scene` (positions and
 * velocities as they were added): the host-side mesh preparation is copied, not redone. No
 * reference counterpart; it instances one mesh many times (served populations). */
GVBD_API int gvbd_scene_clone(gvbd_context* ctx, int32_t scene, int32_t* scene_out);
GVBD_API int gvbd_scene_add_tet_mesh(gvbd_context* ctx, int32_t scene, const gvbd_tet_mesh_desc* desc, int32_t* mesh_out);
/* VBDClothSimulationFramework mesh load + TriMeshFEM::initialize (Framework/BaseClothSimPhsicsFramework.cpp,
 * TriMesh/TriMesh.cpp:15-111) */
GVBD_API int gvbd_scene_add_cloth_mesh(gvbd_context* ctx, int32_t scene, const gvbd_cloth_mesh_desc* desc, int32_t* mesh_out);
GVBD_API int gvbd_scene_initialize(gvbd_context* ctx, int32_t scene);

/* Deformers: prescribed motion of selected vertices, which join the mesh's fixed points
 * (VBDBaseDeformer / VBDClothBaseDeformer; loadDeformers, VBD/VBDDeformer.h:181-222 and
 * VBDCloth/VBDClothDeformer.h:199-243, read from the "Deformers" array of a scene's parameter
 * file). Added before gvbd_scene_initialize; the constants of each vertex are taken from the
 * mesh's positions as added. Applied once per substep at the scene time t (float, advanced by dt
 * at the start of every substep): on tet meshes after the initial step, on cloth meshes before
 * detection and the initial guess, as the two reference frameworks do.
 *   Rotator (DeformerRotator, VBDDeformer.h:45-131; ClothDeformerRotator,
 *   VBDClothDeformer.h:44-128): p = root + R(theta) RootP, root = ((P - center) . axis) axis,
 *   RootP = P - root, theta = angular_velocity * min(t, end_time) (end_time > 0); the axis is
 *   used as given.
 *   Translator on tet meshes (DeformerTranslater, VBDDeformer.h:134-178): p = P + speed min(t,
 *   end_time) (end_time > 0); on cloth meshes (ClothDeformerTranslater, VBDClothDeformer.h:
 *   130-196): p += speed dt in every substep with start_time <= t and (t < end_time or
 *   end_time < 0). */
#define GVBD_DEFORMER_ROTATOR 0
#define GVBD_DEFORMER_TRANSLATOR 1
typedef struct gvbd_deformer_desc {
    int32_t kind;                 /* GVBD_DEFORMER_ROTATOR / GVBD_DEFORMER_TRANSLATOR */
    float center[3];              /* Rotator */
    float axis[3];                /* Rotator */
    float speed[3];               /* Translator */
    float start_time;             /* Translator on cloth meshes (deformationStartTime) */
    double angular_velocity;      /* Rotator */
    double end_time;              /* rotationEndTime / deformationEndTime; -1: none */
    int32_t num_meshes;           /* selectedMeshes (scene mesh indices) */
    const int32_t* meshes;
    const int32_t* num_vertices;  /* per selected mesh */
    const int32_t* vertices;      /* selectedVertices, concatenated */
} gvbd_deformer_desc;
GVBD_API int gvbd_scene_add_deformer(gvbd_context* ctx, int32_t scene, const gvbd_deformer_desc* desc);

/* Vertex-face attachments of cloth vertices (VertexFaceAttachmentConstraint,
 * VBDCloth/VBDTriMeshConstraints.h:31-49; loadConstraints, VBDCloth/VBDClothPhysics.cpp:148-224,
 * read from PhysicsParams.VertexFaceAttachmentConstraintParams). Each ties vertex `vertex` of
 * cloth mesh `mesh` to the point bary[0] x(face[0]) + bary[1] x(face[1]) + bary[2] x(face[2]) of
 * cloth mesh `target_mesh`, or, with target_mesh = -1, to the point `bary`. Its solve adds
 * f -= k b (x_read - target), H += k b^2 I (accumulateConstraintForceAndHessian,
 * VBDClothPhysics.cpp:1465-1498): b = 1 on the vertex; each face vertex j with bary[j] != 0 gets
 * b = -bary[j] and reads x_read from the vertex-side mesh at its own index, as the reference
 * does. Added before gvbd_scene_initialize. */
typedef struct gvbd_attachment_desc {
    int32_t mesh, vertex;  /* vertexConstraintVertexInfos pair */
    int32_t target_mesh;   /* vertexConstraintTargetFacesInfos: mesh (-1: fixed point), face */
    int32_t face[3];
    float bary[3];         /* vertexConstraintBarys */
    float stiffness;       /* vertexConstraintStiffness (1e4 when the file gives none) */
} gvbd_attachment_desc;
GVBD_API int gvbd_scene_add_attachments(gvbd_context* ctx, int32_t scene, const gvbd_attachment_desc* list, int32_t n);

/* Scene time (BasePhysicFramework::curTime, Framework/BasePhysicsFramework.h:100): 0 at
 * creation, + dt per substep; gvbd_scene_load_state sets it from the recovery state. */
GVBD_API int gvbd_scene_get_time(gvbd_context* ctx, int32_t scene, float* cur_time);
GVBD_API int gvbd_scene_set_time(gvbd_context* ctx, int32_t scene, float cur_time);
/* Models.json + Parameters.json; "${REPO_ROOT}" in paths is replaced by repo_root (may be NULL). */
GVBD_API int gvbd_scene_load_json(gvbd_context* ctx, const char* models_json, const char* params_json,
                         const char* repo_root, int32_t* scene_out);

GVBD_API int gvbd_scene_num_meshes(gvbd_context* ctx, int32_t scene, int32_t* n);
GVBD_API int gvbd_scene_mesh_info(gvbd_context* ctx, int32_t scene, int32_t mesh, int32_t* num_vertices,
                         int32_t* num_tets);
GVBD_API int gvbd_scene_num_parallel_groups(gvbd_context* ctx, int32_t scene, int32_t* n);
/* interleaved (mesh, vertex) pairs; pairs may be NULL to query the count */
GVBD_API int gvbd_scene_parallel_group(gvbd_context* ctx, int32_t scene, int32_t group, int32_t* pairs, int32_t* count);
/* rest-pose constants of one mesh: dminv 9*T (column-major), volume T, mass N (any may be NULL) */
GVBD_API int gvbd_scene_mesh_constants(gvbd_context* ctx, int32_t scene, int32_t mesh, float* dminv, float* volume,
                              float* mass);
/* lanes per vertex (1,2,4,8,16) and block size (64,128,256) of the vertex solve */
GVBD_API int gvbd_scene_set_schedule(gvbd_context* ctx, int32_t scene, int32_t lanes_per_vertex, int32_t block);
/* full schedule (mode, work-queue entry size, persistent grid); drops a recorded graph */
GVBD_API int gvbd_scene_set_schedule_ex(gvbd_context* ctx, int32_t scene, const gvbd_schedule* schedule);
GVBD_API int gvbd_scene_get_schedule(gvbd_context* ctx, int32_t scene, gvbd_schedule* schedule);

/* ---- stepping (asynchronous on the context stream) ---- */
GVBD_API int gvbd_scene_run_step(gvbd_context* ctx, int32_t scene);               /* all substeps of one frame */
GVBD_API int gvbd_scene_run_steps(gvbd_context* ctx, int32_t scene, int32_t n);
GVBD_API int gvbd_scene_apply_initial_step(gvbd_context* ctx, int32_t scene);
GVBD_API int gvbd_scene_solve_parallel_group(gvbd_context* ctx, int32_t scene, int32_t group, int32_t iteration);
/* one VBD iteration: every parallel group in order, in the scene's schedule mode
 * (the iteration body of VBDPhysics::runStepGPU / runStep, VBD/VBDPhysics.cpp:879-945, 1605-1627) */
GVBD_API int gvbd_scene_run_iteration(gvbd_context* ctx, int32_t scene, int32_t iteration);
GVBD_API int gvbd_scene_update_velocities(gvbd_context* ctx, int32_t scene);
/* cloth scenes with handle_collision: VBDClothSimulationFramework::applyCollisionDetection
 * (VBDClothPhysics.cpp:735-819) at the current positions; run_step calls it itself before
 * iteration 0 and after any iteration in which a vertex was truncated to its bound */
GVBD_API int gvbd_scene_collision_detection(gvbd_context* ctx, int32_t scene);
/* counters of a scene with handle_collision since its initialisation (no reference counterpart:
 * the reference prints timings instead, VBDClothPhysics.cpp:389-518). detect_ms is device time of
 * the detection work (BVH update, queries, bounds, face links), timed with events around it. */
typedef struct gvbd_contact_stats {
    int64_t frames;          /* frames (substeps) stepped with collision handling */
    int64_t iterations;      /* solver iterations over those frames */
    int64_t detections;      /* detection passes: one per frame plus one after each truncating iteration */
    int64_t rebuilds;        /* full BVH rebuilds (refits otherwise) */
    double detect_ms;        /* device time of the detection passes */
    int64_t vf_pairs;        /* after the last detection: vertex-face candidates within the query radius */
    int64_t ee_pairs;        /* and edge-edge candidates */
} gvbd_contact_stats;
GVBD_API int gvbd_scene_contact_stats(gvbd_context* ctx, int32_t scene, gvbd_contact_stats* out);
/* penetration contact (tet_contact_model 1), after the last detection: per scene vertex the mesh
 * and boundary face (in that mesh's boundary-face order, -1 none) it is pushed out of and flags
 * (1: a tet contained it at the substep's start, 2: the hit came from CCD); each array holds the
 * scene's vertex count (NULL skips it). counts (5, or NULL): DCD hits, contained vertices with no
 * face, CCD hits, ray walks stopped by a full candidate stack, closest-face searches that went
 * past the nearest 8 faces (all rejected) to a full traversal. Replaces reading
 * VBDPhysics::collisionResultsAll (VBD/VBDPhysics.h:154). */
GVBD_API int gvbd_scene_penetration_results(gvbd_context* ctx, int32_t scene, int32_t* hit_mesh, int32_t* hit_face,
                                            uint8_t* flags, int32_t* counts);
/* Vertex solves since initialisation (or the last reset) whose 3x3 block had no solution: tet
 * vertices then take the gradient step (VBD/VBDPhysics.cpp:2449-2466), cloth vertices keep their
 * position (VBDCloth/VBDClothPhysics.cpp:1126-1150). The reference counts none; this is a diagnostic.
 * Synchronises the context's stream; reset != 0 clears the counter. */
GVBD_API int gvbd_scene_solve_failures(gvbd_context* ctx, int32_t scene, int32_t reset, int64_t* count);
/* VBDPhysics::evaluateConvergenceGPU (VBD/VBDPhysics.cpp:3510-3556): at the current positions and
 * inertia targets, the merit energy (inertia 1/(2 dt^2) sum m |y - x|^2 + stable neo-Hookean
 * elastic energy), its two parts, and the average over the tet vertices of |material + damping +
 * inertia force| (fixed vertices count with zero force). Synchronous; any NULL output is skipped.
 * Between frames it measures how far the last frame's iterate is from the implicit-Euler
 * solution; with the per-iteration hooks it reproduces the reference's per-iteration log. */
GVBD_API int gvbd_scene_evaluate_convergence(gvbd_context* ctx, int32_t scene, double* energy, double* inertia_energy,
                                    double* elastic_energy, double* avg_force_norm);
/* No counterpart in the reference. The same evaluation, with the same results to the bit, for
 * callers that cannot allocate device memory (an evaluation inside an iteration loop that is
 * captured or timed, a serving loop): it calls no hipMalloc / hipFree. Its per-block partial sums
 * and totals live in a fixed per-GPU scratch array of the library (1 MiB, held until the process
 * ends), so a scene may have at most about 33.5 million tets + tet vertices + free tet vertices
 * (larger scenes get GVBD_ERR_INVALID). Precondition: the scene's tet energy records are resident,
 * i.e. gvbd_scene_evaluate_convergence ran once on it at set-up (GVBD_ERR_STATE otherwise).
 * Calls on one GPU share the scratch: calls that overlap in time on one GPU (several contexts or
 * threads) race and must be serialised by the caller. */
GVBD_API int gvbd_scene_evaluate_convergence_noalloc(gvbd_context* ctx, int32_t scene, double* energy,
                                                     double* inertia_energy, double* elastic_energy,
                                                     double* avg_force_norm);
/* capture one frame (substeps x (initial step, iterations x groups, velocities)) into a HIP graph,
 * replayed by run_step(s); enable = 0 drops the graph and launches kernels directly */
GVBD_API int gvbd_scene_record_graph(gvbd_context* ctx, int32_t scene, int32_t enable);

/* ---- state I/O (synchronous) ---- */
GVBD_API int gvbd_scene_get_mesh_state(gvbd_context* ctx, int32_t scene, int32_t mesh, float* positions, float* velocities);
GVBD_API int gvbd_scene_set_mesh_state(gvbd_context* ctx, int32_t scene, int32_t mesh, const float* positions,
                              const float* velocities);

/* ---- scene files (synchronous) ----
 * Recovery states: PhysicsState::toJson / fromJson + BasePhysicFramework::recoverFromState
 * (Framework/BasePhysicsFramework.h:145-319, BasePhysicsFramework.cpp:264-281; the cloth framework's
 * ClothPhysicsState, BaseClothSimPhsicsFramework.h:65-120, has the same layout). Every mesh of the
 * scene in addition order; binary = 1 writes "<stem>.pos" / "<stem>.vel" (float32 xyz) beside
 * the JSON. Loading sets the meshes the file holds (a shorter file leaves the rest as they are,
 * as the reference does); a mesh stored without velocities keeps its current ones. */
GVBD_API int gvbd_scene_save_state(gvbd_context* ctx, int32_t scene, const char* json_path, int32_t frame_id, double cur_time,
                          int32_t binary);
GVBD_API int gvbd_scene_load_state(gvbd_context* ctx, int32_t scene, const char* json_path, int32_t* frame_id,
                          double* cur_time);
/* Frame outputs: the surface of every mesh (a tet mesh's boundary vertices, ascending, and its
 * boundary faces; every vertex and face of a cloth mesh) as ASCII PLY (GAIA::saveAsPly /
 * writeAllToPLY, IO/FileIO.cpp:8-50,119-254): all meshes in one file (all_in_one = 1), or one file
 * per mesh named "<dir>/<stem>_Model_<6-digit mesh index><ext>"; and the float32 surface
 * positions of all meshes back to back (writeAllToBinary, IO/FileIO.cpp:52-117). */
GVBD_API int gvbd_scene_write_ply(gvbd_context* ctx, int32_t scene, const char* path, int32_t all_in_one);
GVBD_API int gvbd_scene_write_binary_frame(gvbd_context* ctx, int32_t scene, const char* path);
/* The same file formats from host arrays (no device): num_meshes meshes of num_vertices[m]
 * vertices, positions / velocities xyz back to back. gvbd_state_read fills num_meshes and
 * num_vertices (capacity max_meshes) and, when positions is not NULL, positions and velocities
 * (capacity max_floats floats each; a mesh stored without velocities reads as zeros). */
GVBD_API int gvbd_state_write(const char* json_path, int32_t frame_id, double cur_time, int32_t binary, int32_t num_meshes,
                     const int32_t* num_vertices, const float* positions, const float* velocities);
GVBD_API int gvbd_state_read(const char* json_path, int32_t max_meshes, int32_t* num_meshes, int32_t* num_vertices,
                    int64_t max_floats, float* positions, float* velocities, int32_t* frame_id, double* cur_time);
GVBD_API int gvbd_write_ply(const char* path, int32_t num_vertices, const float* positions, int32_t num_faces,
                   const int32_t* faces);

/* ---- batched stepping: many scenes per launch ---- */
GVBD_API int gvbd_batch_create(gvbd_context* ctx, const int32_t* scenes, int32_t n, int32_t* batch_out);
GVBD_API int gvbd_batch_destroy(gvbd_context* ctx, int32_t batch);
GVBD_API int gvbd_batch_set_schedule(gvbd_context* ctx, int32_t batch, int32_t lanes_per_vertex, int32_t block);
GVBD_API int gvbd_batch_set_schedule_ex(gvbd_context* ctx, int32_t batch, const gvbd_schedule* schedule);
GVBD_API int gvbd_batch_get_schedule(gvbd_context* ctx, int32_t batch, gvbd_schedule* schedule);
/* work-queue size of one iteration and the persistent grid (blocks) of the current schedule */
GVBD_API int gvbd_batch_queue_info(gvbd_context* ctx, int32_t batch, int32_t* entries, int32_t* grid);
GVBD_API int gvbd_batch_record_graph(gvbd_context* ctx, int32_t batch, int32_t enable);
GVBD_API int gvbd_batch_run_steps(gvbd_context* ctx, int32_t batch, int32_t n);
/* per-iteration hooks over every scene of the batch (as the gvbd_scene_ forms) */
GVBD_API int gvbd_batch_apply_initial_step(gvbd_context* ctx, int32_t batch);
GVBD_API int gvbd_batch_run_iteration(gvbd_context* ctx, int32_t batch, int32_t iteration);
GVBD_API int gvbd_batch_update_velocities(gvbd_context* ctx, int32_t batch);
/* launch only the vertex solve of launch `group` of an iteration (timing of the dominant kernel).
 * With group launches a batch has as many launches per iteration as its scene with the most
 * parallel groups; each scene's groups sit in increasing launches chosen to balance their work,
 * so launch g holds group g of a scene with that many groups and some group <= g of the others.
 * Calling the launches 0 .. n-1 in order is one iteration's solve. */
GVBD_API int gvbd_batch_solve_parallel_group(gvbd_context* ctx, int32_t batch, int32_t group, int32_t iteration);
GVBD_API int gvbd_batch_num_parallel_groups(gvbd_context* ctx, int32_t batch, int32_t* n);
/* algorithmic bytes of one solve launch `group` and its vertex count (gather model: per tet
 * vertex 72 B + 88 B per incident tet, per cloth vertex 60 B + 56 B per face + 116 B per bending
 * stencil) */
GVBD_API int gvbd_batch_group_work(gvbd_context* ctx, int32_t batch, int32_t group, int64_t* bytes, int64_t* vertices);

/* ---- serving loop: continuous batching of step requests ----
 * No reference counterpart (the reference steps one scene per process). Scenes registered with a
 * server are made resident in HBM on their first step request when their device bytes fit the
 * budget; idle resident scenes (no pending request) are copied out to host memory, least recently
 * used first, to make room, and copied back when asked for again (a scene copied out once keeps a
 * page-locked host buffer of its arena's size until it is destroyed). Each tick advances every
 * resident scene with a pending request by one frame in one batched launch sequence (the batch
 * plan's kernels), then admits waiting scenes while that frame runs, then waits for the frame and
 * completes one request per stepped scene. Results equal those of the scene stepped alone. */
typedef struct gvbd_server_stats {
    int64_t budget_bytes;
    int64_t resident_bytes;  /* device bytes of the resident scenes */
    int64_t vertex_solves;   /* free vertices x iterations x substeps, over all completed requests */
    int64_t served;          /* completed step requests */
    int64_t admissions;      /* scenes made resident (first upload or copy back) */
    int64_t evictions;       /* idle scenes copied out to host memory */
    int64_t ticks;
    int32_t resident;        /* resident scenes */
    int32_t pending;         /* step requests not completed yet */
    int32_t waiting;         /* scenes with pending requests that are not resident */
    int32_t last_stepped;    /* scenes advanced by the last tick */
} gvbd_server_stats;

typedef struct gvbd_served_step {
    int64_t tag;       /* the caller's tag of the request */
    int32_t scene;
    int32_t frame;     /* frames this scene has completed under the server, this one included */
    double latency_ms; /* request to completion: the frame's kernels have finished */
    double queue_ms;   /* request to the launch of the frame that served it */
} gvbd_served_step;

/* budget_bytes: device bytes the server's scenes may hold at once (clamped to the device's memory;
 * gvbd_server_get_stats reports the budget in force) */
GVBD_API int gvbd_server_create(gvbd_context* ctx, int64_t budget_bytes, int32_t* server_out);
/* scenes stay alive and resident or not as they are; evicted scenes are copied back first */
GVBD_API int gvbd_server_destroy(gvbd_context* ctx, int32_t server);
GVBD_API int gvbd_server_set_schedule(gvbd_context* ctx, int32_t server, const gvbd_schedule* schedule);
/* registers a created scene (meshes added; initialised or not) without collision handling; all
 * scenes of a server share iterations and substeps */
GVBD_API int gvbd_server_add_scene(gvbd_context* ctx, int32_t server, int32_t scene);
/* hands a scene back to the caller (resident again); it must have no pending request */
GVBD_API int gvbd_server_remove_scene(gvbd_context* ctx, int32_t server, int32_t scene);
/* queues one frame of `scene`; requests of one scene are served in order, one per tick */
GVBD_API int gvbd_server_request_step(gvbd_context* ctx, int32_t server, int32_t scene, int64_t tag);
/* one tick (see above); stats may be NULL */
GVBD_API int gvbd_server_tick(gvbd_context* ctx, int32_t server, gvbd_server_stats* stats);
/* drains up to `capacity` completed requests in completion order */
GVBD_API int gvbd_server_poll(gvbd_context* ctx, int32_t server, gvbd_served_step* out, int32_t capacity, int32_t* count);
GVBD_API int gvbd_server_get_stats(gvbd_context* ctx, int32_t server, gvbd_server_stats* stats);
/* whether a registered scene is resident in HBM (1) or held in host memory / not uploaded yet (0) */
GVBD_API int gvbd_server_scene_resident(gvbd_context* ctx, int32_t server, int32_t scene, int32_t* resident);

/* ---- contact detector (GPU LBVH) ----
 * Replaces struct ClothContactDetector (CollisionDetector/ClothContactDetector.h:133-182): Embree
 * scenes of faces and edges become linear BVHs on the device. Ids are global over the meshes in
 * the order they were added (gvbd_cd_mesh_offsets gives each mesh's first vertex / face / edge);
 * edges are numbered per mesh by first appearance over its faces (eV1 -> eV2 the direction in the
 * first face). Contact records (16 bytes each): */
typedef struct gvbd_vf_contact {
    int32_t face;  /* representative face: lowest face id whose closest point is this primitive */
    int32_t type;  /* primitive of the closest point: 0 vertex, 1 edge, 2 face interior */
    int32_t prim;  /* vertex, edge or face id */
    float d;       /* distance */
} gvbd_vf_contact;
typedef struct gvbd_ee_contact {
    int32_t edge;  /* the other edge */
    float mu1;     /* closest point on the query edge: lerp(eV1, eV2, mu1) */
    float mu2;     /* closest point on the other edge */
    float d;
} gvbd_ee_contact;

/* ClothContactDetector(params) with params.maxQueryDis = max_query_distance; at most
 * vf_capacity (1..16) contacts are kept per vertex and ee_capacity per edge (the reference's
 * VF_CONTACT_PREALLOCATE / EE_CONTACT_PREALLOCATE = 8; more are counted as overflow) */
GVBD_API int gvbd_cd_create(gvbd_context* ctx, float max_query_distance, int32_t vf_capacity, int32_t ee_capacity,
                   int32_t* cd_out);
GVBD_API int gvbd_cd_destroy(gvbd_context* ctx, int32_t cd);
/* ClothContactDetector::initialize(meshes): one call per mesh, then finalize (uploads, builds) */
GVBD_API int gvbd_cd_add_mesh(gvbd_context* ctx, int32_t cd, int32_t num_vertices, const float* positions, int32_t num_faces,
                     const int32_t* faces, int32_t* mesh_out);
GVBD_API int gvbd_cd_finalize(gvbd_context* ctx, int32_t cd);
GVBD_API int gvbd_cd_sizes(gvbd_context* ctx, int32_t cd, int32_t* num_vertices, int32_t* num_faces, int32_t* num_edges);
GVBD_API int gvbd_cd_mesh_offsets(gvbd_context* ctx, int32_t cd, int32_t mesh, int32_t* vertex, int32_t* face, int32_t* edge);
/* edges 4 per edge (eV1, eV2, eV12Next, eV21Next; -1 none), face_edges 3 per face (AB, BC, CA) */
GVBD_API int gvbd_cd_topology(gvbd_context* ctx, int32_t cd, int32_t* edges, int32_t* face_edges);
GVBD_API int gvbd_cd_set_positions(gvbd_context* ctx, int32_t cd, int32_t mesh, const float* positions);
GVBD_API int gvbd_cd_get_positions(gvbd_context* ctx, int32_t cd, int32_t mesh, float* positions);
/* workload generator (no reference counterpart): x = x_initial + step * drift[mesh] + sigma * N(0,1),
 * the noise drawn per (seed, step, vertex, axis); drift holds 3 floats per mesh (may be NULL) */
GVBD_API int gvbd_cd_perturb(gvbd_context* ctx, int32_t cd, uint64_t seed, int32_t step, float sigma, const float* drift);
/* ClothContactDetector::updateBVH: rebuild = 1 re-sorts and rebuilds (RTC_BUILD_QUALITY_LOW),
 * 0 refits the boxes (RTC_BUILD_QUALITY_REFIT) (ClothContactDetector.cpp:1137-1181) */
GVBD_API int gvbd_cd_update_bvh(gvbd_context* ctx, int32_t cd, int32_t rebuild);
/* contactQueryVF for every vertex and contactQueryEE for every edge, with resetFaceContactInfo
 * (VBDClothPhysics.cpp:735-757); record = 1 also keeps the uncapped lists (gvbd_cd_vf_pairs,
 * gvbd_cd_vf_contacts, gvbd_cd_ee_contacts). The trees must hold the current positions:
 * GVBD_ERR_STATE after gvbd_cd_set_positions or gvbd_cd_perturb without gvbd_cd_update_bvh */
GVBD_API int gvbd_cd_query(gvbd_context* ctx, int32_t cd, int32_t record);
/* conservative bounds of applyCollisionDetection (VBDClothPhysics.cpp:775-812); bounds may be NULL */
GVBD_API int gvbd_cd_bounds(gvbd_context* ctx, int32_t cd, float thickness, float relaxation, float* bounds);
/* counts[5]: VF candidate pairs, VF contacts, EE contacts, vertices and edges that overflowed */
GVBD_API int gvbd_cd_counts(gvbd_context* ctx, int32_t cd, int64_t* counts);
/* recorded lists (after gvbd_cd_query(record = 1)): (vertex, face) pairs with d < r; VF contacts
 * as (vertex, type, prim, face); EE contacts as (edge, other edge) with mu (2 each) */
GVBD_API int gvbd_cd_vf_pairs(gvbd_context* ctx, int32_t cd, int32_t* pairs, float* d);
GVBD_API int gvbd_cd_vf_contacts(gvbd_context* ctx, int32_t cd, int32_t* records, float* d);
GVBD_API int gvbd_cd_ee_contacts(gvbd_context* ctx, int32_t cd, int32_t* pairs, float* mu, float* d);
/* capped per-primitive results of the last query: count[nv] and vf_capacity records per vertex;
 * count[ne] and ee_capacity records per edge */
GVBD_API int gvbd_cd_vf_results(gvbd_context* ctx, int32_t cd, int32_t* count, gvbd_vf_contact* records);
GVBD_API int gvbd_cd_ee_results(gvbd_context* ctx, int32_t cd, int32_t* count, gvbd_ee_contact* records);
/* minDisToPrimitives of every vertex (VF) and edge (EE), faceMinDisToVertices of every face */
GVBD_API int gvbd_cd_min_distances(gvbd_context* ctx, int32_t cd, float* vertex, float* face, float* edge);

/* ---- timing on the context stream ---- */
/* copies `bytes` (multiple of 16) between two device buffers `reps` times with a plain
 * streaming kernel and returns the average kernel time: the measured copy bandwidth and
 * the calibration point for profiler byte counters (no reference counterpart) */
GVBD_API int gvbd_stream_copy(gvbd_context* ctx, int64_t bytes, int32_t reps, float* avg_ms);
/* the same measurement by method: 0 the library's streaming kernel (gvbd_stream_copy), 1 the
 * runtime's device-to-device hipMemcpyAsync (no reference counterpart) */
GVBD_API int gvbd_copy_bandwidth(gvbd_context* ctx, int64_t bytes, int32_t reps, int32_t method, float* avg_ms);
GVBD_API int gvbd_event_record(gvbd_context* ctx, int32_t slot);  /* slot 0..63 */
GVBD_API int gvbd_event_elapsed_ms(gvbd_context* ctx, int32_t slot_a, int32_t slot_b, float* ms);

/* ---- rest-pose precompute (host; TetMeshFEM::initialize, TetMesh/TetMeshFEM.cpp:409-429) ----
 * dminv 9*T column-major DmInv, volume T signed det(Dm)/6, mass N lumped vol*density/4 */
GVBD_API int gvbd_tet_mesh_constants(int32_t num_vertices, const float* positions, int32_t num_tets, const int32_t* tets,
                            float density, float* dminv, float* volume, float* mass);

/* ---- cloth rest pose, topology and bending forms (host; TriMesh/TriMesh.cpp:213-244, 481-753,
 * VBDCloth/VBDTriMeshStVK.cpp:50-88). dminv 4*F (column-major 2x2), area F, mass N; edges 6*E as
 * (eV1, eV2, eV12Next, eV21Next, fId1, fId2) with -1 on boundary edges; Q 16*E (row-major 4x4).
 * Pass NULL for outputs not needed; *num_edges receives E (call once with edges = Q = NULL to size). */
GVBD_API int gvbd_cloth_mesh_constants(int32_t num_vertices, const float* positions, int32_t num_faces, const int32_t* faces,
                              float density, float* dminv, float* area, float* mass, int32_t* num_edges,
                              int32_t* edges, float* Q);

/* ---- colouring (host) ---- */
/* cloth vertex graph: mesh edges plus (with_bending) the cross pair of every interior edge
 * (GraphColoring/TriMeshVertexGraph.h:20-45) */
GVBD_API int gvbd_color_tri_mesh(int32_t num_vertices, int32_t num_faces, const int32_t* faces, int32_t with_bending,
                        int32_t algorithm, int32_t* colors, int32_t* num_colors);
GVBD_API int gvbd_color_graph(int32_t num_vertices, int32_t num_edges, const int32_t* edges, int32_t algorithm,
                     int32_t* colors, int32_t* num_colors);
GVBD_API int gvbd_color_tet_mesh(int32_t num_vertices, int32_t num_tets, const int32_t* tets, int32_t algorithm,
                        int32_t* colors, int32_t* num_colors);

#ifdef __cplusplus
}
#endif
#endif /* GAIA_VBD_H */
