/*
 * CPU oracle for the Vertex Block Descent hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is a plain C++ restatement of the
 * reference's CPU (TBB) VBD path, in FP32 (parity) and FP64 (error yardstick).
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and only as the checker or as the timed CPU baseline; the product
 * path (gaia_amd) never links or calls it.
 *
 * Every function cites the reference file:line it restates. Parts that no
 * fixture, closed form or invariant pins are marked "parity unpinned".
 *
 * Flat parameter vector layout (double[OP_COUNT]) shared with the tests.
 */
#pragma once
#include <stdint.h>

enum OraclePhysicsParam {
    OP_GRAVITY_X = 0, OP_GRAVITY_Y, OP_GRAVITY_Z,
    OP_DT,                       /* already timeStep / numSubsteps */
    OP_NUM_SUBSTEPS,
    OP_ITERATIONS,
    OP_LOWER_X, OP_LOWER_Y, OP_LOWER_Z,
    OP_UPPER_X, OP_UPPER_Y, OP_UPPER_Z,
    OP_USE_PLANE_GROUND,
    OP_BOUNDARY_STIFFNESS,
    OP_BOUNDARY_FRICTION_MU,
    OP_BOUNDARY_FRICTION_EPS_V,
    OP_STEP_SIZE,
    OP_STEP_SIZE_GD,
    OP_FRICTION_START_ITER,
    OP_ASSOCIATE_GRAVITY,
    OP_INIT_TYPE_UNUSED,         /* per mesh: OM_INIT_TYPE */
    OP_INIT_RATIO_G,
    OP_CLOTH_BOUNDARY_STIFFNESS, /* VBDClothPhysicsParameters::boundaryContactStiffness */
    OP_CLOTH_BOUNDARY_FRICTION_EPS, /* boundaryFrictionEpsValue */
    OP_DEGENERATE_THRES,         /* degenerateTriangleThres */
    /* cloth contact (VBDClothPhysicsParameters.h:60-71, ClothContactDetectorParameters::maxQueryDis) */
    OP_HANDLE_COLLISION,
    OP_CONTACT_STIFFNESS,
    OP_CONTACT_RADIUS,
    OP_THICKNESS,
    OP_CONSERVATIVE_RELAXATION,
    OP_CONTACT_ENERGY_TYPE,
    OP_MAX_QUERY_DIS,
    /* Chebyshev acceleration of the tet iterations (VBDPhysicsParameters.h:64-65) */
    OP_USE_ACCELERATOR,
    OP_ACCELERATOR_RHO,
    OP_COUNT = 34
};

enum OracleClothMaterialParam {
    OMC_DENSITY = 0,
    OMC_MU, OMC_LAMBDA, OMC_BENDING,
    OMC_DAMP_STVK, OMC_DAMP_BENDING,
    OMC_EXP_VEL_DAMPING,         /* applied per substep as given (no pow(., dt) for cloth) */
    OMC_CONST_VEL_DAMPING,
    OMC_INIT_TYPE,               /* VBDObjectParamsTriMesh::initializationType: 0, 2, 3, 4 */
    OMC_FRICTION_MU,             /* TriMeshParams::frictionDynamic (contact friction, averaged per pair) */
    OMC_FRICTION_EPS_V,          /* TriMeshParams::frictionEpsV */
    OMC_NO_GRAV_ZONE,            /* ObjectParams::hasNoGravZone (Materials.h:34) */
    OMC_NO_GRAV_THRESHOLD,       /* noGravZoneThreshold, compared with the velocity's y (VBDClothPhysics.cpp:2330-2332) */
    OMC_MAX_VELOCITY,            /* maxVelocityMagnitude (Materials.h:35; <= 0: none) */
    OMC_COUNT = 16
};

enum OracleMaterialParam {
    OM_DENSITY = 0,
    OM_MU, OM_LAMBDA,
    OM_DAMP_HYDRO, OM_DAMP_DEV,
    OM_EXP_VEL_DAMPING,          /* raw JSON value; the oracle applies pow(., dt) */
    OM_CONST_VEL_DAMPING,        /* raw JSON value; the oracle multiplies by dt */
    OM_INIT_TYPE,                /* ObjectParamsVBD::initializationType */
    OM_FRICTION_MU,              /* contact friction of the boundary surface (contact scenes) */
    OM_FRICTION_EPS_V,
    OM_NO_GRAV_ZONE,             /* ObjectParams::hasNoGravZone (Materials.h:34) */
    OM_NO_GRAV_THRESHOLD,        /* noGravZoneThreshold, compared with the position's y (VBDPhysics.cpp:2534-2536) */
    OM_MAX_VELOCITY,             /* maxVelocityMagnitude (Materials.h:35; <= 0: none) */
    OM_COUNT = 13
};

/* penetration contact of tet meshes (CollisionDetertionParameters.h:43-80, VBDPhysicsParameters.h:36,70) */
enum OracleCollisionParam {
    OC_ALLOW_DCD = 0,
    OC_ALLOW_CCD,
    OC_SELF_COLLISION,           /* handleSelfCollision */
    OC_FEASIBLE_REGION,          /* checkFeasibleRegion */
    OC_TET_TRAVERSE,             /* checkTetTraverse */
    OC_TRAVERSE_NON_SELF,        /* tetrahedralTraverseForNonSelfIntersection */
    OC_SHIFT_TO_CENTER,          /* shiftQueryPointToCenter */
    OC_STOP_AFTER_PASSING,       /* stopTraversingAfterPassingQueryPoint */
    OC_STIFFNESS,                /* collisionStiffness */
    OC_OFF_HEIGHT,               /* collisionOffHeight */
    OC_FEASIBLE_EPS,             /* feasibleRegionEpsilon */
    OC_RAY_EPS,                  /* rayTriIntersectionEPSILON */
    OC_MAX_SEARCH_MULT,          /* maxSearchDistanceMultiplier */
    OC_CENTER_SHIFT,             /* centerShiftLevel */
    OC_COUNT = 14
};

#ifdef __cplusplus
extern "C" {
#endif

/* --- mesh precompute (TetMeshFEM.cpp:409-429) --- */
int oracle_tet_prepare_f64(int nv, const double* pos, int nt, const int32_t* tets, double density,
                           double* dminv, double* vol, double* mass);
int oracle_tet_prepare_f32(int nv, const float* pos, int nt, const int32_t* tets, float density,
                           float* dminv, float* vol, float* mass);

/* --- per-tet material force / Hessian for one vertex (VBD_NeoHookean.cpp:379-581) --- */
void oracle_neohookean_vertex_f64(const double* x4 /*12: tet vertex positions*/, const double* dminv /*9 col-major*/,
                                  double vol, double mu, double lmbd, int order,
                                  double* force3, double* hess9);

/* --- 3x3 solve (own closed form; CuMatrix::solve3x3_psd_stable is absent: parity unpinned) --- */
int oracle_solve3x3_f64(const double* h9, const double* b3, double* x3);
int oracle_solve3x3_f32(const float* h9, const float* b3, float* x3);

/* --- coloring (mcs.cpp:14-168, gready.cpp:15-117, TetMeshVertexGraph.h:15-41) --- */
int oracle_color_tet_mesh(int nv, int nt, const int32_t* tets, int algorithm /*0 mcs, 1 greedy*/, int32_t* colors);
int oracle_color_graph(int nv, int ne, const int32_t* edges, int algorithm, int32_t* colors);

/* --- tet VBD scene (VBDPhysics.cpp runStep_serialCollisionHandling) --- */
typedef struct OracleTetScene OracleTetScene;

/* precision: 32 or 64 */
OracleTetScene* oracle_tet_scene_create(int precision, const double* physics /*OP_COUNT*/);
void oracle_tet_scene_destroy(OracleTetScene*);
/* colors: per-vertex color id (categories = vertices per color, ascending id) */
int oracle_tet_scene_add_mesh(OracleTetScene*, int nv, const double* pos, const double* vel, int nt,
                              const int32_t* tets, const double* material /*OM_COUNT*/,
                              int nfixed, const int32_t* fixed, const int32_t* colors);
int oracle_tet_scene_finalize(OracleTetScene*);
int oracle_tet_scene_num_groups(OracleTetScene*);
/* group g as interleaved (mesh, vertex) pairs; returns pair count, writes when out != NULL */
int oracle_tet_scene_group(OracleTetScene*, int g, int32_t* out);
int oracle_tet_scene_step(OracleTetScene*, int nsteps, int nthreads);
int oracle_tet_scene_get_state(OracleTetScene*, int mesh, double* pos, double* vel);
int oracle_tet_scene_set_state(OracleTetScene*, int mesh, const double* pos, const double* vel);
// Deformers (VBD/VBDDeformer.h): kind 0 rotator, 1 translator; params o[3], axis[3],
// angularVelocity, endTime, speed[3], startTime; before finalize. Scene time (curTime).
int oracle_tet_scene_add_deformer(OracleTetScene*, int kind, const double* params, int nmeshes, const int32_t* meshes,
                                  const int32_t* counts, const int32_t* verts);
double oracle_tet_scene_time(OracleTetScene*);
/* vertex solves whose 3x3 block had no solution (gradient step taken), a diagnostic count */
long long oracle_tet_scene_solve_failures(OracleTetScene*);
void oracle_tet_scene_set_time(OracleTetScene*, double t);
/* VBDPhysics::evaluateConvergence: out[4] = {energy, inertia energy, elastic energy, avg force norm} */
int oracle_tet_scene_evaluate_convergence(OracleTetScene*, double* out);
/* penetration contact (DCD + CCD, runStepGPU semantics); before finalize */
int oracle_tet_scene_set_collision(OracleTetScene*, const double* params /*OC_COUNT*/);
/* last detection, 6 ints per surface vertex: mesh, vertex, hit mesh, hit face, penetrated, from CCD */
int oracle_tet_scene_collisions(OracleTetScene*, int32_t* out);
int oracle_tet_scene_surface(OracleTetScene*, int mesh, int32_t* faces, int32_t* owner);
/* geometry of the collision model, for pins: strict point-in-tet, the earliest moving point /
 * moving triangle hit (t, barycentrics; returns 1 on a hit), cubic roots in [0, 1] */
int oracle_point_in_tet_f32(const float* p, const float* a, const float* b, const float* c, const float* d);
int oracle_ccd_point_triangle_f64(const double* f0, const double* f1, const double* p0, const double* p1, double* t,
                                  double* bary);
int oracle_cubic_roots01_f64(const double* coef, double* roots);
/* the ray walk of the closest-point check through a tet mesh at positions x (1: target entered) */
int oracle_tet_walk_f64(int nv, const double* x, int nt, const int32_t* tets, const double* origin, const double* dir,
                        double maxDis, int startTet, int startLocal, int target);
/* one penetration contact: x4 = (vertex, face a, b, c) now, xp4 at the step start; order 3 for the
 * vertex side, 0..2 for face vertex `order`; f (3) and H (9, column-major) are overwritten */
void oracle_penetration_force_f64(const double* x4, const double* xp4, int order, double k, double mu, double epsV,
                                  double dt, double* f, double* H);

/* --- cloth: TriMesh precompute + topology + bending forms (TriMesh.cpp, VBDTriMeshStVK.cpp:50-88) ---
 * dminv 4F col-major, area F, mass N, edges 6E (eV1 eV2 eV12Next eV21Next fId1 fId2), Q 16E row-major;
 * pass NULL to skip an output; num_edges receives E */
int oracle_cloth_prepare(int precision, int nv, const double* pos, int nf, const int32_t* faces, double density,
                         double* dminv, double* area, double* mass, int32_t* num_edges, int32_t* edges, double* Q);
/* StVK membrane force/Hessian of one triangle for vertex `order` (VBDTriMeshStVK.cpp:721-911) */
void oracle_stvk_vertex_f64(const double* x3, const double* dminv4, double area, double mu, double lmbd, int order,
                            double* force3, double* hess9);

typedef struct OracleClothScene OracleClothScene;
OracleClothScene* oracle_cloth_scene_create(int precision, const double* physics /*OP_COUNT*/);
void oracle_cloth_scene_destroy(OracleClothScene*);
int oracle_cloth_scene_add_mesh(OracleClothScene*, int nv, const double* pos, const double* vel, int nf,
                                const int32_t* faces, const double* material /*OMC_COUNT*/, int nfixed,
                                const int32_t* fixed, const int32_t* colors);
/* a tet mesh whose boundary surface takes part in contact (scenes with handle_collision) */
int oracle_cloth_scene_add_tet(OracleClothScene*, int nv, const double* pos, const double* vel, int nt,
                               const int32_t* tets, const double* material /*OM_COUNT*/, int nfixed,
                               const int32_t* fixed, const int32_t* colors);
/* a mesh's contact faces (a tet mesh's boundary surface); returns the face count */
int oracle_cloth_scene_surface(OracleClothScene*, int mesh, int32_t* faces);
int oracle_cloth_scene_finalize(OracleClothScene*);
int oracle_cloth_scene_num_groups(OracleClothScene*);
int oracle_cloth_scene_group(OracleClothScene*, int g, int32_t* out);
int oracle_cloth_scene_step(OracleClothScene*, int nsteps, int nthreads);
int oracle_cloth_scene_get_state(OracleClothScene*, int mesh, double* pos, double* vel);
int oracle_cloth_scene_set_state(OracleClothScene*, int mesh, const double* pos, const double* vel);
int oracle_cloth_scene_add_deformer(OracleClothScene*, int kind, const double* params, int nmeshes,
                                    const int32_t* meshes, const int32_t* counts, const int32_t* verts);
double oracle_cloth_scene_time(OracleClothScene*);
/* vertex solves whose 3x3 block had no solution (cloth: no move; tet: gradient step) */
long long oracle_cloth_scene_solve_failures(OracleClothScene*);
void oracle_cloth_scene_set_time(OracleClothScene*, double t);
int oracle_cloth_scene_add_attachments(OracleClothScene*, int n, const int32_t* ints, const double* vals);

/* --- cloth vertex graph: mesh edges plus, with bending, the cross edge of every interior edge
 * (TriMeshVertexGraph.h:20-45) --- */
int oracle_color_tri_mesh(int nv, int nf, const int32_t* faces, int with_bending, int algorithm, int32_t* colors);

#ifdef __cplusplus
}
#endif
