// Scene JSON loading; see scene_loader.h.
#include "scene_loader.h"

#include <array>
#include <cmath>
#include <fstream>
#include <sstream>

#include "json_lite.h"

namespace gvbd {

namespace {

bool read_file(const std::string& path, std::string& out, std::string& err) {
    std::ifstream in(path, std::ios::binary);
    if (!in) {
        err = "cannot open " + path;
        return false;
    }
    std::stringstream ss;
    ss << in.rdbuf();
    out = ss.str();
    return true;
}

bool read_json(const std::string& path, json::Value& v, std::string& err) {
    std::string text;
    if (!read_file(path, text, err)) return false;
    json::Parser p(text);
    if (!p.parse(v, err)) {
        err = path + ": " + err;
        return false;
    }
    return true;
}

std::string substitute_root(std::string s, const std::string& root) {
    const std::string key = "${REPO_ROOT}";
    size_t at;
    while ((at = s.find(key)) != std::string::npos) s.replace(at, key.size(), root);
    for (char& c : s)
        if (c == '\\') c = '/';
    return s;
}

void get_num(const json::Value& o, const char* k, float& dst) {
    const json::Value* v = o.get(k);
    if (v && v->type == json::Value::Number) dst = (float)v->num;
}
void get_int(const json::Value& o, const char* k, int32_t& dst) {
    const json::Value* v = o.get(k);
    if (v && v->type == json::Value::Number) dst = (int32_t)v->num;
    if (v && v->type == json::Value::Bool) dst = v->b ? 1 : 0;
}
void get_vec3(const json::Value& o, const char* k, float* dst) {
    const json::Value* v = o.get(k);
    if (v && v->type == json::Value::Array && v->arr.size() == 3)
        for (int d = 0; d < 3; ++d) dst[d] = (float)v->arr[d].num;
}
std::string get_str(const json::Value& o, const char* k) {
    const json::Value* v = o.get(k);
    return v && v->type == json::Value::String ? v->str : std::string();
}

// Rotation (axis-angle vector), per-axis scale, translation, in that order.
void transform_positions(const json::Value& mj, std::vector<float>& pos, float rotEps) {
    float rot[3] = {0, 0, 0}, scale[3] = {1, 1, 1}, trans[3] = {0, 0, 0};
    get_vec3(mj, "rotation", rot);
    get_vec3(mj, "scale", scale);
    get_vec3(mj, "translation", trans);
    const float angle = std::sqrt(rot[0] * rot[0] + rot[1] * rot[1] + rot[2] * rot[2]);
    float R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};  // row-major
    if (angle > rotEps) {
        const float ax = rot[0] / angle, ay = rot[1] / angle, az = rot[2] / angle;
        const float c = std::cos(angle), s = std::sin(angle), t = 1.f - c;
        const float Rr[9] = {t * ax * ax + c,      t * ax * ay - s * az, t * ax * az + s * ay,
                             t * ax * ay + s * az, t * ay * ay + c,      t * ay * az - s * ax,
                             t * ax * az - s * ay, t * ay * az + s * ax, t * az * az + c};
        for (int i = 0; i < 9; ++i) R[i] = Rr[i];
    }
    for (size_t i = 0; i + 2 < pos.size(); i += 3) {
        float* x = &pos[i];
        const float p0 = x[0], p1 = x[1], p2 = x[2];
        for (int r = 0; r < 3; ++r) x[r] = R[3 * r] * p0 + R[3 * r + 1] * p1 + R[3 * r + 2] * p2;
        for (int r = 0; r < 3; ++r) x[r] = x[r] * scale[r] + trans[r];
    }
}

bool read_colors(const std::string& path, int32_t nv, std::vector<int32_t>& colors, std::string& err) {
    json::Value cats;
    if (!read_json(path, cats, err)) return false;
    colors.assign(nv, -1);
    for (size_t c = 0; c < cats.arr.size(); ++c)
        for (const json::Value& v : cats.arr[c].arr) {
            const int32_t id = (int32_t)v.num;
            if (id < 0 || id >= nv) {
                err = path + ": vertex id out of range";
                return false;
            }
            colors[id] = (int32_t)c;
        }
    for (int32_t c : colors)
        if (c < 0) {
            err = path + ": colouring does not cover every vertex";
            return false;
        }
    return true;
}

}  // namespace

gvbd_physics_params default_physics_params() {
    gvbd_physics_params p{};
    p.gravity[0] = 0.f;
    p.gravity[1] = -10.f;
    p.gravity[2] = 0.f;
    p.time_step = 0.01666666666666666666f;
    p.num_substeps = 20;
    p.iterations = 1;
    for (int d = 0; d < 3; ++d) {
        p.world_lower[d] = -1e7f;
        p.world_upper[d] = 1e7f;
    }
    p.use_plane_ground = 1;
    p.boundary_collision_stiffness = 1e5f;
    p.boundary_friction_dynamic = 0.1f;
    p.boundary_friction_eps_v = 1.0f;
    p.step_size = 1.0f;
    p.step_size_gd = 1e-3f;
    p.friction_start_iter = 0;
    p.associate_gravity_with_inertia = 1;
    p.cloth_boundary_contact_stiffness = 1e4f;
    p.cloth_boundary_friction_eps = 1e-2f;
    p.degenerate_triangle_thres = 1e-6f;
    p.cloth_step_size = 1.0f;
    p.handle_collision = 1;
    p.contact_stiffness = 1e4f;
    p.contact_radius = 1.0f;
    p.thickness = 0.f;
    p.conservative_step_relaxation = 0.45f;
    p.contact_energy_type = 1;
    p.contact_query_radius = 0.4f;  // ClothContactDetectorParameters() sets maxQueryDis = 0.4
    p.contact_bvh_rebuild_steps = 32;
    p.use_accelerator = 0;
    p.accelerator_rho = 0.93f;
    // penetration contact of tet meshes (CollisionDetertionParameters.h:43-80, VBDPhysicsParameters.h:36,70,
    // PhysicsParameters.h:81)
    p.tet_contact_model = 0;
    p.allow_dcd = 1;
    p.allow_ccd = 1;
    p.handle_self_collision = 1;
    p.collision_stiffness = 1e5f;
    p.collision_off_height = 1e7f;
    p.check_feasible_region = 1;
    p.check_tet_traverse = 1;
    p.tet_traverse_non_self = 1;
    p.shift_query_point_to_center = 1;
    p.stop_traversing_after_passing = 1;
    p.feasible_region_epsilon = 1e-2f;
    p.ray_tri_intersection_epsilon = 1e-10f;
    p.max_search_distance_multiplier = 1.8f;
    p.center_shift_level = 0.01f;
    p.dcd_bvh_rebuild_steps = 16;
    return p;
}

bool load_scene_json(const std::string& modelsPath, const std::string& paramsPath, const std::string& repoRoot,
                     LoadedScene& out, std::string& err) {
    json::Value models, params;
    if (!read_json(modelsPath, models, err) || !read_json(paramsPath, params, err)) return false;
    out.params = default_physics_params();
    gvbd_physics_params& p = out.params;
    bool handleCollisionGiven = false;
    if (const json::Value* pp = params.get("PhysicsParams")) {
        get_vec3(*pp, "gravity", p.gravity);
        get_num(*pp, "timeStep", p.time_step);
        get_int(*pp, "numSubsteps", p.num_substeps);
        get_int(*pp, "iterations", p.iterations);
        if (const json::Value* wb = pp->get("worldBounds")) {
            if (wb->type == json::Value::Array && wb->arr.size() == 2 && wb->arr[0].arr.size() == 3 &&
                wb->arr[1].arr.size() == 3) {
                for (int d = 0; d < 3; ++d) {
                    p.world_lower[d] = (float)wb->arr[0].arr[d].num;
                    p.world_upper[d] = (float)wb->arr[1].arr[d].num;
                }
            }
        }
        get_int(*pp, "usePlaneGround", p.use_plane_ground);
        get_num(*pp, "boundaryCollisionStiffness", p.boundary_collision_stiffness);
        get_num(*pp, "boundaryFrictionDynamic", p.boundary_friction_dynamic);
        get_num(*pp, "boundaryFrictionEpsV", p.boundary_friction_eps_v);
        get_num(*pp, "stepSize", p.step_size);
        get_num(*pp, "stepSizeGD", p.step_size_gd);
        get_int(*pp, "frictionStartIter", p.friction_start_iter);
        get_int(*pp, "associateGravityWithInertia", p.associate_gravity_with_inertia);
        // cloth (VBDClothPhysicsParameters::fromJson, VBDCloth/VBDClothPhysicsParameters.h:83-130)
        get_num(*pp, "stepSize", p.cloth_step_size);
        get_num(*pp, "boundaryFrictionEpsValue", p.cloth_boundary_friction_eps);
        handleCollisionGiven = pp->get("handleCollision") != nullptr;
        get_int(*pp, "handleCollision", p.handle_collision);
        get_num(*pp, "contactStiffness", p.contact_stiffness);
        get_num(*pp, "contactRadius", p.contact_radius);
        get_num(*pp, "thickness", p.thickness);
        get_num(*pp, "conservativeStepRelaxation", p.conservative_step_relaxation);
        get_int(*pp, "contactBVHReconstructionSteps", p.contact_bvh_rebuild_steps);
        get_int(*pp, "useAccelerator", p.use_accelerator);
        get_num(*pp, "acceleratorRho", p.accelerator_rho);
        // BaseClothPhsicsFramework::parseRunningParameters (Framework/BaseClothSimPhsicsFramework.cpp:17-28)
        if (const json::Value* cdp = pp->get("ContactDetectorParams")) get_num(*cdp, "maxQueryDis", p.contact_query_radius);
        int32_t useBowl = 0;
        get_int(*pp, "useBowlGround", useBowl);
        if (useBowl) {
            err = "useBowlGround is not supported by this build (plane ground only)";
            return false;
        }
    }
    // the tet framework's collision detection (CollisionDetectionParamters::fromJson,
    // CollisionDetector/CollisionDetertionParameters.h:82-120; BasePhysicFramework::initializeCollisionDetector,
    // Framework/BasePhysicsFramework.cpp:251-262) and its solver constants (VBDPhysicsParameters.h:120-147)
    bool collisionParamsGiven = false;
    if (const json::Value* cp = params.get("CollisionParams")) {
        collisionParamsGiven = true;
        get_int(*cp, "allowDCD", p.allow_dcd);
        get_int(*cp, "allowCCD", p.allow_ccd);
        get_int(*cp, "handleSelfCollision", p.handle_self_collision);
        get_int(*cp, "checkFeasibleRegion", p.check_feasible_region);
        get_int(*cp, "checkTetTraverse", p.check_tet_traverse);
        get_int(*cp, "tetrahedralTraverseForNonSelfIntersection", p.tet_traverse_non_self);
        get_int(*cp, "shiftQueryPointToCenter", p.shift_query_point_to_center);
        get_int(*cp, "stopTraversingAfterPassingQueryPoint", p.stop_traversing_after_passing);
        get_num(*cp, "feasibleRegionEpsilon", p.feasible_region_epsilon);
        get_num(*cp, "rayTriIntersectionEPSILON", p.ray_tri_intersection_epsilon);
        get_num(*cp, "maxSearchDistanceMultiplier", p.max_search_distance_multiplier);
        get_num(*cp, "centerShiftLevel", p.center_shift_level);
        int32_t restPose = 0, loopLess = 0, volumetric = 0, edgeCCD = 0, maxCollisions = 1;
        get_int(*cp, "restPoseCloestPoint", restPose);
        get_int(*cp, "loopLessTraverse", loopLess);
        get_int(*cp, "allowVolumetricCollision", volumetric);
        get_int(*cp, "doEdgeEdgeCCD", edgeCCD);
        get_int(*cp, "maxNumberOfCollisions", maxCollisions);
        if (restPose || loopLess || volumetric || edgeCCD || maxCollisions != 1) {
            err = "CollisionParams: restPoseCloestPoint, loopLessTraverse, allowVolumetricCollision, doEdgeEdgeCCD "
                  "and maxNumberOfCollisions other than 1 are not supported by this build";
            return false;
        }
    }
    if (const json::Value* pp = params.get("PhysicsParams")) {
        get_num(*pp, "collisionStiffness", p.collision_stiffness);
        get_num(*pp, "collisionOffHeight", p.collision_off_height);
        get_int(*pp, "dcdTetMeshSceneBVHRebuildSteps", p.dcd_bvh_rebuild_steps);
        // collisionSolutionType selects between the CPU paths; the device path (runStepGPU) has one
        int32_t intermediate = -1;
        get_int(*pp, "intermediateCollisionIterations", intermediate);
        if (intermediate > 0) {
            err = "intermediateCollisionIterations is not supported by this build (detection once per substep)";
            return false;
        }
    }
    // vertex-face attachments (VertexFaceAttachmentConstraintParams::fromJson,
    // VBDCloth/VBDTriMeshConstraints.h:13-19; loadConstraints, VBDClothPhysics.cpp:148-224)
    if (const json::Value* pp = params.get("PhysicsParams")) {
        if (const json::Value* ac = pp->get("VertexFaceAttachmentConstraintParams")) {
            auto ints = [&](const char* k) {
                std::vector<int32_t> out;
                if (const json::Value* v = ac->get(k))
                    for (const json::Value& e : v->arr) out.push_back((int32_t)e.num);
                return out;
            };
            const std::vector<int32_t> vi = ints("vertexConstraintVertexInfos");
            const std::vector<int32_t> tf = ints("vertexConstraintTargetFacesInfos");
            std::vector<std::array<float, 3>> bary;
            if (const json::Value* v = ac->get("vertexConstraintBarys"))
                for (const json::Value& e : v->arr)
                    bary.push_back({(float)e.arr.at(0).num, (float)e.arr.at(1).num, (float)e.arr.at(2).num});
            std::vector<float> k;
            if (const json::Value* v = ac->get("vertexConstraintStiffness"))
                for (const json::Value& e : v->arr) k.push_back((float)e.num);
            const size_t n = vi.size() / 2;
            if (vi.size() % 2 || tf.size() != 4 * n || bary.size() != n || (!k.empty() && k.size() != n)) {
                err = "VertexFaceAttachmentConstraintParams: array lengths disagree";
                return false;
            }
            for (size_t i = 0; i < n; ++i) {
                gvbd_attachment_desc a{};
                a.mesh = vi[2 * i];
                a.vertex = vi[2 * i + 1];
                a.target_mesh = tf[4 * i];
                for (int j = 0; j < 3; ++j) {
                    a.face[j] = tf[4 * i + 1 + j];
                    a.bary[j] = bary[i][j];
                }
                a.stiffness = k.empty() ? 1e4f : k[i];  // VBDTriMeshConstraints.h:32
                out.attachments.push_back(a);
            }
        }
    }
    // deformers (loadDeformers, VBD/VBDDeformer.h:181-222, VBDCloth/VBDClothDeformer.h:199-243)
    if (const json::Value* dl = params.get("Deformers")) {
        for (const json::Value& dj : dl->arr) {
            LoadedDeformer d;
            const std::string name = get_str(dj, "DeformerName");
            if (name == "Rotator") {
                d.desc.kind = GVBD_DEFORMER_ROTATOR;
                get_vec3(dj, "center", d.desc.center);
                get_vec3(dj, "axis", d.desc.axis);
                if (const json::Value* v = dj.get("angularVelocity")) d.desc.angular_velocity = v->num;
                d.desc.end_time = -1.0;
                if (const json::Value* v = dj.get("rotationEndTime")) d.desc.end_time = v->num;
            } else if (name == "Translator") {
                d.desc.kind = GVBD_DEFORMER_TRANSLATOR;
                get_vec3(dj, "speed", d.desc.speed);
                d.desc.end_time = -1.0;
                if (const json::Value* v = dj.get("deformationEndTime")) d.desc.end_time = v->num;
                get_num(dj, "deformationStartTime", d.desc.start_time);
            } else {
                err = "unrecognised deformer \"" + name + "\" (Rotator, Translator)";
                return false;
            }
            if (const json::Value* v = dj.get("selectedMeshes"))
                for (const json::Value& e : v->arr) d.meshes.push_back((int32_t)e.num);
            if (const json::Value* v = dj.get("selectedVertices"))
                for (const json::Value& l : v->arr) {
                    d.vertices.emplace_back();
                    for (const json::Value& e : l.arr) d.vertices.back().push_back((int32_t)e.num);
                }
            if (d.vertices.size() != d.meshes.size()) {
                err = "deformer \"" + name + "\": selectedVertices needs one list per selected mesh";
                return false;
            }
            out.deformers.push_back(std::move(d));
        }
    }
    const json::Value* list = models.get("Models");
    if (!list || list->type != json::Value::Array) {
        err = modelsPath + ": missing \"Models\" array";
        return false;
    }
    for (const json::Value& mj : list->arr) {
        const std::string material = get_str(mj, "materialName");
        const std::string path = substitute_root(get_str(mj, "path"), repoRoot);
        const std::string colorPath = substitute_root(get_str(mj, "verticesColoringCategoriesPath"), repoRoot);
        float vel0[3] = {0, 0, 0};
        get_vec3(mj, "initialVelocity", vel0);
        std::vector<int32_t> fixed;
        if (const json::Value* fp = mj.get("fixedPoints"))
            for (const json::Value& f : fp->arr) fixed.push_back((int32_t)f.num);
        if (material == "NeoHookean") {
            TetMeshHost m;
            if (!load_t_file(path, m.pos, m.tets, err)) return false;
            m.nv = (int32_t)(m.pos.size() / 3);
            m.nt = (int32_t)(m.tets.size() / 4);
            transform_positions(mj, m.pos, 0.f);
            m.vel.resize(3 * (size_t)m.nv);
            for (int32_t i = 0; i < m.nv; ++i)
                for (int d = 0; d < 3; ++d) m.vel[3 * i + d] = vel0[d];
            get_num(mj, "density", m.mat.density);
            get_num(mj, "miu", m.mat.mu);
            get_num(mj, "lmbd", m.mat.lambda);
            get_num(mj, "dampingHydrostatic", m.mat.dampingHydrostatic);
            get_num(mj, "dampingDeviatoric", m.mat.dampingDeviatoric);
            get_num(mj, "exponentialVelDamping", m.mat.exponentialVelDamping);
            get_num(mj, "constantVelDamping", m.mat.constantVelDamping);
            get_int(mj, "initializationType", m.mat.initializationType);
            get_num(mj, "frictionDynamic", m.mat.frictionDynamic);  // boundary-surface contact friction
            get_num(mj, "frictionEpsV", m.mat.frictionEpsV);
            get_int(mj, "hasNoGravZone", m.mat.hasNoGravZone);  // Materials.h:62-64
            get_num(mj, "noGravZoneThreshold", m.mat.noGravZoneThreshold);
            get_num(mj, "maxVelocityMagnitude", m.mat.maxVelocityMagnitude);
            if (m.mat.initializationType < 0 || m.mat.initializationType > 4) {
                err = "initializationType must be 0..4 for tet meshes in this build";
                return false;
            }
            m.fixed = fixed;
            Graph g = tet_vertex_graph(m.nv, m.nt, m.tets.data());
            if (!colorPath.empty()) {
                if (!read_colors(colorPath, m.nv, m.colors, err)) return false;
                if (!coloring_valid(g, m.colors)) {
                    err = colorPath + ": colouring is not valid for " + path;
                    return false;
                }
            } else {
                m.colors = color_mcs(g);
            }
            out.order.push_back({0, (int)out.tets.size()});
            out.tets.push_back(std::move(m));
        } else if (material == "StVK_triMesh") {
            ClothMeshHost m;
            if (!load_obj_file(path, m.pos, m.faces, err)) return false;
            m.nv = (int32_t)(m.pos.size() / 3);
            m.nf = (int32_t)(m.faces.size() / 3);
            transform_positions(mj, m.pos, 1e-5f);
            m.vel.resize(3 * (size_t)m.nv);
            for (int32_t i = 0; i < m.nv; ++i)
                for (int d = 0; d < 3; ++d) m.vel[3 * i + d] = vel0[d];
            get_num(mj, "density", m.mat.density);
            get_num(mj, "miu", m.mat.mu);
            get_num(mj, "lambda", m.mat.lambda);
            get_num(mj, "bendingStiffness", m.mat.bendingStiffness);
            get_num(mj, "dampingStVK", m.mat.dampingStVK);
            get_num(mj, "dampingBending", m.mat.dampingBending);
            get_num(mj, "exponentialVelDamping", m.mat.exponentialVelDamping);
            get_num(mj, "constantVelDamping", m.mat.constantVelDamping);
            get_int(mj, "initializationType", m.mat.initializationType);
            get_num(mj, "frictionDynamic", m.mat.frictionDynamic);  // TriMesh.cpp:459-460
            get_num(mj, "frictionEpsV", m.mat.frictionEpsV);
            get_int(mj, "hasNoGravZone", m.mat.hasNoGravZone);  // Materials.h:62-64
            get_num(mj, "noGravZoneThreshold", m.mat.noGravZoneThreshold);
            get_num(mj, "maxVelocityMagnitude", m.mat.maxVelocityMagnitude);
            const int t = m.mat.initializationType;
            if (!(t == 0 || t == 2 || t == 3 || t == 4)) {
                err = "cloth initializationType must be 0, 2, 3 or 4 in this build";
                return false;
            }
            m.fixed = fixed;
            if (!colorPath.empty()) {
                if (!read_colors(colorPath, m.nv, m.colors, err)) return false;
            } else {
                m.colors = color_mcs(tri_vertex_graph(m.nv, m.nf, m.faces.data(), true));
            }
            out.order.push_back({1, (int)out.cloths.size()});
            out.cloths.push_back(std::move(m));
        } else {
            err = "materialName \"" + material + "\" is not supported (NeoHookean, StVK_triMesh)";
            return false;
        }
    }
    // handleCollision defaults to true in the cloth framework (VBDClothPhysicsParameters.h:60). The
    // tet framework detects through its CollisionParams (DCD and CCD, both on by default): a tet-only
    // scene with a CollisionParams section, or with handleCollision set, takes the penetration
    // contact model; without either it runs without collision detection
    if (out.cloths.empty()) {
        if (collisionParamsGiven) p.handle_collision = (p.allow_dcd || p.allow_ccd) ? 1 : 0;
        else if (!handleCollisionGiven) p.handle_collision = 0;
        if (p.handle_collision) p.tet_contact_model = 1;
    }
    return true;
}

}  // namespace gvbd
