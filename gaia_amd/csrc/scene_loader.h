// Models.json / Parameters.json scene loading (VBDPhysics::loadRunningparameters
// / VBDClothSimulationFramework parameter parsing + the mesh part of initialize).
#pragma once
#include <string>
#include <utility>
#include <vector>

#include "../../include/gaia_vbd.h"
#include "mesh_prep.h"

namespace gvbd {

// One entry of the parameter file's "Deformers" array (loadDeformers, VBD/VBDDeformer.h:181-222).
struct LoadedDeformer {
    gvbd_deformer_desc desc{};
    std::vector<int32_t> meshes;
    std::vector<std::vector<int32_t>> vertices;
};

struct LoadedScene {
    std::vector<LoadedDeformer> deformers;
    std::vector<gvbd_attachment_desc> attachments;  // PhysicsParams.VertexFaceAttachmentConstraintParams
    gvbd_physics_params params{};
    std::vector<TetMeshHost> tets;
    std::vector<ClothMeshHost> cloths;
    std::vector<std::pair<int, int>> order;  // (0 tet / 1 cloth, index) in Models order
};

// Reference defaults of the hot-path parameters (VBD/VBDPhysicsParameters.h:8-81,
// Parameters/PhysicsParameters.h:18-120, VBDCloth/VBDClothPhysicsParameters.h:15-81).
gvbd_physics_params default_physics_params();

// Parses both files. Paths containing "${REPO_ROOT}" get repoRoot substituted
// (Parser/InputHandler.h:128-184). "NeoHookean" models load ".t" files, "StVK_triMesh"
// models load ".obj" files; both are transformed by rotation (axis-angle), scale and
// translation (TetMesh/TetMeshFEM.cpp:358-375, TriMesh/TriMesh.cpp:115-160). Vertex colours
// come from verticesColoringCategoriesPath when given, otherwise from MCS colouring (of the
// vertex graph with bending edges for cloth).
bool load_scene_json(const std::string& modelsPath, const std::string& paramsPath, const std::string& repoRoot,
                     LoadedScene& out, std::string& err);

}  // namespace gvbd
