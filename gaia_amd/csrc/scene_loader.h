// Models.json / Parameters.json scene loading (VBDPhysics::loadRunningparameters
// + the mesh part of BasePhysicFramework::initialize).
#pragma once
#include <string>
#include <vector>

#include "../../include/gaia_vbd.h"
#include "mesh_prep.h"

namespace gvbd {

struct LoadedScene {
    gvbd_physics_params params{};
    std::vector<TetMeshHost> meshes;
};

// Reference defaults of the hot-path parameters (VBD/VBDPhysicsParameters.h:8-81,
// Parameters/PhysicsParameters.h:18-120).
gvbd_physics_params default_physics_params();

// Parses both files. Paths containing "${REPO_ROOT}" get repoRoot substituted
// (Parser/InputHandler.h:128-184). Meshes are loaded from ".t" files and
// transformed by rotation (axis-angle), scale and translation
// (TetMesh/TetMeshFEM.cpp:358-375); vertex colours come from
// verticesColoringCategoriesPath when given, otherwise from MCS colouring.
bool load_scene_json(const std::string& modelsPath, const std::string& paramsPath, const std::string& repoRoot,
                     LoadedScene& out, std::string& err);

}  // namespace gvbd
