// Scene files in the reference's formats: recovery states (PhysicsState, text or binary) and
// frame outputs (ASCII PLY surfaces, raw float32 surface positions). Host-only code; the scene
// side (downloading positions, building surfaces) lives in vbd_api.cpp.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace gvbd {
namespace io {

struct MeshState {
    std::vector<float> position;    // 3 per vertex
    std::vector<float> velocities;  // 3 per vertex, or empty when the file holds none
};

// PhysicsState::toJson + writeToJsonFile(path, indent 2) (Framework/BasePhysicsFramework.h:258-317):
// text: {"curTime", "format": "text", "frameId", "meshesState": [{"position": [[x, y, z], ...],
// "velocities": [...]}, ...]}; binary: the header names "<stem>.pos" and "<stem>.vel" beside the
// JSON, each the meshes' float32 xyz triples back to back, and lists the vertex counts.
void write_state(const std::string& path, int32_t frameId, double curTime, bool binary,
                 const std::vector<MeshState>& meshes);

// PhysicsState::fromJson (BasePhysicsFramework.h:200-256): "format" defaults to text; the binary
// files are found beside the JSON. A mesh without "velocities" gets an empty vector (the shipped
// initial-state files hold positions only).
void read_state(const std::string& path, int32_t& frameId, double& curTime, std::vector<MeshState>& meshes);

// GAIA::saveAsPly (IO/FileIO.cpp:8-50): ASCII PLY, vertices as "%f %f %f", faces as "3 %d %d %d",
// no face element when there are no faces.
void write_ply(const std::string& path, const std::vector<float>& verts, const std::vector<int32_t>& faces);

// GAIA::writeAllToBinary (IO/FileIO.cpp:52-117): the float32 positions, nothing else.
void write_binary(const std::string& path, const std::vector<float>& verts);

}  // namespace io
}  // namespace gvbd
