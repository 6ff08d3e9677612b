// Host-side mesh preparation for the VBD hot path: rest-pose constants, vertex
// graphs, colouring, parallel groups and the colour-major device records.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "device_types.h"

namespace gvbd {

struct TetMaterial {
    float density = 1.f;
    float mu = 1e4f;
    float lambda = 1e4f;
    float dampingHydrostatic = 0.f;
    float dampingDeviatoric = 0.f;
    float exponentialVelDamping = 1.f;
    float constantVelDamping = 0.f;
    int32_t initializationType = 0;
    float initRatioG = 0.5f;
    float frictionDynamic = 0.1f;  // contact friction of the boundary surface (scenes with collision handling)
    float frictionEpsV = 0.1f;
    // no-gravity zone (Materials.h:34-35): position y above the threshold -> speed capped, no damping
    int32_t hasNoGravZone = 0;
    float noGravZoneThreshold = 0.f;
    float maxVelocityMagnitude = -1.f;
};

struct TetMeshHost {
    int32_t nv = 0, nt = 0;
    std::vector<float> pos, vel;  // 3 * nv, xyz interleaved
    std::vector<int32_t> tets;    // 4 * nt
    TetMaterial mat;
    std::vector<int32_t> fixed;
    std::vector<int32_t> colors;  // colour id per vertex
    // derived
    std::vector<float> dminv, vol, mass;              // 9 * nt (column-major), nt, nv
    std::vector<int32_t> incBegin, incTet, incOrder;  // vertex -> tet CSR, ascending tet id
    std::vector<std::vector<int32_t>> categories;     // vertices per colour, sorted by size desc
};

// TetMeshFEM.cpp:409-429 (DmInv, signed rest volume det(Dm)/6, lumped mass
// vol*density/4 per vertex) and TetMeshFEM.cpp:275-325 (incidence CSR).
// Returns false with a message on degenerate tets or bad indices.
bool prepare_tet_mesh(TetMeshHost& m, std::string& err);

// Unique-neighbour adjacency (CSR) of a vertex graph given by an edge list.
struct Graph {
    std::vector<int32_t> begin, adj;
    int32_t n() const { return (int32_t)begin.size() - 1; }
};
Graph graph_from_edges(int32_t nv, const int32_t* edges, int32_t ne);
// TetMeshVertexGraph.h:15-41: every pair of vertices that share a tet.
Graph tet_vertex_graph(int32_t nv, int32_t nt, const int32_t* tets);

// mcs.cpp:14-168 (maximum cardinality search order, lowest id on ties) and
// gready.cpp:15-117 (minimum current degree, lowest id on ties); both give each
// vertex the smallest colour unused by its coloured neighbours. O((V+E) log V).
std::vector<int32_t> color_mcs(const Graph& g);
std::vector<int32_t> color_greedy(const Graph& g);
int32_t num_colors(const std::vector<int32_t>& colors);
bool coloring_valid(const Graph& g, const std::vector<int32_t>& colors);

// ColoringAlgorithms.cpp:69-80 categories (ascending ids per colour), then the
// size-descending sort of TetMeshFEM.cpp:122-133. The reference uses the unstable
// std::sort; this build uses a stable sort (equal-size colours keep colour order).
std::vector<std::vector<int32_t>> sorted_categories(const std::vector<int32_t>& colors);

// VBDPhysics.cpp:99-122,235-274 and VBDClothPhysics.cpp:72-111: colour c of mesh k goes to
// the currently smallest of the first (number of colours of k) parallel groups; with
// removeUsed[k] a mesh uses each group at most once (tet meshes), without it (cloth meshes)
// two colours of one mesh may share a group. Groups hold interleaved (mesh, vertex) pairs.
std::vector<std::vector<int32_t>> merge_parallel_groups(
    const std::vector<const std::vector<std::vector<int32_t>>*>& meshCategories, const std::vector<bool>& removeUsed);

struct ClothMaterial {
    float density = 1.f;
    float mu = 1e4f;
    float lambda = 1e4f;
    float bendingStiffness = 0.f;
    float dampingStVK = 0.f;
    float dampingBending = 0.f;
    float exponentialVelDamping = 1.f;  // applied per substep as given (cloth does not rescale by dt)
    float constantVelDamping = 0.f;
    int32_t initializationType = 2;     // VBDObjectParamsTriMesh default: previous position
    float frictionDynamic = 0.1f;       // TriMesh.h:41-42 (contact friction)
    float frictionEpsV = 0.1f;
    // no-gravity zone (Materials.h:34-35): velocity y above the threshold -> speed capped, no damping
    int32_t hasNoGravZone = 0;
    float noGravZoneThreshold = 0.f;
    float maxVelocityMagnitude = -1.f;
};

// Edge of a triangle mesh (TriMesh/TriMesh.h:51-60). Edges are numbered in order of first
// appearance when scanning faces (edges v0v1, v1v2, v2v0 of each face); eV1 -> eV2 follows the
// first face that contains the edge (fId1), eV12Next is that face's third vertex; fId2 /
// eV21Next are the second face's (-1 on boundary edges).
struct TriEdge {
    int32_t v1, v2, v12Next, v21Next, f1, f2;
};

struct ClothMeshHost {
    int32_t nv = 0, nf = 0;
    std::vector<float> pos, vel;
    std::vector<int32_t> faces;  // 3 * nf
    ClothMaterial mat;
    std::vector<int32_t> fixed;
    std::vector<int32_t> colors;
    // derived
    std::vector<float> dminv, area, mass;              // 4 * nf (col-major 2x2), nf, nv
    std::vector<TriEdge> edges;
    std::vector<float> Q;                               // 16 per edge (row-major), zero on boundary edges
    std::vector<int32_t> vfBegin, vfFace, vfOrder;      // vertex -> faces (ascending)
    std::vector<int32_t> vbBegin, vbEdge, vbOrder;      // vertex -> relevant bendings (order 0..3)
    std::vector<std::vector<int32_t>> categories;
};

// TriMesh.cpp:213-244 (rest pose from 3D: right-handed in-plane frame, Dm, area = det(Dm)/2,
// mass += rho * area / 3), TriMesh.cpp:481-753 (edges, vertex->faces, relevant bendings:
// incident interior edges ascending with order 0/1, then per incident face (ascending) its
// opposite interior edge with order 2 (face is fId1) or 3) and VBDTriMeshStVK.cpp:50-88
// (Q = 3 K0 K0^T / (|n1| + |n2|)). Fails on degenerate faces or non-manifold / inconsistently
// oriented meshes.
bool prepare_cloth_mesh(ClothMeshHost& m, std::string& err);

// Edges of a triangle mesh numbered by first appearance over the faces (TriMesh.cpp:481-560):
// v1 -> v2 is the edge's direction in its first face f1, v12Next that face's third vertex,
// v21Next / f2 the other face (-1 on boundary edges). faceEdges (optional, 3 per face) receives
// the ids of edges (v0 v1), (v1 v2), (v2 v0) of each face (faces3NeighborEdges order).
// Fails on out-of-range indices and non-manifold or inconsistently oriented meshes.
bool build_tri_edges(int32_t nv, int32_t nf, const int32_t* faces, std::vector<TriEdge>& edges,
                     std::vector<int32_t>* faceEdges, std::string& err);

// Boundary surface of a tet mesh: the faces of exactly one tet, in (tet, face) order (face k of
// a tet leaves out its vertex k), each oriented with its normal away from that vertex. Used as
// the mesh's contact surface in scenes with collision handling. Optionally also the owning tet
// and local face of every boundary face (2 per face) and, per tet and local face k, the tet
// sharing that face (4 per tet, -1 on the boundary).
bool build_tet_surface(int32_t nv, int32_t nt, const int32_t* tets, const float* pos, std::vector<int32_t>& faces,
                       std::string& err, std::vector<int32_t>* owner = nullptr, std::vector<int32_t>* tetNbr = nullptr);
// TriMeshVertexGraph.h:20-45: mesh edges; with bending also the two vertices opposite each
// interior edge.
Graph tri_vertex_graph(int32_t nv, int32_t nf, const int32_t* faces, bool withBending);

// Reads the vertices ("v x y z") and faces ("f a b c", 1-based, "a/b/c" forms allowed) of an .obj file.
bool load_obj_file(const std::string& path, std::vector<float>& pos, std::vector<int32_t>& faces, std::string& err);

// Reads a MeshFrame ".t" file ("Vertex i x y z" / "Tet i a b c d" lines).
bool load_t_file(const std::string& path, std::vector<float>& pos, std::vector<int32_t>& tets, std::string& err);

}  // namespace gvbd
