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

// VBDPhysics.cpp:99-122,235-274: colour c of mesh k goes to the currently smallest
// parallel group; with `removeUsed` a mesh uses each group at most once (tet meshes).
// Groups hold interleaved (mesh, vertex) pairs.
std::vector<std::vector<int32_t>> merge_parallel_groups(const std::vector<const std::vector<std::vector<int32_t>>*>& meshCategories,
                                                        bool removeUsed);

// Reads a MeshFrame ".t" file ("Vertex i x y z" / "Tet i a b c d" lines).
bool load_t_file(const std::string& path, std::vector<float>& pos, std::vector<int32_t>& tets, std::string& err);

}  // namespace gvbd
