// CPU oracle: graph colouring restatement.
// TEST INFRASTRUCTURE ONLY (see oracle.h).
//
//   GraphColoring/TetMeshVertexGraph.h:15-41  vertex graph: every vertex pair sharing a tet
//   GraphColoring/ColoringAlgorithms.cpp:12-25 adjacency lists from the edge list
//   GraphColoring/mcs.cpp:14-168              maximum-cardinality-search order, smallest free colour
//   GraphColoring/gready.cpp:15-117           ordered greedy: min current degree, smallest free colour
//
// The shuffled coloringOrder in mcs.cpp only changes when the remaining-node list
// is compacted, never which node is picked (first max-weight node in ascending
// id order), so the result is deterministic; node 0 is always picked first, so the
// compaction's `> 0` filter never drops an unpicked node.
#include "oracle.h"

#include <algorithm>
#include <map>
#include <set>
#include <utility>
#include <vector>

namespace {

using Adj = std::vector<std::vector<int>>;

Adj adjacency(int nv, const std::vector<std::pair<int, int>>& edges) {
    Adj g(nv);
    for (auto& e : edges) {
        g[e.first].push_back(e.second);
        g[e.second].push_back(e.first);
    }
    return g;
}

std::vector<int> color_mcs(const Adj& g) {
    const int n = (int)g.size();
    std::vector<int> weight(n, 0), order, colors(n, -1);
    std::vector<char> taken(n, 0);
    for (int i = 0; i < n; ++i) {
        int best = -1, bw = -1;
        for (int j = 0; j < n; ++j)
            if (!taken[j] && weight[j] > bw) {
                bw = weight[j];
                best = j;
            }
        order.push_back(best);
        taken[best] = 1;
        for (int nb : g[best]) weight[nb] += 1;
    }
    for (int v : order) {
        std::vector<int> used;
        for (int nb : g[v]) used.push_back(colors[nb]);
        std::sort(used.begin(), used.end());
        used.erase(std::unique(used.begin(), used.end()), used.end());
        int c = 0;
        for (int u : used)
            if (u == c) ++c;
        colors[v] = c;
    }
    return colors;
}

std::vector<int> color_greedy(const Adj& g) {
    const int n = (int)g.size();
    std::vector<int> deg(n), colors(n, -1);
    for (int i = 0; i < n; ++i) deg[i] = (int)g[i].size();
    int maxColor = -1;
    for (int k = 0; k < n; ++k) {
        int node = -1, md = n + 1;
        for (int i = 0; i < n; ++i)
            if (deg[i] != -1 && md > deg[i]) {
                md = deg[i];
                node = i;
            }
        if (maxColor == -1) {
            maxColor = 0;
            colors[node] = 0;
        } else {
            std::vector<char> used(maxColor + 1, 0);
            for (int nb : g[node])
                if (colors[nb] >= 0) used[colors[nb]] = 1;
            int c = -1;
            for (int i = 0; i <= maxColor; ++i)
                if (!used[i]) {
                    c = i;
                    break;
                }
            if (c == -1) c = ++maxColor;
            colors[node] = c;
        }
        deg[node] = -1;
        for (int nb : g[node])
            if (deg[nb] != -1) deg[nb]--;
    }
    return colors;
}

}  // namespace

extern "C" {

int oracle_color_graph(int nv, int ne, const int32_t* edges, int algorithm, int32_t* colors) {
    std::vector<std::pair<int, int>> el;
    for (int i = 0; i < ne; ++i) el.push_back({edges[2 * i], edges[2 * i + 1]});
    Adj g = adjacency(nv, el);
    std::vector<int> c = algorithm == 0 ? color_mcs(g) : color_greedy(g);
    int nc = 0;
    for (int i = 0; i < nv; ++i) {
        colors[i] = c[i];
        nc = std::max(nc, c[i] + 1);
    }
    return nc;
}

int oracle_color_tet_mesh(int nv, int nt, const int32_t* tets, int algorithm, int32_t* colors) {
    std::set<std::pair<int, int>> es;
    for (int t = 0; t < nt; ++t)
        for (int a = 0; a < 4; ++a)
            for (int b = a + 1; b < 4; ++b) {
                int u = tets[4 * t + a], v = tets[4 * t + b];
                if (u == v) continue;
                es.insert({std::min(u, v), std::max(u, v)});
            }
    std::vector<int32_t> flat;
    for (auto& e : es) {
        flat.push_back(e.first);
        flat.push_back(e.second);
    }
    return oracle_color_graph(nv, (int)es.size(), flat.data(), algorithm, colors);
}

// TriMeshVertexGraph.h:20-45: every mesh edge, and with bending the pair of vertices
// opposite each interior edge.
int oracle_color_tri_mesh(int nv, int nf, const int32_t* faces, int with_bending, int algorithm, int32_t* colors) {
    std::map<std::pair<int, int>, std::vector<int>> opposite;
    for (int f = 0; f < nf; ++f)
        for (int k = 0; k < 3; ++k) {
            int a = faces[3 * f + k], b = faces[3 * f + (k + 1) % 3], c = faces[3 * f + (k + 2) % 3];
            opposite[{std::min(a, b), std::max(a, b)}].push_back(c);
        }
    std::vector<int32_t> flat;
    for (auto& kv : opposite) {
        flat.push_back(kv.first.first);
        flat.push_back(kv.first.second);
        if (with_bending && kv.second.size() == 2) {
            flat.push_back(kv.second[0]);
            flat.push_back(kv.second[1]);
        }
    }
    return oracle_color_graph(nv, (int)flat.size() / 2, flat.data(), algorithm, colors);
}

}  // extern "C"
