// Host-side mesh preparation (see mesh_prep.h for the reference citations).
#include "mesh_prep.h"

#include <algorithm>
#include <cmath>
#include <fstream>
#include <queue>
#include <sstream>
#include <utility>

namespace gvbd {

namespace {

// Column-major 3x3 (m[r + 3c]) inverse by adjugate; false when singular.
bool inverse3(const float* m, float* out, float& det) {
    const float c00 = m[4] * m[8] - m[7] * m[5];
    const float c01 = -(m[1] * m[8] - m[7] * m[2]);
    const float c02 = m[1] * m[5] - m[4] * m[2];
    const float d = m[0] * c00 + m[3] * c01 + m[6] * c02;
    if (!(d != 0.f) || !std::isfinite(d)) return false;
    const float inv = 1.f / d;
    out[0] = c00 * inv;
    out[1] = c01 * inv;
    out[2] = c02 * inv;
    out[3] = -(m[3] * m[8] - m[6] * m[5]) * inv;
    out[4] = (m[0] * m[8] - m[6] * m[2]) * inv;
    out[5] = -(m[0] * m[5] - m[3] * m[2]) * inv;
    out[6] = (m[3] * m[7] - m[6] * m[4]) * inv;
    out[7] = -(m[0] * m[7] - m[6] * m[1]) * inv;
    out[8] = (m[0] * m[4] - m[3] * m[1]) * inv;
    det = m[0] * (m[4] * m[8] - m[7] * m[5]) - m[1] * (m[3] * m[8] - m[6] * m[5]) +
          m[2] * (m[3] * m[7] - m[6] * m[4]);
    return true;
}

}  // namespace

bool prepare_tet_mesh(TetMeshHost& m, std::string& err) {
    if ((int64_t)m.pos.size() != 3ll * m.nv || (int64_t)m.tets.size() != 4ll * m.nt) {
        err = "tet mesh: array sizes do not match vertex/tet counts";
        return false;
    }
    for (int32_t t : m.tets)
        if (t < 0 || t >= m.nv) {
            err = "tet mesh: vertex index out of range";
            return false;
        }
    for (int32_t f : m.fixed)
        if (f < 0 || f >= m.nv) {
            err = "tet mesh: fixed point index out of range";
            return false;
        }
    if (m.mat.lambda == 0.f) {
        err = "tet mesh: lambda must be non-zero (a = 1 + mu/lambda)";
        return false;
    }
    m.dminv.assign(9 * (size_t)m.nt, 0.f);
    m.vol.assign(m.nt, 0.f);
    m.mass.assign(m.nv, 0.f);
    for (int32_t t = 0; t < m.nt; ++t) {
        const int32_t* tv = &m.tets[4 * (size_t)t];
        float Dm[9];
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 3; ++r) Dm[r + 3 * c] = m.pos[3 * (size_t)tv[c + 1] + r] - m.pos[3 * (size_t)tv[0] + r];
        float det;
        if (!inverse3(Dm, &m.dminv[9 * (size_t)t], det)) {
            err = "tet mesh: degenerate tet " + std::to_string(t);
            return false;
        }
        const float vol = det / 6.f;
        m.vol[t] = vol;
        for (int k = 0; k < 4; ++k)
            m.mass[tv[k]] = (float)((double)m.mass[tv[k]] + 0.25 * (double)vol * (double)m.mat.density);
    }
    // vertex -> tet incidence, ascending tet id
    m.incBegin.assign(m.nv + 1, 0);
    for (int32_t t = 0; t < m.nt; ++t)
        for (int k = 0; k < 4; ++k) m.incBegin[m.tets[4 * (size_t)t + k] + 1]++;
    for (int32_t i = 0; i < m.nv; ++i) m.incBegin[i + 1] += m.incBegin[i];
    m.incTet.assign(m.incBegin[m.nv], 0);
    m.incOrder.assign(m.incBegin[m.nv], 0);
    std::vector<int32_t> fill(m.incBegin.begin(), m.incBegin.end() - 1);
    for (int32_t t = 0; t < m.nt; ++t) {
        for (int k = 0; k < 4; ++k) {
            const int32_t v = m.tets[4 * (size_t)t + k];
            int ord = k;
            for (int kk = 0; kk < k; ++kk)
                if (m.tets[4 * (size_t)t + kk] == v) {
                    ord = kk;
                    break;
                }
            if (ord != k) continue;  // repeated vertex in one tet: counted once per distinct slot
            m.incTet[fill[v]] = t;
            m.incOrder[fill[v]] = ord;
            fill[v]++;
        }
    }
    return true;
}

Graph graph_from_edges(int32_t nv, const int32_t* edges, int32_t ne) {
    std::vector<std::pair<int32_t, int32_t>> e;
    e.reserve(2 * (size_t)ne);
    for (int32_t i = 0; i < ne; ++i) {
        const int32_t a = edges[2 * i], b = edges[2 * i + 1];
        if (a == b || a < 0 || b < 0 || a >= nv || b >= nv) continue;
        e.push_back({a, b});
        e.push_back({b, a});
    }
    std::sort(e.begin(), e.end());
    e.erase(std::unique(e.begin(), e.end()), e.end());
    Graph g;
    g.begin.assign(nv + 1, 0);
    g.adj.resize(e.size());
    for (size_t i = 0; i < e.size(); ++i) {
        g.begin[e[i].first + 1]++;
        g.adj[i] = e[i].second;
    }
    for (int32_t i = 0; i < nv; ++i) g.begin[i + 1] += g.begin[i];
    return g;
}

Graph tet_vertex_graph(int32_t nv, int32_t nt, const int32_t* tets) {
    // vertex -> tet CSR, then per vertex the union of its tets' vertices (marker dedup)
    std::vector<int32_t> ib(nv + 1, 0), it(4 * (size_t)nt);
    for (size_t k = 0; k < 4 * (size_t)nt; ++k) ib[tets[k] + 1]++;
    for (int32_t i = 0; i < nv; ++i) ib[i + 1] += ib[i];
    std::vector<int32_t> fill(ib.begin(), ib.end() - 1);
    for (int32_t t = 0; t < nt; ++t)
        for (int a = 0; a < 4; ++a) it[fill[tets[4 * (size_t)t + a]]++] = t;
    Graph g;
    g.begin.assign(nv + 1, 0);
    g.adj.reserve(14 * (size_t)nv);
    std::vector<int32_t> mark(nv, -1);
    for (int32_t v = 0; v < nv; ++v) {
        mark[v] = v;
        const size_t start = g.adj.size();
        for (int32_t k = ib[v]; k < ib[v + 1]; ++k) {
            const int32_t* tv = tets + 4 * (size_t)it[k];
            for (int a = 0; a < 4; ++a)
                if (mark[tv[a]] != v) {
                    mark[tv[a]] = v;
                    g.adj.push_back(tv[a]);
                }
        }
        std::sort(g.adj.begin() + start, g.adj.end());
        g.begin[v + 1] = (int32_t)g.adj.size();
    }
    return g;
}

namespace {

int32_t smallest_free_color(const Graph& g, int32_t v, const std::vector<int32_t>& colors, std::vector<int32_t>& mark,
                            int32_t stamp) {
    for (int32_t k = g.begin[v]; k < g.begin[v + 1]; ++k) {
        const int32_t c = colors[g.adj[k]];
        if (c >= 0 && c < (int32_t)mark.size()) mark[c] = stamp;
    }
    int32_t c = 0;
    while (c < (int32_t)mark.size() && mark[c] == stamp) ++c;
    return c;
}

}  // namespace

std::vector<int32_t> color_mcs(const Graph& g) {
    const int32_t n = g.n();
    // max weight first, lowest id on ties: key (weight, -id)
    std::priority_queue<std::pair<int32_t, int32_t>> pq;
    std::vector<int32_t> weight(n, 0), order;
    std::vector<char> taken(n, 0);
    order.reserve(n);
    for (int32_t i = 0; i < n; ++i) pq.push({0, -i});
    while (!pq.empty()) {
        auto top = pq.top();
        pq.pop();
        const int32_t v = -top.second;
        if (taken[v] || top.first != weight[v]) continue;
        taken[v] = 1;
        order.push_back(v);
        for (int32_t k = g.begin[v]; k < g.begin[v + 1]; ++k) {
            const int32_t u = g.adj[k];
            weight[u] += 1;
            if (!taken[u]) pq.push({weight[u], -u});
        }
    }
    std::vector<int32_t> colors(n, -1), mark(64, -1);
    int32_t stamp = 0;
    for (int32_t v : order) {
        int32_t deg = g.begin[v + 1] - g.begin[v];
        if ((int32_t)mark.size() < deg + 1) mark.resize(deg + 1, -1);
        colors[v] = smallest_free_color(g, v, colors, mark, stamp++);
    }
    return colors;
}

std::vector<int32_t> color_greedy(const Graph& g) {
    const int32_t n = g.n();
    // min current degree first, lowest id on ties: key (-degree, -id)
    std::priority_queue<std::pair<int32_t, int32_t>> pq;
    std::vector<int32_t> deg(n), colors(n, -1), mark(64, -1);
    std::vector<char> done(n, 0);
    for (int32_t i = 0; i < n; ++i) {
        deg[i] = g.begin[i + 1] - g.begin[i];
        pq.push({-deg[i], -i});
    }
    int32_t stamp = 0;
    while (!pq.empty()) {
        auto top = pq.top();
        pq.pop();
        const int32_t v = -top.second;
        if (done[v] || -top.first != deg[v]) continue;
        done[v] = 1;
        const int32_t d = g.begin[v + 1] - g.begin[v];
        if ((int32_t)mark.size() < d + 1) mark.resize(d + 1, -1);
        colors[v] = smallest_free_color(g, v, colors, mark, stamp++);
        for (int32_t k = g.begin[v]; k < g.begin[v + 1]; ++k) {
            const int32_t u = g.adj[k];
            if (!done[u]) {
                deg[u]--;
                pq.push({-deg[u], -u});
            }
        }
    }
    return colors;
}

int32_t num_colors(const std::vector<int32_t>& colors) {
    int32_t nc = 0;
    for (int32_t c : colors) nc = std::max(nc, c + 1);
    return nc;
}

bool coloring_valid(const Graph& g, const std::vector<int32_t>& colors) {
    if ((int32_t)colors.size() != g.n()) return false;
    for (int32_t v = 0; v < g.n(); ++v) {
        if (colors[v] < 0) return false;
        for (int32_t k = g.begin[v]; k < g.begin[v + 1]; ++k)
            if (colors[g.adj[k]] == colors[v]) return false;
    }
    return true;
}

std::vector<std::vector<int32_t>> sorted_categories(const std::vector<int32_t>& colors) {
    std::vector<std::vector<int32_t>> cat(num_colors(colors));
    for (int32_t v = 0; v < (int32_t)colors.size(); ++v)
        if (colors[v] >= 0) cat[colors[v]].push_back(v);
    std::stable_sort(cat.begin(), cat.end(),
                     [](const std::vector<int32_t>& a, const std::vector<int32_t>& b) { return a.size() > b.size(); });
    return cat;
}

std::vector<std::vector<int32_t>> merge_parallel_groups(
    const std::vector<const std::vector<std::vector<int32_t>>*>& meshCategories, const std::vector<bool>& removeUsed) {
    size_t ng = 0;
    for (auto* c : meshCategories) ng = std::max(ng, c->size());
    std::vector<std::vector<int32_t>> groups(ng);
    for (size_t im = 0; im < meshCategories.size(); ++im) {
        const auto& cats = *meshCategories[im];
        // the available list holds the mesh's own colour count of groups (VBDPhysics.cpp:245-252,
        // VBDClothPhysics.cpp:88-94); groups beyond it are never chosen for this mesh
        std::vector<int32_t> avail(cats.size());
        for (size_t i = 0; i < cats.size(); ++i) avail[i] = (int32_t)i;
        for (size_t c = 0; c < cats.size(); ++c) {
            size_t best = groups[avail[0]].size();
            size_t bestI = 0;
            for (size_t i = 1; i < avail.size(); ++i)
                if (groups[avail[i]].size() < best) {
                    best = groups[avail[i]].size();
                    bestI = i;
                }
            const int32_t gid = avail[bestI];
            if (removeUsed[im]) avail.erase(avail.begin() + bestI);
            for (int32_t v : cats[c]) {
                groups[gid].push_back((int32_t)im);
                groups[gid].push_back(v);
            }
        }
    }
    return groups;
}

namespace {

inline void cross3f(const float* a, const float* b, float* c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}
inline float dot3f(const float* a, const float* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// Edge key -> id via a sorted list of (min, max, first index).
struct EdgeTable {
    std::vector<std::pair<uint64_t, int32_t>> keys;
    static uint64_t key(int32_t a, int32_t b) {
        const uint32_t lo = (uint32_t)std::min(a, b), hi = (uint32_t)std::max(a, b);
        return ((uint64_t)lo << 32) | hi;
    }
    int32_t find(int32_t a, int32_t b) const {
        const uint64_t k = key(a, b);
        auto it = std::lower_bound(keys.begin(), keys.end(), std::make_pair(k, (int32_t)INT32_MIN));
        return (it != keys.end() && it->first == k) ? it->second : -1;
    }
};

}  // namespace

bool build_tri_edges(int32_t nv, int32_t nf, const int32_t* faces, std::vector<TriEdge>& edges,
                     std::vector<int32_t>* faceEdges, std::string& err) {
    EdgeTable table;
    std::vector<std::pair<uint64_t, int32_t>> occ;
    occ.reserve(3 * (size_t)nf);
    for (int32_t f = 0; f < nf; ++f)
        for (int k = 0; k < 3; ++k) {
            const int32_t a = faces[3 * (size_t)f + k], b = faces[3 * (size_t)f + (k + 1) % 3];
            if (a < 0 || a >= nv || b < 0 || b >= nv) {
                err = "triangle mesh: vertex index out of range";
                return false;
            }
            occ.push_back({EdgeTable::key(a, b), 3 * f + k});
        }
    std::stable_sort(occ.begin(), occ.end(),
                     [](const std::pair<uint64_t, int32_t>& a, const std::pair<uint64_t, int32_t>& b) {
                         return a.first < b.first;
                     });
    // first occurrence of each key, in occurrence order
    std::vector<int32_t> firstOcc;
    for (size_t i = 0; i < occ.size(); ++i)
        if (i == 0 || occ[i].first != occ[i - 1].first) firstOcc.push_back(occ[i].second);
    std::sort(firstOcc.begin(), firstOcc.end());
    edges.clear();
    edges.reserve(firstOcc.size());
    for (int32_t o : firstOcc) {
        const int32_t f = o / 3, k = o % 3;
        const int32_t a = faces[3 * (size_t)f + k], b = faces[3 * (size_t)f + (k + 1) % 3],
                      c = faces[3 * (size_t)f + (k + 2) % 3];
        table.keys.push_back({EdgeTable::key(a, b), (int32_t)edges.size()});
        edges.push_back({a, b, c, -1, f, -1});
    }
    std::sort(table.keys.begin(), table.keys.end());
    if (faceEdges) faceEdges->assign(3 * (size_t)nf, -1);
    for (size_t i = 0; i < occ.size(); ++i) {
        const int32_t f = occ[i].second / 3, k = occ[i].second % 3;
        const int32_t a = faces[3 * (size_t)f + k], b = faces[3 * (size_t)f + (k + 1) % 3],
                      c = faces[3 * (size_t)f + (k + 2) % 3];
        const int32_t id = table.find(a, b);
        if (faceEdges) (*faceEdges)[3 * (size_t)f + k] = id;
        TriEdge& e = edges[id];
        if (e.f1 == f) continue;
        if (e.f2 != -1 || e.v1 != b || e.v2 != a) {
            err = "triangle mesh: non-manifold or inconsistently oriented edge (" + std::to_string(a) + ", " +
                  std::to_string(b) + ")";
            return false;
        }
        e.f2 = f;
        e.v21Next = c;
    }
    return true;
}

bool build_tet_surface(int32_t nv, int32_t nt, const int32_t* tets, const float* pos, std::vector<int32_t>& faces,
                       std::string& err, std::vector<int32_t>* owner, std::vector<int32_t>* tetNbr) {
    static const int kOpp[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
    // (sorted vertex triple, 4 t + k) for every tet face; a triple seen once is on the boundary
    struct Key {
        int32_t a, b, c;
        int64_t at;
    };
    std::vector<Key> keys;
    keys.reserve(4 * (size_t)nt);
    for (int32_t t = 0; t < nt; ++t)
        for (int k = 0; k < 4; ++k) {
            int32_t v[3] = {tets[4 * (size_t)t + kOpp[k][0]], tets[4 * (size_t)t + kOpp[k][1]],
                            tets[4 * (size_t)t + kOpp[k][2]]};
            for (int32_t x : v)
                if (x < 0 || x >= nv) {
                    err = "tet mesh: vertex index out of range";
                    return false;
                }
            std::sort(v, v + 3);
            keys.push_back({v[0], v[1], v[2], 4 * (int64_t)t + k});
        }
    std::sort(keys.begin(), keys.end(), [](const Key& p, const Key& q) {
        if (p.a != q.a) return p.a < q.a;
        if (p.b != q.b) return p.b < q.b;
        if (p.c != q.c) return p.c < q.c;
        return p.at < q.at;
    });
    std::vector<int64_t> boundary;
    if (tetNbr) tetNbr->assign(4 * (size_t)nt, -1);
    for (size_t i = 0; i < keys.size();) {
        size_t j = i + 1;
        while (j < keys.size() && keys[j].a == keys[i].a && keys[j].b == keys[i].b && keys[j].c == keys[i].c) ++j;
        if (j - i == 1) boundary.push_back(keys[i].at);
        if (j - i == 2 && tetNbr) {
            (*tetNbr)[keys[i].at] = (int32_t)(keys[i + 1].at / 4);
            (*tetNbr)[keys[i + 1].at] = (int32_t)(keys[i].at / 4);
        }
        i = j;
    }
    std::sort(boundary.begin(), boundary.end());
    faces.clear();
    faces.reserve(3 * boundary.size());
    if (owner) owner->clear();
    for (int64_t at : boundary) {
        const int64_t t = at / 4;
        const int k = (int)(at % 4);
        if (owner) {
            owner->push_back((int32_t)t);
            owner->push_back(k);
        }
        int32_t a = tets[4 * t + kOpp[k][0]], b = tets[4 * t + kOpp[k][1]], c = tets[4 * t + kOpp[k][2]];
        const int32_t o = tets[4 * t + k];
        float e1[3], e2[3], w[3], n[3];
        for (int d = 0; d < 3; ++d) {
            e1[d] = pos[3 * (size_t)b + d] - pos[3 * (size_t)a + d];
            e2[d] = pos[3 * (size_t)c + d] - pos[3 * (size_t)a + d];
            w[d] = pos[3 * (size_t)o + d] - pos[3 * (size_t)a + d];
        }
        cross3f(e1, e2, n);
        if (dot3f(n, w) > 0.f) std::swap(b, c);
        faces.push_back(a);
        faces.push_back(b);
        faces.push_back(c);
    }
    return true;
}

bool prepare_cloth_mesh(ClothMeshHost& m, std::string& err) {
    if ((int64_t)m.pos.size() != 3ll * m.nv || (int64_t)m.faces.size() != 3ll * m.nf) {
        err = "cloth mesh: array sizes do not match vertex/face counts";
        return false;
    }
    for (int32_t v : m.faces)
        if (v < 0 || v >= m.nv) {
            err = "cloth mesh: vertex index out of range";
            return false;
        }
    for (int32_t f : m.fixed)
        if (f < 0 || f >= m.nv) {
            err = "cloth mesh: fixed point index out of range";
            return false;
        }
    // rest pose
    m.dminv.assign(4 * (size_t)m.nf, 0.f);
    m.area.assign(m.nf, 0.f);
    m.mass.assign(m.nv, 0.f);
    for (int32_t f = 0; f < m.nf; ++f) {
        const int32_t* fv = &m.faces[3 * (size_t)f];
        float e1[3], e2[3], n[3];
        for (int d = 0; d < 3; ++d) {
            e1[d] = m.pos[3 * (size_t)fv[1] + d] - m.pos[3 * (size_t)fv[0] + d];
            e2[d] = m.pos[3 * (size_t)fv[2] + d] - m.pos[3 * (size_t)fv[0] + d];
        }
        cross3f(e1, e2, n);
        const float nn = std::sqrt(dot3f(n, n)), l1 = std::sqrt(dot3f(e1, e1));
        if (!(nn > 0.f) || !(l1 > 0.f)) {
            err = "cloth mesh: degenerate face " + std::to_string(f);
            return false;
        }
        const float a1[3] = {e1[0] / l1, e1[1] / l1, e1[2] / l1};
        const float un[3] = {n[0] / nn, n[1] / nn, n[2] / nn};
        float a2[3];
        cross3f(un, a1, a2);
        const float Dm[4] = {dot3f(a1, e1), dot3f(a2, e1), dot3f(a1, e2), dot3f(a2, e2)};
        const float det = Dm[0] * Dm[3] - Dm[2] * Dm[1];
        m.area[f] = det * 0.5f;
        const float inv = 1.f / det;
        float* Di = &m.dminv[4 * (size_t)f];
        Di[0] = Dm[3] * inv;
        Di[1] = -Dm[1] * inv;
        Di[2] = -Dm[2] * inv;
        Di[3] = Dm[0] * inv;
        for (int k = 0; k < 3; ++k) m.mass[fv[k]] += m.mat.density * m.area[f] / 3.f;
    }
    // edges, numbered by first appearance
    std::vector<int32_t> faceEdges;
    if (!build_tri_edges(m.nv, m.nf, m.faces.data(), m.edges, &faceEdges, err)) return false;
    // vertex -> faces
    m.vfBegin.assign(m.nv + 1, 0);
    for (int32_t v : m.faces) m.vfBegin[v + 1]++;
    for (int32_t i = 0; i < m.nv; ++i) m.vfBegin[i + 1] += m.vfBegin[i];
    m.vfFace.assign(m.vfBegin[m.nv], 0);
    m.vfOrder.assign(m.vfBegin[m.nv], 0);
    {
        std::vector<int32_t> fill(m.vfBegin.begin(), m.vfBegin.end() - 1);
        for (int32_t f = 0; f < m.nf; ++f)
            for (int k = 0; k < 3; ++k) {
                const int32_t v = m.faces[3 * (size_t)f + k];
                m.vfFace[fill[v]] = f;
                m.vfOrder[fill[v]] = k;
                fill[v]++;
            }
    }
    // vertex -> incident edges (ascending edge id)
    std::vector<int32_t> veBegin(m.nv + 1, 0), veEdge(2 * m.edges.size());
    for (const TriEdge& e : m.edges) {
        veBegin[e.v1 + 1]++;
        veBegin[e.v2 + 1]++;
    }
    for (int32_t i = 0; i < m.nv; ++i) veBegin[i + 1] += veBegin[i];
    {
        std::vector<int32_t> fill(veBegin.begin(), veBegin.end() - 1);
        for (int32_t id = 0; id < (int32_t)m.edges.size(); ++id) {
            veEdge[fill[m.edges[id].v1]++] = id;
            veEdge[fill[m.edges[id].v2]++] = id;
        }
    }
    m.vbBegin.assign(m.nv + 1, 0);
    m.vbEdge.clear();
    m.vbOrder.clear();
    for (int32_t v = 0; v < m.nv; ++v) {
        for (int32_t k = veBegin[v]; k < veBegin[v + 1]; ++k) {
            const TriEdge& e = m.edges[veEdge[k]];
            if (e.f2 == -1) continue;
            m.vbEdge.push_back(veEdge[k]);
            m.vbOrder.push_back(e.v1 == v ? 0 : 1);
        }
        for (int32_t k = m.vfBegin[v]; k < m.vfBegin[v + 1]; ++k) {
            const int32_t f = m.vfFace[k], o = m.vfOrder[k];
            const int32_t id = faceEdges[3 * (size_t)f + (o + 1) % 3];  // edge (o+1, o+2) of the face
            const TriEdge& e = m.edges[id];
            if (e.f2 == -1) continue;
            m.vbEdge.push_back(id);
            m.vbOrder.push_back(e.f1 == f ? 2 : 3);
        }
        m.vbBegin[v + 1] = (int32_t)m.vbEdge.size();
    }
    // bending quadratic forms
    m.Q.assign(16 * m.edges.size(), 0.f);
    auto cot = [](const float* a, const float* b) {
        float c[3];
        cross3f(a, b, c);
        return dot3f(a, b) / std::sqrt(dot3f(c, c));
    };
    auto face_normal_len = [&](int32_t f) {
        const int32_t* fv = &m.faces[3 * (size_t)f];
        float a[3], b[3], n[3];
        for (int d = 0; d < 3; ++d) {
            a[d] = m.pos[3 * (size_t)fv[1] + d] - m.pos[3 * (size_t)fv[0] + d];
            b[d] = m.pos[3 * (size_t)fv[2] + d] - m.pos[3 * (size_t)fv[0] + d];
        }
        cross3f(a, b, n);
        return std::sqrt(dot3f(n, n));
    };
    for (size_t id = 0; id < m.edges.size(); ++id) {
        const TriEdge& e = m.edges[id];
        if (e.f2 == -1) continue;
        const float* X0 = &m.pos[3 * (size_t)e.v1];
        const float* X1 = &m.pos[3 * (size_t)e.v2];
        const float* X2 = &m.pos[3 * (size_t)e.v12Next];
        const float* X3 = &m.pos[3 * (size_t)e.v21Next];
        float e01[3], e03[3], e02[3], e12[3], e13[3], me01[3];
        for (int d = 0; d < 3; ++d) {
            e01[d] = X1[d] - X0[d];
            e03[d] = X3[d] - X0[d];
            e02[d] = X2[d] - X0[d];
            e12[d] = X2[d] - X1[d];
            e13[d] = X3[d] - X1[d];
            me01[d] = -e01[d];
        }
        const float c01 = cot(e02, e01), c02 = cot(e01, e03), c03 = cot(me01, e12), c04 = cot(me01, e13);
        const float K0[4] = {c03 + c04, c01 + c02, -c01 - c03, -c02 - c04};
        const float A = face_normal_len(e.f1) + face_normal_len(e.f2);
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) m.Q[16 * id + 4 * i + j] = 3.f * K0[i] * K0[j] / A;
    }
    return true;
}

Graph tri_vertex_graph(int32_t nv, int32_t nf, const int32_t* faces, bool withBending) {
    std::vector<std::pair<uint64_t, int32_t>> occ;  // (edge key, opposite vertex)
    occ.reserve(3 * (size_t)nf);
    for (int32_t f = 0; f < nf; ++f)
        for (int k = 0; k < 3; ++k)
            occ.push_back({EdgeTable::key(faces[3 * (size_t)f + k], faces[3 * (size_t)f + (k + 1) % 3]),
                           faces[3 * (size_t)f + (k + 2) % 3]});
    std::sort(occ.begin(), occ.end());
    std::vector<int32_t> edges;
    for (size_t i = 0; i < occ.size();) {
        size_t j = i;
        while (j < occ.size() && occ[j].first == occ[i].first) ++j;
        edges.push_back((int32_t)(occ[i].first >> 32));
        edges.push_back((int32_t)(occ[i].first & 0xffffffffu));
        if (withBending && j - i == 2) {
            edges.push_back(occ[i].second);
            edges.push_back(occ[i + 1].second);
        }
        i = j;
    }
    return graph_from_edges(nv, edges.data(), (int32_t)(edges.size() / 2));
}

bool load_obj_file(const std::string& path, std::vector<float>& pos, std::vector<int32_t>& faces, std::string& err) {
    std::ifstream in(path);
    if (!in) {
        err = "cannot open " + path;
        return false;
    }
    pos.clear();
    faces.clear();
    std::string line, tag;
    while (std::getline(in, line)) {
        std::istringstream ss(line);
        if (!(ss >> tag)) continue;
        if (tag == "v") {
            double x, y, z;
            if (!(ss >> x >> y >> z)) {
                err = "bad v line in " + path;
                return false;
            }
            pos.push_back((float)x);
            pos.push_back((float)y);
            pos.push_back((float)z);
        } else if (tag == "f") {
            std::vector<int32_t> ids;
            std::string tok;
            while (ss >> tok) ids.push_back((int32_t)std::strtol(tok.c_str(), nullptr, 10) - 1);
            if (ids.size() != 3) {
                err = "only triangle faces are supported in " + path;
                return false;
            }
            faces.insert(faces.end(), ids.begin(), ids.end());
        }
    }
    return true;
}

bool load_t_file(const std::string& path, std::vector<float>& pos, std::vector<int32_t>& tets, std::string& err) {
    std::ifstream in(path);
    if (!in) {
        err = "cannot open " + path;
        return false;
    }
    pos.clear();
    tets.clear();
    std::string line, tag;
    while (std::getline(in, line)) {
        std::istringstream ss(line);
        if (!(ss >> tag)) continue;
        if (tag == "Vertex") {
            long id;
            double x, y, z;
            if (!(ss >> id >> x >> y >> z)) {
                err = "bad Vertex line in " + path;
                return false;
            }
            if (id != (long)(pos.size() / 3)) {
                err = "non-consecutive vertex ids in " + path;
                return false;
            }
            pos.push_back((float)x);
            pos.push_back((float)y);
            pos.push_back((float)z);
        } else if (tag == "Tet") {
            long id, a, b, c, d;
            if (!(ss >> id >> a >> b >> c >> d)) {
                err = "bad Tet line in " + path;
                return false;
            }
            tets.push_back((int32_t)a);
            tets.push_back((int32_t)b);
            tets.push_back((int32_t)c);
            tets.push_back((int32_t)d);
        }
    }
    return true;
}

}  // namespace gvbd
