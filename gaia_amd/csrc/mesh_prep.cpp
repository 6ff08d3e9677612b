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
    const std::vector<const std::vector<std::vector<int32_t>>*>& meshCategories, bool removeUsed) {
    size_t ng = 0;
    for (auto* c : meshCategories) ng = std::max(ng, c->size());
    std::vector<std::vector<int32_t>> groups(ng);
    for (size_t im = 0; im < meshCategories.size(); ++im) {
        const auto& cats = *meshCategories[im];
        std::vector<int32_t> avail(ng);
        for (size_t i = 0; i < ng; ++i) avail[i] = (int32_t)i;
        // tet meshes: the available list starts with the mesh's own colour count
        // (VBDPhysics.cpp:245-252); groups beyond it are never chosen.
        if (removeUsed) avail.resize(cats.size());
        for (size_t c = 0; c < cats.size(); ++c) {
            size_t best = groups[avail[0]].size();
            size_t bestI = 0;
            for (size_t i = 1; i < avail.size(); ++i)
                if (groups[avail[i]].size() < best) {
                    best = groups[avail[i]].size();
                    bestI = i;
                }
            const int32_t gid = avail[bestI];
            if (removeUsed) avail.erase(avail.begin() + bestI);
            for (int32_t v : cats[c]) {
                groups[gid].push_back((int32_t)im);
                groups[gid].push_back(v);
            }
        }
    }
    return groups;
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
