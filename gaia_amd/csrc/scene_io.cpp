// Recovery-state and frame-output files in the reference's formats (see scene_io.h).
#include "scene_io.h"

#include <cstdio>
#include <fstream>
#include <sstream>

#include "errors.h"
#include "json_lite.h"

namespace gvbd {
namespace io {
namespace {

// directory part (with the trailing separator) and stem of a path
void split_path(const std::string& path, std::string& dir, std::string& stem) {
    const size_t slash = path.find_last_of("/\\");
    dir = slash == std::string::npos ? std::string() : path.substr(0, slash + 1);
    std::string name = slash == std::string::npos ? path : path.substr(slash + 1);
    const size_t dot = name.find_last_of('.');
    stem = dot == std::string::npos ? name : name.substr(0, dot);
}

// a float value as the JSON number of its double (round-trip digits)
void put_number(std::string& out, float v) {
    char buf[40];
    std::snprintf(buf, sizeof(buf), "%.17g", (double)v);
    out += buf;
}

void put_triples(std::string& out, const std::vector<float>& a, const std::string& pad) {
    out += "[";
    const size_t n = a.size() / 3;
    for (size_t i = 0; i < n; ++i) {
        out += i ? ",\n" : "\n";
        out += pad + "  [\n";
        for (int d = 0; d < 3; ++d) {
            out += pad + "    ";
            put_number(out, a[3 * i + d]);
            out += d < 2 ? ",\n" : "\n";
        }
        out += pad + "  ]";
    }
    out += n ? "\n" + pad + "]" : "]";
}

void write_file(const std::string& path, const std::string& text) {
    std::ofstream f(path, std::ios::binary);
    if (!f) fail(GVBD_ERR_IO, "cannot open " + path + " for writing");
    f << text;
    if (!f) fail(GVBD_ERR_IO, "write failed: " + path);
}

void write_raw(const std::string& path, const std::vector<const std::vector<float>*>& parts) {
    std::ofstream f(path, std::ios::binary);
    if (!f) fail(GVBD_ERR_IO, "cannot open " + path + " for writing");
    for (const auto* p : parts) f.write(reinterpret_cast<const char*>(p->data()), sizeof(float) * p->size());
    if (!f) fail(GVBD_ERR_IO, "write failed: " + path);
}

const json::Value& need(const json::Value& v, const char* key, const std::string& path) {
    const json::Value* x = v.get(key);
    if (!x) fail(GVBD_ERR_IO, std::string("recovery state ") + path + " has no \"" + key + "\"");
    return *x;
}

void read_triples(const json::Value& v, std::vector<float>& out, const std::string& path) {
    if (v.type != json::Value::Array) fail(GVBD_ERR_IO, "recovery state " + path + ": expected an array of triples");
    out.resize(3 * v.arr.size());
    for (size_t i = 0; i < v.arr.size(); ++i) {
        const json::Value& t = v.arr[i];
        if (t.type != json::Value::Array || t.arr.size() != 3)
            fail(GVBD_ERR_IO, "recovery state " + path + ": expected [x, y, z]");
        for (int d = 0; d < 3; ++d) {
            if (t.arr[d].type != json::Value::Number) fail(GVBD_ERR_IO, "recovery state " + path + ": non-number");
            out[3 * i + d] = (float)t.arr[d].num;
        }
    }
}

}  // namespace

void write_state(const std::string& path, int32_t frameId, double curTime, bool binary,
                 const std::vector<MeshState>& meshes) {
    std::string out = "{\n  \"curTime\": ";
    char buf[40];
    std::snprintf(buf, sizeof(buf), "%.17g", curTime);
    out += buf;
    out += ",\n  \"format\": \"";
    out += binary ? "binary" : "text";
    out += "\",\n  \"frameId\": " + std::to_string(frameId) + ",\n  \"meshesState\": ";
    if (binary) {
        std::string dir, stem;
        split_path(path, dir, stem);
        const std::string posFile = stem + ".pos", velFile = stem + ".vel";
        std::vector<const std::vector<float>*> pos, vel;
        std::string counts;
        for (size_t m = 0; m < meshes.size(); ++m) {
            if (meshes[m].velocities.size() != meshes[m].position.size())
                fail(GVBD_ERR_INVALID, "binary recovery state needs velocities for every mesh");
            pos.push_back(&meshes[m].position);
            vel.push_back(&meshes[m].velocities);
            counts += (m ? ",\n      " : "\n      ") + std::to_string(meshes[m].position.size() / 3);
        }
        write_raw(dir + posFile, pos);
        write_raw(dir + velFile, vel);
        out += "{\n    \"nVerts\": [" + counts + (meshes.empty() ? "]" : "\n    ]") + ",\n    \"position_file\": \"" +
               posFile + "\",\n    \"velocities_file\": \"" + velFile + "\"\n  },\n  \"nMeshes\": " +
               std::to_string(meshes.size()) + "\n}";
    } else {
        out += "[";
        for (size_t m = 0; m < meshes.size(); ++m) {
            out += m ? ",\n    {\n      \"position\": " : "\n    {\n      \"position\": ";
            put_triples(out, meshes[m].position, "      ");
            if (!meshes[m].velocities.empty()) {
                out += ",\n      \"velocities\": ";
                put_triples(out, meshes[m].velocities, "      ");
            }
            out += "\n    }";
        }
        out += meshes.empty() ? "]\n}" : "\n  ]\n}";
    }
    write_file(path, out);
}

void read_state(const std::string& path, int32_t& frameId, double& curTime, std::vector<MeshState>& meshes) {
    std::ifstream f(path, std::ios::binary);
    if (!f) fail(GVBD_ERR_IO, "cannot open recovery state " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    json::Value root;
    std::string err;
    if (!json::Parser(ss.str()).parse(root, err)) fail(GVBD_ERR_IO, path + ": " + err);
    if (root.type != json::Value::Object) fail(GVBD_ERR_IO, path + ": not a JSON object");
    frameId = (int32_t)need(root, "frameId", path).num;
    curTime = need(root, "curTime", path).num;
    std::string format = "text";
    if (const json::Value* fm = root.get("format")) format = fm->str;
    meshes.clear();
    const json::Value& ms = need(root, "meshesState", path);
    if (format == "binary") {
        const int nMeshes = (int)need(root, "nMeshes", path).num;
        const json::Value& nv = need(ms, "nVerts", path);
        if (nv.type != json::Value::Array || (int)nv.arr.size() < nMeshes)
            fail(GVBD_ERR_IO, "recovery state " + path + ": nVerts shorter than nMeshes");
        std::string dir, stem;
        split_path(path, dir, stem);
        std::ifstream pf(dir + need(ms, "position_file", path).str, std::ios::binary);
        std::ifstream vf(dir + need(ms, "velocities_file", path).str, std::ios::binary);
        if (!pf || !vf) fail(GVBD_ERR_IO, "recovery state " + path + ": cannot open its .pos / .vel files");
        meshes.resize(nMeshes);
        for (int m = 0; m < nMeshes; ++m) {
            const size_t n = 3 * (size_t)nv.arr[m].num;
            meshes[m].position.resize(n);
            meshes[m].velocities.resize(n);
            pf.read(reinterpret_cast<char*>(meshes[m].position.data()), sizeof(float) * n);
            vf.read(reinterpret_cast<char*>(meshes[m].velocities.data()), sizeof(float) * n);
            if (!pf || !vf) fail(GVBD_ERR_IO, "recovery state " + path + ": .pos / .vel files too short");
        }
    } else if (format == "text") {
        if (ms.type != json::Value::Array) fail(GVBD_ERR_IO, "recovery state " + path + ": meshesState is not a list");
        meshes.resize(ms.arr.size());
        for (size_t m = 0; m < ms.arr.size(); ++m) {
            read_triples(need(ms.arr[m], "position", path), meshes[m].position, path);
            if (const json::Value* v = ms.arr[m].get("velocities")) {
                read_triples(*v, meshes[m].velocities, path);
                if (meshes[m].velocities.size() != meshes[m].position.size())
                    fail(GVBD_ERR_IO, "recovery state " + path + ": velocities and positions differ in length");
            }
        }
    } else {
        fail(GVBD_ERR_IO, "recovery state " + path + ": unknown format \"" + format + "\"");
    }
}

void write_ply(const std::string& path, const std::vector<float>& verts, const std::vector<int32_t>& faces) {
    FILE* fp = std::fopen(path.c_str(), "w");
    if (!fp) fail(GVBD_ERR_IO, "cannot open " + path + " for writing");
    const size_t nv = verts.size() / 3, nf = faces.size() / 3;
    std::fprintf(fp, "ply\nformat ascii 1.0\ncomment Mocap generated\nelement vertex %zu\n", nv);
    std::fprintf(fp, "property float x\nproperty float y\nproperty float z\n");
    if (nf) std::fprintf(fp, "element face %zu \nproperty list uchar int vertex_indices\n", nf);
    std::fprintf(fp, "end_header\n");
    for (size_t i = 0; i < nv; ++i) std::fprintf(fp, "%f %f %f\n", verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]);
    for (size_t i = 0; i < nf; ++i) std::fprintf(fp, "3 %d %d %d\n", faces[3 * i], faces[3 * i + 1], faces[3 * i + 2]);
    const bool ok = std::ferror(fp) == 0;
    std::fclose(fp);
    if (!ok) fail(GVBD_ERR_IO, "write failed: " + path);
}

void write_binary(const std::string& path, const std::vector<float>& verts) { write_raw(path, {&verts}); }

}  // namespace io
}  // namespace gvbd
