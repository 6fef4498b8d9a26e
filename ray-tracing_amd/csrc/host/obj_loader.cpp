// Wavefront OBJ loading as load_obj does it (obj_loader.h:10-141): "v" and
// "vn" records, "f" records with 1-based or negative (relative) indices in
// the forms p, p/t, p/t/n and p//n, polygons fanned around their first
// corner, zero-area triangles dropped, one normal per corner (the file's,
// or the face normal when the corner has none), then every vertex moved so
// the bounding box of the referenced vertices is centred on the origin and
// scaled so its longest side is 1.  Every float expression keeps the
// reference's operation order (the host code is built without contraction).
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "host.hpp"

namespace lrt {
namespace {

struct V3 {
    float x, y, z;
};
V3 sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 cross3(V3 a, V3 b) { return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
float len2(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }

// the next whitespace-separated token of a line, as istream >> string reads it
bool next_token(const char*& s, std::string& tok) {
    while (*s == ' ' || *s == '\t' || *s == '\v' || *s == '\f' || *s == '\r' || *s == '\n') ++s;
    if (!*s) return false;
    const char* b = s;
    while (*s && !(*s == ' ' || *s == '\t' || *s == '\v' || *s == '\f' || *s == '\r' || *s == '\n')) ++s;
    tok.assign(b, (size_t)(s - b));
    return true;
}

// istream >> float x3 into values that start at 0: a failed extraction sets
// its value to 0 and stops the rest
void read3(const char* s, float out[3]) {
    out[0] = out[1] = out[2] = 0.0f;
    std::string tok;
    for (int k = 0; k < 3; ++k) {
        if (!next_token(s, tok)) return;
        char* end = nullptr;
        errno = 0;
        const float v = std::strtof(tok.c_str(), &end);
        if (end == tok.c_str()) { out[k] = 0.0f; return; }
        out[k] = v;
        if (*end) return;  // trailing characters: the next extraction starts there and fails
    }
}

// std::stoi: leading whitespace, sign, digits; false where it would throw
bool stoi_like(const std::string& s, int& v) {
    const char* c = s.c_str();
    char* end = nullptr;
    errno = 0;
    const long r = std::strtol(c, &end, 10);
    if (end == c || errno == ERANGE || r > 2147483647L || r < -2147483647L - 1L) return false;
    v = (int)r;
    return true;
}

int obj_index(int index, int count) { return index < 0 ? count + index : index - 1; }

}  // namespace

bool load_obj(const char* path, std::vector<float>& positions, std::vector<uint32_t>& indices,
              std::vector<float>& corner_normals, std::string& error) {
    positions.clear();
    indices.clear();
    corner_normals.clear();
    FILE* f = std::fopen(path, "rb");
    if (!f) {
        error = std::string("cannot open ") + path;
        return false;
    }
    std::string data;
    char buf[1 << 16];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof(buf), f)) > 0) data.append(buf, got);
    std::fclose(f);

    std::vector<V3> pos, nrm;
    std::vector<uint32_t> idx;
    std::vector<V3> cn;
    std::vector<uint32_t> face_p;
    std::vector<int> face_n;
    std::string line, tag, corner;
    size_t at = 0;
    while (at < data.size()) {
        size_t eol = data.find('\n', at);
        if (eol == std::string::npos) eol = data.size();
        line.assign(data, at, eol - at);
        at = eol + 1;
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (line.empty() || line[0] == '#') continue;
        const char* s = line.c_str();
        if (!next_token(s, tag)) continue;
        if (tag == "v" || tag == "vn") {
            float v[3];
            read3(s, v);
            (tag == "v" ? pos : nrm).push_back(V3{v[0], v[1], v[2]});
        } else if (tag == "f") {
            face_p.clear();
            face_n.clear();
            while (next_token(s, corner)) {
                int pi = 0, ni = 0;
                const size_t first = corner.find('/');
                if (first == std::string::npos) {
                    if (!stoi_like(corner, pi)) { error = "bad face index: " + corner; return false; }
                } else {
                    if (!stoi_like(corner.substr(0, first), pi)) { error = "bad face index: " + corner; return false; }
                    const size_t second = corner.find('/', first + 1);
                    if (second != std::string::npos && second + 1 < corner.size())
                        if (!stoi_like(corner.substr(second + 1), ni)) { error = "bad normal index: " + corner; return false; }
                }
                const int p = obj_index(pi, (int)pos.size());
                if (p < 0 || p >= (int)pos.size()) { error = "face index out of range: " + corner; return false; }
                face_p.push_back((uint32_t)p);
                if (ni == 0) {
                    face_n.push_back(-1);
                } else {
                    const int n = obj_index(ni, (int)nrm.size());
                    if (n < 0 || n >= (int)nrm.size()) { error = "normal index out of range: " + corner; return false; }
                    face_n.push_back(n);
                }
            }
            if (face_p.size() < 3) { error = "face with fewer than 3 corners"; return false; }
            for (size_t i = 1; i + 1 < face_p.size(); ++i) {
                const uint32_t tri[3] = {face_p[0], face_p[i], face_p[i + 1]};
                const int tn[3] = {face_n[0], face_n[i], face_n[i + 1]};
                const V3 g = cross3(sub(pos[tri[1]], pos[tri[0]]), sub(pos[tri[2]], pos[tri[0]]));
                if (len2(g) <= 0.0f) continue;
                const float l = std::sqrt(len2(g));
                const V3 unit{g.x / l, g.y / l, g.z / l};
                for (int k = 0; k < 3; ++k) {
                    idx.push_back(tri[k]);
                    cn.push_back(tn[k] < 0 ? unit : nrm[(size_t)tn[k]]);
                }
            }
        }
    }
    if (idx.empty()) {
        error = "no triangles";
        return false;
    }
    // center_and_unitize (obj_loader.h:15-41)
    V3 lo = pos[idx[0]], hi = lo;
    for (uint32_t i : idx) {
        const V3& p = pos[i];
        lo.x = std::min(lo.x, p.x); lo.y = std::min(lo.y, p.y); lo.z = std::min(lo.z, p.z);
        hi.x = std::max(hi.x, p.x); hi.y = std::max(hi.y, p.y); hi.z = std::max(hi.z, p.z);
    }
    const V3 c{(lo.x + hi.x) * 0.5f, (lo.y + hi.y) * 0.5f, (lo.z + hi.z) * 0.5f};
    const V3 e = sub(hi, lo);
    const float longest = std::max(e.x, std::max(e.y, e.z));
    if (longest <= 0.0f) {
        error = "degenerate bounding box";
        return false;
    }
    const float scale = 1.0f / longest;
    positions.resize(pos.size() * 3);
    for (size_t i = 0; i < pos.size(); ++i) {
        const V3 d = sub(pos[i], c);
        positions[3 * i] = d.x * scale;
        positions[3 * i + 1] = d.y * scale;
        positions[3 * i + 2] = d.z * scale;
    }
    indices = std::move(idx);
    corner_normals.resize(cn.size() * 3);
    for (size_t i = 0; i < cn.size(); ++i) {
        corner_normals[3 * i] = cn[i].x;
        corner_normals[3 * i + 1] = cn[i].y;
        corner_normals[3 * i + 2] = cn[i].z;
    }
    return true;
}

}  // namespace lrt
