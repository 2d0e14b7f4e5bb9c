// hs_mesh.cpp — host mesh module: the reference's `.in` reader and nParts = 1 cell /
// region typing, kept on the host in front of the C ABI.
//
// Grammar (mesh.cpp:1616-1893, CYCASMesh::readMesh mesh.cpp:2384-2488):
//   sections start at the first occurrence of the keywords xgrid, ygrid, zgrid, mobs,
//   walls, holes and run to the next keyword position; text after '=' is read line by
//   line with everything from ';' to the end of the line removed (sliceInputString
//   mesh.cpp:27-40); each line yields `value separator` pairs, so a value needs a
//   trailing separator character to count (quirk Q24).  mobs lines are 1-based inclusive
//   cell boxes i0,i1,j0,j1,k0,k1 (typing mesh.cpp:102-110); walls lines need 8 ints,
//   holes lines 13 (parsed, inert as in the reference).
#include <algorithm>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/hsolver.h"
#include "../../include/hsolver_mesh.h"

struct hs_mesh {
    std::vector<double> x, y, z;
    std::vector<std::vector<int>> mobs, walls, holes;
    std::vector<uint8_t> cell_type;
    int nfluid_regions = 0, nsolid_regions = 0;
};

namespace {

std::string slice_comments(const std::string& line) {
    std::string out;
    bool skip = false;
    for (char ch : line) {
        if (ch == ';') skip = true;
        else if (ch == '\n') skip = false;
        else if (!skip) out += ch;
    }
    return out;
}

template <typename T>
void read_pairs(const std::string& text, std::vector<T>& flat) {
    std::istringstream is(text);
    std::string line;
    while (std::getline(is, line)) {
        std::istringstream ls(slice_comments(line));
        T v;
        char sep;
        while (ls >> v >> sep) flat.push_back(v);
    }
}

template <typename T>
void read_rows(const std::string& text, std::vector<std::vector<T>>& rows, size_t need) {
    std::istringstream is(text);
    std::string line;
    while (std::getline(is, line)) {
        std::istringstream ls(slice_comments(line));
        std::vector<T> r;
        T v;
        char sep;
        while (ls >> v >> sep) r.push_back(v);
        if (!r.empty() && (need == 0 || r.size() == need)) rows.push_back(r);
    }
}

// text of a section: from after the '=' that follows its keyword to the next keyword
std::string section(const std::string& content, size_t pos, const std::vector<size_t>& all) {
    if (pos == std::string::npos) return "";
    size_t end = content.size();
    for (size_t p : all)
        if (p != std::string::npos && p > pos && p < end) end = p;
    const size_t eq = content.find('=', pos);
    if (eq == std::string::npos || eq >= end) return "";
    return content.substr(eq + 1, end - eq - 1);
}

// 6-connected components of one cell class, honouring wall faces as in
// getConnectedRegions mesh.cpp:298-454 (a wall face cuts the adjacency).
int count_regions(const hs_mesh& m, uint8_t cls, const std::vector<uint8_t>& xwall,
                  const std::vector<uint8_t>& ywall, const std::vector<uint8_t>& zwall) {
    const long nx = (long)m.x.size() - 1, ny = (long)m.y.size() - 1, nz = (long)m.z.size() - 1;
    std::vector<int> lab(nx * ny * nz, -1);
    std::vector<long> st;
    int n = 0;
    auto id = [&](long i, long j, long k) { return i + nx * (j + ny * k); };
    for (long c0 = 0; c0 < nx * ny * nz; ++c0) {
        if (m.cell_type[c0] != cls || lab[c0] >= 0) continue;
        st.assign(1, c0);
        lab[c0] = n;
        while (!st.empty()) {
            const long c = st.back();
            st.pop_back();
            const long i = c % nx, j = (c / nx) % ny, k = c / (nx * ny);
            auto go = [&](long a, long b, long d, bool cut) {
                if (a < 0 || a >= nx || b < 0 || b >= ny || d < 0 || d >= nz || cut) return;
                const long q = id(a, b, d);
                if (m.cell_type[q] == cls && lab[q] < 0) { lab[q] = n; st.push_back(q); }
            };
            go(i - 1, j, k, xwall[i + (nx + 1) * (j + ny * k)]);
            go(i + 1, j, k, xwall[i + 1 + (nx + 1) * (j + ny * k)]);
            go(i, j - 1, k, ywall[i + nx * (j + (ny + 1) * k)]);
            go(i, j + 1, k, ywall[i + nx * (j + 1 + (ny + 1) * k)]);
            go(i, j, k - 1, zwall[i + nx * (j + ny * k)]);
            go(i, j, k + 1, zwall[i + nx * (j + ny * (k + 1))]);
        }
        ++n;
    }
    return n;
}

}  // namespace

extern "C" {

int hs_mesh_read(const char* path, hs_mesh** out) {
    if (!path || !out) return HS_E_ARG;
    std::ifstream f(path);
    if (!f.is_open()) return HS_E_IO;
    std::ostringstream ss;
    ss << f.rdbuf();
    const std::string content = ss.str();
    return hs_mesh_parse(content.c_str(), out);
}

int hs_mesh_parse(const char* text, hs_mesh** out) {
    if (!text || !out) return HS_E_ARG;
    const std::string content(text);
    const char* keys[6] = {"xgrid", "ygrid", "zgrid", "mobs", "walls", "holes"};
    std::vector<size_t> pos(6);
    for (int i = 0; i < 6; ++i) pos[i] = content.find(keys[i]);
    if (pos[0] == std::string::npos || pos[1] == std::string::npos || pos[2] == std::string::npos) return HS_E_IO;
    hs_mesh* m = new hs_mesh();
    read_pairs(section(content, pos[0], pos), m->x);
    read_pairs(section(content, pos[1], pos), m->y);
    read_pairs(section(content, pos[2], pos), m->z);
    read_rows(section(content, pos[3], pos), m->mobs, 0);
    read_rows(section(content, pos[4], pos), m->walls, 8);
    read_rows(section(content, pos[5], pos), m->holes, 13);
    const long nx = (long)m->x.size() - 1, ny = (long)m->y.size() - 1, nz = (long)m->z.size() - 1;
    if (nx < 1 || ny < 1 || nz < 1) { delete m; return HS_E_IO; }
    // initCellsType mesh.cpp:84-114
    m->cell_type.assign(nx * ny * nz, 0);
    for (long k = 1; k <= nz; ++k)
        for (long j = 1; j <= ny; ++j)
            for (long i = 1; i <= nx; ++i)
                for (const auto& b : m->mobs) {
                    if (b.size() < 6) continue;
                    if (i >= b[0] && i <= b[1] && j >= b[2] && j <= b[3] && k >= b[4] && k <= b[5]) {
                        m->cell_type[(i - 1) + nx * ((j - 1) + ny * (k - 1))] = 1;
                        break;
                    }
                }
    // wall planes (initFacesType mesh.cpp:168-237): only cut region adjacency here
    std::vector<uint8_t> xw((nx + 1) * ny * nz, 0), yw(nx * (ny + 1) * nz, 0), zw(nx * ny * (nz + 1), 0);
    for (const auto& w : m->walls) {
        const int i0 = w[0], i1 = w[1], j0 = w[2], j1 = w[3], k0 = w[4], k1 = w[5];
        if (i0 == i1 && i0 >= 1 && i0 <= nx + 1)
            for (long k = std::max(1, k0); k <= std::min<long>(nz, k1); ++k)
                for (long j = std::max(1, j0); j <= std::min<long>(ny, j1); ++j) xw[(i0 - 1) + (nx + 1) * ((j - 1) + ny * (k - 1))] = 1;
        if (j0 == j1 && j0 >= 1 && j0 <= ny + 1)
            for (long k = std::max(1, k0); k <= std::min<long>(nz, k1); ++k)
                for (long i = std::max(1, i0); i <= std::min<long>(nx, i1); ++i) yw[(i - 1) + nx * ((j0 - 1) + (ny + 1) * (k - 1))] = 1;
        if (k0 == k1 && k0 >= 1 && k0 <= nz + 1)
            for (long j = std::max(1, j0); j <= std::min<long>(ny, j1); ++j)
                for (long i = std::max(1, i0); i <= std::min<long>(nx, i1); ++i) zw[(i - 1) + nx * ((j - 1) + ny * (k0 - 1))] = 1;
    }
    m->nfluid_regions = count_regions(*m, 0, xw, yw, zw);
    m->nsolid_regions = count_regions(*m, 1, xw, yw, zw);
    *out = m;
    return HS_OK;
}

int hs_mesh_info(const hs_mesh* m, int32_t* nx, int32_t* ny, int32_t* nz, int32_t* nmobs, int32_t* nwalls,
                 int32_t* nfluid_regions, int32_t* nsolid_regions) {
    if (!m) return HS_E_ARG;
    if (nx) *nx = (int32_t)m->x.size() - 1;
    if (ny) *ny = (int32_t)m->y.size() - 1;
    if (nz) *nz = (int32_t)m->z.size() - 1;
    if (nmobs) *nmobs = (int32_t)m->mobs.size();
    if (nwalls) *nwalls = (int32_t)m->walls.size();
    if (nfluid_regions) *nfluid_regions = m->nfluid_regions;
    if (nsolid_regions) *nsolid_regions = m->nsolid_regions;
    return HS_OK;
}

int hs_mesh_grid(const hs_mesh* m, hs_grid* g) {
    if (!m || !g) return HS_E_ARG;
    g->nx = (int32_t)m->x.size() - 1;
    g->ny = (int32_t)m->y.size() - 1;
    g->nz = (int32_t)m->z.size() - 1;
    g->xv = m->x.data();
    g->yv = m->y.data();
    g->zv = m->z.data();
    g->cell_type = m->cell_type.data();
    return HS_OK;
}

int hs_mesh_box(const hs_mesh* m, int32_t which, int32_t* box6) {
    if (!m || !box6 || which < 0 || which >= (int32_t)m->mobs.size() || m->mobs[which].size() < 6) return HS_E_ARG;
    for (int i = 0; i < 6; ++i) box6[i] = m->mobs[which][i];
    return HS_OK;
}

void hs_mesh_free(hs_mesh* m) { delete m; }

}  // extern "C"
