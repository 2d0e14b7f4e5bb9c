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
//
// Cylindrical y form (mesh.cpp:1644-1700, 2413-2450): with no `ygrid` keyword the y
// section is `yl(1)=a, ..., nyr(1)=n, ..., yl(2)=b,`, each value read as std::stoi reads
// the text between the key and the next comma.  The n points are (b - a) / n * i for
// i = 0..n-1 in integer arithmetic (quirk Q25: the step truncates and the points start at
// 0, not at a).  b - a = 360 makes y periodic (perioFlag, mesh.cpp:66-69, 145-165,
// 322-327): n cells instead of n - 1, no y wall faces, and region adjacency wraps from
// row 1 to row n.  The face numbering keeps the reference's formulas with NyP = n, so the
// north face of row n is the south face of row 1 one plane up (quirk Q26, mesh.h:213-215),
// which the processor files inherit.  The solver-side view appends one point, so row n
// has the width of the others (the last-cell width fallback, fluid_solver.h:694-697); the
// solver's y boundaries stay walls (fluid_solver.h has no periodic y condition).
//
// processor0/ files (the decomposer at one part and the solver-side reader):
//   writer  decomposeMesh mesh.cpp:735-1581 with nParts = 1, writeProcFile :2277-2374
//   reader  CYCASMesh::readCells / readFaces / readBoundaries mesh.cpp:2507-2642
#include <sys/stat.h>

#include <algorithm>
#include <cerrno>
#include <stdexcept>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/hsolver.h"
#include "../../include/hsolver_mesh.h"

// One class's share of a processor file set: the cell list, the face list (internal
// faces first, then each boundary group's faces by global face id) and the groups.
struct ProcFace {
    int64_t gid, owner, nb;  // global face id; owner / neighbour = positions in `cells`
};
struct ProcBound {
    std::string name;
    int64_t start, n;
};
struct ProcPart {
    bool present = false;
    int region = 0, group = 0;
    std::vector<int64_t> cells;  // 1-based global cell indices
    std::vector<ProcFace> faces;
    std::vector<ProcBound> bounds;
};

struct hs_mesh {
    std::vector<double> x, y, z;  // vertices as the solver takes them (n + 1 per direction)
    int y_span = 0;               // yl(2) - yl(1) of the cylindrical form, 0 for ygrid
    bool periodic_y = false;      // y_span == 360
    std::vector<std::vector<int>> mobs, walls, holes;
    std::vector<uint8_t> cell_type;
    // per global face (x faces, then y faces, then z faces, x fastest in each): 1 where a
    // walls line makes the face a wall
    std::vector<uint8_t> wall_line;
    int nfluid_regions = 0, nsolid_regions = 0;
    bool have_parts = false;  // part[] read from processor files
    ProcPart part[2];         // fluid, mobs
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

// Global face numbering of the structured grid (mesh.h:209-219, 0-based): the x faces
// (i = 1..nx+1) first, then the y faces, then the z faces, x fastest in each.  Cells are
// numbered from 1, x fastest (mesh.h:201-203).
// With periodic y the reference counts NyP = Ny y-face planes (setDims mesh.cpp:60-76).
struct FaceNum {
    long nx, ny, nz, nyp, nxf, nyf, nzf;
    bool periodic;
    explicit FaceNum(const hs_mesh& m)
        : nx((long)m.x.size() - 1), ny((long)m.y.size() - 1), nz((long)m.z.size() - 1),
          nyp(m.periodic_y ? ny : ny + 1), nxf((nx + 1) * ny * nz), nyf(nx * nyp * nz), nzf(nx * ny * (nz + 1)),
          periodic(m.periodic_y) {}
    long ncells() const { return nx * ny * nz; }
    long nfaces() const { return nxf + nyf + nzf; }
    long fx(long i, long j, long k) const { return (i - 1) + (nx + 1) * ((j - 1) + ny * (k - 1)); }
    long fy(long i, long j, long k) const { return nxf + (i - 1) + nx * ((j - 1) + nyp * (k - 1)); }
    long fz(long i, long j, long k) const { return nxf + nyf + (i - 1) + nx * ((j - 1) + ny * (k - 1)); }
    // faces of 1-based cell c: west, east, south, north, bottom, top (getFaceIndexes mesh.h:245-253)
    void of_cell(long c, long f[6]) const {
        const long q = c - 1, i = q % nx + 1, j = (q / nx) % ny + 1, k = q / (nx * ny) + 1;
        f[0] = fx(i, j, k);
        f[1] = fx(i + 1, j, k);
        f[2] = fy(i, j, k);
        f[3] = fy(i, j + 1, k);
        f[4] = fz(i, j, k);
        f[5] = fz(i, j, k + 1);
    }
    // the domain's outer faces, typed Wall (initFacesType mesh.cpp:116-167); periodic y
    // types its first and last y-face planes Internal (:145-165)
    bool outer(long f) const {
        if (f < nxf) {
            const long i = f % (nx + 1);
            return i == 0 || i == nx;
        }
        if (f < nxf + nyf) {
            if (periodic) return false;
            const long j = ((f - nxf) / nx) % (ny + 1);
            return j == 0 || j == ny;
        }
        const long k = (f - nxf - nyf) / (nx * ny);
        return k == 0 || k == nz;
    }
};

// 6-connected components of one cell class, honouring wall faces as in
// getConnectedRegions mesh.cpp:298-454 (a wall face cuts the adjacency).  An edge is
// taken from a cell across that cell's own face, so with periodic y, where row n's north
// face is another cell's face, adjacency can hold in one direction only; the walk keeps
// that, and starts from the lowest unlabelled index as the reference's BFS does.
int count_regions(const hs_mesh& m, uint8_t cls) {
    const FaceNum fn(m);
    const long n = fn.ncells();
    std::vector<int> lab(n, -1);
    std::vector<long> st;
    int regions = 0;
    const long step[6] = {-1, 1, -fn.nx, fn.nx, -fn.nx * fn.ny, fn.nx * fn.ny};
    const long wrap = fn.nx * (fn.ny - 1);  // row 1 <-> row n (mesh.cpp:322-327)
    long f[6];
    for (long c0 = 0; c0 < n; ++c0) {
        if (m.cell_type[c0] != cls || lab[c0] >= 0) continue;
        st.assign(1, c0);
        lab[c0] = regions;
        while (!st.empty()) {
            const long c = st.back();
            st.pop_back();
            fn.of_cell(c + 1, f);
            const long j = (c / fn.nx) % fn.ny;
            for (int l = 0; l < 6; ++l) {
                if (fn.outer(f[l]) || m.wall_line[f[l]]) continue;
                long q = c + step[l];
                if (fn.periodic && l == 2 && j == 0) q = c + wrap;
                if (fn.periodic && l == 3 && j == fn.ny - 1) q = c - wrap;
                if (q < 0 || q >= n) continue;
                if (m.cell_type[q] == cls && lab[q] < 0) { lab[q] = regions; st.push_back(q); }
            }
        }
        ++regions;
    }
    return regions;
}

// decomposeMesh (mesh.cpp:735-1581) at nParts = 1 for one class, whose cells form one
// region: region 0, group 0, the cells in global order (:1195-1212).  A face touched by
// two of the part's cells and not a wall is internal, owned by the earlier cell in the
// list (:1255-1303).  The others are grouped by name (:1304-1372): "wall" for wall faces,
// "ib_0_0" for faces to a cell of the other class (processor ids are all 0 at one part).
// Groups follow in name order, each by global face id; a face listed twice in a group
// keeps its later owner and neighbour 0 (the map insertion and StFace default, :1374-1389,
// mesh.h:47-54).
void build_part(const hs_mesh& m, uint8_t cls, ProcPart& p) {
    const FaceNum fn(m);
    const long n = fn.ncells();
    p = ProcPart();
    for (long c = 1; c <= n; ++c)
        if (m.cell_type[c - 1] == cls) p.cells.push_back(c);
    if (p.cells.empty()) return;
    p.present = true;
    std::vector<int64_t> seen(fn.nfaces(), -1);
    long f[6];
    for (size_t k = 0; k < p.cells.size(); ++k) {
        fn.of_cell(p.cells[k], f);
        for (int l = 0; l < 6; ++l) {
            if (seen[f[l]] == -1) seen[f[l]] = 0;
            else if (seen[f[l]] == 0) seen[f[l]] = (int64_t)k;
        }
    }
    const long step[6] = {-1, 1, -fn.nx, fn.nx, -fn.nx * fn.ny, fn.nx * fn.ny};
    std::map<std::string, std::vector<ProcFace>> groups;
    for (size_t k = 0; k < p.cells.size(); ++k) {
        fn.of_cell(p.cells[k], f);
        for (int l = 0; l < 6; ++l) {
            const long g = f[l];
            const bool wall = fn.outer(g) || m.wall_line[g];
            if (seen[g] > 0 && !wall) {
                p.faces.push_back({g, (int64_t)k, seen[g]});
                seen[g] = -1;
            } else if (wall) {
                groups["wall"].push_back({g, (int64_t)k, 0});
            } else {
                const long nb = p.cells[k] + step[l];
                if (nb > 0 && nb <= n && m.cell_type[nb - 1] != cls) groups["ib_0_0"].push_back({g, (int64_t)k, 0});
            }
        }
    }
    for (const auto& kv : groups) {
        std::map<int64_t, ProcFace> by_id;
        for (const ProcFace& x : kv.second) by_id[x.gid] = x;
        p.bounds.push_back({kv.first, (int64_t)p.faces.size(), (int64_t)by_id.size()});
        for (const auto& e : by_id) p.faces.push_back(e.second);
    }
}

// One value of the cylindrical y form: the text after `key` up to the next comma, read as
// std::stoi reads it (mesh.cpp:1682-1694); false where the reference would throw.
bool cyl_value(const std::string& content, const char* key, int& v) {
    const size_t at = content.find(key);
    if (at == std::string::npos) return false;
    const size_t from = at + std::strlen(key);
    const size_t comma = content.find(',', at);
    try {
        v = std::stoi(content.substr(from, comma == std::string::npos ? std::string::npos : comma - from));
    } catch (const std::logic_error&) {
        return false;
    }
    return true;
}

std::string join(const char* dir, const char* leaf) {
    std::string s = (dir && *dir) ? dir : ".";
    return s + "/" + leaf;
}

// writeProcFile mesh.cpp:2277-2374: fluid section first, then mobs
void write_sections(std::ostream& cf, std::ostream& ff, std::ostream& bf, const ProcPart& p, const char* type) {
    if (!p.present) return;
    cf << "meta " << type << " " << p.cells.size() << " " << p.region << " " << p.group << "\n";
    for (int64_t c : p.cells) cf << c << " ";
    cf << "\n";
    ff << "meta " << type << " " << p.faces.size() << " " << p.region << " " << p.group << "\n";
    for (const ProcFace& x : p.faces) ff << x.gid << " " << x.owner << " " << x.nb << "\n";
    ff << "\n";
    bf << "meta " << type << " " << p.bounds.size() << " " << p.region << " " << p.group << "\n";
    for (const ProcBound& b : p.bounds) bf << b.name << " " << b.start << " " << b.n << "\n";
    bf << "\n";
}

int type_index(const std::string& t) { return t == "fluid" ? 0 : t == "mobs" ? 1 : -1; }

bool read_file(const std::string& path, std::vector<std::string>& lines) {
    std::ifstream f(path);
    if (!f.is_open()) return false;
    std::string line;
    while (std::getline(f, line)) lines.push_back(line);
    return true;
}

// "meta <type> <n> <region> <group>"; false for any other line
bool meta_line(const std::string& line, int& type, int64_t& n, int& region, int& group) {
    if (line.compare(0, 4, "meta") != 0) return false;
    std::istringstream is(line);
    std::string meta, t;
    if (!(is >> meta >> t >> n >> region >> group) || n < 0) { type = -2; return true; }
    type = type_index(t);
    return true;
}

// readCells / readFaces / readBoundaries (mesh.cpp:2507-2642) into parts, checked
// against the grid: every index in range, each face of the listed cells, each group
// inside the face list, one section per class and file, region ids alike across files.
int read_parts(const hs_mesh& m, const char* dir, ProcPart out[2]) {
    const FaceNum fn(m);
    std::vector<std::string> cl, fl, bl;
    if (!read_file(join(dir, "processor0/cells.txt"), cl) || !read_file(join(dir, "processor0/faces.txt"), fl) ||
        !read_file(join(dir, "processor0/boundary.txt"), bl))
        return HS_E_IO;
    ProcPart p[2];
    bool seen_faces[2] = {false, false}, seen_bounds[2] = {false, false};
    int type, region, group;
    int64_t n;
    for (size_t a = 0; a < cl.size(); ++a) {
        if (!meta_line(cl[a], type, n, region, group)) continue;
        if (type == -2 || a + 1 >= cl.size()) return HS_E_IO;
        if (type < 0) { ++a; continue; }
        if (p[type].present) return HS_E_IO;
        p[type].present = true;
        p[type].region = region;
        p[type].group = group;
        std::istringstream is(cl[++a]);
        for (int64_t i = 0; i < n; ++i) {
            int64_t c;
            if (!(is >> c) || c < 1 || c > fn.ncells()) return HS_E_IO;
            p[type].cells.push_back(c);
        }
    }
    for (size_t a = 0; a < fl.size(); ++a) {
        if (!meta_line(fl[a], type, n, region, group)) continue;
        if (type == -2 || a + (size_t)n >= fl.size()) return HS_E_IO;
        if (type < 0) { a += n; continue; }
        if (!p[type].present || seen_faces[type] || p[type].region != region || p[type].group != group) return HS_E_IO;
        seen_faces[type] = true;
        const int64_t nc = (int64_t)p[type].cells.size();
        for (int64_t i = 0; i < n; ++i) {
            std::istringstream is(fl[++a]);
            ProcFace x;
            if (!(is >> x.gid >> x.owner >> x.nb) || x.gid < 0 || x.gid >= fn.nfaces() || x.owner < 0 ||
                x.owner >= nc || x.nb < 0 || x.nb >= nc)
                return HS_E_IO;
            p[type].faces.push_back(x);
        }
    }
    for (size_t a = 0; a < bl.size(); ++a) {
        if (!meta_line(bl[a], type, n, region, group)) continue;
        if (type == -2 || a + (size_t)n >= bl.size()) return HS_E_IO;
        if (type < 0) { a += n; continue; }
        if (!p[type].present || seen_bounds[type] || p[type].region != region || p[type].group != group) return HS_E_IO;
        seen_bounds[type] = true;
        for (int64_t i = 0; i < n; ++i) {
            std::istringstream is(bl[++a]);
            ProcBound b;
            if (!(is >> b.name >> b.start >> b.n) || b.start < 0 || b.n < 0 ||
                b.start + b.n > (int64_t)p[type].faces.size())
                return HS_E_IO;
            p[type].bounds.push_back(b);
        }
    }
    long f[6];
    for (int t = 0; t < 2; ++t) {
        if (!p[t].present) continue;
        if (!seen_faces[t] || !seen_bounds[t]) return HS_E_IO;
        std::vector<uint8_t> in_group(p[t].faces.size(), 0);
        for (const ProcBound& b : p[t].bounds)
            for (int64_t i = b.start; i < b.start + b.n; ++i) in_group[i] = 1;
        for (size_t i = 0; i < p[t].faces.size(); ++i) {
            const ProcFace& x = p[t].faces[i];
            fn.of_cell(p[t].cells[x.owner], f);
            if (std::find(f, f + 6, (long)x.gid) == f + 6) return HS_E_IO;
            if (in_group[i]) continue;
            fn.of_cell(p[t].cells[x.nb], f);  // an internal face is also a face of its neighbour
            if (x.nb == x.owner || std::find(f, f + 6, (long)x.gid) == f + 6) return HS_E_IO;
        }
    }
    out[0] = std::move(p[0]);
    out[1] = std::move(p[1]);
    return HS_OK;
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
    // without ygrid the cylindrical form's nyr(1)= marks where the y section starts
    const bool cylindrical = pos[1] == std::string::npos;
    if (cylindrical) pos[1] = content.find("nyr(1)=");
    if (pos[0] == std::string::npos || pos[1] == std::string::npos || pos[2] == std::string::npos) return HS_E_IO;
    int nyr = 0, yl1 = 0, yl2 = 0;
    if (cylindrical && (!cyl_value(content, "nyr(1)=", nyr) || !cyl_value(content, "yl(1)=", yl1) ||
                        !cyl_value(content, "yl(2)=", yl2) || nyr < 1))
        return HS_E_IO;
    hs_mesh* m = new hs_mesh();
    read_pairs(section(content, pos[0], pos), m->x);
    if (cylindrical) {
        const int step = (yl2 - yl1) / nyr;
        for (int i = 0; i < nyr; ++i) m->y.push_back((double)(step * i));
        m->y_span = yl2 - yl1;
        m->periodic_y = m->y_span == 360;
        if (m->periodic_y) m->y.push_back((double)(step * nyr));
    } else {
        read_pairs(section(content, pos[1], pos), m->y);
    }
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
    // wall planes (initFacesType mesh.cpp:168-237); they cut region adjacency and type
    // faces for the processor files
    const long nyp = m->periodic_y ? ny : ny + 1;
    const long nxf = (nx + 1) * ny * nz, nyf = nx * nyp * nz;
    m->wall_line.assign(nxf + nyf + nx * ny * (nz + 1), 0);
    uint8_t* xw = m->wall_line.data();
    uint8_t* yw = xw + nxf;
    uint8_t* zw = yw + nyf;
    for (const auto& w : m->walls) {
        const int i0 = w[0], i1 = w[1], j0 = w[2], j1 = w[3], k0 = w[4], k1 = w[5];
        if (i0 == i1 && i0 >= 1 && i0 <= nx + 1)
            for (long k = std::max(1, k0); k <= std::min<long>(nz, k1); ++k)
                for (long j = std::max(1, j0); j <= std::min<long>(ny, j1); ++j) xw[(i0 - 1) + (nx + 1) * ((j - 1) + ny * (k - 1))] = 1;
        if (j0 == j1 && j0 >= 1 && j0 <= nyp)
            for (long k = std::max(1, k0); k <= std::min<long>(nz, k1); ++k)
                for (long i = std::max(1, i0); i <= std::min<long>(nx, i1); ++i) yw[(i - 1) + nx * ((j0 - 1) + nyp * (k - 1))] = 1;
        if (k0 == k1 && k0 >= 1 && k0 <= nz + 1)
            for (long j = std::max(1, j0); j <= std::min<long>(ny, j1); ++j)
                for (long i = std::max(1, i0); i <= std::min<long>(nx, i1); ++i) zw[(i - 1) + nx * ((j - 1) + ny * (k0 - 1))] = 1;
    }
    m->nfluid_regions = count_regions(*m, 0);
    m->nsolid_regions = count_regions(*m, 1);
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

int hs_mesh_y_period(const hs_mesh* m, int32_t* span, int32_t* periodic) {
    if (!m) return HS_E_ARG;
    if (span) *span = m->y_span;
    if (periodic) *periodic = m->periodic_y ? 1 : 0;
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

int hs_mesh_write_proc(const hs_mesh* m, const char* dir) {
    if (!m) return HS_E_ARG;
    ProcPart built[2];
    const ProcPart* p = m->part;
    if (!m->have_parts) {
        // one part holds one region of each class: with more, decomposeMesh stops
        // (fluid, mesh.cpp:762-765) or leaves solid regions unwritten or out of range (:1133-1160)
        if (m->nfluid_regions > 1 || m->nsolid_regions > 1) return HS_E_ARG;
        build_part(*m, 0, built[0]);
        build_part(*m, 1, built[1]);
        p = built;
    }
    const std::string d = join(dir, "processor0");
    if (mkdir(d.c_str(), 0755) != 0 && errno != EEXIST) return HS_E_IO;
    std::ofstream cf(d + "/cells.txt"), ff(d + "/faces.txt"), bf(d + "/boundary.txt");
    if (!cf.is_open() || !ff.is_open() || !bf.is_open()) return HS_E_IO;
    write_sections(cf, ff, bf, p[0], "fluid");
    write_sections(cf, ff, bf, p[1], "mobs");
    cf.close();
    ff.close();
    bf.close();
    return (cf.fail() || ff.fail() || bf.fail()) ? HS_E_IO : HS_OK;
}

int hs_mesh_read_proc(hs_mesh* m, const char* dir) {
    if (!m) return HS_E_ARG;
    ProcPart p[2];
    const int rc = read_parts(*m, dir, p);
    if (rc != HS_OK) return rc;
    // cell types from the region lists: the one-rank solver needs every cell exactly once
    std::vector<uint8_t> ct(m->cell_type.size(), 255);
    for (int t = 0; t < 2; ++t)
        for (int64_t c : p[t].cells) {
            if (ct[c - 1] != 255) return HS_E_IO;
            ct[c - 1] = (uint8_t)t;
        }
    if (std::find(ct.begin(), ct.end(), (uint8_t)255) != ct.end()) return HS_E_IO;
    // in place: a grid view taken earlier (hs_mesh_grid) keeps pointing at the types
    std::copy(ct.begin(), ct.end(), m->cell_type.begin());
    m->nfluid_regions = count_regions(*m, 0);
    m->nsolid_regions = count_regions(*m, 1);
    m->part[0] = std::move(p[0]);
    m->part[1] = std::move(p[1]);
    m->have_parts = true;
    return HS_OK;
}

int hs_mesh_proc_part(const hs_mesh* m, int32_t which, int32_t* region, int32_t* group, int64_t* ncells,
                      int64_t* nfaces, int32_t* nbound) {
    if (!m || which < 0 || which > 1) return HS_E_ARG;
    ProcPart built;
    const ProcPart* p = &m->part[which];
    if (!m->have_parts) {
        if ((which == 0 ? m->nfluid_regions : m->nsolid_regions) > 1) return HS_E_ARG;
        build_part(*m, (uint8_t)which, built);
        p = &built;
    }
    if (region) *region = p->region;
    if (group) *group = p->group;
    if (ncells) *ncells = (int64_t)p->cells.size();
    if (nfaces) *nfaces = (int64_t)p->faces.size();
    if (nbound) *nbound = (int32_t)p->bounds.size();
    return HS_OK;
}

int hs_mesh_proc_boundary(const hs_mesh* m, int32_t which, int32_t idx, char* name, int32_t name_cap,
                          int64_t* start, int64_t* nfaces) {
    if (!m || which < 0 || which > 1 || idx < 0) return HS_E_ARG;
    ProcPart built;
    const ProcPart* p = &m->part[which];
    if (!m->have_parts) {
        if ((which == 0 ? m->nfluid_regions : m->nsolid_regions) > 1) return HS_E_ARG;
        build_part(*m, (uint8_t)which, built);
        p = &built;
    }
    if (idx >= (int32_t)p->bounds.size()) return HS_E_ARG;
    const ProcBound& b = p->bounds[idx];
    if (name) {
        if (name_cap < (int32_t)b.name.size() + 1) return HS_E_ARG;
        std::memcpy(name, b.name.c_str(), b.name.size() + 1);
    }
    if (start) *start = b.start;
    if (nfaces) *nfaces = b.n;
    return HS_OK;
}

void hs_mesh_free(hs_mesh* m) { delete m; }

}  // extern "C"
