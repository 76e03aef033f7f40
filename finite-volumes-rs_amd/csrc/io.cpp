// io.cpp -- host mesh formats around the device path: the SU2 reader and the crate's own
// mesh format (src/core/mesh/io.rs), and the PVTU writer (src/post/write_pvtu.rs) with an
// asynchronous variant whose device fields leave on the context stream and whose files
// are written from a host thread while the GPU carries on.
//
// A topology is what Mesh::add_node / add_face / add_cell receive: node coordinates,
// face node lists with their boundary patch, cell face lists.  fvg_mesh_create_from_topology
// hands it to the device build.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "internal.hpp"

struct fvg_topology {
    int dim = 0;
    std::vector<double> nodes;             // n_nodes x dim
    std::vector<int64_t> face_ptr{0};      // n_faces + 1
    std::vector<int32_t> face_nodes;
    std::vector<int32_t> face_patch;       // -1: interior
    std::vector<int64_t> cell_ptr{0};      // n_cells + 1
    std::vector<int32_t> cell_faces;
    std::vector<std::string> patch_names;  // by patch id ("" when the file names none)
    int64_t n_nodes() const { return dim ? (int64_t)nodes.size() / dim : 0; }
    int64_t n_faces() const { return (int64_t)face_ptr.size() - 1; }
    int64_t n_cells() const { return (int64_t)cell_ptr.size() - 1; }
};

namespace fvg {
namespace {

std::vector<std::string> tokens(const std::string& line) {
    std::vector<std::string> out;
    std::istringstream is(line);
    std::string t;
    while (is >> t) out.push_back(t);
    return out;
}

bool parse_i64(const std::string& s, int64_t& v) {
    const char* b = s.data();
    const char* e = b + s.size();
    auto r = std::from_chars(b, e, v);
    return r.ec == std::errc() && r.ptr == e;
}

bool parse_f64(const std::string& s, double& v) {
    char* end = nullptr;
    v = std::strtod(s.c_str(), &end);
    return end && *end == '\0' && !s.empty();
}

// the value after "KEY=" on a line holding KEY= (io.rs reads it with split("=").nth(1))
bool keyed_value(const std::string& line, const char* key, std::string& value) {
    const size_t p = line.find(key);
    if (p == std::string::npos) return false;
    value = line.substr(p + std::strlen(key));
    const size_t a = value.find_first_not_of(" \t\r");
    const size_t b = value.find_last_not_of(" \t\r");
    value = a == std::string::npos ? "" : value.substr(a, b - a + 1);
    return true;
}

int read_error(const std::string& path, int64_t line, const std::string& what) {
    return fail(FVG_E_INVALID, path + ":" + std::to_string(line) + ": " + what);
}

// VtkElement::faces (src/core/mesh/vtk.rs:56-125): local face node lists per element kind
bool vtk_faces(int kind, const std::vector<int64_t>& n, std::vector<std::vector<int64_t>>& faces) {
    auto F = [&](std::initializer_list<int> idx) {
        std::vector<int64_t> f;
        for (int k : idx) f.push_back(n[(size_t)k]);
        faces.push_back(f);
    };
    faces.clear();
    switch (kind) {
    case 3:   // line
        if (n.size() < 2) return false;
        F({0});
        F({1});
        return true;
    case 5:   // triangle
        if (n.size() < 3) return false;
        F({0, 1});
        F({1, 2});
        F({2, 0});
        return true;
    case 9:   // quad
        if (n.size() < 4) return false;
        F({0, 1});
        F({1, 2});
        F({2, 3});
        F({3, 0});
        return true;
    case 10:   // tetrahedron
        if (n.size() < 4) return false;
        F({0, 2, 1});
        F({0, 1, 3});
        F({1, 2, 3});
        F({0, 3, 2});
        return true;
    case 12:   // hexahedron
        if (n.size() < 8) return false;
        F({0, 3, 2, 1});
        F({4, 5, 6, 7});
        F({0, 1, 5, 4});
        F({1, 2, 6, 5});
        F({2, 3, 7, 6});
        F({3, 0, 4, 7});
        return true;
    }
    return false;
}

struct VecHash {
    size_t operator()(const std::vector<int64_t>& v) const {
        size_t h = v.size();
        for (int64_t x : v) h = h * 1000003u ^ (size_t)x;
        return h;
    }
};

// Mesh::read_su2 (io.rs:489-533): nodes (read_su2_nodes :228-260), elements
// (read_su2_elements :263-297), boundary markers numbered 0.. in file order
// (read_su2_boundaries :301-360), faces by compute_vtk_faces (:364-486): faces keyed by
// their sorted node list in order of first appearance, interior faces first, then the
// boundary faces grouped by ascending marker, each group in first-appearance order.
int read_su2(const std::string& path, fvg_topology& t) {
    std::ifstream in(path);
    if (!in) return fail(FVG_E_INVALID, "cannot open " + path);
    std::vector<std::string> lines;
    for (std::string l; std::getline(in, l);) lines.push_back(l);
    int64_t dim = 0;
    for (const auto& l : lines) {
        std::string v;
        if (keyed_value(l, "NDIME=", v)) {
            if (!parse_i64(tokens(v).empty() ? "" : tokens(v)[0], dim) || (dim != 2 && dim != 3))
                return fail(FVG_E_INVALID, path + ": NDIME must be 2 or 3");
            break;
        }
    }
    if (!dim) return fail(FVG_E_INVALID, path + ": no NDIME= line");
    t.dim = (int)dim;
    std::vector<int> kinds;
    std::vector<std::vector<int64_t>> elem_nodes;
    std::unordered_map<std::vector<int64_t>, int, VecHash> bnd_of;   // sorted marker element -> marker id
    int markers = 0;
    for (size_t i = 0; i < lines.size(); ++i) {
        std::string v;
        if (keyed_value(lines[i], "NPOIN=", v)) {
            int64_t np;
            const auto tv = tokens(v);
            if (tv.empty() || !parse_i64(tv[0], np) || np < 0) return read_error(path, (int64_t)i + 1, "bad NPOIN");
            if (!t.nodes.empty()) continue;
            t.nodes.reserve((size_t)(np * dim));
            for (int64_t k = 0; k < np; ++k) {
                const auto tk = i + 1 + (size_t)k < lines.size() ? tokens(lines[i + 1 + (size_t)k]) : std::vector<std::string>();
                if ((int64_t)tk.size() < dim) return read_error(path, (int64_t)(i + 2 + k), "node needs NDIME coordinates");
                for (int d = 0; d < dim; ++d) {
                    double x;
                    if (!parse_f64(tk[(size_t)d], x)) return read_error(path, (int64_t)(i + 2 + k), "bad coordinate");
                    t.nodes.push_back(x);
                }
            }
            i += (size_t)np;
        } else if (keyed_value(lines[i], "NELEM=", v)) {
            int64_t ne;
            const auto tv = tokens(v);
            if (tv.empty() || !parse_i64(tv[0], ne) || ne < 0) return read_error(path, (int64_t)i + 1, "bad NELEM");
            for (int64_t k = 0; k < ne; ++k) {
                const size_t li = i + 1 + (size_t)k;
                const auto tk = li < lines.size() ? tokens(lines[li]) : std::vector<std::string>();
                std::vector<int64_t> vals;
                for (const auto& s : tk) {
                    int64_t x;
                    if (!parse_i64(s, x)) return read_error(path, (int64_t)li + 1, "bad element entry");
                    vals.push_back(x);
                }
                if (vals.empty()) return read_error(path, (int64_t)li + 1, "empty element line");
                kinds.push_back((int)vals[0]);
                elem_nodes.emplace_back(vals.begin() + 1, vals.end());
            }
            i += (size_t)ne;
        } else if (keyed_value(lines[i], "MARKER_TAG=", v)) {
            t.patch_names.push_back(v);
            ++markers;
        } else if (keyed_value(lines[i], "MARKER_ELEMS=", v)) {
            int64_t nm;
            const auto tv = tokens(v);
            if (markers == 0) return read_error(path, (int64_t)i + 1, "MARKER_ELEMS before MARKER_TAG");
            if (tv.empty() || !parse_i64(tv[0], nm) || nm < 0) return read_error(path, (int64_t)i + 1, "bad MARKER_ELEMS");
            for (int64_t k = 0; k < nm; ++k) {
                const size_t li = i + 1 + (size_t)k;
                const auto tk = li < lines.size() ? tokens(lines[li]) : std::vector<std::string>();
                std::vector<int64_t> key;
                for (size_t q = 1; q < tk.size(); ++q) {   // the element kind is dropped (io.rs:327)
                    int64_t x;
                    if (!parse_i64(tk[q], x)) return read_error(path, (int64_t)li + 1, "bad marker element");
                    key.push_back(x);
                }
                std::sort(key.begin(), key.end());
                bnd_of[key] = markers - 1;
            }
            i += (size_t)nm;
        }
    }
    const int64_t nn = t.n_nodes();
    // compute_vtk_faces
    std::unordered_map<std::vector<int64_t>, int64_t, VecHash> id_of;
    std::vector<std::vector<int64_t>> fnodes;
    std::vector<int> fpatch;
    std::vector<std::vector<int64_t>> efaces;
    std::vector<std::vector<int64_t>> lf;
    for (size_t e = 0; e < kinds.size(); ++e) {
        if (!vtk_faces(kinds[e], elem_nodes[e], lf))
            return fail(FVG_E_INVALID, path + ": element " + std::to_string(e) + " has an unsupported kind or too few nodes");
        std::vector<int64_t> ef;
        for (auto& f : lf) {
            for (int64_t x : f)
                if (x < 0 || x >= nn) return fail(FVG_E_INVALID, path + ": element " + std::to_string(e) + " names a node out of range");
            std::vector<int64_t> key = f;
            std::sort(key.begin(), key.end());
            auto it = id_of.find(key);
            int64_t id;
            if (it == id_of.end()) {
                id = (int64_t)fnodes.size();
                auto b = bnd_of.find(key);
                fpatch.push_back(b == bnd_of.end() ? -1 : b->second);
                fnodes.push_back(f);
                id_of.emplace(std::move(key), id);
            } else {
                id = it->second;
            }
            ef.push_back(id);
        }
        efaces.push_back(std::move(ef));
    }
    const int64_t nf = (int64_t)fnodes.size();
    std::vector<int64_t> final_id((size_t)nf), order;
    order.reserve((size_t)nf);
    for (int64_t f = 0; f < nf; ++f)
        if (fpatch[(size_t)f] < 0) order.push_back(f);
    for (int m = 0; m < markers; ++m)
        for (int64_t f = 0; f < nf; ++f)
            if (fpatch[(size_t)f] == m) order.push_back(f);
    for (size_t q = 0; q < order.size(); ++q) final_id[(size_t)order[q]] = (int64_t)q;
    for (int64_t f : order) {
        for (int64_t x : fnodes[(size_t)f]) t.face_nodes.push_back((int32_t)x);
        t.face_ptr.push_back((int64_t)t.face_nodes.size());
        t.face_patch.push_back(fpatch[(size_t)f]);
    }
    for (const auto& ef : efaces) {
        for (int64_t f : ef) t.cell_faces.push_back((int32_t)final_id[(size_t)f]);
        t.cell_ptr.push_back((int64_t)t.cell_faces.size());
    }
    return FVG_OK;
}

// Mesh::read (io.rs:68-211), the crate's own format written by Mesh::write (io.rs:10-64):
// NDIME=, NPATCH= (name id fstart len), NNODES= (coordinates), NFACES= (count, nodes,
// patch or "-", global id, owner rank or "-"), NCELLS= (count, faces, global id, owner
// rank or "-").  Remote ownership (a partition of a distributed mesh) is refused.
int read_native(const std::string& path, fvg_topology& t) {
    std::ifstream in(path);
    if (!in) return fail(FVG_E_INVALID, "cannot open " + path);
    std::string section = "none";
    int64_t to_read = 0, line_no = 0;
    std::map<int64_t, std::string> names;
    for (std::string raw; std::getline(in, raw);) {
        ++line_no;
        const auto tk = tokens(raw);
        if (tk.empty()) continue;
        if (tk[0][0] == 'N') {
            const size_t eq = raw.find('=');
            if (eq == std::string::npos) return read_error(path, line_no, "section header without '='");
            const size_t a = raw.find_first_not_of(" \t");
            section = raw.substr(a, eq - a);
            int64_t val;
            const auto tv = tokens(raw.substr(eq + 1));
            if (tv.empty() || !parse_i64(tv[0], val) || val < 0) return read_error(path, line_no, "bad section count");
            if (section == "NDIME") {
                if (val != 2 && val != 3) return read_error(path, line_no, "NDIME must be 2 or 3");
                t.dim = (int)val;
            }
            to_read = val;
            continue;
        }
        if (section == "none" || to_read == 0) continue;
        if (section == "NNODES") {
            if (!t.dim) return read_error(path, line_no, "NNODES before NDIME");
            for (int d = 0; d < t.dim; ++d) {
                double x = 0.0;
                if (d < (int)tk.size() && !parse_f64(tk[(size_t)d], x)) return read_error(path, line_no, "bad coordinate");
                t.nodes.push_back(x);
            }
        } else if (section == "NFACES") {
            int64_t size;
            if (!parse_i64(tk[0], size) || size < 0 || size > 64 || (int64_t)tk.size() < size + 4)
                return read_error(path, line_no, "bad face line");
            for (int64_t k = 0; k < size; ++k) {
                int64_t x;
                if (!parse_i64(tk[(size_t)(1 + k)], x) || x < 0) return read_error(path, line_no, "bad face node");
                t.face_nodes.push_back((int32_t)x);
            }
            t.face_ptr.push_back((int64_t)t.face_nodes.size());
            const std::string& b = tk[(size_t)(1 + size)];
            int64_t pid = -1;
            if (b != "-" && (!parse_i64(b, pid) || pid < 0 || pid > 65535)) return read_error(path, line_no, "bad boundary tag");
            t.face_patch.push_back((int32_t)pid);
            if (tk[(size_t)(3 + size)] != "-")
                return fail(FVG_E_UNSUPPORTED, path + ":" + std::to_string(line_no) + ": remote-owned face (distributed mesh)");
        } else if (section == "NCELLS") {
            int64_t size;
            if (!parse_i64(tk[0], size) || size < 0 || size > 64 || (int64_t)tk.size() < size + 3)
                return read_error(path, line_no, "bad cell line");
            for (int64_t k = 0; k < size; ++k) {
                int64_t x;
                if (!parse_i64(tk[(size_t)(1 + k)], x) || x < 0) return read_error(path, line_no, "bad cell face");
                t.cell_faces.push_back((int32_t)x);
            }
            t.cell_ptr.push_back((int64_t)t.cell_faces.size());
            if (tk[(size_t)(2 + size)] != "-")
                return fail(FVG_E_UNSUPPORTED, path + ":" + std::to_string(line_no) + ": remote-owned cell (distributed mesh)");
        } else if (section == "NPATCH") {
            int64_t bid;
            if (tk.size() < 4 || !parse_i64(tk[1], bid) || bid < 0) return read_error(path, line_no, "bad patch line");
            names[bid] = tk[0];
        }
        --to_read;
    }
    if (!t.dim) return fail(FVG_E_INVALID, path + ": no NDIME= line");
    for (const auto& kv : names) {
        if (kv.first >= (int64_t)t.patch_names.size()) t.patch_names.resize((size_t)kv.first + 1);
        t.patch_names[(size_t)kv.first] = kv.second;
    }
    return FVG_OK;
}

// Rust's `{}` for f64: the shortest digits that read back to the same value, never in
// exponent form; NaN and the infinities as Rust spells them
void put_f64(std::string& out, double v) {
    if (v != v) {
        out += "NaN";
        return;
    }
    char buf[512];
    auto r = std::to_chars(buf, buf + sizeof(buf), v, std::chars_format::fixed);
    out.append(buf, r.ptr);
}

void put_i64(std::string& out, int64_t v) {
    char buf[24];
    auto r = std::to_chars(buf, buf + sizeof(buf), v);
    out.append(buf, r.ptr);
}

}  // namespace
}  // namespace fvg

// PvtuWriter (src/post/write_pvtu.rs:30-466) for one rank: the mesh data of
// collect_vtu_mesh_data (:161-204) is built once per writer.
struct fvg_pvtu_writer {
    int dim = 0;
    int64_t n_cells = 0;
    std::vector<int64_t> cell_faces{}, cell_faces_starts{0}, cell_nodes{}, cell_nodes_starts{0};
    std::vector<double> nodes;   // written nodes, in order of first appearance
    // asynchronous writes
    std::thread worker;
    int status = FVG_OK;
    std::string error;
    void* pinned = nullptr;
    size_t pinned_cap = 0;
    hipEvent_t ready = nullptr;
};

namespace fvg {
namespace {

// file_without_extension / file_without_directory (write_pvtu.rs:14-24, 416-424), kept
// as the reference scans: backwards from the end, never looking at the first character
std::string without_extension(const std::string& p) {
    if (p.empty()) return p;
    size_t i = p.size() - 1;
    while (i > 0) {
        if (p[i] == '.') break;
        --i;
    }
    return p.substr(0, i);
}

std::string without_directory(const std::string& p) {
    if (p.empty()) return p;
    size_t i = p.size() - 1;
    while (i > 0) {
        if (p[i] == '/' || p[i] == '\\') {
            ++i;
            break;
        }
        --i;
    }
    return p.substr(i);
}

struct FieldView {
    std::string name;
    int ncomp;
    const double* v;   // n_cells x ncomp
};

int write_files(const fvg_pvtu_writer& w, const std::string& filepath, const std::vector<FieldView>& fields) {
    const std::string prefix = without_extension(filepath);
    {   // write_pvtu_parent_file (:69-118)
        std::string s = "<VTKFile type=\"PUnstructuredGrid\">\n  <PUnstructuredGrid GhostLevel=\"0\">\n";
        if (!fields.empty()) {
            s += "    <PCellData>\n";
            for (const auto& f : fields) {
                if (f.ncomp == 1) s += "      <PDataArray Name=\"" + f.name + "\" type=\"Float32\"/>\n";
                else
                    s += "      <PDataArray Name=\"" + f.name + "\" type=\"Float32\" NumberOfComponents=\"" +
                         std::to_string(f.ncomp) + "\"/>\n";
            }
            s += "    </PCellData>\n";
        }
        s += "    <PPoints>\n      <PDataArray type=\"Float32\" NumberOfComponents=\"3\"/>\n    </PPoints>\n";
        s += "    <Piece Source=\"" + without_directory(prefix) + "_0.vtu\"/>\n";
        s += "  </PUnstructuredGrid>\n</VTKFile>\n";
        std::ofstream o(prefix + ".pvtu", std::ios::binary);
        if (!(o << s)) return fail(FVG_E_INVALID, "cannot write " + prefix + ".pvtu");
    }
    // write_vtu_file (:311-413)
    std::string s;
    const int64_t npts = (int64_t)w.nodes.size() / w.dim, nc = w.n_cells;
    s += "<VTKFile type=\"UnstructuredGrid\">\n  <UnstructuredGrid>\n";
    s += "    <Piece NumberOfPoints=\"" + std::to_string(npts) + "\" NumberOfCells=\"" + std::to_string(nc) + "\">\n";
    s += "      <Points>\n        <DataArray type=\"Float32\" Name=\"Points\" NumberOfComponents=\"3\" format=\"ascii\">\n";
    for (int64_t p = 0; p < npts; ++p) {
        for (int d = 0; d < w.dim; ++d) {
            put_f64(s, w.nodes[(size_t)(p * w.dim + d)]);
            s += ' ';
        }
        for (int d = w.dim; d < 3; ++d) s += "0 ";
    }
    s += "\n        </DataArray>\n      </Points>\n      <Cells>\n";
    auto row = [&](int64_t c0, int64_t c1, const std::vector<int64_t>& a) {
        for (int64_t i = c0; i < c1; ++i) {
            put_i64(s, a[(size_t)i]);
            s += ' ';
        }
        s += '\n';
    };
    if (w.dim == 2) {   // write_polygons (:247-308): the cell's nodes walked around its faces
        s += "        <DataArray type=\"Int64\" Name=\"connectivity\" format=\"ascii\">\n";
        for (int64_t c = 0; c < nc; ++c) {
            const int64_t cs = w.cell_faces_starts[(size_t)c], cfs = w.cell_faces[(size_t)cs];
            for (int64_t fi = 0; fi < cfs; ++fi) {
                const int64_t fo = fi == cfs - 1 ? 0 : fi + 1;
                const int64_t i = cs + 1 + fi * 3, o = cs + 1 + fo * 3;
                const int64_t ni0 = w.cell_faces[(size_t)(i + 1)], ni1 = w.cell_faces[(size_t)(i + 2)];
                const int64_t no0 = w.cell_faces[(size_t)(o + 1)], no1 = w.cell_faces[(size_t)(o + 2)];
                int64_t pick;
                if (ni0 == no0 || ni0 == no1) {
                    pick = ni1;
                } else if (ni1 == no0 || ni1 == no1) {
                    pick = ni0;
                } else {   // compare positions (tolerance 1e-10)
                    auto dist = [&](int64_t a, int64_t b) {
                        double x = 0.0;
                        for (int d = 0; d < 2; ++d) {
                            const double t = w.nodes[(size_t)(a * 2 + d)] - w.nodes[(size_t)(b * 2 + d)];
                            x += t * t;
                        }
                        return std::sqrt(x);
                    };
                    pick = (dist(ni0, no0) < 1e-10 || dist(ni0, no1) < 1e-10) ? ni1 : ni0;
                }
                put_i64(s, pick);
                s += ' ';
            }
            s += '\n';
        }
        s += "        </DataArray>\n        <DataArray type=\"Int64\" Name=\"offsets\" format=\"ascii\">\n";
        int64_t off = 0;
        for (int64_t c = 0; c < nc; ++c) {
            off += w.cell_faces[(size_t)w.cell_faces_starts[(size_t)c]];
            put_i64(s, off);
            s += ' ';
        }
        s += "\n        </DataArray>\n        <DataArray type=\"Int64\" Name=\"types\" format=\"ascii\">\n";
        for (int64_t c = 0; c < nc; ++c) s += "7 ";
        s += "\n        </DataArray>\n";
    } else {   // write_polyhedra (:206-245)
        s += "        <DataArray type=\"Int64\" Name=\"connectivity\" format=\"ascii\">\n";
        for (int64_t c = 0; c < nc; ++c) row(w.cell_nodes_starts[(size_t)c], w.cell_nodes_starts[(size_t)c + 1], w.cell_nodes);
        s += "        </DataArray>\n        <DataArray type=\"Int64\" Name=\"offsets\" format=\"ascii\">\n";
        for (int64_t c = 0; c < nc; ++c) {
            put_i64(s, w.cell_nodes_starts[(size_t)c + 1]);
            s += ' ';
        }
        s += "\n        </DataArray>\n        <DataArray type=\"Int64\" Name=\"types\" format=\"ascii\">\n";
        for (int64_t c = 0; c < nc; ++c) s += "42 ";
        s += "\n        </DataArray>\n        <DataArray type=\"Int64\" Name=\"faces\" format=\"ascii\">\n";
        for (int64_t c = 0; c < nc; ++c) row(w.cell_faces_starts[(size_t)c], w.cell_faces_starts[(size_t)c + 1], w.cell_faces);
        s += "        </DataArray>\n        <DataArray type=\"Int64\" Name=\"faceoffsets\" format=\"ascii\">\n";
        for (int64_t c = 0; c < nc; ++c) {
            put_i64(s, w.cell_faces_starts[(size_t)c + 1]);
            s += ' ';
        }
        s += "\n        </DataArray>\n";
    }
    s += "      </Cells>\n";
    if (!fields.empty()) {
        s += "    <CellData>\n";
        for (const auto& f : fields) {
            if (f.ncomp == 1) s += "      <DataArray Name=\"" + f.name + "\" type=\"Float32\">\n";
            else
                s += "      <DataArray Name=\"" + f.name + "\" type=\"Float32\" NumberOfComponents=\"" +
                     std::to_string(f.ncomp) + "\">\n";
            const int64_t cnt = nc * f.ncomp;
            for (int64_t k = 0; k < cnt; ++k) {
                put_f64(s, f.v[k]);
                s += ' ';
            }
            s += "\n      </DataArray>\n";
        }
        s += "    </CellData>\n";
    }
    s += "    </Piece>\n  </UnstructuredGrid>\n</VTKFile>\n";
    std::ofstream o(prefix + "_0.vtu", std::ios::binary);
    if (!(o << s)) return fail(FVG_E_INVALID, "cannot write " + prefix + "_0.vtu");
    return FVG_OK;
}

int check_fields(int32_t n_fields, const char* const* names, const int32_t* ncomps) {
    if (n_fields < 0 || (n_fields > 0 && (!names || !ncomps))) return fail(FVG_E_INVALID, "null field arrays");
    for (int32_t k = 0; k < n_fields; ++k) {
        if (!names[k]) return fail(FVG_E_INVALID, "null field name");
        if (ncomps[k] < 1 || ncomps[k] > 9) return fail(FVG_E_INVALID, "a field has 1 to 9 components");
    }
    return FVG_OK;
}

int join(fvg_pvtu_writer* w) {
    if (w->worker.joinable()) w->worker.join();
    const int s = w->status;
    if (s != FVG_OK) set_error(w->error);
    w->status = FVG_OK;
    return s;
}

}  // namespace
}  // namespace fvg

using namespace fvg;

extern "C" {

int fvg_topology_read_su2(const char* path, fvg_topology** out) {
    if (!path || !out) return fail(FVG_E_INVALID, "null argument");
    *out = nullptr;
    auto t = std::make_unique<fvg_topology>();
    FVG_TRY(read_su2(path, *t));
    *out = t.release();
    return FVG_OK;
}

int fvg_topology_read_native(const char* path, fvg_topology** out) {
    if (!path || !out) return fail(FVG_E_INVALID, "null argument");
    *out = nullptr;
    auto t = std::make_unique<fvg_topology>();
    FVG_TRY(read_native(path, *t));
    *out = t.release();
    return FVG_OK;
}

int fvg_topology_create(int dim, int64_t n_nodes, const double* nodes, int64_t n_faces, const int64_t* face_ptr,
                        const int32_t* face_nodes, const int32_t* face_patch, int64_t n_cells, const int64_t* cell_ptr,
                        const int32_t* cell_faces, fvg_topology** out) {
    if (!out || !nodes || !face_ptr || !face_nodes || !face_patch || !cell_ptr || !cell_faces)
        return fail(FVG_E_INVALID, "null argument");
    if ((dim != 2 && dim != 3) || n_nodes < 0 || n_faces < 0 || n_cells < 0) return fail(FVG_E_INVALID, "bad sizes");
    if (face_ptr[0] != 0 || cell_ptr[0] != 0) return fail(FVG_E_INVALID, "offsets must start at 0");
    for (int64_t i = 0; i < n_faces; ++i)
        if (face_ptr[i + 1] < face_ptr[i]) return fail(FVG_E_INVALID, "face_ptr must not decrease");
    for (int64_t i = 0; i < n_cells; ++i)
        if (cell_ptr[i + 1] < cell_ptr[i]) return fail(FVG_E_INVALID, "cell_ptr must not decrease");
    auto t = std::make_unique<fvg_topology>();
    t->dim = dim;
    t->nodes.assign(nodes, nodes + n_nodes * dim);
    t->face_ptr.assign(face_ptr, face_ptr + n_faces + 1);
    t->face_nodes.assign(face_nodes, face_nodes + face_ptr[n_faces]);
    t->face_patch.assign(face_patch, face_patch + n_faces);
    t->cell_ptr.assign(cell_ptr, cell_ptr + n_cells + 1);
    t->cell_faces.assign(cell_faces, cell_faces + cell_ptr[n_cells]);
    *out = t.release();
    return FVG_OK;
}

int fvg_topology_destroy(fvg_topology* t) {
    delete t;
    return FVG_OK;
}

int fvg_topology_sizes(const fvg_topology* t, int32_t* dim, int64_t* n_nodes, int64_t* n_faces,
                       int64_t* n_face_nodes, int64_t* n_cells, int64_t* n_cell_faces, int32_t* n_patches) {
    if (!t) return fail(FVG_E_INVALID, "null topology");
    if (dim) *dim = t->dim;
    if (n_nodes) *n_nodes = t->n_nodes();
    if (n_faces) *n_faces = t->n_faces();
    if (n_face_nodes) *n_face_nodes = (int64_t)t->face_nodes.size();
    if (n_cells) *n_cells = t->n_cells();
    if (n_cell_faces) *n_cell_faces = (int64_t)t->cell_faces.size();
    if (n_patches) *n_patches = (int32_t)t->patch_names.size();
    return FVG_OK;
}

int fvg_topology_arrays(const fvg_topology* t, const double** nodes, const int64_t** face_ptr,
                        const int32_t** face_nodes, const int32_t** face_patch, const int64_t** cell_ptr,
                        const int32_t** cell_faces) {
    if (!t) return fail(FVG_E_INVALID, "null topology");
    if (nodes) *nodes = t->nodes.data();
    if (face_ptr) *face_ptr = t->face_ptr.data();
    if (face_nodes) *face_nodes = t->face_nodes.data();
    if (face_patch) *face_patch = t->face_patch.data();
    if (cell_ptr) *cell_ptr = t->cell_ptr.data();
    if (cell_faces) *cell_faces = t->cell_faces.data();
    return FVG_OK;
}

int fvg_topology_patch_name(const fvg_topology* t, int32_t patch, const char** name) {
    if (!t || !name) return fail(FVG_E_INVALID, "null argument");
    if (patch < 0 || patch >= (int32_t)t->patch_names.size()) return fail(FVG_E_INVALID, "patch id out of range");
    *name = t->patch_names[(size_t)patch].c_str();
    return FVG_OK;
}

int fvg_mesh_create_from_topology(fvg_ctx* c, const fvg_topology* t, fvg_mesh** out) {
    if (!c || !t || !out) return fail(FVG_E_INVALID, "null argument");
    static const int32_t none = 0;
    return fvg_mesh_create(c, t->dim, t->n_nodes(), t->nodes.data(), t->n_faces(), t->face_ptr.data(),
                           t->face_nodes.empty() ? &none : t->face_nodes.data(),
                           t->face_patch.empty() ? &none : t->face_patch.data(), t->n_cells(), t->cell_ptr.data(),
                           t->cell_faces.empty() ? &none : t->cell_faces.data(), out);
}

int fvg_pvtu_writer_create(const fvg_topology* t, fvg_pvtu_writer** out) {
    if (!t || !out) return fail(FVG_E_INVALID, "null argument");
    *out = nullptr;
    auto w = std::make_unique<fvg_pvtu_writer>();
    w->dim = t->dim;
    w->n_cells = t->n_cells();
    std::vector<int64_t> local((size_t)t->n_nodes(), -1);
    std::vector<int64_t> cni;
    for (int64_t c = 0; c < t->n_cells(); ++c) {
        const int64_t s = t->cell_ptr[(size_t)c], e = t->cell_ptr[(size_t)c + 1];
        w->cell_faces.push_back(e - s);
        cni.clear();
        for (int64_t k = s; k < e; ++k) {
            const int32_t f = t->cell_faces[(size_t)k];
            if (f < 0 || f >= t->n_faces()) return fail(FVG_E_INVALID, "cell names a face out of range");
            const int64_t fs = t->face_ptr[(size_t)f], fe = t->face_ptr[(size_t)f + 1];
            w->cell_faces.push_back(fe - fs);
            for (int64_t q = fs; q < fe; ++q) {
                const int32_t n = t->face_nodes[(size_t)q];
                if (n < 0 || n >= t->n_nodes()) return fail(FVG_E_INVALID, "face names a node out of range");
                if (local[(size_t)n] < 0) {
                    local[(size_t)n] = (int64_t)w->nodes.size() / t->dim;
                    for (int d = 0; d < t->dim; ++d) w->nodes.push_back(t->nodes[(size_t)(n * t->dim + d)]);
                }
                w->cell_faces.push_back(local[(size_t)n]);
                cni.push_back(local[(size_t)n]);
            }
        }
        // the cell's distinct nodes (the reference collects them in a HashSet, whose
        // iteration order is arbitrary; here, in order of first appearance)
        std::vector<int64_t> seen;
        for (int64_t v : cni)
            if (std::find(seen.begin(), seen.end(), v) == seen.end()) seen.push_back(v);
        w->cell_nodes.insert(w->cell_nodes.end(), seen.begin(), seen.end());
        w->cell_nodes_starts.push_back((int64_t)w->cell_nodes.size());
        w->cell_faces_starts.push_back((int64_t)w->cell_faces.size());
    }
    if (t->dim == 2)
        for (int64_t c = 0; c < w->n_cells; ++c) {
            const int64_t cs = w->cell_faces_starts[(size_t)c], cfs = w->cell_faces[(size_t)cs];
            for (int64_t fi = 0; fi < cfs; ++fi)
                if (w->cell_faces[(size_t)(cs + 1 + fi * 3)] != 2)
                    return fail(FVG_E_INVALID, "2D cells need two-node faces");
        }
    *out = w.release();
    return FVG_OK;
}

int fvg_pvtu_write(fvg_pvtu_writer* w, const char* path, int32_t n_fields, const char* const* names,
                   const int32_t* ncomps, const double* const* values) {
    if (!w || !path || (n_fields > 0 && !values)) return fail(FVG_E_INVALID, "null argument");
    FVG_TRY(check_fields(n_fields, names, ncomps));
    FVG_TRY(join(w));
    std::vector<FieldView> f;
    for (int32_t k = 0; k < n_fields; ++k) {
        if (!values[k]) return fail(FVG_E_INVALID, "null field values");
        f.push_back(FieldView{names[k], ncomps[k], values[k]});
    }
    return write_files(*w, path, f);
}

int fvg_pvtu_write_async(fvg_pvtu_writer* w, fvg_ctx* c, const char* path, int32_t n_fields, const char* const* names,
                         const int32_t* ncomps, const fvg_vec* const* values) {
    if (!w || !c || !path || (n_fields > 0 && !values)) return fail(FVG_E_INVALID, "null argument");
    FVG_TRY(check_fields(n_fields, names, ncomps));
    FVG_TRY(join(w));   // one write in flight per writer; its staging is reused
    size_t total = 0;
    for (int32_t k = 0; k < n_fields; ++k) {
        if (!values[k] || values[k]->n != w->n_cells * ncomps[k])
            return fail(FVG_E_INVALID, "a field holds n_cells x ncomp values");
        total += (size_t)values[k]->n;
    }
    if (w->pinned_cap < total * sizeof(double)) {
        if (w->pinned) (void)hipHostFree(w->pinned);
        w->pinned = nullptr;
        w->pinned_cap = 0;
        FVG_HIP(hipHostMalloc(&w->pinned, std::max<size_t>(total, 1) * sizeof(double), hipHostMallocDefault));
        w->pinned_cap = std::max<size_t>(total, 1) * sizeof(double);
    }
    if (!w->ready) FVG_HIP(hipEventCreateWithFlags(&w->ready, hipEventDisableTiming));
    std::vector<FieldView> f;
    double* dst = static_cast<double*>(w->pinned);
    for (int32_t k = 0; k < n_fields; ++k) {
        FVG_HIP(hipMemcpyAsync(dst, values[k]->d, sizeof(double) * (size_t)values[k]->n, hipMemcpyDeviceToHost,
                               c->stream));
        f.push_back(FieldView{names[k], ncomps[k], dst});
        dst += values[k]->n;
    }
    FVG_HIP(hipEventRecord(w->ready, c->stream));
    const std::string p(path);
    w->worker = std::thread([w, p, f]() {
        if (hipEventSynchronize(w->ready) != hipSuccess) {
            w->status = FVG_E_HIP;
            w->error = "waiting for the field copies failed";
            return;
        }
        const int s = write_files(*w, p, f);
        if (s != FVG_OK) {
            w->status = s;
            w->error = fvg_last_error();
        }
    });
    return FVG_OK;
}

int fvg_pvtu_wait(fvg_pvtu_writer* w) {
    if (!w) return fail(FVG_E_INVALID, "null writer");
    return join(w);
}

int fvg_pvtu_writer_destroy(fvg_pvtu_writer* w) {
    if (!w) return FVG_OK;
    (void)join(w);
    if (w->pinned) (void)hipHostFree(w->pinned);
    if (w->ready) (void)hipEventDestroy(w->ready);
    delete w;
    return FVG_OK;
}

}  // extern "C"
