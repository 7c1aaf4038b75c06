// gmsh MSH 4.1 ASCII reader and node-data writer behind the uepm_msh_* entry
// points.  The reference reads the Brillouin-zone mesh through the gmsh API
// (MeshBZ::read_mesh_geometry_from_msh_file, src/BZ_MESH/bz_mesh.cpp:67-191),
// reads band views back (MeshBZ::read_mesh_bands_from_msh_file, :331-425) and
// writes them (MeshBZ::export_selected_bands_to_gmsh, :1426-1545).  gmsh is not
// a dependency here: the format is parsed directly.
//
// What is read: every $Nodes entity block (file order gives the vertex index),
// the 4-node tetrahedra (element type 4) of every 3-D entity block of
// $Elements, and every $NodeData view.  Other sections are skipped.
#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>
#include <memory>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "uepm.h"
#include "uepm_internal.hpp"

using uepm_host::fail;

struct uepm_msh {
    std::vector<uint64_t> tags;
    std::vector<double> xyz;
    std::vector<int32_t> tet;  // 0-based vertex indices, 4 per tetrahedron
    struct View {
        std::string name;
        int ncomp = 0;
        std::vector<double> data;  // n_nodes x ncomp, node order; NaN where the view has no value
    };
    std::vector<View> views;
    // tag -> vertex index
    uint64_t min_tag = 0;
    bool dense = true;
    std::unordered_map<uint64_t, int64_t> sparse;
    int64_t index_of(uint64_t tag) const {
        if (dense) {
            const uint64_t i = tag - min_tag;
            return (tag >= min_tag && i < tags.size()) ? static_cast<int64_t>(i) : -1;
        }
        auto it = sparse.find(tag);
        return it == sparse.end() ? -1 : it->second;
    }
};

namespace {

struct Cursor {
    const char* p;
    const char* end;
    bool eof() const { return p >= end; }
    void skip_ws() {
        while (p < end && (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n')) ++p;
    }
    bool line(std::string& out) {  // the rest of the current line, without the newline
        if (p >= end) return false;
        const char* s = p;
        while (p < end && *p != '\n') ++p;
        const char* e = p;
        if (p < end) ++p;
        while (e > s && (e[-1] == '\r' || e[-1] == ' ' || e[-1] == '\t')) --e;
        out.assign(s, e);
        return true;
    }
    bool u64(uint64_t& v) {
        skip_ws();
        if (p >= end) return false;
        char* q = nullptr;
        errno = 0;
        const unsigned long long x = std::strtoull(p, &q, 10);
        if (q == p || errno) return false;
        p = q;
        v = x;
        return true;
    }
    bool i64(int64_t& v) {
        skip_ws();
        if (p >= end) return false;
        char* q = nullptr;
        errno = 0;
        const long long x = std::strtoll(p, &q, 10);
        if (q == p || errno) return false;
        p = q;
        v = x;
        return true;
    }
    bool f64(double& v) {
        skip_ws();
        if (p >= end) return false;
        char* q = nullptr;
        const double x = std::strtod(p, &q);
        if (q == p) return false;
        p = q;
        v = x;
        return true;
    }
    // the most entries of at least `min_chars` characters that the rest of the input can hold
    uint64_t room(uint64_t min_chars) const { return static_cast<uint64_t>(end - p) / min_chars + 1; }
    void to_line_end() {
        while (p < end && *p != '\n') ++p;
        if (p < end) ++p;
    }
};

std::string unquote(const std::string& s) {
    if (s.size() >= 2 && s.front() == '"' && s.back() == '"') return s.substr(1, s.size() - 2);
    return s;
}

int parse_error(const std::string& path, const char* what) {
    return fail(UEPM_ERR_RUNTIME, "uepm_msh_read: " + path + ": " + what);
}

// Advances past the "$End<name>" line.
bool skip_section(Cursor& c, const std::string& name) {
    const std::string endtag = "$End" + name;
    std::string ln;
    while (c.line(ln))
        if (ln == endtag) return true;
    return false;
}

int read_nodes(Cursor& c, uepm_msh& m, const std::string& path) {
    uint64_t nblocks, nnodes, mintag, maxtag;
    if (!c.u64(nblocks) || !c.u64(nnodes) || !c.u64(mintag) || !c.u64(maxtag))
        return parse_error(path, "bad $Nodes header");
    const size_t base = m.tags.size();
    // a count read from the file reserves no more than the rest of the file can hold
    const uint64_t room = c.room(2);
    m.tags.reserve(base + std::min(nnodes, room));
    m.xyz.reserve(3 * (base + std::min(nnodes, room)));
    for (uint64_t b = 0; b < nblocks; ++b) {
        int64_t dim, etag, param;
        uint64_t n;
        if (!c.i64(dim) || !c.i64(etag) || !c.i64(param) || !c.u64(n)) return parse_error(path, "bad node block header");
        if (dim < 0 || dim > 3) return parse_error(path, "bad node block dimension");
        for (uint64_t i = 0; i < n; ++i) {
            uint64_t t;
            if (!c.u64(t)) return parse_error(path, "truncated node tags");
            m.tags.push_back(t);
        }
        const int extra = param ? static_cast<int>(dim) : 0;
        for (uint64_t i = 0; i < n; ++i) {
            double x, y, z, skip;
            if (!c.f64(x) || !c.f64(y) || !c.f64(z)) return parse_error(path, "truncated node coordinates");
            for (int e = 0; e < extra; ++e)
                if (!c.f64(skip)) return parse_error(path, "truncated parametric coordinates");
            m.xyz.push_back(x);
            m.xyz.push_back(y);
            m.xyz.push_back(z);
        }
    }
    if (m.tags.size() - base != nnodes) return parse_error(path, "node count differs from the $Nodes header");
    if (m.tags.size() > static_cast<size_t>(std::numeric_limits<int32_t>::max()))
        return parse_error(path, "more nodes than 32-bit vertex indices address");
    return skip_section(c, "Nodes") ? UEPM_OK : parse_error(path, "missing $EndNodes");
}

void index_tags(uepm_msh& m) {
    m.sparse.clear();
    if (m.tags.empty()) {
        m.dense = true;
        m.min_tag = 0;
        return;
    }
    bool consecutive = true;
    for (size_t i = 0; i < m.tags.size() && consecutive; ++i) consecutive = m.tags[i] == m.tags[0] + i;
    m.dense = consecutive;
    m.min_tag = m.tags[0];
    if (!consecutive) {
        m.sparse.reserve(m.tags.size() * 2);
        for (size_t i = 0; i < m.tags.size(); ++i) m.sparse.emplace(m.tags[i], static_cast<int64_t>(i));
    }
}

int read_elements(Cursor& c, uepm_msh& m, const std::string& path) {
    uint64_t nblocks, nelem, mintag, maxtag;
    if (!c.u64(nblocks) || !c.u64(nelem) || !c.u64(mintag) || !c.u64(maxtag))
        return parse_error(path, "bad $Elements header");
    c.to_line_end();
    std::string ln;
    for (uint64_t b = 0; b < nblocks; ++b) {
        int64_t dim, etag, type;
        uint64_t n;
        if (!c.i64(dim) || !c.i64(etag) || !c.i64(type) || !c.u64(n)) return parse_error(path, "bad element block header");
        c.to_line_end();
        if (dim != 3) {  // points, lines and surfaces: one element per line, not needed
            for (uint64_t i = 0; i < n; ++i)
                if (!c.line(ln)) return parse_error(path, "truncated element block");
            continue;
        }
        if (type != 4)
            return fail(UEPM_ERR_INVALID_ARG, "uepm_msh_read: " + path + ": volume element type " +
                                                  std::to_string(type) + " (only 4-node tetrahedra are supported)");
        m.tet.reserve(m.tet.size() + 4 * std::min(n, c.room(10)));
        for (uint64_t i = 0; i < n; ++i) {
            uint64_t et, v[4];
            if (!c.u64(et) || !c.u64(v[0]) || !c.u64(v[1]) || !c.u64(v[2]) || !c.u64(v[3]))
                return parse_error(path, "truncated tetrahedron");
            for (int j = 0; j < 4; ++j) {
                const int64_t idx = m.index_of(v[j]);
                if (idx < 0) return parse_error(path, "tetrahedron refers to a node tag that is not in $Nodes");
                m.tet.push_back(static_cast<int32_t>(idx));
            }
        }
    }
    return skip_section(c, "Elements") ? UEPM_OK : parse_error(path, "missing $EndElements");
}

int read_node_data(Cursor& c, uepm_msh& m, const std::string& path) {
    std::string ln;
    uepm_msh::View v;
    int64_t ns, nr, ni;
    if (!c.i64(ns)) return parse_error(path, "bad $NodeData string tags");
    c.to_line_end();
    for (int64_t i = 0; i < ns; ++i) {
        if (!c.line(ln)) return parse_error(path, "truncated $NodeData string tags");
        if (i == 0) v.name = unquote(ln);
    }
    if (!c.i64(nr)) return parse_error(path, "bad $NodeData real tags");
    for (int64_t i = 0; i < nr; ++i) {
        double t;
        if (!c.f64(t)) return parse_error(path, "truncated $NodeData real tags");
    }
    if (!c.i64(ni) || ni < 3) return parse_error(path, "bad $NodeData integer tags");
    int64_t it[8] = {0};
    for (int64_t i = 0; i < ni; ++i) {
        int64_t t;
        if (!c.i64(t)) return parse_error(path, "truncated $NodeData integer tags");
        if (i < 8) it[i] = t;
    }
    const int64_t ncomp = it[1], n = it[2];
    if (ncomp < 1 || ncomp > 9 || n < 0) return parse_error(path, "bad $NodeData component or entry count");
    v.ncomp = static_cast<int>(ncomp);
    v.data.assign(m.tags.size() * ncomp, std::numeric_limits<double>::quiet_NaN());
    for (int64_t i = 0; i < n; ++i) {
        uint64_t tag;
        if (!c.u64(tag)) return parse_error(path, "truncated $NodeData values");
        const int64_t idx = m.index_of(tag);
        if (idx < 0) return parse_error(path, "$NodeData refers to a node tag that is not in $Nodes");
        for (int64_t k = 0; k < ncomp; ++k)
            if (!c.f64(v.data[idx * ncomp + k])) return parse_error(path, "truncated $NodeData values");
    }
    m.views.push_back(std::move(v));
    return skip_section(c, "NodeData") ? UEPM_OK : parse_error(path, "missing $EndNodeData");
}

bool read_file(const char* path, std::string& buf) {
    std::ifstream in(path, std::ios::binary);
    if (!in) return false;
    std::ostringstream ss;
    ss << in.rdbuf();
    buf = ss.str();
    return true;
}

}  // namespace

extern "C" {

int uepm_msh_read(const char* path, uepm_msh** out) {
    if (!path || !out) return fail(UEPM_ERR_INVALID_ARG, "uepm_msh_read: null argument");
    *out = nullptr;
    std::string buf;
    if (!read_file(path, buf)) return fail(UEPM_ERR_RUNTIME, std::string("uepm_msh_read: cannot open ") + path);
    auto m = std::make_unique<uepm_msh>();
    Cursor c{buf.data(), buf.data() + buf.size()};
    std::string ln;
    bool have_format = false, have_nodes = false;
    try {
        while (c.line(ln)) {
            if (ln.empty() || ln[0] != '$') continue;
            const std::string name = ln.substr(1);
            int st = UEPM_OK;
            if (name == "MeshFormat") {
                std::string fl;
                if (!c.line(fl)) return parse_error(path, "truncated $MeshFormat");
                double ver = 0;
                int ftype = -1, dsize = 0;
                if (std::sscanf(fl.c_str(), "%lf %d %d", &ver, &ftype, &dsize) != 3)
                    return parse_error(path, "bad $MeshFormat line");
                if (ver < 4.0 || ver >= 5.0)
                    return fail(UEPM_ERR_INVALID_ARG, std::string("uepm_msh_read: ") + path +
                                                          ": MSH version " + fl + " (only 4.x is supported)");
                if (ftype != 0)
                    return fail(UEPM_ERR_INVALID_ARG,
                                std::string("uepm_msh_read: ") + path + ": binary MSH files are not supported");
                have_format = true;
                if (!skip_section(c, name)) return parse_error(path, "missing $EndMeshFormat");
            } else if (name == "Nodes") {
                if (!have_format) return parse_error(path, "$Nodes before $MeshFormat");
                st = read_nodes(c, *m, path);
                index_tags(*m);
                have_nodes = true;
            } else if (name == "Elements") {
                if (!have_nodes) return parse_error(path, "$Elements before $Nodes");
                st = read_elements(c, *m, path);
            } else if (name == "NodeData") {
                if (!have_nodes) return parse_error(path, "$NodeData before $Nodes");
                st = read_node_data(c, *m, path);
            } else if (name.rfind("End", 0) != 0) {
                if (!skip_section(c, name)) return parse_error(path, ("missing $End" + name).c_str());
            }
            if (st != UEPM_OK) return st;
        }
    } catch (const std::bad_alloc&) {
        return fail(UEPM_ERR_NO_MEMORY, "uepm_msh_read: out of host memory");
    }
    if (!have_format) return parse_error(path, "no $MeshFormat section");
    if (!have_nodes) return parse_error(path, "no $Nodes section");
    *out = m.release();
    return UEPM_OK;
}

int uepm_msh_counts(const uepm_msh* m, int64_t* n_nodes, int64_t* n_tetra, int32_t* n_views) {
    if (!m) return fail(UEPM_ERR_INVALID_ARG, "uepm_msh_counts: null mesh");
    if (n_nodes) *n_nodes = static_cast<int64_t>(m->tags.size());
    if (n_tetra) *n_tetra = static_cast<int64_t>(m->tet.size() / 4);
    if (n_views) *n_views = static_cast<int32_t>(m->views.size());
    return UEPM_OK;
}

int uepm_msh_nodes(const uepm_msh* m, uint64_t* tags, double* xyz) {
    if (!m) return fail(UEPM_ERR_INVALID_ARG, "uepm_msh_nodes: null mesh");
    if (tags) std::memcpy(tags, m->tags.data(), m->tags.size() * sizeof(uint64_t));
    if (xyz) std::memcpy(xyz, m->xyz.data(), m->xyz.size() * sizeof(double));
    return UEPM_OK;
}

int uepm_msh_tetrahedra(const uepm_msh* m, int32_t* vertex_index) {
    if (!m || !vertex_index) return fail(UEPM_ERR_INVALID_ARG, "uepm_msh_tetrahedra: null argument");
    std::memcpy(vertex_index, m->tet.data(), m->tet.size() * sizeof(int32_t));
    return UEPM_OK;
}

int uepm_msh_view_info(const uepm_msh* m, int32_t view, char* name, int32_t name_capacity, int32_t* ncomp) {
    if (!m) return fail(UEPM_ERR_INVALID_ARG, "uepm_msh_view_info: null mesh");
    if (view < 0 || view >= static_cast<int32_t>(m->views.size()))
        return fail(UEPM_ERR_OUT_OF_RANGE, "uepm_msh_view_info: view index out of range");
    const auto& v = m->views[view];
    if (name && name_capacity > 0) {
        const size_t n = std::min<size_t>(v.name.size(), static_cast<size_t>(name_capacity - 1));
        std::memcpy(name, v.name.data(), n);
        name[n] = '\0';
    }
    if (ncomp) *ncomp = v.ncomp;
    return UEPM_OK;
}

int uepm_msh_view_data(const uepm_msh* m, int32_t view, double* data) {
    if (!m || !data) return fail(UEPM_ERR_INVALID_ARG, "uepm_msh_view_data: null argument");
    if (view < 0 || view >= static_cast<int32_t>(m->views.size()))
        return fail(UEPM_ERR_OUT_OF_RANGE, "uepm_msh_view_data: view index out of range");
    const auto& v = m->views[view];
    std::memcpy(data, v.data.data(), v.data.size() * sizeof(double));
    return UEPM_OK;
}

void uepm_msh_free(uepm_msh* m) { delete m; }

int uepm_msh_append_node_view(const char* out_path, const char* mesh_path, const char* name, int32_t ncomp,
                              const uint64_t* tags, int64_t n, const double* data) {
    if (!out_path || !name || !tags || !data || n < 0 || ncomp < 1 || ncomp > 9)
        return fail(UEPM_ERR_INVALID_ARG, "uepm_msh_append_node_view: bad argument");
    std::FILE* probe = std::fopen(out_path, "rb");
    const bool exists = probe != nullptr;
    if (probe) std::fclose(probe);
    std::string head;
    if (!exists) {
        // A new file starts with the source mesh (every section but its data views).
        if (!mesh_path) return fail(UEPM_ERR_INVALID_ARG, "uepm_msh_append_node_view: new file needs the source mesh");
        std::string src;
        if (!read_file(mesh_path, src))
            return fail(UEPM_ERR_RUNTIME, std::string("uepm_msh_append_node_view: cannot open ") + mesh_path);
        Cursor c{src.data(), src.data() + src.size()};
        std::string ln;
        bool skipping = false;
        head.reserve(src.size());
        while (c.line(ln)) {
            if (!skipping && (ln == "$NodeData" || ln == "$ElementData" || ln == "$ElementNodeData")) {
                skipping = true;
                continue;
            }
            if (skipping) {
                if (ln == "$EndNodeData" || ln == "$EndElementData" || ln == "$EndElementNodeData") skipping = false;
                continue;
            }
            head += ln;
            head += '\n';
        }
    }
    std::FILE* f = std::fopen(out_path, exists ? "ab" : "wb");
    if (!f) return fail(UEPM_ERR_RUNTIME, std::string("uepm_msh_append_node_view: cannot write ") + out_path);
    if (!head.empty()) std::fwrite(head.data(), 1, head.size(), f);
    std::fprintf(f, "$NodeData\n1\n\"%s\"\n1\n0\n3\n0\n%d\n%lld\n", name, ncomp, static_cast<long long>(n));
    char line[512];
    for (int64_t i = 0; i < n; ++i) {
        int len = std::snprintf(line, sizeof line, "%llu", static_cast<unsigned long long>(tags[i]));
        for (int k = 0; k < ncomp; ++k)
            len += std::snprintf(line + len, sizeof line - len, " %.17g", data[i * ncomp + k]);
        line[len++] = '\n';
        std::fwrite(line, 1, len, f);
    }
    std::fputs("$EndNodeData\n", f);
    const bool ok = std::fclose(f) == 0;
    return ok ? UEPM_OK : fail(UEPM_ERR_RUNTIME, std::string("uepm_msh_append_node_view: write failed: ") + out_path);
}

}  // extern "C"
