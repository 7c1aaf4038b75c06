// uepm_bzmesh_*: the Brillouin-zone mesh driver of the band solve.  The
// reference's MeshBZ (src/BZ_MESH/bz_mesh.hpp) reads a gmsh tetrahedral mesh
// of the zone, solves the bands at its vertices (BandsOnBZ), fans irreducible-
// wedge results out to the whole zone, writes band views, reads them back and
// integrates the tetrahedron density of states (DOS_MeshBZ).  Host logic here
// follows MeshBZ method by method (file:line at each function); vertex
// positions, band energies and gradients also live on the device, where the
// density of states runs (tetra_dos.hip).  The eigenstates of BZ_States
// (src/BZ_MESH/bz_states.cpp) at the vertices and at vertex + q stay on the
// device too, for the tetrahedron-weighted dielectric function
// (dielectric.hip).
#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <limits>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "uepm_internal.hpp"

#pragma clang fp contract(off)

using uepm_host::DeviceBuffer;
using uepm_host::fail;

namespace {

struct BandGroup {
    int64_t count = 0, start = 0;
};

// Eigenpairs at every vertex (BZ_States::m_eigenvalues_k / m_eigenvectors_k,
// or the k + q pair), device-resident: E nv x nb, X nv x nb x dim complex.
struct VertexStates {
    bool set = false;
    int nb = 0, dim = 0, device = 0;
    DeviceBuffer E, X;
};

}  // namespace

struct uepm_bzmesh {
    uepm_ctx* ctx = nullptr;
    int64_t nv = 0, nt = 0;
    // geometry (host)
    std::vector<double> xyz;  // 1/m (or file units when not normalised)
    std::vector<uint64_t> tags;
    std::vector<int32_t> tet;
    std::vector<uint8_t> tet_in_wedge;
    bool has_wedge = false;
    double lattice_m = 0.0;  // 0: unknown (mesh built from arrays)
    double total_volume = 0.0, reduce_bz_factor = 1.0;
    std::string filename;
    // irreducible wedge -> zone
    std::vector<std::vector<int64_t>> kstar;
    std::vector<int64_t> iw_list;
    // bands (host): energies and gradients per band over all vertices
    std::vector<std::vector<double>> Eband;  // [band][vertex]
    std::vector<std::vector<double>> Gband;  // [band][3 * vertex]
    std::vector<int> band_type;              // 0 valence, 1 conduction
    std::vector<int64_t> band_local;
    BandGroup valence, conduction;
    std::vector<double> min_band, max_band;
    bool host_bands = false;
    double spin_degeneracy = 2.0;
    // device (geometry uploaded on first device use)
    bool dev_geometry = false;
    int nb = 0;            // bands resident on the device
    bool dev_stale = true; // host bands changed since the last upload
    DeviceBuffer pos, dtet, det, wedge, Eb, Gb, stage, window, partial, out;
    // BZ_States: [0] at the vertices, [1] at vertex + q; q in 1/m
    VertexStates states[2];
    double q_shift_si[3] = {0.0, 0.0, 0.0};
};

namespace {

constexpr double kTwoPi = 2.0 * M_PI;

int dos_checks(const uepm_bzmesh* m, int32_t band, int32_t n_points, double e_min, double e_max, int wedge_only,
               const char* who) {
    if (!m) return fail(UEPM_ERR_INVALID_ARG, std::string(who) + ": null mesh");
    if (m->nb == 0 && !m->host_bands) return fail(UEPM_ERR_INVALID_ARG, std::string(who) + ": no bands set on the mesh");
    if (n_points < 2) return fail(UEPM_ERR_INVALID_ARG, std::string(who) + ": n_points must be at least 2");
    if (!std::isfinite(e_min) || !std::isfinite(e_max) || !(e_max > e_min))
        return fail(UEPM_ERR_INVALID_ARG, std::string(who) + ": energy range must be finite with e_max > e_min");
    if (wedge_only && !m->has_wedge)
        return fail(UEPM_ERR_INVALID_ARG, std::string(who) + ": wedge-only DOS needs the tetrahedron wedge flags");
    (void)band;
    return UEPM_OK;
}

// MeshBZ::is_irreducible_wedge (bz_mesh.cpp:1176-1183), k in the mesh's units.
bool in_wedge(const uepm_bzmesh& m, const double* k) {
    const double s = m.lattice_m / kTwoPi;
    const double x = k[0] * s, y = k[1] * s, z = k[2] * s;
    constexpr double eps = 1e-12;
    return (z >= -eps) && (z <= y + eps) && (y <= x + eps) && (x <= 1.0 + eps) && ((x + y + z) <= 1.5 + eps);
}

// Tetra::compute_signed_volume (mesh_tetra.cpp:128-132) of tetrahedron (a,b,c,d).
double signed_volume(const double* P, const int32_t* v) {
    const double* p0 = P + 3 * v[0];
    double e[3][3];
    for (int j = 0; j < 3; ++j)
        for (int c = 0; c < 3; ++c) e[j][c] = P[3 * v[j + 1] + c] - p0[c];
    const double cx = e[1][1] * e[2][2] - e[1][2] * e[2][1];
    const double cy = e[1][2] * e[2][0] - e[1][0] * e[2][2];
    const double cz = e[1][0] * e[2][1] - e[1][1] * e[2][0];
    return (1.0 / 6.0) * (e[0][0] * cx + e[0][1] * cy + e[0][2] * cz);
}

int upload_geometry(uepm_bzmesh& m) {
    if (m.dev_geometry) return UEPM_OK;
    if (!m.ctx) return fail(UEPM_ERR_INVALID_ARG, "uepm_bzmesh: mesh created without a device context");
    HIP_OK(hipSetDevice(m.ctx->device));
    HIP_OK(m.pos.ensure(sizeof(double) * 3 * std::max<int64_t>(m.nv, 1)));
    HIP_OK(m.dtet.ensure(sizeof(int32_t) * 4 * std::max<int64_t>(m.nt, 1)));
    HIP_OK(m.det.ensure(sizeof(double) * std::max<int64_t>(m.nt, 1)));
    if (m.nv) HIP_OK(hipMemcpy(m.pos.p, m.xyz.data(), sizeof(double) * 3 * m.nv, hipMemcpyHostToDevice));
    if (m.nt) HIP_OK(hipMemcpy(m.dtet.p, m.tet.data(), sizeof(int32_t) * 4 * m.nt, hipMemcpyHostToDevice));
    if (m.has_wedge) {
        HIP_OK(m.wedge.ensure(std::max<int64_t>(m.nt, 1)));
        if (m.nt) HIP_OK(hipMemcpy(m.wedge.p, m.tet_in_wedge.data(), m.nt, hipMemcpyHostToDevice));
    }
    HIP_OK(uepm::launch_tetra_det(m.pos.as<double>(), m.dtet.as<int32_t>(), m.nt, m.det.as<double>(), nullptr));
    HIP_OK(hipDeviceSynchronize());
    m.dev_geometry = true;
    return UEPM_OK;
}

// Host bands -> device (band-major), when they changed.
int sync_bands(uepm_bzmesh& m) {
    if (!m.host_bands || !m.dev_stale) return UEPM_OK;
    const int sg = upload_geometry(m);
    if (sg != UEPM_OK) return sg;
    const int nb = static_cast<int>(m.Eband.size());
    if (static_cast<int>(m.Gband.size()) < nb)
        return fail(UEPM_ERR_RUNTIME, "uepm_bzmesh: fewer gradient bands than energy bands at the vertices");
    HIP_OK(hipSetDevice(m.ctx->device));
    const int64_t nv = std::max<int64_t>(m.nv, 1);
    HIP_OK(m.Eb.ensure(sizeof(double) * std::max(nb, 1) * nv));
    HIP_OK(m.Gb.ensure(sizeof(double) * 3 * std::max(nb, 1) * nv));
    for (int b = 0; b < nb; ++b) {
        HIP_OK(hipMemcpy(m.Eb.as<double>() + b * m.nv, m.Eband[b].data(), sizeof(double) * m.nv,
                         hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(m.Gb.as<double>() + 3 * b * m.nv, m.Gband[b].data(), sizeof(double) * 3 * m.nv,
                         hipMemcpyHostToDevice));
    }
    m.nb = nb;
    m.dev_stale = false;
    return UEPM_OK;
}

// MeshBZ::recompute_min_max_energies (bz_mesh.cpp:513-530)
void recompute_min_max(uepm_bzmesh& m) {
    const size_t nb = m.Eband.size();
    m.min_band.assign(nb, std::numeric_limits<double>::max());
    m.max_band.assign(nb, std::numeric_limits<double>::lowest());
    for (size_t b = 0; b < nb; ++b)
        for (double e : m.Eband[b]) {
            if (e < m.min_band[b]) m.min_band[b] = e;
            if (e > m.max_band[b]) m.max_band[b] = e;
        }
}

void classify_append(uepm_bzmesh& m, bool is_valence, int64_t global) {
    if (is_valence) {
        m.band_type.push_back(0);
        m.band_local.push_back(m.valence.count);
        if (m.valence.count == 0) m.valence.start = global;
        m.valence.count++;
    } else {
        m.band_type.push_back(1);
        m.band_local.push_back(m.conduction.count);
        if (m.conduction.count == 0) m.conduction.start = global;
        m.conduction.count++;
    }
}

// Band info after a solve (compute_band_structure_over_mesh, bz_mesh.cpp:1283-1301):
// valence when the band minimum is below 1e-6 eV.
void classify_by_minimum(uepm_bzmesh& m) {
    m.band_type.clear();
    m.band_local.clear();
    m.valence = {};
    m.conduction = {};
    constexpr double eps = 1e-6;
    for (size_t b = 0; b < m.min_band.size(); ++b) classify_append(m, m.min_band[b] < eps, static_cast<int64_t>(b));
}

// apply_same_symmetry_operation (bz_mesh.cpp:1314-1381): the signed
// permutation that takes k_iw to k_bz, found greedily by |component|, applied
// to v; v unchanged when there is none.
void same_symmetry(const double* a, const double* b, const double* v, double* out) {
    constexpr double eps = 1e-8;
    int perm[3] = {-1, -1, -1}, sign[3] = {1, 1, 1};
    bool used[3] = {false, false, false};
    double scale = 0.0;
    for (int i = 0; i < 3; ++i) scale = std::max(scale, std::abs(a[i]));
    for (int i = 0; i < 3; ++i) scale = std::max(scale, std::abs(b[i]));
    if (scale < 1.0) scale = 1.0;
    const double tol = eps * scale;
    for (int jb = 0; jb < 3; ++jb) {
        bool found = false;
        for (int ia = 0; ia < 3; ++ia) {
            if (used[ia] || !(std::abs(std::abs(a[ia]) - std::abs(b[jb])) <= tol)) continue;
            perm[jb] = ia;
            used[ia] = true;
            if (std::abs(a[ia]) > tol)
                sign[jb] = ((b[jb] >= 0.0) ? 1 : -1) * ((a[ia] >= 0.0) ? 1 : -1);
            else
                sign[jb] = (b[jb] >= 0.0) ? 1 : -1;
            found = true;
            break;
        }
        if (!found) {
            out[0] = v[0];
            out[1] = v[1];
            out[2] = v[2];
            return;
        }
    }
    for (int j = 0; j < 3; ++j) out[j] = sign[j] * v[perm[j]];
}

std::string kstar_default_name(const std::string& mesh_path) {
    // "<stem>_kstar_ibz_to_bz.txt" in the working directory (bz_mesh.cpp:203-204)
    std::string base = mesh_path;
    const size_t slash = base.find_last_of('/');
    if (slash != std::string::npos) base = base.substr(slash + 1);
    const size_t dot = base.find_last_of('.');
    if (dot != std::string::npos && dot > 0) base = base.substr(0, dot);
    return base + "_kstar_ibz_to_bz.txt";
}

// The 48 signed permutation matrices of O_h in the order of
// symmetry_ops_full (apps/bz_meshing.cpp:98-145): sorted by their row-major
// entries, then reversed, so the identity comes first.
std::vector<std::array<int, 9>> oh_operations() {
    std::vector<std::array<int, 9>> ops;
    const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    for (const auto& p : perms)
        for (int s = 0; s < 8; ++s) {
            std::array<int, 9> M{};
            for (int r = 0; r < 3; ++r) M[3 * r + p[r]] = (s >> r & 1) ? -1 : 1;
            ops.push_back(M);
        }
    std::sort(ops.begin(), ops.end());
    std::reverse(ops.begin(), ops.end());
    return ops;
}

struct KeyHash {
    size_t operator()(const std::array<uint64_t, 3>& k) const {
        return std::hash<uint64_t>()(k[0] * 0x9E3779B97F4A7C15ull ^ (k[1] + 0x632BE59BD9B4E019ull) * 31 ^ k[2]);
    }
};
std::array<uint64_t, 3> coord_key(double x, double y, double z) {
    std::array<uint64_t, 3> k;
    const double c[3] = {x + 0.0, y + 0.0, z + 0.0};  // -0 -> +0
    std::memcpy(k.data(), c, sizeof c);
    return k;
}

int write_text(const std::string& path, const std::string& text) {
    std::ofstream out(path, std::ios::binary);
    if (!out) return fail(UEPM_ERR_RUNTIME, "uepm_bzmesh: cannot write " + path);
    out << text;
    return out ? UEPM_OK : fail(UEPM_ERR_RUNTIME, "uepm_bzmesh: write failed: " + path);
}

std::string fmt_g8(double v) {
    std::ostringstream s;
    s.precision(8);
    s << v;
    return s.str();
}

}  // namespace

extern "C" {

/* ---- construction -------------------------------------------------------- */

int uepm_bzmesh_create(uepm_ctx* ctx, const double* xyz, int64_t nv, const int32_t* tet, int64_t nt,
                       const uint8_t* tetra_in_wedge, uepm_bzmesh** out) {
    if (!out || (nv > 0 && !xyz) || (nt > 0 && !tet) || nv < 0 || nt < 0)
        return fail(UEPM_ERR_INVALID_ARG, "uepm_bzmesh_create: bad argument");
    *out = nullptr;
    if (nv > INT32_MAX) return fail(UEPM_ERR_INVALID_ARG, "uepm_bzmesh_create: more than 2^31-1 vertices");
    for (int64_t i = 0; i < 4 * nt; ++i)
        if (tet[i] < 0 || tet[i] >= nv)
            return fail(UEPM_ERR_INVALID_ARG, "uepm_bzmesh_create: tetrahedron vertex index out of range");
    for (int64_t i = 0; i < 3 * nv; ++i)
        if (!std::isfinite(xyz[i])) return fail(UEPM_ERR_INVALID_ARG, "uepm_bzmesh_create: non-finite coordinate");
    auto m = std::make_unique<uepm_bzmesh>();
    m->ctx = ctx;
    m->nv = nv;
    m->nt = nt;
    m->xyz.assign(xyz, xyz + 3 * nv);
    m->tet.assign(tet, tet + 4 * nt);
    m->tags.resize(nv);
    for (int64_t i = 0; i < nv; ++i) m->tags[i] = static_cast<uint64_t>(i + 1);
    m->has_wedge = tetra_in_wedge != nullptr;
    if (m->has_wedge) m->tet_in_wedge.assign(tetra_in_wedge, tetra_in_wedge + nt);
    for (int64_t t = 0; t < nt; ++t) m->total_volume += std::fabs(signed_volume(m->xyz.data(), &m->tet[4 * t]));
    *out = m.release();
    return UEPM_OK;
}

// MeshBZ::read_mesh_geometry_from_msh_file (bz_mesh.cpp:67-191) up to, not
// including, its final load_kstar_ibz_to_bz call (uepm_bzmesh_load_kstar).
int uepm_bzmesh_read_msh(uepm_ctx* ctx, const char* path, double lattice_constant_m, int normalize_by_fourier_factor,
                         uepm_bzmesh** out) {
    if (!path || !out || !(lattice_constant_m > 0.0))
        return fail(UEPM_ERR_INVALID_ARG, "uepm_bzmesh_read_msh: bad argument");
    *out = nullptr;
    uepm_msh* f = nullptr;
    int st = uepm_msh_read(path, &f);
    if (st != UEPM_OK) return st;
    std::unique_ptr<uepm_msh, void (*)(uepm_msh*)> guard(f, uepm_msh_free);
    int64_t nn = 0, ntf = 0;
    uepm_msh_counts(f, &nn, &ntf, nullptr);
    if (ntf == 0) return fail(UEPM_ERR_RUNTIME, "ElementTags is zero when the mesh was imported... Abort.");
    auto m = std::make_unique<uepm_bzmesh>();
    m->ctx = ctx;
    m->filename = path;
    m->lattice_m = lattice_constant_m;
    m->nv = nn;
    m->tags.resize(nn);
    m->xyz.resize(3 * nn);
    uepm_msh_nodes(f, m->tags.data(), m->xyz.data());
    const double fourier_factor = 2.0 * M_PI / lattice_constant_m;
    const double factor = normalize_by_fourier_factor ? fourier_factor : 1.0;
    for (auto& c : m->xyz) c = factor * c;
    std::vector<int32_t> raw(4 * ntf);
    uepm_msh_tetrahedra(f, raw.data());
    // degenerate tetrahedra (|signed volume| < 1e-12) are dropped (bz_mesh.cpp:139-149)
    constexpr double vol_threshold = 1e-12;
    m->tet.reserve(raw.size());
    for (int64_t t = 0; t < ntf; ++t) {
        const int32_t* v = &raw[4 * t];
        const double vol = signed_volume(m->xyz.data(), v);
        if (std::abs(vol) < vol_threshold) continue;
        double bary[3];
        for (int c = 0; c < 3; ++c)
            bary[c] = (m->xyz[3 * v[0] + c] + m->xyz[3 * v[1] + c] + m->xyz[3 * v[2] + c] + m->xyz[3 * v[3] + c]) / 4.0;
        m->tet_in_wedge.push_back(in_wedge(*m, bary) ? 1 : 0);
        m->tet.insert(m->tet.end(), v, v + 4);
        m->total_volume += std::fabs(vol);
    }
    m->nt = static_cast<int64_t>(m->tet.size() / 4);
    m->has_wedge = true;
    const double Vcell = std::pow(lattice_constant_m, 3) / 4.0;
    const double VBZ_theory = std::pow(2.0 * M_PI, 3) / Vcell;
    m->reduce_bz_factor = m->total_volume / VBZ_theory;
    *out = m.release();
    return UEPM_OK;
}

int uepm_bzmesh_destroy(uepm_bzmesh* m) {
    if (!m) return UEPM_OK;
    if (m->ctx && m->dev_geometry) (void)hipSetDevice(m->ctx->device);
    delete m;
    return UEPM_OK;
}

int uepm_bzmesh_counts(const uepm_bzmesh* m, int64_t* n_vertices, int64_t* n_tetra, int32_t* n_bands) {
    if (!m) return fail(UEPM_ERR_INVALID_ARG, "uepm_bzmesh_counts: null mesh");
    if (n_vertices) *n_vertices = m->nv;
    if (n_tetra) *n_tetra = m->nt;
    if (n_bands) *n_bands = m->host_bands ? static_cast<int32_t>(m->Eband.size()) : m->nb;
    return UEPM_OK;
}

int uepm_bzmesh_geometry(const uepm_bzmesh* m, double* xyz, int32_t* tet, uint8_t* tetra_in_wedge, uint64_t* tags,
                         double* total_volume, double* reduce_bz_factor) {
    if (!m) return fail(UEPM_ERR_INVALID_ARG, "uepm_bzmesh_geometry: null mesh");
    if (xyz) std::memcpy(xyz, m->xyz.data(), sizeof(double) * m->xyz.size());
    if (tet) std::memcpy(tet, m->tet.data(), sizeof(int32_t) * m->tet.size());
    if (tetra_in_wedge && m->has_wedge) std::memcpy(tetra_in_wedge, m->tet_in_wedge.data(), m->tet_in_wedge.size());
    if (tags) std::memcpy(tags, m->tags.data(), sizeof(uint64_t) * m->tags.size());
    if (total_volume) *total_volume = m->total_volume;
    if (reduce_bz_factor) *reduce_bz_factor = m->reduce_bz_factor;
    return UEPM_OK;
}

/* ---- irreducible wedge --------------------------------------------------- */

// MeshBZ::load_kstar_ibz_to_bz (bz_mesh.cpp:198-272).
int uepm_bzmesh_load_kstar(uepm_bzmesh* m, const char* path) {
    if (!m) return fail(UEPM_ERR_INVALID_ARG, "uepm_bzmesh_load_kstar: null mesh");
    if (!(m->lattice_m > 0.0))
        return fail(UEPM_ERR_INVALID_ARG, "uepm_bzmesh_load_kstar: mesh has no lattice constant (not read from a file)");
    const std::string name = path ? std::string(path) : kstar_default_name(m->filename);
    std::ifstream in(name);
    if (!in) return fail(UEPM_ERR_RUNTIME, "load_kstar_file: can't find kstar_ibz_to_bz file");
    struct Row {
        int64_t iw;
        std::vector<int64_t> ids;
    };
    std::vector<Row> rows;
    int64_t max_iw = 0;
    std::string line;
    while (std::getline(in, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream iss(line);
        int64_t iw, cnt;
        if (!(iss >> iw >> cnt) || iw < 0 || cnt < 0) continue;
        std::vector<int64_t> ids(cnt);
        for (int64_t k = 0; k < cnt; ++k) iss >> ids[k];
        max_iw = std::max(max_iw, iw);
        rows.push_back({iw, std::move(ids)});
    }
    std::vector<std::vector<int64_t>> kstar(max_iw + 1);
    std::vector<int64_t> iw_list;
    int64_t count = 0;
    for (auto& r : rows) {
        if (r.iw >= m->nv || !in_wedge(*m, &m->xyz[3 * r.iw]))
            return fail(UEPM_ERR_RUNTIME, "load_kstar_file: vertex " + std::to_string(r.iw) +
                                              " does not lie in the irreducible wedge, but it should");
        count += static_cast<int64_t>(r.ids.size());
        kstar[r.iw] = std::move(r.ids);
        iw_list.push_back(r.iw);
    }
    if (count != m->nv)
        return fail(UEPM_ERR_RUNTIME, "load_kstar_file: number of k-points in IBZ (" + std::to_string(count) +
                                          ") does not match number of vertices in mesh (" + std::to_string(m->nv) +
                                          ")");
    for (const auto& ids : kstar)
        for (int64_t id : ids)
            if (id < 0 || id >= m->nv)
                return fail(UEPM_ERR_RUNTIME, "load_kstar_file: index " + std::to_string(id) + " out of range");
    m->kstar = std::move(kstar);
    m->iw_list = std::move(iw_list);
    return UEPM_OK;
}

// Writes the k-star file for this mesh: each wedge vertex (in vertex order)
// with its orbit under the 48 operations of O_h, deduplicated in operation
// order, as apps/bz_meshing.cpp:600-660 writes it for the meshes it builds.
// Every orbit point must be a vertex of the mesh (exact coordinates).
int uepm_bzmesh_write_kstar(const uepm_bzmesh* m, const char* path) {
    if (!m || !path) return fail(UEPM_ERR_INVALID_ARG, "uepm_bzmesh_write_kstar: bad argument");
    if (!(m->lattice_m > 0.0))
        return fail(UEPM_ERR_INVALID_ARG, "uepm_bzmesh_write_kstar: mesh has no lattice constant (not read from a file)");
    std::unordered_map<std::array<uint64_t, 3>, int64_t, KeyHash> index;
    index.reserve(2 * m->nv);
    for (int64_t v = 0; v < m->nv; ++v) index.emplace(coord_key(m->xyz[3 * v], m->xyz[3 * v + 1], m->xyz[3 * v + 2]), v);
    const auto ops = oh_operations();
    std::string text = "# ops=" + std::to_string(ops.size()) + " symPts=" + std::to_string(m->nv) + "\n";
    for (int64_t v = 0; v < m->nv; ++v) {
        const double* p = &m->xyz[3 * v];
        if (!in_wedge(*m, p)) continue;
        std::vector<int64_t> orbit;
        for (const auto& M : ops) {
            double q[3];
            for (int r = 0; r < 3; ++r) q[r] = M[3 * r] * p[0] + M[3 * r + 1] * p[1] + M[3 * r + 2] * p[2];
            auto it = index.find(coord_key(q[0], q[1], q[2]));
            if (it == index.end())
                return fail(UEPM_ERR_RUNTIME, "uepm_bzmesh_write_kstar: the image of vertex " + std::to_string(v) +
                                                  " under a point-group operation is not a mesh vertex");
            if (std::find(orbit.begin(), orbit.end(), it->second) == orbit.end()) orbit.push_back(it->second);
        }
        text += std::to_string(v) + " " + std::to_string(orbit.size());
        for (int64_t id : orbit) text += " " + std::to_string(id);
        text += "\n";
    }
    return write_text(path, text);
}

int uepm_bzmesh_wedge_vertices(const uepm_bzmesh* m, int64_t* ids, int64_t* count) {
    if (!m || !count) return fail(UEPM_ERR_INVALID_ARG, "uepm_bzmesh_wedge_vertices: bad argument");
    *count = static_cast<int64_t>(m->iw_list.size());
    if (ids) std::memcpy(ids, m->iw_list.data(), sizeof(int64_t) * m->iw_list.size());
    return UEPM_OK;
}

/* ---- bands over the mesh ------------------------------------------------- */

// The k-points compute_band_structure_over_mesh solves at (bz_mesh.cpp:1236-1252):
// the wedge vertices in k-star file order, or every vertex, in reduced units.
int uepm_bzmesh_kpoints(const uepm_bzmesh* m, int use_iwedge, double* k, int64_t* count) {
    if (!m || !count) return fail(UEPM_ERR_INVALID_ARG, "uepm_bzmesh_kpoints: bad argument");
    if (!(m->lattice_m > 0.0)) return fail(UEPM_ERR_INVALID_ARG, "uepm_bzmesh_kpoints: mesh has no lattice constant");
    if (use_iwedge && m->iw_list.empty())
        return fail(UEPM_ERR_RUNTIME, "No vertices in irreducible wedge. Load the k-star file first.");
    const int64_t n = use_iwedge ? static_cast<int64_t>(m->iw_list.size()) : m->nv;
    *count = n;
    if (!k) return UEPM_OK;
    const double s = m->lattice_m / (2.0 * M_PI);
    for (int64_t i = 0; i < n; ++i) {
        const int64_t v = use_iwedge ? m->iw_list[i] : i;
        for (int c = 0; c < 3; ++c) k[3 * i + c] = m->xyz[3 * v + c] * s;
    }
    return UEPM_OK;
}

// Energies E (n_used x nb) and gradients G (n_used x nb x 3) at the k-points
// of uepm_bzmesh_kpoints, already adjusted: set at the vertices, fanned out
// from the wedge (distribute_energies_from_iw_wedge_to_full_bz,
// bz_mesh.cpp:1383-1424), then band extrema and valence/conduction info
// (:1280-1301).
int uepm_bzmesh_set_mesh_bands(uepm_bzmesh* m, int use_iwedge, const double* E, const double* G, int32_t nb) {
    if (!m || !E || !G || nb < 1) return fail(UEPM_ERR_INVALID_ARG, "uepm_bzmesh_set_mesh_bands: bad argument");
    if (use_iwedge && m->iw_list.empty())
        return fail(UEPM_ERR_RUNTIME, "No vertices in irreducible wedge. Cannot distribute energies to full BZ.");
    const int64_t n = use_iwedge ? static_cast<int64_t>(m->iw_list.size()) : m->nv;
    m->Eband.assign(nb, std::vector<double>(m->nv, 0.0));
    m->Gband.assign(nb, std::vector<double>(3 * m->nv, 0.0));
    std::vector<uint8_t> set(m->nv, 0);
    for (int64_t i = 0; i < n; ++i) {
        const int64_t v = use_iwedge ? m->iw_list[i] : i;
        set[v] = 1;
        for (int b = 0; b < nb; ++b) {
            m->Eband[b][v] = E[i * nb + b];
            for (int c = 0; c < 3; ++c) m->Gband[b][3 * v + c] = G[3 * (i * nb + b) + c];
        }
    }
    if (use_iwedge) {
        for (int64_t iw : m->iw_list) {
            const double* piw = &m->xyz[3 * iw];
            for (int64_t id : m->kstar[iw]) {
                if (id == iw) continue;
                set[id] = 1;
                const double* pbz = &m->xyz[3 * id];
                for (int b = 0; b < nb; ++b) {
                    m->Eband[b][id] = m->Eband[b][iw];
                    same_symmetry(piw, pbz, &m->Gband[b][3 * iw], &m->Gband[b][3 * id]);
                }
            }
        }
    }
    for (int64_t v = 0; v < m->nv; ++v)
        if (!set[v]) return fail(UEPM_ERR_RUNTIME, "uepm_bzmesh_set_mesh_bands: vertex " + std::to_string(v) +
                                                       " received no band energies");
    m->host_bands = true;
    m->dev_stale = true;
    recompute_min_max(*m);
    classify_by_minimum(*m);
    return UEPM_OK;
}

// compute_band_structure_over_mesh (bz_mesh.cpp:1230-1303) with a plan that
// carries UEPM_GRADIENTS: solve at the mesh k-points, AdjustValues(false)
// (BandStructure.cpp:203-226), then uepm_bzmesh_set_mesh_bands.
int uepm_bzmesh_compute_bands(uepm_bzmesh* m, uepm_plan* plan, int use_iwedge, double* band_gap) {
    if (!m || !plan) return fail(UEPM_ERR_INVALID_ARG, "uepm_bzmesh_compute_bands: bad argument");
    if (!(plan->flags & UEPM_GRADIENTS))
        return fail(UEPM_ERR_INVALID_ARG, "uepm_bzmesh_compute_bands: the plan must carry UEPM_GRADIENTS");
    int64_t n = 0;
    int st = uepm_bzmesh_kpoints(m, use_iwedge, nullptr, &n);
    if (st != UEPM_OK) return st;
    std::vector<double> k(3 * n);
    uepm_bzmesh_kpoints(m, use_iwedge, k.data(), &n);
    const int nb = plan->nb;
    std::vector<double> E(n * nb), G(3 * n * nb);
    st = uepm_solve(plan, k.data(), n, E.data(), nullptr, G.data());
    if (st != UEPM_OK) return st;
    double gap = 0.0;
    st = uepm_adjust_values(E.data(), n, nb, (plan->flags & UEPM_SOC) ? 1 : 0, 0, &gap);
    if (st != UEPM_OK) return st;
    if (band_gap) *band_gap = gap;
    return uepm_bzmesh_set_mesh_bands(m, use_iwedge, E.data(), G.data(), nb);
}

int uepm_bzmesh_bands(const uepm_bzmesh* m, double* E, double* G, int32_t* nb) {
    if (!m || !nb) return fail(UEPM_ERR_INVALID_ARG, "uepm_bzmesh_bands: bad argument");
    if (!m->host_bands) return fail(UEPM_ERR_INVALID_ARG, "uepm_bzmesh_bands: no host bands on the mesh");
    const int n = static_cast<int>(m->Eband.size());
    *nb = n;
    for (int64_t v = 0; v < m->nv; ++v)
        for (int b = 0; b < n; ++b) {
            if (E) E[v * n + b] = m->Eband[b][v];
            if (G)
                for (int c = 0; c < 3; ++c)
                    G[3 * (v * n + b) + c] = b < static_cast<int>(m->Gband.size()) ? m->Gband[b][3 * v + c] : 0.0;
        }
    return UEPM_OK;
}

int uepm_bzmesh_band_info(const uepm_bzmesh* m, int32_t band, int32_t* is_conduction, int64_t* local_index,
                          double* min_energy, double* max_energy) {
    if (!m) return fail(UEPM_ERR_INVALID_ARG, "uepm_bzmesh_band_info: null mesh");
    if (!m->host_bands || band < 0 || band >= static_cast<int32_t>(m->band_type.size()))
        return fail(UEPM_ERR_OUT_OF_RANGE, "uepm_bzmesh_band_info: band index out of range");
    if (is_conduction) *is_conduction = m->band_type[band];
    if (local_index) *local_index = m->band_local[band];
    if (min_energy) *min_energy = m->min_band[band];
    if (max_energy) *max_energy = m->max_band[band];
    return UEPM_OK;
}

int uepm_bzmesh_band_groups(const uepm_bzmesh* m, int64_t* n_valence, int64_t* valence_start, int64_t* n_conduction,
                            int64_t* conduction_start) {
    if (!m) return fail(UEPM_ERR_INVALID_ARG, "uepm_bzmesh_band_groups: null mesh");
    if (n_valence) *n_valence = m->valence.count;
    if (valence_start) *valence_start = m->valence.start;
    if (n_conduction) *n_conduction = m->conduction.count;
    if (conduction_start) *conduction_start = m->conduction.start;
    return UEPM_OK;
}

int uepm_bzmesh_set_bands(uepm_bzmesh* m, const double* E, const double* G, int32_t nb) {
    if (!m || !E || !G || nb < 1)
        return fail(UEPM_ERR_INVALID_ARG, "uepm_bzmesh_set_bands: bad argument (energies and gradients needed)");
    for (int64_t i = 0; i < m->nv * nb; ++i)
        if (!std::isfinite(E[i])) return fail(UEPM_ERR_INVALID_ARG, "uepm_bzmesh_set_bands: non-finite energy");
    for (int64_t i = 0; i < 3 * m->nv * nb; ++i)
        if (!std::isfinite(G[i])) return fail(UEPM_ERR_INVALID_ARG, "uepm_bzmesh_set_bands: non-finite gradient");
    m->Eband.assign(nb, std::vector<double>(m->nv));
    m->Gband.assign(nb, std::vector<double>(3 * m->nv));
    for (int64_t v = 0; v < m->nv; ++v)
        for (int b = 0; b < nb; ++b) {
            m->Eband[b][v] = E[v * nb + b];
            for (int c = 0; c < 3; ++c) m->Gband[b][3 * v + c] = G[3 * (v * nb + b) + c];
        }
    m->host_bands = true;
    m->dev_stale = true;
    recompute_min_max(*m);
    classify_by_minimum(*m);
    return UEPM_OK;
}

int uepm_bzmesh_set_bands_device(uepm_bzmesh* m, const double* d_E, const double* d_G, int32_t ld, int32_t first,
                                 int32_t nb, void* stream) {
    if (!m || !d_E || !d_G || nb < 1 || first < 0 || ld < first + nb)
        return fail(UEPM_ERR_INVALID_ARG, "uepm_bzmesh_set_bands_device: bad argument (energies and gradients needed)");
    const int sg = upload_geometry(*m);
    if (sg != UEPM_OK) return sg;
    HIP_OK(hipSetDevice(m->ctx->device));
    const int64_t nv = std::max<int64_t>(m->nv, 1);
    HIP_OK(m->Eb.ensure(sizeof(double) * nb * nv));
    HIP_OK(m->Gb.ensure(sizeof(double) * 3 * nb * nv));
    HIP_OK(uepm::launch_bands_to_band_major(d_E, d_G, m->nv, ld, first, nb, m->Eb.as<double>(), m->Gb.as<double>(),
                                            static_cast<hipStream_t>(stream)));
    m->nb = nb;
    m->host_bands = false;  // the device copy is now the only one
    m->dev_stale = false;
    return UEPM_OK;
}

/* ---- density of states --------------------------------------------------- */

int uepm_bzmesh_dos_device(uepm_bzmesh* m, int32_t band, int32_t n_points, double e_min, double e_max,
                           double reduce_bz_factor, double spin_degeneracy, int wedge_only, double* d_energies,
                           double* d_dos, void* stream) {
    int st = dos_checks(m, band, n_points, e_min, e_max, wedge_only, "uepm_bzmesh_dos_device");
    if (st != UEPM_OK) return st;
    if (!d_dos) return fail(UEPM_ERR_INVALID_ARG, "uepm_bzmesh_dos_device: null output");
    st = upload_geometry(*m);
    if (st == UEPM_OK) st = sync_bands(*m);
    if (st != UEPM_OK) return st;
    if (band < 0 || band >= m->nb) return fail(UEPM_ERR_OUT_OF_RANGE, "uepm_bzmesh_dos: band index out of range");
    HIP_OK(hipSetDevice(m->ctx->device));
    const int slices = uepm::tetra_dos_slices(m->nt, n_points);
    HIP_OK(m->window.ensure(sizeof(int2) * std::max<int64_t>(m->nt, 1)));
    HIP_OK(m->partial.ensure(sizeof(double) * slices * n_points));
    const double* Eb = m->Eb.as<double>() + static_cast<int64_t>(band) * m->nv;
    const double* Gb = m->Gb.as<double>() + 3 * static_cast<int64_t>(band) * m->nv;
    HIP_OK(uepm::launch_tetra_dos(m->pos.as<double>(), m->dtet.as<int32_t>(), m->det.as<double>(),
                                  wedge_only ? m->wedge.as<uint8_t>() : nullptr, m->nt, Eb, Gb, n_points, e_min, e_max,
                                  reduce_bz_factor, spin_degeneracy, wedge_only ? 48.0 : 1.0, m->window.as<int2>(),
                                  m->partial.as<double>(), d_energies, d_dos, static_cast<hipStream_t>(stream)));
    return UEPM_OK;
}

int uepm_bzmesh_dos(uepm_bzmesh* m, int32_t band, int32_t n_points, double e_min, double e_max,
                    double reduce_bz_factor, double spin_degeneracy, int wedge_only, double* energies, double* dos) {
    int st = dos_checks(m, band, n_points, e_min, e_max, wedge_only, "uepm_bzmesh_dos");
    if (st != UEPM_OK) return st;
    if (!dos) return fail(UEPM_ERR_INVALID_ARG, "uepm_bzmesh_dos: null output");
    if (!m->ctx) return fail(UEPM_ERR_INVALID_ARG, "uepm_bzmesh: mesh created without a device context");
    HIP_OK(hipSetDevice(m->ctx->device));
    HIP_OK(m->out.ensure(sizeof(double) * 2 * n_points));
    double* dE = m->out.as<double>();
    double* dD = dE + n_points;
    st = uepm_bzmesh_dos_device(m, band, n_points, e_min, e_max, reduce_bz_factor, spin_degeneracy, wedge_only, dE, dD,
                                nullptr);
    if (st != UEPM_OK) return st;
    HIP_OK(hipDeviceSynchronize());
    if (energies) HIP_OK(hipMemcpy(energies, dE, sizeof(double) * n_points, hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(dos, dD, sizeof(double) * n_points, hipMemcpyDeviceToHost));
    return UEPM_OK;
}

// MeshBZ::compute_dos_band_at_band_auto (bz_mesh.cpp:889-916): the band's
// extrema widened by 0.1 eV, mesh reduction factor and spin degeneracy 2.
// use_interp selects Tetra::interpolate_dos_at_energy_per_band, which returns
// 0 in the reference (mesh_tetra.cpp:578-581): the DOS is then all zeros.
int uepm_bzmesh_dos_auto(uepm_bzmesh* m, int32_t band, int32_t n_points, int use_interp, int use_iw,
                         double* energies, double* dos) {
    if (!m || !dos) return fail(UEPM_ERR_INVALID_ARG, "uepm_bzmesh_dos_auto: bad argument");
    if (!m->host_bands || band < 0 || band >= static_cast<int32_t>(m->min_band.size()))
        return fail(UEPM_ERR_OUT_OF_RANGE, "Band index out of range in compute_dos_band_at_band_auto.");
    const double margin_energy = 0.1;
    const double e_min = m->min_band[band] - margin_energy;
    const double e_max = m->max_band[band] + margin_energy;
    if (use_interp) {
        if (n_points < 2) return fail(UEPM_ERR_INVALID_ARG, "uepm_bzmesh_dos_auto: n_points must be at least 2");
        const double step = (e_max - e_min) / (n_points - 1);
        for (int i = 0; i < n_points; ++i) {
            if (energies) energies[i] = e_min + static_cast<double>(i) * step;
            dos[i] = 0.0;
        }
        return UEPM_OK;
    }
    return uepm_bzmesh_dos(m, band, n_points, e_min, e_max, m->reduce_bz_factor, m->spin_degeneracy, use_iw, energies,
                           dos);
}

/* ---- band views in and out ----------------------------------------------- */

// MeshBZ::export_selected_bands_to_gmsh (bz_mesh.cpp:1426-1545): views
// "band_<i>" (and "grad_band_<i>") appended to out_path, valence bands first
// (from the highest down when highest_valence_as_band0), then conduction.
int uepm_bzmesh_export_msh(const uepm_bzmesh* m, const char* out_path, int64_t n_valence, int64_t n_conduction,
                           int highest_valence_as_band0, const char* model_msh_path, int write_gradients) {
    if (!m || !out_path) return fail(UEPM_ERR_INVALID_ARG, "uepm_bzmesh_export_msh: bad argument");
    if (!m->host_bands) return fail(UEPM_ERR_INVALID_ARG, "uepm_bzmesh_export_msh: no bands on the mesh");
    if (n_valence < 0 || n_conduction < 0) return fail(UEPM_ERR_INVALID_ARG, "uepm_bzmesh_export_msh: negative count");
    std::vector<int64_t> order;
    if (n_valence > 0) {
        const int64_t v_start = m->valence.start, v_last = v_start + m->valence.count - 1;
        if (highest_valence_as_band0)
            for (int64_t g = v_last; g >= v_last - n_valence + 1; --g) order.push_back(g);
        else
            for (int64_t g = v_start; g < v_start + n_valence; ++g) order.push_back(g);
    }
    for (int64_t g = 0; g < n_conduction; ++g) {
        if (g >= m->conduction.count) return fail(UEPM_ERR_OUT_OF_RANGE, "Local conduction band index out of range.");
        order.push_back(m->conduction.start + g);
    }
    const int64_t nbE = static_cast<int64_t>(m->Eband.size()), nbG = static_cast<int64_t>(m->Gband.size());
    const std::string model = model_msh_path ? model_msh_path : "";
    auto ends = [&](const char* sfx) {
        const size_t n = std::strlen(sfx);
        return model.size() >= n && model.compare(model.size() - n, n, sfx) == 0;
    };
    const bool from_file = model.size() > 4 && (ends(".msh") || ends(".msh2") || ends(".msh4"));
    int64_t out_idx = 0;
    std::vector<double> buf;
    for (int64_t g : order) {
        if (g < 0 || g >= nbE) return fail(UEPM_ERR_RUNTIME, "Vertex band vector too small for requested band index.");
        const std::string name = "band_" + std::to_string(out_idx++);
        int st = uepm_msh_append_node_view(out_path, from_file ? model.c_str() : nullptr, name.c_str(), 1,
                                           m->tags.data(), m->nv, m->Eband[g].data());
        if (st != UEPM_OK) return st;
        if (write_gradients) {
            if (g >= nbG) return fail(UEPM_ERR_RUNTIME, "Vertex gradient vector too small for requested band index.");
            const std::string gname = "grad_band_" + std::to_string(out_idx - 1);
            st = uepm_msh_append_node_view(out_path, from_file ? model.c_str() : nullptr, gname.c_str(), 3,
                                           m->tags.data(), m->nv, m->Gband[g].data());
            if (st != UEPM_OK) return st;
        }
    }
    return UEPM_OK;
}

// MeshBZ::export_energies_and_gradients_to_vtk (bz_mesh.cpp:1647-1673) through
// export_to_vtk (:1580-1645): legacy ASCII unstructured grid, 8 significant
// digits, fields in std::map (name) order.
int uepm_bzmesh_export_vtk(const uepm_bzmesh* m, const char* path) {
    if (!m || !path) return fail(UEPM_ERR_INVALID_ARG, "uepm_bzmesh_export_vtk: bad argument");
    if (!m->host_bands) return fail(UEPM_ERR_INVALID_ARG, "uepm_bzmesh_export_vtk: no bands on the mesh");
    std::ostringstream out;
    out << "# vtk DataFile Version 4.2\nBZ mesh export\nASCII\nDATASET UNSTRUCTURED_GRID\n";
    out << "POINTS " << m->nv << " double\n";
    out.precision(8);
    for (int64_t v = 0; v < m->nv; ++v) out << m->xyz[3 * v] << " " << m->xyz[3 * v + 1] << " " << m->xyz[3 * v + 2] << "\n";
    out << "CELLS " << m->nt << " " << m->nt * 5 << "\n";
    for (int64_t t = 0; t < m->nt; ++t)
        out << "4 " << m->tet[4 * t] << " " << m->tet[4 * t + 1] << " " << m->tet[4 * t + 2] << " " << m->tet[4 * t + 3]
            << "\n";
    out << "CELL_TYPES " << m->nt << "\n";
    for (int64_t t = 0; t < m->nt; ++t) out << "10\n";
    std::map<std::string, int64_t> scalars, vectors;
    for (size_t b = 0; b < m->Eband.size(); ++b) scalars["band_energy_" + std::to_string(b)] = static_cast<int64_t>(b);
    for (size_t b = 0; b < m->Gband.size(); ++b) vectors["band_grad_" + std::to_string(b)] = static_cast<int64_t>(b);
    if (!scalars.empty() || !vectors.empty()) {
        out << "POINT_DATA " << m->nv << "\n";
        for (const auto& [name, b] : scalars) {
            out << "SCALARS " << name << " double 1\nLOOKUP_TABLE default\n";
            out.precision(8);
            for (double e : m->Eband[b]) out << e << "\n";
        }
        for (const auto& [name, b] : vectors) {
            out << "VECTORS " << name << " double\n";
            out.precision(8);
            const auto& g = m->Gband[b];
            for (int64_t v = 0; v < m->nv; ++v) out << g[3 * v] << " " << g[3 * v + 1] << " " << g[3 * v + 2] << "\n";
        }
    }
    return write_text(path, out.str());
}

// MeshBZ::read_mesh_bands_from_msh_file (bz_mesh.cpp:331-425): 1-component
// views are band energies (valence when the first vertex's value is <= 0.1 eV),
// 3-component views gradients; then keep_only_bands (:638-681, which removes
// energies but keeps every gradient, as the reference does), the conduction
// shift (:727-783) and the positive-valence option (:714-725).
int uepm_bzmesh_read_bands_msh(uepm_bzmesh* m, const char* path, int32_t n_conduction, int32_t n_valence,
                               int auto_shift_conduction_band, int set_positive_valence_band) {
    if (!m || !path) return fail(UEPM_ERR_INVALID_ARG, "uepm_bzmesh_read_bands_msh: bad argument");
    uepm_msh* f = nullptr;
    int st = uepm_msh_read(path, &f);
    if (st != UEPM_OK) return st;
    std::unique_ptr<uepm_msh, void (*)(uepm_msh*)> guard(f, uepm_msh_free);
    int64_t nn = 0;
    int32_t nviews = 0;
    uepm_msh_counts(f, &nn, nullptr, &nviews);
    if (nn != m->nv)
        return fail(UEPM_ERR_INVALID_ARG, "The number of energy values does not match the number of vertices. Abort.");
    m->Eband.clear();
    m->Gband.clear();
    m->band_type.clear();
    m->band_local.clear();
    m->valence = {};
    m->conduction = {};
    std::vector<double> data;
    for (int32_t vw = 0; vw < nviews; ++vw) {
        int32_t ncomp = 0;
        uepm_msh_view_info(f, vw, nullptr, 0, &ncomp);
        data.resize(static_cast<size_t>(nn) * ncomp);
        uepm_msh_view_data(f, vw, data.data());
        if (ncomp == 1) {
            classify_append(*m, data[0] <= 0.1, static_cast<int64_t>(m->Eband.size()));
            m->Eband.push_back(data);
        } else if (ncomp == 3) {
            m->Gband.push_back(data);
        } else {
            return fail(UEPM_ERR_RUNTIME, "read_mesh_bands_from_msh_file: unsupported number of components per view: " +
                                              std::to_string(ncomp));
        }
    }
    m->host_bands = true;
    m->dev_stale = true;
    recompute_min_max(*m);
    const int64_t total = static_cast<int64_t>(m->Eband.size());
    if (n_valence >= 0 || n_conduction >= 0) {
        const int64_t keep_v = n_valence >= 0 ? n_valence : m->valence.count;
        const int64_t keep_c = n_conduction >= 0 ? n_conduction : m->conduction.count;
        if (keep_v + keep_c > total) return fail(UEPM_ERR_RUNTIME, "Cannot keep more bands than available.");
        while (m->valence.count > keep_v) {
            const int64_t idx = m->valence.start + m->valence.count - 1;
            m->Eband.erase(m->Eband.begin() + idx);
            m->band_type.erase(m->band_type.begin() + idx);
            m->band_local.erase(m->band_local.begin() + idx);
            m->valence.count--;
            m->conduction.start--;
        }
        while (m->conduction.count > keep_c) {
            const int64_t idx = m->conduction.start + m->conduction.count - 1;
            m->Eband.erase(m->Eband.begin() + idx);
            m->band_type.erase(m->band_type.begin() + idx);
            m->band_local.erase(m->band_local.begin() + idx);
            m->conduction.count--;
        }
        recompute_min_max(*m);
    }
    if (auto_shift_conduction_band && m->conduction.count > 0) {
        double min_conduction = std::numeric_limits<double>::max();
        for (int64_t i = 0; i < m->conduction.count; ++i)
            min_conduction = std::min(min_conduction, m->min_band[m->conduction.start + i]);
        const double band_gap = min_conduction - 0.0;
        if (!(band_gap < 0.1)) {
            const double shift_amount = -band_gap;
            for (int64_t i = 0; i < m->conduction.count; ++i)
                for (double& e : m->Eband[m->conduction.start + i]) e = e + shift_amount;
            recompute_min_max(*m);
        }
    }
    if (set_positive_valence_band)
        for (int64_t i = 0; i < m->valence.count; ++i)
            for (double& e : m->Eband[m->valence.start + i])
                if (e < 0.0) e = std::fabs(e);
    return UEPM_OK;
}

// MeshBZ::apply_scissor (bz_mesh.cpp:444-461): conduction bands + value.
int uepm_bzmesh_apply_scissor(uepm_bzmesh* m, double value) {
    if (!m || !m->host_bands) return fail(UEPM_ERR_INVALID_ARG, "uepm_bzmesh_apply_scissor: no bands on the mesh");
    for (size_t b = 0; b < m->Eband.size(); ++b)
        if (m->band_type[b] == 1)
            for (double& e : m->Eband[b]) e = e + value;
    m->dev_stale = true;
    recompute_min_max(*m);
    return UEPM_OK;
}

// BandStructure::AdjustValues / FindBandGap (src/EPP/BandStructure.cpp:203-262)
// over E (nk x nb, in place).
int uepm_adjust_values(double* E, int64_t nk, int32_t nb, int soc, int min_cond_to_zero, double* band_gap) {
    if ((nk > 0 && !E) || nk < 0 || nb < 0) return fail(UEPM_ERR_INVALID_ARG, "uepm_adjust_values: bad argument");
    double maxVal = 0.0, minCond = 0.0, gap = 0.0;
    bool found = false;
    if (nk > 0 && nb >= 2) {
        double fallback = std::numeric_limits<double>::lowest();
        for (int lo = 2; lo + 1 < nb; ++lo) {
            maxVal = std::numeric_limits<double>::lowest();
            minCond = std::numeric_limits<double>::infinity();
            for (int64_t k = 0; k < nk; ++k) {
                maxVal = std::max(maxVal, E[k * nb + lo]);
                minCond = std::min(minCond, E[k * nb + lo + 1]);
            }
            if (lo == 3) fallback = maxVal;
            if (maxVal + 0.35 < minCond) {
                found = true;
                break;
            }
        }
        if (!found) maxVal = fallback;
    }
    if (found) gap = minCond - maxVal;
    const int nval = soc ? 8 : 4;
    for (int64_t k = 0; k < nk; ++k)
        for (int b = 0; b < nb; ++b) {
            double& e = E[k * nb + b];
            if (b < nval)
                e -= maxVal;
            else if (min_cond_to_zero)
                e -= minCond;
            else
                e -= maxVal;
        }
    if (band_gap) *band_gap = gap;
    return UEPM_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- BZ_States

namespace {
constexpr int kStatesValence = 4;   // index_first_conduction_band (bz_states.cpp:97)
constexpr int64_t kStatesChunk = 8192;  // vertices per dielectric pass
constexpr double kQe = 1.602176634e-19;     // C (src/Common/physical_constants.hpp:20)
constexpr double kEps0 = 8.8541878128e-12;  // F/m (physical_constants.hpp:31)

// Eigen's default stream output of a dense matrix (IOFormat(): stream
// precision, columns aligned to the widest coefficient, " " between
// coefficients, "\n" between rows), coefficient (i, j) from at(i, j).
template <class T, class At>
void write_eigen_dense(std::ostream& os, int64_t rows, int64_t cols, At at) {
    size_t width = 0;
    for (int64_t j = 0; j < cols; ++j)
        for (int64_t i = 0; i < rows; ++i) {
            std::ostringstream one;
            one.copyfmt(os);
            one << static_cast<T>(at(i, j));
            width = std::max(width, one.str().size());
        }
    for (int64_t i = 0; i < rows; ++i) {
        for (int64_t j = 0; j < cols; ++j) {
            if (j) os << ' ';
            os.width(static_cast<std::streamsize>(width));
            os << static_cast<T>(at(i, j));
        }
        if (i + 1 < rows) os << '\n';
    }
}
}  // namespace

extern "C" {

// BZ_States::compute_eigenstates / compute_shifted_eigenstates
// (bz_states.cpp:24-85): k = position / (2 pi / a) per vertex, plus q_shift
// (reduced units) for the shifted set; eigenpairs from the plan's batched
// solve, kept on the device.
int uepm_bzmesh_compute_states(uepm_bzmesh* m, uepm_plan* p, const double* q_shift) {
    if (!m || !p) return fail(UEPM_ERR_INVALID_ARG, "uepm_bzmesh_compute_states: null mesh or plan");
    if (!(p->flags & UEPM_VECTORS) || (p->flags & UEPM_SOC))
        return fail(UEPM_ERR_INVALID_ARG, "uepm_bzmesh_compute_states: the plan needs UEPM_VECTORS and no UEPM_SOC");
    const int which = q_shift ? 1 : 0;
    VertexStates& st = m->states[which];
    st.set = false;
    HIP_OK(hipSetDevice(p->ctx->device));
    const double normalization_factor = 2.0 * M_PI / p->P.a;
    const int64_t nv = m->nv;
    std::vector<double> k(3 * std::max<int64_t>(nv, 1));
    for (int64_t i = 0; i < nv; ++i)
        for (int c = 0; c < 3; ++c) {
            double kc = (m->xyz[3 * i + c] * 1.0) / normalization_factor;
            if (q_shift) kc += q_shift[c];
            k[3 * i + c] = kc;
        }
    DeviceBuffer dk;
    HIP_OK(dk.ensure(sizeof(double) * k.size()));
    HIP_OK(st.E.ensure(sizeof(double) * (size_t)std::max<int64_t>(nv, 1) * p->nb));
    HIP_OK(st.X.ensure(sizeof(double2) * (size_t)std::max<int64_t>(nv, 1) * p->nb * p->dim));
    HIP_OK(hipMemcpy(dk.p, k.data(), sizeof(double) * 3 * nv, hipMemcpyHostToDevice));
    const int rc = uepm_solve_device(p, dk.as<double>(), nv, st.E.as<double>(), st.X.as<double>(), nullptr, nullptr);
    if (rc != UEPM_OK) return rc;
    HIP_OK(hipDeviceSynchronize());
    st.nb = p->nb;
    st.dim = p->dim;
    st.device = p->ctx->device;
    st.set = true;
    if (q_shift)
        for (int c = 0; c < 3; ++c) m->q_shift_si[c] = q_shift[c] * normalization_factor;
    return UEPM_OK;
}

int uepm_bzmesh_states(const uepm_bzmesh* m, int shifted, double* evals, double* evecs, int32_t* nb, int32_t* dim) {
    if (!m) return fail(UEPM_ERR_INVALID_ARG, "uepm_bzmesh_states: null mesh");
    const VertexStates& st = m->states[shifted ? 1 : 0];
    if (!st.set) return fail(UEPM_ERR_INVALID_ARG, "uepm_bzmesh_states: eigenstates not computed");
    if (nb) *nb = st.nb;
    if (dim) *dim = st.dim;
    if (evals || evecs) HIP_OK(hipSetDevice(st.device));
    if (evals && m->nv)
        HIP_OK(hipMemcpy(evals, st.E.p, sizeof(double) * (size_t)m->nv * st.nb, hipMemcpyDeviceToHost));
    if (evecs && m->nv)
        HIP_OK(hipMemcpy(evecs, st.X.p, sizeof(double2) * (size_t)m->nv * st.nb * st.dim, hipMemcpyDeviceToHost));
    return UEPM_OK;
}

// BZ_States::compute_dielectric_function (bz_states.cpp:95-164): every
// tetrahedron adds |V|/4 times the sums of its four vertices, so a vertex
// enters with the weight w_v = sum of |V_t| / 4 over its tetrahedra (taken in
// tetrahedron order); the weighted vertex sums run on the device.  Then
// eps(E) = 1 + (q_e^2 / (eps_0 |q|^2) / q_e) (2 / (2 pi)^3) S(E), q in 1/m.
int uepm_bzmesh_states_dielectric(uepm_bzmesh* m, const double* energies, int32_t n_energies, double eta_smearing,
                                  double* eps, double* sums, double* total_volume) {
    if (!m || n_energies < 0 || (n_energies > 0 && (!energies || !eps)))
        return fail(UEPM_ERR_INVALID_ARG, "uepm_bzmesh_states_dielectric: bad argument");
    const VertexStates& a = m->states[0];
    const VertexStates& b = m->states[1];
    if (!a.set || !b.set)
        return fail(UEPM_ERR_INVALID_ARG,
                    "uepm_bzmesh_states_dielectric: compute the eigenstates at k and at k + q first");
    if (a.nb != b.nb || a.dim != b.dim || a.device != b.device)
        return fail(UEPM_ERR_INVALID_ARG, "uepm_bzmesh_states_dielectric: the two eigenstate sets differ in shape");
    if (a.nb <= kStatesValence)
        return fail(UEPM_ERR_INVALID_ARG, "uepm_bzmesh_states_dielectric: needs more than 4 bands (4 valence bands)");
    const int64_t nv = m->nv;
    std::vector<double> w(std::max<int64_t>(nv, 1), 0.0);
    double tv = 0.0;
    for (int64_t t = 0; t < m->nt; ++t) {
        const int32_t* v = m->tet.data() + 4 * t;
        const double vol = std::fabs(signed_volume(m->xyz.data(), v));
        tv += vol;
        for (int j = 0; j < 4; ++j) w[v[j]] += vol * 0.25;
    }
    if (total_volume) *total_volume = tv;
    std::vector<double> S(std::max<int32_t>(n_energies, 1), 0.0);
    if (n_energies > 0 && nv > 0) {
        HIP_OK(hipSetDevice(a.device));
        const int nb = a.nb, dim = a.dim;
        const int64_t chunk = std::min<int64_t>(nv, kStatesChunk);
        const int npair = (nb - kStatesValence) * kStatesValence;
        DeviceBuffer dw, dwt, dd, part, sum, den;
        HIP_OK(dw.ensure(sizeof(double) * nv));
        HIP_OK(dwt.ensure(sizeof(double) * npair * chunk));
        HIP_OK(dd.ensure(sizeof(double) * npair * chunk));
        HIP_OK(part.ensure(sizeof(double) * (size_t)uepm::dielectric_slices(npair * chunk) * n_energies));
        HIP_OK(sum.ensure(sizeof(double) * n_energies));
        HIP_OK(den.ensure(sizeof(double) * n_energies));
        HIP_OK(hipMemcpy(dw.p, w.data(), sizeof(double) * nv, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(den.p, energies, sizeof(double) * n_energies, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(sum.p, S.data(), sizeof(double) * n_energies, hipMemcpyHostToDevice));  // zeros
        for (int64_t c0 = 0; c0 < nv; c0 += chunk) {
            const int64_t n = std::min<int64_t>(chunk, nv - c0);
            HIP_OK(uepm::launch_dielectric_accumulate(
                a.X.as<double2>() + c0 * nb * (int64_t)dim, b.X.as<double2>() + c0 * nb * (int64_t)dim,
                a.E.as<double>() + c0 * nb, b.E.as<double>() + c0 * nb, dw.as<double>() + c0, nb, dim, n,
                den.as<double>(), n_energies, eta_smearing, dwt.as<double>(), dd.as<double>(), part.as<double>(),
                sum.as<double>(), nullptr));
        }
        HIP_OK(hipDeviceSynchronize());
        HIP_OK(hipMemcpy(S.data(), sum.p, sizeof(double) * n_energies, hipMemcpyDeviceToHost));
    }
    const double* q = m->q_shift_si;
    const double len = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
    const double q_squared = len * len;
    const double coulomb_prefactor_eV = (kQe * kQe) / (kEps0 * q_squared) / kQe;
    const double prefactor = coulomb_prefactor_eV * (2.0 / std::pow(2.0 * M_PI, 3));
    for (int32_t e = 0; e < n_energies; ++e) {
        if (sums) sums[e] = S[e];
        eps[e] = 1.0 + prefactor * S[e];
    }
    return UEPM_OK;
}

// BZ_States::export_dielectric_function (bz_states.cpp:172-179):
// "<prefix>_dielectric_function.csv", default stream formatting.
int uepm_write_dielectric_csv(const char* prefix, const double* energies, const double* eps, int32_t n_energies) {
    if (!prefix || n_energies < 0 || (n_energies > 0 && (!energies || !eps)))
        return fail(UEPM_ERR_INVALID_ARG, "uepm_write_dielectric_csv: bad argument");
    std::ofstream f(std::string(prefix) + "_dielectric_function.csv");
    if (!f) return fail(UEPM_ERR_RUNTIME, std::string("cannot open ") + prefix + "_dielectric_function.csv");
    f << "Energy (eV),Dielectric function" << std::endl;
    for (int32_t e = 0; e < n_energies; ++e) f << energies[e] << "," << eps[e] << std::endl;
    return f ? UEPM_OK : fail(UEPM_ERR_RUNTIME, "uepm_write_dielectric_csv: write failed");
}

// BZ_States::export_full_eigenstates (bz_states.cpp:181-201): replaces
// <root>/eigenstates with eigenvalues/{eigenvalues,shiftedeigenvalues}_<k>.txt
// (one row) and eigenvectors/{eigenvectors,shiftedeigenvectors}_<k>.txt
// (dim x nb, "(re,im)" coefficients), in Eigen's default stream format.
int uepm_bzmesh_export_states(const uepm_bzmesh* m, const char* root) {
    if (!m) return fail(UEPM_ERR_INVALID_ARG, "uepm_bzmesh_export_states: null mesh");
    const VertexStates& a = m->states[0];
    const VertexStates& b = m->states[1];
    if (!a.set || !b.set)
        return fail(UEPM_ERR_INVALID_ARG, "uepm_bzmesh_export_states: compute the eigenstates at k and at k + q first");
    namespace fs = std::filesystem;
    const fs::path base = fs::path(root && *root ? root : ".") / "eigenstates";
    std::error_code ec;
    fs::remove_all(base, ec);
    fs::create_directory(base, ec);
    fs::create_directory(base / "eigenvectors", ec);
    fs::create_directory(base / "eigenvalues", ec);
    if (ec) return fail(UEPM_ERR_RUNTIME, "uepm_bzmesh_export_states: cannot create " + base.string());
    const int64_t nv = m->nv;
    std::vector<double> E[2], X[2];
    for (int s = 0; s < 2; ++s) {
        const VertexStates& st = m->states[s];
        E[s].resize((size_t)nv * st.nb);
        X[s].resize((size_t)nv * st.nb * st.dim * 2);
        HIP_OK(hipSetDevice(st.device));
        if (nv) {
            HIP_OK(hipMemcpy(E[s].data(), st.E.p, sizeof(double) * E[s].size(), hipMemcpyDeviceToHost));
            HIP_OK(hipMemcpy(X[s].data(), st.X.p, sizeof(double) * X[s].size(), hipMemcpyDeviceToHost));
        }
    }
    const char* vname[2] = {"eigenvalues_", "shiftedeigenvalues_"};
    const char* xname[2] = {"eigenvectors_", "shiftedeigenvectors_"};
    for (int64_t k = 0; k < nv; ++k) {
        for (int s = 0; s < 2; ++s) {
            const int nbs = m->states[s].nb, dims = m->states[s].dim;
            std::ofstream fe(base / "eigenvalues" / (std::string(vname[s]) + std::to_string(k) + ".txt"));
            const double* ev = E[s].data() + k * nbs;
            write_eigen_dense<double>(fe, 1, nbs, [&](int64_t, int64_t j) { return ev[j]; });
            fe << std::endl;
            std::ofstream fx(base / "eigenvectors" / (std::string(xname[s]) + std::to_string(k) + ".txt"));
            const double* xv = X[s].data() + k * nbs * (int64_t)dims * 2;
            write_eigen_dense<std::complex<double>>(fx, dims, nbs, [&](int64_t i, int64_t j) {
                return std::complex<double>(xv[2 * (j * dims + i)], xv[2 * (j * dims + i) + 1]);
            });
            fx << std::endl;
            if (!fe || !fx) return fail(UEPM_ERR_RUNTIME, "uepm_bzmesh_export_states: write failed");
        }
    }
    return UEPM_OK;
}

}  // extern "C"
