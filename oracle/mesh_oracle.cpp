// Oracle for the Brillouin-zone mesh consumers of the band solve.  TEST
// INFRASTRUCTURE ONLY: tests, smoke() and bench.py's CPU baseline call it as
// the checker; the product path never does.
//
// Restated from the reference, in its operation order (no contraction):
//   orc_tetra_dos      MeshBZ::compute_dos_at_energy_and_band (src/BZ_MESH/bz_mesh.cpp:844-862)
//                      over Tetra::compute_tetra_dos_energy_band (src/BZ_MESH/mesh_tetra.cpp:477-525),
//                      compute_band_iso_energy_surface (:339-395), order_cyclic (:402-450),
//                      polygon_area (:462-474), compute_barycentric_coordinates (:173-191),
//                      interpolate_gradient_energy_at_band (:214-230), sorted slots (:287-315),
//                      signed volume (:128-132); energy grid of compute_dos_band_at_band_auto
//                      (bz_mesh.cpp:889-916).
//   orc_same_symmetry  apply_same_symmetry_operation (bz_mesh.cpp:1314-1381).
//
// Pinned (tests/test_mesh_host.py): the DOS of linear bands on cube meshes
// against closed-form cross-section areas, the integral of the DOS against
// the mesh volume, tetrahedra cut in all three iso-surface cases, the wedge
// factor; the symmetry map against the 48 signed permutations.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <vector>

namespace {

struct vec3 {
    double x = 0, y = 0, z = 0;
};
vec3 operator+(vec3 a, vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
vec3 operator-(vec3 a, vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
vec3 operator*(double s, vec3 a) { return {s * a.x, s * a.y, s * a.z}; }
vec3 operator*(vec3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
vec3 operator/(vec3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
double dot(vec3 a, vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
vec3 cross(vec3 a, vec3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
double norm(vec3 a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
double triple(vec3 a, vec3 b, vec3 c) { return dot(a, cross(b, c)); }

struct Tetra {
    std::array<vec3, 4> p;
    std::array<vec3, 4> g;
    std::array<double, 4> e;
    std::array<vec3, 3> edge;
    double signed_volume;
};

std::array<int, 4> sorted_slots(std::array<double, 4> v) {
    std::array<int, 4> s = {0, 1, 2, 3};
    auto cs = [&](int a, int b) {
        if (v[a] > v[b]) {
            std::swap(v[a], v[b]);
            std::swap(s[a], s[b]);
        }
    };
    cs(0, 1);
    cs(2, 3);
    cs(0, 2);
    cs(1, 3);
    cs(1, 2);
    return s;
}

vec3 euclid(const Tetra& T, std::array<double, 4> b, const std::array<int, 4>& idx) {
    return b[0] * T.p[idx[0]] + b[1] * T.p[idx[1]] + b[2] * T.p[idx[2]] + b[3] * T.p[idx[3]];
}

std::vector<vec3> iso_surface(const Tetra& T, double E) {
    const auto s = sorted_slots(T.e);
    const double e0 = T.e[s[0]], e1 = T.e[s[1]], e2 = T.e[s[2]], e3 = T.e[s[3]];
    if (E < e1 && E >= e0) {
        const double lU = (E - e0) / (e1 - e0), lV = (E - e0) / (e2 - e0), lW = (E - e0) / (e3 - e0);
        return {euclid(T, {1.0 - lU, lU, 0.0, 0.0}, s), euclid(T, {1.0 - lV, 0.0, lV, 0.0}, s),
                euclid(T, {1.0 - lW, 0.0, 0.0, lW}, s)};
    }
    if (E < e2 && E >= e1) {
        const double lU = (E - e0) / (e2 - e0), lV = (E - e0) / (e3 - e0);
        const double lW = (e2 - E) / (e2 - e1), lX = (E - e1) / (e3 - e1);
        return {euclid(T, {1.0 - lU, 0.0, lU, 0.0}, s), euclid(T, {1.0 - lV, 0.0, 0.0, lV}, s),
                euclid(T, {0.0, lW, 1.0 - lW, 0.0}, s), euclid(T, {0.0, 1.0 - lX, 0.0, lX}, s)};
    }
    const double lU = (e3 - E) / (e3 - e2), lV = (e3 - E) / (e3 - e1), lW = (e3 - E) / (e3 - e0);
    return {euclid(T, {0.0, 0.0, lU, 1.0 - lU}, s), euclid(T, {0.0, lV, 0.0, 1.0 - lV}, s),
            euclid(T, {lW, 0.0, 0.0, 1.0 - lW}, s)};
}

std::vector<vec3> order_cyclic(const std::vector<vec3>& pts) {
    if (pts.size() < 3) return pts;
    vec3 c = std::accumulate(pts.begin(), pts.end(), vec3{0, 0, 0});
    c = c / static_cast<double>(pts.size());
    vec3 n = cross(pts[1] - pts[0], pts[2] - pts[0]);
    const double len = norm(n);
    if (len > 0.0) n = n / len;
    vec3 u = pts[0] - c;
    const double ulen = norm(u);
    if (ulen > 0.0)
        u = u / ulen;
    else
        u = vec3{1, 0, 0};
    const vec3 v = cross(n, u);
    std::vector<std::pair<double, vec3>> wa;
    for (const auto& p : pts) {
        const vec3 d = p - c;
        wa.push_back({std::atan2(dot(d, v), dot(d, u)), p});
    }
    std::stable_sort(wa.begin(), wa.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    std::vector<vec3> out;
    for (auto& w : wa) out.push_back(w.second);
    return out;
}

double polygon_area(const std::vector<vec3>& pts) {
    const auto o = order_cyclic(pts);
    if (o.size() < 3) return 0.0;
    vec3 sum{0, 0, 0};
    for (size_t i = 0; i < o.size(); ++i) sum = sum + cross(o[i], o[(i + 1) % o.size()]);
    return 0.5 * norm(sum);
}

vec3 interp_gradient(const Tetra& T, vec3 k) {
    const vec3 v1 = k - T.p[0];
    const double det = 6.0 * T.signed_volume;
    const double l2 = triple(v1, T.edge[1], T.edge[2]) / det;
    const double l3 = triple(v1, T.edge[2], T.edge[0]) / det;
    const double l4 = triple(v1, T.edge[0], T.edge[1]) / det;
    const double l1 = 1.0 - l2 - l3 - l4;
    const double b[4] = {l1, l2, l3, l4};
    vec3 g{0, 0, 0};
    for (int i = 0; i < 4; ++i) g = g + T.g[i] * b[i];
    return g;
}

double tetra_dos(const Tetra& T, double E) {
    const double mn = *std::min_element(T.e.begin(), T.e.end());
    const double mx = *std::max_element(T.e.begin(), T.e.end());
    if (E < mn || E > mx) return 0.0;
    const auto iso = iso_surface(T, E);
    if (iso.size() < 3) return 0.0;
    const double A = polygon_area(iso);
    if (!(A > 0.0)) return 0.0;
    double inv = 0.0;
    for (const auto& k : iso) inv += 1.0 / std::max(norm(interp_gradient(T, k)), 1e-12);
    const double avg = inv / static_cast<double>(iso.size());
    constexpr double pref = 1.0 / (8.0 * M_PI * M_PI * M_PI);
    return pref * (A * avg);
}

}  // namespace

extern "C" {

// E: nv x nb, G: nv x nb x 3 (vertex-major).  Serial sum over tetrahedra in
// list order per energy, as the reference; energies run in parallel.
int orc_tetra_dos(const double* pos, int64_t nv, const int32_t* tet, int64_t nt, const uint8_t* in_wedge,
                  const double* E, const double* G, int nb, int band, int n_points, double e_min, double e_max,
                  double reduce_bz_factor, double spin_degeneracy, int wedge_only, double* energies, double* dos,
                  int nthreads) {
    if (n_points < 2 || band < 0 || band >= nb) return 1;
    std::vector<Tetra> T(static_cast<size_t>(nt));
    for (int64_t t = 0; t < nt; ++t) {
        for (int j = 0; j < 4; ++j) {
            const int64_t v = tet[4 * t + j];
            if (v < 0 || v >= nv) return 1;
            T[t].p[j] = {pos[3 * v], pos[3 * v + 1], pos[3 * v + 2]};
            T[t].e[j] = E[v * nb + band];
            const double* g = G + 3 * (v * nb + band);
            T[t].g[j] = {g[0], g[1], g[2]};
        }
        T[t].edge = {T[t].p[1] - T[t].p[0], T[t].p[2] - T[t].p[0], T[t].p[3] - T[t].p[0]};
        T[t].signed_volume = (1.0 / 6.0) * triple(T[t].edge[0], T[t].edge[1], T[t].edge[2]);
    }
    const double step = (e_max - e_min) / (n_points - 1);
#pragma omp parallel for schedule(dynamic) num_threads(nthreads > 0 ? nthreads : 1)
    for (int i = 0; i < n_points; ++i) {
        const double energy = e_min + static_cast<double>(i) * step;
        double total = 0.0;
        for (int64_t t = 0; t < nt; ++t) {
            if (wedge_only && !in_wedge[t]) continue;
            total += tetra_dos(T[t], energy);
        }
        total *= reduce_bz_factor;
        total *= spin_degeneracy;
        if (wedge_only) total *= 48.0;
        if (energies) energies[i] = energy;
        dos[i] = total;
    }
    return 0;
}

// Signed permutation taking k_iw to k_bz, applied to v; v unchanged when
// there is none.
void orc_same_symmetry(const double* k_iw, const double* k_bz, const double* v, double* out) {
    constexpr double eps = 1e-8;
    const double a[3] = {k_iw[0], k_iw[1], k_iw[2]}, b[3] = {k_bz[0], k_bz[1], k_bz[2]};
    int perm[3] = {-1, -1, -1}, sign[3] = {1, 1, 1};
    bool used[3] = {false, false, false};
    double scale = 0.0;
    for (double x : a) scale = std::max(scale, std::abs(x));
    for (double x : b) scale = std::max(scale, std::abs(x));
    if (scale < 1.0) scale = 1.0;
    const double tol = eps * scale;
    for (int jb = 0; jb < 3; ++jb) {
        bool found = false;
        for (int ia = 0; ia < 3; ++ia) {
            if (used[ia]) continue;
            if (std::abs(std::abs(a[ia]) - std::abs(b[jb])) <= tol) {
                perm[jb] = ia;
                used[ia] = true;
                if (std::abs(a[ia]) > tol)
                    sign[jb] = ((b[jb] >= 0.0) ? 1 : -1) * ((a[ia] >= 0.0) ? 1 : -1);
                else
                    sign[jb] = (b[jb] >= 0.0) ? 1 : -1;
                found = true;
                break;
            }
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

}  // extern "C"
