// sanitize_parsers.cpp — the host-side parsers (gmsh MSH 4.1 reader and
// node-data writer in csrc/msh_io.cpp, the YAML reader and material loader in
// epp/yaml_reader.cpp and epp/Material.cpp) driven over valid files and over
// truncated and corrupted copies of them.  Built by tests/test_parsers_sanitized.py
// with AddressSanitizer and UndefinedBehaviorSanitizer (g++, host code only); a
// memory or UB error aborts the run.  Every malformed input must come back as an
// error code (MSH) or a std::runtime_error (YAML), never as a crash.
//
// Usage: sanitize_parsers <mesh.msh> <work-dir> <materials.yaml>...
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "Material.h"
#include "uepm.h"
#include "yaml_reader.h"

namespace uepm_host {
std::string& last_error() {
    thread_local std::string s;
    return s;
}
}  // namespace uepm_host

namespace {

int g_failures = 0;

void check(bool ok, const std::string& what) {
    if (!ok) {
        std::printf("FAIL: %s\n", what.c_str());
        ++g_failures;
    }
}

std::string slurp(const std::string& path) {
    std::ifstream in(path, std::ios::binary);
    std::ostringstream ss;
    ss << in.rdbuf();
    return ss.str();
}

void spit(const std::string& path, const std::string& text) {
    std::ofstream out(path, std::ios::binary | std::ios::trunc);
    out << text;
}

// splitmix64: deterministic corruption positions and bytes
uint64_t mix(uint64_t& s) {
    uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

// Reads `path`; on success touches every accessor (sizes from the counts).
int read_and_walk(const std::string& path, int64_t* nodes_out = nullptr, int64_t* tets_out = nullptr,
                  int32_t* views_out = nullptr) {
    uepm_msh* m = nullptr;
    const int rc = uepm_msh_read(path.c_str(), &m);
    if (rc != UEPM_OK) {
        check(m == nullptr, "failed read leaves the handle null: " + path);
        return rc;
    }
    int64_t nn = 0, nt = 0;
    int32_t nv = 0;
    check(uepm_msh_counts(m, &nn, &nt, &nv) == UEPM_OK, "counts");
    std::vector<uint64_t> tags(nn);
    std::vector<double> xyz(3 * nn);
    std::vector<int32_t> tet(4 * nt);
    check(uepm_msh_nodes(m, tags.data(), xyz.data()) == UEPM_OK, "nodes");
    if (nt) check(uepm_msh_tetrahedra(m, tet.data()) == UEPM_OK, "tetrahedra");
    for (int32_t v = 0; v < nv; ++v) {
        char name[8];  // shorter than most names: the copy must truncate
        int32_t nc = 0;
        check(uepm_msh_view_info(m, v, name, sizeof name, &nc) == UEPM_OK && std::strlen(name) < sizeof name,
              "view info");
        std::vector<double> data(nn * (int64_t)nc);
        check(uepm_msh_view_data(m, v, data.data()) == UEPM_OK, "view data");
    }
    check(uepm_msh_view_info(m, nv, nullptr, 0, nullptr) == UEPM_ERR_OUT_OF_RANGE, "view index past the end");
    check(uepm_msh_view_data(m, -1, xyz.data()) != UEPM_OK, "negative view index");
    for (int32_t t = 0; t < 4 * nt; ++t) check(tet[t] >= 0 && tet[t] < nn, "vertex index in range");
    uepm_msh_free(m);
    if (nodes_out) *nodes_out = nn;
    if (tets_out) *tets_out = nt;
    if (views_out) *views_out = nv;
    return UEPM_OK;
}

// A small valid mesh: 5 nodes (sparse tags), two tetrahedra in two entity blocks,
// a 2-D block that is skipped, an unknown section, then two views appended by the writer.
std::string small_mesh(const std::string& dir) {
    const std::string base = dir + "/small.msh";
    spit(base,
         "$MeshFormat\n4.1 0 8\n$EndMeshFormat\n"
         "$Comments\nanything $Nodes here is skipped\n$EndComments\n"
         "$Nodes\n2 5 10 50\n3 1 0 3\n10\n20\n30\n0 0 0\n1 0 0\n0 1 0\n3 2 0 2\n40\n50\n0 0 1\n1 1 1\n$EndNodes\n"
         "$Elements\n3 3 1 3\n2 1 2 1\n1 10 20 30\n3 1 4 1\n1 10 20 30 40\n3 2 4 1\n2 20 30 40 50\n$EndElements\n");
    const std::string out = dir + "/small_views.msh";
    std::remove(out.c_str());
    const uint64_t tags[5] = {10, 20, 30, 40, 50};
    const double a[5] = {1.5, -2.0, 3.25, 1e300, -0.0};
    const double b[10] = {1, 2, 3, 4, 5, 6, 7, 8, 9, 10};
    check(uepm_msh_append_node_view(out.c_str(), base.c_str(), "band_0", 1, tags, 5, a) == UEPM_OK, "write view 1");
    check(uepm_msh_append_node_view(out.c_str(), base.c_str(), "a much longer view name", 2, tags, 5, b) == UEPM_OK,
          "append view 2");
    return out;
}

void msh_cases(const std::string& fixture, const std::string& dir) {
    // the reference's mesh, read and copied through the writer
    int64_t nn = 0, nt = 0;
    int32_t nv = 0;
    check(read_and_walk(fixture, &nn, &nt, &nv) == UEPM_OK && nn > 0 && nt > 0, "fixture mesh reads");
    {
        uepm_msh* m = nullptr;
        check(uepm_msh_read(fixture.c_str(), &m) == UEPM_OK, "fixture read for the writer");
        std::vector<uint64_t> tags(nn);
        std::vector<double> xyz(3 * nn);
        uepm_msh_nodes(m, tags.data(), xyz.data());
        uepm_msh_free(m);
        const std::string out = dir + "/fixture_views.msh";
        std::remove(out.c_str());
        check(uepm_msh_append_node_view(out.c_str(), fixture.c_str(), "xyz", 3, tags.data(), nn, xyz.data()) ==
                  UEPM_OK,
              "fixture view write");
        int32_t nv2 = 0;
        check(read_and_walk(out, nullptr, nullptr, &nv2) == UEPM_OK && nv2 == 1, "fixture view read back");
    }
    // writer argument checks
    const uint64_t t1 = 1;
    const double d1 = 0;
    check(uepm_msh_append_node_view((dir + "/x.msh").c_str(), nullptr, "v", 1, &t1, 1, &d1) == UEPM_ERR_INVALID_ARG,
          "new file without a source mesh");
    check(uepm_msh_append_node_view((dir + "/x.msh").c_str(), fixture.c_str(), "v", 10, &t1, 1, &d1) ==
              UEPM_ERR_INVALID_ARG,
          "ncomp above 9");
    check(uepm_msh_append_node_view((dir + "/x.msh").c_str(), (dir + "/absent.msh").c_str(), "v", 1, &t1, 1, &d1) ==
              UEPM_ERR_RUNTIME,
          "missing source mesh");
    check(uepm_msh_read((dir + "/absent.msh").c_str(), nullptr) == UEPM_ERR_INVALID_ARG, "null out handle");
    uepm_msh* none = nullptr;
    check(uepm_msh_read((dir + "/absent.msh").c_str(), &none) == UEPM_ERR_RUNTIME && !none, "missing file");

    // the small mesh with views: valid, then every truncation and seeded corruptions
    const std::string good_path = small_mesh(dir);
    int64_t sn = 0, st = 0;
    int32_t sv = 0;
    check(read_and_walk(good_path, &sn, &st, &sv) == UEPM_OK && sn == 5 && st == 2 && sv == 2, "small mesh reads");
    const std::string good = slurp(good_path);
    const std::string bad = dir + "/bad.msh";
    int n_ok = 0, n_err = 0;
    for (size_t cut = 0; cut < good.size(); ++cut) {
        spit(bad, good.substr(0, cut));
        (read_and_walk(bad) == UEPM_OK ? n_ok : n_err)++;
    }
    const char alphabet[] = "0123456789 -+.eE\n$\"abcNodesEnd\t\r";
    uint64_t seed = 12345;
    for (int trial = 0; trial < 4000; ++trial) {
        std::string s = good;
        const int edits = 1 + (int)(mix(seed) % 4);
        for (int e = 0; e < edits; ++e) {
            const size_t pos = mix(seed) % s.size();
            const uint64_t op = mix(seed) % 3;
            const char ch = alphabet[mix(seed) % (sizeof alphabet - 1)];
            if (op == 0)
                s[pos] = ch;
            else if (op == 1)
                s.insert(pos, 1, ch);
            else
                s.erase(pos, 1 + mix(seed) % 8);
            if (s.empty()) s = "$";
        }
        spit(bad, s);
        (read_and_walk(bad) == UEPM_OK ? n_ok : n_err)++;
    }
    // oversized counts must fail cleanly (allocation refused or data truncated)
    spit(bad, "$MeshFormat\n4.1 0 8\n$EndMeshFormat\n$Nodes\n1 99999999999999 1 99999999999999\n3 1 0 "
              "99999999999999\n1\n$EndNodes\n");
    check(read_and_walk(bad) != UEPM_OK, "huge node count");
    spit(bad, "$MeshFormat\n4.1 0 8\n$EndMeshFormat\n$Nodes\n1 1 1 1\n3 1 0 1\n1\n0 0 0\n$EndNodes\n"
              "$Elements\n1 1 1 1\n3 1 4 -5\n$EndElements\n");
    check(read_and_walk(bad) != UEPM_OK, "negative element count");
    spit(bad, "$MeshFormat\n4.1 0 8\n$EndMeshFormat\n$Nodes\n1 1 1 1\n3 1 0 1\n1\n0 0 0\n$EndNodes\n"
              "$NodeData\n1\n\"v\"\n1\n0\n3\n0\n2147483647\n1\n1 1\n$EndNodeData\n");
    check(read_and_walk(bad) != UEPM_OK, "huge component count");
    std::printf("msh: %d truncated/corrupted inputs, %d read, %d refused\n", n_ok + n_err, n_ok, n_err);
}

void yaml_cases(const std::vector<std::string>& files, const std::string& dir) {
    using uepm::pseudopotential::Materials;
    namespace yaml = uepm::pseudopotential::yaml;
    for (const auto& f : files) {
        Materials m;
        m.load_material_parameters(f);
        check(!m.materials.empty(), "materials loaded from " + f);
    }
    const std::string good = slurp(files.at(0));
    const std::string bad = dir + "/bad.yaml";
    int n_ok = 0, n_err = 0;
    auto attempt = [&](const std::string& text) {
        try {
            auto root = yaml::parse(text);
            spit(bad, text);
            Materials m;
            m.load_material_parameters(bad);
            ++n_ok;
        } catch (const std::runtime_error&) {
            ++n_err;
        }
    };
    for (size_t cut = 0; cut < good.size(); cut += 7) attempt(good.substr(0, cut));
    const char alphabet[] = " :-#\n\t\"'abcV3S0.5e-";
    uint64_t seed = 777;
    for (int trial = 0; trial < 3000; ++trial) {
        std::string s = good;
        const int edits = 1 + (int)(mix(seed) % 4);
        for (int e = 0; e < edits; ++e) {
            const size_t pos = mix(seed) % s.size();
            const uint64_t op = mix(seed) % 3;
            const char ch = alphabet[mix(seed) % (sizeof alphabet - 1)];
            if (op == 0)
                s[pos] = ch;
            else if (op == 1)
                s.insert(pos, 1, ch);
            else
                s.erase(pos, 1 + mix(seed) % 8);
            if (s.empty()) s = "-";
        }
        attempt(s);
    }
    for (const char* text : {"", "---\n", "- \n", ":\n", "a:\n  - b: 1\n   c: 2\n", "a: 'unterminated\n",
                             "a:\n\tb: 1\n", "a: 1\na: 2\n", "- - - -\n", "materials:\n  - symbol: Si\n"}) {
        attempt(text);
    }
    // spin-orbit periods out of the unsigned range are refused, not converted
    for (const char* period : {"-4", "1e30", "2.5", "nan"}) {
        std::string s = good;
        const size_t at = s.find("period_anion:");
        if (at == std::string::npos) break;
        const size_t eol = s.find('\n', at);
        s.replace(at, eol - at, std::string("period_anion: ") + period);
        const int before = n_err;
        attempt(s);
        check(n_err == before + 1, std::string("period_anion ") + period + " refused");
    }
    std::printf("yaml: %d inputs, %d loaded, %d refused\n", n_ok + n_err, n_ok, n_err);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s <mesh.msh> <work-dir> <materials.yaml>...\n", argv[0]);
        return 2;
    }
    msh_cases(argv[1], argv[2]);
    std::vector<std::string> yml(argv + 3, argv + argc);
    yaml_cases(yml, argv[2]);
    std::printf("%s (%d failures)\n", g_failures ? "FAILED" : "ALL PASSED", g_failures);
    return g_failures ? 1 : 0;
}
