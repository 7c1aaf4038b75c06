"""Pins the oracle (oracle/epm_oracle.cpp) before it is trusted as the checker.

Each test checks one function of the oracle against something independent of
its own code: the reference's known-answer tables and the reference's own
sources compiled into oracle/_ref (through the committed fixtures under
tests/golden/), LAPACK, numerical quadrature, closed forms and symmetry
invariants.  All CPU-only.
"""
from __future__ import annotations

import json
import math
import os

import numpy as np
import pytest
from scipy import integrate, special

from ultimateepm_amd._abi import GRADIENTS, NONLOCAL, SOC, VECTORS
from ultimateepm_amd.material import PARAMS_DIR, Material, load

pytestmark = pytest.mark.filterwarnings("ignore::scipy.integrate.IntegrationWarning")

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# CODATA 2018 (src/Common/physical_constants.hpp:18-41)
HBAR = 6.62607015e-34 / (2 * math.pi)
ME = 9.1093837015e-31
QE = 1.602176634e-19
BOHR = 5.29177210903e-11


def _json(name):
    with open(os.path.join(GOLDEN, name)) as fh:
        return json.load(fh)


def _doctest_approx(a, b):
    eps = 100 * np.finfo(np.float32).eps
    return abs(a - b) <= eps * (1 + max(abs(a), abs(b)))


# ---------------------------------------------------------------- Bessel ----

def test_bessel_known_answer_tables(oracle):
    """tests/epp/test_bessel_functions.cpp:22-65, the reference's own tables."""
    t = _json("bessel_known_answers.json")
    i = 0
    for nu in t["nu"]:
        for x in t["x"]:
            assert _doctest_approx(oracle.lib().orc_generalized_bessel_cylindrical(nu, x), t["cylindrical"][i]), (nu, x)
            assert _doctest_approx(oracle.lib().orc_generalized_bessel(nu, x), t["spherical"][i]), (nu, x)
            i += 1


def test_bessel_matches_reference_build(oracle):
    """generalized_bessel and i_2 from bessel_func.hpp compiled in oracle/_ref: bit-identical."""
    t = _json("ref_functions.json")
    for nu, x, v in t["generalized_bessel"]:
        assert oracle.lib().orc_generalized_bessel(nu, x) == v, (nu, x)
    for x, v in t["i2"]:
        assert oracle.lib().orc_i2(x) == v, x


@pytest.mark.parametrize("l", [0, 1, 2, 3])
def test_spherical_bessel_vs_scipy(oracle, l):
    """j_l(x) = sqrt(pi/2x) J_{l+1/2}(x): the textbook spherical Bessel function."""
    for x in (0.05, 0.3, 1.0, 2.7, 6.0, 11.0, 25.0):
        assert oracle.lib().orc_generalized_bessel(float(l), x) == pytest.approx(special.spherical_jn(l, x), rel=1e-12)
    # nu = -1: j_{-1}(x) = cos(x)/x
    for x in (0.1, 1.0, 4.0):
        assert oracle.lib().orc_generalized_bessel(-1.0, x) == pytest.approx(math.cos(x) / x, rel=1e-12)


def test_i2_is_modified_spherical_bessel(oracle):
    """bessel_func.hpp:47-52 is i_2(x), the modified spherical Bessel function of order 2."""
    for x in (0.5, 1.0, 2.0, 5.0, 12.0):
        assert oracle.lib().orc_i2(x) == pytest.approx(special.spherical_in(2, x), rel=1e-12)
    # the reference formula cancels for small x; it keeps that rounding, and stays within it
    for x in (1e-2, 5e-2, 0.2):
        assert abs(oracle.lib().orc_i2(x) - special.spherical_in(2, x)) <= 4e-16 * 3 / x ** 3 * 4
    assert oracle.lib().orc_i2(0.0) == 0.0


# ------------------------------------------------------- non-local terms ----

def _radial_quad(l, k1, k2, R):
    f = lambda x: special.spherical_jn(l, k1 * R * x) * special.spherical_jn(l, k2 * R * x) * x * x
    return R ** 3 * integrate.quad(f, 0.0, 1.0, epsabs=0, epsrel=1e-13, limit=200)[0]


@pytest.mark.parametrize("l,R", [(0, 1.06e-10), (2, 1.22e-10), (0, 1.5e-10), (2, 2.0e-10)])
def test_F_l_is_the_square_well_overlap(oracle, l, R):
    """F_l(K1, K2) = int_0^R j_l(K1 r) j_l(K2 r) r^2 dr (Material.cpp:87-109), all three branches."""
    rng = np.random.default_rng(7 + l)
    s = 2 * math.pi / 5.43e-10
    for _ in range(12):  # generic branch
        K1, K2 = rng.normal(size=3) * s * 2, rng.normal(size=3) * s * 2
        ref = _radial_quad(l, np.linalg.norm(K1), np.linalg.norm(K2), R)
        assert oracle.F_l(K1, K2, R, l) == pytest.approx(ref, rel=1e-9, abs=1e-12 * R ** 3)
    for _ in range(6):  # equal norms (rotated vectors): the closed-form limit branch
        K1 = rng.normal(size=3) * s * 2
        Q, _r = np.linalg.qr(rng.normal(size=(3, 3)))
        K2 = Q @ K1
        n = np.linalg.norm(K1)
        assert abs(np.linalg.norm(K2) - n) <= 1e-4
        assert oracle.F_l(K1, K2, R, l) == pytest.approx(_radial_quad(l, n, n, R), rel=1e-10, abs=1e-13 * R ** 3)
    # one vector zero: generic branch with j_l(0)
    K1 = np.array([0.3, -1.2, 0.8]) * s
    assert oracle.F_l(K1, np.zeros(3), R, l) == pytest.approx(_radial_quad(l, np.linalg.norm(K1), 0.0, R),
                                                              rel=1e-10, abs=1e-13 * R ** 3)
    # both zero: R^3/3 for l = 0, 0 otherwise
    assert oracle.F_l(np.zeros(3), np.zeros(3), R, l) == pytest.approx(R ** 3 / 3 if l == 0 else 0.0, rel=1e-15)


def test_F2_gauss_is_the_gaussian_well_overlap(oracle):
    """i_2(R^2 K1 K2 / 2) exp(-(K1^2+K2^2) R^2 / 4) = 4/(sqrt(pi) R^3) int exp(-r^2/R^2) j_2 j_2 r^2 dr."""
    R = 1.22e-10
    rng = np.random.default_rng(3)
    s = 2 * math.pi / 5.65e-10
    for _ in range(10):
        K1, K2 = rng.normal(size=3) * s * 1.5, rng.normal(size=3) * s * 1.5
        k1, k2 = np.linalg.norm(K1), np.linalg.norm(K2)
        f = lambda x: math.exp(-x * x) * special.spherical_jn(2, k1 * R * x) * special.spherical_jn(2, k2 * R * x) * x * x
        I = integrate.quad(f, 0.0, 12.0, epsabs=0, epsrel=1e-13, limit=400)[0]
        ref = 4.0 / math.sqrt(math.pi) * I
        assert oracle.F2_gauss(K1, K2, R) == pytest.approx(ref, rel=1e-8, abs=1e-16)


def test_nonlocal_element_structure(oracle):
    """Material.cpp:139-208: the element is Hermitian in (K1, K2) and real for a zero structure phase."""
    rng = np.random.default_rng(11)
    for sym in ("Si", "Ge", "GaAs", "InP"):
        m = load(sym)
        G = oracle.basis(6)
        for _ in range(10):
            k = rng.uniform(-1, 1, 3)
            i, j = rng.integers(0, len(G), 2)
            K1, K2 = k + G[i], k + G[j]
            z12, z21 = oracle.nl_element(m, K1, K2), oracle.nl_element(m, K2, K1)
            assert z12 == pytest.approx(z21.conjugate(), rel=1e-9, abs=1e-12)
        # the diagonal of a square-well l=0 material with K=0 reduces to 4pi/Omega * A0 * R^3/3
    m = load("Si")
    a = m.lattice_constant * 1e-10
    kF = (96 * math.pi ** 2) ** (1 / 3) / a
    A0 = m.non_local.alpha_0_anion + HBAR ** 2 / (2 * ME * QE) * m.non_local.beta_0_anion * (0 - kF ** 2)
    expect = 4 * math.pi / (a ** 3 / 8) * A0 * m.non_local.R0_anion ** 3 / 3
    assert oracle.nl_element(m, np.zeros(3), np.zeros(3)).real == pytest.approx(expect, rel=1e-13)


# ----------------------------------------------------------- spin-orbit ----

@pytest.mark.parametrize("period", [2, 3, 4])
def test_soc_radial_factor_is_slater_p_orbital_transform(oracle, period):
    """B_p(kappa) = 3 int r^(p+1) e^-r j_1(kappa r) dr / (kappa (p+2)!): the np Slater orbital (SpinOrbitFunctional.cpp:20-49)."""
    xi = 4.6
    for kappa in (0.05, 0.3, 1.0, 1.7, 2.5, 4.0):
        I = integrate.quad(lambda r: r ** (period + 1) * math.exp(-r) * special.spherical_jn(1, kappa * r),
                           0, 90, limit=400, epsabs=1e-15, epsrel=1e-12)[0]
        expect = 3 * I / (kappa * math.factorial(period + 2))
        K = kappa * xi / BOHR
        assert oracle.B_radial(period, K, xi) == pytest.approx(expect, rel=1e-7, abs=1e-12)
    assert oracle.B_radial(period, 0.0, xi) == 1.0


def test_soc_unsupported_period_raises(oracle):
    with pytest.raises(RuntimeError, match="Unsupported period"):
        oracle.B_radial(5, 1e10, 4.0)


def test_soc_block_is_pauli_form(oracle):
    """Block = (-i l_s cos(G.tau) + l_a sin(G.tau)) (K x K').sigma with Pauli matrices (SpinOrbitFunctional.cpp:100-138)."""
    sx = np.array([[0, 1], [1, 0]], dtype=complex)
    sy = np.array([[0, -1j], [1j, 0]])
    sz = np.array([[1, 0], [0, -1]], dtype=complex)
    rng = np.random.default_rng(5)
    for sym in ("Ge", "GaAs"):
        m = load(sym)
        a = m.lattice_constant * 1e-10
        so = m.spin_orbit
        G = oracle.basis(6)
        for _ in range(8):
            k = rng.uniform(-1, 1, 3)
            i, j = rng.integers(0, len(G), 2)
            K, Kp = k + G[i], k + G[j]
            Kphys, Kpphys = np.linalg.norm(K) * 2 * math.pi / a, np.linalg.norm(Kp) * 2 * math.pi / a
            l1 = so.mu * oracle.B_radial(so.period_cation, Kphys, so.radial_extent_cation) * \
                oracle.B_radial(so.period_cation, Kpphys, so.radial_extent_cation)
            l2 = so.alpha * so.mu * oracle.B_radial(so.period_anion, Kphys, so.radial_extent_anion) * \
                oracle.B_radial(so.period_anion, Kpphys, so.radial_extent_anion)
            gt = math.pi / 4 * float(np.sum(G[i] - G[j]))
            phase = -1j * (l1 + l2) / 2 * math.cos(gt) + (l1 - l2) / 2 * math.sin(gt)
            c = np.cross(K, Kp)
            expect = phase * (c[0] * sx + c[1] * sy + c[2] * sz)
            got = oracle.soc_block(m, K, Kp, G[i].astype(float), G[j].astype(float))
            assert np.allclose(got, expect, rtol=1e-12, atol=1e-14 * abs(so.mu))
            back = oracle.soc_block(m, Kp, K, G[j].astype(float), G[i].astype(float))
            assert np.allclose(back, got.conj().T, rtol=1e-12, atol=1e-16)


# ------------------------------------------------------------ assembly ----

def test_local_V_matches_reference_build(oracle):
    """Pseudopotential::GetValue compiled from the reference sources, every |dG|^2 <= 24 difference vector."""
    t = _json("ref_functions.json")["local_V"]
    V = t["VSA"]
    m = Material("test", t["a"] * 1e10, tuple(V[:4]), tuple(V[4:]))
    G = np.array(t["G"], dtype=np.int32)
    # H(k=0) off-diagonal (i, 0) with G_0 = 0 holds V(G_i - 0)
    Gb = np.vstack([np.zeros((1, 3), dtype=np.int32), G[np.any(G != 0, axis=1)]])
    H = oracle.assemble(m, Gb, np.zeros(3), nl=False, soc=False, herm=False)
    ref = {tuple(g): complex(v[0], v[1]) for g, v in zip(t["G"], t["V"])}
    for i in range(1, len(Gb)):
        assert H[i, 0] == ref[tuple(Gb[i])], tuple(Gb[i])
    assert H[0, 0] == 0.0  # V(0) = 0 and |k+G|^2 = 0


def test_kpath_matches_reference_build(oracle, uepm):
    """SymmetryPoints::GeneratePoints compiled from the reference sources: bit-identical, oracle and product host code."""
    from ultimateepm_amd._abi import UepmError

    for p in _json("ref_functions.json")["kpath"]:
        if p["count"] < 0:  # the reference throws (fewer points than path nodes)
            with pytest.raises(ValueError):
                oracle.kpath(p["labels"], p["n"])
            with pytest.raises(UepmError):
                uepm.kpath(p["labels"], p["n"])
            continue
        pts, pos = oracle.kpath(p["labels"], p["n"])
        assert len(pts) == p["count"]
        assert np.array_equal(pts, np.array(p["points"]).reshape(-1, 3))
        assert list(pos[: len(p["labels"])]) == p["sym_positions"]
        pts2, pos2 = uepm.kpath(p["labels"], p["n"])
        assert np.array_equal(pts2, pts) and list(pos2) == p["sym_positions"]


N_TABLE = {6: 59, 7: 65, 8: 89, 9: 113, 10: 137, 11: 169, 12: 181, 13: 229, 14: 259}
CUTOFF = {6: 12, 7: 16, 8: 19, 9: 20, 10: 24, 11: 27, 12: 32, 13: 35, 14: 36}


@pytest.mark.parametrize("n", sorted(N_TABLE))
def test_basis_shells(oracle, uepm, n):
    """BandStructure.cpp:26-57: every fcc reciprocal vector with |G|^2 <= cutoff, in (i, j, k) loop order."""
    G = oracle.basis(n)
    assert len(G) == N_TABLE[n]
    b = np.array([[-1, 1, 1], [1, -1, 1], [1, 1, -1]])
    ijk = [(i, j, k) for i in range(-7, 8) for j in range(-7, 8) for k in range(-7, 8)]
    brute = [tuple(i * b[0] + j * b[1] + k * b[2]) for i, j, k in ijk
             if np.sum((i * b[0] + j * b[1] + k * b[2]) ** 2) <= CUTOFF[n]]
    assert [tuple(g) for g in G] == brute
    assert np.array_equal(uepm.basis_vectors(n), G)


def _zero_material(a=5.43):
    return Material("free", a, (0.0,) * 4, (0.0,) * 4)


def test_free_electron_limit(oracle):
    """All V = 0: E = hbar^2/(2 m q) (2 pi/a)^2 |k+G|^2, sorted (Hamiltonian.cpp:61-74)."""
    m = _zero_material()
    G = oracle.basis(8)
    rng = np.random.default_rng(2)
    ks = rng.uniform(-1, 1, (5, 3))
    ev, vec, gr = oracle.solve(m, G, GRADIENTS | VECTORS, 20, ks)
    kin = HBAR ** 2 / (2 * ME * QE) * (2 * math.pi / 5.43e-10) ** 2
    for q, k in enumerate(ks):
        e = np.sort(kin * np.sum((k + G) ** 2, axis=1))[:20]
        assert np.allclose(ev[q], e, rtol=1e-13, atol=1e-12)
        # gradient of a free-electron level: 2 kin (k+G) a/(2 pi)
        order = np.argsort(kin * np.sum((k + G) ** 2, axis=1))[:20]
        expect = 2 * kin * (k + G[order]) * 5.43e-10 / (2 * math.pi)
        assert np.allclose(gr[q], expect, rtol=1e-10, atol=1e-22)


def test_assembly_hermitian_before_symmetrisation(oracle):
    """SetMatrix leaves H Hermitian up to rounding of the non-local F_l (Hamiltonian.cpp:50-123)."""
    for sym, soc in (("Si", False), ("Ge", True), ("GaAs", True), ("InP", False)):
        m = load(sym)
        G = oracle.basis(6)
        k = np.array([0.13, -0.41, 0.27])
        H = oracle.assemble(m, G, k, nl=True, soc=soc, herm=False)
        assert np.abs(H - H.conj().T).max() <= 1e-9 * np.abs(H).max()


# -------------------------------------------------------- eigensolver ----

def test_eigenvalues_vs_lapack_goldens(oracle):
    """LAPACK zheevd on the same H(k) stands in for Eigen's SelfAdjointEigenSolver (1e-10 eV)."""
    z = np.load(os.path.join(GOLDEN, "eig_lapack.npz"))
    ncase = len([f for f in z.files if f.endswith("_meta")])
    assert ncase >= 6
    for c in range(ncase):
        sym, f, n, nl, soc, nb = z[f"case{c}_meta"]
        m = load(str(sym), os.path.join(PARAMS_DIR, str(f)))
        flags = (NONLOCAL if nl == "1" else 0) | (SOC if soc == "1" else 0)
        ev, _, _ = oracle.solve(m, oracle.basis(int(n)), flags, int(nb), z[f"case{c}_k"])
        assert np.abs(ev - z[f"case{c}_evals"]).max() <= 1e-10, (sym, n)


@pytest.mark.parametrize("seed", [1, 2])
def test_eigensolver_random_hermitian(oracle, seed):
    """Random dense Hermitian matrices (N = 9, 15, 64, 200): eigenvalues vs LAPACK, residual, orthonormality."""
    rng = np.random.default_rng(seed)
    for n in (1, 2, 9, 15, 64, 200):
        A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
        H = (A + A.conj().T) * 3
        w, V = oracle.eigh(H, vectors=True)
        assert np.allclose(w, np.linalg.eigvalsh(H), atol=1e-11 * max(1.0, np.abs(w).max()))
        assert np.abs(H @ V - V * w).max() <= 1e-11 * max(1.0, np.abs(w).max())
        assert np.abs(V.conj().T @ V - np.eye(n)).max() <= 1e-12
        # degenerate spectrum: repeated eigenvalues keep an orthonormal basis
    Q, _ = np.linalg.qr(rng.normal(size=(30, 30)) + 1j * rng.normal(size=(30, 30)))
    d = np.repeat([-2.0, 0.5, 0.5000000001, 3.0, 7.0, 9.0], 5)
    H = (Q * d) @ Q.conj().T
    w, V = oracle.eigh(H, vectors=True)
    assert np.allclose(w, d, atol=1e-12)
    assert np.abs(V.conj().T @ V - np.eye(30)).max() <= 1e-12


# ---------------------------------------------------------- invariants ----

def _oh_ops():
    import itertools

    ops = []
    for perm in itertools.permutations(range(3)):
        for signs in itertools.product((1, -1), repeat=3):
            M = np.zeros((3, 3))
            for r, c in enumerate(perm):
                M[r, c] = signs[r]
            ops.append(M)
    return ops


@pytest.mark.parametrize("sym,soc", [("Si", False), ("GaAs", False), ("Ge", True)])
def test_cubic_symmetry_of_bands(oracle, sym, soc):
    """E(S k) = E(k) for the 48 operations of O_h (point group plus time reversal for zincblende)."""
    m = load(sym)
    G = oracle.basis(8 if not soc else 6)
    k = np.array([0.31, 0.17, -0.07])
    flags = NONLOCAL | (SOC if soc else 0)
    ks = np.array([S @ k for S in _oh_ops()])
    ev, _, _ = oracle.solve(m, G, flags, 12, ks, nthreads=8)
    assert np.abs(ev - ev[0]).max() <= 1e-9


def test_kramers_pairs_only_with_inversion(oracle):
    """Si/Ge (diamond): every SOC level is a Kramers pair; GaAs (zincblende) is spin-split off the symmetry lines."""
    k = np.array([0.23, 0.41, 0.05])
    for sym in ("Ge", "GaAs"):
        m = load(sym)
        ev, _, _ = oracle.solve(m, oracle.basis(8), NONLOCAL | SOC, 16, k[None])
        split = np.abs(ev[0, 0::2] - ev[0, 1::2]).max()
        if sym == "Ge":
            assert split <= 1e-9
        else:
            assert split >= 1e-4


def test_zero_soc_doubles_the_spectrum(oracle):
    """Si has mu_soc = 0 in materials.yaml: the spinor spectrum is the scalar spectrum twice over."""
    m = load("Si")
    assert m.spin_orbit.mu == 0.0
    G = oracle.basis(8)
    ks = np.array([[0.1, 0.2, 0.3], [0.5, 0.5, 0.5]])
    e1, _, _ = oracle.solve(m, G, NONLOCAL, 8, ks)
    e2, _, _ = oracle.solve(m, G, NONLOCAL | SOC, 16, ks)
    assert np.abs(e2[:, 0::2] - e1).max() <= 1e-10 and np.abs(e2[:, 1::2] - e1).max() <= 1e-10


def test_spin_orbit_splitting_order_at_gamma(oracle):
    """Ge at Gamma: the j=3/2 quartet is the valence top and the split-off doublet lies below it."""
    m = load("Ge")
    ev, _, _ = oracle.solve(m, oracle.basis(8), NONLOCAL | SOC, 10, np.zeros((1, 3)))
    top = ev[0, 4:8]
    assert np.ptp(top) <= 1e-9
    assert np.abs(ev[0, 2] - ev[0, 3]) <= 1e-9
    assert ev[0, 3] < top[0] - 1e-3


def test_gamma25_triple_degeneracy(oracle):
    """Si at Gamma without SOC: the valence-band top is threefold (Gamma_25')."""
    ev, _, _ = oracle.solve(load("Si"), oracle.basis(10), NONLOCAL, 6, np.zeros((1, 3)))
    assert np.ptp(ev[0, 1:4]) <= 1e-9 and ev[0, 4] - ev[0, 3] > 0.5


def test_gradient_vs_finite_difference(oracle):
    """Hellmann-Feynman gradient (Hamiltonian.cpp:146-194) vs central differences of E(k), local potential."""
    m = load("Si", os.path.join(PARAMS_DIR, "materials-local-cohen.yaml"))
    G = oracle.basis(7)
    k0 = np.array([0.21, -0.13, 0.37])
    ev, _, gr = oracle.solve(m, G, GRADIENTS | VECTORS, 8, k0[None])
    h = 1e-5
    a = m.lattice_constant * 1e-10
    for d in range(3):
        kp, km = k0.copy(), k0.copy()
        kp[d] += h
        km[d] -= h
        ep, _, _ = oracle.solve(m, G, 0, 8, kp[None])
        em, _, _ = oracle.solve(m, G, 0, 8, km[None])
        fd = (ep[0] - em[0]) / (2 * h) * a / (2 * math.pi)
        assert np.allclose(gr[0, :, d], fd, rtol=1e-5, atol=1e-6 * a)


def test_adjust_values_shifts_to_valence_maximum(oracle):
    """AdjustValues (BandStructure.cpp:203-262): the highest valence level over the path becomes 0."""
    m = load("Si", os.path.join(PARAMS_DIR, "materials-local-cohen.yaml"))
    ks, _ = oracle.kpath("LGXWKULWXKG", 60)
    ev, _, _ = oracle.solve(m, oracle.basis(8), 0, 8, ks)
    vbm = ev[:, 3].max()
    shift, E = oracle.adjust_values(ev.copy(), soc=False)
    assert np.allclose(E, ev - vbm, atol=1e-13)
    assert abs(E[:, 3].max()) <= 1e-13


def test_bz_sampler(oracle, uepm):
    """The seeded first-BZ sampler: product host code == numpy restatement; every point in the first BZ."""
    k1 = uepm.bz_sample(20261202, 4000)
    k2 = oracle.splitmix64_bz_sample(20261202, 4000)
    assert np.array_equal(k1, k2)
    a = np.abs(k1)
    assert (a <= 1).all() and (a.sum(1) <= 1.5).all()
    for k in k1[:50]:
        assert oracle.lib().orc_in_first_bz(np.ascontiguousarray(k).ctypes.data_as(
            __import__("ctypes").POINTER(__import__("ctypes").c_double))) == 1
    assert np.abs(k1.mean(0)).max() < 0.03
    assert not np.array_equal(uepm.bz_sample(1, 100), uepm.bz_sample(2, 100))
