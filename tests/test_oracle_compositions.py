"""Pins the composition of the oracle's non-local and spin-orbit terms from
their definitions, and shows that the checks catch plausible mistakes.

The radial pieces (F_l, the gaussian F_2, the spin-orbit radial factors) are
pinned one by one in test_oracle_pins.py.  This file checks what the oracle
builds from them (oracle/epm_oracle.cpp `nl_element`, `soc_block`,
`assemble`) against an evaluation written from the physics, not from the
oracle's expressions:

* non-local (src/EPP/Material.cpp:139-208): for each species s,
      V_s(K1, K2) = (4 pi / Omega) sum_l (2l + 1) A_l P_l(cos t) int w_l(r) j_l(K1 r) j_l(K2 r) r^2 dr
  with Omega = a^3 / 8, w_0 the square well of radius R0, w_2 the square well
  of radius R2 or the gaussian exp(-r^2 / R2^2), A_0 = alpha_0 +
  (hbar^2 / 2 m_e q_e) beta_0 (|K1||K2| - k_F^2), k_F^3 = 3 pi^2 (32 / a^3)
  (four valence electrons per atom, eight atoms per cubic cell) and A_2
  constant; the integrals by adaptive quadrature, P_l from scipy.  The
  element is the two-site structure sum 1/2 [V_anion e^{+i th} +
  V_cation e^{-i th}], th = 2 pi tau.(G_i - G_j) / a, tau = a/8 (1, 1, 1):
  the site signs are the reference's (Material.cpp:198-207).
* spin-orbit (src/EPP/SpinOrbitFunctional.cpp:50-138): the per-site sum
      1/2 sum_s (-i) lambda_s(K, K') ((K x K') . sigma) e^{i s_s th}
  with lambda_cation = mu B_c(|K|) B_c(|K'|) on the +th site and
  lambda_anion = alpha mu B_a(|K|) B_a(|K'|) on the -th site (the
  reference's lambda_1 / lambda_2 and the signs of its phase, :100-138),
  B_p the j_1 transform of the Slater np orbital by quadrature, K and K'
  dimensionless in the cross product (:129-130), sigma the Pauli matrices.
* assembly (src/EPP/Hamiltonian.cpp:50-123): H = kinetic + local + non-local
  on the diagonal spin blocks and the spin-orbit 2x2 blocks at
  [[up-up, up-down], [down-up, down-down]].

The reference puts the non-local anion well on the +th site and the
spin-orbit cation term on the +th site; both are kept as written there.

Mutation checks: each perturbation below is applied to the text of
oracle/epm_oracle.cpp, compiled into a temporary library (CPU only) and must
fail the definition check, so the check is sharp enough to catch it.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import math
import os
import shutil
import subprocess
import tempfile
from concurrent.futures import ThreadPoolExecutor
from functools import lru_cache

import numpy as np
import pytest
from scipy import integrate, special

from ultimateepm_amd.material import load

pytestmark = pytest.mark.filterwarnings("ignore::scipy.integrate.IntegrationWarning")

ORACLE_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")

# CODATA 2018 (src/Common/physical_constants.hpp:18-41)
HBAR = 6.62607015e-34 / (2 * math.pi)
ME = 9.1093837015e-31
QE = 1.602176634e-19
BOHR = 5.29177210903e-11
KIN = HBAR ** 2 / (2 * ME * QE)  # eV m^2

PAULI = (np.array([[0, 1], [1, 0]], dtype=complex), np.array([[0, -1j], [1j, 0]]),
         np.array([[1, 0], [0, -1]], dtype=complex))


# ------------------------------------------------------------ definitions ----

def _overlap(l, k1, k2, R, well):
    """int_0^inf w(r) j_l(k1 r) j_l(k2 r) r^2 dr for a square (r < R) or gaussian exp(-r^2/R^2) well."""
    if well == "square":
        f = lambda x: special.spherical_jn(l, k1 * R * x) * special.spherical_jn(l, k2 * R * x) * x * x
        return R ** 3 * integrate.quad(f, 0.0, 1.0, epsabs=0, epsrel=1e-13, limit=200)[0]
    f = lambda x: math.exp(-x * x) * special.spherical_jn(l, k1 * R * x) * special.spherical_jn(l, k2 * R * x) * x * x
    return R ** 3 * integrate.quad(f, 0.0, 9.0, epsabs=0, epsrel=1e-13, limit=400)[0]


def _species_potential(mat, species, K1, K2):
    """V_s(K1, K2) in eV for physical K (1/m), from the partial-wave definition."""
    nl = mat.non_local
    a = mat.lattice_constant * 1e-10
    omega = a ** 3 / 8
    k_F = (3 * math.pi ** 2 * 32 / a ** 3) ** (1 / 3)
    k1, k2 = float(np.linalg.norm(K1)), float(np.linalg.norm(K2))
    cos_t = float(np.dot(K1, K2) / (k1 * k2)) if k1 * k2 > 0 else 1.0
    alpha = getattr(nl, f"alpha_0_{species}")
    beta = getattr(nl, f"beta_0_{species}")
    A2 = getattr(nl, f"A2_{species}")
    R0 = getattr(nl, f"R0_{species}")
    R2 = getattr(nl, f"R2_{species}")
    V = 0.0
    if R0 != 0.0:  # l = 0: always the square well of radius R0
        A0 = alpha + KIN * beta * (k1 * k2 - k_F ** 2)
        V += 1 * A0 * special.eval_legendre(0, cos_t) * _overlap(0, k1, k2, R0, "square")
    if A2 != 0.0:
        V += 5 * A2 * special.eval_legendre(2, cos_t) * _overlap(2, k1, k2, R2, nl.well_type)
    return 4 * math.pi / omega * V


def nonlocal_definition(mat, K1n, K2n):
    """Non-local element for reduced K1 = k + G_i, K2 = k + G_j."""
    a = mat.lattice_constant * 1e-10
    K1, K2 = np.asarray(K1n) * 2 * math.pi / a, np.asarray(K2n) * 2 * math.pi / a
    th = 2 * math.pi * float(np.dot(np.full(3, 1 / 8), np.asarray(K1n) - np.asarray(K2n)))
    Va = _species_potential(mat, "anion", K1, K2)
    Vc = _species_potential(mat, "cation", K1, K2)
    return 0.5 * (Va * np.exp(1j * th) + Vc * np.exp(-1j * th))


@lru_cache(maxsize=None)
def _slater_radial(period, K, xi):
    """B_p at |K| (1/m): 3 int r^(p+1) e^-r j_1(kappa r) dr / (kappa (p+2)!), kappa = |K| a_0 / xi."""
    kappa = K * BOHR / xi
    if kappa == 0.0:
        return 1.0
    I = integrate.quad(lambda r: r ** (period + 1) * math.exp(-r) * special.spherical_jn(1, kappa * r), 0, 120,
                       limit=500, epsabs=0, epsrel=1e-12)[0]
    return 3 * I / (kappa * math.factorial(period + 2))


def soc_definition(mat, Kn, Kpn, G, Gp):
    """2x2 spin-orbit block for reduced K = k + G, K' = k + G'."""
    so = mat.spin_orbit
    a = mat.lattice_constant * 1e-10
    s = 2 * math.pi / a
    k, kp = float(np.linalg.norm(Kn)) * s, float(np.linalg.norm(Kpn)) * s
    lam_c = so.mu * _slater_radial(so.period_cation, k, so.radial_extent_cation) * \
        _slater_radial(so.period_cation, kp, so.radial_extent_cation)
    lam_a = so.alpha * so.mu * _slater_radial(so.period_anion, k, so.radial_extent_anion) * \
        _slater_radial(so.period_anion, kp, so.radial_extent_anion)
    th = 2 * math.pi * float(np.dot(np.full(3, 1 / 8), np.asarray(G, float) - np.asarray(Gp, float)))
    c = np.cross(Kn, Kpn)
    cs = c[0] * PAULI[0] + c[1] * PAULI[1] + c[2] * PAULI[2]
    return 0.5 * (-1j) * cs * (lam_c * np.exp(1j * th) + lam_a * np.exp(-1j * th))


def hamiltonian_definition(mat, G, k, nl, soc):
    """SetMatrix's H(k) before hermitisation, element by element from the definitions."""
    a = mat.lattice_constant * 1e-10
    N = len(G)
    KG = k[None, :] + G
    H = np.zeros((N, N), dtype=complex)
    th_unit = 2 * math.pi / 8
    for i in range(N):
        for j in range(N):
            d = G[i] - G[j]
            g2 = int(d @ d)
            idx = {3: 0, 4: 1, 8: 2, 11: 3}.get(g2)
            if idx is not None:
                th = th_unit * int(d.sum())
                H[i, j] = mat.V_S[idx] * math.cos(th) + 1j * mat.V_A[idx] * math.sin(th)
            if nl:
                H[i, j] += nonlocal_definition(mat, KG[i], KG[j])
        H[i, i] += KIN * (2 * math.pi / a) ** 2 * float(KG[i] @ KG[i])
    if not soc:
        return H
    out = np.zeros((2 * N, 2 * N), dtype=complex)
    out[:N, :N] = H
    out[N:, N:] = H
    for i in range(N):
        for j in range(N):
            b = soc_definition(mat, KG[i], KG[j], G[i], G[j])
            out[i, j] += b[0, 0]
            out[i, N + j] += b[0, 1]
            out[N + i, j] += b[1, 0]
            out[N + i, N + j] += b[1, 1]
    return out


# ------------------------------------------------------------- samples ----

def _pairs(rng, G, count):
    """Random (K1, K2) = (k + G_i, k + G_j), plus equal-norm, diagonal and zero-vector cases at k = 0."""
    out = []
    for _ in range(count):
        k = rng.uniform(-1, 1, 3)
        i, j = rng.integers(0, len(G), 2)
        out.append((k + G[i], k + G[j], G[i], G[j]))
    norms = np.einsum("ij,ij->i", G, G)
    zero = np.zeros(3)
    for shell in (3, 4, 8, 11):  # equal norms (the second branch of F_l), non-trivial angles
        idx = np.flatnonzero(norms == shell)
        i, j = idx[0], idx[len(idx) // 2]
        out.append((G[i].astype(float), G[j].astype(float), G[i], G[j]))
    g = G[np.flatnonzero(norms == 3)[0]]
    out.append((g.astype(float), g.astype(float), g, g))  # diagonal: cos t = 1
    out.append((zero, g.astype(float), zero.astype(int), g))  # one vector zero
    out.append((zero, zero, zero.astype(int), zero.astype(int)))  # both zero: R^3/3 limit
    return out


NL_MATERIALS = ("Si", "Ge", "GaAs", "InP", "GaP", "ZnSe")


def _nl_worst(nl_fn, materials=NL_MATERIALS, count=10, seed=2024):
    """Largest |oracle - definition| / (1e-9 |definition| + 1e-12 eV) over the samples."""
    from oracle import oracle as O

    G = O.basis(6)
    worst = 0.0
    for sym in materials:
        mat = load(sym)
        rng = np.random.default_rng(seed + len(sym))
        for K1, K2, _g1, _g2 in _pairs(rng, G, count):
            ref = nonlocal_definition(mat, K1, K2)
            got = nl_fn(mat, K1, K2)
            worst = max(worst, abs(got - ref) / (1e-9 * abs(ref) + 1e-12))
    return worst


def _soc_material():
    """GaP with a spin-orbit strength: cation period 3 / anion period 4, different extents and alpha != 1."""
    m = load("GaP")
    return dataclasses.replace(m, spin_orbit=dataclasses.replace(m.spin_orbit, mu=0.0009))


def _soc_worst(soc_fn, count=8, seed=99):
    from oracle import oracle as O

    G = O.basis(6)
    worst = 0.0
    for mat in (load("Ge"), load("GaAs"), load("ZnSe"), _soc_material()):
        rng = np.random.default_rng(seed + int(mat.lattice_constant * 100))
        scale = abs(mat.spin_orbit.mu) * (1 + abs(mat.spin_orbit.alpha))
        for K, Kp, g, gp in _pairs(rng, G, count):
            ref = soc_definition(mat, K, Kp, g, gp)
            got = soc_fn(mat, K, Kp, np.asarray(g, float), np.asarray(gp, float))
            worst = max(worst, float(np.abs(got - ref).max()) / (1e-10 * scale))
    return worst


# --------------------------------------------------------------- tests ----

def test_nonlocal_element_from_definition(oracle):
    """The oracle's non-local element equals the partial-wave definition (square and gaussian wells, l = 0 and 2,
    the beta term at K != 0, both sites) to 1e-9 relative."""
    assert _nl_worst(oracle.nl_element) <= 1.0


def test_soc_block_from_definition(oracle):
    """The oracle's spin-orbit block equals the per-site sum with Slater radial factors by quadrature."""
    assert _soc_worst(oracle.soc_block) <= 1.0


@pytest.mark.parametrize("sym,nl,soc", [("Si", True, False), ("GaAs", True, True), ("InP", True, False),
                                        ("Ge", True, True)])
def test_assembled_hamiltonian_from_definition(oracle, sym, nl, soc):
    """Every element of SetMatrix's H(k) (kinetic, local, non-local, spin-orbit blocks): n = 3 (N = 15) without
    spin-orbit, n = 2 (N = 9, dim 18) with it."""
    mat = load(sym)
    G = oracle.basis(2 if soc else 3)
    rng = np.random.default_rng(len(sym) + 10 * nl + soc)
    for k in (rng.uniform(-1, 1, 3), np.zeros(3)):
        ref = hamiltonian_definition(mat, G, k, nl, soc)
        got = oracle.assemble(mat, G, k, nl=nl, soc=soc, herm=False)
        scale = np.abs(ref).max()
        assert np.abs(got - ref).max() <= 1e-11 * scale, (sym, k)


# --------------------------------------------------- mutation checks ----

# (name, [(original text, replacement)], which check must fail)
MUTANTS = [
    ("gaussian l=2 prefactor 5 pi^1.5 -> 4 pi^1.5",
     [("V2 = 5.0 * std::pow(kPi, 1.5) * (std::pow(m.R2_anion, 3.0)",
       "V2 = 4.0 * std::pow(kPi, 1.5) * (std::pow(m.R2_anion, 3.0)")], "nl"),
    ("sign of the antisymmetric sin term",
     [("std::sin(2.0 * kPi * Gtau) * Va", "-std::sin(2.0 * kPi * Gtau) * Va")], "nl"),
    ("beta term with |K1|^2 instead of |K1||K2|",
     [("m.beta_0_anion * (n1 * n2 - std::pow(kF, 2.0))", "m.beta_0_anion * (n1 * n1 - std::pow(kF, 2.0))")], "nl"),
    ("square-well l=2 without P2 (cation)",
     [("V2 = V_pre * m.A2_cation * (2 * 2 + 1) * leg2 * F2;", "V2 = V_pre * m.A2_cation * (2 * 2 + 1) * F2;")], "nl"),
    ("l=2 square well at the l=0 radius (anion)",
     [("F_l(K1, K2, m.R2_anion, 2)", "F_l(K1, K2, m.R0_anion, 2)")], "nl"),
    ("lambda_s and lambda_a swapped",
     [("const cplx phase(lanti * std::sin(Gtau), -lsym * std::cos(Gtau));",
       "const cplx phase(lsym * std::sin(Gtau), -lanti * std::cos(Gtau));")], "soc"),
    ("sign of the spin-orbit sin term",
     [("const cplx phase(lanti * std::sin(Gtau), -lsym * std::cos(Gtau));",
       "const cplx phase(-lanti * std::sin(Gtau), -lsym * std::cos(Gtau));")], "soc"),
    ("alpha applied to the cation instead of the anion",
     [("const double l1 = m.mu_soc * B_radial", "const double l1 = m.alpha_soc * m.mu_soc * B_radial"),
      ("const double l2 = m.alpha_soc * m.mu_soc * B_radial", "const double l2 = m.mu_soc * B_radial")], "soc"),
    ("cation radial factor with the anion period",
     [("m.mu_soc * B_radial(m.period_cation, lK, m.radial_extent_cation)",
       "m.mu_soc * B_radial(m.period_anion, lK, m.radial_extent_cation)")], "soc"),
    ("z component of K x K' reversed",
     [("const double a00 = K.x * Kp.y - K.y * Kp.x;", "const double a00 = K.y * Kp.x - K.x * Kp.y;")], "soc"),
]


class _Mutant:
    """ctypes view of a mutated oracle build: nl_element and soc_block only."""

    def __init__(self, path):
        self.L = C.CDLL(path)
        D, P = C.POINTER(C.c_double), C.c_void_p
        self.L.orc_nl_element.argtypes = [P, D, D, D]
        self.L.orc_soc_block.argtypes = [P, D, D, D, D, D]

    @staticmethod
    def _d(a):
        return np.ascontiguousarray(a, dtype=np.float64)

    def nl_element(self, mat, K1, K2):
        out = np.zeros(2)
        cm = mat.to_c()
        a, b = self._d(K1), self._d(K2)
        self.L.orc_nl_element(C.byref(cm), a.ctypes.data_as(C.POINTER(C.c_double)),
                              b.ctypes.data_as(C.POINTER(C.c_double)), out.ctypes.data_as(C.POINTER(C.c_double)))
        return complex(out[0], out[1])

    def soc_block(self, mat, K, Kp, G, Gp):
        out = np.zeros(8)
        cm = mat.to_c()
        args = [self._d(v) for v in (K, Kp, G, Gp)]
        self.L.orc_soc_block(C.byref(cm), *[v.ctypes.data_as(C.POINTER(C.c_double)) for v in args],
                             out.ctypes.data_as(C.POINTER(C.c_double)))
        return (out[0::2] + 1j * out[1::2]).reshape(2, 2)


@pytest.fixture(scope="module")
def mutant_builds():
    """Every mutant of oracle/epm_oracle.cpp compiled into its own library (g++, CPU)."""
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    src = open(os.path.join(ORACLE_DIR, "epm_oracle.cpp")).read()
    tmp = tempfile.mkdtemp(prefix="oracle_mutants_")

    def build(i):
        text = src
        for old, new in MUTANTS[i][1]:
            assert text.count(old) == 1, f"mutation site not found once: {old!r}"
            text = text.replace(old, new)
        cpp = os.path.join(tmp, f"m{i}.cpp")
        with open(cpp, "w") as fh:
            fh.write(text)
        so = os.path.join(tmp, f"m{i}.so")
        subprocess.run(["g++", "-std=c++17", "-O1", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
                        f"-I{ORACLE_DIR}", "-o", so, cpp], check=True, capture_output=True)
        return so

    with ThreadPoolExecutor(max_workers=4) as ex:
        libs = list(ex.map(build, range(len(MUTANTS))))
    yield libs
    shutil.rmtree(tmp, ignore_errors=True)


@pytest.mark.parametrize("i", range(len(MUTANTS)), ids=[m[0] for m in MUTANTS])
def test_definition_checks_catch_mutant(mutant_builds, i):
    m = _Mutant(mutant_builds[i])
    kind = MUTANTS[i][2]
    worst = _nl_worst(m.nl_element, count=4) if kind == "nl" else _soc_worst(m.soc_block, count=4)
    assert worst > 1e3, f"the {kind} definition check does not see mutant '{MUTANTS[i][0]}' (worst {worst:.3g})"
