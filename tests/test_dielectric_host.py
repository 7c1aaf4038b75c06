"""Dielectric function on the CPU: the oracle pinned by the free-electron
limit, the reference's bilinear overlap and q -> 0 orthogonality; the host
merge and Kramers-Kronig steps of the C ABI against their formulas."""
from __future__ import annotations

import math

import numpy as np

from oracle import oracle
from ultimateepm_amd import api
from ultimateepm_amd.material import Material, load

HBAR, ME, QE = 6.62607015e-34 / (2 * math.pi), 9.1093837015e-31, 1.602176634e-19  # physical_constants.hpp


def _lorentz(d, E, eta):
    return (d - E) / ((d - E) ** 2 + eta ** 2) + (d + E) / ((d + E) ** 2 + eta ** 2)


def test_free_electron_limit_closed_form():
    """V = 0: eigenvectors are plane waves, so an overlap is 1 when the
    conduction level at k+q and the valence level at k sit on the same G,
    else 0; the energies are hbar^2/(2 m q)(2 pi/a)^2 |k+G|^2."""
    m = Material("free", 5.43, (0.0,) * 4, (0.0,) * 4)
    G = oracle.basis(6)
    nb = 10
    rng = np.random.default_rng(4)
    ks = rng.uniform(-0.8, 0.8, (6, 3))
    qs = np.array([[0.1, 0.05, 0.02], [0.0, 0.0, 0.0]])
    en = np.linspace(0.0, 30.0, 61)
    eta = 0.05
    got = oracle.dielectric(m, G, False, nb, ks, qs, en, eta)
    kin = HBAR ** 2 / (2 * ME * QE) * (2 * math.pi / 5.43e-10) ** 2
    for iq, q in enumerate(qs):
        qq = q if np.linalg.norm(q) > 1e-15 else np.full(3, 1e-15)
        want = np.zeros_like(en)
        for k in ks:
            ek = kin * np.sum((k + G) ** 2, axis=1)
            ekq = kin * np.sum((k + qq + G) ** 2, axis=1)
            ok, oq = np.argsort(ek)[:nb], np.argsort(ekq)[:nb]
            for c in range(4, nb):
                for v in range(4):
                    if oq[c] == ok[v]:
                        want += _lorentz(ekq[oq[c]] - ek[ok[v]], en, eta)
        assert np.allclose(got[iq], want, rtol=1e-10, atol=1e-12 * np.abs(want).max())
        if iq == 1:
            assert not want.any() and np.abs(got[iq]).max() < 1e-20  # q = 0: no level changes its G


def test_bilinear_overlap_of_the_reference():
    """GaAs (complex Hermitian H): the weight is |sum_i x_c,i x_v,i|^2, not |<c|v>|^2."""
    m = load("GaAs")
    G = oracle.basis(6)
    nb = 8
    ks = np.array([[0.13, -0.21, 0.34], [0.5, 0.1, -0.2], [-0.3, 0.6, 0.05]])
    q = np.array([[0.07, 0.02, -0.04]])
    en = np.linspace(0.0, 15.0, 31)
    eta = 0.1
    got = oracle.dielectric(m, G, True, nb, ks, q, en, eta)[0]
    from ultimateepm_amd._abi import NONLOCAL, VECTORS

    ek, xk, _ = oracle.solve(m, G, NONLOCAL | VECTORS, nb, ks)
    eq, xq, _ = oracle.solve(m, G, NONLOCAL | VECTORS, nb, ks + q)
    bil, ses = np.zeros_like(en), np.zeros_like(en)
    for i in range(len(ks)):
        for c in range(4, nb):
            for v in range(4):
                L = _lorentz(eq[i, c] - ek[i, v], en, eta)
                bil += abs(np.sum(xq[i, c] * xk[i, v])) ** 2 * L
                ses += abs(np.vdot(xq[i, c], xk[i, v])) ** 2 * L
    assert np.allclose(got, bil, rtol=1e-11, atol=1e-13 * np.abs(bil).max())
    assert np.abs(ses - bil).max() > 1e-3 * np.abs(bil).max()  # the two forms differ for complex H


def test_q_zero_real_symmetric_hamiltonian():
    """Si (V_A = 0, equal species): H is real symmetric, so at q -> 0 every
    interband weight is an orthogonality residue."""
    m = load("Si")
    G = oracle.basis(6)
    ks = np.array([[0.11, 0.23, -0.37], [0.6, -0.1, 0.2]])
    en = np.linspace(0.0, 10.0, 11)
    S = oracle.dielectric(m, G, True, 8, ks, np.zeros((1, 3)), en, 0.05)
    assert np.abs(S).max() < 1e-20


def test_merge_and_kramers_kronig_formulas():
    q = np.array([[0.1, 0.2, 0.0], [0.0, 0.0, 0.0]])
    S = np.array([[1.0, 2.0, 3.0, 4.0], [0.5, 0.25, 0.0, -1.0]])
    eps = api.dielectric_merge(q, S, 8)
    q2 = np.array([math.pow(math.sqrt(0.1 * 0.1 + 0.2 * 0.2 + 0.0), 2), 1e-14])
    want = 1.0 + (2 * math.pi / q2[:, None]) * (S * (1.0 / 8))
    assert np.array_equal(eps, want)
    n = 5
    er = np.array([[1.0, 2.0, 0.5, 3.0, 1.5]])
    K = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            if i != j:
                K[i, j] = (i / (j * j - i * i)) * (-2.0 / math.pi)
    assert np.allclose(api.kramers_kronig(er), (K @ (er[0] - 1.0))[None, :], rtol=1e-15, atol=1e-15)
    assert api.kramers_kronig(er)[0, 0] == 0.0  # row 0 of K is zero
