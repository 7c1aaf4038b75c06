"""BZ_States on the CPU: the oracle's tetrahedron-weighted dielectric function
pinned against the single-k dielectric sums (themselves pinned by the free-
electron limit) at vertex k-points computed in numpy (position / (2 pi / a),
plus q), combined with |V|/4 weights from numpy determinants, and the
prefactor against the SI constants; the host-side checks and the CSV writer
of the C ABI."""
from __future__ import annotations

import gzip
import os
import shutil

import numpy as np
import pytest

from oracle import oracle
from ultimateepm_amd import mesh
from ultimateepm_amd._abi import UepmError
from ultimateepm_amd.material import PARAMS_DIR, load

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "bz_mesh.msh.gz")
MAT = load("Si", os.path.join(PARAMS_DIR, "materials-cohen.yaml"))  # apps/fullstates.cpp:45
A = MAT.get_lattice_constant_meter()
QE, EPS0 = 1.602176634e-19, 8.8541878128e-12  # physical_constants.hpp:20,31


@pytest.fixture(scope="module")
def geometry(tmp_path_factory):
    d = tmp_path_factory.mktemp("states")
    p = d / "bz_mesh.msh"
    with gzip.open(FIXTURE, "rb") as src, open(p, "wb") as dst:
        shutil.copyfileobj(src, dst)
    m = mesh.MeshBZ.read_mesh_geometry_from_msh_file(None, str(p), A)
    g = m.geometry()
    m.close()
    return g


def sub_mesh(g, first: int, count: int):
    """Tetrahedra [first, first + count) with their vertices renumbered."""
    tet = g["tetrahedra"][first:first + count]
    ids, inv = np.unique(tet.ravel(), return_inverse=True)
    return g["xyz"][ids], inv.reshape(-1, 4).astype(np.int32)


def test_tetrahedron_sum_is_the_weighted_vertex_sums(geometry):
    xyz, tet = sub_mesh(geometry, 500, 30)
    G = oracle.basis(6)
    nb, eta = 8, 0.05
    q = np.array([0.05, 0.02, 0.01])
    en = np.linspace(0.0, 12.0, 49)
    eps, sums, vol = oracle.bz_states_dielectric(MAT, G, False, nb, xyz, tet, q, en, eta)
    k = xyz / (2.0 * np.pi / A)
    per_vertex = np.stack([oracle.dielectric(MAT, G, False, nb, kv[None, :], q[None, :], en, eta)[0] for kv in k])
    e = xyz[tet[:, 1:]] - xyz[tet[:, :1]]
    vols = np.abs(np.linalg.det(e)) / 6.0
    want = np.zeros_like(en)
    for t, v in zip(tet, vols):
        want += v / 4.0 * per_vertex[t].sum(axis=0)
    scale = np.abs(want).max()
    assert scale > 0.0
    assert np.abs(sums - want).max() <= 1e-12 * scale
    assert vol == pytest.approx(vols.sum(), rel=1e-13)
    q_si = q * (2.0 * np.pi / A)
    pref = QE / (EPS0 * np.dot(q_si, q_si)) * 2.0 / (2.0 * np.pi) ** 3
    assert np.allclose(eps, 1.0 + pref * sums, rtol=1e-13, atol=0.0)


def test_csv_writer_matches_the_stream_format(tmp_path):
    """export_dielectric_function: header, then "E,eps" with the stream's
    default six significant digits."""
    en = np.array([0.0, 0.01, 1.5, 19.99, 20.0])
    eps = np.array([12.3456789, 1e-7, -3.25, 1234567.0, 1.0])
    prefix = str(tmp_path / "TEST_DIELECTRIC_FUNCTION_")
    st = mesh.lib().uepm_write_dielectric_csv(prefix.encode(), mesh._d(en), mesh._d(eps), len(en))
    assert st == 0
    text = open(prefix + "_dielectric_function.csv").read()
    want = "Energy (eV),Dielectric function\n" + "".join(f"{a:g},{b:g}\n" for a, b in zip(en, eps))
    assert text == want
    assert "1.23457e+06" in text and "1e-07" in text


def test_states_calls_check_their_arguments(geometry):
    xyz, tet = sub_mesh(geometry, 0, 4)
    m = mesh.BZStates.from_arrays(None, xyz, tet)
    with pytest.raises(UepmError):
        m.compute_dielectric_function(np.zeros(3), 0.1)  # no eigenstates yet
    with pytest.raises(UepmError):
        m.eigen_states()
    with pytest.raises(UepmError):
        m.export_full_eigenstates(None)
    assert mesh.lib().uepm_bzmesh_compute_states(m._h, None, None) != 0
    assert m.compute_fermi_level(1e21, 300.0) == 0.0 and m.get_dielectric_function([0, 0, 0], 1.0) == 1.0
    m.close()
