"""BZ_States on the GPU: eigenstates at the mesh vertices and at vertex + q
(uepm_bzmesh_compute_states) against the plan's own solve, the tetrahedron-
weighted dielectric function against the oracle, the precision form, the
eigenstate files, and the apps/fullstates.cpp run over the whole example mesh."""
from __future__ import annotations

import gzip
import os
import shutil
import time

import numpy as np
import pytest

from oracle import oracle
from ultimateepm_amd import api, mesh
from ultimateepm_amd.material import PARAMS_DIR, load

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "bz_mesh.msh.gz")
MAT = load("Si", os.path.join(PARAMS_DIR, "materials-cohen.yaml"))  # apps/fullstates.cpp:45
A = MAT.get_lattice_constant_meter()
Q = np.array([0.05, 0.02, 0.01])
# Relative to the largest |S|: weights are squared overlaps of eigenvectors
# from two FP64 solvers (~1e-12 apart), summed in another order.
REL_TOL = 1e-8


@pytest.fixture(scope="module")
def msh_path(tmp_path_factory):
    d = tmp_path_factory.mktemp("statesgpu")
    p = d / "bz_mesh.msh"
    with gzip.open(FIXTURE, "rb") as src, open(p, "wb") as dst:
        shutil.copyfileobj(src, dst)
    return str(p)


@pytest.fixture(scope="module")
def geometry(msh_path):
    m = mesh.MeshBZ.read_mesh_geometry_from_msh_file(None, msh_path, A)
    g = m.geometry()
    m.close()
    return g


def sub_mesh(xyz, tet):
    ids, inv = np.unique(tet.ravel(), return_inverse=True)
    return xyz[ids], inv.reshape(-1, 4).astype(np.int32)


def gapped_tetrahedra(g, G, nb, count, start=0):
    """The first `count` tetrahedra from `start` whose vertices have level nb
    apart from level nb - 1 at k and at k + Q: the sum over conduction levels
    4 .. nb-1 is then independent of the basis chosen in degenerate levels."""
    tet = g["tetrahedra"][start:start + 8 * count]
    ids = np.unique(tet.ravel())
    k = g["xyz"][ids] / (2.0 * np.pi / A)
    ok = np.ones(len(ids), dtype=bool)
    for kk in (k, k + Q):
        ev, _, _ = oracle.solve(MAT, G, 0, nb + 1, kk, nthreads=8)
        ok &= (ev[:, nb] - ev[:, nb - 1]) > 1e-5
    good = set(ids[ok].tolist())
    keep = [t for t in tet if all(int(v) in good for v in t)][:count]
    assert len(keep) == count
    return sub_mesh(g["xyz"], np.array(keep))


def test_states_are_the_plan_solves(gpu_ctx, geometry):
    print("case: states of 200 tetrahedra, Si n=8 nb=12, k and k+q", flush=True)
    xyz, tet = sub_mesh(geometry["xyz"], geometry["tetrahedra"][:200])
    plan = api.Plan(gpu_ctx, MAT, api.basis_vectors(8), api.VECTORS, 12)
    m = mesh.BZStates.from_arrays(gpu_ctx, xyz, tet)
    m.compute_eigenstates(plan)
    m.compute_shifted_eigenstates(Q, plan)
    k = xyz / (2.0 * np.pi / A)
    for shifted, kk in ((False, k), (True, k + Q)):
        E, X = m.eigen_states(shifted)
        Ep, Xp, _ = plan.solve(kk)
        assert np.array_equal(E, Ep) and np.array_equal(X, Xp.reshape(X.shape))
    m.close()
    plan.close()


@pytest.mark.parametrize("shells,nb", [(8, 12), (6, 8)])
def test_dielectric_matches_the_oracle(gpu_ctx, geometry, shells, nb):
    print(f"case: mesh dielectric Si n={shells} nb={nb}, 120 tetrahedra", flush=True)
    G = api.basis_vectors(shells)
    xyz, tet = gapped_tetrahedra(geometry, G, nb, 120, start=3000)
    en = np.linspace(0.0, 20.0, 201)
    eta = 0.05
    plan = api.Plan(gpu_ctx, MAT, G, api.VECTORS, nb)
    m = mesh.BZStates.from_arrays(gpu_ctx, xyz, tet)
    m.compute_eigenstates(plan)
    m.compute_shifted_eigenstates(Q, plan)
    eps = m.compute_dielectric_function(en, eta)
    eps_o, sums_o, vol_o = oracle.bz_states_dielectric(MAT, G, False, nb, xyz, tet, Q, en, eta)
    scale = np.abs(sums_o).max()
    assert scale > 0.0
    assert np.abs(m.dielectric_sums - sums_o).max() <= REL_TOL * scale
    assert m.total_volume_integrated == vol_o
    assert np.abs(eps - eps_o).max() <= REL_TOL * np.abs(eps_o - 1.0).max()
    m.close()
    plan.close()


def test_precision_form_states(gpu_ctx, geometry):
    print("case: mesh dielectric, precision form vs FP64", flush=True)
    G = api.basis_vectors(8)
    xyz, tet = gapped_tetrahedra(geometry, G, 12, 80, start=9000)
    en = np.linspace(0.0, 15.0, 64)
    out = []
    for flags in (api.VECTORS, api.VECTORS | api.MIXED):
        plan = api.Plan(gpu_ctx, MAT, G, flags, 12)
        m = mesh.BZStates.from_arrays(gpu_ctx, xyz, tet)
        m.compute_eigenstates(plan)
        m.compute_shifted_eigenstates(Q, plan)
        m.compute_dielectric_function(en, 0.1)
        out.append(m.dielectric_sums.copy())
        m.close()
        plan.close()
    assert np.abs(out[1] - out[0]).max() <= 1e-5 * np.abs(out[0]).max()


def test_eigenstate_files(gpu_ctx, geometry, tmp_path):
    print("case: export_full_eigenstates of 3 tetrahedra", flush=True)
    xyz, tet = sub_mesh(geometry["xyz"], geometry["tetrahedra"][:3])
    plan = api.Plan(gpu_ctx, MAT, api.basis_vectors(4), api.VECTORS, 6)
    m = mesh.BZStates.from_arrays(gpu_ctx, xyz, tet)
    m.compute_eigenstates(plan)
    m.compute_shifted_eigenstates(Q, plan)
    stale = tmp_path / "eigenstates" / "stale.txt"
    stale.parent.mkdir()
    stale.write_text("x")
    m.export_full_eigenstates(str(tmp_path))
    assert not stale.exists()
    dim = plan.dim
    for shifted, (vname, xname) in ((False, ("eigenvalues_", "eigenvectors_")),
                                    (True, ("shiftedeigenvalues_", "shiftedeigenvectors_"))):
        E, X = m.eigen_states(shifted)
        for k in range(len(xyz)):
            row = (tmp_path / "eigenstates" / "eigenvalues" / f"{vname}{k}.txt").read_text()
            vals = np.array(row.split(), dtype=float)
            assert np.allclose(vals, E[k], rtol=1e-5, atol=1e-5)
            lines = (tmp_path / "eigenstates" / "eigenvectors" / f"{xname}{k}.txt").read_text().splitlines()
            assert len(lines) == dim and len({len(line) for line in lines}) == 1  # aligned columns
            got = np.array([[complex(*map(float, tok.strip("()").split(","))) for tok in line.split()]
                            for line in lines])
            assert got.shape == (dim, 6)
            assert np.allclose(got, X[k].T, atol=1e-5)
    m.close()
    plan.close()


def test_fullstates_run_over_the_example_mesh(gpu_ctx, msh_path, geometry, tmp_path):
    """apps/fullstates.cpp: Si (materials-cohen), n = 10, 12 bands, q = (3e-14, 0, 0),
    energies 0 .. 20 eV in steps of 0.01, eta = 0.05, over all 14705 vertices."""
    print("case: fullstates over 14705 vertices, Si n=10 nb=12, energies 0..20 eV step 0.01", flush=True)
    m = mesh.BZStates.read_mesh_geometry_from_msh_file(gpu_ctx, msh_path, A)
    plan = api.Plan(gpu_ctx, MAT, api.basis_vectors(10), api.VECTORS, 12)
    t0 = time.perf_counter()
    m.compute_eigenstates(plan)
    m.compute_shifted_eigenstates(np.array([3e-14, 0.0, 0.0]), plan)
    t1 = time.perf_counter()
    en = []
    e = 0.0
    while e <= 20.0 + 0.01:  # the reference's accumulating loop (fullstates.cpp:88-90)
        en.append(e)
        e += 0.01
    eps = m.compute_dielectric_function(np.array(en), 0.05)
    t2 = time.perf_counter()
    print(f"states {t1 - t0:.3f} s, dielectric {t2 - t1:.3f} s, eps[0] = {eps[0]:.6g}", flush=True)
    assert len(eps) == len(en) and np.all(np.isfinite(eps))
    assert m.total_volume_integrated == pytest.approx(geometry["total_volume"], rel=1e-12)
    m.export_dielectric_function(str(tmp_path / "TEST_DIELECTRIC_FUNCTION_"))
    lines = (tmp_path / "TEST_DIELECTRIC_FUNCTION__dielectric_function.csv").read_text().splitlines()
    assert lines[0] == "Energy (eV),Dielectric function" and len(lines) == len(en) + 1
    m.close()
    plan.close()
