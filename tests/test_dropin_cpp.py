"""The drop-in C++ EPP classes (ultimateepm_amd/epp/, libuepm_epp.so) driven by
tests/cpp/test_dropin.cpp, built by the package build into
ultimateepm_amd/_build/test_dropin.  The GPU leg compares what the classes
return with the oracle."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

from ultimateepm_amd._abi import GRADIENTS, NONLOCAL, SOC, VECTORS
from ultimateepm_amd.material import PARAMS_DIR, load

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "ultimateepm_amd", "_build", "test_dropin")
MATS = os.path.join(PARAMS_DIR, "materials.yaml")
COHEN = os.path.join(PARAMS_DIR, "materials-local-cohen.yaml")


def _bin():
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} is missing: run the package build first")
    return BIN


def test_dropin_host_side():
    files = [MATS]
    ref_file = "/root/reference/parameter_files/materials.yaml"  # the reference's own layout, when present
    if os.path.exists(ref_file):
        files.append(ref_file)
    r = subprocess.run([_bin(), "cpu", *files], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "cpu checks: ok" in r.stdout


@pytest.mark.gpu
def test_dropin_band_structures_vs_oracle(oracle, tmp_path):
    print("case: drop-in C++ classes, test_dropin gpu", flush=True)
    r = subprocess.run([_bin(), "gpu", MATS, COHEN, str(tmp_path)], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    # 1. path, local Cohen-Bergstresser Si, n=10, 16 bands (raw, before AdjustValues)
    ks, _ = oracle.kpath("LGXWKULWXKG", 200)
    ev_ref, _, _ = oracle.solve(load("Si", COHEN), oracle.basis(10), 0, 16, ks, nthreads=8)
    ev = np.loadtxt(tmp_path / "path_raw.txt")
    assert np.abs(ev - ev_ref).max() <= 1e-9
    files = [f for f in os.listdir(tmp_path) if f.startswith("EPM_Si_nb_bands_16_path_LGXWKULWXKG_size_basis_137")]
    assert len(files) == 1
    # 6. DielectricFunction: raw sums against the oracle, the merged CSV
    kd = np.loadtxt(tmp_path / "dielectric_k.txt")
    raw = np.loadtxt(tmp_path / "dielectric_raw.txt")
    qs = np.array([[0.0, 0.0, 0.0], [0.05, 0.02, 0.0]])
    en = 0.2 * np.arange(101)
    ref = oracle.dielectric(load("Si"), oracle.basis(8), True, 16, kd, qs, en, 0.05)
    assert np.abs(raw[1] - ref[1]).max() <= 1e-8 * np.abs(ref[1]).max()
    assert np.abs(raw[0]).max() < 1e-12 and np.abs(ref[0]).max() < 1e-12  # q -> 0, real-symmetric H
    csv = np.loadtxt(tmp_path / "Si__00001_0.050000_0.020000_0.000000.csv", delimiter=",", skiprows=1)
    eps = 1.0 + (2 * np.pi / (0.05 ** 2 + 0.02 ** 2)) * raw[1] / len(kd)
    assert np.allclose(csv[:, 1], eps, rtol=1e-5)
    text = (tmp_path / files[0]).read_text().splitlines()
    assert text[0] == "# Material Si" and text[1] == "# NBands 16" and text[2] == "# Nonlocal No"
    assert text[5].startswith("band_0,band_1") and len(text) == 6 + 200
    # 2. Ge non-local + SOC, n=8, 2x8 bands, gradients
    kl = np.array([[0.1, 0.2, 0.3], [-0.4, 0.05, 0.6], [0.5, 0.5, 0.5], [0, 0, 0]])
    ev_ref, _, gr_ref = oracle.solve(load("Ge"), oracle.basis(8), NONLOCAL | SOC | GRADIENTS | VECTORS, 16, kl)
    ev = np.loadtxt(tmp_path / "ge_soc.txt")
    assert np.abs(ev - ev_ref).max() <= 1e-9
    gr = np.loadtxt(tmp_path / "ge_soc_grad.txt").reshape(4, 16, 3)
    # gradients of degenerate levels depend on the basis chosen inside the degenerate space; compare non-degenerate
    a = 5.65e-10
    for q in range(4):
        for b in range(16):
            gaps = np.abs(ev_ref[q] - ev_ref[q, b])
            gaps[b] = np.inf
            if b % 2 == 0 and gaps.min() > 1e-6 and np.sort(gaps)[0] > 1e-6:
                assert np.allclose(gr[q, b], gr_ref[q, b], atol=1e-8 * a * 10)
    # 3. per-k Hamiltonian, GaAs non-local n=6 (N=59): all eigenpairs
    raw = np.loadtxt(tmp_path / "gaas_all.txt", max_rows=1)
    vec = np.loadtxt(tmp_path / "gaas_all.txt", skiprows=1, max_rows=59 * 59)
    V = (vec[:, 0] + 1j * vec[:, 1]).reshape(59, 59).T  # column j = eigenvector j
    k = np.array([0.13, -0.27, 0.41])
    m = load("GaAs")
    G = oracle.basis(6)
    H = oracle.assemble(m, G, k, nl=True, soc=False, herm=True)
    w = np.linalg.eigvalsh(H)
    assert np.abs(raw - w).max() <= 1e-9
    assert np.abs(H @ V - V * raw).max() <= 1e-8
    assert np.abs(V.conj().T @ V - np.eye(59)).max() <= 1e-10


@pytest.mark.gpu
def test_dropin_mesh_bands_and_dos(gpu_ctx, tmp_path):
    """BandsOnBZ and DOS_MeshBZ call sequences through the drop-in MeshBZ,
    against the same steps through the C ABI."""
    import gzip
    import shutil

    from ultimateepm_amd import api, mesh

    msh = str(tmp_path / "bz_mesh.msh")
    with gzip.open(os.path.join(ROOT, "tests", "golden", "bz_mesh.msh.gz"), "rb") as src, open(msh, "wb") as dst:
        shutil.copyfileobj(src, dst)
    print("case: drop-in MeshBZ, test_dropin mesh", flush=True)
    r = subprocess.run([_bin(), "mesh", COHEN, msh, str(tmp_path)], capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr
    si = load("Si", COHEN)
    m = mesh.MeshBZ.read_mesh_geometry_from_msh_file(gpu_ctx, msh, si.get_lattice_constant_meter())
    plan = api.Plan(gpu_ctx, si, api.basis_vectors(10), GRADIENTS, 20)
    m.compute_band_structure_over_mesh(plan, use_iwedge=False)
    E, G = m.bands()
    E_full = np.loadtxt(tmp_path / "mesh_E_full.txt")
    E_iw = np.loadtxt(tmp_path / "mesh_E_iw.txt")
    assert np.array_equal(E_full, E)
    assert np.abs(E_iw - E).max() <= 1e-9
    views = mesh.MshFile(str(tmp_path / "bz_mesh_bands.msh")).views
    assert [v[0] for v in views[:4]] == ["band_0", "grad_band_0", "band_1", "grad_band_1"] and len(views) == 32
    assert np.array_equal(views[0][1], E[:, 3]) and np.array_equal(views[1][1], G[:, 3])
    assert np.array_equal(views[8][1], E[:, 4])  # first conduction band after the 4 valence ones
    vtk = open(tmp_path / "bz_mesh_bands.vtk").read()
    assert vtk.count("SCALARS band_energy_") == 20 and vtk.count("VECTORS band_grad_") == 20
    # DOS_MeshBZ: the same steps through the C ABI
    d = mesh.MeshBZ.read_mesh_geometry_from_msh_file(gpu_ctx, msh, si.get_lattice_constant_meter())
    d.read_mesh_bands_from_msh_file(str(tmp_path / "bz_mesh_bands.msh"), -1, -1, True, False)
    d.apply_scissor(1.12)
    csv = np.loadtxt(tmp_path / "DOS_bz_mesh_bands.csv", delimiter=",", skiprows=1)
    header = open(tmp_path / "DOS_bz_mesh_bands.csv").readline().strip().split(",")
    assert header[:2] == ["energy_band_0", "dos_band_0"] and header[-1] == "DumbColumn"
    for b in range(16):
        en, dos = d.compute_dos_band_at_band_auto(b, 200)
        assert np.allclose(csv[:, 2 * b], en, rtol=1e-5, atol=1e-5)
        assert np.allclose(csv[:, 2 * b + 1], dos, rtol=1e-5, atol=1e-6 * dos.max())
    plan.close()
    m.close()
    d.close()


@pytest.mark.gpu
def test_dropin_fullstates(gpu_ctx, tmp_path):
    """apps/fullstates.cpp through the drop-in BZ_States, against the same
    steps through the C ABI (same kernels, so the same bits)."""
    import gzip
    import shutil

    from ultimateepm_amd import api, mesh

    cohen = os.path.join(PARAMS_DIR, "materials-cohen.yaml")
    msh = str(tmp_path / "bz_mesh.msh")
    with gzip.open(os.path.join(ROOT, "tests", "golden", "bz_mesh.msh.gz"), "rb") as src, open(msh, "wb") as dst:
        shutil.copyfileobj(src, dst)
    print("case: drop-in BZ_States, test_dropin states", flush=True)
    r = subprocess.run([_bin(), "states", cohen, msh, str(tmp_path)], capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr
    si = load("Si", cohen)
    m = mesh.BZStates.read_mesh_geometry_from_msh_file(gpu_ctx, msh, si.get_lattice_constant_meter())
    plan = api.Plan(gpu_ctx, si, api.basis_vectors(10), VECTORS, 12)
    m.compute_eigenstates(plan)
    m.compute_shifted_eigenstates(np.array([3e-14, 0.0, 0.0]), plan)
    E, _ = m.eigen_states(vectors=False)
    assert np.array_equal(np.loadtxt(tmp_path / "states_E.txt"), E)
    en = []
    e = 0.0
    while e <= 20.0 + 0.01:
        en.append(e)
        e += 0.01
    eps = m.compute_dielectric_function(np.array(en), 0.05)
    csv = np.loadtxt(tmp_path / "TEST_DIELECTRIC_FUNCTION__dielectric_function.csv", delimiter=",", skiprows=1)
    assert csv.shape == (len(en), 2)
    assert np.allclose(csv[:, 0], en, rtol=1e-5, atol=1e-6)
    assert np.allclose(csv[:, 1], eps, rtol=1e-5, atol=0.0)
    assert "Total volume (k-space) integrated" in r.stdout and "EPS[0] = " in r.stdout
    plan.close()
    m.close()
