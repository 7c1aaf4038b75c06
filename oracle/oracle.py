"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker / baseline and never as
the thing measured or shipped.

The oracle restates the reference's per-k band solve; every function names the
reference file:line it follows in oracle/epm_oracle.cpp.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libuepm_ref.so")

_lib = None
_ref = None


def build(with_ref: bool = True) -> None:
    """Compile liboracle.so (and oracle/_ref when the reference tree exists)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    if with_ref and os.path.isdir("/root/reference/src/EPP"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build(with_ref=False)
        L = C.CDLL(LIB)
        D, P, I32 = C.POINTER(C.c_double), C.c_void_p, C.POINTER(C.c_int32)
        L.orc_last_error.restype = C.c_char_p
        L.orc_basis.argtypes = [C.c_int, I32, C.c_int]
        L.orc_kpath.argtypes = [C.c_char_p, C.c_uint, D, C.POINTER(C.c_uint)]
        L.orc_generalized_bessel.restype = C.c_double
        L.orc_generalized_bessel.argtypes = [C.c_double, C.c_double]
        L.orc_generalized_bessel_cylindrical.restype = C.c_double
        L.orc_generalized_bessel_cylindrical.argtypes = [C.c_double, C.c_double]
        L.orc_i2.restype = C.c_double
        L.orc_i2.argtypes = [C.c_double]
        L.orc_nl_element.argtypes = [P, D, D, D]
        L.orc_assemble.argtypes = [P, I32, C.c_int, D, C.c_int, C.c_int, C.c_int, D]
        L.orc_eigh.argtypes = [C.c_int, D, D, D]
        L.orc_tridiagonal.argtypes = [C.c_int, D, D, D]
        L.orc_solve.argtypes = [P, I32, C.c_int, C.c_uint, C.c_int, D, C.c_int64, D, D, D, C.c_int]
        L.orc_min_norm_gap.restype = C.c_double
        L.orc_min_norm_gap.argtypes = [P, I32, C.c_int, D]
        L.orc_in_first_bz.argtypes = [D]
        L.orc_adjust_values.restype = C.c_double
        L.orc_adjust_values.argtypes = [D, C.c_int64, C.c_int, C.c_int, C.c_int]
        L.orc_ryd_to_ev.restype = C.c_double
        L.orc_F_l.restype = C.c_double
        L.orc_F_l.argtypes = [D, D, C.c_double, C.c_int]
        L.orc_F2_gauss.restype = C.c_double
        L.orc_F2_gauss.argtypes = [D, D, C.c_double]
        L.orc_B_radial.restype = C.c_double
        L.orc_B_radial.argtypes = [C.c_int, C.c_double, C.c_double]
        L.orc_soc_block.argtypes = [P, D, D, D, D, D]
        U8 = C.POINTER(C.c_uint8)
        L.orc_tetra_dos.argtypes = [D, C.c_int64, I32, C.c_int64, U8, D, D, C.c_int, C.c_int, C.c_int, C.c_double,
                                    C.c_double, C.c_double, C.c_double, C.c_int, D, D, C.c_int]
        L.orc_same_symmetry.argtypes = [D, D, D, D]
        L.orc_dielectric.argtypes = [P, I32, C.c_int, C.c_int, C.c_int, D, C.c_int64, D, C.c_int64, D, C.c_int,
                                     C.c_double, D, C.c_int]
        L.orc_bz_states_dielectric.argtypes = [P, I32, C.c_int, C.c_int, C.c_int, D, C.c_int64, I32, C.c_int64, D, D,
                                               C.c_int, C.c_double, D, D, D, C.c_int]
        _lib = L
    return _lib


def ref_lib():
    """The reference's own standalone-compilable functions (oracle/_ref), or None."""
    global _ref
    if _ref is None and os.path.exists(REF_LIB):
        R = C.CDLL(REF_LIB)
        D = C.POINTER(C.c_double)
        R.ref_generalized_bessel.restype = C.c_double
        R.ref_generalized_bessel.argtypes = [C.c_double, C.c_double]
        R.ref_generalized_bessel_cylindrical.restype = C.c_double
        R.ref_generalized_bessel_cylindrical.argtypes = [C.c_double, C.c_double]
        R.ref_i2.restype = C.c_double
        R.ref_i2.argtypes = [C.c_double]
        R.ref_local_V.argtypes = [D, C.POINTER(C.c_int), C.c_double, D]
        R.ref_kpath.argtypes = [C.c_char_p, C.c_uint, D, C.POINTER(C.c_uint)]
        _ref = R
    return _ref


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _i(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def basis(n: int) -> np.ndarray:
    cnt = lib().orc_basis(n, None, 0)
    if cnt < 0:
        raise RuntimeError("BandStructure::Initialize: GenerateBasisVectors failed")
    G = np.zeros((cnt, 3), dtype=np.int32)
    lib().orc_basis(n, _i(G), cnt)
    return G


def kpath(labels: str, n: int):
    out = np.zeros((n, 3))
    pos = np.zeros(len(labels), dtype=np.uint32)
    cnt = lib().orc_kpath(labels.encode(), n, _d(out), pos.ctypes.data_as(C.POINTER(C.c_uint)))
    if cnt < 0:
        raise ValueError("invalid path")
    return out[:cnt], pos


def assemble(mat, G: np.ndarray, k, nl: bool, soc: bool, herm: bool = True) -> np.ndarray:
    G = np.ascontiguousarray(G, dtype=np.int32)
    N = len(G)
    dim = 2 * N if soc else N
    H = np.zeros((dim, dim), dtype=np.complex128, order="F")
    kk = np.ascontiguousarray(k, dtype=np.float64)
    cm = mat.to_c()
    r = lib().orc_assemble(C.byref(cm), _i(G), N, _d(kk), int(nl), int(soc), int(herm),
                           H.ctypes.data_as(C.POINTER(C.c_double)))
    if r < 0:
        raise RuntimeError(lib().orc_last_error().decode())
    return H


def eigh(H: np.ndarray, vectors: bool = False):
    H = np.asfortranarray(H, dtype=np.complex128)
    n = H.shape[0]
    w = np.zeros(n)
    V = np.zeros((n, n), dtype=np.complex128, order="F") if vectors else None
    r = lib().orc_eigh(n, H.ctypes.data_as(C.POINTER(C.c_double)), _d(w),
                       V.ctypes.data_as(C.POINTER(C.c_double)) if vectors else None)
    if r != 0:
        raise RuntimeError("Eigenvalue decomposition failed")
    return (w, V) if vectors else w


def solve(mat, G: np.ndarray, flags: int, nb: int, k: np.ndarray, nthreads: int = 1):
    """BandStructure::Compute_parallel restated: (evals, evecs|None, grads|None)."""
    from ultimateepm_amd import _abi

    G = np.ascontiguousarray(G, dtype=np.int32)
    k = np.ascontiguousarray(k, dtype=np.float64).reshape(-1, 3)
    nk = len(k)
    N = len(G)
    dim = 2 * N if flags & _abi.SOC else N
    ev = np.zeros((nk, nb))
    want_vec = bool(flags & (_abi.VECTORS | _abi.GRADIENTS))
    vec = np.zeros((nk, nb, dim), dtype=np.complex128) if want_vec else None
    gr = np.zeros((nk, nb, 3)) if flags & _abi.GRADIENTS else None
    cm = mat.to_c()
    r = lib().orc_solve(C.byref(cm), _i(G), N, flags, nb, _d(k), nk, _d(ev),
                        vec.ctypes.data_as(C.POINTER(C.c_double)) if want_vec else None,
                        _d(gr) if gr is not None else None, nthreads)
    if r != 0:
        raise RuntimeError(lib().orc_last_error().decode())
    return ev, vec, gr


def min_norm_gap(mat, G, k) -> float:
    G = np.ascontiguousarray(G, dtype=np.int32)
    kk = np.ascontiguousarray(k, dtype=np.float64)
    cm = mat.to_c()
    return lib().orc_min_norm_gap(C.byref(cm), _i(G), len(G), _d(kk))


def nl_element(mat, K1, K2) -> complex:
    out = np.zeros(2)
    cm = mat.to_c()
    lib().orc_nl_element(C.byref(cm), _d(np.ascontiguousarray(K1, dtype=np.float64)),
                         _d(np.ascontiguousarray(K2, dtype=np.float64)), _d(out))
    return complex(out[0], out[1])


def adjust_values(E: np.ndarray, soc: bool, min_cond_to_zero: bool = False) -> float:
    E = np.ascontiguousarray(E, dtype=np.float64)
    return lib().orc_adjust_values(_d(E), E.shape[0], E.shape[1], int(soc), int(min_cond_to_zero)), E


def splitmix64_bz_sample(seed: int, n: int) -> np.ndarray:
    """The seeded first-BZ sampler (spec in include/uepm.h), in numpy."""
    M = (1 << 64) - 1
    state = seed & M
    out = []
    while len(out) < n:
        v = []
        for _ in range(3):
            state = (state + 0x9E3779B97F4A7C15) & M
            z = state
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
            z = z ^ (z >> 31)
            u = (z >> 11) * (1.0 / 9007199254740992.0)
            v.append(2.0 * u - 1.0)
        if abs(v[0]) <= 1 and abs(v[1]) <= 1 and abs(v[2]) <= 1 and abs(v[0]) + abs(v[1]) + abs(v[2]) <= 1.5:
            out.append(v)
    return np.array(out, dtype=np.float64).reshape(-1, 3)


def F_l(K1, K2, R: float, l: int) -> float:
    """Radial overlap F_l for physical wave vectors K1, K2 [1/m] and well radius R [m]."""
    return lib().orc_F_l(_d(np.ascontiguousarray(K1, dtype=np.float64)),
                         _d(np.ascontiguousarray(K2, dtype=np.float64)), R, l)


def F2_gauss(K1, K2, R: float) -> float:
    return lib().orc_F2_gauss(_d(np.ascontiguousarray(K1, dtype=np.float64)),
                              _d(np.ascontiguousarray(K2, dtype=np.float64)), R)


def B_radial(period: int, K: float, xi: float) -> float:
    v = lib().orc_B_radial(period, K, xi)
    if v == -1e300:
        raise RuntimeError("SpinOrbitCorrection: Unsupported period.")
    return v


def soc_block(mat, K, Kp, G, Gp) -> np.ndarray:
    """2x2 spin-orbit block [[s00, s01], [s10, s11]] for reduced K = k+G, Kp = k+G'."""
    out = np.zeros(8)
    cm = mat.to_c()
    args = [_d(np.ascontiguousarray(v, dtype=np.float64)) for v in (K, Kp, G, Gp)]
    lib().orc_soc_block(C.byref(cm), *args, _d(out))
    z = out[0::2] + 1j * out[1::2]
    return z.reshape(2, 2)


def tetra_dos(pos, tet, E, G, band: int, n_points: int, e_min: float, e_max: float, reduce_bz_factor: float = 1.0,
              spin_degeneracy: float = 2.0, in_wedge=None, wedge_only: bool = False, nthreads: int = 8):
    """Tetrahedron DOS of one band (mesh_oracle.cpp): (energies, dos), each n_points long.

    pos (nv, 3) in 1/m, tet (nt, 4) vertex indices, E (nv, nb) eV, G (nv, nb, 3) eV m."""
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    tet = np.ascontiguousarray(tet, dtype=np.int32)
    E = np.ascontiguousarray(E, dtype=np.float64)
    G = np.ascontiguousarray(G, dtype=np.float64)
    wedge = None if in_wedge is None else np.ascontiguousarray(in_wedge, dtype=np.uint8)
    en, dos = np.zeros(n_points), np.zeros(n_points)
    st = lib().orc_tetra_dos(_d(pos), len(pos), _i(tet), len(tet),
                             wedge.ctypes.data_as(C.POINTER(C.c_uint8)) if wedge is not None else None,
                             _d(E), _d(G), E.shape[1], band, n_points, e_min, e_max, reduce_bz_factor,
                             spin_degeneracy, int(wedge_only), _d(en), _d(dos), nthreads)
    if st != 0:
        raise ValueError("orc_tetra_dos: bad arguments")
    return en, dos


def same_symmetry(k_iw, k_bz, v) -> np.ndarray:
    """apply_same_symmetry_operation restated (mesh_oracle.cpp)."""
    out = np.zeros(3)
    lib().orc_same_symmetry(*[_d(np.ascontiguousarray(a, dtype=np.float64)) for a in (k_iw, k_bz, v)], _d(out))
    return out


def dielectric(mat, G, non_local: bool, nb: int, k, q, energies, eta: float, nthreads: int = 8) -> np.ndarray:
    """compute_dielectric_function restated (dielectric_oracle.cpp): S[q][E], nq x n_energies."""
    G = np.ascontiguousarray(G, dtype=np.int32)
    k = np.ascontiguousarray(k, dtype=np.float64).reshape(-1, 3)
    q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1, 3)
    en = np.ascontiguousarray(energies, dtype=np.float64)
    out = np.zeros((len(q), len(en)))
    cm = mat.to_c()
    st = lib().orc_dielectric(C.byref(cm), _i(G), len(G), int(non_local), nb, _d(k), len(k), _d(q), len(q), _d(en),
                              len(en), eta, _d(out), nthreads)
    if st != 0:
        raise RuntimeError(f"orc_dielectric failed ({st})")
    return out


def bz_states_dielectric(mat, G, non_local: bool, nb: int, xyz, tet, q, energies, eta: float, nthreads: int = 8):
    """BZ_States eigenstates + compute_dielectric_function restated (dielectric_oracle.cpp):
    (eps, sums, total_volume) over the mesh (xyz in 1/m, tet nt x 4, q reduced)."""
    G = np.ascontiguousarray(G, dtype=np.int32)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    tet = np.ascontiguousarray(tet, dtype=np.int32).reshape(-1, 4)
    q = np.ascontiguousarray(q, dtype=np.float64).reshape(3)
    en = np.ascontiguousarray(energies, dtype=np.float64)
    eps, sums, vol = np.zeros(len(en)), np.zeros(len(en)), C.c_double()
    cm = mat.to_c()
    st = lib().orc_bz_states_dielectric(C.byref(cm), _i(G), len(G), int(non_local), nb, _d(xyz), len(xyz), _i(tet),
                                        len(tet), _d(q), _d(en), len(en), eta, _d(eps), _d(sums), C.byref(vol),
                                        nthreads)
    if st != 0:
        raise RuntimeError(f"orc_bz_states_dielectric failed ({st})")
    return eps, sums, vol.value
