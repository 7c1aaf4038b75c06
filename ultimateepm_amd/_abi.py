"""ctypes view of the C ABI declared in include/uepm.h.

The shared library ``libuepm.so`` (HIP kernels + host C++) is built in-tree
by ``ultimateepm_amd.build``.  Loading fails loudly when it is missing: there
is no CPU fallback for the band solve.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libuepm.so")

UEPM_OK = 0
UEPM_ERR_INVALID_ARG = 1
UEPM_ERR_RUNTIME = 2
UEPM_ERR_OUT_OF_RANGE = 3
UEPM_ERR_DEVICE = 4
UEPM_ERR_NOT_CONVERGED = 5
UEPM_ERR_NO_MEMORY = 6

NONLOCAL = 0x1
SOC = 0x2
VECTORS = 0x4
GRADIENTS = 0x8
MIXED = 0x10


class uepm_material(C.Structure):
    """Mirror of ``uepm_material`` (include/uepm.h)."""

    _fields_ = [
        ("lattice_constant_angstrom", C.c_double),
        ("V_S", C.c_double * 4),
        ("V_A", C.c_double * 4),
        ("nl_populated", C.c_int32),
        ("nl_well_type", C.c_int32),
        ("alpha_0_cation", C.c_double),
        ("beta_0_cation", C.c_double),
        ("A2_cation", C.c_double),
        ("alpha_0_anion", C.c_double),
        ("beta_0_anion", C.c_double),
        ("A2_anion", C.c_double),
        ("R0_cation", C.c_double),
        ("R2_cation", C.c_double),
        ("R0_anion", C.c_double),
        ("R2_anion", C.c_double),
        ("soc_populated", C.c_int32),
        ("period_anion", C.c_int32),
        ("period_cation", C.c_int32),
        ("radial_extent_anion", C.c_double),
        ("radial_extent_cation", C.c_double),
        ("alpha_soc", C.c_double),
        ("mu_soc", C.c_double),
    ]


class uepm_config(C.Structure):
    """Mirror of ``uepm_config``: 0 fields mean automatic."""

    _fields_ = [("tridiag", C.c_int32), ("tri_block", C.c_int32), ("bt_bands", C.c_int32), ("max_chunk", C.c_int32),
                ("lanes", C.c_int32)]


class uepm_batch_item(C.Structure):
    _fields_ = [("plan", C.c_void_p), ("d_k", C.c_void_p), ("nk", C.c_int64), ("d_evals", C.c_void_p),
                ("d_evecs", C.c_void_p), ("d_grads", C.c_void_p)]


class uepm_request(C.Structure):
    _fields_ = [("material", C.POINTER(uepm_material)), ("nearest_neighbors", C.c_int32), ("flags", C.c_uint32),
                ("nb", C.c_int32), ("k", C.POINTER(C.c_double)), ("nk", C.c_int64)]


class uepm_server_stats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("kpoints", C.c_uint64),
                ("evictions", C.c_uint64), ("used_bytes", C.c_uint64), ("inflight", C.c_int32),
                ("queued", C.c_int32)]


class UepmError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"uepm error {code}: {msg}")
        self.code = code


_lib = None


def lib() -> C.CDLL:
    """Load libuepm.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "(the band solve has no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    P, I32, I64, D = C.c_void_p, C.c_int32, C.c_int64, C.POINTER(C.c_double)
    L.uepm_last_error.restype = C.c_char_p
    L.uepm_ctx_create.argtypes = [C.c_int, C.POINTER(P)]
    L.uepm_ctx_destroy.argtypes = [P]
    L.uepm_basis_vectors.argtypes = [C.c_int, C.POINTER(I32), C.c_int, C.POINTER(C.c_int)]
    L.uepm_kpath.argtypes = [C.c_char_p, C.c_uint, C.POINTER(C.c_double), C.POINTER(C.c_uint), C.POINTER(C.c_int)]
    L.uepm_bz_sample.argtypes = [C.c_uint64, I64, C.POINTER(C.c_double)]
    L.uepm_plan_create.argtypes = [P, C.POINTER(uepm_material), C.POINTER(I32), C.c_int, C.c_uint, C.c_int, C.POINTER(P)]
    L.uepm_plan_destroy.argtypes = [P]
    L.uepm_plan_dim.argtypes = [P]
    L.uepm_solve.argtypes = [P, D, I64, D, D, D]
    L.uepm_solve_device.argtypes = [P, P, I64, P, P, P, P]
    L.uepm_plan_status.argtypes = [P, P, C.POINTER(I64), C.POINTER(I64)]
    L.uepm_solve_device_noalloc.argtypes = [P, P, I64, P, P, P, P]
    L.uepm_assemble_device.argtypes = [P, P, I64, P, P]
    L.uepm_plan_set_timing.argtypes = [P, C.c_int]
    L.uepm_plan_last_timing.argtypes = [P, C.c_int, D, C.POINTER(I64), C.POINTER(C.c_char_p)]
    U64 = C.c_uint64
    L.uepm_device_alloc.argtypes = [P, U64, C.POINTER(P)]
    L.uepm_device_free.argtypes = [P, P]
    L.uepm_memcpy_htod.argtypes = [P, P, P, U64]
    L.uepm_memcpy_dtoh.argtypes = [P, P, P, U64]
    L.uepm_memset_device.argtypes = [P, P, C.c_int, U64]
    L.uepm_memcpy_dtod.argtypes = [P, P, P, U64, P]
    L.uepm_stream_create.argtypes = [P, C.POINTER(P)]
    L.uepm_stream_destroy.argtypes = [P, P]
    L.uepm_stream_synchronize.argtypes = [P, P]
    L.uepm_device_synchronize.argtypes = [P]
    L.uepm_plan_get_config.argtypes = [P, C.POINTER(uepm_config)]
    L.uepm_plan_set_config.argtypes = [P, C.POINTER(uepm_config)]
    L.uepm_plan_tune.argtypes = [P, P, I64, P, C.POINTER(uepm_config)]
    L.uepm_plan_workspace_bytes.argtypes = [P, I64, C.POINTER(U64)]
    L.uepm_plan_reserve.argtypes = [P, I64]
    L.uepm_batch_create.argtypes = [P, C.POINTER(uepm_batch_item), C.c_int, C.c_int, C.POINTER(P)]
    L.uepm_batch_launch.argtypes = [P, P]
    L.uepm_batch_destroy.argtypes = [P]
    L.uepm_server_create.argtypes = [P, U64, C.c_int, C.POINTER(P)]
    L.uepm_server_submit.argtypes = [P, C.POINTER(uepm_request), C.POINTER(U64)]
    L.uepm_server_wait.argtypes = [P, U64, D, D, D]
    L.uepm_server_stats_get.argtypes = [P, C.POINTER(uepm_server_stats)]
    L.uepm_server_destroy.argtypes = [P]
    CP, U8, PI64, PI32 = C.c_char_p, C.POINTER(C.c_uint8), C.POINTER(I64), C.POINTER(I32)
    PU64 = C.POINTER(U64)
    L.uepm_msh_read.argtypes = [CP, C.POINTER(P)]
    L.uepm_msh_counts.argtypes = [P, PI64, PI64, PI32]
    L.uepm_msh_nodes.argtypes = [P, PU64, D]
    L.uepm_msh_tetrahedra.argtypes = [P, PI32]
    L.uepm_msh_view_info.argtypes = [P, I32, C.c_char_p, I32, PI32]
    L.uepm_msh_view_data.argtypes = [P, I32, D]
    L.uepm_msh_free.argtypes = [P]
    L.uepm_msh_free.restype = None
    L.uepm_msh_append_node_view.argtypes = [CP, CP, CP, I32, PU64, I64, D]
    L.uepm_bzmesh_create.argtypes = [P, D, I64, PI32, I64, U8, C.POINTER(P)]
    L.uepm_bzmesh_read_msh.argtypes = [P, CP, C.c_double, C.c_int, C.POINTER(P)]
    L.uepm_bzmesh_destroy.argtypes = [P]
    L.uepm_bzmesh_counts.argtypes = [P, PI64, PI64, PI32]
    L.uepm_bzmesh_geometry.argtypes = [P, D, PI32, U8, PU64, D, D]
    L.uepm_bzmesh_load_kstar.argtypes = [P, CP]
    L.uepm_bzmesh_write_kstar.argtypes = [P, CP]
    L.uepm_bzmesh_wedge_vertices.argtypes = [P, PI64, PI64]
    L.uepm_bzmesh_kpoints.argtypes = [P, C.c_int, D, PI64]
    L.uepm_bzmesh_set_mesh_bands.argtypes = [P, C.c_int, D, D, I32]
    L.uepm_bzmesh_compute_bands.argtypes = [P, P, C.c_int, D]
    L.uepm_bzmesh_bands.argtypes = [P, D, D, PI32]
    L.uepm_bzmesh_band_info.argtypes = [P, I32, PI32, PI64, D, D]
    L.uepm_bzmesh_band_groups.argtypes = [P, PI64, PI64, PI64, PI64]
    L.uepm_bzmesh_set_bands.argtypes = [P, D, D, I32]
    L.uepm_bzmesh_set_bands_device.argtypes = [P, P, P, I32, I32, I32, P]
    L.uepm_bzmesh_dos.argtypes = [P, I32, I32, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, D, D]
    L.uepm_bzmesh_dos_device.argtypes = [P, I32, I32, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, P,
                                         P, P]
    L.uepm_bzmesh_dos_auto.argtypes = [P, I32, I32, C.c_int, C.c_int, D, D]
    L.uepm_bzmesh_export_msh.argtypes = [P, CP, I64, I64, C.c_int, CP, C.c_int]
    L.uepm_bzmesh_export_vtk.argtypes = [P, CP]
    L.uepm_bzmesh_read_bands_msh.argtypes = [P, CP, I32, I32, C.c_int, C.c_int]
    L.uepm_bzmesh_apply_scissor.argtypes = [P, C.c_double]
    L.uepm_adjust_values.argtypes = [D, I64, I32, C.c_int, C.c_int, D]
    L.uepm_bzmesh_compute_states.argtypes = [P, P, D]
    L.uepm_bzmesh_states.argtypes = [P, C.c_int, D, D, PI32, PI32]
    L.uepm_bzmesh_states_dielectric.argtypes = [P, D, I32, C.c_double, D, D, D]
    L.uepm_write_dielectric_csv.argtypes = [CP, D, D, I32]
    L.uepm_bzmesh_export_states.argtypes = [P, CP]
    L.uepm_dielectric.argtypes = [P, D, I64, D, I64, D, I32, C.c_double, D]
    L.uepm_dielectric_merge.argtypes = [D, I64, I32, I64, D]
    L.uepm_kramers_kronig.argtypes = [D, I64, I32, D]
    _lib = L
    return L


def check(code: int) -> None:
    if code != UEPM_OK:
        raise UepmError(code, lib().uepm_last_error().decode())
