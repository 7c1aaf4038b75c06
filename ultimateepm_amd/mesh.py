"""Brillouin-zone mesh driver over the C ABI (uepm_msh_* / uepm_bzmesh_*).

``MeshBZ`` keeps the method names of the reference's ``uepm::mesh_bz::MeshBZ``
(src/BZ_MESH/bz_mesh.hpp) for the band-solve side: reading the gmsh mesh,
the k-star map, bands over the mesh (BandsOnBZ), band views in and out and
the tetrahedron density of states (DOS_MeshBZ).  All logic runs in
libuepm.so; the band solve and the DOS run on the GPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import api
from ._abi import check, lib


def _d(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class MshFile:
    """A gmsh MSH 4.1 ASCII file: nodes, 4-node tetrahedra and NodeData views."""

    def __init__(self, path: str):
        self._h = C.c_void_p()
        check(lib().uepm_msh_read(path.encode(), C.byref(self._h)))
        nn, nt, nv = C.c_int64(), C.c_int64(), C.c_int32()
        check(lib().uepm_msh_counts(self._h, C.byref(nn), C.byref(nt), C.byref(nv)))
        self.tags = np.zeros(nn.value, dtype=np.uint64)
        self.xyz = np.zeros((nn.value, 3))
        check(lib().uepm_msh_nodes(self._h, self.tags.ctypes.data_as(C.POINTER(C.c_uint64)), _d(self.xyz)))
        self.tetrahedra = np.zeros((nt.value, 4), dtype=np.int32)
        if nt.value:
            check(lib().uepm_msh_tetrahedra(self._h, self.tetrahedra.ctypes.data_as(C.POINTER(C.c_int32))))
        self.views = []
        for i in range(nv.value):
            name = C.create_string_buffer(256)
            nc = C.c_int32()
            check(lib().uepm_msh_view_info(self._h, i, name, 256, C.byref(nc)))
            data = np.zeros((nn.value, nc.value))
            check(lib().uepm_msh_view_data(self._h, i, _d(data)))
            self.views.append((name.value.decode(), data if nc.value > 1 else data[:, 0]))
        lib().uepm_msh_free(self._h)
        self._h = None


def append_node_view(out_path: str, mesh_path: str | None, name: str, tags: np.ndarray, data: np.ndarray):
    data = np.ascontiguousarray(data, dtype=np.float64)
    ncomp = 1 if data.ndim == 1 else data.shape[1]
    tags = np.ascontiguousarray(tags, dtype=np.uint64)
    check(lib().uepm_msh_append_node_view(out_path.encode(), mesh_path.encode() if mesh_path else None, name.encode(),
                                          ncomp, tags.ctypes.data_as(C.POINTER(C.c_uint64)), len(tags), _d(data)))


def adjust_values(E: np.ndarray, soc: bool, min_cond_to_zero: bool = False):
    """BandStructure::AdjustValues over E (nk, nb): (shifted copy, band gap)."""
    E = np.array(E, dtype=np.float64, order="C")
    gap = C.c_double()
    check(lib().uepm_adjust_values(_d(E), E.shape[0], E.shape[1], int(soc), int(min_cond_to_zero), C.byref(gap)))
    return E, gap.value


class MeshBZ:
    """uepm::mesh_bz::MeshBZ on the band-solve side (uepm_bzmesh_*)."""

    def __init__(self, ctx: api.Context | None, handle: C.c_void_p):
        self.ctx = ctx
        self._h = handle

    # ---- construction ------------------------------------------------------
    @classmethod
    def read_mesh_geometry_from_msh_file(cls, ctx: api.Context | None, path: str, lattice_constant_m: float,
                                         normalize_by_fourier_factor: bool = True) -> "MeshBZ":
        """ctx None: a host-only mesh (no DOS, no device bands)."""
        h = C.c_void_p()
        check(lib().uepm_bzmesh_read_msh(ctx.handle if ctx else None, path.encode(), lattice_constant_m,
                                         int(normalize_by_fourier_factor), C.byref(h)))
        return cls(ctx, h)

    @classmethod
    def from_arrays(cls, ctx: api.Context | None, xyz: np.ndarray, tet: np.ndarray, in_wedge=None) -> "MeshBZ":
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        tet = np.ascontiguousarray(tet, dtype=np.int32)
        w = None if in_wedge is None else np.ascontiguousarray(in_wedge, dtype=np.uint8)
        h = C.c_void_p()
        check(lib().uepm_bzmesh_create(ctx.handle if ctx else None, _d(xyz), len(xyz), tet.ctypes.data_as(C.POINTER(C.c_int32)),
                                       len(tet), w.ctypes.data_as(C.POINTER(C.c_uint8)) if w is not None else None,
                                       C.byref(h)))
        return cls(ctx, h)

    def close(self):
        if self._h:
            lib().uepm_bzmesh_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- geometry ------------------------------------------------------------
    def counts(self):
        nv, nt, nb = C.c_int64(), C.c_int64(), C.c_int32()
        check(lib().uepm_bzmesh_counts(self._h, C.byref(nv), C.byref(nt), C.byref(nb)))
        return nv.value, nt.value, nb.value

    def geometry(self) -> dict:
        nv, nt, _ = self.counts()
        xyz = np.zeros((nv, 3))
        tet = np.zeros((nt, 4), dtype=np.int32)
        wedge = np.zeros(nt, dtype=np.uint8)
        tags = np.zeros(nv, dtype=np.uint64)
        vol, red = C.c_double(), C.c_double()
        check(lib().uepm_bzmesh_geometry(self._h, _d(xyz), tet.ctypes.data_as(C.POINTER(C.c_int32)),
                                         wedge.ctypes.data_as(C.POINTER(C.c_uint8)),
                                         tags.ctypes.data_as(C.POINTER(C.c_uint64)), C.byref(vol), C.byref(red)))
        return {"xyz": xyz, "tetrahedra": tet, "tetra_in_wedge": wedge, "tags": tags, "total_volume": vol.value,
                "reduce_bz_factor": red.value}

    # ---- irreducible wedge -------------------------------------------------
    def load_kstar_ibz_to_bz(self, path: str | None = None):
        check(lib().uepm_bzmesh_load_kstar(self._h, path.encode() if path else None))

    def write_kstar_ibz_to_bz(self, path: str):
        check(lib().uepm_bzmesh_write_kstar(self._h, path.encode()))

    def wedge_vertices(self) -> np.ndarray:
        n = C.c_int64()
        check(lib().uepm_bzmesh_wedge_vertices(self._h, None, C.byref(n)))
        ids = np.zeros(n.value, dtype=np.int64)
        check(lib().uepm_bzmesh_wedge_vertices(self._h, ids.ctypes.data_as(C.POINTER(C.c_int64)), C.byref(n)))
        return ids

    # ---- bands ---------------------------------------------------------------
    def kpoints(self, use_iwedge: bool = False) -> np.ndarray:
        n = C.c_int64()
        check(lib().uepm_bzmesh_kpoints(self._h, int(use_iwedge), None, C.byref(n)))
        k = np.zeros((n.value, 3))
        check(lib().uepm_bzmesh_kpoints(self._h, int(use_iwedge), _d(k), C.byref(n)))
        return k

    def compute_band_structure_over_mesh(self, plan: api.Plan, use_iwedge: bool = False) -> float:
        gap = C.c_double()
        check(lib().uepm_bzmesh_compute_bands(self._h, plan.handle, int(use_iwedge), C.byref(gap)))
        return gap.value

    def set_mesh_bands(self, E: np.ndarray, G: np.ndarray, use_iwedge: bool = False):
        E = np.ascontiguousarray(E, dtype=np.float64)
        G = np.ascontiguousarray(G, dtype=np.float64)
        check(lib().uepm_bzmesh_set_mesh_bands(self._h, int(use_iwedge), _d(E), _d(G), E.shape[1]))

    def set_bands(self, E: np.ndarray, G: np.ndarray):
        E = np.ascontiguousarray(E, dtype=np.float64)
        G = np.ascontiguousarray(G, dtype=np.float64)
        check(lib().uepm_bzmesh_set_bands(self._h, _d(E), _d(G), E.shape[1]))

    def set_bands_device(self, e_ptr: int, g_ptr: int, ld: int, first: int, nb: int, stream: int = 0):
        check(lib().uepm_bzmesh_set_bands_device(self._h, C.c_void_p(e_ptr), C.c_void_p(g_ptr), ld, first, nb,
                                                 C.c_void_p(stream or None)))

    def bands(self):
        """(E (nv, nb), G (nv, nb, 3)) at the vertices."""
        nb = C.c_int32()
        check(lib().uepm_bzmesh_bands(self._h, None, None, C.byref(nb)))
        nv = self.counts()[0]
        E = np.zeros((nv, nb.value))
        G = np.zeros((nv, nb.value, 3))
        check(lib().uepm_bzmesh_bands(self._h, _d(E), _d(G), C.byref(nb)))
        return E, G

    def band_info(self, band: int) -> dict:
        ic, li, mn, mx = C.c_int32(), C.c_int64(), C.c_double(), C.c_double()
        check(lib().uepm_bzmesh_band_info(self._h, band, C.byref(ic), C.byref(li), C.byref(mn), C.byref(mx)))
        return {"type": "conduction" if ic.value else "valence", "local_index": li.value, "min": mn.value,
                "max": mx.value}

    def band_groups(self) -> dict:
        a, b, c, d = (C.c_int64() for _ in range(4))
        check(lib().uepm_bzmesh_band_groups(self._h, C.byref(a), C.byref(b), C.byref(c), C.byref(d)))
        return {"valence": (a.value, b.value), "conduction": (c.value, d.value)}

    def get_number_bands_total(self) -> int:
        return self.counts()[2]

    # ---- density of states -------------------------------------------------
    def dos(self, band: int, n_points: int, e_min: float, e_max: float, reduce_bz_factor: float = 1.0,
            spin_degeneracy: float = 2.0, wedge_only: bool = False):
        en, dos = np.zeros(n_points), np.zeros(n_points)
        check(lib().uepm_bzmesh_dos(self._h, band, n_points, e_min, e_max, reduce_bz_factor, spin_degeneracy,
                                    int(wedge_only), _d(en), _d(dos)))
        return en, dos

    def dos_device(self, band: int, n_points: int, e_min: float, e_max: float, reduce_bz_factor: float,
                   spin_degeneracy: float, wedge_only: bool, energies_ptr: int, dos_ptr: int, stream: int = 0):
        check(lib().uepm_bzmesh_dos_device(self._h, band, n_points, e_min, e_max, reduce_bz_factor, spin_degeneracy,
                                           int(wedge_only), C.c_void_p(energies_ptr or None), C.c_void_p(dos_ptr),
                                           C.c_void_p(stream or None)))

    def compute_dos_band_at_band_auto(self, band: int, nb_points: int, use_interp: bool = False,
                                      use_iw: bool = False):
        en, dos = np.zeros(nb_points), np.zeros(nb_points)
        check(lib().uepm_bzmesh_dos_auto(self._h, band, nb_points, int(use_interp), int(use_iw), _d(en), _d(dos)))
        return en, dos

    # ---- files -----------------------------------------------------------------
    def export_selected_bands_to_gmsh(self, out_path: str, nb_valence: int, nb_conduction: int,
                                      highest_valence_as_band0: bool = True, model_msh_path: str | None = None,
                                      write_gradients: bool = True):
        check(lib().uepm_bzmesh_export_msh(self._h, out_path.encode(), nb_valence, nb_conduction,
                                           int(highest_valence_as_band0),
                                           model_msh_path.encode() if model_msh_path else None, int(write_gradients)))

    def export_energies_and_gradients_to_vtk(self, path: str):
        check(lib().uepm_bzmesh_export_vtk(self._h, path.encode()))

    def read_mesh_bands_from_msh_file(self, path: str, nb_conduction_bands: int = -1, nb_valence_bands: int = -1,
                                      auto_shift_conduction_band: bool = False,
                                      set_positive_valence_band: bool = False):
        check(lib().uepm_bzmesh_read_bands_msh(self._h, path.encode(), nb_conduction_bands, nb_valence_bands,
                                               int(auto_shift_conduction_band), int(set_positive_valence_band)))

    def apply_scissor(self, value: float):
        check(lib().uepm_bzmesh_apply_scissor(self._h, value))


class BZStates(MeshBZ):
    """uepm::mesh_bz::BZ_States (src/BZ_MESH/bz_states.hpp): eigenstates at the
    mesh vertices and at vertex + q, kept on the device, and the tetrahedron-
    weighted dielectric function (bz_states.cpp:95-164).  The plan passed to
    the eigenstate calls plays the reference's Hamiltonian(material, basis)
    with SetMatrix(k, false): UEPM_VECTORS, no SOC (local as in the reference
    unless the plan says otherwise), nb = set_nb_bands."""

    def __init__(self, ctx: api.Context | None, handle: C.c_void_p):
        super().__init__(ctx, handle)
        self.list_energies = np.zeros(0)
        self.dielectric_function_real = np.zeros(0)
        self.dielectric_sums = np.zeros(0)
        self.total_volume_integrated = 0.0

    def compute_eigenstates(self, plan: api.Plan):
        check(lib().uepm_bzmesh_compute_states(self._h, plan.handle, None))

    def compute_shifted_eigenstates(self, q_shift, plan: api.Plan):
        q = np.ascontiguousarray(q_shift, dtype=np.float64).reshape(3)
        check(lib().uepm_bzmesh_compute_states(self._h, plan.handle, _d(q)))

    def eigen_states(self, shifted: bool = False, vectors: bool = True):
        """(E nv x nb, X nv x nb x dim complex or None) of the k (or k + q) set."""
        nb, dim = C.c_int32(), C.c_int32()
        check(lib().uepm_bzmesh_states(self._h, int(shifted), None, None, C.byref(nb), C.byref(dim)))
        nv = self.counts()[0]
        E = np.zeros((nv, nb.value))
        X = np.zeros((nv, nb.value, dim.value), dtype=np.complex128) if vectors else None
        check(lib().uepm_bzmesh_states(self._h, int(shifted), _d(E),
                                       X.ctypes.data_as(C.POINTER(C.c_double)) if vectors else None, None, None))
        return E, X

    def compute_dielectric_function(self, energies, eta_smearing: float) -> np.ndarray:
        en = np.ascontiguousarray(energies, dtype=np.float64)
        eps, sums, vol = np.zeros(len(en)), np.zeros(len(en)), C.c_double()
        check(lib().uepm_bzmesh_states_dielectric(self._h, _d(en), len(en), eta_smearing, _d(eps), _d(sums),
                                                  C.byref(vol)))
        self.list_energies, self.dielectric_function_real, self.dielectric_sums = en, eps, sums
        self.total_volume_integrated = vol.value
        return eps

    def export_dielectric_function(self, prefix: str):
        check(lib().uepm_write_dielectric_csv(prefix.encode(), _d(self.list_energies),
                                              _d(self.dielectric_function_real), len(self.list_energies)))

    def export_full_eigenstates(self, root: str | None = None):
        check(lib().uepm_bzmesh_export_states(self._h, root.encode() if root else None))

    def compute_fermi_level(self, doping_concentration: float, temperature: float) -> float:
        return 0.0  # a stub returning 0 in the reference (bz_states.cpp:166-169)

    def get_dielectric_function(self, q, energy: float) -> complex:
        return 1.0 + 0.0j  # a stub returning 1 in the reference (bz_states.hpp:87-90)
