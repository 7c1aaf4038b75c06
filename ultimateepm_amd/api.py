"""Python view of the C ABI (include/uepm.h) for tests, bench and scripts.

Everything here calls libuepm.so; nothing computes bands on the CPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from ._abi import GRADIENTS, MIXED, NONLOCAL, SOC, VECTORS, check, lib  # noqa: F401


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def basis_vectors(nearest_neighbors: int) -> np.ndarray:
    """BandStructure::GenerateBasisVectors (src/EPP/BandStructure.cpp:26-57)."""
    L = lib()
    cnt = C.c_int(0)
    check(L.uepm_basis_vectors(nearest_neighbors, None, 0, C.byref(cnt)))
    G = np.zeros((cnt.value, 3), dtype=np.int32)
    check(L.uepm_basis_vectors(nearest_neighbors, G.ctypes.data_as(C.POINTER(C.c_int32)), cnt.value, C.byref(cnt)))
    return G


def kpath(labels: str, n_points: int):
    """SymmetryPoints::GeneratePoints (src/EPP/SymmetryPoints.cpp:47-147)."""
    L = lib()
    out = np.zeros((max(n_points, 1), 3))
    pos = np.zeros(max(len(labels), 1), dtype=np.uint32)
    cnt = C.c_int(0)
    check(L.uepm_kpath(labels.encode(), n_points, _dp(out), pos.ctypes.data_as(C.POINTER(C.c_uint)), C.byref(cnt)))
    return out[: cnt.value].copy(), pos[: len(labels)].copy()


def bz_sample(seed: int, n: int) -> np.ndarray:
    """Seeded uniform first-BZ sample (DielectricFunction.cpp:35-57 test, splitmix64 draws)."""
    out = np.zeros((n, 3))
    check(lib().uepm_bz_sample(C.c_uint64(seed), n, _dp(out)))
    return out


class Context:
    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        check(lib().uepm_ctx_create(device, C.byref(self._h)))

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            lib().uepm_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def synchronize(self):
        check(lib().uepm_device_synchronize(self._h))

    def copy(self, dst: int, src: int, nbytes: int, stream: int = 0):
        """Device-to-device copy (both pointers 16-byte aligned), asynchronous on `stream`."""
        check(lib().uepm_memcpy_dtod(self._h, C.c_void_p(dst), C.c_void_p(src), nbytes, C.c_void_p(stream)))

    def stream(self) -> "Stream":
        return Stream(self)


class Stream:
    def __init__(self, ctx: Context):
        self.ctx = ctx
        self._h = C.c_void_p()
        check(lib().uepm_stream_create(ctx.handle, C.byref(self._h)))

    @property
    def handle(self) -> int:
        return self._h.value or 0

    def synchronize(self):
        check(lib().uepm_stream_synchronize(self.ctx.handle, self._h))

    def close(self):
        if self._h:
            lib().uepm_stream_destroy(self.ctx.handle, self._h)
            self._h = C.c_void_p()


class DeviceArray:
    """A device buffer with a numpy shape/dtype, owned by the library's allocator."""

    def __init__(self, ctx: Context, shape, dtype=np.float64):
        self.ctx = ctx
        self.shape = tuple(int(s) for s in np.atleast_1d(shape)) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        self._p = C.c_void_p()
        check(lib().uepm_device_alloc(ctx.handle, self.nbytes, C.byref(self._p)))

    @classmethod
    def from_host(cls, ctx: Context, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(ctx, a.shape, a.dtype)
        d.upload(a)
        return d

    @property
    def ptr(self) -> int:
        return self._p.value or 0

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        check(lib().uepm_memcpy_htod(self.ctx.handle, self._p, a.ctypes.data_as(C.c_void_p), self.nbytes))

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        check(lib().uepm_memcpy_dtoh(self.ctx.handle, out.ctypes.data_as(C.c_void_p), self._p, self.nbytes))
        return out

    def zero(self):
        check(lib().uepm_memset_device(self.ctx.handle, self._p, 0, self.nbytes))

    def free(self):
        if self._p:
            lib().uepm_device_free(self.ctx.handle, self._p)
            self._p = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Plan:
    """One material + basis + flags + band count (include/uepm.h: uepm_plan)."""

    def __init__(self, ctx: Context, material, G: np.ndarray, flags: int, nb: int):
        self._G = np.ascontiguousarray(G, dtype=np.int32)
        self._mat = material.to_c() if hasattr(material, "to_c") else material
        self._h = C.c_void_p()
        self.flags = flags
        self.nb = nb
        check(lib().uepm_plan_create(ctx.handle, C.byref(self._mat), self._G.ctypes.data_as(C.POINTER(C.c_int32)),
                                     len(self._G), flags, nb, C.byref(self._h)))
        self.dim = lib().uepm_plan_dim(self._h)

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            lib().uepm_plan_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def solve(self, k: np.ndarray, vectors: bool | None = None, gradients: bool | None = None):
        """Host-buffer solve: (evals[nk, nb], evecs[nk, nb, dim] | None, grads[nk, nb, 3] | None)."""
        k = np.ascontiguousarray(k, dtype=np.float64).reshape(-1, 3)
        nk = len(k)
        if vectors is None:
            vectors = bool(self.flags & VECTORS)
        if gradients is None:
            gradients = bool(self.flags & GRADIENTS)
        ev = np.zeros((nk, self.nb))
        vec = np.zeros((nk, self.nb, self.dim), dtype=np.complex128) if vectors else None
        gr = np.zeros((nk, self.nb, 3)) if gradients else None
        rc = lib().uepm_solve(self._h, _dp(k), nk, _dp(ev),
                              vec.ctypes.data_as(C.POINTER(C.c_double)) if vectors else None,
                              _dp(gr) if gradients else None)
        if rc == _abi.UEPM_ERR_NOT_CONVERGED:
            # the other k-points are solved: the error carries every output
            err = _abi.UepmError(rc, lib().uepm_last_error().decode())
            err.result = (ev, vec, gr)
            raise err
        check(rc)
        return ev, vec, gr

    def solve_device(self, k_ptr: int, nk: int, evals_ptr: int, evecs_ptr: int = 0, grads_ptr: int = 0,
                     stream: int = 0):
        check(lib().uepm_solve_device(self._h, C.c_void_p(k_ptr), nk, C.c_void_p(evals_ptr),
                                      C.c_void_p(evecs_ptr or None), C.c_void_p(grads_ptr or None),
                                      C.c_void_p(stream or None)))

    def status(self, stream: int = 0):
        """(failed k-points, first failed index or -1) of the last solve, after waiting for `stream`."""
        n, first = C.c_int64(0), C.c_int64(0)
        rc = lib().uepm_plan_status(self._h, C.c_void_p(stream or None), C.byref(n), C.byref(first))
        if rc not in (_abi.UEPM_OK, _abi.UEPM_ERR_NOT_CONVERGED):
            check(rc)
        return n.value, first.value

    def assemble_device(self, k_ptr: int, nk: int, H_ptr: int, stream: int = 0):
        check(lib().uepm_assemble_device(self._h, C.c_void_p(k_ptr), nk, C.c_void_p(H_ptr), C.c_void_p(stream or None)))

    def set_timing(self, on: bool):
        check(lib().uepm_plan_set_timing(self._h, int(on)))

    def config(self) -> dict:
        c = _abi.uepm_config()
        check(lib().uepm_plan_get_config(self._h, C.byref(c)))
        return {"tridiag": c.tridiag, "tri_block": c.tri_block, "bt_bands": c.bt_bands, "max_chunk": c.max_chunk,
                "lanes": c.lanes}

    def set_config(self, tridiag: int = 0, tri_block: int = 0, bt_bands: int = 0, max_chunk: int = 0, lanes: int = 0):
        c = _abi.uepm_config(tridiag, tri_block, bt_bands, max_chunk, lanes)
        check(lib().uepm_plan_set_config(self._h, C.byref(c)))

    def tune(self, k_ptr: int, nk: int, stream: int = 0) -> dict:
        """Autotuner: times the candidate configurations on device k-points, keeps the fastest."""
        c = _abi.uepm_config()
        check(lib().uepm_plan_tune(self._h, C.c_void_p(k_ptr), nk, C.c_void_p(stream), C.byref(c)))
        return {"tridiag": c.tridiag, "tri_block": c.tri_block, "bt_bands": c.bt_bands, "max_chunk": c.max_chunk,
                "lanes": c.lanes}

    def workspace_bytes(self, nk: int) -> int:
        b = C.c_uint64(0)
        check(lib().uepm_plan_workspace_bytes(self._h, nk, C.byref(b)))
        return b.value

    def reserve(self, nk: int):
        check(lib().uepm_plan_reserve(self._h, nk))

    def last_timing(self):
        """{kernel name: (ms, launches)} of the last solve_device call."""
        out = {}
        st = 0
        while True:
            ms = C.c_double()
            n = C.c_int64()
            name = C.c_char_p()
            r = lib().uepm_plan_last_timing(self._h, st, C.byref(ms), C.byref(n), C.byref(name))
            if r != 0:
                break
            out[name.value.decode()] = (ms.value, n.value)
            st += 1
        return out


class Batch:
    """One hipGraph over many (plan, k-set) items (uepm_batch_*)."""

    def __init__(self, ctx: Context, items, branches: int = 4):
        """items: iterable of (plan, k_ptr, nk, evals_ptr[, evecs_ptr[, grads_ptr]])."""
        arr = (_abi.uepm_batch_item * len(items))()
        for i, it in enumerate(items):
            plan, kp, nk, ep = it[:4]
            vp = it[4] if len(it) > 4 else 0
            gp = it[5] if len(it) > 5 else 0
            arr[i] = _abi.uepm_batch_item(plan.handle.value, kp, nk, ep, vp or None, gp or None)
        self._h = C.c_void_p()
        self._keep = list(items)
        check(lib().uepm_batch_create(ctx.handle, arr, len(items), branches, C.byref(self._h)))

    def launch(self, stream: int = 0):
        check(lib().uepm_batch_launch(self._h, C.c_void_p(stream)))

    def close(self):
        if self._h:
            lib().uepm_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching request server (uepm_server_*)."""

    def __init__(self, ctx: Context, hbm_budget_bytes: int, max_inflight: int = 8):
        self._h = C.c_void_p()
        check(lib().uepm_server_create(ctx.handle, hbm_budget_bytes, max_inflight, C.byref(self._h)))
        self._pending = {}

    def submit(self, material, nearest_neighbors: int, flags: int, nb: int, k: np.ndarray) -> int:
        k = np.ascontiguousarray(k, dtype=np.float64).reshape(-1, 3)
        cm = material.to_c() if hasattr(material, "to_c") else material
        req = _abi.uepm_request(C.pointer(cm), nearest_neighbors, flags, nb, _dp(k), len(k))
        t = C.c_uint64(0)
        check(lib().uepm_server_submit(self._h, C.byref(req), C.byref(t)))
        self._pending[t.value] = (len(k), nb)
        return t.value

    def wait(self, ticket: int):
        """(eigenvalues (nk, nb), latency_s, queued_s) of a submitted request."""
        nk, nb = self._pending.pop(ticket)
        ev = np.zeros((nk, nb))
        lat, que = C.c_double(0), C.c_double(0)
        rc = lib().uepm_server_wait(self._h, ticket, _dp(ev), C.byref(lat), C.byref(que))
        if rc == _abi.UEPM_ERR_NOT_CONVERGED:
            err = _abi.UepmError(rc, lib().uepm_last_error().decode())
            err.result = (ev, lat.value, que.value)
            raise err
        check(rc)
        return ev, lat.value, que.value

    def stats(self) -> dict:
        st = _abi.uepm_server_stats()
        check(lib().uepm_server_stats_get(self._h, C.byref(st)))
        return {f: getattr(st, f) for f, _ in st._fields_}

    def close(self):
        if self._h:
            lib().uepm_server_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def dielectric(plan: Plan, k: np.ndarray, q: np.ndarray, energies: np.ndarray, eta: float) -> np.ndarray:
    """DielectricFunction::compute_dielectric_function sums S[q][E] (uepm_dielectric)."""
    k = np.ascontiguousarray(k, dtype=np.float64).reshape(-1, 3)
    q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1, 3)
    en = np.ascontiguousarray(energies, dtype=np.float64)
    out = np.zeros((len(q), len(en)))
    check(lib().uepm_dielectric(plan.handle, _dp(k), len(k), _dp(q), len(q), _dp(en), len(en), eta, _dp(out)))
    return out


def dielectric_merge(q: np.ndarray, S: np.ndarray, total_nk: int) -> np.ndarray:
    """merge_results renormalisation: 1 + (2 pi / |q|^2) S / total_nk (uepm_dielectric_merge)."""
    q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1, 3)
    eps = np.array(S, dtype=np.float64, order="C")
    check(lib().uepm_dielectric_merge(_dp(q), len(q), eps.shape[1], total_nk, _dp(eps)))
    return eps


def kramers_kronig(eps_real: np.ndarray) -> np.ndarray:
    """apply_kramers_kronig: K (eps_real - 1) per q row (uepm_kramers_kronig)."""
    er = np.ascontiguousarray(eps_real, dtype=np.float64)
    out = np.zeros_like(er)
    check(lib().uepm_kramers_kronig(_dp(er), er.shape[0], er.shape[1], _dp(out)))
    return out
