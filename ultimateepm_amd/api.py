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
        check(lib().uepm_solve(self._h, _dp(k), nk, _dp(ev),
                               vec.ctypes.data_as(C.POINTER(C.c_double)) if vectors else None,
                               _dp(gr) if gradients else None))
        return ev, vec, gr

    def solve_device(self, k_ptr: int, nk: int, evals_ptr: int, evecs_ptr: int = 0, grads_ptr: int = 0,
                     stream: int = 0):
        check(lib().uepm_solve_device(self._h, C.c_void_p(k_ptr), nk, C.c_void_p(evals_ptr),
                                      C.c_void_p(evecs_ptr or None), C.c_void_p(grads_ptr or None),
                                      C.c_void_p(stream or None)))

    def assemble_device(self, k_ptr: int, nk: int, H_ptr: int, stream: int = 0):
        check(lib().uepm_assemble_device(self._h, C.c_void_p(k_ptr), nk, C.c_void_p(H_ptr), C.c_void_p(stream or None)))

    def set_timing(self, on: bool):
        check(lib().uepm_plan_set_timing(self._h, int(on)))

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
