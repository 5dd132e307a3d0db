"""ctypes face of the CPU oracle -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker.  Arrays are numpy arrays of shape
(7, NZ, NY, NX) in the reference's SoA order (source/Types.h:31, NodeBlock.h:47).
"""
from __future__ import annotations

import ctypes as C
import os
import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

NVAR = 7
ABSORBING, REFLECTING = 0, 1


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(path)
        for suf, T in (("f", C.c_float), ("d", C.c_double)):
            P = C.c_void_p
            getattr(L, f"orc_create_{suf}").restype = P
            getattr(L, f"orc_create_{suf}").argtypes = [C.c_int] * 3
            getattr(L, f"orc_destroy_{suf}").argtypes = [P]
            getattr(L, f"orc_set_{suf}").argtypes = [P, C.c_int, P]
            getattr(L, f"orc_get_{suf}").argtypes = [P, C.c_int, P]
            getattr(L, f"orc_apply_bc_{suf}").argtypes = [P, P]
            getattr(L, f"orc_get_halo_{suf}").argtypes = [P, C.c_int, P]
            getattr(L, f"orc_set_halo_{suf}").argtypes = [P, C.c_int, P]
            getattr(L, f"orc_rhs_{suf}").argtypes = [P, T, T]
            getattr(L, f"orc_update_{suf}").argtypes = [P, T]
            getattr(L, f"orc_max_sos_{suf}").argtypes = [P, C.c_int]
            getattr(L, f"orc_max_sos_{suf}").restype = T
            getattr(L, f"orc_step_{suf}").argtypes = [P, C.c_double, C.c_double, C.c_double, P, C.c_int, P]
            getattr(L, f"orc_step_{suf}").restype = C.c_double
            getattr(L, f"orc_faces_{suf}").argtypes = [P, C.c_int, P]
            getattr(L, f"orc_face_batch_{suf}").argtypes = [C.c_int, C.c_int, P, P]
            getattr(L, f"orc_weno_{suf}").argtypes = [C.c_int, P, P, P]
            getattr(L, f"orc_weno_raw_{suf}").argtypes = [C.c_int, P, P, P]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_get_threads.restype = C.c_int
        _LIB = L
    return _LIB


def ref_lib():
    """oracle/_ref/libref.so: the reference's own MaxSpeedOfSound.cpp and
    BoundaryConditions.h compiled at a fixed block size (None if absent)."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libref.so")
        if not os.path.exists(path):
            return None
        L = C.CDLL(path)
        L.ref_sizes.argtypes = [C.c_void_p]
        L.ref_max_sos_cpp.argtypes = [C.c_void_p]
        L.ref_max_sos_cpp.restype = C.c_float
        L.ref_apply_bc.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p]
        _REF = L
    return _REF


def _ptr(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


def set_threads(n: int):
    lib().orc_set_threads(int(n))


def threads() -> int:
    return lib().orc_get_threads()


class Block:
    """Reference-layout block: data (u), tmp (RK register R) and six halos."""

    def __init__(self, nx, ny, nz, dtype=np.float32):
        self.shape = (nz, ny, nx)
        self.nx, self.ny, self.nz = nx, ny, nz
        self.dtype = np.dtype(dtype)
        self.suf = "f" if self.dtype == np.float32 else "d"
        self.T = C.c_float if self.suf == "f" else C.c_double
        self._h = getattr(lib(), f"orc_create_{self.suf}")(nx, ny, nz)

    def _f(self, name):
        return getattr(lib(), f"orc_{name}_{self.suf}")

    def __del__(self):
        if getattr(self, "_h", None):
            self._f("destroy")(self._h)
            self._h = None

    def set(self, which: int, arr: np.ndarray):
        a = np.ascontiguousarray(arr, dtype=self.dtype)
        assert a.shape == (NVAR,) + self.shape, a.shape
        self._f("set")(self._h, which, _ptr(a))

    def get(self, which: int) -> np.ndarray:
        out = np.empty((NVAR,) + self.shape, dtype=self.dtype)
        self._f("get")(self._h, which, _ptr(out))
        return out

    @property
    def data(self):
        return self.get(0)

    @data.setter
    def data(self, a):
        self.set(0, a)

    @property
    def tmp(self):
        return self.get(1)

    @tmp.setter
    def tmp(self, a):
        self.set(1, a)

    def halo_size(self, face):
        d = face // 2
        return (3 * self.ny * self.nz, self.nx * 3 * self.nz, self.nx * self.ny * 3)[d]

    def apply_bc(self, bc):
        b = np.ascontiguousarray(bc, dtype=np.int32)
        self._f("apply_bc")(self._h, _ptr(b))

    def get_halo(self, face) -> np.ndarray:
        out = np.empty((NVAR, self.halo_size(face)), dtype=self.dtype)
        self._f("get_halo")(self._h, face, _ptr(out))
        return out

    def set_halo(self, face, arr):
        a = np.ascontiguousarray(arr, dtype=self.dtype).reshape(NVAR, self.halo_size(face))
        self._f("set_halo")(self._h, face, _ptr(a))

    def rhs(self, a, dtinvh):
        self._f("rhs")(self._h, self.T(a), self.T(dtinvh))

    def update(self, b):
        self._f("update")(self._h, self.T(b))

    def max_sos(self, cpu_form=False):
        return self._f("max_sos")(self._h, int(cpu_form))

    def step(self, h, cfl, dt_max, bc, cpu_sos=False):
        b = np.ascontiguousarray(bc, dtype=np.int32)
        sos = C.c_float(0)
        dt = self._f("step")(self._h, h, cfl, dt_max, _ptr(b), int(cpu_sos), C.byref(sos))
        return dt, sos.value

    def faces(self, d):
        """12 face arrays for direction d: 7 fluxes, vel, Gm, Gp, Pm, Pp."""
        fz, fy, fx = self.nz + (d == 2), self.ny + (d == 1), self.nx + (d == 0)
        out = np.empty((12, fz, fy, fx), dtype=self.dtype)
        self._f("faces")(self._h, d, _ptr(out))
        return out


def face_batch(st: np.ndarray, d: int, dtype=np.float32) -> np.ndarray:
    """st: (n, 7, 6) conserved values at stencil points i-3..i+2 -> (n, 12)."""
    st = np.ascontiguousarray(st, dtype=dtype)
    n = st.shape[0]
    out = np.empty((n, 12), dtype=dtype)
    suf = "f" if np.dtype(dtype) == np.float32 else "d"
    getattr(lib(), f"orc_face_batch_{suf}")(n, d, _ptr(st), _ptr(out))
    return out


def weno(in5: np.ndarray, clipped=True, dtype=np.float32):
    """in5: (n, 5) cell values j-2..j+2 -> (minus-state at j+1/2, plus-state at j-1/2)."""
    a = np.ascontiguousarray(in5, dtype=dtype)
    n = a.shape[0]
    m = np.empty(n, dtype=dtype)
    p = np.empty(n, dtype=dtype)
    suf = "f" if np.dtype(dtype) == np.float32 else "d"
    name = f"orc_weno_{suf}" if clipped else f"orc_weno_raw_{suf}"
    getattr(lib(), name)(n, _ptr(a), _ptr(m), _ptr(p))
    return m, p
