"""ctypes binding of libmpcf_hip.so (include/mpcf_hip.h).

Host arrays are numpy float32 arrays of shape (7, NZ, NY, NX): the reference's
seven SoA fields (source/NodeBlock.h:26-29,47-48) in ID3 order (Types.h:31).

The library must be built (cubism-jr-gpu-mpcf_amd/build.py); loading fails
loudly if it is missing -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmpcf_hip.so")

OK, EINVAL, EHIP, ENOMEM, ENUMERIC, EBUSY = 0, 1, 2, 3, 4, 5
BC_ABSORBING, BC_REFLECTING, BC_HALO = 0, 1, 2
FLAG_FAST, FLAG_SOS_CPU, FLAG_SPLIT_XY, FLAG_FUSED_XY, FLAG_P64, FLAG_FP64 = 1, 2, 4, 8, 16, 32
FLAG_NO_PIPELINE = 128
FLAG_E_PAIR = 256
DATA, TMP = 0, 1
BATCH_GROUPS, BATCH_PERSISTENT = 0, 1
NVAR = 7

# LSRK3 coefficients of LSRK3_IntegratorMPI.cpp:46,51,56
LSRK3_A = (0.0, -17.0 / 32, -32.0 / 27)
LSRK3_B = (1.0 / 4, 8.0 / 9, 3.0 / 4)

EXPORTS = [
    "mpcf_last_error", "mpcf_version", "mpcf_create", "mpcf_destroy", "mpcf_set_pipeline", "mpcf_set_stream",
    "mpcf_get_stream", "mpcf_set_bc", "mpcf_upload", "mpcf_download", "mpcf_load_ghosts",
    "mpcf_max_sos", "mpcf_convection", "mpcf_update", "mpcf_process_all", "mpcf_step",
    "mpcf_last_dt", "mpcf_synchronize", "mpcf_query", "mpcf_test_weno", "mpcf_test_face",
    "mpcf_timing_enable", "mpcf_timing_read", "mpcf_ic_shock_tube", "mpcf_ic_shock_bubble",
    "mpcf_ic_bubble_cloud", "mpcf_ic_bubble_collapse", "mpcf_flop_count", "mpcf_host_weno",
    "mpcf_host_face", "mpcf_case_bytes", "mpcf_batch_create", "mpcf_batch_destroy", "mpcf_batch_add",
    "mpcf_batch_remove", "mpcf_batch_download", "mpcf_batch_step", "mpcf_batch_use_graph",
    "mpcf_batch_last_dt", "mpcf_batch_synchronize", "mpcf_batch_get_stream", "mpcf_batch_set_march",
    "mpcf_batch_autotune", "mpcf_batch_query", "mpcf_batch_set_mode", "mpcf_batch_tuning", "mpcf_upload_f64", "mpcf_download_f64",
    "mpcf_last_sos64", "mpcf_compare", "mpcf_integrals", "mpcf_server_create", "mpcf_server_submit",
    "mpcf_server_poll", "mpcf_server_wait", "mpcf_server_stats", "mpcf_server_clock", "mpcf_server_destroy",
    "mpcf_sic_defaults", "mpcf_sic_post_shock", "mpcf_read_cloud_config", "mpcf_read_cloud", "mpcf_ic_sic_cloud",
    "mpcf_stream_channels", "mpcf_stream_out", "mpcf_stream_in", "mpcf_write_xdmf", "mpcf_write_save_info",
    "mpcf_read_save_info", "mpcf_dump", "mpcf_save", "mpcf_restart", "mpcf_post_shock_ratio",
    "mpcf_sod_defaults", "mpcf_ic_sod", "mpcf_sbi_defaults", "mpcf_sbi_config", "mpcf_read_bubbles",
    "mpcf_ic_sbi", "mpcf_ic_static", "mpcf_ic_steady", "mpcf_halo_count", "mpcf_halo_export",
    "mpcf_halo_import", "mpcf_ghosts_read", "mpcf_steps", "mpcf_wavelet_dump", "mpcf_wavelet_read", "mpcf_wavelet_encode", "mpcf_wavelet_decode",
]

# streamer kinds of the dumps (mpcf_stream_out)
STREAM_TENSOR, STREAM_SAVE, STREAM_SCALAR = 0, 1, 2


class SicParams(C.Structure):
    """mpcf_sic_params: the shock-induced cloud collapse parameters
    (Sim_SICCloudMPI.cpp:53-77 for the defaults and option names)."""
    _fields_ = [("rho0", C.c_double), ("u0", C.c_double), ("p0", C.c_double),
                ("rhoB", C.c_double), ("uB", C.c_double), ("pB", C.c_double),
                ("pressure_ratio", C.c_double), ("shock_n", C.c_double * 3), ("shock_s", C.c_double * 3),
                ("g1", C.c_double), ("g2", C.c_double), ("pc1", C.c_double), ("pc2", C.c_double),
                ("eps", C.c_double)]


class SodParams(C.Structure):
    """mpcf_sod_params: the Sod tube of Sim_SodMPI (Sim_SodMPI.cpp:27-56)."""
    _fields_ = [("x0", C.c_double), ("rho1", C.c_double), ("u1", C.c_double), ("p1", C.c_double),
                ("rho2", C.c_double), ("u2", C.c_double), ("p2", C.c_double),
                ("g1", C.c_double), ("g2", C.c_double), ("pc1", C.c_double), ("pc2", C.c_double),
                ("reverse", C.c_int), ("dir", C.c_int)]


class SbiParams(C.Structure):
    """mpcf_sbi_params: the 2D shock-bubble case of Sim_2DSBIMPI
    (Sim_2DSBIMPI.cpp:20-63, 95-108)."""
    _fields_ = [("x0", C.c_double), ("mach", C.c_double), ("rho", C.c_double), ("u", C.c_double),
                ("p", C.c_double), ("rhobubble", C.c_double), ("pbubble", C.c_double),
                ("g1", C.c_double), ("g2", C.c_double), ("pc1", C.c_double), ("pc2", C.c_double),
                ("mollfactor", C.c_int), ("reverse", C.c_int), ("dims", C.c_int * 3)]


# serving policies and the eager-launch flag (mpcf_server_create)
SERVE_ALL, SERVE_FIFO_FILL, SERVE_SRPT_FILL, SERVE_HYBRID, SERVE_FIFO_SLACK = 0, 1, 2, 3, 4
FLAG_SERVE_EAGER = 64

# kernel classes of mpcf_timing_read
KERNEL_CLASSES = ("ghosts", "max_sos", "dt", "sweep_x", "sweep_y", "sweep_z", "update", "other")
# with the fused x+y sweep, "sweep_x" times that kernel and "sweep_y" is unused

_LIB = None


class MpcfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mpcf error {code}: {msg}")
        self.code = code


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run cubism-jr-gpu-mpcf_amd/build.py")
        L = C.CDLL(LIB_PATH)
        P, I, F, D = C.c_void_p, C.c_int, C.c_float, C.c_double
        L.mpcf_last_error.restype = C.c_char_p
        L.mpcf_version.restype = C.c_char_p
        L.mpcf_create.argtypes = [C.POINTER(P), I, I, I, D, C.c_uint]
        L.mpcf_destroy.argtypes = [P]
        L.mpcf_set_pipeline.argtypes = [P, I]
        L.mpcf_set_stream.argtypes = [P, P]
        L.mpcf_get_stream.argtypes = [P]
        L.mpcf_get_stream.restype = P
        L.mpcf_set_bc.argtypes = [P, P]
        L.mpcf_upload.argtypes = [P, I, P]
        L.mpcf_download.argtypes = [P, I, P]
        L.mpcf_load_ghosts.argtypes = [P]
        L.mpcf_max_sos.argtypes = [P, C.POINTER(F)]
        L.mpcf_convection.argtypes = [P, F, F]
        L.mpcf_update.argtypes = [P, F]
        L.mpcf_process_all.argtypes = [P, F, F, F]
        L.mpcf_step.argtypes = [P, D, D, P]
        L.mpcf_last_dt.argtypes = [P, C.POINTER(D), C.POINTER(F)]
        L.mpcf_synchronize.argtypes = [P]
        L.mpcf_query.argtypes = [P, P]
        L.mpcf_test_weno.argtypes = [I, P, P, P, C.c_uint]
        L.mpcf_test_face.argtypes = [I, I, P, P, C.c_uint]
        L.mpcf_flop_count.argtypes = [C.c_uint, P]
        L.mpcf_host_weno.argtypes = [I, P, P, P, C.c_uint]
        L.mpcf_host_face.argtypes = [I, I, P, P, C.c_uint]
        L.mpcf_timing_enable.argtypes = [P, I]
        L.mpcf_timing_read.argtypes = [P, P, P]
        U64 = C.c_ulonglong
        L.mpcf_ic_shock_tube.argtypes = [I, I, I, I, D, U64, P]
        L.mpcf_ic_shock_bubble.argtypes = [I, I, I, U64, P]
        L.mpcf_ic_bubble_cloud.argtypes = [I, I, I, I, U64, P, P]
        L.mpcf_ic_bubble_collapse.argtypes = [I, I, I, D, D, D, D, D, P]
        L.mpcf_case_bytes.argtypes = [I, I, I, P]
        L.mpcf_batch_create.argtypes = [C.POINTER(P), C.c_uint]
        L.mpcf_batch_destroy.argtypes = [P]
        L.mpcf_batch_add.argtypes = [P, I, I, I, D, P, P, P]
        L.mpcf_batch_remove.argtypes = [P, I]
        L.mpcf_batch_download.argtypes = [P, I, I, P]
        L.mpcf_batch_step.argtypes = [P, D, D]
        L.mpcf_batch_use_graph.argtypes = [P, I]
        L.mpcf_batch_last_dt.argtypes = [P, I, C.POINTER(D), C.POINTER(F)]
        L.mpcf_batch_synchronize.argtypes = [P]
        L.mpcf_batch_get_stream.argtypes = [P]
        L.mpcf_batch_get_stream.restype = P
        L.mpcf_batch_set_march.argtypes = [P, I, I, I]
        L.mpcf_batch_autotune.argtypes = [P, I, D, P]
        L.mpcf_batch_set_mode.argtypes = [P, I, I]
        L.mpcf_batch_tuning.argtypes = [P, I, P]
        L.mpcf_batch_query.argtypes = [P, P]
        L.mpcf_upload_f64.argtypes = [P, I, P]
        L.mpcf_download_f64.argtypes = [P, I, P]
        L.mpcf_last_sos64.argtypes = [P, C.POINTER(D)]
        L.mpcf_max_sos_noalloc.argtypes = [P, C.c_longlong, I, I, I, D, D, D, C.c_uint, P, P, P, P]
        L.mpcf_compare.argtypes = [P, P, I, P]
        L.mpcf_integrals.argtypes = [P, I, P]
        LL = C.c_longlong
        L.mpcf_server_create.argtypes = [C.POINTER(P), C.c_uint, LL, I, LL, I, D]
        L.mpcf_server_submit.argtypes = [P, I, I, I, D, P, I, P, P, P]
        L.mpcf_server_poll.argtypes = [P, P, P, I, P]
        L.mpcf_server_wait.argtypes = [P, D]
        L.mpcf_server_stats.argtypes = [P, P]
        L.mpcf_server_clock.argtypes = []
        L.mpcf_server_destroy.argtypes = [P]
        L.mpcf_sic_defaults.argtypes = [P]
        L.mpcf_sic_post_shock.argtypes = [P, P]
        L.mpcf_read_cloud_config.argtypes = [C.c_char_p, P]
        L.mpcf_read_cloud.argtypes = [C.c_char_p, P, I, P]
        L.mpcf_ic_sic_cloud.argtypes = [I, I, I, D, P, P, I, P]
        L.mpcf_post_shock_ratio.argtypes = [P, D, D, D, I, P]
        L.mpcf_sod_defaults.argtypes = [P]
        L.mpcf_ic_sod.argtypes = [I, I, I, D, P, P]
        L.mpcf_sbi_defaults.argtypes = [P]
        L.mpcf_sbi_config.argtypes = [C.c_char_p, P]
        L.mpcf_read_bubbles.argtypes = [C.c_char_p, P, I, P]
        L.mpcf_ic_sbi.argtypes = [I, I, I, D, P, P, I, P]
        L.mpcf_ic_static.argtypes = [I, I, I, P, P]
        L.mpcf_ic_steady.argtypes = [I, I, I, P, P]
        L.mpcf_halo_count.argtypes = [P, I, P]
        L.mpcf_halo_export.argtypes = [P, I, P]
        L.mpcf_halo_import.argtypes = [P, I, P]
        L.mpcf_ghosts_read.argtypes = [P, I, P]
        L.mpcf_steps.argtypes = [P, I, D, D]
        L.mpcf_wavelet_dump.argtypes = [P, I, C.c_char_p, F, I, P]
        L.mpcf_wavelet_read.argtypes = [C.c_char_p, P, LL, P]
        L.mpcf_wavelet_encode.argtypes = [P, I, F, I, P, LL, P]
        L.mpcf_wavelet_decode.argtypes = [P, LL, I, I, P]
        S_ = C.c_char_p
        L.mpcf_stream_channels.argtypes = [I]
        L.mpcf_stream_out.argtypes = [P, I, P]
        L.mpcf_stream_in.argtypes = [P, P]
        L.mpcf_write_xdmf.argtypes = [S_, S_, I, I, I, I, D, I, S_]
        L.mpcf_write_save_info.argtypes = [S_, D, I, I]
        L.mpcf_read_save_info.argtypes = [S_, P, P, P]
        L.mpcf_dump.argtypes = [P, I, S_, S_, I, D]
        L.mpcf_save.argtypes = [P, S_, D, I, I, D]
        L.mpcf_restart.argtypes = [P, S_, P, P, P]
        for name in EXPORTS:
            if name not in ("mpcf_last_error", "mpcf_version", "mpcf_get_stream", "mpcf_batch_get_stream",
                            "mpcf_server_clock", "mpcf_sic_defaults", "mpcf_post_shock_ratio",
                            "mpcf_sod_defaults", "mpcf_sbi_defaults"):
                getattr(L, name).restype = I
        L.mpcf_server_clock.restype = D
        for name in ("mpcf_sic_defaults", "mpcf_post_shock_ratio", "mpcf_sod_defaults", "mpcf_sbi_defaults"):
            getattr(L, name).restype = None
        _LIB = L
    return _LIB


def _check(rc, **values):
    """raise MpcfError for a failed call; `values` (what the call still wrote,
    e.g. dt with MPCF_ENUMERIC) become attributes of the exception"""
    if rc != OK:
        e = MpcfError(rc, lib().mpcf_last_error().decode())
        for k, v in values.items():
            setattr(e, k, v)
        raise e


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _fields(arr):
    ptrs = (C.c_void_p * NVAR)()
    for q in range(NVAR):
        ptrs[q] = arr[q].ctypes.data
    return ptrs


class Block:
    """One resident node block (an mpcf_ctx)."""

    def __init__(self, nx, ny, nz, h, fast=False, sos_cpu=False, bc=None, split_xy=None, p64=False, fp64=False,
                 pipeline=True, epair=False):
        """split_xy: None = the library's default for the build, True/False = force.
        p64: binary32 storage with binary64 pressure recovery and speed-of-sound
        reduction (FAST arithmetic otherwise); fp64: the whole scheme in binary64;
        epair: p64 with E stored as a binary32 pair (hi + lo).
        pipeline: True = the library's choice of z-slabs for mpcf_step, False =
        whole-block sweeps in every stage (MPCF_FLAG_NO_PIPELINE), an int k >= 2
        = k slabs (mpcf_set_pipeline)."""
        self.shape = (nz, ny, nx)
        self.h = float(h)
        p64 = p64 or epair
        fast = fast or p64
        flags = (FLAG_FAST if fast else 0) | (FLAG_SOS_CPU if sos_cpu else 0)
        flags |= (FLAG_P64 if p64 else 0) | (FLAG_FP64 if fp64 else 0) | (0 if pipeline is not False else FLAG_NO_PIPELINE)
        flags |= FLAG_E_PAIR if epair else 0
        if split_xy is True:
            flags |= FLAG_SPLIT_XY
        elif split_xy is False:
            flags |= FLAG_FUSED_XY
        self.fused_xy = (split_xy is False) or (split_xy is None and not fp64)
        h_ = C.c_void_p()
        _check(lib().mpcf_create(C.byref(h_), nx, ny, nz, float(h), flags))
        self._h = h_
        if pipeline is not True and pipeline is not False:
            self.set_pipeline(pipeline)
        if bc is not None:
            self.set_bc(bc)

    def close(self):
        if getattr(self, "_h", None):
            lib().mpcf_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_stream(self, stream):
        """Run on a caller's hipStream_t (its address as an int); None: the
        block's own stream again."""
        _check(lib().mpcf_set_stream(self._h, C.c_void_p(stream) if stream else None))

    def set_pipeline(self, slabs):
        """-1: automatic, 0: whole-block sweeps, k >= 2: k z-slabs on two streams."""
        _check(lib().mpcf_set_pipeline(self._h, int(slabs)))

    def set_bc(self, bc):
        b = np.ascontiguousarray(bc, dtype=np.int32)
        _check(lib().mpcf_set_bc(self._h, _ptr(b)))

    def upload(self, which, arr):
        a = np.ascontiguousarray(arr, dtype=np.float32)
        assert a.shape == (NVAR,) + self.shape, a.shape
        _check(lib().mpcf_upload(self._h, which, _fields(a)))

    def download(self, which):
        out = np.empty((NVAR,) + self.shape, dtype=np.float32)
        _check(lib().mpcf_download(self._h, which, _fields(out)))
        return out

    def upload_f64(self, which, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        assert a.shape == (NVAR,) + self.shape, a.shape
        _check(lib().mpcf_upload_f64(self._h, which, _fields(a)))

    def download_f64(self, which=DATA):
        out = np.empty((NVAR,) + self.shape, dtype=np.float64)
        _check(lib().mpcf_download_f64(self._h, which, _fields(out)))
        return out

    def compare(self, ref, which=DATA):
        """Per-field {l1, linf, l1_ref, linf_ref, rel_l1, rel_linf} of self - ref (binary64)."""
        out = (C.c_double * 28)()
        _check(lib().mpcf_compare(self._h, ref._h, which, C.cast(out, C.c_void_p)))
        res = []
        for q in range(NVAR):
            s1, m1, s2, m2 = out[4 * q:4 * q + 4]
            res.append({"l1": s1, "linf": m1, "l1_ref": s2, "linf_ref": m2,
                        "rel_l1": s1 / s2 if s2 else float(s1 != 0) * float("inf"),
                        "rel_linf": m1 / m2 if m2 else float(m1 != 0) * float("inf")})
        return res

    def integrals(self, which=DATA):
        """h^3 * sum of each field over the interior (binary64)."""
        out = (C.c_double * NVAR)()
        _check(lib().mpcf_integrals(self._h, which, C.cast(out, C.c_void_p)))
        return list(out)

    def stream_out(self, kind=STREAM_TENSOR):
        """The current solution through a dump streamer: (NZ, NY, NX, C) binary32."""
        nc = lib().mpcf_stream_channels(kind)
        out = np.empty(self.shape + (nc,), np.float32)
        _check(lib().mpcf_stream_out(self._h, kind, _ptr(out)))
        return out

    def stream_in(self, arr):
        a = np.ascontiguousarray(arr, dtype=np.float32)
        assert a.shape == self.shape + (7,), a.shape
        _check(lib().mpcf_stream_in(self._h, _ptr(a)))

    def dump(self, kind, dirname, name, step=0):
        _check(lib().mpcf_dump(self._h, kind, str(dirname).encode(), name.encode(), int(step), self.h))

    def save(self, dirname, t, step, fcount=0):
        _check(lib().mpcf_save(self._h, str(dirname).encode(), float(t), int(step), int(fcount), self.h))

    def restart(self, dirname):
        """Read a checkpoint; returns (t, step, fcount)."""
        t, st, fc = C.c_double(), C.c_int(), C.c_int()
        _check(lib().mpcf_restart(self._h, str(dirname).encode(), C.cast(C.byref(t), C.c_void_p),
                                  C.cast(C.byref(st), C.c_void_p), C.cast(C.byref(fc), C.c_void_p)))
        return t.value, st.value, fc.value

    def last_sos64(self):
        s = C.c_double()
        _check(lib().mpcf_last_sos64(self._h, C.byref(s)))
        return s.value

    @property
    def data(self):
        return self.download(DATA)

    @data.setter
    def data(self, a):
        self.upload(DATA, a)

    @property
    def tmp(self):
        return self.download(TMP)

    @tmp.setter
    def tmp(self, a):
        self.upload(TMP, a)

    def load_ghosts(self):
        _check(lib().mpcf_load_ghosts(self._h))

    def max_sos(self):
        s = C.c_float()
        _check(lib().mpcf_max_sos(self._h, C.byref(s)))
        return s.value

    def convection(self, a, dtinvh):
        _check(lib().mpcf_convection(self._h, float(a), float(dtinvh)))

    def update(self, b):
        _check(lib().mpcf_update(self._h, float(b)))

    def process_all(self, a, b, dtinvh):
        _check(lib().mpcf_process_all(self._h, float(a), float(b), float(dtinvh)))

    def step(self, cfl, dt_max=1e30, sync=True):
        if sync:
            dt = C.c_double()
            rc = lib().mpcf_step(self._h, float(cfl), float(dt_max), C.cast(C.byref(dt), C.c_void_p))
            _check(rc, dt=dt.value)
            return dt.value
        _check(lib().mpcf_step(self._h, float(cfl), float(dt_max), None))
        return None

    def steps(self, nsteps, cfl, dt_max=1e30):
        """nsteps LSRK3 steps queued at once (asynchronous; consecutive FAST
        steps overlap); dt of the last one: last_dt()"""
        _check(lib().mpcf_steps(self._h, int(nsteps), float(cfl), float(dt_max)))

    def last_dt(self):
        d, s = C.c_double(), C.c_float()
        rc = lib().mpcf_last_dt(self._h, C.byref(d), C.byref(s))
        _check(rc, dt=d.value, sos=s.value)
        return d.value, s.value

    def synchronize(self):
        _check(lib().mpcf_synchronize(self._h))

    def halo_count(self, face):
        """Elements of one face's halo buffer (7 fields x 3 layers x the face)."""
        n = C.c_longlong()
        _check(lib().mpcf_halo_count(self._h, int(face), C.cast(C.byref(n), C.c_void_p)))
        return n.value

    def halo_export(self, face, dev_ptr):
        """The three boundary cells of `face` into the device buffer at dev_ptr
        (asynchronous on this block's stream)."""
        _check(lib().mpcf_halo_export(self._h, int(face), C.c_void_p(int(dev_ptr))))

    def ghosts_read(self, face, dev_ptr):
        """The ghost band of `face` into the device buffer at dev_ptr, in the
        halo buffer order (asynchronous on this block's stream)."""
        _check(lib().mpcf_ghosts_read(self._h, int(face), C.c_void_p(int(dev_ptr))))

    def halo_import(self, face, dev_ptr):
        """The device buffer at dev_ptr into the ghost band of `face`
        (asynchronous on this block's stream)."""
        _check(lib().mpcf_halo_import(self._h, int(face), C.c_void_p(int(dev_ptr))))

    def wavelet_dump(self, channel, path, threshold=0.0, half=False):
        """Wavelet-compressed dump of StreamerGridPointIterative channel 0..6 to
        <path>.StreamerGridPointIterative.channel<k>; returns (stream bytes,
        file bytes, compression rate)."""
        st = (C.c_double * 3)()
        _check(lib().mpcf_wavelet_dump(self._h, int(channel), str(path).encode(), float(threshold), int(bool(half)),
                                       C.cast(st, C.c_void_p)))
        return tuple(st)

    @property
    def stream(self):
        return lib().mpcf_get_stream(self._h)

    def timing(self, enable=True):
        _check(lib().mpcf_timing_enable(self._h, 1 if enable else 0))

    def timing_read(self):
        """{kernel class: (total ms, launches)} since the last read."""
        ms = (C.c_double * 8)()
        cnt = (C.c_longlong * 8)()
        _check(lib().mpcf_timing_read(self._h, C.cast(ms, C.c_void_p), C.cast(cnt, C.c_void_p)))
        return {k: (ms[i], cnt[i]) for i, k in enumerate(KERNEL_CLASSES)}

    def query(self):
        out = (C.c_longlong * 8)()
        _check(lib().mpcf_query(self._h, C.cast(out, C.c_void_p)))
        return list(out)


def flop_count(fast=False):
    out = (C.c_double * 16)()
    _check(lib().mpcf_flop_count(FLAG_FAST if fast else 0, C.cast(out, C.c_void_p)))
    keys = ["sweep_x", "sweep_y", "sweep_z", "cell_stage", "sos_cell", "cell_step",
            "add", "mul", "div", "sqrt", "dadd", "dmul", "ddiv", "minmax", "cvt"]
    return {k: out[i] for i, k in enumerate(keys)}


def host_weno(in5, fast=False):
    a = np.ascontiguousarray(in5, dtype=np.float32)
    n = a.shape[0]
    m = np.empty(n, np.float32)
    p = np.empty(n, np.float32)
    _check(lib().mpcf_host_weno(n, _ptr(a), _ptr(m), _ptr(p), FLAG_FAST if fast else 0))
    return m, p


def host_face(st, d, fast=False):
    s = np.ascontiguousarray(st, dtype=np.float32)
    n = s.shape[0]
    out = np.empty((n, 12), np.float32)
    _check(lib().mpcf_host_face(n, d, _ptr(s), _ptr(out), FLAG_FAST if fast else 0))
    return out


def test_weno(in5, fast=False):
    a = np.ascontiguousarray(in5, dtype=np.float32)
    n = a.shape[0]
    m = np.empty(n, np.float32)
    p = np.empty(n, np.float32)
    _check(lib().mpcf_test_weno(n, _ptr(a), _ptr(m), _ptr(p), FLAG_FAST if fast else 0))
    return m, p


def test_face(st, d, fast=False, p64=False):
    s = np.ascontiguousarray(st, dtype=np.float32)
    n = s.shape[0]
    out = np.empty((n, 12), np.float32)
    flags = (FLAG_FAST if fast or p64 else 0) | (FLAG_P64 if p64 else 0)
    _check(lib().mpcf_test_face(n, d, _ptr(s), _ptr(out), flags))
    return out


def _new_fields(nx, ny, nz):
    return np.empty((NVAR, nz, ny, nx), dtype=np.float32)


def ic_shock_tube(nx, ny, nz, axis=0, noise=1e-3, seed=0):
    f = _new_fields(nx, ny, nz)
    _check(lib().mpcf_ic_shock_tube(nx, ny, nz, axis, float(noise), seed, _fields(f)))
    return f


def ic_shock_bubble(nx, ny, nz, seed=0):
    f = _new_fields(nx, ny, nz)
    _check(lib().mpcf_ic_shock_bubble(nx, ny, nz, seed, _fields(f)))
    return f


def ic_bubble_cloud(nx, ny, nz, nbubbles=50, seed=0):
    f = _new_fields(nx, ny, nz)
    b = np.zeros((nbubbles, 4))
    _check(lib().mpcf_ic_bubble_cloud(nx, ny, nz, nbubbles, seed, _ptr(b), _fields(f)))
    return f, b


def ic_bubble_collapse(nx, ny, nz, radius, pratio, centre=(0.5, 0.5, 0.5)):
    f = _new_fields(nx, ny, nz)
    _check(lib().mpcf_ic_bubble_collapse(nx, ny, nz, float(radius), float(pratio), *map(float, centre), _fields(f)))
    return f


def sic_params(**kw):
    """SicParams with the reference's defaults, overridden by keyword."""
    p = SicParams()
    lib().mpcf_sic_defaults(C.byref(p))
    for k, v in kw.items():
        if k in ("shock_n", "shock_s"):
            getattr(p, k)[:] = [float(x) for x in v]
        else:
            setattr(p, k, float(v))
    return p


def sic_post_shock(p=None):
    """(rho1, u1, p1, mach, c0) of the shock (Sim_SICCloudMPI.cpp:87-104)."""
    p = p or sic_params()
    out = (C.c_double * 5)()
    _check(lib().mpcf_sic_post_shock(C.byref(p), C.cast(out, C.c_void_p)))
    return tuple(out)


def read_cloud_config(path):
    out = (C.c_double * 11)()
    _check(lib().mpcf_read_cloud_config(str(path).encode(), C.cast(out, C.c_void_p)))
    keys = ("n_shapes", "n_small", "min_rad", "max_rad")
    d = {k: out[i] for i, k in enumerate(keys)}
    d["n_shapes"], d["n_small"] = int(d["n_shapes"]), int(d["n_small"])
    d["seed_s"], d["seed_e"], d["n_sensors"] = tuple(out[4:7]), tuple(out[7:10]), int(out[10])
    return d


def read_cloud(path, cap=1 << 20):
    """cloud.dat spheres as an (n, 4) array of x, y, z, r."""
    buf = np.empty((cap, 4), np.float64)
    n = C.c_int()
    _check(lib().mpcf_read_cloud(str(path).encode(), _ptr(buf), cap, C.cast(C.byref(n), C.c_void_p)))
    return buf[:min(n.value, cap)].copy()


def ic_sic_cloud(nx, ny, nz, h=None, params=None, shapes=None):
    """The shock-induced cloud collapse initial condition (27-point Simpson
    averages of set_IC, Sim_SICCloudMPI.cpp:247-357)."""
    f = _new_fields(nx, ny, nz)
    p = params or sic_params()
    sh = np.ascontiguousarray(np.zeros((0, 4)) if shapes is None else shapes, dtype=np.float64).reshape(-1, 4)
    hh = h if h is not None else 1.0 / max(nx, ny, nz)
    _check(lib().mpcf_ic_sic_cloud(nx, ny, nz, float(hh), C.byref(p), _ptr(sh) if len(sh) else None, len(sh),
                                   _fields(f)))
    return f


def post_shock_ratio(pre, mach, gamma, pc=0.0, real32=False):
    """(rho, u, p) behind a shock of Mach number `mach` running into pre =
    (rho, u, p) (SimTools::getPostShockRatio, Types.h:101-109); real32 rounds
    like the reference's binary32 build."""
    a = (C.c_double * 3)(*map(float, pre))
    out = (C.c_double * 3)()
    lib().mpcf_post_shock_ratio(C.cast(a, C.c_void_p), float(mach), float(gamma), float(pc), int(bool(real32)),
                                C.cast(out, C.c_void_p))
    return tuple(out)


def _set_fields(p, kw, ints):
    for k, v in kw.items():
        if not hasattr(p, k):
            raise TypeError(f"unknown parameter {k!r}")
        if k == "dims":
            p.dims[:] = [int(x) for x in v]
        else:
            setattr(p, k, int(v) if k in ints else float(v))
    return p


def sod_params(**kw):
    """SodParams with the reference's defaults (Sim_SodMPI.cpp:27-40),
    overridden by keyword; config="x"|"y"|"z" sets the tube axis."""
    p = SodParams()
    lib().mpcf_sod_defaults(C.byref(p))
    conf = kw.pop("config", None)
    if conf is not None:
        if conf not in ("x", "y", "z"):
            raise ValueError(f"-config {conf} is not possible")
        p.dir = "xyz".index(conf)
    return _set_fields(p, kw, ("reverse", "dir"))


def ic_sod(nx, ny, nz, h=None, params=None):
    """The Sim_SodMPI initial condition (Sim_SodMPI.cpp:58-95)."""
    f = _new_fields(nx, ny, nz)
    p = params or sod_params()
    hh = h if h is not None else 1.0 / max(nx, ny, nz)
    _check(lib().mpcf_ic_sod(nx, ny, nz, float(hh), C.byref(p), _fields(f)))
    return f


def sbi_config(conf):
    """dims (principal, minor, dummy) for -config xy|xz|yx|yz|zx|zy."""
    d = (C.c_int * 3)()
    _check(lib().mpcf_sbi_config(str(conf).encode(), C.cast(d, C.c_void_p)))
    return tuple(d)


def sbi_params(**kw):
    """SbiParams with the reference's defaults (Sim_2DSBIMPI.cpp:95-108),
    overridden by keyword; config="xy"|... sets dims."""
    p = SbiParams()
    lib().mpcf_sbi_defaults(C.byref(p))
    conf = kw.pop("config", None)
    if conf is not None:
        p.dims[:] = sbi_config(conf)
    return _set_fields(p, kw, ("mollfactor", "reverse", "dims"))


def sbi_bc(dims):
    """Boundary kinds of the 2D shock-bubble case: reflecting on the two faces
    normal to the minor axis dims[1], absorbing elsewhere
    (GPUlab2DSBI_{x,y,z}reflect, Sim_2DSBIMPI.h:27-81, Sim_2DSBIMPI.cpp:66-84)."""
    return [BC_REFLECTING if f // 2 == dims[1] else BC_ABSORBING for f in range(6)]


def sbi_bubble_file(dims):
    """The bubble file the reference reads for dims (Sim_2DSBIMPI.cpp:124-131)."""
    return "xyz"[dims[0]] + "xyz"[dims[1]] + "bubbles.dat"


def read_bubbles(path, cap=1 << 20):
    """A <config>bubbles.dat file as an (n, 4) array of x, y, z, r."""
    buf = np.empty((cap, 4), np.float64)
    n = C.c_int()
    _check(lib().mpcf_read_bubbles(str(path).encode(), _ptr(buf), cap, C.cast(C.byref(n), C.c_void_p)))
    return buf[:min(n.value, cap)].copy()


def ic_sbi(nx, ny, nz, h=None, params=None, bubbles=None):
    """The Sim_2DSBIMPI initial condition (Sim_2DSBIMPI.cpp:86-233);
    bubbles: (n, 4) array of x, y, z, r."""
    f = _new_fields(nx, ny, nz)
    p = params or sbi_params()
    b = np.ascontiguousarray(np.zeros((0, 4)) if bubbles is None else bubbles, dtype=np.float64).reshape(-1, 4)
    hh = h if h is not None else 1.0 / max(nx, ny, nz)
    _check(lib().mpcf_ic_sbi(nx, ny, nz, float(hh), C.byref(p), _ptr(b) if len(b) else None, len(b), _fields(f)))
    return f


def ic_static(nx, ny, nz, dims=(0, 1, 2)):
    """The Sim_StaticIC layout probe (Sim_StaticIC.cpp:76-97)."""
    f = _new_fields(nx, ny, nz)
    d = (C.c_int * 3)(*map(int, dims))
    _check(lib().mpcf_ic_static(nx, ny, nz, C.cast(d, C.c_void_p), _fields(f)))
    return f


def ic_steady(nx, ny, nz, rho=1.0, u=0.0, v=0.0, w=0.0, p=1.0, g=1.4, pc=0.0):
    """The Sim_SteadyStateMPI uniform state (Sim_SteadyStateMPI.cpp:95-134)."""
    f = _new_fields(nx, ny, nz)
    st = (C.c_double * 7)(*map(float, (rho, u, v, w, p, g, pc)))
    _check(lib().mpcf_ic_steady(nx, ny, nz, C.cast(st, C.c_void_p), _fields(f)))
    return f


def wavelet_encode(cube, threshold=0.0, half=False):
    """The compressor's stream (bitset bytes, then survivors) of a cube of edge
    n (a power of two in 8..1024)."""
    a = np.ascontiguousarray(cube, dtype=np.float32)
    n = a.shape[0]
    assert a.shape == (n, n, n)
    nb = C.c_longlong()
    _check(lib().mpcf_wavelet_encode(_ptr(a), n, float(threshold), int(bool(half)), None, 0,
                                     C.cast(C.byref(nb), C.c_void_p)))
    out = np.empty(nb.value, np.uint8)
    _check(lib().mpcf_wavelet_encode(_ptr(a), n, float(threshold), int(bool(half)), _ptr(out), nb.value,
                                     C.cast(C.byref(nb), C.c_void_p)))
    return out.tobytes()


def wavelet_decode(buf, n, half=False):
    b = np.frombuffer(bytes(buf), np.uint8).copy()
    out = np.empty((n, n, n), np.float32)
    _check(lib().mpcf_wavelet_decode(_ptr(b), b.size, int(n), int(bool(half)), _ptr(out)))
    return out


def wavelet_read(path):
    """A wavelet dump read back: the (n, n, n) binary32 channel."""
    n = C.c_int()
    _check(lib().mpcf_wavelet_read(str(path).encode(), None, 0, C.cast(C.byref(n), C.c_void_p)))
    out = np.empty((n.value,) * 3, np.float32)
    _check(lib().mpcf_wavelet_read(str(path).encode(), _ptr(out), out.size, C.cast(C.byref(n), C.c_void_p)))
    return out


def write_save_info(path, t, step, fcount):
    _check(lib().mpcf_write_save_info(str(path).encode(), float(t), int(step), int(fcount)))


def read_save_info(path):
    t, st, fc = C.c_double(), C.c_int(), C.c_int()
    _check(lib().mpcf_read_save_info(str(path).encode(), C.cast(C.byref(t), C.c_void_p),
                                     C.cast(C.byref(st), C.c_void_p), C.cast(C.byref(fc), C.c_void_p)))
    return t.value, st.value, fc.value


def write_xdmf(path, raw_name, nx, ny, nz, channels, h, step, attribute):
    _check(lib().mpcf_write_xdmf(str(path).encode(), raw_name.encode(), nx, ny, nz, channels, float(h), int(step),
                                 attribute.encode()))


def case_bytes(nx, ny, nz):
    b = C.c_longlong()
    _check(lib().mpcf_case_bytes(nx, ny, nz, C.cast(C.byref(b), C.c_void_p)))
    return b.value


class Batch:
    """A batched ensemble (mpcf_batch): cases of uneven sizes advanced together."""

    def __init__(self, fast=False, sos_cpu=False, graph=False, p64=False, mode=None, threads=0):
        """mode: None = the library's default (BATCH_GROUPS), or
        BATCH_GROUPS / BATCH_PERSISTENT; threads: the persistent launches'
        workgroup size (64, 128, 256; 0 = 256)."""
        h_ = C.c_void_p()
        flags = (FLAG_FAST if fast else 0) | (FLAG_SOS_CPU if sos_cpu else 0) | (FLAG_P64 if p64 else 0)
        _check(lib().mpcf_batch_create(C.byref(h_), flags))
        self._h = h_
        self.shapes = {}
        if mode is not None:
            self.set_mode(mode, threads)
        if graph:
            self.use_graph(True)

    def set_mode(self, mode, threads=0):
        _check(lib().mpcf_batch_set_mode(self._h, int(mode), int(threads)))

    def tuning(self, extent):
        """(mode, persistent workgroup size, x+y march length, z march length)
        in force for cases of this extent (0: no tuned length)"""
        out = (C.c_int * 4)()
        _check(lib().mpcf_batch_tuning(self._h, int(extent), C.cast(out, C.c_void_p)))
        return tuple(out)

    def close(self):
        if getattr(self, "_h", None):
            lib().mpcf_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def add(self, fields, h=None, bc=None):
        a = np.ascontiguousarray(fields, dtype=np.float32)
        nz, ny, nx = a.shape[1:]
        h = 1.0 / max(nx, ny, nz) if h is None else h
        cid = C.c_int()
        bcp = None
        if bc is not None:
            bca = np.ascontiguousarray(bc, dtype=np.int32)
            bcp = _ptr(bca)
        _check(lib().mpcf_batch_add(self._h, nx, ny, nz, float(h), bcp, _fields(a), C.cast(C.byref(cid), C.c_void_p)))
        self.shapes[cid.value] = (nz, ny, nx)
        return cid.value

    def remove(self, cid):
        _check(lib().mpcf_batch_remove(self._h, cid))
        self.shapes.pop(cid, None)

    def download(self, cid, which=DATA):
        out = np.empty((NVAR,) + self.shapes[cid], dtype=np.float32)
        _check(lib().mpcf_batch_download(self._h, cid, which, _fields(out)))
        return out

    def step(self, cfl, dt_max=1e30):
        _check(lib().mpcf_batch_step(self._h, float(cfl), float(dt_max)))

    def use_graph(self, enable=True):
        _check(lib().mpcf_batch_use_graph(self._h, 1 if enable else 0))

    def last_dt(self, cid):
        d, s = C.c_double(), C.c_float()
        _check(lib().mpcf_batch_last_dt(self._h, cid, C.byref(d), C.byref(s)))
        return d.value, s.value

    def synchronize(self):
        _check(lib().mpcf_batch_synchronize(self._h))

    def set_march(self, cid, len_xy, len_z):
        _check(lib().mpcf_batch_set_march(self._h, cid, len_xy, len_z))

    def autotune(self, steps=2, cfl=0.3):
        best = C.c_int()
        _check(lib().mpcf_batch_autotune(self._h, steps, float(cfl), C.cast(C.byref(best), C.c_void_p)))
        return best.value

    def query(self):
        out = (C.c_longlong * 6)()
        _check(lib().mpcf_batch_query(self._h, C.cast(out, C.c_void_p)))
        keys = ("cases", "cells", "bytes", "captures", "steps", "tuned")
        return dict(zip(keys, out))


def server_clock():
    """Seconds on the clock the server timestamps completions with."""
    return lib().mpcf_server_clock()


class Server:
    """Continuous-batching request server (mpcf_server): submit cases from this
    thread, the library's engine thread advances them on the device."""

    POLICIES = {"all": SERVE_ALL, "fifo_fill": SERVE_FIFO_FILL, "srpt_fill": SERVE_SRPT_FILL, "hybrid": SERVE_HYBRID,
                "fifo_slack": SERVE_FIFO_SLACK}

    def __init__(self, hbm_bytes, policy="fifo_fill", fill_cells=0, max_live=256, cfl=0.3, fast=True,
                 p64=False, graph=True, sos_cpu=False):
        flags = (FLAG_FAST if fast or p64 else 0) | (FLAG_P64 if p64 else 0) | (0 if graph else FLAG_SERVE_EAGER)
        flags |= FLAG_SOS_CPU if sos_cpu else 0
        h_ = C.c_void_p()
        _check(lib().mpcf_server_create(C.byref(h_), flags, int(hbm_bytes), self.POLICIES[policy],
                                        int(fill_cells), int(max_live), float(cfl)))
        self._h = h_
        self._keep = {}    # request id -> (input, output) arrays, alive until done
        self.results = {}  # request id -> output fields
        self.t_done = {}

    def submit(self, fields, steps, h=None, bc=None):
        a = np.ascontiguousarray(fields, dtype=np.float32)
        nz, ny, nx = a.shape[1:]
        out = np.empty_like(a)
        bcp = None
        if bc is not None:
            b = np.ascontiguousarray(bc, dtype=np.int32)
            bcp = _ptr(b)
        rid = C.c_longlong()
        hh = h if h is not None else 1.0 / max(nx, ny, nz)
        _check(lib().mpcf_server_submit(self._h, nx, ny, nz, float(hh), bcp, int(steps), _fields(a), _fields(out),
                                        C.cast(C.byref(rid), C.c_void_p)))
        self._keep[rid.value] = (a, out)
        return rid.value

    def poll(self):
        """[(id, completion time)] completed since the last poll; results land in self.results."""
        got = []
        cap = 256
        ids = (C.c_longlong * cap)()
        ts = (C.c_double * cap)()
        n = C.c_int()
        while True:
            _check(lib().mpcf_server_poll(self._h, C.cast(ids, C.c_void_p), C.cast(ts, C.c_void_p), cap,
                                          C.cast(C.byref(n), C.c_void_p)))
            for k in range(n.value):
                rid = ids[k]
                got.append((rid, ts[k]))
                self.results[rid] = self._keep.pop(rid)[1]
                self.t_done[rid] = ts[k]
            if n.value < cap:
                return got

    def wait(self, timeout=-1.0):
        _check(lib().mpcf_server_wait(self._h, float(timeout)))
        return self.poll()

    def stats(self):
        out = (C.c_longlong * 8)()
        _check(lib().mpcf_server_stats(self._h, C.cast(out, C.c_void_p)))
        keys = ("submitted", "completed", "steps", "captures", "rebuilds", "peak_active", "peak_bytes",
                "active_cells")
        return dict(zip(keys, list(out)))

    def close(self):
        if getattr(self, "_h", None):
            _check(lib().mpcf_server_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
