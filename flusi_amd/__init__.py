"""flusi_amd -- MI355X implementation of FLUSI's fluid hot path.

Python binding over the C ABI in ``include/flusi_hip.h`` (libflusi_hip.so,
built in-tree by ``flusi_amd.build``).  Method names follow the reference's
routines (``cal_nlk``, ``fluid_timestep``, ``create_mask``, ``cal_drag`` via
``write_integrals``, ``cal_vis``, ``fft``/``ifft``, ``compute_energies_k`` as
``energy_k``).

Array conventions are the reference's single-rank layouts written as C-order
numpy shapes: physical fields ``(3, nz, ny, nx)`` float64, spectral fields
``(3, nx//2+1, ny, nz)`` complex128.  There is no CPU fallback: importing
works without a GPU, but constructing a ``Flusi`` needs the built library and
a visible MI355X.
"""
from __future__ import annotations

import ctypes
import os
from typing import Dict, Optional

import numpy as np

__all__ = ["Flusi", "Ensemble", "FlusiError", "FP64", "FP32_MIXED", "load_library", "INTEGRALS_NV", "EXPORTED",
           "ENSEMBLE_MEMBERS", "ENSEMBLE_BATCHED", "ENSEMBLE_GRAPH"]

FP64 = 0
FP32_MIXED = 1
INTEGRALS_NV = 61
TIMER_SLOTS = 8
ENSEMBLE_MEMBERS = 0   # every member steps on its own
ENSEMBLE_BATCHED = 1   # one launch per pass and batch
ENSEMBLE_GRAPH = 2     # the batched step captured once into a hipGraph and replayed

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libflusi_hip.so")

# name -> (restype, argtypes)
_D = ctypes.POINTER(ctypes.c_double)
_P = ctypes.c_void_p
_I = ctypes.c_int
EXPORTED: Dict[str, tuple] = {
    "flusi_init": (_I, [ctypes.c_char_p, _I, ctypes.POINTER(_P)]),
    "flusi_free": (None, [_P]),
    "flusi_last_error": (ctypes.c_char_p, [_P]),
    "flusi_get_dims": (_I, [_P, ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    "flusi_set_output_dir": (_I, [_P, ctypes.c_char_p]),
    "flusi_init_fields": (_I, [_P]),
    "flusi_set_uk": (_I, [_P, _D]),
    "flusi_get_uk": (_I, [_P, _D]),
    "flusi_get_u": (_I, [_P, _D]),
    "flusi_get_vorticity": (_I, [_P, _D]),
    "flusi_get_pressure": (_I, [_P, ctypes.c_double, _D]),
    "flusi_cal_nlk": (_I, [_P, ctypes.c_double, _I, _D, _D]),
    "flusi_fluid_timestep": (_I, [_P, _D, ctypes.POINTER(_I), _D, _D]),
    "flusi_step": (_I, [_P]),
    "flusi_run": (_I, [_P, _I]),
    "flusi_advance": (_I, [_P, _I]),
    "flusi_save_fields": (_I, [_P]),
    "flusi_dump_runtime_backup": (_I, [_P]),
    "flusi_read_runtime_backup": (_I, [_P, ctypes.c_char_p]),
    "flusi_get_time": (_I, [_P, _D, ctypes.POINTER(_I), _D, _D]),
    "flusi_create_mask": (_I, [_P, ctypes.c_double]),
    "flusi_get_mask": (_I, [_P, _D, _D, ctypes.POINTER(ctypes.c_int16)]),
    "flusi_write_integrals": (_I, [_P, ctypes.c_double, _D]),
    "flusi_cal_drag": (_I, [_P, ctypes.c_double, _D, _D]),
    "flusi_energy_k": (_I, [_P, _D]),
    "flusi_energy_k_noalloc": (_I, [_P, _D]),
    "flusi_cal_vis": (_I, [_P, ctypes.c_double, _D]),
    "flusi_fft": (_I, [_P, _D, _D]),
    "flusi_ifft": (_I, [_P, _D, _D]),
    "flusi_timers": (_I, [_P, _I, _D, ctypes.POINTER(ctypes.c_char_p)]),
    "flusi_autotune": (_I, [_P, _I, ctypes.POINTER(_I), _D]),
    "flusi_adjoint_ncontrols": (_I, [_P, ctypes.POINTER(_I)]),
    "flusi_adjoint_get_controls": (_I, [_P, _D]),
    "flusi_adjoint_forward": (_I, [_P, _D, _I, _I, _D]),
    "flusi_adjoint_backward": (_I, [_P, _D]),
    "flusi_ensemble_create": (_I, [_I, ctypes.POINTER(ctypes.c_char_p), _I, ctypes.POINTER(_P)]),
    "flusi_ensemble_free": (None, [_P]),
    "flusi_ensemble_last_error": (ctypes.c_char_p, [_P]),
    "flusi_ensemble_size": (_I, [_P, ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    "flusi_ensemble_batch": (_I, [_P, _I, ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(_I),
                                  ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    "flusi_ensemble_member": (_I, [_P, _I, ctypes.POINTER(_P)]),
    "flusi_ensemble_init_fields": (_I, [_P]),
    "flusi_ensemble_autotune": (_I, [_P, _I, ctypes.POINTER(_I), _D]),
    "flusi_ensemble_advance": (_I, [_P, _I, _I]),
}

_lib = None


def load_library(path: Optional[str] = None):
    """Load libflusi_hip.so and declare every exported entry point.  Raises
    if the library has not been built (there is no fallback path)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"{p} is missing: build it with `python -m flusi_amd.build` "
                           "(the HIP path has no CPU fallback)")
    lib = ctypes.CDLL(p)
    for name, (res, args) in EXPORTED.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


H5_LIB_PATH = os.path.join(_HERE, "libflusi_h5.so")
_h5 = None


def _h5lib():
    global _h5
    if _h5 is None:
        if not os.path.exists(H5_LIB_PATH):
            raise RuntimeError(f"{H5_LIB_PATH} is missing: build it with `python -m flusi_amd.build` (needs HDF5)")
        lib = ctypes.CDLL(H5_LIB_PATH)
        lib.fh5_field_dims.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(_I), ctypes.POINTER(_I),
                                       ctypes.POINTER(_I), ctypes.c_char_p, _I]
        lib.fh5_read_field.argtypes = [ctypes.c_char_p, ctypes.c_char_p, _I, _I, _I, _P, ctypes.c_char_p, _I]
        lib.fh5_read_attr.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, _P, _I, ctypes.c_char_p, _I]
        _h5 = lib
    return _h5


def read_h5_field(path: str, dset: Optional[str] = None):
    """A field file written by save_fields or a runtime backup: returns the
    (nz, ny, nx) float64 array of dataset `dset` (default: the file name up
    to its first '_', as flusi_hdf5_interface.f90 get_dsetname)."""
    lib = _h5lib()
    if dset is None:
        dset = os.path.basename(path).split("_")[0].split(".")[0]
    err = ctypes.create_string_buffer(512)
    nx, ny, nz = _I(), _I(), _I()
    rc = lib.fh5_field_dims(os.fsencode(path), dset.encode(), ctypes.byref(nx), ctypes.byref(ny), ctypes.byref(nz),
                            err, 512)
    if rc:
        raise FlusiError(rc, err.value.decode())
    out = np.empty((nz.value, ny.value, nx.value))
    rc = lib.fh5_read_field(os.fsencode(path), dset.encode(), nx.value, ny.value, nz.value, out.ctypes.data, err, 512)
    if rc:
        raise FlusiError(rc, err.value.decode())
    return out


def read_h5_attr(path: str, dset: str, name: str, n: int):
    """n values of the attribute `name` of dataset `dset` (as float64)."""
    lib = _h5lib()
    err = ctypes.create_string_buffer(512)
    out = np.empty(n)
    rc = lib.fh5_read_attr(os.fsencode(path), dset.encode(), name.encode(), out.ctypes.data, n, err, 512)
    if rc:
        raise FlusiError(rc, err.value.decode())
    return out


class FlusiError(RuntimeError):
    """A non-zero return of the C ABI: the reference's abort code + message."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"flusi error {code}: {msg}")
        self.code = code


def _dp(a: np.ndarray):
    return a.ctypes.data_as(_D)


class Flusi:
    """One solver context (one GPU, one stream) for a FLUSI .ini file."""

    def __init__(self, ini_path: Optional[str], precision: int = FP64, output_dir: Optional[str] = None,
                 _borrow=None):
        self._lib = load_library()
        self._ctx = _P()
        self._owner = None
        if _borrow is not None:  # a member context owned by an Ensemble
            self._owner, ctx = _borrow
            self._ctx = ctx
        else:
            rc = self._lib.flusi_init(os.fsencode(ini_path), int(precision), ctypes.byref(self._ctx))
            if rc != 0:
                msg = self._lib.flusi_last_error(self._ctx).decode() if self._ctx else ""
                if self._ctx:
                    self._lib.flusi_free(self._ctx)
                    self._ctx = _P()
                raise FlusiError(rc, msg)
        nx, ny, nz = _I(), _I(), _I()
        self._check(self._lib.flusi_get_dims(self._ctx, ctypes.byref(nx), ctypes.byref(ny), ctypes.byref(nz)))
        self.nx, self.ny, self.nz = nx.value, ny.value, nz.value
        self.nxh = self.nx // 2 + 1
        self.precision = precision
        if output_dir:
            self.set_output_dir(output_dir)

    # -- plumbing -----------------------------------------------------------
    def _check(self, rc: int):
        if rc != 0:
            raise FlusiError(rc, self._lib.flusi_last_error(self._ctx).decode())

    def close(self):
        if getattr(self, "_ctx", None):
            if self._owner is None:
                self._lib.flusi_free(self._ctx)
            self._ctx = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def spectral_shape(self):
        return (3, self.nxh, self.ny, self.nz)

    @property
    def physical_shape(self):
        return (3, self.nz, self.ny, self.nx)

    # -- API ------------------------------------------------------------------
    def set_output_dir(self, d: Optional[str]):
        self._check(self._lib.flusi_set_output_dir(self._ctx, os.fsencode(d) if d else None))

    def init_fields(self):
        self._check(self._lib.flusi_init_fields(self._ctx))

    def set_uk(self, uk: np.ndarray):
        a = np.ascontiguousarray(uk, dtype=np.complex128).reshape(self.spectral_shape)
        self._check(self._lib.flusi_set_uk(self._ctx, _dp(a.view(np.float64))))

    def get_uk(self) -> np.ndarray:
        a = np.empty(self.spectral_shape, dtype=np.complex128)
        self._check(self._lib.flusi_get_uk(self._ctx, _dp(a.view(np.float64))))
        return a

    def get_u(self) -> np.ndarray:
        a = np.empty(self.physical_shape)
        self._check(self._lib.flusi_get_u(self._ctx, _dp(a)))
        return a

    def get_vorticity(self) -> np.ndarray:
        a = np.empty(self.physical_shape)
        self._check(self._lib.flusi_get_vorticity(self._ctx, _dp(a)))
        return a

    def get_pressure(self, time: float) -> np.ndarray:
        """save_fields' pressure at `time` (mask regenerated when moving):
        (nz, ny, nx) float64."""
        out = np.empty((self.nz, self.ny, self.nx))
        self._check(self._lib.flusi_get_pressure(self._ctx, float(time), _dp(out)))
        return out

    def cal_nlk(self, time: float, it: int = 0, want_u: bool = False):
        nlk = np.empty(self.spectral_shape, dtype=np.complex128)
        u = np.empty(self.physical_shape) if want_u else None
        self._check(self._lib.flusi_cal_nlk(self._ctx, float(time), int(it), _dp(nlk.view(np.float64)),
                                            _dp(u) if u is not None else None))
        return (nlk, u) if want_u else nlk

    def fluid_timestep(self, time: float, it: int, dt0: float) -> float:
        t, i, d0, d1 = ctypes.c_double(time), _I(it), ctypes.c_double(dt0), ctypes.c_double(0.0)
        self._check(self._lib.flusi_fluid_timestep(self._ctx, ctypes.byref(t), ctypes.byref(i),
                                                   ctypes.byref(d0), ctypes.byref(d1)))
        return d1.value

    def step(self):
        self._check(self._lib.flusi_step(self._ctx))

    def run(self, nsteps: int = 0):
        self._check(self._lib.flusi_run(self._ctx, int(nsteps)))

    def advance(self, nsteps: int):
        self._check(self._lib.flusi_advance(self._ctx, int(nsteps)))

    def save_fields(self):
        """save_fields.f90 save_fields_fsi: HDF5 snapshots of the iSave* fields."""
        self._check(self._lib.flusi_save_fields(self._ctx))

    def dump_runtime_backup(self):
        """runtime_backuping.f90 dump_runtime_backup (runtime_backup0/1.h5)."""
        self._check(self._lib.flusi_dump_runtime_backup(self._ctx))

    def read_runtime_backup(self, path: str):
        """runtime_backuping.f90 read_runtime_backup (state, time, dt0, dt1, n1, it)."""
        self._check(self._lib.flusi_read_runtime_backup(self._ctx, os.fsencode(path)))

    def get_time(self):
        t, i, d0, d1 = ctypes.c_double(), _I(), ctypes.c_double(), ctypes.c_double()
        self._check(self._lib.flusi_get_time(self._ctx, ctypes.byref(t), ctypes.byref(i),
                                             ctypes.byref(d0), ctypes.byref(d1)))
        return t.value, i.value, d0.value, d1.value

    def create_mask(self, time: float):
        self._check(self._lib.flusi_create_mask(self._ctx, float(time)))

    def get_mask(self):
        n = self.nx * self.ny * self.nz
        mask = np.empty((self.nz, self.ny, self.nx))
        us = np.empty((3, self.nz, self.ny, self.nx))
        col = np.empty((self.nz, self.ny, self.nx), dtype=np.int16)
        self._check(self._lib.flusi_get_mask(self._ctx, _dp(mask), _dp(us),
                                             col.ctypes.data_as(ctypes.POINTER(ctypes.c_int16))))
        assert mask.size == n
        return mask, us, col

    def write_integrals(self, time: float) -> np.ndarray:
        out = np.zeros(INTEGRALS_NV)
        self._check(self._lib.flusi_write_integrals(self._ctx, float(time), _dp(out)))
        return out

    def cal_drag(self, time: float, u: Optional[np.ndarray] = None) -> np.ndarray:
        """drag.f90:17-454 cal_drag of the physical velocity u ((3, nz, ny, nx)
        float64; None: the velocity of the current uk) with the current mask:
        the INTEGRALS_NV layout of write_integrals, force slots only."""
        out = np.zeros(INTEGRALS_NV)
        a = None if u is None else np.ascontiguousarray(u, dtype=np.float64).reshape(self.physical_shape)
        self._check(self._lib.flusi_cal_drag(self._ctx, float(time), _dp(a) if a is not None else None, _dp(out)))
        return out

    def energy_k(self) -> float:
        e = ctypes.c_double()
        self._check(self._lib.flusi_energy_k(self._ctx, ctypes.byref(e)))
        return e.value

    def energy_k_noalloc(self) -> float:
        """energy_k without device allocation (scratch in a fixed device array)."""
        e = ctypes.c_double()
        self._check(self._lib.flusi_energy_k_noalloc(self._ctx, ctypes.byref(e)))
        return e.value

    def cal_vis(self, dt: float) -> np.ndarray:
        v = np.empty((self.nxh, self.ny, self.nz))
        self._check(self._lib.flusi_cal_vis(self._ctx, float(dt), _dp(v)))
        return v

    def fft(self, f: np.ndarray) -> np.ndarray:
        a = np.ascontiguousarray(f, dtype=np.float64).reshape(self.nz, self.ny, self.nx)
        out = np.empty((self.nxh, self.ny, self.nz), dtype=np.complex128)
        self._check(self._lib.flusi_fft(self._ctx, _dp(a), _dp(out.view(np.float64))))
        return out

    def ifft(self, fk: np.ndarray) -> np.ndarray:
        a = np.ascontiguousarray(fk, dtype=np.complex128).reshape(self.nxh, self.ny, self.nz)
        out = np.empty((self.nz, self.ny, self.nx))
        self._check(self._lib.flusi_ifft(self._ctx, _dp(a.view(np.float64)), _dp(out)))
        return out

    def timers(self, enable: int = -1):
        ms = np.zeros(TIMER_SLOTS)
        names = (ctypes.c_char_p * TIMER_SLOTS)()
        self._check(self._lib.flusi_timers(self._ctx, int(enable), _dp(ms), names))
        return {names[i].decode(): float(ms[i]) for i in range(TIMER_SLOTS)}

    # -- discrete adjoint ---------------------------------------------------
    def adjoint_controls(self) -> np.ndarray:
        """Current values of the wing kinematics' Fourier coefficients (the
        adjoint's controls)."""
        n = _I()
        self._check(self._lib.flusi_adjoint_ncontrols(self._ctx, ctypes.byref(n)))
        th = np.zeros(n.value)
        self._check(self._lib.flusi_adjoint_get_controls(self._ctx, _dp(th)))
        return th

    def adjoint_forward(self, theta, nsteps: int, every: int = 20) -> float:
        """Mean lift over nsteps RK4 steps from the initial condition with the
        given controls; keeps checkpoints for adjoint_backward."""
        th = np.ascontiguousarray(theta, dtype=np.float64)
        J = ctypes.c_double()
        self._check(self._lib.flusi_adjoint_forward(self._ctx, _dp(th), int(nsteps), int(every), ctypes.byref(J)))
        self._nctl = th.size
        return J.value

    def adjoint_backward(self) -> np.ndarray:
        """dJ/dtheta of the last adjoint_forward."""
        g = np.zeros(getattr(self, "_nctl", 0) or self.adjoint_controls().size)
        self._check(self._lib.flusi_adjoint_backward(self._ctx, _dp(g)))
        return g

    def autotune(self, reps: int = 3):
        """Picks the fastest kernel variant for this grid (cached per grid,
        precision and batch size); returns (variant[6], ms per right-hand side)."""
        v = (_I * 6)()
        ms = ctypes.c_double()
        self._check(self._lib.flusi_autotune(self._ctx, int(reps), v, ctypes.byref(ms)))
        return list(v), ms.value


class Ensemble:
    """Many independent simulations (one .ini each) advanced together; members
    with equal grids and solver parameters form batches whose passes run as
    one launch.  ``member(i)`` is an ordinary ``Flusi`` context for reading
    fields and integrals between advances."""

    def __init__(self, ini_paths, precision: int = FP32_MIXED):
        self._lib = load_library()
        self._e = _P()
        paths = (ctypes.c_char_p * len(ini_paths))(*[os.fsencode(p) for p in ini_paths])
        rc = self._lib.flusi_ensemble_create(len(ini_paths), paths, int(precision), ctypes.byref(self._e))
        if rc != 0:
            msg = self._lib.flusi_ensemble_last_error(self._e).decode() if self._e else ""
            if self._e:
                self._lib.flusi_ensemble_free(self._e)
                self._e = _P()
            raise FlusiError(rc, msg)
        n, nb = _I(), _I()
        self._check(self._lib.flusi_ensemble_size(self._e, ctypes.byref(n), ctypes.byref(nb)))
        self.n, self.nbatches = n.value, nb.value
        self.precision = precision
        self._members = [None] * self.n

    def _check(self, rc: int):
        if rc != 0:
            raise FlusiError(rc, self._lib.flusi_ensemble_last_error(self._e).decode())

    def close(self):
        if getattr(self, "_e", None):
            for m in self._members:
                if m is not None:
                    m._ctx = _P()
            self._lib.flusi_ensemble_free(self._e)
            self._e = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def batches(self, with_lead: bool = False):
        """[(members, nx, ny, nz)] per batch, with the index of the member
        that launches for the batch appended when with_lead is set."""
        out = []
        for b in range(self.nbatches):
            v = [_I() for _ in range(5)]
            self._check(self._lib.flusi_ensemble_batch(self._e, b, *[ctypes.byref(x) for x in v]))
            out.append(tuple(x.value for x in (v if with_lead else v[:4])))
        return out

    def member(self, i: int) -> Flusi:
        if self._members[i] is None:
            ctx = _P()
            self._check(self._lib.flusi_ensemble_member(self._e, int(i), ctypes.byref(ctx)))
            self._members[i] = Flusi(None, self.precision, _borrow=(self, ctx))
        return self._members[i]

    def init_fields(self):
        self._check(self._lib.flusi_ensemble_init_fields(self._e))

    def autotune(self, reps: int = 3):
        v = (_I * (6 * self.nbatches))()
        ms = (ctypes.c_double * self.nbatches)()
        self._check(self._lib.flusi_ensemble_autotune(self._e, int(reps), v, ms))
        return [(list(v[6 * b:6 * b + 6]), ms[b]) for b in range(self.nbatches)]

    def advance(self, nsteps: int, mode: int = ENSEMBLE_GRAPH):
        self._check(self._lib.flusi_ensemble_advance(self._e, int(nsteps), int(mode)))
