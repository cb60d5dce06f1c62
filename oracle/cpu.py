"""ctypes front end of the oracle (test infrastructure only).

Loads ``oracle/liboracle.so`` -- the CPU restatement of the reference's WGSL
solver in oracle.cpp -- for tests/, the smoke check and the benchmark's CPU
baseline.  The library under test never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "liboracle.so")
_lib = None


def set_threads(n: int) -> int:
    """OpenMP worker threads of the oracle's loops; returns the count in force."""
    return int(lib().oracle_set_threads(n))


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        f32p = C.POINTER(C.c_float)
        f64p = C.POINTER(C.c_double)
        L.oracle_set_threads.restype = C.c_int
        L.oracle_set_threads.argtypes = [C.c_int]
        L.oracle_create.restype = C.c_void_p
        L.oracle_create.argtypes = [C.c_int, f32p, f32p, C.c_int]
        L.oracle_destroy.argtypes = [C.c_void_p]
        L.oracle_advance.restype = C.c_int
        L.oracle_advance.argtypes = [C.c_void_p, C.c_longlong]
        L.oracle_record_steps.restype = C.c_int
        L.oracle_record_steps.argtypes = [C.c_void_p, C.c_longlong]
        L.oracle_remaining.restype = C.c_int
        L.oracle_remaining.argtypes = [C.c_void_p]
        L.oracle_time.restype = C.c_float
        L.oracle_time.argtypes = [C.c_void_p]
        L.oracle_get_state.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.oracle_set_state.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.oracle_diagnostics.argtypes = [C.c_void_p, f64p]
        L.oracle_diagnostics_naive_f32.argtypes = [C.c_void_p, f32p, C.POINTER(C.c_uint32)]
        L.oracle_vis_pack.argtypes = [C.c_void_p, f32p]
        L.oracle_prim_from_cons.argtypes = [C.c_int, f32p, f64p, C.c_double, f64p]
        L.oracle_cons_from_prim.argtypes = [f32p, f64p, f64p]
        L.oracle_wave_speeds.argtypes = [f32p, f64p, C.c_int, f64p]
        L.oracle_hll_flux.argtypes = [f32p, f64p, f64p, C.c_int, f64p]
        L.oracle_eos_flux.argtypes = [f64p, f64p, C.c_int, f64p]
        L.oracle_limit_slope.restype = C.c_double
        L.oracle_limit_slope.argtypes = [C.c_double, C.c_double]
        L.oracle_limit_slope_f32.restype = C.c_float
        L.oracle_limit_slope_f32.argtypes = [C.c_float, C.c_float]
        L.oracle_tanh_f64.restype = C.c_double
        L.oracle_tanh_f64.argtypes = [C.c_double]
        L.oracle_tanh_f32.restype = C.c_float
        L.oracle_tanh_f32.argtypes = [C.c_float]
        L.oracle_render.restype = None
        L.oracle_render.argtypes = [f32p, C.POINTER(C.c_uint16), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                    C.POINTER(C.c_uint8)]
        L.oracle_potential.restype = C.c_double
        L.oracle_potential.argtypes = [f32p, C.c_double]
        L.oracle_dpotential.restype = C.c_double
        L.oracle_dpotential.argtypes = [f32p, C.c_double]
        _lib = L
    return _lib


def _f32(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(C.POINTER(C.c_float))


def _f64(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(C.POINTER(C.c_double))


class OracleSim:
    """The reference solver on the CPU (precision 0: f32, 1: f64).

    params: the 24-word SimParams block (float32); events: (n, 5) float32
    rows (time, x, y, z, r0) sorted by time.
    """

    def __init__(self, precision: int, params: np.ndarray, events: np.ndarray):
        self.precision = precision
        self.real = np.float64 if precision == 1 else np.float32
        self._p, pp = _f32(params)
        self._e, ep = _f32(np.asarray(events, dtype=np.float32).reshape(-1, 5))
        u = self._p.view(np.uint32)
        self.grid = (int(u[0]), int(u[1]), int(u[2]))
        self.n = self.grid[0] * self.grid[1] * self.grid[2]
        self._h = lib().oracle_create(precision, pp, ep, len(self._e))

    def close(self):
        if self._h:
            lib().oracle_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def advance(self, n: int) -> int:
        return lib().oracle_advance(self._h, n)

    def record_steps(self, n: int) -> int:
        return lib().oracle_record_steps(self._h, n)

    def remaining(self) -> int:
        return lib().oracle_remaining(self._h)

    @property
    def time(self) -> float:
        return lib().oracle_time(self._h)

    def state(self):
        nx, ny, nz = self.grid
        field = np.empty((nz, ny, nx, 2), dtype=self.real)
        fluid = np.empty((nz, ny, nx, 4), dtype=self.real)
        lib().oracle_get_state(self._h, field.ctypes.data_as(C.c_void_p), fluid.ctypes.data_as(C.c_void_p))
        return field, fluid

    def set_state(self, field, fluid):
        field = np.ascontiguousarray(field, dtype=self.real)
        fluid = np.ascontiguousarray(fluid, dtype=self.real)
        lib().oracle_set_state(self._h, field.ctypes.data_as(C.c_void_p), fluid.ctypes.data_as(C.c_void_p))

    def diagnostics_sums(self) -> np.ndarray:
        """(sum e_total, sum broken, sum kinetic, sum E, max|v|), long double sums."""
        out = np.zeros(5, dtype=np.float64)
        lib().oracle_diagnostics(self._h, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def vis_channels(self) -> np.ndarray:
        """vis.wgsl's four channels per cell, narrowed to f32: shape (n, 4)."""
        out = np.zeros((self.n, 4), dtype=np.float32)
        lib().oracle_vis_pack(self._h, out.ctypes.data_as(C.POINTER(C.c_float)))
        return out

    def diagnostics_naive_f32(self):
        """reduce.wgsl's own f32 result: (sums[4], max_v_bits)."""
        sums = np.zeros(4, dtype=np.float32)
        bits = C.c_uint32()
        lib().oracle_diagnostics_naive_f32(self._h, sums.ctypes.data_as(C.POINTER(C.c_float)), C.byref(bits))
        return sums, int(bits.value)


def prim_from_cons(params, u, phi=0.0, precision=1):
    p, pp = _f32(params)
    uu, up = _f64(u)
    q = np.zeros(4)
    lib().oracle_prim_from_cons(precision, pp, up, phi, q.ctypes.data_as(C.POINTER(C.c_double)))
    return q


def cons_from_prim(params, q):
    p, pp = _f32(params)
    qq, qp = _f64(q)
    u = np.zeros(4)
    lib().oracle_cons_from_prim(pp, qp, u.ctypes.data_as(C.POINTER(C.c_double)))
    return u


def wave_speeds(params, q, axis):
    p, pp = _f32(params)
    qq, qp = _f64(q)
    o = np.zeros(2)
    lib().oracle_wave_speeds(pp, qp, axis, o.ctypes.data_as(C.POINTER(C.c_double)))
    return o


def hll_flux(params, ql, qr, axis):
    p, pp = _f32(params)
    a, ap = _f64(ql)
    b, bp = _f64(qr)
    f = np.zeros(4)
    lib().oracle_hll_flux(pp, ap, bp, axis, f.ctypes.data_as(C.POINTER(C.c_double)))
    return f


def eos_flux(q, u, axis):
    a, ap = _f64(q)
    b, bp = _f64(u)
    f = np.zeros(4)
    lib().oracle_eos_flux(ap, bp, axis, f.ctypes.data_as(C.POINTER(C.c_double)))
    return f


def limit_slope(dm, dp):
    return lib().oracle_limit_slope(dm, dp)


def limit_slope_f32(dm, dp):
    return lib().oracle_limit_slope_f32(dm, dp)


def tanh_f64(x):
    return lib().oracle_tanh_f64(x)


def tanh_f32(x):
    return lib().oracle_tanh_f32(x)


def potential(params, phi):
    p, pp = _f32(params)
    return lib().oracle_potential(pp, phi)


def dpotential(params, phi):
    p, pp = _f32(params)
    return lib().oracle_dpotential(pp, phi)


def eos_words(words, eos=0, eos_gamma=4.0 / 3.0, dof_ratio=1.0):
    """The SimParams words with the oracle's equation-of-state fields set
    (word 3: kind, words 22 and 23: gamma and the dof ratio)."""
    w = np.array(words, dtype=np.float32).copy()
    w.view(np.uint32)[3] = eos
    w[22] = eos_gamma if eos == 1 else 0.0
    w[23] = dof_ratio
    return w


def render(view_words, texels, width: int, height: int) -> np.ndarray:
    """volume.wgsl's fs_main on the CPU: view_words are the 44 floats of the
    view uniform, texels the half-precision channels, shape (nz, ny, nx, 4);
    returns RGBA8 pixels of shape (height, width, 4), row 0 at the top."""
    v, vp = _f32(np.asarray(view_words, dtype=np.float32).reshape(44))
    t = np.ascontiguousarray(texels)
    if t.dtype != np.uint16:
        t = np.ascontiguousarray(t.astype(np.float16)).view(np.uint16)
    nz, ny, nx, c = t.shape
    assert c == 4
    out = np.zeros((height, width, 4), dtype=np.uint8)
    lib().oracle_render(vp, t.ctypes.data_as(C.POINTER(C.c_uint16)), nx, ny, nz, width, height,
                        out.ctypes.data_as(C.POINTER(C.c_uint8)))
    return out
