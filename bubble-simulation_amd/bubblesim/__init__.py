"""Python binding of libbubblesim (the MI355X lattice solver) over its C ABI.

Mirrors the reference's Rust surface (src/physics.rs, src/gpu/sim.rs,
src/config.rs) with the same names and argument meaning: ``Model`` and its
derived quantities, ``make_schedule``, ``Simulation`` (``new`` / ``reset`` /
``set_model`` / ``record_steps`` / diagnostics / ``resize`` /
``nucleate_now``), plus ``Batch`` (many boxes in one launch per RK stage) and
``Server`` (continuous batching of sweep jobs).  Errors raise ``BubbleSimError``
carrying the library's message; the reference returns ``anyhow::Error`` with
the same texts.

The shared library must have been built (``make -C bubble-simulation_amd``);
nothing here falls back to another implementation.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libbubblesim.so")
CLI_PATH = os.path.join(os.path.dirname(_HERE), "bin", "bubble-sim")

BS_OK, BS_ERR_INVALID, BS_ERR_OOM, BS_ERR_DEVICE, BS_ERR_NONFINITE, BS_ERR_NOT_READY, BS_ERR_NOT_FOUND = range(7)
PRECISION_F32, PRECISION_F64 = 0, 1
MATH_FAST, MATH_EXACT = 0, 1
EOS_BAG, EOS_GAMMA_LAW = 0, 1
KERNEL_FUSED, KERNEL_TWO_PASS = 0, 1
SEED_CRITICAL, SEED_CELLS = 0, 1
NUC_SIMULTANEOUS, NUC_EXPONENTIAL, NUC_UNIFORM_STEPS = 0, 1, 2


class BubbleSimError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code
        self.msg = msg


# ---------------------------------------------------------------------------
#  ABI structs (include/bubblesim.h)
# ---------------------------------------------------------------------------
class CModel(C.Structure):
    _fields_ = [("lambda_", C.c_float), ("eps_ratio", C.c_float), ("alpha", C.c_float), ("eta", C.c_float),
                ("phi_b", C.c_float), ("dx", C.c_float), ("cfl", C.c_float), ("seed_kind", C.c_int32),
                ("seed_value", C.c_float)]


class CDerived(C.Structure):
    _fields_ = [(n, C.c_float) for n in (
        "eps", "m2", "delta", "wall_width", "surface_tension", "critical_radius", "e_ref", "t_ref", "a_rad", "dt",
        "seed_radius", "critical_to_wall_ratio", "estimated_wall_velocity", "sound_speed")]


class CEvent(C.Structure):
    _fields_ = [("time", C.c_float), ("pos", C.c_float * 3), ("radius", C.c_float)]


class CSimParams(C.Structure):
    _fields_ = [("nx", C.c_uint32), ("ny", C.c_uint32), ("nz", C.c_uint32), ("pad_u0", C.c_uint32)] + [
        (n, C.c_float) for n in ("dx", "inv_dx", "dt", "eta", "lambda_", "eps", "phi_b", "m2", "delta", "a_rad",
                                 "wall_width", "p_floor", "e_floor", "z_max", "e_ref", "t_ref", "time", "vis_gain",
                                 "pad0", "pad1")]


class CStageParams(C.Structure):
    _fields_ = [("a0", C.c_float), ("a1", C.c_float), ("dt_stage", C.c_float), ("pad", C.c_float)]


class CBubble(C.Structure):
    _fields_ = [("cx", C.c_float), ("cy", C.c_float), ("cz", C.c_float), ("r0", C.c_float)]


class CNucleationBatch(C.Structure):
    _fields_ = [("count", C.c_uint32), ("pad", C.c_uint32 * 3)]


class CDiagnosticsGpu(C.Structure):
    _fields_ = [("sums", C.c_float * 4), ("max_v_bits", C.c_uint32), ("pad", C.c_uint32 * 3)]


class CDiagnostics(C.Structure):
    _fields_ = [("mean_energy", C.c_double), ("broken_fraction", C.c_double), ("mean_kinetic", C.c_double),
                ("mean_fluid_energy", C.c_double), ("max_velocity", C.c_double), ("kinetic_fraction", C.c_double),
                ("cells", C.c_uint64), ("nonfinite", C.c_uint32), ("pad", C.c_uint32)]


class CSpec(C.Structure):
    _fields_ = [("model", CModel), ("grid", C.c_uint32 * 3), ("nucleation", C.c_int32), ("bubbles", C.c_uint32),
                ("nucleation_duration", C.c_float), ("seed", C.c_uint64), ("precision", C.c_int32),
                ("math", C.c_int32), ("kernel", C.c_int32), ("bubble_at_centre", C.c_int32),
                ("uniform_max_step", C.c_uint32), ("eos", C.c_int32), ("time_seed", C.c_uint64),
                ("eos_gamma", C.c_float), ("dof_ratio", C.c_float)]


class CHbmRates(C.Structure):
    _fields_ = [("read_gbps", C.c_double), ("write_gbps", C.c_double), ("copy_gbps", C.c_double)]


SERVER_NO_GRAPHS = 1


class CServerConfig(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_uint64), ("reserve_bytes", C.c_uint64), ("max_active", C.c_uint32),
                ("chunk_steps", C.c_uint32), ("precision", C.c_int32), ("math", C.c_int32),
                ("flags", C.c_uint32), ("pad", C.c_uint32)]


class CJob(C.Structure):
    _fields_ = [("spec", CSpec), ("steps", C.c_uint64), ("keep_state", C.c_int32), ("checksum", C.c_uint32),
                ("init_field", C.c_void_p), ("init_fluid", C.c_void_p)]


class CJobResult(C.Structure):
    _fields_ = [("job_id", C.c_uint64), ("steps", C.c_uint64), ("final_diag", CDiagnostics),
                ("baseline_diag", CDiagnostics), ("baseline_step", C.c_uint64), ("submit_s", C.c_double), ("admit_s", C.c_double), ("finish_s", C.c_double),
                ("latency_s", C.c_double), ("state_checksum", C.c_uint64), ("status", C.c_int32),
                ("pad", C.c_uint32)]


class CServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("admitted", C.c_uint64), ("completed", C.c_uint64),
                ("rejected", C.c_uint64), ("site_updates", C.c_uint64), ("ticks", C.c_uint64),
                ("bytes_in_use", C.c_uint64), ("budget_bytes", C.c_uint64), ("active", C.c_uint32),
                ("queued", C.c_uint32), ("busy_s", C.c_double), ("elapsed_s", C.c_double),
                ("failed", C.c_uint64), ("nonfinite", C.c_uint64)]


class CRenderSettings(C.Structure):
    _fields_ = [("field_mode", C.c_int32), ("colormap", C.c_int32), ("iso_level", C.c_float),
                ("wall_opacity", C.c_float), ("samples_per_cell", C.c_float), ("exposure", C.c_float),
                ("field_gain", C.c_float), ("absorption", C.c_float), ("show_iso", C.c_int32),
                ("show_volume", C.c_int32), ("show_box", C.c_int32), ("background", C.c_float),
                ("clip_axis", C.c_int32), ("clip_pos", C.c_float)]


class CRenderView(C.Structure):
    _fields_ = [("inv_view_proj", C.c_float * 16)] + [
        (n, C.c_float * 4) for n in ("cam_pos", "box_half", "grid", "render0", "render1", "render2", "render3")]


FIELD_PHI, FIELD_SPEED, FIELD_TEMPERATURE, FIELD_KINETIC = range(4)

_P = C.c_void_p
_u32p = C.POINTER(C.c_uint32)
_SIGS = {
    "bs_last_error": (C.c_char_p, []),
    "bs_version": (C.c_char_p, []),
    "bs_struct_size": (C.c_size_t, [C.c_int]),
    "bs_model_default": (None, [C.POINTER(CModel)]),
    "bs_model_validate": (C.c_int, [C.POINTER(CModel)]),
    "bs_model_derive": (C.c_int, [C.POINTER(CModel), C.POINTER(CDerived)]),
    "bs_make_schedule": (C.c_int, [C.c_int32, C.c_uint32, C.c_float, _u32p, C.c_float, C.c_uint64, C.c_uint32,
                                   C.c_uint64, C.c_float, C.POINTER(CEvent)]),
    "bs_pcg64mcg_u64": (None, [C.c_uint64, C.c_uint32, C.POINTER(C.c_uint64)]),
    "bs_pcg64mcg_f32": (None, [C.c_uint64, C.c_uint32, C.POINTER(C.c_float)]),
    "bs_mcg128_from_state": (None, [C.c_uint64, C.c_uint64, C.c_uint32, C.POINTER(C.c_uint64)]),
    "bs_sim_params_fill": (C.c_int, [C.POINTER(CModel), _u32p, C.c_float, C.POINTER(CSimParams)]),
    "bs_rk3_coefficients": (None, [C.POINTER(C.c_float)]),
    "bs_stage_slots": (None, [C.c_uint32, _u32p]),
    "bs_event_steps": (C.c_int, [C.POINTER(CEvent), C.c_uint32, C.c_float, C.POINTER(C.c_uint64)]),
    "bs_spec_validate": (C.c_int, [C.POINTER(CSpec)]),
    "bs_spec_default": (None, [C.POINTER(CSpec)]),
    "bs_parse_grid": (C.c_int, [C.c_char_p, _u32p]),
    "bs_bytes_per_cell": (C.c_uint64, [C.c_int32, C.c_int32]),
    "bs_lattice_bytes": (C.c_uint64, [_u32p, C.c_int32, C.c_int32]),
    "bs_max_cubic_size": (C.c_int, [C.c_uint64, C.c_int32, C.c_int32, _u32p, C.POINTER(C.c_int32)]),
    "bs_lattice_budget": (C.c_uint64, [C.c_uint64, C.c_uint64]),
    "bs_device_count": (C.c_int, []),
    "bs_device_memory": (C.c_int, [C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "bs_device_name": (C.c_int, [C.c_char_p, C.c_size_t]),
    "bs_hbm_probe": (C.c_int, [C.c_uint64, C.c_uint32, C.POINTER(CHbmRates)]),
    "bs_reduce_noalloc": (C.c_int, [C.POINTER(CSimParams), C.c_int32, _P, _P, _P, _P]),
    "bs_sim_create": (C.c_int, [C.POINTER(CSpec), C.POINTER(_P)]),
    "bs_sim_destroy": (None, [_P]),
    "bs_sim_reset": (C.c_int, [_P, C.POINTER(CEvent), C.c_uint32]),
    "bs_sim_set_model": (C.c_int, [_P, C.POINTER(CModel)]),
    "bs_sim_record_steps": (C.c_int, [_P, C.c_uint32]),
    "bs_sim_advance": (C.c_int, [_P, C.c_uint64]),
    "bs_sim_diagnostics": (C.c_int, [_P, C.POINTER(CDiagnostics)]),
    "bs_sim_diagnostics_raw": (C.c_int, [_P, C.POINTER(CDiagnosticsGpu)]),
    "bs_sim_request_diagnostics": (C.c_int, [_P]),
    "bs_sim_poll_diagnostics": (C.c_int, [_P, C.POINTER(CDiagnostics)]),
    "bs_sim_synchronize": (C.c_int, [_P]),
    "bs_sim_resize": (C.c_int, [_P, _u32p]),
    "bs_sim_nucleate_now": (C.c_int, [_P, C.POINTER(C.c_float)]),
    "bs_sim_cell_count": (C.c_uint64, [_P]),
    "bs_sim_remaining_bubbles": (C.c_uint32, [_P]),
    "bs_sim_time": (C.c_float, [_P]),
    "bs_sim_steps": (C.c_uint64, [_P]),
    "bs_sim_nucleated_this_frame": (C.c_int32, [_P]),
    "bs_sim_device_memory_bytes": (C.c_uint64, [_P]),
    "bs_sim_grid": (C.c_int, [_P, _u32p]),
    "bs_sim_model": (C.c_int, [_P, C.POINTER(CModel)]),
    "bs_sim_schedule": (C.c_int, [_P, C.POINTER(CEvent), C.c_uint32, _u32p]),
    "bs_sim_record_visualisation": (C.c_int, [_P]),
    "bs_sim_visualisation_buffer": (C.c_int, [_P, C.POINTER(_P), C.POINTER(C.c_uint64)]),
    "bs_sim_download_visualisation": (C.c_int, [_P, C.POINTER(C.c_uint16), C.c_uint64]),
    "bs_render_settings_default": (C.c_int, [C.c_int32, C.POINTER(CRenderSettings)]),
    "bs_render_view_make": (C.c_int, [_u32p, C.POINTER(CRenderSettings), C.c_uint32, C.c_uint32, C.c_uint32,
                                      C.POINTER(CRenderView)]),
    "bs_sim_render": (C.c_int, [_P, C.POINTER(CRenderView), C.c_uint32, C.c_uint32, C.POINTER(C.c_uint8)]),
    "bs_sim_screenshot": (C.c_int, [_P, C.POINTER(CRenderSettings), C.c_uint32, C.c_uint32, C.c_char_p]),
    "bs_png_write": (C.c_int, [C.c_char_p, C.POINTER(C.c_uint8), C.c_uint32, C.c_uint32]),
    "bs_sim_download": (C.c_int, [_P, _P, _P]),
    "bs_sim_upload": (C.c_int, [_P, _P, _P]),
    "bs_sim_advance_timed": (C.c_int, [_P, C.c_uint64, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "bs_batch_create": (C.c_int, [C.POINTER(CSpec), C.c_uint32, C.POINTER(_P)]),
    "bs_batch_destroy": (None, [_P]),
    "bs_batch_advance": (C.c_int, [_P, C.c_uint64, C.c_int32]),
    "bs_batch_synchronize": (C.c_int, [_P]),
    "bs_batch_size": (C.c_uint32, [_P]),
    "bs_batch_total_cells": (C.c_uint64, [_P]),
    "bs_batch_diagnostics": (C.c_int, [_P, C.c_uint32, C.POINTER(CDiagnostics)]),
    "bs_batch_download": (C.c_int, [_P, C.c_uint32, _P, _P]),
    "bs_batch_upload": (C.c_int, [_P, C.c_uint32, _P, _P]),
    "bs_batch_steps": (C.c_int, [_P, C.c_uint32, C.POINTER(C.c_uint64)]),
    "bs_batch_tile": (C.c_int, [_P, _u32p, _u32p]),
    "bs_batch_plan": (C.c_int, [_P, _u32p]),
    "bs_batch_graph_steps": (C.c_uint64, [_P]),
    "bs_batch_stamps_skipped": (C.c_uint64, [_P]),
    "bs_batch_streams": (C.c_uint32, [_P]),
    "bs_server_create": (C.c_int, [C.POINTER(CServerConfig), C.POINTER(_P)]),
    "bs_server_submit": (C.c_int, [_P, C.POINTER(CJob), C.POINTER(C.c_uint64)]),
    "bs_server_poll": (C.c_int, [_P, C.c_uint64, C.POINTER(CJobResult)]),
    "bs_server_wait": (C.c_int, [_P, C.c_uint64, C.c_double, C.POINTER(CJobResult)]),
    "bs_server_fetch_state": (C.c_int, [_P, C.c_uint64, _P, _P]),
    "bs_server_stats_get": (C.c_int, [_P, C.POINTER(CServerStats)]),
    "bs_server_destroy": (None, [_P]),
}

_lib = None


def lib():
    """The loaded libbubblesim; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise BubbleSimError(BS_ERR_DEVICE, f"{LIB_PATH} is missing: build it with `make -C bubble-simulation_amd`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc: int) -> None:
    if rc != BS_OK:
        raise BubbleSimError(rc, lib().bs_last_error().decode())


def _grid3(g) -> "C.Array":
    if isinstance(g, int):
        g = (g, g, g)
    return (C.c_uint32 * 3)(*[int(v) for v in g])


# ---------------------------------------------------------------------------
#  Physics (physics.rs)
# ---------------------------------------------------------------------------
@dataclass
class Model:
    """Model (physics.rs:44-68).  seed_cells=None means SeedSize::Critical."""
    lambda_: float = 0.5
    eps_ratio: float = 0.3
    alpha: float = 0.1
    eta: float = 0.3
    phi_b: float = 1.0
    dx: float = 1.0
    cfl: float = 0.2
    seed_factor: float = 1.15
    seed_cells: Optional[float] = None

    def to_c(self) -> CModel:
        return CModel(self.lambda_, self.eps_ratio, self.alpha, self.eta, self.phi_b, self.dx, self.cfl,
                      SEED_CELLS if self.seed_cells is not None else SEED_CRITICAL,
                      self.seed_cells if self.seed_cells is not None else self.seed_factor)

    def derived(self) -> CDerived:
        d = CDerived()
        _check(lib().bs_model_derive(C.byref(self.to_c()), C.byref(d)))
        return d

    def validate(self) -> None:
        _check(lib().bs_model_validate(C.byref(self.to_c())))

    def __getattr__(self, name):
        # eps(), m2(), ... as on the Rust Model
        if name in dict(CDerived._fields_):
            return lambda: float(getattr(self.derived(), name))
        raise AttributeError(name)


def make_schedule(mode: int, count: int, radius: float, grid, duration: float, seed: int,
                  uniform_max_step: int = 200, time_seed: int = 0x5EED, dt: float = 0.2) -> np.ndarray:
    """make_schedule (physics.rs:248-296) -> structured array (time, x, y, z, radius)."""
    out = (CEvent * max(count, 1))()
    _check(lib().bs_make_schedule(mode, count, C.c_float(radius), _grid3(grid), C.c_float(duration), seed,
                                  uniform_max_step, time_seed, C.c_float(dt), out))
    return events_to_array(out, count)


def events_to_array(evs, n) -> np.ndarray:
    a = np.zeros((n, 5), dtype=np.float32)
    for i in range(n):
        e = evs[i]
        a[i] = (e.time, e.pos[0], e.pos[1], e.pos[2], e.radius)
    return a


def array_to_events(a: np.ndarray):
    a = np.asarray(a, dtype=np.float32).reshape(-1, 5)
    evs = (CEvent * max(len(a), 1))()
    for i, r in enumerate(a):
        evs[i].time = r[0]
        evs[i].pos[0], evs[i].pos[1], evs[i].pos[2] = r[1], r[2], r[3]
        evs[i].radius = r[4]
    return evs, len(a)


def sim_params(model: Model, grid, time: float = 0.0) -> CSimParams:
    p = CSimParams()
    _check(lib().bs_sim_params_fill(C.byref(model.to_c()), _grid3(grid), C.c_float(time), C.byref(p)))
    return p


def sim_params_words(model: Model, grid, time: float = 0.0) -> np.ndarray:
    """The 96-byte SimParams block as 24 float32 words."""
    p = sim_params(model, grid, time)
    return np.frombuffer(bytes(p), dtype=np.float32).copy()


def pcg64mcg_u64(seed: int, n: int) -> np.ndarray:
    out = (C.c_uint64 * n)()
    lib().bs_pcg64mcg_u64(seed, n, out)
    return np.array(out[:], dtype=np.uint64)


def pcg64mcg_f32(seed: int, n: int) -> np.ndarray:
    out = (C.c_float * n)()
    lib().bs_pcg64mcg_f32(seed, n, out)
    return np.array(out[:], dtype=np.float32)


def mcg128_from_state(state: int, n: int) -> np.ndarray:
    out = (C.c_uint64 * n)()
    lib().bs_mcg128_from_state(state & (2**64 - 1), state >> 64, n, out)
    return np.array(out[:], dtype=np.uint64)


def event_steps(events: np.ndarray, dt: float) -> np.ndarray:
    evs, n = array_to_events(events)
    out = (C.c_uint64 * max(n, 1))()
    _check(lib().bs_event_steps(evs, n, C.c_float(dt), out))
    return np.array(out[:n], dtype=np.uint64)


# ---------------------------------------------------------------------------
#  Spec / Config (config.rs)
# ---------------------------------------------------------------------------
def default_spec() -> CSpec:
    s = CSpec()
    lib().bs_spec_default(C.byref(s))
    return s


def make_spec(grid=192, model: Optional[Model] = None, nucleation: int = NUC_EXPONENTIAL, bubbles: int = 12,
              duration: float = 60.0, seed: int = 1, precision: int = PRECISION_F32, math: int = MATH_FAST,
              kernel: int = KERNEL_FUSED, bubble_at_centre: bool = False, uniform_max_step: int = 200,
              time_seed: int = 0x5EED, eos: int = 0, eos_gamma: float = 4.0 / 3.0, dof_ratio: float = 1.0) -> CSpec:
    s = default_spec()
    s.model = (model or Model()).to_c()
    g = _grid3(grid)
    s.grid[0], s.grid[1], s.grid[2] = g[0], g[1], g[2]
    s.nucleation = nucleation
    s.bubbles = bubbles
    s.nucleation_duration = duration
    s.seed = seed
    s.precision = precision
    s.math = MATH_EXACT if precision == PRECISION_F64 else math
    s.kernel = kernel
    s.bubble_at_centre = 1 if bubble_at_centre else 0
    s.uniform_max_step = uniform_max_step
    s.time_seed = time_seed
    s.eos = eos
    s.eos_gamma = eos_gamma
    s.dof_ratio = dof_ratio
    return s


def validate_spec(spec: CSpec) -> None:
    _check(lib().bs_spec_validate(C.byref(spec)))


def parse_grid(s: str):
    g = (C.c_uint32 * 3)()
    _check(lib().bs_parse_grid(s.encode(), g))
    return tuple(g)


def bytes_per_cell(precision: int = PRECISION_F32, kernel: int = KERNEL_FUSED) -> int:
    return int(lib().bs_bytes_per_cell(precision, kernel))


def lattice_bytes(grid, precision: int = PRECISION_F32, kernel: int = KERNEL_FUSED) -> int:
    return int(lib().bs_lattice_bytes(_grid3(grid), precision, kernel))


def max_cubic_size(budget: int, precision: int = PRECISION_F32, kernel: int = KERNEL_FUSED):
    side = C.c_uint32()
    lim = C.c_int32()
    _check(lib().bs_max_cubic_size(budget, precision, kernel, C.byref(side), C.byref(lim)))
    return int(side.value), bool(lim.value)


def stage_slots(cur: int):
    out = (C.c_uint32 * 6)()
    lib().bs_stage_slots(cur, out)
    return [(out[0], out[1]), (out[2], out[3]), (out[4], out[5])]


def rk3_coefficients():
    out = (C.c_float * 6)()
    lib().bs_rk3_coefficients(out)
    return [(out[0], out[1]), (out[2], out[3]), (out[4], out[5])]


def device_count() -> int:
    return int(lib().bs_device_count())


def device_memory():
    """(free, total) bytes of the current device (hipMemGetInfo)."""
    f, t = C.c_uint64(), C.c_uint64()
    _check(lib().bs_device_memory(C.byref(f), C.byref(t)))
    return int(f.value), int(t.value)


def device_name() -> str:
    buf = C.create_string_buffer(256)
    _check(lib().bs_device_name(buf, 256))
    return buf.value.decode()


def hbm_probe(nbytes: int = 2 << 30, repeats: int = 5) -> dict:
    """Achievable HBM read / write / copy bandwidth of the current device,
    GB/s (bs_hbm_probe)."""
    r = CHbmRates()
    _check(lib().bs_hbm_probe(nbytes, repeats, C.byref(r)))
    return {"read_GBps": r.read_gbps, "write_GBps": r.write_gbps, "copy_GBps": r.copy_gbps}


def diag_dict(d: CDiagnostics) -> dict:
    return {k: getattr(d, k) for k, _ in CDiagnostics._fields_ if k != "pad"}


# ---------------------------------------------------------------------------
#  Simulation (sim.rs:180-1026)
# ---------------------------------------------------------------------------
def render_settings(field_mode: int = FIELD_KINETIC) -> CRenderSettings:
    """RenderSettings::with_field (render.rs:96-105)."""
    st = CRenderSettings()
    _check(lib().bs_render_settings_default(field_mode, C.byref(st)))
    return st


def render_view(grid, settings: Optional[CRenderSettings] = None, width: int = 1280, height: int = 960,
                frame: int = 1) -> CRenderView:
    """The view uniform of the default orbit camera framed on the box (render.rs:288-334)."""
    v = CRenderView()
    settings = settings if settings is not None else render_settings()
    _check(lib().bs_render_view_make(_grid3(grid), C.byref(settings), width, height, frame, C.byref(v)))
    return v


def png_write(path: str, rgba: np.ndarray):
    """The screenshot's PNG writer: rgba of shape (height, width, 4), uint8."""
    rgba = np.ascontiguousarray(rgba, dtype=np.uint8)
    h, w, c = rgba.shape
    assert c == 4
    _check(lib().bs_png_write(os.fsencode(path), rgba.ctypes.data_as(C.POINTER(C.c_uint8)), w, h))


class Simulation:
    """One lattice box on the GPU (Simulation, sim.rs:180)."""

    def __init__(self, spec: CSpec):
        self._h = C.c_void_p()
        self.spec = spec
        _check(lib().bs_sim_create(C.byref(spec), C.byref(self._h)))
        self.real = np.float64 if spec.precision == PRECISION_F64 else np.float32

    @classmethod
    def new(cls, spec: CSpec) -> "Simulation":
        return cls(spec)

    def close(self):
        if self._h:
            lib().bs_sim_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # lifecycle
    def reset(self, schedule: Optional[np.ndarray] = None):
        if schedule is None:
            _check(lib().bs_sim_reset(self._h, None, 0))
        else:
            evs, n = array_to_events(schedule)
            _check(lib().bs_sim_reset(self._h, evs, n))

    def set_model(self, model: Model):
        _check(lib().bs_sim_set_model(self._h, C.byref(model.to_c())))

    def record_steps(self, n: int):
        _check(lib().bs_sim_record_steps(self._h, n))

    def advance(self, n: int):
        _check(lib().bs_sim_advance(self._h, n))

    def advance_timed(self, n: int):
        k = C.c_double()
        w = C.c_double()
        _check(lib().bs_sim_advance_timed(self._h, n, C.byref(k), C.byref(w)))
        return k.value, w.value

    def synchronize(self):
        _check(lib().bs_sim_synchronize(self._h))

    def diagnostics(self) -> CDiagnostics:
        d = CDiagnostics()
        _check(lib().bs_sim_diagnostics(self._h, C.byref(d)))
        return d

    def diagnostics_raw(self) -> CDiagnosticsGpu:
        d = CDiagnosticsGpu()
        _check(lib().bs_sim_diagnostics_raw(self._h, C.byref(d)))
        return d

    def request_diagnostics(self):
        _check(lib().bs_sim_request_diagnostics(self._h))

    def poll_diagnostics(self) -> Optional[CDiagnostics]:
        d = CDiagnostics()
        rc = lib().bs_sim_poll_diagnostics(self._h, C.byref(d))
        if rc == BS_ERR_NOT_READY:
            return None
        _check(rc)
        return d

    def resize(self, grid):
        _check(lib().bs_sim_resize(self._h, _grid3(grid)))
        g = self.grid
        self.spec.grid[0], self.spec.grid[1], self.spec.grid[2] = g

    def nucleate_now(self, pos):
        p = (C.c_float * 3)(*pos)
        _check(lib().bs_sim_nucleate_now(self._h, p))

    # getters
    def cell_count(self) -> int:
        return int(lib().bs_sim_cell_count(self._h))

    def remaining_bubbles(self) -> int:
        return int(lib().bs_sim_remaining_bubbles(self._h))

    @property
    def time(self) -> float:
        return float(lib().bs_sim_time(self._h))

    @property
    def steps(self) -> int:
        return int(lib().bs_sim_steps(self._h))

    @property
    def nucleated_this_frame(self) -> bool:
        return bool(lib().bs_sim_nucleated_this_frame(self._h))

    def device_memory_bytes(self) -> int:
        return int(lib().bs_sim_device_memory_bytes(self._h))

    @property
    def grid(self):
        g = (C.c_uint32 * 3)()
        _check(lib().bs_sim_grid(self._h, g))
        return tuple(g)

    def schedule(self) -> np.ndarray:
        n = C.c_uint32()
        _check(lib().bs_sim_schedule(self._h, None, 0, C.byref(n)))
        evs = (CEvent * max(n.value, 1))()
        _check(lib().bs_sim_schedule(self._h, evs, n.value, C.byref(n)))
        return events_to_array(evs, n.value)

    # visualisation texels (vis.wgsl)
    def record_visualisation(self):
        _check(lib().bs_sim_record_visualisation(self._h))

    def visualisation_buffer(self):
        """(device address, bytes) of the texel buffer."""
        p, n = C.c_void_p(), C.c_uint64()
        _check(lib().bs_sim_visualisation_buffer(self._h, C.byref(p), C.byref(n)))
        return int(p.value or 0), int(n.value)

    def download_visualisation(self) -> np.ndarray:
        """Texels as float16, shape (nz, ny, nx, 4): phi/phi_b, |v|, T/T_ref - 1, (E+p)v^2/e_ref."""
        nx, ny, nz = self.grid
        out = np.zeros((nz, ny, nx, 4), dtype=np.uint16)
        _check(lib().bs_sim_download_visualisation(self._h, out.ctypes.data_as(C.POINTER(C.c_uint16)), out.size))
        return out.view(np.float16)

    # offscreen volume rendering (volume.wgsl, capture.rs)
    def render(self, view: "CRenderView", width: int, height: int) -> np.ndarray:
        """RGBA8 pixels of the last recorded visualisation, shape (height, width, 4), row 0 at the top."""
        out = np.zeros((height, width, 4), dtype=np.uint8)
        _check(lib().bs_sim_render(self._h, C.byref(view), width, height, out.ctypes.data_as(C.POINTER(C.c_uint8))))
        return out

    def screenshot(self, path: str, settings: Optional["CRenderSettings"] = None, size=(1280, 960)):
        """capture::render_to_png: record the visualisation, render frame 1, write an RGBA PNG."""
        settings = settings if settings is not None else render_settings()
        _check(lib().bs_sim_screenshot(self._h, C.byref(settings), size[0], size[1], os.fsencode(path)))

    # state (debug dump / load)
    def download(self):
        nx, ny, nz = self.grid
        field = np.empty((nz, ny, nx, 2), dtype=self.real)
        fluid = np.empty((nz, ny, nx, 4), dtype=self.real)
        _check(lib().bs_sim_download(self._h, field.ctypes.data_as(C.c_void_p), fluid.ctypes.data_as(C.c_void_p)))
        return field, fluid

    def upload(self, field: np.ndarray, fluid: np.ndarray):
        field = np.ascontiguousarray(field, dtype=self.real)
        fluid = np.ascontiguousarray(fluid, dtype=self.real)
        n = self.cell_count()
        assert field.size == 2 * n and fluid.size == 4 * n
        _check(lib().bs_sim_upload(self._h, field.ctypes.data_as(C.c_void_p), fluid.ctypes.data_as(C.c_void_p)))


# ---------------------------------------------------------------------------
#  Batch and Server
# ---------------------------------------------------------------------------
class Batch:
    """Many independent boxes advanced by one fused launch per RK stage."""

    def __init__(self, specs: Sequence[CSpec]):
        arr = (CSpec * len(specs))(*specs)
        self._h = C.c_void_p()
        self.specs = list(specs)
        _check(lib().bs_batch_create(arr, len(specs), C.byref(self._h)))
        self.real = np.float64 if specs[0].precision == PRECISION_F64 else np.float32

    def close(self):
        if self._h:
            lib().bs_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def advance(self, n: int, use_graph: bool = True):
        _check(lib().bs_batch_advance(self._h, n, 1 if use_graph else 0))

    def synchronize(self):
        _check(lib().bs_batch_synchronize(self._h))

    def size(self) -> int:
        return int(lib().bs_batch_size(self._h))

    def total_cells(self) -> int:
        return int(lib().bs_batch_total_cells(self._h))

    def diagnostics(self, box: int) -> CDiagnostics:
        d = CDiagnostics()
        _check(lib().bs_batch_diagnostics(self._h, box, C.byref(d)))
        return d

    def steps(self, box: int) -> int:
        v = C.c_uint64()
        _check(lib().bs_batch_steps(self._h, box, C.byref(v)))
        return int(v.value)

    def tile(self):
        """Tile shape of RK stages 2 and 3."""
        a, b = C.c_uint32(), C.c_uint32()
        _check(lib().bs_batch_tile(self._h, C.byref(a), C.byref(b)))
        return int(a.value), int(b.value)

    def plan(self):
        """{"stage1": (tx, ty, zc), "stages23": (tx, ty, zc)} chosen by the autotuner."""
        out = (C.c_uint32 * 6)()
        _check(lib().bs_batch_plan(self._h, out))
        return {"stage1": tuple(out[0:3]), "stages23": tuple(out[3:6])}

    def graph_steps(self) -> int:
        """Steps advanced so far through replay of a captured hipGraph."""
        return int(lib().bs_batch_graph_steps(self._h))

    def stamps_skipped(self) -> int:
        """Steps advanced without the nucleation stamp launch (no event due)."""
        return int(lib().bs_batch_stamps_skipped(self._h))

    def streams(self) -> int:
        """Stream groups the boxes are spread over."""
        return int(lib().bs_batch_streams(self._h))

    def download(self, box: int):
        s = self.specs[box]
        nx, ny, nz = s.grid[0], s.grid[1], s.grid[2]
        field = np.empty((nz, ny, nx, 2), dtype=self.real)
        fluid = np.empty((nz, ny, nx, 4), dtype=self.real)
        _check(lib().bs_batch_download(self._h, box, field.ctypes.data_as(C.c_void_p),
                                       fluid.ctypes.data_as(C.c_void_p)))
        return field, fluid

    def upload(self, box: int, field: np.ndarray, fluid: np.ndarray):
        s = self.specs[box]
        n = int(s.grid[0]) * int(s.grid[1]) * int(s.grid[2])
        field = np.ascontiguousarray(field, dtype=self.real)
        fluid = np.ascontiguousarray(fluid, dtype=self.real)
        assert field.size == 2 * n and fluid.size == 4 * n
        _check(lib().bs_batch_upload(self._h, box, field.ctypes.data_as(C.c_void_p),
                                     fluid.ctypes.data_as(C.c_void_p)))


class Server:
    """Continuous batching of sweep jobs under an HBM budget."""

    def __init__(self, budget_bytes: int = 0, reserve_bytes: int = 16 << 30, max_active: int = 0,
                 chunk_steps: int = 10, precision: int = PRECISION_F32, math: int = MATH_FAST, flags: int = 0):
        cfg = CServerConfig(budget_bytes, reserve_bytes, max_active, chunk_steps, precision, math, flags, 0)
        self._h = C.c_void_p()
        _check(lib().bs_server_create(C.byref(cfg), C.byref(self._h)))
        self.precision = precision

    def close(self):
        if self._h:
            lib().bs_server_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, spec: CSpec, steps: int, keep_state: bool = False, checksum: bool = False,
               init_state=None) -> int:
        """Queue a job; init_state = (field, fluid) arrays starts it from that
        state instead of the initialised lattice (copied here)."""
        job = CJob(spec, steps, 1 if keep_state else 0, 1 if checksum else 0, None, None)
        if init_state is not None:
            real = np.float64 if self.precision == PRECISION_F64 else np.float32
            f = np.ascontiguousarray(init_state[0], dtype=real)
            u = np.ascontiguousarray(init_state[1], dtype=real)
            n = int(spec.grid[0]) * int(spec.grid[1]) * int(spec.grid[2])
            assert f.size == 2 * n and u.size == 4 * n
            job.init_field = f.ctypes.data
            job.init_fluid = u.ctypes.data
        jid = C.c_uint64()
        _check(lib().bs_server_submit(self._h, C.byref(job), C.byref(jid)))
        return int(jid.value)

    def poll(self, job_id: int) -> Optional[CJobResult]:
        r = CJobResult()
        rc = lib().bs_server_poll(self._h, job_id, C.byref(r))
        if rc == BS_ERR_NOT_READY:
            return None
        _check(rc)
        return r

    def wait(self, job_id: int, timeout_s: float = 600.0) -> CJobResult:
        r = CJobResult()
        rc = lib().bs_server_wait(self._h, job_id, timeout_s, C.byref(r))
        if rc == BS_ERR_NOT_READY:
            raise TimeoutError(f"job {job_id} not done after {timeout_s} s")
        _check(rc)
        return r

    def fetch_state(self, job_id: int, spec: CSpec):
        real = np.float64 if self.precision == PRECISION_F64 else np.float32
        nx, ny, nz = spec.grid[0], spec.grid[1], spec.grid[2]
        field = np.empty((nz, ny, nx, 2), dtype=real)
        fluid = np.empty((nz, ny, nx, 4), dtype=real)
        _check(lib().bs_server_fetch_state(self._h, job_id, field.ctypes.data_as(C.c_void_p),
                                           fluid.ctypes.data_as(C.c_void_p)))
        return field, fluid

    def stats(self) -> CServerStats:
        s = CServerStats()
        _check(lib().bs_server_stats_get(self._h, C.byref(s)))
        return s
