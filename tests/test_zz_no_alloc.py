"""bs_reduce_noalloc, the allocation-free diagnostics reduction, through the
C ABI: on a lattice the caller holds in device memory it gives the sums of
bs_sim_diagnostics on the same state, bit for bit, on two calls in a row
(the second with a smaller lattice, in the other precision, reusing the
library's fixed scratch array)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class _Hip:
    """The few HIP runtime calls the test needs to hold state in device
    memory of its own (the entry point under test allocates nothing)."""

    def __init__(self):
        self.L = C.CDLL("libamdhip64.so")

    def check(self, rc, what):
        assert rc == 0, f"{what} failed: {rc}"

    def malloc(self, n):
        p = C.c_void_p()
        self.check(self.L.hipMalloc(C.byref(p), C.c_size_t(n)), "hipMalloc")
        return p

    def h2d(self, dst, a):
        a = np.ascontiguousarray(a)
        self.check(self.L.hipMemcpy(dst, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1), "hipMemcpy H2D")

    def d2h(self, src, nbytes):
        out = np.zeros(nbytes // 8, dtype=np.float64)
        self.check(self.L.hipMemcpy(out.ctypes.data_as(C.c_void_p), src, C.c_size_t(nbytes), 2), "hipMemcpy D2H")
        return out

    def sync(self):
        self.check(self.L.hipDeviceSynchronize(), "hipDeviceSynchronize")

    def free(self, p):
        self.L.hipFree(p)


def _case(bs, hip, grid, precision, out, steps):
    model = bs.Model()
    r = model.seed_radius()
    sched = np.array([(0.0, grid[0] / 2, grid[1] / 2, grid[2] / 2, r), (0.4, 2.0, 3.0, 5.0, r)], dtype=np.float32)
    sim = bs.Simulation(bs.make_spec(grid=grid, model=model, precision=precision, nucleation=bs.NUC_SIMULTANEOUS,
                                     bubbles=1))
    sim.reset(sched)
    sim.advance(steps)
    want = sim.diagnostics()
    field, fluid = sim.download()
    sim.close()
    df, du = hip.malloc(field.nbytes), hip.malloc(fluid.nbytes)
    try:
        hip.h2d(df, field)
        hip.h2d(du, fluid)
        params = bs.sim_params(model, grid)
        rc = bs.lib().bs_reduce_noalloc(C.byref(params), precision, df, du, out, None)
        assert rc == bs.BS_OK, bs.lib().bs_last_error().decode()
        hip.sync()
        got = hip.d2h(out, 48)
    finally:
        hip.free(df)
        hip.free(du)
    n = float(np.prod(grid))
    print(f"grid {grid} precision {precision}: sums {got[:4]}, max|v| {got[4]}", flush=True)
    # bs_diagnostics' means are the sums over the cell count in double
    assert got[0] / n == want.mean_energy
    assert got[1] / n == want.broken_fraction
    assert got[2] / n == want.mean_kinetic
    assert got[3] / n == want.mean_fluid_energy
    assert got[4] == want.max_velocity
    assert (got[5] != 0) == (want.nonfinite != 0) and want.nonfinite == 0


def test_noalloc_reduction_equals_sim_diagnostics(bs, gpu_available):
    hip = _Hip()
    out = hip.malloc(48)
    try:
        _case(bs, hip, (48, 40, 32), bs.PRECISION_F32, out, 6)
        _case(bs, hip, (24, 16, 16), bs.PRECISION_F64, out, 4)  # smaller, the other precision
    finally:
        hip.free(out)


def test_noalloc_rejects_lattices_outside_the_limits(bs, gpu_available):
    params = bs.sim_params(bs.Model(), (8, 8, 8))
    params.nz = 7
    dummy = C.c_void_p(1)  # never dereferenced: rejected before any launch
    assert bs.lib().bs_reduce_noalloc(C.byref(params), bs.PRECISION_F32, dummy, dummy, dummy, None) == bs.BS_ERR_INVALID
    assert b"below 8" in bs.lib().bs_last_error()
    params.nz = 8
    params.nx = params.ny = 1 << 14  # a 2^28-cell plane
    assert bs.lib().bs_reduce_noalloc(C.byref(params), bs.PRECISION_F32, dummy, dummy, dummy, None) == bs.BS_ERR_INVALID
    assert b"too large" in bs.lib().bs_last_error()
