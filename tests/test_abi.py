"""C ABI: the library loads, exports every declared symbol, and its byte
layouts and host helpers keep the reference's invariants (sim.rs:1053-1117)."""
import ctypes as C
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "bubblesim.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w \*]*?\b(bs_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_every_declared_symbol_is_exported(bs):
    names = declared_functions()
    assert len(names) > 60
    L = bs.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the binding's prototype table covers them all
    assert set(names) <= set(bs._SIGS), sorted(set(names) - set(bs._SIGS))


def test_exported_symbols_via_nm(bs):
    out = subprocess.run(["nm", "-D", "--defined-only", bs.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (bs_\w+)", out))
    assert set(declared_functions()) <= exported


def test_parameter_blocks_have_the_expected_size(bs):
    # parameter_blocks_have_the_expected_size (sim.rs:1070-1078)
    L = bs.lib()
    assert L.bs_struct_size(0) == 96 == C.sizeof(bs.CSimParams)
    assert L.bs_struct_size(1) == 16 == C.sizeof(bs.CStageParams)
    assert L.bs_struct_size(2) == 16 == C.sizeof(bs.CBubble)
    assert L.bs_struct_size(3) == 16 == C.sizeof(bs.CNucleationBatch)
    assert L.bs_struct_size(4) == 32 == C.sizeof(bs.CDiagnosticsGpu)
    assert L.bs_struct_size(5) == C.sizeof(bs.CDiagnostics)
    assert L.bs_struct_size(6) == C.sizeof(bs.CSpec)
    assert L.bs_struct_size(7) == C.sizeof(bs.CModel)
    assert L.bs_struct_size(8) == C.sizeof(bs.CEvent) == 20
    assert L.bs_struct_size(9) == C.sizeof(bs.CDerived)
    assert L.bs_struct_size(10) == C.sizeof(bs.CServerConfig) == 40
    assert L.bs_struct_size(11) == C.sizeof(bs.CJob)
    assert L.bs_struct_size(12) == C.sizeof(bs.CJobResult)
    assert L.bs_struct_size(13) == C.sizeof(bs.CServerStats)


def test_stage_slots_never_alias_output_with_input(bs):
    # stage_slots_never_alias_output_with_input (sim.rs:1057-1068)
    for cur in range(3):
        slots = bs.stage_slots(cur)
        for stage, (i, o) in enumerate(slots):
            assert i != o, f"stage {stage} of cur {cur} writes its input"
            assert cur != o, f"stage {stage} of cur {cur} overwrites U^n"
        assert slots[2][1] == (cur + 1) % 3
    assert bs.stage_slots(0) == [(0, 1), (1, 2), (2, 1)]


def test_rk3_coefficients_are_consistent(bs):
    # rk3_coefficients_are_consistent (sim.rs:1106-1115)
    rk = bs.rk3_coefficients()
    assert rk == [(0.0, 1.0), (0.75, 0.25), (C.c_float(1 / 3).value, C.c_float(2 / 3).value)]
    for a0, a1 in rk:
        assert abs(a0 + a1 - 1.0) < 1e-6
        assert a0 >= 0.0 and a1 > 0.0


def test_lattice_bytes(bs):
    # lattice_bytes_is_96_per_cell (sim.rs:1080-1083): the reference holds
    # 3 x 24 B of state, a 16 B primitive cache and an 8 B vis texture; the
    # fused kernel keeps only the 72 B of state, the two-pass path adds the
    # 16 B primitive cache (the reference's 88 B without the texture)
    assert bs.lattice_bytes((100, 10, 10)) == 10_000 * 72
    assert bs.lattice_bytes((100, 10, 10), bs.PRECISION_F32, bs.KERNEL_TWO_PASS) == 10_000 * 88
    assert bs.lattice_bytes((100, 10, 10), bs.PRECISION_F64) == 10_000 * 144
    assert bs.bytes_per_cell(bs.PRECISION_F64, bs.KERNEL_TWO_PASS) == 176


def test_cap_logic(bs):
    # cap_reports_which_limit_binds / cap_never_goes_below_the_stencil_minimum
    # (sim.rs:1085-1104), with memory as the only ceiling on HIP
    side, lim = bs.max_cubic_size(1_000_000_000)
    assert lim and bs.lattice_bytes((side,) * 3) <= 1_000_000_000
    assert bs.lattice_bytes((side + 1,) * 3) > 1_000_000_000
    assert bs.max_cubic_size(1024) == (8, True)
    side, lim = bs.max_cubic_size(1 << 62)
    assert side == 2048 and not lim
    # lattice_budget (vram.rs:67-69)
    L = bs.lib()
    assert L.bs_lattice_budget(1000, 0) == 850
    assert L.bs_lattice_budget(2**64 - 1, 10) == int(float(2**64 - 1) * 0.85)


def test_sim_params_block(bs):
    # gpu_params (sim.rs:613-644)
    m = bs.Model()
    p = bs.sim_params(m, (64, 48, 32), 1.5)
    d = m.derived()
    assert (p.nx, p.ny, p.nz) == (64, 48, 32)
    assert p.dx == 1.0 and p.inv_dx == 1.0 and p.dt == d.dt and p.eta == C.c_float(0.3).value
    assert p.eps == d.eps and p.m2 == d.m2 and p.delta == d.delta and p.a_rad == d.a_rad
    assert p.wall_width == d.wall_width
    assert p.p_floor == C.c_float(C.c_float(1e-7).value * d.e_ref).value == p.e_floor
    assert p.z_max == C.c_float(0.9999).value and p.e_ref == d.e_ref and p.t_ref == 1.0
    assert p.time == 1.5 and p.vis_gain == 1.0


def test_errors_carry_messages(bs):
    m = bs.Model(cfl=0.6)
    with pytest.raises(bs.BubbleSimError) as e:
        m.validate()
    assert e.value.code == bs.BS_ERR_INVALID
    assert "cfl must lie in (0, 0.5]" in e.value.msg


def test_no_device_fails_loudly_without_gpu(bs):
    if bs.device_count() > 0:
        pytest.skip("a device is present")
    with pytest.raises(bs.BubbleSimError) as e:
        bs.Simulation(bs.make_spec(grid=16))
    assert e.value.code == bs.BS_ERR_DEVICE


def test_hbm_probe_rejects_bad_arguments_before_touching_the_device(bs):
    # argument checks run on the host, so this holds without a GPU
    r = bs.CHbmRates()
    L = bs.lib()
    assert L.bs_hbm_probe(1024, 3, C.byref(r)) == bs.BS_ERR_INVALID
    assert L.bs_hbm_probe(1 << 30, 0, C.byref(r)) == bs.BS_ERR_INVALID
    assert L.bs_hbm_probe(1 << 30, 3, None) == bs.BS_ERR_INVALID
