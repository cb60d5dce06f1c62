"""The bubble-sim front end: the reference's flags (config.rs:13-115), its
validation messages (config.rs:151-178, physics.rs:191-217) and the headless
report (headless.rs:23-116)."""
import os
import re
import subprocess

import pytest


def cli(bs, *args, timeout=300):
    return subprocess.run([bs.CLI_PATH, *args], capture_output=True, text=True, timeout=timeout)


def test_help_lists_reference_flags(bs):
    r = cli(bs, "--help")
    assert r.returncode == 0
    for flag in ("--grid", "--lambda", "--eps-ratio", "--alpha", "--eta", "--cfl", "--seed-factor", "--seed-radius",
                 "--bubbles", "--nucleation", "--nucleation-duration", "--seed", "--steps-per-frame", "--field",
                 "--no-vsync", "--headless", "--report-every", "--screenshot", "--screenshot-size"):
        assert flag in r.stdout


@pytest.mark.parametrize("args,msg", [
    (["--lambda", "50", "--headless", "1"], "cells thick, which the lattice cannot resolve"),
    (["--cfl", "0.6", "--headless", "1"], "cfl must lie in (0, 0.5]"),
    (["--eps-ratio", "1.0", "--headless", "1"], "eps_ratio must lie in (0, 1)"),
    (["--bubbles", "0", "--headless", "1"], "--bubbles must be at least 1"),
    (["--grid", "64,7,64", "--headless", "1"], "grid axis 1 is 7"),
    (["--nucleation-duration", "0", "--headless", "1"], "--nucleation-duration must be positive"),
    (["--steps-per-frame", "0", "--headless", "1"], "--steps-per-frame must be at least 1"),
])
def test_validation_errors(bs, args, msg):
    r = cli(bs, *args)
    assert r.returncode == 1
    assert msg in r.stderr


@pytest.mark.parametrize("args", [["--grid", "1,2"], ["--grid", "abc"], ["--nucleation", "often"], ["--bogus"],
                                  ["--lambda"]])
def test_usage_errors(bs, args):
    r = cli(bs, *args)
    assert r.returncode == 2
    assert "error:" in r.stderr


def test_interactive_window_is_out_of_scope(bs):
    r = cli(bs)
    assert r.returncode == 1 and "--headless" in r.stderr


@pytest.mark.parametrize("size,msg", [
    ("1920", "`1920` is not a size; expected WIDTHxHEIGHT"),
    ("0x600", "width must be positive"),
    ("800x0", "height must be positive"),
    ("axb", "width: invalid digit found in string"),
])
def test_screenshot_size_errors(bs, size, msg):
    # parse_size (config.rs:181-192, the reference's size_parsing test)
    r = cli(bs, "--headless", "1", "--screenshot", "x.png", "--screenshot-size", size)
    assert r.returncode == 2
    assert msg in r.stderr and "--screenshot-size" in r.stderr


def test_field_must_be_a_field_mode(bs):
    r = cli(bs, "--field", "pressure", "--headless", "1")
    assert r.returncode == 2 and "--field" in r.stderr


def test_check_f64_needs_an_f32_run(bs):
    assert "--check-f64" in cli(bs, "--help").stdout
    r = cli(bs, "--precision", "f64", "--check-f64", "--headless", "1")
    assert r.returncode == 2 and "already f64" in r.stderr


ROW = re.compile(r"^ {0,9}\d+ {1,10}\d+\.\d {1,9}\d+\.\d\d% {1,10}\d+\.\d {1,11}\d\.\d{4} {1,11}-?\d\.\d{3}e-?\d+ +"
                 r"(-|\d\.\d{6})$")


@pytest.mark.gpu
def test_headless_report_format(bs, gpu_available, tmp_path):
    dump = tmp_path / "state.bin"
    r = cli(bs, "--grid", "48", "--bubbles", "1", "--nucleation", "simultaneous", "--headless", "40",
            "--report-every", "10", "--dump", str(dump))
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    lines = r.stdout.splitlines()
    assert lines[0].startswith("adapter: ")
    assert re.match(r"^lattice 48x48x48 = 0\.11 M cells \| \d+\.\d\d GB \| dt = 0\.200$", lines[1])
    assert lines[2] == "wall 4.00 cells | R_c 13.3 cells | sigma 0.0833 | eps 0.0125 | e_ref 0.1250"
    assert lines[4] == "      step       time     broken       R_eq      max|v|           K       E/E0"
    rows = lines[5:9]
    for row in rows:
        assert ROW.match(row), row
    assert rows[0].split()[0] == "10" and rows[-1].split()[0] == "40"
    assert rows[0].split()[-1] == "1.000000"  # the only bubble is in at step 0
    assert re.match(r"^40 steps in \d+\.\d\d s  \(\d+\.\d steps/s, \d+\.\d\d G cell-updates/s\)$", lines[-1])
    assert os.path.getsize(dump) == 48 ** 3 * 6 * 4


@pytest.mark.gpu
def test_default_report_every(bs, gpu_available):
    r = cli(bs, "--grid", "32", "--headless", "40")
    assert r.returncode == 0
    rows = [l for l in r.stdout.splitlines() if ROW.match(l)]
    assert len(rows) == 20  # max(steps / 20, 1) = 2 -> 20 reports (main.rs:32)


@pytest.mark.gpu
def test_check_f64_reports_error_and_drift(bs, gpu_available):
    # a headless run's field error and energy drift against the f64 scheme:
    # FastMath f32 against its f64 ExactMath copy after 120 steps (the last
    # bubble is in at step 0)
    r = cli(bs, "--grid", "48", "--bubbles", "2", "--nucleation", "simultaneous", "--headless", "120",
            "--report-every", "20", "--check-f64")
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    out = r.stdout
    assert "f64 check after 120 steps" in out
    num = r"(\d\.\d{3}e[-+]\d\d)"
    m_field = re.search(r"field \(phi, pi\): rel L2 " + num + ", rel Linf " + num, out)
    m_fluid = re.search(r"fluid \(E, Z\):    rel L2 " + num + ", rel Linf " + num, out)
    m_drift = re.search(r"energy drift from step 20: f32 ([-+]\d\.\d{3}e[-+]\d\d), f64 ([-+]\d\.\d{3}e[-+]\d\d) "
                        r"\(largest difference over the reports (\d\.\d{3}e[-+]\d\d)\)", out)
    assert m_field and m_fluid and m_drift, out
    # block relative L2 of FastMath f32 against the f64 scheme: the field
    # within the 1e-5 bar; the fluid block (E-dominated) within 1e-4 here
    # (the bar with its norm is held at the benchmark configurations)
    assert float(m_field.group(1)) < 1e-5 and float(m_fluid.group(1)) < 1e-4
    # both drifts are small, and the f32 one follows the f64 one
    d32, d64, worst = (float(m_drift.group(k)) for k in (1, 2, 3))
    assert abs(d64) < 1e-3 and worst < 1e-5
