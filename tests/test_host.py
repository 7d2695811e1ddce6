"""The C++ host layer on the CPU (no GPU needed): the .cfg reader against the
reference's own configuration files (sims/*.cfg, kept as fixtures under
tests/golden/sims) with every derived constant recomputed here from the
definitions in fileinfo.cc:141-152 and fluid_2d.cc:15-17; its error messages
(fileinfo.cc:11-193, common.cc:11-47); the lift-off analysis on synthetic
height profiles with known minima; and the C++ binding shown in
INTEGRATION.md compiled as a translation unit."""
import os
import re
import struct
import subprocess

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(REPO, "vdropimpact_amd", "_lib", "bin")
GOLDEN = os.path.join(REPO, "tests", "golden", "sims")


@pytest.fixture(scope="module")
def host():
    need = [os.path.join(BIN, b) for b in ("cfg_dump", "tt_analysis", "fluid_test")]
    if not all(os.path.exists(b) for b in need):
        from vdropimpact_amd import build
        build.build()
    return BIN


def cfg_dump(host, path):
    r = subprocess.run([os.path.join(host, "cfg_dump"), path], capture_output=True, text=True)
    return r.returncode, dict(l.split("=", 1) for l in r.stdout.splitlines() if "=" in l), r.stderr


def read_cfg(path):
    """An independent reading of a .cfg: keyword -> list of words."""
    out = {}
    for line in open(path):
        line = line.split("#", 1)[0].split()
        if line:
            out[line[0]] = line[1:]
    return out


@pytest.mark.parametrize("name", ["id_vis32.cfg", "ld_vis100.cfg", "ld_vis32.cfg", "lo_vis10.cfg", "template.cfg"])
def test_reads_the_reference_configurations(host, name, tmp_path):
    path = os.path.join(GOLDEN, name)
    rc, d, err = cfg_dump(host, path)
    assert rc == 0, err
    k = read_cfg(path)
    m, n = int(k["grid_points"][0]), int(k["grid_points"][1])
    nul = 1e-6 * float(k["nul_cSt"][0]) if "nul_cSt" in k else float(k["nul"][0])
    rho, R, V, mug = float(k["rhol"][0]), float(k["R"][0]), float(k["V"][0]), float(k["mug"][0])
    st = mug / (rho * V * R)
    L = float(k["L_nd"][0]) * R * st ** (1 / 3)
    x_sym = "x_sym" in k
    ax, bx = (0.0 if x_sym else -L), L
    by = L * n / m * (1 if x_sym else 2)
    dx = (bx - ax) / m
    tmult = float(k["tmult"][0])
    implicit = "implicit_visc" in k
    mu = nul * rho
    dt_reg = dx * tmult if implicit else dx * dx * rho / mu * tmult * 0.25
    expect = {"m": m, "n": n, "ax": ax, "bx": bx, "by": by, "mu": mu, "rho": rho, "dt_reg": dt_reg,
              "h0": float(k["h0_nd"][0]) * R * st ** (2 / 3),
              "t_end": float(k["t_end_nd"][0]) * R * st ** (2 / 3) / V,
              "alpha": 1 / float(k["gamma"][0]), "sigma": float(k["sigma"][0]), "Pamb": float(k["Pamb"][0]),
              "nframes": int(k["frames"][0]), "ntrace": int(k["tracers"][0])}
    for key, val in expect.items():
        assert float(d[key]) == pytest.approx(val, rel=1e-14), key
    assert int(d["x_sym"]) == x_sym and int(d["implicit_visc"]) == implicit
    assert int(d["mr_time_output"]) == ("mr_time_output" in k)
    bits = {"u": 1, "v": 2, "p": 4, "pb": 8, "pbfull": 16, "w": 32, "h": 64, "fbd": 128, "pg": 256, "hfull": 512,
            "fbdfull": 1024, "pgfull": 2048}
    assert int(d["fflags"]) == sum(bits[w] for w in set(k.get("output", [])))
    assert d["filename"] == path[:-3] + "odr"


def test_output_keyword_sets_every_flag(host, tmp_path):
    text = open(os.path.join(GOLDEN, "template.cfg")).read()
    text = text.replace("output              h", "output u v p pb pbfull w h fbd pg hfull fbdfull pgfull")
    p = tmp_path / "all.cfg"
    p.write_text(text)
    rc, d, _ = cfg_dump(host, str(p))
    assert rc == 0 and int(d["fflags"]) == 4095


BASE = """grid_points 64 32
L_nd 30
h0_nd 4.5
t_end_nd 1
frames 4
tmult 8e-3
nul_cSt 32
rhol 1000
R 1e-3
V 0.5
gamma 1.4
sigma 0.02
mug 1.8e-5
Pamb 1e5
x_sym
implicit_visc
gas_layer_model
"""


@pytest.mark.parametrize("edit,message", [
    (lambda t: t.replace("grid_points 64 32\n", ""), "Error: Grid dimensions not set"),
    (lambda t: t.replace("grid_points 64 32", "grid_points 64 -2"), "Error: Grid dimensions must be positive"),
    (lambda t: t + "frobnicate 3\n", "Keyword 'frobnicate' not understood on line 18"),
    (lambda t: t.replace("L_nd 30", "L_nd 30 40"), "Too many arguments at input line 2"),
    (lambda t: t.replace("grid_points 64 32", "grid_points 64"), "Not enough arguments at input line 1"),
    (lambda t: t.replace("tmult 8e-3\n", ""), "Value of tmult either invalid or not set"),
    (lambda t: t.replace("rhol 1000\n", ""), "Value of rhol either invalid or not set"),
    (lambda t: t.replace("gas_layer_model\n", ""), "Error: Either gas_layer_data or gas_layer_model must be set"),
    (lambda t: t.replace("sigma 0.02\n", ""), "Value of sigma either invalid or not set"),
    (lambda t: t + "output u q\n", "Error: Output type not understood"),
    (lambda t: t + "x_sym extra\n", "Too many arguments at input line 18"),
    (lambda t: t + "nif_center\n", "Error: The non-inertial frame is not available with the GPU step"),
])
def test_configuration_errors_stop_with_the_reference_messages(host, tmp_path, edit, message):
    p = tmp_path / "bad.cfg"
    p.write_text(edit(BASE))
    rc, _d, err = cfg_dump(host, str(p))
    assert rc == 1 and message in err, err


def test_file_name_checks(host, tmp_path):
    p = tmp_path / "x.txt"
    p.write_text(BASE)
    rc, _d, err = cfg_dump(host, str(p))
    assert rc == 1 and "Error: Filename must end in '.cfg'" in err
    rc, _d, err = cfg_dump(host, "cfg")
    assert rc == 1 and "Error: Filename is too short" in err
    rc, _d, err = cfg_dump(host, str(tmp_path / "missing.cfg"))
    assert rc == 1 and "Error opening file" in err


# ---------------------------------------------------------------- lift-off

def write_matrix(path, x, y, rows):
    """Gnuplot binary matrix of floats: header (count, x), then per row (y, values)."""
    with open(path, "wb") as f:
        f.write(struct.pack(f"{len(x) + 1}f", len(x), *x))
        for yy, r in zip(y, rows):
            f.write(struct.pack(f"{len(x) + 1}f", yy, *r))


def test_tt_analysis_finds_the_minimum_and_the_lift_off_frame(host, tmp_path):
    """Synthetic height profiles h = c_k + (x - x_k)^2 / (2 R_k): the raw
    analysis finds the grid point nearest x_k, the smoothed one (cubic
    Hermite through the smoothed values) lands within a grid spacing of x_k
    and at the parabola's minimum plus the smoothing's curvature offset; the
    lift-off frame is the one with the smallest minimum before the height
    rises 0.5 % above it."""
    cfg = tmp_path / "run.cfg"
    cfg.write_text(BASE)
    odr = tmp_path / "run.odr"
    odr.mkdir()
    m = 400
    dx = 2.5e-7
    x = (np.arange(m) + 0.5) * dx
    mins = [4e-6, 2e-6, 1e-6, 6e-7, 4e-7, 3.5e-7, 3.6e-7, 4.5e-7, 8e-7]
    xks = [v + 0.3 * dx for v in (2.0e-5, 2.5e-5, 3.0e-5, 3.3e-5, 3.6e-5, 3.8e-5, 4.0e-5, 4.2e-5, 4.4e-5)]
    Rk = 2e-3
    for k, (c, xk) in enumerate(zip(mins, xks)):
        write_matrix(odr / f"height.{k}", x, [0.0], [c + (x - xk) ** 2 / (2 * Rk)])
    for raw in (True, False):
        r = subprocess.run([os.path.join(host, "tt_analysis")] + (["-r"] if raw else []) + [str(odr)],
                           capture_output=True, text=True, cwd=tmp_path)
        assert r.returncode == 0, r.stderr
        lines = (tmp_path / ("run.ttb" if raw else "run.tta")).read_text().splitlines()
        assert len(lines) == len(mins)
        for line, c, xk in zip(lines, mins, xks):
            k, t, xm, fm = line.split()
            if raw:
                i = int(np.argmin(np.abs(x - xk)))
                assert float(xm) == pytest.approx(float(np.float32(x[i])), rel=1e-6)
            else:
                assert abs(float(xm) - xk) < dx
                # Gaussian smoothing of a parabola adds (sigma^2 / 2) h'' to it
                assert abs(float(fm) - c) < 0.05 * c + 1.0 * (1.2e-6) ** 2 / Rk
        summary = r.stdout.splitlines()
        assert summary[0].endswith(f"{len(mins)} files processed")
        vals = [float(v) for v in summary[1].split()]
        assert vals[0] == pytest.approx(32.0, rel=1e-12)          # viscosity in cSt
        assert vals[1] == pytest.approx(3.5e-7, rel=0.05)         # the minimum height
        assert vals[4] == pytest.approx(100 * 5 / 8, rel=1e-6)    # its frame, in % of the run


# ---------------------------------------------------------------- binding

def test_integration_binding_compiles(tmp_path):
    """Every ```cpp block of INTEGRATION.md, in order, forms one translation
    unit that compiles against include/vdi.h."""
    text = open(os.path.join(REPO, "INTEGRATION.md")).read()
    blocks = re.findall(r"```cpp\n(.*?)```", text, flags=re.S)
    assert len(blocks) >= 3
    src = tmp_path / "binding.cc"
    src.write_text("\n".join(blocks))
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", "-I",
                        os.path.join(REPO, "include"), str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
