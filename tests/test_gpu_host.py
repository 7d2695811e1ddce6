"""The C++ host layer end to end on the GPU: fluid_test runs a .cfg file
(with every output type, tracers and restart files) and each file it writes
is read back here, independently of the host code, and compared with the
FP64 oracle driven through the same frame loop (fluid_2d.cc:40-89: frames of
l = int(interval / dt_reg) + 1 steps of interval / l); then fluid_test -r
resumes from a restart file it wrote, and from one the oracle's state was
written into, and both continuations match the oracle step by step (one step
per frame).

Bars: relative L2 1e-9 for double-precision files (the two sides run the same
V-cycles to the same 1e-10 residual and agree to ~1e-11), float rounding for
the Gnuplot float files."""
import ctypes
import os
import shutil
import struct
import subprocess
import sys

import numpy as np
import pytest

import oracle_lib

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(REPO, "vdropimpact_amd", "_lib", "bin")
# the step's schedules, V(nu1, nu2) on the finest then the restricted levels:
# MAC, FEM, u and v viscous solves (the library's defaults)
STEP_NU = [(3, 2, 3, 2), (3, 2, 2, 2), (4, 2, 0, 1), (4, 2, 0, 1)]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(np.asarray(a, float) - b) / (nb if nb > 0 else 1.0))


def read_matrix(path):
    """Gnuplot binary matrix of floats: (count, x...), then rows (y, values...)."""
    a = np.fromfile(path, dtype=np.float32)
    m = int(a[0])
    x = a[1:m + 1]
    rows = a[m + 1:].reshape(-1, m + 1)
    return x, rows[:, 0], rows[:, 1:]


def read_full(path):
    """count (int), ax, bx (doubles), then count doubles."""
    raw = open(path, "rb").read()
    n = struct.unpack("i", raw[:4])[0]
    ax, bx = struct.unpack("2d", raw[4:20])
    return ax, bx, np.frombuffer(raw[20:], dtype=np.float64, count=n)


def cfg_text(m, n, frames, t_end_nd, ntrace, restart_freq):
    return f"""# host test
grid_points {m} {n}
L_nd 30
h0_nd 4.586010
t_end_nd {t_end_nd!r}
tracers {ntrace}
frames {frames}
tmult 8e-3
restart_freq {restart_freq}
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
mr_time_output
output u v p pb pbfull w h fbd pg hfull fbdfull pgfull
"""


def derived(path):
    r = subprocess.run([os.path.join(BIN, "cfg_dump"), path], capture_output=True, text=True, check=True)
    return dict(l.split("=", 1) for l in r.stdout.splitlines())


def oracle_params(d):
    from vdropimpact_amd.params import VdiParams
    f = {k: float(v) for k, v in d.items() if k not in ("filename",)}
    return VdiParams(m=int(f["m"]), n=int(f["n"]), x_sym=int(f["x_sym"]), implicit_visc=int(f["implicit_visc"]),
                     ax=f["ax"], bx=f["bx"], ay=f["ay"], by=f["by"], rho=f["rho"], mu=f["mu"], dt_reg=f["dt_reg"],
                     alpha=f["alpha"], mug=f["mug"], Pamb=f["Pamb"], h0=f["h0"], R=f["R"], V=f["V"],
                     sigma=f["sigma"])


def c_rand_tracers(p, ntrace):
    """fluid_2d.cc:495-505 with the C library's rand() from its default seed,
    as a fresh fluid_test process draws them."""
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(1)
    RAND_MAX = 2147483647
    pos = np.empty((ntrace, 2))
    for k in range(ntrace):
        pos[k, 0] = p.ax + (p.bx - p.ax) / RAND_MAX * float(libc.rand())
        pos[k, 1] = p.ay + (p.by - p.ay) / RAND_MAX * float(libc.rand())
    return pos


def oracle_state(orc, p):
    """The restart body with tracers, as fluid_2d_sio.cc:97-119 writes it."""
    m, n = p.m, p.n
    u, v, pp = orc.get(oracle_lib.U), orc.get(oracle_lib.V), orc.get(oracle_lib.P)
    fields = np.column_stack([u, v, pp]).ravel()
    tau = orc.get(oracle_lib.GTAU)[2:m + 2]
    vf = oracle_lib.lib().vdo_vframe
    vf.restype = ctypes.c_double
    gas = np.concatenate([[vf(orc.h)], orc.get(oracle_lib.GP), tau, orc.get(oracle_lib.H),
                          orc.get(oracle_lib.HT), orc.get(oracle_lib.PG)])
    return np.concatenate([[orc.time], fields, orc.tracers().ravel(), gas])


def check_frame(odr, k, orc, p):
    """Every file of frame k against the oracle."""
    m, n, ml = p.m, p.n, p.m + 4
    dx, dy = (p.bx - p.ax) / m, (p.by - p.ay) / n
    U, V, P = (orc.padded(a) for a in (oracle_lib.U, oracle_lib.V, oracle_lib.P))
    # float rounding of the file, plus a floor of 1e-9 of the field's
    # physical scale (the liquid starts at rest, where values are noise)
    sv, sp = 1e-9 * p.V, 1e-9 * p.sigma / p.R
    fl = lambda a, floor=0.0: float(np.abs(a).max()) * 2e-7 + floor
    x, y, val = read_matrix(os.path.join(odr, f"u.{k}"))
    np.testing.assert_allclose(x, p.ax + (np.arange(ml) - 1.5) * dx, rtol=1e-6)
    np.testing.assert_allclose(y, p.ay + (np.arange(n + 4) - 1.5) * dy, rtol=1e-6, atol=1e-12)
    np.testing.assert_allclose(val, U, rtol=0, atol=fl(U, sv) + 1e-9 * np.abs(U).max())
    _x, _y, val = read_matrix(os.path.join(odr, f"v.{k}"))
    np.testing.assert_allclose(val, V, rtol=0, atol=fl(V, sv) + 1e-9 * np.abs(V).max())
    nodes = P[2:n + 3, 2:m + 3]
    x, y, val = read_matrix(os.path.join(odr, f"p.{k}"))
    np.testing.assert_allclose(x, p.ax + np.arange(m + 1) * dx, rtol=1e-6)
    np.testing.assert_allclose(val, nodes, rtol=0, atol=fl(nodes, sp) + 1e-9 * np.abs(nodes).max())
    _x, y, val = read_matrix(os.path.join(odr, f"pb.{k}"))
    assert val.shape == (1, m + 1)
    np.testing.assert_allclose(val[0], nodes[0], rtol=0, atol=fl(nodes[0], sp) + 1e-9 * np.abs(nodes[0]).max())
    # vorticity at the nodes from the centred differences of the four cells
    # around each node
    c = lambda A, di, dj: A[1 + dj:n + 2 + dj, 1 + di:m + 2 + di]
    w = 0.5 * ((c(V, 1, 1) - c(V, 0, 1) + c(V, 1, 0) - c(V, 0, 0)) / dx -
               (c(U, 1, 1) + c(U, 0, 1) - c(U, 1, 0) - c(U, 0, 0)) / dy)
    _x, _y, val = read_matrix(os.path.join(odr, f"w.{k}"))
    np.testing.assert_allclose(val, w, rtol=0, atol=fl(w, sv / dx) + 1e-8 * np.abs(w).max())
    ax, bx, pbf = read_full(os.path.join(odr, f"pbfull.{k}"))
    assert (ax, bx) == (p.ax, p.bx)
    assert np.linalg.norm(pbf - nodes[0]) <= 1e-9 * np.linalg.norm(nodes[0]) + sp * np.sqrt(m + 1)
    for stem, full, aid, at_nodes in (("height", "hfull", oracle_lib.H, False), ("fbd", "fbdfull", oracle_lib.FBD, False),
                                      ("pg", "pgfull", oracle_lib.PG, True)):
        ref = orc.get(aid)
        x, y, val = read_matrix(os.path.join(odr, f"{stem}.{k}"))
        np.testing.assert_allclose(x, p.ax + (np.arange(ref.size) + (0 if at_nodes else 0.5)) * dx, rtol=1e-6)
        scale = {"height": p.h0, "fbd": p.V * p.dt_reg, "pg": p.sigma / p.R}[stem]
        np.testing.assert_allclose(val[0], ref, rtol=0, atol=fl(ref, 1e-9 * scale) + 1e-9 * np.abs(ref).max())
        ax, bx, vf = read_full(os.path.join(odr, f"{full}.{k}"))
        assert (ax, bx) == (p.ax, p.bx) and vf.size == ref.size
        floor = 1e-6 * scale * np.sqrt(ref.size) if stem == "fbd" else 1e-9 * scale * np.sqrt(ref.size)
        assert np.linalg.norm(vf - ref) <= 1e-9 * np.linalg.norm(ref) + floor, stem
    tr = np.fromfile(os.path.join(odr, f"trace.{k}"), dtype=np.float32).reshape(-1, 2)
    np.testing.assert_allclose(tr, orc.tracers(), rtol=2e-7, atol=1e-12)


def check_restart(path, orc, p):
    body = np.fromfile(path, dtype=np.float64)
    ref = oracle_state(orc, p)
    assert body.size == ref.size
    assert body[0] == ref[0]
    m, n = p.m, p.n
    nf = 3 * (m + 4) * (n + 4)
    for name, sl in (("fields", slice(1, 1 + nf)), ("tracers", slice(1 + nf, 1 + nf + 2 * orc.ntrace)),
                     ("gas", slice(1 + nf + 2 * orc.ntrace, None))):
        assert np.linalg.norm(body[sl] - ref[sl]) <= 1e-9 * np.linalg.norm(ref[sl]) + 1e-9 * p.V, name


def run_fluid_test(cfg, resume=False):
    args = [os.path.join(BIN, "fluid_test")] + (["-r"] if resume else []) + [cfg]
    r = subprocess.run(args, capture_output=True, text=True, timeout=300, cwd=os.path.dirname(cfg))
    assert r.returncode == 0, r.stderr
    return r.stdout


def test_gpu_fluid_test_outputs_and_restarts_match_oracle(vdi, oracle, tmp_path):
    m, n, frames, ntrace = 64, 32, 6, 40
    # one step per frame: the frame interval is 0.9 dt_reg
    probe = tmp_path / "probe.cfg"
    probe.write_text(cfg_text(m, n, frames, 1.0, ntrace, 2))
    d = derived(str(probe))
    t_end_nd = frames * 0.9 * float(d["dt_reg"]) / float(d["t_end"])
    cfg = tmp_path / "run.cfg"
    cfg.write_text(cfg_text(m, n, frames, t_end_nd, ntrace, 2))
    d = derived(str(cfg))
    p = oracle_params(d)
    interval = float(d["t_end"]) / frames
    l = int(interval / p.dt_reg) + 1
    adt = interval / l
    assert l == 1
    say("fluid_test", m, n, "frames", frames, "steps per frame", l)
    out = run_fluid_test(str(cfg))
    lines = [s for s in out.splitlines() if s and not s.startswith("#")]
    assert len(lines) == frames
    for k, line in enumerate(lines, start=1):
        w = line.split()
        assert int(w[0]) == k and int(w[1]) == l and int(w[4]) > 0 and int(w[6]) > 0
    odr = str(tmp_path / "run.odr")

    orc = oracle.OracleSim(p, rel_tol=1e-10)
    for q, nu in enumerate(STEP_NU):
        orc.set_schedule(q, *nu)
    orc.init_fields()
    orc.set_tracers(c_rand_tracers(p, ntrace))
    check_frame(odr, 0, orc, p)
    states = {}
    for k in range(1, frames + 1):
        for _ in range(l):
            orc.step_forward(adt)
        check_frame(odr, k, orc, p)
        if k % 2 == 0:
            check_restart(os.path.join(odr, f"res.{k}"), orc, p)
        states[k] = oracle_state(orc, p)
        say("  frame", k, "matches the oracle")

    # resume from the GPU's own restart file at frame 2
    os.makedirs(tmp_path / "again.odr")
    shutil.copy(cfg, tmp_path / "again.cfg")
    shutil.copy(os.path.join(odr, "res.2"), tmp_path / "again.odr" / "res.2")
    out = run_fluid_test(str(tmp_path / "again.cfg"), resume=True)
    assert "# Starting from restart file at frame 2" in out
    for k in range(3, frames + 1):
        assert rel(read_full(str(tmp_path / "again.odr" / f"hfull.{k}"))[2],
                   read_full(os.path.join(odr, f"hfull.{k}"))[2]) < 1e-9
    body = np.fromfile(str(tmp_path / "again.odr" / f"res.{frames}"), dtype=np.float64)
    assert np.linalg.norm(body - states[frames]) <= 1e-9 * np.linalg.norm(states[frames])

    # resume from the oracle's state at frame 2, written here in the
    # reference's restart format, and compare every later frame (one step
    # each) with the oracle's continuation
    orc2 = oracle.OracleSim(p, rel_tol=1e-10)
    for q, nu in enumerate(STEP_NU):
        orc2.set_schedule(q, *nu)
    orc2.init_fields()
    orc2.set_tracers(c_rand_tracers(p, ntrace))
    for _ in range(2 * l):
        orc2.step_forward(adt)
    os.makedirs(tmp_path / "fromoracle.odr")
    shutil.copy(cfg, tmp_path / "fromoracle.cfg")
    oracle_state(orc2, p).tofile(str(tmp_path / "fromoracle.odr" / "res.2"))
    # a restart file carries no fbd (gas_layer_solve.cc:326-332): the resumed
    # run accumulates the flow across the boundary from zero
    orc2.set(oracle_lib.FBD, np.zeros(p.m))
    run_fluid_test(str(tmp_path / "fromoracle.cfg"), resume=True)
    for k in range(3, frames + 1):
        for _ in range(l):
            orc2.step_forward(adt)
        check_frame(str(tmp_path / "fromoracle.odr"), k, orc2, p)
    check_restart(str(tmp_path / "fromoracle.odr" / f"res.{frames}"), orc2, p)
    orc.close()
    orc2.close()
