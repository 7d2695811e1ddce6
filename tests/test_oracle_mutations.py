"""Mutation check of the oracle pins: each plausible mistake below (a sign or
index slip, a dropped term, a swapped operand, a wrong boundary copy) is
planted in a copy of oracle/vdo.c, built into a temporary library, and the
pin tests run against it; every one must make at least one pin fail.  The
oracle source itself is never modified."""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(REPO, "oracle", "vdo.c")
PIN_TESTS = ["tests/test_oracle_step_pins.py", "tests/test_oracle_pins.py"]

PLANTS = [
    ("K1 dt term uses v", "AT(ul, i, j) = fu - 0.5 * (dx + dt * fu) * AT(ux, i, j);", "AT(ul, i, j) = fu - 0.5 * (dx + dt * fv) * AT(ux, i, j);"),
    ("K1 uy from x neighbours", "AT(uy, i, j) = ysp * vdo_mono_diff(AT(u, i, j - 2), AT(u, i, j - 1), AT(u, i, j), AT(u, i, j + 1), AT(u, i, j + 2));",
     "AT(uy, i, j) = ysp * vdo_mono_diff(AT(u, i - 2, j), AT(u, i - 1, j), AT(u, i, j), AT(u, i + 1, j), AT(u, i + 2, j));"),
    ("K2 boundary copy ur[-1]", "if (i == 0) AT(ur, -1, j) = AT(ul, 0, j);", "if (i == 0) AT(ur, -1, j) = AT(ur, 0, j);"),
    ("K3 index slip", "AT(c0, i, j) = 0.5 * xsp * (AT(ul, i + 1, j) + AT(ul, i, j)) * (AT(ur, i, j) - AT(ur, i - 1, j));",
     "AT(c0, i, j) = 0.5 * xsp * (AT(ul, i + 1, j) + AT(ul, i, j)) * (AT(ur, i + 1, j) - AT(ur, i, j));"),
    ("K4 pressure sign", "double f_x = -px * rhoinv + viscfac * (uxx + uyy);", "double f_x = px * rhoinv + viscfac * (uxx + uyy);"),
    ("K4 transverse term dropped", "AT(ud, i, j) = fu + 0.5 * (-dy * fuy + dt * (-fc0 - fv * fuy + f_x));", "AT(ud, i, j) = fu + 0.5 * (-dy * fuy + dt * (-fv * fuy + f_x));"),
    ("K5 outflow copy", "else { AT(ul, m, j) = AT(ur, m - 1, j); AT(vl, m, j) = AT(vr, m - 1, j); }\n            if (j < n - 1) vdo_godunov_set(",
     "else { AT(ul, m, j) = AT(ur, m - 1, j); AT(vl, m, j) = AT(vl, m - 1, j); }\n            if (j < n - 1) vdo_godunov_set("),
    ("K6 s_bc factor", "const double macfac = 0.125 * s->rhoinv * dt;", "const double macfac = 0.25 * s->rhoinv * dt;"),
    ("K7 bottom face sign", "for (int i = 0; i < m; i++) AT(vd, i, 0) -= 2 * ysp * (zm[i] - s_bc[i]);", "for (int i = 0; i < m; i++) AT(vd, i, 0) -= 2 * ysp * (zm[i] + s_bc[i]);"),
    ("K8 index slip", "sy * (AT(ud, i, j + 1) - AT(ud, i, j))", "sy * (AT(ud, i + 1, j) - AT(ud, i, j))"),
    ("visco tau index", "double dudy = muinv * 0.5 * (s->gtau_old[i + 2] + s->gtau[i + 2]);", "double dudy = muinv * 0.5 * (s->gtau_old[i + 2] + s->gtau[i + 1]);"),
    ("visco top ghost", "if (j == n - 1) r += vfy * AT(w, i, n);", "if (j == n - 1) r += vfx * AT(w, i, n);"),
    ("K14 right flux sign", "if (j < n) src[m + (size_t)j * me] -= 0.25 * dy * rdt * (AT(c0, m - 1, j) + AT(u, m, j));",
     "if (j < n) src[m + (size_t)j * me] += 0.25 * dy * rdt * (AT(c0, m - 1, j) + AT(u, m, j));"),
    ("K15 x/y swap", "AT(u, i, j) = AT(c0, i, j) - 0.5 * dt * rhoinv * xsp * (q[me + 1] + q[1] - q[me] - q[0]);",
     "AT(u, i, j) = AT(c0, i, j) - 0.5 * dt * rhoinv * xsp * (q[me + 1] - q[1] + q[me] - q[0]);"),
    ("viscous bc outer ghost", "AT(u, c, -2) = AT(u, c, 1) - 3 * s->dy * un_edge;", "AT(u, c, -2) = AT(u, c, 1) - 2 * s->dy * un_edge;"),
    ("viscous bc v ghost", "AT(v, c, -2) = AT(v, c, -1) + s->dy * ux2;", "AT(v, c, -2) = AT(v, c, -1) - s->dy * ux2;"),
    ("Simpson image sign", "*fsu = p * (xms * rm + xps * rp);", "*fsu = p * (xms * rm - xps * rp);"),
    ("inviscid left mirror", "AT(u, -1, j) = -AT(u, m, j); AT(v, -1, j) = AT(v, m, j);", "AT(u, -1, j) = AT(u, m, j); AT(v, -1, j) = AT(v, m, j);"),
    ("Newton x_sym row", "if (i == 0) c[i] += a[i];", ";"),
    ("liquid p x_sym edge", "p[0] = pg[0] + 2 * fac * (h[1] - h[0]) - fac2;", "p[0] = pg[0] + fac * (h[1] - h[0]) - fac2;"),
    ("extrapolation coefficient", "double a = 1. / (g->t0 - g->t_prev), bb = (t - g->t_prev) * a, c = (t - g->t0) * a;",
     "double a = 1. / (g->t0 - g->t_prev), bb = (t - g->t_prev) * a, c = (t - g->t_prev) * a;"),
    ("tracer weights swapped", "tp[0] += dt * ((1 - y) * (AT(u, i, j) * (1 - x) + AT(u, i + 1, j) * x) +",
     "tp[0] += dt * ((1 - y) * (AT(u, i, j) * x + AT(u, i + 1, j) * (1 - x)) +"),
    ("hx x_sym edge", "s->hx[0] = s->P.x_sym ? 0 : s->hx[1];", "s->hx[0] = s->P.x_sym ? s->hx[1] : s->hx[1];"),
]


def _run(plant, tmp):
    name, old, new = plant
    text = open(SRC).read()
    assert text.count(old) == 1, f"{name}: pattern not unique"
    d = os.path.join(tmp, name.replace(" ", "_").replace("/", "_"))
    os.makedirs(d, exist_ok=True)
    csrc = os.path.join(d, "vdo.c")
    open(csrc, "w").write(text.replace(old, new))
    so = os.path.join(d, "libvdo.so")
    subprocess.run(["gcc", "-O2", "-fopenmp", "-fPIC", "-std=c11", "-D_DEFAULT_SOURCE", "-shared",
                    "-I", os.path.join(REPO, "oracle"), "-o", so, csrc, "-lm"], check=True)
    env = dict(os.environ, VDO_ORACLE_SO=so, OMP_NUM_THREADS="1")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", *PIN_TESTS],
                       cwd=REPO, env=env, capture_output=True, text=True)
    return name, r.returncode, r.stdout[-400:]


@pytest.mark.slow
def test_every_planted_mistake_fails_a_pin(tmp_path):
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        results = list(ex.map(lambda p: _run(p, str(tmp_path)), PLANTS))
    missed = [name for name, rc, _ in results if rc == 0]
    assert not missed, f"mistakes no pin catches: {missed}"
