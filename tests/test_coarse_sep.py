"""The exact coarse-level solve by separation of variables (host factors in
vdropimpact_amd/csrc/coarse_sep.hpp), on the CPU: for every system kind and
both transform directions, the pencil built from the operator's parameters
solves the same system as the oracle's level operator (assembled from its
action on unit vectors, a dense solve by numpy), and the generalized
eigenbasis satisfies V^T R V = I."""
import os
import subprocess

import numpy as np
import pytest

import closed_forms as cf

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

DRIVER = r'''
#include <cstdio>
#include <vector>
#include "coarse_sep.hpp"
using namespace vdi::sep;
int main() {
    int full, fem, tdir, m, n, bc[4];
    double a[6];
    if (scanf("%d %d %d %d %d", &full, &fem, &tdir, &m, &n) != 5) return 2;
    for (int k = 0; k < 4; k++) if (scanf("%d", &bc[k]) != 1) return 2;
    for (int k = 0; k < 6; k++) if (scanf("%lf", &a[k]) != 1) return 2;
    Tri Pt, Rt, Sp, Qp;
    if (fem) {
        if (!fem_pencil(a[0], a[1], a[2], a[3], m, n, tdir, Pt, Rt, Sp, Qp)) return 3;
    } else {
        cc_pencil(a[0], a[1], bc[0], bc[1], bc[2], bc[3], m, n, a[4], a[5], tdir, Pt, Rt, Sp, Qp);
    }
    const int T = (int)Pt.d.size(), P = (int)Sp.d.size();
    std::vector<double> f;
    if (!(full ? factor2(Pt, Rt, Sp, Qp, f) : factor(Pt, Rt, Sp, Qp, f))) return 4;
    std::vector<double> b((size_t)T * P), z;
    for (auto &v : b) if (scanf("%lf", &v) != 1) return 2;
    if (full) solve2(T, P, f, b, z);
    else solve(T, P, f, b, z);
    printf("%d %d\n", T, P);
    for (int i = 0; i < T * T; i++) printf("%.17g\n", f[i]);
    for (double v : z) printf("%.17g\n", v);
    return 0;
}
'''


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    d = tmp_path_factory.mktemp("sep")
    src = d / "drv.cc"
    src.write_text(DRIVER)
    exe = d / "drv"
    r = subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-I",
                        os.path.join(REPO, "vdropimpact_amd", "csrc"), str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return str(exe)


def run(driver, fem, tdir, m, n, bc, a, b, full=0):
    inp = f"{full} {fem} {tdir} {m} {n}\n" + " ".join(map(str, bc)) + "\n" + " ".join(repr(float(v)) for v in a) + "\n"
    inp += "\n".join(repr(float(v)) for v in b) + "\n"
    r = subprocess.run([driver], input=inp, capture_output=True, text=True)
    assert r.returncode == 0, (r.returncode, r.stderr)
    vals = r.stdout.split()
    T, P = int(vals[0]), int(vals[1])
    V = np.array(vals[2:2 + T * T], float).reshape(T, T)
    z = np.array(vals[2 + T * T:], float)
    return T, P, V, z


# (kind, m, n, cx, cy): oblong levels both ways round, so that each
# transform direction runs over the shorter and over the longer side
CASES = [
    (cf.MAC, 32, 16, 1 / 0.01 ** 2, 1 / 0.013 ** 2),
    (cf.MAC, 12, 20, 3.0, 2.0),
    (cf.VX, 32, 16, 0.13, 0.09),
    (cf.VY, 16, 24, 0.2, 0.31),
    (cf.FEM, 33, 17, 0.01, 0.013),
    (cf.FEM, 13, 21, 0.02, 0.011),
]


@pytest.mark.parametrize("x_sym", [1, 0])
@pytest.mark.parametrize("kind,m,n,cx,cy", CASES)
@pytest.mark.parametrize("tdir,full", [(0, 0), (1, 0), (0, 1)])
def test_separable_solve_matches_dense_solve(oracle, driver, kind, m, n, cx, cy, x_sym, tdir, full):
    """full = 1: both directions diagonalised (the multi-workgroup solve of
    larger coarsest levels), rows transformed along y as the kernels do."""
    g = oracle.OracleMG(kind, m, n, cx, cy, x_sym)
    N = m * n
    A = np.stack([g.apply(np.eye(N)[c]) for c in range(N)], axis=1)
    rng = np.random.default_rng(m * 31 + n + tdir)
    b = rng.standard_normal(N)
    fem = kind == cf.FEM
    if fem:
        b[:m] = 0.0            # row 0 is the Dirichlet row (handled outside the separable part)
        free = np.arange(m, N)
        z_ref = np.zeros(N)
        z_ref[free] = np.linalg.solve(A[np.ix_(free, free)], b[free])
    else:
        z_ref = np.linalg.solve(A, b)
    # the pencil parameters as the library derives them from the level
    if fem:
        r = cx / cy
        a = (4 / 3 * (r + 1 / r), (r - 2 / r) / 3, (1 / r - 2 * r) / 3, -(r + 1 / r) / 6, 0.0, 0.0)
        bcs = (0, 0, 0, 0)
    elif kind == cf.MAC:
        a, bcs = (0.0, 1.0, 0.0, 0.0, cx, cy), (0, 0, 1, 0)
    else:
        left = (1 if kind == cf.VX else 0) if x_sym else 2
        a, bcs = (1.0, -1.0, 0.0, 0.0, cx, cy), (left, 2, 0, 2)
    rows = n - 1 if fem else n
    bb = b.reshape(n, m)[1:] if fem else b.reshape(n, m)
    # the driver indexes unknowns p + P t: t is the row for tdir 0, the column for tdir 1
    bt = bb.ravel() if tdir == 0 else bb.T.ravel()
    T, P, V, zs = run(driver, int(fem), tdir, m, n, bcs, a, bt, full)
    assert (T, P) == ((rows, m) if tdir == 0 else (m, rows))
    zz = zs.reshape(T, P) if tdir == 0 else zs.reshape(T, P).T
    z = np.concatenate([np.zeros(m), zz.ravel()]) if fem else zz.ravel()
    assert np.linalg.norm(z - z_ref) <= 1e-11 * np.linalg.norm(z_ref)
    # V^T R V = I for the mass-type factor R of the transform direction
    if fem:
        if tdir == 0:
            R = np.diag([4.0] * (T - 1) + [2.0]) + np.diag(np.ones(T - 1), 1) + np.diag(np.ones(T - 1), -1)
        else:
            R = np.diag([2.0] + [4.0] * (T - 2) + [2.0]) + np.diag(np.ones(T - 1), 1) + np.diag(np.ones(T - 1), -1)
    else:
        R = np.eye(T)
    np.testing.assert_allclose(V.T @ R @ V, np.eye(T), atol=1e-12)
