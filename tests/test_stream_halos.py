"""Row halos of the streaming level launches (csrc/mg_stream.hpp), checked on
the CPU by brute force: a Gauss-Seidel sweep restricted to a segment of rows
plus its halo (zeros outside, updates limited to the region a pass can still
get right) must reproduce the global sweep on the rows the launch reads
afterwards, and a halo two rows shorter must not.

Cell-centred levels: red-black passes of the 5-point operator, halo 2 nu + 1
rows, pass p limited to rows [ylo + p, yhi - p].  Nodal levels: the
four-colour passes of the 9-point operator in fem_tile_sweep's order
(colour (p-1)&3: column parity (p-1)&1, row parity ((p-1)>>1)&1), halo
2 nu + 2 rows, pass p limited to rows [ylo + ceil(p/2), yhi - ceil(p/2)].
The stencil weights are arbitrary: exactness is about which values a pass
reads, not about the operator."""
import numpy as np
import pytest


def _sweep(z, b, passes, nodal, ylo=None, yhi=None):
    n, m = z.shape
    for p in range(1, passes + 1):
        if nodal:
            c = (p - 1) & 3
            ci, cj = c & 1, c >> 1
            shrink = (p + 1) // 2
        else:
            colour = (p - 1) & 1
            shrink = p
        for r in range(1, n - 1):
            if ylo is not None and not (ylo + shrink <= r <= yhi - shrink):
                continue
            if nodal and r % 2 != cj:
                continue
            xs = range(ci, m, 2) if nodal else [x for x in range(m) if (x + r) % 2 == colour]
            for x in xs:
                acc = 0.0
                for dy in (-1, 0, 1):
                    for dx in (-1, 0, 1):
                        if (dy, dx) == (0, 0) or (not nodal and dy != 0 and dx != 0):
                            continue
                        if 0 <= x + dx < m:
                            acc += (0.1 + 0.01 * (dy + 2) + 0.003 * (dx + 2)) * z[r + dy, x + dx]
                z[r, x] = (b[r, x] - acc) / 3.0
    return z


def _segment(z0, b, passes, nodal, y0, S, hv):
    zs = np.zeros_like(z0)
    ylo, yhi = y0 - hv, y0 + S + hv - 1
    zs[ylo:yhi + 1] = z0[ylo:yhi + 1]
    return _sweep(zs, b, passes, nodal, ylo, yhi)


@pytest.mark.parametrize("nodal", [False, True])
@pytest.mark.parametrize("nu", [1, 2, 3])
@pytest.mark.parametrize("y0", [20, 22])
def test_stream_row_halo_exact_and_tight(nodal, nu, y0):
    rng = np.random.default_rng(100 * nu + y0 + nodal)
    n, m, S = 60, 12, 8
    passes = 4 * nu if nodal else 2 * nu
    hv = 2 * nu + 2 if nodal else 2 * nu + 1
    z0 = rng.standard_normal((n, m))
    b = rng.standard_normal((n, m))
    zg = _sweep(z0.copy(), b, passes, nodal)
    # rows read afterwards: the residual rows and their neighbours; the
    # residual rows are y0 .. y0 + S - 1 (cell-centred: 2 x 2 restriction
    # blocks inside the segment) or y0 - 1 .. y0 + S - 1 (nodal: full
    # weighting reaches one row below)
    rows = slice(y0 - 2, y0 + S + 1) if nodal else slice(y0 - 1, y0 + S + 1)
    assert np.array_equal(_segment(z0, b, passes, nodal, y0, S, hv)[rows], zg[rows])
    assert not np.array_equal(_segment(z0, b, passes, nodal, y0, S, hv - 2)[rows], zg[rows])
