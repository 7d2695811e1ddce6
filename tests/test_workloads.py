"""Synthetic operator-level inputs: the generator and the right-hand sides."""
import numpy as np

from vdropimpact_amd import workloads as w


def test_splitmix64_reference_sequence():
    """SplitMix64's published outputs for seed 1234567."""
    r = w.SplitMix64(1234567)
    assert [r.next_u64() for _ in range(5)] == [
        6457827717110365317, 3203168211198807973, 9817491932198370423, 4593380528125082431, 16408922859458223821]


def test_modes_are_deterministic_and_in_range():
    a, b = w.fourier_modes(), w.fourier_modes()
    assert a == b and len(a) == 2 and len(a[0]) == 64
    for comp in a:
        for kx, ky, amp, phi in comp:
            assert 1 <= kx <= 16 and 1 <= ky <= 16 and 0 <= phi < 2 * np.pi
    assert w.fourier_modes(seed=1) != a


def test_mac_rhs_is_discrete_divergence_of_smooth_field():
    """For the Fourier field the face divergence converges to div(u) at
    second order."""
    errs = []
    for m in (64, 128):
        n = m // 2
        cu, cv = w.fourier_modes()
        x = (np.arange(m) + 0.5) / m
        y = (np.arange(n) + 0.5) / m
        X, Y = np.meshgrid(x, y)
        div = np.zeros_like(X)
        for kx, ky, amp, phi in cu:
            div += -amp * 2 * np.pi * kx * np.sin(2 * np.pi * (kx * X + ky * Y) + phi)
        for kx, ky, amp, phi in cv:
            div += -amp * 2 * np.pi * ky * np.sin(2 * np.pi * (kx * X + ky * Y) + phi)
        errs.append(np.abs(w.mac_rhs(m, n) - div.ravel()).max())
    assert errs[1] < errs[0] / 3.5


def test_fem_rhs_dirichlet_row_and_sum():
    b = w.fem_rhs(32, 16).reshape(17, 33)
    assert np.all(b[0] == 0)
    assert np.isfinite(b).all()


def test_batch_members_draws():
    cfgs, prms = w.batch_members(64)
    assert len(prms) == 64
    sizes = {(c.m, c.n) for c in cfgs}
    assert sizes <= set(w.BATCH_GRIDS) and len(sizes) == 3
    for c, p in zip(cfgs, prms):
        assert 1e-6 <= c.nul <= 100e-6 and 0.3 <= c.V <= 1.0 and 0.5e-3 <= c.R <= 2e-3
        # h0 = h_min = 5 um for every member (fileinfo.cc:145-146)
        assert abs(p.h0 - 5e-6) < 1e-15
        assert p.m == c.m and p.dt_reg > 0
    assert w.batch_members(64)[0] == cfgs
