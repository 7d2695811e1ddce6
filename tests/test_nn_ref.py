"""The CPU reference of the learned-guess network: its backward pass
against central finite differences (float64, no BF16 rounding), and its
parameter count."""
import numpy as np

import nn_ref


def test_param_count_about_one_million():
    _, n = nn_ref.param_layout()
    assert 0.9e6 < n < 1.2e6


def test_reference_gradient_matches_finite_differences():
    H, W = 8, 16
    rng = np.random.default_rng(3)
    layout, n = nn_ref.param_layout()
    p = np.zeros(n)
    for k, off, cout, cin in layout:
        p[off:off + cout * 9 * cin] = rng.uniform(-1, 1, cout * 9 * cin) * np.sqrt(6 / (9 * cin))
        p[off + cout * 9 * cin: off + cout * 9 * cin + cout] = rng.uniform(-0.1, 0.1, cout)
    x2 = rng.standard_normal((nn_ref.NIN, H * W))
    t = rng.standard_normal(H * W)
    net = nn_ref.Net(H, W, p, emulate_bf16=False)
    loss, g = net.loss_and_grad(x2, t)
    # a sample of weights and biases of the first, a middle and the head layers
    picks = []
    for k, off, cout, cin in [layout[0], layout[5], layout[-1]]:
        picks += [off + 5, off + cout * 9 * cin // 2, off + cout * 9 * cin + 1]
    for q in picks:
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[q] += h
        pm[q] -= h
        lp = nn_ref.Net(H, W, pp, emulate_bf16=False).loss_and_grad(x2, t)[0]
        lm = nn_ref.Net(H, W, pm, emulate_bf16=False).loss_and_grad(x2, t)[0]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[q]) <= 1e-6 * max(1.0, abs(fd)), (q, fd, g[q])


def test_bf16_rounding():
    a = np.array([1.0, 1.00390625, 1.005859375, -3.0000001, 65504.0], np.float32)
    r = nn_ref.bf16(a)
    # 1 + 2^-8 is a tie -> even (1.0); 1 + 1.5 * 2^-8 rounds up to 1 + 2^-7
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0078125
    assert r[3] == -3.0 and r[4] == 65536.0
