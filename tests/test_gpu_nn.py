"""The learned-guess network on the GPU against its CPU reference
(tests/nn_ref.py): forward outputs and parameter gradients, with the
reference rounding to BF16 exactly where the GPU stores BF16; and a short
training run that must lower the loss."""
import sys

import numpy as np
import pytest

import nn_ref

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("H,W", [(64, 64), (64, 128)])
def test_gpu_unet_forward_and_gradient_match_cpu(vdi, H, W):
    say("unet", H, W)
    net = vdi.UNet(H, W, seed=11)
    p = net.get_params()
    rng = np.random.default_rng(5)
    # non-zero biases so that every bias gradient path is exercised
    layout, n = nn_ref.param_layout()
    assert n == net.nparams
    for k, off, cout, cin in layout:
        p[off + cout * 9 * cin: off + cout * 9 * cin + cout] = rng.uniform(-0.05, 0.05, cout)
    net.set_params(p)
    x2 = rng.standard_normal((nn_ref.NIN, H * W)).astype(np.float32)
    t = rng.standard_normal(H * W).astype(np.float32)
    ref = nn_ref.Net(H, W, p.astype(np.float64), emulate_bf16=True)
    y_cpu = ref.forward(x2.astype(np.float64))
    y_gpu = net.forward(x2)
    e_y = rel(y_gpu, y_cpu)
    say("  forward rel L2", e_y)
    assert e_y < 1e-2
    loss_cpu, g_cpu = ref.loss_and_grad(x2.astype(np.float64), t.astype(np.float64))
    loss_gpu, g_gpu = net.gradient(x2, t)
    say("  loss", loss_gpu, loss_cpu)
    assert abs(loss_gpu - loss_cpu) <= 2e-2 * loss_cpu
    e_g = rel(g_gpu, g_cpu)
    say("  gradient rel L2", e_g)
    assert e_g < 3e-2
    worst = 0.0
    for k, off, cout, cin in layout:
        L = cout * 9 * cin + cout
        e = rel(g_gpu[off:off + L], g_cpu[off:off + L])
        worst = max(worst, e)
    say("  worst layer gradient rel L2", worst)
    assert worst < 1e-1
    net.close()


def test_gpu_mfma_conv_exact_on_bf16_values(vdi):
    """With inputs and weights that BF16 represents exactly and sums that
    FP32 holds exactly (small integers), the first layer's pre-activation is
    exact, so the head output matches the reference to FP32 rounding."""
    H, W = 64, 64
    net = vdi.UNet(H, W, seed=3)
    layout, n = nn_ref.param_layout()
    rng = np.random.default_rng(1)
    p = np.zeros(n, np.float32)
    for k, off, cout, cin in layout:
        p[off:off + cout * 9 * cin] = rng.integers(-2, 3, cout * 9 * cin) / 8.0
    net.set_params(p)
    x2 = rng.integers(-3, 4, (nn_ref.NIN, H * W)).astype(np.float32)
    ref = nn_ref.Net(H, W, p.astype(np.float64), emulate_bf16=True)
    y_cpu = ref.forward(x2.astype(np.float64))
    y_gpu = net.forward(x2)
    say("  exact-value forward rel L2", rel(y_gpu, y_cpu))
    assert rel(y_gpu, y_cpu) < 1e-2
    net.close()
