"""Tile shapes and the autotuner: every candidate shape computes the same
V-cycle (bit for bit: the temporally blocked sweeps reproduce a global sweep
whatever the tiling), and a tuned context gives the same steps."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


CASES = [(0, 256, 128, 1e4, 6e3), (1, 257, 129, 0.01, 0.013), (2, 256, 128, 0.13, 0.09)]


@pytest.mark.parametrize("kind,m,n,cx,cy", CASES)
@pytest.mark.parametrize("form", [0, 1])
def test_gpu_every_tile_shape_same_cycle(vdi, kind, m, n, cx, cy, form):
    rng = np.random.default_rng(7 + kind)
    b = rng.standard_normal(m * n)
    ref = vdi.MGSystem(kind, m, n, cx, cy, 1)
    ref.set_form(form)
    ref.set_b(b)
    ref.set_z(np.zeros(m * n))
    ref.vcycles(3)
    z_ref = ref.get_z()
    levels = ref.levels
    for shape in range(vdi.NUM_TILE_SHAPES):
        say("tile shape", vdi.tile_shape(shape), "kind", kind, "form", form)
        g = vdi.MGSystem(kind, m, n, cx, cy, 1)
        g.set_form(form)
        prec = 1 if form else 0
        for lv in range(levels):
            for which in (0, 1):
                try:
                    g.set_tile(prec, which, lv, shape)
                except vdi.VdiError:
                    pass   # level handled by the single-workgroup kernel, or LDS too small
        if form:
            try:
                g.set_tile(1, 2, 0, shape)
            except vdi.VdiError:
                pass
        g.set_b(b)
        g.set_z(np.zeros(m * n))
        g.vcycles(3)
        np.testing.assert_array_equal(g.get_z(), z_ref)
        g.close()
    ref.close()


STREAM_SCHEDULES = [(0, 1), (1, 0), (1, 1), (3, 2), (4, 4)]


@pytest.mark.parametrize("kind,m,n,cx,cy", [(0, 296, 136, 1e4, 6e3), (1, 297, 137, 0.01, 0.013),
                                             (2, 296, 136, 0.13, 0.09)])
@pytest.mark.parametrize("form", [0, 1])
def test_gpu_streaming_launches_every_schedule(vdi, kind, m, n, cx, cy, form):
    """The streaming launches (every sweep count 0..4 their kernels are
    instantiated for, strips and segments that do not divide the level)
    compute the tiled launches' V-cycle bit for bit; in the mixed form this
    includes the level-0 up launch with its FP64 residual."""
    stream = [s for s in range(vdi.NUM_TILE_SHAPES) if vdi.tile_stream(s)]
    assert stream
    rng = np.random.default_rng(11 + kind)
    b = rng.standard_normal(m * n)
    prec = 1 if form else 0
    ns = len(STREAM_SCHEDULES)
    for i, (nu1, nu2) in enumerate(STREAM_SCHEDULES):
        ref = vdi.MGSystem(kind, m, n, cx, cy, 1, nu1, nu2)
        ref.set_form(form)
        ref.set_b(b)
        ref.set_z(np.zeros(m * n))
        ref.vcycles(3)
        z_ref = ref.get_z()
        ref.close()
        # two shapes per schedule, so that every streaming shape meets a
        # schedule other than the default
        for shape in sorted({stream[i % len(stream)], stream[(i + ns) % len(stream)]}):
            say("schedule", (nu1, nu2), "shape", vdi.tile_label(shape), "kind", kind, "form", form)
            g = vdi.MGSystem(kind, m, n, cx, cy, 1, nu1, nu2)
            g.set_form(form)
            for lv in range(g.levels):
                for which in (0, 1):
                    try:
                        g.set_tile(prec, which, lv, shape)
                    except vdi.VdiError:
                        pass   # level handled by the coarsest-level kernels
            if form:
                g.set_tile(1, 2, 0, shape)
            g.set_b(b)
            g.set_z(np.zeros(m * n))
            g.vcycles(3)
            np.testing.assert_array_equal(g.get_z(), z_ref)
            g.close()


@pytest.mark.parametrize("kind,m,n,cx,cy", [(0, 2048, 1024, 1e4, 6e3), (1, 2049, 1025, 0.01, 0.013)])
@pytest.mark.parametrize("form", [0, 1])
def test_gpu_streaming_launches_full_size(vdi, kind, m, n, cx, cy, form):
    """At an operator-level size (strips of 116 / 108 columns and segments of
    16 and 64 rows do not divide the levels) the streaming launches on every
    level, and in the mixed form the streaming level-0 up launch, give the
    tiled launches' V-cycles bit for bit."""
    rng = np.random.default_rng(21 + kind)
    b = rng.standard_normal(m * n)
    prec = 1 if form else 0
    ref = vdi.MGSystem(kind, m, n, cx, cy, 1)
    ref.set_form(form)
    ref.set_b(b)
    ref.set_z(np.zeros(m * n))
    ref.vcycles(2)
    z_ref = ref.get_z()
    levels = ref.levels
    ref.close()
    for ty in (16, 64):
        shape = next(s for s in range(vdi.NUM_TILE_SHAPES) if vdi.tile_stream(s) and vdi.tile_shape(s)[1] == ty)
        say("full size", (m, n), "kind", kind, "form", form, "shape", vdi.tile_label(shape))
        g = vdi.MGSystem(kind, m, n, cx, cy, 1)
        g.set_form(form)
        for lv in range(levels):
            for which in (0, 1):
                try:
                    g.set_tile(prec, which, lv, shape)
                except vdi.VdiError:
                    pass   # level handled by the coarsest-level kernels
        if form:
            g.set_tile(1, 2, 0, shape)
        g.set_b(b)
        g.set_z(np.zeros(m * n))
        g.vcycles(2)
        np.testing.assert_array_equal(g.get_z(), z_ref)
        g.close()


def test_gpu_autotuned_context_same_steps(vdi):
    from vdropimpact_amd.params import PhysicalConfig
    p = PhysicalConfig(m=256, n=128).params()
    a, c = vdi.Fluid2D(p), vdi.Fluid2D(p)
    a.init_fields()
    c.init_fields()
    for _ in range(4):
        a.step_forward(p.dt_reg)
        c.step_forward(p.dt_reg)
    a.autotune(3)
    say("tuned MAC tiles", a.tiles(vdi.MAC))
    for _ in range(8):
        a.step_forward(p.dt_reg)
        c.step_forward(p.dt_reg)
    for aid in (vdi.U, vdi.V, vdi.P, vdi.PG, vdi.H):
        x, y = a.get(aid), c.get(aid)
        assert np.linalg.norm(x - y) <= 1e-9 * max(np.linalg.norm(y), 1e-30)
    a.close()
    c.close()


def test_gpu_step_schedules(vdi):
    """Per-solver V(nu1, nu2): the step's defaults, a larger schedule after
    tuning (tiles whose halo no longer fits LDS move to a default shape), and
    the argument check."""
    from vdropimpact_amd.params import PhysicalConfig
    p = PhysicalConfig(m=256, n=128).params()
    a = vdi.Fluid2D(p)
    assert [a.schedule(k) for k in (vdi.MAC, vdi.FEM, vdi.VX, vdi.VY)] == [(3, 2, 3, 2), (3, 2, 2, 2), (4, 2, 0, 1),
                                                                          (4, 2, 0, 1)]
    a.init_fields()
    for _ in range(3):
        a.step_forward(p.dt_reg)
    a.autotune(3)
    for k in (vdi.MAC, vdi.FEM, vdi.VX, vdi.VY):
        a.set_schedule(k, 4, 4)
        assert a.schedule(k) == (4, 4, 4, 4)
    a.set_schedule(vdi.VX, 4, 2, 1, 1)
    assert a.schedule(vdi.VX) == (4, 2, 1, 1)
    for _ in range(3):
        a.step_forward(p.dt_reg)
    for bad in [(5, 2), (0, 0), (-1, 2), (3, 2, 5, 2), (3, 2, 0, 0)]:
        with pytest.raises(vdi.VdiError):
            a.set_schedule(vdi.MAC, *bad)
    a.close()
