"""The autotuner's search space (tile shape, waves per push workgroup, push-kernel
variant) changes no result bit: the current is summed in integers, so every candidate
gives the same particles and fields, and the tuned choice is purely one of speed."""
import numpy as np
import pytest

import osiris_amd as oa
from osiris_amd import autotune

pytestmark = pytest.mark.gpu


def _two_stream(cfg, nx=(128, 96)):
    sim = oa.Simulation(2, nx, (0.05, 0.05), 0.035, 2, tile=cfg["tile"],
                        push_waves=cfg.get("push_waves", 0), kernel=cfg.get("kernel", 0))
    for k, ufl in enumerate((0.2, -0.2)):
        sid = sim.add_species(-1.0)
        sim.generate(sid, (4, 4), density=0.5, ufl=(ufl, 0, 0), uth=(0.05,) * 3, seed=91 + k)
    return sim


def _magnetized_3d(cfg, nx=(24, 16, 16)):
    sim = oa.Simulation(3, nx, (0.1, 0.1, 0.1), 0.05, 3, tile=cfg["tile"],
                        kernel=cfg.get("kernel", 0))
    sid = sim.add_species(-1.0, oa.OSR_VAY)
    sim.generate(sid, (2, 2, 2), density=1.0, uth=(0.1,) * 3, seed=5)
    sim.fill_fields(b0=(0.0, 0.0, 0.5))
    return sim


def _state(sim, nspec):
    return sim.get_fields(), [sim.get_particles(s) for s in range(nspec)], sim.history()


def _assert_same(a, b):
    (fa, pa, ha), (fb, pb, hb) = a, b
    for x, y in zip(fa, fb):
        assert np.array_equal(x, y)
    for sa, sb in zip(pa, pb):
        for x, y in zip(sa, sb):
            assert np.array_equal(x, y)
    assert np.array_equal(ha[0], hb[0])  # field energies: FP64 sums over the same grid
    # kinetic energies: FP64 partial sums per tile and wave, equal to round-off
    assert np.allclose(ha[1], hb[1], rtol=1e-12, atol=0)


def test_every_2d_candidate_resolves_and_changes_no_result():
    ref = None
    for cfg in autotune.DEFAULT_2D:
        print(f"case {autotune.label(cfg)}", flush=True)
        with _two_stream(cfg) as sim:
            lc = sim.launch_config
            assert lc["tile"] == cfg["tile"] and lc["push_waves"] == cfg["push_waves"]
            assert lc["compiled_tile"]
            sim.run(30)
            st = _state(sim, 2)
        if ref is None:
            ref = st
        else:
            _assert_same(ref, st)


def test_every_3d_candidate_changes_no_result():
    ref = None
    for cfg in autotune.DEFAULT_3D:
        print(f"case {autotune.label(cfg)}", flush=True)
        with _magnetized_3d(cfg) as sim:
            lc = sim.launch_config
            assert lc["tile"] == cfg["tile"]
            assert lc["compiled_tile"] == (cfg["kernel"] == 0)
            sim.run(12)
            st = _state(sim, 1)
        if ref is None:
            ref = st
        else:
            _assert_same(ref, st)


@pytest.mark.parametrize("dims", [2, 3])
def test_tuned_choice_matches_the_default_bit_for_bit(dims):
    make = _two_stream if dims == 2 else _magnetized_3d
    best, ms = autotune.tune(make, f"test-{dims}d", dims=dims, warmup=1, steps=2, cache=None)
    print(f"tuned {autotune.label(best)} from {ms}", flush=True)
    assert autotune.label(best) in ms and len(ms) == len(
        autotune.DEFAULT_2D if dims == 2 else autotune.DEFAULT_3D)
    nsteps, nspec = (30, 2) if dims == 2 else (12, 1)
    default = {"tile": (16, 16) if dims == 2 else (8, 8, 4)}
    out = []
    for cfg in (default, best):
        with make(cfg) as sim:
            sim.run(nsteps)
            out.append(_state(sim, nspec))
    _assert_same(*out)


def test_push_waves_requests_are_checked():
    with pytest.raises(oa.OsrError, match="push_waves must be"):
        _two_stream({"tile": (16, 16), "push_waves": 2})
    with pytest.raises(oa.OsrError, match="no push kernel with 8 waves"):
        _two_stream({"tile": (8, 8), "push_waves": 8})
    with pytest.raises(oa.OsrError, match="no push kernel with 8 waves"):
        _two_stream({"tile": (16, 16), "push_waves": 8, "kernel": 1})


@pytest.mark.parametrize("order,pusher", [(1, oa.OSR_BORIS), (3, oa.OSR_BORIS),
                                          (1, oa.OSR_VAY), (2, oa.OSR_VAY), (3, oa.OSR_VAY)])
def test_waves_change_no_result_every_order_and_pusher(order, pusher):
    """Every compile-time 2D instantiation with 4 and with 8 waves per workgroup (16x16
    and 32x16 tiles) gives the bits of the generic kernel, for each order and pusher."""
    nx, dx, dt = (64, 48), (0.05, 0.05), 0.035
    ref = None
    for tile, waves, kernel in (((16, 16), 0, 1), ((16, 16), 4, 0), ((16, 16), 8, 0),
                                ((32, 16), 4, 0), ((32, 16), 8, 0)):
        print(f"case order={order} pusher={pusher} tile={tile} waves={waves} kernel={kernel}",
              flush=True)
        with oa.Simulation(2, nx, dx, dt, order, tile=tile, push_waves=waves,
                           kernel=kernel) as sim:
            for k, ufl in enumerate((0.2, -0.2)):
                sid = sim.add_species(-1.0, pusher)
                sim.generate(sid, (4, 4), density=0.5, ufl=(ufl, 0, 0), uth=(0.05,) * 3,
                             seed=71 + k)
            sim.fill_fields(b0=(0.1, 0.0, 0.3))
            sim.run(20)
            st = _state(sim, 2)
        if ref is None:
            ref = st
        else:
            _assert_same(ref, st)
