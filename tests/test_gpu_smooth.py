"""Current smoothing on the GPU against the oracle.

* drop-in (osr_smooth_vdf): the FP32 filter of the single-precision build, bit-exact
  against the FP32 oracle (same products, same summation order, no contraction);
* resident simulation (osr_sim_set_smooth): energy histories against the FP64 oracle run
  with the same filter, within the 1e-4 relative tolerance of the other resident tests.
"""
import numpy as np
import pytest

import osiris_amd as oa
from oracle import oracle as O
from tests.pic_inputs import thermal_plasma

pytestmark = pytest.mark.gpu

FILTERS = [
    ("binomial", dict(type=("binomial", "binomial", "binomial"), order=(2, 1, 3))),
    ("compensated", dict(type=("compensated", "compensated", "none"), order=(3, 2, 0))),
    ("custom", dict(type=("custom", "none", "custom"), order=(2, 0, 1))),
    ("digital", dict(type=("digital", "tri8", "5pass"), order=(5, 0, 0),
                     digital_A=(40.0, 0.0, 0.0), digital_fc=(0.3, 0.0, 0.0))),
]


@pytest.mark.parametrize("dims", (2, 3))
@pytest.mark.parametrize("name,spec", FILTERS)
def test_smooth_vdf_bit_exact(dims, name, spec):
    spec = dict(spec)
    spec["type"] = spec["type"][:dims]
    spec["order"] = spec["order"][:dims]
    if name == "custom":
        spec["swfj"] = np.random.default_rng(1).uniform(0.2, 2.0, (3, 15, 3))
    for k in ("digital_A", "digital_fc"):
        if k in spec:
            spec[k] = spec[k][:dims]
    order, coef = oa.smooth_setup(dims, **spec)
    nx = (40, 28) if dims == 2 else (20, 14, 10)
    g = O.Grid(dims, nx, tuple(max(2, n) for n in order[:dims]),
               tuple(max(3, n + 1) for n in order[:dims]))
    rng = np.random.default_rng(7 + dims)
    f = rng.standard_normal(tuple(reversed(g.ld)) + (3,)).astype(np.float32)
    want = f.copy()
    O.smooth(g, want, order, coef.astype(np.float64), "f32")
    got = f.copy()
    print(f"case dims={dims} filter={name} order={order}", flush=True)
    oa.smooth_vdf(dims, got, order, coef)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert not np.array_equal(got, f)


@pytest.mark.parametrize("plain", (False, True))
def test_resident_smoothing_vs_oracle(plain):
    rng = np.random.default_rng(60)
    nx, dx, dt, order, steps = (48, 32), (0.1, 0.1), 0.07, 2, 30
    ix, x, u, q = thermal_plasma(rng, 2, nx, 9, 0.08)
    sorder, coef = oa.smooth_setup(2, ("compensated", "binomial"), (2, 2))
    hist = []
    for smooth in (False, True):
        print(f"case plain={plain} smooth={smooth}", flush=True)
        with oa.Simulation(2, nx, dx, dt, order, plain=plain) as sim:
            sid = sim.add_species(-1.0)
            sim.set_particles(sid, ix, x, u, q)
            if smooth:
                sim.set_smooth(sorder, coef)
            sim.run(steps)
            hist.append(sim.history())
    ref = O.Sim(2, nx, dx, dt, order, smooth_order=tuple(sorder),
                smooth_coef=coef.astype(np.float64))
    ref.add_species(-1.0, 0, ix, x, u, q)
    ref.step(steps)
    fld, par = hist[1]
    rfld, rpar = np.array(ref.fld_hist), np.array(ref.par_hist)
    assert np.max(np.abs(par - rpar) / np.abs(rpar)) <= 1e-4
    fscale = np.maximum(rfld.sum(1, keepdims=True), 1e-300)
    assert np.max(np.abs(fld - rfld) / fscale) <= 1e-4
    # the filter is active: smoothing removes grid-scale field energy
    assert hist[0][0][-1].sum() > 1.05 * fld[-1].sum()


def test_smoothing_switched_off_again():
    rng = np.random.default_rng(61)
    nx, dx, dt = (32, 32), (0.1, 0.1), 0.07
    ix, x, u, q = thermal_plasma(rng, 2, nx, 4, 0.08)
    sorder, coef = oa.smooth_setup(2, ("binomial", "binomial"), (1, 1))
    out = []
    for mode in ("never", "on-off"):
        with oa.Simulation(2, nx, dx, dt, 2) as sim:
            sid = sim.add_species(-1.0)
            sim.set_particles(sid, ix, x, u, q)
            if mode == "on-off":
                sim.set_smooth(sorder, coef)
                sim.set_smooth((0, 0), np.zeros((3, 31), np.float32))
            sim.run(10)
            out.append(sim.history()[0])
    assert np.array_equal(out[0], out[1])
