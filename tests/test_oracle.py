"""Pins of the CPU oracle (runs without a GPU).

1. Golden vectors of the reference's own AVX2 fused kernels (tests/golden/, produced by
   tests/golden/make_ref_vectors.py from the reference sources): the FP64 restatement
   must reproduce the AVX2d kernel to round-off and the FP32 restatement the AVX2
   kernel, cell indices bit-exactly.
2. What the mathematics fixes independently of the restated formulas: B-spline closed
   forms and partition of unity, the longitudinal weights as partial sums of spline
   differences (discrete charge conservation in 1D), ntrim edge cases, Boris |u|
   conservation and rotation angle, Vay's exact E x B drift, exact gather of uniform and
   linear fields, the discrete continuity equation (div E - rho conserved by push +
   deposit + Yee), the Yee numerical dispersion of a vacuum plane wave, energy
   normalisations.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_simd_vectors.npz")


# ------------------------------------------------------------------------------------
# 1. golden vectors of the reference kernels
# ------------------------------------------------------------------------------------
@pytest.mark.parametrize("dims,order", [(d, o) for d in (2, 3) for o in (1, 2, 3)])
@pytest.mark.parametrize("prec", ["d", "s"])
def test_oracle_matches_reference_simd_golden(dims, order, prec):
    G = np.load(GOLD)
    key = f"{dims}d_s{order}"
    dt, dx0, lo = G[key + "_par"]
    lo = int(lo)
    dx = (dx0,) * dims
    nx = (6,) * dims if dims == 2 else (4,) * dims
    g = O.make_grid(dims, nx, order)
    r = np.float64 if prec == "d" else np.float32
    op = "f64" if prec == "d" else "f32"
    ix = G[key + "_in_ix"].copy()
    x = G[key + "_in_x"].astype(r)
    u = G[key + "_in_u"].astype(r)
    q = G[key + "_in_q"].astype(r)
    e = G[key + "_in_e"].astype(r)
    b = G[key + "_in_b"].astype(r)
    j = g.zeros(3, r)
    ene = O.advance_deposit(g, order, 0, -1.0, dx, dt, e, b, j, ix, x, u, q, prec=op)
    assert np.array_equal(ix, G[f"{key}_{prec}_ix"])
    tol = 1e-14 if prec == "d" else 4e-7
    assert np.abs(x - G[f"{key}_{prec}_x"]).max() <= tol
    assert np.abs(u - G[f"{key}_{prec}_u"]).max() <= tol * max(1.0, np.abs(u).max())
    rj = G[f"{key}_{prec}_j"]
    if not (prec == "s" and dims == 3 and order == 3):
        # the reference's FP32 AVX2 3D cubic deposit strides z planes by Dy
        # (simd/os-spec-current-avx2.c:1097); its FP64 twin is compared instead
        assert np.abs(j - rj).max() <= (1e-13 if prec == "d" else 2e-6) * np.abs(rj).max()
    rene = float(G[f"{key}_{prec}_ene"])
    assert abs(ene - rene) <= (1e-13 if prec == "d" else 1e-6) * abs(rene)


@pytest.mark.parametrize("order", [1, 2, 3])
def test_fast_reference_build_computes_the_same_push(order):
    """The -Ofast build of the reference's FP32 2D kernel, the one the CPU baseline times,
    pushes the golden inputs as the -O2 build does: same cells, positions and momenta to
    FP32 reassociation (a few ulp), same current and energy to 2e-6."""
    if not (O.ref_available("s") and O.ref_available("s_fast")):
        pytest.skip("reference AVX2 kernels not built (oracle/_ref)")
    G = np.load(GOLD)
    key = f"2d_s{order}"
    dt, dx0, lo = G[key + "_par"]
    lo = int(lo)
    g = O.make_grid(2, (6, 6), order)
    args = [G[key + "_in_" + a].astype(np.int32 if a == "ix" else np.float32)
            for a in ("ix", "x", "u", "q")]
    e = G[key + "_in_e"].astype(np.float32)
    b = G[key + "_in_b"].astype(np.float32)
    out = [O.ref_vadvance_deposit(2, order, *args, -1.0, e, b, lo, g.zeros(3, np.float32), lo,
                                  (dx0, dx0), dt, prec=p) for p in ("s", "s_fast")]
    (ix0, x0, u0, j0, e0), (ix1, x1, u1, j1, e1) = out
    assert np.array_equal(ix0, ix1)
    assert np.abs(x1 - x0).max() <= 4e-7
    assert np.abs(u1 - u0).max() <= 4e-7 * max(1.0, np.abs(u0).max())
    assert np.abs(j1 - j0).max() <= 2e-6 * np.abs(j0).max()
    assert abs(e1 - e0) <= 1e-6 * abs(e0)


def test_reference_fp32_3d_cubic_deposit_discrepancy_is_documented():
    """The FP32 AVX2 3D cubic reference deposit differs from the FP64 one (known bug of
    the reference); everything else of that kernel agrees."""
    G = np.load(GOLD)
    js, jd = G["3d_s3_s_j"], G["3d_s3_d_j"]
    assert np.abs(js - jd).max() > 1e-3 * np.abs(jd).max()
    assert np.array_equal(G["3d_s3_s_ix"], G["3d_s3_d_ix"])


# ------------------------------------------------------------------------------------
# 2. mathematics
# ------------------------------------------------------------------------------------
def bspline(order, d):
    """Cardinal B-spline of degree `order` at distance d (closed forms)."""
    a = abs(d)
    if order == 1:
        return max(0.0, 1 - a)
    if order == 2:
        if a <= 0.5:
            return 0.75 - a * a
        return 0.5 * (1.5 - a) ** 2 if a < 1.5 else 0.0
    if a <= 1:
        return 2.0 / 3 - a * a + a ** 3 / 2
    return (2 - a) ** 3 / 6 if a < 2 else 0.0


# node positions relative to the cell centre for stencil points LP..UP
NODES = {1: [-0.5, 0.5], 2: [-1.0, 0.0, 1.0], 3: [-1.5, -0.5, 0.5, 1.5]}


@pytest.mark.parametrize("order", (1, 2, 3))
def test_splines_closed_form_and_partition_of_unity(order):
    for x in np.linspace(-0.5, 0.4999, 57):
        s = O.spline(order, x)
        ref = [bspline(order, x - c) for c in NODES[order]]
        assert np.allclose(s, ref, atol=1e-14)
        assert abs(s.sum() - 1) < 1e-14
        # half-cell shifted spline: nodes moved by +1/2 (h = 0) or -1/2 (h = 1)
        h = 1 if x < 0 else 0
        sh = O.splineh(order, x, h)
        refh = [bspline(order, x - c - 0.5 + h) for c in NODES[order]]
        assert np.allclose(sh, refh, atol=1e-14)


@pytest.mark.parametrize("order", (1, 2, 3))
def test_longitudinal_weights_are_charge_conserving(order):
    """wl(k)/qn = sum_{k'<=k} (S(x0)_k' - S(x1)_k'): the 1D discrete continuity
    equation of the deposit."""
    rng = np.random.default_rng(order)
    for _ in range(50):
        x0 = rng.uniform(-0.5, 0.5)
        x1 = np.clip(x0 + rng.uniform(-0.5, 0.5), -0.5, 0.5)
        qn = rng.uniform(0.1, 2.0)
        w = O.wl(order, qn, x0, x1)
        cum = np.cumsum(O.spline(order, x0) - O.spline(order, x1))
        assert np.allclose(w[:-1] / qn, cum[:-1], atol=1e-13)
        assert w[-1] == 0


def test_ntrim_edges():
    assert O.ntrim(-0.5) == 0
    assert O.ntrim(np.nextafter(-0.5, -1)) == -1
    assert O.ntrim(0.5) == 1
    assert O.ntrim(np.nextafter(0.5, 0)) == 0
    assert O.ntrim(0.0) == 0


def test_boris_pure_magnetic_rotation():
    """|u| is conserved and the rotation angle per step is 2*atan(|t|), t = tem*B/gamma."""
    rng = np.random.default_rng(3)
    n = 64
    u = rng.normal(0, 1.0, (n, 3))
    q = np.ones(n)
    bp = rng.normal(0, 2.0, (n, 3))
    ep = np.zeros((n, 3))
    dt, rqm = 0.1, -1.0
    u0 = u.copy()
    O.dudt(0, rqm, dt, u, q, ep, bp)
    assert np.allclose(np.linalg.norm(u, axis=1), np.linalg.norm(u0, axis=1), rtol=1e-14)
    tem = 0.5 * dt / rqm
    gam = np.sqrt(1 + (u0 ** 2).sum(1))
    t = tem * bp / gam[:, None]
    bh = bp / np.linalg.norm(bp, axis=1)[:, None]
    perp0 = u0 - (u0 * bh).sum(1)[:, None] * bh
    perp1 = u - (u * bh).sum(1)[:, None] * bh
    cosang = (perp0 * perp1).sum(1) / np.linalg.norm(perp0, axis=1) / np.linalg.norm(perp1, axis=1)
    ang = np.arccos(np.clip(cosang, -1, 1))
    assert np.allclose(ang, 2 * np.arctan(np.linalg.norm(t, axis=1)), atol=1e-10)


def test_boris_pure_electric_acceleration():
    """Without B: u(n+1) = u(n) + 2*tem*E exactly (two half accelerations)."""
    u = np.array([[0.1, -0.2, 0.3]])
    e = np.array([[0.5, 0.25, -1.0]])
    O.dudt(0, -1.0, 0.1, u, np.ones(1), e, np.zeros((1, 3)))
    assert np.allclose(u, [[0.1, -0.2, 0.3]] + 2 * (0.5 * 0.1 / -1.0) * e[0], atol=1e-15)


@pytest.mark.parametrize("pusher", (0, 1))
def test_exb_drift(pusher):
    """A particle moving at v = E x B / B^2 feels no force: Vay keeps it exactly (to
    round-off) even relativistically; Boris only approximately."""
    ey, bz = 0.6, 1.0
    v = ey / bz  # drift along x
    gam = 1 / np.sqrt(1 - v * v)
    u = np.array([[gam * v, 0.0, 0.0]])
    for _ in range(20):
        O.dudt(pusher, -1.0, 0.2, u, np.ones(1), np.array([[0.0, ey, 0.0]]),
               np.array([[0.0, 0.0, bz]]))
    err = np.abs(u[0] - [gam * v, 0, 0]).max()
    if pusher == 1:
        assert err < 1e-12
    else:
        assert err > 1e-6  # the Boris pusher does not preserve the relativistic drift


def _vay_numpy(u, e, b, rqm, dt, drop_ustar=False):
    """Vay momentum update of spec/os-spec-push.f03:706-769 written out in numpy (FP64):
    half acceleration plus the rotation with the old velocity, the second half
    acceleration, the gamma of the new momentum from the quadratic of the implicit
    rotation, then the rotation.  drop_ustar removes the (u' . tau)^2 term of that gamma
    (a planted mistake, to show the pin below detects it)."""
    tem = 0.5 * dt / rqm
    e, b = tem * e, tem * b
    v = u / np.sqrt(1 + (u * u).sum(-1, keepdims=True))
    w = u + 2 * e + np.cross(v, b)
    ustar = (w * b).sum(-1, keepdims=True)
    sigma = 1 + (w * w).sum(-1, keepdims=True) - (b * b).sum(-1, keepdims=True)
    extra = 0.0 if drop_ustar else ustar * ustar
    gam = np.sqrt(0.5 * (sigma + np.sqrt(sigma * sigma + 4 * ((b * b).sum(-1, keepdims=True) + extra))))
    t = b / gam
    s = 1 / (1 + (t * t).sum(-1, keepdims=True))
    return s * (w + (w * t).sum(-1, keepdims=True) * t + np.cross(w, t))


def test_vay_force_free_with_momentum_along_b():
    """E perpendicular to B and v = E x B / B^2 + v_par b: the Lorentz force vanishes, and
    Vay's implicit rotation keeps u exactly.  The momentum has a component along B, so the
    (u' . tau)^2 term of the gamma solve (os-spec-push.f03:723-731) is exercised: without
    it the gamma, and then u, come out wrong."""
    bvec = np.array([0.3, -0.4, 1.2])
    bh = bvec / np.linalg.norm(bvec)
    evec = 0.5 * np.cross(bh, [1.0, 0.0, 0.0])  # perpendicular to B
    vdrift = np.cross(evec, bvec) / (bvec @ bvec)
    v = vdrift + 0.55 * bh
    assert v @ v < 1
    u0 = v / np.sqrt(1 - v @ v)
    u = u0[None, :].copy()
    for _ in range(20):
        O.dudt(1, -1.0, 0.2, u, np.ones(1), evec[None, :], bvec[None, :])
    assert np.abs(u[0] - u0).max() < 1e-12
    # the numpy statement reproduces the oracle on random inputs ...
    rng = np.random.default_rng(11)
    ur, er, br = rng.normal(0, 1, (3, 32, 3))
    uo = ur.copy()
    O.dudt(1, -1.0, 0.3, uo, np.ones(32), er, br)
    assert np.allclose(_vay_numpy(ur, er, br, -1.0, 0.3), uo, rtol=0, atol=1e-13)
    # ... and with the ustar term dropped it breaks the force-free motion
    up = u0[None, :].copy()
    for _ in range(20):
        up = _vay_numpy(up, evec[None, :], bvec[None, :], -1.0, 0.2, drop_ustar=True)
    assert np.abs(up[0] - u0).max() > 1e-5  # measured 7.6e-4


def test_vay_acceleration_along_b():
    """E, B and u parallel: no rotation, u(n+1) = u(n) + 2 tem E, and the energy sum is the
    time-centred q |u_c|^2 / (gamma_c + 1) of u_c = (u(n) + u(n+1)) / 2
    (os-spec-push.f03:772-792)."""
    bh = np.array([2.0, -1.0, 2.0]) / 3.0
    u0 = 0.7 * bh
    e, b = 0.8 * bh, 1.5 * bh
    u = u0[None, :].copy()
    ene = O.dudt(1, -1.0, 0.1, u, np.ones(1), e[None, :], b[None, :])
    tem = 0.5 * 0.1 / -1.0
    assert np.allclose(u[0], u0 + 2 * tem * e, rtol=0, atol=1e-15)
    uc = 0.5 * (u0 + u[0])
    assert ene == pytest.approx((uc @ uc) / (np.sqrt(1 + uc @ uc) + 1), rel=1e-14)


def test_history_rows_are_time_aligned():
    """Cold Langmuir oscillation (immobile ions, u_x = u0 sin(kx)): energy flows between
    the electrons and E_x at the plasma frequency.  With the reference's step order (push
    with E, B at integer time; the time-centred kinetic energy of row n; field energy of
    the same E, B; then fold and Yee, os-simulation.f03:588-661) the total of row n is
    conserved to O((w_p dt)^2) of the exchanged energy.  A field row taken after the Yee
    step, or kinetic and field rows one step apart, leave an O(w_p dt) oscillation: the
    planted shifts below must fail the bound."""
    nx, dx, dt, ppc = (64, 4), (0.1, 0.1), 0.07, (4, 4)
    cells = np.stack(np.meshgrid(np.arange(1, nx[0] + 1), np.arange(1, nx[1] + 1),
                                 indexing="ij"), -1).reshape(-1, 2)
    ix = np.repeat(cells, ppc[0] * ppc[1], 0).astype(np.int32)
    lat = np.stack(np.meshgrid((np.arange(ppc[0]) + 0.5) / ppc[0] - 0.5,
                               (np.arange(ppc[1]) + 0.5) / ppc[1] - 0.5, indexing="ij"),
                   -1).reshape(-1, 2)
    x = np.tile(lat, (len(cells), 1))
    k = 2 * np.pi / (nx[0] * dx[0])
    u = np.zeros((len(ix), 3))
    u[:, 0] = 0.01 * np.sin(k * (ix[:, 0] - 0.5 + x[:, 0]) * dx[0])
    q = np.full(len(ix), -1.0 / 16)
    s = O.Sim(2, nx, dx, dt, 2)
    s.add_species(-1.0, 0, ix, x, u, q)
    s.step(200)  # about 2.2 plasma periods
    fld = np.array(s.fld_hist).sum(1)
    kin = np.array(s.par_hist)[:, 0]
    exchanged = kin.max() - kin.min()
    assert fld.max() > 0.5 * exchanged  # the energy did flow into the field
    assert np.ptp(fld + kin) < 0.01 * exchanged  # measured 1.6e-3 (w_p dt = 0.07)
    for shifted in (fld[1:] + kin[:-1], fld[:-1] + kin[1:]):
        assert np.ptp(shifted) > 0.1 * exchanged  # measured 0.14


# Yee staggering: along axis a, E_a and B_b (b != a) sit half a cell from E_b, B_a
HALF = {0: ((True, False, False), (False, True, True)),
        1: ((False, True, False), (True, False, True)),
        2: ((False, False, True), (True, True, False))}


@pytest.mark.parametrize("dims,order", [(2, 1), (2, 2), (2, 3), (3, 1), (3, 2), (3, 3)])
def test_gather_uniform_and_linear_fields(dims, order):
    """Uniform fields are interpolated exactly.  A field linear along one axis is
    reproduced exactly (B-splines of every order reproduce linear functions) at the
    particle position, with the Yee staggering: along axis a the half-cell quantities
    (E_a and B_b, b != a) sit half a cell below the full-cell ones.  Full-cell node i sits
    at the particle coordinate ix + x = i - 1/2 for odd orders and i for even orders
    (x in [-1/2, 1/2) from the cell centre)."""
    rng = np.random.default_rng(10 * dims + order)
    nx = (8,) * dims
    g = O.make_grid(dims, nx, order)
    e = g.zeros()
    b = g.zeros()
    e[...] = [0.3, -0.7, 1.1]
    b[...] = [-2.0, 0.5, 0.25]
    n = 200
    ix = np.stack([rng.integers(2, nx[d], n) for d in range(dims)], 1).astype(np.int32)
    x = rng.uniform(-0.5, 0.5, (n, dims))
    ep, bp = O.get_emf(g, order, e, b, ix, x)
    assert np.allclose(ep, [0.3, -0.7, 1.1], atol=1e-14)
    assert np.allclose(bp, [-2.0, 0.5, 0.25], atol=1e-14)
    full = 0.5 if order % 2 else 0.0
    for ax in range(dims):
        node = np.arange(g.ld[ax]) - g.gcl[ax] + 1  # Fortran node index i
        shape = [1] * (dims + 1)
        shape[dims - 1 - ax] = -1
        pos = ix[:, ax] + x[:, ax]
        for comp in range(3):
            e[...] = 0
            b[...] = 0
            e[..., comp] = node.reshape(shape[:-1])
            b[..., comp] = node.reshape(shape[:-1])
            ep, bp = O.get_emf(g, order, e, b, ix, x)
            eh, bh = HALF[ax][0][comp], HALF[ax][1][comp]
            assert np.allclose(ep[:, comp], pos + full - 0.5 * eh, atol=1e-12)
            assert np.allclose(bp[:, comp], pos + full - 0.5 * bh, atol=1e-12)


@pytest.mark.parametrize("dims,order", [(2, 1), (2, 2), (2, 3), (3, 1), (3, 3)])
def test_gauss_law_conserved_by_push_deposit_yee(dims, order):
    """div E - rho is constant over time: the deposit satisfies the discrete continuity
    equation with the DEPOSIT_RHO charge, and the Yee E update adds -dt*J."""
    rng = np.random.default_rng(5 + dims + order)
    nx = (12, 10) if dims == 2 else (6, 5, 4)
    sim = O.Sim(dims, nx, (0.1,) * dims, 0.05, order)
    n = int(np.prod(nx)) * 4
    ix = np.stack([rng.integers(1, nx[d] + 1, n) for d in range(dims)], 1).astype(np.int32)
    x = rng.uniform(-0.5, 0.5, (n, dims))
    u = rng.normal(0, 0.5, (n, 3))
    q = np.full(n, -0.05)
    sim.add_species(-1.0, 0, ix, x, u, q)
    sim.add_species(1.0 / 100, 1, ix, -x * 0.5, u * 0.1, -q)  # a Vay-pushed ion species
    r0 = sim.gauss_residual()
    sim.step(8)
    r1 = sim.gauss_residual()
    assert np.abs(r1 - r0).max() < 1e-12 * max(1.0, np.abs(r0).max())


@pytest.mark.parametrize("axis", (0, 1))
def test_yee_plane_wave_dispersion(axis):
    """A vacuum plane wave E3 = cos(k x - w t) with the Yee dispersion relation
    sin(w dt/2)/dt = sin(k dx/2)/dx is an exact discrete solution."""
    nx = (32, 16) if axis == 0 else (16, 32)
    dx = (0.1, 0.1)
    dt = 0.06
    g = O.make_grid(2, nx, 1)
    L = nx[axis]
    k = 2 * np.pi * 3 / (L * dx[axis])
    w = 2 / dt * np.arcsin(dt / dx[axis] * np.sin(k * dx[axis] / 2))
    e, b = g.zeros(), g.zeros()
    idx = np.arange(g.ld[axis]) - g.gcl[axis]  # 0-based interior index, guards included
    xn = idx * dx[axis]               # E3 node position (integer)
    xh = (idx + 0.5) * dx[axis]       # B1/B2 position (half integer along the axis)
    sgn = 1 if axis == 0 else -1      # dB2/dt = +dE3/dx ; dB1/dt = -dE3/dy
    comp = 1 if axis == 0 else 0
    def put(f, c, prof):
        if axis == 0:
            f[:, :, c] = prof[None, :]
        else:
            f[:, :, c] = prof[:, None]
    put(e, 2, np.cos(k * xn))
    # B stored at integer time: B^0 = B^{-1/2} + (dt/2) dE3/dx
    bm = -sgn * np.cos(k * xh + w * dt / 2)
    b0 = bm + sgn * (dt / 2) * (np.cos(k * (xn + dx[axis])) - np.cos(k * xn)) / dx[axis]
    put(b, comp, b0)
    j = g.zeros()
    nsteps = 40
    for _ in range(nsteps):
        O.update_periodic(g, e, 0)
        O.update_periodic(g, b, 0)
        O.yee_advance(g, dx, dt, e, b, j)
    t = nsteps * dt
    exact = np.cos(k * xn - w * t)
    got = (g.interior(e)[0, :, 2] if axis == 0 else g.interior(e)[:, 0, 2])
    ref = exact[g.gcl[axis]:g.gcl[axis] + nx[axis]]
    assert np.abs(got - ref).max() < 1e-11


def test_field_and_kinetic_energy_normalisation():
    g = O.make_grid(2, (4, 5), 1)
    e, b = g.zeros(), g.zeros()
    e[...] = [1.0, 2.0, 3.0]
    b[...] = [0.5, 0.0, -1.0]
    out = O.field_energy(g, e, b, dvol=0.02)
    ncell = 20
    assert np.allclose(out, 0.5 * 0.02 * ncell * np.array([0.25, 0, 1.0, 1, 4, 9]))
    u = np.array([[3.0, 0.0, 4.0]])
    ke = O.kinetic_energy(u, np.array([2.0]))
    assert abs(ke - 2.0 * (np.sqrt(26.0) - 1)) < 1e-12  # q (gamma - 1)


def test_split_2d_segments_join_and_conserve_vz():
    """Virtual particles tile the trajectory (end of one = start of the next, in the
    frame of its cell), stay inside their cell, and their vz sum to the particle's."""
    rng = np.random.default_rng(9)
    n = 400
    xo = rng.uniform(-0.5, 0.5, (n, 2))
    xn = xo + rng.uniform(-0.9, 0.9, (n, 2))
    dxi = np.array([[O.ntrim(v) for v in row] for row in xn], np.int32)
    ixo = np.full((n, 2), 5, np.int32)
    rg = np.ones(n)
    u = np.zeros((n, 3))
    u[:, 2] = rng.normal(0, 1, n)
    q = np.ones(n)
    vpf, vpi = O.split(2, dxi, xn, ixo, xo, q, rg, u)
    k = 0
    for i in range(n):
        nvp = 1 + (dxi[i, 0] != 0) + (dxi[i, 1] != 0)
        seg = vpf[k:k + nvp]
        cells = vpi[k:k + nvp]
        assert np.all(np.abs(seg[:, :4]) <= 0.5 + 1e-12)
        # global coordinates of the segment ends are continuous
        ends = seg[:, 2:4] + cells
        starts = seg[:, 0:2] + cells
        assert np.allclose(ends[:-1], starts[1:], atol=1e-12)
        assert np.allclose(starts[0], xo[i] + 5) and np.allclose(ends[-1], xn[i] + 5)
        assert abs(seg[:, 5].sum() - u[i, 2]) < 1e-12
        k += nvp
    assert k == len(vpf)
