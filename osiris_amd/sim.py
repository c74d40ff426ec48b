"""Host-side mirror of the reference's species-push / EMF-solver seams over the C ABI.

``advance_deposit`` mirrors the reference's native fused kernel entry point
(FNAME(vadvance_deposit_{2d,3d}_sN), simd/os-spec-push-avx2.c:471-475) with the same
argument meaning: particles in the reference layout are pushed and their current is
added to ``j``.  ``yee_advance`` mirrors t_emf_solver%advance for periodic boundaries
(emf/solver/os-emf-solver-yee.f03:542-568).  ``Simulation`` is the resident path: the
particles and fields stay in HBM and ``run`` advances whole time steps in the order of
os-simulation.f03:498-666 (push + deposit, energy report, current fold, Yee).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import (OSR_BND_LINDMAN, OSR_BND_OPEN, OSR_BND_PERIODIC, OSR_BND_VPML, OSR_BORIS,
                   OSR_PER, OSR_PROP, OSR_SMOOTH, OSR_TENV, OSR_VAY, SimDesc, SmoothParams,
                   ZpulseBox, ZpulseWall, check, lib)


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


def device_count() -> int:
    return lib().osr_device_count()


def copy_probe(nbytes: int, reps: int = 5) -> float:
    """Mean time in ms of a plain device copy of ``nbytes`` bytes (bandwidth probe and
    traffic-counter calibration)."""
    ms = C.c_double()
    check(lib().osr_copy_probe(int(nbytes), int(reps), C.byref(ms)))
    return ms.value


def advance_deposit(dims, order, ix, x, u, q, rqm, e, b, emf_lo, j, j_lo, dx, dt, pusher=OSR_BORIS):
    """Push all particles of the arrays (in place) and ADD their current to ``j``.

    ``e``, ``b``, ``j`` are grids in the reference layout (numpy arrays of shape
    ``(ld_z, ld_y, ld_x, 3)`` or ``(ld_y, ld_x, 3)``) with ``emf_lo`` / ``j_lo`` lower guard
    cells.  Returns the time-centred energy sum (not normalised)."""
    ix = np.ascontiguousarray(ix, np.int32)
    x = np.ascontiguousarray(x, np.float32)
    u = np.ascontiguousarray(u, np.float32)
    q = np.ascontiguousarray(q, np.float32)
    e = np.ascontiguousarray(e, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    assert j.dtype == np.float32 and j.flags["C_CONTIGUOUS"]
    emf_size = np.array(list(reversed(e.shape[:-1])), np.int32)
    j_size = np.array(list(reversed(j.shape[:-1])), np.int32)
    emf_off = np.array([emf_lo] * dims, np.int32)
    j_off = np.array([j_lo] * dims, np.int32)
    dxa = np.array(list(dx), np.float64)
    ene = np.zeros(1, np.float64)
    n = len(q)
    check(lib().osr_advance_deposit(dims, order, pusher, _p(ix), _p(x), _p(u), _p(q), 1, n,
                                    float(rqm), _p(e), _p(b), _p(emf_size), _p(emf_off), _p(j),
                                    _p(j_size), _p(j_off), _p(dxa), float(dt), _p(ene)))
    return ix, x, u, float(ene[0])


def yee_advance(dims, nx, gcl, dx, dt, e, b, jay):
    """Periodic Yee step on guarded reference-layout arrays (in place)."""
    assert e.dtype == np.float32 and b.dtype == np.float32
    jay = np.ascontiguousarray(jay, np.float32)
    ld = np.array(list(reversed(e.shape[:-1])), np.int32)
    check(lib().osr_yee_advance(dims, _p(np.array(nx, np.int32)), _p(ld),
                                _p(np.array(gcl, np.int32)), _p(np.array(dx, np.float64)),
                                float(dt), _p(e), _p(b), _p(jay)))


def smooth_setup(dims, type, order=(0, 0, 0), dx=(1.0, 1.0, 1.0), gamma=1.0, swfj=None,
                 laserkdx_omega0=-1.0, laserkdx_dir=1, digital_A=None, digital_fc=None):
    """The nl_smooth section of the current (vdf/os-vdf-smooth.f90 read_nml_smooth +
    setup_smooth): ``type`` per direction as names ("binomial", "compensated", ...) or
    OSR_SMOOTH codes.  Returns (order[3], coef float32 (3, 31), s(k) at [:, 15 + k])."""
    p = SmoothParams()
    for d in range(3):
        t = type[d] if d < len(type) else "none"
        p.type[d] = OSR_SMOOTH[t] if isinstance(t, str) else int(t)
        p.order[d] = int(order[d]) if d < len(order) else 0
        for k in range(15):
            for m in range(3):
                p.swfj[d][k][m] = (float(swfj[d][k][m]) if swfj is not None and d < len(swfj)
                                   else 1.0)
        huge = -1.7976931348623157e308  # the namelist default -huge(1.0d0)
        p.digital_A[d] = (float(digital_A[d]) if digital_A is not None and d < len(digital_A)
                          else huge)
        p.digital_fc[d] = (float(digital_fc[d]) if digital_fc is not None and d < len(digital_fc)
                           else huge)
    p.laserkdx_omega0 = float(laserkdx_omega0)
    p.laserkdx_dir = int(laserkdx_dir)
    n = np.zeros(3, np.int32)
    coef = np.zeros((3, 31), np.float32)
    dxa = np.array(list(dx) + [1.0] * (3 - len(dx)), np.float64)
    check(lib().osr_smooth_setup(dims, C.byref(p), _p(dxa), float(gamma), _p(n), _p(coef)))
    return [int(v) for v in n], coef


def smooth_vdf(dims, f, order, coef):
    """smooth_vdf on a guarded reference-layout float32 array (in place)."""
    assert f.dtype == np.float32 and f.flags["C_CONTIGUOUS"]
    ld = np.array(list(reversed(f.shape[:-1])), np.int32)
    n = np.array(list(order) + [0] * (3 - len(order)), np.int32)
    check(lib().osr_smooth_vdf(dims, _p(ld), f.shape[-1], _p(n),
                               _p(np.ascontiguousarray(coef, np.float32)), _p(f)))


class Simulation:
    """A resident simulation on the GPU (one per grid; species share the grid)."""

    def __init__(self, dims, nx, dx, dt, order, tile=None, plain=False, kernel=0,
                 xbnd="periodic", reorder=0, vpml_size=0, vpml_diffuse=(False, False),
                 push_waves=0):
        """xbnd: "periodic"; "open" or "lindman" for Lindman absorbing field walls, "vpml"
        for perfectly matched layers of vpml_size cells (0: 8; vpml_diffuse: lower, upper
        wall mix the split B3 pair, 2D), both with absorbing
        particles at both x ends (fused path, nx a multiple of the tile width)."""
        d = SimDesc()
        d.path = 1 if plain else 0
        d.kernel = kernel
        d.xbnd = {"periodic": OSR_BND_PERIODIC, "open": OSR_BND_OPEN,
                  "lindman": OSR_BND_LINDMAN, "vpml": OSR_BND_VPML}[xbnd]
        d.vpml_size = vpml_size
        d.vpml_diffuse[0], d.vpml_diffuse[1] = int(vpml_diffuse[0]), int(vpml_diffuse[1])
        d.reorder = reorder  # in-tile particle order restored every `reorder` steps
        d.push_waves = push_waves  # waves per push workgroup (0: the tile's default)
        d.dims = dims
        for i in range(3):
            d.nx[i] = nx[i] if i < dims else 1
            d.dx[i] = dx[i] if i < dims else 1.0
            d.tile[i] = (tile[i] if tile is not None and i < dims else 0)
        d.dt = dt
        d.order = order
        self.dims, self.nx, self.dx, self.dt, self.order = dims, tuple(nx), tuple(dx), dt, order
        h = C.c_void_p()
        check(lib().osr_sim_create(C.byref(d), C.byref(h)))
        self._h = h
        self.nspecies = 0

    def close(self):
        if getattr(self, "_h", None):
            lib().osr_sim_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- species
    def add_species(self, rqm, pusher=OSR_BORIS):
        sid = C.c_int()
        check(lib().osr_sim_add_species(self._h, float(rqm), int(pusher), C.byref(sid)))
        self.nspecies += 1
        return sid.value

    def set_particles(self, sid, ix, x, u, q):
        ix = np.ascontiguousarray(ix, np.int32)
        x = np.ascontiguousarray(x, np.float32)
        u = np.ascontiguousarray(u, np.float32)
        q = np.ascontiguousarray(q, np.float32)
        check(lib().osr_species_set(self._h, sid, len(q), _p(ix), _p(x), _p(u), _p(q)))

    def generate(self, sid, ppc, density=1.0, ufl=(0, 0, 0), uth=(0, 0, 0), seed=0, random_pos=True):
        ppc = np.array(list(ppc) + [1] * (3 - len(ppc)), np.int32)
        check(lib().osr_species_generate(self._h, sid, _p(ppc), float(density),
                                         _p(np.array(ufl, np.float64)),
                                         _p(np.array(uth, np.float64)), int(seed), int(random_pos)))

    def count(self, sid):
        n = C.c_int64()
        check(lib().osr_species_count(self._h, sid, C.byref(n)))
        return n.value

    def get_particles(self, sid):
        n = self.count(sid)
        ix = np.zeros((n, self.dims), np.int32)
        x = np.zeros((n, self.dims), np.float32)
        u = np.zeros((n, 3), np.float32)
        q = np.zeros(n, np.float32)
        check(lib().osr_species_get(self._h, sid, _p(ix), _p(x), _p(u), _p(q)))
        return ix, x, u, q

    # ---- fields (reference layout)
    def set_fields(self, e=None, b=None, gcl=None):
        ref = e if e is not None else b
        ld = np.array(list(reversed(ref.shape[:-1])), np.int32)
        gcl = np.array(gcl if gcl is not None else [0] * self.dims, np.int32)
        e = None if e is None else np.ascontiguousarray(e, np.float32)
        b = None if b is None else np.ascontiguousarray(b, np.float32)
        check(lib().osr_fields_set(self._h, _p(e), _p(b), _p(ld), _p(gcl)))

    def get_fields(self, gcl=None, gcu=None):
        gcl = list(gcl) if gcl is not None else [0] * self.dims
        gcu = list(gcu) if gcu is not None else [0] * self.dims
        ld = [self.nx[d] + gcl[d] + gcu[d] for d in range(self.dims)]
        e = np.zeros(tuple(reversed(ld)) + (3,), np.float32)
        b = np.zeros_like(e)
        check(lib().osr_fields_get(self._h, _p(e), _p(b), _p(np.array(ld, np.int32)),
                                   _p(np.array(gcl, np.int32))))
        return e, b

    def fill_fields(self, e0=None, b0=None):
        check(lib().osr_fields_fill(self._h, _p(None if e0 is None else np.array(e0, np.float64)),
                                    _p(None if b0 is None else np.array(b0, np.float64))))

    def init_fields(self, sid, ufl1=0.0):
        """Add the initial field of species sid's charge (before the first step):
        the electrostatic field for ufl1 = 0 (2D), the beam field of a species drifting
        along x with fluid momentum ufl1 otherwise (2D and 3D)."""
        check(lib().osr_sim_init_fields(self._h, sid, float(ufl1)))

    def add_zpulse_wall(self, xmin=(0.0, 0.0), **kw):
        """Laser injection from an x wall: keyword arguments as the zpulse_wall input
        section (a0, omega0, phase, pol, pol_type, propagation, tenv_type, tenv_fwhm,
        tenv_rise, tenv_flat, tenv_fall, tenv_duration, tenv_range, per_type, per_center,
        per_w0, per_fwhm, per_focus, per_tem_mode, launch_time), with its defaults."""
        z = ZpulseWall()
        lib().osr_zpulse_wall_defaults(C.byref(z))
        names = {"propagation": OSR_PROP, "tenv_type": OSR_TENV, "per_type": OSR_PER}
        for k, v in kw.items():
            if k not in {f[0] for f in ZpulseWall._fields_} - {"xmin"}:
                raise ValueError(f"unknown zpulse_wall parameter {k}")
            if k in ("per_center", "per_tem_mode"):  # per transverse direction
                v = list(v) if isinstance(v, (tuple, list)) else [v]
                for i, vi in enumerate(v):
                    getattr(z, k)[i] = vi
                continue
            setattr(z, k, names[k][v] if k in names else v)
        for i, v in enumerate(xmin):
            z.xmin[i] = v
        check(lib().osr_sim_add_zpulse_wall(self._h, C.byref(z)))

    def add_zpulse(self, xmin=(0.0, 0.0, 0.0), **kw):
        """An in-box laser pulse along x (the zpulse input section: a0, omega0, phase, pol,
        pol_type, propagation, lon_type, lon_start, lon_fwhm, lon_rise, lon_flat, lon_fall,
        lon_duration, lon_x0, lon_range, per_type, per_center, per_w0, per_fwhm, per_focus,
        per_tem_mode, no_div_corr, with the reference's defaults) launched into the fields
        now, before the first step."""
        z = ZpulseBox()
        lib().osr_zpulse_defaults(C.byref(z))
        names = {"propagation": OSR_PROP, "lon_type": OSR_TENV, "per_type": OSR_PER}
        for k, v in kw.items():
            if k not in {f[0] for f in ZpulseBox._fields_} - {"xmin"}:
                raise ValueError(f"unknown zpulse parameter {k}")
            if k in ("per_center", "per_tem_mode"):  # per transverse direction
                v = list(v) if isinstance(v, (tuple, list)) else [v]
                for i, vi in enumerate(v):
                    getattr(z, k)[i] = vi
                continue
            setattr(z, k, names[k][v] if k in names else (int(v) if k == "no_div_corr" else v))
        for i, v in enumerate(xmin):
            z.xmin[i] = v
        check(lib().osr_sim_add_zpulse(self._h, C.byref(z)))

    def zpulse_column(self, k, t, second=False):
        """(b12, b3, active): what pulse k adds to B2, B3 of its column (plane in 3D, shape
        (nz, ny)) at time t."""
        shape = (self.nx[1],) if self.dims == 2 else (self.nx[2], self.nx[1])
        b12 = np.zeros(shape, np.float32)
        b3 = np.zeros(shape, np.float32)
        act = C.c_int()
        check(lib().osr_sim_zpulse_column(self._h, k, int(second), float(t), _p(b12), _p(b3),
                                          C.byref(act)))
        return b12, b3, bool(act.value)

    def set_smooth(self, order, coef):
        """Smooth the current every step (order per direction, coef from smooth_setup)."""
        n = np.array(list(order) + [0] * (3 - len(order)), np.int32)
        check(lib().osr_sim_set_smooth(self._h, _p(n), _p(np.ascontiguousarray(coef, np.float32))))

    def write_hist(self, dir, names=None, first=0, count=None):
        """HIST/fld_ene, par_ene, parNN_ene rows of the history (reference text formats)."""
        if count is None:
            count = self.steps - first
        arr = None
        if names is not None:
            arr = (C.c_char_p * len(names))(*[str(n).encode() for n in names])
        check(lib().osr_sim_write_hist(self._h, str(dir).encode(), arr, int(first), int(count)))

    def save_field(self, quant, ms_dir, xmin=None):
        """OSIRIS-style MS/FLD/<q>/<q>-<n>.zdf dump of e1..e3 (0..2) or b1..b3 (3..5)."""
        xm = None if xmin is None else np.array(list(xmin) + [0.0] * (3 - len(xmin)), np.float64)
        check(lib().osr_sim_save_field(self._h, int(quant), str(ms_dir).encode(), _p(xm)))

    # ---- stepping and diagnostics
    def run(self, nsteps):
        check(lib().osr_sim_run(self._h, int(nsteps)))

    @property
    def steps(self):
        s = C.c_int64()
        check(lib().osr_sim_steps(self._h, C.byref(s)))
        return s.value

    def history(self, first=0, count=None):
        if count is None:
            count = self.steps - first
        fld = np.zeros((count, 6))
        par = np.zeros((count, max(self.nspecies, 1)))
        check(lib().osr_sim_history(self._h, first, count, _p(fld), _p(par)))
        return fld, par[:, :self.nspecies]

    def gauss(self, want_field=False):
        """max |div E - rho|, and with want_field the residual field."""
        n = int(np.prod(self.nx[:self.dims]))
        res = np.zeros(n) if want_field else None
        m = np.zeros(1)
        check(lib().osr_sim_gauss(self._h, _p(res), None, _p(m)))
        if want_field:
            return float(m[0]), res.reshape(tuple(reversed(self.nx[:self.dims])))
        return float(m[0])

    def charge_density(self):
        """rho on the interior cell corners (FP64), shape (nz, ny, nx) / (ny, nx)."""
        n = int(np.prod(self.nx[:self.dims]))
        rho = np.zeros(n)
        m = np.zeros(1)
        check(lib().osr_sim_gauss(self._h, None, _p(rho), _p(m)))
        return rho.reshape(tuple(reversed(self.nx[:self.dims])))

    @property
    def relayouts(self):
        n = C.c_int64()
        check(lib().osr_sim_relayouts(self._h, C.byref(n)))
        return n.value

    def time_kernels(self, nsteps):
        ms = np.zeros(6)
        check(lib().osr_sim_time_kernels(self._h, int(nsteps), _p(ms)))
        # tile_fill runs beside yee (the field advance with its walls); finalize is the
        # step end after both
        return {"push_deposit": ms[0], "tile_fill": ms[1], "yee": ms[2], "finalize": ms[3],
                "step": ms[4]}

    @property
    def launch_config(self):
        """{'tile', 'push_waves', 'compiled_tile'} the fused push resolved to."""
        t = (C.c_int * 3)()
        w, k = C.c_int(), C.c_int()
        check(lib().osr_sim_launch_config(self._h, t, C.byref(w), C.byref(k)))
        return {"tile": tuple(t[:self.dims]), "push_waves": w.value,
                "compiled_tile": bool(k.value)}

    @property
    def push_waves(self):
        return self.launch_config["push_waves"]

    @property
    def footprint(self):
        n = C.c_int64()
        check(lib().osr_sim_footprint(self._h, C.byref(n)))
        return n.value


__all__ = ["Simulation", "advance_deposit", "yee_advance", "device_count", "OSR_BORIS", "OSR_VAY"]
