"""ctypes front end of the CPU oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg use this
module, and only as the checker.  It wraps two libraries built by ``oracle/Makefile``:

* ``lib/liboracle_f64.so`` / ``lib/liboracle_f32.so`` -- the C restatement of the
  reference Fortran hot path (pic_oracle.c), in FP64 and FP32;
* ``_ref/libosiris_simd_{s,d}.so`` -- the reference's own AVX2 fused push+deposit
  kernels (simd/os-spec-push-avx2{,d}.c), compiled from the reference sources.

Array conventions follow the reference: grids are ``(3, nx+gc, ny+gc[, nz+gc])`` in
Fortran order (component fastest), stored here as numpy arrays of shape
``(ld_z, ld_y, ld_x, 3)`` (C order, same memory); particle arrays are ``ix (N, D)``
int32 with 1-based cell indices, ``x (N, D)``, ``u (N, 3)``, ``q (N,)``.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS: dict = {}


class _Grid(C.Structure):
    _fields_ = [("dims", C.c_int), ("nx", C.c_int * 3), ("gcl", C.c_int * 3),
                ("gcu", C.c_int * 3)]


def _lib(prec: str):
    """Load the oracle restatement for 'f64' or 'f32'."""
    if prec not in _LIBS:
        path = os.path.join(_HERE, "lib", f"liboracle_{prec}.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
        L = C.CDLL(path)
        P = C.c_void_p
        L.orc_advance_deposit.restype = C.c_double
        real = C.c_double if prec == "f64" else C.c_float
        L.orc_advance_deposit.argtypes = [P, C.c_int, C.c_int, real, P, C.c_double, P, P, P,
                                          P, P, P, P, C.c_int]
        L.orc_dudt_boris.restype = C.c_double
        L.orc_dudt_boris.argtypes = [real, C.c_double, P, P, P, P, C.c_int]
        L.orc_dudt_vay.restype = C.c_double
        L.orc_dudt_vay.argtypes = [real, C.c_double, P, P, P, P, C.c_int]
        L.orc_get_emf.argtypes = [P, C.c_int, P, P, P, P, C.c_int, P, P]
        L.orc_spline.argtypes = [C.c_int, real, P]
        L.orc_splineh.argtypes = [C.c_int, real, C.c_int, P]
        L.orc_wl.argtypes = [C.c_int, real, real, real, P]
        L.orc_ntrim.argtypes = [real]
        L.orc_split.argtypes = [C.c_int, C.c_int, P, P, P, P, P, P, P, P, P]
        L.orc_split.restype = C.c_int
        L.orc_update_periodic.argtypes = [P, P, C.c_int, C.c_int]
        L.orc_particles_periodic.argtypes = [P, P, C.c_int]
        L.orc_yee_advance.argtypes = [P, P, C.c_double, P, P, P]
        L.orc_field_energy.argtypes = [P, P, P, C.c_double, P]
        L.orc_kinetic_energy.restype = C.c_double
        L.orc_kinetic_energy.argtypes = [P, P, C.c_int]
        L.orc_deposit_rho.argtypes = [P, C.c_int, P, P, P, C.c_int, P]
        L.orc_div.argtypes = [P, P, P, P]
        L.orc_step.argtypes = [P, C.c_int, P, C.c_double, P, P, P, C.c_int, P, P, P, P, P, P,
                               P, P, P, P, P]
        L.orc_smooth.argtypes = [P, C.c_int, P, P, P]
        L.orc_step_open.argtypes = [P, C.c_int, P, C.c_double, P, P, P, P, P, C.c_int, P, P,
                                    P, P, P, P, P, P, P]
        L.orc_open_boundary_vpml.argtypes = [P, P, P, P]
        L.orc_yee_advance_vpml.argtypes = [P, P, C.c_double, P, P, P, P]
        L.orc_vpml_coefs.argtypes = [P, C.c_double if prec == "f64" else C.c_float]
        L.orc_open_boundary.argtypes = [P, P, P, P]
        L.orc_yee_advance_open.argtypes = [P, P, C.c_double, P, P, P, P]
        L.orc_particles_open.argtypes = [P, P, P, P, P, C.c_int]
        L.orc_particles_open.restype = C.c_int
        L.orc_zpulse_wall_2d.argtypes = [P, P, C.c_int, P, C.c_double, C.c_double, P]
        L.orc_es_solver_2d.argtypes = [P, P, P, P]
        L.orc_update_periodic_from.argtypes = [P, P, C.c_int, C.c_int, C.c_int]
        L.orc_es_solver_2d_beam.argtypes = [P, P, P, C.c_float, C.c_int, P, P]
        L.orc_es_solver_3d_beam.argtypes = [P, P, P, C.c_float, C.c_int, P, P]
        L.orc_zpulse_box_add.argtypes = [P, P, C.c_int, P, P, P]
        L._real = np.float64 if prec == "f64" else np.float32
        _LIBS[prec] = L
    return _LIBS[prec]


def _lindman_struct(real):
    """orc_lindman of pic_oracle.c for the working precision."""
    R = C.c_double if real == np.float64 else C.c_float

    class Lindman(C.Structure):
        _fields_ = [("disp_corr", R), ("dtdxi", R * 3), ("ebl", C.c_void_p),
                    ("bbl", C.c_void_p), ("ebu", C.c_void_p)]
    return Lindman


class LindmanWalls:
    """Lindman x walls of a 2D or 3D grid (new_lindman, emf/boundary/os-emf-lindman.f90:
    190-330): zeroed wall buffers and the coefficients in the working precision."""

    def __init__(self, grid, dx, dt, real):
        ld1 = grid.nx[1] + grid.gcl[1] + grid.gcu[1]
        ld2 = grid.nx[2] + grid.gcl[2] + grid.gcu[2] if grid.dims > 2 else 1
        self.ebl = np.zeros((ld2, ld1, 2, 3), real)
        self.bbl = np.zeros((ld2, ld1, 2, 3), real)
        self.ebu = np.zeros((ld2, ld1, 3, 3), real)
        S = _lindman_struct(real)
        self.c = S()
        r = real
        self.c.disp_corr = r((1.0 - dt / dx[0]) / (1.0 + dt / dx[0]))
        for d in range(len(dx)):
            self.c.dtdxi[d] = r(r(dt / dx[d]) * r(0.5))
        self.c.ebl, self.c.bbl, self.c.ebu = (a.ctypes.data for a in (self.ebl, self.bbl, self.ebu))


VPML_MAX = 64  # ORC_VPML_MAX of pic_oracle.c


def _vpml_struct(real):
    """orc_vpml of pic_oracle.c for the working precision."""
    R = C.c_double if real == np.float64 else C.c_float

    class Vpml(C.Structure):
        _fields_ = [("n", C.c_int), ("gc1", C.c_int), ("gc2", C.c_int), ("gc3", C.c_int),
                    ("gc4", C.c_int), ("diffuse", C.c_int * 2), ("dtdx2", R), ("dtdx3", R),
                    ("coef", R * VPML_MAX * 3 * 2 * 2), ("w", C.c_void_p * 2 * 2)]
    return Vpml


class VpmlWalls:
    """VPML x walls of a 2D or 3D grid (setup_vpml, emf/boundary/os-emf-vpml.f03:270-474):
    both walls zeroed, n cells thick, lower guards max(gcl, 1) and upper guards
    max(gcu, 2) in y (and z); the attenuation coefficients from setup_att_coeff
    (:3586-3689) in the working precision.  e[wl], b[wl] have shape (nz rows, ny rows, n,
    components) (wl 0 lower, 1 upper; 4 split components in 2D, 6 in 3D; one z row in
    2D)."""

    def __init__(self, grid, dx, dt, real, n=8, diffuse=(False, False)):
        if not 4 <= n <= VPML_MAX:
            raise ValueError(f"vpml_bnd_size must be in 4..{VPML_MAX}")
        S = _vpml_struct(real)
        self.c = S()
        self.c.n = n
        if any(diffuse) and grid.dims != 2:
            raise ValueError("vpml_diffuse is a 2D option")
        self.c.diffuse[0], self.c.diffuse[1] = int(diffuse[0]), int(diffuse[1])
        self.c.gc1 = max(grid.gcl[1], 1)
        self.c.gc2 = max(grid.gcu[1], 2)
        self.c.dtdx2 = real(dt / dx[1])
        nrow2 = grid.nx[1] + self.c.gc1 + self.c.gc2
        nrow3, nc = 1, 4
        if grid.dims > 2:
            self.c.gc3 = max(grid.gcl[2], 1)
            self.c.gc4 = max(grid.gcu[2], 2)
            self.c.dtdx3 = real(dt / dx[2])
            nrow3, nc = grid.nx[2] + self.c.gc3 + self.c.gc4, 6
        self.e = [np.zeros((nrow3, nrow2, n, nc), real) for _ in range(2)]
        self.b = [np.zeros((nrow3, nrow2, n, nc), real) for _ in range(2)]
        for wl in range(2):
            self.c.w[wl][0] = self.e[wl].ctypes.data
            self.c.w[wl][1] = self.b[wl].ctypes.data
        L = _lib("f64" if real == np.float64 else "f32")
        L.orc_vpml_coefs(C.byref(self.c), float(real(dt / dx[0])))

    def coefs(self):
        """[wall][E, B][1..3][cell] as an array (2, 2, 3, n)."""
        a = np.ctypeslib.as_array(self.c.coef)
        return a[..., :self.c.n].copy()


class ZpulseParams(C.Structure):
    """orc_zpulse of pic_oracle.c (parameters after read_input_zpulse_wall)."""
    _fields_ = [("a0", C.c_double), ("omega0", C.c_double), ("phase0", C.c_double),
                ("pol", C.c_double), ("pol_type", C.c_int), ("backward", C.c_int),
                ("tenv_type", C.c_int), ("per_type", C.c_int), ("tem_mode", C.c_int * 2),
                ("rise", C.c_double), ("flat", C.c_double),
                ("fall", C.c_double), ("duration", C.c_double), ("range", C.c_double),
                ("center", C.c_double * 2), ("w0", C.c_double), ("focus", C.c_double),
                ("launch_time", C.c_double), ("xmin", C.c_double * 3), ("xmax", C.c_double * 3)]


ZPULSE_DEFAULTS = dict(a0=1.0, omega0=10.0, phase=0.0, pol_type=0, pol=90.0,
                       propagation="forward", tenv_type="polynomial", tenv_fwhm=-1.0,
                       tenv_rise=0.0, tenv_flat=0.0, tenv_fall=0.0, tenv_duration=0.0,
                       tenv_range=None, per_type="plane", per_center=None, per_w0=0.0,
                       per_fwhm=0.0, per_focus=None, per_tem_mode=0, launch_time=0.0)


def zpulse_wall_params(nx, dx, xmin=(0.0, 0.0, 0.0), **kw):
    """read_input_zpulse_wall (zpulse/os-zpulse-wall.f03:196-628) for direction 1 of a 2D
    or 3D box (len(nx)): the reference's defaults (:196-239) updated by kw (per_center
    and per_tem_mode take one value per transverse direction); returns ZpulseParams."""
    unknown = set(kw) - set(ZPULSE_DEFAULTS)
    if unknown:
        raise ValueError(f"unknown zpulse_wall parameters {sorted(unknown)}")
    a = dict(ZPULSE_DEFAULTS, **kw)
    pi_180 = 0.0174532925199432957692369  # os-math.f90:15
    p = ZpulseParams()
    p.a0, p.omega0 = a["a0"], a["omega0"]
    p.phase0 = a["phase"] * pi_180
    if a["pol_type"] not in (-1, 0, 1):
        raise ValueError("pol_type must be -1, 0 or 1")
    p.pol_type = a["pol_type"]
    p.pol = a["pol"] * pi_180
    p.backward = {"forward": 0, "backward": 1}[a["propagation"]]
    dims = len(nx)
    for d in range(dims):
        p.xmin[d] = xmin[d] if d < len(xmin) else 0.0
        p.xmax[d] = p.xmin[d] + nx[d] * dx[d]
    p.tenv_type = {"polynomial": 0, "sin2": 1, "gaussian": 2}[a["tenv_type"]]
    if p.tenv_type < 2:
        if a["tenv_fwhm"] > 0:
            p.rise, p.flat, p.fall = a["tenv_fwhm"] / 2.0, 0.0, a["tenv_fwhm"] / 2.0
        else:
            p.rise, p.flat, p.fall = a["tenv_rise"], a["tenv_flat"], a["tenv_fall"]
    else:
        p.duration = (a["tenv_fwhm"] / np.sqrt(2.0 * np.log(2.0)) if a["tenv_fwhm"] > 0
                      else a["tenv_duration"])
        if a["tenv_range"] is None:
            raise ValueError("a gaussian envelope needs tenv_range")
        p.range = a["tenv_range"]
    p.per_type = {"plane": 0, "gaussian": 1, "hermite": 1}[a["per_type"]]
    if p.per_type == 1:
        modes = np.atleast_1d(a["per_tem_mode"])
        for d in range(dims - 1):
            p.tem_mode[d] = int(modes[d]) if d < len(modes) else 0
        p.w0 = a["per_fwhm"] / np.sqrt(4.0 * np.log(2.0)) if a["per_fwhm"] > 0 else a["per_w0"]
        if p.w0 <= 0 or a["per_focus"] is None:
            raise ValueError("a gaussian beam needs per_w0 (or per_fwhm) and per_focus")
        p.focus = a["per_focus"]
    cen = a["per_center"]
    cen = [None] * (dims - 1) if cen is None else list(np.atleast_1d(cen))
    for d in range(dims - 1):
        c = cen[d] if d < len(cen) else None
        p.center[d] = 0.5 * (p.xmin[d + 1] + p.xmax[d + 1]) if c is None else c
    p.launch_time = a["launch_time"]
    if p.backward:
        p.focus = -p.focus
    return p


class ZpulseBoxParams(C.Structure):
    """orc_zpulse_box of pic_oracle.c (parameters after read_input_zpulse)."""
    _fields_ = [("a0", C.c_double), ("omega0", C.c_double), ("phase0", C.c_double),
                ("pol", C.c_double), ("pol_type", C.c_int), ("backward", C.c_int),
                ("lon_type", C.c_int), ("per_type", C.c_int), ("tem_mode", C.c_int * 2),
                ("div_corr", C.c_int), ("lon_start", C.c_double), ("rise", C.c_double),
                ("flat", C.c_double), ("fall", C.c_double), ("lon_x0", C.c_double),
                ("duration", C.c_double), ("range", C.c_double), ("center", C.c_double * 2),
                ("w0", C.c_double), ("focus", C.c_double), ("xmin", C.c_double * 3)]


ZPULSE_BOX_DEFAULTS = dict(a0=1.0, omega0=10.0, phase=0.0, pol_type=0, pol=90.0,
                           propagation="forward", lon_type="polynomial", lon_start=None,
                           lon_fwhm=-1.0, lon_rise=0.0, lon_flat=0.0, lon_fall=0.0,
                           lon_duration=0.0, lon_x0=None, lon_range=None, per_type="plane",
                           per_center=None, per_w0=0.0, per_fwhm=0.0, per_focus=None,
                           per_tem_mode=0, no_div_corr=False)


def zpulse_box_params(nx, dx, xmin=(0.0, 0.0, 0.0), **kw):
    """read_input_zpulse (zpulse/os-zpulse-std.f03:157-805) for an in-box pulse along x1
    of a 2D or 3D box (len(nx)) without boost, tilt or chirp: the reference's defaults
    (:246-302) updated by kw; plane pulses get no divergence correction (init_zpulse
    :884-887).  Returns ZpulseBoxParams."""
    unknown = set(kw) - set(ZPULSE_BOX_DEFAULTS)
    if unknown:
        raise ValueError(f"unknown zpulse parameters {sorted(unknown)}")
    a = dict(ZPULSE_BOX_DEFAULTS, **kw)
    pi_180 = 0.0174532925199432957692369  # os-math.f90:15
    p = ZpulseBoxParams()
    p.a0, p.omega0 = a["a0"], a["omega0"]
    p.phase0 = a["phase"] * pi_180
    if a["pol_type"] not in (-1, 0, 1):
        raise ValueError("pol_type must be -1, 0 or 1")
    p.pol_type = a["pol_type"]
    p.pol = a["pol"] * pi_180
    p.backward = {"forward": 0, "backward": 1}[a["propagation"]]
    dims = len(nx)
    xmax = [0.0] * 3
    for d in range(dims):
        p.xmin[d] = xmin[d] if d < len(xmin) else 0.0
        xmax[d] = p.xmin[d] + nx[d] * dx[d]
    p.lon_type = {"polynomial": 0, "sin2": 1, "gaussian": 2}[a["lon_type"]]
    if p.lon_type < 2:
        if a["lon_fwhm"] > 0:
            p.rise, p.flat, p.fall = a["lon_fwhm"] / 2.0, 0.0, a["lon_fwhm"] / 2.0
        else:
            p.rise, p.flat, p.fall = a["lon_rise"], a["lon_flat"], a["lon_fall"]
        if a["lon_start"] is None:
            raise ValueError("polynomial and sin2 envelopes need lon_start")
        p.lon_start = a["lon_start"]
    else:
        p.duration = (a["lon_fwhm"] / np.sqrt(2.0 * np.log(2.0)) if a["lon_fwhm"] > 0
                      else a["lon_duration"])
        if a["lon_range"] is None or a["lon_x0"] is None:
            raise ValueError("a gaussian envelope needs lon_range and lon_x0")
        p.range, p.lon_x0 = a["lon_range"], a["lon_x0"]
    p.per_type = {"plane": 0, "gaussian": 1, "hermite": 1}[a["per_type"]]
    if p.per_type == 1:
        modes = np.atleast_1d(a["per_tem_mode"])
        for d in range(dims - 1):
            p.tem_mode[d] = int(modes[d]) if d < len(modes) else 0
        p.w0 = a["per_fwhm"] / np.sqrt(4.0 * np.log(2.0)) if a["per_fwhm"] > 0 else a["per_w0"]
        if p.w0 <= 0 or a["per_focus"] is None:
            raise ValueError("a gaussian beam needs per_w0 (or per_fwhm) and per_focus")
        p.focus = a["per_focus"]
    cen = a["per_center"]
    cen = [None] * (dims - 1) if cen is None else list(np.atleast_1d(cen))
    for d in range(dims - 1):
        c = cen[d] if d < len(cen) else None
        p.center[d] = 0.5 * (p.xmin[d + 1] + xmax[d + 1]) if c is None else c
    p.div_corr = 0 if (p.per_type == 0 or a["no_div_corr"]) else 1
    return p


def zpulse_box_add(grid: "Grid", p, order, dx, e, b, prec="f64"):
    """launch_zpulse (os-zpulse-std.f03:2350-2469) of an in-box pulse: adds the pulse
    (divergence-corrected unless plane or no_div_corr) to the guarded grids e, b."""
    L = _lib(prec)
    g = grid.c()
    dxa = np.array((list(dx) + [1.0, 1.0])[:3], np.float64)
    L.orc_zpulse_box_add(C.byref(g), C.byref(p), int(order), _p(dxa), _p(e), _p(b))


def ref_available(prec: str = "s") -> bool:
    return os.path.exists(os.path.join(_HERE, "_ref", f"libosiris_simd_{prec}.so")) and \
        _cpu_has_avx2()


def _cpu_has_avx2() -> bool:
    try:
        with open("/proc/cpuinfo") as f:
            flags = f.read()
        return " avx2" in flags and " fma" in flags
    except OSError:
        return False


def _ref(prec: str):
    """The reference's AVX2 kernels built by `make -C oracle ref`: prec "s" (FP32) or "d"
    (FP64) with -O2, or "s_fast" (FP32 with -Ofast, the build the CPU baseline times)."""
    key = "ref_" + prec
    if key not in _LIBS:
        path = os.path.join(_HERE, "_ref", f"libosiris_simd_{prec}.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle ref` where the "
                                    "reference sources are present")
        _LIBS[key] = C.CDLL(path)
    return _LIBS[key]


def _aligned(shape, dtype, fill=0, align=64):
    """Zero/fill-initialised array whose data pointer is ``align``-byte aligned (the
    reference SIMD kernels use aligned AVX loads and stores)."""
    dtype = np.dtype(dtype)
    nbytes = int(np.prod(shape)) * dtype.itemsize
    raw = np.empty(nbytes + align, np.uint8)
    off = (-raw.ctypes.data) % align
    a = raw[off:off + nbytes].view(dtype).reshape(shape)
    a[...] = fill
    return a


def _p(a: np.ndarray):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "arrays must be C-contiguous"
    return a.ctypes.data_as(C.c_void_p)


# --------------------------------------------------------------------------------------
# Grid helpers (reference layout)
# --------------------------------------------------------------------------------------
@dataclass
class Grid:
    dims: int
    nx: tuple
    gcl: tuple
    gcu: tuple

    def c(self) -> _Grid:
        g = _Grid()
        g.dims = self.dims
        for d in range(3):
            g.nx[d] = self.nx[d] if d < self.dims else 1
            g.gcl[d] = self.gcl[d] if d < self.dims else 0
            g.gcu[d] = self.gcu[d] if d < self.dims else 0
        return g

    @property
    def ld(self):
        return tuple(self.nx[d] + self.gcl[d] + self.gcu[d] for d in range(self.dims))

    def zeros(self, nc=3, dtype=np.float64):
        return np.zeros(tuple(reversed(self.ld)) + (nc,), dtype=dtype)

    def interior(self, f):
        """View of the interior cells [1..nx] of a grid array."""
        sl = tuple(slice(self.gcl[d], self.gcl[d] + self.nx[d]) for d in reversed(range(self.dims)))
        return f[sl]


def guard_cells(order: int):
    """Guard cells (lower, upper) required by interpolation and deposit of a given order:
    particles/os-particles.f03:321-360 (linear (1,2), quadratic (2,3), cubic (3,4))."""
    return {1: (1, 2), 2: (2, 3), 3: (3, 4)}[order]


def make_grid(dims, nx, order):
    lo, hi = guard_cells(order)
    return Grid(dims, tuple(nx), (lo,) * dims, (hi,) * dims)


# --------------------------------------------------------------------------------------
# Elementary pieces
# --------------------------------------------------------------------------------------
def spline(order, x, prec="f64"):
    L = _lib(prec)
    s = np.zeros(4, L._real)
    L.orc_spline(order, x, _p(s))
    return s[:order + 1]


def splineh(order, x, h, prec="f64"):
    L = _lib(prec)
    s = np.zeros(4, L._real)
    L.orc_splineh(order, x, h, _p(s))
    return s[:order + 1]


def wl(order, qn, x0, x1, prec="f64"):
    L = _lib(prec)
    s = np.zeros(4, L._real)
    L.orc_wl(order, qn, x0, x1, _p(s))
    return s[:order + 1]


def ntrim(x, prec="f64"):
    return _lib(prec).orc_ntrim(x)


def split(dims, dxi, xnew, ixold, xold, q, rgamma, u, prec="f64"):
    L = _lib(prec)
    r = L._real
    n = len(q)
    dxi = np.ascontiguousarray(dxi, np.int32)
    ixold = np.ascontiguousarray(ixold, np.int32)
    xnew, xold = (np.ascontiguousarray(a, r) for a in (xnew, xold))
    q, rgamma, u = (np.ascontiguousarray(a, r) for a in (q, rgamma, u))
    nf = 6 if dims == 2 else 7
    vpf = np.zeros((4 * max(n, 1), nf), r)
    vpi = np.zeros((4 * max(n, 1), dims), np.int32)
    ns = L.orc_split(dims, n, _p(dxi), _p(xnew), _p(ixold), _p(xold), _p(q), _p(rgamma), _p(u),
                     _p(vpf), _p(vpi))
    return vpf[:ns], vpi[:ns]


def get_emf(grid: Grid, order, e, b, ix, x, prec="f64"):
    L = _lib(prec)
    n = len(ix)
    ep = np.zeros((n, 3), L._real)
    bp = np.zeros((n, 3), L._real)
    g = grid.c()
    L.orc_get_emf(C.byref(g), order, _p(e), _p(b), _p(ix), _p(x), n, _p(ep), _p(bp))
    return ep, bp


def dudt(pusher, rqm, dt, u, q, ep, bp, prec="f64"):
    """Momentum update only; returns the un-normalised energy sum. u is updated in place."""
    L = _lib(prec)
    f = L.orc_dudt_boris if pusher == 0 else L.orc_dudt_vay
    ep = np.array(ep, L._real)
    bp = np.array(bp, L._real)
    return f(rqm, dt, _p(u), _p(q), _p(ep), _p(bp), len(q))


def advance_deposit(grid: Grid, order, pusher, rqm, dx, dt, e, b, j, ix, x, u, q, prec="f64"):
    """Fortran advance_deposit_{2d,3d}: updates ix/x/u in place, accumulates j."""
    L = _lib(prec)
    g = grid.c()
    dxa = np.array(list(dx) + [1.0] * (3 - len(dx)), np.float64)
    return L.orc_advance_deposit(C.byref(g), order, pusher, rqm, _p(dxa), dt, _p(e), _p(b),
                                 _p(j), _p(ix), _p(x), _p(u), _p(q), len(q))


def update_periodic(grid: Grid, f, mode, prec="f64"):
    L = _lib(prec)
    g = grid.c()
    L.orc_update_periodic(C.byref(g), _p(f), f.shape[-1], mode)


def particles_periodic(grid: Grid, ix, prec="f64"):
    L = _lib(prec)
    g = grid.c()
    L.orc_particles_periodic(C.byref(g), _p(ix), len(ix))


def smooth(grid: Grid, f, order, coef, prec="f64"):
    """smooth_vdf (vdf/os-vdf-smooth.f90:515-540) in place on a grid array, guard cells
    included; order per direction, coef (3, 31) centred at 15."""
    L = _lib(prec)
    g = grid.c()
    n = np.array(list(order) + [0] * (3 - len(order)), np.int32)
    c = np.ascontiguousarray(np.asarray(coef, np.float64).reshape(3, 31), L._real)
    L.orc_smooth(C.byref(g), f.shape[-1], _p(f), _p(n), _p(c))


def yee_advance(grid: Grid, dx, dt, e, b, j, prec="f64"):
    L = _lib(prec)
    g = grid.c()
    dxa = np.array(list(dx) + [1.0] * (3 - len(dx)), np.float64)
    L.orc_yee_advance(C.byref(g), _p(dxa), dt, _p(e), _p(b), _p(j))


def field_energy(grid: Grid, e, b, dvol, prec="f64"):
    L = _lib(prec)
    out = np.zeros(6)
    g = grid.c()
    L.orc_field_energy(C.byref(g), _p(e), _p(b), dvol, _p(out))
    return out


def kinetic_energy(u, q, prec="f64"):
    return _lib(prec).orc_kinetic_energy(_p(u), _p(q), len(q))


def deposit_rho(grid: Grid, order, ix, x, q, prec="f64"):
    L = _lib(prec)
    rho = grid.zeros(1, L._real)
    g = grid.c()
    L.orc_deposit_rho(C.byref(g), order, _p(ix), _p(x), _p(q), len(q), _p(rho))
    return rho


def div(grid: Grid, dx, f, prec="f64"):
    L = _lib(prec)
    out = grid.zeros(1, L._real)
    g = grid.c()
    dxa = np.array(list(dx) + [1.0] * (3 - len(dx)), np.float64)
    L.orc_div(C.byref(g), _p(dxa), _p(f), _p(out))
    return out


# --------------------------------------------------------------------------------------
# Reference AVX2 kernels (simd/os-spec-push-avx2{,d}.c:471-760 and 3D variant)
# --------------------------------------------------------------------------------------
def _copy_aligned(a, dtype):
    out = _aligned(a.shape, dtype)
    out[...] = a
    return out


def ref_vadvance_deposit(dims, order, ix, x, u, q, rqm, e, b, emf_lo, j, j_lo, dx, dt,
                         prec="s"):
    """Call FNAME(vadvance_deposit_{2d,3d}_s<order>) of the reference on copies.

    ``e``/``b``/``j`` are grids in the reference layout with ``emf_lo``/``j_lo`` lower guard
    cells.  Particle arrays are padded to a multiple of the SIMD width (8 single / 4
    double) because the reference writes dummy particles past i1
    (os-spec-push-avx2.c:523-542).  Returns (ix, x, u, j, ene)."""
    L = _ref(prec)
    single = prec.startswith("s")
    real = np.float32 if single else np.float64
    fn = getattr(L, f"vadvance_deposit_{dims}d_s{order}_")
    n = len(q)
    pad = n + 16
    ixp = _aligned((pad, dims), np.int32, 1)
    ixp[:n] = ix
    xp = _aligned((pad, dims), real)
    xp[:n] = x
    up = _aligned((pad, 3), real)
    up[:n] = u
    qp = _aligned((pad,), real)
    qp[:n] = q
    e = _copy_aligned(e, real)
    b = _copy_aligned(b, real)
    jj = _copy_aligned(j, real)
    emf_size = np.array(list(reversed(e.shape[:-1])), np.int32)
    j_size = np.array(list(reversed(jj.shape[:-1])), np.int32)
    emf_off = np.array([emf_lo] * dims, np.int32)
    j_off = np.array([j_lo] * dims, np.int32)
    i0 = C.c_int(1)
    i1 = C.c_int(n)
    rq = C.c_float(rqm) if single else C.c_double(rqm)
    dxa = np.array(list(dx), np.float64)
    dtc = C.c_double(dt)
    ene = C.c_double(0.0)
    fn(_p(ixp), _p(xp), _p(up), _p(qp), C.byref(i0), C.byref(i1), C.byref(rq), _p(e), _p(b),
       _p(emf_size), _p(emf_off), _p(jj), _p(j_size), _p(j_off), _p(dxa), C.byref(dtc),
       C.byref(ene))
    return ixp[:n].copy(), xp[:n].copy(), up[:n].copy(), jj, ene.value


# --------------------------------------------------------------------------------------
# Whole-simulation driver (os-simulation.f03:498-666 order, periodic)
# --------------------------------------------------------------------------------------
@dataclass
class Species:
    rqm: float
    pusher: int  # 0 Boris, 1 Vay
    ix: np.ndarray
    x: np.ndarray
    u: np.ndarray
    q: np.ndarray


@dataclass
class Sim:
    """FP64 (or FP32) oracle simulation in the reference layout."""
    dims: int
    nx: tuple
    dx: tuple
    dt: float
    order: int
    prec: str = "f64"
    species: list = field(default_factory=list)
    # current smoothing: per-direction filter orders and coefficients (3, 31), centre 15,
    # from oracle.smooth.setup; None = no smoothing
    smooth_order: tuple = None
    smooth_coef: np.ndarray = None
    # x walls: False = periodic; True = absorbing fields and absorbing particles
    # (y, z periodic), the field walls of type xwall: "lindman" or "vpml" (vpml_size
    # cells thick)
    open_x: bool = False
    xwall: str = "lindman"
    vpml_size: int = 8
    vpml_diffuse: tuple = (False, False)  # lower, upper x wall

    def __post_init__(self):
        self.grid = make_grid(self.dims, self.nx, self.order)
        r = _lib(self.prec)._real
        self.sm_order = self.sm_coef = None
        if self.smooth_order is not None:
            # current/os-current-define.f03:196-205: at least n lower and n+1 upper guards
            n = list(self.smooth_order) + [0] * (3 - len(self.smooth_order))
            g = self.grid
            self.grid = Grid(self.dims, g.nx, tuple(max(g.gcl[d], n[d]) for d in range(self.dims)),
                             tuple(max(g.gcu[d], n[d] + 1) for d in range(self.dims)))
            self.sm_order = np.array(n[:3], np.int32)
            self.sm_coef = np.ascontiguousarray(np.asarray(self.smooth_coef, np.float64).reshape(3, 31), r)
        self.e = self.grid.zeros(3, r)
        self.b = self.grid.zeros(3, r)
        self.j = self.grid.zeros(3, r)
        self.walls = None
        self.zpulses = []
        if self.open_x:
            if self.smooth_order is not None:
                raise ValueError("open x walls are implemented without smoothing")
            if self.xwall == "vpml":
                self.walls = VpmlWalls(self.grid, self.dx, self.dt, r, self.vpml_size,
                                       self.vpml_diffuse)
            elif self.xwall == "lindman":
                self.walls = LindmanWalls(self.grid, self.dx, self.dt, r)
            else:
                raise ValueError(f"unknown x wall type {self.xwall!r}")
        self.fld_hist = []
        self.par_hist = []

    @property
    def real(self):
        return _lib(self.prec)._real

    def add_species(self, rqm, pusher, ix, x, u, q):
        """Copies the arrays (the simulation advances its own copies in place)."""
        r = self.real
        self.species.append(Species(rqm, pusher, np.array(ix, np.int32), np.array(x, r),
                                    np.array(u, r), np.array(q, r)))

    def set_interior(self, which, arr):
        """Set interior E or B from an array of shape (nz, ny, nx, 3) (C order)."""
        f = self.e if which == "e" else self.b
        self.grid.interior(f)[...] = arr

    def fill_uniform(self, e0=(0.0, 0.0, 0.0), b0=(0.0, 0.0, 0.0)):
        """Uniform fields on every cell, guard cells included (os-emf-gridval.f03)."""
        self.e[...] = np.asarray(e0, np.float64).astype(self.real)
        self.b[...] = np.asarray(b0, np.float64).astype(self.real)

    def step(self, n=1):
        if self.open_x:
            return self._step_open(n)
        L = _lib(self.prec)
        g = self.grid.c()
        ns = len(self.species)
        dxa = np.array(list(self.dx) + [1.0] * (3 - self.dims), np.float64)
        IA = C.c_void_p * ns
        pus = np.array([s.pusher for s in self.species], np.int32)
        rqm = np.array([s.rqm for s in self.species], self.real)
        npa = np.array([len(s.q) for s in self.species], np.int32)
        ixs = IA(*[s.ix.ctypes.data for s in self.species])
        xs = IA(*[s.x.ctypes.data for s in self.species])
        us = IA(*[s.u.ctypes.data for s in self.species])
        qs = IA(*[s.q.ctypes.data for s in self.species])
        for _ in range(n):
            fe = np.zeros(6)
            pe = np.zeros(ns)
            L.orc_step(C.byref(g), self.order, _p(dxa), self.dt, _p(self.e), _p(self.b),
                       _p(self.j), ns, _p(pus), _p(rqm), ixs, xs, us, qs, _p(npa), _p(fe),
                       _p(pe), _p(self.sm_order), _p(self.sm_coef))
            self.fld_hist.append(fe)
            self.par_hist.append(pe)

    def _step_open(self, n):
        L = _lib(self.prec)
        g = self.grid.c()
        ns = len(self.species)
        dxa = np.array(list(self.dx) + [1.0] * (3 - self.dims), np.float64)
        IA = C.c_void_p * ns
        pus = np.array([s.pusher for s in self.species], np.int32)
        rqm = np.array([s.rqm for s in self.species], self.real)
        npa = np.array([len(s.q) for s in self.species], np.int32)
        ixs = IA(*[s.ix.ctypes.data for s in self.species])
        xs = IA(*[s.x.ctypes.data for s in self.species])
        us = IA(*[s.u.ctypes.data for s in self.species])
        qs = IA(*[s.q.ctypes.data for s in self.species])
        for _ in range(n):
            # zpulse launch at t = n dt before the step (os-simulation.f03:528-531)
            t = len(self.fld_hist) * self.dt
            for zp in self.zpulses:
                L.orc_zpulse_wall_2d(C.byref(g), C.byref(zp), self.order, _p(dxa), self.dt, t,
                                     _p(self.b))
            fe = np.zeros(6)
            pe = np.zeros(ns)
            L.orc_step_open(C.byref(g), self.order, _p(dxa), self.dt, _p(self.e), _p(self.b),
                            _p(self.j), *self._wall_args(), ns, _p(pus), _p(rqm), ixs, xs,
                            us, qs, _p(npa), _p(fe), _p(pe))
            self.fld_hist.append(fe)
            self.par_hist.append(pe)
        for s, k in zip(self.species, npa):  # leading slices stay contiguous
            s.ix, s.x, s.u, s.q = s.ix[:k], s.x[:k], s.u[:k], s.q[:k]

    def _wall_args(self):
        if self.xwall == "vpml":
            return None, C.byref(self.walls.c)
        return C.byref(self.walls.c), None

    def init_fields(self, k, ufl1=0.0):
        """init_fields of species k (spec/os-spec-initfields.f90:38-100): rho deposited
        and folded; ufl1 = 0: the 2D direct-sum ES field added to E (the reference has no
        3D form); otherwise the beam field of a species drifting along x with u1 = ufl1
        (es_solver_2d_beam, es_solver_3d_beam) added to E and B."""
        if self.dims != 2 and ufl1 == 0:
            raise ValueError("the reference's ES solver (ufl(1) = 0) is 2D only")
        sp = self.species[k]
        rho = deposit_rho(self.grid, self.order, sp.ix, sp.x, sp.q, self.prec)
        L = _lib(self.prec)
        g = self.grid.c()
        L.orc_update_periodic_from(C.byref(g), _p(rho), 1, 1, 1 if self.open_x else 0)
        ef = self.grid.zeros(3, self.real)
        dxa = np.array((list(self.dx) + [1.0, 1.0])[:3], np.float64)
        if ufl1 == 0:
            L.orc_es_solver_2d(C.byref(g), _p(dxa), _p(rho), _p(ef))
            self.grid.interior(self.e)[...] += self.grid.interior(ef)
            return ef
        bf = self.grid.zeros(3, self.real)
        beam = L.orc_es_solver_2d_beam if self.dims == 2 else L.orc_es_solver_3d_beam
        beam(C.byref(g), _p(dxa), _p(rho), float(np.float32(ufl1)), int(self.open_x), _p(ef),
             _p(bf))
        self.e += ef  # add(emf%e, ef), add(emf%b, bf): every cell
        self.b += bf
        return ef, bf

    def add_zpulse(self, xmin=(0.0, 0.0, 0.0), **kw):
        """An in-box zpulse section (direction 1) launched now, before the first step:
        its field is added to E and B."""
        p = zpulse_box_params(self.nx, self.dx, xmin, **kw)
        zpulse_box_add(self.grid, p, self.order, self.dx, self.e, self.b, self.prec)
        return p

    def add_zpulse_wall(self, xmin=(0.0, 0.0, 0.0), **kw):
        """A zpulse_wall section (direction 1); needs open x walls (os-zpulse-wall.f03:
        600-608)."""
        if not self.open_x:
            raise ValueError("the launch wall must be open")
        self.zpulses.append(zpulse_wall_params(self.nx, self.dx, xmin, **kw))

    def wall_refresh(self):
        """The boundary update the next step starts with (Lindman restore or VPML wall
        refresh, periodic y), so
        that the guard cells hold what the next push reads."""
        L = _lib(self.prec)
        g = self.grid.c()
        fn = L.orc_open_boundary_vpml if self.xwall == "vpml" else L.orc_open_boundary
        fn(C.byref(g), C.byref(self.walls.c), _p(self.e), _p(self.b))

    def gauss_residual(self):
        """div E - rho (interior nodes) with rho from all species (no ion background)."""
        rho = self.grid.zeros(1, self.real)
        for s in self.species:
            rho += deposit_rho(self.grid, self.order, s.ix, s.x, s.q, self.prec)
        update_periodic(self.grid, rho, 1, self.prec)
        e = self.e.copy()
        update_periodic(self.grid, e, 0, self.prec)
        dv = div(self.grid, self.dx, e, self.prec)
        return self.grid.interior(dv)[..., 0].astype(np.float64) - \
            self.grid.interior(rho)[..., 0].astype(np.float64)


# --------------------------------------------------------------------------------------
# CPU baseline: the reference's AVX2 kernel run the way the reference runs it with
# OpenMP (particles/os-particles.f03:1416-1443): T threads, disjoint particle chunks,
# private current arrays, then a fixed-order reduction.
# --------------------------------------------------------------------------------------
class RefSimdTimer:
    """Times FNAME(vadvance_deposit_2d_s<order>) (FP32 AVX2) on host threads.

    The particles are pushed again and again without the periodic particle boundary,
    so each push starts from the initial cell indices (a reset of ix per thread before
    the call, ~1% of its time): positions and momenta keep evolving, and no particle
    can drift past the guard cells, which the kernel does not check."""

    def __init__(self, order, ix, x, u, q, rqm, e, b, glo, dx, dt, threads, fast=False):
        import threading  # noqa: F401

        self.L = _ref("s_fast" if fast else "s")
        self.fn = getattr(self.L, f"vadvance_deposit_2d_s{order}_")
        self.threads = threads
        n = len(q)
        chunks = np.array_split(np.arange(n), threads)
        self.e = _copy_aligned(e, np.float32)
        self.b = _copy_aligned(b, np.float32)
        self.emf_size = np.array(list(reversed(e.shape[:-1])), np.int32)
        self.off = np.array([glo, glo], np.int32)
        self.dx = np.array(list(dx), np.float64)
        self.dt = C.c_double(dt)
        self.rqm = C.c_float(rqm)
        self.parts = []
        for c in chunks:
            m = len(c)
            pad = m + 16
            ixp = _aligned((pad, 2), np.int32, 1)
            ixp[:m] = ix[c]
            xp = _aligned((pad, 2), np.float32)
            xp[:m] = x[c]
            up = _aligned((pad, 3), np.float32)
            up[:m] = u[c]
            qp = _aligned((pad,), np.float32)
            qp[:m] = q[c]
            jj = _aligned(e.shape, np.float32)
            self.parts.append((m, ixp, xp, up, qp, jj, ixp.copy()))

    def _run(self, k):
        m, ixp, xp, up, qp, jj, ix0 = self.parts[k]
        C.memmove(_p(ixp), _p(ix0), ix0.nbytes)
        C.memset(_p(jj), 0, jj.nbytes)  # each thread zeroes its private current
        i0, i1 = C.c_int(1), C.c_int(m)
        ene = C.c_double(0)
        self.fn(_p(ixp), _p(xp), _p(up), _p(qp), C.byref(i0), C.byref(i1), C.byref(self.rqm),
                _p(self.e), _p(self.b), _p(self.emf_size), _p(self.off), _p(jj),
                _p(self.emf_size), _p(self.off), _p(self.dx), C.byref(self.dt), C.byref(ene))

    def step(self):
        """One push+deposit of every particle on all threads, then the J reduction."""
        import threading

        ts = [threading.Thread(target=self._run, args=(k,)) for k in range(self.threads)]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        j = self.parts[0][5]
        for p in self.parts[1:]:
            j += p[5]
        return j
