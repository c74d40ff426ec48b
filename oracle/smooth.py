"""Current smoothing coefficients (test infrastructure only: the checker).

Restates the set-up half of vdf/os-vdf-smooth.f90; the filter itself is ``orc_smooth`` in
pic_oracle.c (wrapped by ``oracle.smooth``):

* ``read_nml`` -- read_nml_smooth (:183-361): per-direction filter type and order, the
  laserkdx override (:251-280), custom kernels combined at read time (:352-359);
* ``setup`` -- setup_smooth (:366-470): single-pass kernels of every type, then the
  normalisation to unit sum (:458-468);
* ``combine_kernel3`` (:142-172), ``digital_filter`` (:110-137) and ``besseli0``
  (os-math.f90:450-472, the Abramowitz & Stegun 9.8.1/9.8.2 polynomial fits).

Precision follows the reference's single-precision build (p_k_fld = real(4)): scoef is a
p_k_fld array, so every value stored into it is rounded to float32, the normalisation sum
runs in double over those floats and the quotient is rounded to float32 again.  Fortran
literals without a kind suffix are default (single precision) reals; they are written here
as ``_s(literal)``.  With ``kfld=np.float64`` the double-precision build is reproduced.
"""
from __future__ import annotations

import math

import numpy as np

NONE, CUSTOM, BINOMIAL, COMPENSATED, PASS5, TRI6, TRI8 = 0, 1, 2, 3, 4, 5, 6
LASERKDX, DIGITAL = 10, 11
ORDER_MAX = 15  # p_order_max
NCOEF = 2 * ORDER_MAX + 1
PI = 3.14159265358979323846264  # os-math.f90:13


def _s(v):
    """A default-kind (single precision) Fortran literal or expression, as a double."""
    return float(np.float32(v))


def besseli0(x: float) -> float:
    """besseli0_rdoub, os-math.f90:450-472."""
    ax = abs(x)
    if ax < _s(3.75):
        y = (x / _s(3.75)) ** 2
        return _s(1.0) + y * (_s(3.5156229) + y * (_s(3.0899424) + y * (_s(1.2067492) + y * (
            _s(0.2659732) + y * (0.360768e-1 + y * 0.45813e-2)))))
    y = _s(3.75) / ax
    return (math.exp(ax) / math.sqrt(ax)) * (_s(0.39894228) + y * (0.1328592e-1 + y * (
        0.225319e-2 + y * (-0.157565e-2 + y * (0.916281e-2 + y * (-0.2057706e-1 + y * (
            0.2635537e-1 + y * (-0.1647633e-1 + y * 0.392377e-2))))))))


def digital_filter(fc: float, A: float, n: int) -> np.ndarray:
    """digital_filter, os-vdf-smooth.f90:110-137 (2n+1 values, centre n)."""
    if A <= _s(21.0):
        alpha = 0.0
    elif A >= _s(50.0):
        alpha = 0.1102 * (A - 8.7)
    else:
        alpha = 0.5842 * ((A - 21.0) ** 0.4) + 0.07886 * (A - 21.0)
    f = np.zeros(2 * n + 1)
    for i in range(1, n + 1):
        r = float(np.float32(np.float32(i) / np.float32(n)) ** np.float32(2))
        f[n + i] = besseli0(alpha * math.sqrt(1.0 - r)) / besseli0(alpha) * \
            math.sin(i * PI * fc) / (i * PI)
    f[n] = fc
    for i in range(1, n + 1):
        f[n - i] = f[n + i]
    return f


def combine_kernel3(kern3, n: int, nmax: int) -> np.ndarray:
    """combine_kernel3, os-vdf-smooth.f90:142-172: n 3-point kernels kern3[k] = (a, b, c)
    (kern3(1:3, k)) into one 2 nmax + 1 point kernel, in double precision."""
    c = np.zeros(2 * nmax + 1)
    c[nmax] = 1.0
    for k in range(1, n + 1):
        a, b, cc = (float(v) for v in kern3[k - 1])
        f1 = 0.0
        for i in range(nmax - (k - 1), nmax + k + 1):
            c[i - 1] = cc * c[i] + c[i - 1]
            f2 = a * c[i]
            c[i] = b * c[i] + f1
            f1 = f2
    return c


def read_nml(dims, type_, order, swfj=None, laserkdx_omega0=-1.0, laserkdx_dir=1,
             digital_A=None, digital_fc=None, kfld=np.float32):
    """read_nml_smooth after the namelist is read: returns a state dict with the
    per-direction ``type``, ``order``, ``max_order`` and ``scoef`` (dims, 2 max_order + 1)
    (custom kernels already combined), or raises ValueError where the reference aborts."""
    type_ = list(type_)
    order = list(order)
    st = {"type": [NONE] * dims, "order": [0] * dims, "laserkdx_omega0": laserkdx_omega0,
          "laserkdx_dir": laserkdx_dir, "digital_A": [0.0] * dims, "digital_fc": [0.0] * dims}
    if swfj is None:
        swfj = np.ones((dims, ORDER_MAX, 3))
    if type_[0] == LASERKDX:
        st["type"] = [LASERKDX] * dims
        if laserkdx_dir < 1 or laserkdx_dir > dims:
            raise ValueError("laserkdx_dir must be between 1 and dims")
        st["order"] = [2 if i + 1 == laserkdx_dir else 0 for i in range(dims)]
        if laserkdx_omega0 <= 0.0:
            raise ValueError("invalid laserkdx_omega0")
    for i in range(dims):
        t = type_[i]
        if t == NONE:
            st["order"][i] = 0
            st["type"][i] = NONE
        elif t in (CUSTOM, BINOMIAL):
            st["type"][i] = t
            if order[i] < 1 or order[i] > ORDER_MAX:
                raise ValueError("order must be >= 1 and <= p_order_max")
            st["order"][i] = order[i]
        elif t == COMPENSATED:
            st["type"][i] = t
            st["order"][i] = 2 if order[i] < 2 else order[i]
        elif t == PASS5:
            st["type"][i], st["order"][i] = t, 5
        elif t == TRI6:
            st["type"][i], st["order"][i] = t, 3
        elif t == TRI8:
            st["type"][i], st["order"][i] = t, 4
        elif t == LASERKDX:
            pass
        elif t == DIGITAL:
            st["type"][i] = t
            if order[i] < 1:
                raise ValueError("order must be >= 1 for digital smoothing")
            st["order"][i] = order[i]
            st["digital_A"][i] = digital_A[i]
            st["digital_fc"][i] = digital_fc[i]
        else:
            raise ValueError(f"invalid smooth type {t}")
    mo = max(st["order"])
    st["max_order"] = mo
    st["scoef"] = None
    if mo > 0:
        sc = np.zeros((dims, 2 * mo + 1), kfld)
        for i in range(dims):
            if st["type"][i] == CUSTOM:
                sc[i] = combine_kernel3(swfj[i], st["order"][i], mo)
        st["scoef"] = sc
    return st


def setup(st, dx, gamma=1.0, kfld=np.float32):
    """setup_smooth, os-vdf-smooth.f90:366-470, on a read_nml state (scoef in place)."""
    dims, mo = len(st["order"]), st["max_order"]
    sc = st["scoef"]
    if sc is None:
        return st
    c = mo
    if st["type"][0] == LASERKDX:
        d = st["laserkdx_dir"] - 1
        lk = st["laserkdx_omega0"] * dx[d] / (gamma + math.sqrt(gamma ** 2 - 1))
        kern3 = [(1.0, 2.0, 1.0), (-1.0, 4 + 2 * math.cos(lk), -1.0)]
        sc[d] = combine_kernel3(kern3, 2, mo)
    else:
        for i in range(dims):
            t, n = st["type"][i], st["order"][i]
            if t == BINOMIAL:
                sc[i] = combine_kernel3([(1.0, 2.0, 1.0)] * n, n, mo)
            elif t == COMPENSATED:
                # (4.0 + 2.0 * (order - 1)) / (order - 1) in default (single) precision
                comp = float(np.float32(np.float32(4.0 + 2.0 * (n - 1)) / np.float32(n - 1)))
                kern3 = [(1.0, 2.0, 1.0)] * (n - 1) + [(-1.0, comp, -1.0)]
                sc[i] = combine_kernel3(kern3, n, mo)
            elif t == PASS5:
                kern3 = [(1.0, 2.0, 1.0)] * 4 + [(-5.0, 15.0, -5.0)]
                sc[i] = combine_kernel3(kern3, 5, mo)
            elif t == TRI6:
                sc[i, c - 3:c + 4] = [0.015625, -0.09375, 0.234375, 0.6875, 0.234375, -0.09375,
                                      0.015625]
            elif t == TRI8:
                sc[i, c - 4:c + 5] = [-0.00390625, 0.03125, -0.109375, 0.21875, 0.7265625,
                                      0.21875, -0.109375, 0.03125, -0.00390625]
            elif t == DIGITAL:
                n = st["order"][i]
                sc[i, c - n:c + n + 1] = digital_filter(st["digital_fc"][i], st["digital_A"][i], n)
    for i in range(dims):
        n = st["order"][i]
        if n > 0:
            total = 0.0
            for j in range(c - n, c + n + 1):
                total = total + float(sc[i, j])
            for j in range(c - n, c + n + 1):
                sc[i, j] = float(sc[i, j]) / total
    return st


def coefficients(dims, type_, order, dx, gamma=1.0, kfld=np.float32, **kw):
    """read_nml + setup; returns (order[3], coef (3, 31) float64 centred at 15)."""
    st = setup(read_nml(dims, type_, order, kfld=kfld, **kw), dx, gamma, kfld)
    out = np.zeros((3, NCOEF))
    n3 = [0, 0, 0]
    mo = st["max_order"]
    for i in range(dims):
        n = st["order"][i]
        n3[i] = n
        if n > 0:
            out[i, ORDER_MAX - n:ORDER_MAX + n + 1] = st["scoef"][i, mo - n:mo + n + 1]
    return n3, out
