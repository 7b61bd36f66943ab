"""ctypes wrapper around the CPU FP64 oracle (liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
(as the checker) and by bench.py's ``cpu_baseline`` leg.  The product never
imports this module.

Array layouts follow the reference's Fortran arrays (see inca_oracle.h).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")


def build(force: bool = False) -> str:
    path = os.path.join(_HERE, "liborc.so")
    if force or not os.path.exists(path):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", _HERE],
                       check=True)
    return path


_LIB_FMA = None


def lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    path = os.path.join(_HERE, "liborc.so")
    if not os.path.exists(path):
        build()
    _LIB = _configure(C.CDLL(path))
    return _LIB


def lib_fma():
    """The same oracle built with fused multiply-adds wherever the compiler
    contracts them (liborc_fma.so), as the reference's release build does on
    an FMA host: a second FP64 implementation for the roundoff audit."""
    global _LIB_FMA
    if _LIB_FMA is not None:
        return _LIB_FMA
    path = os.path.join(_HERE, "liborc_fma.so")
    if not os.path.exists(path):
        build(force=True)
    _LIB_FMA = _configure(C.CDLL(path))
    return _LIB_FMA


def lib_native(path):
    """The oracle built for the host it runs on (make native NATIVE_OUT=path),
    for the CPU baseline's timing."""
    subprocess.run(["make", "-s", "-B", "-C", _HERE, "native", "NATIVE_OUT=%s" % path],
                   check=True)
    return _configure(C.CDLL(path))


def _configure(L):
    L.orc_dfact.argtypes = [C.c_int]
    L.orc_dfact.restype = C.c_int
    L.orc_normprim.argtypes = [C.c_double, C.c_int, C.c_int, C.c_int]
    L.orc_normprim.restype = C.c_double
    L.orc_gauherm.argtypes = [C.c_int, _dp, _dp]
    L.orc_gauherm.restype = C.c_int
    L.orc_polycoef.argtypes = [C.c_double] * 7 + [C.c_int] * 5 + \
        [C.c_double] * 3 + [_dp]
    L.orc_polycoef.restype = C.c_int
    L.orc_intracule.argtypes = [C.c_int, _dp, _ip, _dp, _dp, _dp, C.c_int,
                                C.c_double, C.c_int64, _ip, _dp, C.c_int64,
                                _dp, _dp, _lp, C.c_int, C.c_int]
    L.orc_intracule.restype = C.c_int
    L.orc_intracule_ex.argtypes = L.orc_intracule.argtypes + [C.c_int]
    L.orc_intracule_ex.restype = C.c_int
    L.orc_gauleg.argtypes = [C.c_int, _dp, _dp]
    L.orc_gridpoints.argtypes = [C.c_int, _dp, _ip, _ip, _dp, _dp, C.c_int,
                                 _dp, C.POINTER(C.c_int64),
                                 C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                 _dp, _dp, _dp, _ip, C.c_int64]
    L.orc_gridpoints.restype = C.c_int
    L.orc_tfvc.argtypes = [C.c_int64, _dp, _ip, C.c_int, _dp, _dp, _dp]
    L.orc_gridpoints2.argtypes = [C.c_int, _dp, _dp, _ip, C.c_int, _dp,
                                  C.POINTER(C.c_int64), C.c_void_p,
                                  C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_int64]
    L.orc_gridpoints2.restype = C.c_int64
    L.orc_gridpoints3.argtypes = [_dp, _dp, _ip, C.c_void_p]
    L.orc_gridpoints3.restype = C.c_int64
    L.orc_centercalc.argtypes = [C.c_int, _ip, _dp, C.c_int, C.c_int, _dp,
                                 _dp, _dp]
    L.orc_centercalc.restype = C.c_int
    L.orc_bs_radius.argtypes = [C.c_int]
    L.orc_bs_radius.restype = C.c_double
    L.orc_smearing.argtypes = [C.c_int64, _dp, _dp, _dp, C.c_double, C.c_int,
                               C.c_double, _dp, _dp, _dp]
    L.orc_onee.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _ip, _dp, _dp, _dp, C.c_int,
                           C.c_void_p, C.c_void_p, C.c_int64, _dp, _dp]
    L.orc_set_arith.argtypes = [C.c_int]
    L.orc_get_arith.restype = C.c_int
    L.orc_fit_rows.argtypes = [C.c_int, _dp, _ip, _dp, _dp, _dp, C.c_int, C.c_double,
                               C.c_int64, _ip, _dp, C.c_void_p, C.c_int64]
    L.orc_fit_rows.restype = C.c_int64
    L.orc_polycoef_rows.argtypes = [C.c_int64, _dp, _dp, _ip]
    L.orc_polycoef_rows.restype = C.c_int64
    L.orc_fit_census.argtypes = [C.c_int, _dp, _ip, _dp, _dp, _dp, C.c_int, C.c_double,
                                 C.c_int64, _ip, _dp, C.c_double, C.c_void_p, C.c_int64]
    L.orc_fit_census.restype = C.c_int64
    return L


FIT_ROW = 17  # ORC_FIT_ROW


class arith:
    """Context manager: run the oracle `L` (default liborc.so) in arithmetic
    mode `mode` (0: repeated multiplication, 1: the reference's libm forms)."""

    def __init__(self, mode, L=None):
        self.L = L or lib()
        self.mode = mode

    def __enter__(self):
        self.prev = self.L.orc_get_arith()
        self.L.orc_set_arith(int(self.mode))
        return self.L

    def __exit__(self, *a):
        self.L.orc_set_arith(self.prev)


def fit_rows(prims, ijkl, d, thresh, normalize=True, L=None):
    """(n, FIT_ROW) inputs of every polynomial fit (axes of degree >= 1) of
    the records passing screen 1: [record, axis, Ri, Rj, Rk, Rl, P_ik, P_jl,
    P, ti, tj, tk, tl, np, sqe, invz, alfijkl]."""
    L = L or lib()
    alpha = _f64(prims["alpha"])
    nprim = len(alpha)
    tmn = _i32(np.asarray(prims["tmn"]).reshape(nprim, 3).T.reshape(-1))
    ijkl = _i32(ijkl).reshape(-1)
    d = _f64(d)
    args = (nprim, alpha, tmn, _f64(prims["x"]), _f64(prims["y"]), _f64(prims["z"]),
            int(normalize), float(thresh), len(d), ijkl, d)
    n = L.orc_fit_rows(*args, None, 0)
    rows = np.zeros((max(n, 1), FIT_ROW))
    m = L.orc_fit_rows(*args, rows.ctypes.data, n)
    assert m == n
    return rows[:n]


def fit_census(prims, ijkl, d, thresh, cmin, normalize=True, L=None, cap=1 << 20):
    """Rows (as fit_rows) of the fits zeroed, or with max |C| >= cmin, in
    arithmetic mode 0 or 1 of library L (OpenMP over records)."""
    L = L or lib()
    alpha = _f64(prims["alpha"])
    nprim = len(alpha)
    tmn = _i32(np.asarray(prims["tmn"]).reshape(nprim, 3).T.reshape(-1))
    ijkl = _i32(ijkl).reshape(-1)
    d = _f64(d)
    rows = np.zeros((cap, FIT_ROW))
    n = L.orc_fit_census(nprim, alpha, tmn, _f64(prims["x"]), _f64(prims["y"]),
                         _f64(prims["z"]), int(normalize), float(thresh), len(d), ijkl, d,
                         float(cmin), rows.ctypes.data, cap)
    assert 0 <= n < cap, n
    return rows[:n]


def polycoef_rows(rows, L=None):
    """Fits of the rows in the current arithmetic mode of `L`: (C (n,22) with
    C[:,0] the highest power, status (n,) with 1 where the reference would
    zero the fit)."""
    L = L or lib()
    rows = _f64(rows)
    n = rows.shape[0]
    Cc = np.zeros((max(n, 1), 22))
    st = np.zeros(max(n, 1), np.int32)
    L.orc_polycoef_rows(n, rows.reshape(-1) if n else np.zeros(FIT_ROW), Cc, st)
    return Cc[:n], st[:n]


class OrcQuartet(C.Structure):
    _fields_ = [("A", C.c_double), ("sqe", C.c_double), ("dx", C.c_double),
                ("dy", C.c_double), ("dz", C.c_double), ("screen1", C.c_double),
                ("screen2", C.c_double), ("L", C.c_int * 3), ("pass1", C.c_int),
                ("pass2", C.c_int), ("polyfail", C.c_int), ("C", (C.c_double * 22) * 3),
                ("E", (C.c_double * 22) * 3)]


def prepare_quartet(prims, q, D, thresh, normalize=True):
    """One record through screens + prefactors + polycoef (orc_prepare_quartet)."""
    L = lib()
    L.orc_prepare_quartet.argtypes = [C.c_int, _dp, _ip, _dp, _dp, _dp, C.c_void_p,
                                      C.c_double, C.c_int, C.c_int, C.c_int, C.c_int,
                                      C.c_double, C.POINTER(OrcQuartet)]
    alpha = _f64(prims["alpha"])
    n = len(alpha)
    tmn = _i32(np.asarray(prims["tmn"]).reshape(n, 3).T.reshape(-1))
    nrm = _f64(normprim(alpha, np.asarray(prims["tmn"]).reshape(n, 3))) if normalize else None
    lim = thresh * (1.0 / (n * (n + 1.0) * 0.5))
    out = OrcQuartet()
    L.orc_prepare_quartet(n, alpha, tmn, _f64(prims["x"]), _f64(prims["y"]), _f64(prims["z"]),
                          None if nrm is None else nrm.ctypes.data, lim, int(q[0]), int(q[1]),
                          int(q[2]), int(q[3]), float(D), C.byref(out))
    return out


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def lebedev_unit(nang: int) -> np.ndarray:
    """Lebedev-Laikov rule with nang points as (nang,4) [x,y,z,w], sum w = 1.

    Taken from scipy.integrate.lebedev_rule (independent of the product's
    own table), weights divided by 4*pi as in the reference's LDnnnn.
    """
    from scipy.integrate import lebedev_rule
    deg_to_order = {6: 3, 14: 5, 26: 7, 38: 9, 50: 11, 74: 13, 86: 15,
                    110: 17, 146: 19, 170: 21, 194: 23, 230: 25, 266: 27,
                    302: 29, 350: 31, 434: 35, 590: 41, 770: 47, 974: 53,
                    1202: 59, 1454: 65, 1730: 71, 2030: 77, 2354: 83,
                    2702: 89, 3074: 95, 3470: 101, 3890: 107, 4334: 113,
                    4802: 119, 5294: 125, 5810: 131}
    x, w = lebedev_rule(deg_to_order[nang])
    out = np.empty((nang, 4))
    out[:, :3] = x.T
    out[:, 3] = w / (4.0 * np.pi)
    return out


def normprim(alpha, tmn):
    L = lib()
    return np.array([L.orc_normprim(float(a), int(t[0]), int(t[1]), int(t[2]))
                     for a, t in zip(alpha, tmn)])


def intracule(prims: dict, ijkl: np.ndarray, d: np.ndarray, r: np.ndarray,
              thresh: float, normalize: bool = True, nthreads: int = 0,
              dynamic: bool = False, exact: bool = False, L=None):
    """Quartet loop of the reference; returns (ivec, counts[4]).
    exact=True evaluates the exact axis polynomials instead of the fits.
    L: the oracle library to run (default liborc.so)."""
    L = L or lib()
    alpha = _f64(prims["alpha"])
    nprim = len(alpha)
    tmn = _i32(np.asarray(prims["tmn"]).reshape(nprim, 3).T.reshape(-1))
    ijkl = _i32(ijkl).reshape(-1)
    d = _f64(d)
    r = _f64(r).reshape(-1)
    G = r.size // 3
    ivec = np.zeros(max(G, 1))
    counts = np.zeros(4, dtype=np.int64)
    L.orc_intracule_ex(nprim, alpha, tmn, _f64(prims["x"]), _f64(prims["y"]),
                       _f64(prims["z"]), int(normalize), float(thresh),
                       len(d), ijkl, d, G, r if G else np.zeros(3), ivec, counts,
                       int(nthreads), int(dynamic), int(exact))
    return ivec[:G], counts


def gauleg(n):
    t = np.zeros(n)
    w = np.zeros(n)
    lib().orc_gauleg(n, t, w)
    return t, w


def gridpoints(cent, nradc, nangc, sfalpha, Ps, nosym=False):
    L = lib()
    cent = _f64(cent).reshape(-1)
    nq = len(nradc)
    ang = np.concatenate([lebedev_unit(int(n)).reshape(-1) for n in nangc])
    cap = int(sum(int(a) * int(b) for a, b in zip(nradc, nangc)))
    r = np.zeros(3 * cap)
    w = np.zeros(cap)
    wv = np.zeros(cap)
    mp = np.zeros(cap, dtype=np.int32)
    mg, rg, rrg = C.c_int64(), C.c_int64(), C.c_int64()
    rc = L.orc_gridpoints(nq, cent, _i32(nradc), _i32(nangc), _f64(sfalpha),
                          _f64(Ps), int(nosym), _f64(ang), C.byref(mg),
                          C.byref(rg), C.byref(rrg), r, w, wv, mp, cap)
    assert rc == 0
    n = rrg.value
    return dict(maxgrid=mg.value, rgrid=rg.value, rrgrid=n,
                r=r[:3 * n].reshape(n, 3), w=w[:n], w_vee=wv[:n], map=mp[:n])


def tfvc(pts, owner, cent, Ps):
    pts = _f64(pts).reshape(-1)
    n = pts.size // 3
    out = np.zeros(max(n, 1))
    lib().orc_tfvc(n, pts, _i32(owner), len(Ps), _f64(cent).reshape(-1),
                   _f64(Ps), out)
    return out[:n]


def gridpoints2(tart, step, nang, nosym=False):
    L = lib()
    tart = _f64(tart).reshape(-1)
    step = _f64(step)
    nang = _i32(nang)
    ang = np.concatenate([lebedev_unit(int(n)).reshape(-1) for n in nang])
    nradi = C.c_int64()
    L.orc_gridpoints2(len(step), tart, step, nang, int(nosym), _f64(ang),
                      C.byref(nradi), None, None, None, None, 0)
    nr = nradi.value
    radi = np.zeros(nr)
    smn = np.zeros(nr, dtype=np.int32)
    cap = int(sum(nang[b] * (int((tart[2*b+1] - tart[2*b]) / step[b]) + 1)
                  for b in range(len(step))))
    r = np.zeros(3 * max(cap, 1))
    w = np.zeros(max(cap, 1))
    n = L.orc_gridpoints2(len(step), tart, step, nang, int(nosym), _f64(ang),
                          C.byref(nradi), radi.ctypes.data, smn.ctypes.data,
                          r.ctypes.data, w.ctypes.data, cap)
    assert n >= 0
    return dict(radi=radi, smn=smn, r=r[:3 * n].reshape(n, 3), w_ang=w[:n])


def gridpoints3(center, step, np_i):
    L = lib()
    np_i = _i32(np_i)
    n = int(np_i[0]) * int(np_i[1]) * int(np_i[2])
    r = np.zeros(3 * n)
    L.orc_gridpoints3(_f64(center), _f64(step), np_i, r.ctypes.data)
    return r.reshape(n, 3)


def centercalc(an, cartes, nohydro=True, betaone=False):
    L = lib()
    an = _i32(an)
    nat = len(an)
    cart = _f64(np.asarray(cartes, dtype=np.float64).reshape(nat, 3).T.reshape(-1))
    nqmax = nat * (nat - 1) + 1
    cent = np.zeros(3 * nqmax)
    Ps = np.zeros(nqmax)
    sf = np.zeros(nqmax)
    nq = L.orc_centercalc(nat, an, cart, int(nohydro), int(betaone), cent, Ps,
                          sf)
    return cent[:3 * nq].reshape(nq, 3), Ps[:nq], sf[:nq]


def bs_radius(z):
    return lib().orc_bs_radius(int(z))


def smearing(coords, weights, density, u_max, n_bins=1000, sigma=0.15):
    coords = _f64(coords).reshape(-1)
    G = coords.size // 3
    u = np.zeros(n_bins)
    P = np.zeros(n_bins)
    Ia = np.zeros(n_bins)
    lib().orc_smearing(G, coords, _f64(weights), _f64(density), float(u_max),
                       int(n_bins), float(sigma), u, P, Ia)
    return u, P, Ia


def onee(kind, index, prims, r, T=None, w=None):
    """One-electron function at points r (n,3): kind 1 primitive `index`, 3
    orbital `index`, 5 density, 6 the reference's Laplacian (orc_onee).
    T (nmo, nprim) coefficients of the unnormalised primitives, w (nmo)."""
    L = lib()
    alpha = _f64(prims["alpha"])
    nprim = len(alpha)
    tmn = _i32(np.asarray(prims["tmn"]).reshape(nprim, 3).T.reshape(-1))
    r = _f64(r).reshape(-1)
    n = r.size // 3
    Tf = None if T is None else np.asfortranarray(T, dtype=np.float64)
    nmo = 0 if T is None else Tf.shape[0]
    wv = None if w is None else _f64(w)
    out = np.zeros(max(n, 1))
    L.orc_onee(int(kind), int(index), nprim, alpha, tmn, _f64(prims["x"]), _f64(prims["y"]),
               _f64(prims["z"]), nmo, None if Tf is None else Tf.ctypes.data,
               None if wv is None else wv.ctypes.data, n, r, out)
    return out[:n]
