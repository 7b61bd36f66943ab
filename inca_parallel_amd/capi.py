"""ctypes binding of libinca_hip.so, the C-ABI declared in include/inca_hip.h.

This is the Python face of the same boundary the Fortran host binds with
ISO_C_BINDING; tests and bench.py drive the engine through it.  Loading
fails loudly when the library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "libinca_hip.so")

INCA_FP64 = 0
INCA_MIXED = 1
INCA_ERANGE = -6

# every symbol include/inca_hip.h declares
SYMBOLS = ["inca_create", "inca_destroy", "inca_last_error", "inca_set_primitives",
           "inca_grid_multicenter", "inca_grid_radial_plot", "inca_grid_cube",
           "inca_grid_points", "inca_get_grid", "inca_get_radial_layout",
           "inca_load_dm2p", "inca_set_quartets", "inca_evaluate", "inca_get_ivec",
           "inca_integrate", "inca_radial_values", "inca_smearing", "inca_get_timings",
           "inca_class_histogram", "inca_set_mixed_tail", "inca_get_mixed_split", "inca_set_prepare_chunk", "inca_release_scratch",
           "inca_set_tiles", "inca_autotune", "inca_get_prepared",
           "inca_set_polynomials", "inca_get_analytic_integral", "inca_set_point_cutoff",
           "inca_set_cube_contraction",
           "inca_get_eval_counts", "inca_set_grouping", "inca_get_groups", "inca_get_eval_flops",
           "inca_release_records", "inca_set_keep_records",
           "inca_prepare", "inca_evaluate_prepared", "inca_evaluate_prepared_noalloc", "inca_autogrid", "inca_bragg_slater", "inca_onee", "inca_batch_create", "inca_batch_destroy",
           "inca_batch_last_error", "inca_batch_evaluate", "inca_batch_get_ivec",
           "inca_batch_autotune", "inca_batch_set_points", "inca_batch_info",
           "inca_mlp_create", "inca_mlp_destroy", "inca_mlp_last_error", "inca_mlp_get_params",
           "inca_mlp_set_params", "inca_mlp_load", "inca_mlp_forward", "inca_mlp_loss_grad", "inca_mlp_adam_step",
           "inca_atom_features", "inca_learned_autogrid",
           "inca_lebedev", "inca_gauleg"]

_lib = None
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64 = C.POINTER(C.c_int64)


class IncaError(RuntimeError):
    pass


def load():
    """Load libinca_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise IncaError("libinca_hip.so not found at %s: run "
                        "`python -m inca_parallel_amd.build`" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.inca_create.argtypes = [C.c_int32, C.POINTER(vp)]
    L.inca_destroy.argtypes = [vp]
    L.inca_last_error.argtypes = [vp]
    L.inca_last_error.restype = C.c_char_p
    L.inca_set_primitives.argtypes = [vp, C.c_int32, _dp, _ip, _dp, _dp, _dp, C.c_int32]
    L.inca_grid_multicenter.argtypes = [vp, C.c_int32, _dp, _ip, _ip, _dp, _dp, C.c_int32,
                                        _i64, _i64, _i64]
    L.inca_grid_radial_plot.argtypes = [vp, C.c_int32, _dp, _dp, _ip, C.c_int32, _i64, _i64]
    L.inca_grid_cube.argtypes = [vp, _dp, _dp, _ip, _i64]
    L.inca_grid_points.argtypes = [vp, C.c_int64, _dp]
    L.inca_get_grid.argtypes = [vp, vp, vp, vp, vp]
    L.inca_get_radial_layout.argtypes = [vp, vp, vp]
    L.inca_load_dm2p.argtypes = [vp, C.c_char_p, _i64, C.POINTER(C.c_int32)]
    L.inca_set_quartets.argtypes = [vp, C.c_int64, vp, C.c_int32]
    L.inca_evaluate.argtypes = [vp, C.c_int32, C.c_double, vp, vp]
    L.inca_get_ivec.argtypes = [vp, vp]
    L.inca_integrate.argtypes = [vp, vp, vp, vp, vp]
    L.inca_radial_values.argtypes = [vp, vp]
    L.inca_smearing.argtypes = [vp, C.c_double, C.c_int32, C.c_double, vp, vp, vp]
    L.inca_get_timings.argtypes = [vp, _dp]
    L.inca_class_histogram.argtypes = [vp, C.c_int32, vp]
    L.inca_set_mixed_tail.argtypes = [vp, C.c_double]
    L.inca_set_prepare_chunk.argtypes = [vp, C.c_int64]
    L.inca_release_scratch.argtypes = [vp]
    L.inca_set_tiles.argtypes = [vp, C.c_int32, C.c_int32]
    L.inca_autotune.argtypes = [vp, vp]
    L.inca_get_prepared.argtypes = [vp, C.c_int64, C.c_int64, vp, vp]
    L.inca_set_polynomials.argtypes = [vp, C.c_int32]
    L.inca_get_mixed_split.argtypes = [vp, vp, vp, vp]
    L.inca_get_analytic_integral.argtypes = [vp, vp]
    L.inca_set_point_cutoff.argtypes = [vp, C.c_double]
    L.inca_set_grouping.argtypes = [vp, C.c_int32]
    L.inca_get_groups.argtypes = [vp, vp, vp, vp]
    L.inca_get_eval_flops.argtypes = [vp, vp]
    L.inca_release_records.argtypes = [vp]
    L.inca_set_keep_records.argtypes = [vp, C.c_int32]
    L.inca_set_cube_contraction.argtypes = [vp, C.c_int32]
    L.inca_get_eval_counts.argtypes = [vp, vp, vp]
    L.inca_prepare.argtypes = [vp, C.c_int32, C.c_double, vp]
    L.inca_evaluate_prepared.argtypes = [vp, C.c_int32, vp]
    L.inca_evaluate_prepared_noalloc.argtypes = [vp, C.c_int32, vp, vp]
    L.inca_autogrid.argtypes = [C.c_int32, vp, vp, C.c_int32, C.c_int32, vp, C.c_int32, vp,
                                vp, vp, vp]
    L.inca_bragg_slater.argtypes = [C.c_int32, vp, vp]
    L.inca_onee.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, vp, vp, vp]
    L.inca_batch_create.argtypes = [C.c_int32, C.c_int32, vp, C.POINTER(vp)]
    L.inca_batch_destroy.argtypes = [vp]
    L.inca_batch_last_error.argtypes = [vp]
    L.inca_batch_last_error.restype = C.c_char_p
    L.inca_batch_evaluate.argtypes = [vp, vp, vp]
    L.inca_batch_get_ivec.argtypes = [vp, C.c_int32, vp]
    L.inca_batch_autotune.argtypes = [vp, vp]
    L.inca_batch_set_points.argtypes = [vp, C.c_int32]
    L.inca_batch_info.argtypes = [vp, vp, vp, vp, vp, vp]
    L.inca_mlp_create.argtypes = [C.c_int32, C.c_uint64, C.POINTER(vp)]
    L.inca_mlp_destroy.argtypes = [vp]
    L.inca_mlp_last_error.argtypes = [vp]
    L.inca_mlp_last_error.restype = C.c_char_p
    L.inca_mlp_get_params.argtypes = [vp, vp]
    L.inca_mlp_set_params.argtypes = [vp, vp]
    L.inca_mlp_forward.argtypes = [vp, C.c_int32, vp, vp]
    L.inca_mlp_load.argtypes = [vp, C.c_char_p]
    L.inca_mlp_loss_grad.argtypes = [vp, C.c_int32, vp, vp, vp, vp]
    L.inca_mlp_adam_step.argtypes = [vp, C.c_float]
    L.inca_atom_features.argtypes = [C.c_int32, vp, vp, C.c_int32, vp, vp, vp, vp]
    L.inca_learned_autogrid.argtypes = [vp, C.c_int32, vp, vp, C.c_int32, vp, vp, vp, C.c_int32,
                                        C.c_int32, vp, vp, vp, vp, vp, vp, vp]
    L.inca_lebedev.argtypes = [C.c_int32, _dp, _dp, _dp, _dp]
    L.inca_gauleg.argtypes = [C.c_int32, _dp, _dp]
    for s in SYMBOLS:
        if s not in ("inca_last_error", "inca_batch_last_error", "inca_mlp_last_error"):
            getattr(L, s).restype = C.c_int
    _lib = L
    return L


def lebedev(n):
    """(n,4) array [x, y, z, w] of the n-point Lebedev-Laikov rule (sum w = 1)."""
    L = load()
    x, y, z, w = (np.zeros(n) for _ in range(4))
    if L.inca_lebedev(int(n), x, y, z, w) != 0:
        raise IncaError("no Lebedev rule with %d points" % n)
    return np.stack([x, y, z, w], axis=1)


def gauleg(n):
    L = load()
    t, w = np.zeros(n), np.zeros(n)
    if L.inca_gauleg(int(n), t, w) != 0:
        raise IncaError("bad Gauss-Legendre order %d" % n)
    return t, w


def _ptr(a):
    return None if a is None else a.ctypes.data


def autogrid(an, cartes, nohydro=True, betaone=False, radius=None):
    """Integration centres (inca_autogrid): cent (nq,3), Ps (nq), sfalpha (nq).
    cartes is (natoms,3) in bohr; radius optional per-atom radii in bohr."""
    L = load()
    an = np.ascontiguousarray(an, np.int32)
    n = len(an)
    cm = np.ascontiguousarray(np.asarray(cartes, np.float64).reshape(n, 3).T.reshape(-1))
    rad = None if radius is None else np.ascontiguousarray(radius, np.float64)
    nq = C.c_int32()
    L.inca_autogrid(n, an.ctypes.data, cm.ctypes.data, int(bool(nohydro)), int(bool(betaone)),
                    _ptr(rad), 0, C.addressof(nq), None, None, None)
    cent, Ps, sf = np.zeros((nq.value, 3)), np.zeros(nq.value), np.zeros(nq.value)
    rc = L.inca_autogrid(n, an.ctypes.data, cm.ctypes.data, int(bool(nohydro)),
                         int(bool(betaone)), _ptr(rad), nq.value, C.addressof(nq),
                         cent.ctypes.data, Ps.ctypes.data, sf.ctypes.data)
    if rc != 0:
        raise IncaError("inca_autogrid failed (%d)" % rc)
    return cent, Ps, sf


MLP_IN, MLP_OUT, MLP_NPARAM = 16, 9, 9993
NRAD_CLASSES = (50, 75, 100, 150)
NANG_CLASSES = (110, 194, 302, 590)


def bragg_slater(an):
    """Bragg-Slater radii in bohr of the atomic numbers `an` (inca_bragg_slater)."""
    L = load()
    an = np.ascontiguousarray(an, np.int32)
    r = np.zeros(max(len(an), 1))
    if L.inca_bragg_slater(len(an), an.ctypes.data, r.ctypes.data) != 0:
        raise IncaError("inca_bragg_slater failed")
    return r[:len(an)]


def _molecule_arrays(an, cartes, alpha, tmn, prim_atom):
    an = np.ascontiguousarray(an, np.int32)
    n = len(an)
    cm = np.ascontiguousarray(np.asarray(cartes, np.float64).reshape(n, 3).T.reshape(-1))
    alpha = np.ascontiguousarray(alpha, np.float64)
    npr = len(alpha)
    tm = np.ascontiguousarray(np.asarray(tmn, np.int32).reshape(npr, 3).T.reshape(-1))
    ra = np.ascontiguousarray(prim_atom, np.int32)
    return an, cm, alpha, tm, ra


def atom_features(an, cartes, alpha, tmn, prim_atom):
    """(natoms,16) float32 inputs of the learned autogrid (inca_atom_features)."""
    L = load()
    an, cm, alpha, tm, ra = _molecule_arrays(an, cartes, alpha, tmn, prim_atom)
    f = np.zeros((len(an), MLP_IN), np.float32)
    rc = L.inca_atom_features(len(an), an.ctypes.data, cm.ctypes.data, len(alpha),
                              alpha.ctypes.data, tm.ctypes.data, ra.ctypes.data, f.ctypes.data)
    if rc != 0:
        raise IncaError("inca_atom_features failed (%d)" % rc)
    return f


class MLP:
    """The learned-autogrid network on one device (inca_mlp_*)."""

    def __init__(self, device=0, seed=1):
        self._L = load()
        self._h = C.c_void_p()
        rc = self._L.inca_mlp_create(int(device), int(seed), C.byref(self._h))
        if rc != 0:
            msg = self._L.inca_mlp_last_error(self._h).decode() if self._h else ""
            self.close()
            raise IncaError("inca_mlp_create failed (%d): %s" % (rc, msg))

    def _ck(self, rc, what):
        if rc != 0:
            raise IncaError("%s failed (%d): %s" % (
                what, rc, self._L.inca_mlp_last_error(self._h).decode()))

    def params(self):
        p = np.zeros(MLP_NPARAM, np.float32)
        self._ck(self._L.inca_mlp_get_params(self._h, p.ctypes.data), "inca_mlp_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == MLP_NPARAM
        self._ck(self._L.inca_mlp_set_params(self._h, p.ctypes.data), "inca_mlp_set_params")

    def load(self, path):
        """Parameters from a .npy file (inca_mlp_load)."""
        self._ck(self._L.inca_mlp_load(self._h, os.fsencode(path)), "inca_mlp_load")

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        out = np.zeros((x.shape[0], MLP_OUT), np.float32)
        self._ck(self._L.inca_mlp_forward(self._h, x.shape[0], x.ctypes.data, out.ctypes.data),
                 "inca_mlp_forward")
        return out

    def loss_grad(self, x, target, want_grad=True):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(target, np.float32)
        loss = C.c_double()
        g = np.zeros(MLP_NPARAM, np.float32) if want_grad else None
        self._ck(self._L.inca_mlp_loss_grad(self._h, x.shape[0], x.ctypes.data, t.ctypes.data,
                                            C.addressof(loss), _ptr(g)), "inca_mlp_loss_grad")
        return loss.value, g

    def adam_step(self, lr):
        self._ck(self._L.inca_mlp_adam_step(self._h, float(lr)), "inca_mlp_adam_step")

    def learned_autogrid(self, an, cartes, alpha, tmn, prim_atom, nohydro=True):
        """cent (nq,3), Ps, sfalpha, nradc, nangc, per-atom radii."""
        an, cm, alpha, tm, ra = _molecule_arrays(an, cartes, alpha, tmn, prim_atom)
        nq = C.c_int32()
        args = (self._h, len(an), an.ctypes.data, cm.ctypes.data, len(alpha), alpha.ctypes.data,
                tm.ctypes.data, ra.ctypes.data, int(bool(nohydro)))
        rc = self._L.inca_learned_autogrid(*args, 0, C.addressof(nq), None, None, None, None,
                                           None, None)
        if rc not in (0, INCA_ERANGE):
            self._ck(rc, "inca_learned_autogrid")
        n = nq.value
        cent, Ps, sf = np.zeros((n, 3)), np.zeros(n), np.zeros(n)
        nr, na = np.zeros(n, np.int32), np.zeros(n, np.int32)
        rad = np.zeros(len(an))
        self._ck(self._L.inca_learned_autogrid(*args, n, C.addressof(nq), cent.ctypes.data,
                                               Ps.ctypes.data, sf.ctypes.data, nr.ctypes.data,
                                               na.ctypes.data, rad.ctypes.data),
                 "inca_learned_autogrid")
        return cent, Ps, sf, nr, na, rad

    def close(self):
        if self._h:
            self._L.inca_mlp_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many prepared contexts evaluated by one captured persistent launch."""

    def __init__(self, engines, device=0):
        self._L = load()
        self.engines_G = [e.G for e in engines]
        arr = (C.c_void_p * len(engines))(*[e._h.value for e in engines])
        self._h = C.c_void_p()
        rc = self._L.inca_batch_create(int(device), len(engines), arr, C.byref(self._h))
        if rc != 0:
            msg = self._L.inca_batch_last_error(self._h).decode() if self._h else ""
            self.close()
            raise IncaError("inca_batch_create failed (%d): %s" % (rc, msg))
        self.nmol = len(engines)

    def _ck(self, rc, what):
        if rc != 0:
            raise IncaError("%s failed (%d): %s" % (
                what, rc, self._L.inca_batch_last_error(self._h).decode()))

    def evaluate(self):
        ti, tv = np.zeros(self.nmol), np.zeros(self.nmol)
        self._ck(self._L.inca_batch_evaluate(self._h, ti.ctypes.data, tv.ctypes.data),
                 "inca_batch_evaluate")
        return ti, tv

    def ivec(self, m):
        iv = np.zeros(max(self.engines_G[m], 1))
        self._ck(self._L.inca_batch_get_ivec(self._h, int(m), iv.ctypes.data),
                 "inca_batch_get_ivec")
        return iv[:self.engines_G[m]]

    def autotune(self):
        b = C.c_int32()
        self._ck(self._L.inca_batch_autotune(self._h, C.addressof(b)), "inca_batch_autotune")
        return b.value

    def set_points(self, points_per_lane):
        self._ck(self._L.inca_batch_set_points(self._h, int(points_per_lane)),
                 "inca_batch_set_points")

    def info(self):
        n, p, nb, r, ms = C.c_int64(), C.c_int32(), C.c_int32(), C.c_int64(), C.c_double()
        self._ck(self._L.inca_batch_info(self._h, C.addressof(n), C.addressof(p),
                                         C.addressof(nb), C.addressof(r), C.addressof(ms)),
                 "inca_batch_info")
        return dict(items=n.value, points_per_lane=p.value, blocks=nb.value,
                    replays=r.value, last_ms=ms.value)

    def close(self):
        if self._h:
            self._L.inca_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Engine:
    """One device context: primitives, a grid, DM2 records, Ivec."""

    def __init__(self, device: int = 0):
        self._L = load()
        self._h = C.c_void_p()
        rc = self._L.inca_create(int(device), C.byref(self._h))
        if rc != 0:
            raise IncaError("inca_create failed (%d): %s" % (
                rc, self._L.inca_last_error(None).decode()))
        self.G = 0
        self.nquad = 0
        self.nradi = 0

    def _ck(self, rc, what):
        if rc != 0:
            raise IncaError("%s failed (%d): %s" % (
                what, rc, self._L.inca_last_error(self._h).decode()))

    def close(self):
        if self._h:
            self._L.inca_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_primitives(self, alpha, tmn, x, y, z, normalize=True):
        alpha = np.ascontiguousarray(alpha, np.float64)
        n = len(alpha)
        tmn_cm = np.ascontiguousarray(np.asarray(tmn, np.int32).reshape(n, 3).T.reshape(-1))
        self._ck(self._L.inca_set_primitives(
            self._h, n, alpha, tmn_cm, np.ascontiguousarray(x, np.float64),
            np.ascontiguousarray(y, np.float64), np.ascontiguousarray(z, np.float64),
            int(bool(normalize))), "inca_set_primitives")

    def grid_multicenter(self, cent, nradc, nangc, sfalpha, Ps, nosym=False):
        cent = np.ascontiguousarray(cent, np.float64).reshape(-1)
        nq = len(nradc)
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        self._ck(self._L.inca_grid_multicenter(
            self._h, nq, cent, np.ascontiguousarray(nradc, np.int32),
            np.ascontiguousarray(nangc, np.int32), np.ascontiguousarray(sfalpha, np.float64),
            np.ascontiguousarray(Ps, np.float64), int(bool(nosym)),
            C.byref(a), C.byref(b), C.byref(c)), "inca_grid_multicenter")
        self.G, self.nquad = c.value, nq
        return a.value, b.value, c.value

    def grid_radial_plot(self, tart, step, nang, nosym=False):
        tart = np.ascontiguousarray(tart, np.float64).reshape(-1)
        nr, ng = C.c_int64(), C.c_int64()
        self._ck(self._L.inca_grid_radial_plot(
            self._h, len(step), tart, np.ascontiguousarray(step, np.float64),
            np.ascontiguousarray(nang, np.int32), int(bool(nosym)),
            C.byref(nr), C.byref(ng)), "inca_grid_radial_plot")
        self.G, self.nradi = ng.value, nr.value
        return nr.value, ng.value

    def grid_cube(self, center, step, np_i):
        n = C.c_int64()
        self._ck(self._L.inca_grid_cube(
            self._h, np.ascontiguousarray(center, np.float64),
            np.ascontiguousarray(step, np.float64), np.ascontiguousarray(np_i, np.int32),
            C.byref(n)), "inca_grid_cube")
        self.G = n.value
        return n.value

    def grid_points(self, r):
        r = np.ascontiguousarray(r, np.float64).reshape(-1)
        self._ck(self._L.inca_grid_points(self._h, r.size // 3, r), "inca_grid_points")
        self.G = r.size // 3

    def get_grid(self):
        G = self.G
        r = np.zeros((G, 3))
        w, wv = np.zeros(G), np.zeros(G)
        mp = np.zeros(G, np.int32)
        self._ck(self._L.inca_get_grid(self._h, _ptr(r), _ptr(w), _ptr(wv), _ptr(mp)),
                 "inca_get_grid")
        return dict(r=r, w=w, w_vee=wv, map=mp)

    def radial_layout(self):
        radi = np.zeros(self.nradi)
        smn = np.zeros(self.nradi, np.int32)
        self._ck(self._L.inca_get_radial_layout(self._h, _ptr(radi), _ptr(smn)),
                 "inca_get_radial_layout")
        return radi, smn

    def load_dm2p(self, path):
        n = C.c_int64()
        fmt = C.c_int32()
        self._ck(self._L.inca_load_dm2p(self._h, os.fsencode(path), C.byref(n), C.byref(fmt)),
                 "inca_load_dm2p")
        return n.value, fmt.value

    def set_quartets(self, records=None, ijkl=None, d=None, device_ptr=None, n=None):
        """records: structured/raw 24-byte records; or ijkl (n,4) int32 + d (n,);
        or device_ptr (int) + n for records already in device memory."""
        if device_ptr is not None:
            self._ck(self._L.inca_set_quartets(self._h, int(n), C.c_void_p(int(device_ptr)), 1),
                     "inca_set_quartets")
            return
        if records is None:
            ijkl = np.asarray(ijkl, np.int32).reshape(-1, 4)
            records = np.empty(len(ijkl), dtype=[("ijkl", "<i4", 4), ("d", "<f8")])
            records["ijkl"] = ijkl
            records["d"] = d
        records = np.ascontiguousarray(records)
        nrec = records.nbytes // 24
        self._ck(self._L.inca_set_quartets(self._h, nrec, records.ctypes.data, 0),
                 "inca_set_quartets")

    def evaluate(self, precision=INCA_FP64, thresh=1e-12, want_ivec=True):
        iv = np.zeros(max(self.G, 1)) if want_ivec else None
        cnt = np.zeros(4, np.int64)
        self._ck(self._L.inca_evaluate(self._h, int(precision), float(thresh), _ptr(iv),
                                       cnt.ctypes.data), "inca_evaluate")
        return (iv[:self.G] if want_ivec else None), cnt

    def prepare(self, precision=INCA_FP64, thresh=1e-12):
        """Pre-pass only (inca_prepare); returns the four counters."""
        cnt = np.zeros(4, np.int64)
        self._ck(self._L.inca_prepare(self._h, int(precision), float(thresh), cnt.ctypes.data),
                 "inca_prepare")
        return cnt

    def evaluate_prepared(self, precision=INCA_FP64, want_ivec=True):
        """Grid loop only, over the quartets already prepared (inca_evaluate_prepared)."""
        iv = np.zeros(max(self.G, 1)) if want_ivec else None
        self._ck(self._L.inca_evaluate_prepared(self._h, int(precision), _ptr(iv)),
                 "inca_evaluate_prepared")
        return iv[:self.G] if want_ivec else None

    def evaluate_prepared_noalloc(self, precision, ivec_device_ptr, stream=None):
        """inca_evaluate_prepared_noalloc: the grid loop into a device array of
        G doubles (an int address), enqueued on `stream` (an int hipStream_t
        address, or None for the context's stream); returns without waiting."""
        self._ck(self._L.inca_evaluate_prepared_noalloc(
            self._h, int(precision), C.c_void_p(int(ivec_device_ptr)),
            C.c_void_p(int(stream)) if stream else None), "inca_evaluate_prepared_noalloc")

    def onee(self, kind, index=0, T=None, w=None):
        """One-electron function on the current grid (inca_onee): kind 1
        primitive `index`, 3 orbital `index`, 5 density, 6 the reference's
        Laplacian; T (nmo, nprim), w (nmo)."""
        out = np.zeros(max(self.G, 1))
        Tf = None if T is None else np.asfortranarray(T, dtype=np.float64)
        wv = None if w is None else np.ascontiguousarray(w, np.float64)
        nmo = 0 if Tf is None else Tf.shape[0]
        self._ck(self._L.inca_onee(self._h, int(kind), int(index), nmo, _ptr(Tf), _ptr(wv),
                                   out.ctypes.data), "inca_onee")
        return out[:self.G]

    def ivec(self):
        iv = np.zeros(max(self.G, 1))
        self._ck(self._L.inca_get_ivec(self._h, iv.ctypes.data), "inca_get_ivec")
        return iv[:self.G]

    def integrate(self):
        ti, tv = C.c_double(), C.c_double()
        ipc, vpc = np.zeros(self.nquad), np.zeros(self.nquad)
        self._ck(self._L.inca_integrate(self._h, C.addressof(ti), C.addressof(tv),
                                        ipc.ctypes.data, vpc.ctypes.data), "inca_integrate")
        return ti.value, tv.value, ipc, vpc

    def radial_values(self):
        out = np.zeros(max(self.nradi, 1))
        self._ck(self._L.inca_radial_values(self._h, out.ctypes.data), "inca_radial_values")
        return out[:self.nradi]

    def smearing(self, u_max, n_bins=1000, sigma=0.15):
        u, P, Ia = np.zeros(n_bins), np.zeros(n_bins), np.zeros(n_bins)
        self._ck(self._L.inca_smearing(self._h, float(u_max), int(n_bins), float(sigma),
                                       u.ctypes.data, P.ctypes.data, Ia.ctypes.data),
                 "inca_smearing")
        return u, P, Ia

    def class_histogram(self, nbins=64):
        h = np.zeros(nbins, np.int64)
        self._ck(self._L.inca_class_histogram(self._h, int(nbins), h.ctypes.data),
                 "inca_class_histogram")
        return h

    def release_scratch(self):
        """Free the records and scratch, keeping grid and prepared quartets."""
        self._ck(self._L.inca_release_scratch(self._h), "inca_release_scratch")

    def release_records(self):
        """Free the prepared records of a grouped preparation (only the group
        evaluation remains; inca_release_records)."""
        self._ck(self._L.inca_release_records(self._h), "inca_release_records")

    def set_keep_records(self, keep):
        """0: build the groups chunk by chunk and keep no prepared records
        (inca_set_keep_records)."""
        self._ck(self._L.inca_set_keep_records(self._h, int(bool(keep))), "inca_set_keep_records")

    def set_prepare_chunk(self, records):
        """Records per pre-pass chunk (inca_set_prepare_chunk)."""
        self._ck(self._L.inca_set_prepare_chunk(self._h, int(records)), "inca_set_prepare_chunk")

    def set_mixed_tail(self, tail):
        self._ck(self._L.inca_set_mixed_tail(self._h, float(tail)), "inca_set_mixed_tail")

    def mixed_split(self):
        a, b, t = C.c_int64(), C.c_int64(), C.c_double()
        self._ck(self._L.inca_get_mixed_split(self._h, C.addressof(a), C.addressof(b),
                                              C.addressof(t)), "inca_get_mixed_split")
        return a.value, b.value, t.value

    def analytic_integral(self):
        """Sum of the prepared terms' integrals over all space."""
        t = C.c_double()
        self._ck(self._L.inca_get_analytic_integral(self._h, C.addressof(t)),
                 "inca_get_analytic_integral")
        return t.value

    def set_point_cutoff(self, log2_rel):
        """0: evaluate every quartet at every point; else skip tiles beyond
        the 2^-log2_rel radius of each quartet."""
        self._ck(self._L.inca_set_point_cutoff(self._h, float(log2_rel)),
                 "inca_set_point_cutoff")

    def set_cube_contraction(self, on):
        """2 (default): cube grids by the per-axis contraction (every quartet
        at every voxel) on the matrix cores; 1: the FP64 range on the vector
        units; 0: term by term like other grids."""
        self._ck(self._L.inca_set_cube_contraction(self._h, int(on)),
                 "inca_set_cube_contraction")

    def eval_counts(self):
        """(reference evals, executed evals) of the last grid loop."""
        a, b = C.c_int64(), C.c_int64()
        self._ck(self._L.inca_get_eval_counts(self._h, C.addressof(a), C.addressof(b)),
                 "inca_get_eval_counts")
        return a.value, b.value

    def set_grouping(self, on):
        """1 (default): evaluate by shell-quartet groups (built at the next
        preparation); 0: record by record (inca_set_grouping)."""
        self._ck(self._L.inca_set_grouping(self._h, int(bool(on))), "inca_set_grouping")

    def groups(self):
        """(groups, FP64-range groups, pool coefficients) of the last preparation."""
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        self._ck(self._L.inca_get_groups(self._h, C.addressof(a), C.addressof(b),
                                         C.addressof(c)), "inca_get_groups")
        return a.value, b.value, c.value

    def eval_flops(self):
        """FP64-equivalent operations of the last grid loop when it ran by groups."""
        f = C.c_double()
        self._ck(self._L.inca_get_eval_flops(self._h, C.addressof(f)), "inca_get_eval_flops")
        return f.value

    def set_polynomials(self, exact):
        self._ck(self._L.inca_set_polynomials(self._h, int(bool(exact))), "inca_set_polynomials")

    def set_tiles(self, p64, p32):
        self._ck(self._L.inca_set_tiles(self._h, int(p64), int(p32)), "inca_set_tiles")

    def autotune(self):
        best = np.zeros(2, np.int32)
        self._ck(self._L.inca_autotune(self._h, best.ctypes.data), "inca_autotune")
        return int(best[0]), int(best[1])

    def prepared(self, first=0, n=None):
        """Prepared quartet records as a (n, stride) float64 array."""
        st = C.c_int32()
        self._ck(self._L.inca_get_prepared(self._h, 0, 0, None, C.addressof(st)),
                 "inca_get_prepared")
        if n is None:
            n = self.mixed_split()[0] + self.mixed_split()[1] - first
        out = np.zeros((max(n, 1), st.value))
        self._ck(self._L.inca_get_prepared(self._h, int(first), int(n), out.ctypes.data,
                                           C.addressof(st)), "inca_get_prepared")
        return out[:n]

    def timings(self):
        t = np.zeros(3)
        self._ck(self._L.inca_get_timings(self._h, t), "inca_get_timings")
        return t
