"""ctypes wrapper over the FP64 CPU oracle (oracle/scx_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by the
cpu_baseline leg of bench.py, as the checker. The product path (scisim_amd) never
imports this package.
"""

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libscx_oracle.so")
_lib = None

_d = ctypes.c_double
_pd = ctypes.POINTER(ctypes.c_double)
_pu32 = ctypes.POINTER(ctypes.c_uint32)
_pi32 = ctypes.POINTER(ctypes.c_int32)


class OrcParams(ctypes.Structure):
    _fields_ = [("dt", _d), ("g", _d * 3), ("mu", _d), ("CoR", _d), ("tol", _d),
                ("max_iters", ctypes.c_uint32), ("eval_every", ctypes.c_uint32),
                ("integrator", ctypes.c_int32), ("order", ctypes.c_int32),
                ("warm_start", ctypes.c_int32), ("staggering", ctypes.c_int32), ("stabilize", ctypes.c_int32),
                ("pad", ctypes.c_int32), ("penetration_threshold", _d)]


class OrcStats(ctypes.Structure):
    _fields_ = [("npairs", ctypes.c_uint64), ("ncontacts", ctypes.c_uint32), ("ncolours", ctypes.c_uint32),
                ("nrounds", ctypes.c_uint32), ("sweeps", ctypes.c_uint32), ("error", _d),
                ("succeeded", ctypes.c_int32), ("pad", ctypes.c_int32)]


def build():
    """Compile the oracle with its Makefile (g++, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_grid_pairs.restype = ctypes.c_int64
        L.orc_grid_pairs.argtypes = [ctypes.c_int, ctypes.c_uint32, _pd, _pd, _pu32, ctypes.c_uint64]
        L.orc_all_pairs.restype = ctypes.c_int64
        L.orc_all_pairs.argtypes = [ctypes.c_int, ctypes.c_uint32, _pd, _pd, _pu32, ctypes.c_uint64]
        L.orc_ccd_coeffs.argtypes = [_pd, _pd, _d, _pd, _pd, _d, _pd]
        L.orc_ccd_happens.restype = ctypes.c_int
        L.orc_ccd_happens.argtypes = [_pd, _pd]
        L.orc_fb_err.restype = _d
        L.orc_fb_err.argtypes = [ctypes.c_int, _d, _pd, _pd]
        L.orc_fb_phi.restype = None
        L.orc_fb_phi.argtypes = [ctypes.c_int, _d, _pd, _pd, _pd]
        L.orc_local_solve.argtypes = [ctypes.c_int, _d, _d, _d, _pd, _pd]
        L.orc_transport.argtypes = [ctypes.c_int, _pd, _pd, _pd, _pd]
        L.orc_transport.restype = None
        L.orc_umap_flow.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _pd, _pd, _d, _pd, _pd, _pd, _pd]
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.c_int]
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        L.orc_clone.restype = ctypes.c_void_p
        L.orc_clone.argtypes = [ctypes.c_void_p]
        L.orc_step.restype = ctypes.c_int
        L.orc_step.argtypes = [ctypes.c_void_p, ctypes.c_int, _pd, _pd, _pd, _pd, _pd, _pd, ctypes.c_int, _pd, _pd,
                               ctypes.POINTER(OrcParams), ctypes.POINTER(OrcStats)]
        L.orc_run.restype = ctypes.c_int
        L.orc_run.argtypes = [ctypes.c_void_p, ctypes.c_int, _pd, _pd, _pd, _pd, _pd, _pd, ctypes.c_int, _pd, _pd,
                              ctypes.POINTER(OrcParams), ctypes.c_int, ctypes.POINTER(OrcStats), _pu32, _pu32]
        L.orc_get_contacts.restype = ctypes.c_int
        L.orc_get_contacts.argtypes = [ctypes.c_void_p, ctypes.c_uint32, _pi32, _pi32, _pi32, _pd, _pd, _pd, _pd,
                                       _pi32]
        L.orc_get_v2.restype = ctypes.c_int
        L.orc_get_v2.argtypes = [ctypes.c_void_p, _pd, ctypes.c_uint64]
        L.orc_get_hist.restype = ctypes.c_int
        L.orc_get_hist.argtypes = [ctypes.c_void_p, _pd, ctypes.c_uint32]
        L.orc_friction_solve.restype = _d
        L.orc_friction_solve.argtypes = [ctypes.c_int, ctypes.c_int, _pd, _pd, ctypes.c_int, _pi32, _pi32, _pd, _pd,
                                         _pd, _pd, _pd, _d, _pd, _pd, _pd, ctypes.c_uint32, ctypes.c_uint32, _d,
                                         ctypes.c_int, _pu32]
        L.orc_colour.restype = ctypes.c_int
        L.orc_colour.argtypes = [ctypes.c_int, ctypes.c_int, _pi32, _pi32, _pi32, _pi32]
        L.orc_num_threads.restype = ctypes.c_int
        L.orc_set_threads.restype = None
        L.orc_set_threads.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def _p(a, t=_pd):
    return None if a is None else a.ctypes.data_as(t)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def grid_pairs(dim, mn, mx):
    """SpatialGridDetector::getPotentialOverlaps on (n, dim) min/max arrays -> (P, 2) uint32."""
    mn, mx = _f64(mn), _f64(mx)
    n = mn.shape[0]
    cap = max(16, 8 * n)
    while True:
        out = np.empty((cap, 2), dtype=np.uint32)
        k = lib().orc_grid_pairs(dim, n, _p(mn), _p(mx), _p(out, _pu32), cap)
        if k >= 0:
            return out[:k].copy()
        cap = -k


def all_pairs(dim, mn, mx):
    """SpatialGridDetector::getPotentialOverlapsAllPairs -> (P, 2) uint32 (lexicographic)."""
    mn, mx = _f64(mn), _f64(mx)
    n = mn.shape[0]
    cap = max(16, 8 * n)
    while True:
        out = np.empty((cap, 2), dtype=np.uint32)
        k = lib().orc_all_pairs(dim, n, _p(mn), _p(mx), _p(out, _pu32), cap)
        if k >= 0:
            return out[:k].copy()
        cap = -k


def ccd_coeffs(q0a, q1a, ra, q0b, q1b, rb):
    c = np.zeros(3)
    lib().orc_ccd_coeffs(_p(_f64(q0a)), _p(_f64(q1a)), ra, _p(_f64(q0b)), _p(_f64(q1b)), rb, _p(c))
    return c


def ccd_happens(c):
    t = ctypes.c_double(-1.0)
    hit = lib().orc_ccd_happens(_p(_f64(c)), ctypes.byref(t))
    return bool(hit), t.value


def fb_err(dim, mu, r, u):
    return lib().orc_fb_err(dim, mu, _p(_f64(r)), _p(_f64(u)))


def fb_phi(dim, mu, r, u):
    """The Fischer-Burmeister vector phi whose squared norm fb_err returns."""
    phi = np.zeros(dim)
    lib().orc_fb_phi(dim, mu, _p(_f64(r)), _p(_f64(u)), _p(phi))
    return phi


def transport(n_old, r_old, basis):
    """Cached world impulse r_old (with its normal n_old) rotated onto the normal of
    `basis` (D x D, columns n s t) and expressed in that basis."""
    B = _f64(basis)
    dim = B.shape[0]
    out = np.zeros(dim)
    lib().orc_transport(dim, _p(_f64(n_old)), _p(_f64(r_old)), _p(np.ascontiguousarray(B.T).ravel()), _p(out))
    return out


def local_solve(dim, mu, wn, wt, bt):
    r = np.zeros(dim)
    lib().orc_local_solve(dim, mu, wn, wt, _p(_f64(bt)), _p(r))
    return r


def umap_flow(dim, integrator, m, g, dt, q0, v0):
    q0, v0, m = _f64(q0), _f64(v0), _f64(m)
    q1, v1 = np.empty_like(q0), np.empty_like(v0)
    lib().orc_umap_flow(dim, integrator, m.size, _p(m), _p(_f64(g)), dt, _p(q0), _p(v0), _p(q1), _p(v1))
    return q1, v1


def colour(nbodies, a, b):
    a = np.ascontiguousarray(a, dtype=np.int32)
    b = np.ascontiguousarray(b, dtype=np.int32)
    col = np.empty(a.size, dtype=np.int32)
    nr = ctypes.c_int32(0)
    nc = lib().orc_colour(nbodies, a.size, _p(a, _pi32), _p(b, _pi32), _p(col, _pi32), ctypes.byref(nr))
    return nc, nr.value, col


def num_threads():
    return lib().orc_num_threads()


def set_threads(n):
    """Threads of the oracle's parallel parts (residual evaluation, per-contact set-up)."""
    lib().orc_set_threads(int(n))


class Sim:
    """Oracle simulation of GeometricImpactFrictionMap::flow steps on circles/spheres."""

    def __init__(self, dim, q, v, r, m, planes_x, planes_n, w=None, inertia=None):
        self.dim = dim
        self.q = _f64(q).ravel().copy()
        self.v = _f64(v).ravel().copy()
        self.r = _f64(r).ravel().copy()
        self.m = _f64(m).ravel().copy()
        self.n = self.r.size
        self.w = _f64(w).ravel().copy() if w is not None else (np.zeros(3 * self.n) if dim == 3 else None)
        self.inertia = _f64(inertia).ravel().copy() if inertia is not None else (
            (0.4 * self.m * self.r * self.r) if dim == 3 else None)
        self.px = _f64(planes_x).ravel().copy()
        self.pn = _f64(planes_n).ravel().copy()
        self.nplanes = self.px.size // dim
        self._ctx = lib().orc_create(dim)

    def clone(self):
        """An independent copy: state, impulse cache and last step."""
        c = Sim.__new__(Sim)
        c.__dict__.update({k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in self.__dict__.items()
                           if k != "_ctx"})
        c._ctx = lib().orc_clone(self._ctx)
        return c

    def __del__(self):
        if getattr(self, "_ctx", None) and _lib is not None:
            _lib.orc_destroy(self._ctx)
            self._ctx = None

    def step(self, dt, g, mu, CoR=0.0, tol=1e-6, max_iters=5000, eval_every=1, integrator=0, order=0,
             warm_start=0, staggering=0, stabilize=0, penetration_threshold=0.0):
        """One map step: staggering 0 geometric, 1 symplectic Euler (stabilize adds the
        min(0, gap)/dt term), 2 stabilized."""
        P = OrcParams(dt=dt, g=(_d * 3)(*(list(g) + [0.0] * (3 - len(g)))), mu=mu, CoR=CoR, tol=tol,
                      max_iters=max_iters, eval_every=eval_every, integrator=integrator, order=order,
                      warm_start=warm_start, staggering=staggering, stabilize=stabilize,
                      penetration_threshold=penetration_threshold)
        S = OrcStats()
        rc = lib().orc_step(self._ctx, self.n, _p(self.q), _p(self.v), _p(self.w), _p(self.r), _p(self.m),
                            _p(self.inertia), self.nplanes, _p(self.px), _p(self.pn), ctypes.byref(P),
                            ctypes.byref(S))
        if rc != 0:
            raise RuntimeError("oracle step failed (%d)" % rc)
        return {"npairs": S.npairs, "ncontacts": S.ncontacts, "ncolours": S.ncolours, "nrounds": S.nrounds,
                "sweeps": S.sweeps, "error": S.error, "succeeded": bool(S.succeeded)}

    def run(self, nsteps, dt, g, mu, CoR=0.0, tol=1e-6, max_iters=5000, eval_every=1, integrator=0, order=0,
            warm_start=0, staggering=0, stabilize=0, penetration_threshold=0.0):
        P = OrcParams(dt=dt, g=(_d * 3)(*(list(g) + [0.0] * (3 - len(g)))), mu=mu, CoR=CoR, tol=tol,
                      max_iters=max_iters, eval_every=eval_every, integrator=integrator, order=order,
                      warm_start=warm_start, staggering=staggering, stabilize=stabilize,
                      penetration_threshold=penetration_threshold)
        S = OrcStats()
        ms = ctypes.c_uint32(0)
        nf = ctypes.c_uint32(0)
        rc = lib().orc_run(self._ctx, self.n, _p(self.q), _p(self.v), _p(self.w), _p(self.r), _p(self.m),
                           _p(self.inertia), self.nplanes, _p(self.px), _p(self.pn), ctypes.byref(P), nsteps,
                           ctypes.byref(S), ctypes.byref(ms), ctypes.byref(nf))
        if rc != 0:
            raise RuntimeError("oracle run failed (%d)" % rc)
        return {"npairs": S.npairs, "ncontacts": S.ncontacts, "sweeps": S.sweeps, "error": S.error,
                "max_sweeps": ms.value, "failures": nf.value}

    def contacts(self):
        D = self.dim
        C = lib().orc_get_contacts(self._ctx, 0, None, None, None, None, None, None, None, None)
        C = -C if C < 0 else C
        a = np.empty(C, np.int32)
        b = np.empty(C, np.int32)
        pl = np.empty(C, np.int32)
        basis = np.empty((C, D, D))
        alpha = np.empty(C)
        beta = np.empty((C, D - 1))
        rw = np.empty((C, D))
        col = np.empty(C, np.int32)
        lib().orc_get_contacts(self._ctx, C, _p(a, _pi32), _p(b, _pi32), _p(pl, _pi32), _p(basis), _p(alpha),
                               _p(beta), _p(rw), _p(col, _pi32))
        return {"a": a, "b": b, "plane": pl, "basis": basis, "alpha": alpha, "beta": beta, "rworld": rw,
                "colour": col}

    def v2(self):
        out = np.empty(self.dim * self.n)
        lib().orc_get_v2(self._ctx, _p(out), out.size)
        return out

    def residual_history(self):
        k = lib().orc_get_hist(self._ctx, None, 0)
        h = np.empty(k)
        lib().orc_get_hist(self._ctx, _p(h), k)
        return h


def friction_solve(dim, m, inertia, a, b, basis, arma, armb, mu, ndotv, CoR, v0, r0, max_iters=5000, eval_every=1,
                   tol=1e-12, order=0):
    """Stand-alone FrictionSolver::solve restatement; returns (r_local, v_out, err, sweeps)."""
    m = _f64(m)
    C = len(a)
    r = _f64(r0).ravel().copy()
    v0 = _f64(v0).ravel()
    vout = np.empty_like(v0)
    sweeps = ctypes.c_uint32(0)
    a = np.ascontiguousarray(a, np.int32)
    b = np.ascontiguousarray(b, np.int32)
    err = lib().orc_friction_solve(dim, m.size, _p(m), _p(_f64(inertia)) if inertia is not None else None, C,
                                   _p(a, _pi32), _p(b, _pi32), _p(_f64(basis)),
                                   _p(_f64(arma)) if arma is not None else None,
                                   _p(_f64(armb)) if armb is not None else None, _p(_f64(mu)),
                                   _p(_f64(ndotv)) if ndotv is not None else None, CoR, _p(v0), _p(r), _p(vout),
                                   max_iters, eval_every, tol, order, ctypes.byref(sweeps))
    return r, vout, err, sweeps.value
