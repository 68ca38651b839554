"""MI355X (gfx950) contact pipeline for circle/sphere scenes, Python view of the C ABI.

The compute path is libscx.so (hand-written HIP kernels, built in-tree by
scisim_amd.build). This module only binds the C ABI declared in include/scx.h with
ctypes; it has no compute fallback: if the library is missing or a call fails, it
raises.
"""

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libscx.so")

SCX_OK = 0
SCX_E_ARG = -1
SCX_E_HIP = -2
SCX_E_CAPACITY = -3
SCX_E_STATE = -4
SCX_E_COLOURS = -5

SYMPLECTIC_EULER = 0
VERLET = 1
FP64 = 0
FP32_POLISH = 1
WARM_NONE = 0
WARM_CACHE = 1
WARM_NORMAL = 2
WARM_TRANSPORT = 3
STAGGER_GEOMETRIC = 0
STAGGER_SYMPLECTIC_EULER = 1
STAGGER_STABILIZED = 2

# Every symbol include/scx.h declares.
EXPORTS = ["scx_version", "scx_last_error", "scx_create", "scx_destroy", "scx_potential_overlaps", "scx_set_bodies",
           "scx_set_planes", "scx_get_state", "scx_step", "scx_run", "scx_get_contacts", "scx_get_v2",
           "scx_friction_solve", "scx_unconstrained_flow", "scx_set_timing", "scx_get_timing", "scx_get_sweep_counters",
           "scx_set_warm_model", "scx_get_warm_model", "scx_get_samples", "scx_warm_model_gradients",
           "scx_train_warm_model", "scx_set_scenes", "scx_get_scene_stats", "scx_set_autotune", "scx_get_autotune",
           "scx_save_checkpoint", "scx_load_checkpoint", "scx_checkpoint_info", "scx_get_contact_forces", "scx_potential_overlaps_noalloc"]

MLP_PARAMS = 19075  # SCX_MLP_PARAMS
WARM_LEARNED = 4

_d = ctypes.c_double
_pd = ctypes.POINTER(ctypes.c_double)
_pu32 = ctypes.POINTER(ctypes.c_uint32)
_pi32 = ctypes.POINTER(ctypes.c_int32)
_pf = ctypes.POINTER(ctypes.c_float)
_vp = ctypes.c_void_p


class ScxConfig(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int32), ("device", ctypes.c_int32), ("max_bodies", ctypes.c_uint32),
                ("max_planes", ctypes.c_uint32), ("max_contacts", ctypes.c_uint32), ("max_pairs", ctypes.c_uint32)]


class ScxStepParams(ctypes.Structure):
    _fields_ = [("dt", _d), ("g", _d * 3), ("mu", _d), ("CoR", _d), ("tol", _d), ("max_iters", ctypes.c_uint32),
                ("eval_every", ctypes.c_uint32), ("integrator", ctypes.c_int32), ("precision", ctypes.c_int32),
                ("warm_start", ctypes.c_int32), ("iteration", ctypes.c_uint32), ("staggering", ctypes.c_int32),
                ("stabilize", ctypes.c_int32), ("penetration_threshold", _d)]


class ScxStepStats(ctypes.Structure):
    _fields_ = [("nentries", ctypes.c_uint64), ("npairs", ctypes.c_uint32), ("nplane_contacts", ctypes.c_uint32),
                ("ncontacts", ctypes.c_uint32), ("ncolours", ctypes.c_uint32), ("sweeps", ctypes.c_uint32),
                ("sweeps_fp32", ctypes.c_uint32), ("error", _d), ("succeeded", ctypes.c_int32),
                ("ncells", ctypes.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ScxSweepCounters(ctypes.Structure):
    _fields_ = [("full_evals", ctypes.c_uint32), ("full_evals_fp32", ctypes.c_uint32),
                ("skipped_evals", ctypes.c_uint32), ("component_launches", ctypes.c_uint32),
                ("component_overflows", ctypes.c_uint32), ("component_width", ctypes.c_uint32),
                ("component_share", _d)]


class ScxContactSet(ctypes.Structure):
    _fields_ = [("ncontacts", ctypes.c_uint32), ("body0", _pi32), ("body1", _pi32), ("bases", _pd), ("arm0", _pd),
                ("arm1", _pd), ("mu", _pd), ("CoR", _pd), ("nrel_extra", _pd), ("drel_extra", _pd)]


_lib = None


class ScxError(RuntimeError):
    pass


def lib():
    """Load libscx.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ScxError("libscx.so not found at %s: run `python -m scisim_amd.build` first" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        L.scx_version.restype = ctypes.c_char_p
        L.scx_last_error.restype = ctypes.c_char_p
        L.scx_last_error.argtypes = [_vp]
        L.scx_create.argtypes = [ctypes.POINTER(ScxConfig), ctypes.POINTER(_vp)]
        L.scx_destroy.argtypes = [_vp]
        L.scx_potential_overlaps.argtypes = [_vp, ctypes.c_uint32, _pd, _pd, ctypes.c_uint64, _pu32,
                                             ctypes.POINTER(ctypes.c_uint64)]
        L.scx_potential_overlaps_noalloc.argtypes = [_vp, ctypes.c_uint32, _pd, _pd, ctypes.c_uint64, _pu32,
                                                     ctypes.POINTER(ctypes.c_uint64)]
        L.scx_set_bodies.argtypes = [_vp, ctypes.c_uint32, _pd, _pd, _pd, _pd, _pd, _pd]
        L.scx_set_planes.argtypes = [_vp, ctypes.c_uint32, _pd, _pd]
        L.scx_set_scenes.argtypes = [_vp, ctypes.c_uint32, _pu32, _pu32, _pd]
        L.scx_get_scene_stats.argtypes = [_vp, ctypes.c_uint32, _pu32, _pd, _pu32]
        L.scx_get_state.argtypes = [_vp, _pd, _pd, _pd]
        L.scx_step.argtypes = [_vp, ctypes.POINTER(ScxStepParams), ctypes.POINTER(ScxStepStats)]
        L.scx_run.argtypes = [_vp, ctypes.POINTER(ScxStepParams), ctypes.c_uint32, ctypes.POINTER(ScxStepStats)]
        L.scx_get_contacts.argtypes = [_vp, ctypes.c_uint32, _pi32, _pi32, _pi32, _pd, _pd, _pd, _pi32, _pu32]
        L.scx_get_v2.argtypes = [_vp, _pd, _pd]
        L.scx_get_contact_forces.argtypes = [_vp, ctypes.c_uint32, _pi32, _pd, _pd, _pd, _pu32]
        L.scx_friction_solve.argtypes = [_vp, ctypes.c_uint32, _pd, _pd, _pd, _pd, ctypes.POINTER(ScxContactSet), _d,
                                         ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int32, _pd, _pd, _pd, _pd, _pd,
                                         _pi32, _pd, _pu32]
        L.scx_unconstrained_flow.argtypes = [_vp, ctypes.c_int32, ctypes.c_uint32, _pd, _pd, _d, _pd, _pd, _pd, _pd]
        L.scx_set_timing.argtypes = [_vp, ctypes.c_int32]
        L.scx_get_timing.argtypes = [_vp, ctypes.c_uint32, ctypes.POINTER(ctypes.c_char_p), _pd, _pu32]
        L.scx_get_sweep_counters.argtypes = [_vp, ctypes.POINTER(ScxSweepCounters)]
        L.scx_set_autotune.argtypes = [_vp, ctypes.c_int32]
        L.scx_get_autotune.argtypes = [_vp, ctypes.c_uint32, _pu32, _pi32, _pi32, _pf, _pi32, _pu32]
        L.scx_set_warm_model.argtypes = [_vp, _pf, ctypes.c_uint32]
        L.scx_get_warm_model.argtypes = [_vp, _pf, ctypes.c_uint32]
        L.scx_get_samples.argtypes = [_vp, ctypes.c_uint32, _pf, _pf, _pu32]
        L.scx_warm_model_gradients.argtypes = [_vp, ctypes.c_uint32, _pf, _pf, _pf, _pd]
        L.scx_train_warm_model.argtypes = [_vp, ctypes.c_uint32, _pf, _pf, ctypes.c_uint32, ctypes.c_uint32,
                                           ctypes.c_float, ctypes.c_uint64, _pd]
        L.scx_save_checkpoint.argtypes = [_vp, ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]
        L.scx_load_checkpoint.argtypes = [_vp, ctypes.c_void_p, ctypes.c_uint64]
        L.scx_checkpoint_info.argtypes = [ctypes.c_void_p, ctypes.c_uint64, _pi32, _pu32, _pu32, _pu32, _pu32,
                                          _pi32]
        _lib = L
    return _lib


def checkpoint_info(image):
    """Validates a checkpoint image (bytes) on the host and returns its contents as a dict;
    raises ScxError if it is malformed or corrupt."""
    buf = ctypes.create_string_buffer(bytes(image), len(image))
    dim, has_model = ctypes.c_int32(0), ctypes.c_int32(0)
    n, npl, nsc, ncache = ctypes.c_uint32(0), ctypes.c_uint32(0), ctypes.c_uint32(0), ctypes.c_uint32(0)
    rc = lib().scx_checkpoint_info(buf, len(image), ctypes.byref(dim), ctypes.byref(n), ctypes.byref(npl),
                                   ctypes.byref(nsc), ctypes.byref(ncache), ctypes.byref(has_model))
    if rc != SCX_OK:
        raise ScxError("scx_checkpoint_info failed (%d)" % rc)
    return {"dim": dim.value, "bodies": n.value, "planes": npl.value, "scenes": nsc.value,
            "cache_entries": ncache.value, "has_model": bool(has_model.value)}


def _p(a, t=_pd):
    return None if a is None else a.ctypes.data_as(t)


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64).ravel()


def step_params(dt, g, mu, CoR=0.0, tol=1e-6, max_iters=5000, eval_every=1, integrator=SYMPLECTIC_EULER,
                precision=FP64, warm_start=WARM_NONE, iteration=1, staggering=None, stabilize=0,
                penetration_threshold=0.0):
    gg = list(g) + [0.0] * (3 - len(g))
    return ScxStepParams(dt=dt, g=(_d * 3)(*gg), mu=mu, CoR=CoR, tol=tol, max_iters=max_iters,
                         eval_every=eval_every, integrator=integrator, precision=precision, warm_start=warm_start,
                         iteration=iteration, staggering=STAGGER_GEOMETRIC if staggering is None else staggering,
                         stabilize=stabilize, penetration_threshold=penetration_threshold)


class Context:
    """One device context (scx_ctx) for 2D circles or 3D spheres."""

    def __init__(self, dim, max_bodies, max_planes=8, device=0, max_contacts=0, max_pairs=0):
        self.dim = dim
        self._h = _vp()
        cfg = ScxConfig(dim=dim, device=device, max_bodies=max_bodies, max_planes=max_planes,
                        max_contacts=max_contacts, max_pairs=max_pairs)
        rc = lib().scx_create(ctypes.byref(cfg), ctypes.byref(self._h))
        if rc != SCX_OK:
            msg = lib().scx_last_error(self._h) if self._h else b"invalid configuration"
            if self._h:
                lib().scx_destroy(self._h)
                self._h = _vp()
            raise ScxError("scx_create failed (%d): %s" % (rc, msg.decode()))
        self.n = 0
        self.nscenes = 0

    def close(self):
        if self._h:
            lib().scx_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != SCX_OK:
            raise ScxError("%s failed (%d): %s" % (what, rc, lib().scx_last_error(self._h).decode()))

    @property
    def handle(self):
        return self._h

    def potential_overlaps(self, aabb_min, aabb_max):
        mn, mx = _f64(aabb_min), _f64(aabb_max)
        n = mn.size // self.dim
        cap = max(16, 8 * n)
        np_ = ctypes.c_uint64(0)
        while True:
            out = np.empty((cap, 2), dtype=np.uint32)
            rc = lib().scx_potential_overlaps(self._h, n, _p(mn), _p(mx), cap, _p(out, _pu32), ctypes.byref(np_))
            if rc == SCX_E_CAPACITY and np_.value > cap:
                cap = int(np_.value)
                continue
            self._check(rc, "scx_potential_overlaps")
            return out[:np_.value].copy()

    def set_bodies(self, q, v, r, m, omega=None, inertia=None):
        q, v, r, m = _f64(q), _f64(v), _f64(r), _f64(m)
        if r.size != self.n:
            self.nscenes = 0
        self.n = r.size
        self._check(lib().scx_set_bodies(self._h, self.n, _p(q), _p(v), _p(_f64(omega)), _p(r), _p(m),
                                         _p(_f64(inertia))), "scx_set_bodies")

    def set_planes(self, x, n):
        x, n = _f64(x), _f64(n)
        if x.size // self.dim != getattr(self, "nplanes", 0):
            self.nscenes = 0
        self.nplanes = x.size // self.dim
        self._check(lib().scx_set_planes(self._h, self.nplanes, _p(x), _p(n)), "scx_set_planes")

    def set_scenes(self, body_off, plane_off, mu):
        """Ensemble of independent scenes: scene k owns bodies body_off[k]:body_off[k+1] and
        planes plane_off[k]:plane_off[k+1], with friction coefficient mu[k]; an empty
        body_off returns to one scene."""
        bo = np.ascontiguousarray(body_off, dtype=np.uint32)
        po = np.ascontiguousarray(plane_off, dtype=np.uint32)
        mu = _f64(mu)
        ns = max(0, bo.size - 1)
        self.nscenes = ns
        self._check(lib().scx_set_scenes(self._h, ns, _p(bo, _pu32), _p(po, _pu32), _p(mu)), "scx_set_scenes")

    def scene_stats(self):
        """(contacts, residual, sweeps to the tolerance) per scene of the last step of an
        ensemble."""
        ns = self.nscenes
        cnt = np.zeros(ns, dtype=np.uint32)
        err = np.zeros(ns)
        sw = np.zeros(ns, dtype=np.uint32)
        self._check(lib().scx_get_scene_stats(self._h, ns, _p(cnt, _pu32), _p(err), _p(sw, _pu32)),
                    "scx_get_scene_stats")
        return cnt, err, sw

    def get_state(self):
        q = np.empty(self.dim * self.n)
        v = np.empty(self.dim * self.n)
        w = np.empty(3 * self.n) if self.dim == 3 else None
        self._check(lib().scx_get_state(self._h, _p(q), _p(v), _p(w)), "scx_get_state")
        return q, v, w

    def step(self, params):
        st = ScxStepStats()
        self._check(lib().scx_step(self._h, ctypes.byref(params), ctypes.byref(st)), "scx_step")
        return st.as_dict()

    def run(self, params, nsteps):
        st = ScxStepStats()
        self._check(lib().scx_run(self._h, ctypes.byref(params), nsteps, ctypes.byref(st)), "scx_run")
        return st.as_dict()

    def contacts(self):
        D = self.dim
        nc = ctypes.c_uint32(0)
        rc = lib().scx_get_contacts(self._h, 0, None, None, None, None, None, None, None, ctypes.byref(nc))
        C = nc.value
        if rc != SCX_OK and rc != SCX_E_CAPACITY:
            self._check(rc, "scx_get_contacts")
        a = np.empty(C, np.int32)
        b = np.empty(C, np.int32)
        pl = np.empty(C, np.int32)
        basis = np.empty((C, D, D))
        alpha = np.empty(C)
        beta = np.empty((C, D - 1))
        col = np.empty(C, np.int32)
        self._check(lib().scx_get_contacts(self._h, C, _p(a, _pi32), _p(b, _pi32), _p(pl, _pi32), _p(basis),
                                           _p(alpha), _p(beta), _p(col, _pi32), ctypes.byref(nc)),
                    "scx_get_contacts")
        return {"a": a, "b": b, "plane": pl, "basis": basis, "alpha": alpha, "beta": beta, "colour": col}

    def contact_forces(self):
        """The last step's contacts as the reference's force output datasets
        (scx_get_contact_forces), shaped as the reference writes them (one column per
        contact): collision_count, collision_indices (2, C; a static plane as -plane - 2),
        collision_points, collision_normals, collision_forces (dim, C)."""
        C = self.contacts()["a"].size
        D = self.dim
        idx = np.empty((C, 2), np.int32)
        pts, nrm, frc = np.empty((C, D)), np.empty((C, D)), np.empty((C, D))
        cnt = ctypes.c_uint32(0)
        self._check(lib().scx_get_contact_forces(self._h, C, _p(idx, _pi32), _p(pts), _p(nrm), _p(frc),
                                                 ctypes.byref(cnt)), "scx_get_contact_forces")
        return {"collision_count": int(cnt.value), "collision_indices": np.ascontiguousarray(idx.T),
                "collision_points": np.ascontiguousarray(pts.T), "collision_normals": np.ascontiguousarray(nrm.T),
                "collision_forces": np.ascontiguousarray(frc.T)}

    def v2(self):
        v2 = np.empty(self.dim * self.n)
        w2 = np.empty(3 * self.n) if self.dim == 3 else None
        self._check(lib().scx_get_v2(self._h, _p(v2), _p(w2)), "scx_get_v2")
        return v2, w2

    def friction_solve(self, m, v0, body0, body1, bases, mu, inertia=None, omega0=None, arm0=None, arm1=None,
                       CoR=None, alpha=None, beta=None, tol=1e-12, max_iters=5000, eval_every=1, precision=FP64):
        """FrictionSolver::solve on an explicit active set; returns a dict of outputs."""
        D = self.dim
        m, v0 = _f64(m), _f64(v0)
        n = m.size
        C = len(body0)
        b0 = np.ascontiguousarray(body0, np.int32)
        b1 = np.ascontiguousarray(body1, np.int32)
        bases, mu = _f64(bases), _f64(mu)
        arm0, arm1, CoR = _f64(arm0), _f64(arm1), _f64(CoR)
        cs = ScxContactSet(ncontacts=C, body0=_p(b0, _pi32), body1=_p(b1, _pi32), bases=_p(bases), arm0=_p(arm0),
                           arm1=_p(arm1), mu=_p(mu), CoR=_p(CoR), nrel_extra=None, drel_extra=None)
        al = np.zeros(C) if alpha is None else _f64(alpha).copy()
        be = np.zeros((D - 1) * C) if beta is None else _f64(beta).copy()
        v2 = np.empty(D * n)
        w2 = np.empty(3 * n) if D == 3 else None
        f = np.empty((2 if D == 2 else 6) * n)
        ok = ctypes.c_int32(0)
        err = ctypes.c_double(0.0)
        sw = ctypes.c_uint32(0)
        self._check(lib().scx_friction_solve(self._h, n, _p(m), _p(_f64(inertia)), _p(v0), _p(_f64(omega0)),
                                             ctypes.byref(cs), tol, max_iters, eval_every, precision, _p(al), _p(be),
                                             _p(v2), _p(w2), _p(f), ctypes.byref(ok), ctypes.byref(err),
                                             ctypes.byref(sw)), "scx_friction_solve")
        return {"alpha": al, "beta": be.reshape(C, D - 1), "v2": v2, "omega2": w2, "f": f,
                "succeeded": bool(ok.value), "error": err.value, "sweeps": sw.value}

    def unconstrained_flow(self, integrator, m, g, dt, q0, v0):
        m, q0, v0 = _f64(m), _f64(q0), _f64(v0)
        gg = _f64(list(g) + [0.0] * (3 - len(g)))
        q1, v1 = np.empty_like(q0), np.empty_like(v0)
        self._check(lib().scx_unconstrained_flow(self._h, integrator, m.size, _p(m), _p(gg), dt, _p(q0), _p(v0),
                                                 _p(q1), _p(v1)), "scx_unconstrained_flow")
        return q1, v1

    def set_timing(self, on):
        self._check(lib().scx_set_timing(self._h, 1 if on else 0), "scx_set_timing")

    def timing(self):
        cnt = ctypes.c_uint32(0)
        names = (ctypes.c_char_p * 32)()
        ms = np.zeros(32)
        self._check(lib().scx_get_timing(self._h, 32, names, _p(ms), ctypes.byref(cnt)), "scx_get_timing")
        return [(names[i].decode(), float(ms[i])) for i in range(min(cnt.value, 32))]

    def sweep_counters(self):
        """Counters of the last step's sweeps (scx_sweep_counters): full residual evaluations
        (all phases and the FP32 phase's), FP32 evaluations replaced by the check of the
        listed contacts, sweep launches that ran the component phase or overflowed it, the
        widest colour of a workgroup's component list, and the share of the contacts the next
        step's component phase may hold."""
        x = ScxSweepCounters()
        self._check(lib().scx_get_sweep_counters(self._h, ctypes.byref(x)), "scx_get_sweep_counters")
        return {"full": x.full_evals, "full_fp32": x.full_evals_fp32, "skipped": x.skipped_evals,
                "component_launches": x.component_launches, "component_overflows": x.component_overflows,
                "component_width": x.component_width, "component_share": x.component_share}

    def set_autotune(self, on):
        """Online autotuning of the sweep launch width (results do not depend on it)."""
        self._check(lib().scx_set_autotune(self._h, 1 if on else 0), "scx_set_autotune")

    def autotune_table(self):
        """The tuner's classes: list of dicts with key, exponent, blocks, us_per_sweep, settled."""
        cnt = ctypes.c_uint32(0)
        self._check(lib().scx_get_autotune(self._h, 0, None, None, None, None, None, ctypes.byref(cnt)),
                    "scx_get_autotune")
        k = cnt.value
        keys = np.zeros(max(k, 1), dtype=np.uint32)
        ex, bl, st = (np.zeros(max(k, 1), dtype=np.int32) for _ in range(3))
        us = np.zeros(max(k, 1), dtype=np.float32)
        self._check(lib().scx_get_autotune(self._h, k, _p(keys, _pu32), _p(ex, _pi32), _p(bl, _pi32), _p(us, _pf),
                                           _p(st, _pi32), ctypes.byref(cnt)), "scx_get_autotune")
        return [{"key": int(keys[i]), "contacts_half_octave": int(keys[i] & 0xFF), "dim": 3 if keys[i] >> 8 & 1 else 2,
                 "phase": "fp64" if keys[i] >> 9 & 1 else "fp32", "ensemble": bool(keys[i] >> 10 & 1),
                 "exponent": int(ex[i]), "blocks": int(bl[i]), "us_per_sweep": float(us[i]),
                 "settled": bool(st[i])} for i in range(k)]

    # learned warm start
    def save_checkpoint(self):
        """The context's state as a checkpoint image (bytes), scx_save_checkpoint."""
        need = ctypes.c_uint64(0)
        self._check(lib().scx_save_checkpoint(self._h, None, 0, ctypes.byref(need)), "scx_save_checkpoint")
        buf = ctypes.create_string_buffer(need.value)
        self._check(lib().scx_save_checkpoint(self._h, buf, need.value, ctypes.byref(need)), "scx_save_checkpoint")
        return buf.raw[:need.value]

    def load_checkpoint(self, image):
        """Restores a checkpoint image written by save_checkpoint (scx_load_checkpoint)."""
        info = checkpoint_info(image)
        buf = ctypes.create_string_buffer(bytes(image), len(image))
        self._check(lib().scx_load_checkpoint(self._h, buf, len(image)), "scx_load_checkpoint")
        self.n = info["bodies"]
        self.nscenes = info["scenes"]

    def set_warm_model(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32).ravel()
        self._check(lib().scx_set_warm_model(self._h, p.ctypes.data_as(_pf), p.size), "scx_set_warm_model")

    def warm_model(self):
        p = np.empty(MLP_PARAMS, np.float32)
        self._check(lib().scx_get_warm_model(self._h, p.ctypes.data_as(_pf), p.size), "scx_get_warm_model")
        return p

    def samples(self):
        """(features [C][16], targets [C][3]) of the last step."""
        cnt = ctypes.c_uint32(0)
        lib().scx_get_samples(self._h, 0, None, None, ctypes.byref(cnt))
        C = cnt.value
        X = np.empty((C, 16), np.float32)
        T = np.empty((C, 3), np.float32)
        self._check(lib().scx_get_samples(self._h, C, X.ctypes.data_as(_pf), T.ctypes.data_as(_pf),
                                          ctypes.byref(cnt)), "scx_get_samples")
        return X, T

    def warm_model_gradients(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        g = np.empty(MLP_PARAMS, np.float32)
        loss = ctypes.c_double(0.0)
        self._check(lib().scx_warm_model_gradients(self._h, X.shape[0], X.ctypes.data_as(_pf), T.ctypes.data_as(_pf),
                                                   g.ctypes.data_as(_pf), ctypes.byref(loss)),
                    "scx_warm_model_gradients")
        return g, loss.value

    def train_warm_model(self, X, T, epochs, batch, lr, seed=1):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        loss = ctypes.c_double(0.0)
        self._check(lib().scx_train_warm_model(self._h, X.shape[0], X.ctypes.data_as(_pf), T.ctypes.data_as(_pf),
                                               epochs, batch, lr, seed, ctypes.byref(loss)), "scx_train_warm_model")
        return loss.value
