"""Seeded synthetic scenes for the benchmark workloads (wrapper of libscx_scenes.so)."""

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libscx_scenes.so")
_lib = None
_pd = ctypes.POINTER(ctypes.c_double)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            raise RuntimeError("libscx_scenes.so not built: run `python -m scisim_amd.build`")
        L = ctypes.CDLL(_LIB)
        L.scn_lattice2d.argtypes = [ctypes.c_uint32, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_uint64, _pd, _pd]
        L.scn_rsa2d.argtypes = [ctypes.c_uint32, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_uint64,
                                ctypes.c_uint64, _pd, _pd, _pd]
        L.scn_rsa3d.argtypes = [ctypes.c_uint32, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                ctypes.c_double, ctypes.c_double, ctypes.c_uint64, ctypes.c_uint64, _pd, _pd, _pd]
        L.scn_deposit3d.argtypes = [ctypes.c_uint32, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_uint64, ctypes.c_double,
                                    _pd, _pd, _pd]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_pd)


def box_planes_2d(half_x, half_y):
    """4 static planes bounding [-hx, hx] x [-hy, hy], normals pointing inwards."""
    x = np.array([[-half_x, 0.0], [half_x, 0.0], [0.0, -half_y], [0.0, half_y]])
    n = np.array([[1.0, 0.0], [-1.0, 0.0], [0.0, 1.0], [0.0, -1.0]])
    return x, n


def box_planes_3d(half_base, floor=0.0):
    """Floor plus 4 walls of a box [-b, b] x [floor, inf) x [-b, b], normals inwards."""
    x = np.array([[0.0, floor, 0.0], [-half_base, 0.0, 0.0], [half_base, 0.0, 0.0], [0.0, 0.0, -half_base],
                  [0.0, 0.0, half_base]])
    n = np.array([[0.0, 1.0, 0.0], [1.0, 0.0, 0.0], [-1.0, 0.0, 0.0], [0.0, 0.0, 1.0], [0.0, 0.0, -1.0]])
    return x, n


def ball2d_lattice(m=64, side=10.0, r=0.05, vstd=0.1, seed=1):
    """4096 equal circles (m=64): jittered lattice in [-5, 5]^2, m = 1, v ~ N(0, 0.1^2), 4 planes."""
    n = m * m
    q = np.empty(2 * n)
    v = np.empty(2 * n)
    rc = lib().scn_lattice2d(m, side, r, vstd, seed, _p(q), _p(v))
    if rc != 0:
        raise ValueError("lattice spacing too small for the radius")
    px, pn = box_planes_2d(side / 2, side / 2)
    return {"dim": 2, "q": q, "v": v, "r": np.full(n, r), "m": np.ones(n), "px": px, "pn": pn}


def ball2d_rsa(n=1048576, rmin=0.02, rmax=0.06, phi=0.55, seed=1, max_attempts=1000000):
    """Circles with U[rmin, rmax] radii, density 1 (m = pi r^2), random sequential packing to area fraction phi."""
    q = np.empty(2 * n)
    r = np.empty(n)
    side = ctypes.c_double(0.0)
    rc = lib().scn_rsa2d(n, rmin, rmax, phi, seed, max_attempts, _p(q), _p(r), ctypes.byref(side))
    if rc != 0:
        raise RuntimeError("RSA packing failed at circle %d" % (-rc - 1))
    L = side.value
    px, pn = box_planes_2d(L / 2, L / 2)
    return {"dim": 2, "q": q, "v": np.zeros(2 * n), "r": r, "m": np.pi * r * r, "px": px, "pn": pn, "side": L}


def spheres_rsa(n=262144, rmed=0.05, sigma=0.25, base=None, phi=0.3, seed=1, density=1.0, max_attempts=1000000):
    """Spheres with log-normal radii in a box (floor + 4 walls), random sequential packing to volume fraction phi.

    m = density 4/3 pi r^3, I = 2/5 m r^2 (rigidbody3d/Geometry/RigidBodySphere.cpp:10-18).
    """
    if base is None:
        # base so that the settled pile (volume fraction ~0.6) is roughly as tall as wide
        vol = n * 4.0 / 3.0 * np.pi * rmed ** 3 * np.exp(4.5 * sigma * sigma)
        base = (vol / 0.6) ** (1.0 / 3.0)
    q = np.empty(3 * n)
    r = np.empty(n)
    height = ctypes.c_double(0.0)
    rc = lib().scn_rsa3d(n, rmed, sigma, rmed * np.exp(-5 * sigma), rmed * np.exp(5 * sigma), base, phi, seed,
                         max_attempts, _p(q), _p(r), ctypes.byref(height))
    if rc != 0:
        raise RuntimeError("RSA packing failed at sphere %d" % (-rc - 1))
    m = density * (4.0 * np.pi * r * r * r / 3.0)
    px, pn = box_planes_3d(base / 2)
    return {"dim": 3, "q": q, "v": np.zeros(3 * n), "r": r, "m": m, "inertia": 2.0 * m * r * r / 5.0, "px": px,
            "pn": pn, "base": base, "height": height.value}


def spheres_pile(n=262144, rmed=0.05, sigma=0.25, base=None, seed=1, density=1.0, sink=1e-9, rmin=None, rmax=None):
    """A poured pile at rest: spheres dropped one by one onto the floor of a box (floor +
    4 walls) and rolled downhill into a pocket of their supports (about six contacts per
    sphere). Radii log-normal (median rmed, sigma), or uniform in [rmin, rmax] when both
    are given. Each sphere overlaps its supports by `sink` so that its contacts are active
    from the first step. v = 0."""
    uniform = rmin is not None and rmax is not None
    if base is None:
        # base so that the pile (volume fraction ~0.58) is roughly as tall as wide
        if uniform:
            r3 = (rmax ** 4 - rmin ** 4) / (4.0 * (rmax - rmin))  # E[r^3] of U[rmin, rmax]
        else:
            r3 = rmed ** 3 * np.exp(4.5 * sigma * sigma)
        vol = n * 4.0 / 3.0 * np.pi * r3
        base = (vol / 0.58) ** (1.0 / 3.0)
    q = np.empty(3 * n)
    r = np.empty(n)
    height = ctypes.c_double(0.0)
    lo, hi = (rmin, rmax) if uniform else (rmed * np.exp(-5 * sigma), rmed * np.exp(5 * sigma))
    rc = lib().scn_deposit3d(n, rmed, sigma, lo, hi, 1 if uniform else 0, base, seed, sink, _p(q), _p(r),
                             ctypes.byref(height))
    if rc != 0:
        raise RuntimeError("deposition failed (%d)" % rc)
    m = density * (4.0 * np.pi * r * r * r / 3.0)
    px, pn = box_planes_3d(base / 2)
    return {"dim": 3, "q": q, "v": np.zeros(3 * n), "r": r, "m": m, "inertia": 2.0 * m * r * r / 5.0, "px": px,
            "pn": pn, "base": base, "height": height.value}


def ensemble_piles(nscenes=512, nmin=500, nmax=20000, rmin=0.03, rmax=0.07, mu_lo=0.1, mu_hi=0.8, seed=1,
                   workers=16):
    """Independent poured piles for an ensemble: body counts log-uniform in [nmin, nmax],
    radii uniform in [rmin, rmax], friction coefficients uniform in [mu_lo, mu_hi], all
    seeded. Every pile sits in its own box at the same coordinates (the library keeps
    scenes apart). Returns the concatenated scene (q, v, r, m, inertia, px, pn), the body
    and plane offsets, mu per scene and the body counts."""
    from concurrent.futures import ThreadPoolExecutor
    rng = np.random.default_rng(seed)
    counts = np.exp(rng.uniform(np.log(nmin), np.log(nmax), nscenes)).round().astype(np.int64)
    mus = rng.uniform(mu_lo, mu_hi, nscenes)
    seeds = rng.integers(1, 2 ** 62, nscenes)
    with ThreadPoolExecutor(max_workers=workers) as ex:  # the generator releases the GIL
        piles = list(ex.map(lambda k: spheres_pile(n=int(counts[k]), rmin=rmin, rmax=rmax, seed=int(seeds[k])),
                            range(nscenes)))
    cat = {k: np.concatenate([p[k] for p in piles]) for k in ("q", "v", "r", "m", "inertia")}
    cat["px"] = np.concatenate([p["px"] for p in piles])
    cat["pn"] = np.concatenate([p["pn"] for p in piles])
    cat["dim"] = 3
    body_off = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint32)
    plane_off = np.arange(nscenes + 1, dtype=np.uint32) * 5
    return cat, body_off, plane_off, mus, counts
