"""CPU checks of the C ABI library: it loads, exports every symbol include/scx.h
declares, and rejects invalid configurations before touching a device."""

import ctypes
import os
import re

import numpy as np
import pytest

import scisim_amd as scx

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "scx.h")


def declared_symbols():
    src = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const char\*|int|void)\s+(scx_\w+)\s*\(", src, re.M)))


def test_header_and_python_binding_agree():
    assert declared_symbols() == sorted(scx.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = scx.lib()
    for name in declared_symbols():
        assert hasattr(lib, name), name
    assert lib.scx_version().startswith(b"scx")


def test_invalid_config_is_rejected_on_the_host():
    lib = scx.lib()
    h = ctypes.c_void_p()
    for cfg in (scx.ScxConfig(dim=4, device=0, max_bodies=10, max_planes=4),
                scx.ScxConfig(dim=2, device=0, max_bodies=0, max_planes=4),
                scx.ScxConfig(dim=3, device=0, max_bodies=10, max_planes=(1 << 24) + 1)):
        assert lib.scx_create(ctypes.byref(cfg), ctypes.byref(h)) == scx.SCX_E_ARG
        assert not h
    assert lib.scx_create(None, ctypes.byref(h)) == scx.SCX_E_ARG


def test_null_context_is_rejected():
    lib = scx.lib()
    assert lib.scx_step(None, None, None) == scx.SCX_E_ARG
    assert lib.scx_set_timing(None, 1) == scx.SCX_E_ARG
    assert lib.scx_last_error(None) == b"null context"


def test_scene_generators_are_deterministic_and_non_overlapping():
    from scisim_amd import scenes
    a = scenes.ball2d_rsa(n=3000, seed=9)
    b = scenes.ball2d_rsa(n=3000, seed=9)
    assert np.array_equal(a["q"], b["q"]) and np.array_equal(a["r"], b["r"])
    q = a["q"].reshape(-1, 2)
    d = np.sqrt(((q[:, None, :] - q[None, :, :]) ** 2).sum(-1))
    rs = a["r"][:, None] + a["r"][None, :]
    np.fill_diagonal(d, np.inf)
    assert np.all(d >= rs)
    assert abs(np.pi * (a["r"] ** 2).sum() / a["side"] ** 2 - 0.55) < 1e-9
    s = scenes.spheres_rsa(n=2000, seed=3)
    q = s["q"].reshape(-1, 3)
    d = np.sqrt(((q[:, None, :] - q[None, :, :]) ** 2).sum(-1))
    np.fill_diagonal(d, np.inf)
    assert np.all(d >= s["r"][:, None] + s["r"][None, :])
    assert np.all(q[:, 1] - s["r"] >= 0.0)


def test_poured_pile_rests_on_its_supports():
    from scisim_amd import scenes
    sink = 1e-9
    a = scenes.spheres_pile(n=1500, seed=4, sink=sink)
    b = scenes.spheres_pile(n=1500, seed=4, sink=sink)
    assert np.array_equal(a["q"], b["q"]) and np.array_equal(a["r"], b["r"])
    q, r = a["q"].reshape(-1, 3), a["r"]
    d = np.sqrt(((q[:, None, :] - q[None, :, :]) ** 2).sum(-1))
    np.fill_diagonal(d, np.inf)
    gap = d - (r[:, None] + r[None, :])
    # overlaps only by the requested sink (plus rounding)
    assert gap.min() >= -sink * (1 + 1e-3) - 1e-12
    # inside the walls, above the floor (up to the sink)
    assert np.all(np.abs(q[:, [0, 2]]) <= a["base"] / 2 - r[:, None] + 1e-12)
    assert np.all(q[:, 1] - r >= -sink * (1 + 1e-3))
    # spheres rest on the floor or on a sphere below their centre (the rolling of a few
    # is cut short by its trial budget; the first simulated steps settle those)
    touching = gap <= 1e-7
    below = q[None, :, 1] < q[:, None, 1]
    supported = (q[:, 1] - r <= 1e-7) | np.any(touching & below, axis=1)
    assert supported.mean() > 0.99
    # a pile, not a gas: about four contacts per sphere or more before settling
    contacts = (touching.sum() + (q[:, 1] - r <= 1e-7).sum()) / len(r)
    assert contacts > 3.5


def test_ensemble_generator_offsets_and_ranges():
    """ensemble_piles: seeded, offsets consistent with the concatenated arrays, body counts,
    radii and friction coefficients in their ranges, five planes per scene, unit-density
    masses and sphere inertias, and the same result with one worker thread or several."""
    from scisim_amd import scenes
    kw = dict(nscenes=6, nmin=100, nmax=400, rmin=0.03, rmax=0.07, mu_lo=0.1, mu_hi=0.8, seed=5, workers=3)
    s, bo, po, mus, counts = scenes.ensemble_piles(**kw)
    s2, bo2, _, mus2, _ = scenes.ensemble_piles(**dict(kw, workers=1))
    assert np.array_equal(s["q"], s2["q"]) and np.array_equal(bo, bo2) and np.array_equal(mus, mus2)
    assert bo[0] == 0 and np.array_equal(np.diff(bo.astype(np.int64)), counts)
    assert np.all((counts >= 100) & (counts <= 400))
    n = int(bo[-1])
    assert s["q"].size == 3 * n and s["r"].size == n and s["m"].size == n and s["inertia"].size == n
    assert np.all((s["r"] >= 0.03) & (s["r"] <= 0.07))
    assert np.all((mus >= 0.1) & (mus <= 0.8))
    assert np.array_equal(po, np.arange(7, dtype=np.uint32) * 5)
    assert np.asarray(s["px"]).reshape(-1, 3).shape[0] == 30
    # masses and inertias of unit-density spheres
    assert np.allclose(s["m"], 4.0 / 3.0 * np.pi * s["r"] ** 3, rtol=1e-12)
    assert np.allclose(s["inertia"], 0.4 * s["m"] * s["r"] ** 2, rtol=1e-12)


@pytest.mark.parametrize("lang,header", [("c", "scx.h"), ("c++", "scx.h"), ("c++", "scx_scisim.hpp")])
def test_public_headers_compile_on_their_own(lang, header):
    """The C ABI header is valid C99 and C++17 by itself, and the reference-shaped C++
    layer compiles with nothing but its own includes (no HIP or torch headers needed)."""
    import shutil
    import subprocess
    cc = "gcc" if lang == "c" else "g++"
    if shutil.which(cc) is None:
        pytest.skip(cc + " not found")
    std = "-std=c99" if lang == "c" else "-std=c++17"
    src = '#include "%s"\n' % header
    p = subprocess.run([cc, std, "-fsyntax-only", "-Wall", "-Wextra", "-Werror", "-I",
                        os.path.join(ROOT, "include"), "-x", lang, "-"], input=src, text=True, capture_output=True)
    assert p.returncode == 0, p.stderr
