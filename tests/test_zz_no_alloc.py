"""scx_potential_overlaps_noalloc, the broad phase on the library's file-scope scratch
array, called through the C ABI: two calls in a row (the second with fewer boxes, so it
must not see the first call's leftovers) give exactly the pairs of
scx_potential_overlaps on the same boxes, and inputs beyond its stated limits are refused
with SCX_E_CAPACITY."""

import ctypes

import numpy as np
import pytest

import scisim_amd as scx

pytestmark = pytest.mark.gpu


def _boxes(dim, n, seed, extent):
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.0, extent, (n, dim))
    h = rng.uniform(0.02, 0.08, (n, dim))
    return np.ascontiguousarray(c - h), np.ascontiguousarray(c + h)


def _noalloc(ctx, mn, mx, cap):
    n = mn.shape[0]
    out = np.full((max(cap, 1), 2), 0xFFFFFFFF, dtype=np.uint32)
    npairs = ctypes.c_uint64(0)
    rc = scx.lib().scx_potential_overlaps_noalloc(ctx.handle, n, mn.ctypes.data_as(scx._pd),
                                                  mx.ctypes.data_as(scx._pd), cap,
                                                  out.ctypes.data_as(scx._pu32), ctypes.byref(npairs))
    return rc, out[:npairs.value].copy() if rc == scx.SCX_OK else None, npairs.value


@pytest.mark.parametrize("dim", [2, 3])
def test_noalloc_overlaps_match_the_allocating_entry_point_on_two_calls(dim):
    ctx = scx.Context(dim, 64, max_planes=4)
    try:
        for k, (n, extent) in enumerate([(60000, 20.0 if dim == 3 else 60.0), (7000, 6.0 if dim == 3 else 18.0)]):
            mn, mx = _boxes(dim, n, seed=10 * dim + k, extent=extent)
            ref = ctx.potential_overlaps(mn, mx)
            rc, got, P = _noalloc(ctx, mn, mx, cap=len(ref) + 10)
            print("dim", dim, "call", k, "boxes", n, "pairs", P, "reference", len(ref), flush=True)
            assert rc == scx.SCX_OK, scx.lib().scx_last_error(ctx.handle)
            assert P == len(ref) > 0
            assert np.array_equal(got, ref)
    finally:
        ctx.close()


def test_noalloc_overlaps_refuse_inputs_beyond_the_limits():
    ctx = scx.Context(3, 64, max_planes=4)
    try:
        mn, mx = _boxes(3, 65537, seed=1, extent=30.0)
        rc, _, _ = _noalloc(ctx, mn, mx, cap=1 << 20)
        assert rc == scx.SCX_E_CAPACITY
        # too little room for the pairs: the count is reported, nothing is written
        mn, mx = _boxes(3, 5000, seed=2, extent=4.0)
        ref = ctx.potential_overlaps(mn, mx)
        rc, _, P = _noalloc(ctx, mn, mx, cap=len(ref) - 1)
        assert rc == scx.SCX_E_CAPACITY and P == len(ref)
        # the context's own broad phase is untouched by the no-allocation calls
        assert np.array_equal(ctx.potential_overlaps(mn, mx), ref)
    finally:
        ctx.close()
