"""CPU checks of the checkpoint image validation (scx_checkpoint_info, host only): a
well-formed image is accepted and described, and every malformed one (bad magic,
version, dimension, counts, truncation, a flipped payload byte) is refused before
anything would reach a device. The images are built here from the layout stated in
include/scx.h and scx_api.hip (header, then the arrays in order)."""

import struct

import numpy as np
import pytest

import scisim_amd as scx

MLP = scx.MLP_PARAMS


def fnv1a(b):
    h = 1469598103934665603
    for x in b:
        h ^= x
        h = (h * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def image(dim=3, n=5, nplanes=2, nscenes=0, scene_planes=0, cache=3, layout=1, model=False, seed=0,
          payload_extra=b"", version=1, magic=b"SCXCKPT1"):
    rng = np.random.default_rng(seed)
    parts = [rng.normal(size=dim * n), rng.normal(size=dim * n), rng.normal(size=3 * n), rng.uniform(0.1, 1, n),
             rng.uniform(0.1, 1, n), rng.uniform(0.1, 1, n), rng.normal(size=dim * nplanes),
             rng.normal(size=dim * nplanes)]
    if nscenes:
        parts += [np.zeros(n, np.uint32), np.arange(nscenes + 1, dtype=np.uint32), np.full(nscenes, 0.4),
                  np.zeros(3 * nscenes)]
    if cache:
        parts += [np.arange(cache, dtype=np.uint64), rng.normal(size=6 * cache)]
    if model:
        parts += [rng.normal(size=MLP).astype(np.float32), np.zeros(MLP, np.float32), np.zeros(MLP, np.float32)]
    payload = b"".join(np.ascontiguousarray(p).tobytes() for p in parts) + payload_extra
    hdr = struct.pack("<8sIiIIIIIiIIQQ", magic, version, dim, n, nplanes, nscenes, scene_planes, cache, layout,
                      int(model), 0, len(payload), fnv1a(payload))
    assert len(hdr) == 64
    return hdr + payload


@pytest.mark.parametrize("kw", [dict(), dict(dim=2, cache=0), dict(nscenes=2, scene_planes=1, model=True),
                                dict(nplanes=0, cache=0, layout=0)])
def test_well_formed_images_are_described(kw):
    img = image(**kw)
    info = scx.checkpoint_info(img)
    assert info["dim"] == kw.get("dim", 3)
    assert info["bodies"] == 5
    assert info["planes"] == kw.get("nplanes", 2)
    assert info["scenes"] == kw.get("nscenes", 0)
    assert info["cache_entries"] == kw.get("cache", 3)
    assert info["has_model"] == kw.get("model", False)


def test_malformed_images_are_refused():
    good = image()
    bad = [good[:63], good[:-1], good + b"\0", b"XCXCKPT1" + good[8:], image(version=2), image(dim=4),
           image(n=0, cache=0), image(layout=2), image(payload_extra=b"\0" * 8)]
    flipped = bytearray(good)
    flipped[100] ^= 0x10
    bad.append(bytes(flipped))
    for img in bad:
        with pytest.raises(scx.ScxError):
            scx.checkpoint_info(img)
