"""Table image headers on the CPU: sizes, the documented header, and the
checks every load makes before touching the device (magic, version,
geometry, size, checksum)."""
import ctypes as C
import struct

import numpy as np
import pytest

import gpuquotientfilters_amd as G
import oracle as O
from image_util import fnv1a64, make_image


def test_image_sizes():
    L = G.lib()
    assert L.gqf_image_bytes(G.IMAGE_SQF_GPU, 20, 5) == 64 + O.sqf_table_bytes(20, 5)
    assert L.gqf_image_bytes(G.IMAGE_SQF_CPU, 20, 5) == 64 + 4 * O.lib().orc_sqf_cpu_words(20, 5)
    assert L.gqf_image_bytes(G.IMAGE_RSQF, 12, 5) == 64 + L.gqf_rsqf_ref_bytes(12, 5)
    assert L.gqf_image_bytes(9, 12, 5) == 0


def test_header_fields_and_checksum():
    assert fnv1a64(b"") == 0xCBF29CE484222325 and fnv1a64(b"a") == 0xAF63DC4C8601EC8C  # FNV-1a 64 test vectors
    q, r = 6, 5
    payload = np.arange(O.sqf_table_bytes(q, r), dtype=np.uint8).tobytes()
    img = make_image(G.IMAGE_SQF_GPU, q, r, payload, hash_kind=1)
    assert G.image_info(img) == (G.IMAGE_SQF_GPU, q, r, 1)


def _status(img):
    b = np.frombuffer(img, dtype=np.uint8)
    return G.lib().gqf_image_info(b.ctypes.data, b.size, None, None, None, None)


def test_bad_images_are_refused():
    q, r = 6, 5
    payload = bytes(O.sqf_table_bytes(q, r))
    good = make_image(G.IMAGE_SQF_GPU, q, r, payload)
    assert _status(good) == G.OK
    assert _status(b"XQFTABLE" + good[8:]) == G.ERR_INVALID                      # magic
    assert _status(make_image(G.IMAGE_SQF_GPU, q, r, payload, version=2)) == G.ERR_UNSUPPORTED
    assert _status(make_image(G.IMAGE_SQF_GPU, q, r + 1, payload)) == G.ERR_INVALID  # size vs geometry
    assert _status(make_image(5, q, r, payload)) == G.ERR_INVALID                 # layout
    assert _status(good[:-1]) == G.ERR_INVALID                                    # truncated
    bad = bytearray(good)
    bad[64 + 3] ^= 1
    assert _status(bytes(bad)) == G.ERR_INVALID                                   # checksum
    # a load refuses a bad image before any device call
    fake = G._SQF(q, r, 1, C.c_void_p(0x1000))
    b = np.frombuffer(bytes(bad), dtype=np.uint8)
    assert G.lib().gqf_sqf_image_load(fake, b.ctypes.data, b.size) == G.ERR_INVALID
    other = make_image(G.IMAGE_SQF_GPU, q + 1, r, bytes(O.sqf_table_bytes(q + 1, r)))
    b = np.frombuffer(other, dtype=np.uint8)
    assert G.lib().gqf_sqf_image_load(fake, b.ctypes.data, b.size) == G.ERR_INVALID  # q mismatch


def test_hash_mismatch_is_refused_before_any_device_call():
    """An image records the SQF hash its table was built with; loading it on a
    thread that queries with the other hash would turn every stored key into a
    miss, so the load fails (checked before the device is touched)."""
    q, r = 6, 5
    fake = G._SQF(q, r, 1, C.c_void_p(0x1000))
    fnv = np.frombuffer(make_image(G.IMAGE_SQF_GPU, q, r, bytes(O.sqf_table_bytes(q, r)), hash_kind=G.HASH_FNV),
                        dtype=np.uint8)
    assert G.sqf_hash() == G.HASH_AP
    assert G.lib().gqf_sqf_image_load(fake, fnv.ctypes.data, fnv.size) == G.ERR_INVALID
    assert "hash" in G.lib().gqf_last_error().decode()
    fake_r = G._RSQF(q, G.RSQF.calcNumBlocksGPU(q), C.c_void_p(0x1000), r)
    rs = np.frombuffer(make_image(G.IMAGE_RSQF, q, r, bytes(G.lib().gqf_rsqf_ref_bytes(q, r)), hash_kind=G.HASH_FNV,
                                  units=G.RSQF.calcNumBlocksGPU(q)), dtype=np.uint8)
    assert G.lib().gqf_rsqf_image_load(fake_r, rs.ctypes.data, rs.size) == G.ERR_INVALID
