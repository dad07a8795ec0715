"""Table images written independently of the library (struct + FNV-1a 64), to
check that its reader and writer follow the documented header."""
import struct

import numpy as np

MAGIC = b"GQFTABLE"


def fnv1a64(data):
    h = 0xCBF29CE484222325
    for b in data:  # small payloads only
        h = ((h ^ b) * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    return h


def fnv1a64_np(payload):
    """Same hash, vectorised enough for table-sized payloads."""
    h = 0xCBF29CE484222325
    p = np.frombuffer(payload, dtype=np.uint8)
    for b in p.tolist():
        h = ((h ^ b) * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    return h


def make_image(layout, q, r, payload, hash_kind=0, units=0, version=1, checksum=None):
    payload = bytes(payload)
    if checksum is None:
        checksum = fnv1a64_np(payload)
    head = MAGIC + struct.pack("<IIIIII", version, layout, q, r, hash_kind, 0)
    head += struct.pack("<QQQQ", len(payload), checksum, units, 0)
    assert len(head) == 64
    return head + payload
