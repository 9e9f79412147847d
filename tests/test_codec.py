"""Configuration formats (host-only entry points of libym_amd.so).

The canonical SU(2) text codec (ym_su2_config_encode / _decode) against the
oracle's restatement of lib.rs:587-694 (oracle/ref_codec.py), the
reference's round-trip fixture (main.rs:13-24) and the bit patterns of the
identity; the binary chain checkpoint (ym_checkpoint_*) by exact round trips
and rejection of corrupted or truncated data.
"""
import math
import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "oracle"))
import ref_codec as RC  # noqa: E402


def status_of(fn, *a):
    import ym_amd
    try:
        fn(*a)
    except ym_amd.YmError as e:
        return e.status
    except RC.OracleError as e:
        return e.status
    return 0


def identity_aos(n):
    a = np.zeros((int(np.prod(n)) * 4, 4))
    a[:, 0] = 1.0
    return a


def random_su2(count, seed):
    rng = np.random.default_rng(seed)
    q = rng.standard_normal((count, 4))
    q /= np.linalg.norm(q, axis=1)[:, None]
    return q


def test_identity_encoding_is_fixed_by_ieee_bits(ym):
    n = (2, 2, 3, 2)
    text = ym.codec.encode_su2_config(n, identity_aos(n))
    lines = text.split("\n")
    assert lines[0] == "YM-SU2-CONFIG-V1" and lines[1] == "extents=2,2,3,2" and lines[2] == "link_count=96"
    assert lines[3:-1] == ["3ff0000000000000,0000000000000000,0000000000000000,0000000000000000"] * 96
    assert lines[-1] == "" and len(text) == len("YM-SU2-CONFIG-V1\nextents=2,2,3,2\nlink_count=96\n") + 96 * 68


def test_reference_round_trip_fixture(ym):
    """main.rs:13-24: one traceless link (0, 1, 0, 0) at the origin, direction
    0, on 3^4: encode -> decode -> encode is exact."""
    n = (3, 3, 3, 3)
    a = identity_aos(n)
    a[0] = (0.0, 1.0, 0.0, 0.0)
    text = ym.codec.encode_su2_config(n, a)
    assert text == RC.encode(n, [tuple(r) for r in a])
    ext, back = ym.codec.decode_su2_config(text)
    assert ext == n and np.array_equal(back.reshape(-1, 4), a)
    assert ym.codec.encode_su2_config(ext, back) == text
    assert RC.decode(text) == (n, [tuple(r) for r in a])


def test_random_configurations_product_vs_oracle(ym):
    for n, seed in (((2, 3, 2, 4), 1), ((4, 4, 4, 4), 2)):
        q = random_su2(int(np.prod(n)) * 4, seed)
        text = ym.codec.encode_su2_config(n, q)
        assert text == RC.encode(n, [tuple(r) for r in q])
        ext, back = ym.codec.decode_su2_config(text)
        assert ext == n and np.array_equal(back.reshape(-1, 4), q)  # bit for bit
        assert RC.decode(text)[0] == n


def test_decode_rejections_match_oracle(ym):
    n = (2, 2, 2, 2)
    good = ym.codec.encode_su2_config(n, random_su2(64, 3))
    lines = good.split("\n")

    def mutate(i, new):
        m = list(lines)
        m[i] = new
        return "\n".join(m)

    bad_unit = "3ff0000000000001,3f50000000000000,0000000000000000,0000000000000000"
    nan_line = "7ff8000000000000,0000000000000000,0000000000000000,0000000000000000"
    cases = {
        "wrong header": (mutate(0, "YM-SU2-CONFIG-V2"), 18),
        "five extents": (mutate(1, "extents=2,2,2,2,2"), 18),
        "three extents": (mutate(1, "extents=2,2,2"), 18),
        "extent below two": (mutate(1, "extents=1,2,2,2").replace("link_count=64", "link_count=32"), 1),
        "plus sign": (mutate(1, "extents=+2,2,2,2"), 18),
        "space": (mutate(1, "extents= 2,2,2,2"), 18),
        "count too large": (mutate(2, "link_count=%d" % ((1 << 20) + 1)), 2),
        "count mismatch": (mutate(2, "link_count=63"), 18),
        "uppercase hex": (mutate(3, lines[3].upper()), 18),
        "short token": (mutate(3, lines[3][1:]), 18),
        "five components": (mutate(3, lines[3] + ",0000000000000000"), 18),
        "non-unit": (mutate(3, bad_unit), 15),
        "nan": (mutate(3, nan_line), 13),
        "extra line": (good + "extra\n", 18),
        "blank line": (good + "\n", 18),
        "missing newline": (good[:-1], 18),
        "missing link": ("\n".join(lines[:-2]) + "\n", 18),
        "crlf": (good.replace("\n", "\r\n"), 18),
        "empty": ("", 18),
    }
    for name, (text, code) in cases.items():
        assert status_of(ym.codec.decode_su2_config, text) == code, name
        assert status_of(RC.decode, text) == code, name
    # a lattice above the format's bound cannot be encoded; a non-unit link fails
    big = (32, 32, 34, 8)
    assert status_of(ym.codec.encode_su2_config, big, identity_aos(big)) == 2
    a = identity_aos(n)
    a[5, 0] = 1.0 + 1e-11
    assert status_of(ym.codec.encode_su2_config, n, a) == 15
    assert status_of(RC.encode, n, [tuple(r) for r in a]) == 15
    a[5, 0] = math.nan
    assert status_of(ym.codec.encode_su2_config, n, a) == 15


@pytest.mark.parametrize("group,prec", [(3, 32), (3, 64), (2, 32), (2, 64)])
def test_checkpoint_round_trip_and_corruption(ym, group, prec):
    import ctypes as C
    L = ym.lib()
    n = (2, 3, 2, 4)
    nc = 18 if group == 3 else 4
    rng = np.random.default_rng(group * 100 + prec)
    aos = rng.standard_normal(int(np.prod(n)) * 4 * nc)
    if prec == 32:
        aos = aos.astype(np.float32).astype(np.float64)
    ext = (C.c_int64 * 4)(*n)
    size = C.c_int64()
    assert L.ym_checkpoint_bytes(group, prec, ext, C.byref(size)) == 0
    assert size.value == 96 + aos.size * prec // 8
    buf = C.create_string_buffer(size.value)
    ln = C.c_int64()
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    key, nxt = 0xFEEDFACECAFEBEEF, 12345
    assert L.ym_checkpoint_encode(group, prec, ext, dp(aos), key, nxt, buf, size.value, C.byref(ln)) == 0
    raw = buf.raw
    g, p = C.c_int32(), C.c_int32()
    e2 = (C.c_int64 * 4)()
    k2, s2 = C.c_uint64(), C.c_uint64()
    assert L.ym_checkpoint_header(raw, len(raw), C.byref(g), C.byref(p), e2, C.byref(k2), C.byref(s2)) == 0
    assert (g.value, p.value, tuple(e2), k2.value, s2.value) == (group, prec, n, key, nxt)
    back = np.zeros_like(aos)
    assert L.ym_checkpoint_decode(raw, len(raw), dp(back), back.size) == 0
    assert np.array_equal(back, aos)
    # any flipped byte (header or payload) and any truncation is rejected
    for pos in (0, 17, 40, 70, 90, 96, len(raw) // 2, len(raw) - 1):
        bad = bytearray(raw)
        bad[pos] ^= 0x10
        assert L.ym_checkpoint_decode(bytes(bad), len(bad), dp(back), back.size) == 18, pos
    assert L.ym_checkpoint_decode(raw[:-1], len(raw) - 1, dp(back), back.size) == 18
    assert L.ym_checkpoint_decode(raw[:50], 50, dp(back), back.size) == 18
    assert L.ym_checkpoint_decode(raw, len(raw), dp(back), back.size - 1) == 96


def test_checkpoint_rejects_extents_whose_volume_wraps(ym):
    """A 96-byte header with four extents of 2^32 (4 x 2^128 sites wraps to 0
    links and an empty payload) and a matching FNV-1a checksum must not be
    accepted; neither may a volume above the 2^30 sites a lattice can have."""
    import ctypes as C
    import struct
    L = ym.lib()

    def fnv(b, h=0xCBF29CE484222325):
        for x in b:
            h = ((h ^ x) * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
        return h

    good = C.create_string_buffer(96 + 16 * 4 * 4 * 8)
    n = (2, 2, 2, 2)
    aos = np.tile(np.array([1.0, 0.0, 0.0, 0.0]), 16 * 4)
    ln = C.c_int64()
    assert L.ym_checkpoint_encode(2, 64, (C.c_int64 * 4)(*n), aos.ctypes.data_as(C.POINTER(C.c_double)), 1, 2, good,
                                  len(good), C.byref(ln)) == 0
    head = bytearray(good.raw[:96])
    for ext, links in (((1 << 32,) * 4, 0), (((1 << 15), (1 << 15), 2, 2), 4 << 32)):
        h = bytearray(head)
        h[24:56] = struct.pack("<4q", *ext)
        h[72:88] = struct.pack("<QQ", links & 0xFFFFFFFFFFFFFFFF, (links * 32) & 0xFFFFFFFFFFFFFFFF)
        h[88:96] = struct.pack("<Q", fnv(bytes(h[:88])))
        g, p = C.c_int32(), C.c_int32()
        e2 = (C.c_int64 * 4)()
        k2, s2 = C.c_uint64(), C.c_uint64()
        st = L.ym_checkpoint_header(bytes(h), 96, C.byref(g), C.byref(p), e2, C.byref(k2), C.byref(s2))
        assert st != 0, ext
        size = C.c_int64()
        assert L.ym_checkpoint_bytes(2, 64, (C.c_int64 * 4)(*ext), C.byref(size)) == 2
