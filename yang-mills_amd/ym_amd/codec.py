"""Configuration formats over libym_amd.so (host-only entry points).

``to_canonical_config`` / ``from_canonical_config`` restate the reference's
YM-SU2-CONFIG-V1 text codec (lib.rs:587-694) for SU(2) lattices;
``save_checkpoint`` / ``load_checkpoint`` write and read the binary chain
checkpoint (links in storage precision + key + next sweep), from which a
chain continues bit for bit because the device RNG is counter-based.
"""
import ctypes as C

import numpy as np

from ._lib import check, lib


def _ext(extents):
    return (C.c_int64 * 4)(*[int(v) for v in extents])


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def encode_su2_config(extents, aos):
    """AoS FP64 SU(2) links (4 per link) -> canonical text (lib.rs:587-614)."""
    a = np.ascontiguousarray(aos, dtype=np.float64).reshape(-1)
    n = C.c_int64()
    check(lib().ym_su2_config_encode(_ext(extents), _dp(a), None, 0, C.byref(n)), "ym_su2_config_encode")
    buf = C.create_string_buffer(n.value)
    check(lib().ym_su2_config_encode(_ext(extents), _dp(a), buf, n.value, C.byref(n)), "ym_su2_config_encode")
    return buf.raw[:n.value].decode("ascii")


def decode_su2_config(text):
    """Canonical text -> (extents, AoS FP64 links) (lib.rs:616-694)."""
    raw = text.encode("utf-8") if isinstance(text, str) else bytes(text)
    ext = (C.c_int64 * 4)()
    check(lib().ym_su2_config_decode(raw, len(raw), ext, None, 0), "ym_su2_config_decode")
    extents = tuple(int(v) for v in ext)
    aos = np.zeros(4 * 4 * int(np.prod(extents)))
    check(lib().ym_su2_config_decode(raw, len(raw), ext, _dp(aos), aos.size), "ym_su2_config_decode")
    return extents, aos


def to_canonical_config(lattice):
    """Lattice4::to_canonical_config for a device SU(2) lattice."""
    if lattice.GROUP != 2:
        raise ValueError("the canonical text format holds SU(2) configurations")
    return encode_su2_config(lattice.extents, lattice.download())


def from_canonical_config(text, precision=64):
    """Lattice4::from_canonical_config into a new device lattice."""
    from .lattice import Lattice4
    extents, aos = decode_su2_config(text)
    lat = Lattice4(extents, precision)
    lat.upload(aos)
    return lat


def checkpoint_bytes(lattice, key, next_sweep):
    """Binary checkpoint of a lattice and its chain position."""
    aos = np.ascontiguousarray(lattice.download())
    size = C.c_int64()
    check(lib().ym_checkpoint_bytes(lattice.GROUP, lattice.precision, _ext(lattice.extents), C.byref(size)),
          "ym_checkpoint_bytes")
    buf = C.create_string_buffer(size.value)
    n = C.c_int64()
    check(lib().ym_checkpoint_encode(lattice.GROUP, lattice.precision, _ext(lattice.extents), _dp(aos),
                                     C.c_uint64(int(key)), C.c_uint64(int(next_sweep)), buf, size.value,
                                     C.byref(n)), "ym_checkpoint_encode")
    return buf.raw[:n.value]


def lattice_from_checkpoint(data):
    """(lattice, key, next_sweep) from checkpoint bytes."""
    from .lattice import Lattice4, Su3Lattice4
    raw = bytes(data)
    g, pr = C.c_int32(), C.c_int32()
    ext = (C.c_int64 * 4)()
    key, nxt = C.c_uint64(), C.c_uint64()
    check(lib().ym_checkpoint_header(raw, len(raw), C.byref(g), C.byref(pr), ext, C.byref(key), C.byref(nxt)),
          "ym_checkpoint_header")
    cls = Su3Lattice4 if g.value == 3 else Lattice4
    lat = cls(tuple(ext), pr.value)
    aos = np.zeros(lat.link_count() * cls.NCOMP)
    check(lib().ym_checkpoint_decode(raw, len(raw), _dp(aos), aos.size), "ym_checkpoint_decode")
    lat.upload(aos)
    return lat, key.value, nxt.value


def save_checkpoint(path, lattice, key, next_sweep):
    data = checkpoint_bytes(lattice, key, next_sweep)
    with open(path, "wb") as f:
        f.write(data)
    return len(data)


def load_checkpoint(path):
    with open(path, "rb") as f:
        return lattice_from_checkpoint(f.read())
