"""CPU oracle (test infrastructure only) for the SU(2) canonical text codec.

Restates Lattice4::to_canonical_config / from_canonical_config
(reference/rust/src/lib.rs:587-694, format id and tolerances lib.rs:20-23,
Su2::from_preserved_unit_components lib.rs:163-177, Lattice4::new_identity
lib.rs:242-256) as the checker of libym_amd.so's ym_su2_config_encode /
ym_su2_config_decode.  Nothing in the product imports this file.

Pinned by the reference's own round-trip fixture (main.rs:13-24: a 3^4
lattice with one traceless link encodes, decodes and re-encodes exactly) and
by the IEEE-754 bit patterns of the identity (1.0 = 3ff0000000000000) that
fix every line of an identity configuration (tests/test_codec.py).
Errors are raised as OracleError with the product's status numbers.
"""
import struct

FORMAT_ID = "YM-SU2-CONFIG-V1"
TOL = 2.0e-12
MAX_LINKS = 1 << 20

STATUS = {"NonFiniteSu2Component": 13, "NonUnitSu2": 15, "InvalidLatticeExtent": 1, "LatticeSizeOverflow": 19,
          "LatticeTooLarge": 2, "InvalidConfigurationEncoding": 18}


class OracleError(Exception):
    def __init__(self, name):
        super().__init__(name)
        self.name = name
        self.status = STATUS[name]


def _bits(x):
    return struct.unpack("<Q", struct.pack("<d", x))[0]


def _from_bits(b):
    return struct.unpack("<d", struct.pack("<Q", b))[0]


def _link_count(extents):  # lib.rs:242-256
    if any(n < 2 for n in extents):
        raise OracleError("InvalidLatticeExtent")
    sites = 1
    for n in extents:
        sites *= n
        if sites >= 1 << 64:
            raise OracleError("LatticeSizeOverflow")
    links = sites * 4
    if links >= 1 << 64:
        raise OracleError("LatticeSizeOverflow")
    if links > MAX_LINKS:
        raise OracleError("LatticeTooLarge")
    return links


def _norm_sq(c):
    s = 0.0
    for v in c:
        s += v * v
    return s


def encode(extents, links):  # lib.rs:587-614; links: list of 4-tuples
    worst = 0.0
    for c in links:
        r = abs(_norm_sq(c) - 1.0)
        worst = max(worst, r) if r == r else float("inf")
    if worst > TOL:
        raise OracleError("NonUnitSu2")
    out = [FORMAT_ID, "extents=%d,%d,%d,%d" % tuple(extents), "link_count=%d" % len(links)]
    for c in links:
        out.append(",".join("%016x" % _bits(v) for v in c))
    return "\n".join(out) + "\n"


def _usize(s):  # Rust usize::from_str
    t = s[1:] if s.startswith("+") else s
    if not t or not all("0" <= ch <= "9" for ch in t):
        raise OracleError("InvalidConfigurationEncoding")
    v = int(t)
    if v >= 1 << 64:
        raise OracleError("InvalidConfigurationEncoding")
    return v


def _lines(text):  # str::lines
    parts = text.split("\n")
    if parts and parts[-1] == "":
        parts.pop()
    return [p[:-1] if p.endswith("\r") else p for p in parts]


def decode(text):  # lib.rs:616-694
    lines = iter(_lines(text))
    enc = OracleError("InvalidConfigurationEncoding")
    if next(lines, None) != FORMAT_ID:
        raise enc
    el = next(lines, None)
    if el is None or not el.startswith("extents="):
        raise enc
    ep = el[len("extents="):].split(",")
    if len(ep) < 4:
        raise enc
    extents = [_usize(p) for p in ep[:4]]
    if len(ep) > 4:
        raise enc
    cl = next(lines, None)
    if cl is None or not cl.startswith("link_count="):
        raise enc
    declared = _usize(cl[len("link_count="):])
    if declared > MAX_LINKS:
        raise OracleError("LatticeTooLarge")
    n = _link_count(extents)
    if declared != n:
        raise enc
    links = []
    for _ in range(n):
        line = next(lines, None)
        if line is None:
            raise enc
        parts = line.split(",")
        comps = []
        for k in range(4):
            if k >= len(parts):
                raise enc
            tok = parts[k]
            if len(tok) != 16 or not all(ch in "0123456789abcdef" for ch in tok):
                raise enc
            comps.append(_from_bits(int(tok, 16)))
        if len(parts) > 4:
            raise enc
        if not all(v == v and abs(v) != float("inf") for v in comps):
            raise OracleError("NonFiniteSu2Component")
        if any(abs(v) > 1.0 + TOL for v in comps):
            raise OracleError("NonUnitSu2")
        ns = _norm_sq(comps)
        if abs(ns - 1.0) > TOL:
            raise OracleError("NonUnitSu2")
        links.append(tuple(comps))
    if next(lines, None) is not None:
        raise enc
    if encode(extents, links) != text:
        raise enc
    return tuple(extents), links
