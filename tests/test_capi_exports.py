"""The C-ABI libraries load and export every declared symbol; host-side
(non-GPU) entry points agree with the oracle bit for bit."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle_lib as O
from conftest import REPO


def declared(header):
    with open(header) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ym[a-z]?_[a-z0-9_]+)\s*\(", text)))


def test_product_library_exports_every_declared_symbol(ym):
    lib = ym.lib()
    names = declared(os.path.join(REPO, "include", "ym_amd.h"))
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the Python mirror binds the whole ABI
    from ym_amd._lib import SIGNATURES
    assert set(names) == set(SIGNATURES), set(names) ^ set(SIGNATURES)


def test_oracle_library_exports_every_declared_symbol():
    lib = O.lib()
    for h in ("ym_oracle.h", "dev_exact.h"):
        for n in declared(os.path.join(REPO, "oracle", h)):
            assert hasattr(lib, n), n


def test_product_does_not_link_the_oracle():
    so = os.path.join(REPO, "yang-mills_amd", "libym_amd.so")
    with open(so, "rb") as f:
        blob = f.read()
    assert b"libym_oracle" not in blob and b"ymo_" not in blob and b"ymd_" not in blob


def test_host_statistics_match_oracle(ym):
    rng = np.random.default_rng(4)
    L = O.lib()
    for n, lag in ((16, 4), (200, 24), (97, 30)):
        v = np.cumsum(rng.standard_normal(n)) * 0.01 + 0.6
        d, ac = ym.diagnostics(v, lag)
        od = O.ChainDiagnostics()
        oac = np.zeros(lag + 1)
        assert L.ymo_diagnostics(O.dptr(v), n, lag, C.byref(od), O.dptr(oac)) == 0
        assert (d.mean, d.variance, d.tau_int, d.ess, d.standard_error, d.positive_window_lag) == \
            (od.mean, od.variance, od.tau_int, od.ess, od.standard_error, od.positive_window_lag)
        assert np.array_equal(ac, oac)
    for m, c in ((0x3C00198212230101, 0), (0x2C02198105010002, 3), (7, 1 << 40)):
        assert ym.derive_su3_chain_seed(m, c) == L.ymo_derive_su3_chain_seed(m, c)
        assert ym.derive_chain_seed(m, c) == L.ymo_derive_chain_seed(m, c)


def test_diagnostics_error_codes(ym):
    with pytest.raises(ym.YmError) as e:
        ym.diagnostics([1.0], 0)
    assert e.value.status == 32
    with pytest.raises(ym.YmError) as e:
        ym.diagnostics([1.0, 2.0, 3.0], 3)
    assert e.value.status == 34
    with pytest.raises(ym.YmError) as e:
        ym.diagnostics([1.0, float("nan")], 0)
    assert e.value.status == 37


def test_device_calls_fail_loudly_without_a_gpu(ym):
    from conftest import gpu_available
    if gpu_available():
        pytest.skip("a GPU is present")
    with pytest.raises(ym.YmError) as e:
        ym.Su3Lattice4((4, 4, 4, 4), 32)
    assert e.value.status == 97
