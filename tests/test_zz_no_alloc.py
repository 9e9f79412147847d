"""The allocation-free plaquette entry point against ym_average_plaquette.

ym_average_plaquette_noalloc runs the same kernels with its partials in a
fixed library-wide device array instead of the handle's buffer, so the two
results must agree bit for bit; the array keeps the previous call's
contents, so a smaller lattice right after a larger one must not see stale
partials.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _plaquettes(ym, lat):
    lib = ym.lib()
    a, b = C.c_double(), C.c_double()
    assert lib.ym_average_plaquette_noalloc(lat.handle, C.byref(a)) == 0
    assert lib.ym_average_plaquette(lat.handle, C.byref(b)) == 0
    return a.value, b.value


@pytest.mark.parametrize("group,prec", [(3, 32), (3, 64), (2, 32), (2, 64)])
def test_noalloc_plaquette_matches_allocating_entry_point(ym, group, prec):
    cls = ym.Su3Lattice4 if group == 3 else ym.Lattice4
    big = cls((16, 16, 16, 32), prec)
    small = cls((4, 4, 6, 8), prec)
    big.hot_start(ym.derive_su3_chain_seed(7, group * 100 + prec))
    small.hot_start(ym.derive_su3_chain_seed(8, group * 100 + prec))
    print(f"[noalloc] group {group} prec {prec}: 16^3x32 then 4x4x6x8", flush=True)
    for lat in (big, small):
        a, b = _plaquettes(ym, lat)
        assert np.float64(a).tobytes() == np.float64(b).tobytes(), (a, b)
        assert -1.0 < a < 1.0
    # a second round in the same order: the array holds the small lattice's
    # partials when the large one is measured again, and vice versa
    for lat in (big, small):
        a, b = _plaquettes(ym, lat)
        assert np.float64(a).tobytes() == np.float64(b).tobytes(), (a, b)


def test_noalloc_plaquette_identity_and_limit(ym):
    lib = ym.lib()
    lat = ym.Su3Lattice4((8, 8, 8, 8), 32)
    v = C.c_double()
    assert lib.ym_average_plaquette_noalloc(lat.handle, C.byref(v)) == 0
    assert v.value == 1.0
    assert lib.ym_average_plaquette_noalloc(lat.handle, None) == 96
    # 2^23 + 8 slots per parity: one tile beyond the fixed workspace
    over = ym.Lattice4((2, 2, 2, (1 << 21) + 2), 32)
    assert lib.ym_average_plaquette_noalloc(over.handle, C.byref(v)) == 95
    # the allocating entry point still measures it
    assert lib.ym_average_plaquette(over.handle, C.byref(v)) == 0
    assert v.value == 1.0
