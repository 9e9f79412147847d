"""Configuration formats with device lattices.

A chain resumed from a binary checkpoint (links + key + next sweep)
continues bit for bit like the uninterrupted chain (SU(3) Metropolis and
SU(2) heat-bath/over-relaxation, FP32 and FP64 storage); a device SU(2)
configuration round-trips through the canonical YM-SU2-CONFIG-V1 text
exactly, and its text equals the oracle's encoding of the same links.
"""
import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "oracle"))
import ref_codec as RC  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [32, 64])
def test_su3_chain_resumes_bit_for_bit(ym, tmp_path, prec):
    n = (8, 8, 6, 8)
    key = ym.derive_su3_chain_seed(0xC0FFEE, 7)
    straight = ym.Su3Lattice4(n, prec)
    straight.hot_start(key)
    p0, a0 = straight.metropolis_sweeps(6.0, 0.35, key, 0, 6, 8)
    part = ym.Su3Lattice4(n, prec)
    part.hot_start(key)
    p1, a1 = part.metropolis_sweeps(6.0, 0.35, key, 0, 3, 8)
    path = str(tmp_path / "chain.ckpt")
    part.save_checkpoint(path, key, 3)
    part.close()
    resumed, k, nxt = ym.Su3Lattice4.load_checkpoint(path)
    assert (k, nxt, resumed.extents, resumed.precision) == (key, 3, n, prec)
    p2, a2 = resumed.metropolis_sweeps(6.0, 0.35, k, nxt, 3, 8)
    assert (p1 + p2, a1 + a2) == (p0, a0)
    assert np.array_equal(resumed.download(), straight.download())


@pytest.mark.parametrize("prec", [32, 64])
def test_su2_chain_resumes_bit_for_bit(ym, tmp_path, prec):
    n = (6, 6, 8, 8)
    key = ym.derive_chain_seed(0xBEEF, 3)

    def update(lat, sweep):
        lat.heatbath_sweeps(2.3, key, sweep, 1)
        lat.overrelax_sweeps(3)

    straight = ym.Lattice4(n, prec)
    straight.hot_start(key)
    for s in range(5):
        update(straight, s)
    part = ym.Lattice4(n, prec)
    part.hot_start(key)
    for s in range(2):
        update(part, s)
    path = str(tmp_path / "su2.ckpt")
    part.save_checkpoint(path, key, 2)
    resumed, k, nxt = ym.Lattice4.load_checkpoint(path)
    for s in range(nxt, 5):
        update(resumed, s)
    assert np.array_equal(resumed.download(), straight.download())


def test_su2_canonical_text_round_trip(ym):
    n = (4, 4, 4, 6)
    key = ym.derive_chain_seed(0x2B00201211160101, 0)
    lat = ym.Lattice4(n, 64)
    lat.hot_start(key)
    lat.heatbath_sweeps(2.3, key, 0, 2)
    aos = lat.download()
    text = lat.to_canonical_config()
    assert text == RC.encode(n, [tuple(r) for r in aos.reshape(-1, 4)])
    back = ym.Lattice4.from_canonical_config(text)
    assert back.extents == n and np.array_equal(back.download(), aos)
    assert back.to_canonical_config() == text
    assert abs(back.average_plaquette() - lat.average_plaquette()) == 0.0
    # FP32 storage breaks the 2e-12 unit invariant of the format
    lat32 = ym.Lattice4(n, 32)
    lat32.hot_start(key)
    with pytest.raises(ym.YmError) as e:
        lat32.to_canonical_config()
    assert e.value.status == 15
