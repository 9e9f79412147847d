"""Gauge invariance of the device observables.

Every observable the library measures is a function of traces of closed
loops (plaquette, averaged Wilson loops, Polyakov loop, the A1++ timeslice
sums) and spatial stout smearing is gauge covariant, so a random gauge
transformation U_mu(x) -> g(x) U_mu(x) g(x + mu)^dagger of the links must
leave each of them unchanged: to rounding in FP64 storage, to the storage
rounding in FP32.  This pins the device kernels by the mathematics alone
(a wrong neighbour, orientation or conjugation anywhere in a loop breaks
it), independently of the oracle.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _neighbours(n):
    """Lexicographic index of x + mu for every site (x3 fastest), per mu."""
    idx = np.arange(int(np.prod(n))).reshape(n)
    return [np.roll(idx, -1, axis=mu).ravel() for mu in range(4)]


def _random_su3(rng, count):
    z = (rng.standard_normal((count, 3, 3)) + 1j * rng.standard_normal((count, 3, 3))) / np.sqrt(2.0)
    q, r = np.linalg.qr(z)
    d = np.diagonal(r, axis1=1, axis2=2)
    q = q * (d / np.abs(d))[:, None, :]
    return q / (np.linalg.det(q) ** (1.0 / 3.0))[:, None, None]


def _gauge_su3(n, aos, g):
    V = int(np.prod(n))
    a = aos.reshape(V, 4, 3, 3, 2)
    U = a[..., 0] + 1j * a[..., 1]
    out = np.empty_like(U)
    nb = _neighbours(n)
    for mu in range(4):
        out[:, mu] = g @ U[:, mu] @ np.conj(np.transpose(g[nb[mu]], (0, 2, 1)))
    res = np.empty_like(a)
    res[..., 0] = out.real
    res[..., 1] = out.imag
    return res.reshape(-1)


# SU(2) links are quaternions (a0, a1, a2, a3) = a0 I - i a.sigma with the
# product (a0 b0 - a.b, a0 b + b0 a + a x b) (lib.rs:106-130).
def _qmul(a, b):
    a0, av = a[..., :1], a[..., 1:]
    b0, bv = b[..., :1], b[..., 1:]
    return np.concatenate([a0 * b0 - np.sum(av * bv, axis=-1, keepdims=True),
                           a0 * bv + b0 * av + np.cross(av, bv)], axis=-1)


def _qdag(a):
    return np.concatenate([a[..., :1], -a[..., 1:]], axis=-1)


def _gauge_su2(n, aos, g):
    V = int(np.prod(n))
    U = aos.reshape(V, 4, 4)
    out = np.empty_like(U)
    nb = _neighbours(n)
    for mu in range(4):
        out[:, mu] = _qmul(_qmul(g, U[:, mu]), _qdag(g[nb[mu]]))
    return out.reshape(-1)


def _observables_su3(ym, lat):
    w = lat.wilson_loops(3, 3)
    p = [lat.polyakov_loop(d) for d in range(4)]
    a = lat.timeslice_a1pp()
    sm = ym.spatial_stout_smear(lat, 0.1, 2)
    b = sm.timeslice_a1pp()
    return {"plaquette": np.array([lat.average_plaquette()]), "wilson_re": w[0], "wilson_im": w[1],
            "polyakov": np.array([v[0] for v in p]), "polyakov_abs": np.array([v[1] for v in p]),
            "a1pp_plaq": a[0], "a1pp_rect": a[1], "stout_plaquette": np.array([sm.average_plaquette()]),
            "stout_a1pp_plaq": b[0], "stout_a1pp_rect": b[1]}


@pytest.mark.parametrize("prec", [64, 32])
def test_su3_observables_are_gauge_invariant(ym, prec):
    n = (6, 6, 6, 8)
    key = ym.derive_su3_chain_seed(0x6A0E, 1)
    src = ym.Su3Lattice4.new_identity(n, 64)  # cold start and a few sweeps: loops of order 0.1 - 0.7
    src.metropolis_sweeps(6.0, 0.35, key, 0, 6, 8)
    links = src.download()
    g = _random_su3(np.random.default_rng(7), int(np.prod(n)))
    moved = _gauge_su3(n, links, g)
    a, b = ym.Su3Lattice4(n, prec), ym.Su3Lattice4(n, prec)
    a.upload(links)
    b.upload(moved)
    oa, ob = _observables_su3(ym, a), _observables_su3(ym, b)
    tol = 2e-13 if prec == 64 else 2e-6
    for k in oa:
        assert np.max(np.abs(oa[k] - ob[k])) <= tol, (k, np.max(np.abs(oa[k] - ob[k])))
    assert oa["plaquette"][0] > 0.3 and np.max(np.abs(oa["wilson_re"])) > 0.01


@pytest.mark.parametrize("prec", [64, 32])
def test_su2_observables_are_gauge_invariant(ym, prec):
    n = (6, 6, 6, 8)
    src = ym.Lattice4.new_identity(n, 64)
    src.markov_sweeps(2.3, 0x5EED, 0, 4)
    links = src.download()
    rng = np.random.default_rng(11)
    g = rng.standard_normal((int(np.prod(n)), 4))
    g /= np.linalg.norm(g, axis=1, keepdims=True)
    moved = _gauge_su2(n, links, g)
    a, b = ym.Lattice4(n, prec), ym.Lattice4(n, prec)
    a.upload(links)
    b.upload(moved)
    tol = 2e-13 if prec == 64 else 2e-6
    assert abs(a.average_plaquette() - b.average_plaquette()) <= tol
    wa, wb = a.wilson_loops(3, 3), b.wilson_loops(3, 3)
    assert np.max(np.abs(wa[0] - wb[0])) <= tol and np.max(np.abs(wa[1] - wb[1])) <= tol
    for d in range(4):
        pa, pb = a.polyakov_loop(d), b.polyakov_loop(d)
        assert abs(pa[0] - pb[0]) <= tol and abs(pa[1] - pb[1]) <= tol
    assert a.average_plaquette() > 0.3
