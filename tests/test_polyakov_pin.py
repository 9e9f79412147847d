"""Pin of the reference-exact SU(3) Polyakov loop (oracle/ref_su3.c,
ymo_su3_polyakov_loop, restating su3_lattice.rs:185-212).

* Centre-slice transform (tests/su3_center.rs:97-120, su3_center.rs:139-174):
  multiplying every direction-d link of one slice x_d = s by the centre
  element z = e^{+-2 pi i/3} multiplies the Polyakov loop at every start
  point by z and leaves the action unchanged -- on the identity lattice
  (the reference's test, 2e-15) and on hot-start configurations.  A loop
  that misses a link, runs along another direction or stops short of the
  full extent fails this for some slice.
* Gauge invariance: P(x) is unchanged by U_mu(x) -> g(x) U_mu(x) g(x+mu)^+;
  a product taken in the wrong order is not.
* Closed form: with identity links off direction d and diagonal links
  diag(e^{i a(x)}, e^{i b(x)}, e^{-i(a(x)+b(x))}) along d, the loop from a
  base point is (1/3) sum_k exp(i sum_t theta_k), which fixes the trace
  normalisation and the extent of the product.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O

L = O.lib()
Z = {1: complex(-0.5, np.sqrt(3.0) / 2.0), 2: complex(-0.5, -np.sqrt(3.0) / 2.0)}


def hot_su3(n, key):
    vh = L.ymd_slots_per_parity(O.ext(n))
    buf = np.zeros(8 * 18 * vh)
    assert L.ymd_su3_hot_start(O.ext(n), buf.ctypes.data_as(C.c_void_p), 64, key) == 0
    aos = np.zeros(int(np.prod(n)) * 4 * 18)
    assert L.ymd_su3_from_device(O.ext(n), buf.ctypes.data_as(C.c_void_p), 64, O.dptr(aos)) == 0
    return aos


def centre_transform(n, links, direction, slice_, element):
    """Links of direction `direction` on x_direction = slice_ multiplied on the left by z I."""
    out = links.copy().reshape(-1, 4, 9, 2)
    z = Z[element]
    for i, x in enumerate(np.ndindex(*n)):
        if x[direction] == slice_:
            m = out[i, direction, :, 0] + 1j * out[i, direction, :, 1]
            m = z * m
            out[i, direction, :, 0] = m.real
            out[i, direction, :, 1] = m.imag
    return np.ascontiguousarray(out.reshape(-1))


def polyakov(n, links, start, d):
    a, b = C.c_double(), C.c_double()
    st = L.ymo_su3_polyakov_loop(O.ext(n), O.dptr(links), (C.c_int64 * 4)(*start), d, C.byref(a), C.byref(b))
    assert st == 0, st
    return complex(a.value, b.value)


def action(n, links, beta=2.0):
    a = C.c_double()
    assert L.ymo_su3_wilson_action(O.ext(n), O.dptr(links), beta, C.byref(a)) == 0
    return a.value


def test_centre_slice_on_identity_lattice():
    n = (3, 3, 3, 3)
    links = O.su3_identity(n)
    start = (2, 1, 2, 1)
    before = polyakov(n, links, start, 0)
    assert before == 1.0
    t = centre_transform(n, links, 0, 0, 1)
    after = polyakov(n, t, start, 0)
    assert abs(after - Z[1] * before) <= 2e-15
    assert action(n, links) == 0.0
    assert abs(action(n, t)) <= 2e-12


@pytest.mark.parametrize("direction", [0, 1, 2, 3])
def test_centre_slice_on_hot_configuration(direction):
    n = (4, 3, 4, 6)
    links = hot_su3(n, 0x5EED0000 + direction)
    s_before = action(n, links, 5.7)
    for element in (1, 2):
        for slice_ in range(n[direction]):
            t = centre_transform(n, links, direction, slice_, element)
            assert abs(action(n, t, 5.7) - s_before) <= 1e-10 * abs(s_before)
            for start in [(0, 0, 0, 0), (1, 2, 3, 5), (3, 1, 0, 2)]:
                before = polyakov(n, links, start, direction)
                after = polyakov(n, t, start, direction)
                assert abs(after - Z[element] * before) <= 1e-14, (direction, slice_, start)
        # a slice transformed twice with the two elements cancels
        t = centre_transform(n, centre_transform(n, links, direction, 1, 1), direction, 1, 2)
        assert abs(polyakov(n, t, (0, 0, 0, 0), direction) - polyakov(n, links, (0, 0, 0, 0), direction)) <= 1e-14


def test_polyakov_gauge_invariant():
    n = (3, 4, 3, 4)
    links = hot_su3(n, 0xC0FFEE)
    g = np.zeros(int(np.prod(n)) * 18)
    for i, x in enumerate(np.ndindex(*n)):
        assert L.ymo_su3_deterministic_site_gauge((C.c_int64 * 4)(*x), O.dptr(g[i * 18:(i + 1) * 18])) == 0
    t = np.zeros_like(links)
    assert L.ymo_su3_gauge_transform(O.ext(n), O.dptr(links), O.dptr(g), O.dptr(t)) == 0
    for d in range(4):
        for start in np.ndindex(*n):
            assert abs(polyakov(n, links, start, d) - polyakov(n, t, start, d)) <= 2e-11


def diagonal_field(n, d, seed):
    """Identity links off direction d; diag(e^{ia}, e^{ib}, e^{-i(a+b)}) along d."""
    rng = np.random.default_rng(seed)
    links = O.su3_identity(n).reshape(-1, 4, 9, 2)
    ang = rng.uniform(-np.pi, np.pi, size=(int(np.prod(n)), 2))
    for i in range(len(ang)):
        a, b = ang[i]
        ph = np.exp(1j * np.array([a, b, -(a + b)]))
        m = np.diag(ph)
        links[i, d, :, 0] = m.real.reshape(9)
        links[i, d, :, 1] = m.imag.reshape(9)
    return np.ascontiguousarray(links.reshape(-1)), ang


def closed_form(n, ang, d, start):
    """(1/3) sum_k exp(i sum_t theta_k(start + t d))."""
    tot = np.zeros(3)
    x = list(start)
    for _ in range(n[d]):
        i = np.ravel_multi_index(tuple(x), n)
        a, b = ang[i]
        tot += [a, b, -(a + b)]
        x[d] = (x[d] + 1) % n[d]
    return np.exp(1j * tot).sum() / 3.0


@pytest.mark.parametrize("d", [0, 1, 2, 3])
def test_polyakov_closed_form_diagonal_links(d):
    n = (3, 4, 5, 6)
    links, ang = diagonal_field(n, d, 40 + d)
    for start in [(0, 0, 0, 0), (2, 3, 4, 5), (1, 0, 2, 3)]:
        assert abs(polyakov(n, links, start, d) - closed_form(n, ang, d, start)) <= 1e-13
    # constant phases: P = (1/3)(e^{iNa} + e^{iNb} + e^{-iN(a+b)}) everywhere
    a, b = 0.3, -1.1
    const = O.su3_identity(n).reshape(-1, 4, 9, 2)
    m = np.diag(np.exp(1j * np.array([a, b, -(a + b)])))
    const[:, d, :, 0] = m.real.reshape(9)
    const[:, d, :, 1] = m.imag.reshape(9)
    const = np.ascontiguousarray(const.reshape(-1))
    N = n[d]
    want = (np.exp(1j * N * a) + np.exp(1j * N * b) + np.exp(-1j * N * (a + b))) / 3.0
    assert abs(polyakov(n, const, (1, 1, 1, 1), d) - want) <= 1e-13
