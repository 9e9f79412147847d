"""Pins of the numpy oracle of the learned proposal (oracle/learned_oracle.py)
by mathematics: finite differences, round trips and closed forms.  No
reference counterpart exists for this part."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import learned_oracle as LO  # noqa: E402


def random_eps(rng, scale):
    return rng.normal(size=8) * scale


def test_haar_jacobian_matches_finite_difference_volume():
    rng = np.random.default_rng(1)
    for scale in (0.05, 0.5, 1.2):
        for _ in range(5):
            eps = random_eps(rng, scale)
            w = np.linalg.eigvalsh(LO.from_coords(eps))
            if np.max(w) - np.min(w) > 5.5:
                continue
            assert abs(LO.haar_log_jacobian(w) - LO.numerical_log_jacobian(eps)) < 1e-6


def test_haar_jacobian_special_values():
    assert LO.haar_log_jacobian([0, 0, 0]) == 0.0
    # one pair at separation pi: sinc^2(pi/2) = 4/pi^2, the other two pairs at pi/2
    x = [np.pi / 2, -np.pi / 2, 0.0]
    expect = 2 * np.log(2 / np.pi) + 4 * np.log(np.sin(np.pi / 4) / (np.pi / 4))
    assert abs(LO.haar_log_jacobian(x) - expect) < 1e-14


def test_gell_mann_coordinates_round_trip_and_normalisation():
    for a in range(8):
        for b in range(8):
            assert abs(np.trace(LO.GELL_MANN[a] @ LO.GELL_MANN[b]) - 2 * (a == b)) < 1e-15
    rng = np.random.default_rng(2)
    eps = rng.normal(size=8)
    assert np.allclose(LO.coords(LO.from_coords(eps)), eps, atol=1e-14)


def test_log_inverts_exp_on_the_chart():
    rng = np.random.default_rng(3)
    for scale in (1e-6, 0.1, 0.8):
        for _ in range(10):
            eps = random_eps(rng, scale)
            X = LO.from_coords(eps)
            if np.max(np.abs(np.linalg.eigvalsh(X))) > 3.0:
                continue
            r = LO.su3_log(LO.expi(X))
            assert r is not None
            assert np.allclose(r[0], X, atol=1e-12)


def test_frame_is_special_unitary_and_covariant():
    rng = np.random.default_rng(4)

    def haar():
        z = (rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))) / np.sqrt(2)
        q, r = np.linalg.qr(z)
        q = q * (np.diag(r) / np.abs(np.diag(r)))
        return q / np.linalg.det(q) ** (1 / 3)

    A = sum(haar() for _ in range(6))
    W = LO.frame(A)
    assert np.allclose(W @ W.conj().T, np.eye(3), atol=1e-13)
    assert abs(np.linalg.det(W) - 1) < 1e-13
    # A -> g A h^dagger gives W -> h W g^dagger; features are invariant
    g, h = haar(), haar()
    A2 = g @ A @ h.conj().T
    assert np.allclose(LO.frame(A2), h @ W @ g.conj().T, atol=1e-12)
    assert np.allclose(LO.features(A2), LO.features(A), atol=1e-12)


def test_loss_gradient_matches_finite_differences():
    rng = np.random.default_rng(5)
    p = rng.normal(size=LO.N_PARAMS) * 0.2
    X = rng.normal(size=(40, 12))
    T = rng.normal(size=(40, 8)) * 0.5
    valid = (rng.random(40) > 0.2).astype(float)
    _, g = LO.mlp_loss_grad(p, X, T, valid)
    idx = rng.choice(LO.N_PARAMS, 40, replace=False)
    for i in idx:
        e = np.zeros_like(p)
        e[i] = 1e-6
        fd = (LO.mlp_loss_grad(p + e, X, T, valid)[0] - LO.mlp_loss_grad(p - e, X, T, valid)[0]) / 2e-6
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd))


def test_adam_first_step_is_lr_times_sign():
    p = np.array([1.0, -2.0, 3.0])
    g = np.array([0.5, -1e-3, 0.0])
    p1, m, v = LO.adam_step(p, g, np.zeros(3), np.zeros(3), 1, 0.01)
    # bias-corrected first step: m_hat / sqrt(v_hat) = g / |g|
    assert np.allclose(p1[:2], p[:2] - 0.01 * np.sign(g[:2]), atol=1e-7)
    assert p1[2] == p[2]


def test_bf16_rounding_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.14159265], dtype=np.float32)
    y = LO.bf16(x)
    assert y[0] == 1.0
    assert y[1] == 1.0                 # tie -> even
    assert y[2] == 1.0 + 2 ** -6       # tie -> even (up)
    assert y[3] == -2.5
    assert abs(y[4] - 3.140625) < 1e-7
