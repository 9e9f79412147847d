"""ORACLE -- TEST INFRASTRUCTURE ONLY.

numpy restatement of the learned SU(3) link proposal, used only by tests to
check the HIP kernels.  There is no reference counterpart (the reference
samples with plain single-generator Metropolis, su3_ensemble.rs:336-391), so
every function here is pinned by mathematics instead, in
tests/test_learned_oracle.py:
  * haar_log_jacobian: against a finite-difference volume of the exponential
    map (the 8x8 Jacobian of eps -> exp(i X(eps)), left-translated);
  * su3_log / chart: log(exp(i X)) = X on the chart, U = exp(i X) W;
  * frame: W is special unitary and maximises nothing by assumption -- it is
    checked as the SVD polar factor with the determinant phase removed;
  * mlp_loss_grad: against central finite differences in float64;
  * adam_step: the published update (Kingma & Ba, 2015) written out.
Parity of this module with the reference: unpinned (no reference).
"""
import numpy as np

N_IN, N_HID, N_OUT = 12, 64, 16
OFF = {}
_o = 0
for _name, _n in (("W1", N_HID * N_IN), ("b1", N_HID), ("W2", N_HID * N_HID), ("b2", N_HID),
                  ("W3", N_OUT * N_HID), ("b3", N_OUT)):
    OFF[_name] = (_o, _n)
    _o += _n
N_PARAMS = _o
LS_MIN, LS_MAX = -6.0, 1.0

GELL_MANN = np.zeros((8, 3, 3), dtype=np.complex128)
GELL_MANN[0][0, 1] = GELL_MANN[0][1, 0] = 1
GELL_MANN[1][0, 1] = -1j
GELL_MANN[1][1, 0] = 1j
GELL_MANN[2][0, 0], GELL_MANN[2][1, 1] = 1, -1
GELL_MANN[3][0, 2] = GELL_MANN[3][2, 0] = 1
GELL_MANN[4][0, 2] = -1j
GELL_MANN[4][2, 0] = 1j
GELL_MANN[5][1, 2] = GELL_MANN[5][2, 1] = 1
GELL_MANN[6][1, 2] = -1j
GELL_MANN[6][2, 1] = 1j
GELL_MANN[7] = np.diag([1, 1, -2]) / np.sqrt(3)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    g = {k: p[o:o + n] for k, (o, n) in OFF.items()}
    return (g["W1"].reshape(N_HID, N_IN), g["b1"], g["W2"].reshape(N_HID, N_HID), g["b2"],
            g["W3"].reshape(N_OUT, N_HID), g["b3"])


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return u.view(np.float32)


def mlp_forward(p, X, bf16_round=False):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    X = np.asarray(X, dtype=np.float64)
    if bf16_round:
        q = lambda a: bf16(a).astype(np.float64)  # noqa: E731
        X, W1, W2, W3 = q(X), q(W1), q(W2), q(W3)
        h1 = q(np.tanh(X @ W1.T + b1))
        h2 = q(np.tanh(h1 @ W2.T + b2))
    else:
        h1 = np.tanh(X @ W1.T + b1)
        h2 = np.tanh(h1 @ W2.T + b2)
    return h2 @ W3.T + b3


def mlp_loss_grad(p, X, T, valid):
    """Mean over valid rows of sum_a 0.5 (t_a - mu_a)^2 e^{-2 ls_a} + ls_a,
    ls = clip(out[8:], LS_MIN, LS_MAX); gradient with respect to p."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    X, T = np.asarray(X, np.float64), np.asarray(T, np.float64)
    w = np.asarray(valid, np.float64)
    h1 = np.tanh(X @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    out = h2 @ W3.T + b3
    mu, raw = out[:, :8], out[:, 8:]
    ls = np.clip(raw, LS_MIN, LS_MAX)
    iv = np.exp(-2 * ls)
    e = T - mu
    n = w.sum()
    loss = (w * (0.5 * e * e * iv + ls).sum(axis=1)).sum() / n
    d = np.zeros_like(out)
    d[:, :8] = -e * iv
    inside = (raw > LS_MIN) & (raw < LS_MAX)
    d[:, 8:] = np.where(inside, 1 - e * e * iv, 0.0)
    d *= w[:, None] / n
    gW3 = d.T @ h2
    gb3 = d.sum(0)
    dz2 = (d @ W3) * (1 - h2 ** 2)
    gW2 = dz2.T @ h1
    gb2 = dz2.sum(0)
    dz1 = (dz2 @ W2) * (1 - h1 ** 2)
    gW1 = dz1.T @ X
    gb1 = dz1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, g


def adam_step(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def from_coords(eps):
    return np.einsum("a,aij->ij", np.asarray(eps, np.float64), GELL_MANN) / 2


def coords(X):
    return np.real(np.einsum("ij,aji->a", X, GELL_MANN))


def expi(X):
    w, V = np.linalg.eigh(X)
    return (V * np.exp(1j * w)) @ V.conj().T


def su3_log(Y):
    """Principal log of a special unitary Y: (X Hermitian with eigenvalues in
    (-pi, pi), phases) or None when the phases do not sum to zero."""
    # Y is normal: the Schur form is diagonal
    w, V = np.linalg.eig(Y)
    # re-orthonormalise eigenvectors (degenerate eigenvalues)
    V, _ = np.linalg.qr(V)
    z = np.einsum("ik,ij,jk->k", V.conj(), Y, V)
    ph = np.angle(z)
    if abs(ph.sum()) > 1e-9 or np.max(np.abs(ph)) >= np.pi - 1e-6:
        return None
    return (V * ph) @ V.conj().T, ph


def frame(A):
    """W(A) = e^{-i psi/3} polar(A^dagger), polar from the SVD."""
    Uo, s, Vh = np.linalg.svd(A.conj().T)
    Z = Uo @ Vh
    psi = np.angle(np.linalg.det(Z))
    return np.exp(-1j * psi / 3) * Z


def features(A):
    s = np.sort(np.linalg.svd(A, compute_uv=False))[::-1] / 3.0
    psi = np.angle(np.linalg.det(A))
    return np.array([s[0], s[1], s[2], s[0] * s[1], s[0] * s[2], s[1] * s[2], s[0] * s[1] * s[2],
                     (s[0] + s[1] + s[2]) / 3, np.cos(psi), np.sin(psi), np.cos(psi / 3), np.sin(psi / 3)])


def reunitarize(U):
    """Gram-Schmidt of rows 0 and 1, row 2 = conj(row0 x row1)."""
    r0 = U[0] / np.linalg.norm(U[0])
    r1 = U[1] - np.vdot(r0, U[1]) * r0
    r1 = r1 / np.linalg.norm(r1)
    r2 = np.conj(np.cross(r0, r1))
    return np.array([r0, r1, r2])


def chart(U, A):
    W = frame(A)
    r = su3_log(U @ W.conj().T)
    if r is None:
        return W, None, None
    X, ph = r
    return W, coords(X), ph


def haar_log_jacobian(phases):
    x = np.asarray(phases, np.float64)
    acc = 0.0
    for j in range(3):
        for k in range(j + 1, 3):
            d = 0.5 * (x[j] - x[k])
            acc += 2 * np.log(np.sinc(d / np.pi))
    return acc


def numerical_log_jacobian(eps, h=1e-6):
    """log |det| of the 8x8 map d eps -> coordinates of exp(iX)^dagger d exp(iX),
    by central differences, relative to the same at eps = 0."""
    U0 = expi(from_coords(eps))
    J = np.zeros((8, 8))
    for a in range(8):
        e1 = np.array(eps, np.float64)
        e2 = np.array(eps, np.float64)
        e1[a] += h
        e2[a] -= h
        dU = (expi(from_coords(e1)) - expi(from_coords(e2))) / (2 * h)
        Om = U0.conj().T @ dU  # i * Hermitian
        J[:, a] = coords(-1j * Om)
    return np.log(abs(np.linalg.det(J)))
