"""Learned SU(3) link proposal (include/ym_amd.h, ym_mlp_* and
ym_su3_learned_sweeps).

The plain reference sampler is single-generator Metropolis
(su3_ensemble.rs:336-391).  ``Mlp`` holds a 12 -> 64 -> 64 -> 16 tanh network
on the device together with a dataset of (staple features, chart
coordinates) pairs collected from plain chains; ``learned_sweeps`` runs
checkerboard sweeps in which every link gets one independence proposal
exp(i X(eps)) W(A), eps ~ N(mu(A), sigma(A)^2), accepted with the exact
Metropolis-Hastings ratio, followed by plain Metropolis hits.
"""
import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import check, lib

N_PARAMS = 6032
N_FEATURES = 12
N_OUT = 16


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Mlp:
    def __init__(self, seed=1, dataset_capacity=0):
        h = C.c_void_p()
        check(lib().ym_mlp_create(int(seed), int(dataset_capacity), C.byref(h)), "ym_mlp_create")
        self._h = h

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None):
            lib().ym_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self):
        p = np.zeros(N_PARAMS, dtype=np.float32)
        check(lib().ym_mlp_get_params(self._h, _fp(p)), "ym_mlp_get_params")
        return p

    def set_params(self, p):
        a = np.ascontiguousarray(np.asarray(p, dtype=np.float32).reshape(-1))
        if a.size != N_PARAMS:
            raise ValueError("6032 parameters")
        check(lib().ym_mlp_set_params(self._h, _fp(a)), "ym_mlp_set_params")

    def dataset_size(self):
        n, cap = C.c_int64(), C.c_int64()
        check(lib().ym_mlp_dataset_size(self._h, C.byref(n), C.byref(cap)), "ym_mlp_dataset_size")
        return n.value, cap.value

    def clear_pairs(self):
        check(lib().ym_mlp_clear_pairs(self._h), "ym_mlp_clear_pairs")

    def collect_pairs(self, lattice):
        t = C.c_int64()
        check(lib().ym_mlp_collect_pairs(self._h, lattice.handle, C.byref(t)), "ym_mlp_collect_pairs")
        return t.value

    def download_pairs(self, first, count):
        f = np.zeros((count, N_FEATURES), dtype=np.float32)
        t = np.zeros((count, 8), dtype=np.float32)
        v = np.zeros(count, dtype=np.int32)
        check(lib().ym_mlp_download_pairs(self._h, int(first), int(count), _fp(f), _fp(t),
                                          v.ctypes.data_as(C.POINTER(C.c_int32))), "ym_mlp_download_pairs")
        return f, t, v

    def upload_pairs(self, features, targets, valid):
        f = np.ascontiguousarray(np.asarray(features, dtype=np.float32))
        t = np.ascontiguousarray(np.asarray(targets, dtype=np.float32))
        v = np.ascontiguousarray(np.asarray(valid, dtype=np.int32))
        check(lib().ym_mlp_upload_pairs(self._h, _fp(f), _fp(t), v.ctypes.data_as(C.POINTER(C.c_int32)), v.size),
              "ym_mlp_upload_pairs")

    def forward(self, features, bf16=False):
        f = np.ascontiguousarray(np.asarray(features, dtype=np.float32).reshape(-1, N_FEATURES))
        out = np.zeros((f.shape[0], N_OUT), dtype=np.float32)
        check(lib().ym_mlp_forward(self._h, _fp(f), f.shape[0], 1 if bf16 else 0, _fp(out)), "ym_mlp_forward")
        return out

    def loss_grad(self, first, count):
        loss = C.c_float()
        g = np.zeros(N_PARAMS, dtype=np.float32)
        check(lib().ym_mlp_loss_grad(self._h, int(first), int(count), C.byref(loss), _fp(g)), "ym_mlp_loss_grad")
        return loss.value, g

    def train(self, epochs, batch=8192, lr=1e-3, seed=7):
        losses = np.zeros(epochs, dtype=np.float32)
        ms = C.c_float()
        check(lib().ym_mlp_train(self._h, int(epochs), int(batch), float(lr), int(seed), _fp(losses), C.byref(ms)),
              "ym_mlp_train")
        return losses, ms.value


@dataclass
class LearnedCounts:
    learned_proposals: int
    learned_accepts: int
    metro_accepts: int
    device_ms: float


def learned_sweeps(lattice, mlp, beta, key, first_sweep, n_sweeps, metro_hits=1, angle=0.35, bf16=False):
    lp, la_, ma = C.c_uint64(0), C.c_uint64(0), C.c_uint64(0)
    ms = C.c_float()
    check(lib().ym_su3_learned_sweeps(lattice.handle, mlp.handle, float(beta), int(key), int(first_sweep),
                                      int(n_sweeps), int(metro_hits), float(angle), 1 if bf16 else 0, C.byref(lp),
                                      C.byref(la_), C.byref(ma), C.byref(ms)), "ym_su3_learned_sweeps")
    return LearnedCounts(lp.value, la_.value, ma.value, ms.value)


@dataclass
class LinkChart:
    staple: np.ndarray     # (3, 3) complex
    frame: np.ndarray      # (3, 3) complex, W(A)
    eps: np.ndarray        # (8,)
    phases: np.ndarray     # (3,)
    features: np.ndarray   # (12,)
    frame_ok: bool
    chart_ok: bool
    log_jacobian: float


def _c3(v):
    return (v[0::2] + 1j * v[1::2]).reshape(3, 3)


def link_chart(lattice, x, mu):
    out = np.zeros(62)
    xs = (C.c_int64 * 4)(*[int(v) for v in x])
    check(lib().ym_su3_link_chart(lattice.handle, xs, int(mu), out.ctypes.data_as(C.POINTER(C.c_double))),
          "ym_su3_link_chart")
    return LinkChart(_c3(out[0:18]), _c3(out[18:36]), out[36:44].copy(), out[44:47].copy(), out[47:59].copy(),
                     bool(out[59]), bool(out[60]), float(out[61]))


@dataclass
class MhStep:
    """One learned MH step as the sweep kernel evaluates it (ym_su3_learned_probe)."""
    eps: np.ndarray            # chart coordinates of the link
    mu: np.ndarray             # network means
    sigma: np.ndarray          # network standard deviations
    eps_proposal: np.ndarray   # proposed coordinates
    log_q_current: float
    log_q_proposal: float
    action_term: float         # (beta/3) Re Tr((U' - U) A)
    log_ratio: float
    accept_uniform: float
    in_chart: bool
    accepted: bool
    chart_ok: bool
    proposal: np.ndarray       # (3, 3) complex U'


def learned_probe(lattice, mlp, beta, key, sweep, x, mu):
    out = np.zeros(58)
    xs = (C.c_int64 * 4)(*[int(v) for v in x])
    check(lib().ym_su3_learned_probe(lattice.handle, mlp.handle, float(beta), int(key), int(sweep), xs, int(mu),
                                     out.ctypes.data_as(C.POINTER(C.c_double))), "ym_su3_learned_probe")
    return MhStep(out[0:8].copy(), out[8:16].copy(), out[16:24].copy(), out[24:32].copy(), float(out[32]),
                  float(out[33]), float(out[34]), float(out[35]), float(out[36]), bool(out[37]), bool(out[38]),
                  bool(out[39]), _c3(out[40:58]))
