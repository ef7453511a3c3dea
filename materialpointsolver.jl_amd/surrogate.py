"""The learned return map: host-side handle of libmpm_hip's MLP (mpm_mlp_* in
include/mpm_hip.h).  No reference counterpart: MaterialPointSolver.jl has no
surrogate constitutive model.

``Surrogate`` owns one network: initialisation (seeded), input / output
normalisation from a sample set, training on the GPU (the library's own
forward and backward matrix-core kernels with Adam), inference, gradients,
saving and loading (``.npz``).

Sample layout (row-major FP32): inputs (n, 16) = the trial stress as its
principal values p1 >= p2 >= p3 over the scale S = max|p| + c cos(phi) and its
principal frame as a unit quaternion (x, y, z; w >= 0), the strain increment
[xx,yy,zz,xy,yz,xz] times 2 G / S, eps_q, friction angle, c / S, dilation
angle (``mc_features``); outputs (n, 7) = scores of the four branches of the
principal-space Mohr-Coulomb return (main plane, edge s1 = s2, edge s2 = s3,
apex; training targets one-hot) and three zeros.
"""
from __future__ import annotations

import numpy as np

from . import _capi as capi

NIN, NH, NOUT = 16, 128, 7
SHAPES = [("W1", (NH, NIN)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)),
          ("W3", (NH, NH)), ("b3", (NH,)), ("W4", (NOUT, NH)), ("b4", (NOUT,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)
NTOTAL = NPARAM + NIN + NIN + NOUT


BRANCHES = ("main plane", "edge s1 = s2", "edge s2 = s3", "apex")


def _quat_xyz(V):
    """Unit quaternion (x, y, z; w >= 0) of rotation matrices V (n, 3, 3)."""
    n = V.shape[0]
    out = np.zeros((n, 4))
    tr = V[:, 0, 0] + V[:, 1, 1] + V[:, 2, 2]
    for k in range(n):
        m = V[k]
        if tr[k] > 0:
            s = np.sqrt(tr[k] + 1.0) * 2
            q = (0.25 * s, (m[2, 1] - m[1, 2]) / s, (m[0, 2] - m[2, 0]) / s, (m[1, 0] - m[0, 1]) / s)
        elif m[0, 0] > m[1, 1] and m[0, 0] > m[2, 2]:
            s = np.sqrt(1.0 + m[0, 0] - m[1, 1] - m[2, 2]) * 2
            q = ((m[2, 1] - m[1, 2]) / s, 0.25 * s, (m[0, 1] + m[1, 0]) / s, (m[0, 2] + m[2, 0]) / s)
        elif m[1, 1] > m[2, 2]:
            s = np.sqrt(1.0 + m[1, 1] - m[0, 0] - m[2, 2]) * 2
            q = ((m[0, 2] - m[2, 0]) / s, (m[0, 1] + m[1, 0]) / s, 0.25 * s, (m[1, 2] + m[2, 1]) / s)
        else:
            s = np.sqrt(1.0 + m[2, 2] - m[0, 0] - m[1, 1]) * 2
            q = ((m[1, 0] - m[0, 1]) / s, (m[0, 2] + m[2, 0]) / s, (m[1, 2] + m[2, 1]) / s, 0.25 * s)
        out[k] = q
    out[out[:, 0] < 0] *= -1.0
    return out[:, 1:]


def mc_features(sig, de, eq, phi, c, psi, Gs):
    """The learned return's 16 inputs (as the device computes them) for trial
    stresses sig (n, 6, [xx,yy,zz,xy,yz,xz]), strain increments de (n, 6),
    plastic strains eq (n,) and the strength in effect (phi, c, psi) and shear
    modulus Gs (scalars or (n,)).  Returns (X (n, 16), p (n, 3) principal
    stresses, descending)."""
    sig = np.atleast_2d(np.asarray(sig, dtype=np.float64))
    n = sig.shape[0]
    T = np.zeros((n, 3, 3))
    T[:, 0, 0], T[:, 1, 1], T[:, 2, 2] = sig[:, 0], sig[:, 1], sig[:, 2]
    T[:, 0, 1] = T[:, 1, 0] = sig[:, 3]
    T[:, 1, 2] = T[:, 2, 1] = sig[:, 4]
    T[:, 0, 2] = T[:, 2, 0] = sig[:, 5]
    w, V = np.linalg.eigh(T)
    p = w[:, ::-1]
    V = V[:, :, ::-1].copy()
    V[np.linalg.det(V) < 0, :, 2] *= -1.0
    q3 = _quat_xyz(V)
    phi = np.broadcast_to(np.asarray(phi, dtype=np.float64), (n,))
    c = np.broadcast_to(np.asarray(c, dtype=np.float64), (n,))
    psi = np.broadcast_to(np.asarray(psi, dtype=np.float64), (n,))
    Gs = np.broadcast_to(np.asarray(Gs, dtype=np.float64), (n,))
    S = np.maximum(np.abs(p[:, 0]), np.abs(p[:, 2])) + c * np.cos(phi) + 1e-30
    X = np.concatenate([p / S[:, None], q3, np.asarray(de, dtype=np.float64) * (2 * Gs / S)[:, None],
                        np.asarray(eq, dtype=np.float64)[:, None] * np.ones((n, 1)),
                        phi[:, None], (c / S)[:, None], psi[:, None]], axis=1)
    return X.astype(np.float32), p


def init_params(seed=0):
    """He-style uniform initialisation of the weights, zero biases, identity
    normalisation."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, shape in SHAPES:
        if name.startswith("W"):
            fan_in = shape[1]
            lim = np.sqrt(6.0 / fan_in) * (0.5 if name == "W4" else 1.0)
            parts.append(rng.uniform(-lim, lim, size=shape).ravel())
        else:
            parts.append(np.zeros(shape).ravel())
    parts += [np.zeros(NIN), np.ones(NIN), np.ones(NOUT)]
    return np.concatenate(parts).astype(np.float32)


class Surrogate:
    def __init__(self, batch_capacity=8192, seed=0, params=None):
        lib = capi.load()
        if lib.mpm_mlp_param_count() != NTOTAL:
            raise ImportError("libmpm_hip.so MLP layout mismatch")
        self.lib = lib
        h = capi.C.c_void_p()
        rc = lib.mpm_mlp_create(int(batch_capacity), capi.C.byref(h))
        if rc != capi.MPM_OK:
            raise capi.MPMError(rc, "mpm_mlp_create failed")
        self.h = h
        self.capacity = int(batch_capacity)
        self.set_params(init_params(seed) if params is None else params)

    # -- parameters ---------------------------------------------------------
    def _check(self, rc):
        if rc != capi.MPM_OK:
            msg = self.lib.mpm_mlp_last_error(self.h)
            raise capi.MPMError(rc, msg.decode() if msg else "")

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != NTOTAL:
            raise ValueError(f"expected {NTOTAL} parameters")
        self._check(self.lib.mpm_mlp_set_params(self.h, capi.fptr(p), NTOTAL))

    def params(self):
        p = np.zeros(NTOTAL, dtype=np.float32)
        self._check(self.lib.mpm_mlp_get_params(self.h, capi.fptr(p), NTOTAL))
        return p

    def fit_normalisation(self, X, Y):
        """Standardise inputs and scale outputs from a sample set (keeps weights).
        Sets of more than 4M rows are summarised from an evenly strided subset
        of about 4M rows."""
        step = max(1, len(X) // 4_000_000)
        X = np.asarray(X[::step], dtype=np.float64)
        Y = np.asarray(Y[::step], dtype=np.float64)
        p = self.params()
        mean = X.mean(0)
        std = X.std(0)
        std[std < 1e-12 * (1 + np.abs(mean))] = 1.0
        scale = np.sqrt((Y ** 2).mean(0))
        scale[scale <= 0] = 1.0
        p[NPARAM:NPARAM + NIN] = mean
        p[NPARAM + NIN:NPARAM + 2 * NIN] = 1.0 / std
        p[NPARAM + 2 * NIN:] = scale
        self.set_params(p)

    # -- compute -------------------------------------------------------------
    def forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        n = X.shape[0]
        Y = np.zeros((n, NOUT), dtype=np.float32)
        self._check(self.lib.mpm_mlp_forward(self.h, capi.fptr(X), n, capi.fptr(Y)))
        return Y

    def grad(self, X, Y):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.ascontiguousarray(Y, dtype=np.float32)
        g = np.zeros(NPARAM, dtype=np.float32)
        loss = capi.C.c_float(0.0)
        self._check(self.lib.mpm_mlp_grad(self.h, capi.fptr(X), capi.fptr(Y), X.shape[0],
                                          capi.fptr(g), capi.C.byref(loss)))
        return float(loss.value), g

    def train(self, X, Y, steps=2000, batch=4096, lr=1e-3, seed=1):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.ascontiguousarray(Y, dtype=np.float32)
        hist = np.zeros(max(int(steps), 1), dtype=np.float32)
        self._check(self.lib.mpm_mlp_train(self.h, capi.fptr(X), capi.fptr(Y), X.shape[0],
                                           int(steps), int(batch), float(lr), int(seed),
                                           capi.fptr(hist)))
        return hist[:steps]

    def bench(self, n, reps=20):
        ms = capi.C.c_double(0.0)
        self._check(self.lib.mpm_mlp_bench(self.h, int(n), int(reps), capi.C.byref(ms)))
        return ms.value

    # -- persistence ----------------------------------------------------------
    def save(self, path):
        np.savez(path, params=self.params())

    @classmethod
    def load(cls, path, batch_capacity=8192):
        with np.load(path, allow_pickle=False) as f:
            return cls(batch_capacity=batch_capacity, params=f["params"])

    def close(self):
        if getattr(self, "h", None):
            self.lib.mpm_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # flops of one forward evaluation (multiply-adds x 2)
    FLOPS_PER_SAMPLE = 2 * (NH * NIN + NH * NH + NH * NH + NOUT * NH)
