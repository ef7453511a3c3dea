"""The learned return map's matrix-core kernels against the FP64 MLP oracle
(oracle/mlp_oracle.py) and the build-defined 3D Mohr-Coulomb return.

Tolerances: BF16 operands carry 8 significant bits (relative rounding 2^-9),
so the forward pass is compared with an oracle that rounds the same operands
(inputs, weights, activations) to BF16 at 5e-3 relative L2 (FP32 summation
order can flip single activations by one BF16 ulp), and with the unrounded
FP64 network at 5e-2.  Gradients use BF16 deltas as well: relative L2 per
layer <= 6e-2 and cosine >= 0.998 against the FP64 gradient.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mlo():
    import mlp_oracle
    return mlp_oracle


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, float) - b) / max(np.linalg.norm(b), 1e-300))


def _random_params(mpm, seed):
    S = mpm.surrogate
    rng = np.random.default_rng(seed)
    p = S.init_params(seed).astype(np.float64)
    o = 0
    for name, shape in S.SHAPES:
        k = int(np.prod(shape))
        if name.startswith("b"):
            p[o:o + k] = rng.normal(0, 0.2, k)
        o += k
    p[o:o + 16] = rng.normal(0, 1, 16)
    p[o + 16:o + 32] = rng.uniform(0.5, 2, 16)
    p[o + 32:] = rng.uniform(0.5, 2, 7)
    return p.astype(np.float32)


def test_forward_matches_oracle(mpm, mlo):
    S = mpm.Surrogate(batch_capacity=1024)
    try:
        for seed, n in ((3, 1000), (4, 64), (5, 4133)):
            p = _random_params(mpm, seed)
            S.set_params(p)
            rng = np.random.default_rng(seed + 10)
            q = mlo.unpack(p)
            X = (q["x_mean"] + rng.normal(0, 1, (n, 16)) / q["x_istd"]).astype(np.float32)
            y = S.forward(X)
            ref_r = mlo.forward(p.astype(np.float64), X.astype(np.float64), rounding=True)
            ref = mlo.forward(p.astype(np.float64), X.astype(np.float64))
            print(f"n={n}: rel L2 vs BF16-rounded oracle {_rel(y, ref_r):.2e}, vs FP64 {_rel(y, ref):.2e}",
                  flush=True)
            assert np.all(np.isfinite(y))
            assert _rel(y, ref_r) <= 5e-3
            assert _rel(y, ref) <= 5e-2
    finally:
        S.close()


def test_gradient_matches_oracle(mpm, mlo):
    S = mpm.Surrogate(batch_capacity=512)
    try:
        p = _random_params(mpm, 7)
        S.set_params(p)
        rng = np.random.default_rng(8)
        q = mlo.unpack(p)
        n = 300
        X = (q["x_mean"] + rng.normal(0, 1, (n, 16)) / q["x_istd"]).astype(np.float32)
        Y = (rng.normal(0, 1, (n, 7)) * q["y_scale"]).astype(np.float32)
        loss, g = S.grad(X, Y)
        lref, gref = mlo.loss_and_grad(p.astype(np.float64), X.astype(np.float64), Y.astype(np.float64))
        print(f"loss {loss:.6f} vs {lref:.6f}", flush=True)
        assert loss == pytest.approx(lref, rel=2e-2)
        o = 0
        for name, shape in mlo.SHAPES:
            k = int(np.prod(shape))
            a, b = g[o:o + k].astype(np.float64), gref[o:o + k]
            cos = float(a @ b / (np.linalg.norm(a) * np.linalg.norm(b) + 1e-300))
            print(f"  {name}: rel L2 {_rel(a, b):.2e} cos {cos:.6f}", flush=True)
            assert _rel(a, b) <= 6e-2 and cos >= 0.998, name
            o += k
    finally:
        S.close()


def _mc_samples(orc, n, seed):
    """Return-mapping samples of the oracle's 3D Mohr-Coulomb over random trial
    stresses and materials (plastic states only)."""
    rng = np.random.default_rng(seed)
    K, G = 7.0e5, 3.2308e5
    X, Y = [], []
    while len(X) < n:
        s = rng.normal(-2e4, 1.2e4, 6)
        s[3:] = rng.normal(0, 6e3, 3)
        de = rng.normal(0, 2e-4, 6)
        phi = math.radians(rng.uniform(20, 40))
        c = float(np.exp(rng.normal(math.log(1e3), 0.5)))
        psi = 0.0
        eq = float(abs(rng.normal(0, 0.02)))
        out, eq2, ek, kind = orc.mc3_point(s, eq, K, G, phi, psi, c)
        if kind == 0:
            continue
        X.append(np.concatenate([s, de, [eq, phi, c, psi]]))
        Y.append(np.concatenate([out - s, [eq2 - eq]]))
    return np.array(X, np.float32), np.array(Y, np.float32)


def test_training_learns_the_return_map(mpm, orc):
    X, Y = _mc_samples(orc, 24000, 1)
    Xt, Yt = _mc_samples(orc, 3000, 2)
    S = mpm.Surrogate(batch_capacity=2048, seed=0)
    try:
        S.fit_normalisation(X, Y)
        err0 = _rel(S.forward(Xt), Yt)
        hist = S.train(X, Y, steps=1500, batch=1024, lr=2e-3, seed=3)
        err1 = _rel(S.forward(Xt), Yt)
        print(f"held-out rel L2 of the correction: untrained {err0:.3f} -> trained {err1:.3f}; "
              f"loss {hist[:20].mean():.4f} -> {hist[-20:].mean():.4f}", flush=True)
        assert np.all(np.isfinite(hist))
        assert hist[-50:].mean() < 0.2 * hist[:50].mean()
        assert err1 < 0.35 and err1 < 0.4 * err0
    finally:
        S.close()


def _mc_column(mpm, h, eps="FP32", nsteps=400):
    return mpm.scenes.column_collapse_3d(h=h, width=0.16, height=0.08, box_xy=0.4, box_z=0.12,
                                         nsteps=nsteps, eps=eps, constitutive="mohrcoulomb",
                                         cohesion=200.0, psi_deg=0.0)


def test_return_log_matches_exact_return(mpm, orc):
    args, grid, mp, attr, bc = _mc_column(mpm, 0.02)
    m = mpm.HIPModel(grid, mp, attr, bc, sort_interval=10)
    try:
        m.return_log(200000, 1)
        m.run(m.step_args(args), float(args.ΔT), 0.0, 60)
        X, Y = m.read_return_log()
    finally:
        m.close()
    print(f"logged {len(X)} samples", flush=True)
    assert len(X) > 1000
    K, G = float(attr.Ks[0]), float(attr.Gs[0])
    rng = np.random.default_rng(0)
    worst = 0.0
    for r in rng.choice(len(X), 300, replace=False):
        x = X[r].astype(np.float64)
        out, eq1, ek, kind = orc.mc3_point(x[:6], x[12], K, G, x[13], x[15], x[14])
        assert kind != 0
        scale = np.abs(x[:6]).max() + x[14]
        worst = max(worst, np.abs(Y[r, :6] - (out - x[:6])).max() / scale)
        assert abs(Y[r, 6] - (eq1 - x[12])) <= 2e-3 * abs(eq1 - x[12]) + 1e-9
    print(f"worst stress-correction difference / |σ|: {worst:.2e}", flush=True)
    assert worst <= 1e-4


def test_learned_return_in_the_step(mpm):
    # plain run with logging -> training -> run with the learned return
    nsteps = 300
    args, grid, mp, attr, bc = _mc_column(mpm, 0.01, nsteps=nsteps)
    sa_dt = float(args.ΔT)
    x0 = np.asarray(mp.ξ, dtype=np.float64)
    plain = mpm.HIPModel(grid, mp, attr, bc, sort_interval=10)
    try:
        plain.return_log(2_000_000, 1)
        plain.run(plain.step_args(args), sa_dt, 0.0, nsteps)
        X, Y = plain.read_return_log()
        ref = plain.download()
    finally:
        plain.close()
    print(f"{mp.np} particles, {len(X)} return samples", flush=True)
    S = mpm.Surrogate(batch_capacity=4096, seed=0)
    U = mpm.Surrogate(batch_capacity=64, seed=0)
    try:
        S.fit_normalisation(X, Y)
        U.set_params(S.params())  # same normalisation, untrained weights
        hist = S.train(X, Y, steps=3000, batch=4096, lr=2e-3, seed=5)
        print(f"training loss {hist[:20].mean():.4f} -> {hist[-20:].mean():.5f}", flush=True)
        results = {}
        for name, net in (("trained", S), ("untrained", U)):
            m = mpm.HIPModel(grid, mp, attr, bc, sort_interval=10)
            try:
                m.set_surrogate(net, tol=1e-3)
                m.run(m.step_args(args), sa_dt, 0.0, nsteps)
                got = m.download()
                plastic, fb = m.surrogate_stats()
            finally:
                m.close()
            run_ref = mpm.scenes.runout(ref["xi"], x0)
            run_got = mpm.scenes.runout(got["xi"], x0)
            d = _rel(got["xi"] - x0, ref["xi"] - x0)
            results[name] = (plastic, fb, run_got / run_ref - 1.0, d)
            print(f"{name}: yielding particle-steps {plastic}, fallbacks {fb} "
                  f"({fb / max(plastic, 1):.3f}), runout vs exact {run_got / run_ref - 1:+.2e}, "
                  f"displacement rel L2 {d:.2e}", flush=True)
    finally:
        S.close()
        U.close()
    tr, un = results["trained"], results["untrained"]
    assert tr[0] > 0 and un[0] > 0
    # what the trained network gains: it serves yielding particles the untrained one cannot
    assert tr[1] / tr[0] < un[1] / un[0]
    # and the run stays on the exact solution at the runout tolerance of FP32 mode
    assert abs(tr[2]) <= 5e-3
