"""The learned return map's matrix-core kernels against the FP64 MLP oracle
(oracle/mlp_oracle.py) and the build-defined 3D Mohr-Coulomb return.  The
network scores the branches of the principal-space return; the step takes the
best-scoring branch where its conditions hold (else the exact return).

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


def _mc_samples(mpm, orc, n, seed):
    """Learned-return samples over random trial stresses and materials (plastic
    states only): the 16 inputs (mc_features) and the oracle's return branch as
    a one-hot over the first four outputs."""
    rng = np.random.default_rng(seed)
    K, G = 7.0e5, 3.2308e5
    S, D, E, kinds, mats = [], [], [], [], []
    while len(S) < n:
        s = rng.normal(-2e4, 1.2e4, 6)
        s[3:] = rng.normal(0, 6e3, 3)
        phi = math.radians(rng.uniform(20, 40))
        psi = math.radians(rng.uniform(0, 8))
        c = float(np.exp(rng.normal(math.log(1e3), 0.5)))
        eq = float(abs(rng.normal(0, 0.02)))
        kind = orc.mc3_point(s, eq, K, G, phi, psi, c)[3]
        if kind == 0:
            continue
        S.append(s)
        D.append(rng.normal(0, 2e-4, 6))
        E.append(eq)
        mats.append((phi, c, psi))
        kinds.append(kind)
    mats = np.array(mats)
    X, _ = mpm.surrogate.mc_features(np.array(S), np.array(D), np.array(E), mats[:, 0], mats[:, 1],
                                     mats[:, 2], G)
    Y = np.zeros((n, 7), np.float32)
    Y[np.arange(n), np.array(kinds) - 1] = 1.0
    return X, Y


def _accuracy(pred, Y):
    return float(np.mean(np.argmax(pred[:, :4], 1) == np.argmax(Y[:, :4], 1)))


def test_training_learns_the_return_branch(mpm, orc):
    X, Y = _mc_samples(mpm, orc, 24000, 1)
    Xt, Yt = _mc_samples(mpm, orc, 3000, 2)
    freq = Y[:, :4].mean(0)
    print("branch frequencies (main, edge 12, edge 23, apex):", np.round(freq, 3), flush=True)
    S = mpm.Surrogate(batch_capacity=2048, seed=0)
    try:
        S.fit_normalisation(X, Y)
        acc0 = _accuracy(S.forward(Xt), Yt)
        hist = S.train(X, Y, steps=1500, batch=1024, lr=2e-3, seed=3)
        acc1 = _accuracy(S.forward(Xt), Yt)
        print(f"held-out branch accuracy: untrained {acc0:.3f} -> trained {acc1:.3f}; "
              f"loss {hist[:20].mean():.4f} -> {hist[-20:].mean():.4f}", flush=True)
        assert np.all(np.isfinite(hist))
        assert hist[-50:].mean() < 0.5 * hist[:50].mean()
        # better than always answering the most frequent branch
        assert acc1 > 0.9 and acc1 > freq.max() + 0.02
    finally:
        S.close()


def _mc_column(mpm, h, eps="FP32", nsteps=400):
    return mpm.scenes.column_collapse_3d(h=h, width=0.16, height=0.08, box_xy=0.4, box_z=0.12,
                                         nsteps=nsteps, eps=eps, constitutive="mohrcoulomb",
                                         cohesion=200.0, psi_deg=0.0)


def test_return_log_matches_exact_return(mpm, orc):
    # a logged sample's inputs hold the trial stress's principal values and c
    # in units of S; the principal-space return is homogeneous in (p, c), so
    # the oracle's return on them takes the logged branch
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
    Dt, Dd = K + 1.333333 * G, K - 0.666667 * G
    assert np.all(Y[:, 4:] == 0) and np.all(Y[:, :4].sum(1) == 1)
    rng = np.random.default_rng(0)
    agree = 0
    rows = rng.choice(len(X), 300, replace=False)
    for r in rows:
        x = X[r].astype(np.float64)
        assert np.linalg.norm(x[3:6]) <= 1.0 + 1e-6  # a unit quaternion's vector part
        phi, c, psi = x[13], x[14], x[15]
        p, kind = orc.mc3_principal(x[:3], Dt, Dd, math.sin(phi), math.cos(phi), math.sin(psi), c)
        agree += kind == int(np.argmax(Y[r, :4])) + 1
    print(f"logged branch = oracle branch in {agree} of {len(rows)}", flush=True)
    # FP32 features carry ~1e-7 of rounding: a state on a branch boundary may flip
    assert agree >= len(rows) - 3


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
    assert tr[1] / tr[0] < 0.5
    # and the run stays on the exact solution at the runout tolerance of FP32 mode
    assert abs(tr[2]) <= 5e-3
