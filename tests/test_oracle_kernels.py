"""Oracle pins for each restated kernel of the step: conservation laws,
Newton's law on the grid, affine-field reproduction, elastic closed forms and
plastic-return residuals.  Each check fails for a plausible slip (dropped
term, wrong sign, wrong index or transposed operand)."""
import math

import numpy as np
import pytest


def _block(mpm, dim, n=3, h=0.1, seed=0, constitutive="druckerprager", **kw):
    """A jittered block of particles in the middle of a small padded grid."""
    rng = np.random.default_rng(seed)
    if dim == 2:
        args, grid, mp, attr, bc = mpm.scenes.column_collapse_2d(
            h=h, width=n * h, height=n * h, box_x=10 * h, box_y=10 * h, pad=2, dt=1e-4,
            nsteps=1, constitutive=constitutive, **kw)
    else:
        args, grid, mp, attr, bc = mpm.scenes.column_collapse_3d(
            h=h, width=n * h, height=n * h, box_xy=10 * h, box_z=10 * h, pad=2, eps="FP64",
            nsteps=1, constitutive=constitutive, **kw)
    shift = np.full(dim, 3.0 * h)
    if dim == 3:
        shift[:2] = 0.0
    mp.ξ[...] = mp.ξ + shift + rng.uniform(-0.2 * h, 0.2 * h, size=mp.ξ.shape)
    return args, grid, mp, attr, bc


def _rand_state(o, seed=1):
    rng = np.random.default_rng(seed)
    a = o.a
    a["vs"][...] = rng.normal(size=a["vs"].shape)
    a["sig"][...] = rng.normal(scale=1e3, size=a["sig"].shape)
    return rng


@pytest.mark.parametrize("dim", [2, 3])
def test_p2g_conservation(mpm, orc, dim):
    o = orc.Oracle(*_block(mpm, dim))
    _rand_state(o)
    G = -9.8
    o.resetgrid()
    o.resetmp()
    o.p2g(G)
    a = o.a
    m_tot = a["ms"].sum()
    assert abs(a["g_ms"].sum() - m_tot) < 1e-13 * m_tot
    assert np.allclose(a["g_ps"].sum(axis=0), a["ps"].sum(axis=0), rtol=1e-12, atol=1e-12)
    f = a["g_fs"].sum(axis=0)
    # internal forces sum to zero (sum of basis gradients is zero); gravity on the last axis
    assert np.allclose(f[:-1], 0.0, atol=1e-9)
    assert abs(f[-1] - G * m_tot) < 1e-9
    # first moment: sum_n m_n x_n = sum_p m_p x_p (linear reproduction)
    assert np.allclose(a["g_ms"] @ a["gxi"], a["ms"] @ a["xi"], rtol=1e-12)


@pytest.mark.parametrize("dim", [2, 3])
def test_p2g_body_force_extension(mpm, orc, dim):
    # the body force on the other axes (extension; zero reproduces the reference)
    # adds m_p g_a to axis a only, ramped with Ti / Te like gravity
    o = orc.Oracle(*_block(mpm, dim))
    _rand_state(o)
    o.m.gravity_x, o.m.gravity_y, o.m.Te = 2.5, (-1.5 if dim == 3 else 0.0), 0.4
    for Ti, fac in ((0.1, 0.25), (0.4, 1.0)):
        o.m.Ti_now = Ti
        o.resetgrid()
        o.resetmp()
        o.p2g(-9.8)
        f = o.a["g_fs"].sum(axis=0)
        m_tot = o.a["ms"].sum()
        assert abs(f[0] - 2.5 * fac * m_tot) < 1e-9
        if dim == 3:
            assert abs(f[1] + 1.5 * fac * m_tot) < 1e-9
        assert abs(f[-1] + 9.8 * m_tot) < 1e-9


@pytest.mark.parametrize("dim", [2, 3])
def test_p2g_internal_force_is_divergence(mpm, orc, dim):
    # a uniform stress field exerts no net force on interior nodes and pulls the
    # block's boundary nodes outward / inward with total moment sum_n x_n f_n = -V σ
    o = orc.Oracle(*_block(mpm, dim))
    a = o.a
    ns = 4 if dim == 2 else 6
    s = np.array([1.0, 2.0, 3.0, 0.5, 0.25, -0.75][:ns]) * 1e3
    a["sig"][...] = s
    a["vs"][...] = 0.0
    o.resetgrid()
    o.resetmp()
    o.p2g(0.0)
    M = a["gxi"].T @ a["g_fs"]  # sum_n x_n f_n^T  (dim x dim)
    V = a["omega"].sum()
    S = np.zeros((dim, dim))
    if dim == 2:
        S[0, 0], S[1, 1], S[0, 1] = s[0], s[1], s[3]
        S[1, 0] = s[3]
    else:
        S[0, 0], S[1, 1], S[2, 2] = s[0], s[1], s[2]
        S[0, 1] = S[1, 0] = s[3]
        S[1, 2] = S[2, 1] = s[4]
        S[0, 2] = S[2, 0] = s[5]
    assert np.allclose(M, -V * S, rtol=1e-10, atol=1e-10)


def test_solvegrid_newton_law_and_bc_quirk(mpm, orc):
    o = orc.Oracle(*_block(mpm, 2))
    a, m = o.a, o.m
    ni = o.ni
    rng = np.random.default_rng(4)
    dt = 1e-3
    for va in (0, 1):
        m.va = va
        m.zeta_s = 0.0
        a["g_ms"][...] = rng.uniform(0.5, 2.0, ni)
        a["g_ms"][:5] = 0.0                    # massless nodes
        a["g_ms"][5] = 1e-17                   # below eps(Float64): treated as massless
        P = rng.normal(size=(ni, 2))
        Fv = rng.normal(size=(ni, 2))
        a["g_ps"][...] = P
        a["g_fs"][...] = Fv
        a["vy_idx"][...] = 0
        bcx = a["vx_idx"]
        bcx[...] = 0
        bcx[10] = 1
        a["vx_val"][...] = 0.0
        a["vx_val"][10] = 0.5
        o.solvegrid(dt)
        with np.errstate(divide="ignore"):
            msd = np.where(a["g_ms"] < 2.220446049250313e-16, 0.0, 1.0 / a["g_ms"])
        free = np.ones(ni, bool)
        free[10] = False
        if va == 0:
            # Newton: vs = p/m, a = f/m, vsT = vs + a dt
            assert np.allclose(a["g_vs"][free], (P * msd[:, None])[free], rtol=1e-14)
            assert np.allclose(a["g_as"][free], (Fv * msd[:, None])[free], rtol=1e-14)
            assert np.allclose(a["g_vsT"][free], (P * msd[:, None] + Fv * msd[:, None] * dt)[free],
                               rtol=1e-13, atol=1e-15)
            # BC written into the momentum first: vs = val / m (utils_OS.jl:604-607)
            assert a["g_vs"][10, 0] == 0.5 * msd[10]
            assert a["g_vsT"][10, 0] == 0.5 and a["g_as"][10, 0] == 0.5
        else:
            assert np.allclose(a["g_vsT"][free], ((P + Fv * dt) * msd[:, None])[free], rtol=1e-13,
                               atol=1e-15)
            assert a["g_vs"][10, 0] == P[10, 0] * msd[10]
            assert a["g_vsT"][10, 0] == 0.5
        assert np.all(a["g_vs"][:6] == 0.0)
        assert np.all(a["g_ps"] == 0.0)        # momentum cleared for the MUSL re-map


def test_solvegrid_damping_dissipates(mpm, orc):
    o = orc.Oracle(*_block(mpm, 3))
    a, m = o.a, o.m
    ni = o.ni
    rng = np.random.default_rng(5)
    m.zeta_s = 0.3
    a["g_ms"][...] = 1.0
    V = rng.normal(size=(ni, 3))
    Fv = rng.normal(size=(ni, 3))
    a["g_ps"][...] = V
    a["g_fs"][...] = Fv
    a["vx_idx"][...] = 0
    a["vy_idx"][...] = 0
    a["vz_idx"][...] = 0
    o.solvegrid(1.0)
    Fd = a["g_as"]
    # local damping: each component is reduced by ζ|f| against the velocity sign
    want = Fv - 0.3 * np.linalg.norm(Fv, axis=1)[:, None] * np.sign(V)
    assert np.allclose(Fd, want, rtol=1e-14)
    # it never adds power: F_d . v <= f . v
    assert np.all(np.sum(Fd * V, axis=1) <= np.sum(Fv * V, axis=1) + 1e-14)


@pytest.mark.parametrize("dim", [2, 3])
def test_g2p_affine_field(mpm, orc, dim):
    o = orc.Oracle(*_block(mpm, dim))
    a = o.a
    rng = np.random.default_rng(6)
    A = rng.normal(scale=1e-3, size=(dim, dim))
    b = rng.normal(scale=1e-3, size=dim)
    F0 = np.eye(dim) + rng.normal(scale=1e-2, size=(dim, dim))
    nf = dim * dim
    a["F"][...] = F0.reshape(1, nf)            # row-major per particle
    o.initmp()
    a["g_dus"][...] = a["gxi"] @ A.T + b       # Δu(x) = A x + b
    dt = 0.01
    eps0 = a["eps_s"].copy()
    o.g2p(dt)
    # uGIMP reproduces linear fields: ΔF = ∇Δu = A exactly (to rounding)
    dF = a["dF"].reshape(-1, dim, dim)
    assert np.allclose(dF, A, atol=1e-15)
    Fn = np.einsum("ij,jk->ik", np.eye(dim) + A, F0)
    assert np.allclose(a["F"].reshape(-1, dim, dim), Fn, atol=1e-14)
    J = np.linalg.det(Fn)
    assert np.allclose(a["omega"], J * a["omega0"], rtol=1e-13)
    assert np.allclose(a["rho"], a["rho0"] / J, rtol=1e-13)
    S = 0.5 * (A + A.T)
    if dim == 2:
        want = [A[0, 0], A[1, 1], 0.0, A[0, 1] + A[1, 0]]
        rate = math.sqrt(S[0, 0] ** 2 + S[1, 1] ** 2 + 2 * S[0, 1] ** 2) / dt
    else:
        want = [A[0, 0], A[1, 1], A[2, 2], A[0, 1] + A[1, 0], A[1, 2] + A[2, 1], A[0, 2] + A[2, 0]]
        rate = math.sqrt(np.sum(S * S)) / dt
    assert np.allclose(a["deps_s"], want, atol=1e-15)
    assert np.allclose(a["eps_s"] - eps0, want, atol=1e-15)
    assert np.allclose(a["eps_v"], rate, rtol=1e-12)


@pytest.mark.parametrize("dim", [2, 3])
def test_remap_rigid_translation_and_momentum(mpm, orc, dim):
    o = orc.Oracle(*_block(mpm, dim))
    a, m = o.a, o.m
    v0 = np.array([0.3, -0.2, 0.1][:dim])
    a["vs"][...] = v0
    o.resetgrid()
    o.resetmp()
    x0 = a["xi"].copy()
    a["g_vsT"][...] = v0
    a["g_vs"][...] = v0
    a["g_as"][...] = 0.0
    dt = 1e-3
    for va in (0, 1):
        m.va = va
        a["xi"][...] = x0
        a["vs"][...] = v0
        o.dm1(dt)
        assert np.allclose(a["xi"], x0 + v0 * dt, atol=1e-15)
        assert np.allclose(a["vs"], v0, atol=1e-15)
    a["g_ps"][...] = 0.0
    o.dm2()
    assert np.allclose(a["g_ps"].sum(axis=0), a["ps"].sum(axis=0), rtol=1e-13)


def _mat_block(mpm, dim, **kw):
    sc = _block(mpm, dim, n=2, **kw)
    return sc


def test_liE_isotropic_closed_form(mpm, orc):
    for dim in (2, 3):
        args, grid, mp, attr, bc = _mat_block(mpm, dim, constitutive="linearelastic")
        o = orc.Oracle(args, grid, mp, attr, bc)
        a = o.a
        K, G = float(attr.Ks[0]), float(attr.Gs[0])
        rng = np.random.default_rng(dim)
        ns = 4 if dim == 2 else 6
        s0 = rng.normal(scale=100, size=ns)
        a["sig"][...] = s0
        e = rng.normal(scale=1e-4, size=ns)
        if dim == 2:
            e[2] = 0.0      # plane strain: Δϵzz is never written
        a["deps_s"][...] = e
        a["dF"][...] = 0.0  # no spin
        o.liE()
        tr = e[0] + e[1] + e[2]
        want = s0.copy()
        for k in range(3):
            want[k] += K * tr + 2 * G * (e[k] - tr / 3)
        for k in range(3, ns):
            want[k] += G * e[k]             # engineering shear strain
        # the reference's truncated constants (1.333333, 0.666667) -> 1e-6 agreement
        assert np.allclose(a["sig"], want, rtol=2e-6, atol=1e-9)
        sm = a["sig"][:, :3].sum(axis=1) * 0.333333
        assert np.allclose(a["sigm"], sm, rtol=1e-15)


def test_liE_spin_preserves_invariants(mpm, orc):
    args, grid, mp, attr, bc = _mat_block(mpm, 3, constitutive="linearelastic")
    o = orc.Oracle(args, grid, mp, attr, bc)
    a = o.a
    s0 = np.array([-300.0, -120.0, -50.0, 40.0, -25.0, 10.0])
    a["sig"][...] = s0
    a["deps_s"][...] = 0.0
    w = 1e-4
    W = np.array([[0, -w, 0.3 * w], [w, 0, -0.5 * w], [-0.3 * w, 0.5 * w, 0]])
    a["dF"][...] = W.reshape(1, 9)
    o.liE()

    def mat(s):
        return np.array([[s[0], s[3], s[5]], [s[3], s[1], s[4]], [s[5], s[4], s[2]]])

    S0, S1 = mat(s0), mat(a["sig"][0])
    assert abs(np.trace(S1) - np.trace(S0)) < 1e-9
    e0, e1 = np.linalg.eigvalsh(S0), np.linalg.eigvalsh(S1)
    assert np.allclose(e0, e1, atol=300 * 4 * w * w * 10)
    # first-order Jaumann rotation: S1 - S0 = ω S0 - S0 ω with ω = ½(ΔF^T - ΔF)
    # (ωxy = ½(∂Δu_y/∂x - ∂Δu_x/∂y), linearelastic.jl:90-92)
    om = 0.5 * (W.T - W)
    assert np.allclose(S1 - S0, om @ S0 - S0 @ om, atol=1e-10)
    assert not np.allclose(S1 - S0, -(om @ S0 - S0 @ om), atol=1e-6)


def _dp_invariants(s, dim):
    sm = (s[0] + s[1] + s[2]) * 0.333333
    d = np.array([s[0] - sm, s[1] - sm, s[2] - sm])
    q = 0.5 * np.sum(d * d) + np.sum(np.asarray(s[3:]) ** 2)
    return sm, math.sqrt(q)


@pytest.mark.parametrize("dim", [2, 3])
def test_dpP_return(mpm, orc, dim):
    phi, c = math.radians(30.0), 2000.0
    args, grid, mp, attr, bc = _mat_block(mpm, dim, phi_deg=30.0, cohesion=c, sigma_t=500.0)
    o = orc.Oracle(args, grid, mp, attr, bc)
    a = o.a
    K, G = float(attr.Ks[0]), float(attr.Gs[0])
    kphi = 6 * c * math.cos(phi) / (1.732051 * (3 + math.sin(phi)))
    qphi = 6 * math.sin(phi) / (1.732051 * (3 + math.sin(phi)))
    st = min(500.0, kphi / qphi)
    n = o.np
    ns = 4 if dim == 2 else 6
    rng = np.random.default_rng(9)
    for trial in range(20):
        s = rng.normal(scale=4000.0, size=(n, ns))
        s[:, :3] -= rng.uniform(0, 8000.0, size=(n, 1))
        a["sig"][...] = s
        a["dF"][...] = 0.0
        a["deps_s"][...] = 0.0
        o.liE()  # fills σm and sij from σ without changing it (zero increment)
        pre = a["sig"].copy()
        eq0 = a["eps_q"].copy()
        o.dpP()
        for i in range(n):
            sm0, t0 = _dp_invariants(pre[i], dim)
            sm1, t1 = _dp_invariants(a["sig"][i], dim)
            fs0 = t0 + qphi * sm0 - kphi
            ab = math.sqrt(1 + qphi * qphi) - qphi
            BF = (t0 - (kphi - qphi * st)) - ab * (sm0 - st)
            if (sm0 < st and fs0 > 0) or (sm0 >= st and BF > 0):
                # shear return onto the cone; ψ = 0 keeps the mean stress
                # (the reference's 0.333333 leaves a 1e-6-relative trace)
                scale = abs(sm0) + t0 + kphi
                assert abs(sm1 - sm0) < 2e-6 * scale
                if sm0 <= kphi / qphi:
                    assert abs(t1 + qphi * sm1 - kphi) < 2e-6 * scale
                else:
                    # beyond the apex the reference still scales the deviator by
                    # (kϕ - qϕ σm) / τ, which is negative there
                    assert abs(t1 - abs(kphi - qphi * sm0)) < 2e-6 * scale
                assert a["eps_q"][i] > eq0[i]
            elif sm0 >= st and BF <= 0:
                # tension apex: mean stress returned to the cut-off, deviator kept
                scale = abs(sm0) + t0 + kphi
                assert abs(sm1 - st) < 2e-6 * scale
                assert abs(t1 - t0) < 2e-6 * scale
            else:
                assert np.allclose(a["sig"][i], pre[i], rtol=0, atol=0)
                assert a["eps_q"][i] == eq0[i]


def test_mcP_2d_return(mpm, orc):
    phi, c = math.radians(28.0), 1500.0
    args, grid, mp, attr, bc = _mat_block(mpm, 2, phi_deg=28.0, cohesion=c,
                                          constitutive="mohrcoulomb")
    o = orc.Oracle(args, grid, mp, attr, bc)
    a = o.a
    n = o.np
    rng = np.random.default_rng(11)
    for trial in range(20):
        s = rng.normal(scale=5000.0, size=(n, 4))
        s[:, :3] -= rng.uniform(-2000, 6000.0, size=(n, 1))
        a["sig"][...] = s
        pre = s.copy()
        o.mcP()
        for i in range(n):
            s0 = pre[i]
            tau0 = math.sqrt(0.25 * (s0[0] - s0[1]) ** 2 + s0[3] ** 2)
            sg0 = 0.5 * (s0[0] + s0[1])
            f0 = tau0 + sg0 * math.sin(phi) - c * math.cos(phi)
            s1 = a["sig"][i]
            tau1 = math.sqrt(0.25 * (s1[0] - s1[1]) ** 2 + s1[3] ** 2)
            sg1 = 0.5 * (s1[0] + s1[1])
            if f0 <= 0:
                assert np.array_equal(s1, s0)
            elif sg0 > c / math.tan(phi):
                assert s1[0] == s1[1] == c / math.tan(phi) and s1[3] == 0.0
            else:
                # back on the Mohr-Coulomb line with the in-plane centre kept
                assert abs(tau1 + sg1 * math.sin(phi) - c * math.cos(phi)) < 1e-7 * c
                assert abs(sg1 - sg0) < 1e-9 * max(1, abs(sg0))
            assert s1[2] == s0[2]


def test_vollock_cell_mean(mpm, orc):
    args, grid, mp, attr, bc = _block(mpm, 2, n=4)
    o = orc.Oracle(args, grid, mp, attr, bc)
    a = o.a
    rng = np.random.default_rng(12)
    a["sig"][...] = rng.normal(scale=100, size=a["sig"].shape)
    a["omega"][...] = rng.uniform(0.5, 1.5, o.np) * a["omega0"]
    a["dF"][...] = 0.0
    a["deps_s"][...] = 0.0
    o.resetgrid()
    o.resetmp()
    o.liE()
    pre = a["sig"].copy()
    sm = a["sigm"].copy()
    o.vollock()
    cells = a["p2c"]
    assert len(np.unique(cells)) > 1
    for c in np.unique(cells):
        idx = cells == c
        # σ = s + σ̄ with s the pre-lock deviator and σ̄ the volume-weighted cell
        # mean of σm (utils_OS.jl:1162-1223); shear components untouched
        mean = np.sum(a["omega"][idx] * sm[idx]) / np.sum(a["omega"][idx])
        for k in range(3):
            assert np.allclose(a["sig"][idx, k] - pre[idx, k], mean - sm[idx], atol=1e-12)
        assert np.array_equal(a["sig"][idx, 3], pre[idx, 3])
        # after the lock every particle of the cell carries the same mean stress
        # up to the 1e-6 trace the truncated 1/3 leaves
        assert np.allclose(a["sigm"][idx], mean, rtol=0, atol=1e-6 * np.abs(pre[idx]).max())
