"""The build-defined 3D Mohr-Coulomb return of the oracle, pinned without the
reference (whose 3D mcP! body is commented out: no arithmetic to match).

Checks: the Jacobi eigen-solver against numpy's eigh; the returned stress is
admissible on all six Mohr-Coulomb planes and lies on the surface; for
associative flow (psi = phi) it is the closest admissible stress in the
complementary-energy norm, found independently by enumerating active sets of
the six planes; for non-associative flow the correction lies in the cone of
the active planes' elastic flow directions; apex; rotation invariance; the
plastic-strain bookkeeping of a main-plane return in closed form.
"""
import itertools
import math

import numpy as np
import pytest

K, G = 7.0e5, 3.2308e5
DT, DD = K + 1.333333 * G, K - 0.666667 * G
D3 = np.array([[DT, DD, DD], [DD, DT, DD], [DD, DD, DT]])


def _tensor(s):
    return np.array([[s[0], s[3], s[5]], [s[3], s[1], s[4]], [s[5], s[4], s[2]]])


def _voigt(T):
    return np.array([T[0, 0], T[1, 1], T[2, 2], T[0, 1], T[1, 2], T[0, 2]])


def _planes(sphi, cphi, c):
    """The six Mohr-Coulomb planes g . p <= b over unordered principal stresses."""
    rows, b = [], []
    for i, j in itertools.permutations(range(3), 2):
        g = np.zeros(3)
        g[i] += 1.0 + sphi
        g[j] += -(1.0 - sphi)
        rows.append(g)
        b.append(2.0 * c * cphi)
    return np.array(rows), np.array(b)


def _closest_point(p, sphi, cphi, c):
    """min (q-p)^T D^-1 (q-p) s.t. the six planes, by active-set enumeration."""
    Gm, b = _planes(sphi, cphi, c)
    best, bestq = np.inf, None
    Dinv = np.linalg.inv(D3)
    if np.all(Gm @ p <= b + 1e-9):
        return p
    for k in (1, 2, 3):
        for act in itertools.combinations(range(6), k):
            A = Gm[list(act)]
            M = A @ D3 @ A.T
            if abs(np.linalg.det(M)) < 1e-12 * np.linalg.norm(M) ** k:
                continue
            lam = np.linalg.solve(M, A @ p - b[list(act)])
            if np.any(lam < -1e-9):
                continue
            q = p - D3 @ A.T @ lam
            if np.any(Gm @ q > b + 1e-6 * (1 + np.abs(b))):
                continue
            e = (q - p) @ Dinv @ (q - p)
            if e < best:
                best, bestq = e, q
    return bestq


def _rand_stress(rng, scale=1e4):
    return rng.normal(0.0, scale, 6) + np.array([-1, -1, -1, 0, 0, 0]) * scale


def test_jacobi_eigen_matches_eigh(orc):
    rng = np.random.default_rng(1)
    cases = [rng.normal(size=6) * 10 ** rng.uniform(-3, 6) for _ in range(200)]
    cases += [np.array([-3.0, -3.0, -7.0, 0, 0, 0]), np.array([5.0, 5.0, 5.0, 0, 0, 0]),
              np.array([-1e4, -1e4, -3e4, 0, 0, 0]), np.zeros(6)]
    for s in cases:
        lam, V = orc.sym3_eig(s)
        T = _tensor(s)
        ref = np.linalg.eigvalsh(T)[::-1]
        scale = max(np.abs(ref).max(), 1e-300)
        assert np.allclose(lam, ref, rtol=0, atol=1e-13 * scale)
        assert np.allclose(V @ V.T, np.eye(3), atol=1e-13)
        assert np.allclose(V @ np.diag(lam) @ V.T, T, rtol=0, atol=1e-13 * scale)


@pytest.mark.parametrize("phi_deg,psi_deg,c", [(30, 0, 0.0), (25, 5, 5e3), (35, 0, 1e3), (20, 20, 2e3)])
def test_admissible_and_on_surface(orc, phi_deg, psi_deg, c):
    phi, psi = math.radians(phi_deg), math.radians(psi_deg)
    sphi, cphi = math.sin(phi), math.cos(phi)
    Gm, b = _planes(sphi, cphi, c)
    rng = np.random.default_rng(2)
    kinds = set()
    for _ in range(400):
        s = _rand_stress(rng)
        out, eq, ek, kind = orc.mc3_point(s, 0.0, K, G, phi, psi, c)
        kinds.add(kind)
        p = np.linalg.eigvalsh(_tensor(out))
        fmax = np.max(Gm @ p - b)
        scale = 1e-9 * (np.abs(s).max() + c)
        assert fmax <= scale
        if kind == 0:
            assert np.array_equal(out, s) and ek == 0.0 and eq == 0.0
        elif kind != 4:
            assert fmax >= -scale  # on the surface
    assert {0, 1}.issubset(kinds) and len(kinds) >= 3


def test_associative_return_is_the_closest_point(orc):
    phi = math.radians(30)
    sphi, cphi = math.sin(phi), math.cos(phi)
    c = 2e3
    rng = np.random.default_rng(3)
    seen = set()
    for _ in range(300):
        p = np.sort(rng.normal(-1e4, 1.5e4, 3))[::-1]
        q, kind = orc.mc3_principal(p, DT, DD, sphi, cphi, sphi, c)
        seen.add(kind)
        ref = _closest_point(p, sphi, cphi, c)
        assert np.allclose(np.sort(q), np.sort(ref), rtol=0, atol=1e-7 * (np.abs(p).max() + c)), (p, q, ref, kind)
    assert {0, 1, 2, 3, 4} <= seen


def test_nonassociative_correction_in_flow_cone(orc):
    phi, psi = math.radians(32), math.radians(4)
    sphi, cphi, spsi = math.sin(phi), math.cos(phi), math.sin(psi)
    c = 1e3
    Gm, b = _planes(sphi, cphi, c)
    N = []
    for i, j in itertools.permutations(range(3), 2):
        n = np.zeros(3)
        n[i] += 1.0 + spsi
        n[j] += -(1.0 - spsi)
        N.append(n)
    N = np.array(N)
    rng = np.random.default_rng(4)
    for _ in range(300):
        p = np.sort(rng.normal(-1e4, 1.5e4, 3))[::-1]
        q, kind = orc.mc3_principal(p, DT, DD, sphi, cphi, spsi, c)
        if kind in (0, 4):
            continue
        tol = 1e-9 * (np.abs(p).max() + c)
        active = np.where(np.abs(Gm @ q - b) <= tol)[0]
        assert active.size >= 1
        dl = np.linalg.solve(D3, p - q)  # plastic strain increment
        lam, *_ = np.linalg.lstsq(N[active].T, dl, rcond=None)
        assert np.linalg.norm(N[active].T @ lam - dl) <= 1e-9 * np.linalg.norm(dl)
        assert np.all(lam >= -1e-9 * np.abs(lam).max())


def test_apex_and_triaxial_closed_forms(orc):
    phi = math.radians(30)
    c = 1e3
    # hydrostatic tension beyond the apex returns to the apex c cot(phi)
    out, eq, ek, kind = orc.mc3_point(np.array([5e3, 5e3, 5e3, 0, 0, 0]), 0.0, K, G, phi, 0.0, c)
    assert kind == 4
    assert np.allclose(out[:3], c / math.tan(phi), rtol=1e-12) and np.allclose(out[3:], 0.0)
    # cohesionless triaxial compression at failure: s1 / s3 = (1 - sin) / (1 + sin)
    s3 = -3e4
    s1 = s3 * (1 - math.sin(phi)) / (1 + math.sin(phi))
    out, _, _, kind = orc.mc3_point(np.array([s1 * 1.001, s1 * 1.001, s3, 0, 0, 0]), 0.0, K, G,
                                    phi, 0.0, 0.0)
    assert kind == 0
    out, _, _, kind = orc.mc3_point(np.array([s1 * 0.9, s1 * 0.9, s3, 0, 0, 0]), 0.0, K, G,
                                    phi, 0.0, 0.0)
    assert kind == 2
    p = np.sort(out[:3])
    assert abs(p[2] / p[0] - (1 - math.sin(phi)) / (1 + math.sin(phi))) < 1e-12


def test_rotation_invariance(orc):
    rng = np.random.default_rng(5)
    phi, psi, c = math.radians(28), math.radians(3), 3e3
    for _ in range(50):
        s = _rand_stress(rng)
        A = rng.normal(size=(3, 3))
        R, _ = np.linalg.qr(A)
        sr = _voigt(R @ _tensor(s) @ R.T)
        o1, eq1, _, k1 = orc.mc3_point(s, 0.0, K, G, phi, psi, c)
        o2, eq2, _, k2 = orc.mc3_point(sr, 0.0, K, G, phi, psi, c)
        ref = _voigt(R @ _tensor(o1) @ R.T)
        assert np.allclose(o2, ref, rtol=0, atol=1e-8 * np.abs(s).max())


def test_main_plane_strain_bookkeeping(orc):
    # principal-aligned trial on the main plane: Δσ = -γ D N, so
    # eps_k = tr(D^-1 Δσ) = -γ (N1 + N2 + N3) = -2 γ sin(psi) and
    # eps_q = sqrt(0.666667 γ² (N1² + N3²))
    phi, psi, c = math.radians(30), math.radians(10), 1e3
    s = np.array([-1e3, -1.2e4, -4e4, 0, 0, 0])
    out, eq, ek, kind = orc.mc3_point(s, 0.0, K, G, phi, psi, c)
    assert kind == 1
    spsi = math.sin(psi)
    Nv = np.array([1 + spsi, 0.0, -(1 - spsi)])
    gam = (s[0] - out[0]) / (D3 @ Nv)[0]
    assert np.allclose(s[:3] - out[:3], gam * (D3 @ Nv), rtol=1e-10)
    assert ek == pytest.approx(-2 * gam * spsi, rel=1e-9)
    assert eq == pytest.approx(math.sqrt(0.666667 * gam ** 2 * (Nv[0] ** 2 + Nv[2] ** 2)), rel=1e-9)
    # cohesion softening c = max(c + Hp eps_q, cr): a larger eps_q with Hp < 0 lowers c
    o_soft, _, _, _ = orc.mc3_point(s, 0.5, K, G, phi, psi, c, Hp=-1e3, cr=100.0)
    o_hard, _, _, _ = orc.mc3_point(s, 0.0, K, G, phi, psi, 500.0)
    assert np.allclose(o_soft, o_hard, rtol=1e-12)


# --- softening (mohr_coulomb.tcc:8-25 linear rule; one peak/residual strain pair) ---
PHI, PHIR, PSI = math.radians(30.0), math.radians(20.0), math.radians(6.0)
COH, CR, HP = 10.0e3, 2.0e3, -400.0e3  # residual plastic strain (CR - COH) / HP = 0.02


@pytest.mark.parametrize("eq, phi_deg, psi_deg, c", [
    (0.0, 30.0, 6.0, 10.0e3),      # peak
    (0.005, 27.5, 4.5, 8.0e3),     # a quarter of the way: t = 0.25
    (0.01, 25.0, 3.0, 6.0e3),      # half way
    (0.02, 20.0, 0.0, 2.0e3),      # residual reached
    (0.05, 20.0, 0.0, 2.0e3),      # past it: residual values stay
])
def test_softening_rule_hand_values(orc, eq, phi_deg, psi_deg, c):
    phi, psi, coh = orc.mc_soften(eq, PHI, PSI, COH, HP, CR, PHIR)
    assert math.degrees(phi) == pytest.approx(phi_deg, abs=1e-12)
    assert math.degrees(psi) == pytest.approx(psi_deg, abs=1e-12)
    assert coh == pytest.approx(c, rel=1e-14)


@pytest.mark.parametrize("Hp, cr, phir", [
    (0.0, 2.0e3, PHIR),        # no cohesion softening modulus
    (400.0e3, 2.0e3, PHIR),    # hardening cohesion
    (HP, 10.0e3, PHIR),        # cohesion already residual: no strain interval
    (HP, 2.0e3, 0.0),          # the reference's default phir = 0: friction stays
])
def test_softening_off_keeps_phi_and_psi(orc, Hp, cr, phir):
    for eq in (0.0, 0.01, 0.1):
        phi, psi, coh = orc.mc_soften(eq, PHI, PSI, COH, Hp, cr, phir)
        assert (phi, psi) == (PHI, PSI)
        assert coh == max(COH + Hp * eq, cr)  # the 2D mcP! cohesion rule


@pytest.mark.parametrize("eq", [0.0, 0.004, 0.013, 0.03])
def test_softened_return_is_the_return_with_softened_parameters(orc, eq):
    # the softening only sets the strength the return uses, from the step-start
    # plastic strain: the same return with those parameters fixed gives the
    # same stress and strains
    rng = np.random.default_rng(11)
    phi_s, psi_s, c_s = orc.mc_soften(eq, PHI, PSI, COH, HP, CR, PHIR)
    for _ in range(20):
        s = rng.normal(scale=4.0e4, size=6) - np.array([3e4, 3e4, 3e4, 0, 0, 0])
        a = orc.mc3_point(s, eq, K, G, PHI, PSI, COH, Hp=HP, cr=CR, phir=PHIR)
        b = orc.mc3_point(s, eq, K, G, phi_s, psi_s, c_s, Hp=0.0, cr=c_s, phir=0.0)
        np.testing.assert_array_equal(a[0], b[0])
        assert a[1:] == b[1:]
        if a[3] != 0:
            # on the softened surface: max over pairs of the Mohr-Coulomb function
            lam = np.linalg.eigvalsh(_tensor(a[0]))
            f = (lam[2] - lam[0]) + (lam[2] + lam[0]) * math.sin(phi_s) - 2 * c_s * math.cos(phi_s)
            assert abs(f) <= 1e-7 * (abs(lam).max() + c_s)
