"""Synthetic scenes: granular column collapses built through the user API.

``column_collapse_2d`` follows the reference's 2D validation script
(test/validation/2d_case.jl:17-122): a Drucker-Prager column on a rigid floor,
2x2 particles per cell at cell quarter-points, no-slip floor and slip side
walls applied by node coordinate.  ``column_collapse_3d`` is its 3D
counterpart with 2x2x2 particles per cell.  The grid is padded by ``pad``
ghost cells on every side because the uGIMP stencil reaches one node beyond a
particle's cell and the reference has no bounds checks.

Default sizes are the plane-strain column of 12,800 particles (0.2 m x 0.1 m,
h = 2.5 mm, 0.5 m x 0.12 m box, dt = 2e-5 s, 1 s) and the 3D column of
2,048,000 particles (0.4 x 0.4 x 0.2 m, h = 5 mm, 1.2 x 1.2 x 0.25 m box,
dt = 0.2 h / c_p, 20,000 steps).
"""
from __future__ import annotations

import math

import numpy as np

from .types import (UserArgs2D, UserArgs3D, UserGrid2D, UserGrid3D, UserParticle2D,
                    UserParticle3D, UserProperty, UserVBoundary2D, UserVBoundary3D,
                    julia_range, meshbuilder)

# Granular material of the collapse cases (2d_case.jl:23-30 with E = 0.84 MPa):
RHO = 2650.0
NU = 0.3
E = 0.84e6
KS = E / (3.0 * (1.0 - 2.0 * NU))   # 7.0e5
GS = E / (2.0 * (1.0 + NU))         # 3.2308e5
PHI_DEG = 19.8


def p_wave_speed(Ks=KS, Gs=GS, rho=RHO):
    return math.sqrt((Ks + 4.0 / 3.0 * Gs) / rho)


def _lattice(lo, hi, step):
    """Quarter-point lattice lo+step/2 : step : hi-step/2 (2d_case.jl:77-78)."""
    return julia_range(lo + step / 2, step, hi - step / 2)


def column_collapse_2d(*, h=0.0025, width=0.2, height=0.1, box_x=0.5, box_y=0.12, pad=2,
                       dt=2e-5, Ttol=None, nsteps=50_000, eps="FP64", constitutive="druckerprager",
                       phi_deg=PHI_DEG, cohesion=0.0, psi_deg=0.0, sigma_t=0.0, Hp=0.0, cr=0.0,
                       va="a", MVL=False, zeta=0.0, Te=0.0, gravity=-9.8, basis="uGIMP",
                       sort_interval=None, device="HIP", project_name="collapse2d",
                       project_path="."):
    """Plane-strain granular column collapse; returns (args, grid, mp, attr, bc)."""
    if Ttol is None:
        Ttol = dt * nsteps
    nic = 9 if basis == "uGIMP" else 4
    args = UserArgs2D(Ttol=Ttol, Te=Te, ΔT=dt, FLIP=1, PIC=0, constitutive=constitutive,
                      basis=basis, MVL=MVL, device=device, coupling="OS", scheme="MUSL",
                      va=va, gravity=gravity, ζs=zeta, project_name=project_name,
                      project_path=project_path, ϵ=eps, progressbar=False,
                      sort_interval=sort_interval)
    grid = UserGrid2D(ϵ=eps, x1=-pad * h, x2=box_x + pad * h, y1=-pad * h, y2=box_y + pad * h,
                      dx=h, dy=h, NIC=nic)
    lp = h / 2
    pts = meshbuilder(_lattice(0.0, width, lp), _lattice(0.0, height, lp))
    mp = UserParticle2D(ϵ=eps, NIC=nic, dx=lp, dy=lp, ξ=pts, ρs=np.full(len(pts), RHO))
    attr = UserProperty(ϵ=eps, nid=np.ones(mp.np), ν=[NU], Es=[E], Gs=[GS], Ks=[KS],
                        σt=[sigma_t], ϕ=[math.radians(phi_deg)], ϕr=[0], ψ=[math.radians(psi_deg)],
                        c=[cohesion], cr=[cr], Hp=[Hp])
    gx, gy = grid.ξ[:, 0], grid.ξ[:, 1]
    vx_idx = ((gx <= 0.0) | (gx >= box_x) | (gy <= 0.0)).astype(np.int64)
    vy_idx = (gy <= 0.0).astype(np.int64)
    bc = UserVBoundary2D(ϵ=eps, vx_s_idx=vx_idx, vx_s_val=np.zeros(grid.ni),
                         vy_s_idx=vy_idx, vy_s_val=np.zeros(grid.ni))
    return args, grid, mp, attr, bc


def column_collapse_3d(*, h=0.005, width=0.4, height=0.2, box_xy=1.2, box_z=0.25, pad=2,
                       dt=None, cfl=0.2, Ttol=None, nsteps=20_000, eps="FP32",
                       constitutive="druckerprager", phi_deg=PHI_DEG, cohesion=0.0, psi_deg=0.0,
                       sigma_t=0.0, va="a", MVL=False, zeta=0.0, Te=0.0, gravity=-9.8,
                       basis="uGIMP", sort_interval=None, device="HIP",
                       project_name="collapse3d", project_path="."):
    """3D granular column collapse centred in x and y; returns (args, grid, mp, attr, bc).
    Floor z <= 0 no-slip; walls at x, y = +-box_xy/2 slip (normal velocity fixed)."""
    if dt is None:
        dt = cfl * h / p_wave_speed()
    if Ttol is None:
        Ttol = dt * nsteps
    nic = 27 if basis == "uGIMP" else 8
    half = box_xy / 2
    args = UserArgs3D(Ttol=Ttol, Te=Te, ΔT=dt, FLIP=1, PIC=0, constitutive=constitutive,
                      basis=basis, MVL=MVL, device=device, coupling="OS", scheme="MUSL",
                      va=va, gravity=gravity, ζs=zeta, project_name=project_name,
                      project_path=project_path, ϵ=eps, progressbar=False,
                      sort_interval=sort_interval)
    grid = UserGrid3D(ϵ=eps, x1=-half - pad * h, x2=half + pad * h, y1=-half - pad * h,
                      y2=half + pad * h, z1=-pad * h, z2=box_z + pad * h, dx=h, dy=h, dz=h,
                      NIC=nic)
    lp = h / 2
    w2 = width / 2
    pts = meshbuilder(_lattice(-w2, w2, lp), _lattice(-w2, w2, lp), _lattice(0.0, height, lp))
    mp = UserParticle3D(ϵ=eps, NIC=nic, dx=lp, dy=lp, dz=lp, ξ=pts, ρs=np.full(len(pts), RHO))
    attr = UserProperty(ϵ=eps, nid=np.ones(mp.np), ν=[NU], Es=[E], Gs=[GS], Ks=[KS],
                        σt=[sigma_t], ϕ=[math.radians(phi_deg)], ϕr=[0], ψ=[math.radians(psi_deg)],
                        c=[cohesion], cr=[0], Hp=[0])
    gx, gy, gz = grid.ξ[:, 0], grid.ξ[:, 1], grid.ξ[:, 2]
    floor = gz <= 0.0
    vx_idx = ((gx <= -half) | (gx >= half) | floor).astype(np.int64)
    vy_idx = ((gy <= -half) | (gy >= half) | floor).astype(np.int64)
    vz_idx = floor.astype(np.int64)
    z = np.zeros(grid.ni)
    bc = UserVBoundary3D(ϵ=eps, vx_s_idx=vx_idx, vx_s_val=z, vy_s_idx=vy_idx, vy_s_val=z,
                         vz_s_idx=vz_idx, vz_s_val=z)
    return args, grid, mp, attr, bc


def runout(xi: np.ndarray, xi0: np.ndarray, axis=0, threshold=None) -> float:
    """Largest coordinate along `axis` reached by particles that moved more than
    `threshold` (default: 1e-4 of the domain) -- the collapse runout distance."""
    d = np.linalg.norm(xi - xi0, axis=1)
    thr = threshold if threshold is not None else 1e-6
    moved = d > thr
    if not np.any(moved):
        return float(np.max(xi0[:, axis]))
    return float(np.max(xi[moved, axis]))


def kinetic_energy(ms: np.ndarray, vs: np.ndarray) -> float:
    return float(0.5 * np.sum(ms * np.sum(vs * vs, axis=1)))
