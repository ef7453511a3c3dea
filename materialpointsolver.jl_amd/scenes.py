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
                       phir_deg=0.0, va="a", MVL=False, zeta=0.0, Te=0.0, gravity=-9.8, basis="uGIMP",
                       scheme="MUSL", sort_interval=None, device="HIP", project_name="collapse2d",
                       project_path="."):
    """Plane-strain granular column collapse; returns (args, grid, mp, attr, bc)."""
    if Ttol is None:
        Ttol = dt * nsteps
    nic = 9 if basis == "uGIMP" else 4
    args = UserArgs2D(Ttol=Ttol, Te=Te, ΔT=dt, FLIP=1, PIC=0, constitutive=constitutive,
                      basis=basis, MVL=MVL, device=device, coupling="OS", scheme=scheme,
                      va=va, gravity=gravity, ζs=zeta, project_name=project_name,
                      project_path=project_path, ϵ=eps, progressbar=False,
                      sort_interval=sort_interval)
    grid = UserGrid2D(ϵ=eps, x1=-pad * h, x2=box_x + pad * h, y1=-pad * h, y2=box_y + pad * h,
                      dx=h, dy=h, NIC=nic)
    lp = h / 2
    pts = meshbuilder(_lattice(0.0, width, lp), _lattice(0.0, height, lp))
    mp = UserParticle2D(ϵ=eps, NIC=nic, dx=lp, dy=lp, ξ=pts, ρs=np.full(len(pts), RHO))
    attr = UserProperty(ϵ=eps, nid=np.ones(mp.np), ν=[NU], Es=[E], Gs=[GS], Ks=[KS],
                        σt=[sigma_t], ϕ=[math.radians(phi_deg)], ϕr=[math.radians(phir_deg)],
                        ψ=[math.radians(psi_deg)], c=[cohesion], cr=[cr], Hp=[Hp])
    gx, gy = grid.ξ[:, 0], grid.ξ[:, 1]
    vx_idx = ((gx <= 0.0) | (gx >= box_x) | (gy <= 0.0)).astype(np.int64)
    vy_idx = (gy <= 0.0).astype(np.int64)
    bc = UserVBoundary2D(ϵ=eps, vx_s_idx=vx_idx, vx_s_val=np.zeros(grid.ni),
                         vy_s_idx=vy_idx, vy_s_val=np.zeros(grid.ni))
    return args, grid, mp, attr, bc


def column_collapse_3d(*, h=0.005, width=0.4, height=0.2, box_xy=1.2, box_z=0.25, pad=2,
                       dt=None, cfl=0.2, Ttol=None, nsteps=20_000, eps="FP32",
                       constitutive="druckerprager", phi_deg=PHI_DEG, cohesion=0.0, psi_deg=0.0,
                       sigma_t=0.0, Hp=0.0, cr=0.0, phir_deg=0.0, va="a", MVL=False, zeta=0.0,
                       Te=0.0, gravity=-9.8,
                       basis="uGIMP", scheme="MUSL", sort_interval=None, device="HIP",
                       project_name="collapse3d", project_path="."):
    """3D granular column collapse centred in x and y; returns (args, grid, mp, attr, bc).
    Floor z <= 0 no-slip; walls at x, y = +-box_xy/2 slip (normal velocity fixed).
    Hp, cr and phir_deg are the Mohr-Coulomb softening parameters (cohesion
    softening modulus and residual cohesion and friction)."""
    if dt is None:
        dt = cfl * h / p_wave_speed()
    if Ttol is None:
        Ttol = dt * nsteps
    nic = 27 if basis == "uGIMP" else 8
    half = box_xy / 2
    args = UserArgs3D(Ttol=Ttol, Te=Te, ΔT=dt, FLIP=1, PIC=0, constitutive=constitutive,
                      basis=basis, MVL=MVL, device=device, coupling="OS", scheme=scheme,
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
                        σt=[sigma_t], ϕ=[math.radians(phi_deg)], ϕr=[math.radians(phir_deg)],
                        ψ=[math.radians(psi_deg)], c=[cohesion], cr=[cr], Hp=[Hp])
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


def ensemble_2d(*, n_instances=256, seed=2024, h=0.0025, np_range=(5_000, 60_000),
                aspect_range=(0.5, 4.0), phi_deg_range=(20.0, 40.0), cohesion_median=1.0e3,
                cohesion_sigma_ln=0.5, runout_factor=3.0, gap_cells=4, row_width=6.0,
                dt=2e-5, nsteps=20_000, eps="FP32", va="a", device="HIP"):
    """A parameter-uncertainty ensemble of independent plane-strain column
    collapses packed side by side into one background grid.

    Instance k (seeded): column aspect ratio H/W ~ U(aspect_range), particle
    count ~ U(np_range) at 2x2 particles per cell, Drucker-Prager friction
    ~ U(phi_deg_range), cohesion ~ LogNormal(ln cohesion_median,
    cohesion_sigma_ln), its own material id.  Its box has a no-slip floor, a
    slip wall behind the column and a slip wall runout_factor * H + W ahead
    of it.  Boxes are placed in rows of at most row_width metres with
    gap_cells empty cells between them, so no uGIMP stencil (one node beyond
    a particle's cell) reaches two instances and the instances advance
    independently in one set of kernel launches.

    Returns (args, grid, mp, attr, bc, layout); layout holds per-instance
    particle ranges (caller order), box origins and lengths, and parameters.
    """
    rng = np.random.default_rng(seed)
    lp = h / 2
    inst = []
    for k in range(n_instances):
        a = rng.uniform(*aspect_range)
        n = rng.uniform(*np_range)
        cells = n / 4.0
        wc = max(2, int(round(math.sqrt(cells / a))))
        hc = max(2, int(round(wc * a)))
        Lc = wc + int(math.ceil(runout_factor * hc))
        inst.append(dict(wc=wc, hc=hc, Lc=Lc, phi=math.radians(rng.uniform(*phi_deg_range)),
                         c=float(math.exp(rng.normal(math.log(cohesion_median), cohesion_sigma_ln)))))
    # shelf packing in rows of at most row_width (cells), tallest first per row
    order = sorted(range(n_instances), key=lambda k: -inst[k]["hc"])
    maxw = int(row_width / h)
    x, y, row_h = gap_cells, gap_cells, 0
    for k in order:
        b = inst[k]
        if x + b["Lc"] + gap_cells > maxw and x > gap_cells:
            y += row_h + 2 + gap_cells
            x, row_h = gap_cells, 0
        b["cx"], b["cy"] = x, y
        x += b["Lc"] + gap_cells
        row_h = max(row_h, b["hc"])
    ncx = maxw
    ncy = y + row_h + 2 + gap_cells
    grid = UserGrid2D(ϵ=eps, x1=0.0, x2=ncx * h, y1=0.0, y2=ncy * h, dx=h, dy=h, NIC=9)
    pts, nid, ranges = [], [], []
    start = 0
    for k, b in enumerate(inst):
        x0, y0 = b["cx"] * h, b["cy"] * h
        px = _lattice(x0, x0 + b["wc"] * h, lp)
        py = _lattice(y0, y0 + b["hc"] * h, lp)
        p = meshbuilder(px, py)
        pts.append(p)
        nid.append(np.full(len(p), k + 1, dtype=np.int64))
        ranges.append((start, start + len(p)))
        start += len(p)
    pts = np.concatenate(pts)
    mp = UserParticle2D(ϵ=eps, NIC=9, dx=lp, dy=lp, ξ=pts, ρs=np.full(len(pts), RHO))
    nm = n_instances
    attr = UserProperty(ϵ=eps, nid=np.concatenate(nid), ν=[NU] * nm, Es=[E] * nm, Gs=[GS] * nm,
                        Ks=[KS] * nm, σt=[0.0] * nm, ϕ=[b["phi"] for b in inst], ϕr=[0.0] * nm,
                        ψ=[0.0] * nm, c=[b["c"] for b in inst], cr=[0.0] * nm, Hp=[0.0] * nm)
    # boundary conditions by node coordinate, instance by instance, on the
    # (column, row) view of the reference node order (rows from the top)
    nnx, nny = grid.nnx, grid.nny
    vx_idx = np.zeros(grid.ni, dtype=np.int64)
    vy_idx = np.zeros(grid.ni, dtype=np.int64)
    vx2 = vx_idx.reshape(nnx, nny)
    vy2 = vy_idx.reshape(nnx, nny)
    for b in inst:
        xa, xb = b["cx"], b["cx"] + b["Lc"]
        ya, yb = b["cy"], b["cy"] + b["hc"] + 2
        cols = np.arange(max(xa - 2, 0), min(xb + 2, nnx - 1) + 1)
        iys = np.arange(max(ya - 2, 0), min(yb, nny - 1) + 1)
        rows = nny - 1 - iys
        gxs, gys = np.meshgrid(cols, iys, indexing="ij")
        floor = gys <= ya
        walls = (gxs <= xa) | (gxs >= xb)
        sub = np.ix_(cols, rows)
        vx2[sub] |= (floor | walls).astype(np.int64)
        vy2[sub] |= floor.astype(np.int64)
    bc = UserVBoundary2D(ϵ=eps, vx_s_idx=vx_idx, vx_s_val=np.zeros(grid.ni),
                         vy_s_idx=vy_idx, vy_s_val=np.zeros(grid.ni))
    args = UserArgs2D(Ttol=dt * nsteps, Te=0.0, ΔT=dt, FLIP=1, PIC=0,
                      constitutive="druckerprager", basis="uGIMP", MVL=False, device=device,
                      coupling="OS", scheme="MUSL", va=va, gravity=-9.8, ζs=0.0,
                      project_name="ensemble2d", project_path=".", ϵ=eps, progressbar=False)
    layout = dict(h=h, ranges=ranges,
                  origin=[(b["cx"] * h, b["cy"] * h) for b in inst],
                  width=[b["wc"] * h for b in inst], height=[b["hc"] * h for b in inst],
                  length=[b["Lc"] * h for b in inst], phi=[b["phi"] for b in inst],
                  cohesion=[b["c"] for b in inst])
    return args, grid, mp, attr, bc, layout


def ensemble_instance(layout, k, *, eps="FP32", dt=2e-5, nsteps=20_000, device="HIP", pad=2):
    """Instance k of an ensemble_2d layout as a standalone scene on its own
    tight grid (same particles, material and walls): the sequential baseline."""
    h = layout["h"]
    lp = h / 2
    W, H, L = layout["width"][k], layout["height"][k], layout["length"][k]
    grid = UserGrid2D(ϵ=eps, x1=-pad * h, x2=L + pad * h, y1=-pad * h, y2=H + 2 * h + pad * h,
                      dx=h, dy=h, NIC=9)
    pts = meshbuilder(_lattice(0.0, W, lp), _lattice(0.0, H, lp))
    mp = UserParticle2D(ϵ=eps, NIC=9, dx=lp, dy=lp, ξ=pts, ρs=np.full(len(pts), RHO))
    attr = UserProperty(ϵ=eps, nid=np.ones(mp.np), ν=[NU], Es=[E], Gs=[GS], Ks=[KS], σt=[0.0],
                        ϕ=[layout["phi"][k]], ϕr=[0.0], ψ=[0.0], c=[layout["cohesion"][k]],
                        cr=[0.0], Hp=[0.0])
    gx = np.rint(grid.ξ[:, 0] / h).astype(np.int64)
    gy = np.rint(grid.ξ[:, 1] / h).astype(np.int64)
    Lc = int(round(L / h))
    vx_idx = ((gx <= 0) | (gx >= Lc) | (gy <= 0)).astype(np.int64)
    vy_idx = (gy <= 0).astype(np.int64)
    bc = UserVBoundary2D(ϵ=eps, vx_s_idx=vx_idx, vx_s_val=np.zeros(grid.ni),
                         vy_s_idx=vy_idx, vy_s_val=np.zeros(grid.ni))
    args = UserArgs2D(Ttol=dt * nsteps, Te=0.0, ΔT=dt, FLIP=1, PIC=0,
                      constitutive="druckerprager", basis="uGIMP", MVL=False, device=device,
                      coupling="OS", scheme="MUSL", va="a", gravity=-9.8, ζs=0.0,
                      project_name="instance", project_path=".", ϵ=eps, progressbar=False)
    return args, grid, mp, attr, bc


def gaussian_random_field(nx, ny, dx, corr_len, std, seed):
    """Seeded stationary Gaussian random field on an nx x ny grid of spacing dx
    with squared-exponential covariance std^2 exp(-r^2 / corr_len^2).

    Spectral synthesis: white noise filtered by the square root of the
    covariance's spectral density exp(-k^2 corr_len^2 / 4), on a grid padded by
    3 corr_len so the periodic FFT does not correlate opposite edges; the
    sample is then shifted to zero mean and scaled to the requested standard
    deviation.  (The reference's grf_gc! draws from Julia's global RNG and
    min-max rescales; this one is reproducible from `seed`.)"""
    pad = int(math.ceil(3.0 * corr_len / dx))
    Nx, Ny = nx + 2 * pad, ny + 2 * pad
    rng = np.random.default_rng(seed)
    w = rng.standard_normal((Nx, Ny))
    kx = 2.0 * np.pi * np.fft.fftfreq(Nx, dx)
    ky = 2.0 * np.pi * np.fft.rfftfreq(Ny, dx)
    k2 = kx[:, None] ** 2 + ky[None, :] ** 2
    f = np.fft.irfft2(np.fft.rfft2(w) * np.exp(-k2 * corr_len ** 2 / 8.0), s=(Nx, Ny))
    f = f[pad:pad + nx, pad:pad + ny]
    f = f - f.mean()
    return f * (std / f.std())


def _bilinear(field, x0, dx, y0, dy, x, y):
    fx = np.asarray(x, dtype=np.float64)
    fy = np.asarray(y, dtype=np.float64)
    fx = np.clip((fx - x0) / dx, 0.0, field.shape[0] - 1.0)
    fy = np.clip((fy - y0) / dy, 0.0, field.shape[1] - 1.0)
    i = np.minimum(fx.astype(np.int64), field.shape[0] - 2)
    j = np.minimum(fy.astype(np.int64), field.shape[1] - 2)
    tx, ty = fx - i, fy - j
    return ((1 - tx) * (1 - ty) * field[i, j] + tx * (1 - ty) * field[i + 1, j] +
            (1 - tx) * ty * field[i, j + 1] + tx * ty * field[i + 1, j + 1])


def slope_3d(*, h=0.5, length=400.0, width=200.0, angle_deg=35.0, depth=10.0,
             footprint=(250.0, 125.0), footprint_x0=20.0, corr_len=20.0, grf_std=2.0,
             phi_deg_range=(25.0, 35.0), cohesion_median=5.0e3, cohesion_sigma_ln=0.5,
             rho=2000.0, E_mod=10.0e6, nu=0.3, seed=7, headroom=25.0, pad=2, eps="FP32",
             cfl=0.2, k0_steps=2000, zeta_k0=0.1, runout_time=30.0, device="HIP"):
    """Synthetic 3D slope failure in a slope-aligned frame (x downslope, y
    across, z normal to the mean plane), Mohr-Coulomb.

    The rigid bed is the plane z = 0 perturbed by a seeded Gaussian random field
    (correlation corr_len, standard deviation grf_std); every grid node at or
    below it is fixed.  A depth-thick sliding mass covers the footprint
    (2x2x2 particles per cell, columns following the bed).  Each cell of the
    initial mass gets its own material: friction ~ U(phi_deg_range), cohesion
    ~ LogNormal(ln cohesion_median, cohesion_sigma_ln), drawn from `seed`.
    Gravity acts along -z cos(angle) and +x sin(angle) (args.ext carries the
    downslope component).  Walls: x = 0 and x = length fix vx, y = 0 and y =
    width fix vy.  The returned args describe the K0 phase (k0_steps damped
    elastic steps with a gravity ramp, Te = k0_steps dt); `runout_args` in the
    layout is the undamped runout phase.  The defaults give 20,000,000
    particles (250 x 125 m footprint, 10 m deep, h = 0.5 m).  The grid reaches
    `headroom` above the highest bed point plus the depth.  The mass reaches
    the x = length wall after about 12 s of runout and piles up there to about
    45 m above the bed (measured at h = 2 m): a runout past that needs
    headroom >= 60.

    Returns (args, grid, mp, attr, bc, layout)."""
    rng = np.random.default_rng(seed + 1)
    lp = h / 2
    g = 9.8
    ang = math.radians(angle_deg)
    # bed on the node lattice
    nx_n = int(round(length / h)) + 1 + 2 * pad
    ny_n = int(round(width / h)) + 1 + 2 * pad
    x1, y1 = -pad * h, -pad * h
    bed = gaussian_random_field(nx_n, ny_n, h, corr_len, grf_std, seed)
    zmin = math.floor((bed.min() - 2 * h) / h) * h - pad * h
    zmax = bed.max() + depth + headroom
    grid = UserGrid3D(ϵ=eps, x1=x1, x2=x1 + (nx_n - 1) * h, y1=y1, y2=y1 + (ny_n - 1) * h,
                      z1=zmin, z2=zmin + math.ceil((zmax - zmin) / h) * h, dx=h, dy=h, dz=h,
                      NIC=27)
    # particles: lp lattice in x, y over the footprint; 2 * depth / h layers above the bed
    fx0, fy0 = footprint_x0, (width - footprint[1]) / 2.0
    px = fx0 + lp / 2 + lp * np.arange(int(round(footprint[0] / lp)))
    py = fy0 + lp / 2 + lp * np.arange(int(round(footprint[1] / lp)))
    nz = int(round(depth / lp))
    PX, PY = np.meshgrid(px, py, indexing="ij")
    zb = _bilinear(bed, x1, h, y1, h, PX.ravel(), PY.ravel())
    ncol = zb.size
    pts = np.empty((ncol * nz, 3), dtype=np.float64)
    pts[:, 0] = np.repeat(PX.ravel(), nz)
    pts[:, 1] = np.repeat(PY.ravel(), nz)
    pts[:, 2] = np.repeat(zb, nz) + lp / 2 + lp * np.tile(np.arange(nz), ncol)
    mp = UserParticle3D(ϵ=eps, NIC=27, dx=lp, dy=lp, dz=lp, ξ=pts,
                        ρs=np.full(len(pts), rho))
    # one material per initial cell
    cell = np.floor((pts - np.array([x1, y1, zmin])) / h).astype(np.int64)
    key = cell[:, 0] + nx_n * (cell[:, 1] + ny_n * cell[:, 2])
    uniq, nid = np.unique(key, return_inverse=True)
    nm = len(uniq)
    phi = np.radians(rng.uniform(*phi_deg_range, size=nm))
    coh = np.exp(rng.normal(math.log(cohesion_median), cohesion_sigma_ln, size=nm))
    Ks = E_mod / (3 * (1 - 2 * nu))
    Gs = E_mod / (2 * (1 + nu))
    z = np.zeros(nm)
    attr = UserProperty(ϵ=eps, nid=nid + 1, ν=np.full(nm, nu), Es=np.full(nm, E_mod),
                        Gs=np.full(nm, Gs), Ks=np.full(nm, Ks), σt=z, ϕ=phi, ϕr=z, ψ=z,
                        c=coh, cr=z, Hp=z)
    # rigid bed and walls by node coordinate (reference 3D node order: y fastest)
    gx, gy, gz = grid.ξ[:, 0], grid.ξ[:, 1], grid.ξ[:, 2]
    bed_at_node = _bilinear(bed, x1, h, y1, h, gx, gy)
    fixed = gz <= bed_at_node
    vx_idx = (fixed | (gx <= 0.0) | (gx >= length)).astype(np.int64)
    vy_idx = (fixed | (gy <= 0.0) | (gy >= width)).astype(np.int64)
    vz_idx = fixed.astype(np.int64)
    zv = np.zeros(grid.ni)
    bc = UserVBoundary3D(ϵ=eps, vx_s_idx=vx_idx, vx_s_val=zv, vy_s_idx=vy_idx, vy_s_val=zv,
                         vz_s_idx=vz_idx, vz_s_val=zv)
    dt = cfl * h / p_wave_speed(Ks, Gs, rho)
    gvec = dict(gravity_x=g * math.sin(ang))
    common = dict(ΔT=dt, FLIP=1, PIC=0, constitutive="mohrcoulomb", basis="uGIMP", MVL=False,
                  device=device, coupling="OS", scheme="MUSL", va="a",
                  gravity=-g * math.cos(ang), project_path=".", ϵ=eps, progressbar=False,
                  ext=gvec)
    Te = k0_steps * dt
    args = UserArgs3D(Ttol=Te, Te=Te, ζs=zeta_k0, project_name="slope3d_k0", **common)
    runout_args = UserArgs3D(Ttol=Te + runout_time, Te=Te, ζs=0.0, project_name="slope3d",
                             **common)
    layout = dict(bed=bed, dt=dt, k0_steps=k0_steps, runout_args=runout_args, nmat=nm,
                  footprint=(fx0, fy0, footprint[0], footprint[1]), seed=seed)
    return args, grid, mp, attr, bc, layout
