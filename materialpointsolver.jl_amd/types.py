"""Host-side model structs: a Python mirror of MaterialPointSolver.jl's user types.

Each constructor keeps the reference's keyword names, defaults and checks
(src/types/modelargs.jl, grid.jl, particle.jl, property.jl, boundary.jl).
Julia symbols become strings (``:uGIMP`` -> ``"uGIMP"``); Julia column-major
``(rows, k)`` matrices become Fortran-ordered numpy arrays of the same shape, so
each column is one contiguous stream, exactly the memory the reference's
device plugins copy.  Greek keyword names are kept (Python folds ``ϵ`` to
``ε`` and ``ϕ`` to ``φ`` when it reads identifiers, so either spelling works).

Deliberate differences, all outside the step's arithmetic:

* ``UserArgs*`` does not delete and recreate ``project_path/project_name``
  (the reference does, modelargs.jl:68-69); the folder is created only when
  output is requested.
* The host grid keeps only what the device needs from it (node coordinates
  ``ξ`` and the ``p2nD`` table); the nodal field arrays live on the device.
* The host particle struct keeps the input state and the fields
  ``device2host!`` writes back; the per-slot basis arrays (``Nij``, ``∂N``,
  ``p2n``) are not materialised because the device recomputes the basis from
  the positions in every pass.
* Property vectors given with one entry (the reference's ``=[0]`` defaults)
  are broadcast to the number of materials instead of being indexed out of
  bounds.
"""
from __future__ import annotations

from fractions import Fraction

import numpy as np


class ArgumentError(ValueError):
    """Julia's ArgumentError (invalid constructor input)."""


def _types(eps: str):
    """T1/T2 of the reference's ``ϵ`` keyword (modelargs.jl:65-66)."""
    if eps not in ("FP64", "FP32"):
        eps = "FP64"
    return (np.int32, np.float32, eps) if eps == "FP32" else (np.int64, np.float64, eps)


# --------------------------------------------------------------------------- Args
class Args:
    """Args2D / Args3D (src/types/modelargs.jl:31-58, 122-149)."""

    _fields = ("Ttol", "Te", "ΔT", "time_step", "FLIP", "PIC", "constitutive", "basis",
               "animation", "hdf5", "hdf5_step", "MVL", "device", "coupling", "scheme", "va",
               "progressbar", "gravity", "ζs", "ζw", "αT", "iter_num", "end_time",
               "start_time", "project_name", "project_path", "ext")

    def __init__(self, dim, T1, T2, eps, **kw):
        self.dim = dim
        self.T1, self.T2, self.ϵ = T1, T2, eps
        for k, v in kw.items():
            setattr(self, k, v)

    def __repr__(self):
        return (f"DeviceArgs{self.dim}D(project={self.project_name!r}, precision={self.ϵ}, "
                f"constitutive={self.constitutive}, basis={self.basis}, MVL={self.MVL}, "
                f"device={self.device})")


_DEV_SET = ("CPU", "CUDA", "ROCm", "oneAPI", "Metal", "HIP")


def _user_args(dim, *, Ttol, Te=0, ΔT, time_step="fixed", FLIP=1, PIC=0, constitutive,
               basis="uGIMP", animation=False, hdf5=False, hdf5_step=1, MVL=False,
               device="HIP", coupling="OS", scheme="MUSL", va="a", progressbar=False,
               gravity=-9.8, ζs=0, ζw=0, αT=0.5, iter_num=0, end_time=0, start_time=0,
               project_name, project_path, ext=0, ϵ="FP64", sort_interval=None):
    T1, T2, eps = _types(ϵ)
    # parameter checks of modelargs.jl:76-91 / 167-182
    if not 0 < Ttol:
        raise ValueError("Simulation time cannot be ≤0 s.")
    if not ΔT <= Ttol:
        raise ValueError(f"Time step cannot be >{Ttol}s.")
    if not 0 <= Te <= Ttol:
        raise ValueError(f"Elastic loading time must in 0~{Ttol}s.")
    if basis not in ("uGIMP", "linear", "gslinear"):
        raise ValueError(f"Cannot find {basis} basis function.")
    if (FLIP + PIC) != 1:
        raise ValueError("FLIP+PIC must be 1.")
    if not 0 <= αT <= 1:
        raise ValueError("Time step factor error.")
    if not 0 <= ζs <= 1:
        raise ValueError("Out of range 0≤ζs≤1.")
    if not 0 <= ζw <= 1:
        raise ValueError("Out of range 0≤ζw≤1.")
    if project_name == "":
        raise ValueError("Empty project name.")
    if coupling not in ("OS", "TS"):
        raise ValueError("Coupling mode is wrong.")
    if device not in _DEV_SET:
        raise ValueError(f"Cannot find {device} device.")
    if time_step not in ("fixed", "auto"):
        raise ValueError(f"{time_step} time step is not allowed.")
    if va not in ("v", "a"):
        raise ValueError(f"Cannot find {va} velocity update mode.")
    if animation and not hdf5:
        hdf5 = True  # "HDF5 forced ON due to the animation"
    if hdf5 and hdf5_step <= 0:
        raise ValueError("HDF5 step cannot be ≤0.")
    return Args(dim, T1, T2, eps, Ttol=T2(Ttol), Te=T2(Te), ΔT=T2(ΔT), time_step=time_step,
                FLIP=T2(FLIP), PIC=T2(PIC), constitutive=constitutive, basis=basis,
                animation=animation, hdf5=hdf5, hdf5_step=int(hdf5_step), MVL=bool(MVL),
                device=device, coupling=coupling, scheme=scheme, va=va,
                progressbar=progressbar, gravity=T2(gravity), ζs=T2(ζs), ζw=T2(ζw),
                αT=T2(αT), iter_num=int(iter_num), end_time=float(end_time),
                start_time=float(start_time), project_name=project_name,
                project_path=project_path, ext=ext, sort_interval=sort_interval)


def UserArgs2D(**kw):
    """UserArgs2D (src/types/modelargs.jl:61-98)."""
    return _user_args(2, **kw)


def UserArgs3D(**kw):
    """UserArgs3D (src/types/modelargs.jl:152-189)."""
    return _user_args(3, **kw)


# --------------------------------------------------------------------------- ranges
def _rat(x: float):
    """Small-denominator rational that reproduces x exactly as a float, if any."""
    if x == 0.0:
        return Fraction(0)
    f = Fraction(x).limit_denominator(1 << 24)
    return f if float(f) == x else None


def julia_range(a: float, step: float, b: float) -> np.ndarray:
    """Elements of the Julia float range ``a:step:b``.

    Julia evaluates ranges of decimal-looking floats through exact rationals
    (``Base.(:)`` for IEEEFloat, base/twiceprecision.jl), so element k is the
    correctly rounded value of (start_n + k*step_n)/den.  That is restated with
    Python's Fraction; other inputs fall back to ``a + k*step``.
    """
    a, step, b = float(a), float(step), float(b)
    ra, rs, rb = _rat(a), _rat(step), _rat(b)
    if ra is not None and rs is not None and rb is not None and rs != 0:
        n = int((rb - ra) // rs) + 1
        return np.array([float(ra + k * rs) for k in range(max(n, 0))], dtype=np.float64)
    n = int(np.floor((b - a) / step + 1e-10)) + 1
    return a + np.arange(max(n, 0), dtype=np.float64) * step


# --------------------------------------------------------------------------- Grid
class Grid:
    """Grid2D / Grid3D (src/types/grid.jl:33-75, 170-217), host side."""

    def __init__(self, dim, T1, T2, eps, **kw):
        self.dim = dim
        self.T1, self.T2, self.ϵ = T1, T2, eps
        for k, v in kw.items():
            setattr(self, k, v)

    def __repr__(self):
        return (f"DeviceGrid{self.dim}D(NIC={self.NIC}, ϵ={self.ϵ}, nc={self.nc:.2e}, "
                f"ni={self.ni:.2e})")


def UserGrid2D(*, ϵ="FP64", phase=1, x1, x2, y1, y2, dx, dy, NIC=9, ext=0):
    """UserGrid2D (src/types/grid.jl:79-145): nodes in columns from x1, each column
    from the top (y descending), node n = row + (col-1)*nny."""
    if not x1 < x2:
        raise ArgumentError("x1 should be less than x2")
    if not y1 < y2:
        raise ArgumentError("y1 should be less than y2")
    if not (dx > 0 and dy > 0):
        raise ArgumentError("dx and dy should be positive")
    phase = phase if phase in (1, 2) else 1
    NIC = NIC if NIC in (4, 9) else 9
    T1, T2, eps = _types(ϵ)
    vx = julia_range(x1, dx, x2)
    vy = julia_range(y1, dy, y2)[::-1]
    nnx, nny = len(vx), len(vy)
    ni = nnx * nny
    xi = np.empty((ni, 2), dtype=np.float64, order="F")
    xi[:, 0] = np.repeat(vx, nny)
    xi[:, 1] = np.tile(vy, nnx)
    ncx, ncy = nnx - 1, nny - 1
    if NIC == 4:
        p2nD = np.array([[0, 1], [0, 0], [nny, 1], [nny, 0]], dtype=np.int64)
    else:
        p2nD = np.array([[-nny, 2], [-nny, 1], [-nny, 0], [-nny, -1],
                         [0, 2], [0, 1], [0, 0], [0, -1],
                         [nny, 2], [nny, 1], [nny, 0], [nny, -1],
                         [2 * nny, 2], [2 * nny, 1], [2 * nny, 0], [2 * nny, -1]], dtype=np.int64)
    return Grid(2, T1, T2, eps, phase=phase, x1=T2(x1), x2=T2(x2), y1=T2(y1), y2=T2(y2),
                dx=T2(dx), dy=T2(dy), nnx=nnx, nny=nny, ni=ni, NIC=NIC,
                ξ=xi.astype(T2, order="F"), ncx=ncx, ncy=ncy, nc=ncx * ncy,
                p2nD=np.asfortranarray(p2nD), ext=ext)


# The 64 uGIMP candidate offsets of UserGrid3D (src/types/grid.jl:289-301):
# (layer, column, row, constant) offsets, y-groups slowest.
_P2ND_3D = [
    (-1, -1, -1, 0), (-1, 0, -1, 0), (0, 0, -1, 0), (0, -1, -1, 0), (0, -2, -1, 0), (-1, -2, -1, 0),
    (-2, -2, -1, 0), (-2, -1, -1, 0), (-2, 0, -1, 0), (-2, 1, -1, 0), (-1, 1, -1, 0), (0, 1, -1, 0),
    (1, 1, -1, 0), (1, 0, -1, 0), (1, -1, -1, 0), (1, -2, -1, 0), (-1, -1, -1, 1), (-1, 0, -1, 1),
    (0, 0, -1, 1), (0, -1, -1, 1), (0, -2, -1, 1), (-1, -2, -1, 1), (-2, -2, -1, 1), (-2, -1, -1, 1),
    (-2, 0, -1, 1), (-2, 1, -1, 1), (-1, 1, -1, 1), (0, 1, -1, 1), (1, 1, -1, 1), (1, 0, -1, 1),
    (1, -1, -1, 1), (1, -2, -1, 1), (-1, -1, 0, 1), (-1, 0, 0, 1), (0, 0, 0, 1), (0, -1, 0, 1),
    (0, -2, 0, 1), (-1, -2, 0, 1), (-2, -2, 0, 1), (-2, -1, 0, 1), (-2, 0, 0, 1), (-2, 1, 0, 1),
    (-1, 1, 0, 1), (0, 1, 0, 1), (1, 1, 0, 1), (1, 0, 0, 1), (1, -1, 0, 1), (1, -2, 0, 1),
    (-1, -1, 0, 2), (-1, 0, 0, 2), (0, 0, 0, 2), (0, -1, 0, 2), (0, -2, 0, 2), (-1, -2, 0, 2),
    (-2, -2, 0, 2), (-2, -1, 0, 2), (-2, 0, 0, 2), (-2, 1, 0, 2), (-1, 1, 0, 2), (0, 1, 0, 2),
    (1, 1, 0, 2), (1, 0, 0, 2), (1, -1, 0, 2), (1, -2, 0, 2),
]


def UserGrid3D(*, ϵ="FP64", phase=1, x1, x2, y1, y2, z1, z2, dx, dy, dz, NIC=27, ext=0):
    """UserGrid3D (src/types/grid.jl:221-309): node n = iy + (ix-1)*nny + (iz-1)*nny*nnx,
    y fastest, all axes ascending."""
    if not x1 < x2:
        raise ArgumentError("x1 should be less than x2")
    if not y1 < y2:
        raise ArgumentError("y1 should be less than y2")
    if not z1 < z2:
        raise ArgumentError("z1 should be less than z2")
    if not (dx > 0 and dy > 0 and dz > 0):
        raise ArgumentError("dx, dy, and dz should be positive")
    phase = phase if phase in (1, 2) else 1
    NIC = NIC if NIC in (8, 27) else 27
    T1, T2, eps = _types(ϵ)
    vx = julia_range(x1, dx, x2)
    vy = julia_range(y1, dy, y2)
    vz = julia_range(z1, dz, z2)
    nnx, nny, nnz = len(vx), len(vy), len(vz)
    ni = nnx * nny * nnz
    xi = np.empty((ni, 3), dtype=np.float64, order="F")
    xi[:, 0] = np.tile(np.repeat(vx, nny), nnz)
    xi[:, 1] = np.tile(vy, nnx * nnz)
    xi[:, 2] = np.repeat(vz, nnx * nny)
    if NIC == 8:
        p2nD = np.array([[-1, -1, -1], [-1, 0, -1], [0, 0, -1], [0, -1, -1],
                         [-1, -1, 0], [-1, 0, 0], [0, 0, 0], [0, -1, 0]], dtype=np.int64)
    else:
        p2nD = np.array(_P2ND_3D, dtype=np.int64)
    ncx, ncy, ncz = nnx - 1, nny - 1, nnz - 1
    return Grid(3, T1, T2, eps, phase=phase, x1=T2(x1), x2=T2(x2), y1=T2(y1), y2=T2(y2),
                z1=T2(z1), z2=T2(z2), dx=T2(dx), dy=T2(dy), dz=T2(dz), nnx=nnx, nny=nny,
                nnz=nnz, ni=ni, NIC=NIC, ξ=xi.astype(T2, order="F"), ncx=ncx, ncy=ncy,
                ncz=ncz, nc=ncx * ncy * ncz, p2nD=np.asfortranarray(p2nD), ext=ext)


# --------------------------------------------------------------------------- Particle
class Particle:
    """Particle2D / Particle3D (src/types/particle.jl:33-81, 160-210), host side."""

    def __init__(self, dim, T1, T2, eps, **kw):
        self.dim = dim
        self.T1, self.T2, self.ϵ = T1, T2, eps
        for k, v in kw.items():
            setattr(self, k, v)

    def __repr__(self):
        return f"DeviceParticle{self.dim}D(NIC={self.NIC}, ϵ={self.ϵ}, np={self.np:.2e})"


def _user_particle(dim, ϵ, phase, NIC, spacing, ξ, ρs):
    T1, T2, eps = _types(ϵ)
    ξ = np.asarray(ξ, dtype=np.float64)
    if ξ.ndim != 2 or ξ.shape[1] != dim:
        raise ArgumentError(f"ξ should have {dim} columns")
    n = ξ.shape[0]
    vol = float(np.prod(spacing))
    ns, nf = (4, 4) if dim == 2 else (6, 9)
    F = np.zeros((n, nf), dtype=T2, order="F")
    for k in ((0, 3) if dim == 2 else (0, 4, 8)):
        F[:, k] = 1
    rho = np.broadcast_to(np.asarray(ρs, dtype=T2), (n,)).copy()
    z1 = lambda: np.zeros(n, dtype=T2)
    zk = lambda k: np.zeros((n, k), dtype=T2, order="F")
    kw = dict(phase=phase if phase in (1, 2) else 1, np=n, NIC=NIC,
              ξ=np.asfortranarray(ξ.astype(T2)), ξ0=np.asfortranarray(ξ.astype(T2)),
              σm=z1(), ϵq=z1(), ϵk=z1(), ϵv=z1(), Ω=np.full(n, vol, dtype=T2),
              Ω0=np.full(n, vol, dtype=T2), ms=z1(), ρs=rho, ρs0=rho.copy(), σij=zk(ns),
              ϵijs=zk(ns), vs=zk(dim), F=F)
    if dim == 2:
        kw.update(dx=T2(spacing[0]), dy=T2(spacing[1]))
    else:
        kw.update(dx=T2(spacing[0]), dy=T2(spacing[1]), dz=T2(spacing[2]))
    return Particle(dim, T1, T2, eps, **kw)


def UserParticle2D(*, ϵ="FP64", phase=1, NIC=9, dx, dy, ξ, n=(0,), ρs, ρw=(0,), ext=0):
    """UserParticle2D (src/types/particle.jl:85-138): Ω0 = dx*dy, F = I."""
    if not (dx > 0 and dy > 0):
        raise ArgumentError("dx and dy should be positive")
    NIC = NIC if NIC in (4, 9) else 9
    return _user_particle(2, ϵ, phase, NIC, (dx, dy), ξ, ρs)


def UserParticle3D(*, ϵ="FP64", phase=1, NIC=27, dx, dy, dz, ξ, n=(0,), ρs, ρw=(0,), ext=0):
    """UserParticle3D (src/types/particle.jl:214-271): Ω0 = dx*dy*dz, F = I."""
    if not (dx > 0 and dy > 0 and dz > 0):
        raise ArgumentError("dx, dy, and dz should be positive")
    NIC = NIC if NIC in (8, 27) else 27
    return _user_particle(3, ϵ, phase, NIC, (dx, dy, dz), ξ, ρs)


# --------------------------------------------------------------------------- Property
class Property:
    """Property (src/types/property.jl:31-52), host side."""

    def __init__(self, T1, T2, eps, **kw):
        self.T1, self.T2, self.ϵ = T1, T2, eps
        for k, v in kw.items():
            setattr(self, k, v)

    @property
    def nmat(self):
        return len(self.Es)


def UserProperty(*, ϵ="FP64", nid, ν, Es, Gs, Ks, Kw=(0,), k=(0,), σt=(0,), ϕ=(0,), ϕr=(0,),
                 ψ=(0,), c=(0,), cr=(0,), Hp=(0,), ext=0):
    """UserProperty (src/types/property.jl:57-72)."""
    T1, T2, eps = _types(ϵ)
    if not (len(Es) == len(Gs) == len(Ks) == len(ν)):
        raise ArgumentError("The length of Es, Gs, Ks, and ν must be the same.")
    nid_f = np.asarray(nid, dtype=np.float64)
    if not np.all(nid_f == np.round(nid_f)):
        raise ArgumentError("nid must hold integer material ids")
    nid_i = nid_f.astype(np.int64)
    if len(np.unique(nid_i)) != len(Es):
        raise ArgumentError("nid layer must be the same as the length of properties.")
    nm = len(Es)

    def vec(v, name):
        a = np.asarray(v, dtype=np.float64).ravel()
        if a.size == 1 and nm > 1:
            a = np.full(nm, a[0])
        if a.size != nm:
            raise ArgumentError(f"{name} must have one entry per material ({nm})")
        return a.astype(T2)

    return Property(T1, T2, eps, tmp1=0, tmp2=0.0, nid=nid_i.astype(T1), ν=vec(ν, "ν"),
                    Es=vec(Es, "Es"), Gs=vec(Gs, "Gs"), Ks=vec(Ks, "Ks"), Kw=vec(Kw, "Kw"),
                    k=vec(k, "k"), σt=vec(σt, "σt"), ϕ=vec(ϕ, "ϕ"), ϕr=vec(ϕr, "ϕr"),
                    ψ=vec(ψ, "ψ"), c=vec(c, "c"), cr=vec(cr, "cr"), Hp=vec(Hp, "Hp"), ext=ext)


# --------------------------------------------------------------------------- Boundary
class VBoundary:
    """VBoundary2D / VBoundary3D (src/types/boundary.jl:115-131, 169-190)."""

    def __init__(self, dim, T1, T2, eps, **kw):
        self.dim = dim
        self.T1, self.T2, self.ϵ = T1, T2, eps
        for k, v in kw.items():
            setattr(self, k, v)


def _ivec(a, T1):
    a = np.asarray(a, dtype=np.float64).ravel()
    return a.astype(T1)


def UserVBoundary2D(*, ϵ="FP64", vx_s_idx, vx_s_val, vy_s_idx, vy_s_val, vx_w_idx=(0,),
                    vx_w_val=(0,), vy_w_idx=(0,), vy_w_val=(0,), smlength=0, tmp1=0, tmp2=0,
                    ext=0):
    """UserVBoundary2D (src/types/boundary.jl:136-151)."""
    T1, T2, eps = _types(ϵ)
    if not (len(vx_s_idx) == len(vx_s_val) and len(vy_s_idx) == len(vy_s_val) and
            len(vx_w_idx) == len(vx_w_val) and len(vy_w_idx) == len(vy_w_val)):
        raise ArgumentError("The length of vx_s_idx, vx_s_val, vy_s_idx, vy_s_val, vx_w_idx, "
                            "vx_w_val, vy_w_idx, vy_s_val should be the same.")
    return VBoundary(2, T1, T2, eps, vx_s_idx=_ivec(vx_s_idx, T1),
                     vx_s_val=np.asarray(vx_s_val, dtype=T2).ravel(),
                     vy_s_idx=_ivec(vy_s_idx, T1), vy_s_val=np.asarray(vy_s_val, dtype=T2).ravel(),
                     vx_w_idx=_ivec(vx_w_idx, T1), vx_w_val=np.asarray(vx_w_val, dtype=T2).ravel(),
                     vy_w_idx=_ivec(vy_w_idx, T1), vy_w_val=np.asarray(vy_w_val, dtype=T2).ravel(),
                     smlength=T2(smlength), tmp1=tmp1, tmp2=tmp2, ext=ext)


def UserVBoundary3D(*, ϵ="FP64", vx_s_idx, vx_s_val, vy_s_idx, vy_s_val, vz_s_idx, vz_s_val,
                    vx_w_idx=(0,), vx_w_val=(0,), vy_w_idx=(0,), vy_w_val=(0,), vz_w_idx=(0,),
                    vz_w_val=(0,), smlength=0, tmp1=0, tmp2=0, ext=0):
    """UserVBoundary3D (src/types/boundary.jl:194-213)."""
    T1, T2, eps = _types(ϵ)
    if not (len(vx_s_idx) == len(vx_s_val) and len(vy_s_idx) == len(vy_s_val) and
            len(vz_s_idx) == len(vz_s_val) and len(vz_w_idx) == len(vz_w_val) and
            len(vx_w_idx) == len(vx_w_val) and len(vy_w_idx) == len(vy_w_val)):
        raise ArgumentError("The length of vx_s_idx, vx_s_val, vy_s_idx, vy_s_val, vz_s_idx, "
                            "vz_s_val, vx_w_idx, vx_w_val, vy_w_idx, vy_s_val, vz_w_idx, "
                            "vz_w_val should be the same.")
    f = lambda a: np.asarray(a, dtype=T2).ravel()
    return VBoundary(3, T1, T2, eps, vx_s_idx=_ivec(vx_s_idx, T1), vx_s_val=f(vx_s_val),
                     vy_s_idx=_ivec(vy_s_idx, T1), vy_s_val=f(vy_s_val),
                     vz_s_idx=_ivec(vz_s_idx, T1), vz_s_val=f(vz_s_val),
                     vx_w_idx=_ivec(vx_w_idx, T1), vx_w_val=f(vx_w_val),
                     vy_w_idx=_ivec(vy_w_idx, T1), vy_w_val=f(vy_w_val),
                     vz_w_idx=_ivec(vz_w_idx, T1), vz_w_val=f(vz_w_val),
                     smlength=T2(smlength), tmp1=tmp1, tmp2=tmp2, ext=ext)


# --------------------------------------------------------------------------- meshbuilder
def meshbuilder(*axes) -> np.ndarray:
    """meshbuilder (src/toolkits/meshbuilder.jl:16-37): lattice points, y fastest,
    then x, then z.  Each axis is an array (use julia_range for a:step:b)."""
    if len(axes) == 2:
        x, y = (np.asarray(a, dtype=np.float64) for a in axes)
        return np.column_stack([np.repeat(x, len(y)), np.tile(y, len(x))])
    if len(axes) == 3:
        x, y, z = (np.asarray(a, dtype=np.float64) for a in axes)
        nx, ny, nz = len(x), len(y), len(z)
        return np.column_stack([np.tile(np.repeat(x, ny), nz), np.tile(y, nx * nz),
                                np.repeat(z, nx * ny)])
    raise ArgumentError("meshbuilder takes 2 or 3 axes")
