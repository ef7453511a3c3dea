"""ctypes driver of liboracle.so, the FP64 CPU restatement of the reference step.

TEST INFRASTRUCTURE ONLY: tests, the smoke check and bench.py's cpu_baseline
leg use it as the checker / CPU baseline.  The product path never imports it.

``Oracle(args, grid, mp, attr, bc)`` copies the host-mirror structs into the
reference's full per-particle layout (including the stored per-slot basis
arrays Nij / dN / p2n) and exposes the whole step (``procedure``), the
submit_work! loop (``run``) and each restated kernel for unit tests.

Pinning status of each restated function (see tests/test_oracle_*.py):
  oracle_fld / oracle_cld     pinned: brute force with exact rationals + the
                              Julia manual's example fld(6.0, 0.1) == 59
  oracle_ugimp_basis          pinned: closed form at branch points, continuity,
                              partition of unity, derivative by finite difference
  oracle_linear_basis         pinned: closed form
  oracle_p2c / oracle_p2n     pinned: brute-force nearest-node scan on small
                              grids (2D and 3D), candidate set == support set
  oracle_resetmp              pinned: partition of unity, sum N x_n = x_p,
                              sum dN = 0 for interior particles
  oracle_p2g                  pinned: mass / momentum / gravity sums exact,
                              internal force sums to zero, one-particle hand case
  oracle_solvegrid, _dm3      pinned: closed form per node incl. BC-on-momentum,
                              damping sign and the massless cut
  oracle_dm1, _dm2            pinned: rigid translation, momentum conservation
  oracle_g2p                  pinned: affine velocity field -> exact ΔF, det F
  oracle_liE                  pinned: closed-form isotropic increment, spin
                              rotation preserves invariants to O(ω²)
  oracle_dpP                  pinned: post-return yield residual == 0, elastic
                              states unchanged, tension apex
  oracle_mcP (2D)             pinned: post-return f <= 0, closed-form apex
  oracle_mcP (3D) and         build-defined (the reference's 3D mcP! is a no-op):
  oracle_sym3_eig,            pinned by eigh, f <= 0 on all six planes, the
  oracle_mc3_*                energy-norm closest point (associative flow) found
                              by active-set enumeration, flow-cone membership
                              (non-associative), apex, rotation invariance;
                              parity with the reference unpinned
  oracle_vollock              pinned: cell-mean identity
  whole step                  pinned: elastic column under gravity vs
                              σzz = -ρ g (H - z); experimental deposit profile
                              of the 2D collapse (tolerance level)
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int64)


class OracleModel(C.Structure):
    _fields_ = [
        ("dim", C.c_int32), ("nic", C.c_int32), ("basis", C.c_int32), ("va", C.c_int32),
        ("constitutive", C.c_int32), ("mvl", C.c_int32), ("nthreads", C.c_int32),
        ("pad0", C.c_int32),
        ("flip", C.c_double), ("pic", C.c_double), ("zeta_s", C.c_double),
        ("gravity", C.c_double), ("Te", C.c_double),
        ("nnx", C.c_int64), ("nny", C.c_int64), ("nnz", C.c_int64), ("ncx", C.c_int64),
        ("ncy", C.c_int64), ("ncz", C.c_int64), ("ni", C.c_int64), ("nc", C.c_int64),
        ("x1", C.c_double), ("y1", C.c_double), ("z1", C.c_double),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
        ("gxi", _D), ("p2nD", _I), ("p2nD_rows", C.c_int64),
        ("g_ms", _D), ("g_ps", _D), ("g_fs", _D), ("g_vs", _D), ("g_vsT", _D), ("g_as", _D),
        ("g_dus", _D), ("g_sigm", _D), ("g_omega", _D),
        ("np", C.c_int64), ("lpx", C.c_double), ("lpy", C.c_double), ("lpz", C.c_double),
        ("xi", _D), ("vs", _D), ("ps", _D), ("sig", _D), ("sij", _D), ("eps_s", _D),
        ("deps_s", _D), ("dF", _D), ("F", _D), ("sigm", _D), ("eps_q", _D), ("eps_k", _D),
        ("eps_v", _D), ("omega", _D), ("omega0", _D), ("ms", _D), ("rho", _D), ("rho0", _D),
        ("p2c", _I), ("p2n", _I), ("Nij", _D), ("dNx", _D), ("dNy", _D), ("dNz", _D),
        ("nid", _I), ("Ks", _D), ("Gs", _D), ("c", _D), ("phi", _D), ("psi", _D),
        ("sigt", _D), ("Hp", _D), ("cr", _D), ("phir", _D),
        ("vx_idx", _I), ("vy_idx", _I), ("vz_idx", _I), ("vx_val", _D), ("vy_val", _D),
        ("vz_val", _D),
        ("gravity_x", C.c_double), ("gravity_y", C.c_double), ("Ti_now", C.c_double),
        ("scheme", C.c_int32), ("pad1", C.c_int32),
    ]


_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} missing: run `make -C oracle`")
        lib = C.CDLL(LIB_PATH)
        P = C.POINTER(OracleModel)
        for name in ("oracle_resetgrid", "oracle_resetmp", "oracle_initmp", "oracle_dm2",
                     "oracle_liE", "oracle_dpP", "oracle_mcP", "oracle_vollock"):
            getattr(lib, name).argtypes = [P]
            getattr(lib, name).restype = None
        for name in ("oracle_solvegrid", "oracle_dm1", "oracle_dm3", "oracle_g2p",
                     "oracle_p2g"):
            getattr(lib, name).argtypes = [P, C.c_double]
            getattr(lib, name).restype = None
        lib.oracle_procedure.argtypes = [P, C.c_double, C.c_double]
        lib.oracle_procedure.restype = None
        lib.oracle_run.argtypes = [P, C.c_double, C.c_double, C.c_int64]
        lib.oracle_run.restype = C.c_double
        lib.oracle_fld.argtypes = [C.c_double, C.c_double]
        lib.oracle_fld.restype = C.c_double
        lib.oracle_cld.argtypes = [C.c_double, C.c_double]
        lib.oracle_cld.restype = C.c_double
        lib.oracle_p2c.argtypes = [P, C.c_double, C.c_double, C.c_double]
        lib.oracle_p2c.restype = C.c_int64
        lib.oracle_p2n.argtypes = [P, C.c_int64, C.c_int64]
        lib.oracle_p2n.restype = C.c_int64
        for name in ("oracle_ugimp_basis",):
            getattr(lib, name).argtypes = [C.c_double, C.c_double, C.c_double, _D, _D]
            getattr(lib, name).restype = None
        lib.oracle_linear_basis.argtypes = [C.c_double, C.c_double, _D, _D]
        lib.oracle_linear_basis.restype = None
        lib.oracle_sym3_eig.argtypes = [_D, _D, _D]
        lib.oracle_sym3_eig.restype = None
        lib.oracle_mc3_principal.argtypes = [_D] + [C.c_double] * 6
        lib.oracle_mc3_principal.restype = C.c_int
        lib.oracle_mc3_point.argtypes = [_D, _D, _D] + [C.c_double] * 8
        lib.oracle_mc3_point.restype = C.c_int
        lib.oracle_mc_soften.argtypes = [C.c_double] * 7 + [_D] * 3
        lib.oracle_mc_soften.restype = None
        lib.oracle_struct_size.restype = C.c_int
        if lib.oracle_struct_size() != C.sizeof(OracleModel):
            raise ImportError("oracle_model layout mismatch")
        _lib = lib
    return _lib


def fld(x, y):
    return load().oracle_fld(float(x), float(y))


def cld(x, y):
    return load().oracle_cld(float(x), float(y))


def ugimp_basis(d, h, lp):
    N, dN = C.c_double(), C.c_double()
    load().oracle_ugimp_basis(float(d), float(h), float(lp), C.byref(N), C.byref(dN))
    return N.value, dN.value


def linear_basis(d, h):
    N, dN = C.c_double(), C.c_double()
    load().oracle_linear_basis(float(d), float(h), C.byref(N), C.byref(dN))
    return N.value, dN.value


def sym3_eig(s6):
    """Eigenvalues (descending) and eigenvectors (columns) of [xx,yy,zz,xy,yz,xz]."""
    s = np.ascontiguousarray(s6, dtype=np.float64)
    lam = np.zeros(3)
    V = np.zeros(9)
    load().oracle_sym3_eig(s.ctypes.data_as(_D), lam.ctypes.data_as(_D), V.ctypes.data_as(_D))
    return lam, V.reshape(3, 3)


def mc3_principal(p3, Dt, Dd, sphi, cphi, spsi, c):
    """Principal-space Mohr-Coulomb return; returns (p_new, kind)."""
    p = np.ascontiguousarray(p3, dtype=np.float64).copy()
    kind = load().oracle_mc3_principal(p.ctypes.data_as(_D), Dt, Dd, sphi, cphi, spsi, c)
    return p, kind


def mc3_point(s6, eq, Ks, Gs, phi, psi, c, Hp=0.0, cr=0.0, phir=0.0):
    """Build-defined 3D Mohr-Coulomb on one stress; returns (s, eq, ek, kind)."""
    s = np.ascontiguousarray(s6, dtype=np.float64).copy()
    q, k = C.c_double(eq), C.c_double(0.0)
    kind = load().oracle_mc3_point(s.ctypes.data_as(_D), C.byref(q), C.byref(k), Ks, Gs, phi,
                                   psi, c, Hp, cr, phir)
    return s, q.value, k.value, kind


def mc_soften(eq, phi, psi, c, Hp, cr, phir):
    """(phi, psi, c) in effect at plastic strain eq (3D Mohr-Coulomb softening)."""
    o = [C.c_double(0.0) for _ in range(3)]
    load().oracle_mc_soften(eq, phi, psi, c, Hp, cr, phir, *[C.byref(v) for v in o])
    return tuple(v.value for v in o)


_CONST = {"linearelastic": 0, "druckerprager": 1, "mohrcoulomb": 2}


class Oracle:
    """FP64 reference-layout model built from the host-mirror structs."""

    def __init__(self, args, grid, mp, attr, bc, nthreads=1):
        self.lib = load()
        D = grid.dim
        n = int(mp.np)
        ni = int(grid.ni)
        nc = int(grid.nc)
        nic = int(mp.NIC)
        ns, nf = (4, 4) if D == 2 else (6, 9)
        f = lambda a: np.asfortranarray(np.array(a, dtype=np.float64, copy=True))
        i = lambda a: np.ascontiguousarray(np.array(a, dtype=np.int64, copy=True))
        zf = lambda *s: np.zeros(s, dtype=np.float64, order="F")
        self.D, self.np, self.ni, self.nic = D, n, ni, nic
        a = self.a = dict(
            gxi=f(grid.ξ), p2nD=np.asfortranarray(np.array(grid.p2nD, dtype=np.int64)),
            g_ms=zf(ni), g_ps=zf(ni, D), g_fs=zf(ni, D), g_vs=zf(ni, D), g_vsT=zf(ni, D),
            g_as=zf(ni, D), g_dus=zf(ni, D), g_sigm=zf(max(nc, 1)), g_omega=zf(max(nc, 1)),
            xi=f(mp.ξ), vs=f(mp.vs), ps=zf(n, D), sig=f(mp.σij), sij=zf(n, ns), eps_s=f(mp.ϵijs),
            deps_s=zf(n, ns), dF=zf(n, nf), F=f(mp.F), sigm=f(mp.σm), eps_q=f(mp.ϵq),
            eps_k=f(mp.ϵk), eps_v=f(mp.ϵv), omega=f(mp.Ω), omega0=f(mp.Ω0), ms=f(mp.ms),
            rho=f(mp.ρs), rho0=f(mp.ρs0), p2c=np.zeros(n, dtype=np.int64),
            p2n=np.zeros((n, nic), dtype=np.int64, order="F"), Nij=zf(n, nic), dNx=zf(n, nic),
            dNy=zf(n, nic), dNz=zf(n, nic) if D == 3 else zf(1, 1),
            nid=i(attr.nid), Ks=f(attr.Ks), Gs=f(attr.Gs), c=f(attr.c), phi=f(attr.ϕ),
            psi=f(attr.ψ), sigt=f(attr.σt), Hp=f(attr.Hp), cr=f(attr.cr), phir=f(attr.ϕr),
            vx_idx=i(bc.vx_s_idx), vy_idx=i(bc.vy_s_idx),
            vz_idx=i(bc.vz_s_idx) if D == 3 else np.zeros(1, dtype=np.int64),
            vx_val=f(bc.vx_s_val), vy_val=f(bc.vy_s_val),
            vz_val=f(bc.vz_s_val) if D == 3 else zf(1))
        m = self.m = OracleModel()
        m.dim, m.nic = D, nic
        m.basis = 1 if grid.NIC in (4, 8) else 0
        m.va = 0 if args.va == "a" else 1
        m.scheme = {"MUSL": 0, "USL": 1, "USF": 2}[args.scheme]
        m.constitutive = _CONST[args.constitutive]
        m.mvl = 1 if args.MVL else 0
        m.nthreads = int(nthreads)
        m.flip, m.pic, m.zeta_s = float(args.FLIP), float(args.PIC), float(args.ζs)
        m.gravity, m.Te = float(args.gravity), float(args.Te)
        ext = args.ext if isinstance(args.ext, dict) else {}
        m.gravity_x = float(ext.get("gravity_x", 0.0))
        m.gravity_y = float(ext.get("gravity_y", 0.0))
        m.nnx, m.nny = int(grid.nnx), int(grid.nny)
        m.nnz = int(grid.nnz) if D == 3 else 1
        m.ncx, m.ncy = int(grid.ncx), int(grid.ncy)
        m.ncz = int(grid.ncz) if D == 3 else 1
        m.ni, m.nc = ni, nc
        m.x1, m.y1, m.dx, m.dy = float(grid.x1), float(grid.y1), float(grid.dx), float(grid.dy)
        m.z1 = float(grid.z1) if D == 3 else 0.0
        m.dz = float(grid.dz) if D == 3 else 1.0
        m.p2nD_rows = a["p2nD"].shape[0]
        m.np = n
        m.lpx, m.lpy = float(mp.dx), float(mp.dy)
        m.lpz = float(mp.dz) if D == 3 else 1.0
        for k, v in a.items():
            if v.dtype == np.int64:
                setattr(m, k, v.ctypes.data_as(_I))
            else:
                setattr(m, k, v.ctypes.data_as(_D))
        self._p = C.byref(m)
        # initmpstatus! before the loop (src/solver.jl:45): the basis that the
        # first USF step's G2P reads
        self.initmp()

    # whole step / loop
    def procedure(self, dt, Ti):
        self.lib.oracle_procedure(self._p, float(dt), float(Ti))

    def run(self, dt, Ti, nsteps):
        return self.lib.oracle_run(self._p, float(dt), float(Ti), int(nsteps))

    # individual restated kernels
    def resetgrid(self):
        self.lib.oracle_resetgrid(self._p)

    def resetmp(self):
        self.lib.oracle_resetmp(self._p)

    def initmp(self):
        self.lib.oracle_initmp(self._p)

    def p2g(self, G):
        self.lib.oracle_p2g(self._p, float(G))

    def solvegrid(self, dt):
        self.lib.oracle_solvegrid(self._p, float(dt))

    def dm1(self, dt):
        self.lib.oracle_dm1(self._p, float(dt))

    def dm2(self):
        self.lib.oracle_dm2(self._p)

    def dm3(self, dt):
        self.lib.oracle_dm3(self._p, float(dt))

    def g2p(self, dt):
        self.lib.oracle_g2p(self._p, float(dt))

    def liE(self):
        self.lib.oracle_liE(self._p)

    def dpP(self):
        self.lib.oracle_dpP(self._p)

    def mcP(self):
        self.lib.oracle_mcP(self._p)

    def vollock(self):
        self.lib.oracle_vollock(self._p)

    def p2c(self, x, y, z=0.0):
        return self.lib.oracle_p2c(self._p, float(x), float(y), float(z))

    def p2n(self, p2c, iy):
        return self.lib.oracle_p2n(self._p, int(p2c), int(iy))

    def state(self):
        """The device2host! field set (copies), keyed like HIPModel.download()."""
        a = self.a
        return dict(xi=a["xi"].copy(), vs=a["vs"].copy(), sigma=a["sig"].copy(),
                    sigma_m=a["sigm"].copy(), omega=a["omega"].copy(), eps_q=a["eps_q"].copy(),
                    eps_k=a["eps_k"].copy(), eps_v=a["eps_v"].copy(), eps_s=a["eps_s"].copy(),
                    ms=a["ms"].copy())
