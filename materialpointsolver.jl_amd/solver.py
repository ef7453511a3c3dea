"""Solver entry and the HIP device plugin, mirroring MaterialPointSolver.jl.

* ``materialpointsolver``  -- materialpointsolver! (src/MaterialPointSolver.jl:61-75)
* ``submit_work``          -- submit_work! (src/solver.jl:37-131)
* ``procedure``            -- procedure!(..., Val(:OS), Val(:MUSL), Val(:a|:v))
                              (src/solvers/OS_MUSL.jl:12-102), one libmpm_hip step
* ``host2device`` / ``device2host`` / ``clean_device`` / ``getBackend`` / ``Tpeak``
                           -- the device plugin of ext/AMDExt/devicehelpfunc_amd.jl,
                              for the new device symbol ``"HIP"``

Julia dispatches these on ``Val(args.device)``; here the device string is an
argument.  Only ``"HIP"`` has an implementation: every other device raises.
"""
from __future__ import annotations

import os
import time

import numpy as np

from . import _capi as capi


class HIPModel:
    """Owner of one libmpm_hip context (the device copy of grid, mp, attr, bc)."""

    def __init__(self, grid, mp, attr, bc, sort_interval=0):
        lib = capi.load()
        self.lib = lib
        self.dim = grid.dim
        self.np = int(mp.np)
        if grid.dim != mp.dim or grid.dim != bc.dim:
            raise ValueError("grid, mp and bc must have the same dimension")
        if grid.ϵ != mp.ϵ:
            raise ValueError("grid and mp must use the same ϵ")
        cfg = capi.MPMConfig()
        cfg.dim = grid.dim
        cfg.precision = capi.MPM_FP32 if mp.ϵ == "FP32" else capi.MPM_FP64
        lin = grid.NIC in (4, 8)
        if (mp.NIC in (4, 8)) != lin:
            raise ValueError("grid.NIC and mp.NIC disagree on the basis")
        cfg.basis = capi.MPM_LINEAR if lin else capi.MPM_UGIMP
        cfg.sort_interval = int(sort_interval or 0)
        cfg.np = self.np
        cfg.nmat = int(attr.nmat)
        cfg.nnx, cfg.nny = int(grid.nnx), int(grid.nny)
        cfg.nnz = int(grid.nnz) if grid.dim == 3 else 1
        cfg.x1, cfg.y1 = float(grid.x1), float(grid.y1)
        cfg.dx, cfg.dy = float(grid.dx), float(grid.dy)
        cfg.lpx, cfg.lpy = float(mp.dx), float(mp.dy)
        if grid.dim == 3:
            cfg.z1, cfg.dz, cfg.lpz = float(grid.z1), float(grid.dz), float(mp.dz)
        else:
            cfg.z1, cfg.dz, cfg.lpz = 0.0, 1.0, 1.0
        self.cfg = cfg
        ctx = capi.C.c_void_p()
        rc = lib.mpm_create(capi.C.byref(cfg), capi.C.byref(ctx))
        if rc != capi.MPM_OK:
            raise capi.MPMError(rc, lib.mpm_last_error(None).decode())
        self.ctx = ctx
        try:
            self._upload(grid, mp, attr, bc)
        except Exception:
            self.close()
            raise

    def _check(self, rc):
        return capi.check(self.lib, self.ctx, rc)

    def _upload_particles(self, arrs):
        if arrs["nid"].size != self.np:
            raise ValueError("attr.nid must have one entry per particle")
        nv = 4 if self.dim == 2 else 6
        cols = dict(xi=self.dim, vs=self.dim, sigma=nv, eps_s=nv, F=self.dim * self.dim)
        for k in ("xi", "vs", "sigma", "eps_s", "F", "omega", "omega0", "rho", "rho0", "eps_q",
                  "eps_k"):
            if arrs[k].shape[0] != self.np:
                raise ValueError(f"{k} must have np = {self.np} rows")
            want = cols.get(k, 1)
            got = arrs[k].shape[1] if arrs[k].ndim == 2 else 1
            if got != want or arrs[k].size != self.np * want:
                raise ValueError(f"{k} must be np x {want} (got shape {arrs[k].shape})")
        d = capi.dptr
        self._check(self.lib.mpm_upload_particles(
            self.ctx, d(arrs["xi"]), d(arrs["vs"]), d(arrs["sigma"]), d(arrs["eps_s"]),
            d(arrs["F"]), d(arrs["omega"]), d(arrs["omega0"]), d(arrs["rho"]), d(arrs["rho0"]),
            d(arrs["eps_q"]), d(arrs["eps_k"]), capi.iptr(arrs["nid"])))

    @staticmethod
    def snapshot_state(group, Ω0, nid):
        """Particle arrays for mpm_upload_particles from one snapshot group (the
        datasets of src/solver.jl:60-83: coords, velocity_s, stress, strain_s,
        eqstrain, ekstrain, volume, mass_s, time).  The reference saves neither
        F nor ρs: det F is re-derived as volume / Ω0 (the step only reads det F)
        and the densities as mass_s / volume and mass_s / Ω0.  Returns
        (arrays, time)."""
        g = group if isinstance(group, dict) else dict(np.load(group))
        f64 = lambda a: np.asfortranarray(np.asarray(a, dtype=np.float64))
        x = f64(g["coords"])
        n, dim = x.shape
        vol = np.asarray(g["volume"], dtype=np.float64).reshape(n)
        vol0 = np.asarray(Ω0, dtype=np.float64).reshape(n)
        ms = np.asarray(g["mass_s"], dtype=np.float64).reshape(n)
        F = np.zeros((n, dim * dim))
        F[:, 0] = vol / vol0                 # det F = J, row-major diag(J, 1[, 1])
        F[:, dim + 1] = 1.0
        if dim == 3:
            F[:, 8] = 1.0
        arrs = dict(xi=x, vs=f64(g["velocity_s"]), sigma=f64(g["stress"]),
                    eps_s=f64(g["strain_s"]), F=f64(F), omega=f64(vol), omega0=f64(vol0),
                    rho=f64(ms / vol), rho0=f64(ms / vol0), eps_q=f64(g["eqstrain"]),
                    eps_k=f64(g["ekstrain"]),
                    nid=np.ascontiguousarray(np.asarray(nid, dtype=np.int64)))
        return arrs, float(np.asarray(g["time"]))

    def upload_snapshot(self, group, Ω0, nid=None):
        """Resume from a snapshot group written by SnapshotWriter (the reference's
        HDF5 group layout): re-uploads the particle state and returns the
        group's time, the Ti to continue from.  Ω0 is the particles' initial
        volume (mp.Ω0 of the host struct; the reference's snapshot does not hold
        it), nid defaults to the material ids uploaded at creation.  A USF run's
        pending displacement is not part of a snapshot (nor of the
        reference's): the next USF step starts from a zero increment."""
        arrs, Ti = self.snapshot_state(group, Ω0, self._nid if nid is None else nid)
        self._upload_particles(arrs)
        return Ti

    def _upload(self, grid, mp, attr, bc):
        f64 = lambda a: np.asfortranarray(np.asarray(a, dtype=np.float64))
        i64 = lambda a: np.ascontiguousarray(np.asarray(a, dtype=np.int64))
        arrs = dict(
            xi=f64(mp.ξ), vs=f64(mp.vs), sigma=f64(mp.σij), eps_s=f64(mp.ϵijs), F=f64(mp.F),
            omega=f64(mp.Ω), omega0=f64(mp.Ω0), rho=f64(mp.ρs), rho0=f64(mp.ρs0),
            eps_q=f64(mp.ϵq), eps_k=f64(mp.ϵk), nid=i64(attr.nid))
        self._upload_particles(arrs)
        self._nid = arrs["nid"]
        m = [f64(a) for a in (attr.Ks, attr.Gs, attr.ϕ, attr.ψ, attr.c, attr.σt, attr.Hp, attr.cr,
                              attr.ϕr)]
        if any(a.size != int(attr.nmat) for a in m):
            raise ValueError(f"material arrays must have attr.nmat = {int(attr.nmat)} entries")
        self._check(self.lib.mpm_upload_materials(self.ctx, *[capi.dptr(a) for a in m]))
        ni = int(grid.ni)
        bx = [i64(bc.vx_s_idx), f64(bc.vx_s_val), i64(bc.vy_s_idx), f64(bc.vy_s_val)]
        if self.dim == 3:
            bx += [i64(bc.vz_s_idx), f64(bc.vz_s_val)]
        for a in bx:
            if a.size != ni:
                raise ValueError(f"boundary arrays must have grid.ni = {ni} entries")
        if self.dim == 2:
            bx += [None, None]
        self._check(self.lib.mpm_upload_bc(
            self.ctx, capi.iptr(bx[0]), capi.dptr(bx[1]), capi.iptr(bx[2]), capi.dptr(bx[3]),
            capi.iptr(bx[4]), capi.dptr(bx[5])))

    @staticmethod
    def step_args(args):
        a = capi.MPMArgs()
        if args.coupling != "OS" or args.scheme not in capi.SCHEME:
            raise NotImplementedError("the HIP path implements coupling=:OS with the "
                                      "MUSL, USL and USF schemes")
        if args.constitutive not in capi.CONSTITUTIVE:
            raise NotImplementedError(f"constitutive {args.constitutive} has no HIP kernels")
        a.scheme = capi.SCHEME[args.scheme]
        a.va = capi.MPM_VA_A if args.va == "a" else capi.MPM_VA_V
        a.constitutive = capi.CONSTITUTIVE[args.constitutive]
        a.mvl = 1 if args.MVL else 0
        a.flip, a.pic = float(args.FLIP), float(args.PIC)
        a.zeta_s, a.gravity, a.Te = float(args.ζs), float(args.gravity), float(args.Te)
        # body force on the other axes: an extension carried in args.ext (the
        # reference's user-extension slot); absent = the reference's gravity
        ext = args.ext if isinstance(args.ext, dict) else {}
        a.gravity_x = float(ext.get("gravity_x", 0.0))
        a.gravity_y = float(ext.get("gravity_y", 0.0))
        return a

    def step(self, sargs, dt, Ti):
        self._check(self.lib.mpm_step(self.ctx, capi.C.byref(sargs), float(dt), float(Ti)))

    def run(self, sargs, dt, Ti, nsteps, graph=False):
        """nsteps steps from Ti with the submit_work! clock; returns the final Ti.
        graph=True replays captured HIP graphs of the steps between sorts."""
        out = capi.C.c_double(0.0)
        fn = self.lib.mpm_run_graph if graph else self.lib.mpm_run
        self._check(fn(self.ctx, capi.C.byref(sargs), float(dt), float(Ti), int(nsteps),
                       capi.C.byref(out)))
        return out.value

    def sync(self):
        self._check(self.lib.mpm_sync(self.ctx))

    def download(self):
        """The device2host! field set in the caller's particle order (FP64), keyed
        xi, vs, sigma, sigma_m, omega, eps_q, eps_k, eps_v, eps_s, ms."""
        n, D = self.np, self.dim
        ns = 4 if D == 2 else 6
        out = dict(xi=np.zeros((n, D), order="F"), vs=np.zeros((n, D), order="F"),
                   sigma=np.zeros((n, ns), order="F"), sigma_m=np.zeros(n), omega=np.zeros(n),
                   eps_q=np.zeros(n), eps_k=np.zeros(n), eps_v=np.zeros(n),
                   eps_s=np.zeros((n, ns), order="F"), ms=np.zeros(n))
        p = capi.dptr
        self._check(self.lib.mpm_download_particles(
            self.ctx, p(out["xi"]), p(out["vs"]), p(out["sigma"]), p(out["sigma_m"]),
            p(out["omega"]), p(out["eps_q"]), p(out["eps_k"]), p(out["eps_v"]),
            p(out["eps_s"]), p(out["ms"])))
        return out

    def count_active_nodes(self, noalloc=False):
        """Grid nodes with a non-zero weight from some particle.  noalloc=True
        uses mpm_count_active_nodes_noalloc (no device allocation; grids of at
        most 2**25 nodes)."""
        out = np.zeros(1, dtype=np.int64)
        fn = self.lib.mpm_count_active_nodes_noalloc if noalloc else self.lib.mpm_count_active_nodes
        self._check(fn(self.ctx, capi.iptr(out)))
        return int(out[0])

    def set_option(self, option, value):
        """Runtime option: "sort_interval" (>= 1), "tiles" (bool), "scatter"
        (P2G #1's tile scatter form: "rounds", "rounds_layered", "owned",
        "owned_layers" or "pairs"; None = automatic: "pairs" in FP32 3D uGIMP,
        else "rounds"), "remap_scatter" (the re-map momentum scatter's form,
        None = automatic: "owned" in FP32 3D uGIMP, else as "scatter"),
        "fused_remap" (bool: the one-pass MUSL re-map; off by default),
        "tile_partials" (bool: tile windows stored and summed by the grid solves
        instead of flushed with atomics) or
        "profile_kernels" (kernel names that profile_enable times; empty or
        None = every kernel)."""
        if option == "profile_kernels":
            names = [self.lib.mpm_profile_kernel_name(k).decode()
                     for k in range(self.lib.mpm_profile_kernel_count())]
            mask = 0
            for name in value or ():
                if name not in names:
                    raise ValueError(f"unknown kernel {name!r}; kernels: {names}")
                mask |= 1 << names.index(name)
            self._check(self.lib.mpm_set_option(self.ctx, capi.MPM_OPT_PROFILE_MASK, mask))
            return
        if option in ("scatter", "remap_scatter"):
            if value is None:
                code = -1  # automatic
            elif value in capi.SCATTER:
                code = capi.SCATTER[value]
            else:
                raise ValueError(f"{option} form {value!r}; forms: {sorted(capi.SCATTER)}")
            opt = capi.MPM_OPT_SCATTER if option == "scatter" else capi.MPM_OPT_REMAP_SCATTER
            self._check(self.lib.mpm_set_option(self.ctx, opt, code))
            return
        code = {"sort_interval": capi.MPM_OPT_SORT_INTERVAL, "tiles": capi.MPM_OPT_TILES,
                "fused_remap": capi.MPM_OPT_FUSED_REMAP,
                "tile_partials": capi.MPM_OPT_TILE_PARTIALS}[option]
        self._check(self.lib.mpm_set_option(self.ctx, code, int(value)))

    # -- learned 3D Mohr-Coulomb return ----------------------------------------
    def return_log(self, capacity, stride=1):
        """Start logging exact 3D Mohr-Coulomb return-mapping samples."""
        self._check(self.lib.mpm_return_log(self.ctx, int(capacity), int(stride)))

    def read_return_log(self, max_rows=None):
        """Logged samples as (X (n, 16), Y (n, 7)) float32 arrays."""
        cap = int(max_rows) if max_rows is not None else 1 << 26
        rows = np.zeros(1, dtype=np.int64)
        self._check(self.lib.mpm_return_log_read(self.ctx, None, None, cap, capi.iptr(rows)))
        n = int(rows[0])
        X = np.zeros((n, 16), dtype=np.float32)
        Y = np.zeros((n, 7), dtype=np.float32)
        if n:
            self._check(self.lib.mpm_return_log_read(self.ctx, capi.fptr(X), capi.fptr(Y), n,
                                                     capi.iptr(rows)))
        return X, Y

    def run_logged(self, sargs, dt, Ti, nsteps, chunk=200, capacity=2_000_000, stride=1,
                   graph=False):
        """Run nsteps with the exact return, logging return-mapping samples in
        chunks of `chunk` steps: each chunk logs up to `capacity` samples (every
        stride-th yielding particle in the current particle order, until the
        chunk's log is full), so the samples cover the whole run rather than its
        first steps.  Returns (Ti, X, Y) with X (n, 16) and Y (n, 7) float32."""
        xs, ys = [], []
        done = 0
        while done < nsteps:
            k = min(chunk, nsteps - done)
            self.return_log(capacity, stride)
            Ti = self.run(sargs, dt, Ti, k, graph=graph)
            X, Y = self.read_return_log()
            xs.append(X)
            ys.append(Y)
            done += k
        self.return_log(0, 1)
        X = np.concatenate(xs) if xs else np.zeros((0, 16), np.float32)
        Y = np.concatenate(ys) if ys else np.zeros((0, 7), np.float32)
        return Ti, X, Y

    def set_surrogate(self, surrogate, tol=1e-3):
        """Use a trained Surrogate as the 3D Mohr-Coulomb return (None detaches)."""
        h = surrogate.h if surrogate is not None else None
        self._check(self.lib.mpm_set_surrogate(self.ctx, h, float(tol)))
        self._surrogate = surrogate  # keep the network alive while attached

    def surrogate_stats(self):
        """(yielding particle-steps, fallbacks to the exact return) since attaching."""
        out = np.zeros(2, dtype=np.int64)
        self._check(self.lib.mpm_surrogate_stats(self.ctx, capi.iptr(out[:1]), capi.iptr(out[1:])))
        return int(out[0]), int(out[1])

    def profile_enable(self, on=True):
        self._check(self.lib.mpm_profile_enable(self.ctx, 1 if on else 0))

    def profile_read(self):
        k = self.lib.mpm_profile_kernel_count()
        ms = np.zeros(k)
        cnt = np.zeros(k, dtype=np.int64)
        self._check(self.lib.mpm_profile_read(self.ctx, capi.dptr(ms), capi.iptr(cnt), k))
        names = [self.lib.mpm_profile_kernel_name(i).decode() for i in range(k)]
        return {names[i]: (float(ms[i]), int(cnt[i])) for i in range(k)}

    def close(self):
        if getattr(self, "ctx", None) is not None:
            self.lib.mpm_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _DeviceHandle:
    """dev_grid / dev_mp / dev_attr / dev_bc returned by host2device: views of one model."""

    def __init__(self, kind, model, host):
        self.kind, self.model, self.host = kind, model, host

    def __getattr__(self, name):
        return getattr(self.host, name)


def _require_hip(device):
    if device != "HIP":
        raise NotImplementedError(
            f"device :{device} is not provided by this package; it implements the :HIP device "
            "(libmpm_hip.so on MI355X).  There is no CPU fallback.")


def host2device(grid, mp, attr, bc, device, sort_interval=None):
    """host2device(grid, mp, attr, bc, Val(:HIP)) -- ext/AMDExt/devicehelpfunc_amd.jl:17-37."""
    _require_hip(device)
    if sort_interval is None:
        sort_interval = 10 if mp.np >= 65536 else 0
    model = HIPModel(grid, mp, attr, bc, sort_interval=sort_interval)
    return (_DeviceHandle("grid", model, grid), _DeviceHandle("mp", model, mp),
            _DeviceHandle("attr", model, attr), _DeviceHandle("bc", model, bc))


def device2host(args, mp, dev_mp, device, verbose=False):
    """device2host!(args, mp, dev_mp, Val(:HIP)) -- devicehelpfunc_amd.jl:39-69: copies
    σm, vs, ξ, Ω, σij, ϵk, ϵq, ϵv, ϵijs, ms into the host struct in place."""
    _require_hip(device)
    out = dev_mp.model.download()
    pairs = ((mp.σm, "sigma_m"), (mp.vs, "vs"), (mp.ξ, "xi"), (mp.Ω, "omega"),
             (mp.σij, "sigma"), (mp.ϵk, "eps_k"), (mp.ϵq, "eps_q"), (mp.ϵv, "eps_v"),
             (mp.ϵijs, "eps_s"), (mp.ms, "ms"))
    for dst, key in pairs:
        dst[...] = out[key].astype(dst.dtype)
    if verbose:
        print("[▼ I/O: downloading from :HIP → host]")


def clean_device(dev_grid, dev_mp, dev_attr, dev_bc, device):
    """clean_device!(..., Val(:HIP)) -- devicehelpfunc_amd.jl:71-102."""
    _require_hip(device)
    dev_mp.model.close()


def getBackend(device):
    """getBackend(Val(:HIP)): the synchronisation handle of the device."""
    _require_hip(device)
    return "HIP"


def KAsync(dev_handle):
    """KAsync(getBackend(Val(:HIP))) -- wait for the device, report kernel flags."""
    dev_handle.model.sync()


def procedure(args, grid, mp, attr, bc, ΔT, Ti, coupling, scheme, va):
    """procedure!(args, grid, mp, attr, bc, ΔT, Ti, Val(:OS), Val(scheme), Val(va))
    -- src/solvers/OS_MUSL.jl:12-102, OS_USL.jl:102-144,325-366 and
    OS_USF.jl:11-97, as one device step."""
    if coupling != "OS" or scheme not in capi.SCHEME:
        raise NotImplementedError("the HIP path implements Val(:OS) with Val(:MUSL), "
                                  "Val(:USL) and Val(:USF)")
    if not isinstance(mp, _DeviceHandle):
        raise TypeError("procedure needs the device structs returned by host2device")
    sargs = HIPModel.step_args(args)
    sargs.scheme = capi.SCHEME[scheme]
    sargs.va = 0 if va == "a" else 1
    mp.model.step(sargs, ΔT, Ti)


class SnapshotWriter:
    """The reference's HDF5 snapshot layout (src/solver.jl:53-115) in NumPy
    containers (h5py is not available here): one ``group<k>.npz`` per snapshot
    with the datasets stress, strain_s, eqstrain, ekstrain, eqrate, coords,
    velocity_s, volume, mass_s, time, and ``meta.npz`` with FILE_NUM,
    grid_coords, mp_coords0, nid and the vbc_* arrays, under
    ``<project_path>/<project_name>/<project_name>_snapshots/``.  Arrays keep
    the reference's (np, k) layout and dataset names."""

    def __init__(self, args):
        self.dir = os.path.join(args.project_path, args.project_name,
                                f"{args.project_name}_snapshots")
        if os.path.isdir(self.dir):
            for f in os.listdir(self.dir):
                if f.endswith(".npz"):
                    os.remove(os.path.join(self.dir, f))
        os.makedirs(self.dir, exist_ok=True)

    def group(self, k, mp, Ti):
        np.savez(os.path.join(self.dir, f"group{k}.npz"), stress=mp.σij, strain_s=mp.ϵijs,
                 eqstrain=mp.ϵq, ekstrain=mp.ϵk, eqrate=mp.ϵv, coords=mp.ξ, velocity_s=mp.vs,
                 volume=mp.Ω, mass_s=mp.ms, time=np.asarray(Ti))

    def close(self, file_num, grid, mp, attr, bc):
        extra = {}
        if grid.dim == 3:
            extra = dict(vbc_zs_idx=bc.vz_s_idx, vbc_zs_val=bc.vz_s_val)
        np.savez(os.path.join(self.dir, "meta.npz"), FILE_NUM=np.asarray(file_num),
                 grid_coords=grid.ξ, mp_coords0=mp.ξ0, nid=attr.nid, vbc_xs_idx=bc.vx_s_idx,
                 vbc_xs_val=bc.vx_s_val, vbc_ys_idx=bc.vy_s_idx, vbc_ys_val=bc.vy_s_val, **extra)


def submit_work(args, grid, mp, attr, bc, workflow):
    """submit_work! (src/solver.jl:37-131) for the HIP device: upload, the
    ``while Ti < Ttol`` loop with ``Ti += ΔT`` in T2, snapshots every
    ``hdf5_step`` steps when ``args.hdf5`` (SnapshotWriter), final sync,
    download, free.  (The reference's host-side initmpstatus! only fills the
    per-slot basis arrays, which the device recomputes every step.)"""
    T2 = args.T2
    Ti = T2(0.0)
    ΔT = T2(args.ΔT)
    dev_grid, dev_mp, dev_attr, dev_bc = host2device(grid, mp, attr, bc, args.device,
                                                     sort_interval=args.sort_interval)
    try:
        writer = SnapshotWriter(args) if args.hdf5 else None
        group_id, switch = 1, 0
        args.start_time = time.time()
        while Ti < args.Ttol:
            if writer is not None and (switch == args.hdf5_step or switch == 0):
                device2host(args, mp, dev_mp, args.device)
                writer.group(group_id, mp, Ti)
                switch, group_id = 0, group_id + 1
            workflow(args, dev_grid, dev_mp, dev_attr, dev_bc, ΔT, Ti, args.coupling,
                     args.scheme, args.va)
            Ti = T2(Ti + ΔT)
            switch += 1
            args.iter_num += 1
        args.end_time = time.time()
        if writer is not None:
            writer.close(group_id, grid, mp, attr, bc)
        KAsync(dev_mp)
        device2host(args, mp, dev_mp, args.device, verbose=args.progressbar)
    finally:
        clean_device(dev_grid, dev_mp, dev_attr, dev_bc, args.device)


def info_print(args, grid, mp):
    if args.progressbar:
        print(f"[ MaterialPointSolver (HIP) ] {args.project_name}: {args.constitutive}, "
              f"np={mp.np}, ni={grid.ni}, ϵ={args.ϵ}")


def perf(args):
    """perf (src/toolkits/terminaltxt.jl:67-85): iterations per second of the run."""
    wall = max(args.end_time - args.start_time, 1e-12)
    it_s = args.iter_num / wall
    if args.progressbar:
        print(f"[ Performance ] iterations: {args.iter_num}, wall time: {wall:.3f} s, "
              f"{it_s:.2f} it/s")
    return it_s


def materialpointsolver(args, grid, mp, attr, bc, workflow=procedure):
    """materialpointsolver! (src/MaterialPointSolver.jl:61-75)."""
    if args.animation:
        raise NotImplementedError("animation output is not implemented in this package yet")
    info_print(args, grid, mp)
    submit_work(args, grid, mp, attr, bc, workflow)
    perf(args)


def Tpeak(device="HIP", datatype="FP32", ID=0, nbytes=1 << 30, reps=10):
    """Tpeak(Val(:HIP)) -- triad A = B + C bandwidth (devicehelpfunc_amd.jl:104-142).
    Returns GB/s (1e9 B/s); the reference prints GiB/s."""
    _require_hip(device)
    lib = capi.load()
    out = capi.C.c_double(0.0)
    capi.check(lib, None, lib.mpm_tpeak(int(nbytes), int(reps), capi.C.byref(out)))
    return out.value
