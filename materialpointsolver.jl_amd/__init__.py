"""MI355X-native explicit MPM timestep behind MaterialPointSolver.jl's API.

The package mirrors the reference's user-facing API for the single-phase MUSL
path (src/MaterialPointSolver.jl, src/solver.jl, src/types/*.jl) on top of
libmpm_hip.so, a gfx950 HIP library with a C ABI (include/mpm_hip.h).  The
device is ``"HIP"``; no other device and no CPU fallback is provided.

The folder name carries a dot, so load it with importlib, e.g.::

    spec = importlib.util.spec_from_file_location(
        "materialpointsolver_amd", ".../materialpointsolver.jl_amd/__init__.py",
        submodule_search_locations=[".../materialpointsolver.jl_amd"])
"""
from .types import (ArgumentError, UserArgs2D, UserArgs3D, UserGrid2D, UserGrid3D,
                    UserParticle2D, UserParticle3D, UserProperty, UserVBoundary2D,
                    UserVBoundary3D, julia_range, meshbuilder)
from .solver import (HIPModel, KAsync, Tpeak, clean_device, device2host, getBackend,
                     host2device, materialpointsolver, perf, procedure, submit_work)
from . import scenes
from .surrogate import Surrogate
from .autotune import SCATTER_FORMS, autotune

__all__ = ["ArgumentError", "UserArgs2D", "UserArgs3D", "UserGrid2D", "UserGrid3D",
           "UserParticle2D", "UserParticle3D", "UserProperty", "UserVBoundary2D",
           "UserVBoundary3D", "julia_range", "meshbuilder", "HIPModel", "KAsync", "Tpeak",
           "clean_device", "device2host", "getBackend", "host2device", "materialpointsolver",
           "perf", "procedure", "submit_work", "scenes", "Surrogate", "autotune", "SCATTER_FORMS"]
