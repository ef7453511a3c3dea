"""C ABI checks that need no GPU: the library loads, exports every entry point
include/mpm_hip.h declares, and rejects bad configurations before touching HIP."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "mpm_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mpm_[a-z_0-9]+)\s*\(", src)))


def test_every_declared_symbol_is_exported(mpm):
    lib = mpm._capi.load()
    names = _declared()
    assert len(names) >= 16
    for n in names:
        assert hasattr(lib, n), n
    # and the ctypes prototypes cover exactly the declared set
    assert sorted(mpm._capi.exported_symbols()) == names


def test_abi_version_and_kernel_names(mpm):
    lib = mpm._capi.load()
    assert lib.mpm_abi_version() == 2
    k = lib.mpm_profile_kernel_count()
    names = [lib.mpm_profile_kernel_name(i).decode() for i in range(k)]
    assert names[:5] == ["k_p2g", "k_grid1", "k_g2p1_p2g2", "k_grid2", "k_g2p2"]
    assert lib.mpm_profile_kernel_name(k) == b""


def _cfg(mpm, **kw):
    c = mpm._capi.MPMConfig()
    c.dim, c.precision, c.basis, c.sort_interval = 3, 1, 0, 10
    c.np, c.nmat = 100, 1
    c.nnx = c.nny = c.nnz = 10
    c.dx = c.dy = c.dz = 0.1
    c.lpx = c.lpy = c.lpz = 0.05
    for k, v in kw.items():
        setattr(c, k, v)
    return c


@pytest.mark.parametrize("bad,code,msg", [
    (dict(dim=4), -1, "dim"),
    (dict(precision=7), -1, "precision"),
    (dict(basis=2), -2, "gslinear"),
    (dict(np=0), -1, "np"),
    (dict(nmat=0), -1, "nmat"),
    (dict(nnx=2), -1, "3 nodes"),
    (dict(dx=0.0), -1, "spacing"),
    (dict(lpx=0.2), -1, "uGIMP"),
    (dict(sort_interval=-1), -1, "sort_interval"),
])
def test_create_rejects_bad_config(mpm, bad, code, msg):
    lib = mpm._capi.load()
    ctx = C.c_void_p()
    rc = lib.mpm_create(C.byref(_cfg(mpm, **bad)), C.byref(ctx))
    assert rc == code
    assert not ctx.value
    assert msg in lib.mpm_last_error(None).decode()


def test_null_arguments(mpm):
    lib = mpm._capi.load()
    assert lib.mpm_create(None, None) == -1
    assert lib.mpm_step(None, None, 1e-3, 0.0) == -1
    assert lib.mpm_sync(None) == -1
    lib.mpm_destroy(None)  # no-op
    assert lib.mpm_tpeak(10, 1, None) == -1


def test_python_constants_match_the_header(mpm):
    # every MPM_* constant the ctypes layer defines carries the header's value
    src = open(os.path.join(ROOT, "include", "mpm_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    vals = {n: int(v) for n, v in re.findall(r"\b(MPM_[A-Z0-9_]+)\s*=\s*(-?\d+)", src)}
    vals.update({n: int(v) for n, v in re.findall(r"#define\s+(MPM_[A-Z0-9_]+)\s+(-?\d+)", src)})
    capi = mpm._capi
    shared = [n for n in vals if hasattr(capi, n)]
    assert {"MPM_OPT_PROFILE_MASK", "MPM_USF", "MPM_OK"} <= set(shared)
    for n in shared:
        assert getattr(capi, n) == vals[n], n
    # the error-code names and the constitutive ids the host layer maps
    assert {vals[n]: n for n in capi.ERRORS.values()} == capi.ERRORS
    for name, code in capi.CONSTITUTIVE.items():
        assert vals["MPM_" + name.upper()] == code, name
