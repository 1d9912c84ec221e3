"""The OpenSPH-side adapter (opensph_amd/adapter: GpuSymmetricSolver for ISolver, GpuPredictorCorrector for
ITimeStepping, GpuGravity for IGravity) built against the reference's own headers and linked with the
reference core and the HIP library (oracle/Makefile target `adapter`): the library loads and exports its
driver.  Builds where the reference sources are present (not on the GPU box, which runs the built
library in tests/test_gpu_adapter.py)."""
import ctypes as C
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_CORE = "/root/reference/core"
ADAPTER_SO = os.path.join(ROOT, "oracle", "_ref", "libosph_adapter.so")


@pytest.mark.skipif(not os.path.isdir(REF_CORE), reason="reference sources not present")
def test_adapter_builds_links_and_loads():
    r = subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "adapter", "REF=/root/reference"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    lib = C.CDLL(ADAPTER_SO)
    for name in ("adapter_create", "adapter_destroy", "adapter_step", "adapter_integrate", "adapter_read",
                 "adapter_error", "adapter_save", "adjoint_run"):
        assert hasattr(lib, name), name
    # the adapter's classes are compiled in, and it resolves the C ABI from libosph_gpu.so
    syms = subprocess.run(["nm", "-D", "-C", ADAPTER_SO], capture_output=True, text=True).stdout
    for cls in ("Sph::GpuSymmetricSolver::integrate", "Sph::GpuPredictorCorrector::stepParticles",
                "Sph::GpuGravity::evalSelfGravity", "Sph::uploadStorage", "Sph::downloadStorage",
                "Sph::GpuAdjointRun::forward", "Sph::GpuAdjointRun::backward"):
        assert cls in syms, cls
    for abi in ("osph_step_pc", "osph_integrate", "osph_set_quantity", "osph_get_quantity", "osph_gravity_eval_self",
                "osph_adjoint_forward", "osph_adjoint_backward"):
        assert f"U {abi}" in syms, abi
    deps = subprocess.run(["ldd", ADAPTER_SO], capture_output=True, text=True).stdout
    assert "libosph_gpu.so" in deps and "libosph_ref.so" in deps and "not found" not in deps, deps
