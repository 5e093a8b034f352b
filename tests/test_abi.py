"""CPU checks of the drop-in boundary: the C-ABI library loads and exports
every entry point that include/veritas_amd.h declares; the host library loads
and exports its C entry points.  No compute call is made (no GPU here)."""
import ctypes
import os
import re
import subprocess

import pytest

import veritas_amd as va

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(va.HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(vx_[a-z0-9_]+)\s*\(", src)))


@pytest.fixture(scope="module")
def built():
    if not (os.path.exists(va.KERNEL_LIB) and os.path.exists(va.HOST_LIB)):
        va.build()
    return True


def test_header_declares_the_stage_entry_points():
    names = declared_functions()
    for n in ("vx_ctx_create", "vx_fct_sub0", "vx_fct_sub1", "vx_fct_sub2", "vx_fct_sub3",
              "vx_fct_stage_walls", "vx_ghost_zero", "vx_moments_patch", "vx_moments_assemble",
              "vx_poisson", "vx_maxwell_stage", "vx_cfl_bound", "vx_neutralize"):
        assert n in names


def test_kernel_library_exports_every_declared_symbol(built):
    lib = ctypes.CDLL(va.KERNEL_LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_kernel_library_is_gfx950_code_object(built):
    out = subprocess.run(["/opt/rocm/llvm/bin/llvm-readelf", "-n", va.KERNEL_LIB],
                         stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True).stdout
    # the offload bundle is embedded; check the bundle id string instead
    data = open(va.KERNEL_LIB, "rb").read()
    assert b"gfx950" in data


def test_host_library_exports_entry_points(built):
    L = va.host_lib()
    for n in ("vxh_create", "vxh_destroy", "vxh_advance", "vxh_get", "vxh_time_advance",
              "vxh_stage_moments", "vxh_stage_poisson", "vxh_stage_vlasov_sub"):
        assert hasattr(L, n)


def test_no_cpu_fallback_without_gpu(built):
    """Creating a run on a machine without a GPU fails loudly."""
    import torch  # noqa: F401  (only to know whether a GPU is present)
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    with pytest.raises(va.VeritasError):
        va.Run(va.landau())
