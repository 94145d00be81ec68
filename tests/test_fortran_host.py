"""The Fortran side of the boundary: the ISO_C_BINDING module binds exactly
the C-ABI's entry points, and the Fortran host driver runs the hot path on
the GPU in update_ocean_model order and conserves tracer content."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MOD = os.path.join(ROOT, "mom5_amd", "fortran", "mom_gpu_mod.F90")
DRIVER = os.path.join(ROOT, "mom5_amd", "ocean_gpu_driver")


def test_fortran_bindings_name_declared_c_functions():
    from test_abi import declared_symbols
    txt = open(MOD).read()
    bound = set(re.findall(r'bind\(C,\s*name="(mom_gpu_\w+)"\)', txt))
    assert bound, "no bindings found"
    assert bound <= declared_symbols(), bound - declared_symbols()
    for must in ("mom_gpu_horz_advect_tracer", "mom_gpu_invtri", "mom_gpu_update_ocean_barotropic",
                 "mom_gpu_vert_diffuse_implicit", "mom_gpu_create"):
        assert must in bound


@pytest.mark.gpu
def test_fortran_host_driver_runs_and_conserves():
    assert os.path.exists(DRIVER), "build the Fortran driver first (mom5_amd.build.build_fortran)"
    p = subprocess.run([DRIVER], capture_output=True, text=True, timeout=120)
    print(p.stdout)
    print(p.stderr)
    assert p.returncode == 0
    assert "content conserved" in p.stdout
