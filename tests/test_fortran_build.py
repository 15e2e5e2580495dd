"""The Fortran host builds (CPU): the modules the reference's main program
uses exist with the reference's names, and where the reference sources are
mounted its unmodified main.f90 has been compiled and linked against them
and the HIP library (lib/channel_fd_ref, made by build())."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "channel-flow_amd", "lib")
MODS = os.path.join(ROOT, "channel-flow_amd", "_build", "fortran")
REF_MAIN = "/root/reference/main.f90"


def test_reference_module_names_exist():
    if not os.path.isdir(MODS):
        pytest.skip("Fortran host not built here (channel-flow_amd/fortran/build.sh)")
    have = {f[:-4] for f in os.listdir(MODS) if f.endswith(".mod")}
    need = {"global", "input_output", "initialization", "time_integration", "monitor", "statistics",
            "finalization", "equations", "boundary_conditions", "projection", "pressure", "mass_flow",
            "subgrid", "eqwm", "wallmodel"}
    assert need <= have, need - have


@pytest.mark.skipif(not os.path.exists(REF_MAIN), reason="reference sources not mounted")
def test_reference_main_links_against_the_modules():
    exe = os.path.join(LIB, "channel_fd_ref")
    assert os.path.exists(exe), "build() did not produce lib/channel_fd_ref"
    out = subprocess.run(["nm", "-D", "--undefined-only", exe], capture_output=True, text=True).stdout
    # the reference driver reaches the library only through the modules' C-ABI calls
    assert "cf_create" in out and "cf_rk3_step" in out and "cf_statistics" in out
    assert os.path.getmtime(exe) >= os.path.getmtime(REF_MAIN)
