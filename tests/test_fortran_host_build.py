"""The Fortran host program is built against libndsp_hip.so and resolves it
through its rpath (no GPU needed)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "ndspmhd_amd", "fortran", "ndsp_run")


def test_host_program_links_library():
    if not os.path.exists(EXE):
        from ndspmhd_amd.build import build_fortran
        build_fortran()
    out = subprocess.run(["ldd", EXE], capture_output=True, text=True).stdout
    line = [l for l in out.splitlines() if "libndsp_hip.so" in l]
    assert line and "not found" not in line[0], out
