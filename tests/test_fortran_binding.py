"""The Fortran side of the boundary: projectracer_amd/fortran/projectracer_gpu.f90
(the ISO_C_BINDING interface block a reference maintainer adds) compiles,
links against libprojectracer_gpu.so, and sees the same struct layout and
defaults as the C header.  CPU only: no call here reaches the GPU."""
import ctypes as C
import shutil
import subprocess

import pytest

from projectracer_amd import _abi

FLANG = shutil.which("flang") or shutil.which("gfortran")
pytestmark = pytest.mark.skipif(FLANG is None, reason="no Fortran compiler in this image")

PROG = """
program binding_probe
 use projectracer_gpu
 implicit none
 type(pt_params) :: p
 type(pt_solve_info) :: info
 type(pt_stats_row) :: row
 type(c_ptr) :: ctx
 integer(c_int) :: ierr
 call pt_default_params(p)
 write(*,'(a,4i8)') "ints ", p%nx, p%nz, p%grids, p%nstageS
 write(*,'(a,4es25.17)') "reals ", p%tolerance, p%dampingS, p%c_2, p%courant_stream
 write(*,'(a,3i6)') "sizes ", int(c_sizeof(p)), int(c_sizeof(info)), int(c_sizeof(row))
 write(*,'(a,3i8,es25.17)') "tail ", p%buffer, p%equalize, p%RKC1, p%dist_factor
 p%nx = 2
 ierr = pt_create(p, ctx)
 write(*,'(a,i4)') "invalid ", ierr
 write(*,'(a,a)') "message ", pt_error_message()
end program
"""


def test_fortran_interface_compiles_links_and_matches_header(tmp_path):
    lib = _abi.LIB_PATH
    if not lib.exists():
        pytest.skip("libprojectracer_gpu.so not built")
    mod_src = _abi.PKG_DIR / "fortran" / "projectracer_gpu.f90"
    (tmp_path / "probe.f90").write_text(PROG)
    subprocess.run([FLANG, "-c", str(mod_src), "-o", "mod.o"], cwd=tmp_path, check=True)
    subprocess.run([FLANG, "probe.f90", "mod.o", f"-L{lib.parent}", "-lprojectracer_gpu",
                    f"-Wl,-rpath,{lib.parent}", "-o", "probe"], cwd=tmp_path, check=True)
    out = subprocess.run([str(tmp_path / "probe")], capture_output=True, text=True, timeout=120).stdout
    rows = {line.split()[0]: line.split(None, 1)[1] for line in out.splitlines() if line.strip()}
    p = _abi.default_params()
    assert [int(v) for v in rows["ints"].split()] == [p.nx, p.nz, p.grids, p.nstageS]
    reals = [float(v) for v in rows["reals"].split()]
    assert reals == [p.tolerance, p.dampingS, p.c_2, p.courant_stream]
    assert [int(v) for v in rows["sizes"].split()] == [C.sizeof(_abi.PtParams), C.sizeof(_abi.PtSolveInfo),
                                                        C.sizeof(_abi.PtStatsRow)]
    # the struct's last fields, where an added field would shift a mismatched layout
    tail = rows["tail"].split()
    assert [int(v) for v in tail[:3]] == [p.buffer, p.equalize, p.RKC1] == [10, 0, 1]
    assert float(tail[3]) == p.dist_factor
    assert int(rows["invalid"]) == _abi.PT_ERR_INVALID
    assert "nx" in rows["message"]


@pytest.mark.parametrize("option", ["splitting=1", "Chebyshev=1", "non_Newtonian=1", "plastic=1", "iheattype=1"])
def test_fortran_program_refuses_options_outside_the_subset(tmp_path, option):
    """projectracer_amd/fortran/pt_steps.f90 reads the reference's namelist
    file; an option that pt_params cannot carry and the GPU path does not
    build stops the program (code 2) before pt_create, with the supported
    values in the message.  No device is needed."""
    from projectracer_amd import cases
    from projectracer_amd.build import FORTRAN_PROG
    from _support import write_inputs_nml
    if not FORTRAN_PROG.exists():
        pytest.skip(f"{FORTRAN_PROG} not built")
    write_inputs_nml(tmp_path / "inputs.nml", cases.isoviscous(nx=64, nz=64, tpc=20).params)
    txt = (tmp_path / "inputs.nml").read_text()
    name = option.split("=")[0]
    default = f"{name}=0"
    assert default in txt
    (tmp_path / "inputs.nml").write_text(txt.replace(default, option))
    (tmp_path / "steps.nml").write_text("&steps\nnrun=1, precision=0\n/\n")
    proc = subprocess.run([str(FORTRAN_PROG)], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert proc.returncode == 2, proc.stdout + proc.stderr
    assert "outside the supported subset" in proc.stdout
