"""The Fortran side of the boundary: mocca_amd/fortran/mocca_amd_c.f90
declares every C entry point with bind(C); small Fortran programs built
against it drive the library the way shims of the reference's modules would.
Needs amdflang (ROCm); skipped without it."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle_lib import golden, kin_coefs, relerr

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FC = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
LIBDIR = os.path.join(ROOT, "mocca_amd")

pytestmark = pytest.mark.skipif(not os.path.exists(FC), reason="amdflang not available")


def _build(tmp_path, program):
    mod = os.path.join(ROOT, "mocca_amd", "fortran", "mocca_amd_c.f90")
    src = os.path.join(ROOT, "tests", "fortran", program + ".f90")
    exe = str(tmp_path / program)
    subprocess.run([FC, "-c", mod, "-o", str(tmp_path / "mocca_amd_c.o"),
                    "-module-dir", str(tmp_path)], check=True, cwd=tmp_path)
    subprocess.run([FC, src, str(tmp_path / "mocca_amd_c.o"), "-I", str(tmp_path), "-o", exe,
                    "-L" + LIBDIR, "-lmocca_amd", "-Wl,-rpath," + LIBDIR],
                   check=True, cwd=tmp_path)
    return exe


def test_binding_compiles_and_reports_errors(tmp_path):
    exe = _build(tmp_path, "abi_errors")
    out = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "FORTRAN_ABI_OK" in out.stdout
    assert "at least 6 points" in out.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["ev8_todd_fd", "ev8_trc_lag"])
def test_fortran_driver_grad_desc_vs_reference(tmp_path, case):
    g = golden(case)
    m = g["meta"]
    N = m["nx"] * m["ny"] * m["nz"]
    exe = _build(tmp_path, "graddesc_driver")
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    with open(fin, "wb") as f:
        np.array([m["nx"], m["ny"], m["nz"], int(m["trc"]), int(m["lagrange"]), m["nwt"]],
                 "<i4").tofile(f)
        np.array([m["dx"], g["consts"][6], m["dt"] / g["consts"][5], m["momentum"]],
                 "<f8").tofile(f)
        np.asarray(kin_coefs(g), "<f8").tofile(f)
        np.asarray(g["qn"], "<i4").tofile(f)
        np.asarray(g["occ"], "<f8").tofile(f)
        np.asarray(g["psi"], "<f8").tofile(f)
        for k, n in (("U", 2), ("B", 2), ("NablaB", 6), ("W", 18), ("DivW", 6), ("S", 6),
                     ("A", 6), ("DivA", 2)):
            a = np.asarray(g[k], "<f8").reshape(-1)
            assert a.size == n * N
            a.tofile(f)
    print(f"fortran driver: case {case}", flush=True)
    out = subprocess.run([exe, fin, fout], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "FORTRAN_GRADDESC_OK" in out.stdout
    res = np.fromfile(fout, "<f8")
    M = m["nwt"]
    psi = res[:M * 4 * N]
    E = res[M * 4 * N:M * 4 * N + M]
    rho = res[M * 4 * N + 2 * M:]
    assert relerr(psi, g["step1_psi"].reshape(-1)) < 1e-10
    assert relerr(E, g["step1_ed"][:, 0]) < 1e-11
    # densities of the stepped states against the oracle's of the golden states
    assert rho.size == 2 * N and np.all(np.isfinite(rho))
