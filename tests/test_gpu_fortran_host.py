"""GPU: a Fortran host (tests/fortran/pede_host.f90, built against include/mpgpu_mod.f90 by
__graft_entry__.build()) reads a Mille file, runs loopn and the solve through the C-ABI, and gets
the results the Python binding gets from the same file: the Fortran side of the boundary in use
(pede.f90 would make these calls in loop2 and xloopn, src/pede.f90:8501-9009)."""
import os
import subprocess

import numpy as np
import pytest

import mpgpu

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fortran", "_build", "pede_host")


@pytest.mark.parametrize("metsol,matsto", [(mpgpu.MINRES, mpgpu.SPARSE), (mpgpu.INVERSION, mpgpu.FULL)])
def test_fortran_host_matches_python(tmp_path, metsol, matsto):
    assert os.path.exists(EXE), "tests/fortran/_build/pede_host missing: run __graft_entry__.build()"
    rp, ii, gg, _ = mpgpu.pixel_records(47, 3, 4, 3, 2000, 3, 6)
    path = str(tmp_path / "data.bin")
    mpgpu.write_mille(path, rp, ii, gg)
    r = subprocess.run([EXE, path, str(metsol), str(matsto), "5"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.split("\n")
    head = lines[0].split()
    nrec, ntot, nvar, ncon = (int(v) for v in head[:4])
    chi2, ndf, method, istop = float(head[4]), int(head[5]), int(head[6]), int(head[7])
    corr = {int(l.split()[0]): float(l.split()[1]) for l in lines[1:] if l.strip()}

    P = mpgpu.Pede(matsto=matsto, metsol=metsol, mreqenf=5, mcount=1)
    assert P.read_mille(path) == nrec == len(rp) - 1
    P.build_structure()
    st = P.loopn()
    x, sst = P.solve()
    lab, var = P.labels()[:2]
    assert (ntot, nvar, ncon) == (P.ntot, P.nvar, P.ncon)
    assert ndf == st["ndf"] and method == sst["method"] and istop == sst["istop"]
    # two processes: the same arithmetic up to the order of the FP64 atomic additions of the
    # vector (1e-16 relative), which the solve may amplify to ~1e-13
    assert abs(chi2 - st["chi2"]) <= 1e-12 * abs(st["chi2"])
    exp = {int(l): x[v - 1] for l, v in zip(lab, var) if v > 0}
    assert sorted(corr) == sorted(exp)
    scale = max(abs(v) for v in exp.values())
    assert max(abs(corr[k] - exp[k]) for k in exp) <= 1e-10 * scale
