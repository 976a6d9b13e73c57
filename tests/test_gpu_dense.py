"""Dense direct solves on the device through the C-ABI.

'method decomposition' is pede's root-free LDL^T from the last row (chdec2/chslv2,
src/mpnum.f90:891-987): solution, rank and the largest / smallest pivot against the oracle's
restatement, which is bit-identical to the reference routines (tests/test_oracle_ref.py).  A zero
row (a parameter whose derivatives are all zero) is a rank defect that chdec2 reports and skips:
its correction is zero and the rest is the solution of the reduced system.  The LAPACK methods
stop on it with exit code 29 (dpptrf, src/mplapack.f90:144-156).  The FP32 factor with FP64
conjugate-gradient refinement (method 11) must actually run (method 11 reported, refinement
converged) and is judged by max |dp| / sigma_p against the FP64 solution.

FP64 results agree with the oracle to 1e-8 relative (the FP64 bar of the path): the blocked
factor sums in another order than chdec2, which moves the solution by about cond * eps."""
import numpy as np
import pytest

import mpgpu
from helpers import packed_to_dense, rel_err

pytestmark = pytest.mark.gpu


def _dense_problem(seed, nl=10, nx=5, ny=4, ntr=20000, zero_label=None, **opt):
    """Dense detector (every track crosses every layer); loose pre-sigmas on all parameters
    remove the global modes.  zero_label: a label whose derivatives are set to zero."""
    rp, ii, gg, _ = mpgpu.dense_records(seed, nl, nx, ny, ntr)
    labels = mpgpu.pixel_labels(nl, nx, ny)
    if zero_label is not None:
        gg = gg.copy()
        gg[ii == zero_label] = 0.0
        labels = labels[labels != zero_label]
    presig = np.array([0.1 if l % 10 <= 3 else 0.01 for l in labels])
    opt.setdefault("mreqenf", 5)
    P = mpgpu.Pede(matsto=mpgpu.FULL, mcount=1, **opt)
    P.add_records(rp, ii, gg)
    P.set_steering_parameters(labels, np.zeros(len(labels)), presig)
    P.build_structure()
    P.loopn()
    return P


def _oracle_chdec2(orc, P):
    n = P.nall
    g, nrank, evmax, evmin = orc.chdec2(P.matrix_packed(), n)
    b, _ = P.vector()
    return orc.chslv2(g, b), nrank, evmax, evmin


def test_decomposition_matches_chdec2(orc):
    P = _dense_problem(31, metsol=mpgpu.DECOMPOSITION)
    xo, nrank, evmax, evmin = _oracle_chdec2(orc, P)
    x, st = P.solve()
    assert st["method"] == mpgpu.DECOMPOSITION
    assert st["nrank"] == nrank == P.nall
    assert abs(st["evmax"] - evmax) <= 1e-10 * evmax and abs(st["evmin"] - evmin) <= 1e-10 * evmin
    assert rel_err(x, xo) < 1e-8
    assert st["rel_residual"] < 1e-12


def test_decomposition_rank_defect(orc):
    """chdec2 counts only positive pivots and multiplies by the raw (zero) pivot in chslv2."""
    zl = 3 * 100000 + 7 * 10 + 2      # layer 3, module 7, dv
    # zero derivatives are not entries: 'entries 0' keeps the parameter variable
    P = _dense_problem(32, zero_label=zl, metsol=mpgpu.DECOMPOSITION, mreqenf=0)
    lab = P.labels()[0]
    k = int(np.nonzero(lab == zl)[0][0])
    assert P.labels()[1][k] > 0
    A = packed_to_dense(P.matrix_packed(), P.nall)
    iv = P.labels()[1][k] - 1
    assert not A[iv].any()
    xo, nrank, evmax, evmin = _oracle_chdec2(orc, P)
    x, st = P.solve()
    assert nrank == P.nall - 1 and st["nrank"] == nrank
    assert x[iv] == 0.0 and xo[iv] == 0.0
    assert rel_err(x, xo) < 1e-8
    Q = _dense_problem(32, zero_label=zl, metsol=mpgpu.LAPACK_PACKED, mreqenf=0)
    with pytest.raises(mpgpu.MpgpuError) as e:
        Q.solve()
    assert e.value.code == 29


def test_fp32_refinement_runs(orc):
    """Method 11 on the dense problem: the FP32 factor on the matrix cores, FP64 CG refinement
    to the backward-error test of LAPACK's DSPOSV; max |dp|/sigma_p against chdec2/chslv2."""
    P = _dense_problem(33, metsol=mpgpu.CHOLESKY_FP32_IR, fp32_min_n=0)
    xo, nrank, _, _ = _oracle_chdec2(orc, P)
    x, st = P.solve()
    assert st["method"] == mpgpu.CHOLESKY_FP32_IR and st["istop"] == 1, st
    assert 1 <= st["itn"] <= 10, st
    A = packed_to_dense(P.matrix_packed(), P.nall)
    sig = np.sqrt(np.diag(np.linalg.inv(A)))
    dps = np.max(np.abs(x - xo) / sig)
    print(f"FP32+IR: {st['itn']} CG steps, max |dp|/sigma_p = {dps:.2e}, rel. residual {st['rel_residual']:.2e}")
    assert dps < 1e-6
    assert rel_err(x, xo) < 1e-8


def test_fp32_refinement_nonconvergence_falls_back():
    """With one refinement step allowed and a tolerance far below what one step reaches, the
    refinement reports non-convergence (istop 3) and the FP64 factor solves the system."""
    P = _dense_problem(34, metsol=mpgpu.CHOLESKY_FP32_IR, refine_steps=1, refine_tol=1e-30, fp32_min_n=0)
    x, st = P.solve()
    assert st["method"] == mpgpu.DECOMPOSITION and st["istop"] == 3 and st["itn"] == 1
    Q = _dense_problem(34, metsol=mpgpu.DECOMPOSITION)
    x64, _ = Q.solve()
    assert rel_err(x, x64) < 1e-8


def test_fp32_size_rule():
    """Below fp32_min_n parameters (default 16384) method 11 solves with the FP64 factor, which
    is the faster one there; it reports method 3, istop 4, and the decomposition's solution."""
    P = _dense_problem(36, metsol=mpgpu.CHOLESKY_FP32_IR)
    assert P.nvar < 16384
    x, st = P.solve()
    assert st["method"] == mpgpu.DECOMPOSITION and st["istop"] == 4 and st["itn"] == 0, st
    P.set_options(metsol=mpgpu.DECOMPOSITION)          # the same matrix, the FP64 method
    x64, st64 = P.solve()
    assert st64["method"] == mpgpu.DECOMPOSITION and np.array_equal(x, x64)


@pytest.mark.parametrize("n_layers,nx,ny", [(3, 1, 1), (4, 11, 3)])
def test_decomposition_sizes(orc, n_layers, nx, ny):
    """Sizes below one factor block (18) and not a multiple of it (792 = 6 blocks + 24)."""
    P = _dense_problem(35, nl=n_layers, nx=nx, ny=ny, ntr=3000, metsol=mpgpu.DECOMPOSITION)
    xo, nrank, evmax, evmin = _oracle_chdec2(orc, P)
    x, st = P.solve()
    assert st["nrank"] == nrank
    assert rel_err(x, xo) < 1e-8
