"""GPU parity: the HIP path through the C-ABI against the CPU oracle on the same inputs.

Tolerances: matrix, vector, chi^2 and corrections agree to 1e-10 relative (same FP64
arithmetic, only the summation order of FP64 atomics differs); counters, reject counts,
ndf and index maps are exact."""
import numpy as np
import pytest

import labels as LB
import mpgpu
from helpers import packed_to_dense, rel_err

pytestmark = pytest.mark.gpu

TOY_FIXED = [100 * p + 10 * m + k for p in (1, 10) for m in range(5) for k in (1, 2)]


def _pede(rp, ii, gg, steer=None, presig=None, cons=None, **opt):
    P = mpgpu.Pede(**opt)
    P.add_records(rp, ii, gg)
    if steer is not None:
        P.set_steering_parameters(steer, np.zeros(len(steer)), presig)
    if cons:
        P.add_constraints(cons)
    P.build_structure()
    return P


def _check_maps(P, rp, ii, gg, steer=(), presig=None, mreqenf=25, mcount=0):
    lab, var, cnt, gsl = P.labels()
    l2, v2, c2, g2 = LB.label_maps(rp, ii, gg.astype(float), steer_labels=steer, presigma=presig, mreqenf=mreqenf,
                                   mcount=mcount)
    assert np.array_equal(lab, l2)
    assert np.array_equal(var, v2)
    assert np.array_equal(cnt, c2)
    assert np.array_equal(gsl, g2)
    return lab, var


def _compare_pass(P, orc, rp, ii, gg, lab, var, pre_weight=None, **kw):
    nv = P.nvar
    s_gpu = P.loopn(icalcm=kw.get("icalcm", 1), nloopn=kw.get("nloopn", 1), chicut=kw.get("chicut", 0.0))
    mat, vec, cnt, s_orc = orc.loopn(rp, ii, gg.astype(np.float64), lab, var, P.parameters(), nv,
                                     pre_weight=pre_weight, **kw)
    v_gpu, c_gpu = P.vector()
    assert s_gpu["nrej"] == s_orc["nrej"]
    assert s_gpu["ndf"] == s_orc["ndf"]
    assert abs(s_gpu["chi2"] - s_orc["chi2"]) <= 1e-11 * abs(s_orc["chi2"])
    assert np.array_equal(c_gpu, cnt)
    assert rel_err(v_gpu[:nv], vec) < 1e-10
    if kw.get("icalcm", 1) == 1:
        m_gpu = P.matrix_packed()
        n = P.nall
        A = packed_to_dense(m_gpu, n)[:nv, :nv]
        assert rel_err(A, packed_to_dense(mat, nv)) < 1e-10
    return s_gpu, mat, vec


@pytest.mark.parametrize("matsto", [mpgpu.FULL, mpgpu.SPARSE])
def test_toy_pass(orc, matsto):
    rp, ii, gg, truth = mpgpu.toy_records(42, 10000)
    pres = [-1.0] * len(TOY_FIXED)
    P = _pede(rp, ii, gg, TOY_FIXED, pres, matsto=matsto, mreqenf=10, metsol=1 if matsto == 1 else 4)
    lab, var = _check_maps(P, rp, ii, gg, TOY_FIXED, pres, mreqenf=10)
    _compare_pass(P, orc, rp, ii, gg, lab, var)


@pytest.mark.parametrize("matsto", [mpgpu.FULL, mpgpu.SPARSE])
def test_pixel_pass(orc, matsto):
    rp, ii, gg, _ = mpgpu.pixel_records(11, 4, 6, 3, 3000, 3, 6)
    P = _pede(rp, ii, gg, matsto=matsto, mreqenf=5, mcount=1, metsol=1 if matsto == 1 else 4)
    lab, var = _check_maps(P, rp, ii, gg, mreqenf=5, mcount=1)
    _compare_pass(P, orc, rp, ii, gg, lab, var)


def test_two_phase_build_in_batches(orc, monkeypatch):
    """Group-block storage with an update buffer far smaller than the pass (8 MB): the records
    are taken in several batches, each followed by its block gather, which adds the batch's sums
    of matrix and vector values to the blocks.  Matrix, vector, chi^2, ndf, rejects and counters
    against the oracle; the batched pass equals the one-batch pass within FP64 rounding."""
    rp, ii, gg, _ = mpgpu.pixel_records(19, 4, 6, 3, 3000, 3, 6)
    opt = dict(matsto=mpgpu.SPARSE, mreqenf=5, mcount=1, metsol=4)
    P1 = _pede(rp, ii, gg, **opt)
    s1 = P1.loopn()
    assert s1["n_gather"] == 1
    monkeypatch.setenv("MPGPU_UPDATE_BUFFER_GB", "0.0078125")
    P = _pede(rp, ii, gg, **opt)
    lab, var = _check_maps(P, rp, ii, gg, mreqenf=5, mcount=1)
    s, _, _ = _compare_pass(P, orc, rp, ii, gg, lab, var)
    assert s["n_gather"] >= 2
    n = P.nall
    assert rel_err(packed_to_dense(P.matrix_packed(), n), packed_to_dense(P1.matrix_packed(), n)) < 1e-13
    assert rel_err(P.vector()[0], P1.vector()[0]) < 1e-13
    P.close()
    P1.close()


def test_downweighting_pass(orc):
    rp, ii, gg, _ = mpgpu.toy_records(5, 4000)
    gg = gg.copy()
    recs = LB.decode(rp, ii, gg)
    for r in range(0, 4000, 17):
        gg[recs[r][min(3, len(recs[r]) - 1)][0]] += 0.2
    P = _pede(rp, ii, gg, matsto=mpgpu.FULL, mreqenf=1, lhuber=3, dwcut=0.2)
    lab, var = _check_maps(P, rp, ii, gg, mreqenf=1)
    par = np.zeros(P.ntot)
    par[::3] = 1e-3
    P.set_parameters(par)
    for nloopn, chicut in ((1, 0.0), (2, 4.0), (3, 1.5)):
        _compare_pass(P, orc, rp, ii, gg, lab, var, nloopn=nloopn, lhuber=3, chicut=chicut, dwcut=0.2)


@pytest.mark.parametrize("matsto", [mpgpu.FULL, mpgpu.SPARSE])
def test_rank_deficient_local_fits(orc, matsto):
    """Local systems without full rank: in every 5th record local parameter 3 never occurs (an
    empty column), in every 7th its derivative is ten times that of parameter 1 in every
    equation (proportional columns, the pivot search stops at rank 3).  pede rejects both
    (nrank != nalc); rejects, chi^2, ndf and the update must agree with the oracle."""
    rp, ii, gg, _ = mpgpu.pixel_records(13, 4, 6, 3, 3000, 3, 6)
    ii = ii.copy()
    gg = gg.copy()
    recs = LB.decode(rp, ii, gg)
    nmod = 0
    for r, eqs in enumerate(recs):
        if r % 5 and r % 7:
            continue
        nmod += 1
        for (sL, sG, lG) in eqs:
            for w in range(sL + 1, sG):
                if ii[w] == 3:
                    if r % 5 == 0:
                        ii[w] = 4 if all(ii[x] != 4 for x in range(sL + 1, sG)) else ii[w]
                    else:
                        gg[w] = 10.0 * gg[sL + 1]
    P = _pede(rp, ii, gg, matsto=matsto, mreqenf=5, mcount=1, metsol=1 if matsto == 1 else 4)
    lab, var = _check_maps(P, rp, ii, gg, mreqenf=5, mcount=1)
    s_gpu, _, _ = _compare_pass(P, orc, rp, ii, gg, lab, var)
    assert s_gpu["nrej"][1] >= nmod * 9 // 10


@pytest.mark.parametrize("matsto", [mpgpu.FULL, mpgpu.SPARSE])
def test_brokenlines_band_pass(orc, matsto):
    """Records whose local fits take pede's bordered-band path (sqmibb, band 3 + border 4):
    the GPU band kernel's matrix, vector, chi^2, ndf, reject counts and counters against the
    oracle, which follows sqmibb step by step."""
    rp, ii, gg = mpgpu.brokenlines_records(31, 30, 4, 3, 1500, 8, 30)
    P = _pede(rp, ii, gg, matsto=matsto, mreqenf=1, mcount=1)
    lab, var = _check_maps(P, rp, ii, gg, mreqenf=1, mcount=1)
    s_gpu, mat, vec = _compare_pass(P, orc, rp, ii, gg, lab, var)
    assert s_gpu["naccept"] > 1400


def _band_side(recs, ii, r):
    """bordered-band choice of one record (loopbf, src/pede.f90:3815-3874): 0 full, 1 border
    first, 2 border last"""
    nalc = max(int(ii[w]) for (sL, sG, lG) in recs[r] for w in range(sL + 1, sG))
    if nalc <= 5:
        return 0
    ibh = [0] * (2 * nalc + 2)
    for (sL, sG, lG) in recs[r]:
        loc = [int(ii[w]) for w in range(sL + 1, sG)]
        for j in range(len(loc)):
            for k in range(j + 1):
                a, b = loc[j], loc[k]
                d = abs(a - b) + 1
                ibh[d] = max(ibh[d], min(a, b))
                ibh[nalc + d] = max(ibh[nalc + d], min(nalc + 1 - a, nalc + 1 - b))
    icmn, side = nalc ** 3, 0
    for sd in (1, 2):
        if side:
            break
        kbdr = 0
        for k in range(nalc, 1, -1):
            kbdr = max(kbdr, ibh[k if sd == 1 else k + nalc])
            t = k - 2 + kbdr + 1
            icost = 6 * (nalc - kbdr) * t * t + 2 * kbdr ** 3
            if icost < icmn:
                icmn, side = icost, sd
    return side


@pytest.mark.parametrize("matsto", [mpgpu.FULL, mpgpu.SPARSE])
def test_brokenlines_band_border_last(orc, matsto):
    """The same broken-lines records with their local parameters numbered backwards: the border
    is then at the end of the local system and pede's cost model picks the border-last
    factorization (sqmibb2, src/mpnum.f90:3390-3628).  The GPU band kernel's pass against the
    oracle, and equal to the pass of the records as generated (a renumbering of the local
    parameters does not change the fit)."""
    rp, ii, gg = mpgpu.brokenlines_records(33, 24, 3, 3, 1000, 8, 24)
    recs = LB.decode(rp, ii, gg)
    # local index j -> nalc + 1 - j, the local words of each equation kept ascending (pede adds
    # them into the packed lower triangle assuming that order, src/pede.f90:3818-3824)
    rev, gr = ii.copy(), gg.copy()
    for eqs in recs:
        nalc = max(int(ii[w]) for (sL, sG, lG) in eqs for w in range(sL + 1, sG))
        for (sL, sG, lG) in eqs:
            loc = sorted((nalc + 1 - int(ii[w]), gg[w]) for w in range(sL + 1, sG))
            for q, (j, d) in enumerate(loc):
                rev[sL + 1 + q], gr[sL + 1 + q] = j, d
    sides = [_band_side(recs, rev, r) for r in range(len(recs))]
    assert sides.count(2) > 900, sides.count(2)
    assert all(_band_side(recs, ii, r) == 1 for r in range(0, len(recs), 50))
    P = _pede(rp, rev, gr, matsto=matsto, mreqenf=1, mcount=1)
    lab, var = _check_maps(P, rp, rev, gr, mreqenf=1, mcount=1)
    s_rev, mat_rev, vec_rev = _compare_pass(P, orc, rp, rev, gr, lab, var)
    Q = _pede(rp, ii, gg, matsto=matsto, mreqenf=1, mcount=1)
    s_fwd, mat_fwd, vec_fwd = _compare_pass(Q, orc, rp, ii, gg, lab, var)
    assert s_rev["ndf"] == s_fwd["ndf"] and s_rev["naccept"] == s_fwd["naccept"] > 900
    assert abs(s_rev["chi2"] - s_fwd["chi2"]) <= 1e-9 * abs(s_fwd["chi2"])
    assert rel_err(vec_rev, vec_fwd) < 1e-9
    nv = P.nvar
    assert rel_err(packed_to_dense(mat_rev, nv), packed_to_dense(mat_fwd, nv)) < 1e-9


def test_brokenlines_band_downweighting(orc):
    """Band-path records with outliers: down-weighting iterations (lhuber 3), chi^2 cut and
    parameters away from zero, against the oracle."""
    rp, ii, gg = mpgpu.brokenlines_records(32, 24, 3, 3, 1200, 8, 24)
    gg = gg.copy()
    recs = LB.decode(rp, ii, gg)
    for r in range(0, 1200, 13):
        gg[recs[r][min(5, len(recs[r]) - 1)][0]] += 0.05
    P = _pede(rp, ii, gg, matsto=mpgpu.SPARSE, mreqenf=1, mcount=1, lhuber=3, dwcut=0.2)
    lab, var = _check_maps(P, rp, ii, gg, mreqenf=1, mcount=1)
    par = np.zeros(P.ntot)
    par[::5] = 2e-4
    P.set_parameters(par)
    for nloopn, chicut in ((1, 0.0), (2, 60.0), (3, 25.0)):
        _compare_pass(P, orc, rp, ii, gg, lab, var, nloopn=nloopn, lhuber=3, chicut=chicut, dwcut=0.2)


def test_inversion_matches_sqminl(orc):
    rp, ii, gg, truth = mpgpu.toy_records(42, 10000)
    pres = [-1.0] * len(TOY_FIXED)
    P = _pede(rp, ii, gg, TOY_FIXED, pres, matsto=mpgpu.FULL, mreqenf=10, metsol=mpgpu.INVERSION)
    lab, var = _check_maps(P, rp, ii, gg, TOY_FIXED, pres, mreqenf=10)
    P.loopn()
    A = P.matrix_packed()
    b, _ = P.vector()
    x, st = P.solve()
    inv, xo, nr = orc.sqminl(A, b)
    assert st["nrank"] == nr == P.nall
    assert rel_err(x, xo) < 1e-10
    err, gcor = P.errors()
    e_or = np.sqrt(np.abs(np.array([inv[(i * i + i) // 2 - 1] for i in range(1, P.nvar + 1)], dtype=np.float32)))
    assert np.allclose(err, e_or, rtol=1e-6)
    assert st["rel_residual"] < 1e-10


def _pixel_cons_problem(seed=21, nl=4, nx=5, ny=3, ntr=4000):
    rp, ii, gg, _ = mpgpu.pixel_records(seed, nl, nx, ny, ntr, 3, 6)
    cons = mpgpu.pixel_constraints(nl, nx, ny)
    # loose pre-sigmas on dw and the rotations remove the null modes that straight tracks leave
    # (coherent tilt of all layers, rotation about the beam axis)
    gam = [(l + 1) * 100000 + (m + 1) * 10 + k for l in range(nl) for m in range(nx * ny) for k in (3, 4, 5, 6)]
    return rp, ii, gg, cons, gam


def _kkt_reference(P):
    A = packed_to_dense(P.matrix_packed(), P.nall)
    b, _ = P.vector()
    return np.linalg.solve(A, b), A, b


@pytest.mark.parametrize("mbandw", [0, 6])
def test_minres_sparse_constraints(orc, mbandw):
    """MINRES on the group-block matrix with Lagrange rows.  An iterative solve is judged the
    way pede's own MINRES is (src/minresModule.f90:591-600): relative residual at the
    tolerance, the solution's deviation from the direct solve small against the parameter
    errors (max |dp|/sigma_p), and the iteration count of the oracle's MINRES on the same
    matrix (without preconditioner, so that the run is fully specified)."""
    rp, ii, gg, cons, gam = _pixel_cons_problem()
    P = _pede(rp, ii, gg, gam, mpgpu.pixel_presigma(gam), cons, matsto=mpgpu.SPARSE, metsol=mpgpu.MINRES, mreqenf=5,
              mcount=1, mbandw=mbandw, mrestl=1e-12, withmultipliers=1)
    assert P.ncon == 12
    P.loopn()
    x, st = P.solve()
    xr, A, b = _kkt_reference(P)
    assert st["istop"] in (1, 2)
    assert np.linalg.norm(A @ x - b) / np.linalg.norm(b) < 1e-9
    sig = np.sqrt(np.abs(np.diag(np.linalg.inv(A))[:P.nvar]))
    assert np.max(np.abs(x[:P.nvar] - xr[:P.nvar]) / sig) < 1e-4
    if mbandw > 0:
        # the oracle's MINRES with the reference's band preconditioner (lltdec/equdecs/equslvs,
        # pinned bitwise to src/mpnum.f90) on the same matrix: same stop code, iteration count
        # within a few iterations, so the GPU preconditioner is the same operator
        msolve = _oracle_band_precond(orc, A, P.nvar, P.ncon, mbandw)
        xo, so, io, _ = orc.minres(P.nall, lambda v: A @ v, msolve, b, 1e-12, precon=True)
        assert so == st["istop"]
        assert abs(io - st["itn"]) <= max(3, io // 20), (io, st["itn"])
        assert np.max(np.abs(x[:P.nvar] - xo[:P.nvar]) / sig) < 1e-4
    if mbandw == 0:
        # Without preconditioner this KKT system (condition ~1e15) does not converge, and
        # iterates of runs that differ only in rounding drift apart after some 50 iterations.
        # The recurrences are compared instead on the first 30 iterations: GPU iterate against
        # the oracle's MINRES restatement on the same matrix.
        xo, so, io, _ = orc.minres(P.nall, lambda v: A @ v, None, b, 1e-12, itnlim=30, precon=False)
        P2 = _pede(rp, ii, gg, gam, mpgpu.pixel_presigma(gam), cons, matsto=mpgpu.SPARSE, metsol=mpgpu.MINRES,
                   mreqenf=5, mcount=1, mbandw=-1, mrestl=1e-12, itnlim=30, withmultipliers=1)
        P2.loopn()
        x2, st2 = P2.solve()
        assert (st2["istop"], st2["itn"]) == (so, io) == (5, 30)
        assert np.max(np.abs(x2[:P.nvar] - xo[:P.nvar]) / sig) < 1e-6


@pytest.mark.parametrize("mbandw", [0, 6])
def test_minres_repeatable(mbandw):
    """MINRES runs its iterations from a captured graph of eight iterations (replayed until the
    device-side stop code is set): solving the same system three times, with eager work and
    synchronisation in between, gives bitwise the same solution, iteration count and stop code."""
    rp, ii, gg, cons, gam = _pixel_cons_problem(seed=23)
    P = _pede(rp, ii, gg, gam, mpgpu.pixel_presigma(gam), cons, matsto=mpgpu.SPARSE, metsol=mpgpu.MINRES, mreqenf=5,
              mcount=1, mbandw=mbandw, mrestl=1e-12, itnlim=203, withmultipliers=1)
    P.loopn()
    runs = []
    for _ in range(3):
        x, st = P.solve()
        P.avprod(x)                     # eager work between the solves (the matrix is not rebuilt:
        P.vector()                      # the vector's FP64 atomics would change its last bits)
        runs.append((x.copy(), st["itn"], st["istop"]))
    for x, itn, istop in runs[1:]:
        assert np.array_equal(x, runs[0][0])
        assert (itn, istop) == runs[0][1:]
    assert runs[0][1] <= 203


@pytest.mark.parametrize("lprecm", [0, 1])
def test_band_factor_chunked_equals_sequential(monkeypatch, lprecm):
    """The band preconditioner's factorization (lltdec, src/mpnum.f90:2155-2230) by chunks of
    rows in parallel gives the sequential factor bit for bit: with the same matrix and
    right-hand side, MINRES returns bitwise the same solution, iteration count and stop code
    whichever factorization built its preconditioner (4,320 parameters: above the size where
    the chunked one is used)."""
    nl, nx, ny = 8, 10, 9
    rp, ii, gg, _ = mpgpu.pixel_records(27, nl, nx, ny, 20000, 3, 6)
    cons = mpgpu.pixel_constraints(nl, nx, ny)
    gam = [(l + 1) * 100000 + (m + 1) * 10 + k for l in range(nl) for m in range(nx * ny) for k in (3, 4, 5, 6)]
    runs, b = [], None
    for seq in (False, True):
        if seq:
            monkeypatch.setenv("MPGPU_SEQUENTIAL_LLT", "1")
        else:
            monkeypatch.delenv("MPGPU_SEQUENTIAL_LLT", raising=False)
        P = _pede(rp, ii, gg, gam, mpgpu.pixel_presigma(gam), cons, matsto=mpgpu.SPARSE, metsol=mpgpu.MINRES,
                  mreqenf=5, mcount=1, mbandw=6, lprecm=lprecm, mrestl=1e-12, withmultipliers=1)
        assert P.nvar >= 4096
        P.loopn()
        if b is None:
            b = P.vector()[0]
        x, st = P.solve(rhs=b)
        runs.append((x, st["itn"], st["istop"]))
        P.close()
    assert runs[0][1] > 10
    assert np.array_equal(runs[0][0], runs[1][0]) and runs[0][1:] == runs[1][1:]


@pytest.mark.parametrize("mbandw,mrmode", [(-1, 0), (-1, 2), (6, 0), (6, 2)])
def test_minresqlp_sparse_constraints(orc, mbandw, mrmode):
    """MINRES-QLP (mresmode 0: QR steps, QLP once ill-conditioned; 2: QLP only) on the
    group-block KKT matrix against the oracle's restatement (pinned bitwise to
    src/minresqlpModule.f90) driven with the same matrix and preconditioner: same stop code,
    iteration count within a few, solution within 1e-4 of the parameter errors."""
    rp, ii, gg, cons, gam = _pixel_cons_problem()
    itnlim = 2000 if mbandw > 0 else 30      # unpreconditioned: first 30 iterations (see MINRES test)
    P = _pede(rp, ii, gg, gam, mpgpu.pixel_presigma(gam), cons, matsto=mpgpu.SPARSE, metsol=mpgpu.MINRES_QLP, mreqenf=5,
              mcount=1, mbandw=mbandw, mrestl=1e-12, mrmode=mrmode, itnlim=itnlim, withmultipliers=1)
    P.loopn()
    x, st = P.solve()
    xr, A, b = _kkt_reference(P)
    sig = np.sqrt(np.abs(np.diag(np.linalg.inv(A))[:P.nvar]))
    if mbandw > 0:
        assert np.linalg.norm(A @ x - b) / np.linalg.norm(b) < 1e-8
        assert np.max(np.abs(x[:P.nvar] - xr[:P.nvar]) / sig) < 1e-4
    msolve = _oracle_band_precond(orc, A, P.nvar, P.ncon, mbandw) if mbandw > 0 else None
    trancond = 1.0 if mrmode == 2 else 1.0e7
    xo, so, io, _ = orc.minresqlp(P.nall, lambda v: A @ v, msolve, b, 1e-12, itnlim=itnlim, precon=mbandw > 0,
                                  trancond=trancond)
    assert so == st["istop"], (so, st["istop"])
    assert abs(io - st["itn"]) <= max(3, io // 20), (io, st["itn"])
    assert np.max(np.abs(x[:P.nvar] - xo[:P.nvar]) / sig) < (1e-4 if mbandw > 0 else 1e-6)


def _oracle_band_precond(orc, A, n, m, mb):
    """pede's band preconditioner (vmprep india layout, constraint blocks of overlapping ranges,
    equdecs/equslvs) built by the oracle for the dense KKT matrix A."""
    import ctypes as C
    india = np.zeros(n, np.int32)
    for i in range(1, min(mb, n) + 1):
        india[i - 1] = (i * i + i) // 2
    for i in range(min(mb, n) + 1, n + 1):
        india[i - 1] = india[i - 2] + mb
    band = np.zeros(india[-1])
    for i in range(1, n + 1):
        for j in range(max(1, i - mb + 1), i + 1):
            band[india[i - 1] - i + j - 1] = A[i - 1, j - 1]
    rows = [np.nonzero(A[n + j, :n])[0] for j in range(m)]
    first = [int(r[0]) + 1 for r in rows]
    last = [int(r[-1]) + 1 for r in rows]
    l, kpart, j = [], [], 0
    while j < m:
        lo, hi, k = first[j], last[j], j + 1
        while k < m and first[k] <= hi:
            hi, lo, k = max(hi, last[k]), min(lo, first[k]), k + 1
        nin = hi - lo + 1 + mb - 1
        l += [j + 1, k, lo - 1, nin]
        for q in range(j, k):
            row = np.zeros(nin)
            row[: hi - lo + 1] = A[n + q, lo - 1:hi]
            kpart.append(row)
        j = k
    l = np.array(l, np.int32)
    nb = len(l) // 4
    cc = np.concatenate([band, np.zeros(m * (m + 1) // 2), np.concatenate(kpart)])
    ind = np.concatenate([india, np.zeros(m, np.int32)])
    n1, n2 = C.c_int32(), C.c_int32()
    P = lambda a, t=C.c_double: a.ctypes.data_as(C.POINTER(t))
    orc.lib.orc_equdecs(n, m, nb, 0, P(cc), P(ind, C.c_int32), P(l, C.c_int32), C.byref(n1), C.byref(n2))

    def msolve(v):
        y = np.ascontiguousarray(v, dtype=np.float64).copy()
        orc.lib.orc_equslvs(n, m, nb, P(cc), P(ind, C.c_int32), P(l, C.c_int32), P(y))
        return y
    return msolve


def _dense_constraints(P, cons):
    """constraint matrix over the variable parameters and residuals d = rhs - C p (addcst,
    src/pede.f90:730-761) from the constraint blocks and the current parameters"""
    lab, var, _, _ = P.labels()
    pos = {int(l): i for i, l in enumerate(lab)}
    par = P.parameters()
    C, d = np.zeros((len(cons), P.nvar)), np.zeros(len(cons))
    for j, (r, _, labs, facs) in enumerate(cons):
        d[j] = r
        for l, f in zip(labs, facs):
            t = pos[int(l)]
            d[j] -= f * par[t]
            if var[t] > 0:
                C[j, var[t] - 1] += f
    return C, d


def _oracle_elimination(orc, A, b, C, d, mb, rtol, qlp=False, trancond=1.0e7):
    """pede's MINRES with eliminated constraints (mminrs / mminrsqlp, src/pede.f90:8270-8330,
    8370-8430) from the oracle's QL routines (pinned bitwise to src/mpqldec.f90): QL
    decomposition in pede's constraint layout, reduced system (Q^T A Q)_ff, right-hand side
    Q^T (b - A Q [0; y]) with L^T y = d, pede's band (or diagonal) preconditioner of Q^T A Q
    (qlpssq) decomposed by equdecs (or precons), the oracle's MINRES / MINRES-QLP."""
    import ctypes as CT
    import oracle
    n, m = A.shape[0], C.shape[0]
    nf = n - m
    cons = [(list(np.nonzero(C[j])[0] + 1), list(C[j][np.nonzero(C[j])[0]])) for j in range(m)]
    lay = oracle.ql_layout(n, cons)
    q = orc.qldecb(n, lay, aligned=True)
    Q = orc.qlmlq(q, np.eye(n), False)
    y = orc.qlbsub(q, d[lay["order"]])
    e = np.zeros(n)
    e[nf:] = y
    bred = (Q.T @ (b - A @ (Q @ e)))[:nf]
    M = (Q.T @ A @ Q)[:nf, :nf]
    P = lambda a, t=CT.c_double: a.ctypes.data_as(CT.POINTER(t))
    w = max(mb, 1)
    band = np.concatenate([A[i, max(0, i - w + 1):i + 1] for i in range(n)])
    bq = orc.qlpssq(q, A, band, w, True)[:sum(min(i + 1, w) for i in range(nf))].copy()
    if mb > 0:
        india = np.cumsum([min(i + 1, mb) for i in range(nf)]).astype(np.int32)
        dummy = np.zeros(4, np.int32)
        n1, n2 = CT.c_int32(), CT.c_int32()
        orc.lib.orc_equdecs(nf, 0, 0, 0, P(bq), P(india, CT.c_int32), P(dummy, CT.c_int32), CT.byref(n1), CT.byref(n2))

        def msolve(v):
            yv = np.ascontiguousarray(v, dtype=np.float64).copy()
            orc.lib.orc_equslvs(nf, 0, 0, P(bq), P(india, CT.c_int32), P(dummy, CT.c_int32), P(yv))
            return yv
    else:
        cu = np.where(bq > 0, 1.0 / np.sqrt(np.where(bq > 0, bq, 1.0)), 0.0)

        def msolve(v):
            return cu * cu * v
    if qlp:
        xo, so, io, _ = orc.minresqlp(nf, lambda v: M @ v, msolve, bred, rtol, precon=True, trancond=trancond)
    else:
        xo, so, io, _ = orc.minres(nf, lambda v: M @ v, msolve, bred, rtol, precon=True)
    return Q @ np.concatenate([xo, y]), so, io


@pytest.mark.parametrize("matsto,metsol,mbandw,rhs", [
    (mpgpu.SPARSE, mpgpu.MINRES, 6, 0.0), (mpgpu.SPARSE, mpgpu.MINRES, 0, 0.0), (mpgpu.FULL, mpgpu.MINRES, 6, 0.0),
    (mpgpu.SPARSE, mpgpu.MINRES_QLP, 6, 0.0), (mpgpu.SPARSE, mpgpu.MINRES, 6, 2e-3)])
def test_minres_elimination(orc, matsto, metsol, mbandw, rhs):
    """Constraints eliminated by QL decomposition (pede's default constraint mode): the solve
    works on nvar - ncon parameters.  Against the dense KKT solve of the same matrix (the
    constrained least-squares solution), the constraints C x = d met, and against the oracle's
    elimination MINRES (see _oracle_elimination): same stop code, iteration count within a few,
    solution within 1e-4 of the parameter errors.  rhs != 0 gives non-zero constraint residuals,
    so that the solve for the eliminated components (L^T y = d) matters."""
    rp, ii, gg, cons, gam = _pixel_cons_problem()
    cons = [(rhs * (j + 1), s, l, f) for j, (_, s, l, f) in enumerate(cons)]
    P = _pede(rp, ii, gg, gam, mpgpu.pixel_presigma(gam), cons, matsto=matsto, metsol=metsol, mreqenf=5, mcount=1,
              mbandw=mbandw, mrestl=1e-12, withmultipliers=0)
    assert P.ncon == 0 and P.nall == P.nvar
    P.loopn()
    x, st = P.solve()
    A = packed_to_dense(P.matrix_packed(), P.nvar)
    b, _ = P.vector()
    C, d = _dense_constraints(P, cons)
    m = C.shape[0]
    K = np.block([[A, C.T], [C, np.zeros((m, m))]])
    xr = np.linalg.solve(K, np.concatenate([b, d]))[:P.nvar]
    sig = np.sqrt(np.abs(np.diag(np.linalg.inv(K))[:P.nvar]))
    if metsol == mpgpu.MINRES:
        assert st["istop"] in (1, 2)
    assert st["rel_residual"] < 1e-9
    assert np.abs(C @ x - d).max() < 1e-10 * max(1.0, np.abs(C).max() * np.abs(x).max())
    assert np.max(np.abs(x - xr) / sig) < 1e-4
    xo, so, io = _oracle_elimination(orc, A, b, C, d, mbandw, 1e-12, qlp=metsol == mpgpu.MINRES_QLP)
    assert so == st["istop"], (so, st["istop"])
    assert abs(io - st["itn"]) <= max(3, io // 20), (io, st["itn"])
    assert np.max(np.abs(x - xo) / sig) < 1e-4


@pytest.mark.parametrize("fortran", [False, True])
def test_mille_file_plain_and_gzip(tmp_path, fortran):
    """Mille binary files are read plain or gzip-compressed (pede reads C files through the
    Mille library, which accepts both, src/pede.f90:1689): the same records written plain, as a
    .gz file and read back through mpgpu_read_mille_file give the pass of the in-memory records."""
    import gzip
    rp, ii, gg, _ = mpgpu.pixel_records(31, 3, 4, 3, 1500, 3, 6)
    plain = str(tmp_path / "r.bin")
    mpgpu.write_mille(plain, rp, ii, gg, fortran=fortran)
    with open(plain, "rb") as fi, gzip.open(plain + ".gz", "wb") as fo:
        fo.write(fi.read())
    res = []
    for src in (None, plain, plain + ".gz"):
        P = mpgpu.Pede(matsto=mpgpu.SPARSE, metsol=mpgpu.MINRES, mreqenf=5, mcount=1)
        if src is None:
            P.add_records(rp, ii, gg)
        else:
            assert P.read_mille(src, fortran=fortran) == len(rp) - 1
        P.build_structure()
        st = P.loopn()
        res.append((st["chi2"], st["ndf"], st["nrej"], st["naccept"], P.labels()[0].tolist()))
    assert res[0] == res[1] == res[2]


def test_inversion_with_multipliers():
    rp, ii, gg, cons, gam = _pixel_cons_problem(seed=22, ntr=3000)
    P = _pede(rp, ii, gg, gam, mpgpu.pixel_presigma(gam), cons, matsto=mpgpu.FULL, metsol=mpgpu.INVERSION, mreqenf=5, mcount=1)
    P.loopn()
    xr, A, b = _kkt_reference(P)
    x, st = P.solve()
    assert st["nrank"] == P.nall
    assert rel_err(x, xr) < 1e-9


@pytest.mark.parametrize("metsol", [mpgpu.DECOMPOSITION, mpgpu.CHOLESKY_FP32_IR])
def test_cholesky(metsol):
    rp, ii, gg, cons, gam = _pixel_cons_problem(seed=23, nl=6, nx=8, ny=4, ntr=20000)
    P = _pede(rp, ii, gg, gam, mpgpu.pixel_presigma(gam), cons, matsto=mpgpu.FULL, metsol=metsol, mreqenf=5, mcount=1,
              fp32_min_n=0)
    P.loopn()
    xr, A, b = _kkt_reference(P)
    x, st = P.solve()
    # the method that actually ran: the FP32 factor with converged FP64 refinement
    assert st["method"] == metsol, st
    if metsol == mpgpu.CHOLESKY_FP32_IR:
        assert st["istop"] == 1 and 1 <= st["itn"] <= 20, st
    tol = 1e-9 if metsol == mpgpu.DECOMPOSITION else 1e-8
    assert rel_err(x[:P.nvar], xr[:P.nvar]) < tol
    # max |dp| / sigma_p for the mixed-precision factorisation
    Ainv = np.linalg.inv(A)
    sig = np.sqrt(np.abs(np.diag(Ainv)[:P.nvar]))
    dps = np.max(np.abs(x[:P.nvar] - xr[:P.nvar]) / sig)
    print(f"method {st['method']}: {st['itn']} refinement steps, max |dp|/sigma_p = {dps:.2e}")
    assert dps < 1e-6


@pytest.mark.parametrize("metsol", [mpgpu.LAPACK_PACKED, mpgpu.LAPACK_UNPACKED])
@pytest.mark.parametrize("constrained", [True, False])
def test_lapack_with_errors(metsol, constrained):
    """'method fullLAPACK' / 'unpackedLAPACK' with 'LAPACKwitherrors' (mdptrf / mdutrf,
    src/mplapack.f90:41-355; inverse from the factor after the iterations,
    src/pede.f90:9017-9120).  The solution against numpy's solve of the same matrix, and the
    errors sign * sqrt|Cov_ii| and global correlations (prtglo, src/pede.f90:4536-4546) with Cov
    the parameter block of the inverse KKT matrix: with eliminated constraints pede forms
    Q [Abar^-1 0; 0 0] Q^T, the same matrix (the first nvar - ncon columns of Q span the null
    space of C).  Constrained: the pixel problem with its 18 constraints; unconstrained: a
    pre-sigma on every parameter.  The errors also equal those of 'method inversion'."""
    rp, ii, gg, cons, gam = _pixel_cons_problem(seed=25, nl=6, nx=8, ny=4, ntr=20000)
    if constrained:
        steer, presig = gam, mpgpu.pixel_presigma(gam)
    else:
        steer = [(l + 1) * 100000 + (m + 1) * 10 + k for l in range(6) for m in range(32) for k in range(1, 7)]
        presig, cons = np.full(len(steer), 0.05), None
    P = _pede(rp, ii, gg, steer, presig, cons, matsto=mpgpu.FULL, metsol=metsol, mreqenf=5, mcount=1,
              lapack_errors=1)
    P.loopn()
    xr, A, b = _kkt_reference(P)
    x, st = P.solve()
    nv = P.nvar
    assert st["method"] == metsol and P.ncon == (len(cons) if constrained else 0)
    assert rel_err(x[:nv], xr[:nv]) < 1e-9
    cov = np.diag(np.linalg.inv(A))[:nv]
    e, g = P.errors()
    e_ref = (np.sign(cov) * np.sqrt(np.abs(cov).astype(np.float32))).astype(np.float64)
    assert np.max(np.abs(e - e_ref) / np.abs(e_ref)) < 1e-6
    g2 = 1.0 - 1.0 / (cov * np.diag(A)[:nv])
    g_ref = np.where((g2 >= 0) & (g2 <= 1), np.sqrt(np.clip(g2, 0, 1)).astype(np.float32), -1.0)
    assert np.max(np.abs(g - g_ref)) < 1e-5
    Q = _pede(rp, ii, gg, steer, presig, cons, matsto=mpgpu.FULL, metsol=mpgpu.INVERSION, mreqenf=5, mcount=1)
    Q.loopn()
    Q.solve()
    e2, g2 = Q.errors()
    assert np.max(np.abs(e - e2) / np.abs(e2)) < 1e-6 and np.max(np.abs(g - g2)) < 1e-5


def test_preconditioner_kept_until_matrix_changes():
    """pede builds the MINRES preconditioner only after a matrix pass (src/pede.f90:8292-8306);
    the library keeps it until the matrix changes.  With one right-hand side for every solve
    (the vector's FP64 atomics differ in the last bits between passes): two solves on one
    matrix agree bit for bit, and after mupdat the solve equals, bit for bit, that of a fresh
    context given the same update (a stale band factor would take another iteration path)."""
    rp, ii, gg, cons, gam = _pixel_cons_problem(seed=26)

    def make():
        P = _pede(rp, ii, gg, gam, mpgpu.pixel_presigma(gam), cons, matsto=mpgpu.SPARSE, metsol=mpgpu.MINRES, mreqenf=5,
                  mcount=1, mbandw=6, mrestl=1e-12, withmultipliers=1)
        P.loopn()
        return P
    U = ([1, 2], [1, 2], [5.0e3, 5.0e3])
    P = make()
    b = P.vector()[0].copy()
    x1, s1 = P.solve(b)
    x2, s2 = P.solve(b)
    assert np.array_equal(x1, x2) and s1["itn"] == s2["itn"]
    P.mupdat(*U)
    x3, _ = P.solve(b)
    Q = make()
    Q.mupdat(*U)
    assert np.array_equal(Q.matrix_packed(), P.matrix_packed())
    x4, _ = Q.solve(b)
    assert np.array_equal(x3, x4) and not np.array_equal(x3, x1)


def test_avprod_matches_dense():
    rp, ii, gg, cons, gam = _pixel_cons_problem(seed=24, ntr=2000)
    for matsto in (mpgpu.FULL, mpgpu.SPARSE):
        P = _pede(rp, ii, gg, gam, mpgpu.pixel_presigma(gam), cons, matsto=matsto, metsol=4, mreqenf=5, mcount=1,
                  withmultipliers=1)
        P.loopn()
        A = packed_to_dense(P.matrix_packed(), P.nall)
        x = np.random.default_rng(1).normal(size=P.nall)
        assert rel_err(P.avprod(x), A @ x) < 1e-12


def test_mupdat_and_errors_codes():
    rp, ii, gg, _ = mpgpu.toy_records(3, 500)
    P = _pede(rp, ii, gg, matsto=mpgpu.FULL, mreqenf=1)
    P.loopn()
    before = P.matrix_packed().copy()
    P.mupdat([1, 3], [1, 2], [2.5, -1.0])
    after = P.matrix_packed()
    d = after - before
    assert abs(d[0] - 2.5) < 1e-12 and abs(d[4] + 1.0) < 1e-12
    with pytest.raises(mpgpu.MpgpuError) as e:
        P.mupdat([P.nall + 5], [1], [1.0])
    assert e.value.code == 23


def _duplicate_first_global(rp, ii, gg, every=3):
    """Copy of the records where the first global word of the first measurement of every
    `every`-th record appears twice (the same label and derivative repeated in one equation)."""
    recs = LB.decode(rp, ii, gg)
    out_i, out_g, ptr = [], [], [0]
    for r, eqs in enumerate(recs):
        w0, w1 = int(rp[r]), int(rp[r + 1])
        I, G = list(ii[w0:w1]), list(gg[w0:w1])
        if r % every == 0 and eqs:
            sL, sG, lG = eqs[0]
            if lG > sG:
                k = sG + 1 - w0
                I.insert(lG + 1 - w0, I[k])
                G.insert(lG + 1 - w0, G[k])
        out_i += I
        out_g += G
        ptr.append(len(out_i))
    return np.array(ptr, np.int64), np.array(out_i, np.int32), np.array(out_g, np.float32)


@pytest.mark.parametrize("matsto", [mpgpu.FULL, mpgpu.SPARSE])
def test_label_maps_repeated_labels(orc, matsto):
    """Counting per equation (pede's default, mcount 0) with a label repeated inside an equation:
    pede keeps the repeat in the sorted equation list (src/pede.f90:2259-2270), which counts
    both entries and splits the parameter group at that label (pargrp, 2357-2386).  Index maps
    against the checker (itself pinned to pargrp's procedure), and the data pass against the
    oracle: matrix, vector, chi^2, ndf, rejects and counters."""
    rp, ii, gg, _ = mpgpu.pixel_records(31, 4, 5, 3, 3000, 3, 6)
    rp, ii, gg = _duplicate_first_global(rp, ii, gg)
    P = _pede(rp, ii, gg, matsto=matsto, mreqenf=1, mcount=0)
    lab, var = _check_maps(P, rp, ii, gg, mreqenf=1, mcount=0)
    gsl = P.labels()[3]
    assert len(set(gsl.tolist())) > len(set((lab // 10).tolist()))   # some module groups split
    _, mat, _ = _compare_pass(P, orc, rp, ii, gg, lab, var)
    # The diagonal element of a repeated label: pede adds w d_j d_k over the word pairs k <= j
    # of the equation (src/pede.f90:4196-4207), w (d1^2 + d1 d2 + d2^2) for two entries, where
    # the summed derivative gives w (d1 + d2)^2.  The difference is far above the tolerance, so
    # the comparison above tells the two apart (and catches a lost G contribution as well).
    pos = {int(l): i for i, l in enumerate(lab)}
    extra = np.zeros(P.nvar)
    for eqs in LB.decode(rp, ii, gg):
        for (sL, sG, lG) in eqs:
            seen = {}
            for w in range(sG + 1, lG + 1):
                seen.setdefault(int(ii[w]), []).append(float(gg[w]))
            for l, ds in seen.items():
                if var[pos[l]] > 0:
                    for j in range(len(ds)):
                        for k in range(j):
                            extra[var[pos[l]] - 1] += ds[j] * ds[k] / float(gg[sG]) ** 2
    assert np.abs(extra).max() > 1e-4 * np.abs(packed_to_dense(mat, P.nvar)).max()


@pytest.mark.parametrize("matsto", [mpgpu.FULL, mpgpu.SPARSE])
def test_repeated_labels_band_pass(orc, matsto):
    """Bordered-band local fits (sqmibb) over records with a label repeated inside an equation:
    the band kernel's G rows sum both entries (the word pairs of one element summed by one
    thread in the reference's order); the pass against the oracle."""
    rp, ii, gg = mpgpu.brokenlines_records(35, 24, 3, 3, 900, 8, 24)
    rp, ii, gg = _duplicate_first_global(rp, ii, gg, every=2)
    recs = LB.decode(rp, ii, gg)
    assert sum(_band_side(recs, ii, r) > 0 for r in range(len(recs))) > len(recs) // 2
    P = _pede(rp, ii, gg, matsto=matsto, mreqenf=1, mcount=0)
    lab, var = _check_maps(P, rp, ii, gg, mreqenf=1, mcount=0)
    _compare_pass(P, orc, rp, ii, gg, lab, var)


def test_scaleerrors(orc):
    """'scaleerrors f1 f2' multiplies the error of every measurement with global derivatives by
    f1 and of every other one by f2 in FP64 when the records are prepared (src/mppar.f90:312-316):
    the pass equals the oracle's pass over records whose error words were scaled the same way."""
    rp, ii, gg = mpgpu.brokenlines_records(33, 6, 4, 3, 800, 4, 6)
    f1, f2 = 1.7, 0.6
    P = _pede(rp, ii, gg, matsto=mpgpu.SPARSE, mreqenf=1, mcount=1, scale_errors=(f1, f2))
    lab, var = _check_maps(P, rp, ii, gg, mreqenf=1, mcount=1)
    gs = gg.astype(np.float64)
    for eqs in LB.decode(rp, ii, gg):
        for (sL, sG, lG) in eqs:
            gs[sG] *= f1 if lG > sG else f2
    st = P.loopn()
    mat, vec, cnt, so = orc.loopn(rp, ii, gs, lab, var, P.parameters(), P.nvar)
    assert st["nrej"] == so["nrej"] and st["ndf"] == so["ndf"]
    assert abs(st["chi2"] - so["chi2"]) <= 1e-11 * abs(so["chi2"])
    v, _ = P.vector()
    assert rel_err(v[:P.nvar], vec) < 1e-10
    A = packed_to_dense(P.matrix_packed(), P.nall)[:P.nvar, :P.nvar]
    assert rel_err(A, packed_to_dense(mat, P.nvar)) < 1e-10


def test_empty_and_redundant_constraints():
    """A constraint on parameters that are all fixed is empty: without 'skipemptycons' pede ends
    without a solution (exit 5, src/pede.f90:1177-1190), with it the constraint is dropped.  With
    'resolveredundancycons' a group of constraints with as many constraints as parameters fixes
    those parameters (index -4) and drops the constraints (src/pede.f90:1015-1045, 5965-5966)."""
    rp, ii, gg, _ = mpgpu.pixel_records(34, 4, 5, 3, 3000, 3, 6)
    fixed = [100000 + 10 + k for k in range(1, 7)]               # layer 1, module 1: fixed
    empty = [(0.0, 0.0, fixed[:2], [1.0, 1.0])]
    red = [(0.0, 0.0, [200011], [1.0]), (0.0, 0.0, [200012, 200013], [1.0, 1.0]),
           (0.0, 0.0, [200012, 200013], [1.0, -1.0])]           # 3 constraints on 3 parameters
    ok = [(0.0, 0.0, [300011, 300021], [1.0, 1.0])]

    def make(cons, **kw):
        P = mpgpu.Pede(matsto=mpgpu.FULL, metsol=mpgpu.INVERSION, mreqenf=1, mcount=1, withmultipliers=1, **kw)
        P.add_records(rp, ii, gg)
        P.set_steering_parameters(fixed, np.zeros(6), -np.ones(6))
        P.add_constraints(cons)
        return P
    with pytest.raises(mpgpu.MpgpuError) as e:
        make(empty + ok).build_structure()
    assert e.value.code == 5
    P = make(empty + ok, skip_empty_cons=1)
    P.build_structure()
    assert P.ncon == 1
    Q = make(red + ok, resolve_redundancy_cons=1)
    Q.build_structure()
    lab, var = Q.labels()[:2]
    pos = {int(l): i for i, l in enumerate(lab)}
    assert [var[pos[l]] for l in (200011, 200012, 200013)] == [-4, -4, -4]
    assert var[pos[200014]] > 0 and Q.ncon == 1
    R = make(red + ok)
    R.build_structure()
    assert R.ncon == 4


def test_redundant_constraint_groups_follow_grpcon():
    """pede's constraint grouping (grpcon, src/pede.f90:938-1060) is greedy in (first, last)
    parameter order: a constraint joins the group of its first related constraint with a group,
    sets the group of its parameters and re-marks its related constraints.  With the tix ranks
    p1 < p2 < p3 < p5 < p6 < p8 and d{p1,p5}, c{p2,p6}, c2{p2,p3}, c3{p5,p8}, c4{p6,p8} the
    procedure ends with groups {d, c, c3, c4} and {c2}, and parameter groups p1, p5, p6, p8 -> 1,
    p2, p3 -> 2 (p6 re-set by c4).  Group 1 counts 4 constraints and 4 parameters of its own in its
    range: with 'resolveredundancycons' p1, p5, p6, p8 become fixed (-4) and the four constraints
    are dropped; c2 stays.  (Connected components would give one group of 5 constraints on 6
    parameters: nothing resolved.)"""
    rp, ii, gg, _ = mpgpu.pixel_records(35, 4, 5, 3, 3000, 3, 6)
    P0 = _pede(rp, ii, gg, matsto=mpgpu.FULL, metsol=mpgpu.INVERSION, mreqenf=1, mcount=1)
    lab = P0.labels()[0]
    P0.close()
    k0 = len(lab) // 2
    p1, p2, p3, p5, p6, p8 = (int(lab[k0 + d]) for d in (0, 1, 2, 4, 5, 7))
    cons = [(0.0, 0.0, [p1, p5], [1.0, 1.0]),     # d
            (0.0, 0.0, [p2, p6], [1.0, -1.0]),    # c
            (0.0, 0.0, [p2, p3], [1.0, 1.0]),     # c2
            (0.0, 0.0, [p5, p8], [1.0, 2.0]),     # c3
            (0.0, 0.0, [p6, p8], [2.0, 1.0])]     # c4
    Q = mpgpu.Pede(matsto=mpgpu.FULL, metsol=mpgpu.INVERSION, mreqenf=1, mcount=1, withmultipliers=1,
                   resolve_redundancy_cons=1)
    Q.add_records(rp, ii, gg)
    Q.add_constraints(cons)
    Q.build_structure()
    lab, var = Q.labels()[:2]
    pos = {int(l): i for i, l in enumerate(lab)}
    assert [int(var[pos[p]]) for p in (p1, p5, p6, p8)] == [-4] * 4
    assert var[pos[p2]] > 0 and var[pos[p3]] > 0 and Q.ncon == 1
    Q.close()


@pytest.mark.parametrize("matsto", [mpgpu.FULL, mpgpu.SPARSE])
def test_cauchy_downweighting_closed_form(matsto):
    """The device's last outlier iteration with lhuber > 3 (Cauchy weights and chi^2,
    src/pede.f90:3988-3998, 4106-4131) against the closed form of tests/test_oracle_loop.py."""
    from test_oracle_loop import cauchy_expected, cauchy_records
    rp, ii, gg = cauchy_records()
    P = _pede(rp, ii, gg, matsto=matsto, mreqenf=1, lhuber=4)
    st = P.loopn(nloopn=2)
    ex = cauchy_expected(r=float(np.float32(5.03) - np.float32(5.0)), sigma=float(np.float32(0.01)))
    assert st["ndf"] == 5 and st["nrej"] == [0] * 6
    assert abs(st["chi2"] - ex["chi2"]) < 1e-9 * ex["chi2"]
    assert abs(P.matrix_packed()[0] - ex["mat"]) < 1e-9 * ex["mat"]
    assert abs(abs(P.vector()[0][0]) - ex["vec"]) < 1e-9 * ex["vec"]
