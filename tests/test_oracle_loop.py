"""Pin the oracle's record loop (loopbf restatement) and label maps on the CPU.

* Schur identity (doc/draftManual.md:291-330): solving the reduced global system built by
  the record loop must equal the global part of the joint least-squares solution over
  globals and every record's local parameters.
* Per record, the Millepede update must equal G^T W G - G^T W L (L^T W L)^-1 L^T W G.
* Label maps: sorted labels, variable indices, entry counts, group partition invariants.
"""
import os
import tempfile

import numpy as np
import pytest

import labels as LB
import mpgpu
from helpers import packed_to_dense


def toy(ntracks=300, seed=3):
    rp, ii, gg, truth = mpgpu.toy_records(seed, ntracks)
    return rp, ii, gg.astype(np.float64), truth


def joint_lsq(rp, ii, gg, var, labels, nvar):
    """Explicit joint least squares over variable globals + all locals of every record."""
    recs = LB.decode(rp, ii, gg)
    pos = {int(l): k for k, l in enumerate(labels)}
    rows, cols, vals, y, w = [], [], [], [], []
    nloc_off = nvar
    r = 0
    for eqs in recs:
        nalc = 0
        for (sL, sG, lG) in eqs:
            for q in range(sL + 1, sG):
                nalc = max(nalc, int(ii[q]))
        if len(eqs) <= nalc:      # ndf <= 0: pede rejects the record (src/pede.f90:4080-4089)
            continue
        for (sL, sG, lG) in eqs:
            for q in range(sL + 1, sG):
                rows.append(r); cols.append(nloc_off + int(ii[q]) - 1); vals.append(gg[q])
            for q in range(sG + 1, lG + 1):
                v = var[pos[int(ii[q])]]
                if v > 0:
                    rows.append(r); cols.append(v - 1); vals.append(gg[q])
            y.append(gg[sL]); w.append(1.0 / gg[sG] ** 2)
            r += 1
        nloc_off += nalc
    X = np.zeros((r, nloc_off))
    np.add.at(X, (np.array(rows), np.array(cols)), np.array(vals))
    sw = np.sqrt(np.array(w))
    sol, *_ = np.linalg.lstsq(X * sw[:, None], np.array(y) * sw, rcond=None)
    return sol[:nvar]


def test_schur_identity_toy(orc):
    rp, ii, gg, truth = toy()
    # planes 1 and 10 fixed by pre-sigma (-1), as the toy configuration steers
    fixed = [100 * p + 10 * m + k for p in (1, 10) for m in range(5) for k in (1, 2)]
    labels, var, cnt, gsl = LB.label_maps(rp, ii, gg, steer_labels=fixed, presigma=[-1.0] * len(fixed), mreqenf=1)
    nvar = int(var.max())
    par = np.zeros(len(labels))
    mat, vec, c, st = orc.loopn(rp, ii, gg, labels, var, par, nvar)
    assert st["nrej"][:3] == [0, 0, 0] and st["nrej"][4:] == [0, 0] and st["nband"] == 0
    A = packed_to_dense(mat, nvar)
    dp = np.linalg.solve(A, vec)
    ref = joint_lsq(rp, ii, gg, var, labels, nvar)
    assert np.allclose(dp, ref, rtol=1e-8, atol=1e-12 * np.abs(ref).max())
    # the fit recovers the simulated misalignments of the free planes to within errors
    Ai = np.linalg.inv(A)
    t = np.array([truth[k] for k, lab in enumerate(mpgpu.toy_labels()) if var[np.searchsorted(labels, lab)] > 0])
    pulls = (dp - t) / np.sqrt(np.diag(Ai))
    assert np.abs(pulls).max() < 6.0


def closed_form_record(ii, gg, eqs, var, pos, nvar):
    """Dense restatement of one record's contribution: C = C1 - G Gamma^-1 G^T,
    b = G W (y - L q), chi2 = sum W (y - L q)^2 with q = Gamma^-1 L^T W y."""
    nalc = max(int(ii[q]) for (sL, sG, lG) in eqs for q in range(sL + 1, sG))
    Lm = np.zeros((len(eqs), nalc))
    Gm = np.zeros((len(eqs), nvar))
    y = np.zeros(len(eqs))
    W = np.zeros(len(eqs))
    for e, (sL, sG, lG) in enumerate(eqs):
        for q in range(sL + 1, sG):
            Lm[e, int(ii[q]) - 1] = gg[q]
        for q in range(sG + 1, lG + 1):
            v = var[pos[int(ii[q])]]
            if v > 0:
                Gm[e, v - 1] += gg[q]
        y[e] = gg[sL]
        W[e] = 1.0 / gg[sG] ** 2
    Gam = Lm.T @ (W[:, None] * Lm)
    C1 = Gm.T @ (W[:, None] * Gm)
    GL = Gm.T @ (W[:, None] * Lm)
    q = np.linalg.solve(Gam, Lm.T @ (W * y))
    res = y - Lm @ q
    return C1 - GL @ np.linalg.solve(Gam, GL.T), Gm.T @ (W * res), float(np.sum(W * res * res)), len(eqs) - nalc


def check_records_closed_form(orc, rp, ii, gg, labels, var, nvar, records, **kw):
    pos = {int(l): k for k, l in enumerate(labels)}
    recs = LB.decode(rp, ii, gg)
    nband = 0
    for r in records:
        sub_rp = np.array([0, rp[r + 1] - rp[r]], np.int64)
        sl = slice(int(rp[r]), int(rp[r + 1]))
        mat, vec, _, st = orc.loopn(sub_rp, ii[sl], gg[sl], labels, var, np.zeros(len(labels)), nvar, **kw)
        expect, b, chi2, ndf = closed_form_record(ii, gg, recs[r], var, pos, nvar)
        A = packed_to_dense(mat, nvar)
        assert np.allclose(A, expect, rtol=1e-9, atol=1e-9 * np.abs(expect).max())
        assert np.allclose(vec, b, rtol=1e-9, atol=1e-9 * max(1e-300, np.abs(b).max()))
        assert abs(st["chi2"] - chi2) < 1e-9 * max(1.0, chi2)
        assert st["ndf"] == ndf
        nband += st["nband"]
    return nband


def test_record_update_closed_form(orc):
    """One record at a time: matrix contribution = C1 - G Gamma^-1 G^T explicitly."""
    rp, ii, gg, _ = toy(ntracks=5, seed=11)
    labels, var, cnt, gsl = LB.label_maps(rp, ii, gg, mreqenf=0)
    nvar = int(var.max())
    check_records_closed_form(orc, rp, ii, gg, labels, var, nvar, range(5))


def test_chi2_cut_and_downweighting_paths(orc):
    """Second pass with Huber down-weighting and a chi^2 cut runs and rejects consistently."""
    rp, ii, gg, _ = toy(ntracks=400, seed=5)
    gg = gg.copy()
    # inject outliers: shift a few measurements by 20 sigma
    recs = LB.decode(rp, ii, gg)
    for r in range(0, 400, 37):
        sL = recs[r][3][0]
        gg[sL] += 0.3
    labels, var, cnt, gsl = LB.label_maps(rp, ii, gg, mreqenf=1)
    nvar = int(var.max())
    par = np.zeros(len(labels))
    _, _, _, s1 = orc.loopn(rp, ii, gg, labels, var, par, nvar, icalcm=0, nloopn=1, chicut=3.0)
    _, _, _, s2 = orc.loopn(rp, ii, gg, labels, var, par, nvar, icalcm=0, nloopn=2, lhuber=3, chicut=3.0)
    assert s1["nrej"][5] > 0
    assert s2["chi2"] < s1["chi2"] + 1e-9


def test_label_maps_invariants():
    rp, ii, gg, _ = mpgpu.pixel_records(7, 3, 4, 3, 200, 3, 5)
    for mcount in (0, 1):
        labels, var, cnt, gsl = LB.label_maps(rp, ii, gg.astype(float), mreqenf=5, mcount=mcount)
        assert np.all(np.diff(labels) > 0)
        v = var[var > 0]
        assert np.array_equal(v, np.arange(1, len(v) + 1))
        assert np.all(var[cnt < 5] == -2)
        # each equation's (or record's) label set is a union of whole groups
        recs = LB.decode(rp, ii, gg)
        pos = {int(l): k for k, l in enumerate(labels)}
        grp = np.searchsorted(np.unique(gsl), gsl)
        for eqs in recs[:50]:
            sets = [[pos[int(ii[q])] for q in range(sG + 1, lG + 1)] for (sL, sG, lG) in eqs]
            if mcount:
                sets = [sorted(set(sum(sets, [])))]
            for s in sets:
                for g in set(grp[s]):
                    members = set(np.nonzero(grp == g)[0])
                    assert members <= set(s)
    # pixel modules: with countrecords the 6 globals of a module form one group
    labels, var, cnt, gsl = LB.label_maps(rp, ii, gg.astype(float), mreqenf=0, mcount=1)
    _, sizes = np.unique(gsl, return_counts=True)
    assert np.all(sizes % 6 == 0)


@pytest.mark.parametrize("fortran", [False, True])
def test_mille_file_roundtrip(orc, fortran):
    rp, ii, gg, _ = mpgpu.toy_records(1, 50)
    with tempfile.TemporaryDirectory() as d:
        f = os.path.join(d, "t.bin")
        mpgpu.write_mille(f, rp, ii, gg, fortran=fortran)
        rp2, ii2, gg2 = orc.read_mille(f, fortran=fortran)
        assert np.array_equal(rp2, rp) and np.array_equal(ii2, ii)
        assert np.array_equal(gg2, gg.astype(np.float64))
        # C-binary layout (python/readMilleBinary.py:186-213): int32 nr, nw floats, nw ints
        raw = np.fromfile(f, dtype=np.int32)
        off = 1 if fortran else 0
        assert raw[off] == 2 * (rp[1] - rp[0])


def test_record_update_closed_form_band_path(orc):
    """Records whose local fits take the bordered-band path (sqmibb): each record's matrix,
    vector, chi2 and ndf equal the dense closed form, and every record used the band path.
    (The joint problem of these records is rank deficient -- per-plane track offsets absorb
    module shifts -- so the check is per record.)"""
    rp, ii, gg = mpgpu.brokenlines_records(4, 12, 3, 2, 12, 8, 12)
    gg = gg.astype(np.float64)
    labels, var, cnt, gsl = LB.label_maps(rp, ii, gg, mreqenf=0)
    nvar = int(var.max())
    nband = check_records_closed_form(orc, rp, ii, gg, labels, var, nvar, range(12))
    assert nband == 12


def test_band_and_full_local_fit_agree(orc):
    """The same records with the band check switched off (localfit ... 0, full sqminv path)
    give the same matrix and vector to rounding."""
    rp, ii, gg = mpgpu.brokenlines_records(5, 10, 2, 2, 60, 6, 10)
    gg = gg.astype(np.float64)
    labels, var, cnt, gsl = LB.label_maps(rp, ii, gg, mreqenf=1, mcount=1)
    nvar = int(var.max())
    m1, v1, c1, s1 = orc.loopn(rp, ii, gg, labels, var, np.zeros(len(labels)), nvar, lfitbb=1)
    m2, v2, c2, s2 = orc.loopn(rp, ii, gg, labels, var, np.zeros(len(labels)), nvar, lfitbb=0)
    assert s1["nband"] > 0 and s2["nband"] == 0
    assert np.allclose(m1, m2, rtol=1e-7, atol=1e-7 * np.abs(m2).max())
    assert np.allclose(v1, v2, rtol=1e-7, atol=1e-7 * np.abs(v2).max())
    assert abs(s1["chi2"] - s2["chi2"]) < 1e-8 * s2["chi2"]


def test_loopn_blocks_equals_packed(orc):
    """The multi-threaded pass into a group-block layout (the CPU baseline of the timing) adds
    the same updates as the single-threaded packed pass: pixel records, groups of the 6
    parameters of each module, a block for every pair of groups."""
    rp, ii, gg, _ = mpgpu.pixel_records(5, 4, 3, 2, 1500, 3, 6)
    lab, var, cnt, gsl = LB.label_maps(rp, ii, gg.astype(float), mreqenf=1, mcount=1)
    nv = int(var.max())
    par = np.zeros(len(lab))
    mat, vec, c, so = orc.loopn(rp, ii, gg.astype(np.float64), lab, var, par, nv)
    ng = nv // 6
    gfirst = np.arange(ng + 1, dtype=np.int32) * 6 + 1
    brow = np.zeros(ng + 1, np.int64)
    bcol, boff, off = [], [], 0
    for a in range(ng):
        for b in range(a + 1):
            bcol.append(b)
            boff.append(off)
            off += 36
        brow[a + 1] = len(bcol)
    blk, v2, c2, s2 = orc.loopn_blocks(rp, ii, gg.astype(np.float64), lab, var, par, nv,
                                       (gfirst, brow, np.array(bcol), np.array(boff), off), nthreads=4)
    assert s2["nrej"] == so["nrej"] and s2["ndf"] == so["ndf"] and abs(s2["chi2"] - so["chi2"]) < 1e-9 * so["chi2"]
    assert np.array_equal(c2, c) and np.allclose(v2, vec, rtol=1e-12, atol=0)
    A = packed_to_dense(mat, nv)
    B = np.zeros((nv, nv))
    k = 0
    for a in range(ng):
        for b in range(a + 1):
            B[6 * a:6 * a + 6, 6 * b:6 * b + 6] = blk[boff[k]:boff[k] + 36].reshape(6, 6)
            k += 1
    L = np.tril(B)
    assert np.abs(np.tril(A) - L).max() <= 1e-12 * np.abs(A).max()


def cauchy_records(nrec=5, M=5.0, r=0.03, sigma=0.01, label=11):
    """Records with one local parameter (an offset) and two measurements M +- r of error sigma;
    the first measurement also depends on one global parameter (derivative 1).  By symmetry
    the local fit leaves residuals +-r whatever the weights."""
    I = [0, 0, 1, 0, label, 0, 1, 0]
    G = [0.0, M + r, 1.0, sigma, 1.0, M - r, 1.0, sigma]
    rp = np.arange(nrec + 1, dtype=np.int64) * len(I)
    return rp, np.tile(np.array(I, np.int32), nrec), np.tile(np.array(G, np.float32), nrec)


def cauchy_expected(nrec=5, r=0.03, sigma=0.01):
    """Closed form of the last (Cauchy) outlier iteration (lhuber > 3, src/pede.f90:3988-3998,
    4106-4131): each measurement gets weight (1/sigma^2) / (1 + (r/sigma/c)^2) with c = 2.3849
    and chi^2 contribution c^2 log(1 + (r/sigma/c)^2); the matrix element of the global is
    w - w^2 / (2 w) = w / 2 per record, its vector element w r (Schur complement of the offset)."""
    c = 2.3849
    t = (r / sigma / c) ** 2
    w = (1.0 / sigma ** 2) / (1.0 + t)
    # chi^2 sum in float rounding of pede's REAL(4) r/sigma is not involved: all in FP64
    return dict(chi2=nrec * 2 * c * c * np.log1p(t), mat=nrec * w / 2, vec=nrec * w * r)


@pytest.mark.parametrize("lhuber", [4, 6])
def test_cauchy_downweighting_closed_form(orc, lhuber):
    rp, ii, gg = cauchy_records()
    lab, var, cnt, gsl = LB.label_maps(rp, ii, gg.astype(float), mreqenf=1)
    mat, vec, c, so = orc.loopn(rp, ii, gg.astype(np.float64), lab, var, np.zeros(1), 1, nloopn=2, lhuber=lhuber)
    # the records are FP32: r and sigma are the FP32 values of the file
    r32 = float(np.float32(5.03) - np.float32(5.0))
    ex = cauchy_expected(r=abs(r32), sigma=float(np.float32(0.01)))
    assert so["ndf"] == 5 and so["nrej"] == [0] * 6
    assert abs(so["chi2"] - ex["chi2"]) < 1e-9 * ex["chi2"]
    assert abs(mat[0] - ex["mat"]) < 1e-9 * ex["mat"]
    assert abs(abs(vec[0]) - ex["vec"]) < 1e-9 * ex["vec"]
    # with lhuber = 3 the last iteration is a Huber one: a different chi^2
    _, _, _, sh = orc.loopn(rp, ii, gg.astype(np.float64), lab, var, np.zeros(1), 1, nloopn=2, lhuber=3)
    assert abs(sh["chi2"] - ex["chi2"]) > 1e-3 * ex["chi2"]
