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


def test_record_update_closed_form(orc):
    """One record at a time: matrix contribution = C1 - G Gamma^-1 G^T explicitly."""
    rp, ii, gg, _ = toy(ntracks=5, seed=11)
    labels, var, cnt, gsl = LB.label_maps(rp, ii, gg, mreqenf=0)
    nvar = int(var.max())
    pos = {int(l): k for k, l in enumerate(labels)}
    recs = LB.decode(rp, ii, gg)
    for r in range(5):
        sub_rp = np.array([0, rp[r + 1] - rp[r]], np.int64)
        sl = slice(int(rp[r]), int(rp[r + 1]))
        mat, vec, _, st = orc.loopn(sub_rp, ii[sl], gg[sl], labels, var, np.zeros(len(labels)), nvar)
        eqs = recs[r]
        nalc = max(int(ii[q]) for (sL, sG, lG) in eqs for q in range(sL + 1, sG))
        Lm = np.zeros((len(eqs), nalc))
        Gm = np.zeros((len(eqs), nvar))
        y = np.zeros(len(eqs))
        W = np.zeros(len(eqs))
        for e, (sL, sG, lG) in enumerate(eqs):
            for q in range(sL + 1, sG):
                Lm[e, int(ii[q]) - 1] = gg[q]
            for q in range(sG + 1, lG + 1):
                Gm[e, var[pos[int(ii[q])]] - 1] += gg[q]
            y[e] = gg[sL]
            W[e] = 1.0 / gg[sG] ** 2
        Gam = Lm.T @ (W[:, None] * Lm)
        C1 = Gm.T @ (W[:, None] * Gm)
        GL = Gm.T @ (W[:, None] * Lm)
        expect = C1 - GL @ np.linalg.solve(Gam, GL.T)
        q = np.linalg.solve(Gam, Lm.T @ (W * y))
        res = y - Lm @ q
        assert np.allclose(packed_to_dense(mat, nvar), expect, rtol=1e-9, atol=1e-6)
        assert np.allclose(vec, Gm.T @ (W * res), rtol=1e-9, atol=1e-9)
        assert abs(st["chi2"] - np.sum(W * res * res)) < 1e-9 * max(1.0, st["chi2"])
        assert st["ndf"] == len(eqs) - nalc


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
