"""GPU: the index maps and the group-block structure built on the device, record decoding
(special data, FP64 payloads, malformed records).

The block structure is compared exactly with a direct restatement: the blocks are the group
pairs (a >= b) that the variables of some record couple, plus every diagonal block, in
row-major order of the lower block triangle, and the offsets are the running sum of the block
sizes (loop2's pair marking, src/pede.f90:6889-6908, in group-block form).  Record decoding
follows decodeNextMeasurement (src/pede.f90:2428-2468): special data (src/mille.f90:150-191)
is skipped, and a record that breaks the format ends the build with exit code 20."""
import numpy as np
import pytest

import labels as LB
import mpgpu
from helpers import packed_to_dense, rel_err

pytestmark = pytest.mark.gpu


def _build(rp, ii, gg, **opt):
    P = mpgpu.Pede(**opt)
    P.add_records(rp, ii, gg)
    P.build_structure()
    return P


@pytest.mark.parametrize("mcount", [0, 1])
def test_block_structure_matches_restatement(mcount):
    rp, ii, gg, _ = mpgpu.pixel_records(17, 5, 6, 4, 4000, 3, 7)
    P = _build(rp, ii, gg, matsto=mpgpu.SPARSE, metsol=mpgpu.MINRES, mreqenf=3, mcount=mcount)
    lab, var, cnt, gsl = P.labels()
    gfirst, brow, bcol, boff, nstored = P.block_structure()
    ng = len(gfirst) - 1
    assert ng > 1 and gfirst[0] == 1 and gfirst[-1] == P.nvar + 1
    pos = {int(l): t for t, l in enumerate(lab)}
    pairs = {(g, g) for g in range(ng)}
    for eqs in LB.decode(rp, ii, gg):
        vs = {int(var[pos[int(ii[w])]]) for (sL, sG, lG) in eqs for w in range(sG + 1, lG + 1)}
        gs = sorted({int(np.searchsorted(gfirst, v, side="right")) - 1 for v in vs if v > 0})
        pairs.update((a, b) for i, a in enumerate(gs) for b in gs[:i + 1])
    exp = sorted(pairs)
    size = np.diff(gfirst)
    exp_brow = np.searchsorted(np.array([a for a, _ in exp]), np.arange(ng + 1), side="left")
    exp_boff = np.concatenate([[0], np.cumsum([size[a] * size[b] for a, b in exp])])
    assert len(bcol) == len(exp)
    assert np.array_equal(bcol, [b for _, b in exp])
    assert np.array_equal(brow, exp_brow)
    assert np.array_equal(boff, exp_boff[:-1]) and nstored == exp_boff[-1]


def _with_special(rp, ii, gg, every=3, seed=5):
    """Insert a special-data block (0, 0.0), (0, -nsp), nsp words after the start word of every
    `every`-th record; the special integers include 0, negative and huge values."""
    rng = np.random.default_rng(seed)
    out_p, out_i, out_g = [0], [], []
    for r in range(len(rp) - 1):
        w0, w1 = int(rp[r]), int(rp[r + 1])
        ri, rg = list(ii[w0:w1]), list(gg[w0:w1])
        if r % every == 0:
            nsp = int(rng.integers(1, 6))
            sp_i = [int(x) for x in rng.choice([0, -7, 3, 2 ** 30, 123456789], nsp)]
            sp_g = list(rng.normal(size=nsp))
            ri = ri[:1] + [0, 0] + sp_i + ri[1:]
            rg = rg[:1] + [0.0, -float(nsp)] + sp_g + rg[1:]
        out_i += ri
        out_g += rg
        out_p.append(len(out_i))
    return np.array(out_p, np.int64), np.array(out_i, np.int32), np.array(out_g, gg.dtype)


@pytest.mark.parametrize("matsto", [mpgpu.FULL, mpgpu.SPARSE])
def test_special_data_words_are_skipped(matsto):
    rp, ii, gg, _ = mpgpu.pixel_records(23, 4, 5, 3, 2000, 3, 6)
    sp, si, sg = _with_special(rp, ii, gg)
    assert len(si) > len(ii)
    opt = dict(matsto=matsto, metsol=mpgpu.MINRES, mreqenf=2, mcount=1)
    P, Q = _build(rp, ii, gg, **opt), _build(sp, si, sg, **opt)
    for a, b in zip(P.labels(), Q.labels()):
        assert np.array_equal(a, b)
    s1, s2 = P.loopn(), Q.loopn()
    assert s1["chi2"] == s2["chi2"] and s1["ndf"] == s2["ndf"] and s1["nrej"] == s2["nrej"]
    assert rel_err(P.vector()[0], Q.vector()[0]) < 1e-12
    assert rel_err(P.matrix_packed(), Q.matrix_packed()) < 1e-12


@pytest.mark.parametrize("matsto", [mpgpu.FULL, mpgpu.SPARSE])
def test_fp64_payload_matches_fp32(orc, matsto):
    """The same records as FP64 payload (the values are FP32 numbers, so both payloads carry the
    same values): index maps equal, and the pass equals the FP32 pass and the oracle's."""
    rp, ii, gg, _ = mpgpu.pixel_records(29, 4, 5, 3, 2000, 3, 6)
    opt = dict(matsto=matsto, metsol=mpgpu.MINRES, mreqenf=2, mcount=1)
    P, Q = _build(rp, ii, gg, **opt), _build(rp, ii, gg.astype(np.float64), **opt)
    for a, b in zip(P.labels(), Q.labels()):
        assert np.array_equal(a, b)
    s1, s2 = P.loopn(), Q.loopn()
    assert s1["ndf"] == s2["ndf"] and s1["nrej"] == s2["nrej"]
    assert abs(s1["chi2"] - s2["chi2"]) <= 1e-12 * abs(s1["chi2"])
    n = P.nall
    assert rel_err(packed_to_dense(P.matrix_packed(), n), packed_to_dense(Q.matrix_packed(), n)) < 1e-12
    lab, var = Q.labels()[:2]
    mat, vec, cnt, so = orc.loopn(rp, ii, gg.astype(np.float64), lab, var, Q.parameters(), Q.nvar)
    assert abs(s2["chi2"] - so["chi2"]) <= 1e-11 * abs(so["chi2"])
    assert rel_err(packed_to_dense(Q.matrix_packed(), n)[:Q.nvar, :Q.nvar], packed_to_dense(mat, Q.nvar)) < 1e-10


def test_malformed_records_report_exit_code_20():
    rp, ii, gg, _ = mpgpu.pixel_records(3, 3, 4, 3, 300, 3, 5)
    opt = dict(matsto=mpgpu.SPARSE, metsol=mpgpu.MINRES)
    # record 41 (1-based) cut after its first measurement word: no error word follows
    r = 40
    w0, w1 = int(rp[r]), int(rp[r + 1])
    keep = np.r_[0:w0 + 2, w1:len(ii)]
    rp3 = rp.copy()
    rp3[r + 1:] -= w1 - (w0 + 2)
    with pytest.raises(mpgpu.MpgpuError) as e:
        _build(rp3, ii[keep], gg[keep], **opt)
    assert e.value.code == 20 and "record 41" in str(e.value)
    # record 8: a local index below 1
    sL, sG, lG = LB.decode(rp, ii, gg)[7][0]
    assert sG > sL + 1
    ii4 = ii.copy()
    ii4[sL + 1] = -1
    with pytest.raises(mpgpu.MpgpuError) as e:
        _build(rp, ii4, gg, **opt)
    assert e.value.code == 20 and "record 8" in str(e.value)
    # a global label <= 0
    ii5 = ii.copy()
    ii5[lG] = -3
    with pytest.raises(mpgpu.MpgpuError) as e:
        _build(rp, ii5, gg, **opt)
    assert e.value.code == 20


def test_block_product_variants_identical(monkeypatch):
    """The autotuner chooses among product variants with the same arithmetic: every variant
    gives the bit-identical product (so the timing-dependent choice never changes a result),
    equal to the dense product of the stored matrix."""
    rp, ii, gg, _ = mpgpu.pixel_records(41, 4, 6, 3, 4000, 3, 7)
    x = np.random.default_rng(3).normal(size=1)
    ys = {}
    for v in (2, 4, 8, 16 + 2, 16 + 4, 16 + 8, 32 + 2, 32 + 4, 32 + 8, -1):
        if v >= 0:
            monkeypatch.setenv("MPGPU_SPMV_VARIANT", str(v))
        else:
            monkeypatch.delenv("MPGPU_SPMV_VARIANT", raising=False)
        P = _build(rp, ii, gg, matsto=mpgpu.SPARSE, metsol=mpgpu.MINRES, mreqenf=2, mcount=1)
        P.loopn()
        if x.size != P.nall:
            x = np.random.default_rng(3).normal(size=P.nall)
        ys[v] = P.avprod(x)
        tun = P.tuning()
        assert ("fixed variant" in tun) if v >= 0 else ("chosen" in tun or "default variant" in tun)
        if v == -1:
            n = P.nall
            A = packed_to_dense(P.matrix_packed(), n)
            assert rel_err(ys[v], A @ x) < 1e-12
        P.close()
    for v, y in ys.items():
        assert np.array_equal(y, ys[-1]), v


@pytest.mark.parametrize("mcount", [0, 1])
def test_record_summary_matches_restatement(mcount):
    """The loop1 / loop2 counts of millepede.log (equations, equations with globals, global
    words, most equations / largest local index / most variable parameters per record, records
    without variable parameters, groups over all and over variable parameters) against the
    checker's restatement, with special data and a zero derivative in the records."""
    rp, ii, gg, _ = mpgpu.pixel_records(31, 4, 5, 3, 3000, 3, 7)
    sp, si, sg = _with_special(rp, ii, gg)
    sg = sg.copy()
    eq0 = LB.decode(sp, si, sg)[4][0]
    sg[eq0[2]] = 0.0                          # a zero global derivative (still a global word)
    P = _build(sp, si, sg, matsto=mpgpu.SPARSE, metsol=mpgpu.MINRES, mreqenf=5, mcount=mcount)
    lab, var, cnt, gsl = P.labels()
    got = P.record_summary()
    exp = LB.record_summary(sp, si, sg, lab, var)
    for k, v in exp.items():
        assert got[k] == v, (k, got[k], v)
    assert got["nrec"] == len(sp) - 1 and got["ntot"] == len(lab) and got["nvar"] == P.nvar
    assert got["fp64_payload"] == 0
    assert got["ngroup_tot"] == len(set(gsl[cnt > 0].tolist()))
    assert got["ngroup_var"] == len(set(gsl[var > 0].tolist())) == len(P.block_structure()[0]) - 1
    P.close()


def test_iterated_entries_cut():
    """'entries n1 n2 n3' with counting per equation (loop1i, src/pede.f90:6164-6265): the case
    whose outcome is fixed by construction (tests/test_oracle_labels.py), then a pixel detector
    with a cut that fixes modules: every flag equals the checker's restatement."""
    import test_oracle_labels as TL
    rp, ii, gg = TL._eq_records(TL.ITEREN_EQS)
    P = _build(rp, ii, gg.astype(np.float32), matsto=mpgpu.FULL, metsol=mpgpu.INVERSION, mreqenf=2, mcount=0, iteren=3)
    lab, var = P.labels()[:2]
    assert {int(l): (1 if v > 0 else int(v)) for l, v in zip(lab, var)} == TL.ITEREN_EXPECT
    P.close()
    rp, ii, gg, _ = mpgpu.pixel_records(37, 4, 6, 4, 1500, 3, 6)
    for mreqenf, iteren in ((40, 80), (60, 1000)):
        P = _build(rp, ii, gg, matsto=mpgpu.SPARSE, metsol=mpgpu.MINRES, mreqenf=mreqenf, mcount=0, iteren=iteren)
        lab, var, cnt, gsl = P.labels()
        elab, evar, ecnt, egsl = LB.label_maps(rp, ii, gg, mreqenf=mreqenf, mcount=0, iteren=iteren)
        assert np.array_equal(lab, elab) and np.array_equal(cnt, ecnt)
        assert np.array_equal(var, evar), np.nonzero(var != evar)
        assert (var == -3).any() and (var == -2).any() and (var > 0).any(), (mreqenf, iteren)
        P.close()
