"""GPU: record ingest.  Records go to the device as they are added: arrays straight from the
caller's memory, Mille files (src/pede.f90:1714-2075, read through zlib) through two pinned
staging buffers (4 Mi words each) whose copies overlap the reading of the next chunk.  Every
route must give the same device records, so the structure and the pass agree exactly: one large
call against many small ones, a file read of several staging chunks against the in-memory records, FP32 and FP64 records mixed in one file (the values held so far are
widened when the first FP64 record arrives), and a truncated file whose records are dropped
while the earlier ones stay."""
import numpy as np
import pytest

import mpgpu

pytestmark = pytest.mark.gpu

CHUNK = 4 << 20


def _pass(P):
    P.build_structure()
    st = P.loopn()
    vec, cnt = P.vector()
    x = np.random.default_rng(5).normal(size=P.nall)
    return (st["chi2"], st["ndf"], st["nrej"], st["naccept"]), P.labels()[0], vec, cnt, P.avprod(x)


def _same(a, b):
    """pass sums, labels, counters and the matrix product exactly; the vector (FP64 atomics in
    the record kernel) to 1e-12"""
    assert a[0] == b[0]
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[3], b[3]) and np.array_equal(a[4], b[4])
    assert np.abs(a[2] - b[2]).max() <= 1e-12 * np.abs(a[2]).max()


def _opts():
    return dict(matsto=mpgpu.SPARSE, metsol=mpgpu.MINRES, mreqenf=5, mcount=1)


def _write_c_file(path, rec_ptr, inder, glder, fp64_recs):
    """C-format Mille file (src/mille.f90:208-219 layout: word count, values, indices); the
    records in fp64_recs carry FP64 values (negative count)."""
    with open(path, "wb") as f:
        for r in range(len(rec_ptr) - 1):
            w0, w1 = int(rec_ptr[r]), int(rec_ptr[r + 1])
            nw = w1 - w0
            if r in fp64_recs:
                np.array([-2 * nw], np.int32).tofile(f)
                np.asarray(glder[w0:w1], np.float64).tofile(f)
            else:
                np.array([2 * nw], np.int32).tofile(f)
                np.asarray(glder[w0:w1], np.float32).tofile(f)
            np.asarray(inder[w0:w1], np.int32).tofile(f)


def test_staged_ingest_equals_direct(tmp_path):
    rp, ii, gg, _ = mpgpu.pixel_records(61, 6, 8, 6, 60000, 8, 30)
    assert rp[-1] > 2 * CHUNK                      # three staging chunks at least
    P = mpgpu.Pede(**_opts())
    P.add_records(rp, ii, gg)                      # one call
    a = _pass(P)
    P.close()
    Q = mpgpu.Pede(**_opts())
    step = 2000
    for r0 in range(0, len(rp) - 1, step):         # many small calls (the device arrays grow)
        r1 = min(r0 + step, len(rp) - 1)
        w0, w1 = int(rp[r0]), int(rp[r1])
        Q.add_records(rp[r0:r1 + 1] - w0, ii[w0:w1], gg[w0:w1])
    b = _pass(Q)
    Q.close()
    _same(a, b)
    path = str(tmp_path / "big.bin")
    mpgpu.write_mille(path, rp, ii, gg)
    R = mpgpu.Pede(**_opts())
    assert R.read_mille(path) == len(rp) - 1       # file reader through the staging buffers
    _same(a, _pass(R))
    R.close()


def test_mixed_fp32_fp64_records(tmp_path):
    rp, ii, gg, _ = mpgpu.pixel_records(62, 4, 5, 3, 3000, 3, 7)
    nrec = len(rp) - 1
    ref = mpgpu.Pede(**_opts())
    ref.add_records(rp, ii, gg.astype(np.float64))
    a = _pass(ref)
    ref.close()
    # FP32 records, then FP64 ones, then FP32 again, in one file
    path = str(tmp_path / "mixed.bin")
    _write_c_file(path, rp, ii, gg, set(range(nrec // 3, 2 * nrec // 3)))
    P = mpgpu.Pede(**_opts())
    assert P.read_mille(path) == nrec
    _same(a, _pass(P))
    P.close()
    # FP32 then FP64 across calls of the direct route
    Q = mpgpu.Pede(**_opts())
    h = nrec // 2
    w = int(rp[h])
    Q.add_records(rp[:h + 1], ii[:w], gg[:w])
    Q.add_records(rp[h:] - w, ii[w:], gg[w:].astype(np.float64))
    _same(a, _pass(Q))
    Q.close()


def test_truncated_file_keeps_earlier_records(tmp_path):
    rp, ii, gg, _ = mpgpu.pixel_records(63, 4, 5, 3, 2000, 3, 7)
    good = str(tmp_path / "good.bin")
    mpgpu.write_mille(good, rp, ii, gg)
    data = open(good, "rb").read()
    bad = str(tmp_path / "bad.bin")
    with open(bad, "wb") as f:
        f.write(data[:len(data) // 2 + 3])         # cut inside a record
    ref = mpgpu.Pede(**_opts())
    ref.read_mille(good)
    a = _pass(ref)
    ref.close()
    P = mpgpu.Pede(**_opts())
    assert P.read_mille(good) == len(rp) - 1
    with pytest.raises(mpgpu.MpgpuError) as e:
        P.read_mille(bad)
    assert e.value.code == 18
    _same(a, _pass(P))
    P.close()
