"""Serving engine: concurrent jobs on one GPU give the same result as the job run alone, and
every job reports its latency."""
import numpy as np
import pytest

import mpgpu

pytestmark = pytest.mark.gpu


def _job(seed, nl, nx, ny, ntr):
    rp, ii, gg, _ = mpgpu.pixel_records(seed, nl, nx, ny, ntr, 4, 8, wx=100.0 / nx, wy=100.0 / ny)
    weak = mpgpu.pixel_weak_labels(nl, nx, ny)
    params = (np.array(weak), np.zeros(len(weak)), np.array(mpgpu.pixel_presigma(weak)))
    return (rp, ii, gg), params, mpgpu.pixel_constraints(nl, nx, ny)


def test_alone_vs_in_the_mix():
    rec, par, cons = _job(1, 6, 4, 3, 3000)
    opt = dict(matsto=mpgpu.FULL, metsol=mpgpu.DECOMPOSITION, mreqenf=5, mcount=1)
    S = mpgpu.Server(nworkers=4, budget_bytes=8 << 30)
    j = S.submit(rec, par, cons, **opt)
    rc, st, lab0, x0 = S.wait(j, capacity=100000)
    assert rc == 0 and st["nvar"] > 0 and st["latency_ms"] > 0
    others = [_job(10 + k, 5 + k % 3, 3 + k % 2, 3, 1500 + 200 * k) for k in range(6)]
    ids = []
    for k, (r2, p2, c2) in enumerate(others):
        o2 = dict(opt) if k % 2 == 0 else dict(matsto=mpgpu.SPARSE, metsol=mpgpu.MINRES, mbandw=6, mreqenf=5, mcount=1,
                                                mrestl=1e-10)
        ids.append(S.submit(r2, p2, c2, **o2))
        if k == 2:
            jm = S.submit(rec, par, cons, **opt)
    for i in ids:
        rc2, st2, _, _ = S.wait(i)
        assert rc2 == 0, st2
    rc, st, lab1, x1 = S.wait(jm, capacity=100000)
    assert rc == 0
    assert np.array_equal(lab0, lab1)
    # the two runs differ only in the order of the FP64 atomic additions of the matrix build
    # (1e-16 relative per element), amplified by the condition of the system (~1e8 here)
    assert np.max(np.abs(x1 - x0)) <= 1e-7 * np.max(np.abs(x0))
    S.close()


def test_group_block_job_bitwise_alone_vs_in_the_mix():
    """Group-block storage has no FP64 atomics on the path of a pixel job: the matrix and the
    vector are summed per block in processing order (the vector values ride in the records'
    updates), the product, the preconditioner and the MINRES reductions use fixed orders, and
    the product variants the autotuner picks between are bit-identical.  The same job alone and
    among concurrent jobs gives the bit-identical solution."""
    rec, par, cons = _job(3, 6, 4, 3, 3000)
    opt = dict(matsto=mpgpu.SPARSE, metsol=mpgpu.MINRES, mbandw=6, mreqenf=5, mcount=1, mrestl=1e-12)
    S = mpgpu.Server(nworkers=4, budget_bytes=8 << 30)
    rc, st, lab0, x0 = S.wait(S.submit(rec, par, cons, **opt), capacity=100000)
    assert rc == 0 and st["nvar"] > 0
    others = [_job(20 + k, 5 + k % 3, 3 + k % 2, 3, 1500 + 300 * k) for k in range(6)]
    ids = []
    for k, (r2, p2, c2) in enumerate(others):
        ids.append(S.submit(r2, p2, c2, **opt))
        if k == 2:
            jm = S.submit(rec, par, cons, **opt)
    for i in ids:
        rc2, st2, _, _ = S.wait(i)
        assert rc2 == 0, st2
    rc, st, lab1, x1 = S.wait(jm, capacity=100000)
    assert rc == 0
    assert np.array_equal(lab0, lab1)
    assert np.array_equal(x0, x1)
    S.close()


def test_budget_rejects_oversized_job():
    """A job larger than the whole budget ends with status 30 (the worker that takes it sizes
    it); the server keeps serving the next ones."""
    rec, par, cons = _job(2, 4, 3, 3, 500)
    S = mpgpu.Server(nworkers=1, budget_bytes=1 << 20)
    j = S.submit(rec, par, cons, matsto=mpgpu.FULL, metsol=mpgpu.DECOMPOSITION, mreqenf=5, mcount=1)
    rc, st, _, _ = S.wait(j)
    assert rc == 30 and st["status"] == 30
    S.close()
    S = mpgpu.Server(nworkers=2, budget_bytes=8 << 30)
    j = S.submit(rec, par, cons, matsto=mpgpu.FULL, metsol=mpgpu.DECOMPOSITION, mreqenf=5, mcount=1)
    rc, st, lab, x = S.wait(j, capacity=10000)
    assert rc == 0 and st["nvar"] == len(lab) > 0
    S.close()
