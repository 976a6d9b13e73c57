"""End-to-end drop-in check against the reference's CI golden files.

The CI job of the reference (.gitlab-ci.yml:86-100) generates 100,000 wire-chamber records
with pede's own generator (src/mptest1.f90, built from the reference sources into
oracle/_ref/mptest1_gen), solves with CI/mp2str.txt and compares millepede.res with
CI/refFiles/millepede.res (kept here as tests/golden/ci_wirechamber_millepede.res).  The CI
steering ends with 'method fullLAPACK 3 0.8' (the Cholesky factorization of mplapack.f90);
test_ci_wirechamber_res uses 'method inversion' with the same iteration count and convergence
limit and Lagrange multipliers instead of elimination for the two constraints (the
constrained least-squares solution is the same), test_ci_wirechamber_fulllapack the CI's own
method.
Parameters are compared at the 5 significant digits the file holds; entry counts exactly.  The
residual monitoring of the CI steering ('monitorresiduals': normalized residuals of the first
local-fit iteration per measurement group, after the first and the last pass) must reproduce
CI/refFiles/millepede.mon (kept as tests/golden/ci_wirechamber_millepede.mon) line for line.
"""
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

import mpgpu

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GEN = os.path.join(ROOT, "oracle", "_ref", "mptest1_gen")
GOLD = os.path.join(ROOT, "tests", "golden", "ci_wirechamber_millepede.res")
GOLD_MON = os.path.join(ROOT, "tests", "golden", "ci_wirechamber_millepede.mon")
GOLD_END = os.path.join(ROOT, "tests", "golden", "ci_wirechamber_millepede.end")
GOLD_LOG = os.path.join(ROOT, "tests", "golden", "ci_wirechamber_millepede_log.txt")

STEER = """fortranfiles
mp2con.txt
mp2tst.bin
Cfiles
method inversion 3 0.8
entries 50 10 2
outlierdownweighting 3
dwfractioncut 0.1
matiter 1
printcounts
countrecords
monitorresiduals
end
"""


def parse_res(path):
    out = {}
    for line in open(path).read().splitlines()[1:]:
        t = line.split()
        if not t:
            continue
        out[int(t[0])] = ([float(x) for x in t[1:-1]], int(t[-1]))
    return out


def _ulp(tok):
    """one unit in the last printed digit of a Fortran E or F field"""
    t = tok.lstrip("-")
    if "E" in t:
        mant, ex = t.split("E")
        return 10.0 ** (int(ex) - len(mant.split(".")[1]))
    return 10.0 ** -len(t.split(".")[1]) if "." in t else 0.0


def compare_log(log, gold):
    """The iteration table, line-search tables and chi^2 summary of millepede.log against the
    CI golden (from the table header to the down-weighting line): the same lines and text;
    integers equal; printed reals within 1.5 units of their last digit, plus 1e-9 relative for
    the 17-digit chi^2 sums (summed in another order).  In a line-search table (ptlprt) F(x),
    a difference of two chi^2 sums of 8e6, also within 1e-8; the slopes F'(X) within 1e-6
    relative, and within 1e-6 of the first slope where they are that small (the zero crossing,
    whose digits are rounding noise).  Columns 70-76 (hh mm ss) are skipped."""
    def section(lines):
        i = next(k for k, l in enumerate(lines) if l.startswith(" it fc   fcn_value"))
        j = next(k for k, l in enumerate(lines) if "with correction for down-weighting" in l)
        return lines[i:j + 1]
    a, b = section(log), section(gold)
    assert len(a) == len(b), (len(a), len(b))
    scale = 0.0
    for la, lb in zip(a, b):
        if lb.rstrip().endswith(("FMS", " F", "   S", " end")) and len(lb) >= 76:
            la, lb = la[:69] + la[76:], lb[:69] + lb[76:]
        ta, tb = la.split(), lb.split()
        assert len(ta) == len(tb), (la, lb)
        table = len(tb) >= 6 and tb[0].isdigit() and re.fullmatch(r"-?\d+\.\d{6}", tb[1]) is not None
        col = {}
        if table:
            first = tb[-1] == "ratio"
            f_at, s_at = (2, 3) if first else (3, 4)
            if first:
                scale = abs(float(tb[s_at]))
            col = {f_at: "F", s_at: "slope"}
        for k, (x, y) in enumerate(zip(ta, tb)):
            try:
                fy = float(y)
            except ValueError:
                assert x == y, (la, lb)
                continue
            if "." not in y and "E" not in y:
                assert x == y, (la, lb)
                continue
            tol = max(1.5 * _ulp(y), 1e-9 * abs(fy))
            if col.get(k) == "F":
                tol += 1e-8
            elif col.get(k) == "slope":
                tol += 1e-6 * abs(fy) + (1e-6 * scale if abs(fy) < 1e-5 * scale else 0.0)
            elif table:
                tol += 1e-6
            assert abs(float(x) - fy) <= tol, (x, y, la, lb)


def compare_log_sizes(log, gold):
    """The structure and setup sections of millepede.log against the CI golden, line for line
    (trailing blanks aside): loop1's counts of labels, parameters, groups, records, equations
    and global derivatives; loop2's size table, solution method and storage, the rank of the
    constraints, the chi^2 cut table; the outlier and solution-algorithm tables; the feasibility
    check and the rank of the first factorization -- everything from 'LOOP1: starting' up to the
    iteration table except pede's read cache (NCACHE, cache splitting), which this path does not
    have.  Then the loop count after the chi^2 summary."""
    def section(lines):
        i = next(k for k, l in enumerate(lines) if l.startswith(" LOOP1: starting"))
        j = next(k for k, l in enumerate(lines) if l.startswith(" it fc   fcn_value"))
        return [l.rstrip() for l in lines[i:j] if not l.startswith("   NCACHE") and "cache splitting" not in l]
    a, b = section(log), section(gold)
    diff = [(x, y) for x, y in zip(a, b) if x != y]
    assert len(a) == len(b) and not diff, (len(a), len(b), diff[:5])
    loops = [l for l in log if l.startswith(" In total 2 +")]
    assert loops == [l for l in gold if l.startswith(" In total 2 +")]


@pytest.mark.skipif(not os.path.exists(GEN), reason="reference generator not built (oracle/_ref)")
def test_ci_wirechamber_res():
    with tempfile.TemporaryDirectory() as d:
        subprocess.run([GEN, "100000"], cwd=d, check=True, capture_output=True)
        with open(os.path.join(d, "steer.txt"), "w") as f:
            f.write(STEER)
        st = mpgpu.run_steering(os.path.join(d, "steer.txt"), d)
        assert st["exit_code"] == 0
        assert st["nrec"] == 100000 and st["ntot"] == 200 and st["ncon"] == 2
        got = parse_res(os.path.join(d, "millepede.res"))
        mon = open(os.path.join(d, "millepede.mon")).read().splitlines()
        assert open(os.path.join(d, "millepede.end")).read() == open(GOLD_END).read()
    gold_mon = open(GOLD_MON).read().splitlines()
    assert len(mon) == len(gold_mon) == 202
    diff = [(a, b) for a, b in zip(mon, gold_mon) if a != b]
    assert not diff, diff[:5]
    gold = parse_res(GOLD)
    assert sorted(got) == sorted(gold)
    bad = []
    for lab, (vals, cnt) in gold.items():
        gv, gc = got[lab]
        # 5 significant digits: compare to half a unit in the 5th digit (plus rounding slack)
        if abs(gv[0] - vals[0]) > 1.2e-4 * max(abs(vals[0]), 1e-6):
            bad.append((lab, vals[0], gv[0]))
        assert gc == cnt, (lab, gc, cnt)
    assert not bad, bad[:10]
    # final chi^2 of the CI log (CI/refFiles/millepede.log:153-155)
    assert abs(st["final_chi2"] - 8130954.4875403922) < 1e-6 * 8130954.4875403922


@pytest.mark.skipif(not os.path.exists(GEN), reason="reference generator not built (oracle/_ref)")
def test_errlabels_match_inversion(capfd):
    """'errlabels' (solglo, src/pede.f90:537-612): for selected labels the MINRES run prints the
    error sqrt|(A^-1)_ii| from one solve of A x = e_i.  On the CI wire chamber (constraints with
    multipliers, as the inversion uses) the printed errors (4 significant digits) agree with the
    error column that 'method inversion' writes to millepede.res."""
    labs = [12, 14, 100, 210, 501, 550]
    with tempfile.TemporaryDirectory() as d:
        subprocess.run([GEN, "20000"], cwd=d, check=True, capture_output=True)
        # without down-weighting the matrix does not depend on the parameters, which the two
        # methods reach by different routes
        base = STEER.replace("outlierdownweighting 3\ndwfractioncut 0.1\n", "")
        with open(os.path.join(d, "inv.txt"), "w") as f:
            f.write(base)
        st = mpgpu.run_steering(os.path.join(d, "inv.txt"), d)
        assert st["exit_code"] == 0
        inv = parse_res(os.path.join(d, "millepede.res"))
        capfd.readouterr()
        steer = base.replace("method inversion 3 0.8", "method sparseMINRES 3 0.8\nmrestol 1.0D-10\nwithmultipliers\n"
                             "bandwidth 6\nerrlabels " + " ".join(map(str, labs)))
        with open(os.path.join(d, "mr.txt"), "w") as f:
            f.write(steer)
        st = mpgpu.run_steering(os.path.join(d, "mr.txt"), d)
        assert st["exit_code"] == 0
        out = capfd.readouterr().out.splitlines()
    i = next(k for k, l in enumerate(out) if "label     parameter    presigma      differ       Error gcor^2   iit" in l)
    rows = {int(l.split()[0]): l for l in out[i + 2:i + 2 + len(labs)]}
    seen = 0
    for lab in labs:
        if lab not in inv or len(inv[lab][0]) < 4:
            continue                 # fixed or unknown label: no line
        line = rows[lab]
        err = float(line[12 + 36:12 + 48])
        assert abs(err - inv[lab][0][3]) <= 2e-3 * abs(inv[lab][0][3]), (lab, err, inv[lab][0][3])
        seen += 1
    assert seen >= 3


@pytest.mark.skipif(not os.path.exists(GEN), reason="reference generator not built (oracle/_ref)")
@pytest.mark.parametrize("method", ["sparseMINRES", "fullMINRES", "sparseMINRES-QLP"])
def test_ci_wirechamber_minres_elimination(method):
    """The CI wire chamber solved by MINRES with the two constraints eliminated (pede's default
    constraint mode, QL decomposition): the parameters of CI/refFiles/millepede.res at its 5
    significant digits and the entry counts, as for the inversion."""
    with tempfile.TemporaryDirectory() as d:
        subprocess.run([GEN, "100000"], cwd=d, check=True, capture_output=True)
        steer = STEER.replace("method inversion 3 0.8", f"method {method} 3 0.8\nmrestol 1.0D-10")
        with open(os.path.join(d, "steer.txt"), "w") as f:
            f.write(steer)
        st = mpgpu.run_steering(os.path.join(d, "steer.txt"), d)
        assert st["exit_code"] == 0
        assert st["ncon"] == 0              # no multiplier rows: the constraints are eliminated
        got = parse_res(os.path.join(d, "millepede.res"))
    gold = parse_res(GOLD)
    assert sorted(got) == sorted(gold)
    bad = [(lab, vals[0], got[lab][0][0]) for lab, (vals, cnt) in gold.items()
           if abs(got[lab][0][0] - vals[0]) > 1.2e-4 * max(abs(vals[0]), 1e-6)]
    assert not bad, bad[:10]
    assert all(got[lab][1] == cnt for lab, (vals, cnt) in gold.items())


@pytest.mark.skipif(not os.path.exists(GEN), reason="reference generator not built (oracle/_ref)")
def test_ci_wirechamber_verbatim_steering():
    """CI/mp2str.txt exactly as the reference's CI job runs it (tests/golden/ci_mp2str.txt): six
    method lines of which the last, 'fullLAPACK 3 0.8', applies; 'bandwidth 6 1',
    'maxlocalcond 11', 'skipemptycons', 'resolveredundancycons', 'printcounts 2' and 'threads 24'
    included.  All four goldens: millepede.res (parameters to 5 digits, counts exactly),
    millepede.mon line for line, millepede.end, the iteration table of millepede.log, and the
    final chi^2 of CI/refFiles/millepede.log:153 to 1e-9 relative."""
    with tempfile.TemporaryDirectory() as d:
        subprocess.run([GEN, "100000"], cwd=d, check=True, capture_output=True)
        steer = os.path.join(d, "mp2str.txt")
        with open(os.path.join(ROOT, "tests", "golden", "ci_mp2str.txt")) as src, open(steer, "w") as dst:
            dst.write(src.read())
        st = mpgpu.run_steering(steer, d)
        assert st["exit_code"] == 0
        assert st["nrec"] == 100000 and st["ntot"] == 200 and st["ncon"] == 2
        got = parse_res(os.path.join(d, "millepede.res"))
        mon = open(os.path.join(d, "millepede.mon")).read().splitlines()
        log = open(os.path.join(d, "millepede.log")).read().splitlines()
        # 'LOOP1: 200 is total number nTotalParGroups of label/parameter groups'
        # (CI/refFiles/millepede.log:17): every wire-chamber parameter is its own group
        P = mpgpu.Pede(matsto=mpgpu.FULL, metsol=mpgpu.INVERSION, mreqenf=50, mcount=1)
        assert P.read_mille(os.path.join(d, "mp2tst.bin"), fortran=True) == 100000
        P.build_structure()
        assert len(set(P.labels()[3].tolist())) == 200
        P.close()
        assert open(os.path.join(d, "millepede.end")).read() == open(GOLD_END).read()
    gold = parse_res(GOLD)
    assert sorted(got) == sorted(gold)
    bad = [(lab, vals[0], got[lab][0][0]) for lab, (vals, cnt) in gold.items()
           if abs(got[lab][0][0] - vals[0]) > 1.2e-4 * max(abs(vals[0]), 1e-6)]
    assert not bad, bad[:10]
    assert all(got[lab][1] == cnt for lab, (vals, cnt) in gold.items())
    gold_mon = open(GOLD_MON).read().splitlines()
    diff = [(a, b) for a, b in zip(mon, gold_mon) if a != b]
    assert len(mon) == len(gold_mon) and not diff, diff[:5]
    compare_log(log, open(GOLD_LOG).read().splitlines())
    compare_log_sizes(log, open(GOLD_LOG).read().splitlines())
    rel = abs(st["final_chi2"] - 8130954.4875403922) / 8130954.4875403922
    print(f"final chi^2 {st['final_chi2']!r}: relative difference to the CI log {rel:.2e}")
    assert rel <= 1e-9


@pytest.mark.skipif(not os.path.exists(GEN), reason="reference generator not built (oracle/_ref)")
def test_ci_wirechamber_fulllapack():
    """The CI's own method, 'method fullLAPACK 3 0.8' (mdptrf, src/mplapack.f90:41-179), here
    with 'LAPACKwitherrors': parameters and counts of CI/refFiles/millepede.res, the iteration
    table, line-search tables and chi^2 summary of CI/refFiles/millepede.log, and the error and
    global-correlation columns equal to those 'method inversion' writes for the same data (both
    print 5 significant digits)."""
    with tempfile.TemporaryDirectory() as d:
        subprocess.run([GEN, "100000"], cwd=d, check=True, capture_output=True)
        steer = STEER.replace("method inversion 3 0.8", "method fullLAPACK 3 0.8\nLAPACKwitherrors")
        with open(os.path.join(d, "lp.txt"), "w") as f:
            f.write(steer)
        st = mpgpu.run_steering(os.path.join(d, "lp.txt"), d)
        assert st["exit_code"] == 0
        got = parse_res(os.path.join(d, "millepede.res"))
        log = open(os.path.join(d, "millepede.log")).read().splitlines()
        with open(os.path.join(d, "inv.txt"), "w") as f:
            f.write(STEER)
        st2 = mpgpu.run_steering(os.path.join(d, "inv.txt"), d)
        assert st2["exit_code"] == 0
        inv = parse_res(os.path.join(d, "millepede.res"))
    gold = parse_res(GOLD)
    assert sorted(got) == sorted(gold)
    bad = [(lab, vals[0], got[lab][0][0]) for lab, (vals, cnt) in gold.items()
           if abs(got[lab][0][0] - vals[0]) > 1.2e-4 * max(abs(vals[0]), 1e-6)]
    assert not bad, bad[:10]
    assert all(got[lab][1] == cnt for lab, (vals, cnt) in gold.items())
    assert abs(st["final_chi2"] - 8130954.4875403922) < 1e-6 * 8130954.4875403922
    compare_log(log, open(GOLD_LOG).read().splitlines())
    nerr = 0
    for lab, (vals, _) in inv.items():
        if len(vals) < 4:
            continue
        g = got[lab][0]
        assert len(g) == len(vals)
        for a, b in zip(g[3:], vals[3:]):
            assert abs(a - b) <= 2.5e-4 * max(abs(b), 1e-6), (lab, g, vals)
        nerr += 1
    assert nerr >= 150
