"""Pure-Python restatement of pede's label bookkeeping for small inputs (loop1,
src/pede.f90:5777-6146; decodeNextMeasurement 2428-2468; updateParGroups/pargrp 2157-2404).

TEST INFRASTRUCTURE ONLY (checker for the library's index maps).

Group partition: pede builds parameter groups incrementally (pargrp) while reading; for each
sorted list (an equation's labels, duplicates kept, or with countrecords a record's distinct
labels) the partition gets a boundary at the list's first element, after its last, and at
both sides of every step that is not +1 (a repeated label is such a step).  That rule is what
this module computes.  tests/test_oracle_labels.py pins it against a step-by-step restatement of
pargrp's incremental procedure on random lists with gaps and repeats, and checks maximality by
brute force (for repeat-free lists: every list is a union of whole groups, and no two adjacent
groups can be merged without breaking that).
"""
import numpy as np


def decode(rec_ptr, inder, glder):
    """Per record: list of equations (sL, sG, lG) as absolute word indices (0-based)."""
    recs = []
    for r in range(len(rec_ptr) - 1):
        w0, w1 = int(rec_ptr[r]), int(rec_ptr[r + 1])
        nw = w1 - w0
        I = lambda k: int(inder[w0 + k - 1])
        V = lambda k: float(glder[w0 + k - 1])
        eqs = []
        last, end = 1, nw
        while last < end:
            last += 1
            while I(last) != 0:
                last += 1
            sL = last
            last += 1
            while I(last) != 0:
                last += 1
            sG = last
            if sL + 1 == sG and V(sG) < 0.0:
                last += int(round(-V(sG)))
                continue
            while last < end and I(last + 1) != 0:
                last += 1
            eqs.append((w0 + sL - 1, w0 + sG - 1, w0 + last - 1))
        recs.append(eqs)
    return recs


def label_maps(rec_ptr, inder, glder, steer_labels=(), presigma=None, mreqenf=25, mcount=0, extra_labels=(),
               iteren=0):
    """labels (sorted), var_index (>0 variable, -1 fixed by pre-sigma, -2 entries cut, -3
    iterated entries cut), counters, group start label per label.

    Iterated entries cut (loop1i, src/pede.f90:6164-6265; only with counting per equation and
    iteren > mreqenf): the global words of every equation without a parameter fixed by the
    entries cut are counted again (zero derivatives included); a variable parameter with fewer
    than iteren entries and fewer than mreqenf such counts is fixed (-3)."""
    recs = decode(rec_ptr, inder, glder)
    labs = set(int(x) for x in steer_labels) | set(int(x) for x in extra_labels)
    for eqs in recs:
        for (sL, sG, lG) in eqs:
            for w in range(sG + 1, lG + 1):
                labs.add(int(inder[w]))
    labels = np.array(sorted(labs), dtype=np.int32)
    pos = {int(l): i for i, l in enumerate(labels)}
    n = len(labels)
    cnt = np.zeros(n, np.int64)
    bnd = np.zeros(n + 1, bool)
    bnd[0] = True

    def mark(lst):
        if not lst:
            return
        lst = sorted(lst)
        bnd[lst[0]] = True
        for a, b in zip(lst[:-1], lst[1:]):
            if b != a + 1:
                bnd[a + 1] = True
                bnd[b] = True
        bnd[lst[-1] + 1] = True

    for eqs in recs:
        recl = []
        for (sL, sG, lG) in eqs:
            el = [pos[int(inder[w])] for w in range(sG + 1, lG + 1) if glder[w] != 0.0]
            if mcount == 0:
                for t in el:
                    cnt[t] += 1
                mark(el)
            recl += el
        if mcount != 0:
            u = sorted(set(recl))
            for t in u:
                cnt[t] += 1
            mark(u)
    ps = np.zeros(n)
    if presigma is not None:
        for l, s in zip(steer_labels, presigma):
            ps[pos[int(l)]] = s
    var = np.zeros(n, np.int32)
    k = 0
    for t in range(n):
        if ps[t] < 0:
            var[t] = -1
        elif cnt[t] < mreqenf:
            var[t] = -2
        else:
            k += 1
            var[t] = k
    if mcount == 0 and iteren > mreqenf:
        low = var == -2
        newc = np.zeros(n, np.int64)
        for eqs in recs:
            for (sL, sG, lG) in eqs:
                words = [pos[int(inder[w])] for w in range(sG + 1, lG + 1)]
                if not any(low[t] for t in words):
                    for t in words:
                        newc[t] += 1
        k = 0
        for t in range(n):
            if var[t] > 0:
                if newc[t] >= mreqenf or cnt[t] >= iteren:
                    k += 1
                    var[t] = k
                else:
                    var[t] = -3
    gsl = np.zeros(n, np.int32)
    start = 0
    for t in range(n):
        if bnd[t]:
            start = t
        gsl[t] = labels[start]
    return labels, var, cnt, gsl


def record_summary(rec_ptr, inder, glder, labels, var):
    """The counts pede's first and second data loops write to millepede.log: equations, those
    with global derivatives and their global words (prepareRecords, src/mppar.f90:345-351), and
    per record the equations, the largest local index and the distinct variable parameters of
    all global words (loop2, src/pede.f90:6683-6838)."""
    recs = decode(rec_ptr, inder, glder)
    pos = {int(l): i for i, l in enumerate(labels)}
    out = dict(neq=0, neq_glob=0, nglob_deriv=0, max_eq_rec=0, max_loc_rec=0, max_glob_rec=0, nempty_rec=0)
    for eqs in recs:
        out["neq"] += len(eqs)
        out["max_eq_rec"] = max(out["max_eq_rec"], len(eqs))
        vs = set()
        for (sL, sG, lG) in eqs:
            if lG > sG:
                out["neq_glob"] += 1
                out["nglob_deriv"] += lG - sG
            for w in range(sL + 1, sG):
                out["max_loc_rec"] = max(out["max_loc_rec"], int(inder[w]))
            vs.update(int(var[pos[int(inder[w])]]) for w in range(sG + 1, lG + 1))
        nv = len([v for v in vs if v > 0])
        out["max_glob_rec"] = max(out["max_glob_rec"], nv)
        out["nempty_rec"] += nv == 0
    return out
