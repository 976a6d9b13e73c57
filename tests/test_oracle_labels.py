"""Parameter groups of the label checker (oracle/labels.py) against pede's own procedure.

pargrp (src/pede.f90:2311-2404) builds the groups incrementally, list by list, in reading
order; pargrp_steps below restates that procedure step by step (build phase, then the three
split rules).  The checker's closed rule -- boundaries at a list's first element, after its
last, and at both sides of every step that is not +1 -- must give the same partition for any
sequence of sorted lists with gaps and repeated labels.  For repeat-free lists the partition is
also checked by brute force to be the coarsest one in which every list is a union of whole
groups."""
import numpy as np

import labels as LB


def pargrp_steps(lists, n):
    """Group start (index) of each of n indices after feeding the lists (each sorted,
    0-based indices) through pargrp in order; indices never listed keep -1.  Restates
    src/pede.f90:2327-2402 with globalParLabelIndex(3,:) as `start` (0 = not in a group is
    represented by -1, labels by their indices)."""
    start = [-1] * n
    last = n - 1
    for lst in lists:
        if not lst:
            continue
        ltgbi, lstart = -2, -2       # 'ltgbi=-1' / 'lstart=-1' in 1-based terms
        for li in lst:               # build up groups (2327-2342)
            istart = start[li]
            if istart == -1:
                if li != ltgbi + 1:
                    start[li] = li
                elif lstart == -1:
                    start[li] = start[ltgbi]
                else:
                    start[li] = li
            ltgbi, lstart = li, istart
        li = lst[0]                  # start inside group? (2345-2355)
        istart, jstart = start[li], li
        if istart != jstart:
            while start[li] == istart:
                start[li] = jstart
                li += 1
                if li > last:
                    break
        ltgbi = lst[0]               # not neighbours anymore (2357-2386)
        for li in lst[1:]:
            if li != ltgbi + 1:
                lstart = start[ltgbi]
                jt = ltgbi + 1
                jstart = jt
                while jt <= last and start[jt] == lstart:
                    start[jt] = jstart
                    jt += 1
                    if jt > last:
                        break
                    if jt == li:
                        jstart = jt
                jt = li
                istart, jstart = start[jt], jt
                if istart != jstart:
                    while start[jt] == istart:
                        start[jt] = jstart
                        jt += 1
                        if jt > last:
                            break
            ltgbi = li
        li = lst[-1]                 # end inside group? (2388-2399)
        if li < last:
            istart = start[li]
            li += 1
            jstart = li
            while start[li] == istart:
                start[li] = jstart
                li += 1
                if li > last:
                    break
    return start


def _records(lists):
    """One record per list: measurement (1.0), one local (1.0, 1), sigma, globals with labels
    10 + index (derivative 1.0)."""
    rec_ptr, I, G = [0], [], []
    for lst in lists:
        I += [0, 0, 1, 0] + [10 + k for k in lst]
        G += [0.0, 1.0, 1.0, 1.0] + [1.0] * len(lst)
        rec_ptr.append(len(I))
    return np.array(rec_ptr), np.array(I, np.int32), np.array(G)


def _checker_starts(lists, n, mcount=0):
    rp, ii, gg = _records(lists)
    labels, var, cnt, gsl = LB.label_maps(rp, ii, gg, mreqenf=0, mcount=mcount, extra_labels=[10 + k for k in range(n)])
    pos = {int(l): i for i, l in enumerate(labels)}
    return [pos[int(g)] for g in gsl], cnt


def test_repeated_label_closes_group():
    # [5,5,6] -> {5}, {6}  (the repeat is a step that is not +1)
    starts, cnt = _checker_starts([[5, 5, 6]], 8)
    assert starts[5] == 5 and starts[6] == 6
    assert pargrp_steps([[5, 5, 6]], 8)[5:7] == [5, 6]
    assert cnt[5] == 2 and cnt[6] == 1            # each list element is an entry (2328-2329)
    # without the repeat the two stay together
    assert _checker_starts([[5, 6]], 8)[0][5:7] == [5, 5]


def test_closed_rule_equals_pargrp_procedure():
    rng = np.random.default_rng(11)
    for case in range(400):
        n = int(rng.integers(2, 40))
        lists = []
        for _ in range(int(rng.integers(1, 12))):
            k = int(rng.integers(1, min(n, 10) + 1))
            base = int(rng.integers(0, n))
            if rng.random() < 0.5:           # a run of consecutive indices, maybe with repeats
                lst = [min(n - 1, base + j) for j in range(k)]
            else:
                lst = list(rng.integers(0, n, size=k))
            if rng.random() < 0.3:
                lst.append(lst[int(rng.integers(0, len(lst)))])
            lists.append(sorted(int(v) for v in lst))
        used = sorted({v for l in lists for v in l})
        ref = pargrp_steps(lists, n)
        got, _ = _checker_starts(lists, n)
        # the checker also groups indices that never appear (they are absent from every list);
        # compare on the indices the lists use
        for v in used:
            assert ref[v] == got[v], (case, lists, v, ref, got)


def test_partition_is_coarsest_union_of_groups():
    rng = np.random.default_rng(12)
    for case in range(200):
        n = int(rng.integers(2, 30))
        lists = []
        for _ in range(int(rng.integers(1, 8))):
            base = int(rng.integers(0, n))
            k = int(rng.integers(1, 6))
            lists.append(sorted({min(n - 1, base + j) for j in range(k)} |
                                ({int(rng.integers(0, n))} if rng.random() < 0.3 else set())))
        starts, _ = _checker_starts(lists, n)
        groups = {}
        for i, s in enumerate(starts):
            groups.setdefault(s, []).append(i)
        order = sorted(groups)

        def unions(gs):
            for lst in lists:
                S = set(lst)
                for g in gs:
                    inter = S & set(g)
                    if inter and inter != set(g):
                        return False
            return True
        parts = [groups[s] for s in order]
        assert unions(parts)
        for a in range(len(parts) - 1):     # merging any two neighbours breaks a list
            merged = parts[:a] + [parts[a] + parts[a + 1]] + parts[a + 2:]
            assert not unions(merged), (lists, parts, a)


def _eq_records(eqs):
    """One record per equation: measurement 1.0, local 1 (1.0), sigma 1.0, then the global
    (label, derivative) pairs."""
    rec_ptr, I, G = [0], [], []
    for glob in eqs:
        I += [0, 0, 1, 0] + [l for l, _ in glob]
        G += [0.0, 1.0, 1.0, 1.0] + [d for _, d in glob]
        rec_ptr.append(len(I))
    return np.array(rec_ptr), np.array(I, np.int32), np.array(G)


# Equations built so that the iterated entries cut (loop1i, src/pede.f90:6164-6265) has a known
# outcome with mreqenf 2, iteren 3 (labels A=1 B=2 C=3 D=4 F=6 K=11, L1..L5=21..25):
#   B and L1..L5 have one entry: fixed by the entries cut (-2);
#   A has 2 entries, one of them next to B: recounted once, 1 < 2 and 2 < iteren -> -3;
#   F has 3 entries, all next to a fixed label: recounted 0 times, but 3 >= iteren -> variable;
#   K has 2 nonzero entries next to fixed labels and 2 zero derivatives in clean equations:
#   zero derivatives are not entries but are recounted (2 >= 2) -> variable.
ITEREN_EQS = [[(1, 1.0), (2, 1.0)], [(1, 1.0), (3, 1.0)], [(3, 1.0), (4, 1.0)], [(3, 1.0), (4, 1.0)],
              [(6, 1.0), (21, 1.0)], [(6, 1.0), (22, 1.0)], [(6, 1.0), (23, 1.0)], [(11, 1.0), (24, 1.0)],
              [(11, 1.0), (25, 1.0)], [(11, 0.0), (3, 1.0)], [(11, 0.0), (4, 1.0)]]
ITEREN_EXPECT = {1: -3, 2: -2, 3: 1, 4: 1, 6: 1, 11: 1, 21: -2, 22: -2, 23: -2, 24: -2, 25: -2}


def test_iterated_entries_cut_known_case():
    rp, ii, gg = _eq_records(ITEREN_EQS)
    labels, var, cnt, _ = LB.label_maps(rp, ii, gg, mreqenf=2, mcount=0, iteren=3)
    got = {int(l): (1 if v > 0 else int(v)) for l, v in zip(labels, var)}
    assert got == ITEREN_EXPECT
    assert [int(c) for c in cnt] == [2, 1, 4, 3, 3, 2, 1, 1, 1, 1, 1]
    # no iteration with counting per record, or with iteren <= mreqenf: A stays variable
    for kw in (dict(mcount=1, iteren=3), dict(mcount=0, iteren=2)):
        labels, var, _, _ = LB.label_maps(rp, ii, gg, mreqenf=2, **kw)
        assert var[list(labels).index(1)] > 0
