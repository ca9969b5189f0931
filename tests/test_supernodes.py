"""Relaxed supernodes of the multifrontal LU (host only).

The library's partition (topology.cpp) equals the oracle's (oracle.c) on
circuits of every family, and is checked against the rule it implements,
from the filled pattern alone (numpy, no library code):
- a supernode's columns form a chain of the elimination tree, and every
  column's structure lies within the later columns and the last column's
  rows, so the front holds every entry of the factor;
- its explicit zeros are at most a fifth of its L entries (front order
  <= 512, at most 16 pivots), or there are none;
- every boundary is one the rule had to draw: the column below it is not a
  child of the first column above, or joining it would break the bound;
- on the mesh and the logic family the relaxed partition holds fewer fronts
  than the fundamental one and fewer update-block entries (sum of m^2).
"""
from __future__ import annotations

import numpy as np
import pytest

from karam_amd import _abi, netlists


def _structure(T):
    rp = T.export(_abi.NR_EXPORT_FILL_ROW_PTR)
    ci = T.export(_abi.NR_EXPORT_FILL_COL_IDX)
    n = len(rp) - 1
    st = []
    for j in range(n):
        row = ci[rp[j]:rp[j + 1]]
        st.append(np.sort(row[row > j]))
    par = np.array([s[0] if len(s) else -1 for s in st])
    cnt = np.array([len(s) for s in st])
    nch = np.bincount(par[par >= 0], minlength=n)
    return st, par, cnt, nch


def _ents(c, r):
    return c * r - c * (c + 1) // 2


def _check_rule(firsts, st, par, cnt, nch):
    n = len(st)
    bounds = list(firsts) + [n]
    for s in range(len(firsts)):
        f, last = bounds[s], bounds[s + 1] - 1
        c, r = last - f + 1, last - f + 1 + cnt[last]
        rows = set(range(f, last + 1)) | set(st[last].tolist())
        zeros = 0
        for k in range(f, last):
            assert par[k] == k + 1, (f, last, k)
            assert set(st[k].tolist()) <= rows, (f, last, k)
            zeros += (last - k + cnt[last]) - cnt[k]
        fundamental = all(cnt[k] == cnt[k + 1] + 1 and nch[k + 1] == 1 for k in range(f, last))
        assert zeros == 0 or (5 * zeros <= _ents(c, r) and r <= 512 and c <= 16), (f, last, zeros)
        if fundamental:
            assert zeros == 0
        # the boundary below f: joining column f-1 is refused by the rule
        if f > 0:
            j = f
            if par[j - 1] != j:
                continue
            zadd = (last - j + 1 + cnt[last]) - cnt[j - 1]
            c2, r2 = last - j + 2, last - j + 2 + cnt[last]
            assert not (zadd == 0 and nch[j] == 1), j
            assert r2 > 512 or c2 > 16 or 5 * (zeros + zadd) > _ents(c2, r2), j


def _fundamental(par, cnt, nch):
    n = len(par)
    return [0] + [j for j in range(1, n) if not (par[j - 1] == j and cnt[j - 1] == cnt[j] + 1 and nch[j] == 1)]


def circuits():
    return {
        "rdmesh16": netlists.rd_mesh(16, seed=2),
        "rdmesh40": netlists.rd_mesh(40, seed=3),
        "logic2000": netlists.random_logic(2000, seed=1, width=64, p_inv=0.3, p_nand3=0.1),
        "hightw": netlists.random_logic(1000, seed=3, spread=32, window=6, width=48),
        "chain200": netlists.inverter_chain(200),
        "inv16": netlists.inverter_chain(16),
    }


@pytest.mark.parametrize("name", list(circuits()))
def test_partition_equals_oracle_and_rule(name):
    from karam_amd import nr
    from oracle import oracle as O
    nl = circuits()[name]
    T = nr.Topology(nl)
    To = O.Topology(nl, ordering=T.export(_abi.NR_EXPORT_ORDERING_RAW))
    firsts = T.export(_abi.NR_EXPORT_SN_FIRST)
    np.testing.assert_array_equal(To.export(_abi.NR_EXPORT_SN_FIRST), firsts)
    assert len(firsts) == T.info["n_supernodes"] and firsts[0] == 0
    st, par, cnt, nch = _structure(T)
    _check_rule(firsts, st, par, cnt, nch)


@pytest.mark.parametrize("name", ["rdmesh40", "logic2000"])
def test_relaxed_partition_holds_fewer_fronts(name):
    from karam_amd import nr
    nl = circuits()[name]
    T = nr.Topology(nl)
    st, par, cnt, nch = _structure(T)
    n = len(st)

    def sum_m2(firsts):
        lasts = np.array(list(firsts[1:]) + [n]) - 1
        return int((cnt[lasts].astype(np.int64) ** 2).sum())

    relaxed, fund = T.export(_abi.NR_EXPORT_SN_FIRST), _fundamental(par, cnt, nch)
    assert len(relaxed) < len(fund)
    assert sum_m2(relaxed) < sum_m2(fund)
