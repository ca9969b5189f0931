"""Writes tests/golden/logic20k_ordering.npz: the oracle's literal Min-Fill
ordering (the O(n^2 d^2) restatement of TreewidthAnalyzer::analyze,
infrastructure/treewidth_analyzer.cpp:14-141, on the factor graph of
infrastructure/factor_graph.cpp:9-162) of the random-logic netlist of the
20,000-node configuration, so that tests need not recompute it (about a
quarter of an hour on one core here).

    python tests/golden/make_logic20k.py

The netlist is netlists.random_logic(**LOGIC20K) (karam_amd/netlists.py).
Keys: "ordering_raw" (int32), "treewidth", "nnz_f", "seconds".
"""
from __future__ import annotations

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

# NAND2 / NAND3 / INV logic, fan-in 1-3, about 17 gate levels of 640 gates
LOGIC20K = dict(n_target=20000, seed=1, width=640, p_inv=0.3, p_nand3=0.1)


def main():
    from karam_amd import _abi, netlists
    from oracle import oracle as O
    nl = netlists.random_logic(**LOGIC20K)
    t = time.time()
    T = O.Topology(nl)
    dt = time.time() - t
    np.savez_compressed(os.path.join(HERE, "logic20k_ordering.npz"),
                        ordering_raw=T.export(_abi.NR_EXPORT_ORDERING_RAW).astype(np.int32),
                        treewidth=np.int32(T.treewidth), nnz_f=np.int64(T.nnz_f), seconds=np.float64(dt))
    print(f"n={nl.num_nodes} mos={len(nl.mos_w)} treewidth={T.treewidth} nnz_f={T.nnz_f} in {dt:.0f} s", flush=True)


if __name__ == "__main__":
    main()
