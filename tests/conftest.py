"""Shared test set-up: repository root on sys.path, the `gpu` marker, and
loaders for the golden fixtures under tests/golden/."""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built HIP library")


def load_golden(name: str) -> dict:
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as d:
        return {k: d[k] for k in d.files}


def netlist_from_golden(fx: dict):
    """Rebuild the Netlist a topology fixture was generated from."""
    from karam_amd.netlists import ModelCard, Netlist

    def card(s):
        name, typ, kp, vto, lam = str(s).split("|")
        f = lambda x: None if x == "" else float(x)  # noqa: E731
        return ModelCard(name, typ or None, f(kp), f(vto), f(lam))

    kw = {k[3:]: fx[k] for k in fx if k.startswith("nl_") and k != "nl_models"}
    return Netlist(num_nodes=int(fx["num_nodes"]), ambient_temp_K=float(fx["ambient_temp_K"]),
                   models=[card(s) for s in fx["nl_models"]], **kw)


TOPO_CASES = ["inv16", "rdmesh8", "rdmesh16", "mixed1", "mixed2"]


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle as O
    O.lib()
    return O
