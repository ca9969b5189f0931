"""The learned seeder on the GPU: its BF16-MFMA forward and backward passes
against the FP64 restatement (oracle/mlp.py), and what the trained network
gains over the plain GMIN-stepping start on Monte-Carlo instances.

Tolerances: against the restatement that rounds every matrix-product
operand to BF16 as the kernels do, loss within 1e-3 relative and the
gradient within 2e-2 relative (norm); against the unrounded FP64 network,
5e-2 (the BF16 error budget).  Seeded solves meet the plain path's bar:
KCL <= 1e-9 A, and voltages and iterations equal to the oracle started from
the same seed (1e-9 rel + 1e-12 V)."""
from __future__ import annotations

import numpy as np
import pytest

from karam_amd import _abi, netlists

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from karam_amd import nr
    return nr.Context(0)


def test_loss_and_gradient_match_fp64_restatement(ctx):
    from karam_amd import nr
    from oracle import mlp
    F, O, H, NH, B = 64, 100, 64, 2, 96
    s = nr.Seeder(ctx, F, O, hidden=H, n_hidden=NH, init_seed=3)
    dims = mlp.layer_dims(F, O, H, NH)
    rng = np.random.default_rng(5)
    p = (rng.normal(0, 1, s.params().size) * 0.2).astype(np.float32)
    s.set_params(p)
    np.testing.assert_array_equal(s.params(), p)
    x = rng.normal(size=(B, F)).astype(np.float32)
    y = rng.normal(size=(B, O)).astype(np.float32)
    loss, g = s.loss_grad(x, y)
    loss_q, g_q = mlp.loss_grad(p.astype(np.float64), x.astype(np.float64), y.astype(np.float64), O, dims,
                                round_bf16=True)
    loss_f, g_f = mlp.loss_grad(p.astype(np.float64), x.astype(np.float64), y.astype(np.float64), O, dims)
    assert abs(loss - loss_q) <= 1e-3 * loss_q
    assert np.linalg.norm(g - g_q) <= 2e-2 * np.linalg.norm(g_q)
    assert abs(loss - loss_f) <= 5e-2 * loss_f
    assert np.linalg.norm(g - g_f) <= 5e-2 * np.linalg.norm(g_f)
    # every layer's weight and bias gradient, not only the whole vector
    o = 0
    for din, dout in dims:
        for size in (din * dout, dout):
            seg = slice(o, o + size)
            assert np.linalg.norm(g[seg] - g_q[seg]) <= 3e-2 * np.linalg.norm(g_q[seg]) + 1e-7
            o += size


FAMILIES = {
    # the configured Monte-Carlo spread: the operating points move little, so a
    # non-learned seed at the last GMIN step is already close
    "rd_is_5pct": dict(),
    # the current sources spread by 40 %: every node moves with them
    "sources_40pct": dict(sigma_isrc=0.4),
}


@pytest.mark.parametrize("family", list(FAMILIES))
def test_trained_seeder_cuts_newton_iterations(ctx, family):
    """16x16 R-diode mesh: train on 4096 solved instances, then solve 512
    unseen instances seeded at the last GMIN step with one polish iteration.
    Seeded answers lie within the voltage tolerance (1e-9 rel + 1e-12 V) of
    the plain FP64 answers.  Against the seed without learning at the same
    step (the mean of the training solutions): never more iterations, and
    on the family whose sources spread, clearly fewer."""
    from karam_amd import nr
    from oracle import oracle as O
    kw = FAMILIES[family]
    nl = netlists.rd_mesh(16, seed=2)
    n, F = nl.num_nodes, 64
    T = nr.Topology(nl, ctx)
    Ntr, Nte = 4096, 512
    ftr = netlists.mc_fields(nl, Ntr, seed=700000, **kw)
    bt = nr.Batch(ctx, T, Ntr)
    bt.set_fields(ftr)
    assert bt.solve_dc()["n_converged"] == Ntr
    s = nr.Seeder(ctx, F, n, hidden=256, n_hidden=3, init_seed=1)
    s.record_solution(netlists.mc_features(nl, ftr, F), bt.voltages())
    assert s.has_history() and s.history() == Ntr
    loss0 = s.train(epochs=1, batch=256, lr=0.0)       # the untrained network's loss
    loss = s.train(epochs=40, batch=256, lr=1e-3)
    assert loss < 0.7 * loss0
    fte = netlists.mc_fields(nl, Nte, seed=900000, **kw)
    xte = netlists.mc_features(nl, fte, F)
    te = nr.Batch(ctx, T, Nte)
    te.set_fields(fte)
    plain = te.solve_dc()
    it_plain = te.stats()["iterations"]
    V_plain = te.voltages()
    te.set_seed(np.tile(bt.voltages().mean(axis=0), (Nte, 1)))
    te.solve_dc(seeded_start_step=18, polish=1)
    st_mean = te.stats()
    te.set_seed_from(s, xte)
    seeded = te.solve_dc(seeded_start_step=18, polish=1)
    st = te.stats()
    V = te.voltages()
    assert plain["n_converged"] == Nte and seeded["n_converged"] == Nte
    assert st["kcl"].max() <= 1e-9 and (st["seeded"] == 1).all()
    m_net, m_mean = st["iterations"].mean(), st_mean["iterations"].mean()
    print(f"{family}: seeder loss {loss0:.3f} -> {loss:.3f}; iterations plain {it_plain.mean():.2f}, "
          f"learned seed {m_net:.2f} (max {st['iterations'].max()}), training-mean seed {m_mean:.2f}")
    assert m_net <= 7 and m_net * 10 <= it_plain.mean()
    assert m_net <= m_mean
    if kw:
        assert m_net <= m_mean - 0.5
    # within the voltage tolerance of the plain answers: two Newton paths to
    # the same root part at the FP64 floor of this mesh (the KCL settles near
    # 1e-11 A, about half the tolerance), so a few instances may sit just
    # above it
    e = (np.abs(V - V_plain) / (1e-9 * np.abs(V_plain) + 1e-12)).max(axis=1)
    print(f"{family}: seeded vs plain FP64: max {e.max():.3f} of the tolerance, {(e <= 1).mean():.4f} within")
    assert (e <= 1.0).mean() >= 0.99 and e.max() <= 3.0
    # the device-side seed equals the host prediction; the oracle started from
    # that seed follows the same iterations to the same voltages
    seed_v = s.predict_seed(xte)
    pick = [0, 77, 311]
    sub = {k: v[pick] for k, v in fte.items()}
    Vo, so = O.solve_dc_batch(nl, O.Topology(nl), sub, 3, seeds=seed_v[pick], start_step=18, polish=1)
    for j, b in enumerate(pick):
        assert so["iterations"][j] == st["iterations"][b]
        assert np.all(np.abs(V[b] - Vo[j]) <= 1e-9 * np.abs(Vo[j]) + 1e-12)


def test_plain_fallback_after_seeded_failure(ctx):
    """Instances whose seed fails (here: a wild seed with a 2-iteration cap)
    are solved again from the plain start, bit-identically to a plain solve
    with the same options; good seeds keep their seeded result."""
    from karam_amd import nr
    nl = netlists.rd_mesh(16, seed=2)
    B = 8
    f = netlists.mc_fields(nl, B, seed=31)
    T = nr.Topology(nl, ctx)
    bt = nr.Batch(ctx, T, B)
    bt.set_fields(f)
    bt.solve_dc()
    V = bt.voltages()
    bt.solve_dc(max_iter=2)
    V2, st2 = bt.voltages(), bt.stats()
    seed = V.copy()
    seed[4:] = 1e3
    bt.set_seed(seed)
    bt.solve_dc(seeded_start_step=18, max_iter=2, fallback_plain=0)
    st = bt.stats()
    assert np.all(st["seeded"] == 1)
    assert np.all(st["status"][:4] == _abi.NR_STATUS_CONVERGED) and np.all(st["iterations"][:4] == 1)
    assert np.all(st["status"][4:] == _abi.NR_STATUS_NOT_CONVERGED)
    bt.solve_dc(seeded_start_step=18, max_iter=2, fallback_plain=1)
    Vf, st = bt.voltages(), bt.stats()
    np.testing.assert_array_equal(st["seeded"], [1, 1, 1, 1, 0, 0, 0, 0])
    assert np.all(np.abs(Vf[:4] - V[:4]) <= 1e-9 * np.abs(V[:4]) + 1e-12)
    np.testing.assert_array_equal(Vf[4:], V2[4:])
    np.testing.assert_array_equal(st["status"][4:], st2["status"][4:])
    np.testing.assert_array_equal(st["iterations"][4:], st2["iterations"][4:] + 2)
