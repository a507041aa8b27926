"""GPU checks of learned r-RESPA (no reference counterpart): the BF16-MFMA network against the oracle's FP64
forward and backward, the descriptors against the oracle's, and the learned dynamics against plain PME and against
the stale slow force it replaces (the ablation)."""
import numpy as np
import pytest

import caps_amd as ca
import helpers as H
import orcpy as O

pytestmark = pytest.mark.gpu
NP = 128 * 96 + 128 + 128 * 128 + 128 + 3 * 128 + 3


def _params(rng, scale=1.0):
    P = [rng.normal(0, np.sqrt(1 / 96), (128, 96)), rng.normal(0, 0.1, 128), rng.normal(0, np.sqrt(1 / 128), (128, 128)),
         rng.normal(0, 0.1, 128), rng.normal(0, np.sqrt(1 / 128), (3, 128)), rng.normal(0, 0.1, 3)]
    return [p * scale for p in P]


def _flat(P):
    return np.concatenate([p.reshape(-1) for p in P]).astype(np.float32)


def _bf16(a):
    """round to the nearest BF16 (ties to even), as the kernels store weights and inputs"""
    b = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) >> 16 << 16
    return b.astype(np.uint32).view(np.float32)


def _evaluator():
    s = H.melt_case(chains=2)
    return ca.Evaluator(s.ff, ca.EnergyOptions(cutoff=10.0, coulomb=False, parts=1, tail=False))


def test_network_forward_matches_the_oracle():
    print("case mlp forward", flush=True)
    rng = np.random.default_rng(0)
    P = _params(rng)
    m = 700   # three workgroups of 256 rows, the last one partial
    x = rng.normal(0, 1, (m, 96))
    ev = _evaluator()
    y = ev.learned_forward(_flat(P), x)
    # the oracle at the BF16-rounded weights and inputs: what is left is the BF16 rounding of the activations
    Pb = [_bf16(p).astype(np.float64) if k % 2 == 0 else p for k, p in enumerate(P)]
    yo = np.zeros((m, 3))
    O.orc().orc_mlp_forward(*[np.ascontiguousarray(p, np.float64).reshape(-1) for p in Pb], _bf16(x).astype(np.float64).reshape(-1), m,
                            yo.reshape(-1))
    err = np.sqrt(((y - yo) ** 2).sum() / (yo ** 2).sum())
    print("forward rel RMS", err, flush=True)
    assert err < 1e-2


def test_network_gradient_matches_the_oracle():
    print("case mlp gradient", flush=True)
    rng = np.random.default_rng(1)
    P = _params(rng)
    x, t = rng.normal(0, 1, (512, 96)), rng.normal(0, 1, (512, 3))
    ev = _evaluator()
    loss, g = ev.learned_gradient(_flat(P), x, t)
    G = [np.zeros(p.size) for p in P]
    Pb = [_bf16(p).astype(np.float64) if k % 2 == 0 else p for k, p in enumerate(P)]
    lo = O.orc().orc_mlp_backward(*[np.ascontiguousarray(p, np.float64).reshape(-1) for p in Pb], _bf16(x).astype(np.float64).reshape(-1),
                                  np.ascontiguousarray(t).reshape(-1), 512, *G)
    assert loss == pytest.approx(lo, rel=2e-2)
    off = 0
    for k, p in enumerate(P):
        gk = g[off:off + p.size]
        off += p.size
        err = np.sqrt(((gk - G[k]) ** 2).sum() / (G[k] ** 2).sum())
        print("grad block", k, "rel RMS", err, flush=True)
        assert err < 5e-2, k


def test_descriptors_match_the_oracle():
    print("case descriptors", flush=True)
    s = H.melt_case(chains=16)   # L ≈ 25 Å: three 8 Å cells per axis
    ev = ca.Evaluator(s.ff, ca.EnergyOptions(cutoff=10.0, coulomb=False, parts=1, tail=False))
    ev.set_state(s.x, None, s.cell)
    d = ev.learned_descriptors()
    do = np.zeros(s.ff.n * 96)
    O.orc().orc_descriptors(np.ascontiguousarray(s.x, np.float64).reshape(-1), np.ascontiguousarray(s.ff.charge), s.ff.n,
                            np.ascontiguousarray(s.cell, np.float64), do)
    do = do.reshape(-1, 96)
    err = np.sqrt(((d - do) ** 2).sum() / (do ** 2).sum())
    print("descriptor rel RMS", err, flush=True)
    assert err < 1e-5


def test_descriptors_triclinic_partial_wave():
    # a sheared cell five 4 Å cells wide, rows that wrap at every face, and 1100 atoms: the last workgroup is partly
    # empty
    print("case descriptors triclinic", flush=True)
    s = H.triclinic_case(n=1100)
    ev = ca.Evaluator(s.ff, ca.EnergyOptions(cutoff=8.0, coulomb=False, parts=1, tail=False))
    ev.set_state(s.x, None, s.cell)
    d = ev.learned_descriptors()
    do = np.zeros(s.ff.n * 96)
    O.orc().orc_descriptors(np.ascontiguousarray(s.x, np.float64).reshape(-1), np.ascontiguousarray(s.ff.charge), s.ff.n,
                            np.ascontiguousarray(s.cell, np.float64), do)
    do = do.reshape(-1, 96)
    err = np.sqrt(((d - do) ** 2).sum() / (do ** 2).sum())
    print("descriptor rel RMS", err, flush=True)
    assert err < 1e-5
    # the same positions give the same bits
    assert np.array_equal(d, ev.learned_descriptors())


def test_learned_respa_runs_and_judges_itself():
    # 200 frames of plain PME train the network on (ΔD, ΔF/q) pairs four steps apart; then the slow force comes from the
    # network between full PME steps every 4 steps, each full step judging the prediction against PME and against the
    # stale force held since the last full step (the ablation)
    print("case learned r-RESPA dynamics", flush=True)
    s = H.melt_case(chains=16)
    op = ca.EnergyOptions(cutoff=10.0, coulomb=True, electrostatics="pme", pme_order=4, pme_spacing=1.0, tail=False, parts=5)
    ev = ca.Evaluator(s.ff, op)
    r = ca.run_dynamics(s, ca.DynamicsOptions(dt=0.5, steps=600, temperature=300, thermostat="langevin", seed=3, thermo_every=4,
                                              new_velocities=True, energy=op, learned_k=4, learned_tol=1.0, learned_train=200),
                        evaluator=ev)
    rows = [row for row in r.thermo if row["learned_rmse"] > 0]
    lr = np.array([row["learned_rmse"] for row in rows])
    st = np.array([row["learned_rmse_stale"] for row in rows])
    print("learned steps", r.learned_steps, "full", r.full_steps, "fallbacks", r.fallbacks, flush=True)
    print("judged relative force RMSE: learned %.3e, stale %.3e (mean of %d judgements)" % (lr.mean(), st.mean(), len(lr)), flush=True)
    assert r.learned_steps > 0 and r.full_steps >= 200 and len(lr) > 10
    assert r.judgments >= 1 and r.fallbacks == 0
    assert np.isfinite(s.x).all() and np.isfinite(lr).all()
    # what the network gains over the method without it: a smaller slow-force error than the stale force
    assert lr.mean() < st.mean()
    # with a tolerance below every judged error the run falls back to plain PME
    s2 = H.melt_case(chains=16)
    r2 = ca.run_dynamics(s2, ca.DynamicsOptions(dt=0.5, steps=60, temperature=300, thermostat="langevin", seed=3, thermo_every=4,
                                                new_velocities=True, energy=op, learned_k=4, learned_tol=1e-9, learned_train=0),
                         evaluator=ev)
    print("tight tolerance: learned", r2.learned_steps, "full", r2.full_steps, "judgments", r2.judgments, "fallbacks", r2.fallbacks,
          flush=True)
    # each run starts unjudged: its first full step (step 1) is a reference, the first judgment comes k steps later; after
    # each failed judgment the run waits 4, 8, 16, ... plain steps, takes a reference full step and judges k steps later:
    # judgments at steps 5, 13, 25 and 45 of these 60, every one a fallback, and no step predicted
    assert r2.judgments == 4 and r2.fallbacks == 4 and r2.learned_steps == 0
    judged = sorted({row["learned_rmse"] for row in r2.thermo if row["learned_rmse"] > 0})
    assert len(judged) == 4 and min(judged) > 1e-9
