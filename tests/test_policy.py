"""The learned direction policy: MLP gradients (GPU vs the FP64 oracle, the
oracle vs finite differences), training, the probe, and the two claims the
policy rests on -- it never changes an output, and what it gains over the
alpha/beta rule, measured on held-out roots."""
import sys

import numpy as np
import pytest

from oracle import mlp_oracle as mo


def say(m):
    print(m, flush=True)
    sys.stdout.flush()


def _data(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 1, (n, 16)).astype(np.float32) * 3 + 1
    y = rng.integers(0, 2, n).astype(np.int32)
    w = rng.random(n).astype(np.float32) + 0.1
    p = np.zeros(mo.NW, np.float32)
    p[:mo.NTRAIN] = rng.normal(0, 0.3, mo.NTRAIN)
    p[mo.NTRAIN:mo.NTRAIN + 16] = x.mean(0)
    p[mo.NTRAIN + 16:] = x.std(0) + 0.5
    return x, y, w, p


# ---------------------------------------------------------------- CPU
def test_oracle_gradient_vs_finite_differences():
    x, y, w, p = _data(40, 1)
    p = p.astype(np.float64)
    _, g = mo.loss_grad(x, y, w, p)
    rng = np.random.default_rng(2)
    idx = np.concatenate([rng.choice(mo.NTRAIN, 60, replace=False), [mo.NTRAIN - 1, mo.NTRAIN - 2]])
    eps = 1e-6
    for k in idx:
        e = np.zeros_like(p)
        e[k] = eps
        fd = (mo.loss(x, y, w, p + e) - mo.loss(x, y, w, p - e)) / (2 * eps)
        assert abs(fd - g[k]) <= 1e-6 + 1e-5 * abs(fd), (k, fd, g[k])
    # planted mistake: a gradient without the ReLU mask disagrees with the
    # finite differences on W1 entries of inactive units
    W1, b1, W2, b2, mu, sd = mo.unpack(p)
    xs, a, h, z = mo.forward(x, p)
    assert (a < 0).any()
    wn = w / w.sum()
    zm = z - z.max(axis=1, keepdims=True)
    pr = np.exp(zm) / np.exp(zm).sum(axis=1, keepdims=True)
    oh = np.zeros_like(pr)
    oh[np.arange(len(y)), y] = 1.0
    wrong_gW1 = (((wn[:, None] * (pr - oh)) @ W2)).T @ xs       # no (a > 0) mask
    diff = np.abs(wrong_gW1.ravel() - g[:mo.H * mo.F]).max()
    assert diff > 1e-3


def test_heuristic_choice_mirrors_the_rule():
    from lgmi.policy import heuristic_choice
    rec = np.zeros((4, 20))
    # after top-down: m_F = 2^10-1, m_U = 2^12-1 -> m_F > m_U/15 -> bottom-up
    rec[0, [1, 2, 8, 18]] = [10, 12, 0, 3]
    # after top-down, m_F small -> stay top-down
    rec[1, [1, 2, 8, 18]] = [4, 12, 0, 3]
    # after bottom-up, shrinking and |F| <= nv/18 -> top-down
    rec[2, [5, 7, 8, 18]] = [0.01, -1, 1, 5]
    # after bottom-up, growing -> stay bottom-up
    rec[3, [5, 7, 8, 18]] = [0.01, 1, 1, 5]
    assert heuristic_choice(rec).tolist() == [1, 0, 0, 1]


# ---------------------------------------------------------------- GPU
@pytest.mark.gpu
def test_gpu_gradient_and_logits_match_oracle():
    from lgmi import policy
    for n, seed in ((1, 3), (37, 4), (5000, 5)):
        x, y, w, p = _data(n, seed)
        loss, g = policy.loss_grad(x, y, w, p)
        lo, go = mo.loss_grad(x, y, w, p.astype(np.float64))
        say(f"n={n} loss gpu {loss:.6f} oracle {lo:.6f}")
        assert abs(loss - lo) <= 1e-5 * max(1.0, abs(lo))
        scale = np.abs(go).max()
        assert np.allclose(g, go, rtol=1e-3, atol=1e-5 * scale)
        assert (g[mo.NTRAIN:] == 0).all()
        z = policy.logits(x, p)
        _, _, _, zo = mo.forward(x, p.astype(np.float64))
        assert np.allclose(z, zo, rtol=1e-4, atol=1e-4)
    # unit weights (w = NULL) equal explicit ones
    x, y, w, p = _data(100, 6)
    l1, g1 = policy.loss_grad(x, y, None, p)
    l2, g2 = policy.loss_grad(x, y, np.ones(100, np.float32), p)
    assert abs(l1 - l2) < 1e-6 and np.allclose(g1, g2, atol=1e-7)


@pytest.mark.gpu
def test_gpu_training_reduces_loss_and_is_reproducible():
    from lgmi import policy
    rng = np.random.default_rng(7)
    x = rng.normal(0, 1, (3000, 16)).astype(np.float32)
    y = (x[:, 1] - x[:, 2] + 0.3 * x[:, 5] > 0).astype(np.int32)   # learnable rule
    w = np.ones(3000, np.float32)
    p0 = policy.init_params(x, seed=1)
    l0, _ = policy.loss_grad(x, y, w, p0)
    p1, l1 = policy.train(x, y, w, steps=400, lr=1e-2, params=p0)
    p2, l2 = policy.train(x, y, w, steps=400, lr=1e-2, params=p0)
    say(f"loss {l0:.4f} -> {l1:.4f}")
    assert l1 < 0.25 * l0
    assert (p1 == p2).all() and l1 == l2
    acc = (policy.logits(x, p1).argmax(1) == y).mean()
    assert acc > 0.95
    lo, _ = mo.loss_grad(x, y, w, p1.astype(np.float64))
    assert abs(lo - l1) < 1e-4


@pytest.mark.gpu
def test_probe_records_and_state_restore(oracle):
    import lgmi
    from lgmi import policy, traversals
    g = lgmi.generators.rmat(14, 16, seed=3)
    rp, col = g.csr()
    root = int(np.argmax(np.diff(rp)))
    depth = policy.bfs_levels(g, root)
    rec = policy.probe(g, root, range(depth))
    assert len(rec) == depth
    assert (rec[:, 16] > 0).all() and (rec[:, 17] > 0).all()
    assert rec[:, 18].tolist() == list(range(depth))
    # a normal BFS afterwards is still exact
    assert (traversals._parents0(g, root + 1, 0) == oracle.bfs_parents(rp, col, root)).all()


@pytest.mark.gpu
def test_learned_policy_outputs_identical_and_gain(oracle):
    """Train on probes of some roots, evaluate on held-out roots: outputs
    are bit-identical to the alpha/beta run, and the per-level decision cost
    (time of the chosen direction, measured on the identical state) of the
    network is below the heuristic's and no better than the best choice."""
    import lgmi
    from lgmi import policy
    graphs = [lgmi.generators.rmat(16, 16, seed=11), lgmi.generators.barabasi_albert(60_000, 8, seed=12),
              lgmi.generators.rmat(14, 16, seed=13, directed=True)]
    tr, te = [], []
    rng = np.random.default_rng(0)
    for g in graphs:
        deg = np.diff(g.csr()[0])
        roots = rng.choice(np.nonzero(deg > 0)[0], size=48, replace=False)
        tr.append(policy.collect(g, roots[:32]))
        te.append((g, roots[32:], policy.collect(g, roots[32:])))
    x, y, w = policy.labels(tr)
    params, loss = policy.train(x, y, w, steps=1500, lr=3e-3)
    tot = {"heuristic": 0.0, "learned": 0.0, "best": 0.0}
    for g, roots, rec in te:
        tot["heuristic"] += policy.decision_cost(rec, policy.heuristic_choice(rec))
        tot["learned"] += policy.decision_cost(rec, policy.logits(rec[:, :16], params).argmax(1))
        tot["best"] += policy.decision_cost(rec, (rec[:, 17] < rec[:, 16]).astype(int))
        th, tl, same = policy.evaluate(g, roots[:6], params)
        assert same, "learned policy changed an output"
    say(f"held-out decision cost (us): {tot}; gain {tot['heuristic'] / tot['learned']:.3f}x")
    assert tot["best"] <= tot["learned"] + 1e-6
    # the network must beat the rule it replaces on roots it never saw
    assert tot["learned"] < tot["heuristic"]
