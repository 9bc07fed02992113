"""GPU tests of the BF16 MFMA step-scale model and the learned minimiser, through the C ABI.

The model's reference is a numpy restatement that applies the kernels' BF16 roundings (inputs,
weights, the first hidden layer, and the dz operands of the weight-gradient products) with FP64
arithmetic elsewhere.  The numpy gradient formula itself is pinned by central finite differences of
the unrounded FP64 loss, so a wrong MFMA operand layout or a wrong backward term fails one of them.
"""
import numpy as np
import pytest

import lima_amd as L

pytestmark = pytest.mark.gpu

from tests.mlp_ref import (NI, O_W1, O_B1, O_W2, O_B2, O_W3, O_B3, bf16, forward_ref, grad_ref,
                            sample_problem)


def test_mlp_forward_matches_reference():
    m = L.Mlp(seed=7)
    p = m.params()
    x, _ = sample_problem(1000)
    s_gpu = m.forward(x)
    s_ref, _ = forward_ref(p, x)
    d = np.abs(s_gpu - s_ref)
    print(f"forward: median |ds| {np.median(d):.2e}, max {d.max():.2e}")
    assert np.median(d) < 2e-5
    assert d.max() < 2e-2   # a BF16 rounding of one hidden value may flip between FP32 and FP64 sums
    assert np.all((s_gpu > 0) & (s_gpu < 4))


def test_mlp_gradient_matches_reference():
    m = L.Mlp(seed=11)
    rng = np.random.default_rng(1)
    p = m.params() + rng.normal(0, 0.05, m.n_params).astype(np.float32)
    m.set_params(p)
    x, t = sample_problem(333)       # not a multiple of 16: exercises the padded tail
    g_gpu, loss_gpu = m.gradient(x, t)
    g_ref, loss_ref = grad_ref(m.params(), x, t)
    assert loss_gpu == pytest.approx(loss_ref, rel=1e-3)
    for name, a, b in (("W1", O_W1, O_B1), ("b1", O_B1, O_W2), ("W2", O_W2, O_B2), ("b2", O_B2, O_W3), ("w3", O_W3, O_B3),
                       ("b3", O_B3, O_B3 + 1)):
        rel = np.linalg.norm(g_gpu[a:b] - g_ref[a:b]) / max(np.linalg.norm(g_ref[a:b]), 1e-12)
        print(f"{name}: rel err {rel:.2e}")
        assert rel < 2e-2, name


def test_mlp_training_fits_a_learnable_target():
    rng = np.random.default_rng(5)
    x = rng.normal(0, 1, (8192, NI)).astype(np.float32)
    w = rng.normal(0, 0.4, NI)
    t = (4.0 / (1.0 + np.exp(-(x @ w)))).astype(np.float32)
    m = L.Mlp(seed=1)
    losses, sec = m.train(x, t, epochs=8, batch=256, lr=3e-3, seed=2)
    print("losses", losses, f"{sec:.2f}s")
    assert losses[-1] < 0.1 * losses[0]
    s = m.forward(x[:512])
    assert np.sqrt(np.mean((s - t[:512]) ** 2)) < 0.3


def _images_hist(pos_nm, q, L_nm, r=1.0, bins=14):
    """Brute-force neighbour histograms: every periodic image within r, bins of r/bins per charge sign."""
    n = len(pos_nm)
    h = np.zeros((n, 2 * bins), np.int64)
    shifts = np.array([[a, b, c] for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1)], np.float64) * L_nm
    for i in range(n):
        d = pos_nm[None, :, :] - pos_nm[i][None, None, :] + shifts[:, None, :]
        dist = np.sqrt((d ** 2).sum(-1))
        dist[13, i] = 99.0
        for s_idx in range(27):
            m = dist[s_idx] <= r
            b = np.minimum(bins - 1, (dist[s_idx][m] * bins / r).astype(int)) + np.where(q[m] < 0, bins, 0)
            np.add.at(h[i], b, 1)
    return h


def test_neighbour_histogram_features_match_brute_force():
    box = L.gen_em_system(300, 1000, 0.02)
    b = L.Batch([box], L.default_params())
    b.record(100.0, 0, 0.01, atom_stride=1, max_states=1)     # features of the starting state
    f = b.features(0)[:, 4:]
    st = b.state(0)
    li = box.layout_index()
    origo = st["compound_origo"]
    coord = st["particle_coord"].reshape(-1, 3)
    comp = li // 32
    pos = np.mod((origo[comp].astype(np.int64) * 100000000 + coord[li]).astype(np.float64) * 1e-8, box.box_nm[0])
    q = box.charge.reshape(-1)[li]
    h = _images_hist(pos, q, box.box_nm[0])
    counts = np.rint(f * 8).astype(np.int64)
    mism = np.abs(counts - h)
    print(f"histogram mismatches: {np.count_nonzero(mism)} of {mism.size}")
    assert mism.max() <= 1 and np.count_nonzero(mism) <= 0.002 * mism.size


def test_histogram_counted_in_the_pair_pass_matches_brute_force():
    """In learned runs the histograms are counted inside the pair pass of the evaluation before a
    feature refresh.  With one feature refresh per iteration and one iteration, the features are
    those counted while evaluating the starting state."""
    box = L.gen_em_system(700, 1001, 0.02)
    start = L.Batch([box], L.default_params())
    start.minimize(100.0, 0, 0.01)                   # evaluates the starting state only
    st = start.state(0)
    b = L.Batch([box], L.default_params())
    m = L.Mlp(seed=3)                                # scales not all equal: proposals are made
    b.minimize_learned(m, 100.0, 1, 0.01, feature_interval=1)
    f = b.features(0)[:, 4:]
    li = box.layout_index()
    comp = li // 32
    coord = st["particle_coord"].reshape(-1, 3)
    pos = np.mod((st["compound_origo"][comp].astype(np.int64) * 100000000 + coord[li]).astype(np.float64) * 1e-8,
                 box.box_nm[0])
    h = _images_hist(pos, box.charge.reshape(-1)[li], box.box_nm[0])
    counts = np.rint(f * 8).astype(np.int64)
    mism = np.abs(counts - h)
    print(f"histogram mismatches: {np.count_nonzero(mism)} of {mism.size}, total count {h.sum()}")
    assert h.sum() > 0
    assert mism.max() <= 1 and np.count_nonzero(mism) <= 0.002 * mism.size


def test_learned_minimisation_converges_with_fallback():
    boxes = [L.gen_em_system(n, s, 0.02) for n, s in ((300, 1010), (900, 1011), (1700, 1012))]
    p = L.default_params()
    plain = L.Batch(boxes, p)
    rp, wp = plain.minimize(100.0, 5000, 0.01)
    learned = L.Batch(boxes, p)
    m = L.Mlp(seed=3)   # untrained: s ~ 2 everywhere; the plain fallback keeps it correct
    rl, wl = learned.minimize_learned(m, 100.0, 5000, 0.01, feature_interval=4)
    cnt = learned.counters()
    for k in range(len(boxes)):
        print(f"system {k}: plain {rp[k].iterations} it E {rp[k].final_energy:.6e} | learned {rl[k].iterations} it "
              f"{cnt[k, 0]} evals {cnt[k, 1]} learned-accepted E {rl[k].final_energy:.6e}")
        assert rl[k].converged == 1 and rl[k].final_fmax < 100.0
        assert cnt[k, 0] >= rl[k].iterations
        assert rl[k].final_energy == pytest.approx(rp[k].final_energy, rel=1e-2)


def test_recording_and_training_set():
    boxes = [L.gen_em_system(n, s, 0.02) for n, s in ((300, 2000), (600, 2001))]
    b = L.Batch(boxes, L.default_params())
    states = b.record(100.0, 400, 0.01, atom_stride=8, max_states=500)
    x, t = b.training_set(lookahead=1)
    print(f"{states} states, {len(t)} samples, target mean {t.mean():.3f}")
    assert states > 100 and len(t) > 1000
    assert np.all((t >= 0.05) & (t <= 3.95))
    assert np.all(np.isfinite(x))
    assert np.all(np.abs(x[:, :3]) <= 1.0 + 1e-6) and np.all(x[:, 3] <= 1.0 + 1e-6)
    m = L.Mlp(seed=9)
    losses, _ = m.train(x, t, epochs=3, batch=512, lr=1e-3, seed=4)
    assert losses[-1] < losses[0]


def test_constant_scale_model_reproduces_plain_minimiser():
    """A model whose output is the same scale for every atom proposes the plain step's direction
    with the plain step's largest displacement, so the learned path takes the plain step itself
    (one evaluation per iteration, no fallback): the run is the plain minimiser's, bit for bit,
    while the histogram and inference passes still run."""
    boxes = [L.gen_em_system(n, s, 0.02) for n, s in ((400, 1020), (1100, 1021))]
    p = L.default_params()
    m = L.Mlp(seed=0)
    m.set_params(np.zeros(m.n_params, np.float32))      # 4 sigmoid(0) = 2 for every atom
    for iters, tol in ((25, 1e-3), (5000, 100.0)):
        plain = L.Batch(boxes, p)
        rp, _ = plain.minimize(tol, iters, 0.01)
        learned = L.Batch(boxes, p)
        rl, _ = learned.minimize_learned(m, tol, iters, 0.01, feature_interval=4, grow_on_learned=True)
        cnt_p, cnt_l = plain.counters(), learned.counters()
        ph = learned.phase_times()
        for k in range(len(boxes)):
            print(f"max_iter {iters} system {k}: plain {rp[k].iterations} it {rp[k].accepted} acc E {rp[k].final_energy:.10e} | "
                  f"constant-scale {rl[k].iterations} it {rl[k].accepted} acc {cnt_l[k, 1]} learned E {rl[k].final_energy:.10e}")
            assert rl[k].iterations == rp[k].iterations
            assert rl[k].accepted == rp[k].accepted
            assert rl[k].final_energy == rp[k].final_energy
            assert rl[k].converged == rp[k].converged
            assert cnt_l[k, 0] == cnt_p[k, 0]        # force evaluations: no fallback evaluations
            assert cnt_l[k, 1] == 0                  # no learned proposal taken
            assert ph[k, 6] > 0.0                    # inference ran
            if iters == 5000:
                assert rl[k].converged == 1
