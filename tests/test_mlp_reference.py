"""CPU pin of the restated step-scale model: its backward pass against central finite differences
of the unrounded FP64 loss."""
import numpy as np

from tests.mlp_ref import grad_ref, sample_problem, O_B3, O_W3, O_B1, O_B2


def test_gradient_formula_matches_finite_differences():
    """CPU pin of the restated backward pass (no rounding, FP64)."""
    rng = np.random.default_rng(0)
    p = rng.normal(0, 0.2, O_B3 + 1)
    x, t = sample_problem(40)
    g, _ = grad_ref(p, x, t, rounded=False)
    for k in list(rng.choice(O_B3 + 1, 60, replace=False)) + [O_B3, O_W3 + 5, O_B1 + 7, O_B2 + 3]:
        pp, pm = p.copy(), p.copy()
        pp[k] += 1e-6
        pm[k] -= 1e-6
        fd = (grad_ref(pp, x, t, False)[1] - grad_ref(pm, x, t, False)[1]) / 2e-6
        assert abs(fd - g[k]) < 1e-5 * max(1.0, abs(g[k])), (k, fd, g[k])
