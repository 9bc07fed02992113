"""lima_mlp_forward_noalloc, the allocation-free variant of lima_mlp_forward: it stages features and
step scales in the library's fixed per-GPU array instead of allocating.  Its scales equal the
allocating entry point's bit for bit (the same kernel on the same rows), on consecutive calls whose
sizes shrink (the array keeps the larger call's rows, which the smaller call must not read), and it
rejects more rows than the array holds.
"""
import numpy as np
import pytest

import lima_amd as L

pytestmark = pytest.mark.gpu


def _features(n, seed):
    """Rows shaped like the minimiser's features: a unit force direction, a log-magnitude term and
    28 neighbour counts / 8."""
    rng = np.random.default_rng(seed)
    x = np.zeros((n, 32), np.float32)
    f = rng.normal(size=(n, 3))
    f /= np.abs(f).max()
    x[:, :3] = f
    x[:, 3] = 1.0 + np.log10(np.maximum(np.linalg.norm(f, axis=1), 1e-6)) / 6.0
    x[:, 4:] = rng.poisson(3.0, size=(n, 28)) / 8.0
    return x


def test_noalloc_forward_equals_allocating_forward():
    m = L.Mlp(seed=11)
    sizes = [(L.Mlp.NOALLOC_ROWS, 1), (1237, 2), (5, 3)]     # full array, then smaller calls
    for n, seed in sizes:
        x = _features(n, seed)
        ref = m.forward(x)
        got = m.forward_noalloc(x)
        print(f"{n} rows: scales in [{got.min():.4f}, {got.max():.4f}], identical: {np.array_equal(got, ref)}", flush=True)
        assert got.shape == (n,) and np.all(np.isfinite(got))
        assert np.array_equal(got, ref)
    # other parameters, same array: still the allocating entry point's results
    rng = np.random.default_rng(5)
    m.set_params((m.params() + rng.normal(scale=0.05, size=m.n_params)).astype(np.float32))
    x = _features(3000, 4)
    assert np.array_equal(m.forward_noalloc(x), m.forward(x))


def test_noalloc_forward_rejects_rows_beyond_the_array():
    m = L.Mlp(seed=3)
    x = np.zeros((L.Mlp.NOALLOC_ROWS + 1, 32), np.float32)
    with pytest.raises(L.LimaError, match="exceed the static staging array"):
        m.forward_noalloc(x)
    # the refused call left the entry point usable
    x = _features(64, 9)
    assert np.array_equal(m.forward_noalloc(x), m.forward(x))
