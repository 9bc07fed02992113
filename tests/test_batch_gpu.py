"""GPU tests of the batched minimiser (lima_batch_*) through the C ABI, against the FP64 oracle and
against the single-system engine."""
import numpy as np
import pytest

import lima_amd as L
from oracle.oracle import Oracle
from tests.helpers import rel_rms, layout_forces

pytestmark = pytest.mark.gpu

# sizes cover boxes shorter than 2 rc (several images per pair) and longer ones
SIZES = [(300, 1000), (700, 1001), (1300, 1002), (2600, 1003)]


def em_boxes(specs):
    return [L.gen_em_system(n, seed, 0.02) for n, seed in specs]


def shape_tuning(threads, rb):
    t = L.BatchTuning()
    t.n_buckets = 1
    t.bucket_max_atoms[0] = 8192
    t.threads[0] = threads
    t.pair_unroll[0] = rb
    return t


@pytest.fixture(scope="module")
def batch4():
    boxes = em_boxes(SIZES)
    b = L.Batch(boxes, L.default_params())
    return boxes, b


def test_batch_initial_forces_match_oracle(batch4):
    boxes, b = batch4
    b.reset()
    res, _ = b.minimize(100.0, 0, 0.01)        # max_iter 0: evaluate the starting state only
    p = L.default_params()
    for k, box in enumerate(boxes):
        print(f"system {k}: {box.n_particles} atoms, box {box.box_nm[0]:.3f} nm", flush=True)
        f_ref, _, e_ref = Oracle(box, p).forces()
        err = rel_rms(layout_forces(box, b.state(k)), f_ref)
        e_rel = abs(res[k].final_energy - e_ref.potential) / abs(e_ref.potential)
        print(f"  force rel RMS {err:.2e}, potential rel {e_rel:.2e}", flush=True)
        assert err < 1e-5
        assert e_rel < 1e-6
        assert res[k].iterations == 0


@pytest.mark.parametrize("threads,rb", [(256, 1), (512, 4), (1024, 2)])
def test_batch_forces_independent_of_shape(batch4, threads, rb):
    boxes, b = batch4
    b.set_tuning(shape_tuning(1024, 1))
    b.reset()
    ref_res, _ = b.minimize(100.0, 0, 0.01)
    ref = [b.state(k)["particle_force"].copy() for k in range(len(boxes))]
    b.set_tuning(shape_tuning(threads, rb))
    b.reset()
    res, _ = b.minimize(100.0, 0, 0.01)
    p = L.default_params()
    for k in range(len(boxes)):
        f = b.state(k)["particle_force"]
        d = np.abs(f - ref[k]).max() / np.abs(ref[k]).max()
        f_ref, _, _ = Oracle(boxes[k], p).forces()
        err = rel_rms(layout_forces(boxes[k], b.state(k)), f_ref)
        print(f"system {k}: max |dF| / max |F| vs shape (1024, 1) {d:.2e}; rel RMS vs oracle {err:.2e}", flush=True)
        assert err < 1e-5
        # same pairs in the same order for every shape; only the instruction schedule of the
        # unrolled pair loop differs, which moves the last bits
        assert d < 1e-6
        assert res[k].final_energy == pytest.approx(ref_res[k].final_energy, rel=1e-9)


def test_batch_minimize_converges_like_the_engine(batch4):
    boxes, b = batch4
    b.set_tuning(shape_tuning(1024, 2))
    b.reset()
    res, wall = b.minimize(100.0, 5000, 0.01)
    p = L.default_params()
    for k, box in enumerate(boxes):
        e = L.Engine(box, p)
        e0 = e.compute_forces().potential
        r = e.minimize(100.0, 5000, 0.01)
        print(f"system {k}: batch {res[k].iterations} it E {res[k].final_energy:.6e} fmax {res[k].final_fmax:.1f} | "
              f"engine {r.iterations} it E {r.final_energy:.6e}", flush=True)
        assert res[k].converged == 1 and res[k].final_fmax < 100.0
        assert res[k].final_energy < e0
        # two FP32 paths take the same first steps and then drift apart (the accept / reject
        # sequence amplifies rounding; the engine's float atomics also vary run to run), ending in
        # neighbouring minima of the rough landscape
        assert res[k].final_energy == pytest.approx(r.final_energy, rel=1e-2)
    print(f"batch wall {wall:.3f}s", flush=True)


def test_batch_minimize_matches_oracle_small():
    box = L.gen_em_system(300, 1000, 0.02)
    b = L.Batch([box], L.default_params())
    res, _ = b.minimize(100.0, 3000, 0.01)
    o = Oracle(box, L.default_params())
    r = o.minimize(tol=100.0, max_iter=3000, h0=0.01)
    print(f"batch {res[0].iterations} it E {res[0].final_energy:.6e} | oracle {r.iterations} it E {r.final_energy:.6e}")
    assert res[0].converged == 1 and r.converged == 1
    assert res[0].final_energy == pytest.approx(r.final_energy, rel=1e-3)


@pytest.mark.parametrize("n,seed", [(300, 1000), (700, 1001)])
def test_batch_first_iterations_match_oracle(n, seed):
    """The first SD iterations follow the FP64 oracle's trajectory: same accept / reject sequence
    and the same energy after 30 iterations, to FP32 rounding of the forces and steps measured on
    the starting energy scale (the energy itself falls through zero on the way)."""
    box = L.gen_em_system(n, seed, 0.02)
    b = L.Batch([box], L.default_params())
    res, _ = b.minimize(100.0, 30, 0.01)
    o = Oracle(box, L.default_params())
    e0 = o.forces()[2].potential
    r = o.minimize(tol=100.0, max_iter=30, h0=0.01)
    print(f"n={n}: E0 {e0:.6e} | batch acc {res[0].accepted} E {res[0].final_energy:.8e} Fmax {res[0].final_fmax:.2f} | "
          f"oracle acc {r.accepted} E {r.final_energy:.8e} Fmax {r.final_fmax:.2f}")
    assert res[0].iterations == r.iterations == 30
    assert res[0].accepted == r.accepted
    assert abs(res[0].final_energy - r.final_energy) < 5e-6 * abs(e0)
    # Fmax is the force on one atom, often in a steep clash where a few Coord units of step
    # rounding move it by 1e-3
    assert res[0].final_fmax == pytest.approx(r.final_fmax, rel=1e-2)


def test_batch_is_deterministic(batch4):
    boxes, b = batch4
    b.set_tuning(shape_tuning(512, 2))
    out = []
    for _ in range(2):
        b.reset()
        res, _ = b.minimize(100.0, 200, 0.01)
        out.append(([r.final_energy for r in res], [r.iterations for r in res], b.state(0)["particle_coord"].copy()))
    assert out[0][0] == out[1][0] and out[0][1] == out[1][1]
    assert np.array_equal(out[0][2], out[1][2])


def test_batch_autotune_picks_valid_shapes(batch4):
    boxes, b = batch4
    t = b.autotune(3)
    for k in range(t.n_buckets):
        assert t.threads[k] in (256, 512, 1024) and t.pair_unroll[k] in (1, 2, 4)
    print(t.as_list())
    # tuning leaves the starting positions in place
    b.reset()
    res, _ = b.minimize(100.0, 0, 0.01)
    f_ref, _, _ = Oracle(boxes[0], L.default_params()).forces()
    assert rel_rms(layout_forces(boxes[0], b.state(0)), f_ref) < 1e-5


def test_batch_rejects_oversized_systems():
    box = L.gen_em_system(300, 1000, 0.02)
    p = L.default_params()
    p.pme_grid[0] = 40
    with pytest.raises(L.LimaError, match="PME grid"):
        L.Batch([box], p)
