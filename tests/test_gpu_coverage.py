"""GPU parity for the paths the first parity file leaves open: the engine's steepest descent
against the oracle's, trajectories of LIMA's single-site tinymols, data-log frames against the
oracle's state at the logged steps, the 64-bit list path of boxes wider than ~9.4 nm, the
headline chain + water box and the 1M-atom water box against the oracle, and the rollback that
undoes steps after a pair-list overflow found late.

Tolerances as in test_gpu_parity.py: relative RMS force error <= 1e-5, trajectories within 1e-5 nm
RMSD, neighbour lists bit-exact.
"""
import numpy as np
import pytest

import lima_amd as L
from oracle.oracle import Oracle
from tests.helpers import rel_rms, layout_forces, layout_pot, unwrapped_positions

pytestmark = pytest.mark.gpu

FORCE_TOL = 1e-5


def _velocities(box, state):
    li = box.layout_index()
    v = np.concatenate([state["particle_velocity"].reshape(-1, 3), state["tinymol_velocity"][:box.n_tinymols]])
    return v[li].astype(np.float64)


def _oracle_velocities(box, orc):
    return _velocities(box, orc.state())


def test_minimize_follows_oracle():
    """lima_minimize (Programs::EnergyMinimize's entry, steepest descent with FP64 acceptance) and the
    oracle's minimiser take the same accept / refuse decisions over 30 iterations from a clashing
    start, and end at the same energy and positions."""
    box = L.gen_em_system(600, 1004, 0.02)
    p = L.default_params(dt_fs=1.0)
    eng = L.Engine(box, p)
    rg = eng.minimize(1e-6, 30, 0.01)
    orc = Oracle(box, p)
    ro = orc.minimize(tol=1e-6, max_iter=30, h0=0.01)
    print(f"GPU: {rg.iterations} it, {rg.accepted} accepted, E {rg.final_energy:.6e}, fmax {rg.final_fmax:.2f}; "
          f"oracle: {ro.iterations} it, {ro.accepted} accepted, E {ro.final_energy:.6e}, fmax {ro.final_fmax:.2f}")
    assert rg.iterations == ro.iterations == 30
    assert rg.accepted == ro.accepted and 5 <= ro.accepted < 30
    assert abs(rg.final_energy - ro.final_energy) <= 1e-6 * abs(ro.final_energy)
    d = (unwrapped_positions(box, eng.state()) - orc.positions()) * 1e-8
    rmsd = float(np.sqrt((d ** 2).sum(1).mean()))
    print("positions RMSD nm", rmsd)
    assert rmsd < 1e-5
    # the state left behind is the minimised one, with forces evaluated there
    f_ref, _, e_ref = Oracle(L.StateBox(box, eng.state()), p).forces()
    assert rel_rms(layout_forces(box, eng.state()), f_ref) <= FORCE_TOL


@pytest.mark.parametrize("name", ["lj_tinymol_fluid", "water_with_tinymols"])
def test_tinymol_trajectory_matches_oracle(name):
    """LIMA's single-site LJ solvent (TinymolIntegrationLoggingAndTransferout,
    EngineKernels.cuh:591-713): 20 velocity-Verlet steps of a pure tinymol fluid (2,000 argon-like
    particles) and of 700 waters with 20 tinymols; positions and velocities of every particle
    against the oracle."""
    if name == "lj_tinymol_fluid":
        box = L.gen_large_water(0, 2000, 0.34, 0.99, 39.948, 7, 120.0)
        p = L.default_params(dt_fs=2.0, enable_electrostatics=0)
    else:
        box = L.gen_large_water(700, 20, 0.3, 0.5, 39.948, 5, 300.0)
        p = L.default_params(dt_fs=2.0)
    assert box.n_tinymols > 0
    eng = L.Engine(box, p)
    orc = Oracle(box, p)
    eng.step(20)
    orc.step(20)
    st = eng.state()
    d = (unwrapped_positions(box, st) - orc.positions()) * 1e-8
    nt = box.n_tinymols
    rmsd = float(np.sqrt((d ** 2).sum(1).mean()))
    rmsd_t = float(np.sqrt((d[-nt:] ** 2).sum(1).mean()))
    vg, vo = _velocities(box, st), _oracle_velocities(box, orc)
    verr = float(np.sqrt(((vg - vo) ** 2).sum() / (vo ** 2).sum()))
    print(f"{name}: {nt} tinymols, 20-step RMSD {rmsd:.3e} nm (tinymols {rmsd_t:.3e}), velocity rel RMS {verr:.3e}")
    assert rmsd < 1e-5 and rmsd_t < 1e-5
    assert verr < 1e-4


def test_log_frames_match_oracle():
    """Data-log frames (LogCompoundData / LogSolventData, EngineUtils.cuh:143-173) at logged steps
    0, 5, 10, 15 against the oracle stepped alongside: the frame's force, potE and energy totals
    are those of the oracle's evaluation at that step, its positions the oracle's after the step."""
    box = L.gen_large_water(700, 20, 0.3, 0.5, 39.948, 5, 300.0)
    p = L.default_params(dt_fs=1.0, data_logging_interval=5, log_capacity=8)
    eng = L.Engine(box, p)
    orc = Oracle(box, p)
    eng.step(16)
    log = eng.read_log()
    assert list(log["step"]) == [0, 5, 10, 15]
    li = box.layout_index()
    L_nm = np.asarray(box.box_nm)
    f = 0
    for k in range(16):
        if k % 5 == 0:
            f_ref, pot_ref, e_ref = orc.forces()
        orc.step(1)
        if k % 5 != 0:
            continue
        err = rel_rms(log["force"][f][li], f_ref[li])
        x_ref = orc.positions() * 1e-8
        d = log["position"][f][li].astype(np.float64) - x_ref
        d -= L_nm * np.round(d / L_nm)
        print(f"frame {f} (step {k}): force rel RMS {err:.3e}, max |dx| {np.abs(d).max():.3e} nm, "
              f"E {log['energy'][f].potential:.6e} / {e_ref.potential:.6e}")
        assert err <= FORCE_TOL
        assert np.abs(d).max() < 2e-5
        assert np.allclose(log["pot_energy"][f][li], pot_ref[li], rtol=1e-4, atol=1e-4 * np.abs(pot_ref[li]).max())
        assert abs(log["energy"][f].potential - e_ref.potential) <= 1e-6 * abs(e_ref.potential)
        f += 1
    assert f == 4


def _sorted_entries(atoms, off, ent, msk):
    """Cluster list as one sorted int64 array of (cluster, j particle, shift code, mask) keys."""
    counts = np.diff(off)
    cl = np.repeat(np.arange(len(counts), dtype=np.int64), counts)
    key = (cl << 40) | ((ent.astype(np.int64) & 0xFFFFFF) << 16) | ((ent.astype(np.int64) >> 24) << 8) | msk.astype(np.int64)
    return np.sort(key)


def test_wide_box_list_and_forces_match_oracle():
    """An 11 nm water box (132,000 atoms): coordinate differences no longer fit 32 bits, so the list
    build takes its 64-bit path (k_nlist<false>).  The list is bit-exact against the oracle and the
    forces match it."""
    box = L.gen_large_water(44000, 0, 0.3, 0.5, 39.948, 11, 300.0)
    assert box.box_nm[0] > 10.9
    p = L.default_params(dt_fs=2.0)
    eng = L.Engine(box, p)
    g = eng.neighbor_list()
    orc = Oracle(box, p)
    o = orc.neighbor_list()
    assert np.array_equal(g[0], o[0]), "cluster membership differs"
    assert np.array_equal(g[1], o[1]), "entry counts per cluster differ"
    assert np.array_equal(_sorted_entries(*g), _sorted_entries(*o))
    eng.compute_forces()
    f_ref, _, _ = orc.forces()
    err = rel_rms(layout_forces(box, eng.state()), f_ref)
    print(f"wide box {box.box_nm[0]:.2f} nm, {box.n_particles} atoms: {len(g[2])} entries, force rel RMS {err:.3e}")
    assert err <= FORCE_TOL


def test_rollback_after_pair_list_overflow_is_exact():
    """Steps issued past a list rebuild whose pair list overflowed are undone: the engine restores
    the state of that rebuild, grows the capacity and runs them again.  An engine started with the
    exact capacity its first list needs overflows within a few rebuilds of hot water; its
    trajectory (step graph replay, which the reallocation forces to capture again) equals, bit for
    bit, that of an engine whose capacity never overflows."""
    box = L.gen_water_lattice(10, 3.104, 0, 600.0)
    probe = L.Engine(box, L.default_params(dt_fs=2.0, nlist_interval=10))
    need = probe.status().list_max_entries     # slots the fullest cluster of the first list uses
    del probe
    ref = L.Engine(box, L.default_params(dt_fs=2.0, nlist_interval=10, nlist_capacity=4 * need))
    tight = L.Engine(box, L.default_params(dt_fs=2.0, nlist_interval=10, nlist_capacity=need))
    assert tight.status().list_capacity == need
    steps = 0
    while tight.status().capacity_rollbacks == 0 and steps < 400:
        tight.step(20)
        ref.step(20)
        steps += 20
    st_t, st_r = tight.state(), ref.state()
    s_t, s_r = tight.status(), ref.status()
    print(f"capacity {need} entries: {s_t.capacity_rollbacks} rollbacks within {steps} steps; reference rollbacks "
          f"{s_r.capacity_rollbacks}")
    assert s_t.capacity_rollbacks >= 1 and s_r.capacity_rollbacks == 0
    assert s_t.current_step == s_r.current_step == steps and s_t.n_list_rebuilds == s_r.n_list_rebuilds
    assert np.array_equal(unwrapped_positions(box, st_t), unwrapped_positions(box, st_r))
    assert np.array_equal(_velocities(box, st_t), _velocities(box, st_r))


def test_chain_water_c1_forces_match_oracle():
    """The headline box: 128 chains of 50 atoms in two leaflets plus 12,000 waters (42,400 atoms) in
    8 x 8 x 7 nm, PME 72 x 72 x 60."""
    box = L.gen_chain_water()
    assert box.n_particles == 42400
    p = L.default_params(dt_fs=2.0)
    eng = L.Engine(box, p)
    e_gpu = eng.compute_forces()
    st = eng.state()
    f_ref, pot_ref, e_ref = Oracle(box, p).forces()
    err = rel_rms(layout_forces(box, st), f_ref)
    print(f"chain + water: N={box.n_particles} rel RMS force error {err:.3e}, E {e_gpu.potential:.6e} / {e_ref.potential:.6e}")
    assert err <= FORCE_TOL
    assert abs(e_gpu.potential - e_ref.potential) <= 1e-6 * abs(e_ref.potential)


def test_water_1m_forces_match_oracle():
    """The 1M-atom box: 333,333 waters and one neutral LJ tinymol in a 21.5 nm cube, PME 180^3 (the
    64-bit list path and the largest PME grid)."""
    box = L.gen_large_water()
    assert box.n_particles == 1000000
    p = L.default_params(dt_fs=2.0)
    eng = L.Engine(box, p)
    e_gpu = eng.compute_forces()
    f_gpu = layout_forces(box, eng.state())
    del eng
    f_ref, _, e_ref = Oracle(box, p).forces()
    err = rel_rms(f_gpu, f_ref)
    print(f"water 1M: rel RMS force error {err:.3e}, E {e_gpu.potential:.6e} / {e_ref.potential:.6e}", flush=True)
    assert err <= FORCE_TOL
    assert abs(e_gpu.potential - e_ref.potential) <= 1e-6 * abs(e_ref.potential)


def test_list_stays_complete_when_atoms_move_fast():
    """Water at 900 K with lists meant to live 40 steps: intervals whose two largest displacements
    since the build exceed the skin are undone and run again with shorter lists, so every step
    evaluates every pair inside the cutoff.  The pair kernel's count of in-cutoff pairs at each
    checked step lies between the oracle's brute-force counts at cutoff -+ 2e-6 nm (the kernel's
    FP32 distances of pairs that close to the cutoff may round either way; the band holds ~5 of
    the ~1.08M pairs), and no kept interval breaks the bound."""
    box = L.gen_water_lattice(10, 3.104, 0, 900.0)
    p = L.default_params(dt_fs=2.0, nlist_interval=40, nlist_skin_nm=0.15, energy_interval=1)
    eng = L.Engine(box, p)
    rc = float(p.cutoff_nm)
    checked = 0
    for k in range(120):
        st = eng.state()
        eng.step(1)
        if k % 12 != 11:
            continue
        n_gpu = eng.pair_count()          # pairs inside the cutoff evaluated at the positions st
        orc = Oracle(L.StateBox(box, st), p)
        lo, hi = orc.pair_count(rc - 2e-6), orc.pair_count(rc + 2e-6)
        assert lo <= n_gpu <= hi, (k, lo, n_gpu, hi)
        checked += 1
    s = eng.status()
    print(f"{checked} steps checked; list interval now {s.list_interval} (max 40), {s.list_stale_rollbacks} stale "
          f"intervals undone, {s.list_violations} kept intervals over the bound, {s.n_list_rebuilds} rebuilds", flush=True)
    assert s.list_stale_rollbacks >= 1 and s.list_interval < 40
    assert s.list_violations == 0


def test_coarse_pme_grid_matches_oracle():
    """A coarse PME grid (0.3 nm spacing: 12^3 on the 3.104 nm water box, tiles of 6 grid points
    holding ~375 charges each, where a 0.12 nm grid's tiles hold ~50): forces and energy of the start
    state and a 10-step trajectory against the oracle, which derives the same grid."""
    box = L.gen_water_lattice(10, 3.104, 0, 300.0)
    p = L.default_params(dt_fs=1.0, pme_spacing_nm=0.3)
    eng = L.Engine(box, p)
    orc = Oracle(box, p)
    assert list(orc.pme_dims()) == [12, 12, 12]
    e_gpu = eng.compute_forces()
    f_ref, _, e_ref = orc.forces()
    err = rel_rms(layout_forces(box, eng.state()), f_ref)
    eng.step(10)
    orc.step(10)
    d = (unwrapped_positions(box, eng.state()) - orc.positions()) * 1e-8
    rmsd = float(np.sqrt((d ** 2).sum(1).mean()))
    print(f"PME 12^3: force rel RMS {err:.3e}, E {e_gpu.potential:.6e} / {e_ref.potential:.6e}, 10-step RMSD {rmsd:.3e} nm",
          flush=True)
    assert err <= FORCE_TOL
    assert abs(e_gpu.potential - e_ref.potential) <= 1e-6 * abs(e_ref.potential)
    assert rmsd < 1e-5
