"""GPU parity: the HIP path against the FP64 oracle on the same seeded boxes.

Tolerances: relative RMS force error <= 1e-5 (north-star bar for FP32 forces with the
oracle in FP64), energies relative 1e-5 of the summed magnitude, neighbour lists bit-exact.
"""
import numpy as np
import pytest

import lima_amd as L
from oracle.oracle import Oracle
from tests.helpers import rel_rms, layout_forces, layout_pot, canonical_entries, unwrapped_positions, min_image, max_list_age_error

pytestmark = pytest.mark.gpu

FORCE_TOL = 1e-5


def _boxes():
    return {
        "water_c0": (lambda: L.gen_water_lattice(10, 3.104, 0, 300.0), 1.0),
        "em_300": (lambda: L.gen_em_system(300, 1000, 0.02), 1.0),
        "em_1200": (lambda: L.gen_em_system(1200, 1003, 0.0), 1.0),
        "tinymol_water": (lambda: L.gen_large_water(700, 20, 0.3, 0.5, 39.948, 5, 300.0), 2.0),
    }


@pytest.mark.parametrize("name", list(_boxes()))
def test_forces_match_oracle(name):
    make, dt = _boxes()[name]
    box = make()
    p = L.default_params(dt_fs=dt)
    eng = L.Engine(box, p)
    e_gpu = eng.compute_forces()
    st = eng.state()
    f_gpu = layout_forces(box, st)
    f_ref, pot_ref, e_ref = Oracle(box, p).forces()
    err = rel_rms(f_gpu, f_ref)
    print(f"{name}: N={box.n_particles} rel RMS force error {err:.3e}")
    assert err <= FORCE_TOL, err
    scale = abs(e_ref.lj) + abs(e_ref.coulomb_sr) + abs(e_ref.bonded) + abs(e_ref.pme_recip)
    for k in ("lj", "coulomb_sr", "bonded", "pme_recip"):
        assert abs(getattr(e_gpu, k) - getattr(e_ref, k)) <= 1e-5 * scale + 1e-3, (k, getattr(e_gpu, k), getattr(e_ref, k))
    pot_gpu = layout_pot(box, st)
    li = box.layout_index()
    assert np.allclose(pot_gpu[li], pot_ref[li], rtol=1e-4, atol=1e-4 * np.abs(pot_ref[li]).max())


@pytest.mark.parametrize("grid", [None, (100, 100, 16), (27, 27, 4)], ids=["tiled_27", "tiled_100x100x16", "untiled"])
def test_pme_paths_match_oracle(grid):
    """Both spreading paths: tiled (per-tile LDS blocks; the default 27^3 grid of the water box and
    a 100 x 100 x 16 grid with different tile edges per axis) and untiled (no tile edge in 5..12
    divides 4, so charges are spread with global atomics).  Reciprocal energy and forces against
    the oracle."""
    box = L.gen_water_lattice(10, 3.104, 0, 300.0)
    p = L.default_params(dt_fs=1.0) if grid is None else L.default_params(dt_fs=1.0, pme_grid=grid)
    eng = L.Engine(box, p)
    e_gpu = eng.compute_forces()
    f_gpu = layout_forces(box, eng.state())
    f_ref, _, e_ref = Oracle(box, p).forces()
    err = rel_rms(f_gpu, f_ref)
    print(f"grid {grid}: rel RMS force error {err:.3e}, pme_recip gpu {e_gpu.pme_recip:.2f} oracle {e_ref.pme_recip:.2f}")
    assert err <= FORCE_TOL
    assert abs(e_gpu.pme_recip - e_ref.pme_recip) <= 1e-5 * abs(e_ref.pme_recip)


@pytest.mark.parametrize("name", ["water_c0", "em_300"])
def test_neighbor_list_bit_exact(name):
    make, dt = _boxes()[name]
    box = make()
    p = L.default_params(dt_fs=dt)
    eng = L.Engine(box, p)
    g = eng.neighbor_list()
    o = Oracle(box, p).neighbor_list()
    assert np.array_equal(g[0], o[0]), "cluster membership differs"
    assert np.array_equal(g[1], o[1]), "entry counts per cluster differ"
    assert canonical_entries(*g) == canonical_entries(*o)


def test_md_trajectory_matches_oracle():
    box = L.gen_water_lattice(10, 3.104, 0, 300.0)
    p = L.default_params(dt_fs=1.0, energy_interval=1)   # energies on every step: the last one is read below
    eng = L.Engine(box, p)
    orc = Oracle(box, p)
    eng.step(20)
    orc.step(20)
    xg = unwrapped_positions(box, eng.state())
    xo = orc.positions()
    d = (xg - xo) * 1e-8
    rmsd = float(np.sqrt((d ** 2).sum(1).mean()))
    print("20-step RMSD nm", rmsd)
    assert rmsd < 1e-5
    eg, eo = eng.energy(), orc.forces()[2]
    assert eg.step == 19
    assert abs(eg.kinetic - eo.kinetic) < 1e-3 * abs(eo.kinetic) + 1.0


def test_every_bonded_term_type_matches_oracle():
    """Bonds, angles with Urey-Bradley, dihedrals, impropers and 1-4 pairs on a 40-particle helix
    (two compounds, 32 + 8) shifted across the periodic boundary, plus a second copy: the
    atom-centric bonded pass against the oracle, forces and energy (the generated workloads carry
    no impropers or 1-4 pairs)."""
    from lima_amd.boxdesc import CustomBox
    n = 40
    k = np.arange(n)
    helix = np.stack([0.3 * np.cos(1.7 * k), 0.3 * np.sin(1.7 * k), 0.045 * k], 1)
    pos = np.concatenate([helix + [2.8, 2.9, 1.2], helix + [1.0, 1.2, 0.3]])
    mol = np.repeat([0, 1], n)
    terms = dict(bonds=[], angles=[], dihedrals=[], impropers=[], pairs=[])
    for o in (0, n):
        terms["bonds"] += [(o + i, o + i + 1, 0.15, 2e5) for i in range(n - 1)]
        terms["angles"] += [(o + i, o + i + 1, o + i + 2, 111.0, 400.0, 0.25, 2e4) for i in range(n - 2)]
        terms["dihedrals"] += [(o + i, o + i + 1, o + i + 2, o + i + 3, 15.0, 3.0, 3) for i in range(n - 3)]
        terms["impropers"] += [(o + i + 1, o + i, o + i + 2, o + i + 3, 5.0, 150.0) for i in range(0, n - 3, 5)]
        terms["pairs"] += [(o + i, o + i + 3, 0.32, 0.6) for i in range(n - 3)]
    q = np.where(k % 2 == 0, 0.3, -0.3)
    box = CustomBox([3.1] * 3, pos, np.concatenate([q, q]), np.full(2 * n, 0.33), np.full(2 * n, 0.5),
                    np.full(2 * n, 12.0), molecule=mol, **terms)
    p = L.default_params(dt_fs=1.0)
    eng = L.Engine(box, p)
    e_gpu = eng.compute_forces()
    f_gpu = layout_forces(box, eng.state())
    f_ref, pot_ref, e_ref = Oracle(box, p).forces()
    err = rel_rms(f_gpu, f_ref)
    print(f"bonded chain: rel RMS force error {err:.3e}, bonded E gpu {e_gpu.bonded:.3f} oracle {e_ref.bonded:.3f}")
    assert err <= FORCE_TOL
    assert abs(e_gpu.bonded - e_ref.bonded) <= 1e-5 * abs(e_ref.bonded)
    pot_gpu = layout_pot(box, eng.state())
    li = box.layout_index()
    assert np.allclose(pot_gpu[li], pot_ref[li], rtol=1e-4, atol=1e-4 * np.abs(pot_ref[li]).max())


def _velocities(box, state):
    li = box.layout_index()
    v = np.concatenate([state["particle_velocity"].reshape(-1, 3), state["tinymol_velocity"][:box.n_tinymols]])
    return v[li].astype(np.float64)


def test_thermostat_trajectory_matches_oracle():
    """Thermostat on, measured every 5 steps: temperatures, positions and velocities follow the
    oracle.  Over 20 steps the scale (1 + 0.0002 per step) changes velocities by ~0.3%, far above
    the tolerance, so a thermostat that is not applied, or applied at the wrong steps, fails."""
    box = L.gen_water_lattice(10, 3.104, 0, 300.0)
    p = L.default_params(dt_fs=1.0, apply_thermostat=1, steps_per_temperature_measurement=5, data_logging_interval=1)
    eng = L.Engine(box, p)
    orc = Oracle(box, p)
    assert np.isnan(eng.status().current_temperature)
    eng.step(20)
    orc.step(20)
    scale_o, T_o = orc.thermostat()
    T_g = eng.status().current_temperature
    print(f"T gpu {T_g:.4f} oracle {T_o:.4f}, oracle scale {scale_o:.7f}")
    assert scale_o != 1.0
    assert T_g == pytest.approx(T_o, rel=1e-5)
    st = eng.state()
    d = (unwrapped_positions(box, st) - orc.positions()) * 1e-8
    assert float(np.sqrt((d ** 2).sum(1).mean())) < 1e-5
    vg, vo = _velocities(box, st), _velocities(box, orc.state())
    err = float(np.sqrt(((vg - vo) ** 2).sum() / (vo ** 2).sum()))
    print("velocity rel RMS", err)
    assert err < 2e-4


@pytest.mark.parametrize("name", ["chain_water_c1", "water_c0_hot_2fs"])
def test_force_error_at_max_list_age(name):
    """Lists hold the pairs within r_c + skin of each cluster's box at the build and live until the
    two largest displacements since then could add up to the skin (water H reach ~7 nm/ps once these
    boxes heat to 500-800 K, so the engine shortens the interval and reruns a list interval that
    went past it).  The step that uses the oldest list of an interval matches the FP64 oracle at the
    same positions within the 1e-5 force bar (measured ~6e-7 on the chain + water box, ~2e-6 on the
    hot water box at 2 fs)."""
    if name == "chain_water_c1":
        box = L.gen_chain_water()
        p = L.default_params(dt_fs=2.0)
        eng = L.Engine(box, p)
        eng.minimize(1000.0, 400, 0.01)
        eng.step(300)
    else:
        box = L.gen_water_lattice(10, 3.104, 0, 300.0)
        p = L.default_params(dt_fs=2.0)
        eng = L.Engine(box, p)
        eng.step(400)
    errs = [max_list_age_error(eng, box, p) for _ in range(2)]
    st = eng.status()
    print(name, "skin", round(p.nlist_skin_nm, 3), "interval", st.list_interval, "of at most", p.nlist_interval,
          f"({st.list_stale_rollbacks} intervals rerun)", "force error at the oldest list", ", ".join(f"{e:.3e}" for e in errs))
    assert st.list_violations == 0
    assert max(errs) <= FORCE_TOL


def test_graph_replay_matches_eager_launches():
    """The MD step replayed as a HIP graph gives the trajectory of the eager launches, bit for bit:
    every force buffer is written by plain stores of the one wave that owns the slot and summed in a
    fixed order, so neither launch mode nor the PME stream changes a rounding.  The run crosses
    three list rebuilds."""
    box = L.gen_em_system(1200, 1003, 0.0)
    runs = []
    for mode in (0, 1, 3):                       # graph replay, eager launches, eager on one stream
        eng = L.Engine(box, L.default_params(dt_fs=1.0, launch_mode=mode, nlist_interval=10))
        eng.step(35)                             # crosses three list rebuilds
        st = eng.state()
        runs.append((unwrapped_positions(box, st), _velocities(box, st), eng.compute_forces().potential, eng.status()))
    (xg, vg, eg, sg) = runs[0]
    for (xe, ve, ee, se) in runs[1:]:
        assert sg.current_step == se.current_step == 35 and sg.n_list_rebuilds == se.n_list_rebuilds
        print("max |dx| Coord units", np.abs(xg - xe).max(), "E graph", eg, "eager", ee)
        assert np.array_equal(xg, xe)
        assert np.array_equal(vg, ve)
        assert eg == ee


def test_two_runs_are_bit_identical():
    """1,000 steps of the same box twice give bit-identical positions and velocities (no float
    atomics in any force sum)."""
    box = L.gen_water_lattice(10, 3.104, 0, 300.0)
    out = []
    for _ in range(2):
        eng = L.Engine(box, L.default_params(dt_fs=1.0))
        eng.step(1000)
        st = eng.state()
        out.append((unwrapped_positions(box, st), _velocities(box, st)))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])


def test_minimize_decreases_energy():
    box = L.gen_em_system(300, 1000, 0.02)
    p = L.default_params(dt_fs=1.0)
    eng = L.Engine(box, p)
    e0 = eng.compute_forces().potential
    r = eng.minimize(100.0, 2000, 0.01)
    print("iters", r.iterations, "fmax", r.final_fmax, "E", r.final_energy, "E0", e0)
    assert r.final_energy < e0
    assert r.converged == 1
