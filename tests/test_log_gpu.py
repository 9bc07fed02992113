"""GPU data log (LIMA's DatabuffersDeviceController + offloadLoggingData) and the energy analysis
of a logged NVE run.

A logged frame must equal the engine state that lima_get_state reads after the same step (both
come from the same device arrays; the state itself is pinned against the FP64 oracle by the
parity tests): the step's forces and potE bit for bit, positions in LIMA's convention (origo +
Coord in FP32, tinymols wrapped) and |v| to one FP32 rounding (the device contracts the
multiply-adds to FMAs, as LIMA's build does; numpy rounds each operation).  Energies of a frame
equal lima_get_energy of that step.
"""
import numpy as np
import pytest

import lima_amd as L

pytestmark = pytest.mark.gpu


def _expected_frame(box, st):
    """LIMA buffer layout [n_compounds*32 + n_tinymols] from a state dict, in FP32 as LIMA computes
    GetAbsolutePositionNM (origo + Coord * 1e-8) and |v|."""
    nc, nt = box.n_compounds, box.n_tinymols
    P = nc * 32 + nt
    pos = np.zeros((P, 3), np.float32)
    pot = np.zeros(P, np.float32)
    spd = np.zeros(P, np.float32)
    frc = np.zeros((P, 3), np.float32)
    f32 = np.float32
    for c in range(nc):
        n = box.n_particles_per_compound[c]
        sl = slice(c * 32, c * 32 + n)
        pos[sl] = st["compound_origo"][c].astype(f32) + st["particle_coord"][c, :n].astype(f32) * f32(1e-8)
        pot[sl] = st["particle_pot_energy"][c, :n]
        frc[sl] = st["particle_force"][c, :n]
        v = st["particle_velocity"][c, :n].astype(f32)
        spd[sl] = np.sqrt((v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1]) + v[:, 2] * v[:, 2])
    if nt:
        sl = slice(nc * 32, nc * 32 + nt)
        pos[sl] = st["tinymol_block"][:nt].astype(f32) + st["tinymol_coord"][:nt].astype(f32) * f32(1e-8)
        pot[sl] = st["tinymol_pot_energy"][:nt]
        frc[sl] = st["tinymol_force"][:nt]
        v = st["tinymol_velocity"][:nt].astype(f32)
        spd[sl] = np.sqrt((v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1]) + v[:, 2] * v[:, 2])
    return pos, pot, spd, frc


@pytest.mark.parametrize("name", ["em_1200", "tinymol_water"])
def test_log_frames_equal_state_after_the_step(name):
    box = (L.gen_em_system(1200, 1003, 0.0) if name == "em_1200"
           else L.gen_large_water(700, 20, 0.3, 0.5, 39.948, 5, 300.0))
    eng = L.Engine(box, L.default_params(dt_fs=1.0, data_logging_interval=5, log_capacity=4))
    expected = []
    for k in range(12):                       # logged steps 0, 5, 10
        eng.step(1)
        if k % 5 == 0:
            expected.append((k, _expected_frame(box, eng.state()), eng.energy()))
    info = eng.log_info()
    assert (info.frames_logged, info.first_held, info.capacity, info.interval) == (3, 0, 4, 5)
    assert info.particles_per_frame == box.n_compounds * 32 + box.n_tinymols
    log = eng.read_log()
    assert list(log["step"]) == [0, 5, 10]
    for f, (step, (pos, pot, spd, frc), en) in enumerate(expected):
        print(name, "frame", f, "step", step, "max |dx|", np.abs(log["position"][f] - pos).max())
        assert np.all(np.abs(log["position"][f] - pos) <= np.spacing(np.abs(pos)))
        assert np.all(np.abs(log["speed"][f] - spd) <= 2 * np.spacing(spd))
        assert np.array_equal(log["pot_energy"][f], pot)
        assert np.array_equal(log["force"][f], frc)
        assert log["energy"][f].potential == en.potential and log["energy"][f].kinetic == en.kinetic
        assert log["energy"][f].potential_lima == en.potential_lima


def test_log_ring_and_graph_replay():
    """Logging inside multi-step calls (graph replay between the logged steps) keeps the eager
    trajectory; the ring holds the last log_capacity frames and refuses older ones."""
    box = L.gen_em_system(1200, 1003, 0.0)
    logs = []
    for mode in (0, 1):                       # graph replay, eager launches
        eng = L.Engine(box, L.default_params(dt_fs=1.0, data_logging_interval=5, log_capacity=3, launch_mode=mode,
                                            nlist_interval=10))
        eng.step(23)                          # logged steps 0, 5, 10, 15, 20; crosses two list rebuilds
        info = eng.log_info()
        assert (info.frames_logged, info.first_held) == (5, 2)
        with pytest.raises(L.LimaError):
            eng.read_log(1, 1)
        with pytest.raises(L.LimaError):
            eng.read_log(4, 2)
        logs.append(eng.read_log())
    g, e = logs
    assert list(g["step"]) == list(e["step"]) == [10, 15, 20]
    # force sums are in a fixed order, so both launch modes log the same bits
    assert np.array_equal(g["position"], e["position"])
    assert np.array_equal(g["force"], e["force"])
    for a, b in zip(g["energy"], e["energy"]):
        assert a.potential == b.potential and a.kinetic == b.kinetic


def test_nve_stability_by_energy_analysis():
    """LIMA judges MD runs by the energy analysis of the logged steps (TestUtils.h:293-320,
    loadAndRunBasicSimulation: variance coefficient <= 0.001 and |energy gradient| <= 1e-7 by
    default).  The 1000-water box at 300 K, dt = 1 fs, 300 NVE steps logged every 5 steps must pass
    that bar, and its standard total energy (potential + kinetic) must stay within 1 % of the
    kinetic energy of its starting value."""
    box = L.gen_water_lattice(10, 3.104, 0, 300.0)
    eng = L.Engine(box, L.default_params(dt_fs=1.0, data_logging_interval=5, log_capacity=64))
    eng.step(301)
    log = eng.read_log()
    assert len(log["step"]) == 61
    pot = np.array([en.potential_lima for en in log["energy"]])
    kin = np.array([en.kinetic for en in log["energy"]])
    tot = np.array([en.total for en in log["energy"]])
    r = L.analyze_energy(pot, kin)
    print(f"VC {r.variance_coefficient:.3e} gradient {r.energy_gradient:.3e} mean {r.mean_energy:.6e} "
          f"total drift {tot[-1] - tot[0]:.1f} J/mol, kinetic {kin.mean():.1f} J/mol, max |dE| {np.abs(tot - tot[0]).max():.1f}")
    assert r.n_entries == 59
    assert r.variance_coefficient <= 1e-3 and abs(r.energy_gradient) <= 1e-7
    assert np.abs(tot - tot[0]).max() <= 0.01 * kin.mean()
