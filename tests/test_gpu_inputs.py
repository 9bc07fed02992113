"""Electrostatics switched off (LIMA's `enable_electrostatics = false`, Simulation.cpp:63-86) and
rejected inputs, through the C-ABI.

With electrostatics off the step has no short-range Coulomb and no PME (Engine.cu:333 skips the
PME call), so forces are LJ + bonded only; the FP64 oracle with the same flag is the reference
(relative RMS force error <= 1e-5).  Invalid parameters must raise LimaError with a message
naming the problem, and must leave the library usable for the next engine."""
import numpy as np
import pytest

import lima_amd as L
from oracle.oracle import Oracle
from tests.helpers import rel_rms, layout_forces, unwrapped_positions

pytestmark = pytest.mark.gpu


# Chain + water systems, relaxed and perturbed.  Not the water lattice: with LJ on O only, the O
# atoms on a perfect lattice feel LJ forces that cancel by symmetry and H carries no LJ, so without
# Coulomb its total force is FP32 rounding noise (RMS ~1 kJ/mol/nm against single pair forces of
# ~1e3) and a relative error of it measures nothing.
@pytest.mark.parametrize("make", [lambda: L.gen_em_system(600, 1010, 0.0),
                                  lambda: L.gen_em_system(1500, 1011, 0.02)], ids=["em_600", "em_1500_perturbed"])
def test_forces_without_electrostatics_match_oracle(make):
    box = make()
    p = L.default_params(dt_fs=1.0, enable_electrostatics=0)
    eng = L.Engine(box, p)
    e_gpu = eng.compute_forces()
    f_ref, _, e_ref = Oracle(box, p).forces()
    err = rel_rms(layout_forces(box, eng.state()), f_ref)
    rms_ref = float(np.sqrt((f_ref ** 2).mean()))
    print(f"electrostatics off: N={box.n_particles} force RMS {rms_ref:.3e}, rel RMS force error {err:.3e}", flush=True)
    assert rms_ref > 10.0, "the system must carry forces for the relative error to mean anything"
    assert err <= 1e-5, err
    assert e_gpu.coulomb_sr == 0.0 and e_gpu.pme_recip == 0.0
    assert e_ref.coulomb_sr == 0.0 and e_ref.pme_recip == 0.0
    scale = abs(e_ref.lj) + abs(e_ref.bonded)
    for k in ("lj", "bonded"):
        assert abs(getattr(e_gpu, k) - getattr(e_ref, k)) <= 1e-5 * scale + 1e-3, k
    # the trajectory stays finite and the list machinery runs without PME
    eng.step(50)
    eng.synchronize()                     # raises if a force went non-finite
    assert np.isfinite(eng.compute_forces().total)
    assert eng.status().critical_error_occured == 0


@pytest.mark.parametrize("override, message", [
    ({"cutoff_nm": 0.0}, "cutoff must be positive"),
    ({"cutoff_nm": 3.5}, "box edge must exceed cutoff"),
    ({"launch_mode": 7}, "launch_mode"),
    ({"log_capacity": -1}, "log_capacity"),
], ids=["zero_cutoff", "cutoff_beyond_box", "bad_launch_mode", "negative_log_capacity"])
def test_bad_parameters_are_rejected(override, message):
    box = L.gen_water_lattice(10, 3.104, 0, 300.0)   # 3.104 nm edges
    with pytest.raises(L.LimaError, match=message):
        L.Engine(box, L.default_params(dt_fs=1.0, **override))
    # the failed creation left nothing behind that breaks the next engine
    eng = L.Engine(box, L.default_params(dt_fs=1.0))
    assert np.isfinite(eng.compute_forces().total)


def test_trajectory_without_electrostatics_matches_oracle():
    """20 velocity-Verlet steps of a chain + water system with electrostatics off: the integer Coord
    positions of the HIP path and of the FP64 oracle stay within 1e-5 nm RMSD."""
    box = L.gen_em_system(600, 1010, 0.0)
    p = L.default_params(dt_fs=1.0, enable_electrostatics=0)
    eng = L.Engine(box, p)
    orc = Oracle(box, p)
    eng.step(20)
    orc.step(20)
    d = (unwrapped_positions(box, eng.state()) - orc.positions()) * 1e-8
    rmsd = float(np.sqrt((d ** 2).sum(1).mean()))
    print(f"electrostatics off: 20-step RMSD {rmsd:.3e} nm", flush=True)
    assert rmsd < 1e-5, rmsd
