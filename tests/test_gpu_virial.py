"""GPU parity of the FP64 virial tensor (lima_energy.virial: W = -dU/d(strain), xx yy zz xy xz yz)
against the oracle's, whose virial the CPU tests pin by strain finite differences, the
sum x (x) F identity of finite-range terms and a direct Ewald sum.

The engine sums per-wave FP64 records of k_pair (-1/2 S d (x) d per evaluated pair), k_local's
intra pairs and bonded lanes ((x_atom - x_ref) (x) F per term incidence) and the PME z pass
(E_m (delta - 2 (1/k^2 + 1/4 kappa^2) k k) per wave vector).  Pair products are FP32, so the bar is
relative to the largest per-term virial (the components cancel in the total): 2e-5.
"""
import numpy as np
import pytest

import lima_amd as L
from oracle.oracle import Oracle, NB_CUTOFF, NB_INTRA, BONDED, PME

pytestmark = pytest.mark.gpu

VIR_TOL = 2e-5


def _term_scale(orc):
    """Largest |component| of the per-term virials (cutoff pairs, intra pairs, bonded, PME)."""
    return max(np.abs(np.array(orc.forces(c)[2].virial)).max() for c in (NB_CUTOFF, NB_INTRA, BONDED, PME))


def _bonded_chains():
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
    return CustomBox([3.1] * 3, pos, np.concatenate([q, q]), np.full(2 * n, 0.33), np.full(2 * n, 0.5),
                     np.full(2 * n, 12.0), molecule=mol, **terms)


BOXES = {
    "water_c0": lambda: L.gen_water_lattice(10, 3.104, 0, 300.0),
    "em_1200": lambda: L.gen_em_system(1200, 1003, 0.0),
    "tinymol_water": lambda: L.gen_large_water(700, 20, 0.3, 0.5, 39.948, 5, 300.0),
    "all_bonded_terms": _bonded_chains,
    "chain_water_c1": lambda: L.gen_chain_water(),
}


@pytest.mark.parametrize("name", list(BOXES))
def test_virial_matches_oracle(name):
    box = BOXES[name]()
    p = L.default_params(dt_fs=2.0)
    eng = L.Engine(box, p)
    e = eng.compute_forces()
    orc = Oracle(box, p)
    eo = orc.forces()[2]
    Wg, Wo = np.array(e.virial), np.array(eo.virial)
    scale = _term_scale(orc)
    err = np.abs(Wg - Wo).max() / scale
    print(f"{name}: W gpu {np.array2string(Wg, precision=6)}\n   oracle {np.array2string(Wo, precision=6)}\n"
          f"   max |diff| / term scale {err:.3e} (scale {scale:.4e})", flush=True)
    assert scale > 0
    assert err <= VIR_TOL


def test_virial_along_a_trajectory_and_in_log_frames():
    """Energy evaluations of the step loop (energy_interval = 1) and data-log frames carry the virial
    of their step: the frame logged at step 10 matches the oracle's virial of the step-10 state,
    and graph replay (launch mode 0) equals eager launches (mode 1) bit for bit."""
    box = L.gen_water_lattice(10, 3.104, 0, 300.0)
    ref = L.Engine(box, L.default_params(dt_fs=2.0))
    ref.step(10)
    st10 = ref.state()
    w_cf = np.array(ref.compute_forces().virial)
    frames = {}
    for mode in (0, 1):
        p = L.default_params(dt_fs=2.0, energy_interval=1, data_logging_interval=5, log_capacity=4, launch_mode=mode)
        eng = L.Engine(box, p)
        eng.step(11)
        log = eng.read_log()
        assert list(log["step"]) == [0, 5, 10]
        frames[mode] = np.array(log["energy"][2].virial)
    assert np.array_equal(frames[0], frames[1])
    orc = Oracle(L.StateBox(box, st10), L.default_params(dt_fs=2.0))
    Wo = np.array(orc.forces()[2].virial)
    scale = _term_scale(orc)
    e_frame, e_cf = np.abs(frames[0] - Wo).max() / scale, np.abs(w_cf - Wo).max() / scale
    print(f"step 10: frame virial error {e_frame:.3e}, compute_forces virial error {e_cf:.3e} (term scale {scale:.4e})",
          flush=True)
    assert e_frame <= VIR_TOL and e_cf <= VIR_TOL
