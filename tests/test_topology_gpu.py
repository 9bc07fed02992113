"""GPU parity on systems read from .gro / .top files: LIMA's DMPC lipid (Slipids 2020 forcefield:
CHARMM multi-term dihedrals, Urey-Bradley angles, impropers, 1-4 pairs) in vacuum and in TIP3P
water, against the FP64 oracle on the same box (relative RMS force error <= 1e-5)."""
from pathlib import Path

import numpy as np
import pytest

import lima_amd as L
from oracle.oracle import Oracle
from tests.helpers import rel_rms, layout_forces

pytestmark = pytest.mark.gpu

GOLD = Path(__file__).resolve().parent / "golden" / "topology"
FF_DIRS = [GOLD / "forcefields"]


def _solvated(tmp_path, spacing=0.31, clearance=0.3):
    """DMPC.gro plus TIP3P waters on a lattice of the 5 nm box, away from the lipid."""
    lines = (GOLD / "DMPC.gro").read_text().splitlines()
    lipid = lines[2:-1]
    xl = np.array([[float(l[20 + 8 * d:28 + 8 * d]) for d in range(3)] for l in lipid])
    box = 5.0
    n = int(box / spacing)
    g = (np.arange(n) + 0.5) * (box / n)
    sites = np.array(np.meshgrid(g, g, g, indexing="ij")).reshape(3, -1).T
    d = sites[:, None, :] - xl[None, :, :]
    d -= box * np.round(d / box)
    keep = sites[np.sqrt((d ** 2).sum(-1)).min(1) > clearance]
    a = np.deg2rad(104.52)
    h1, h2 = np.array([0.09572, 0, 0]), 0.09572 * np.array([np.cos(a), np.sin(a), 0])
    out = lines[:2] + lipid
    k = len(lipid)
    for w, o in enumerate(keep):
        for name, p in (("OW", o), ("HW1", o + h1), ("HW2", o + h2)):
            k += 1
            out.append(f"{2 + w:<5d}{'SOL':<5s}{name:<5s}{k % 100000:>5d}{p[0]:8.3f}{p[1]:8.3f}{p[2]:8.3f}")
    out[1] = str(k)
    out.append(lines[-1])
    (tmp_path / "solv.gro").write_text("\n".join(out) + "\n")
    (tmp_path / "solv.top").write_text(
        '#include "%s"\n#define _FF_CHARMM\n#include "Slipids_2020.ff/tip3p.itp"\n\n'
        "[ system ]\nDMPC in TIP3P\n\n[ molecules ]\nDMPC 1\nSOL %d\n" % (GOLD / "DMPC.itp", len(keep)))
    return tmp_path / "solv.gro", tmp_path / "solv.top", len(keep)


def _check(box, tag):
    p = L.default_params(dt_fs=1.0)
    eng = L.Engine(box, p)
    e_gpu = eng.compute_forces()
    f_ref, _, e_ref = Oracle(box, p).forces()
    err = rel_rms(layout_forces(box, eng.state()), f_ref)
    print(f"{tag}: N={box.n_particles} rel RMS force error {err:.3e}; bonded {e_gpu.bonded:.6e} / {e_ref.bonded:.6e}, "
          f"potential {e_gpu.potential:.6e} / {e_ref.potential:.6e}", flush=True)
    assert err <= 1e-5, err
    scale = abs(e_ref.lj) + abs(e_ref.coulomb_sr) + abs(e_ref.bonded) + abs(e_ref.pme_recip)
    for k in ("lj", "coulomb_sr", "bonded", "pme_recip"):
        assert abs(getattr(e_gpu, k) - getattr(e_ref, k)) <= 1e-5 * scale + 1e-3, k
    eng.step(20)
    eng.synchronize()                     # raises if a force went non-finite
    assert np.isfinite(eng.compute_forces().total)


def test_dmpc_vacuum_matches_oracle(tmp_path):
    top = tmp_path / "dmpc.top"
    top.write_text('#include "%s"\n\n[ system ]\nDMPC in vacuum\n\n[ molecules ]\nDMPC 1\n' % (GOLD / "DMPC.itp"))
    s = L.load_system(GOLD / "DMPC.gro", top, FF_DIRS)
    _check(s.box(), "DMPC vacuum")


@pytest.mark.parametrize("keep", [True, False], ids=["water_as_compounds", "water_as_lima_tinymols"])
def test_dmpc_in_water_matches_oracle(tmp_path, keep):
    gro, top, nw = _solvated(tmp_path)
    s = L.load_system(gro, top, FF_DIRS, keep_small_molecules=keep)
    i = s.info()
    assert i.n_tinymols == (0 if keep else nw)
    _check(s.box(), f"DMPC + {nw} TIP3P ({'compounds' if keep else 'tinymols'})")
