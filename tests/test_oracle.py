"""CPU tests that pin the FP64 oracle (oracle/lima_oracle.cpp) independently of its own formulas.

The reference's fixtures live outside its repository (LIMA_data, TestUtils.h:18-20), so each
oracle function is pinned by a closed form, a textbook routine or an invariant that a plausible
mistake (dropped term, wrong sign, transposed index, wrong unit) fails:

  Coulomb constant   k e^2 / (1 nm) = 2.307078e-10 N per unit-charge pair (ElectrostaticsTests.h:19-26)
  LJ pair            4 eps ((s/r)^12 - (s/r)^6) and its derivative (ForceCorrectness.h:336-403)
  bonded terms       closed-form energies + forces equal to -dE/dx by central differences
  nonbonded          forces equal -dE/dx; net force zero
  FFT                numpy.fft.fftn
  PME spread         sum(rho) V_cell = sum(q) (B-spline partition of unity)
  PME total          direct Ewald sum within LIMA's own PME tolerances (ElectrostaticsTests.h:385-392)
  neighbour list     expanded cluster list == brute-force O(N^2) pair set
  integrator         harmonic bond frequency sqrt(k/mu)/2pi within 1% (ForceCorrectness.h:169-230)
  minimiser          the GROMACS steep rule in closed form on two bonds (accept / refuse sequence,
                     energies, max |F|); monotone energy; a stretched bond relaxes to b0
  thermostat         force-free gas: velocities v0 s^k with LIMA's clamp and measurement gating
  improper           axis geometry: closed-form energy, rotational gradient on i and l, zero net force
"""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest
from scipy.special import erfc

import lima_amd as L
from lima_amd.boxdesc import CustomBox, DEG_TO_RAD_L
from oracle.oracle import Oracle, fft3d, NB_CUTOFF, NB_INTRA, BONDED, PME, ALL

ROOT = Path(__file__).resolve().parent.parent
N_A = 6.02214076e23
E_KCMOL = float(np.float32(1.602176634e-19 * N_A / 1000.0))


def params(**kw):
    p = L.default_params(dt_fs=1.0)
    for k, v in kw.items():
        if k == "pme_grid":
            for d in range(3):
                p.pme_grid[d] = v[d]
        else:
            setattr(p, k, v)
    return p


def pair_box(r, q=(0.0, 0.0), sig=(0.0, 0.0), eps=(0.0, 0.0), L_nm=6.0, same_molecule=False):
    pos = [[1.0, 1.0, 1.0], [1.0 + r, 1.0, 1.0]]
    mol = [0, 0] if same_molecule else [0, 1]
    return CustomBox([L_nm] * 3, pos, q, sig, eps, [12.0, 12.0], molecule=mol)


# ---------------------------------------------------------------------------------------------
def test_coulomb_constant_and_erfc_pair_energy():
    """Two unit charges 0.6 nm apart in a 6 nm box, short-range part only: E = k q^2 erfc(kr)/r."""
    r = 0.6
    b = pair_box(r, q=(1.0, 1.0))
    p = params(enable_electrostatics=1)
    f, pot, e = Oracle(b, p).forces(NB_CUTOFF)
    f = f[b.layout_index()]
    # k from the reference's own check: k e^2 / 1 nm = 2.307078e-10 N per pair
    k_si_per_pair = 2.307078e-10                             # N at 1 nm: k e^2 = F r^2 (J m)
    k_jmol_nm = k_si_per_pair * 1e-18 * N_A / 1e-9           # J/mol * nm per e^2 (= 138.9 kJ/mol nm)
    kappa = 3.0 / 1.2
    expect = k_jmol_nm * erfc(kappa * r) / r
    assert e.coulomb_sr == pytest.approx(expect, rel=2e-6)
    # repulsive: force on particle 0 points away from particle 1 (-x)
    assert f[0, 0] < 0 and f[1, 0] > 0


def test_lj_pair_closed_form():
    sig, eps_kj, r = 0.34, 0.8, 0.37
    b = pair_box(r, sig=(sig, sig), eps=(eps_kj, eps_kj))
    f, pot, e = Oracle(b, params(enable_electrostatics=0)).forces(NB_CUTOFF)
    f, pot = f[b.layout_index()], pot[b.layout_index()]
    eps = eps_kj * 1000.0
    s6 = (sig / r) ** 6
    e_ref = 4 * eps * (s6 * s6 - s6)
    dEdr = 4 * eps * (-12 * s6 * s6 + 6 * s6) / r
    assert e.lj == pytest.approx(e_ref, rel=1e-6)
    assert f[1, 0] == pytest.approx(-dEdr, rel=1e-6)          # force on 1 = -dE/dr along +x
    assert f[0, 0] == pytest.approx(dEdr, rel=1e-6)
    assert pot[0] == pytest.approx(0.5 * e_ref, rel=1e-6)     # LIMA splits the pair energy


def test_lj_hard_cutoff_is_inclusive_and_not_shifted():
    sig, eps = 0.34, 0.8
    inside = Oracle(pair_box(1.1999, sig=(sig, sig), eps=(eps, eps)), params(enable_electrostatics=0)).forces(NB_CUTOFF)[2]
    outside = Oracle(pair_box(1.2001, sig=(sig, sig), eps=(eps, eps)), params(enable_electrostatics=0)).forces(NB_CUTOFF)[2]
    s6 = (sig / 1.1999) ** 6
    assert inside.lj == pytest.approx(4 * eps * 1000 * (s6 * s6 - s6), rel=1e-6)
    assert outside.lj == 0.0


def test_intra_pairs_have_no_cutoff_and_respect_exclusions():
    # 4-particle chain: 0-1-2-3 bonded, pairs (0,3) are 1-4 -> excluded; a far 5th atom in the
    # same molecule (bonded via 3-4) interacts with 0 beyond the cutoff (LIMA immediate kernel)
    pos = [[1, 1, 1], [1.15, 1, 1], [1.3, 1.0, 1], [1.45, 1, 1], [2.9, 1, 1]]
    n = 5
    b = CustomBox([6.0] * 3, pos, [0] * n, [0.3] * n, [0.5] * n, [12.0] * n, molecule=[0] * n,
                  bonds=[(0, 1, 0.15, 1e5), (1, 2, 0.15, 1e5), (2, 3, 0.15, 1e5), (3, 4, 1.45, 1e3)])
    f, pot, e = Oracle(b, params(enable_electrostatics=0)).forces(NB_INTRA)
    eps = 500.0
    expect = 0.0
    for i in range(n):
        for j in range(i + 1, n):
            if j - i <= 1:
                continue          # bonded pairs 0-1, 1-2, 2-3, 3-4 excluded
            r = np.linalg.norm(np.subtract(pos[j], pos[i]))
            s6 = (float(np.float32(0.3)) / r) ** 6      # parameters are stored in float32
            expect += 4 * eps * (s6 * s6 - s6)
    assert e.lj == pytest.approx(expect, rel=1e-9)


# ---------------------------------------------------------------------------------------------
def _fd_check(make_box, comps, energy_of, h=1e-5, rtol=1e-5):
    pos0 = make_box.pos.copy()
    b = make_box.build(pos0)
    f, _, _ = Oracle(b, make_box.p).forces(comps)
    li = b.layout_index()
    f = f[li]
    worst = 0.0
    for a in range(len(pos0)):
        for d in range(3):
            pp = pos0.copy(); pp[a, d] += h
            pm = pos0.copy(); pm[a, d] -= h
            ep = energy_of(Oracle(make_box.build(pp), make_box.p).forces(comps)[2])
            em = energy_of(Oracle(make_box.build(pm), make_box.p).forces(comps)[2])
            g = -(ep - em) / (2 * h)
            worst = max(worst, abs(g - f[a, d]) / (np.abs(f).max() + 1e-12))
    assert worst < rtol, worst
    return f


class _Maker:
    def __init__(self, pos, p, **kw):
        self.pos = np.asarray(pos, float)
        self.p = p
        self.kw = kw

    def build(self, pos):
        n = len(pos)
        kw = dict(self.kw)
        q = kw.pop("q", np.zeros(n))
        sig = kw.pop("sig", np.zeros(n))
        eps = kw.pop("eps", np.zeros(n))
        return CustomBox([5.0] * 3, pos, q, sig, eps, np.full(n, 12.0), **kw)


BASE4 = np.array([[1.0, 1.0, 1.0], [1.15, 1.02, 1.01], [1.2, 1.16, 1.05], [1.35, 1.2, 1.12]])


def _helix_chain(n=40):
    k = np.arange(n)
    return np.stack([0.5 + 0.3 * np.cos(1.7 * k), 0.5 + 0.3 * np.sin(1.7 * k), 0.5 + 0.045 * k], 1)


@pytest.mark.parametrize("components", [BONDED, NB_INTRA])
def test_chain_forces_invariant_across_the_periodic_boundary(components):
    """A 40-particle chain is two compounds (32 + 8).  Translated so that its compounds' key
    particles wrap to opposite sides of the box, every bonded term and bonded-compound pair must
    see the same geometry: LIMA moves a neighbour compound to its nearest image (hyperpos,
    EngineKernels.cuh:235-236).  Unwrapped origo differences would stretch the bonds that join the
    two compounds by a whole box length."""
    pos = _helix_chain()
    n = len(pos)
    bonds = [(i, i + 1, 0.15, 2e5) for i in range(n - 1)]
    angles = [(i, i + 1, i + 2, 111.0, 400.0, 0.0, 0.0) for i in range(n - 2)]
    dihedrals = [(i, i + 1, i + 2, i + 3, 0.0, 2.0, 3) for i in range(n - 3)]
    q = np.where(np.arange(n) % 2 == 0, 0.3, -0.3)

    def forces(shift):
        b = CustomBox([3.0] * 3, pos + shift, q, np.full(n, 0.33), np.full(n, 0.5), np.full(n, 12.0),
                      bonds=bonds, angles=angles, dihedrals=dihedrals)
        o = Oracle(b, params())
        f, _, e = o.forces(components)
        return f[b.gid], e, b

    f0, e0, b0 = forces(np.zeros(3))
    f1, e1, b1 = forces(np.array([2.3, 0.0, 1.6]))   # compound 0's key near x = 2.8, compound 1 wraps past 3.0
    assert len(set(map(tuple, b1.origo.tolist()))) == 2
    assert np.abs(b1.origo[0] - b1.origo[1]).max() >= 2   # the two origos sit on opposite sides of the box
    scale = np.abs(f0).max()
    assert scale > 0
    assert np.abs(f1 - f0).max() < 1e-6 * scale
    tot0 = e0.bonded + e0.lj + e0.coulomb_sr
    tot1 = e1.bonded + e1.lj + e1.coulomb_sr
    assert tot1 == pytest.approx(tot0, rel=1e-9, abs=1e-9)


@pytest.mark.parametrize("term", ["bond", "angle", "urey_bradley", "dihedral", "dihedral_phase", "pair14"])
def test_bonded_forces_are_energy_gradients(term):
    rng = np.random.default_rng(3)
    pos = BASE4 + rng.normal(0, 0.02, (4, 3))
    kw = {"bond": dict(bonds=[(0, 1, 0.14, 2e5)]),
          "angle": dict(angles=[(0, 1, 2, 100.0, 400.0, 0.0, 0.0)]),
          "urey_bradley": dict(angles=[(0, 1, 2, 100.0, 0.0, 0.22, 3e4)]),
          "dihedral": dict(dihedrals=[(0, 1, 2, 3, 0.0, 4.0, 3)]),
          "dihedral_phase": dict(dihedrals=[(0, 1, 2, 3, 30.0, 4.0, 2)]),
          "pair14": dict(pairs=[(0, 3, 0.32, 0.7)], dihedrals=[(0, 1, 2, 3, 0.0, 0.0, 3)])}[term]
    m = _Maker(pos, params(enable_electrostatics=0), **kw)
    f = _fd_check(m, BONDED, lambda e: e.bonded)
    assert np.abs(f.sum(0)).max() < 1e-6 * np.abs(f).max()       # Newton's third law


def test_bonded_closed_form_energies():
    # bond stretched by 0.01 nm: 1/2 k dx^2 ; angle at 90 deg vs 100: 1/2 k dtheta^2 ;
    # dihedral in a planar trans geometry (phi = 180 deg): k_lima (1 + cos(3 pi)) = 0
    pos = [[1, 1, 1], [1.16, 1, 1], [1.16, 1.15, 1], [1.31, 1.15, 1]]
    b = CustomBox([5.0] * 3, pos, [0] * 4, [0] * 4, [0] * 4, [12.0] * 4,
                  bonds=[(0, 1, 0.15, 2e5)], angles=[(0, 1, 2, 100.0, 400.0, 0.0, 0.0)])
    e = Oracle(b, params(enable_electrostatics=0)).forces(BONDED)[2]
    dth = np.pi / 2 - np.float32(100.0) * np.float32(np.float32(3.14159) / np.float32(180.0))
    assert e.bonded == pytest.approx(0.5 * 2e5 * 1000 * 0.01 ** 2 + 0.5 * 400 * 1000 * dth ** 2, rel=1e-5)
    pos_t = [[1, 1, 1], [1.1, 1.1, 1], [1.25, 1.1, 1], [1.35, 1.2, 1]]   # zig-zag in a plane: trans
    bt = CustomBox([5.0] * 3, pos_t, [0] * 4, [0] * 4, [0] * 4, [12.0] * 4, dihedrals=[(0, 1, 2, 3, 0.0, 4.0, 3)])
    assert Oracle(bt, params(enable_electrostatics=0)).forces(BONDED)[2].bonded == pytest.approx(0.0, abs=1e-6)
    pos_c = [[1, 1, 1], [1.1, 1.1, 1], [1.25, 1.1, 1], [1.35, 1.0, 1]]   # cis: phi = 0 -> k (1 + 1)
    bc = CustomBox([5.0] * 3, pos_c, [0] * 4, [0] * 4, [0] * 4, [12.0] * 4, dihedrals=[(0, 1, 2, 3, 0.0, 4.0, 3)])
    assert Oracle(bc, params(enable_electrostatics=0)).forces(BONDED)[2].bonded == pytest.approx(2 * 4.0 * 1000 / 2, rel=1e-6)


def _improper_box(phi, theta=0.0, a1=0.12, a2=0.12, r=0.13, psi0_deg=0.0, k_kj=300.0):
    """j and k on an axis along z, i at radius r and angle theta around it, l at radius r and
    angle phi.  LIMA's plane normals (j - i) x (k - i) and (j - l) x (k - l) then point along the
    tangents of i and l, so the improper angle is the rotation theta - phi between the half planes
    (negative for l ahead of i, by LIMA's sign rule n1 . (l - i) > 0)."""
    c = np.array([1.5, 1.5, 1.5])
    pos = np.array([[r * np.cos(theta), r * np.sin(theta), 0.0], [0.0, 0.0, -a1], [0.0, 0.0, a2],
                    [r * np.cos(phi), r * np.sin(phi), 0.0]]) + c
    return CustomBox([4.0] * 3, pos, [0] * 4, [0] * 4, [0] * 4, [12.0] * 4, impropers=[(0, 1, 2, 3, psi0_deg, k_kj)])


def _improper_eval(**kw):
    b = _improper_box(**kw)
    f, _, e = Oracle(b, params(enable_electrostatics=0)).forces(BONDED)
    return f[b.gid], e.bonded


@pytest.mark.parametrize("phi,psi0_deg", [(0.3, 0.0), (1.0, 20.0), (2.0, -15.0)])
def test_improper_closed_form_and_rotational_gradient(phi, psi0_deg):
    """LIMA's improper (ForceComputations.cuh:214-249) is a heuristic force split, not the gradient
    of its energy, so it cannot be pinned by finite differences in general.  In the axis geometry
    above it is pinned independently: the energy is 1/2 k (psi - psi0)^2 with psi = theta - phi
    (geometry, not LIMA's formula), the forces on i and l along their rotation tangents equal
    -dE/ds (central differences of the FP64 energy under rotation about the axis), the four forces
    sum to zero, and the symmetric j / k pair shares the residual equally."""
    k_kj, r = 300.0, 0.13
    f, e = _improper_eval(phi=phi, psi0_deg=psi0_deg, k_kj=k_kj, r=r)
    psi0 = float(np.float32(psi0_deg) * DEG_TO_RAD_L)
    assert e == pytest.approx(0.5 * k_kj * 1000.0 * (-phi - psi0) ** 2, rel=1e-6)
    h = 1e-2   # E is exactly quadratic in the rotation angle, so central differences are exact;
    #            a large step keeps the 1e-8 nm Coord rounding of the rotated positions negligible
    dE_l = (_improper_eval(phi=phi + h, psi0_deg=psi0_deg, k_kj=k_kj, r=r)[1]
            - _improper_eval(phi=phi - h, psi0_deg=psi0_deg, k_kj=k_kj, r=r)[1]) / (2 * h)
    dE_i = (_improper_eval(phi=phi, theta=h, psi0_deg=psi0_deg, k_kj=k_kj, r=r)[1]
            - _improper_eval(phi=phi, theta=-h, psi0_deg=psi0_deg, k_kj=k_kj, r=r)[1]) / (2 * h)
    t_l = np.array([-np.sin(phi), np.cos(phi), 0.0])
    t_i = np.array([0.0, 1.0, 0.0])
    scale = np.abs(f).max()
    assert f[3] @ t_l == pytest.approx(-dE_l / r, rel=1e-5, abs=1e-6 * scale)
    assert f[0] @ t_i == pytest.approx(-dE_i / r, rel=1e-5, abs=1e-6 * scale)
    assert np.abs(f.sum(0)).max() < 1e-9 * scale
    assert np.allclose(f[1], f[2], rtol=1e-9, atol=1e-9 * scale)


def test_improper_residual_goes_mostly_to_the_nearer_axis_atom():
    """With j nearer to i than k, j takes the larger share of the residual -(F_i + F_l), both
    shares parallel to it (lever split, ForceComputations.cuh:241-244)."""
    f, _ = _improper_eval(phi=1.0, a1=0.08, a2=0.2)
    res = -(f[0] + f[3])
    assert np.linalg.norm(f[1]) > np.linalg.norm(f[2]) > 0
    for fj in (f[1], f[2]):
        assert np.linalg.norm(np.cross(fj, res)) < 1e-9 * np.linalg.norm(fj) * np.linalg.norm(res)
        assert fj @ res > 0
    assert np.allclose(f[1] + f[2], res, rtol=1e-9, atol=1e-9 * np.abs(res).max())


def test_nonbonded_forces_are_energy_gradients():
    rng = np.random.default_rng(5)
    n = 8
    pos = rng.uniform(0.5, 2.5, (n, 3))
    q = rng.uniform(-0.6, 0.6, n)
    q -= q.mean()
    m = _Maker(pos, params(enable_electrostatics=1, cutoff_nm=1.2), q=q, sig=np.full(n, 0.3), eps=np.full(n, 0.5),
               molecule=np.arange(n))
    # avoid pairs near the hard cutoff where the energy is discontinuous
    d = np.linalg.norm(pos[:, None] - pos[None], axis=-1)
    if np.any(np.abs(d - 1.2) < 0.01):
        pytest.skip("pair at the cutoff")
    _fd_check(m, NB_CUTOFF, lambda e: e.lj + e.coulomb_sr, rtol=1e-5)


# ---------------------------------------------------------------------------------------------
@pytest.mark.parametrize("shape", [(6, 10, 15), (8, 9, 10), (27, 12, 5)])
def test_oracle_fft_matches_numpy(shape):
    rng = np.random.default_rng(1)
    a = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    assert np.allclose(fft3d(a), np.fft.fftn(a), atol=1e-9 * np.abs(a).sum())
    assert np.allclose(fft3d(a, inverse=True), np.fft.ifftn(a) * a.size, atol=1e-9 * np.abs(a).sum())


def test_pme_spread_conserves_charge():
    rng = np.random.default_rng(2)
    n = 20
    pos = rng.uniform(0, 3.0, (n, 3))
    q = rng.uniform(-1, 1, n)
    b = CustomBox([3.0] * 3, pos, q, np.zeros(n), np.zeros(n), np.full(n, 12.0), molecule=np.arange(n))
    o = Oracle(b, params(enable_electrostatics=1, pme_grid=(20, 25, 30)))
    rho, _ = o.pme_grids()
    vcell = (3.0 / 20) * (3.0 / 25) * (3.0 / 30)
    q_stored = b.q.astype(np.float64).sum()      # kC/mol, as held in the box (float32)
    assert rho.sum() * vcell == pytest.approx(q_stored, rel=1e-9)


def _ewald_direct(pos, q_kc, Lb, kappa, kc, kmax=12, nimg=2):
    """Textbook Ewald sum (real + reciprocal + self) in FP64 for a cubic box."""
    n = len(pos)
    e_real = 0.0
    f = np.zeros((n, 3))
    rng_img = range(-nimg, nimg + 1)
    for i in range(n):
        for j in range(n):
            for a in rng_img:
                for b in rng_img:
                    for c in rng_img:
                        if i == j and a == 0 and b == 0 and c == 0:
                            continue
                        d = pos[j] + Lb * np.array([a, b, c]) - pos[i]
                        r = np.linalg.norm(d)
                        if r > 4.5:
                            continue
                        e_real += 0.5 * kc * q_kc[i] * q_kc[j] * erfc(kappa * r) / r
                        sc = erfc(kappa * r) + 2 * kappa / np.sqrt(np.pi) * r * np.exp(-kappa ** 2 * r * r)
                        f[i] -= kc * q_kc[i] * q_kc[j] * d / r ** 3 * sc
    V = Lb ** 3
    e_rec = 0.0
    ks = np.arange(-kmax, kmax + 1)
    for kx in ks:
        for ky in ks:
            for kz in ks:
                if kx == 0 and ky == 0 and kz == 0:
                    continue
                kv = 2 * np.pi * np.array([kx, ky, kz]) / Lb
                k2 = kv @ kv
                S = np.sum(q_kc * np.exp(1j * pos @ kv))
                pref = 4 * np.pi / k2 * np.exp(-k2 / (4 * kappa ** 2)) / V * kc
                e_rec += 0.5 * pref * abs(S) ** 2
                # F_i = -dE/dx_i = pref q_i Im(exp(i k.x_i) S*) k
                f += (pref * q_kc * np.imag(np.exp(1j * pos @ kv) * np.conj(S)))[:, None] * kv[None, :]
    e_self = -kappa / np.sqrt(np.pi) * kc * np.sum(q_kc ** 2)
    return e_real + e_rec + e_self, f


KC_JMOL = 8.9875517873681764e9 / 1e-9 / N_A * 1e6     # k_C in J/mol nm per (kC/mol)^2


def _six_charges():
    rng = np.random.default_rng(11)
    n, Lb = 6, 3.0
    pos = rng.uniform(0.3, 2.7, (n, 3))
    q = np.array([1, -1, 1, -1, 0.5, -0.5])
    b = CustomBox([Lb] * 3, pos, q, np.zeros(n), np.zeros(n), np.full(n, 12.0), molecule=np.arange(n))
    q_kc = b.q.reshape(-1)[b.layout_index()].astype(np.float64)
    e_ref, f_ref = _ewald_direct(pos, q_kc, Lb, 2.5, KC_JMOL)
    return b, e_ref, f_ref


def _pme_total(b, K, divide):
    o = Oracle(b, params(enable_electrostatics=1, cutoff_nm=1.2, pme_grid=(K, K, K)))
    o.set_pme_divide_modulus(divide)
    f, _, e = o.forces(NB_CUTOFF | PME)
    f = f[b.layout_index()]
    return e.coulomb_sr + e.pme_recip + e.pme_self_ewald, f


def _max_rel_force_err(f, f_ref):
    return float((np.linalg.norm(f - f_ref, axis=1) / np.linalg.norm(f_ref, axis=1)).max())


def test_pme_divided_form_matches_direct_ewald():
    """With the textbook deconvolution (Green's function divided by the spline transform) the
    oracle's spread + FFT + Green's function + gather reproduce a direct Ewald sum: energy to
    1e-5 at K=32, forces to 2% with the O(h^2) convergence of the finite-difference gather."""
    b, e_ref, f_ref = _six_charges()
    e32, f32 = _pme_total(b, 32, True)
    _, f64 = _pme_total(b, 64, True)
    assert e32 == pytest.approx(e_ref, rel=1e-5)
    err32, err64 = _max_rel_force_err(f32, f_ref), _max_rel_force_err(f64, f_ref)
    assert err32 < 0.02
    assert 3.0 < err32 / err64 < 6.0


def test_pme_lima_form_converges_to_ewald_at_second_order():
    """LIMA multiplies by the spline transform (PME.cuh:471-491); that is an O(h^2) error in the
    Green's function, so the energy error must fall ~4x per halving of the grid spacing toward
    the same Ewald limit, and stay within LIMA's 7% force / 0.1% Newton bars at K=32."""
    b, e_ref, f_ref = _six_charges()
    e32, f32 = _pme_total(b, 32, False)
    e64, _ = _pme_total(b, 64, False)
    d32, d64 = abs(e32 - e_ref) / abs(e_ref), abs(e64 - e_ref) / abs(e_ref)
    assert d32 < 0.05
    assert 3.0 < d32 / d64 < 5.0
    assert _max_rel_force_err(f32, f_ref) < 0.07
    assert np.abs(f32.sum(0)).max() < 1e-3 * np.abs(f32).max()


# the six two-charge geometries of ElectrostaticsTests.h:318-325 (30 nm box, unit charges -1 / +1,
# expected force = direct Coulomb of the minimum image + one mirror image along mirrorDir)
LIMA_TWO_CHARGE = [
    ((6.025, 10.025, 10.0), (8.025, 10.025, 10.025), (-1, 0, 0)),
    ((4.0, 10.0, 10.0), (10.07, 10.025, 10.025), (-1, 0, 0)),
    ((4.025, 11.025, 9.025), (10.025, 10.025, 10.025), (-1, 0, 0)),
    ((9.6, 10.0, 10.0), (10.0, 10.0, 10.0), (-1, 0, 0)),
    ((0.1, 10.0, 10.0), (5.6, 10.0, 10.0), (-1, 0, 0)),
    ((0.1, 10.0, 10.0), (28.0, 10.0, 10.0), (0, 0, 0)),
]


@pytest.mark.parametrize("p0,p1,mirror", LIMA_TWO_CHARGE)
def test_pme_two_charges_within_lima_tolerances(p0, p1, mirror):
    """ElectrostaticsTests.h:300-396 at LIMA's 10 grid points per nm (PME.cuh:16): total
    (short range + PME) force within 7% of Coulomb + mirror; forces equal and opposite to 0.1%."""
    Lb = 30.0
    b = CustomBox([Lb] * 3, [p0, p1], [-1.0, 1.0], [0, 0], [0, 0], [12.0, 12.0], molecule=[0, 1])
    o = Oracle(b, params(enable_electrostatics=1, pme_grid=(300, 300, 300)))
    f = o.forces(NB_CUTOFF | PME)[0][b.layout_index()]
    q = b.q.reshape(-1)[b.layout_index()].astype(np.float64)
    x0, x1 = np.array(p0), np.array(p1)
    d = x1 - x0
    d -= Lb * np.round(d / Lb)
    dm = x1 + Lb * np.array(mirror) - x0
    f_exp = -KC_JMOL * q[0] * q[1] * (d / np.linalg.norm(d) ** 3 + dm / np.linalg.norm(dm) ** 3)
    assert np.linalg.norm(f[0] - f_exp) / np.linalg.norm(f_exp) < 0.07
    assert np.linalg.norm(f[0] + f[1]) / np.linalg.norm(f[0]) < 1e-3


# ---------------------------------------------------------------------------------------------
def _expand(box, atoms, off, ent, msk, radius):
    """Cluster list -> (set of (i, j, shift) pairs with i < j that lie within `radius` (integer
    test), number of (entry, slot) items inside the radius)."""
    o = Oracle(box, params())
    pos = o.positions()
    Lq = np.round(box.box_nm * 1e8).astype(np.int64)
    W = np.mod(pos, Lq)
    R = int(round(radius * 1e8))
    out = set()
    n = 0
    for I in range(len(off) - 1):
        for k in range(off[I], off[I + 1]):
            j = int(ent[k] & 0xFFFFFF)
            code = int(ent[k] >> 24)
            s = np.array([code % 3 - 1, (code // 3) % 3 - 1, code // 9 - 1])
            for sl in range(8):
                if not (msk[k] >> sl) & 1:
                    continue
                i = int(atoms[I * 8 + sl])
                d = W[j] + s * Lq - W[i]
                if int((d * d).sum()) <= R * R:
                    out.add((i, j, code) if i < j else (j, i, 26 - code))
                    n += 1
    return out, n


@pytest.mark.parametrize("which", ["water_small", "em_system"])
def test_cluster_list_covers_exactly_the_bruteforce_pairs(which):
    box = L.gen_water_lattice(6, 1.9, 7, 0.0) if which == "water_small" else L.gen_em_system(300, 4, 0.02)
    p = params()
    o = Oracle(box, p)
    atoms, off, ent, msk = o.neighbor_list()
    radius = 1.2 + 0.1
    got, n_items = _expand(box, atoms, off, ent, msk, radius)
    pi, pj, ps = o.pairs(radius)
    brute = set(zip(pi.tolist(), pj.tolist(), ps.tolist()))
    assert got == brute
    # a full list: every pair (and image) inside the radius appears exactly twice, once per side
    assert n_items == 2 * len(brute)
    assert len(ent) > 0


def test_cluster_list_in_a_wide_box_matches_numpy_bruteforce():
    """A 12 nm box (Coord differences past 2^31, squared distances of shifted images past 2^63):
    400 LJ particles at random positions, a quarter of them within 0.5 nm of a box face.  The
    expanded cluster list equals a brute-force pair set computed here in numpy (per-component
    screening keeps every square inside int64), and the oracle's own brute force agrees."""
    rng = np.random.default_rng(12)
    Lb = 12.0
    pos = rng.uniform(0.0, Lb, (400, 3))
    edge = rng.integers(0, 3, 100)
    pos[:100, :][np.arange(100), edge] = rng.uniform(0.0, 0.5, 100) * np.where(rng.random(100) < 0.5, 1, -1) % Lb
    box = CustomBox([Lb] * 3, pos, np.zeros(400), np.full(400, 0.3), np.full(400, 0.5), np.full(400, 12.0),
                    molecule=np.arange(400))
    p = params(enable_electrostatics=0)
    o = Oracle(box, p)
    atoms, off, ent, msk = o.neighbor_list()
    radius = 1.2 + p.nlist_skin_nm
    got, n_items = _expand(box, atoms, off, ent, msk, radius)
    W = np.mod(o.positions(), np.round(box.box_nm * 1e8).astype(np.int64))
    Lq = int(round(Lb * 1e8))
    R = int(round(radius * 1e8))
    brute = set()
    for code in range(27):
        s = np.array([code % 3 - 1, (code // 3) % 3 - 1, code // 9 - 1], np.int64)
        d = W[None, :, :] + s * Lq - W[:, None, :]          # d[i, j] = w_j + s L - w_i
        near = np.all(np.abs(d) <= R, axis=2)
        d2 = np.where(near[..., None], d, 0)
        ok = near & ((d2 * d2).sum(2) <= R * R)
        for i, j in zip(*np.nonzero(ok)):
            if i < j:
                brute.add((int(i), int(j), code))
    assert len(brute) > 50
    assert got == brute
    assert n_items == 2 * len(brute)
    pi, pj, ps = o.pairs(radius)
    assert set(zip(pi.tolist(), pj.tolist(), ps.tolist())) == brute
    # the entries themselves: (cluster, particle, image) with the image within R of the cluster's
    # bounding box, from the list definition evaluated here (no self pair; no intra pairs exist)
    ncl = len(off) - 1
    want = set()
    for I in range(ncl):
        mem = [int(a) for a in atoms[I * 8:(I + 1) * 8] if a >= 0]
        lo, hi = W[mem].min(0), W[mem].max(0)
        for code in range(27):
            s = np.array([code % 3 - 1, (code // 3) % 3 - 1, code // 9 - 1], np.int64)
            q = W + s * Lq
            g = np.maximum(np.maximum(lo - q, q - hi), 0)
            near = np.all(g <= R, axis=1)
            g = np.where(near[:, None], g, 0)
            inside = near & ((g * g).sum(1) <= R * R)
            for b in np.nonzero(inside)[0]:
                if code == 13 and len(mem) == 1 and mem[0] == b:
                    continue
                want.add((I, int(b), code))
    have = set()
    for I in range(ncl):
        for k in range(off[I], off[I + 1]):
            have.add((I, int(ent[k] & 0xFFFFFF), int(ent[k] >> 24)))
    assert have == want


# ---------------------------------------------------------------------------------------------
def test_integrator_bond_frequency():
    k_kj, b0 = 3e5, 0.1
    m1, m2 = 12.0, 16.0
    pos = [[1.0, 1.0, 1.0], [1.0 + b0 + 0.002, 1.0, 1.0]]
    b = CustomBox([4.0] * 3, pos, [0, 0], [0, 0], [0, 0], [m1, m2], bonds=[(0, 1, b0, k_kj)])
    p = params(enable_electrostatics=0, dt_fs=0.5)
    o = Oracle(b, p)
    xs = []
    for _ in range(600):
        o.step(1)
        x = o.positions()
        xs.append((x[1, 0] - x[0, 0]) * 1e-8)
    xs = np.array(xs) - b0
    crossings = np.where(np.diff(np.sign(xs)) > 0)[0]
    period_steps = np.diff(crossings).mean()
    mu = m1 * m2 / (m1 + m2) / 1000.0
    f_ref = np.sqrt(k_kj * 1000.0 / mu) / (2 * np.pi)        # 1/ns
    f_sim = 1.0 / (period_steps * 0.5e-6)
    assert f_sim == pytest.approx(f_ref, rel=0.01)


def test_integrator_conserves_energy_water():
    """Velocity Verlet (EngineUtils.cuh:31-41) conserves E = PE(x(t)) + KE(v(t)) for LJ + bonded
    water (no PME: LIMA's reciprocal forces are not the gradient of its PME energy).  After a step
    the oracle holds x(t+dt) and v(t), so PE is taken from the force call before each step and KE
    from the velocity that step computes."""
    box = L.gen_water_lattice(4, 1.3, 9, 300.0)
    li, m = box.layout_index(), box.masses_global()

    def kinetic(o):
        st = o.state()
        v = np.concatenate([st["particle_velocity"].reshape(-1, 3), st["tinymol_velocity"][:box.n_tinymols]])[li]
        return 0.5 * np.sum(m * (v.astype(np.float64) ** 2).sum(1))

    o = Oracle(box, params(dt_fs=0.5, cutoff_nm=0.6, enable_electrostatics=0))
    o.step(5)
    tot = []
    for _ in range(300):
        pe = o.forces()[2].potential
        o.step(1)
        tot.append(pe + kinetic(o))
    tot = np.array(tot)
    ke = kinetic(o)
    assert abs(tot[-1] - tot[0]) < 1e-2 * ke
    assert tot.std() < 5e-3 * ke


def _free_gas(T_target, n=8, seed=3):
    """n LJ-free tinymols (epsilon 0, no charges): forces vanish, so velocities change only
    through the thermostat scale.  Velocities [nm/ns] drawn and rescaled so that
    2 KE / ((3n - 3) k_B N_A) equals T_target exactly (in FP64)."""
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0.5, 5.5, (n, 3))
    m_g = 18.0
    v = rng.normal(0.0, 1.0, (n, 3))
    ke = 0.5 * (m_g / 1000.0) * (v ** 2).sum()
    kB = float(np.float32(1.38066e-23))
    v *= np.sqrt(T_target * (3 * n - 3) * kB * N_A / 2.0 / ke)
    v = v.astype(np.float32).astype(np.float64)
    box = CustomBox([6.0] * 3, pos, [0.0] * n, [0.3] * n, [0.0] * n, [m_g] * n, velocity=v, tinymol=[True] * n)
    T = 2.0 * 0.5 * (np.float32(m_g / 1000.0) * (v ** 2).sum()) / ((3 * n - 3) * kB * N_A)
    return box, v, T


def test_thermostat_clamps_and_gates_like_lima():
    """LIMA's thermostat (Thermostat.cuh:56-103, Engine.cu:162-174): the temperature is measured
    after steps that are multiples of both steps_per_temperature_measurement and
    5 * data_logging_interval, from T = 2 KE / (dof k_B N_A), dof = 3N - 3; the velocity scale
    clamp(310 / T, 1 -+ 0.001 / every) multiplies every later step's velocities.  On a force-free
    gas the velocities after k steps are therefore v0 * s^(k - first measurement step)."""
    box, v0, T0 = _free_gas(1000.0)
    p = params(dt_fs=1.0, enable_electrostatics=0, apply_thermostat=1, steps_per_temperature_measurement=10,
               data_logging_interval=1)
    o = Oracle(box, p)
    o.step(9)
    assert np.isnan(o.thermostat()[1])
    o.step(16)                                   # measured at steps 10 and 20
    s_expected = float(np.float32(1.0) - np.float32(np.float32(0.001) / np.float32(10.0)))
    scale, T = o.thermostat()
    assert scale == s_expected
    v = o.state()["tinymol_velocity"][:8].astype(np.float64)
    assert np.allclose(v, v0 * s_expected ** 15, rtol=1e-6, atol=0)
    assert T == pytest.approx(T0 * s_expected ** 20, rel=1e-5)
    # 5 * data_logging_interval = 15 does not divide 10 or 20: no measurement before step 30
    o2 = Oracle(box, params(dt_fs=1.0, enable_electrostatics=0, apply_thermostat=1,
                            steps_per_temperature_measurement=10, data_logging_interval=3))
    o2.step(25)
    assert np.isnan(o2.thermostat()[1]) and o2.thermostat()[0] == 1.0
    assert np.allclose(o2.state()["tinymol_velocity"][:8], v0, rtol=1e-6, atol=0)
    # measured but not applied
    o3 = Oracle(box, params(dt_fs=1.0, enable_electrostatics=0, apply_thermostat=0,
                            steps_per_temperature_measurement=5, data_logging_interval=1))
    o3.step(5)
    assert o3.thermostat()[0] == 1.0 and o3.thermostat()[1] == pytest.approx(T0, rel=1e-5)


def test_thermostat_unclamped_scale_is_target_over_temperature():
    box, v0, T0 = _free_gas(309.9)
    o = Oracle(box, params(dt_fs=1.0, enable_electrostatics=0, apply_thermostat=1,
                           steps_per_temperature_measurement=1, data_logging_interval=1))
    o.step(5)                                    # first measurement at step 5 (buffer of 5 steps)
    scale, T = o.thermostat()
    assert T == pytest.approx(T0, rel=1e-5)
    assert scale == pytest.approx(310.0 / T0, rel=1e-6)
    assert abs(scale - 1.0) < 0.001


def _sd_closed_form(pos_q, bonds, kb, b0, h0, n_iter, grow=1.2, shrink=0.2, per_atom=False):
    """Steepest descent on harmonic bonds, derived here from the GROMACS steep rule: every atom moves
    by h F_i / max_j |F_j| (per component, truncated to integer Coord units of 1e-8 nm); the trial is
    kept when the potential decreases, after which h grows by `grow`, otherwise h shrinks by
    `shrink`.  per_atom=True normalises each atom by its own |F_i| instead (a wrong rule the test
    must tell apart).  Returns the accepted energy [J/mol] and max |F| [kJ/mol/nm] after each
    iteration, as the oracle's trace holds them."""
    def energy_forces(p):
        F = np.zeros((len(p), 3))
        E = 0.0
        for a, b in bonds:
            d = (p[a] - p[b]).astype(np.float64) * 1e-8
            r = np.sqrt((d * d).sum())
            E += 0.5 * kb * (r - b0) ** 2
            f = -kb * (r - b0) * d / r
            F[a] += f
            F[b] -= f
        return E, F
    pos = pos_q.copy()
    E, F = energy_forces(pos)
    h = h0
    te, tf = [], []
    for _ in range(n_iter):
        norm = np.sqrt((F * F).sum(1))
        scale = norm[:, None] if per_atom else norm.max()
        trial = pos + np.trunc(h * F / scale * 1e8).astype(np.int64)
        Et, Ft = energy_forces(trial)
        if Et < E:
            pos, E, F = trial, Et, Ft
            h *= grow
        else:
            h *= shrink
        te.append(E)
        tf.append(np.sqrt((F * F).sum(1)).max() / 1000.0)
    return np.array(te), np.array(tf)


def test_steepest_descent_follows_the_steep_rule():
    """Three atoms A-B-C with bonds A-B along x (stretched 0.06 nm) and B-C along y (0.02 nm): B's
    force has two components and differs in size from A's and C's, so the step normalisation by
    the largest |F| (not per atom, not per component) decides every displacement.  The oracle's
    accept / reject sequence, energies and max |F| over 40 iterations equal the closed-form steep
    rule (h x 1.2 after an accepted trial, x 0.2 after a refused one); swapped factors or per-atom
    normalisation give a different sequence, so they would fail."""
    b0 = float(np.float32(0.15))
    kb = float(np.float32(2e5 * 1000.0))
    pos_nm = np.array([[1.0 - (0.15 + 0.06), 1.0, 1.0], [1.0, 1.0, 1.0], [1.0, 1.0 + 0.15 + 0.02, 1.0]])
    box = CustomBox([4.0] * 3, pos_nm, [0, 0, 0], [0, 0, 0], [0, 0, 0], [12.0] * 3, bonds=[(0, 1, 0.15, 2e5), (1, 2, 0.15, 2e5)])
    o = Oracle(box, params(enable_electrostatics=0))
    pos_q = o.positions()
    n_iter = 40
    r, te, tf = o.minimize(tol=1e-9, max_iter=n_iter, h0=0.01, trace=True)
    assert r.iterations == n_iter
    ce, cf = _sd_closed_form(pos_q, [(0, 1), (1, 2)], kb, b0, 0.01, n_iter)
    acc_oracle = np.diff(np.concatenate([[np.inf], te])) < 0
    acc_closed = np.diff(np.concatenate([[np.inf], ce])) < 0
    assert acc_oracle.sum() >= 5 and (~acc_oracle).sum() >= 5, "the run must both accept and refuse trials"
    assert np.array_equal(acc_oracle, acc_closed)
    assert np.allclose(te, ce, rtol=1e-6, atol=1e-3)
    assert np.allclose(tf, cf, rtol=1e-6, atol=1e-6)
    assert r.accepted == int(acc_closed.sum())
    # the rules a plausible mistake would implement instead differ from the oracle's trace
    for kw in (dict(grow=0.2, shrink=1.2), dict(per_atom=True), dict(grow=1.2, shrink=0.5)):
        we, _ = _sd_closed_form(pos_q, [(0, 1), (1, 2)], kb, b0, 0.01, n_iter, **kw)
        assert not np.allclose(we, te, rtol=1e-6, atol=1e-3), kw


def test_minimizer_monotone_and_bond_reaches_b0():
    pos = [[1.0, 1.0, 1.0], [1.25, 1.0, 1.0]]
    b = CustomBox([4.0] * 3, pos, [0, 0], [0, 0], [0, 0], [12.0, 12.0], bonds=[(0, 1, 0.15, 2e5)])
    o = Oracle(b, params(enable_electrostatics=0))
    r, te, tf = o.minimize(tol=1e-3, max_iter=2000, h0=0.01, trace=True)
    assert np.all(np.diff(te) <= 0)
    x = o.positions()
    assert abs((x[1, 0] - x[0, 0]) * 1e-8 - 0.15) < 1e-6
    box = L.gen_em_system(300, 1000, 0.02)
    o = Oracle(box, params())
    r, te, tf = o.minimize(tol=100.0, max_iter=300, h0=0.01, trace=True)
    assert np.all(np.diff(te) <= 0) and te[-1] < te[0]


# ---------------------------------------------------------------------------------------------
def test_generators_are_deterministic():
    a = L.gen_water_lattice(10, 3.104, 0, 300.0)
    b = L.gen_water_lattice(10, 3.104, 0, 300.0)
    c = L.gen_water_lattice(10, 3.104, 1, 300.0)
    assert a.hash == b.hash and a.hash != c.hash
    e1 = L.gen_em_system(1000, 1002, 0.02)
    e2 = L.gen_em_system(1000, 1002, 0.02)
    assert e1.hash == e2.hash
    sizes = L.em_batch_sizes(256)
    assert min(sizes) >= 300 and max(sizes) <= 3000 and len(set(sizes)) > 100


def test_water_box_density_and_neutrality():
    box = L.gen_water_lattice(10, 3.104, 0, 300.0)
    assert box.n_particles == 3000
    q = box.charge.sum()
    assert abs(q) < 1e-3
    mass_g = box.mass.sum() * 1000.0
    rho = mass_g / (6.02214076e23 * np.prod(box.box_nm) * 1e-21)
    assert rho == pytest.approx(1.0, rel=0.01)


def test_c_abi_library_exports_every_declared_symbol():
    header = (ROOT / "include" / "lima_hip.h").read_text()
    names = set(re.findall(r"\b(lima_[a-z_]+)\s*\(", header))
    lib_path = ROOT / "lima_amd" / "lib" / "liblima_hip.so"
    if not lib_path.exists():
        pytest.skip("HIP library not built")
    lib = ctypes.CDLL(str(lib_path))
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.parametrize("header", ["lima_gen.h", "lima_io.h", "lima_topo.h"])
def test_host_library_exports_every_declared_symbol(header):
    """liblima_host.so serves the generator, analysis / trajectory and topology headers."""
    text = (ROOT / "include" / header).read_text()
    names = set(re.findall(r"\b(lima_[a-z_]+)\s*\(", text))
    assert names
    lib = L.load_host()
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
