"""Build a lima_box_desc from numpy arrays (hand-made systems: tests, small examples).

Particles are given in nm / e / kJ units; the builder converts them to LIMA's stored units
(charges kC/mol, epsilon and force constants J/mol, masses kg/mol; Forcefield.cpp:342-347,
Bodies.cpp:9-47) and packs them into compounds of <= 32 particles with BondedParticlesLUT
exclusions, the same way lima_gen.cpp does for generated boxes.
"""
import ctypes as C

import numpy as np

from ._abi import BoxDesc

E_TO_KCMOL = np.float32(1.602176634e-19 * 6.02214076e23 / 1000.0)
DEG_TO_RAD_L = np.float32(np.float32(3.14159) / np.float32(180.0))


class CustomBox:
    """Keeps every array alive for as long as the descriptor is used."""

    def __init__(self, box_nm, pos_nm, charge_e, sigma_nm, eps_kj, mass_g, molecule=None, bonds=(), angles=(),
                 dihedrals=(), impropers=(), pairs=(), velocity=None, tinymol=None):
        pos = np.asarray(pos_nm, np.float64).reshape(-1, 3)
        n = len(pos)
        molecule = np.zeros(n, np.int64) if molecule is None else np.asarray(molecule)
        tinymol = np.zeros(n, bool) if tinymol is None else np.asarray(tinymol, bool)
        vel = np.zeros((n, 3)) if velocity is None else np.asarray(velocity, np.float64).reshape(-1, 3)
        # compounds: consecutive particles of the same molecule, <= 32 each
        comp_of = np.full(n, -1, np.int64)
        slot_of = np.zeros(n, np.int64)
        comps = []
        i = 0
        order = [k for k in range(n) if not tinymol[k]]
        while i < len(order):
            j = i
            while j < len(order) and molecule[order[j]] == molecule[order[i]] and j - i < 32:
                j += 1
            for s, k in enumerate(order[i:j]):
                comp_of[k] = len(comps)
                slot_of[k] = s
            comps.append(order[i:j])
            i = j
        nc = len(comps)
        tin = [k for k in range(n) if tinymol[k]]
        nt = len(tin)
        self.gid = np.where(comp_of >= 0, comp_of * 32 + slot_of, -1)
        L = np.asarray(box_nm, np.float64)
        Lq = np.round(L * 1e8).astype(np.int64)
        self.cn = np.array([len(c) for c in comps], np.int32)
        self.key = np.zeros(nc, np.int32)
        self.origo = np.zeros((nc, 3), np.int32)
        self.coord = np.zeros((nc, 32, 3), np.int32)
        self.vel = np.zeros((nc, 32, 3), np.float32)
        self.q = np.zeros((nc, 32), np.float32)
        self.sig = np.zeros((nc, 32), np.float32)
        self.eps = np.zeros((nc, 32), np.float32)
        self.mass = np.zeros((nc, 32), np.float32)
        self.lut = np.zeros((nc, 5, 32), np.uint32)
        for c, members in enumerate(comps):
            P = pos[members]
            cen = P.mean(0)
            k = int(np.argmin(((P - cen) ** 2).sum(1)))
            self.key[c] = k
            ku = np.round(P[k] * 1e8).astype(np.int64)
            kw = np.mod(ku, Lq)
            shift = kw - ku
            org = np.floor_divide(kw, 100000000)
            self.origo[c] = org
            for s, m in enumerate(members):
                self.coord[c, s] = np.round(pos[m] * 1e8).astype(np.int64) + shift - org * 100000000
                self.vel[c, s] = vel[m]
                self.q[c, s] = np.float32(charge_e[m]) * E_TO_KCMOL
                self.sig[c, s] = sigma_nm[m]
                self.eps[c, s] = np.float32(eps_kj[m]) * np.float32(1000.0)
                self.mass[c, s] = mass_g[m] / 1000.0
            for s in range(32):
                self.lut[c, 2, s] |= np.uint32(1 << s)
        self.tblock = np.zeros((max(nt, 1), 3), np.int32)
        self.tcoord = np.zeros((max(nt, 1), 3), np.int32)
        self.tvel = np.zeros((max(nt, 1), 3), np.float32)
        self.tsig = np.zeros(max(nt, 1), np.float32)
        self.teps = np.zeros(max(nt, 1), np.float32)
        self.tmass = np.zeros(max(nt, 1), np.float32)
        for t, m in enumerate(tin):
            u = np.mod(np.round(pos[m] * 1e8).astype(np.int64), Lq)
            blk = np.floor_divide(u, 100000000)
            self.tblock[t] = blk
            self.tcoord[t] = u - blk * 100000000
            self.tvel[t] = vel[m]
            self.tsig[t] = sigma_nm[m]
            self.teps[t] = np.float32(eps_kj[m]) * np.float32(1000.0)
            self.tmass[t] = mass_g[m] / 1000.0

        def mark(ids):
            for a in ids:
                for b in ids:
                    if a == b:
                        continue
                    ca, cb = comp_of[a], comp_of[b]
                    if abs(ca - cb) > 2:
                        raise ValueError("bonded term spans compounds more than 2 apart")
                    self.lut[ca, cb - ca + 2, slot_of[a]] |= np.uint32(1 << int(slot_of[b]))

        def terms(lst, na, conv):
            ids = np.zeros((max(len(lst), 1), na), np.int32)
            prm = np.zeros((max(len(lst), 1), len(conv)), np.float32)
            for k, t in enumerate(lst):
                atoms = [int(x) for x in t[:na]]
                mark(atoms)
                ids[k] = [self.gid[a] for a in atoms]
                prm[k] = [f(x) for f, x in zip(conv, t[na:])]
            return len(lst), ids, prm

        kilo = lambda x: np.float32(x) * np.float32(1000.0)
        deg = lambda x: np.float32(x) * DEG_TO_RAD_L
        same = lambda x: np.float32(x)
        self.nb, self.bid, self.bp = terms(bonds, 2, [same, kilo])
        self.na, self.aid, self.ap = terms(angles, 3, [deg, kilo, same, kilo])
        self.nd, self.did, self.dp = terms(dihedrals, 4, [deg, lambda x: np.float32(x) * np.float32(1000.0) / np.float32(2.0), same])
        self.ni, self.iid, self.ip = terms(impropers, 4, [deg, kilo])
        self.npr, self.pid, self.pp = terms(pairs, 2, [same, kilo])

        d = BoxDesc()
        for k in range(3):
            d.box_nm[k] = float(Lq[k]) / 1e8
        d.n_compounds, d.n_tinymols = nc, nt
        ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))
        fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
        d.compound_n_particles = ip(self.cn)
        d.compound_origo = ip(self.origo)
        d.compound_key_particle = ip(self.key)
        d.particle_coord = ip(self.coord)
        d.particle_velocity = fp(self.vel)
        d.particle_charge = fp(self.q)
        d.particle_sigma = fp(self.sig)
        d.particle_epsilon = fp(self.eps)
        d.particle_mass = fp(self.mass)
        d.bonded_lut = self.lut.ctypes.data_as(C.POINTER(C.c_uint32))
        d.tinymol_block = ip(self.tblock)
        d.tinymol_coord = ip(self.tcoord)
        d.tinymol_velocity = fp(self.tvel)
        d.tinymol_sigma = fp(self.tsig)
        d.tinymol_epsilon = fp(self.teps)
        d.tinymol_mass = fp(self.tmass)
        d.n_bonds, d.bond_ids, d.bond_params = self.nb, ip(self.bid), fp(self.bp)
        d.n_angles, d.angle_ids, d.angle_params = self.na, ip(self.aid), fp(self.ap)
        d.n_dihedrals, d.dihedral_ids, d.dihedral_params = self.nd, ip(self.did), fp(self.dp)
        d.n_impropers, d.improper_ids, d.improper_params = self.ni, ip(self.iid), fp(self.ip)
        d.n_pairs, d.pair_ids, d.pair_params = self.npr, ip(self.pid), fp(self.pp)
        self.desc = d
        self.n_compounds, self.n_tinymols = nc, nt
        self.n_particles = n
        self.box_nm = Lq / 1e8
        self.n_particles_per_compound = self.cn

    def layout_index(self):
        """LIMA layout index of each input particle (input order)."""
        out = np.zeros(self.n_particles, np.int64)
        tcount = 0
        for k in range(self.n_particles):
            if self.gid[k] >= 0:
                out[k] = self.gid[k]
            else:
                out[k] = self.n_compounds * 32 + tcount
                tcount += 1
        return out
