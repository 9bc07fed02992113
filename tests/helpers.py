"""Shared helpers for the parity tests (oracle comparisons in LIMA layout)."""
import numpy as np


def rel_rms(f_gpu, f_ref):
    """||F_gpu - F_ref||_2 / ||F_ref||_2 over all particles and components."""
    d = np.asarray(f_gpu, np.float64) - np.asarray(f_ref, np.float64)
    return float(np.sqrt((d ** 2).sum()) / np.sqrt((np.asarray(f_ref, np.float64) ** 2).sum()))


def layout_forces(box, state):
    """Force array in the [n_compounds*32 + n_tinymols, 3] layout of the oracle."""
    nc, nt = box.n_compounds, box.n_tinymols
    f = np.zeros((nc * 32 + nt, 3))
    f[:nc * 32] = state["particle_force"].reshape(-1, 3)
    if nt:
        f[nc * 32:] = state["tinymol_force"][:nt]
    return f


def layout_pot(box, state):
    nc, nt = box.n_compounds, box.n_tinymols
    p = np.zeros(nc * 32 + nt)
    p[:nc * 32] = state["particle_pot_energy"].reshape(-1)
    if nt:
        p[nc * 32:] = state["tinymol_pot_energy"][:nt]
    return p


def unwrapped_positions(box, state):
    """Unwrapped int64 Coord positions in global particle order from a state dict."""
    out = []
    for c in range(box.n_compounds):
        org = state["compound_origo"][c].astype(np.int64) * 100000000
        for s in range(box.n_particles_per_compound[c]):
            out.append(org + state["particle_coord"][c, s].astype(np.int64))
    for t in range(box.n_tinymols):
        out.append(state["tinymol_block"][t].astype(np.int64) * 100000000 + state["tinymol_coord"][t].astype(np.int64))
    return np.array(out, np.int64)


def min_image(d, box_nm):
    L = np.round(np.asarray(box_nm) * 1e8).astype(np.int64)
    return d - L * np.round(d / L).astype(np.int64)


def canonical_entries(atoms, off, ent, msk):
    """Set of (i-cluster member tuple, j particle, shift code, mask) for a cluster list."""
    out = set()
    ncl = len(off) - 1
    for I in range(ncl):
        members = tuple(int(a) for a in atoms[I * 8:(I + 1) * 8])
        for k in range(off[I], off[I + 1]):
            out.add((members, int(ent[k] & 0xFFFFFF), int(ent[k] >> 24), int(msk[k])))
    return out


def max_list_age_error(eng, box, params):
    """Relative RMS force error against the FP64 oracle of a step whose neighbour list is
    list_interval - 1 steps old (the oldest list any step uses): steps to the next rebuild, then to
    the last step before the following one, and compares that step's forces with the oracle's at
    the same positions."""
    import lima_amd as L
    from oracle.oracle import Oracle
    n0 = eng.status().n_list_rebuilds
    while eng.status().n_list_rebuilds == n0:
        eng.step(1)                      # this call rebuilt the list at its start
    interval = eng.status().list_interval   # lists now live this many steps (at most nlist_interval)
    if interval > 2:
        eng.step(interval - 2)
    st = eng.state()                     # the next step evaluates forces here with the oldest list
    eng.step(1)
    f_gpu = layout_forces(box, eng.state())
    f_ref, _, _ = Oracle(L.StateBox(box, st), params).forces()
    return rel_rms(f_gpu, f_ref)
