"""lima_amd -- MI355X-native LIMA force / minimization path.

Python access to the two shared libraries built in-tree:

* ``lib/liblima_hip.so``  -- the product: HIP/CDNA4 kernels behind the C ABI of
  ``include/lima_hip.h`` (stands in for LIMA_ENGINE's Engine and EM entry points).
* ``lib/liblima_host.so`` -- host C++: the synthetic box generators of ``include/lima_gen.h``.

Nothing here falls back to a CPU implementation: if the HIP library is missing,
``load_hip()`` raises.
"""
import ctypes as C
import os
from pathlib import Path

import numpy as np

from ._abi import BoxDesc, SimParams, Energy, RunStatus, StateOut, EmResult, Timings, BatchTuning, LogInfo, LogOut, EnergyAnalysis, SystemInfo

LIB_DIR = Path(__file__).resolve().parent / "lib"
REPO_ROOT = Path(__file__).resolve().parent.parent

_host = None
_hip = None


class LimaError(RuntimeError):
    pass


def load_host():
    global _host
    if _host is None:
        path = LIB_DIR / "liblima_host.so"
        if not path.exists():
            raise LimaError(f"{path} not built; run __graft_entry__.build()")
        lib = C.CDLL(str(path))
        pp = C.POINTER(C.c_void_p)
        lib.lima_gen_water_lattice.argtypes = [C.c_int32, C.c_double, C.c_uint64, C.c_double, pp]
        lib.lima_gen_chain_water.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int32, C.c_int32, C.c_int32,
                                             C.c_uint64, C.c_double, pp]
        lib.lima_gen_em_system.argtypes = [C.c_int32, C.c_uint64, C.c_double, pp]
        lib.lima_gen_large_water.argtypes = [C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double,
                                             C.c_uint64, C.c_double, pp]
        for f in ("lima_gen_water_lattice", "lima_gen_chain_water", "lima_gen_em_system", "lima_gen_large_water"):
            getattr(lib, f).restype = C.c_int
        lib.lima_box_describe.argtypes = [C.c_void_p, C.POINTER(BoxDesc)]
        lib.lima_box_free.argtypes = [C.c_void_p]
        lib.lima_box_hash.argtypes = [C.c_void_p]
        lib.lima_box_hash.restype = C.c_uint64
        lib.lima_box_n_particles.argtypes = [C.c_void_p]
        lib.lima_box_n_particles.restype = C.c_int64
        lib.lima_gen_last_error.restype = C.c_char_p
        lib.lima_default_params.argtypes = [C.POINTER(SimParams), C.c_float, C.c_int64]
        dp = C.POINTER(C.c_double)
        lib.lima_analyze_energy.argtypes = [dp, dp, C.c_int64, C.POINTER(EnergyAnalysis)]
        lib.lima_analyze_energy.restype = C.c_int
        lib.lima_write_trr.argtypes = [C.c_char_p, C.c_int64, C.c_int64, dp, C.POINTER(C.c_float),
                                       C.POINTER(C.c_int64), C.c_int32]
        lib.lima_write_trr.restype = C.c_int
        lib.lima_io_last_error.restype = C.c_char_p
        lib.lima_load_system.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(C.c_char_p), C.c_int32, C.c_int32,
                                         C.POINTER(C.c_void_p)]
        lib.lima_load_system.restype = C.c_int
        lib.lima_system_free.argtypes = [C.c_void_p]
        lib.lima_system_get_info.argtypes = [C.c_void_p, C.POINTER(SystemInfo)]
        lib.lima_system_get_info.restype = C.c_int
        lib.lima_system_make_box.argtypes = [C.c_void_p, C.POINTER(C.c_void_p)]
        lib.lima_system_make_box.restype = C.c_int
        lib.lima_system_layout.argtypes = [C.c_void_p, C.POINTER(C.c_int32)]
        lib.lima_system_layout.restype = C.c_int
        lib.lima_system_write_gro.argtypes = [C.c_void_p, C.POINTER(C.c_float), C.c_char_p]
        lib.lima_system_write_gro.restype = C.c_int
        lib.lima_topo_last_error.restype = C.c_char_p
        _host = lib
    return _host


def load_hip():
    """Load the HIP product library.  Raises if it has not been built."""
    global _hip
    if _hip is None:
        path = LIB_DIR / "liblima_hip.so"
        if not path.exists():
            raise LimaError(f"{path} not built; the HIP path has no CPU fallback")
        lib = C.CDLL(str(path))
        vp = C.c_void_p
        sig = {
            "lima_create": ([C.POINTER(BoxDesc), C.POINTER(SimParams), C.POINTER(vp)], C.c_int),
            "lima_destroy": ([vp], C.c_int),
            "lima_step": ([vp, C.c_int64], C.c_int),
            "lima_synchronize": ([vp], C.c_int),
            "lima_compute_forces": ([vp, C.POINTER(Energy)], C.c_int),
            "lima_get_energy": ([vp, C.POINTER(Energy)], C.c_int),
            "lima_minimize": ([vp, C.c_double, C.c_int64, C.c_double, C.POINTER(EmResult)], C.c_int),
            "lima_get_state": ([vp, C.POINTER(StateOut)], C.c_int),
            "lima_get_status": ([vp, C.POINTER(RunStatus)], C.c_int),
            "lima_get_neighbor_counts": ([vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)], C.c_int),
            "lima_get_neighbor_list": ([vp, C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_uint32),
                                        C.POINTER(C.c_uint8)], C.c_int),
            "lima_get_timings": ([vp, C.POINTER(Timings)], C.c_int),
            "lima_set_timing": ([vp, C.c_int32], C.c_int),
            "lima_get_pair_count": ([vp], C.c_int64),
            "lima_get_log_info": ([vp, C.POINTER(LogInfo)], C.c_int),
            "lima_read_log": ([vp, C.c_int64, C.c_int64, C.POINTER(LogOut)], C.c_int),
            "lima_batch_create": ([C.POINTER(BoxDesc), C.c_int32, C.POINTER(SimParams), C.POINTER(vp)], C.c_int),
            "lima_batch_destroy": ([vp], C.c_int),
            "lima_batch_autotune": ([vp, C.c_int32, C.POINTER(BatchTuning)], C.c_int),
            "lima_batch_set_tuning": ([vp, C.POINTER(BatchTuning)], C.c_int),
            "lima_batch_get_tuning": ([vp, C.POINTER(BatchTuning)], C.c_int),
            "lima_batch_minimize": ([vp, C.c_double, C.c_int64, C.c_double, C.POINTER(EmResult), C.POINTER(C.c_double)],
                                    C.c_int),
            "lima_batch_get_state": ([vp, C.c_int32, C.POINTER(StateOut)], C.c_int),
            "lima_batch_reset": ([vp], C.c_int),
            "lima_batch_get_phase_times": ([vp, C.POINTER(C.c_double)], C.c_int),
            "lima_batch_minimize_learned": ([vp, vp, C.c_double, C.c_int64, C.c_double, C.c_int32, C.c_int32,
                                             C.POINTER(EmResult), C.POINTER(C.c_double)], C.c_int),
            "lima_batch_record": ([vp, C.c_double, C.c_int64, C.c_double, C.c_int32, C.c_int32, C.POINTER(C.c_int64)], C.c_int),
            "lima_batch_training_set": ([vp, C.c_int32, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int64,
                                         C.POINTER(C.c_int64)], C.c_int),
            "lima_batch_get_counters": ([vp, C.POINTER(C.c_int64)], C.c_int),
            "lima_batch_get_features": ([vp, C.c_int32, C.POINTER(C.c_float)], C.c_int),
            "lima_mlp_n_params": ([], C.c_int32),
            "lima_mlp_create": ([C.c_uint64, C.POINTER(vp)], C.c_int),
            "lima_mlp_destroy": ([vp], C.c_int),
            "lima_mlp_get_params": ([vp, C.POINTER(C.c_float)], C.c_int),
            "lima_mlp_set_params": ([vp, C.POINTER(C.c_float)], C.c_int),
            "lima_mlp_forward": ([vp, C.POINTER(C.c_float), C.c_int64, C.POINTER(C.c_float)], C.c_int),
            "lima_mlp_forward_noalloc": ([vp, C.POINTER(C.c_float), C.c_int64, C.POINTER(C.c_float)], C.c_int),
            "lima_mlp_gradient": ([vp, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int64, C.POINTER(C.c_float),
                                   C.POINTER(C.c_double)], C.c_int),
            "lima_mlp_train": ([vp, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int64, C.c_int32, C.c_int32, C.c_float,
                                C.c_uint64, C.POINTER(C.c_float), C.POINTER(C.c_double)], C.c_int),
        }
        for name, (args, res) in sig.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        lib.lima_last_error.restype = C.c_char_p
        _hip = lib
    return _hip


def _io_check(lib, st):
    if st != 0:
        raise LimaError(f"status {st}: {lib.lima_io_last_error().decode()}")


def analyze_energy(pot, kin):
    """LIMA's energy analysis (SimAnalysis::analyzeEnergy) of the per-logged-step potential and
    kinetic energy series; see include/lima_io.h."""
    lib = load_host()
    p = np.ascontiguousarray(pot, np.float64)
    k = np.ascontiguousarray(kin, np.float64)
    if p.shape != k.shape or p.ndim != 1:
        raise ValueError("pot and kin must be 1-D series of one length")
    out = EnergyAnalysis()
    dp = C.POINTER(C.c_double)
    _io_check(lib, lib.lima_analyze_energy(p.ctypes.data_as(dp), k.ctypes.data_as(dp), len(p), C.byref(out)))
    return out


def write_trr(path, positions, box_nm, steps=None, host_order=False):
    """Positions-only TRR trajectory (MDFiles::TrrFile::Dump): positions [frames, atoms, 3] nm."""
    lib = load_host()
    x = np.ascontiguousarray(positions, np.float32)
    if x.ndim != 3 or x.shape[2] != 3:
        raise ValueError("positions must be [frames, atoms, 3]")
    b = np.ascontiguousarray(box_nm, np.float64)
    st = None if steps is None else np.ascontiguousarray(steps, np.int64)
    _io_check(lib, lib.lima_write_trr(str(path).encode(), x.shape[0], x.shape[1], b.ctypes.data_as(C.POINTER(C.c_double)),
                                      x.ctypes.data_as(C.POINTER(C.c_float)),
                                      None if st is None else st.ctypes.data_as(C.POINTER(C.c_int64)),
                                      1 if host_order else 0))


class System:
    """A .gro + .top system built into LIMA compounds (include/lima_topo.h)."""

    def __init__(self, gro, top, include_dirs=(), keep_small_molecules=False):
        self.lib = load_host()
        dirs = [str(d).encode() for d in include_dirs]
        arr = (C.c_char_p * max(len(dirs), 1))(*dirs)
        h = C.c_void_p()
        st = self.lib.lima_load_system(str(gro).encode(), str(top).encode(), arr, len(dirs),
                                       1 if keep_small_molecules else 0, C.byref(h))
        if st != 0:
            raise LimaError(f"status {st}: {self.lib.lima_topo_last_error().decode()}")
        self._h = h

    def _check(self, st):
        if st != 0:
            raise LimaError(f"status {st}: {self.lib.lima_topo_last_error().decode()}")

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                self.lib.lima_system_free(self._h)
                self._h = None
        except Exception:
            pass

    def info(self):
        i = SystemInfo()
        self._check(self.lib.lima_system_get_info(self._h, C.byref(i)))
        return i

    def box(self):
        """A new Box (lima_box) holding the system."""
        h = C.c_void_p()
        self._check(self.lib.lima_system_make_box(self._h, C.byref(h)))
        return Box(h.value)

    def layout(self):
        """.gro atom index of every entry of LIMA's buffer layout (-1 for unused slots)."""
        i = self.info()
        out = np.zeros(i.n_compounds * 32 + i.n_tinymols, np.int32)
        self._check(self.lib.lima_system_layout(self._h, out.ctypes.data_as(C.POINTER(C.c_int32))))
        return out

    def write_gro(self, positions, path):
        x = np.ascontiguousarray(positions, np.float32)
        self._check(self.lib.lima_system_write_gro(self._h, x.ctypes.data_as(C.POINTER(C.c_float)), str(path).encode()))


def load_system(gro, top, include_dirs=(), keep_small_molecules=False):
    return System(gro, top, include_dirs, keep_small_molecules)


def default_params(dt_fs=2.0, n_steps=1000, **overrides):
    p = SimParams()
    load_host().lima_default_params(C.byref(p), dt_fs, n_steps)
    for k, v in overrides.items():
        if k == "pme_grid":
            for d in range(3):
                p.pme_grid[d] = v[d]
        else:
            setattr(p, k, v)
    return p


def _arr(ptr, n, dtype):
    if n == 0 or not ptr:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


class Box:
    """Owning handle of a generated box (lima_box*) plus numpy copies of its arrays."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle)
        self.desc = BoxDesc()
        load_host().lima_box_describe(self._h, C.byref(self.desc))
        d = self.desc
        nc, nt = d.n_compounds, d.n_tinymols
        self.n_compounds, self.n_tinymols = nc, nt
        self.box_nm = np.array(list(d.box_nm))
        self.n_particles_per_compound = _arr(d.compound_n_particles, nc, np.int32)
        self.charge = _arr(d.particle_charge, nc * 32, np.float64).reshape(nc, 32)
        self.mass = _arr(d.particle_mass, nc * 32, np.float64).reshape(nc, 32)
        self.n_particles = int(load_host().lima_box_n_particles(self._h))
        self.hash = int(load_host().lima_box_hash(self._h))

    def __del__(self):
        try:
            if self._h:
                load_host().lima_box_free(self._h)
                self._h = None
        except Exception:
            pass

    # ---- global particle order (compound particles, then tinymols) <-> LIMA layout ----
    def layout_index(self):
        """Index into the [n_compounds*32 + n_tinymols] LIMA layout for each global particle."""
        idx = [c * 32 + s for c in range(self.n_compounds) for s in range(self.n_particles_per_compound[c])]
        idx += [self.n_compounds * 32 + t for t in range(self.n_tinymols)]
        return np.array(idx, dtype=np.int64)

    def masses_global(self):
        m = self.mass.reshape(-1)
        li = self.layout_index()
        out = np.empty(len(li))
        nc32 = self.n_compounds * 32
        for k, i in enumerate(li):
            out[k] = m[i] if i < nc32 else 0.0
        if self.n_tinymols:
            tm = _arr(self.desc.tinymol_mass, self.n_tinymols, np.float64)
            out[len(li) - self.n_tinymols:] = tm
        return out


def _gen(fn, *args):
    h = C.c_void_p()
    st = fn(*args, C.byref(h))
    if st != 0:
        raise LimaError(load_host().lima_gen_last_error().decode())
    return Box(h.value)


def gen_water_lattice(n_per_dim=10, box_nm=3.104, seed=0, temperature=300.0):
    return _gen(load_host().lima_gen_water_lattice, n_per_dim, box_nm, seed, temperature)


def gen_chain_water(box=(8.0, 8.0, 7.0), n_chains=128, chain_len=50, n_waters=12000, seed=1, temperature=300.0):
    return _gen(load_host().lima_gen_chain_water, box[0], box[1], box[2], n_chains, chain_len, n_waters, seed, temperature)


def gen_em_system(target_atoms, seed, noise_nm=0.02):
    return _gen(load_host().lima_gen_em_system, target_atoms, seed, noise_nm)


def gen_large_water(n_waters=333333, n_extra=1, extra_sigma=0.3, extra_eps_kj=0.5, extra_mass=39.948, seed=4,
                    temperature=300.0):
    return _gen(load_host().lima_gen_large_water, n_waters, n_extra, extra_sigma, extra_eps_kj, extra_mass, seed,
                temperature)


def em_batch_sizes(n_systems=256, lo=300, hi=3000, seed=1000):
    """System sizes drawn log-uniform in [lo, hi] with the project RNG convention (SplitMix64)."""
    sizes = []
    x = seed & 0xFFFFFFFFFFFFFFFF
    for _ in range(n_systems):
        x = (x + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        z ^= z >> 31
        u = (z >> 11) / float(1 << 53)
        sizes.append(int(round(lo * (hi / lo) ** u)))
    return sizes


def _state_arrays(nc, nt):
    """numpy buffers for a lima_state_out of a box with nc compounds and nt tinymols."""
    out = {
        "compound_origo": np.zeros((nc, 3), np.int32), "particle_coord": np.zeros((nc, 32, 3), np.int32),
        "particle_velocity": np.zeros((nc, 32, 3), np.float32), "particle_force": np.zeros((nc, 32, 3), np.float32),
        "particle_pot_energy": np.zeros((nc, 32), np.float32),
        "tinymol_block": np.zeros((max(nt, 1), 3), np.int32), "tinymol_coord": np.zeros((max(nt, 1), 3), np.int32),
        "tinymol_velocity": np.zeros((max(nt, 1), 3), np.float32), "tinymol_force": np.zeros((max(nt, 1), 3), np.float32),
        "tinymol_pot_energy": np.zeros(max(nt, 1), np.float32),
    }
    st = StateOut()
    names = [f[0] for f in StateOut._fields_]
    for k, a in out.items():
        setattr(st, k, a.ctypes.data_as(StateOut._fields_[names.index(k)][1]))
    return out, st


class Engine:
    """Handle of a lima_engine (Engine.cuh:78-99) on the GPU."""

    def __init__(self, box, params):
        self.lib = load_hip()
        self.box = box
        h = C.c_void_p()
        self._check(self.lib.lima_create(C.byref(box.desc), C.byref(params), C.byref(h)))
        self._h = h

    def _check(self, st):
        if st != 0:
            raise LimaError(f"status {st}: {self.lib.lima_last_error().decode()}")

    def close(self):
        if getattr(self, "_h", None):
            self.lib.lima_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n=1):
        """Issues n MD steps (lima_step returns before the device has run them)."""
        self._check(self.lib.lima_step(self._h, n))

    def synchronize(self):
        """Waits for the issued steps; raises what the device met (lima_synchronize)."""
        self._check(self.lib.lima_synchronize(self._h))

    def compute_forces(self):
        e = Energy()
        self._check(self.lib.lima_compute_forces(self._h, C.byref(e)))
        return e

    def energy(self):
        e = Energy()
        self._check(self.lib.lima_get_energy(self._h, C.byref(e)))
        return e

    def minimize(self, fmax_tol_kj=100.0, max_iter=10000, initial_step_nm=0.01):
        r = EmResult()
        self._check(self.lib.lima_minimize(self._h, fmax_tol_kj, max_iter, initial_step_nm, C.byref(r)))
        return r

    def status(self):
        s = RunStatus()
        self._check(self.lib.lima_get_status(self._h, C.byref(s)))
        return s

    def state(self):
        """Forces, potE, velocities, positions in LIMA layout (numpy)."""
        out, st = _state_arrays(self.box.n_compounds, self.box.n_tinymols)
        self._check(self.lib.lima_get_state(self._h, C.byref(st)))
        return out

    def neighbor_list(self):
        ncl, nent = C.c_int64(), C.c_int64()
        self._check(self.lib.lima_get_neighbor_counts(self._h, C.byref(ncl), C.byref(nent)))
        atoms = np.zeros(ncl.value * 8, np.int32)
        off = np.zeros(ncl.value + 1, np.int64)
        ent = np.zeros(max(nent.value, 1), np.uint32)
        msk = np.zeros(max(nent.value, 1), np.uint8)
        self._check(self.lib.lima_get_neighbor_list(
            self._h, atoms.ctypes.data_as(C.POINTER(C.c_int32)), off.ctypes.data_as(C.POINTER(C.c_int64)),
            ent.ctypes.data_as(C.POINTER(C.c_uint32)), msk.ctypes.data_as(C.POINTER(C.c_uint8))))
        return atoms, off, ent[:nent.value], msk[:nent.value]

    def set_timing(self, on=True):
        self._check(self.lib.lima_set_timing(self._h, 1 if on else 0))

    def timings(self):
        t = Timings()
        self._check(self.lib.lima_get_timings(self._h, C.byref(t)))
        return t

    def pair_count(self):
        return int(self.lib.lima_get_pair_count(self._h))

    def log_info(self):
        i = LogInfo()
        self._check(self.lib.lima_get_log_info(self._h, C.byref(i)))
        return i

    def read_log(self, first=None, n=None):
        """Data-log frames (LIMA buffer layout, P = n_compounds * 32 + n_tinymols entries per frame):
        dict of step [n], energy [n] (lima_energy), position [n, P, 3], pot_energy [n, P],
        speed [n, P], force [n, P, 3].  Defaults: every frame still held."""
        info = self.log_info()
        if first is None:
            first = info.first_held
        if n is None:
            n = info.frames_logged - first
        P = info.particles_per_frame
        out = {"step": np.zeros(max(n, 1), np.int64), "position": np.zeros((max(n, 1), P, 3), np.float32),
               "pot_energy": np.zeros((max(n, 1), P), np.float32), "speed": np.zeros((max(n, 1), P), np.float32),
               "force": np.zeros((max(n, 1), P, 3), np.float32)}
        energies = (Energy * max(n, 1))()
        lo = LogOut()
        lo.step = out["step"].ctypes.data_as(C.POINTER(C.c_int64))
        lo.energy = C.cast(energies, C.POINTER(Energy))
        for k in ("position", "pot_energy", "speed", "force"):
            setattr(lo, k, out[k].ctypes.data_as(C.POINTER(C.c_float)))
        self._check(self.lib.lima_read_log(self._h, first, n, C.byref(lo)))
        res = {k: v[:n] for k, v in out.items()}
        res["energy"] = [energies[k] for k in range(n)]
        return res


class Batch:
    """Handle of a lima_batch: many independent systems minimised together on the GPU."""

    def __init__(self, boxes, params):
        self.lib = load_hip()
        self.boxes = list(boxes)
        arr = (BoxDesc * len(self.boxes))()
        for k, b in enumerate(self.boxes):
            arr[k] = b.desc
        self._descs = arr
        h = C.c_void_p()
        self._check(self.lib.lima_batch_create(arr, len(self.boxes), C.byref(params), C.byref(h)))
        self._h = h

    def _check(self, st):
        if st != 0:
            raise LimaError(f"status {st}: {self.lib.lima_last_error().decode()}")

    def close(self):
        if getattr(self, "_h", None):
            self.lib.lima_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def autotune(self, trial_iterations=4):
        t = BatchTuning()
        self._check(self.lib.lima_batch_autotune(self._h, trial_iterations, C.byref(t)))
        return t

    def tuning(self):
        t = BatchTuning()
        self._check(self.lib.lima_batch_get_tuning(self._h, C.byref(t)))
        return t

    def set_tuning(self, t):
        self._check(self.lib.lima_batch_set_tuning(self._h, C.byref(t)))

    def minimize(self, fmax_tol_kj=100.0, max_iter=10000, initial_step_nm=0.01):
        """Returns (list of EmResult, wall seconds of the whole batch)."""
        res = (EmResult * len(self.boxes))()
        wall = C.c_double()
        self._check(self.lib.lima_batch_minimize(self._h, fmax_tol_kj, max_iter, initial_step_nm, res, C.byref(wall)))
        return list(res), wall.value

    def state(self, k):
        b = self.boxes[k]
        out, st = _state_arrays(b.n_compounds, b.n_tinymols)
        self._check(self.lib.lima_batch_get_state(self._h, k, C.byref(st)))
        return out

    def reset(self):
        self._check(self.lib.lima_batch_reset(self._h))

    def minimize_learned(self, model, fmax_tol_kj=100.0, max_iter=10000, initial_step_nm=0.01, feature_interval=4,
                         grow_on_learned=True):
        """Learned proposals with the plain step as fallback.  Returns (results, wall seconds).

        With grow_on_learned (the default) h grows by 1.2 after every accepted step, learned or plain,
        so a model with a constant scale reproduces the plain minimiser step for step."""
        res = (EmResult * len(self.boxes))()
        wall = C.c_double()
        self._check(self.lib.lima_batch_minimize_learned(self._h, model._h, fmax_tol_kj, max_iter, initial_step_nm,
                                                         feature_interval, 1 if grow_on_learned else 0, res, C.byref(wall)))
        return list(res), wall.value

    def record(self, fmax_tol_kj=100.0, max_iter=10000, initial_step_nm=0.01, atom_stride=16, max_states=2000):
        n = C.c_int64()
        self._check(self.lib.lima_batch_record(self._h, fmax_tol_kj, max_iter, initial_step_nm, atom_stride, max_states,
                                               C.byref(n)))
        return n.value

    def training_set(self, lookahead=1):
        n = C.c_int64()
        self._check(self.lib.lima_batch_training_set(self._h, lookahead, None, None, 0, C.byref(n)))
        x = np.zeros((n.value, 32), np.float32)
        t = np.zeros(n.value, np.float32)
        fp = C.POINTER(C.c_float)
        self._check(self.lib.lima_batch_training_set(self._h, lookahead, x.ctypes.data_as(fp), t.ctypes.data_as(fp), n.value,
                                                     C.byref(n)))
        return x, t

    def counters(self):
        """[n_systems, 3]: force evaluations, accepted learned proposals, recorded states."""
        out = np.zeros((len(self.boxes), 3), np.int64)
        self._check(self.lib.lima_batch_get_counters(self._h, out.ctypes.data_as(C.POINTER(C.c_int64))))
        return out

    def features(self, k):
        out = np.zeros((self.boxes[k].n_particles, 32), np.float32)
        self._check(self.lib.lima_batch_get_features(self._h, k, out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def phase_times(self):
        """[n_systems, 7] seconds: cutoff tile pairs, per-atom sums + intra + bonded, PME, reductions,
        trial steps, feature histograms, scale inference."""
        out = np.zeros((len(self.boxes), 7))
        self._check(self.lib.lima_batch_get_phase_times(self._h, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out


class Mlp:
    """Handle of a lima_mlp: the per-atom step-scale model (BF16 MFMA inference and training)."""

    NOALLOC_ROWS = 65536        # LIMA_MLP_NOALLOC_ROWS

    def __init__(self, seed=0):
        self.lib = load_hip()
        h = C.c_void_p()
        self._check(self.lib.lima_mlp_create(seed, C.byref(h)))
        self._h = h
        self.n_params = int(self.lib.lima_mlp_n_params())

    def _check(self, st):
        if st != 0:
            raise LimaError(f"status {st}: {self.lib.lima_last_error().decode()}")

    def close(self):
        if getattr(self, "_h", None):
            self.lib.lima_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        p = np.zeros(self.n_params, np.float32)
        self._check(self.lib.lima_mlp_get_params(self._h, p.ctypes.data_as(C.POINTER(C.c_float))))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        self._check(self.lib.lima_mlp_set_params(self._h, p.ctypes.data_as(C.POINTER(C.c_float))))

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        out = np.zeros(len(x), np.float32)
        fp = C.POINTER(C.c_float)
        self._check(self.lib.lima_mlp_forward(self._h, x.ctypes.data_as(fp), len(x), out.ctypes.data_as(fp)))
        return out

    def forward_noalloc(self, x):
        """lima_mlp_forward_noalloc: the same scales, staged in the library's fixed per-GPU array
        (at most NOALLOC_ROWS rows) instead of memory allocated per call."""
        x = np.ascontiguousarray(x, np.float32)
        out = np.zeros(len(x), np.float32)
        fp = C.POINTER(C.c_float)
        self._check(self.lib.lima_mlp_forward_noalloc(self._h, x.ctypes.data_as(fp), len(x), out.ctypes.data_as(fp)))
        return out

    def gradient(self, x, t):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        g = np.zeros(self.n_params, np.float32)
        loss = C.c_double()
        fp = C.POINTER(C.c_float)
        self._check(self.lib.lima_mlp_gradient(self._h, x.ctypes.data_as(fp), t.ctypes.data_as(fp), len(x),
                                               g.ctypes.data_as(fp), C.byref(loss)))
        return g, loss.value

    def train(self, x, t, epochs=4, batch=1024, lr=1e-3, seed=0):
        """Returns (mean loss per epoch, seconds)."""
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        losses = np.zeros(epochs, np.float32)
        sec = C.c_double()
        fp = C.POINTER(C.c_float)
        self._check(self.lib.lima_mlp_train(self._h, x.ctypes.data_as(fp), t.ctypes.data_as(fp), len(x), epochs, batch, lr, seed,
                                            losses.ctypes.data_as(fp), C.byref(sec)))
        return losses, sec.value


class StateBox:
    """A box descriptor whose positions / velocities come from an engine state (LIMA layout),
    with every other field shared with `base`.  Keeps the arrays alive."""

    def __init__(self, base, state):
        self.base = base
        self.desc = BoxDesc()
        C.pointer(self.desc)[0] = base.desc
        self._origo = np.ascontiguousarray(state["compound_origo"], np.int32)
        self._coord = np.ascontiguousarray(state["particle_coord"], np.int32)
        self._vel = np.ascontiguousarray(state["particle_velocity"], np.float32)
        self._tb = np.ascontiguousarray(state["tinymol_block"], np.int32)
        self._tc = np.ascontiguousarray(state["tinymol_coord"], np.int32)
        self._tv = np.ascontiguousarray(state["tinymol_velocity"], np.float32)
        ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))
        fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
        self.desc.compound_origo = ip(self._origo)
        self.desc.particle_coord = ip(self._coord)
        self.desc.particle_velocity = fp(self._vel)
        if base.n_tinymols:
            self.desc.tinymol_block = ip(self._tb)
            self.desc.tinymol_coord = ip(self._tc)
            self.desc.tinymol_velocity = fp(self._tv)
        for k in ("n_compounds", "n_tinymols", "box_nm", "n_particles", "n_particles_per_compound"):
            setattr(self, k, getattr(base, k))

    def layout_index(self):
        return self.base.layout_index()
