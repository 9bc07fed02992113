"""ctypes wrapper of the FP64 oracle (oracle/_build/liblima_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, as the reference the tests compare with and as the CPU baseline.  The product (lima_amd) never imports it.
"""
import ctypes as C
from pathlib import Path

import numpy as np

from lima_amd._abi import Energy, EmResult, StateOut, BoxDesc, SimParams

ORACLE_DIR = Path(__file__).resolve().parent
LIB_PATH = ORACLE_DIR / "_build" / "liblima_oracle.so"

ALL = 15
NB_CUTOFF, NB_INTRA, BONDED, PME = 1, 2, 4, 8

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} not built; run `make -C oracle`")
        L = C.CDLL(str(LIB_PATH))
        vp = C.c_void_p
        d = C.POINTER(C.c_double)
        sig = {
            "lima_oracle_create": ([C.POINTER(BoxDesc), C.POINTER(SimParams), C.POINTER(vp)], C.c_int),
            "lima_oracle_destroy": ([vp], None),
            "lima_oracle_set_threads": ([C.c_int], None),
            "lima_oracle_forces": ([vp, C.c_int, d, d, C.POINTER(Energy)], C.c_int),
            "lima_oracle_step": ([vp, C.c_int64], C.c_int),
            "lima_oracle_thermostat": ([vp, C.POINTER(C.c_float)], C.c_int),
            "lima_oracle_minimize": ([vp, C.c_double, C.c_int64, C.c_double, C.POINTER(EmResult), d, d], C.c_int),
            "lima_oracle_get_state": ([vp, C.POINTER(StateOut)], C.c_int),
            "lima_oracle_get_positions": ([vp, C.POINTER(C.c_int64)], C.c_int),
            "lima_oracle_n_particles": ([vp], C.c_int64),
            "lima_oracle_neighbor_counts": ([vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)], C.c_int),
            "lima_oracle_neighbor_list": ([vp, C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_uint32),
                                           C.POINTER(C.c_uint8)], C.c_int),
            "lima_oracle_pair_count": ([vp, C.c_double], C.c_int64),
            "lima_oracle_pairs": ([vp, C.c_double, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32)], C.c_int),
            "lima_oracle_pme_grid_dims": ([vp, C.POINTER(C.c_int32)], C.c_int),
            "lima_oracle_set_pme_divide_modulus": ([vp, C.c_int], None),
            "lima_oracle_pme_grids": ([vp, d, d], C.c_int),
            "lima_oracle_fft3d": ([C.c_int, C.c_int, C.c_int, d, d, d, d, C.c_int], C.c_int),
        }
        for n, (a, r) in sig.items():
            f = getattr(L, n)
            f.argtypes = a
            f.restype = r
        L.lima_oracle_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def set_threads(n):
    lib().lima_oracle_set_threads(n)


def fft3d(a, inverse=False):
    """Complex 3D FFT of a [nz, ny, nx] array (forward sign -1, unnormalised)."""
    a = np.ascontiguousarray(a, dtype=np.complex128)
    nz, ny, nx = a.shape
    re = np.ascontiguousarray(a.real)
    im = np.ascontiguousarray(a.imag)
    ro = np.zeros_like(re)
    io = np.zeros_like(re)
    lib().lima_oracle_fft3d(nx, ny, nz, _dp(re), _dp(im), _dp(ro), _dp(io), 1 if inverse else 0)
    return ro + 1j * io


class Oracle:
    def __init__(self, box, params):
        self.box = box
        h = C.c_void_p()
        st = lib().lima_oracle_create(C.byref(box.desc), C.byref(params), C.byref(h))
        if st != 0:
            raise RuntimeError(lib().lima_oracle_last_error().decode())
        self._h = h
        self.n = int(lib().lima_oracle_n_particles(h))

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib().lima_oracle_destroy(self._h)
        except Exception:
            pass

    @property
    def layout_size(self):
        return self.box.n_compounds * 32 + self.box.n_tinymols

    def forces(self, components=ALL):
        """Returns (force [layout,3], pot [layout], Energy) in LIMA layout."""
        f = np.zeros((self.layout_size, 3))
        p = np.zeros(self.layout_size)
        e = Energy()
        lib().lima_oracle_forces(self._h, components, _dp(f), _dp(p), C.byref(e))
        return f, p, e

    def step(self, n):
        lib().lima_oracle_step(self._h, n)

    def thermostat(self):
        """(velocity scale applied by the integrator, last measured temperature K or NaN)."""
        out = (C.c_float * 2)()
        lib().lima_oracle_thermostat(self._h, out)
        return float(out[0]), float(out[1])

    def minimize(self, tol=100.0, max_iter=1000, h0=0.01, trace=False):
        r = EmResult()
        te = np.zeros(max(max_iter, 1))
        tf = np.zeros(max(max_iter, 1))
        lib().lima_oracle_minimize(self._h, tol, max_iter, h0, C.byref(r), _dp(te) if trace else None,
                                   _dp(tf) if trace else None)
        if trace:
            return r, te[:r.iterations], tf[:r.iterations]
        return r

    def positions(self):
        p = np.zeros((self.n, 3), np.int64)
        lib().lima_oracle_get_positions(self._h, p.ctypes.data_as(C.POINTER(C.c_int64)))
        return p

    def state(self):
        nc, nt = self.box.n_compounds, self.box.n_tinymols
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
        lib().lima_oracle_get_state(self._h, C.byref(st))
        return out

    def neighbor_list(self):
        ncl, nent = C.c_int64(), C.c_int64()
        lib().lima_oracle_neighbor_counts(self._h, C.byref(ncl), C.byref(nent))
        atoms = np.zeros(ncl.value * 8, np.int32)
        off = np.zeros(ncl.value + 1, np.int64)
        ent = np.zeros(max(nent.value, 1), np.uint32)
        msk = np.zeros(max(nent.value, 1), np.uint8)
        lib().lima_oracle_neighbor_list(self._h, atoms.ctypes.data_as(C.POINTER(C.c_int32)),
                                        off.ctypes.data_as(C.POINTER(C.c_int64)),
                                        ent.ctypes.data_as(C.POINTER(C.c_uint32)),
                                        msk.ctypes.data_as(C.POINTER(C.c_uint8)))
        return atoms, off, ent[:nent.value], msk[:nent.value]

    def pair_count(self, radius_nm):
        """Unordered non-intra particle pairs (over all images) with distance <= radius_nm, by brute force."""
        return int(lib().lima_oracle_pair_count(self._h, radius_nm))

    def pairs(self, radius_nm):
        n = lib().lima_oracle_pair_count(self._h, radius_nm)
        pi = np.zeros(max(n, 1), np.int32)
        pj = np.zeros(max(n, 1), np.int32)
        ps = np.zeros(max(n, 1), np.int32)
        ip = C.POINTER(C.c_int32)
        lib().lima_oracle_pairs(self._h, radius_nm, pi.ctypes.data_as(ip), pj.ctypes.data_as(ip), ps.ctypes.data_as(ip))
        return pi[:n], pj[:n], ps[:n]

    def set_pme_divide_modulus(self, on=True):
        lib().lima_oracle_set_pme_divide_modulus(self._h, 1 if on else 0)

    def pme_dims(self):
        d = np.zeros(3, np.int32)
        lib().lima_oracle_pme_grid_dims(self._h, d.ctypes.data_as(C.POINTER(C.c_int32)))
        return tuple(int(x) for x in d)

    def pme_grids(self):
        kx, ky, kz = self.pme_dims()
        rho = np.zeros((kz, ky, kx))
        phi = np.zeros((kz, ky, kx))
        lib().lima_oracle_pme_grids(self._h, _dp(rho), _dp(phi))
        return rho, phi
