"""Host tools either side of the force path: LIMA's energy analysis and TRR trajectory output.

The analysis is pinned by closed forms (an arithmetic progression has a known mean, population
standard deviation and least-squares slope; a constant series has zero spread), by LIMA's
dropping of the first and last logged entry, and by its row definitions.  The TRR writer is pinned
by decoding its bytes with an independent reader of the GROMACS TRR layout (XDR records: magic
1993, the version string "GMX_trn_file", 13 int32 header fields, time and lambda, box, positions).
"""
import struct

import numpy as np
import pytest

import lima_amd as L


@pytest.mark.parametrize("a,b,c,grad_tol", [(-500.0, 2.0, 100.0, 1e-6), (-5.0e5, 12.5, 2.0e4, 2e-3)],
                         ids=["exact_in_fp32", "large_offset"])
def test_analysis_of_an_arithmetic_progression(a, b, c, grad_tol):
    """With small integers every FP32 sum of the regression is exact, so the slope is b to rounding;
    with a large offset LIMA's FP32 sums (n sum(xy) - sum(x) sum(y)) cancel to ~1e-3 relative."""
    n = 41                                   # entries 1 .. 39 are analysed
    i = np.arange(n, dtype=np.float64)
    pot, kin = a + b * i, np.full(n, c)
    r = L.analyze_energy(pot, kin)
    m = n - 2
    assert r.n_entries == m
    mean = a + c + b * (n - 1) / 2.0         # mean over entries 1 .. n-2
    assert abs(r.mean_energy - mean) <= 1e-6 * abs(mean)
    sd = abs(b) * np.sqrt((m * m - 1) / 12.0)
    assert abs(r.total.stddev - sd) <= 1e-4 * sd
    assert abs(r.pot.stddev - sd) <= 1e-4 * sd
    assert r.kin.stddev == 0.0
    assert abs(r.energy_gradient - b / mean) <= grad_tol * abs(b / mean)
    assert abs(r.variance_coefficient - sd / abs(mean)) <= 1e-4 * sd / abs(mean)
    assert r.total.min == np.float32(a + c + b) and r.total.max == np.float32(a + c + b * (n - 2))
    first, last = a + c + b, a + c + b * (n - 2)
    assert abs(r.total.change - (last - first) / first) <= 1e-5 * abs((last - first) / first)


def test_analysis_drops_first_and_last_entry():
    pot = np.array([1e9, 1.0, 2.0, 3.0, -1e9])
    r = L.analyze_energy(pot, np.zeros(5))
    assert r.n_entries == 3
    assert (r.pot.min, r.pot.max) == (1.0, 3.0)
    assert r.pot.change == 2.0                    # (3 - 1) / 1
    assert r.mean_energy == 2.0
    assert abs(r.pot.stddev - np.sqrt(2.0 / 3.0)) < 1e-7
    assert abs(r.energy_gradient - 0.5) < 1e-7    # slope 1 over mean 2


def test_analysis_constant_and_short_series():
    r = L.analyze_energy(np.full(10, -3.0), np.full(10, 1.0))
    assert r.variance_coefficient == 0.0 and r.energy_gradient == 0.0 and r.total.stddev == 0.0
    assert r.mean_energy == -2.0
    for n in (0, 1, 2):
        r = L.analyze_energy(np.zeros(n), np.zeros(n))
        assert r.n_entries == 0 and r.mean_energy == 0.0
    z = L.analyze_energy(np.zeros(6), np.zeros(6))
    assert z.variance_coefficient == 0.0        # 0 when mean and deviation are both 0


def _read_trr(raw, endian):
    """Frames of a positions-only TRR file: list of (step, natoms, box 3x3, positions [natoms, 3])."""
    frames, o = [], 0
    i32 = lambda: struct.unpack_from(endian + "i", raw, o)[0]
    while o < len(raw):
        magic = i32(); o += 4
        assert magic == 1993
        slen = i32(); o += 4
        olen = i32(); o += 4
        assert slen == olen + 1
        version = raw[o:o + olen].decode(); o += olen + (4 - olen % 4) % 4
        assert version == "GMX_trn_file"
        hdr = struct.unpack_from(endian + "13i2f", raw, o); o += 13 * 4 + 8
        ir, e, box_size, vir, pres, top, sym, x_size, v_size, f_size, natoms, step, nre = hdr[:13]
        assert (ir, e, vir, pres, top, sym, v_size, f_size, nre) == (0,) * 9
        assert hdr[13] == 0.0 and hdr[14] == 0.0          # time, lambda
        assert box_size == 36 and x_size == 12 * natoms
        box = np.array(struct.unpack_from(endian + "9f", raw, o)).reshape(3, 3); o += 36
        x = np.array(struct.unpack_from(endian + f"{3 * natoms}f", raw, o), np.float32).reshape(natoms, 3); o += 12 * natoms
        frames.append((step, natoms, box, x))
    return frames


@pytest.mark.parametrize("host_order", [False, True], ids=["xdr", "lima_host_order"])
def test_trr_layout(tmp_path, host_order):
    rng = np.random.default_rng(7)
    x = rng.uniform(0, 3, size=(3, 5, 3)).astype(np.float32)
    box = (3.104, 3.2, 2.9)
    path = tmp_path / "traj.trr"
    L.write_trr(path, x, box, steps=[0, 50, 100], host_order=host_order)
    raw = path.read_bytes()
    assert len(raw) == 3 * (120 + 12 * 5)         # header 84 B + box 36 B + 12 B per atom
    frames = _read_trr(raw, "<" if host_order else ">")
    assert [f[0] for f in frames] == [0, 50, 100]
    for k, (step, natoms, b, xx) in enumerate(frames):
        assert natoms == 5
        assert np.array_equal(np.diag(b), np.float32(box)) and np.count_nonzero(b - np.diag(np.diag(b))) == 0
        assert np.array_equal(xx, x[k])


def test_trr_defaults_and_errors(tmp_path):
    x = np.zeros((2, 1, 3), np.float32)
    L.write_trr(tmp_path / "a.trr", x, (1, 1, 1))
    assert [f[0] for f in _read_trr((tmp_path / "a.trr").read_bytes(), ">")] == [0, 1]   # frame index as step
    with pytest.raises(L.LimaError, match="extension"):
        L.write_trr(tmp_path / "a.xtc", x, (1, 1, 1))
