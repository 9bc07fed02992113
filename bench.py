#!/usr/bin/env python3
"""bench.py -- ns/day of the MI355X-native LIMA step (force path + integrator) on one GPU.

Workload: the chain + water box of the mixed-precision config (8 x 8 x 7 nm, 128 chains of 50
atoms in two leaflets along z, 12,000 waters filling the rest, 42,400 atoms; seed 1), cutoff
1.2 nm, PME spacing 0.12 nm (72 x 72 x 60 grid), dt = 2 fs, FP32 forces with FP64 energy
reductions.  The generated structure is relaxed with steepest descent and equilibrated with 1,000 MD steps
(both untimed), W warm-up steps run, then K steps are timed: lima_step only issues work, so the host clock runs from an idle
device to lima_synchronize after the K steps.

The line also carries: the same K steps issued one lima_step(h, 1) call at a time (LIMA's
per-step loop); per-phase GPU times from HIP events and the rooflines of k_pair (FP32 VALU), the
PME chain and the list rebuild (HBM); the FP64 oracle on the host CPU; the force, energy and
virial errors against the oracle for every config; the 1M-atom water box; and the batched and
learned minimisation.

Prints one JSON line (stdout).  Progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

FP32_PEAK_TFLOPS = 157.3           # MI355X published dense FP32 peak
HBM_PEAK_GBS = 8000.0              # MI355X published HBM3E bandwidth
FLOPS_PER_PAIR = 64                # algorithmic FP32 flops per pair inside the cutoff
COPY_CALIBRATION_BYTES = 512 << 20


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def load_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` (k_pair, pme_chain per step, nlist_rebuild per rebuild) from the
    committed counter summary profiles/traffic_<workload>.txt (line 'traffic_bytes_per_launch <kernel> = N',
    FETCH_SIZE and WRITE_SIZE scaled by a copy of known size), or None."""
    path = ROOT / "profiles" / f"traffic_{workload}.txt"
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        head, _, val = line.partition("=")
        if head.split() == ["traffic_bytes_per_launch", kernel]:
            try:
                return int(float(val.split()[0]))
            except (ValueError, IndexError):
                return None
    return None


def calibrate_copy(repeats=5):
    """A plain device-to-device copy of known size (512 MiB read + 512 MiB written per launch):
    the FETCH_SIZE / WRITE_SIZE counters read on it give the factor that converts those counters
    into bytes for the kernels measured."""
    import torch
    src = torch.ones(COPY_CALIBRATION_BYTES // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(repeats):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": COPY_CALIBRATION_BYTES, "launches": repeats}), flush=True)
    return 0


def em_bench(args):
    """Time to Fmax < 100 kJ/mol/nm for the batch of independent systems: plain steepest descent,
    then learned proposals (model trained on plain trajectories of fresh systems), same start."""
    import numpy as np
    import lima_amd as L
    t = time.time()
    sizes = L.em_batch_sizes(args.em_systems, seed=1000)
    boxes = [L.gen_em_system(int(n), 1000 + k, 0.02) for k, n in enumerate(sizes)]
    natoms = int(sum(b.n_particles for b in boxes))
    log(f"em: generated {len(boxes)} systems ({natoms} atoms) in {time.time() - t:.1f}s")
    p = L.default_params()
    plain = L.Batch(boxes, p)
    tune = plain.autotune(3)
    log(f"em: autotuned shapes {tune.as_list()}")
    res_p, wall_p = plain.minimize(100.0, args.em_max_iter, 0.01)
    cnt_p = plain.counters()
    log(f"em: plain {wall_p:.3f}s, {sum(r.iterations for r in res_p)} iterations")
    # training data: plain trajectories of fresh systems
    t = time.time()
    tsizes = L.em_batch_sizes(args.em_train_systems, seed=2000)
    tboxes = [L.gen_em_system(int(n), 2000 + k, 0.02) for k, n in enumerate(tsizes)]
    tb = L.Batch(tboxes, p)
    tb.set_tuning(tune)
    states = tb.record(100.0, args.em_max_iter, 0.01, atom_stride=args.em_stride, max_states=4000)
    x, tgt = tb.training_set(lookahead=1)
    rec_s = time.time() - t
    model = L.Mlp(seed=1)
    losses, train_s = model.train(x, tgt, epochs=args.em_epochs, batch=1024, lr=1e-3, seed=3)
    log(f"em: {states} recorded states, {len(tgt)} samples ({rec_s:.1f}s), trained {args.em_epochs} epochs in "
        f"{train_s:.2f}s, loss {losses[0]:.4f} -> {losses[-1]:.4f}")
    learned = L.Batch(boxes, p)
    learned.set_tuning(tune)
    res_l, wall_l = learned.minimize_learned(model, 100.0, args.em_max_iter, 0.01, feature_interval=4,
                                            grow_on_learned=True)
    cnt_l = learned.counters()
    log(f"em: learned {wall_l:.3f}s, {sum(r.iterations for r in res_l)} iterations")
    # control: the same learned path with a model whose output is a constant scale; such a model
    # proposes the plain step, which the path then takes as such (one evaluation per iteration), so
    # the control's extra time is the path's overhead (feature histograms, inference)
    const = L.Mlp(seed=0)
    prm = np.zeros(const.n_params, np.float32)
    const.set_params(prm)     # 4 sigmoid(0) = 2 everywhere
    control = L.Batch(boxes, p)
    control.set_tuning(tune)
    res_c, wall_c = control.minimize_learned(const, 100.0, args.em_max_iter, 0.01, feature_interval=4,
                                              grow_on_learned=True)
    cnt_c = control.counters()
    log(f"em: constant-scale control {wall_c:.3f}s, {sum(r.iterations for r in res_c)} iterations")
    ep = np.array([r.final_energy for r in res_p])
    el = np.array([r.final_energy for r in res_l])
    rel = (el - ep) / np.abs(ep)

    def side(res, wall, cnt):
        return {"seconds": round(wall, 4), "converged": int(sum(r.converged for r in res)),
                "iterations_total": int(sum(r.iterations for r in res)),
                "iterations_max": int(max(r.iterations for r in res)),
                "force_evaluations_total": int(cnt[:, 0].sum()),
                "slowest_system_seconds": round(max(r.wall_seconds for r in res), 4)}

    return {
        "workload": f"{len(boxes)} chain+water systems, 300-3000 atoms log-uniform (seeds 1000..{999 + len(boxes)}), "
                    f"{natoms} atoms, 0.02 nm noise, Fmax < 100 kJ/mol/nm, h0 0.01 nm",
        "plain": side(res_p, wall_p, cnt_p),
        "learned": dict(side(res_l, wall_l, cnt_l), learned_proposals_accepted=int(cnt_l[:, 1].sum())),
        "constant_scale_control": dict(side(res_c, wall_c, cnt_c), learned_proposals_accepted=int(cnt_c[:, 1].sum())),
        "speedup_time_to_fmax": round(wall_p / wall_l, 3) if wall_l > 0 else None,
        "final_energy_total_rel_diff": float(f"{(el.sum() - ep.sum()) / abs(ep.sum()):.3e}"),
        "final_energy_per_system": {"median_abs_rel_diff": float(f"{np.median(np.abs(rel)):.3e}"),
                                    "within_0.1pct": int(np.sum(np.abs(rel) < 1e-3)),
                                    "learned_lower": int(np.sum(el < ep))},
        "training": {"systems": len(tboxes), "seeds": f"2000..{1999 + len(tboxes)}", "recorded_states": int(states),
                     "samples": int(len(tgt)), "atom_stride": args.em_stride, "target": "secant step estimate to the next accepted state, log-mapped onto (0.05, 3.95)", "epochs": args.em_epochs,
                     "record_seconds": round(rec_s, 2), "train_seconds": round(train_s, 3),
                     "loss_first_last": [round(float(losses[0]), 5), round(float(losses[-1]), 5)]},
        "autotuned_shapes": tune.as_list(),
    }


def pme_grid_dims(box_nm, spacing=0.12):
    """Grid points per axis from the PME spacing: ceil(L / h), at least 4, rounded up to 2^a 3^b 5^c
    (the rule lima_create applies when no explicit grid is given)."""
    import math

    def smooth(n):
        for q in (2, 3, 5):
            while n % q == 0:
                n //= q
        return n == 1
    out = []
    for L in box_nm:
        n = max(4, math.ceil(L / spacing - 1e-9))
        while not smooth(n):
            n += 1
        out.append(n)
    return out


# ---------------------------------------------------------------------------------------------
# host CPU description and the FP64 oracle baseline


def host_cpu():
    """lscpu model name and sockets, and the physical cores behind the CPUs this process may use."""
    info = {"model": None, "sockets": None}
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() == "Model name":
                info["model"] = v.strip()
            elif k.strip() == "Socket(s)":
                info["sockets"] = int(v.strip())
    except Exception:
        pass
    cpus = sorted(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else list(range(os.cpu_count() or 1))
    cores = {}
    for c in cpus:
        try:
            base = f"/sys/devices/system/cpu/cpu{c}/topology"
            key = (open(f"{base}/physical_package_id").read().strip(), open(f"{base}/core_id").read().strip())
        except OSError:
            key = ("?", str(c))
        cores.setdefault(key, []).append(c)
    info["cpus_allowed"] = len(cpus)
    info["physical_cores_allowed"] = len(cores)
    info["_cores"] = cores
    return info


def cpu_baseline_c1(sbox, p, dt_fs, threads, cpu, warm=5, timed=50, repeats=3, one_thread_steps=2):
    """The FP64 oracle on the host: `repeats` runs of `timed` MD steps after `warm` warm-up steps on
    `threads` threads (median ms/step), and a bounded one-thread sample (median of 3 single-step
    timings after one warm-up step)."""
    from oracle.oracle import Oracle, set_threads
    set_threads(threads)
    orc = Oracle(sbox, p)
    runs = []
    for r in range(repeats):
        orc.step(warm)
        t = time.perf_counter()
        orc.step(timed)
        runs.append((time.perf_counter() - t) / timed)
        log(f"cpu baseline run {r + 1}/{repeats}: {runs[-1] * 1e3:.1f} ms/step on {threads} threads")
    ms_all = sorted(runs)[len(runs) // 2] * 1e3
    set_threads(1)
    orc1 = Oracle(sbox, p)
    orc1.step(1)
    one = []
    for r in range(3):
        t = time.perf_counter()
        orc1.step(one_thread_steps)
        one.append((time.perf_counter() - t) / one_thread_steps)
    ms_one = sorted(one)[1] * 1e3
    set_threads(threads)
    log(f"cpu baseline: {ms_all:.1f} ms/step on {threads} threads, {ms_one:.0f} ms/step on 1 thread")
    nsd = lambda ms: round(1e3 / ms * dt_fs * 86400.0 / 1e6, 4)
    return {"value": nsd(ms_all), "unit": "ns/day", "cores": threads, "kind": "port",
            "sample": f"FP64 oracle (g++ -O3 -march=native, OpenMP) MD steps of the same 42,400-atom state: "
                      f"{repeats} runs of {timed} steps after {warm} warm-up steps each on {threads} threads, median "
                      f"{ms_all:.1f} ms/step; one thread: median of 3 x {one_thread_steps} steps after 1 warm-up, "
                      f"{ms_one:.0f} ms/step (a bounded sample; 55 steps x 3 on one thread would take ~10 min)",
            "cpu_model": cpu["model"], "sockets": cpu["sockets"], "cpus_allowed": cpu["cpus_allowed"],
            "physical_cores_allowed": cpu["physical_cores_allowed"], "threads": threads,
            "ms_per_step": round(ms_all, 2), "runs_ms_per_step": [round(x * 1e3, 2) for x in runs],
            "one_thread": {"value": nsd(ms_one), "unit": "ns/day", "ms_per_step": round(ms_one, 1),
                           "runs_ms_per_step": [round(x * 1e3, 1) for x in one]}}


# ---------------------------------------------------------------------------------------------
# accuracy against the FP64 oracle


def engine_accuracy(eng, box, p, list_age=True, virial=True):
    """Relative RMS force error and potential-energy error of the engine's current state against the
    FP64 oracle (and, when asked, of the step that uses the oldest list of an interval)."""
    import lima_amd as L
    from oracle.oracle import Oracle
    from tests.helpers import rel_rms, layout_forces, max_list_age_error
    out = {}
    if list_age:
        out["force_rel_rms_vs_fp64_max_list_age"] = float(f"{max_list_age_error(eng, box, p):.3e}")
    st = eng.state()
    sbox = L.StateBox(box, st)
    e_gpu = eng.compute_forces()
    f_gpu = layout_forces(sbox, eng.state())
    f_ref, _, e_ref = Oracle(sbox, p).forces()
    out["force_rel_rms_vs_fp64"] = float(f"{rel_rms(f_gpu, f_ref):.3e}")
    out["potential_rel_err_vs_fp64"] = float(f"{abs(e_gpu.potential - e_ref.potential) / abs(e_ref.potential):.3e}")
    if virial:   # relative to the largest per-term virial component (cutoff, intra, bonded, PME)
        import numpy as np
        scale = max(np.abs(np.array(Oracle(sbox, p).forces(c)[2].virial)).max() for c in (1, 2, 4, 8))
        err = np.abs(np.array(e_gpu.virial) - np.array(e_ref.virial)).max() / scale
        out["virial_max_abs_err_rel_term_scale"] = float(f"{err:.3e}")
    return out, sbox


def c0_accuracy(args):
    """Config 0 (1,000 waters, 3.104 nm, 1 fs): forces and energy of the start state and after 1,000
    steps against the oracle."""
    import lima_amd as L
    box = L.gen_water_lattice(10, 3.104, 0, 300.0)
    p = L.default_params(dt_fs=1.0)
    eng = L.Engine(box, p)
    a0, _ = engine_accuracy(eng, box, p, list_age=False)
    eng.step(1000)
    a1, _ = engine_accuracy(eng, box, p, list_age=False)
    return {"start": a0, "after_1000_steps": a1}


def c2_accuracy(n_systems=8):
    """Config 2 (batched minimisation): the first systems of the batch, start forces and energies
    from the batch kernel against the oracle, and the energy the batch reaches against the
    oracle's energy of the same final state."""
    import numpy as np
    import lima_amd as L
    from oracle.oracle import Oracle
    from tests.helpers import rel_rms, layout_forces
    sizes = L.em_batch_sizes(256, seed=1000)[:n_systems]
    boxes = [L.gen_em_system(int(n), 1000 + k, 0.02) for k, n in enumerate(sizes)]
    p = L.default_params()
    b = L.Batch(boxes, p)
    res0, _ = b.minimize(100.0, 0, 0.01)
    ferr, eerr = [], []
    for k, box in enumerate(boxes):
        f_ref, _, e_ref = Oracle(box, p).forces()
        ferr.append(rel_rms(layout_forces(box, b.state(k)), f_ref))
        eerr.append(abs(res0[k].final_energy - e_ref.potential) / abs(e_ref.potential))
    b.reset()
    res1, _ = b.minimize(100.0, 20000, 0.01)
    efin = []
    for k, box in enumerate(boxes):
        e_ref = Oracle(L.StateBox(box, b.state(k)), p).forces()[2]
        efin.append(abs(res1[k].final_energy - e_ref.potential) / abs(e_ref.potential))
    return {"systems": n_systems, "atoms": [int(x.n_particles) for x in boxes],
            "start_force_rel_rms_vs_fp64_max": float(f"{max(ferr):.3e}"),
            "start_potential_rel_err_vs_fp64_max": float(f"{max(eerr):.3e}"),
            "final_potential_rel_err_vs_fp64_max": float(f"{max(efin):.3e}")}


# ---------------------------------------------------------------------------------------------
# the MD legs


def timed_window(eng, steps):
    """Host clock around `steps` steps issued by one lima_step call, device idle at both ends."""
    eng.synchronize()
    t0 = time.perf_counter()
    eng.step(steps)
    eng.synchronize()
    return time.perf_counter() - t0


def per_call_window(eng, steps):
    """Host clock around `steps` calls of lima_step(h, 1), as LIMA's Environment::run loop drives
    Engine::step() (Environment.cpp:196-213), device idle at both ends."""
    eng.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.step(1)
    eng.synchronize()
    return time.perf_counter() - t0


def phase_times(eng, steps, nlist_interval):
    """Per-phase GPU times from HIP events recorded on the streams each phase runs on."""
    r0 = eng.status().n_list_rebuilds
    eng.set_timing(True)
    eng.step(steps)
    tm = eng.timings()
    eng.set_timing(False)
    rebuilds = eng.status().n_list_rebuilds - r0
    n = max(tm.steps_timed, 1)
    return {"pair": tm.pair_ms / n, "intra_bonded": (tm.intra_ms + tm.bonded_ms) / n, "pme": tm.pme_ms / n,
            "pack_integrate": tm.integrate_ms / n, "step_total": tm.total_ms / n,
            "nlist_rebuild_per_step": tm.nlist_ms / n}, tm.nlist_ms / max(rebuilds, 1), rebuilds


def rooflines(box, eng, phases, rebuild_ms, workload, K):
    import numpy as np
    pairs = eng.pair_count()
    pair_ms = phases["pair"]
    tf = pairs * FLOPS_PER_PAIR / (pair_ms * 1e-3) / 1e12 if pair_ms > 0 else 0.0
    roof = {"kernel": "k_pair", "bound": "fp32 valu", "pipe": "FP32 vector ALU (no matrix instructions)",
            "achieved": round(tf, 3), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(tf / FP32_PEAK_TFLOPS, 5),
            "traffic": load_traffic("k_pair", workload),
            "work": f"{pairs} pairs inside 1.2 nm x {FLOPS_PER_PAIR} flop / {pair_ms * 1e3:.2f} us per launch"}
    n_q = int(np.count_nonzero(box.charge.reshape(-1)[box.layout_index()[box.layout_index() < box.n_compounds * 32]]))
    n_grid = K[0] * K[1] * K[2]
    pme_bytes = 48 * n_q + 34 * n_grid
    pme_gbs = pme_bytes / (phases["pme"] * 1e-3) / 1e9 if phases["pme"] > 0 else 0.0
    roof_pme = {"kernel": "PME chain per step (k_spread_tiles, k_fft_x_fwd_tiled, 3 x k_fft_strided, k_fft_x_inv, "
                          "k_gather_tiles)", "bound": "hbm", "achieved": round(pme_gbs, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(pme_gbs / HBM_PEAK_GBS, 5), "traffic": load_traffic("pme_chain", workload),
                "work": f"48 B x {n_q} charges + 34 B x {n_grid} grid points ({K[0]}x{K[1]}x{K[2]}) = "
                        f"{pme_bytes / 1e6:.2f} MB / {phases['pme'] * 1e3:.1f} us"}
    st = eng.status()
    entries = int(st.list_entries)
    nl_bytes = 36 * box.n_particles + 4 * entries
    nl_gbs = nl_bytes / (rebuild_ms * 1e-3) / 1e9 if rebuild_ms > 0 else 0.0
    roof_nl = {"kernel": "list rebuild (columns, sort, cluster bounds, k_nlist)", "bound": "hbm",
               "achieved": round(nl_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(nl_gbs / HBM_PEAK_GBS, 5),
               "traffic": load_traffic("nlist_rebuild", workload),
               "work": f"36 B x {box.n_particles} atoms + 4 B x {entries} stored cluster entries = {nl_bytes / 1e6:.2f} MB "
                       f"/ {rebuild_ms * 1e3:.1f} us per rebuild"}
    return roof, roof_pme, roof_nl, pairs


def water_1m_leg(args):
    """Config 4: 333,333 3-site waters plus one neutral LJ particle in a 21.5 nm cube, PME 180^3,
    dt = 2 fs: ns/day, k_pair against the FP32 peak, PME and the list rebuild against HBM, and the
    accuracy against the oracle."""
    import lima_amd as L
    t = time.time()
    box = L.gen_large_water()
    dt_fs = 2.0
    p = L.default_params(dt_fs=dt_fs, n_steps=args.steps_1m, energy_interval=0, launch_mode=args.launch_mode)
    eng = L.Engine(box, p)
    r = eng.minimize(1000.0, args.em_iters_1m, 0.01)
    eng.step(args.equil_steps_1m)   # equilibration, untimed
    eng.step(args.warmup_1m)
    log(f"water_1m: {box.n_particles} atoms, {r.iterations} SD iterations, set up in {time.time() - t:.1f}s")
    elapsed = timed_window(eng, args.steps_1m)
    ms = elapsed * 1e3 / args.steps_1m
    nsd = args.steps_1m / elapsed * dt_fs * 86400.0 / 1e6
    log(f"water_1m: {args.steps_1m} steps, {ms:.4f} ms/step, {nsd:.1f} ns/day")
    phases, rebuild_ms, _ = phase_times(eng, args.timing_steps_1m, p.nlist_interval)
    eng.compute_forces()
    K = pme_grid_dims(box.box_nm)
    roof, roof_pme, roof_nl, pairs = rooflines(box, eng, phases, rebuild_ms, "water_1m", K)
    acc = None
    if not args.no_accuracy:
        t = time.time()
        acc, _ = engine_accuracy(eng, box, p, list_age=False, virial=False)
        log(f"water_1m accuracy {acc} ({time.time() - t:.1f}s)")
    return {"value": round(nsd, 3), "unit": "ns/day", "steps": args.steps_1m, "warmup": args.warmup_1m,
            "ms_per_step": round(ms, 5),
            "config": {"workload": "water_1m", "n_atoms": box.n_particles, "box_nm": [float(x) for x in box.box_nm],
                       "waters": 333333, "extra_lj_particles": 1, "dt_fs": dt_fs, "cutoff_nm": 1.2, "pme_grid": K},
            "roofline": roof, "roofline_pme": roof_pme, "roofline_nlist": roof_nl, "accuracy": acc,
            "phases_ms_per_step": {k: round(v, 5) for k, v in phases.items()}, "pairs_in_cutoff": pairs}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=500)
    ap.add_argument("--em-iters", type=int, default=400)
    ap.add_argument("--equil-steps", type=int, default=1000,
                    help="untimed MD steps after the relaxation, before the warm-up")
    ap.add_argument("--equil-steps-1m", type=int, default=200)
    ap.add_argument("--timing-steps", type=int, default=1000)
    ap.add_argument("--steps-1m", type=int, default=300)
    ap.add_argument("--warmup-1m", type=int, default=60)
    ap.add_argument("--timing-steps-1m", type=int, default=100)
    ap.add_argument("--em-iters-1m", type=int, default=200)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-1m", action="store_true", help="skip the 1M-atom water box leg")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run the copy kernel used to calibrate the HBM byte counters")
    ap.add_argument("--no-em", action="store_true", help="skip the batched-minimisation measurement")
    ap.add_argument("--per-call-steps", type=int, default=1000,
                    help="steps of the per-call comparison (at least --steps)")
    ap.add_argument("--launch-mode", type=int, default=0,
                    help="lima_simparams.launch_mode: 0 step graph replay with k_local and PME on a second stream; "
                         "+1 eager launches, +2 PME in line on one stream")
    ap.add_argument("--em-systems", type=int, default=256)
    ap.add_argument("--em-train-systems", type=int, default=100)
    ap.add_argument("--em-max-iter", type=int, default=20000)
    ap.add_argument("--em-stride", type=int, default=64, help="record every k-th atom for training")
    ap.add_argument("--em-epochs", type=int, default=4)
    ap.add_argument("--only", default=None, choices=["chain_water_c1", "water_1m"],
                    help="run one MD leg alone (profiling)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: this path has no multi-GPU mode (--gpus must be 1)", file=sys.stderr)
        return 2
    if args.calibrate_copy:
        return calibrate_copy()
    if args.only == "water_1m":
        print(json.dumps(water_1m_leg(args)), flush=True)
        return 0

    import lima_amd as L

    t = time.time()
    box = L.gen_chain_water()
    log(f"generated chain+water box: {box.n_particles} atoms, {box.n_compounds} compounds in {time.time() - t:.2f}s")
    dt_fs = 2.0
    p = L.default_params(dt_fs=dt_fs, n_steps=args.steps, energy_interval=0, launch_mode=args.launch_mode)
    eng = L.Engine(box, p)
    t = time.time()
    r = eng.minimize(1000.0, args.em_iters, 0.01)
    log(f"relaxed: {r.iterations} SD iterations, fmax {r.final_fmax:.1f} kJ/mol/nm, {time.time() - t:.1f}s")
    # equilibration MD (untimed, part of preparing the workload): the relaxed structure heats up and
    # the list lifetime settles at the longest interval that keeps the lists complete
    t = time.time()
    eng.step(args.equil_steps)
    eng.synchronize()
    s0 = eng.status()
    log(f"equilibrated {args.equil_steps} steps in {time.time() - t:.2f}s: lists live {s0.list_interval} steps, "
        f"{s0.list_stale_rollbacks} intervals rerun")
    t = time.time()
    eng.step(args.warmup)
    eng.synchronize()
    log(f"warm-up {args.warmup} steps in {time.time() - t:.2f}s")

    host0 = eng.timings()
    elapsed = timed_window(eng, args.steps)
    host1 = eng.timings()
    host_issue_us = (host1.host_issue_ms - host0.host_issue_ms) * 1e3 / max(host1.host_steps - host0.host_steps, 1)
    ms_per_step = elapsed * 1e3 / args.steps
    ns_day = args.steps / elapsed * dt_fs * 86400.0 / 1e6
    log(f"timed {args.steps} steps: {ms_per_step:.4f} ms/step, {ns_day:.1f} ns/day; host issue {host_issue_us:.1f} us/step")
    # the per-call path against the multi-step path, on a sample long enough that list rebuilds and
    # the occasional rerun of an interval land in both alike
    n_pc = max(args.steps, args.per_call_steps)
    ms_multi = timed_window(eng, n_pc) * 1e3 / n_pc
    pc_ms = per_call_window(eng, n_pc) * 1e3 / n_pc
    log(f"per-call: {n_pc} x lima_step(h, 1): {pc_ms:.4f} ms/step ({pc_ms / ms_multi:.3f} x one {n_pc}-step call, "
        f"{ms_multi:.4f} ms/step)")
    stt = eng.status()
    log(f"list: {stt.n_list_rebuilds} rebuilds, lists live {stt.list_interval} steps, {stt.list_stale_rollbacks} intervals "
        f"rerun, {stt.list_violations} kept intervals over the bound, last max |dx| {stt.max_displacement_nm:.4f} nm")

    phases, rebuild_ms, rebuilds = phase_times(eng, args.timing_steps, p.nlist_interval)
    log("phase ms/step: " + ", ".join(f"{k} {v:.4f}" for k, v in phases.items()) + f"; {rebuild_ms:.4f} ms per rebuild")
    e_end = eng.compute_forces()
    workload = "chain_water_c1"
    roof, roof_pme, roof_nl, pairs = rooflines(box, eng, phases, rebuild_ms, workload, pme_grid_dims(box.box_nm))
    stt = eng.status()
    list_slots = stt.list_entries * 8

    accuracy = {}
    cpu = None
    if not (args.no_accuracy and args.no_cpu_baseline):
        from oracle.oracle import set_threads
        info = host_cpu()
        threads = min(int(os.environ.get("OMP_NUM_THREADS", "0")) or info["physical_cores_allowed"],
                      info["physical_cores_allowed"])
        set_threads(threads)
        sbox = None
        if not args.no_accuracy:
            t = time.time()
            accuracy["chain_water_c1"], sbox = engine_accuracy(eng, box, p)
            log(f"C1 accuracy {accuracy['chain_water_c1']} ({time.time() - t:.1f}s)")
        if not args.no_cpu_baseline:
            if sbox is None:
                sbox = L.StateBox(box, eng.state())
            cpu = cpu_baseline_c1(sbox, p, dt_fs, threads, info)
        if not args.no_accuracy:
            t = time.time()
            accuracy["water_c0"] = c0_accuracy(args)
            accuracy["em_batch_c2"] = c2_accuracy()
            log(f"C0 / C2 accuracy ({time.time() - t:.1f}s): {accuracy['water_c0']}, {accuracy['em_batch_c2']}")

    w1m = None
    if not args.no_1m:
        w1m = water_1m_leg(args)
        if w1m.get("accuracy"):
            accuracy["water_1m_c4"] = w1m["accuracy"]

    em = None
    if not args.no_em:
        em = em_bench(args)

    out = {
        "metric": "ns/day (42k-atom box); time to Fmax<100 kJ/mol/nm, learned vs plain EM",
        "value": round(ns_day, 3),
        "unit": "ns/day",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": workload, "n_atoms": box.n_particles, "box_nm": [float(x) for x in box.box_nm],
                   "chains": "128 x 50", "waters": (box.n_particles - 6400) // 3, "dt_fs": dt_fs, "cutoff_nm": 1.2,
                   "pme_spacing_nm": 0.12, "pme_grid": pme_grid_dims(box.box_nm), "nlist_interval_max": p.nlist_interval,
                   "equilibration_steps": args.equil_steps,
                   "skin_nm": round(p.nlist_skin_nm, 4), "energy_reductions": "fp64",
                   "launch": ("eager" if args.launch_mode & 1 else "step graph") +
                             (", one stream" if args.launch_mode & 2 else ", k_local and PME on a second stream")},
        "roofline": roof,
        "roofline_pme": roof_pme,
        "roofline_nlist": roof_nl,
        "cpu_baseline": cpu,
        "per_call": {"ms_per_step": round(pc_ms, 5), "value": round(1e3 / pc_ms * dt_fs * 86400.0 / 1e6, 3),
                     "ratio_to_multistep": round(pc_ms / ms_multi, 4), "multistep_ms_per_step": round(ms_multi, 5),
                     "sample": f"{n_pc} lima_step(h, 1) calls against one {n_pc}-step call, both after the timed window"},
        "accuracy": accuracy or None,
        "phases_ms_per_step": {k: round(v, 5) for k, v in phases.items()},
        "nlist_rebuild_ms": round(rebuild_ms, 5),
        "host_issue_us_per_step": round(host_issue_us, 2),
        "pairs_in_cutoff": pairs,
        "pair_slots_evaluated": list_slots,
        "cluster_pair_efficiency": round(2 * pairs / list_slots, 4) if list_slots else None,
        "list_rebuilds": stt.n_list_rebuilds,
        "list_interval": stt.list_interval,
        "list_intervals_rerun": stt.list_stale_rollbacks,
        "list_kept_intervals_over_bound": stt.list_violations,
        "final_energy_kj": round(e_end.total / 1000.0, 3),
        "water_1m": w1m,
        "em_batch": em,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
