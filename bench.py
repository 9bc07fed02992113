#!/usr/bin/env python3
"""bench.py -- ns/day of the MI355X-native LIMA step (force path + integrator) on one GPU.

Workload: the chain + water box of the mixed-precision config (8 x 8 x 7 nm, 128 chains of 50
atoms in two leaflets along z, waters filling the rest; seed 1), cutoff 1.2 nm, PME spacing
0.12 nm (72 x 72 x 60 grid), dt = 2 fs, FP32 forces with FP64 energy reductions.  The
generated structure is relaxed with steepest descent (untimed), W warm-up steps run, then K
steps are timed.  lima_step returns after the device has finished, so the host clock around
it brackets exactly the K steps.

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
FLOPS_PER_PAIR = 64                # algorithmic FP32 flops per pair inside the cutoff
COPY_CALIBRATION_BYTES = 512 << 20


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def load_traffic(kernel, workload):
    """HBM bytes per launch of `kernel`, from the committed counter summary
    profiles/traffic_<kernel>_<workload>.txt (line 'traffic_bytes_per_launch = N'), or None."""
    path = ROOT / "profiles" / f"traffic_{kernel}_{workload}.txt"
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch"):
            try:
                return int(float(line.split("=")[1]))
            except ValueError:
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
    # control: the same learned path with a model whose output is a constant scale, which makes every
    # proposal the plain step; it separates the path's overhead from the effect of the trained scales
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


def large_bench(args):
    """The 1M-atom plain-dynamics workload: 333,333 3-site waters plus one neutral LJ particle in a
    21.5 nm cube (seed 4), cutoff 1.2 nm, PME 180^3 (0.12 nm), dt = 2 fs.  Reports ns/day, per-phase
    kernel times, the pair kernel against the FP32 peak and the PME chain against HBM bandwidth."""
    import numpy as np
    import lima_amd as L
    t = time.time()
    box = L.gen_large_water()
    log(f"generated 1M-atom water box: {box.n_particles} atoms, {box.box_nm[0]:.3f} nm in {time.time() - t:.1f}s")
    dt_fs = 2.0
    p = L.default_params(dt_fs=dt_fs, n_steps=args.steps, energy_interval=0, launch_mode=args.launch_mode)
    eng = L.Engine(box, p)
    t = time.time()
    r = eng.minimize(1000.0, args.em_iters, 0.01)
    log(f"relaxed: {r.iterations} SD iterations, fmax {r.final_fmax:.1f} kJ/mol/nm, {time.time() - t:.1f}s")
    t = time.time()
    eng.step(args.warmup)
    log(f"warm-up {args.warmup} steps in {time.time() - t:.2f}s")
    eng.synchronize()
    t0 = time.perf_counter()
    eng.step(args.steps)
    eng.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed * 1e3 / args.steps
    ns_day = args.steps / elapsed * dt_fs * 86400.0 / 1e6
    log(f"timed {args.steps} steps: {ms_per_step:.4f} ms/step, {ns_day:.1f} ns/day")
    pairs = eng.pair_count()
    eng.set_timing(True)
    eng.step(args.timing_steps)
    tm = eng.timings()
    eng.set_timing(False)
    n = max(tm.steps_timed, 1)
    phases = {"pair": tm.pair_ms / n, "intra": tm.intra_ms / n, "bonded": tm.bonded_ms / n, "pme": tm.pme_ms / n,
              "pack_integrate": tm.integrate_ms / n, "step_total": tm.total_ms / n}
    log("phase ms/step: " + ", ".join(f"{k} {v:.4f}" for k, v in phases.items()))
    charges = box.charge.reshape(-1)
    n_q = int(np.count_nonzero(charges))
    K = pme_grid_dims(box.box_nm)
    n_grid = K[0] * K[1] * K[2]
    pme_bytes = 48 * n_q + 34 * n_grid
    pme_gbs = pme_bytes / (phases["pme"] * 1e-3) / 1e9
    pair_tf = pairs * FLOPS_PER_PAIR / (phases["pair"] * 1e-3) / 1e12
    roofline = {"kernel": "k_pair", "bound": "mfma", "pipe": "fp32 vector ALU (no matrix instructions)",
                "achieved": round(pair_tf, 3), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(pair_tf / FP32_PEAK_TFLOPS, 5), "traffic": None,
                "work": f"{pairs} pairs inside 1.2 nm x {FLOPS_PER_PAIR} flop / {phases['pair'] * 1e3:.1f} us"}
    roofline_pme = {"kernel": "PME chain (bin, spread, FFT passes, gather)", "bound": "hbm",
                    "achieved": round(pme_gbs, 1), "peak": 8000.0, "unit": "GB/s", "frac": round(pme_gbs / 8000.0, 5),
                    "work": f"48 B x {n_q} charges + 34 B x {n_grid} grid points ({K[0]}x{K[1]}x{K[2]}) = "
                            f"{pme_bytes / 1e6:.1f} MB / {phases['pme'] * 1e3:.1f} us"}
    accuracy = cpu = None
    if not (args.no_accuracy and args.no_cpu_baseline):
        from oracle.oracle import Oracle, set_threads
        from tests.helpers import rel_rms, layout_forces
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        set_threads(threads)
        if not args.no_accuracy:
            from tests.helpers import max_list_age_error
            stale = max_list_age_error(eng, box, p)
            log(f"force rel RMS at list age {p.nlist_interval - 1} (the oldest list a step uses): {stale:.2e}")
        sbox = L.StateBox(box, eng.state())
        if not args.no_accuracy:
            e_gpu = eng.compute_forces()
            f_gpu = layout_forces(sbox, eng.state())
            t = time.time()
            f_ref, _, e_ref = Oracle(sbox, p).forces()
            err = rel_rms(f_gpu, f_ref)
            e_rel = abs(e_gpu.potential - e_ref.potential) / abs(e_ref.potential)
            accuracy = {"force_rel_rms_vs_fp64": float(f"{err:.3e}"), "potential_rel_err_vs_fp64": float(f"{e_rel:.3e}"),
                        "force_rel_rms_vs_fp64_max_list_age": float(f"{stale:.3e}")}
            log(f"accuracy vs FP64 oracle: force rel RMS {err:.2e}, potential rel {e_rel:.2e} ({time.time() - t:.1f}s)")
        if not args.no_cpu_baseline:
            orc = Oracle(sbox, p)
            t = time.perf_counter()
            orc.step(1)
            dtc = time.perf_counter() - t
            cpu = {"value": round(1 / dtc * dt_fs * 86400.0 / 1e6, 5), "unit": "ns/day", "cores": threads, "kind": "port",
                   "sample": f"1 FP64 oracle MD step of the same box ({dtc:.1f}s)"}
            log(f"cpu baseline: {cpu}")
    return {
        "metric": "ns/day (1M-atom water box)",
        "value": round(ns_day, 3), "unit": "ns/day", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "water_1m", "n_atoms": box.n_particles, "box_nm": [float(x) for x in box.box_nm],
                   "waters": 333333, "extra_lj_particles": 1, "dt_fs": dt_fs, "cutoff_nm": 1.2, "pme_grid": K,
                   "nlist_interval": p.nlist_interval, "skin_nm": round(p.nlist_skin_nm, 4), "energy_reductions": "fp64"},
        "roofline": roofline, "roofline_pme": roofline_pme, "cpu_baseline": cpu, "accuracy": accuracy,
        "phases_ms_per_step": {k: round(v, 5) for k, v in phases.items()}, "pairs_in_cutoff": pairs,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=500)
    ap.add_argument("--em-iters", type=int, default=400)
    ap.add_argument("--timing-steps", type=int, default=1000)
    ap.add_argument("--cpu-steps", type=int, default=8, help="oracle MD steps in the CPU baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run the copy kernel used to calibrate the HBM byte counters")
    ap.add_argument("--no-em", action="store_true", help="skip the batched-minimisation measurement")
    ap.add_argument("--launch-mode", type=int, default=0,
                    help="lima_simparams.launch_mode: 0 step graph replay with PME on a second stream; "
                         "+1 eager launches, +2 PME in line on one stream")
    ap.add_argument("--em-systems", type=int, default=256)
    ap.add_argument("--em-train-systems", type=int, default=100)
    ap.add_argument("--em-max-iter", type=int, default=20000)
    ap.add_argument("--em-stride", type=int, default=64, help="record every k-th atom for training")
    ap.add_argument("--em-epochs", type=int, default=4)
    ap.add_argument("--workload", default="chain_water_c1", choices=["chain_water_c1", "water_1m"],
                    help="chain_water_c1: the headline 42k-atom-config box (default); water_1m: the 1M-atom water box "
                         "(uses --steps/--warmup/--timing-steps/--em-iters; defaults 300/100/100/200 when not given)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: this path has no multi-GPU mode (--gpus must be 1)", file=sys.stderr)
        return 2
    if args.calibrate_copy:
        return calibrate_copy()
    if args.workload == "water_1m":
        given = set(a.split("=")[0] for a in sys.argv[1:])
        for flag, attr, val in (("--steps", "steps", 300), ("--warmup", "warmup", 100),
                                ("--timing-steps", "timing_steps", 100), ("--em-iters", "em_iters", 200)):
            if flag not in given:
                setattr(args, attr, val)
        print(json.dumps(large_bench(args)), flush=True)
        return 0

    import lima_amd as L

    t = time.time()
    box = L.gen_chain_water()
    log(f"generated chain+water box: {box.n_particles} atoms, {box.n_compounds} compounds in {time.time() - t:.2f}s")
    dt_fs = 2.0
    p = L.default_params(dt_fs=dt_fs, n_steps=args.steps, energy_interval=0, launch_mode=args.launch_mode)
    eng = L.Engine(box, p)
    log("engine created")
    t = time.time()
    r = eng.minimize(1000.0, args.em_iters, 0.01)
    log(f"relaxed: {r.iterations} SD iterations, fmax {r.final_fmax:.1f} kJ/mol/nm, {time.time() - t:.1f}s")
    t = time.time()
    eng.step(args.warmup)
    log(f"warm-up {args.warmup} steps in {time.time() - t:.2f}s")

    host0 = eng.timings()
    t0 = time.perf_counter()
    eng.step(args.steps)
    eng.synchronize()
    t1 = time.perf_counter()
    host1 = eng.timings()
    host_issue_us = (host1.host_issue_ms - host0.host_issue_ms) * 1e3 / max(host1.host_steps - host0.host_steps, 1)
    elapsed = t1 - t0
    ms_per_step = elapsed * 1e3 / args.steps
    ns_day = args.steps / elapsed * dt_fs * 86400.0 / 1e6
    log(f"timed {args.steps} steps: {ms_per_step:.4f} ms/step, {ns_day:.1f} ns/day; host issue {host_issue_us:.1f} us/step")
    e_end = eng.compute_forces()
    pairs = eng.pair_count()
    stt = eng.status()
    list_slots = stt.list_entries * 8
    log(f"list: {stt.n_list_rebuilds} rebuilds, {stt.list_violations} intervals with 2 max|dx| > skin, "
        f"last max |dx| {stt.max_displacement_nm:.4f} nm")

    # per-phase kernel times with HIP events recorded on the engine's stream (no host syncs)
    eng.set_timing(True)
    eng.step(args.timing_steps)
    tm = eng.timings()
    eng.set_timing(False)
    n = max(tm.steps_timed, 1)
    pair_ms = tm.pair_ms / n
    phases = {"pair": pair_ms, "intra": tm.intra_ms / n, "bonded": tm.bonded_ms / n, "pme": tm.pme_ms / n,
              "pack_integrate": tm.integrate_ms / n, "step_total": tm.total_ms / n,
              "nlist_rebuild_per_step": tm.nlist_ms / n}
    log("phase ms/step: " + ", ".join(f"{k} {v:.4f}" for k, v in phases.items()))
    flops = pairs * FLOPS_PER_PAIR
    achieved = flops / (pair_ms * 1e-3) / 1e12 if pair_ms > 0 else 0.0
    workload = "chain_water_c1"
    roofline = {"kernel": "k_pair", "bound": "mfma", "pipe": "fp32 vector ALU (no matrix instructions)",
                "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 5),
                "traffic": load_traffic("k_pair", workload),
                "work": f"{pairs} pairs inside 1.2 nm x {FLOPS_PER_PAIR} flop / {pair_ms * 1e3:.2f} us"}

    accuracy = None
    cpu = None
    if not (args.no_accuracy and args.no_cpu_baseline):
        from oracle.oracle import Oracle, set_threads
        import numpy as np
        from tests.helpers import rel_rms, layout_forces
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        set_threads(threads)
        if not args.no_accuracy:
            from tests.helpers import max_list_age_error
            stale = max_list_age_error(eng, box, p)
            log(f"force rel RMS at list age {p.nlist_interval - 1} (the oldest list a step uses): {stale:.2e}")
        st = eng.state()
        sbox = L.StateBox(box, st)
        if not args.no_accuracy:
            e_gpu = eng.compute_forces()
            f_gpu = layout_forces(sbox, eng.state())
            t = time.time()
            orc = Oracle(sbox, p)
            f_ref, _, e_ref = orc.forces()
            err = rel_rms(f_gpu, f_ref)
            e_rel = abs(e_gpu.potential - e_ref.potential) / abs(e_ref.potential)
            accuracy = {"force_rel_rms_vs_fp64": float(f"{err:.3e}"), "potential_rel_err_vs_fp64": float(f"{e_rel:.3e}"),
                        "force_rel_rms_vs_fp64_max_list_age": float(f"{stale:.3e}")}
            log(f"accuracy vs FP64 oracle: force rel RMS {err:.2e}, potential rel {e_rel:.2e} ({time.time() - t:.1f}s)")
        if not args.no_cpu_baseline:
            orc = Oracle(sbox, p)
            orc.step(1)   # warm-up
            t = time.perf_counter()
            orc.step(args.cpu_steps)
            dtc = time.perf_counter() - t
            cpu = {"value": round(args.cpu_steps / dtc * dt_fs * 86400.0 / 1e6, 4), "unit": "ns/day", "cores": threads,
                   "kind": "port",
                   "sample": f"{args.cpu_steps} FP64 oracle MD steps of the same box after 1 warm-up step "
                             f"({dtc:.1f}s, {dtc / args.cpu_steps * 1e3:.0f} ms/step)"}
            log(f"cpu baseline: {cpu}")

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
                   "pme_spacing_nm": 0.12, "nlist_interval": p.nlist_interval, "skin_nm": round(p.nlist_skin_nm, 4),
                   "energy_reductions": "fp64",
                   "launch": ("eager" if args.launch_mode & 1 else "step graph") +
                             (", one stream" if args.launch_mode & 2 else ", PME on a second stream")},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "accuracy": accuracy,
        "phases_ms_per_step": {k: round(v, 5) for k, v in phases.items()},
        "host_issue_us_per_step": round(host_issue_us, 2),
        "pairs_in_cutoff": pairs,
        "pair_slots_evaluated": list_slots,
        "cluster_pair_efficiency": round(2 * pairs / list_slots, 4) if list_slots else None,
        "list_rebuilds": stt.n_list_rebuilds,
        "list_stale_intervals": stt.list_violations,
        "final_energy_kj": round(e_end.total / 1000.0, 3),
        "em_batch": em,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
