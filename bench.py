"""Benchmark: FDTD room-acoustics step on one MI355X.

A step is one leapfrog time step of the furnished room (256 x 192 x 128
cells, 4 IIR wall materials, 12 seeded furniture boxes, 12-cell PML on one
face, band-limited noise source, 16 receivers), FP32, physical boundary
mode, run through the C ABI.  Prints one JSON line; its extra legs are the
48-room batch (batched), the learned surrogate over uniform positions
(surrogate) and over a loudspeaker zone (surrogate_zone), and the
640 x 480 x 320 room in every storage precision and with two steps per pass
(large).

  python bench.py [--steps K] [--warmup W] [--precision fp32|fp16|bf16|fp64]
                  [--compat] [--no-cpu-baseline] [--no-tune] [--no-batched] [--no-surrogate]
                  [--no-zone] [--no-large] [--large-precisions LIST] [--copy-probe BYTES]
                  [--ie224 SECONDS] [--mlp-train STEPS] [--mlp-inference ROWS]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "anc-field_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0


def progress(msg):
    """One line on stderr as each part of the run ends (stdout carries only
    the JSON result)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)
BYTES_PER_CELL = {"fp32": 12, "fp64": 24, "fp16": 6, "bf16": 6}


def host_cores():
    """(threads to use, CPUs visible): every CPU this process may run on
    (its affinity set), limited to the CPU share the host grants it (the
    cgroup cpu.max quota, and OMP_NUM_THREADS when the host sets it)."""
    visible = len(os.sched_getaffinity(0))
    share = visible
    try:
        quota, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if quota != "max":
            share = min(share, max(1, int(int(quota) // int(period))))
    except (OSError, ValueError):
        pass
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        share = min(share, int(omp))
    return share, visible


def cpu_baseline(w, budget_s=12.0):
    """Time the CPU oracle (OpenMP C restatement) on the same room for a
    bounded number of steps on every host core granted to this process."""
    import numpy as np
    import orc
    cores, visible = host_cores()
    o = orc.from_workload(w, compat=False, threads=cores)
    o.run(5)  # first touch
    steps, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < budget_s:
        o.run(10)
        steps += 10
    dt = time.perf_counter() - t0
    o.close()
    return {"value": round(w.cells * steps / dt / 1e9, 4), "unit": "Gcell-updates/s", "cores": cores,
            "cpus_visible": visible, "kind": "port",
            "sample": f"{w.name} {w.dims[0]}x{w.dims[1]}x{w.dims[2]}, {steps} steps after 5 warm-up steps, "
                      f"FP32 oracle (C, OpenMP) on {cores} host threads"}


def configs_leg():
    """The single-room configurations over their full runs: the 64x48x32
    shoebox (8-cell PML on x+, Ricker source, 4 receivers, 2000 steps) and
    the furnished room (10 000 steps).  FP32 throughput, FP32 receiver error
    against FP64, and for the shoebox the CPU oracle over the same 2000 steps
    on this host's cores, its traces compared with the GPU's bit for bit."""
    import numpy as np
    import orc
    from anc_field_py import _afx, rooms
    out = {}
    for name, w in (("shoebox_64x48x32_2000_steps", rooms.shoebox_small(steps=2000)),
                    ("furnished_256x192x128_10000_steps", rooms.furnished(steps=10000))):
        tr, gps, tuned = {}, 0.0, None
        for prec in ("fp32", "fp64"):
            s = rooms.make_sim(w, _afx.Sim, precision=prec)
            try:
                if prec == "fp32":
                    # as a user would run it: autotuned (cached when the main leg tuned this room)
                    tuned = list(s.autotune(30)) + [s.time_block()[0], s.tile_split()]
                s.run(1)
                s.synchronize()
                t0 = time.perf_counter()
                s.run(w.steps - 1)
                s.synchronize()
                dt = time.perf_counter() - t0
                tr[prec] = s.receivers()
            finally:
                s.close()
            if prec == "fp32":
                gps = w.cells * (w.steps - 1) / dt / 1e9
            progress(f"configs: {name} {prec} done")
        a, b = tr["fp32"].astype(np.float64), tr["fp64"].astype(np.float64)
        ent = {"cells": w.cells, "steps": w.steps, "fp32_gcell_per_s": round(gps, 1),
               "fp32_autotuned_tile_steps_per_pass_split": tuned,
               "fp32_hbm_roofline_frac_algorithmic": round(gps * 12 / HBM_PEAK_GBS, 4),
               "rel_l2_fp32_vs_fp64": float(np.linalg.norm(a - b) / np.linalg.norm(b))}
        if w.cells < 10 ** 6:
            cores = host_cores()[0]
            o = orc.from_workload(w, compat=False, threads=cores)
            t0 = time.perf_counter()
            o.run(w.steps)
            dt = time.perf_counter() - t0
            ent["cpu_oracle_gcell_per_s"] = round(w.cells * w.steps / dt / 1e9, 3)
            ent["cpu_cores"] = cores
            ent["gpu_bitexact_vs_cpu_oracle"] = bool(np.array_equal(np.asarray(o.receivers(), np.float32), tr["fp32"]))
            o.close()
        out[name] = ent
    return out


def batched_leg(steps):
    """Batched identification: 48 seeded rooms (64^3 .. 256x192x128, 2-4 IIR
    materials, one Ricker source and 8 receivers each) in one handle, against
    the same 48 rooms run one after another in FP32; reduced-precision
    storage errors are stated against the FP64 run of the same batch."""
    from anc_field_py import batch, rooms
    ws = rooms.batch_rooms(steps=steps)
    n_ir = sum(len(w.mic_cells) for w in ws)
    cells = sum(w.cells for w in ws)
    seq, t_seq = batch.run_sequential(ws, steps, "fp32")
    # the reference's way of stepping: separate interior / boundary / source
    # launches every step, issued eagerly (simulation.hpp:358-431 issues ~29)
    seq_split, t_split = batch.run_sequential(ws, steps, "fp32", fused=False, graph=False)
    res = {"rooms": len(ws), "steps": steps, "impulse_responses": n_ir, "cells": cells,
           "sequential_fp32": {"ir_per_s": round(n_ir / t_seq, 2), "wall_s": round(t_seq, 3),
                               "kind": "one fused launch per step, graph replay"},
           "sequential_fp32_split_eager": {"ir_per_s": round(n_ir / t_split, 2), "wall_s": round(t_split, 3),
                                           "kind": "separate launches per step, no graphs",
                                           "bitexact_vs_fused": bool((seq_split == seq).all())}}
    runs = {}
    progress("batched: sequential runs done")
    for prec in ("fp32", "fp16", "bf16", "fp64"):
        runs[prec] = batch.run_batched(ws, steps, prec, tune=True)
        progress(f"batched: {prec} done")
    ref = runs["fp64"][0]
    for prec, (tr, t, tile) in runs.items():
        res[f"batched_{prec}"] = {
            "ir_per_s": round(n_ir / t, 2), "wall_s": round(t, 3), "speedup_vs_sequential_fp32": round(t_seq / t, 2),
            "speedup_vs_sequential_split_eager": round(t_split / t, 2),
            "gcell_per_s": round(cells * steps / t / 1e9, 1), "autotuned_tile": list(tile[:3]), "steps_per_pass": int(tile[3]),
            "rel_l2_vs_fp64": batch.rel_l2(tr, ref)}
    res["batched_fp32"]["bitexact_vs_sequential"] = bool((runs["fp32"][0] == seq).all())
    # how the reduced-precision error grows with the run length: relative L2
    # against FP64 over the first n steps of every trace
    import numpy as np
    for prec in ("fp32", "fp16", "bf16"):
        tr = np.asarray(runs[prec][0], np.float64)
        res[f"batched_{prec}"]["rel_l2_vs_fp64_first_steps"] = {
            str(n): batch.rel_l2(tr[:, :n], ref[:, :n]) for n in (256, 512, 1024, 2048, 4096) if n <= steps}
    # the same batch stepped in groups whose fields fit the Infinity Cache
    for prec in ("fp32", "fp16"):
        tr, t, groups = batch.run_grouped(ws, steps, prec, tune=True)
        res[f"batched_cache_groups_{prec}"] = {
            "ir_per_s": round(n_ir / t, 2), "wall_s": round(t, 3), "speedup_vs_sequential_fp32": round(t_seq / t, 2),
            "gcell_per_s": round(cells * steps / t / 1e9, 1), "groups": len(groups),
            "rel_l2_vs_fp64": batch.rel_l2(tr, ref)}
        if prec == "fp32":
            res[f"batched_cache_groups_{prec}"]["bitexact_vs_sequential"] = bool((tr == seq).all())
        progress(f"batched: cache groups {prec} done")
    return res


def surrogate_leg(train_steps, eager=False):
    """Learned surrogate in the furnished room (generator seed fixed): 64
    sources x 64 receivers = 4096 FDTD impulse responses for training and
    8 x 64 = 512 held out, positions uniform over free cells, each trace
    low-passed at 2 kHz and decimated by 4 to 1024 taps.  The 6 x 512
    Fourier-feature MLP (64 frequencies) trains on the BF16 MFMA kernels,
    batch 512 (15 000 steps, learning rate 2e-3 to 1e-5 on a cosine; batch
    256 over 30 000 steps at 1e-3 fits the same, in 1.75x the time).
    Held-out queries go through the validated-bound check and
    fall back to FDTD, which judges the network's answers."""
    import numpy as np
    from anc_field_py import rooms, surrogate
    w = rooms.furnished(seed=2, steps=4096)
    rng = np.random.Generator(np.random.PCG64(5))
    src = surrogate.free_cells(w, rng, 72)
    rcv = [surrogate.free_cells(w, rng, 64) for _ in range(72)]
    sur = surrogate.Surrogate(w, seed=1, batch=512, flags=1 if eager else 0)  # 1: AFX_MLP_NO_GRAPH
    try:
        t = time.perf_counter()
        ir = sur.simulate(src, rcv)
        progress("surrogate: FDTD responses done")
        t_fdtd = time.perf_counter() - t
        ps, pr = surrogate.pairs(src[:64], rcv[:64])
        hs, hr = surrogate.pairs(src[64:], rcv[64:])
        tr_ir, ho_ir = ir[:len(ps)], ir[len(ps):]
        loss = sur.fit(sur.encode(ps, pr), tr_ir, train_steps, 2e-3, 1e-5, seed=3)
        st = sur.net.stats()
        e_tr = sur.validate(ps, pr, tr_ir)
        fwd = sur.net.stats()
        # inference with queries and responses resident in HBM: 65 536 rows
        # (2 chunks of 32 768) through afx_mlp_forward_device, best of three
        from anc_field_py import _afx
        rows_dev = 65536
        xr = np.random.default_rng(0).uniform(0, 1, (rows_dev, 6)).astype(np.float32)
        with _afx.DeviceBuffer(xr.nbytes) as xd, _afx.DeviceBuffer(rows_dev * sur.taps * 4) as yd:
            xd.upload(xr)
            inf_ms = 1e30
            for _ in range(3):
                sur.net.forward_device(xd, rows_dev, yd)
                inf_ms = min(inf_ms, sur.net.stats().last_forward_gpu_ms)
        cfg = sur.net.cfg
        dims = [2 * cfg.in_dim * cfg.n_freq] + [cfg.width] * (cfg.depth - 1) + [cfg.out_dim]
        flop_row = 2.0 * sum(a * b for a, b in zip(dims[:-1], dims[1:]))
        inf_tflops = flop_row * rows_dev / (inf_ms * 1e-3) / 1e12
        e_ho = surrogate.rel_l2(sur.predict(hs, hr), ho_ir)
        e_nn = surrogate.rel_l2(surrogate.nearest_pair_baseline(ps, pr, tr_ir, hs, hr), ho_ir)
        runs = sur.fdtd_runs
        t = time.perf_counter()
        ans, fb, _ = sur.query(np.concatenate([ps[:512], hs]), np.concatenate([pr[:512], hr]))
        t_query = time.perf_counter() - t
        truth = np.concatenate([tr_ir[:512], ho_ir])
        e_ans = surrogate.rel_l2(ans, truth)
        train_s = st.last_train_gpu_ms * 1e-3
        return {
            "train_pairs": int(len(ps)), "heldout_pairs": int(len(hs)), "taps": sur.taps, "bound": sur.bound,
            "fdtd_impulse_responses_per_s": round(len(ir) / t_fdtd, 1), "fdtd_s": round(t_fdtd, 2),
            "net": {"layers": int(sur.net.cfg.depth), "width": int(sur.net.cfg.width), "fourier_freqs":
                    int(sur.net.cfg.n_freq), "params": int(st.params), "batch": int(sur.net.cfg.batch)},
            "train_steps": int(train_steps), "train_gpu_s": round(train_s, 3),
            "train_tflops": round(st.flop_per_step * train_steps / train_s / 1e12, 2),
            "mfma_bf16_peak_tflops": 2500.0,
            "loss_first_100": float(np.mean(loss[:100])), "loss_last_100": float(np.mean(loss[-100:])),
            "train_rel_l2_median": float(np.median(e_tr)), "train_within_bound": float(np.mean(e_tr <= sur.bound)),
            "heldout_rel_l2_median": float(np.median(e_ho)), "heldout_within_bound": float(np.mean(e_ho <= sur.bound)),
            "nearest_training_pair_heldout_rel_l2_median": float(np.median(e_nn)),
            "net_forward_impulse_responses_per_s": round(fwd.last_forward_rows / (fwd.last_forward_ms * 1e-3), 1),
            "inference_device": {"rows": rows_dev, "gpu_ms": round(inf_ms, 3),
                                 "impulse_responses_per_s": round(rows_dev / (inf_ms * 1e-3), 1),
                                 "tflops": round(inf_tflops, 1), "frac_of_bf16_peak": round(inf_tflops / 2500.0, 4)},
            "query": {"pairs": int(len(ans)), "answered_by_net": int((~fb).sum()), "fdtd_fallbacks": int(fb.sum()),
                      "fdtd_source_runs": int(sur.fdtd_runs - runs), "wall_s": round(t_query, 3),
                      "max_rel_l2_vs_fdtd": float(np.max(e_ans))},
        }
    finally:
        sur.close()


def zone_leg(train_steps, n_val=32, n_test=32, members=2):
    """Learned surrogate over a loudspeaker zone of the furnished room: a
    9^3-cell box of candidate source cells, 16 fixed receivers.  FDTD from
    the box's lattice of every 2nd cell (125 sources) trains an ensemble of
    two 6 x 512 networks (fmax 16 cycles per room length); n_val held-out box
    sources calibrate the ensemble-spread threshold (0.8 x the bound); n_test
    others are queried.  Reported against trilinear interpolation of the same
    lattice responses, the answer the data gives without a network."""
    import numpy as np
    from anc_field_py import rooms, surrogate
    w = rooms.furnished(seed=2, steps=4096)
    rng = np.random.Generator(np.random.PCG64(7))
    size, step = 9, 2
    lo = surrogate.zone_box(w, rng, size)
    mics = surrogate.free_cells(w, rng, 16)
    zone = surrogate.SourceZone(lo, size, step, mics)
    sur = surrogate.Surrogate(w, seed=1, members=members, freq_min=0.25, freq_max=16.0)
    try:
        t = time.perf_counter()
        loss = surrogate.fit_zone(sur, zone, train_steps, 1e-3, 1e-5, seed=3)
        progress("surrogate zone: lattice FDTD and training done")
        t_fit = time.perf_counter() - t
        cells = lo + np.stack(np.meshgrid(*[np.arange(size)] * 3, indexing="ij"), -1).reshape(-1, 3)[:, ::-1]
        off = cells[np.any((cells - lo) % step, axis=1)]
        pick = off[rng.permutation(len(off))]
        val, test = pick[:n_val], pick[n_val:n_val + n_test]
        e_val, t_val, sp_val = surrogate.calibrate_zone(sur, zone, val)
        progress("surrogate zone: calibration done")
        ts, tr_ = surrogate.pairs(test, [mics] * len(test))
        truth = sur.simulate(test, [mics] * len(test))
        net = sur.predict(ts, tr_)
        tri = zone.interpolate(ts, np.tile(np.arange(len(mics)), len(test)))
        runs = sur.fdtd_runs
        t = time.perf_counter()
        ans, fb, _ = sur.query(ts, tr_)
        t_query = time.perf_counter() - t
        e_net, e_tri, e_ans = (surrogate.rel_l2(a, truth) for a in (net, tri, ans))
        acc = ~fb
        src_acc = int(np.sum(~fb.reshape(len(test), -1).any(axis=1)))
        return {
            "zone_cells": size ** 3, "lattice_sources": int(len(zone.lattice())), "receivers": int(len(mics)),
            "members": len(sur.nets), "train_steps_per_member": int(train_steps), "fit_wall_s": round(t_fit, 1),
            "loss_last_100": float(np.mean(loss[-100:])),
            "validation": {"sources": int(len(val)), "ensemble_rel_l2_median": float(np.median(e_val)),
                           "trilinear_rel_l2_median": float(np.median(t_val)),
                           "spread_threshold": zone.delta, "margin": 0.8},
            "test": {"sources": int(len(test)), "pairs": int(len(ts)),
                     "ensemble_rel_l2_median": float(np.median(e_net)),
                     "ensemble_within_bound": float(np.mean(e_net <= sur.bound)),
                     "trilinear_rel_l2_median": float(np.median(e_tri)),
                     "trilinear_within_bound": float(np.mean(e_tri <= sur.bound)),
                     "sources_answered_by_net": src_acc, "fdtd_source_runs": int(sur.fdtd_runs - runs),
                     "pairs_answered_by_net": int(acc.sum()),
                     "max_rel_l2_of_net_answers": float(e_ans[acc].max()) if acc.any() else None,
                     "max_rel_l2_of_all_answers": float(e_ans.max()), "query_wall_s": round(t_query, 3)},
        }
    finally:
        sur.close()


def large_leg(steps, precisions=("fp64", "fp32", "fp16", "bf16", "fp32_two_steps_per_pass",
                                 "fp64_two_steps_per_pass")):
    """Large room, reduced-precision storage: 640 x 480 x 320 cells (98 M),
    dx 0.01 m, 6 IIR materials, 16-cell PML on x+ and y+, band-limited
    noise, 32 receivers.  The same run in FP64, FP32, FP16 and BF16 storage;
    receiver-trace error of each against FP64.  FP32 and FP64 once more with
    two steps per pass (time-step blocking), whose traces must equal the
    single-step ones bit for bit.  A name ending in `one_step_per_pass` runs
    the default tile with one step per pass and no autotuning."""
    import numpy as np
    from anc_field_py import _afx, rooms
    w = rooms.large(steps=steps)
    res = {"cells": w.cells, "steps": steps, "receivers": int(len(w.mic_cells))}
    traces = {}
    for prec in precisions:
        two = prec.endswith("two_steps_per_pass")
        one = prec.endswith("one_step_per_pass")
        base = prec.split("_")[0]
        s = rooms.make_sim(w, _afx.Sim, precision=base)
        try:
            tuned = None
            if two or one:
                s.set_time_block(2 if two else 1)
            elif base in ("fp16", "bf16", "fp32"):
                t0 = time.perf_counter()
                tuned = list(s.autotune(8)) + [s.time_block()[0]]
                split = s.tile_split()
                tune_s = time.perf_counter() - t0
            s.run(1)
            s.synchronize()
            t0 = time.perf_counter()
            s.run(steps - 1)
            s.synchronize()
            dt = time.perf_counter() - t0
            traces[prec] = s.receivers().astype(np.float64)
            pairs = s.time_block()[1]
        finally:
            s.close()
        res[prec] = {"gcell_per_s": round(w.cells * (steps - 1) / dt / 1e9, 1), "wall_s": round(dt, 2)}
        if tuned is not None:
            res[prec]["autotuned_tile_and_steps_per_pass"] = tuned
            res[prec]["tile_split"] = split
            res[prec]["tune_s"] = round(tune_s, 1)
        progress(f"large: {prec} done")
        if two:
            res[prec]["pair_tiles"] = int(pairs)
        if two or one:
            if base in traces:
                res[prec][f"bitexact_vs_{base}"] = bool(np.array_equal(traces[prec], traces[base]))
    if "fp64" in traces:
        ref = traces["fp64"]
        for prec in ("fp32", "fp16", "bf16"):
            if prec in traces:
                res[prec]["rel_l2_vs_fp64"] = float(np.linalg.norm(traces[prec] - ref) / np.linalg.norm(ref))
    return res


def traffic_for(name):
    """HBM bytes per launch of the dominant kernel from the committed
    rocprofv3 counter summary, if one exists for this workload."""
    path = os.path.join(ROOT, "profiles", f"traffic_{name}.json")
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f).get("bytes_per_launch")
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # 20 000 timed steps (about 0.4 s) after 2000 warm-up steps: the GPU's
    # clocks settle over the first few thousand steps (3000 timed steps after
    # 300 read 3-4 % lower)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--warmup", type=int, default=2000)
    ap.add_argument("--precision", default="fp32", choices=sorted(BYTES_PER_CELL))
    ap.add_argument("--compat", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="keep the default tile shape (no autotuning)")
    ap.add_argument("--profile-steps", type=int, default=300)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--batched-steps", type=int, default=4096)
    ap.add_argument("--ie224", type=float, default=0.0, metavar="SECONDS",
                    help="only time the ie224 field simulation for SECONDS of audio (32 kHz steps)")
    ap.add_argument("--mlp-train", type=int, default=0, metavar="STEPS",
                    help="only time STEPS surrogate training steps at batch 512 on synthetic data")
    ap.add_argument("--mlp-inference", type=int, default=0, metavar="ROWS",
                    help="only time the surrogate network's forward pass on ROWS device-resident queries")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run a plain device copy of BYTES (HBM counter calibration) and print its rate")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--no-large", action="store_true")
    ap.add_argument("--large-steps", type=int, default=20000)
    ap.add_argument("--large-precisions", default="fp64,fp32,fp16,bf16,fp32_two_steps_per_pass,fp64_two_steps_per_pass",
                    help="comma-separated runs of the large-room leg")
    ap.add_argument("--surrogate-steps", type=int, default=15000)
    ap.add_argument("--surrogate-eager", action="store_true", help="train without captured graphs")
    ap.add_argument("--no-zone", action="store_true")
    ap.add_argument("--no-configs", action="store_true", help="skip the full single-room runs")
    ap.add_argument("--zone-steps", type=int, default=10000)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              f"(the multi-GPU slab decomposition is a design, not an implementation)", file=sys.stderr)
        sys.exit(2)

    # torch is not imported: it bundles its own HIP runtime, which must not
    # share a process with the one libafx.so links.  The timed region is
    # bracketed by device-wide synchronisation through the library instead.
    from anc_field_py import _afx, rooms

    if args.copy_probe:
        print(json.dumps({"copy_bytes": args.copy_probe, "reps": 3,
                          "gbps": round(_afx.copy_probe(args.copy_probe, reps=3), 1)}), flush=True)
        return
    if args.ie224:
        # The reference's own timed run (examples/ie224-noise-results/report.md:3,9):
        # the ie224 room (416x160x576 cells from its Kerkythea model, here the
        # test fixture copy), white noise at the loudspeaker, 7 receivers, the
        # reference's constants and compat mode, FP32; it reports about 10
        # minutes for 10 s of audio on a GTX 1070.
        import numpy as np
        here = os.path.dirname(os.path.abspath(__file__))
        model = os.path.join(here, "tests", "golden", "ie224", "model.xml")
        steps = int(np.float32(args.ie224) * np.float32(32000))  # main.cpp:117-121
        sig = np.random.default_rng(0).standard_normal(steps).astype(np.float32)
        mics = [[4, 1.6, 5.3], [4, 1.6, 4.5], [4, 1.6, 6.5], [3, 1.6, 5.5], [5, 1.6, 5.5], [4, 0.6, 5.5],
                [4, 2.6, 5.5]]
        s = _afx.Sim()
        s.load_kerkythea(model)
        s.add_source([4, 1.6, 1.0], sig)
        for m in mics:
            s.add_receiver(m)
        s.run(0)
        st = s.stats()
        cells = st.nx * st.ny * st.nz
        tuned = None
        if not args.no_tune:  # tile shape and steps per pass; the state is kept
            tuned = list(s.autotune(30)) + [s.time_block()[0]]
        s.synchronize()
        t0 = time.perf_counter()
        done = 0
        while done < steps:  # progress every 32 000 steps
            k = min(32000, steps - done)
            s.run(k)
            s.synchronize()
            done += k
            progress(f"ie224: {done}/{steps} steps")
        wall = time.perf_counter() - t0
        tr = s.receivers()
        s.close()
        print(json.dumps({"ie224_audio_s": args.ie224, "steps": steps, "cells": cells, "wall_s": round(wall, 2),
                          "gcell_per_s": round(cells * steps / wall / 1e9, 1), "finite": bool(np.all(np.isfinite(tr))),
                          "tuned_tile_and_steps_per_pass": tuned,
                          "reference": "about 10 minutes for 10 s of audio on a GTX 1070"}), flush=True)
        return
    if args.mlp_train:
        # the 6x512 network at batch 512 on 4096 synthetic pairs; with
        # --surrogate-eager the steps are launched one by one instead of
        # replayed from captured graphs
        import numpy as np
        rng = np.random.default_rng(0)
        n = 4096
        with _afx.Mlp(seed=3, batch=512, flags=1 if args.surrogate_eager else 0) as m:
            xs = rng.uniform(0, 1, (n, 6)).astype(np.float32)
            ts = (0.05 * rng.standard_normal((n, m.cfg.out_dim))).astype(np.float32)
            m.set_data(xs, ts, np.full(n, 1.0 / n, np.float32))
            m.train(64, 2e-3, 1e-5)  # warm-up, graph capture
            m.train(args.mlp_train, 2e-3, 1e-5)
            st = m.stats()
        per = st.last_train_gpu_ms / args.mlp_train
        print(json.dumps({"mlp_train_steps": args.mlp_train, "batch": 512, "gpu_ms_per_step": round(per, 4),
                          "tflops": round(st.flop_per_step / (per * 1e-3) / 1e12, 1)}), flush=True)
        return
    if args.mlp_inference:
        # the 6x512 network (random weights) on uniform queries in HBM, best of six
        import numpy as np
        rows = args.mlp_inference
        xq = np.random.default_rng(0).uniform(0, 1, (rows, 6)).astype(np.float32)
        with _afx.Mlp(seed=3) as m, _afx.DeviceBuffer(xq.nbytes) as xd, \
                _afx.DeviceBuffer(rows * m.cfg.out_dim * 4) as yd:
            xd.upload(xq)
            best = float("inf")
            for _ in range(6):
                m.forward_device(xd, rows, yd)
                best = min(best, m.stats().last_forward_gpu_ms)
        print(json.dumps({"mlp_inference_rows": rows, "gpu_ms": round(best, 4),
                          "impulse_responses_per_s": round(rows / (best * 1e-3), 1)}), flush=True)
        return

    total = args.warmup + args.steps + args.profile_steps
    w = rooms.furnished(steps=total)
    sim = rooms.make_sim(w, _afx.Sim, compat=args.compat, precision=args.precision)
    sim.run(args.warmup)
    tile = None if args.no_tune else sim.autotune(30)
    sim.synchronize()
    t0 = time.perf_counter()
    sim.run(args.steps)
    sim.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    gcells = w.cells * args.steps / elapsed / 1e9
    st = sim.stats()
    # Dominant kernel: with the fused step kernel there is one launch per step
    # (plus one 1-thread counter kernel per 30-step graph), so its average
    # launch duration is the device time of the timed region, taken from HIP
    # events recorded on the library's stream, over the number of steps.
    k_ms = st.last_run_gpu_ms / st.last_run_steps
    prof = sim.profile(args.profile_steps)  # eager launches, one event pair per kernel
    bpc = BYTES_PER_CELL[args.precision]
    achieved = bpc * w.cells / (k_ms * 1e-3) / 1e9 if k_ms > 0 else 0.0
    kernel = "k_fused" if st.fused else "k_interior"
    # a plain device copy moving the same bytes (read half, write half): the
    # practical rate for this working-set size
    copy_bytes = (bpc * w.cells // 2) // 16 * 16
    copy_gbps = _afx.copy_probe(copy_bytes, reps=50)
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_for(f"{w.name}_{args.precision}"),
            "kernel": kernel, "kernel_ms": round(k_ms, 5),
            "algorithmic_bytes_per_launch": bpc * w.cells,
            "plain_copy_same_bytes": {"gbps": round(copy_gbps, 1),
                                      "us": round(2 * copy_bytes / (copy_gbps * 1e9) * 1e6, 2)}}
    out = {
        "metric": "Gcell-updates/s",
        "value": round(gcells, 2),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp32": "f32", "fp64": "f64", "fp16": "f16-storage/f32", "bf16": "bf16-storage/f32"}[args.precision],
        "data": "synthetic (seeded room generator)",
        "config": {"workload": f"furnished room {w.dims[0]}x{w.dims[1]}x{w.dims[2]} cells, dx {w.dx} m, "
                               f"4 IIR materials, 12 boxes, 12-cell PML x+, 16 receivers",
                   "mode": "compat" if args.compat else "physical",
                   "boundary_cells": int(st.boundary), "fused_step_kernel": bool(st.fused),
                   "tile": [st.tile_x, st.tile_y, st.tile_z], "steps_per_pass": sim.time_block()[0],
                   "tile_split": sim.tile_split(),
                   "autotuned": tile is not None,
                   "profile": {"step_ms": round(prof.step_ms, 5),
                               "kernel_ms": [round(x, 5) for x in prof.kernel_ms],
                               "launches": list(prof.launches)}},
        "roofline": roof,
    }
    sim.close()
    progress("main leg done")
    if not args.no_configs:
        out["configs"] = configs_leg()
    if not args.no_batched:
        out["batched"] = batched_leg(args.batched_steps)
    if not args.no_surrogate:
        out["surrogate"] = surrogate_leg(args.surrogate_steps, args.surrogate_eager)
        progress("surrogate done")
    if not args.no_zone:
        out["surrogate_zone"] = zone_leg(args.zone_steps)
        progress("surrogate zone done")
    if not args.no_large:
        out["large"] = large_leg(args.large_steps, tuple(args.large_precisions.split(",")))
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(w)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
