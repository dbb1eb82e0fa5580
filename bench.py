"""Benchmark of the multislice STEM hot path on one MI355X: the HAADF scan workload.

Workload (the headline HAADF configuration): SrTiO3 tiled 8 x 8 x 20 (31.24 x 31.24 x 78.1 Angstrom, 6,400 atoms),
512 x 512 grid, 40 slices of 1.9525 Angstrom, 200 keV probe with a 20 mrad soft aperture (C10 = C30 = 0), a 64 x 64
GridScan over one 3.905 Angstrom cell, annular detector 50-200 mrad, complex64 waves, 1024 probe positions per launch
with the slice loop captured in a hipGraph and replayed.

One step = one full 64 x 64 scan (4,096 positions x 40 slices = 163,840 position-slices) with the prepared
potential resident in HBM and the HAADF image left in HBM. `value` = position-slices per second over the K timed
steps of the headline plan: 1024 positions per graph replay, run by the library as tiles of 128 positions (64 on each
of two streams) whose waves stay in the 256 MiB Infinity Cache between passes.

`roofline` is the HBM roofline of the same scan with the waves streamed through HBM: 1024 positions per kernel launch
(`--streamed-tile`), graph-replayed. Its dominant pass is read three ways: algorithmic bytes (16 N per position and
pass, the full complex64 wave read and written), the compact bytes it stores (the columns inside the bandlimit), and
the L2-to-fabric bytes of the committed rocprofv3 counter summary; `step` checks the scan's algorithmic bytes per
second against the peak. `llc_resident` gives the headline plan's kernels, whose algorithmic rate is an effective
bandwidth above what HBM could deliver, not an HBM fraction; `compare_no_graph` times the headline plan with its
kernels launched one by one. Kernel times come from HIP events attached to the
launches of one more scan of the same tiles run on one stream. The CPU baseline runs the float64 oracle on a bounded
sample of positions on the host's cores (one warm-up, five runs, median and min-max), and the same sample gives the
HIP path's error against float64.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# measured HBM bytes per launch of the dominant kernel: rocprofv3 FETCH_SIZE and WRITE_SIZE passes over this
# workload, scaled by what the same counters read for a device copy of known size (the summary holds the commands)
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "column_pass_haadf_512_traffic.md")


def measured_traffic(kernel: str, batch: int):
    """The `traffic` block of the committed counter summary for `kernel` at this workload and positions per launch,
    or None."""
    try:
        text = open(TRAFFIC_SUMMARY).read()
    except OSError:
        return None
    start = text.find("```json")
    if start < 0:
        return None
    end = text.find("```", start + 7)
    data = json.loads(text[start + 7:end])
    for entry in data.get("entries", []):
        if entry["kernel"] == kernel and int(entry["positions_per_launch"]) == batch:
            return float(entry["hbm_bytes_per_launch"])
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=None,
                    help="positions per launch (one graph replay); 0: the library's choice (default: 1024 for the "
                         "HAADF scan, the library's choice for frozen phonons)")
    ap.add_argument("--tile", type=int, default=0,
                    help="positions per kernel launch within a batch; 0: the library's choice (tiles that stay in "
                         "the Infinity Cache)")
    ap.add_argument("--streamed-tile", type=int, default=1024,
                    help="also time the plan whose kernel launches carry this many positions, streaming the waves "
                         "through HBM: the roofline entry (0: skip)")
    ap.add_argument("--only", default="all", choices=["all", "headline", "streamed"],
                    help="run only one of the two plans (for a profiler)")
    ap.add_argument("--no-graph", action="store_true", help="launch the slice loop's kernels one by one")
    ap.add_argument("--generic", action="store_true", help="generic LDS passes instead of the fast passes")
    ap.add_argument("--col-tile", type=int, default=0, help="columns per 512-point column-pass block (8, 16)")
    ap.add_argument("--streams", type=int, default=0,
                    help="streams each batch is split over (1-4; 0 = the library's choice, 2 for this scan)")
    ap.add_argument("--cpu-sample", type=int, default=32, help="positions of the float64 CPU baseline sample")
    ap.add_argument("--cpu-runs", type=int, default=5, help="timed runs of the CPU baseline (after one warm-up)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: min(8, cores), the reference's default")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", default="haadf",
                    choices=["haadf", "frozen_phonons", "precision_study", "surrogate_training"],
                    help="haadf (the headline scan), frozen_phonons (16 configurations of uneven slice counts, "
                         "pixelated 128 x 128 float32 crop + centre of mass), precision_study (the HAADF geometry "
                         "at grid --grid, 4096 positions in one launch) or surrogate_training (Adam steps of the "
                         "HAADF surrogate, batch 16)")
    ap.add_argument("--vector-units", action="store_true",
                    help="surrogate_training: the convolutions on the vector units instead of the matrix cores")
    ap.add_argument("--grid", type=int, default=384, help="precision_study: grid points per axis (256, 384, 509, "
                    "512, 1024)")
    ap.add_argument("--precision", default="float32", choices=["float32", "split_bf16"],
                    help="precision_study: transforms in float32 or split-bfloat16 on the matrix cores")
    ap.add_argument("--fp-scan", type=int, default=128, help="frozen_phonons: scan points per axis")
    ap.add_argument("--pipeline", type=int, default=-1, help="frozen_phonons: 1 lane pipeline, 0 one potential at a time")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="first run one device-to-device copy of BYTES (the reference kernel that scales the "
                         "FETCH_SIZE / WRITE_SIZE counters of a profile)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        sys.stderr.write("bench.py measures one GPU: this build runs the scan on a single MI355X and has no "
                         "multi-GPU path (positions would shard with no exchange); run with --gpus 1.\n")
        sys.exit(2)

    import numpy as np
    import torch

    import simulatokit_amd as sk
    from simulatokit_amd.workloads import HaadfScan

    if not sk.hip_is_available():
        raise SystemExit("no HIP device: " + sk._native.last_error())
    if args.calibrate_copy:
        src = torch.ones(args.calibrate_copy // 4, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    if args.batch is None:
        args.batch = 1024 if args.workload == "haadf" else 0
    if args.workload == "frozen_phonons":
        print(json.dumps(frozen_phonon_bench(args, sk, torch)), flush=True)
        return
    if args.workload == "precision_study":
        print(json.dumps(precision_study_bench(args, sk, torch)), flush=True)
        return
    if args.workload == "surrogate_training":
        print(json.dumps(surrogate_training_bench(args, torch)), flush=True)
        return

    w = HaadfScan(batch=args.batch)
    st = w.structure
    pot = sk.Potential(st, w.n, w.n, slice_thickness=w.slice_thickness)
    beam = sk.ElectronBeam(w.energy)
    probe = sk.Probe(beam, pot.grid, sk.Aperture(w.semiangle))
    scan = sk.GridScan((0.0, 0.0), (sk.workloads.SRTIO3_A, sk.workloads.SRTIO3_A), w.scan_n, w.scan_n)
    det = sk.AnnularDetector(w.inner, w.outer)
    S = pot.slice_count
    npos = w.scan_n * w.scan_n
    N = w.n * w.n
    graph = not args.no_graph

    def plan(tile, use_graph=graph):
        return sk.ExecutionOptions(batch=args.batch, tile=tile, use_graph=use_graph, fast_path=not args.generic,
                                   col_tile=args.col_tile, streams=args.streams)

    t0 = time.perf_counter()
    prepared = pot.prepare(beam)
    prepare_ms = (time.perf_counter() - t0) * 1e3
    out = torch.zeros(npos, dtype=torch.float64, device="cuda:0")
    positions = scan.positions

    def step(o):
        sk.STEM.scan(probe, [prepared], positions, [det], o, device_outputs=[out.data_ptr()])

    def timed(o):
        for _ in range(args.warmup):
            step(o)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step(o)
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    def kernel_times(o):
        """Mean device time per launch of each pass, from HIP events attached to the launches of one more scan of
        the same tiles run on one stream (so each launch runs alone; graph replay does not change the kernels)."""
        sk.set_timing(True)
        step(dataclasses.replace(o, use_graph=False))
        sk.set_timing(False)
        return sk.last_timing()

    def roofline_of(tm, rate):
        """Three readings of one plan's dominant pass: algorithmic bytes (16 N per position read + written, the
        full complex64 wave), the compact bytes the pass stores (the kept columns only), and the measured L2-to-
        fabric bytes of the committed counter profile at this tile; plus the step-level check."""
        B = int(tm["batch"])
        col_ms, row_ms = tm["col_ms"], tm["row_ms"]
        dom = "column pass" if col_ms >= row_ms else "row pass"
        dom_ms = max(col_ms, row_ms)
        alg = 16.0 * N * B
        stored = tm["col_bytes"]
        measured = measured_traffic(dom, B)
        gbs = lambda b: b / (dom_ms * 1e-3) / 1e9
        step_gbs = 32.0 * N * rate / 1e9  # algorithmic bytes per position-slice times the scan's rate
        r = {"bound": "hbm", "achieved": gbs(alg), "peak": HBM_PEAK_GBS, "unit": "GB/s",
             "frac": gbs(alg) / HBM_PEAK_GBS, "traffic": measured, "kernel": dom, "kernel_ms": dom_ms,
             "row_ms": row_ms, "col_ms": col_ms, "positions_per_kernel_launch": B,
             "algorithmic_bytes_per_launch": alg,
             "stored": {"bytes_per_launch": stored, "GBs": gbs(stored), "frac": gbs(stored) / HBM_PEAK_GBS},
             "measured": None if measured is None else {"bytes_per_launch": measured, "GBs": gbs(measured),
                                                        "frac": gbs(measured) / HBM_PEAK_GBS},
             "step": {"GBs": step_gbs, "frac": step_gbs / HBM_PEAK_GBS,
                      "ok": bool(step_gbs <= HBM_PEAK_GBS and gbs(alg) <= HBM_PEAK_GBS)}}
        return r

    result_plans = {}
    # the plan whose waves stream through HBM: each kernel launch carries the whole batch
    if args.streamed_tile > 0 and args.only in ("all", "streamed"):
        o = plan(args.streamed_tile)
        el = timed(o)
        rate = npos * S * args.steps / el
        result_plans["streamed"] = (o, rate, el / args.steps * 1e3, kernel_times(o))
    if args.only in ("all", "headline"):
        o = plan(args.tile)
        el = timed(o)
        rate = npos * S * args.steps / el
        result_plans["headline"] = (o, rate, el / args.steps * 1e3, kernel_times(o))
        if graph and args.only == "all":
            el = timed(plan(args.tile, use_graph=False))
            result_plans["no_graph"] = (None, npos * S * args.steps / el, el / args.steps * 1e3)
    haadf = out.cpu().numpy()
    copy_gbs = copy_bandwidth(torch)
    if "headline" in result_plans:
        o, value, ms_per_step, tm_head = result_plans["headline"]
    else:
        o, value, ms_per_step, tm_head = result_plans["streamed"]

    result = {
        "metric": "probe-positions*slices/sec (HAADF scan) + fraction of HBM roofline at <=1e-4 rel err",
        "value": value,
        "unit": "probe-positions*slices/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic",
        "config": {"workload": w.name, "grid": [w.n, w.n], "slices": S, "positions": npos,
                   "batch": int(tm_head["group"]), "graph": bool(o.use_graph),
                   "positions_per_kernel_launch": int(tm_head["batch"]),
                   "detector": f"annular {w.inner:g}-{w.outer:g} mrad", "energy_keV": w.energy / 1e3,
                   "semiangle_mrad": w.semiangle, "fast_path": not args.generic,
                   # the library splits this scan's tiles over 2 streams unless --streams says otherwise
                   "streams": 1 if args.generic else (args.streams or 2),
                   "parallelism": "single GPU"},
        "prepare_ms": prepare_ms,
        "scan_pos_slices": npos * S,
    }
    if "streamed" in result_plans:
        so, srate, sms, stm = result_plans["streamed"]
        roof = roofline_of(stm, srate)
        roof.update({"copy_GBs": copy_gbs, "plan": {
            "what": f"{int(stm['group'])} positions per launch and per kernel launch, graph-replayed slice loop: "
                    "the waves stream through HBM between passes",
            "value": srate, "ms_per_step": sms, "graph": bool(so.use_graph)}})
        result["roofline"] = roof
    if "headline" in result_plans:
        llc = roofline_of(tm_head, value)
        result["llc_resident"] = {
            "what": f"the headline plan's kernels: {int(tm_head['batch'])} positions per kernel launch, whose waves "
                    "stay in the 256 MiB Infinity Cache between passes; algorithmic bytes over their time are an "
                    "effective bandwidth, not an HBM fraction",
            "kernel": llc["kernel"], "kernel_ms": llc["kernel_ms"], "row_ms": llc["row_ms"], "col_ms": llc["col_ms"],
            "effective_GBs": llc["achieved"], "effective_x_hbm_peak": llc["frac"],
            "step_effective_GBs": llc["step"]["GBs"], "stored": llc["stored"], "measured": llc["measured"]}
        if "no_graph" in result_plans:
            _, ng_rate, ng_ms = result_plans["no_graph"]
            result["compare_no_graph"] = {
                "what": "the headline plan with the slice loop's kernels launched one by one instead of replayed "
                        "from a hipGraph", "value": ng_rate, "ms_per_step": ng_ms}

    if not args.no_cpu_baseline:
        result["cpu_baseline"], result["rel_err_vs_fp64"] = cpu_baseline(args, w, st, pot, positions, haadf, S)

    print(json.dumps(result), flush=True)


def frozen_phonon_bench(args, sk, torch):
    """4D-STEM with frozen phonons: 16 configurations of a seeded 40 x 40 A random cell (2000 atoms, Z from
    {8, 14, 22, 38, 79}, sigma 0.08 A) cut to thicknesses uniform in 40-120 A, 512 x 512, 1 A slices, so the slice
    counts differ; a fp_scan x fp_scan scan over the cell into a 128 x 128 float32 crop of the pattern and the
    centre of mass, averaged over the configurations. One step = the whole scan through every configuration."""
    import numpy as np
    from simulatokit_amd.workloads import uneven_frozen_phonon_stack

    structures, T = uneven_frozen_phonon_stack()
    pots = [sk.Potential(st, 512, 512, slice_thickness=1.0) for st in structures]
    probe = sk.Probe(sk.ElectronBeam(200e3), pots[0].grid, sk.Aperture(20.0))
    t0 = time.perf_counter()
    prepared = [p.prepare(probe.beam) for p in pots]
    prepare_ms = (time.perf_counter() - t0) * 1e3
    n = args.fp_scan
    scan = sk.GridScan((0.0, 0.0), (40.0, 40.0), n, n)
    dets = [sk.PixelatedDetector(crop=(128, 128), dtype="float32"), sk.CenterOfMassDetector("cutoff", "mrad")]
    npos = n * n
    out_pix = torch.empty((npos, 128, 128), dtype=torch.float32, device="cuda:0")
    out_com = torch.empty((npos, 2), dtype=torch.float64, device="cuda:0")
    opt = sk.ExecutionOptions(batch=args.batch, pipeline=args.pipeline, streams=args.streams)

    def step():
        sk.STEM.scan(probe, prepared, scan, dets, opt, device_outputs=[out_pix.data_ptr(), out_com.data_ptr()])

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    slices = [p.slice_count for p in prepared]
    work = npos * sum(slices)
    return {"metric": "probe-positions*slices/sec (frozen-phonon 4D-STEM, uneven slice counts)",
            "value": work * args.steps / elapsed, "unit": "probe-positions*slices/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": elapsed / args.steps * 1e3,
            "higher_is_better": True, "dtype": "c64", "data": "synthetic",
            "config": {"workload": f"frozen_phonons_16cfg_512sq_{n}x{n}", "configurations": len(prepared),
                       "slices": slices, "positions": npos, "batch": args.batch or "library", "pipeline": args.pipeline,
                       "streams": args.streams or "library",
                       "detectors": "pixelated 128x128 float32 + centre of mass"},
            "prepare_ms": prepare_ms, "scan_pos_slices": work}


def precision_study_bench(args, sk, torch):
    """The HAADF geometry at one grid length: SrTiO3 8 x 8 cells (40 slices of 2 A; 20 at 509, whose Bluestein
    transforms cost 5x), 200 keV, 20 mrad, a 64 x 64 scan over one cell, all 4096 positions in ONE launch, annular
    50 mrad to 200 mrad (to 100 and 150 mrad at 256 and 384 points, where 200 mrad exceeds the largest simulated
    angle). One step = one scan of the 4096 positions, potential prepared beforehand; the line also carries the
    device time of one row and one column launch (averaged over the slices of a timed scan)."""
    from simulatokit_amd.workloads import SRTIO3_A, srtio3

    n = args.grid
    st = srtio3(8, 8, 10 if n == 509 else 20)
    pot = sk.Potential(st, n, n, slice_thickness=2.0)
    probe = sk.Probe(sk.ElectronBeam(200e3), pot.grid, sk.Aperture(20.0))
    prepared = pot.prepare(probe.beam)
    scan = sk.GridScan((0.0, 0.0), (SRTIO3_A, SRTIO3_A), 64, 64)
    outer = {256: 100.0, 384: 150.0}.get(n, 200.0)
    det = sk.AnnularDetector(50.0, outer)
    npos = 64 * 64
    out = torch.zeros(npos, dtype=torch.float64, device="cuda:0")
    opt = sk.ExecutionOptions(batch=npos, tile=npos, col_tile=args.col_tile, precision=args.precision)

    def step():
        sk.STEM.scan(probe, [prepared], scan, [det], opt, device_outputs=[out.data_ptr()])

    for _ in range(max(1, args.warmup)):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    sk.set_timing(True)
    step()
    sk.set_timing(False)
    tm = sk.last_timing()
    S = prepared.slice_count
    prepared.release()
    return {"metric": "probe-positions*slices/sec (HAADF scan, precision-study geometry)",
            "value": npos * S * args.steps / elapsed, "unit": "probe-positions*slices/s", "n_gpus": 1,
            "steps": args.steps, "warmup": max(1, args.warmup), "ms_per_step": elapsed / args.steps * 1e3,
            "higher_is_better": True, "dtype": "c64" if args.precision == "float32" else "split-bf16/f32",
            "data": "synthetic",
            "config": {"workload": f"precision_study_srtio3_{n}sq_64x64", "grid": n, "slices": S, "positions": npos,
                       "batch": npos, "precision": args.precision, "col_tile": args.col_tile or "library",
                       "detector": f"annular 50-{outer:.0f} mrad"},
            "row_ms": tm["row_ms"], "col_ms": tm["col_ms"]}


def surrogate_training_bench(args, torch):
    """Training of the HAADF surrogate network as specified (six 5 x 5 layers, 2 -> 32 -> ... -> 32 -> 1 channels,
    bfloat16 weights, 128 x 128 input, 96 x 96 output): one step = one Adam step on a batch of 16 (forward, relative
    squared error, backward, update, with the 8-symmetry augmentation), through HaadfSurrogate.train. The inputs and
    labels are seeded random arrays of the network's shapes (the step's arithmetic does not depend on their values);
    64 cells, so an epoch is 4 steps; `--steps` epochs are timed after `--warmup`."""
    import numpy as np

    from simulatokit_amd import surrogate as SG

    rng = np.random.default_rng(7)
    cells = 64
    data = SG.Dataset(list(range(cells)), rng.normal(size=(cells, 2, SG.INPUT_N, SG.INPUT_N)).astype(np.float32),
                      np.abs(rng.normal(size=(cells, SG.SCAN_N, SG.SCAN_N))).astype(np.float32) * 0.01)
    model = SG.HaadfSurrogate(seed=0, matrix_cores=not args.vector_units)
    model.train(data, epochs=max(1, args.warmup))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    model.train(data, epochs=args.steps)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    steps = args.steps * (cells // 16)
    return {"metric": "surrogate training steps/sec (batch 16)", "value": steps / elapsed, "unit": "steps/s",
            "n_gpus": 1, "steps": steps, "warmup": max(1, args.warmup) * (cells // 16),
            "ms_per_step": elapsed / steps * 1e3, "higher_is_better": True, "dtype": "bf16 weights, f32 activations",
            "data": "synthetic",
            "config": {"workload": "haadf_surrogate_training_6x32_128sq", "channels": SG.CHANNELS, "batch": 16,
                       "convolutions": "vector units" if args.vector_units else "matrix cores"}}


def copy_bandwidth(torch, nbytes: int = 1 << 30, reps: int = 5) -> float:
    """The device-to-device copy rate of this GPU (read + write bytes per second, GB/s): the practical ceiling of a
    pass that reads and writes every byte once, measured beside the published 8 TB/s peak."""
    src = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda:0").fill_(1.0)
    dst = torch.empty_like(src)
    dst.copy_(src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del src, dst
    return 2.0 * nbytes / (ms * 1e-3) / 1e9


def host_cpu():
    """The host's CPU model, the logical CPUs it reports, and the share this process may use (the affinity mask,
    capped by OMP_NUM_THREADS when the environment sets it, as a shared GPU box does)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    logical = os.cpu_count() or 1
    try:
        share = len(os.sched_getaffinity(0))
    except AttributeError:
        share = logical
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        share = min(share, int(omp))
    return model, logical, share


def cpu_baseline(args, w, st, pot, positions, haadf, S):
    """The float64 oracle (the reference's CPU path restated) on the first `cpu_sample` positions of the scan, timed
    as the reference times its benchmarks (one warm-up, then five runs; median and min-max;
    Tests/SimulatoKitParityTests/Detector/BenchmarkTableTests.swift:25-33) with the reference's threading
    (min(8, cores) workers, one position each; MetalMultislice.swift:28-47) and again on every core of the host's
    share. Only the position loop is timed (the transmission functions are prepared first), and the rate is
    position-slices per second of the sample, which scales linearly with positions as the reference's own
    subset benchmarks assume (BenchmarkTests.swift:46-47,68). The same sample gives the HIP path's error."""
    import numpy as np

    from oracle import oracle as O

    model, logical, share = host_cpu()
    k = min(args.cpu_sample, len(positions))
    V, _ = O.potential(w.n, w.n, st.numbers, st.positions, st.box, pot.slice_thicknesses)

    def runs(threads):
        secs, ref = [], None
        for r in range(1 + args.cpu_runs):
            ref = O.stem_scan([V], [pot.slice_thicknesses], (st.box[0], st.box[1]), w.energy, w.semiangle,
                              positions[:k], [O.det_annular(w.inner, w.outer)], threads=threads)[0]
            if r > 0:  # the first run warms up
                secs.append(O.last_scan_seconds())
        rates = sorted(k * S / t for t in secs)
        return {"threads": threads, "median": float(np.median(rates)), "min": rates[0], "max": rates[-1],
                "runs": len(rates), "seconds_per_run": float(np.median(secs))}, ref

    ref_threads = args.cpu_threads or min(8, share)
    base, ref = runs(ref_threads)
    every = runs(share)[0] if share != ref_threads else base
    got = haadf[:k]
    rel_l2 = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
    rel_max = float(np.abs(got - ref).max() / np.abs(ref).max())
    out = {"value": base["median"], "unit": "probe-positions*slices/s", "cores": ref_threads, "kind": "port",
           "sample": f"{k} of {len(positions)} positions x {S} slices of the same scan (float64 oracle, transmission "
                     f"functions prepared before the timed loop); one warm-up then {args.cpu_runs} runs, median, on "
                     f"{ref_threads} threads (the reference's min(8, cores)); {base['seconds_per_run']:.1f} s per run",
           "min": base["min"], "max": base["max"], "runs": base["runs"],
           "all_cores": every, "cpu_model": model, "host_logical_cpus": logical, "host_cpu_share": share}
    return out, {"haadf_rel_l2": rel_l2, "haadf_rel_max": rel_max, "positions": k}


if __name__ == "__main__":
    main()
