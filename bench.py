#!/usr/bin/env python3
"""Benchmark of BZTrellisQdc.ir_interpolate_at on one MI355X.

Workload (default): the learned-proposal configuration -- hexagonal P6_3/mmc,
a=b=3.2 A, c=5.2 A, 2 atoms, 6 modes with complex eigenvectors, trellis node
volume 1e-5 of the irreducible volume, 1e8 Q uniform in [-4,4]^3 rlu,
inputs and outputs resident in HBM.  One step = one ir_interpolate_at pass over
all Q.  The headline path is the binned FP32 pipeline with the learned
tetrahedron proposal (FP64 fold-in, indices and weights; FP32/complex64 sums);
the binned pipeline without the proposal, the fused tile kernel, the binned
FP64 path and the plain FP64 kernel are timed beside it, and the oracle on the
host cores.  The line also carries configuration 2 (64 uneven grids, 1e7 Q,
one autotuned graph-replayed batch) under "batch_config2"; `--config 2`
makes that batch the line itself, with its per-grid baselines.

Prints one JSON line.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def table_bytes(t):
    keys = ["prim_plane_a", "prim_plane_b", "prim_plane_c", "prim_normals", "prim_taus", "prim_tau_lens",
            "outer_plane_a", "outer_plane_b", "outer_plane_c", "wedge_normals", "rotations", "inverse_index",
            "node_type", "node_index", "cube_vertices", "poly_tet_offset", "tet_vertices", "tet_circumsphere",
            "tet_volume", "vertices"]
    return int(sum(np.asarray(t[k]).nbytes for k in keys) + sum(np.asarray(k).nbytes for k in t["knots"]))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", type=int, default=4)
    ap.add_argument("--n-q", type=float, default=None)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline (profiling runs)")
    ap.add_argument("--no-learned", action="store_true", help="config 4: skip training the tetrahedron proposal")
    ap.add_argument("--no-batch", action="store_true", help="config 4: skip the configuration-2 batch leg")
    ap.add_argument("--batch-schedule", type=str, default=None,
                    help="config 2: tile,blocks_per_cu,tiles_per_grab,order,binned_min instead of autotuning "
                         "(counter runs serialise kernels, which changes what the autotuner picks)")
    ap.add_argument("--no-sort", action="store_true", help="config 4: skip the mode-sort leg")
    ap.add_argument("--sort-only", action="store_true",
                    help="only the mode-sort leg on the config-4 grid (profiling runs)")
    ap.add_argument("--paths", default="all",
                    help="config 4: comma-separated paths to time beside the headline (binned32, tile32, binned64; "
                         "'none' for the headline alone, as the byte-counter runs use it)")
    ap.add_argument("--train-pairs", type=float, default=2e7, help="config 4: (q, tetrahedron) training pairs")
    ap.add_argument("--train-epochs", type=int, default=3)
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only copy this many MiB device-to-device once (calibrates the HBM byte counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the path would shard by Q with no exchange between GPUs)", file=sys.stderr)
        sys.exit(2)

    if args.copy_probe_mib:
        import torch
        x = torch.ones(args.copy_probe_mib << 20, dtype=torch.uint8, device="cuda")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        y.copy_(x)
        torch.cuda.synchronize()
        log(f"copied {args.copy_probe_mib} MiB device-to-device")
        return

    from brille_amd import capi
    from brille_amd import synthetic as S

    lib = capi.lib()
    if lib.brille_device_count() < 1:
        raise SystemExit("no HIP device")
    cfg = args.config
    if cfg == 2:
        return bench_batch(args, capi, S)
    c = S.CONFIGS[cfg]
    n = int(args.n_q if args.n_q is not None else c["n_q"])
    t0 = time.time()
    g = S.grid(cfg)
    T = g.tables()
    log(f"grid cfg{cfg}: {T['vertices'].shape[0]} vertices, {T['node_type'].size} nodes, "
        f"{T['tet_volume'].size} tetrahedra, built in {time.time() - t0:.1f}s")
    if args.sort_only:
        print(json.dumps(sort_leg(args, capi, S, T)), flush=True)
        return
    dev = capi.Grid(T)
    B = int(T["branches"])
    sv = int(sum(T["values_elements"]))
    sc = int(sum(T["vectors_elements"]))
    Q = S.q_points(cfg, n)
    dQ = capi.DeviceArray.from_host(Q)
    vs, cs = dev.out_shapes(n)
    dv = capi.DeviceArray(vs, np.float64)
    dc = capi.DeviceArray(cs, np.complex128)
    stream = C.c_void_p()
    capi.check(lib.brille_stream_create(C.byref(stream)))

    def run(opts, iters):
        times = capi.KernelTimes()
        capi.check(lib.brille_time_ir_interpolate_at(dev.h, dQ.ptr, n, C.byref(opts), dv.ptr, dc.ptr, iters,
                                                     stream, C.byref(times)))
        return times

    def timed(opts, label):
        # one checked call first: any failing point aborts the benchmark
        o = capi.RunOpts(*[getattr(opts, f) for f, _ in capi.RunOpts._fields_])
        o.check = 1
        capi.check(lib.brille_ir_interpolate_at(dev.h, dQ.ptr, n, C.byref(o), dv.ptr, dc.ptr, None, stream))
        for _ in range(args.warmup):
            run(opts, 1)
        capi.check(lib.brille_device_sync())
        tot, parts = [], []
        w0 = time.perf_counter()
        for s in range(args.steps):
            t = run(opts, 1)
            tot.append(t.total_ms)
            parts.append([t.ms[k] for k in range(t.n_kernels)])
        capi.check(lib.brille_device_sync())
        wall = (time.perf_counter() - w0) * 1000.0 / args.steps
        names = [t.names[k].value.decode() for k in range(t.n_kernels)]
        stages = {nm: round(float(v), 4) for nm, v in zip(names, np.mean(parts, axis=0))}
        log(f"{label}: {wall:.3f} ms/step wall, kernels {np.mean(tot):.3f} ms {stages}, {n / wall * 1e3:.4g} Q/s")
        return dict(value=n / wall * 1e3, ms_per_step=wall, kernel_ms=float(np.mean(tot)), stages=stages)

    results = {}
    learned = None
    if cfg == 4 and not args.no_learned:
        topts = capi.TrainOpts(int(args.train_pairs), 2026, args.train_epochs, 8192, 2e-3)
        rep = capi.TrainReport()
        capi.check(lib.brille_proposal_train(dev.h, C.byref(topts), C.byref(rep)))
        learned = dict(pairs=int(rep.n_pairs), epochs=args.train_epochs, final_loss=float(rep.final_loss),
                       held_out_hit_rate=float(rep.train_hit_rate), tetrahedra=int(rep.n_classes),
                       train_ms=float(rep.ms))
        log(f"proposal trained: {learned}")
        results["learned+binned fp32"] = timed(capi.RunOpts(0, capi.FP32, 1, capi.MEM_DEVICE, capi.KERNEL_BINNED, 0),
                                               "learned+binned fp32")
        tried, hits = C.c_uint64(), C.c_uint64()
        capi.check(lib.brille_proposal_stats(dev.h, C.byref(tried), C.byref(hits)))
        learned["eval_proposals"] = int(tried.value)
        learned["eval_verified"] = int(hits.value)
        learned["eval_hit_rate"] = hits.value / max(1, tried.value)
    want = set(("binned32", "tile32", "binned64") if args.paths == "all" else
               [x.strip() for x in args.paths.split(",") if x.strip() and x.strip() != "none"])
    if "binned32" in want or "learned+binned fp32" not in results:
        results["binned fp32"] = timed(capi.RunOpts(0, capi.FP32, 0, capi.MEM_DEVICE, capi.KERNEL_BINNED, 0),
                                       "binned fp32")
    if "tile32" in want:
        results["tile fp32"] = timed(capi.RunOpts(0, capi.FP32, 0, capi.MEM_DEVICE, capi.KERNEL_TILE, 0), "tile fp32")
    if "binned64" in want:
        results["binned fp64"] = timed(capi.RunOpts(0, capi.FP64, 0, capi.MEM_DEVICE, capi.KERNEL_BINNED, 0),
                                       "binned fp64")
    plain_res = None
    if not args.no_plain and args.paths == "all":
        plain_res = timed(capi.RunOpts(0, capi.FP64, 0, capi.MEM_DEVICE, capi.KERNEL_PLAIN, 0), "plain fp64")
        plain_res.update(dtype="f64", bytes_per_q=24 + B * sv * 8 + B * sc * 16)
    capi.check(lib.brille_stream_destroy(stream))
    head_name = "learned+binned fp32" if "learned+binned fp32" in results else "binned fp32"
    # FP32 error of the headline path against the FP64 path (itself at 1e-12 of
    # the oracle) on the first 1e6 points of the workload
    m = min(n, 1_000_000)
    v32, c32 = dev.ir_interpolate_at(Q[:m], precision=capi.FP32, kernel=capi.KERNEL_BINNED,
                                     use_proposal=learned is not None)
    v64, c64 = dev.ir_interpolate_at(Q[:m], precision=capi.FP64, kernel=capi.KERNEL_TILE)
    vscale = np.max(np.abs(v64), axis=0, keepdims=True)
    cnorm = np.sqrt(np.sum(np.abs(c64.reshape(m, B, -1)) ** 2, axis=2, keepdims=True))
    fp32_error = dict(sample=m, eigenvalue=float(np.max(np.abs(v32 - v64) / vscale)),
                      eigenvector=float(np.max(np.abs(c32 - c64).reshape(m, B, -1) / cnorm)),
                      measure="max |FP32 - FP64| over the mode's value scale / eigenvector norm")
    log(f"fp32 error vs fp64: {fp32_error}")
    head = results[head_name]
    wall_fast = head["ms_per_step"]
    k_fast = head["kernel_ms"]

    bytes_per_q = 24 + B * sv * 4 + B * sc * 8
    tb = table_bytes(T)
    alg = n * bytes_per_q + tb
    # The path is one pipeline of kernels per call (fold, bucketing, sums): the
    # algorithmic bytes of the call (Q in, outputs out, tables once) over the
    # HIP-event time of all its kernels.  The largest stage is the sums kernel.
    achieved = alg / (k_fast * 1e-3) / 1e9
    per_q = measured_traffic(cfg, head_name)
    # a call runs in passes of 2^24 points (one k_bin_sums dispatch each); with
    # several passes the next pass is prepared while the sums run, so a
    # stage's busy time can exceed its share of the call
    passes = -(-n // (1 << 24))
    sums_ms = head["stages"].get("k_bin_sums")
    roofline = dict(bound="hbm", achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s", frac=achieved / HBM_PEAK_GBS,
                    traffic=(per_q * n if per_q else None), bytes_per_q=bytes_per_q, table_bytes=tb,
                    traffic_bytes_per_q=per_q,
                    timed=f"{head_name}: all kernels of one call, HIP events on the call's stream",
                    kernel="binned pipeline (k_bin_fold, k_scan_*, k_bin_scatter, k_bucket_perm, "
                           "k_bin_propose when learned, k_bin_sums)",
                    dominant_kernel=dict(name="k_bin_sums<float, Gamma layout>", ms=sums_ms,
                                         dispatches_per_call=passes,
                                         avg_dispatch_ms=(sums_ms / passes if sums_ms else None),
                                         share=(sums_ms / k_fast if sums_ms else None),
                                         output_bytes_per_q=B * sv * 4 + B * sc * 8,
                                         output_gbs=(n * (B * sv * 4 + B * sc * 8) / (sums_ms * 1e-3) / 1e9
                                                     if sums_ms else None)),
                    stages=head["stages"])
    cpu = None if args.no_cpu else cpu_baseline(T, cfg, c, args.cpu_seconds)
    out = {
        "metric": "Q-points/s ir_interpolate_at, 1 MI355X: learned+batched vs plain, %HBM peak",
        "value": n / wall_fast * 1e3,
        "unit": "Q-points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall_fast,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"config{cfg}: {c['name']}, node volume {c['volume_fraction']:g} x V_IR, "
                               f"{n:.3g} Q in [-{c['box']},{c['box']}]^3, FP64 fold-in/weights, FP32/c64 sums; "
                               f"headline path: {head_name}",
                   "n_q": n, "vertices": int(T["vertices"].shape[0]), "tetrahedra": int(T["tet_volume"].size),
                   "modes": B, "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "plain": plain_res,
        "speedup_vs_plain": (plain_res["ms_per_step"] / wall_fast) if plain_res else None,
        # the same comparison against the strongest baselines, not only the literal FP64 kernel
        "speedup_vs_binned_no_proposal": (results["binned fp32"]["ms_per_step"] / wall_fast
                                          if "binned fp32" in results else None),
        "speedup_vs_tile_fp32": (results["tile fp32"]["ms_per_step"] / wall_fast if "tile fp32" in results else None),
        "speedup_vs_binned_fp64": (results["binned fp64"]["ms_per_step"] / wall_fast
                                   if "binned fp64" in results else None),
        "paths": results,
        "learned": learned,
        "fp32_error": fp32_error,
    }
    if not args.no_sort:
        out["sort_config4"] = sort_leg(args, capi, S, T)
    if not args.no_batch:
        # configuration 2 beside the headline, timed in the same run: the
        # config-4 buffers are released first (the batch outputs take 152 GB)
        del dv, dc, dQ, dev
        res, _, _ = batch_leg(args, capi, S, with_baselines=False)
        out["batch_config2"] = res
    print(json.dumps(out), flush=True)


def batch_leg(args, capi, S, with_baselines):
    """Configuration 2: 64 uneven grids (vertex counts log-uniform in 2e3-2e5,
    no payload cap: each grid's payload is generated and filled in turn,
    keeping the FP32 rows), 1e7 Q, one persistent graph launch with the
    autotuned schedule; with_baselines adds one tile-kernel call per grid and
    the plain FP64 kernel grid by grid."""
    import ctypes as C
    lib = capi.lib()
    n_q = int(args.n_q) if (args.n_q is not None and args.config == 2) else 10_000_000
    t0 = time.time()
    specs = S.batch_specs(payload_cap=None)
    Q, offsets = S.batch_queries(n_q=n_q)
    devs, V, Bm, per_q_out, table_total = [], [], [], [], 0
    for spec in specs:
        g = S.batch_grid(spec)
        vals, vecs = S.batch_payload(g, spec)
        t = S.batch_layout_tables(g, spec)
        dev = capi.Grid(t)
        keep = capi.KEEP_BOTH if with_baselines else capi.KEEP_FP32
        dev.fill_ex(t, vals.ctypes.data, vecs.ctypes.data, capi.MEM_HOST, capi.FP64, keep)
        devs.append(dev)
        V.append(int(vals.shape[0]))
        Bm.append(spec["modes"])
        per_q_out.append(spec["modes"] * 4 + spec["modes"] * 3 * spec["atoms"] * 8)
        table_total += table_bytes(t)
        del vals, vecs
    log(f"64 grids built and filled in {time.time() - t0:.1f}s: V {min(V)}..{max(V)} (sum {sum(V)}), "
        f"modes {min(Bm)}..{max(Bm)}")
    batch = capi.Batch(devs)
    dQ = capi.DeviceArray.from_host(Q)
    vals, vecs = batch.outputs(offsets, capi.FP32)
    if getattr(args, "batch_schedule", None):
        tp, bpc, grab, order, bmin = [int(float(x)) for x in args.batch_schedule.split(",")]
        batch.set_schedule(tp, bpc, grab, order, binned_min=bmin)
        capi.check(batch.run(dQ, offsets, vals, vecs, precision=capi.FP32))
    else:
        capi.check(batch.run(dQ, offsets, vals, vecs, precision=capi.FP32, autotune=True))
    sch = batch.schedule()
    log(f"autotuned schedule: {sch}")
    for _ in range(args.warmup):
        capi.check(batch.run(dQ, offsets, vals, vecs, precision=capi.FP32))
    capi.check(lib.brille_device_sync())
    # a step is one batch call as a user makes it (slots staged, graph replayed,
    # error records checked); the graph replay alone is timed with HIP events
    w0 = time.perf_counter()
    for _ in range(args.steps):
        capi.check(batch.run(dQ, offsets, vals, vecs, precision=capi.FP32))
    capi.check(lib.brille_device_sync())
    wall = (time.perf_counter() - w0) * 1000.0 / args.steps
    kms = [batch.time(dQ, offsets, vals, vecs, precision=capi.FP32, iters=1) for _ in range(max(1, args.steps // 2))]
    counts = np.diff(offsets.astype(np.int64))
    alg = float((counts * (24 + np.array(per_q_out))).sum()) + table_total
    achieved = alg / (np.mean(kms) * 1e-3) / 1e9
    per_q = measured_traffic(2, "batch fp32", summary="batch") if n_q == 10_000_000 else None
    res = dict(value=n_q / wall * 1e3, unit="Q-points/s", ms_per_step=wall, batch_ms=float(np.mean(kms)),
               n_q=n_q, vertices_total=int(sum(V)), vertices=[min(V), max(V)], modes=[min(Bm), max(Bm)],
               schedule=sch, output_gb=float((counts * np.array(per_q_out)).sum() / 1e9),
               roofline=dict(bound="hbm", achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s",
                             frac=achieved / HBM_PEAK_GBS, traffic=(per_q * n_q if per_q else None),
                             traffic_bytes_per_q=per_q,
                             kernel="k_batch<float> graph replay, then the binned pipelines of the grids "
                                    "given at least binned_min points (HIP events around both)"))
    log(f"batched fp32: {wall:.3f} ms/step wall, batch {np.mean(kms):.3f} ms, {achieved:.0f} GB/s")
    if with_baselines:
        seq = batch.time(dQ, offsets, vals, vecs, precision=capi.FP32, kernel=capi.KERNEL_TILE, iters=1)
        res["per_grid_tile_ms"] = seq
        res["speedup_vs_per_grid_calls"] = seq / np.mean(kms)
        if not args.no_plain:
            # FP64 outputs of all 64 grids do not fit one device beside the FP32
            # ones: the plain baseline runs grid by grid into one buffer sized
            # for the largest grid
            del vals, vecs
            vs_max = max(int(np.prod(devs[g].out_shapes(int(counts[g]))[0])) for g in range(len(devs)))
            cs_max = max(int(np.prod(devs[g].out_shapes(int(counts[g]))[1])) for g in range(len(devs)))
            dv64 = capi.DeviceArray((max(vs_max, 1),), np.float64)
            dc64 = capi.DeviceArray((max(cs_max, 1),), np.complex128)
            o = capi.RunOpts(0, capi.FP64, 0, capi.MEM_DEVICE, capi.KERNEL_PLAIN, 0)
            plain = 0.0
            for g, dev in enumerate(devs):
                if counts[g] == 0:
                    continue
                t = capi.KernelTimes()
                qp = C.c_void_p(dQ.ptr.value + int(offsets[g]) * 24)
                capi.check(lib.brille_time_ir_interpolate_at(dev.h, qp, int(counts[g]), C.byref(o), dv64.ptr,
                                                             dc64.ptr, 1, None, C.byref(t)))
                plain += float(t.total_ms)
            res["per_grid_plain_fp64_ms"] = plain
            res["speedup_vs_plain"] = plain / wall
            del dv64, dc64
        log(f"per-grid tile calls {seq:.3f} ms; per-grid plain fp64 {res.get('per_grid_plain_fp64_ms')} ms")
    return res, Q, offsets


def bench_batch(args, capi, S):
    """`--config 2`: the batch as the benchmark line itself."""
    res, Q, offsets = batch_leg(args, capi, S, with_baselines=True)
    out = {
        "metric": "Q-points/s ir_interpolate_at, 1 MI355X: learned+batched vs plain, %HBM peak",
        "value": res["value"],
        "unit": "Q-points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"config2: 64 grids (space groups 221/139/62/14/194 cycling), V {res['vertices'][0]}.."
                               f"{res['vertices'][1]}, 3..72 modes, {res['n_q']:.3g} Q split Dirichlet(0.3), "
                               "one persistent graph launch",
                   "n_q": res["n_q"], "vertices_total": res["vertices_total"], "schedule": res["schedule"],
                   "parallelism": "single GPU"},
        "roofline": dict(res["roofline"], traffic=None),
        "cpu_baseline": None,
        "per_grid_tile_ms": res.get("per_grid_tile_ms"),
        "per_grid_plain_fp64_ms": res.get("per_grid_plain_fp64_ms"),
        "speedup_vs_plain": res.get("speedup_vs_plain"),
        "speedup_vs_per_grid_calls": res.get("speedup_vs_per_grid_calls"),
    }
    print(json.dumps(out), flush=True)


def sort_leg(args, capi, S, T):
    """The mode sort (fill(..., sort=True)) on the benchmark grid: a payload
    whose modes never cross, shuffled at every vertex; one brille_grid_sort
    timed after a warm-up sort, and the oracle's sort (C, OpenMP) on a bounded
    sample of the same pairs for the CPU rate."""
    t = dict(T)
    values, vectors, _ = S.shuffled_modes(t, seed=4)
    t["values"], t["vectors"] = values, vectors
    dev = capi.Grid(t)
    dev.sort()  # warm: work space, code objects
    lib = capi.lib()
    capi.check(lib.brille_device_sync())
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        dev.sort()
    capi.check(lib.brille_device_sync())
    ms = (time.perf_counter() - t0) * 1e3 / reps
    keys, index, perms = dev.permutations()
    n_pairs = len(keys) // 2
    res = dict(ms=ms, keys=int(len(keys)), pairs=int(n_pairs), pairs_per_s=n_pairs / ms * 1e3,
               non_identity=int(len(perms) - 1), modes=int(t["branches"]),
               workload="config-4 grid, 121 200 vertices, 6 modes shuffled at every vertex")
    if not args.no_cpu:
        from oracle import oracle as O
        hw = host_cpu()
        V = np.uint64(t["vertices"].shape[0])
        i = (keys // V).astype(np.uint32)
        j = (keys % V).astype(np.uint32)
        fwd = np.flatnonzero(i < j)
        sample = fwd[: min(len(fwd), 400_000)]
        pairs = np.stack([i[sample], j[sample]], axis=1)
        c0 = time.perf_counter()
        rc, row, col = O.sort_pairs(t, pairs, nthreads=hw["threads"])
        cs = time.perf_counter() - c0
        assert rc == 0
        # the sample's permutations agree with the device table
        pos = np.searchsorted(keys, pairs[:, 0].astype(np.uint64) * V + pairs[:, 1].astype(np.uint64))
        assert np.array_equal(perms[index[pos]], row.astype(np.uint32))
        res["cpu_baseline"] = dict(pairs_per_s=len(pairs) / cs, sample=f"{len(pairs)} pairs of the same table",
                                   threads=hw["threads"], kind="port")
        res["speedup_vs_cpu"] = res["pairs_per_s"] / res["cpu_baseline"]["pairs_per_s"]
    log(f"sort: {res}")
    del dev
    return res


def measured_traffic(cfg, path_name, summary="binned"):
    """HBM bytes per Q of one call of a path, from the committed rocprofv3
    --pmc summary (FETCH_SIZE + WRITE_SIZE of every kernel of the call, scaled
    by the copy-kernel calibration, over Q); None when there is no summary for
    this config and path."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{summary}_config{cfg}.md")
    if not os.path.exists(path):
        return None
    key = "traffic_bytes_per_q[" + path_name + "]:"
    for line in open(path):
        if line.startswith(key):
            return float(line.split(":", 1)[1])
    return None


def host_cpu():
    """Host CPU model, the thread counts the process sees and the threads used:
    16 at most, the host-thread share a one-GPU job is given on the GPU box
    (more may be visible there, but they serve other jobs)."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return dict(model=model, nproc=os.cpu_count(), usable=usable,
                omp_num_threads=os.environ.get("OMP_NUM_THREADS"), threads=min(16, usable))


# the reference's one published measurement of this path: 1e5 Q in 19.56 s on
# 6 OpenMP threads (stage (a) 7.14 s, (b)+(c1) 9.13 s, (c2) 3.29 s), La2Zr2O7
# grid, brille 0.4.3, unspecified CPU (validation/profiling.md:85-93)
REFERENCE_PUBLISHED = dict(value=1e5 / 19.56, unit="Q-points/s", threads=6,
                           stages_s={"(a) ir_moveinto": 7.14, "(b)+(c1) locate + sums": 9.13, "(c2) rotation": 3.29},
                           source="validation/profiling.md:85-93 (1e5 Q, La2Zr2O7 grid, 6 threads)")


def cpu_baseline(T, cfg, c, seconds):
    """The oracle (C, OpenMP, the reference's loop structure: serial fold-in,
    dynamic-scheduled location and sums, static-scheduled rotation) on a
    bounded sample of the same workload, with its stage times."""
    from brille_amd import synthetic as S
    from oracle import oracle as O
    hw = host_cpu()
    cores = hw["threads"]
    probe = S.q_points(cfg, 20000, stream=7)
    tp = time.perf_counter()
    O.ir_interpolate_at(T, probe, nthreads=cores)
    rate = len(probe) / (time.perf_counter() - tp)
    m = int(min(max(rate * seconds, 20000), 5e6))
    sample = S.q_points(cfg, m, stream=8)
    tc = time.perf_counter()
    rc, *_ = O.ir_interpolate_at(T, sample, nthreads=cores)
    cpu_s = time.perf_counter() - tc
    st = O.stage_seconds()
    log(f"cpu baseline: {m / cpu_s:.4g} Q/s on {cores} threads, stages {st}")
    return dict(value=m / cpu_s, unit="Q-points/s", cores=cores, kind="port",
                sample=f"{m} Q of the same workload (uniform in [-{c['box']},{c['box']}]^3), FP64, {cores} OpenMP threads",
                cpu=hw,
                stages_s={"(a) ir_moveinto (serial)": round(float(st[0]), 4),
                          "(b)+(c1) locate + sums": round(float(st[1]), 4),
                          "(c2) rotation": round(float(st[2]), 4)},
                note="the timed code is this build's C restatement of the reference's loops (the reference needs "
                     "HighFive to compile); it is much faster than the reference's one published figure, "
                     "given beside it",
                reference_published=REFERENCE_PUBLISHED)


if __name__ == "__main__":
    main()
