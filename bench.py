#!/usr/bin/env python3
"""Benchmark of the heat-method geodesic + wiring-metric hot path on one MI355X.

Metric: cortical sources per second on the 164k-vertex folded surface (icosphere
L7, R = 70 mm, SH folding to degree 32, 0.3h tangential jitter, 10% masked cap:
148,165 cortical vertices, 295,882 faces), 12 area scales 0.002-0.05 and 3
samples between scales. One step = one batch of `--streams` RHS blocks of
`--block` sources (the blocks the device keeps in flight, one per HIP stream),
each through heat solve -> fused divergence -> Poisson solve -> metric kernel,
with outputs copied back. The mesh, operators and factors are resident in HBM before the
timed region; the default run times every cortical source once, in each
precision form (plain FP64; FP32 Poisson sweeps refined against the FP64
residual; split-BF16x3 Poisson sweeps refined the same way).

Also reported:
* the end-to-end latency of one whole hemisphere, measured (surface in host
  memory -> device mesh preparation -> analysis + both factorisations -> every
  source -> outputs in host memory), per form;
* the serving stream (40 hemisphere requests, Poisson arrivals with mean gap
  5 s, levels 5/6/7 at 30/40/30%, caps 5-20%): p50 / p95 / p99 latency;
* the roofline of the dominant kernel family, and per launch class (leaf level,
  wide levels, narrow levels) from HIP events around every launch;
* accuracy against the CPU FP64 oracle on a seeded sample and against the FP64
  form over every source, each labelled with the sources it covers;
* the CPU FP64 oracle (SuperLU + C metric loop) timed on the host's cores, one
  process per core, on a bounded seeded sample of sources.
"""

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "sources/s @164k-vertex mesh, batched FP32+IR vs plain FP64; p95 latency/hemi"
HBM_PEAK_GBS = 8000.0
PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3}
BOX_CPU_SHARE = 16  # host cores one GPU's job may use on the GPU hosts
WORKLOAD_NAMES = {
    0: "icosphere L4 sphere R=50mm, closed (2,562 vertices)",
    1: "folded icosphere L5 R=70mm SH<=16, 10% cap",
    2: "folded icosphere L6 R=70mm SH<=16, 10% cap",
    3: "folded icosphere L7 R=70mm SH<=32, 0.3h jitter, 10% cap (163,842-vertex surface)",
}
CPU_KEEP_FIELDS = 4  # distance fields per CPU worker kept for the distance accuracy check


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def prepare(workload):
    from fastcorticalwiring_amd import mesh_prep as mp
    from fastcorticalwiring_amd import meshes

    s = meshes.config_surface(workload)
    V, F, M = mp.validate_mesh_inputs(s.vertices, s.faces, s.cortex_mask)
    Vs, Fs, s2o, o2s = mp.build_cortex_submesh(V, F, M)
    fa, fl, va = mp.face_geometry(Vs, Fs)
    prep = dict(surface=s, V=Vs, F=Fs, vertex_areas=va, health_h=float(np.median(fl)),
                width=mp.default_bin_width(Vs, Fs), boundary=mp.boundary_indices(Fs),
                order=mp.bfs_order_all(Vs.shape[0], Fs), n_full=V.shape[0])
    prep["targets"] = np.array(meshes.TWELVE_SCALES) * float(va.sum())
    prep["fracs"] = np.array([0.25, 0.5, 0.75])
    return prep


def bench_prepare_surface(workload):
    """Only what the tuner needs of a surface: the submesh and its BFS order."""
    return prepare(workload)


def cpu_sample(n, n_sources, seed=0):
    return np.random.default_rng(seed).choice(n, size=min(n_sources, n), replace=False)


# --------------------------------------------------------------------------------------------
# CPU baseline leg (the oracle): one process per core, each timing its share of the sample
# --------------------------------------------------------------------------------------------
def cpu_worker(workload, n_sources, rank, nworkers, out_path):
    from oracle import analysis as oa
    from oracle import geometry as og
    from oracle import heat as oh

    prep = prepare(workload)
    s = prep["surface"]
    t0 = time.perf_counter()
    mesh = oa.OracleMesh(s.vertices, s.faces, s.cortex_mask)
    H = oh.HeatOracle(mesh.V, mesh.F)
    t_setup = time.perf_counter() - t0
    src = np.array_split(cpu_sample(mesh.n, n_sources), nworkers)[rank]
    T = prep["targets"]
    fr = prep["fracs"]
    keep = dict(D=[], radii=[], perim=[], msd_u=[], msd_w=[], dist_b=[], flag=[])
    t0 = time.perf_counter()
    for i, sidx in enumerate(src.tolist()):  # one source at a time, the reference loop's shape
        d = H.compute_distance(sidx)  # (core_analysis.py:1550-1619)
        neg, viol, bad = og.field_health(d, mesh.V, sidx, mesh.health_h)
        if i < CPU_KEEP_FIELDS:
            keep["D"].append(d)
        keep["flag"].append(bad)
        if bad:
            for key in ("radii", "perim"):
                keep[key].append(np.full(T.size, np.nan))
            for key in ("msd_u", "msd_w", "dist_b"):
                keep[key].append(np.nan)
            continue
        if mesh.boundary.size:
            bd = d[mesh.boundary]
            keep["dist_b"].append(float(np.min(bd[np.isfinite(bd)])))
        else:
            keep["dist_b"].append(np.inf)
        valid = (d > mesh.eps) & np.isfinite(d)
        keep["msd_u"].append(np.mean(d[valid]))
        keep["msd_w"].append(float((d[valid] * mesh.vertex_areas[valid]).sum() / mesh.vertex_areas[valid].sum()))
        g = og.per_source_geometry(d, mesh.F, mesh.lean_area, mesh.g11, mesh.g12, mesh.g22, mesh.width, T, fr)
        keep["radii"].append(g["radii"][0])
        keep["perim"].append(g["perim"][0])
    dt = time.perf_counter() - t0
    np.savez(out_path + ".npz", src=src, health_h=mesh.health_h, **{k: np.asarray(v) for k, v in keep.items()})
    with open(out_path, "w") as f:
        json.dump(dict(sources=int(src.size), seconds=dt, setup_seconds=t_setup), f)
    log(f"bench: CPU process {rank}: {src.size} sources in {dt:.1f}s (factorisation {t_setup:.1f}s)")


def start_cpu_leg(args):
    cores = min(len(os.sched_getaffinity(0)), BOX_CPU_SHARE)
    nworkers = args.cpu_workers or max(1, cores - 1)  # one core left for the GPU host thread
    outdir = os.path.join(ROOT, "gpurun_out" if os.path.isdir(os.path.join(ROOT, "gpurun_out")) else ".")
    env = dict(os.environ, OMP_NUM_THREADS="1", OPENBLAS_NUM_THREADS="1", MKL_NUM_THREADS="1")
    procs = []
    for r in range(nworkers):
        out = os.path.join(outdir, f".cpu_baseline_{os.getpid()}_{r}.json")
        cmd = [sys.executable, os.path.abspath(__file__), "--cpu-worker", out, "--cpu-rank", str(r),
               "--cpu-workers", str(nworkers), "--workload", str(args.workload), "--cpu-sources",
               str(args.cpu_sources)]
        procs.append((subprocess.Popen(cmd, env=env), out))
    return dict(procs=procs, nworkers=nworkers, cores_seen=len(os.sched_getaffinity(0)))


def finish_cpu_leg(leg, n_sources):
    parts, arrays = [], []
    for p, out in leg["procs"]:
        rc = p.wait()
        if rc != 0 or not os.path.exists(out):
            log(f"bench: CPU worker failed (rc {rc})")
            return None, None
        with open(out) as f:
            parts.append(json.load(f))
        with np.load(out + ".npz") as z:
            arrays.append({k: z[k] for k in z.files})
        os.remove(out)
        os.remove(out + ".npz")
    total = sum(p["sources"] for p in parts)
    rate = sum(p["sources"] / p["seconds"] for p in parts if p["sources"])
    try:
        import potpourri3d  # noqa: F401  (the reference's production solver)
        pp3d = "importable but not timed"
    except ImportError:
        pp3d = "unavailable offline (not installed; no package index)"
    cpu = dict(value=rate, unit="sources/s", cores=leg["nworkers"], kind="port",
               sample=(f"{total} seeded sources of the same mesh over {leg['nworkers']} processes, one per core "
                       f"(OMP/BLAS 1 thread; {leg['cores_seen']} cores in the affinity set, {BOX_CPU_SHARE} the "
                       f"host share): per source SuperLU (MMD on A^T+A, symmetric mode) FP64 heat + Poisson "
                       f"solves, C fused divergence, numpy health/MSD/boundary, C lean geometry; rate = sum of the "
                       f"processes' rates; factorisation of both systems "
                       f"{np.mean([p['setup_seconds'] for p in parts]):.1f}s per process excluded"),
               potpourri3d=pp3d, seconds_per_process=float(np.mean([p["seconds"] for p in parts])))
    ref = {k: np.concatenate([a[k] for a in arrays]) for k in ("src", "radii", "perim", "msd_u", "msd_w", "dist_b")}
    ref["D_src"] = np.concatenate([a["src"][:CPU_KEEP_FIELDS] for a in arrays])
    ref["D"] = np.concatenate([a["D"] for a in arrays if a["D"].size])
    ref["health_h"] = float(arrays[0]["health_h"])
    return cpu, ref


# --------------------------------------------------------------------------------------------
# GPU measurement
# --------------------------------------------------------------------------------------------
def run_gpu(prep, precision, block, steps, warmup, profile_blocks, streams, check_src, check_d_src, leaf_size=64,
            tile_rows=0):
    from fastcorticalwiring_amd.engine import HeatContext

    t0 = time.perf_counter()
    ctx = HeatContext(prep["V"], prep["F"], block=block, precision=precision, streams=streams, leaf_size=leaf_size,
                      tile_rows=tile_rows)
    ctx.set_metrics(prep["targets"], prep["fracs"], prep["boundary"], prep["vertex_areas"],
                    health_h=prep["health_h"], bin_width=prep["width"])
    setup = time.perf_counter() - t0
    order = prep["order"]
    k = ctx.block * streams  # sources per step: one block per stream
    if steps is None:
        src = order
        steps = (order.size + k - 1) // k
    else:
        src = np.resize(order, steps * k)
    if warmup:
        ctx.wiring(np.resize(order, warmup * k))
    t0 = time.perf_counter()
    res = ctx.wiring(src)
    wall = time.perf_counter() - t0
    issue = ctx.stage_times()["issue"]  # host seconds spent issuing the timed blocks
    outputs = dict(src=src, radii=res["radii"], perim=res["perimeters"], msd_u=res["scalars"][:, 4],
                   msd_w=res["scalars"][:, 5], dist_b=res["scalars"][:, 3])
    chk = ctx.wiring(check_src)
    outputs["check"] = dict(src=check_src, radii=chk["radii"], perim=chk["perimeters"], msd_u=chk["scalars"][:, 4],
                            msd_w=chk["scalars"][:, 5], dist_b=chk["scalars"][:, 3],
                            D_src=check_d_src, D=ctx.distances(check_d_src))
    out = dict(precision=precision, sources=int(src.size), steps=int(steps), seconds=wall,
               sources_per_s=src.size / wall, ms_per_step=1e3 * wall / steps, setup_s=setup, host_issue_s=issue,
               flagged=int(np.sum(res["scalars"][:, 0] > 0)), band_overflow=int(np.sum(res["scalars"][:, 8] > 0)),
               info=ctx.info, outputs=outputs, all_sources=bool(np.array_equal(src, order)))
    # distance fields handed to the caller's host memory (fcw_distance_block: solve, shift, one
    # device-to-host copy of k x n doubles per block), the adapter's path; PCIe-inclusive, reported
    # beside the resident-input rate, never as it
    nblk, kb = 4, ctx.block
    ctx.distance_block(order[:kb])
    t1 = time.perf_counter()
    for b in range(nblk):
        ctx.distance_block(np.resize(order, (b + 2) * kb)[(b + 1) * kb:])
    out["distances_to_host_sources_per_s"] = nblk * kb / (time.perf_counter() - t1)
    # per-launch timing with HIP events on the context's streams
    ctx.set_profiling(True)
    ctx.wiring(np.resize(order, profile_blocks * kb))
    out["kernel_ms"] = ctx.kernel_times()
    out["kernel_log"] = ctx.kernel_log()
    out["stage_s"] = ctx.stage_times()
    out["profile_blocks"] = profile_blocks
    ctx.close()
    return out


def hemisphere_latency(prep, precision, block, streams, leaf_size=64, tile_rows=0):
    """One whole hemisphere from the surface in host memory to every source's outputs in host
    memory: device mesh preparation, analysis and both factorisations, all sources, copy back."""
    from fastcorticalwiring_amd.serving import HemisphereRequest, ServingEngine

    s = prep["surface"]
    req = HemisphereRequest(0, 0.0, s.vertices, s.faces, s.cortex_mask)
    eng = ServingEngine(max_active=1, block=block, precision=precision, keep_outputs=False, streams=streams,
                        chunk_blocks=1 << 20, leaf_size=leaf_size, tile_rows=tile_rows)
    r = eng.run([req])[0]
    if r.error:
        raise RuntimeError(r.error)
    return dict(latency_s=r.latency_s, prepare_s=r.ready_s - r.admitted_s, sources=r.n_sources,
                end_to_end_sources_per_s=r.n_sources / r.latency_s)


def _rel(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    ok = np.isfinite(a) & np.isfinite(b)
    if not ok.any():
        return None
    return float(np.max(np.abs(a[ok] - b[ok]) / np.maximum(np.abs(b[ok]), 1e-300)))


def accuracy(out, ref, health_h=None):
    """Max relative differences of one form's outputs against a reference (the CPU oracle's sample,
    or the FP64 form), with the number of sources compared; distances as compare_fastcw_lean.py:235-239
    (relative where the reference exceeds 5h, and max abs / max)."""
    acc = {"sources": int(np.asarray(ref["msd_u"]).size),
           "msd_u_rel": _rel(out["msd_u"], ref["msd_u"]), "msd_w_rel": _rel(out["msd_w"], ref["msd_w"]),
           "dist_to_boundary_rel": _rel(out["dist_b"], ref["dist_b"]),
           "radius_rel": _rel(out["radii"], ref["radii"]), "perimeter_rel": _rel(out["perim"], ref["perim"])}
    if "D" in out and "D" in ref and np.asarray(ref["D"]).size:
        D, R = np.asarray(out["D"]), np.asarray(ref["D"])
        acc["distance_fields"] = int(R.shape[0])
        acc["distance_max_abs_over_max"] = float(np.max(np.abs(D - R)) / np.max(R))
        far = R > 5.0 * health_h
        acc["distance_rel_far"] = float(np.max(np.abs(D[far] - R[far]) / R[far]))
    return acc


def _launch_class(level, wgs):
    if level == 0:
        return "leaves"
    return "wide" if wgs >= 1024 else "narrow"


def rooflines(gpu, block):
    """Achieved rate per kernel family and, for the sweeps, per (system, level class), from the
    HIP-event log of the profiling pass (full blocks of `block` sources)."""
    info = gpu["info"]
    n, m = info["n"], info["m"]
    log_ = gpu["kernel_log"]
    nb = gpu["profile_blocks"]
    fam, cls = {}, {}
    for rec in log_:
        kind = str(rec["kind"])
        ms = float(rec["ms"])
        if kind in ("fwd_gemm", "bwd_gemm", "fwd_gemm32", "bwd_gemm32"):
            work = 2.0 * float(rec["entries"]) * block  # 2 flops per factor entry per right-hand side
            peak = PEAK_TFLOPS["fp32" if kind.endswith("32") else "fp64"]
            unit, bound = "TFLOP/s", "mfma"
            key = f"{kind}.{'heat' if rec['system'] == 0 else 'poisson'}.{_launch_class(rec['level'], rec['wgs'])}"
            c = cls.setdefault(key, dict(ms=0.0, work=0.0, launches=0, peak=peak, unit=unit, bound=bound))
            c["ms"] += ms
            c["work"] += work
            c["launches"] += 1
        elif kind == "divergence":
            work = 2.0 * n * block * 8  # heat solution read, Poisson right-hand side written
            peak, unit, bound = HBM_PEAK_GBS, "GB/s", "hbm"
        elif kind == "chunk_stats":
            work = 1.0 * n * block * 8  # the field read once
            peak, unit, bound = HBM_PEAK_GBS, "GB/s", "hbm"
        elif kind == "metrics":
            work = block * n * 8 + m * (3 * 4 + 4 * 8) + n * (3 * 8 + 8)  # field + face and vertex data
            peak, unit, bound = HBM_PEAK_GBS, "GB/s", "hbm"
        elif kind == "residual":
            work = n * block * (8 + 8 + 4) + 12.0 * 7 * n  # x and b read, FP32 r written, CSR once
            peak, unit, bound = HBM_PEAK_GBS, "GB/s", "hbm"
        else:
            continue
        f = fam.setdefault(kind, dict(ms=0.0, work=0.0, launches=0, peak=peak, unit=unit, bound=bound))
        f["ms"] += ms
        f["work"] += work
        f["launches"] += 1

    def fin(d):
        out = {}
        for k, v in d.items():
            scale = 1e12 if v["unit"] == "TFLOP/s" else 1e9
            ach = v["work"] / (v["ms"] * 1e-3) / scale if v["ms"] > 0 else 0.0
            out[k] = dict(bound=v["bound"], achieved=ach, peak=v["peak"], unit=v["unit"], frac=ach / v["peak"],
                          avg_launch_ms=v["ms"] / max(v["launches"], 1), launches=v["launches"],
                          ms_per_block=v["ms"] / nb, work_per_launch=v["work"] / max(v["launches"], 1))
        return out

    return fin(fam), fin(cls)


def load_traffic(kernel):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if present."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            data = json.load(f)
        return data.get(kernel, {}).get("bytes_per_launch")
    except Exception:
        return None


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="batches of --streams RHS blocks timed (default: every cortical source once)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", type=int, default=3)
    ap.add_argument("--block", type=int, default=256, help="RHS block width when not autotuned")
    ap.add_argument("--leaf-size", type=int, default=64, help="nested-dissection leaf size when not autotuned")
    ap.add_argument("--autotune", default="on", choices=("on", "off"),
                    help="time block widths {128,256,512} x leaf sizes {32,64,96,128} first and use the fastest")
    ap.add_argument("--streams", type=int, default=8, help="RHS blocks in flight (one HIP stream each)")
    ap.add_argument("--precision", default="auto", help="fp64 | fp32_ir | bf16x3_ir | auto (every form)")
    ap.add_argument("--cpu-baseline", default="on", choices=("on", "off"))
    ap.add_argument("--cpu-sources", type=int, default=512)
    ap.add_argument("--cpu-workers", type=int, default=0, help="CPU processes (0: host share minus one)")
    ap.add_argument("--profile-blocks", type=int, default=4)
    ap.add_argument("--latency", default="on", choices=("on", "off"),
                    help="measure one whole hemisphere end to end per form")
    ap.add_argument("--serving-requests", type=int, default=40,
                    help="hemisphere requests of the seeded serving stream replayed at their arrival times (0 = skip)")
    ap.add_argument("--serving-precision", default="fp32_ir")
    ap.add_argument("--serving-time-scale", type=float, default=1.0)
    ap.add_argument("--calibrate-bytes", type=int, default=0,
                    help="run one plain copy kernel of this many bytes first (byte-counter calibration)")
    ap.add_argument("--cpu-worker", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--cpu-rank", type=int, default=0, help=argparse.SUPPRESS)
    args = ap.parse_args()

    if args.cpu_worker:
        cpu_worker(args.workload, args.cpu_sources, args.cpu_rank, args.cpu_workers, args.cpu_worker)
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(sources shard across GPUs with no exchange; see the design notes).", file=sys.stderr)
        return 2

    leg = start_cpu_leg(args) if args.cpu_baseline == "on" else None
    if leg:
        log(f"bench: CPU baseline leg started, {leg['nworkers']} processes")
    log(f"bench: preparing workload {args.workload}")
    prep = prepare(args.workload)
    n = prep["V"].shape[0]
    log(f"bench: {n} cortical vertices, {prep['F'].shape[0]} faces")
    from fastcorticalwiring_amd import _lib

    if args.calibrate_bytes:
        _lib.check(_lib.lib().fcw_calibration_copy(int(args.calibrate_bytes), 0))
    forms = ["fp64", "fp32_ir", "bf16x3_ir"] if args.precision == "auto" else [args.precision]
    tuned = None
    block, leaf, tile = args.block, args.leaf_size, 0
    if args.autotune == "on":
        from fastcorticalwiring_amd import autotune

        log("bench: autotuning block width, leaf size and sweep tile height")
        best_cfg, trials = autotune.autotune(prep["V"], prep["F"], blocks=(128, 256, 512), leaf_sizes=(32, 64, 96, 128),
                                             precision=forms[0], streams=args.streams, order=prep["order"],
                                             n_sources=min(prep["order"].size, 16384), log=log, tile_rows=(0, 64, 32))
        block, leaf, tile = best_cfg["block"], best_cfg["leaf_size"], best_cfg["tile_rows"]
        tuned = dict(block=block, leaf_size=leaf, tile_rows=tile, precision=forms[0], trials=trials)
    check_src = cpu_sample(n, args.cpu_sources)  # the CPU leg's sample, checked on the GPU as well
    nw = leg["nworkers"] if leg else 1
    check_d_src = np.concatenate([p[:CPU_KEEP_FIELDS] for p in np.array_split(check_src, nw)])
    results = {}
    for form in forms:
        log(f"bench: timing {form}")
        results[form] = run_gpu(prep, form, block, args.steps, args.warmup, args.profile_blocks,
                                args.streams, check_src, check_d_src, leaf, tile)
        r = results[form]
        log(f"bench: {form}: {r['sources_per_s']:.1f} sources/s ({r['ms_per_step']:.2f} ms/block), "
            f"setup {r['setup_s']:.2f}s")
    cpu, oracle_out = (None, None)
    if leg:
        log("bench: waiting for the CPU baseline leg")
        cpu, oracle_out = finish_cpu_leg(leg, args.cpu_sources)
    latency = {}
    if args.latency == "on":
        for form in forms:
            latency[form] = hemisphere_latency(prep, form, block, args.streams, leaf, tile)
            log(f"bench: {form}: one hemisphere end to end {latency[form]['latency_s']:.2f}s")
    best = max(results.values(), key=lambda r: r["sources_per_s"])
    serving = None
    if args.serving_requests > 0:
        from fastcorticalwiring_amd.serving import ServingEngine, stream_requests

        from fastcorticalwiring_amd import autotune
        from fastcorticalwiring_amd import meshes as _meshes

        log(f"bench: serving stream, {args.serving_requests} requests, {args.serving_precision}")
        reqs = stream_requests(args.serving_requests)
        # block width and leaf size per size class, tuned before the stream on one surface of each level
        tuner = autotune.Tuner(precision=args.serving_precision, streams=4, blocks=(256, 512), leaf_sizes=(64, 96),
                               log=log)
        for lvl in sorted({r.meta["level"] for r in reqs}):
            wl = {5: 1, 6: 2, 7: 3}[lvl]
            sp = prep if wl == args.workload else bench_prepare_surface(wl)
            tuner.choose(sp["V"], sp["F"], sp["order"])
        def done(r):
            log(f"bench: request {r.rid} ({r.n_sources} sources) done, latency {r.latency_s:.2f}s"
                + (f", error {r.error}" if r.error else ""))

        eng = ServingEngine(max_active=3, block=block, precision=args.serving_precision, keep_outputs=False,
                            on_done=done, tuner=tuner)
        sres = eng.run(reqs, time_scale=args.serving_time_scale)
        serving = ServingEngine.summary(sres)
        serving.update(precision=args.serving_precision, time_scale=args.serving_time_scale,
                       tuned={str(k): v for k, v in tuner.choices.items()},
                       levels=[r.meta["level"] for r in reqs],
                       stream="folded icospheres L5/L6/L7 (30/40/30%), caps 5-20%, Poisson arrivals mean gap 5 s")
        log(f"bench: serving p50 {serving['p50_s']:.2f}s p95 {serving['p95_s']:.2f}s "
            f"p99 {serving['p99_s']:.2f}s, {serving['hemispheres_per_hour']:.0f} hemispheres/hour")
    accuracy_report = {}
    for form, r in results.items():
        rep = {}
        if oracle_out is not None and np.array_equal(oracle_out["src"], r["outputs"]["check"]["src"]):
            chk = dict(r["outputs"]["check"])
            ref = dict(oracle_out)
            if not np.array_equal(chk["D_src"], ref["D_src"]):
                chk.pop("D")
            rep["vs_oracle"] = accuracy(chk, ref, oracle_out["health_h"])
        if form != "fp64" and "fp64" in results:
            a, b = r["outputs"], results["fp64"]["outputs"]
            if np.array_equal(a["src"], b["src"]):
                rep["vs_fp64"] = accuracy(a, b)
                rep["vs_fp64"]["all_cortical_sources"] = bool(r["all_sources"])
        accuracy_report[form] = rep
    fams, classes = rooflines(best, best["info"]["block"])
    dom = max(fams, key=lambda k: fams[k]["ms_per_block"])
    roof = dict(fams[dom])
    roof["traffic"] = load_traffic(dom)
    fp64_rate = results["fp64"]["sources_per_s"] if "fp64" in results else None
    line = {
        "metric": METRIC,
        "value": best["sources_per_s"],
        "unit": "sources/s",
        "n_gpus": 1,
        "steps": best["steps"],
        "warmup": args.warmup,
        "ms_per_step": best["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32_ir": "f32+f64ir", "bf16x3_ir": "bf16x3+f64ir"}[best["precision"]],
        "data": "synthetic",
        "config": {
            "workload": WORKLOAD_NAMES.get(args.workload, str(args.workload)),
            "cortical_sources": int(prep["order"].size),
            "surface_vertices": int(prep["n_full"]),
            "faces": int(prep["F"].shape[0]),
            "block": int(best["info"]["block"]),
            "sources_per_step": int(best["info"]["block"]) * int(args.streams),
            "leaf_size": int(leaf),
            "tile_rows": int(tile) if tile else "chosen per tree level",
            "streams": int(args.streams),
            "precision": best["precision"],
            "scales": 12,
            "samples_between_scales": 3,
        },
        "roofline": {"kernel": dom, "bound": roof["bound"], "achieved": roof["achieved"], "peak": roof["peak"],
                     "unit": roof["unit"], "frac": roof["frac"], "traffic": roof["traffic"],
                     "avg_launch_ms": roof["avg_launch_ms"]},
        "cpu_baseline": None if cpu is None else {k: cpu[k] for k in ("value", "unit", "cores", "kind", "sample")},
        "speedup_vs_fp64": None if fp64_rate is None else best["sources_per_s"] / fp64_rate,
        "forms": {k: {"sources_per_s": v["sources_per_s"], "ms_per_step": v["ms_per_step"],
                      "speedup_vs_fp64": None if fp64_rate is None else v["sources_per_s"] / fp64_rate,
                      "timed_sources": v["sources"], "all_cortical_sources": v["all_sources"],
                      "setup_s": v["setup_s"], "hemisphere": latency.get(k),
                      "distances_to_host_sources_per_s": v["distances_to_host_sources_per_s"],
                      "host_issue_s": v["host_issue_s"], "timed_s": v["seconds"], "flagged": v["flagged"],
                      "band_overflow": v["band_overflow"],
                      # the library's own split of the context creation inside setup_s
                      "setup_breakdown_s": dict(zip(("analyse", "assemble", "factor", "upload"),
                                                    v["info"]["setup_seconds"]))}
                  for k, v in results.items()},
        "hemisphere_latency_s": (latency.get(best["precision"]) or {}).get("latency_s"),
        "accuracy": accuracy_report,
        "serving": serving,
        "kernel_families": {k: {kk: v[kk] for kk in ("achieved", "unit", "frac", "avg_launch_ms", "launches",
                                                     "ms_per_block")} for k, v in fams.items()},
        "sweep_classes": {k: {kk: v[kk] for kk in ("achieved", "unit", "frac", "avg_launch_ms", "launches",
                                                   "ms_per_block")} for k, v in classes.items()},
        "stage_s_per_block": {k: v / best["profile_blocks"] for k, v in best["stage_s"].items()
                              if k not in ("wall", "issue")},
    }
    if tuned is not None:
        line["autotune"] = tuned
    if cpu is not None:
        line["cpu_baseline_detail"] = {k: cpu[k] for k in ("potpourri3d", "seconds_per_process")}
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
