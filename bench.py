#!/usr/bin/env python3
"""Benchmark of the heat-method geodesic + wiring-metric hot path on one MI355X.

Metric: cortical sources per second on the 164k-vertex folded surface (icosphere
L7, R = 70 mm, SH folding to degree 32, 0.3h tangential jitter, 10% masked cap:
148,165 cortical vertices, 295,882 faces), 12 area scales 0.002-0.05 and 3
samples between scales. One step = one RHS block of `--block` sources through
heat solve -> fused divergence -> Poisson solve -> shift -> metric kernel, with
outputs copied back. The mesh, operators and factors are resident in HBM before
the timed region; the default run times every cortical source once.

Also reported: the plain FP64 form and (when built) FP32+IR form side by side,
setup time (analysis + assembly + factorisation), the end-to-end hemisphere
latency, the roofline of the dominant kernel from HIP events around each of its
launches, and the CPU FP64 oracle (SuperLU + C metric loop, one core) timed on a
bounded sample of sources in a concurrent subprocess.
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
FP64_PEAK_TFLOPS = 78.6
WORKLOAD_NAMES = {
    0: "icosphere L4 sphere R=50mm, closed (2,562 vertices)",
    1: "folded icosphere L5 R=70mm SH<=16, 10% cap",
    2: "folded icosphere L6 R=70mm SH<=16, 10% cap",
    3: "folded icosphere L7 R=70mm SH<=32, 0.3h jitter, 10% cap (163,842-vertex surface)",
}


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


# --------------------------------------------------------------------------------------------
# CPU baseline leg (the oracle), run as a subprocess of this script
# --------------------------------------------------------------------------------------------
def cpu_worker(workload, n_sources, out_path, seed=0):
    from oracle import analysis as oa
    from oracle import geometry as og
    from oracle import heat as oh

    prep = prepare(workload)
    s = prep["surface"]
    t0 = time.perf_counter()
    mesh = oa.OracleMesh(s.vertices, s.faces, s.cortex_mask)
    H = oh.HeatOracle(mesh.V, mesh.F)
    t_setup = time.perf_counter() - t0
    rng = np.random.default_rng(seed)
    src = rng.choice(mesh.n, size=min(n_sources, mesh.n), replace=False)
    T = prep["targets"]
    fr = prep["fracs"]
    keep = dict(D=[], radii=[], perim=[], msd_u=[], msd_w=[], flag=[])
    t0 = time.perf_counter()
    for sidx in src.tolist():  # one source at a time, the reference loop's shape (core_analysis.py:1550-1619)
        d = H.compute_distance(sidx)
        neg, viol, bad = og.field_health(d, mesh.V, sidx, mesh.health_h)
        keep["D"].append(d)
        keep["flag"].append(bad)
        if bad:
            for key in ("radii", "perim"):
                keep[key].append(np.full(T.size, np.nan))
            keep["msd_u"].append(np.nan)
            keep["msd_w"].append(np.nan)
            continue
        if mesh.boundary.size:
            bd = d[mesh.boundary]
            _ = np.min(bd[np.isfinite(bd)])
        valid = (d > mesh.eps) & np.isfinite(d)
        keep["msd_u"].append(np.mean(d[valid]))
        keep["msd_w"].append(float((d[valid] * mesh.vertex_areas[valid]).sum() / mesh.vertex_areas[valid].sum()))
        g = og.per_source_geometry(d, mesh.F, mesh.lean_area, mesh.g11, mesh.g12, mesh.g22, mesh.width, T, fr)
        keep["radii"].append(g["radii"][0])
        keep["perim"].append(g["perim"][0])
    dt = time.perf_counter() - t0
    np.savez(out_path + ".npz", src=src, health_h=mesh.health_h, **{k: np.asarray(v) for k, v in keep.items()})
    res = dict(value=src.size / dt, unit="sources/s", cores=1, kind="port",
               sample=f"{src.size} seeded sources of the same mesh, one at a time: SuperLU (MMD on A^T+A, symmetric mode) FP64 "
                      f"heat + Poisson solves, C fused divergence, numpy health/MSD, C lean geometry; "
                      f"factorisation of both systems {t_setup:.1f}s excluded",
               seconds=dt, setup_seconds=t_setup)
    with open(out_path, "w") as f:
        json.dump(res, f)


# --------------------------------------------------------------------------------------------
# GPU measurement
# --------------------------------------------------------------------------------------------
def run_gpu(prep, precision, block, steps, warmup, profile_blocks, streams=8, check_src=None):
    from fastcorticalwiring_amd.engine import HeatContext

    t0 = time.perf_counter()
    ctx = HeatContext(prep["V"], prep["F"], block=block, precision=precision, streams=streams)
    ctx.set_metrics(prep["targets"], prep["fracs"], prep["boundary"], prep["vertex_areas"],
                    health_h=prep["health_h"], bin_width=prep["width"])
    setup = time.perf_counter() - t0
    order = prep["order"]
    k = ctx.block
    n_src = order.size
    if steps is None:
        src = order
        steps = (n_src + k - 1) // k
    else:
        src = np.resize(order, steps * k)
    if warmup:
        ctx.wiring(np.resize(order, warmup * k))
    t0 = time.perf_counter()
    res = ctx.wiring(src)
    wall = time.perf_counter() - t0
    issue = ctx.stage_times()["issue"]  # host seconds spent issuing the timed blocks
    flags = int(np.sum(res["scalars"][:, 0] > 0))
    ovf = int(np.sum(res["scalars"][:, 8] > 0))
    # outputs kept for the accuracy report (not timed): all sources' metrics in `src` order, and
    # the distance fields of the CPU leg's sample sources
    outputs = dict(src=src, radii=res["radii"], perim=res["perimeters"], msd_u=res["scalars"][:, 4],
                   msd_w=res["scalars"][:, 5])
    if check_src is not None:
        chk = ctx.wiring(check_src)
        outputs["check"] = dict(src=check_src, D=ctx.distances(check_src), radii=chk["radii"],
                                perim=chk["perimeters"], msd_u=chk["scalars"][:, 4], msd_w=chk["scalars"][:, 5])
    out = dict(precision=precision, sources=int(src.size), steps=int(steps), seconds=wall,
               sources_per_s=src.size / wall, ms_per_step=1e3 * wall / steps, setup_s=setup,
               hemisphere_latency_s=setup + wall * (n_src / src.size),
               end_to_end_sources_per_s=n_src / (setup + wall * (n_src / src.size)), host_issue_s=issue,
               flagged=flags, band_overflow=ovf,
               info=ctx.info, outputs=outputs)
    # distance fields handed to the caller's host memory (fcw_distance_block: solve, shift, one
    # device-to-host copy of k x n doubles per block), the adapter's path; PCIe-inclusive, reported
    # beside the resident-input rate, never as it
    nblk = 4
    ctx.distance_block(order[:k])
    t1 = time.perf_counter()
    for b in range(nblk):
        ctx.distance_block(np.resize(order, (b + 2) * k)[(b + 1) * k:])
    out["distances_to_host_sources_per_s"] = nblk * k / (time.perf_counter() - t1)
    # per-kernel timing with HIP events on the context's stream
    ctx.set_profiling(True)
    ctx.wiring(np.resize(order, profile_blocks * k))
    out["kernel_ms"] = ctx.kernel_times()
    out["stage_s"] = ctx.stage_times()
    out["profile_blocks"] = profile_blocks
    ctx.close()
    return out


def _rel(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    ok = np.isfinite(a) & np.isfinite(b)
    if not ok.any():
        return None
    return float(np.max(np.abs(a[ok] - b[ok]) / np.maximum(np.abs(b[ok]), 1e-300)))


def accuracy(out, ref, health_h=None):
    """Max relative differences of one form's outputs against a reference (the CPU oracle's sample,
    or the FP64 form over all sources); distances as compare_fastcw_lean.py:235-239 (relative where
    the reference exceeds 5h, and max abs / max)."""
    acc = {"msd_u_rel": _rel(out["msd_u"], ref["msd_u"]), "msd_w_rel": _rel(out["msd_w"], ref["msd_w"]),
           "radius_rel": _rel(out["radii"], ref["radii"]), "perimeter_rel": _rel(out["perim"], ref["perim"])}
    if "D" in out and "D" in ref:
        D, R = np.asarray(out["D"]), np.asarray(ref["D"])
        acc["distance_max_abs_over_max"] = float(np.max(np.abs(D - R)) / np.max(R))
        far = R > 5.0 * health_h
        acc["distance_rel_far"] = float(np.max(np.abs(D[far] - R[far]) / R[far]))
    return acc


def roofline(gpu, prep, block):
    info = gpu["info"]
    kt = gpu["kernel_ms"]
    nb = gpu["profile_blocks"]
    n, m = info["n"], info["m"]
    # algorithmic work per launch
    fam = {}
    tot_ms, cnt = kt["metrics"]
    # metric kernel: each field read once (n doubles per source), face and vertex data once per launch
    bytes_per_launch = block * n * 8 + m * (3 * 4 + 4 * 8) + n * (3 * 8 + 8)
    fam["metrics_kernel"] = dict(bound="hbm", total_ms=tot_ms, launches=cnt,
                                 work_per_launch=bytes_per_launch, unit="GB/s", peak=HBM_PEAK_GBS)
    for name, key in (("fwd_gemm_kernel", "fwd_gemm"), ("bwd_gemm_kernel", "bwd_gemm")):
        tms, c = kt[key]
        # 2 flops per factor nonzero per RHS per sweep (8 nnz(L) per source over 2 systems x 2 sweeps)
        flops_total = nb * 2 * 2.0 * info["factor_nnz"] * block
        fam[name] = dict(bound="mfma", total_ms=tms, launches=c, work_per_launch=flops_total / max(c, 1),
                         unit="TFLOP/s", peak=FP64_PEAK_TFLOPS)
    dom = max(fam, key=lambda k: fam[k]["total_ms"])
    out = {}
    for name, f in fam.items():
        avg_ms = f["total_ms"] / max(f["launches"], 1)
        if f["unit"] == "GB/s":
            ach = f["work_per_launch"] / (avg_ms * 1e-3) / 1e9
        else:
            ach = f["work_per_launch"] / (avg_ms * 1e-3) / 1e12
        out[name] = dict(bound=f["bound"], achieved=ach, peak=f["peak"], unit=f["unit"], frac=ach / f["peak"],
                         avg_launch_ms=avg_ms, launches=f["launches"], traffic=None)
    traffic = load_traffic(dom)
    if traffic is not None:
        out[dom]["traffic"] = traffic
    return dom, out


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
    ap.add_argument("--steps", type=int, default=None, help="RHS blocks timed (default: every cortical source once)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", type=int, default=3)
    ap.add_argument("--block", type=int, default=256)
    ap.add_argument("--streams", type=int, default=8, help="RHS blocks in flight (one HIP stream each)")
    ap.add_argument("--precision", default="auto", help="fp64 | fp32_ir | auto (every form available)")
    ap.add_argument("--cpu-baseline", default="on", choices=("on", "off"))
    ap.add_argument("--cpu-sources", type=int, default=24)
    ap.add_argument("--profile-blocks", type=int, default=4)
    ap.add_argument("--serving-requests", type=int, default=12,
                    help="hemisphere requests of the seeded serving stream replayed at their arrival times (0 = skip)")
    ap.add_argument("--serving-time-scale", type=float, default=1.0)
    ap.add_argument("--calibrate-bytes", type=int, default=0,
                    help="run one plain copy kernel of this many bytes first (byte-counter calibration)")
    ap.add_argument("--cpu-worker", default=None, help=argparse.SUPPRESS)
    args = ap.parse_args()

    if args.cpu_worker:
        cpu_worker(args.workload, args.cpu_sources, args.cpu_worker)
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(sources shard across GPUs with no exchange; see the design notes).", file=sys.stderr)
        return 2

    cpu_proc = None
    cpu_out = os.path.join(ROOT, "gpurun_out" if os.path.isdir(os.path.join(ROOT, "gpurun_out")) else ".",
                           f".cpu_baseline_{os.getpid()}.json")
    if args.cpu_baseline == "on":
        env = dict(os.environ, OMP_NUM_THREADS="1", OPENBLAS_NUM_THREADS="1", MKL_NUM_THREADS="1")
        cpu_proc = subprocess.Popen([sys.executable, os.path.abspath(__file__), "--cpu-worker", cpu_out,
                                     "--workload", str(args.workload), "--cpu-sources", str(args.cpu_sources)],
                                    env=env)
    log(f"bench: preparing workload {args.workload}")
    prep = prepare(args.workload)
    log(f"bench: {prep['V'].shape[0]} cortical vertices, {prep['F'].shape[0]} faces")
    from fastcorticalwiring_amd import _lib

    if args.calibrate_bytes:
        _lib.check(_lib.lib().fcw_calibration_copy(int(args.calibrate_bytes), 0))
    forms = ["fp64", "fp32_ir"] if args.precision == "auto" else [args.precision]
    # the CPU leg's sample sources (same seed and rule as cpu_worker), checked on the GPU as well
    check_src = np.random.default_rng(0).choice(prep["V"].shape[0], size=min(args.cpu_sources, prep["V"].shape[0]),
                                                replace=False)
    results = {}
    for form in forms:
        try:
            log(f"bench: timing {form}")
            results[form] = run_gpu(prep, form, args.block, args.steps, args.warmup, args.profile_blocks,
                                    args.streams, check_src=check_src)
            r = results[form]
            log(f"bench: {form}: {r['sources_per_s']:.1f} sources/s ({r['ms_per_step']:.2f} ms/block), "
                f"setup {r['setup_s']:.2f}s")
        except _lib.FcwError as exc:
            if args.precision == "auto" and form != "fp64":
                log(f"bench: {form} unavailable: {exc}")
                continue
            raise
    best = max(results.values(), key=lambda r: r["sources_per_s"])
    dom, roofs = roofline(best, prep, best["info"]["block"])
    serving = None
    if args.serving_requests > 0:
        from fastcorticalwiring_amd.serving import ServingEngine, stream_requests

        log(f"bench: serving stream, {args.serving_requests} requests")
        reqs = stream_requests(args.serving_requests)
        eng = ServingEngine(max_active=3, block=args.block, precision=best["precision"], keep_outputs=False)
        sres = eng.run(reqs, time_scale=args.serving_time_scale)
        serving = ServingEngine.summary(sres)
        serving.update(time_scale=args.serving_time_scale,
                       levels=[r.meta["level"] for r in reqs],
                       stream="folded icospheres L5/L6/L7 (30/40/30%), caps 5-20%, Poisson arrivals mean gap 5 s")
        log(f"bench: serving p50 {serving['p50_s']:.2f}s p95 {serving['p95_s']:.2f}s "
            f"p99 {serving['p99_s']:.2f}s, {serving['hemispheres_per_hour']:.0f} hemispheres/hour")
    cpu = None
    oracle_out = None
    if cpu_proc is not None:
        log("bench: waiting for the CPU baseline leg")
        rc = cpu_proc.wait()
        if rc == 0 and os.path.exists(cpu_out):
            with open(cpu_out) as f:
                cpu = json.load(f)
            os.remove(cpu_out)
        if rc == 0 and os.path.exists(cpu_out + ".npz"):
            with np.load(cpu_out + ".npz") as z:
                oracle_out = {k: z[k] for k in z.files}
            os.remove(cpu_out + ".npz")
    accuracy_report = {}
    for form, r in results.items():
        rep = {}
        if oracle_out is not None and np.array_equal(oracle_out["src"], r["outputs"]["check"]["src"]):
            rep["vs_oracle_sample"] = accuracy(r["outputs"]["check"], oracle_out, float(oracle_out["health_h"]))
        if form != "fp64" and "fp64" in results:
            a, b = r["outputs"], results["fp64"]["outputs"]
            if np.array_equal(a["src"], b["src"]):
                rep["vs_fp64_all_sources"] = accuracy(a, b)
            rep["vs_fp64_sample"] = accuracy(r["outputs"]["check"], results["fp64"]["outputs"]["check"],
                                             prep["health_h"])
        accuracy_report[form] = rep
    roof = dict(roofs[dom])
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
        "dtype": "f64" if best["precision"] == "fp64" else "f32+f64ir",
        "data": "synthetic",
        "config": {
            "workload": WORKLOAD_NAMES.get(args.workload, str(args.workload)),
            "cortical_sources": int(prep["order"].size),
            "surface_vertices": int(prep["n_full"]),
            "faces": int(prep["F"].shape[0]),
            "block": int(best["info"]["block"]),
            "streams": int(args.streams),
            "precision": best["precision"],
            "scales": 12,
            "samples_between_scales": 3,
        },
        "roofline": {"kernel": dom, "bound": roof["bound"], "achieved": roof["achieved"], "peak": roof["peak"],
                     "unit": roof["unit"], "frac": roof["frac"], "traffic": roof["traffic"],
                     "avg_launch_ms": roof["avg_launch_ms"]},
        "cpu_baseline": None if cpu is None else {k: cpu[k] for k in ("value", "unit", "cores", "kind", "sample")},
        "forms": {k: {"sources_per_s": v["sources_per_s"], "ms_per_step": v["ms_per_step"], "setup_s": v["setup_s"],
                      "hemisphere_latency_s": v["hemisphere_latency_s"],
                      "end_to_end_sources_per_s": v["end_to_end_sources_per_s"],
                      "distances_to_host_sources_per_s": v["distances_to_host_sources_per_s"],
                      "host_issue_s": v["host_issue_s"], "timed_s": v["seconds"], "flagged": v["flagged"],
                      # the library's own split of the context creation inside setup_s
                      "setup_breakdown_s": dict(zip(("analyse", "assemble", "factor", "upload"),
                                                    v["info"]["setup_seconds"]))}
                  for k, v in results.items()},
        "hemisphere_latency_s": best["hemisphere_latency_s"],
        "accuracy": accuracy_report,
        "serving": serving,
        "other_kernels": {k: {kk: v[kk] for kk in ("achieved", "unit", "frac", "avg_launch_ms", "launches")}
                          for k, v in roofs.items() if k != dom},
        "stage_s_per_block": {k: v / best["profile_blocks"] for k, v in best["stage_s"].items() if k != "wall"},
    }
    if "fp64" in results and best["precision"] != "fp64":
        line["speedup_vs_fp64"] = best["sources_per_s"] / results["fp64"]["sources_per_s"]
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
