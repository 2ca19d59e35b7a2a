"""Batch runner: many FreeSurfer subjects through the wiring analysis on one MI355X.

The reference's run_pipeline.py (run_pipeline.py:20-347) starts one fastcw.py process per subject,
pinned to a NUMA-balanced core, in a pool of ``-j`` workers. Here every (subject, hemisphere) is
a job in this process and all of them share the GPU: ``-j`` hemispheres are on the device at once,
and a new one is admitted only while the active ones' estimated device memory plus its own fits
the HBM budget (the serving loop's admission, :class:`serving.Admission`). Each job runs in a worker
thread with its own context and streams (ctypes releases the GIL during device calls), so their
kernels overlap on the device while the next surfaces are read and analysed on the host.

Arguments, the per-subject log (START / CMD / FINISH lines in ``<log-dir>/<subject>.log``),
``--dry-run`` and ``--overwrite`` follow the reference; its interpreter and script paths and the
NUMA pinning have no counterpart. Usage::

    python -m fastcorticalwiring_amd.pipeline subjects.txt --subjects-dir $SUBJECTS_DIR --output-label run1 -j 3
"""

from __future__ import annotations

import argparse
import datetime
import logging
import os
import shlex
import sys
import threading
from concurrent.futures import ThreadPoolExecutor

from . import cli
from .serving import Admission, estimate_device_bytes

__all__ = ["build_parser", "job_argv", "run_pipeline", "main"]


def _utc_now_iso():
    return datetime.datetime.now(datetime.timezone.utc).isoformat(timespec="seconds")


def build_parser():
    """The reference's arguments (run_pipeline.py:196-302) plus the device admission controls."""
    from .meshes import TWELVE_SCALES

    p = argparse.ArgumentParser(description="Cortical wiring costs for many subjects on one MI355X")
    p.add_argument("subject_list", help="Path to text file containing subject IDs")
    p.add_argument("--subjects-dir", default=os.environ.get("SUBJECTS_DIR"), help="Path to FreeSurfer SUBJECTS_DIR")
    p.add_argument("--hemispheres", nargs="+", default=["lh", "rh"])
    p.add_argument("--surf-type", default="pial")
    p.add_argument("--custom-label", default=None)
    p.add_argument("--mask", default=None)
    p.add_argument("--no-mask", action="store_true", default=False)
    p.add_argument("--output-label", required=True)
    p.add_argument("--output-dir", default=None)
    p.add_argument("--output-format", choices=["auto", "csv", "mgh", "gii"], default="auto")
    p.add_argument("--engine", default="mi355x_potpourri")
    p.add_argument("--engine-kw", action="append", default=[], help="key=value (repeatable)")
    p.add_argument("--diffusion-length-mm", type=float, default=None)
    p.add_argument("--sample", type=float, default=None)
    p.add_argument("--sample-kind", choices=["frac", "count"], default="frac")
    p.add_argument("--sample-method", choices=["stratified", "random", "fps"], default=None)
    p.add_argument("--vertex-list", default=None)
    p.add_argument("--scale", nargs="+", type=float, default=list(TWELVE_SCALES))  # run_pipeline.py:248-267
    p.add_argument("--eps", type=float, default=1e-6)
    p.add_argument("--n-samples-between-scales", type=int, default=3)
    p.add_argument("--boundary-cap-fraction", type=cli._parse_optional_float, default=None)
    p.add_argument("--verbose-timing", action="store_true", default=False)
    p.add_argument("-j", "--jobs", type=int, default=3, help="Hemispheres on the device at once")
    p.add_argument("--hbm-budget-gb", type=float, default=200.0,
                   help="Device memory the active hemispheres' contexts may take together (estimated)")
    p.add_argument("--overwrite", action="store_true")
    p.add_argument("--dry-run", action="store_true", help="Print the jobs without running them")
    p.add_argument("--log-dir", default="logs_fastcw", help="Directory for per-subject logs")
    # accepted for command-line compatibility with the reference, without effect
    p.add_argument("--use-robust", action="store_true", default=False, help=argparse.SUPPRESS)
    p.add_argument("--batch-size", type=int, default=None, help=argparse.SUPPRESS)
    p.add_argument("--area-tol", type=float, default=None, help=argparse.SUPPRESS)
    return p


def job_argv(args, subject, hemi):
    """The single-surface command line equivalent to one job (logged, and printed by --dry-run)."""
    argv = [args.subjects_dir, subject, "--hemispheres", hemi, "--surf-type", args.surf_type,
            "--output-label", args.output_label, "--output-format", args.output_format, "--engine", args.engine]
    if args.custom_label:
        argv += ["--custom-label", args.custom_label]
    if args.mask:
        argv += ["--mask", args.mask]
    if args.no_mask:
        argv.append("--no-mask")
    if args.overwrite:
        argv.append("--overwrite")
    if args.output_dir:
        argv += ["--output-dir", args.output_dir]
    if args.diffusion_length_mm is not None:
        argv += ["--diffusion-length-mm", str(args.diffusion_length_mm)]
    if args.sample is not None:
        argv += ["--sample", str(args.sample), "--sample-kind", args.sample_kind]
    if args.sample_method is not None:
        argv += ["--sample-method", args.sample_method]
    if args.vertex_list:
        argv += ["--vertex-list", args.vertex_list]
    if args.verbose_timing:
        argv.append("--verbose-timing")
    argv += ["--scale", *[str(s) for s in args.scale], "--eps", str(args.eps),
             "--n-samples-between-scales", str(args.n_samples_between_scales),
             "--boundary-cap-fraction", "none" if args.boundary_cap_fraction is None else str(args.boundary_cap_fraction)]
    for item in args.engine_kw:
        argv += ["--engine-kw", item]
    return argv


def _engine_kwargs(args):
    engine = cli.ENGINE_ALIASES.get(str(args.engine).lower(), str(args.engine).lower())
    if engine not in cli.ENGINES:
        raise ValueError(f"--engine must be one of {', '.join(cli.ENGINES)} (or potpourri)")
    kw = cli.parse_engine_kwargs(args.engine_kw)
    if args.diffusion_length_mm is not None:
        if "diffusion_length_mm" in kw and float(kw["diffusion_length_mm"]) != float(args.diffusion_length_mm):
            raise ValueError("Conflicting diffusion lengths from --diffusion-length-mm and --engine-kw.")
        kw["diffusion_length_mm"] = float(args.diffusion_length_mm)
    kw = cli.engine_kwargs_for(engine, kw)
    # the active hemispheres share the device's hardware queues
    kw.setdefault("streams", max(2, 8 // max(1, args.jobs)))
    return engine, kw


def run_pipeline(argv=None):
    """Run every (subject, hemisphere) of the list; returns [(subject, hemi, status, paths written)]
    with status "ok", "skipped" (outputs exist, no --overwrite), "failed: <reason>" or "dry-run"."""
    args = build_parser().parse_args(argv)
    logging.basicConfig(level=logging.INFO, format="%(asctime)s [%(levelname)s] %(message)s",
                        handlers=[logging.StreamHandler(sys.stdout)])
    log = logging.getLogger("fastcorticalwiring_amd.pipeline")
    if args.use_robust:
        print("NOTE: --use-robust is deprecated and ignored; the default engine always uses robust operators.")
    if args.batch_size is not None:
        print("NOTE: --batch-size is deprecated and ignored; sources are solved in device blocks.")
    if args.area_tol is not None:
        print("NOTE: --area-tol is deprecated and ignored; radii are solved exactly.")
    if not args.subjects_dir:
        raise SystemExit("SUBJECTS_DIR is not set and was not provided via --subjects-dir.")
    if not os.path.exists(args.subject_list):
        raise SystemExit(f"Subject list '{args.subject_list}' not found.")
    cli.validate_output_label(args.output_label)
    engine, kw = _engine_kwargs(args)
    with open(args.subject_list, "r", encoding="utf-8") as f:
        subjects = [line.strip() for line in f if line.strip()]
    os.makedirs(args.log_dir, exist_ok=True)
    jobs = [(s, h) for s in subjects for h in args.hemispheres]
    mode = "DRY RUN" if args.dry_run else "EXECUTION"
    log.info(f"Initializing {mode} for {len(subjects)} subjects ({len(jobs)} hemispheres), "
             f"{args.jobs} on the device at once.")
    admission = Admission(max(1, args.jobs), int(args.hbm_budget_gb * (1 << 30)))
    block = int(kw.get("block", 256))
    streams = int(kw["streams"])
    log_lock = threading.Lock()

    def write_log(subject, line):
        with log_lock, open(os.path.join(args.log_dir, f"{subject}.log"), "a", encoding="utf-8") as f:
            f.write(f"[{_utc_now_iso()}] {line}\n")

    def run_job(subject, hemi):
        cmd = "python -m fastcorticalwiring_amd " + " ".join(shlex.quote(a) for a in job_argv(args, subject, hemi))
        if args.dry_run:
            log.info(f"[DRY RUN - {subject} {hemi}] {cmd}")
            return subject, hemi, "dry-run", []
        write_log(subject, f"START subject={subject} hemi={hemi}")
        write_log(subject, f"CMD {cmd}")

        def admit(n_cortex):
            need = estimate_device_bytes(n_cortex, 2 * n_cortex, block, streams)
            admission.admit(need)
            return lambda: admission.release(need)

        try:
            written = cli.run_single_surface(
                standard="freesurfer", surface_path=None, mask_path=args.mask, output_dir=args.output_dir,
                output_basename=None, output_label=args.output_label, subject_dir=args.subjects_dir,
                subject_id=subject, hemi=hemi, surf_type=args.surf_type, custom_label=args.custom_label,
                no_mask=args.no_mask, output_format=args.output_format, engine=engine, engine_kwargs=dict(kw),
                scale=args.scale, eps=args.eps, overwrite=args.overwrite,
                n_samples_between_scales=args.n_samples_between_scales,
                boundary_cap_fraction=args.boundary_cap_fraction, verbose_timing=args.verbose_timing,
                sample=args.sample, sample_kind=args.sample_kind, sample_method=args.sample_method,
                vertex_list=args.vertex_list, admit=admit)
        except Exception as exc:  # reported per job; the other jobs go on
            write_log(subject, f"FINISH subject={subject} hemi={hemi} status=failed error={type(exc).__name__}: {exc}")
            log.error(f"FAILED: {subject} {hemi}: {type(exc).__name__}: {exc}")
            return subject, hemi, f"failed: {type(exc).__name__}: {exc}", []
        status = "ok" if written else "skipped"
        write_log(subject, f"FINISH subject={subject} hemi={hemi} status={status}")
        log.info(f"FINISHED: {subject} {hemi} ({status})")
        return subject, hemi, status, written

    # host work (reading, submesh, analysis, factorisation) of the next jobs overlaps the device
    # work of the admitted ones: a few more workers than device slots
    with ThreadPoolExecutor(max_workers=max(1, args.jobs) + 2) as ex:
        results = list(ex.map(lambda sh: run_job(*sh), jobs))
    log.info("Pipeline complete.")
    return results


def main():  # pragma: no cover - console entry
    run_pipeline()


if __name__ == "__main__":  # pragma: no cover
    main()
