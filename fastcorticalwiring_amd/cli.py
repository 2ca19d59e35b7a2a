"""Command line: load a surface, run the wiring analysis on the MI355X, write the results.

The reference's ``fastcw.py`` CLI (fastcw.py:10-160, 435-744) with its arguments, naming and
output rules, driving :class:`fastcorticalwiring_amd.analysis.WiringAnalysis`:

    python -m fastcorticalwiring_amd --surface lh.pial --mask lh.cortex.label --output-label run1
    python -m fastcorticalwiring_amd SUBJECTS_DIR bert --output-label run1 --hemispheres lh rh

Engines: ``mi355x_potpourri`` (default; the potpourri engine's contract -- robust operators,
t = diffusion_length_mm^2, default 0.7 mm) and ``mi355x_heat`` (plain operators, t_coef or
diffusion_length_mm). The reference's ``potpourri`` is accepted as a name for the first.
"""

from __future__ import annotations

import argparse
import os
import re
import sys
import warnings

import numpy as np

from . import inputs, mesh_prep, outputs

ENGINES = ("mi355x_potpourri", "mi355x_heat")
ENGINE_ALIASES = {"potpourri": "mi355x_potpourri"}


def validate_output_label(label):
    """A simple directory name (fastcw.py:15-23)."""
    text = str(label or "").strip()
    if not text:
        raise ValueError("--output-label is required and cannot be empty.")
    if text in {".", ".."} or os.path.basename(text) != text or os.path.isabs(text):
        raise ValueError("--output-label must be a simple directory name, not a path.")
    if not re.match(r"^[A-Za-z0-9._-]+$", text):
        raise ValueError("--output-label may only contain letters, numbers, dots, underscores, and hyphens.")
    return text


def metric_names_for_scales(scales):
    names = ["msd_unweighted", "msd_weighted"]
    for s in scales:
        tok = mesh_prep.scale_token(s)
        names += [f"radius_{tok}", f"perimeter_{tok}"]
    return tuple(names + ["dist_to_boundary"])


def resolve_output_kinds(output_format, standard):
    """auto: csv + mgh (FreeSurfer) or csv + gii (fsLR); mgh / gii only for their standard."""
    fmt, std = str(output_format).lower(), str(standard).lower()
    if fmt == "auto":
        return ("csv", "mgh") if std == "freesurfer" else ("csv", "gii")
    if fmt == "csv":
        return ("csv",)
    if fmt == "mgh":
        if std != "freesurfer":
            raise ValueError("--output-format mgh requires --standard freesurfer.")
        return ("csv", "mgh")
    if fmt == "gii":
        if std != "fslr":
            raise ValueError("--output-format gii requires --standard fslr.")
        return ("csv", "gii")
    raise ValueError(f"Unsupported output format: {output_format}")


def resolve_naming(metadata, output_dir, output_basename=None, suffix=""):
    """(csv filename, scalar stem with {metric}) as fastcw.py:54-75."""
    out_abs = os.path.abspath(output_dir)
    sd, sid, hemi, st = (metadata.get(k) for k in ("subject_dir", "subject_id", "hemi", "surf_type"))
    if metadata.get("legacy_mode") and sd and sid and hemi and st:
        surf_dir = os.path.abspath(os.path.join(sd, sid, "surf"))
        csv = f"{sid}_{hemi}_{st}{suffix}_wiring_costs.csv"
        if out_abs == surf_dir or os.path.dirname(out_abs) == surf_dir:
            return csv, f"{hemi}.{st}{suffix}.{{metric}}"
        return csv, f"{sid}_{hemi}_{st}{suffix}.{{metric}}"
    base = output_basename or metadata.get("output_basename") or inputs.infer_output_basename(metadata.get("surface_path"))
    return f"{base}{suffix}_wiring_costs.csv", f"{base}{suffix}.{{metric}}"


def sampling_suffix(method, sample_frac=None, sample_count=None):
    if sample_frac is not None:
        return f"_sample-{method}-frac{format(float(sample_frac) * 100.0, 'g').replace('.', 'p')}p"
    if sample_count is not None:
        return f"_sample-{method}-n{int(sample_count)}"
    return ""


def expected_output_files(output_dir, kinds, csv_filename, scalar_stem, metric_names):
    out = []
    if "csv" in kinds:
        out.append(os.path.join(output_dir, csv_filename))
    for kind, ext in (("mgh", ".mgh"), ("gii", ".shape.gii")):
        if kind in kinds:
            out += [os.path.join(output_dir, scalar_stem.format(metric=m) + ext) for m in metric_names]
    out.append(os.path.join(output_dir, outputs.sample_npz_filename(csv_filename)))
    return out


def parse_engine_kwargs(values):
    """Repeated key=value strings; true/false, int and float values coerced (fastcw.py:435-461)."""
    out = {}
    for item in values or []:
        if "=" not in item:
            raise ValueError(f"Invalid --engine-kw '{item}'. Expected key=value.")
        key, raw = item.split("=", 1)
        key = key.strip()
        if not key:
            raise ValueError(f"Invalid --engine-kw '{item}'. Empty key.")
        text = raw.strip()
        low = text.lower()
        if low in {"true", "false"}:
            out[key] = low == "true"
            continue
        try:
            out[key] = float(text) if ("." in text or "e" in low) else int(text)
        except ValueError:
            out[key] = text
    return out


def _parse_optional_float(raw):
    text = str(raw).strip().lower()
    return None if text in {"none", "null", "off", "false", "disable", "disabled"} else float(raw)


def engine_kwargs_for(engine, kw):
    """WiringAnalysis kwargs for an engine name (the potpourri contract checked here)."""
    from . import distance_engines as de

    kw = dict(kw)
    if engine == "mi355x_potpourri":
        kw.pop("allow_eigen_fallback", None)
        if "use_robust" in kw and not bool(kw.pop("use_robust")):
            raise ValueError("use_robust=False is not supported; robust mode is always used.")
        if "t_coef" in kw:
            raise ValueError("Set diffusion_length_mm instead of t_coef.")
        L = float(kw.get("diffusion_length_mm", de.DEFAULT_DIFFUSION_LENGTH_MM))
        if not np.isfinite(L) or L < de.MIN_DIFFUSION_LENGTH_MM:
            raise ValueError(f"diffusion_length_mm must be finite and at least {de.MIN_DIFFUSION_LENGTH_MM} mm.")
        if L < de.RECOMMENDED_MIN_DIFFUSION_LENGTH_MM:
            warnings.warn(f"diffusion_length_mm={L} is below the validated {de.RECOMMENDED_MIN_DIFFUSION_LENGTH_MM} mm.",
                          RuntimeWarning)
        kw["diffusion_length_mm"] = L
        kw["use_robust"] = True
    return kw


def run_single_surface(*, standard, surface_path, mask_path, output_dir, output_basename, output_label, subject_dir,
                       subject_id, hemi, surf_type, custom_label, no_mask, output_format, engine, engine_kwargs,
                       scale, eps, overwrite, n_samples_between_scales, boundary_cap_fraction, verbose_timing=False,
                       sample=None, sample_kind="frac", sample_method=None, vertex_list=None, admit=None):
    """One surface from loading to written files (fastcw.py:159-348); returns the paths written
    ([] when outputs exist and overwrite is off). admit(n_cortical_vertices), when given, is called
    before the device work starts and returns the function to call when it has ended (the batch
    runner's admission)."""
    from .analysis import WiringAnalysis

    V, F, mask, meta = inputs.load_surface_and_mask(standard=standard, surface_path=surface_path, mask_path=mask_path,
                                                    subject_dir=subject_dir, subject_id=subject_id, hemi=hemi,
                                                    surf_type=surf_type, custom_label=custom_label, no_mask=no_mask)
    if output_dir is None:
        if meta.get("legacy_mode") and subject_dir and subject_id:
            if output_label is None:
                raise ValueError("--output-label is required when writing positional FreeSurfer outputs.")
            output_dir = os.path.join(subject_dir, subject_id, "surf", validate_output_label(output_label))
        else:
            base = os.path.dirname(os.path.abspath(meta["surface_path"])) or os.getcwd()
            output_dir = os.path.join(base, validate_output_label(output_label)) if output_label is not None else base
    scales = mesh_prep.normalize_scales(scale)
    sample_frac = sample_count = None
    if sample is not None:
        kind = str(sample_kind or "frac").lower()
        if kind == "frac":
            sample_frac = float(sample)
            if not (0.0 < sample_frac <= 1.0):
                raise ValueError("sample_frac must satisfy 0 < sample_frac <= 1.")
        elif kind == "count":
            if abs(float(sample) - round(float(sample))) > 1e-9:
                raise ValueError("sample-kind=count requires an integer-valued --sample.")
            sample_count = int(round(float(sample)))
            if sample_count <= 0:
                raise ValueError("sample_count must be a positive integer.")
        else:
            raise ValueError(f"Unknown sample_kind: {sample_kind}")
    method = str(sample_method or "stratified").lower()
    if method not in {"stratified", "random", "fps"}:
        raise ValueError(f"Unknown sample_method: {sample_method}")
    suffix = "_subset" if vertex_list is not None else sampling_suffix(method, sample_frac, sample_count)
    kinds = resolve_output_kinds(output_format, standard)
    csv_filename, stem = resolve_naming(meta, output_dir, output_basename, suffix)
    existing = [f for f in expected_output_files(output_dir, kinds, csv_filename, stem, metric_names_for_scales(scales))
                if os.path.exists(f)]
    if existing and not overwrite:
        print("ERROR: Output files already exist:")
        for p in existing:
            print(f"  - {p}")
        print("Use --overwrite or a different --output-dir.")
        return []
    subset = None
    if vertex_list:
        if not os.path.exists(vertex_list):
            raise FileNotFoundError(f"--vertex-list file not found: {vertex_list}")
        subset = np.atleast_1d(np.loadtxt(vertex_list, dtype=int)).astype(np.int64).tolist()
    elif sample_frac is not None or sample_count is not None:
        subset = inputs.select_vertex_subset(V, mask, sample_frac=sample_frac, sample_count=sample_count,
                                             sample_method=method)
    release = admit(int(np.count_nonzero(mask))) if admit is not None else None
    try:
        an = WiringAnalysis(V, F, mask, engine_kwargs=engine_kwargs, eps=eps, metadata=meta)
        an.compute_all_wiring_costs(scale=scales, vertex_subset=subset,
                                    n_samples_between_scales=n_samples_between_scales,
                                    boundary_cap_fraction=boundary_cap_fraction, verbose=verbose_timing)
        an.close()  # device memory is given back before the admission slot
    finally:
        if release is not None:
            release()
    return outputs.save_analysis_outputs(an, output_dir, kinds, csv_filename, stem)


def build_parser():
    from .analysis import WiringAnalysis

    p = argparse.ArgumentParser(description="Cortical wiring costs on one MI355X (heat-method geodesics)")
    p.add_argument("subject_dir", nargs="?", help="FreeSurfer subjects directory (positional mode)")
    p.add_argument("subject_id", nargs="?", help="Subject ID (positional mode)")
    p.add_argument("--standard", choices=["freesurfer", "fslr"], default="freesurfer")
    p.add_argument("--surface", default=None)
    p.add_argument("--mask", default=None)
    p.add_argument("--output-format", choices=["auto", "csv", "mgh", "gii"], default="auto")
    p.add_argument("--output-basename", default=None)
    p.add_argument("--no-mask", action="store_true", default=False)
    p.add_argument("--engine", default="mi355x_potpourri",
                   help="mi355x_potpourri (robust, default; 'potpourri' is accepted) or mi355x_heat")
    p.add_argument("--engine-kw", action="append", default=[], help="key=value (repeatable)")
    p.add_argument("--diffusion-length-mm", type=float, default=None)
    p.add_argument("--use-robust", action="store_true", default=False, help=argparse.SUPPRESS)
    p.add_argument("--allow-eigen-fallback", action="store_true", default=False, help=argparse.SUPPRESS)
    p.add_argument("--output-dir", default=None)
    p.add_argument("--output-label", default=None)
    p.add_argument("--hemispheres", nargs="+", default=["lh", "rh"])
    p.add_argument("--hemi", default="lh")
    p.add_argument("--surf-type", default="pial")
    p.add_argument("--custom-label", default=None)
    p.add_argument("--overwrite", action="store_true", default=False)
    p.add_argument("--scale", nargs="+", type=float, default=list(WiringAnalysis.DEFAULT_SCALES))
    p.add_argument("--eps", type=float, default=1e-6)
    p.add_argument("--n-samples-between-scales", type=int, default=3)
    p.add_argument("--boundary-cap-fraction", type=_parse_optional_float, default=None)
    p.add_argument("--verbose-timing", action="store_true", default=False)
    p.add_argument("--sample", type=float, default=None)
    p.add_argument("--sample-kind", choices=["frac", "count"], default="frac")
    p.add_argument("--sample-method", choices=["stratified", "random", "fps"], default=None)
    p.add_argument("--vertex-list", default=None)
    return p


def run_cli(argv=None):
    args = build_parser().parse_args(argv)
    engine = ENGINE_ALIASES.get(str(args.engine).lower(), str(args.engine).lower())
    if engine not in ENGINES:
        raise ValueError(f"--engine must be one of {', '.join(ENGINES)} (or potpourri)")
    kw = parse_engine_kwargs(args.engine_kw)
    if args.diffusion_length_mm is not None:
        if "diffusion_length_mm" in kw and float(kw["diffusion_length_mm"]) != float(args.diffusion_length_mm):
            raise ValueError("Conflicting diffusion lengths from --diffusion-length-mm and --engine-kw.")
        kw["diffusion_length_mm"] = float(args.diffusion_length_mm)
    kw = engine_kwargs_for(engine, kw)
    common = dict(output_format=args.output_format, engine=engine, engine_kwargs=kw, scale=args.scale, eps=args.eps,
                  overwrite=args.overwrite, n_samples_between_scales=args.n_samples_between_scales,
                  boundary_cap_fraction=args.boundary_cap_fraction, verbose_timing=args.verbose_timing,
                  sample=args.sample, sample_kind=args.sample_kind, sample_method=args.sample_method,
                  vertex_list=args.vertex_list, custom_label=args.custom_label, no_mask=args.no_mask)
    written = []
    if args.surface is not None:
        written = run_single_surface(standard=args.standard, surface_path=args.surface, mask_path=args.mask,
                                     output_dir=args.output_dir, output_basename=args.output_basename,
                                     output_label=args.output_label, subject_dir=args.subject_dir,
                                     subject_id=args.subject_id, hemi=args.hemi, surf_type=args.surf_type, **common)
    elif args.subject_dir is not None and args.subject_id is not None:
        if args.standard != "freesurfer":
            raise ValueError("Positional FreeSurfer mode only supports --standard freesurfer.")
        if args.output_dir is None:
            validate_output_label(args.output_label)
        for hemi in args.hemispheres:
            written += run_single_surface(standard="freesurfer", surface_path=None, mask_path=args.mask,
                                          output_dir=args.output_dir, output_basename=args.output_basename,
                                          output_label=args.output_label, subject_dir=args.subject_dir,
                                          subject_id=args.subject_id, hemi=hemi, surf_type=args.surf_type, **common)
    else:
        raise ValueError("Specify either positional arguments <subject_dir> <subject_id> "
                         "or explicit --surface with --standard.")
    if written:
        print("Wrote outputs:")
        for p in written:
            print(f"  - {p}")
    print("Analysis complete!")
    return written


def main():  # pragma: no cover - console entry
    try:
        run_cli()
    except (ValueError, FileNotFoundError) as exc:
        print(f"error: {exc}", file=sys.stderr)
        sys.exit(2)


if __name__ == "__main__":  # pragma: no cover
    main()
