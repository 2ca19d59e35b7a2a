"""The batch runner (run_pipeline.py:20-347 on one GPU): argument and job handling and the device
admission on the CPU, and one run over two small FreeSurfer subjects on the GPU."""

import os
import threading
import time

import numpy as np
import pytest

from conftest import rel_err
from fastcorticalwiring_amd import inputs, meshes, pipeline
from fastcorticalwiring_amd.serving import Admission


def test_dry_run_lists_every_hemisphere(tmp_path):
    lst = tmp_path / "subjects.txt"
    lst.write_text("bert\n\nernie\n")
    logs = tmp_path / "logs"
    res = pipeline.run_pipeline([str(lst), "--subjects-dir", str(tmp_path), "--output-label", "run1", "--dry-run",
                                 "--log-dir", str(logs), "--sample", "0.1", "--engine-kw", "block=128"])
    assert [(s, h, st) for s, h, st, _ in res] == [("bert", "lh", "dry-run"), ("bert", "rh", "dry-run"),
                                                   ("ernie", "lh", "dry-run"), ("ernie", "rh", "dry-run")]
    args = pipeline.build_parser().parse_args([str(lst), "--subjects-dir", "/s", "--output-label", "run1"])
    argv = pipeline.job_argv(args, "bert", "rh")
    assert argv[:4] == ["/s", "bert", "--hemispheres", "rh"] and "--boundary-cap-fraction" in argv
    assert args.scale == list(meshes.TWELVE_SCALES) and args.jobs == 3
    with pytest.raises(SystemExit):
        pipeline.run_pipeline([str(tmp_path / "missing.txt"), "--subjects-dir", "/s", "--output-label", "x"])
    with pytest.raises(ValueError):
        pipeline.run_pipeline([str(lst), "--subjects-dir", "/s", "--output-label", "a/b", "--dry-run"])


def test_admission_bounds_active_jobs_and_memory():
    adm = Admission(max_active=2, hbm_budget_bytes=100)
    peak = {"active": 0, "bytes": 0}
    lock = threading.Lock()

    def job(need):
        adm.admit(need)
        with lock:
            peak["active"] = max(peak["active"], adm.active)
            peak["bytes"] = max(peak["bytes"], adm.in_use)
        time.sleep(0.02)
        adm.release(need)

    ths = [threading.Thread(target=job, args=(n,)) for n in (60, 30, 50, 20, 70, 10, 120)]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    assert peak["active"] <= 2
    assert peak["bytes"] <= 120  # the oversized request ran alone
    assert adm.active == 0 and adm.in_use == 0


@pytest.mark.gpu
def test_pipeline_two_subjects(tmp_path):
    subjects = tmp_path / "subjects"
    for k, sid in enumerate(("bert", "ernie")):
        V, F = meshes.icosphere(4)
        V = V * (50.0 + 5.0 * k)
        d = subjects / sid
        (d / "surf").mkdir(parents=True)
        (d / "label").mkdir()
        inputs.write_freesurfer_surface(str(d / "surf" / "lh.pial"), V, F)
        V32 = V.astype(np.float32)
        cortex = np.flatnonzero(V32[:, 2] < 0.8 * V32[:, 2].max())
        lines = ["#!ascii label", str(cortex.size)] + [f"{i} 0 0 0 0" for i in cortex]
        (d / "label" / "lh.cortex.label").write_text("\n".join(lines) + "\n")
    lst = tmp_path / "list.txt"
    lst.write_text("bert\nernie\n")
    logs = tmp_path / "logs"
    argv = [str(lst), "--subjects-dir", str(subjects), "--hemispheres", "lh", "--output-label", "run1", "-j", "2",
            "--log-dir", str(logs), "--scale", "0.01", "0.05"]
    print("pipeline: two subjects", flush=True)
    res = pipeline.run_pipeline(argv)
    assert [st for _, _, st, _ in res] == ["ok", "ok"], res
    for sid in ("bert", "ernie"):
        out = subjects / sid / "surf" / "run1"
        assert (out / f"{sid}_lh_pial_wiring_costs.csv").exists()
        assert (out / "lh.pial.radius_0.05.mgh").exists()
        text = (logs / f"{sid}.log").read_text()
        assert "START subject=" in text and "status=ok" in text
    again = pipeline.run_pipeline(argv)
    assert [st for _, _, st, _ in again] == ["skipped", "skipped"]


@pytest.mark.gpu
def test_pipeline_files_match_the_oracle_loop(tmp_path):
    """The batch runner's written files for a subject whose surface has a non-manifold fin, with the
    default engine (robust operators on the tufted cover; the diffusion length set to 2.5 mm, about the
    mesh spacing, so t = 6.25 mm squared: distance_engines.py:75-125) against the oracle loop (core_analysis.py:1441-1632 restated) run on the robust oracle's fields:
    the CSV table, the MGH overlays and the NPZ arrays hold the same per-vertex values (float32, as the
    reference writes them) within 1e-5 relative, the same flags and the same sample counts."""
    import csv

    import test_intrinsic as ti
    from fastcorticalwiring_amd import outputs
    from oracle import analysis as oa
    from oracle import intrinsic as oi

    V, F = ti.fin_mesh(level=4, height=3.0, fins=1)
    V = V.astype(np.float32).astype(np.float64)  # as the surface file stores them
    subjects = tmp_path / "subjects"
    d = subjects / "bert"
    (d / "surf").mkdir(parents=True)
    (d / "label").mkdir()
    inputs.write_freesurfer_surface(str(d / "surf" / "lh.pial"), V, F)
    cortex = np.flatnonzero(V[:, 2] < 0.8 * V[:, 2].max())
    lines = ["#!ascii label", str(cortex.size)] + [f"{i} 0 0 0 0" for i in cortex]
    (d / "label" / "lh.cortex.label").write_text("\n".join(lines) + "\n")
    (tmp_path / "list.txt").write_text("bert\n")
    scales = (0.01, 0.05)
    argv = [str(tmp_path / "list.txt"), "--subjects-dir", str(subjects), "--hemispheres", "lh", "--output-label",
            "run1", "-j", "1", "--log-dir", str(tmp_path / "logs"), "--scale", "0.01", "0.05",
            "--boundary-cap-fraction", "0.1", "--diffusion-length-mm", "2.5"]
    print("pipeline: fin subject", flush=True)
    res = pipeline.run_pipeline(argv)
    assert [st for _, _, st, _ in res] == ["ok"], res
    out = d / "surf" / "run1"

    mask = np.zeros(V.shape[0], bool)
    mask[cortex] = True
    mesh = oa.OracleMesh(V, F, mask)
    R = oi.RobustHeatOracle(mesh.V, mesh.F, t=2.5 ** 2)
    assert R.cover
    print("oracle loop:", mesh.n, "sources", flush=True)
    ref = oa.wiring_costs(mesh, R.compute_distance_batch, scales, boundary_cap_fraction=0.1)

    def close(a, b, what):
        a = np.asarray(a, np.float64)
        b = np.asarray(b, np.float64)
        assert np.array_equal(np.isnan(a), np.isnan(b)), what
        fin = np.isfinite(b)
        assert np.array_equal(np.isfinite(a), fin), what
        assert np.max(rel_err(a[fin], b[fin]), initial=0.0) < 1e-5, what

    npz = outputs.load_sampled_pairs(str(out / "bert_lh_pial_wiring_samples.npz"))
    assert np.array_equal(npz["field_health_flag"], ref["field_flag"])
    assert np.mean(ref["field_flag"][cortex] == 0) > 0.9
    for key in ("msd_unweighted", "msd_weighted", "dist_to_boundary"):
        close(npz[key], ref[key], key)
    for i, sc in enumerate(scales):
        close(npz[f"radius_{sc:g}"], ref["radius"][i], f"radius {sc}")
        close(npz[f"perimeter_{sc:g}"], ref["perimeter"][i], f"perimeter {sc}")
        # the overlay stores non-finite values as 0 (io_utils.py:389-410)
        mgh = outputs.read_mgh(str(out / f"lh.pial.radius_{sc:g}.mgh"))[0]
        r = ref["radius"][i].copy()
        r[~np.isfinite(r)] = 0.0
        close(np.ravel(mgh), r, f"mgh radius {sc}")
    assert np.array_equal(npz["sample_indptr"], ref["sample_indptr"])
    close(npz["sample_radii_flat"], ref["sample_radii"], "sample radii")
    close(npz["sample_areas_flat"], ref["sample_areas"], "sample areas")
    with open(out / "bert_lh_pial_wiring_costs.csv") as fh:
        rows = list(csv.DictReader(fh))
    assert len(rows) == V.shape[0]
    close([float(r["msd_weighted"]) for r in rows], ref["msd_weighted"], "csv msd_weighted")
    close([float(r["radius_0.05"]) for r in rows], ref["radius"][1], "csv radius 0.05")
