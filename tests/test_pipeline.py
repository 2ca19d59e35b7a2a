"""The batch runner (run_pipeline.py:20-347 on one GPU): argument and job handling and the device
admission on the CPU, and one run over two small FreeSurfer subjects on the GPU."""

import os
import threading
import time

import numpy as np
import pytest

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
