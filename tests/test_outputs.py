"""On-disk result writers (io_utils.py:300-493, fastcw.py:101-156): CSV columns and value
format, MGH header layout and geometry, GIFTI array encoding, NPZ schema keys."""

import gzip
import json
import math
import os
import struct
import types

import numpy as np
import pytest

from fastcorticalwiring_amd import outputs as io


def _fake_analysis(n=7):
    rng = np.random.default_rng(3)
    mask = np.ones(n, bool)
    mask[[1, 5]] = False
    msd_u = rng.random(n).astype(np.float32)
    msd_u[~mask] = np.nan
    msd_w = (msd_u * 1.5).astype(np.float32)
    scales = (0.002, 0.05)
    rad = {float(s): (rng.random(n) * 10).astype(np.float32) for s in scales}
    per = {float(s): (rng.random(n) * 50).astype(np.float32) for s in scales}
    for s in scales:
        rad[s][~mask] = np.nan
        per[s][~mask] = np.nan
    counts = np.where(mask, 3, 0)
    indptr = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    a = types.SimpleNamespace(
        cortex_mask_full=mask, n_vertices_full=n, msd_unweighted=msd_u, msd_weighted=msd_w,
        dist_to_boundary=np.full(n, np.inf, np.float32), radius_function=rad, perimeter_function=per,
        active_scales=scales, sample_indptr=indptr, sample_radii_flat=np.arange(indptr[-1], dtype=np.float32),
        sample_areas_flat=np.arange(indptr[-1], dtype=np.float32) * 2, n_samples_between_scales=3,
        boundary_cap_fraction=None, sub_to_orig=np.flatnonzero(mask).astype(np.int32),
        field_flag=np.zeros(n, np.uint8), metadata={})

    def metric_arrays():
        out = {"msd_unweighted": msd_u, "msd_weighted": msd_w, "dist_to_boundary": a.dist_to_boundary}
        for s in scales:
            out[f"radius_{format(s, 'g')}"] = rad[s]
            out[f"perimeter_{format(s, 'g')}"] = per[s]
        return out

    a.get_metric_arrays = metric_arrays
    a.provenance = lambda: {"engine": "mi355x_heat", "t_coef": 1.0}
    return a


def test_csv_columns_and_values(tmp_path):
    a = _fake_analysis()
    p = io.save_results_csv(tmp_path, "x_wiring_costs.csv", a.cortex_mask_full, a.msd_unweighted, a.msd_weighted,
                            a.radius_function, a.perimeter_function)
    lines = open(p).read().splitlines()
    assert lines[0] == ("vertex_id,is_cortex,msd_unweighted,msd_weighted,radius_0.002,perimeter_0.002,"
                        "radius_0.05,perimeter_0.05")
    assert len(lines) == 1 + 7
    row1 = lines[2].split(",")  # vertex 1 is not cortical: NaN metrics written as Python's 'nan'
    assert row1[:2] == ["1", "0"] and row1[2] == "nan"
    row0 = lines[1].split(",")
    assert row0[2] == str(float(a.msd_unweighted[0])) and row0[4] == str(float(a.radius_function[0.002][0]))
    with pytest.raises(ValueError):
        io.save_results_csv(tmp_path, "y.csv", a.cortex_mask_full, a.msd_unweighted[:3], a.msd_weighted,
                            a.radius_function, a.perimeter_function)
    with pytest.raises(ValueError):
        io.save_results_csv(tmp_path, "y.csv", a.cortex_mask_full, a.msd_unweighted, a.msd_weighted,
                            a.radius_function, {0.002: a.perimeter_function[0.002]})


def test_mgh_layout_and_default_geometry(tmp_path):
    a = _fake_analysis()
    metrics = {"msd_unweighted": a.msd_unweighted, "allnan": np.full(7, np.nan)}
    paths = io.save_results_mgh(tmp_path, "lh.{metric}.mgh", metrics, 7)
    assert [os.path.basename(p) for p in paths] == ["lh.msd_unweighted.mgh"]  # all-NaN metric skipped
    raw = open(paths[0], "rb").read()
    assert len(raw) == 284 + 4 * 7 + 16
    assert struct.unpack(">7i", raw[:28]) == (1, 7, 1, 1, 1, 3, 0)
    vals, hdr = io.read_mgh(paths[0])
    expect = a.msd_unweighted.copy()
    expect[~np.isfinite(expect)] = 0
    assert np.array_equal(vals, expect)
    assert hdr["good_ras"] == 1 and hdr["delta"] == (1.0, 1.0, 1.0)
    # columns of the direction cosines are the default affine's voxel axes
    assert np.array_equal(hdr["mdc"], io.DEFAULT_MGH_AFFINE[:3, :3])
    # vox2ras rebuilt from the header (M = Mdc diag(delta), P0 = c_ras - M dims/2) is the affine
    M = hdr["mdc"] * np.array(hdr["delta"])
    P0 = np.array(hdr["c_ras"]) - M.dot(np.array([7, 1, 1]) / 2.0)
    assert np.allclose(P0, 0.0) and np.allclose(hdr["c_ras"], (-3.5, 0.5, -0.5))


def test_mgh_template_geometry_is_kept(tmp_path):
    th = 0.3
    A = np.eye(4)
    A[:3, :3] = np.array([[math.cos(th), -math.sin(th), 0], [math.sin(th), math.cos(th), 0], [0, 0, 1]]) * [
        0.5, 2.0, 1.0]
    A[:3, 3] = [10.0, -4.0, 2.5]
    tpl = tmp_path / "tpl.mgz"
    with gzip.open(tpl, "wb") as f:
        f.write(io._mgh_bytes(np.zeros(11, np.float32), A))
    assert np.allclose(io.mgh_template_affine(tpl), A, atol=1e-6)
    paths = io.save_results_mgh(tmp_path, "{metric}.mgh", {"m": np.arange(5.0)}, 5, template_mgz_path=str(tpl))
    _, hdr = io.read_mgh(paths[0])
    M = hdr["mdc"] * np.array(hdr["delta"])
    assert np.allclose(M, A[:3, :3], atol=1e-6)
    assert np.allclose(np.array(hdr["c_ras"]) - M.dot(np.array([5, 1, 1]) / 2.0), A[:3, 3], atol=1e-5)


def test_gifti_encoding_round_trip(tmp_path):
    a = _fake_analysis()
    paths = io.save_results_gifti(tmp_path, "lh.{metric}.shape.gii", {"radius_0.05": a.radius_function[0.05],
                                                                      "none": np.full(7, np.nan)}, 7)
    assert [os.path.basename(p) for p in paths] == ["lh.radius_0.05.shape.gii"]
    vals, name = io.read_gifti(paths[0])
    assert name == "radius_0.05"
    assert np.array_equal(vals, a.radius_function[0.05], equal_nan=True)
    import xml.etree.ElementTree as ET

    da = ET.parse(paths[0]).getroot().find("DataArray")
    assert da.get("Intent") == "NIFTI_INTENT_SHAPE" and da.get("Dim0") == "7" and da.get("Dimensionality") == "1"
    with pytest.raises(ValueError):
        io.save_results_gifti(tmp_path, "{metric}.gii", {"bad": np.zeros(3)}, 7)


def test_npz_schema_and_sample_loader(tmp_path):
    a = _fake_analysis()
    p = io.save_analysis_npz(tmp_path, "x_wiring_samples.npz", a)
    d = io.load_sampled_pairs(p)
    assert int(d["fastcw_output_schema_version"]) == 3
    assert str(d["sample_storage_layout"]) == "csr_flat_v1"
    assert d["sample_indptr"].dtype == np.int64 and d["sample_radii_flat"].dtype == np.float32
    assert np.array_equal(d["sample_scales_solved"], [0.002, 0.05])
    assert math.isnan(float(d["boundary_cap_fraction"])) and int(d["n_samples_between_scales"]) == 3
    assert json.loads(str(d["provenance_json"]))["engine"] == "mi355x_heat"
    assert d["field_health_flag"].dtype == np.uint8 and d["cortex_mask"].dtype == bool
    for k in ("msd_unweighted", "radius_0.002", "perimeter_0.05", "dist_to_boundary", "sub_to_orig"):
        assert k in d
    bad = tmp_path / "bad.npz"
    np.savez(bad, sample_indptr=np.zeros(2))
    with pytest.raises(ValueError):
        io.load_sampled_pairs(bad)


def test_output_step_names(tmp_path):
    assert io.sample_npz_filename("lh_wiring_costs.csv") == "lh_wiring_samples.npz"
    assert io.sample_npz_filename("out.csv") == "out_wiring_samples.npz"
    a = _fake_analysis()
    paths = io.save_analysis_outputs(a, tmp_path, {"csv", "mgh", "gii"}, "lh_wiring_costs.csv", "lh.{metric}")
    names = sorted(os.path.basename(p) for p in paths)
    assert "lh_wiring_costs.csv" in names and "lh_wiring_samples.npz" in names
    assert "lh.radius_0.05.mgh" in names and "lh.msd_weighted.shape.gii" in names
