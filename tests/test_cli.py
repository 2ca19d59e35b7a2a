"""The command line (fastcw.py:10-160, 435-744): naming and output rules, kwargs, sampling
(CPU), and one end-to-end run on a small FreeSurfer subject (GPU)."""

import os

import numpy as np
import pytest

from fastcorticalwiring_amd import cli, inputs, meshes, outputs


def test_output_rules_and_naming():
    assert cli.resolve_output_kinds("auto", "freesurfer") == ("csv", "mgh")
    assert cli.resolve_output_kinds("auto", "fslr") == ("csv", "gii")
    with pytest.raises(ValueError):
        cli.resolve_output_kinds("gii", "freesurfer")
    meta = dict(legacy_mode=True, subject_dir="/s", subject_id="bert", hemi="lh", surf_type="pial")
    assert cli.resolve_naming(meta, "/s/bert/surf/run1") == ("bert_lh_pial_wiring_costs.csv", "lh.pial.{metric}")
    assert cli.resolve_naming(meta, "/elsewhere") == ("bert_lh_pial_wiring_costs.csv", "bert_lh_pial.{metric}")
    meta2 = dict(output_basename="lh.midthickness", surface_path="/x/lh.midthickness.surf.gii")
    assert cli.resolve_naming(meta2, "/o", suffix="_subset") == ("lh.midthickness_subset_wiring_costs.csv",
                                                                "lh.midthickness_subset.{metric}")
    assert cli.sampling_suffix("stratified", sample_frac=0.05) == "_sample-stratified-frac5p"
    assert cli.sampling_suffix("random", sample_count=200) == "_sample-random-n200"
    for bad in ("", "..", "a/b", "x y"):
        with pytest.raises(ValueError):
            cli.validate_output_label(bad)
    assert cli.metric_names_for_scales([0.002, 0.05])[2:4] == ("radius_0.002", "perimeter_0.002")


def test_engine_kwargs():
    assert cli.parse_engine_kwargs(["a=1", "b=2.5", "c=true", "d=x"]) == {"a": 1, "b": 2.5, "c": True, "d": "x"}
    with pytest.raises(ValueError):
        cli.parse_engine_kwargs(["nokey"])
    kw = cli.engine_kwargs_for("mi355x_potpourri", {})
    assert kw == {"diffusion_length_mm": 0.7, "use_robust": True}
    with pytest.raises(ValueError, match="instead of t_coef"):
        cli.engine_kwargs_for("mi355x_potpourri", {"t_coef": 1.0})
    with pytest.raises(ValueError, match="at least"):
        cli.engine_kwargs_for("mi355x_potpourri", {"diffusion_length_mm": 0.3})
    assert cli.engine_kwargs_for("mi355x_heat", {"t_coef": 2.0}) == {"t_coef": 2.0}


def test_vertex_subsets():
    mask = np.zeros(1000, bool)
    mask[100:900] = True
    V = np.random.default_rng(0).random((1000, 3))
    for method in ("stratified", "random", "fps"):
        s = inputs.select_vertex_subset(V, mask, sample_count=50, sample_method=method, random_state=1)
        assert len(s) == 50 and len(set(s)) == 50 and all(mask[i] for i in s)
    s = inputs.select_vertex_subset(V, mask, sample_frac=0.1, sample_method="stratified", random_state=1)
    assert len(s) == 80
    # stratified: every bin of the cortical order contributes its share
    bins = np.array_split(np.flatnonzero(mask), 8)
    counts = [len(set(s) & set(b.tolist())) for b in bins]
    assert max(counts) - min(counts) <= 1
    assert inputs.select_vertex_subset(V, mask) is None
    # fps: each pick is the cortical vertex farthest from the picks before it
    f = inputs.farthest_point_sampling(V, 5, mask)
    C = np.flatnonzero(mask)
    for i in range(1, 5):
        dmin = np.min(np.stack([np.sum((V[C] - V[f[j]]) ** 2, axis=1) for j in range(i)]), axis=0)
        assert C[int(np.argmax(dmin))] == f[i]


@pytest.mark.gpu
def test_cli_end_to_end_on_a_freesurfer_subject(tmp_path):
    from fastcorticalwiring_amd.analysis import WiringAnalysis

    V, F = meshes.icosphere(4)
    V = V * 50.0
    subj = tmp_path / "subjects" / "bert"
    (subj / "surf").mkdir(parents=True)
    (subj / "label").mkdir()
    inputs.write_freesurfer_surface(str(subj / "surf" / "lh.pial"), V, F)
    V32 = V.astype(np.float32).astype(np.float64)
    cortex = np.flatnonzero(V32[:, 2] < 0.8 * 50.0)
    lines = ["#!ascii label", str(cortex.size)] + [f"{i} 0 0 0 0" for i in cortex]
    (subj / "label" / "lh.cortex.label").write_text("\n".join(lines) + "\n")
    print("cli: running", flush=True)
    written = cli.run_cli([str(tmp_path / "subjects"), "bert", "--hemispheres", "lh", "--output-label", "run1",
                           "--sample", "64", "--sample-kind", "count", "--sample-method", "random"])
    out_dir = subj / "surf" / "run1"
    names = sorted(os.path.basename(p) for p in written)
    assert "bert_lh_pial_sample-random-n64_wiring_costs.csv" in names
    assert "lh.pial_sample-random-n64.radius_0.002.mgh" in names
    npz = outputs.load_sampled_pairs(str(out_dir / "bert_lh_pial_sample-random-n64_wiring_samples.npz"))
    assert int(npz["fastcw_output_schema_version"]) == 3
    # the CLI sampled 64 cortical vertices (unseeded, as the reference); the same analysis run
    # directly on those vertices gives the same arrays
    subset = np.flatnonzero(np.isfinite(npz["msd_weighted"]))
    assert subset.size == 64 and np.all(np.isin(subset, cortex))
    mask = np.zeros(V.shape[0], bool)
    mask[cortex] = True
    an = WiringAnalysis(V32, F, mask, engine_kwargs={"diffusion_length_mm": 0.7, "use_robust": True})
    an.compute_all_wiring_costs(scale=None, vertex_subset=subset.tolist())
    assert np.array_equal(npz["radius_0.002"], an.radius_function[0.002], equal_nan=True)
    assert np.array_equal(npz["msd_weighted"], an.msd_weighted, equal_nan=True)
    # a second run refuses to overwrite
    again = cli.run_cli([str(tmp_path / "subjects"), "bert", "--hemispheres", "lh", "--output-label", "run1",
                         "--sample", "64", "--sample-kind", "count", "--sample-method", "random"])
    assert again == []
