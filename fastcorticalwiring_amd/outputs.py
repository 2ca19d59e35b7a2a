"""On-disk results: CSV table, FreeSurfer MGH and GIFTI overlays, compressed NPZ with samples.

Same files, names and contents as the reference's writers (io_utils.py:300-493) and its
output step (fastcw.py:101-156), written from the arrays of
:class:`fastcorticalwiring_amd.analysis.WiringAnalysis` (or any object with the reference
analysis's attributes). nibabel is not needed: the MGH and GIFTI files are written directly
from their published layouts.

* MGH (FreeSurfer "MGH format"): a 284-byte big-endian header -- version 1, the four
  dimensions (width, height, depth, frames), data type (3 = float), dof, the "good RAS" flag,
  voxel sizes, the 3 x 3 direction cosines (columns x, y, z) and the centre RAS -- then the
  voxels as big-endian float32 (first dimension fastest), then the four float32 scan
  parameters (TR, flip angle, TE, TI). Without a template the geometry is the reference's
  default affine [[-1, 0, 0, 0], [0, 0, 1, 0], [0, -1, 0, 0], [0, 0, 0, 1]] (io_utils.py:383-386),
  centre = affine[:3, :3] . (dims / 2) + affine[:3, 3].
* GIFTI 1.0: one ``DataArray`` per file, intent NIFTI_INTENT_SHAPE, float32, little endian,
  GZipBase64Binary encoding (zlib stream, base64), the metric name in the array's metadata.
"""

from __future__ import annotations

import base64
import json
import os
import struct
import zlib
from xml.sax.saxutils import escape

import numpy as np

__all__ = ["save_results_csv", "save_results_mgh", "save_results_gifti", "save_analysis_npz",
           "load_sampled_pairs", "read_mgh", "read_gifti", "mgh_template_affine", "sample_npz_filename",
           "save_analysis_outputs"]

MGH_HEADER_BYTES = 284
MGH_FLOAT = 3
DEFAULT_MGH_AFFINE = np.array([[-1.0, 0.0, 0.0, 0.0], [0.0, 0.0, 1.0, 0.0], [0.0, -1.0, 0.0, 0.0],
                               [0.0, 0.0, 0.0, 1.0]])
NIFTI_INTENT_SHAPE = 2005


# --------------------------------------------------------------------------------------------
# CSV (io_utils.py:300-356)
# --------------------------------------------------------------------------------------------
def save_results_csv(output_dir, csv_filename, cortex_mask, msd_unweighted, msd_weighted, radius_function,
                     perimeter_function):
    """Write the per-vertex table: vertex_id, is_cortex, msd_unweighted, msd_weighted, then
    radius_<scale>, perimeter_<scale> per scale (scale formatted with "g"); values as Python
    floats. Raises ValueError on mismatched lengths or scale keys, as io_utils.py:313-334."""
    os.makedirs(output_dir, exist_ok=True)
    path = os.path.join(output_dir, csv_filename)
    n = int(len(cortex_mask))
    if len(msd_unweighted) != n or len(msd_weighted) != n:
        raise ValueError("MSD arrays must match cortex_mask length for CSV output.")
    if not isinstance(radius_function, dict) or not isinstance(perimeter_function, dict):
        raise ValueError("radius_function and perimeter_function must be dicts keyed by scale.")
    if set(radius_function.keys()) != set(perimeter_function.keys()):
        raise ValueError("radius_function and perimeter_function must share the same scale keys.")
    keys = list(radius_function.keys())
    cols = [[str(i) for i in range(n)], [str(int(bool(b))) for b in np.asarray(cortex_mask)],
            _float_strings(msd_unweighted), _float_strings(msd_weighted)]
    header = ["vertex_id", "is_cortex", "msd_unweighted", "msd_weighted"]
    for key in keys:
        r = np.asarray(radius_function[key])
        p = np.asarray(perimeter_function[key])
        if r.shape[0] != n or p.shape[0] != n:
            raise ValueError("All metric arrays must match cortex_mask length for CSV output.")
        tok = format(float(key), "g")
        header += [f"radius_{tok}", f"perimeter_{tok}"]
        cols += [_float_strings(r), _float_strings(p)]
    with open(path, "w", encoding="utf-8") as f:
        f.write(",".join(header) + "\n")
        f.writelines(",".join(row) + "\n" for row in zip(*cols))
    return path


def _float_strings(a):
    # str(float(x)) of each element, as the reference writes them (float32 widened exactly)
    return [str(x) for x in np.asarray(a).astype(np.float64).tolist()]


# --------------------------------------------------------------------------------------------
# MGH (io_utils.py:359-412)
# --------------------------------------------------------------------------------------------
def _mgh_bytes(values, affine):
    v = np.asarray(values, dtype=np.float32).reshape(-1)
    n = v.size
    dims = np.array([n, 1, 1], dtype=np.float64)
    delta = np.sqrt(np.sum(affine[:3, :3] ** 2, axis=0))
    mdc = affine[:3, :3] / delta
    c_ras = affine[:3, :3].dot(dims / 2.0) + affine[:3, 3]
    hdr = struct.pack(">7i", 1, n, 1, 1, 1, MGH_FLOAT, 0) + struct.pack(">h", 1)
    hdr += struct.pack(">3f", *delta)
    hdr += struct.pack(">9f", *mdc.T.reshape(-1))  # x_r x_a x_s, y_r y_a y_s, z_r z_a z_s
    hdr += struct.pack(">3f", *c_ras)
    hdr += b"\0" * (MGH_HEADER_BYTES - len(hdr))
    return hdr + v.astype(">f4").tobytes() + struct.pack(">4f", 0.0, 0.0, 0.0, 0.0)


def read_mgh(path):
    """(values, header dict) of a one-frame float MGH file written by :func:`save_results_mgh`."""
    with open(path, "rb") as f:
        raw = f.read()
    version, w, h, d, nframes, dtype, dof = struct.unpack(">7i", raw[:28])
    (good,) = struct.unpack(">h", raw[28:30])
    delta = struct.unpack(">3f", raw[30:42])
    mdc = np.array(struct.unpack(">9f", raw[42:78])).reshape(3, 3).T
    c_ras = struct.unpack(">3f", raw[78:90])
    if version != 1 or dtype != MGH_FLOAT:
        raise ValueError("not a float MGH volume")
    count = w * h * d * nframes
    vals = np.frombuffer(raw, dtype=">f4", count=count, offset=MGH_HEADER_BYTES).astype(np.float32)
    return vals, dict(version=version, dims=(w, h, d, nframes), type=dtype, dof=dof, good_ras=good,
                      delta=delta, mdc=mdc, c_ras=c_ras, footer=struct.unpack(">4f", raw[MGH_HEADER_BYTES + 4 * count:]))


def mgh_template_affine(path):
    """Voxel-to-RAS affine of an MGH / MGZ volume (its direction cosines, voxel sizes and centre):
    M = Mdc diag(delta), translation = c_ras - M (dims / 2)."""
    import gzip

    opener = gzip.open if str(path).endswith((".mgz", ".gz")) else open
    with opener(path, "rb") as f:
        raw = f.read(MGH_HEADER_BYTES)
    version, w, h, d = struct.unpack(">4i", raw[:16])
    (good,) = struct.unpack(">h", raw[28:30])
    if version != 1:
        raise ValueError(f"{path}: not an MGH volume")
    if not good:
        return DEFAULT_MGH_AFFINE.copy()
    delta = np.array(struct.unpack(">3f", raw[30:42]), dtype=np.float64)
    mdc = np.array(struct.unpack(">9f", raw[42:78]), dtype=np.float64).reshape(3, 3).T
    c_ras = np.array(struct.unpack(">3f", raw[78:90]), dtype=np.float64)
    A = np.eye(4)
    A[:3, :3] = mdc * delta
    A[:3, 3] = c_ras - A[:3, :3].dot(np.array([w, h, d], dtype=np.float64) / 2.0)
    return A


def save_results_mgh(output_dir, filename_template, metrics, n_vertices, template_mgz_path=None):
    """One MGH overlay per metric, filename_template.format(metric=name); metrics with no finite
    value are skipped and non-finite values written as 0 (io_utils.py:389-410). With a template
    volume its vox-to-RAS affine is kept (and the centre recomputed for the overlay's
    dimensions), otherwise the reference's default affine."""
    affine = (mgh_template_affine(template_mgz_path) if template_mgz_path and os.path.exists(template_mgz_path)
              else DEFAULT_MGH_AFFINE)
    os.makedirs(output_dir, exist_ok=True)
    written = []
    for name, arr in metrics.items():
        a = np.asarray(arr)
        if a.shape[0] != int(n_vertices):
            raise ValueError(f"Metric '{name}' length mismatch for MGH output: {a.shape[0]} vs {n_vertices}")
        if not np.any(np.isfinite(a)):
            continue
        payload = a.astype(np.float32).reshape(-1).copy()
        payload[~np.isfinite(payload)] = 0
        path = os.path.join(output_dir, filename_template.format(metric=name))
        with open(path, "wb") as f:
            f.write(_mgh_bytes(payload, affine))
        written.append(path)
    return written


# --------------------------------------------------------------------------------------------
# GIFTI (io_utils.py:415-438)
# --------------------------------------------------------------------------------------------
def _gifti_text(values, name):
    v = np.asarray(values, dtype="<f4").reshape(-1)
    data = base64.b64encode(zlib.compress(v.tobytes())).decode("ascii")
    return ('<?xml version="1.0" encoding="UTF-8"?>\n'
            '<!DOCTYPE GIFTI SYSTEM "http://www.nitrc.org/frs/download.php/115/gifti.dtd">\n'
            '<GIFTI Version="1.0" NumberOfDataArrays="1">\n'
            '<MetaData/>\n<LabelTable/>\n'
            f'<DataArray Intent="NIFTI_INTENT_SHAPE" DataType="NIFTI_TYPE_FLOAT32" '
            f'ArrayIndexingOrder="RowMajorOrder" Dimensionality="1" Dim0="{v.size}" '
            'Encoding="GZipBase64Binary" Endian="LittleEndian" ExternalFileName="" ExternalFileOffset="">\n'
            f'<MetaData>\n<MD>\n<Name><![CDATA[Name]]></Name>\n<Value><![CDATA[{escape(name)}]]></Value>\n</MD>\n'
            '</MetaData>\n'
            f'<Data>{data}</Data>\n'
            '</DataArray>\n</GIFTI>\n')


def read_gifti(path):
    """(values, metric name) of a one-array GIFTI file written by :func:`save_results_gifti`."""
    import xml.etree.ElementTree as ET

    root = ET.parse(path).getroot()
    da = root.find("DataArray")
    if da is None or da.get("Encoding") != "GZipBase64Binary" or da.get("DataType") != "NIFTI_TYPE_FLOAT32":
        raise ValueError("not a float32 GZipBase64Binary GIFTI array")
    raw = zlib.decompress(base64.b64decode(da.find("Data").text))
    vals = np.frombuffer(raw, dtype="<f4" if da.get("Endian") == "LittleEndian" else ">f4").astype(np.float32)
    name = None
    for md in da.find("MetaData").findall("MD"):
        if md.find("Name").text == "Name":
            name = md.find("Value").text
    return vals, name


def save_results_gifti(output_dir, filename_template, metrics, n_vertices):
    """One GIFTI shape overlay per metric (float32, NaN kept); metrics with no finite value are
    skipped (io_utils.py:421-438)."""
    os.makedirs(output_dir, exist_ok=True)
    written = []
    for name, arr in metrics.items():
        a = np.asarray(arr, dtype=np.float32)
        if a.shape[0] != int(n_vertices):
            raise ValueError(f"Metric '{name}' length mismatch for GIFTI output: {a.shape[0]} vs {n_vertices}")
        if not np.any(np.isfinite(a)):
            continue
        path = os.path.join(output_dir, filename_template.format(metric=name))
        with open(path, "w", encoding="utf-8") as f:
            f.write(_gifti_text(a, name))
        written.append(path)
    return written


# --------------------------------------------------------------------------------------------
# NPZ with CSR samples (io_utils.py:441-493)
# --------------------------------------------------------------------------------------------
def save_analysis_npz(output_dir, npz_filename, analysis, n_samples_between_scales=None):
    """Scalar metrics plus the CSR (radius, area) samples, schema 3 when the analysis reports its
    provenance (field_health_flag, provenance_json), else 2 -- the keys of io_utils.py:446-478."""
    os.makedirs(output_dir, exist_ok=True)
    path = os.path.join(output_dir, npz_filename)
    m = analysis.n_samples_between_scales if n_samples_between_scales is None else n_samples_between_scales
    cap = getattr(analysis, "boundary_cap_fraction", None)
    payload = dict(analysis.get_metric_arrays())
    payload.update({
        "fastcw_output_schema_version": np.asarray(2, dtype=np.int32),
        "sample_storage_layout": np.asarray("csr_flat_v1"),
        "sample_storage_note": np.asarray(
            "sample_radii_flat/sample_areas_flat are CSR-style rows indexed by sample_indptr[v]:sample_indptr[v+1]."),
        "sample_radii_flat": np.asarray(analysis.sample_radii_flat, dtype=np.float32),
        "sample_areas_flat": np.asarray(analysis.sample_areas_flat, dtype=np.float32),
        "sample_indptr": np.asarray(analysis.sample_indptr, dtype=np.int64),
        "sample_scales_solved": np.asarray(analysis.active_scales, dtype=np.float64),
        "n_samples_between_scales": np.asarray(int(m), dtype=np.int32),
        "boundary_cap_fraction": np.asarray(np.nan if cap is None else float(cap), dtype=np.float64),
        "cortex_mask": np.asarray(analysis.cortex_mask_full, dtype=bool),
        "sub_to_orig": np.asarray(analysis.sub_to_orig, dtype=np.int32),
    })
    if hasattr(analysis, "field_flag"):
        payload["field_health_flag"] = np.asarray(analysis.field_flag, dtype=np.uint8)
    if hasattr(analysis, "provenance"):
        payload["fastcw_output_schema_version"] = np.asarray(3, dtype=np.int32)
        payload["provenance_json"] = np.asarray(json.dumps(analysis.provenance(), sort_keys=True))
    np.savez_compressed(path, **payload)
    return path


def load_sampled_pairs(npz_path):
    """The NPZ's arrays; ValueError if the CSR sample arrays are missing (io_utils.py:481-493)."""
    data = np.load(npz_path, allow_pickle=False)
    missing = [k for k in ("sample_radii_flat", "sample_areas_flat", "sample_indptr") if k not in data]
    if missing:
        raise ValueError(f"NPZ file is missing sampled-pair arrays: {', '.join(missing)}")
    return data


# --------------------------------------------------------------------------------------------
# The CLI's output step (fastcw.py:101-156)
# --------------------------------------------------------------------------------------------
def sample_npz_filename(csv_filename):
    """<stem>_wiring_samples.npz beside <stem>_wiring_costs.csv (fastcw.py:101-108)."""
    name = str(csv_filename)
    if name.endswith("_wiring_costs.csv"):
        return name[: -len("_wiring_costs.csv")] + "_wiring_samples.npz"
    root, _ = os.path.splitext(name)
    return f"{root}_wiring_samples.npz"


def save_analysis_outputs(analysis, output_dir, output_kinds, csv_filename, scalar_stem):
    """Write the requested kinds ("csv", "mgh", "gii") and always the samples NPZ; returns the
    paths written."""
    metrics = analysis.get_metric_arrays()
    written = []
    if "csv" in output_kinds:
        written.append(save_results_csv(output_dir, csv_filename, analysis.cortex_mask_full, analysis.msd_unweighted,
                                        analysis.msd_weighted, analysis.radius_function, analysis.perimeter_function))
    if "mgh" in output_kinds:
        written.extend(save_results_mgh(output_dir, f"{scalar_stem}.mgh", metrics, analysis.n_vertices_full,
                                        getattr(analysis, "metadata", {}).get("mgh_template_path")))
    if "gii" in output_kinds:
        written.extend(save_results_gifti(output_dir, f"{scalar_stem}.shape.gii", metrics,
                                          analysis.n_vertices_full))
    written.append(save_analysis_npz(output_dir, sample_npz_filename(csv_filename), analysis))
    return written
