"""Surface / mask loaders (io_utils.py:23-297) on files built byte by byte from the formats'
layouts, and the reference's resolution rules for FreeSurfer subjects and fsLR inputs."""

import base64
import os
import struct
import zlib

import numpy as np
import pytest

from fastcorticalwiring_amd import inputs, meshes


def _fs_surface_bytes(V, F, comment=b"created by test on today"):
    return (b"\xff\xff\xfe" + comment + b"\n\n" + struct.pack(">2i", len(V), len(F))
            + np.asarray(V, ">f4").tobytes() + np.asarray(F, ">i4").tobytes())


def _label_text(idx, n_total):
    lines = ["#!ascii label  , from subject test vox2ras=TkReg", str(len(idx))]
    lines += [f"{i}  1.000  2.000  3.000 0.0000000000" for i in idx]
    return "\n".join(lines) + "\n"


def _annot_bytes(labels, entries):
    """labels: per-vertex annotation value (-1 = none); entries: [(name, (r, g, b))], version 2 table."""
    out = struct.pack(">i", len(labels))
    for v, lab in enumerate(labels):
        out += struct.pack(">2i", v, lab)
    out += struct.pack(">i", 1) + struct.pack(">i", -2) + struct.pack(">i", len(entries))
    ct = b"colortable.txt\0"
    out += struct.pack(">i", len(ct)) + ct + struct.pack(">i", len(entries))
    for k, (name, (r, g, b)) in enumerate(entries):
        nb = name.encode() + b"\0"
        out += struct.pack(">i", k) + struct.pack(">i", len(nb)) + nb + struct.pack(">4i", r, g, b, 0)
    return out


def _gifti_surface_text(V, F):
    pts = base64.b64encode(zlib.compress(np.asarray(V, "<f4").tobytes())).decode()
    tri = " ".join(str(x) for x in np.asarray(F).ravel())
    return ('<?xml version="1.0" encoding="UTF-8"?>\n<GIFTI Version="1.0" NumberOfDataArrays="2">\n'
            f'<DataArray Intent="NIFTI_INTENT_POINTSET" DataType="NIFTI_TYPE_FLOAT32" ArrayIndexingOrder="RowMajorOrder" '
            f'Dimensionality="2" Dim0="{len(V)}" Dim1="3" Encoding="GZipBase64Binary" Endian="LittleEndian">'
            f'<Data>{pts}</Data></DataArray>\n'
            f'<DataArray Intent="NIFTI_INTENT_TRIANGLE" DataType="NIFTI_TYPE_INT32" ArrayIndexingOrder="RowMajorOrder" '
            f'Dimensionality="2" Dim0="{len(F)}" Dim1="3" Encoding="ASCII" Endian="LittleEndian">'
            f'<Data>{tri}</Data></DataArray>\n</GIFTI>\n')


def _gifti_scalar_text(x):
    data = base64.b64encode(np.asarray(x, "<f4").tobytes()).decode()
    return ('<?xml version="1.0" encoding="UTF-8"?>\n<GIFTI Version="1.0" NumberOfDataArrays="1">\n'
            f'<DataArray Intent="NIFTI_INTENT_SHAPE" DataType="NIFTI_TYPE_FLOAT32" ArrayIndexingOrder="RowMajorOrder" '
            f'Dimensionality="1" Dim0="{len(x)}" Encoding="Base64Binary" Endian="LittleEndian">'
            f'<Data>{data}</Data></DataArray>\n</GIFTI>\n')


@pytest.fixture
def sphere():
    V, F = meshes.icosphere(2)
    return (V * 50).astype(np.float32).astype(np.float64), F.astype(np.int32)


def test_freesurfer_surface_layout(tmp_path, sphere):
    V, F = sphere
    p = tmp_path / "lh.pial"
    p.write_bytes(_fs_surface_bytes(V, F))
    V2, F2 = inputs.read_freesurfer_surface(str(p))
    assert np.array_equal(V2, V) and np.array_equal(F2, F)
    q = tmp_path / "lh.white"
    inputs.write_freesurfer_surface(str(q), V, F)
    assert q.read_bytes()[:3] == b"\xff\xff\xfe"
    V3, F3 = inputs.read_freesurfer_surface(str(q))
    assert np.array_equal(V3, V) and np.array_equal(F3, F)
    bad = tmp_path / "bad"
    bad.write_bytes(b"\x00\x01\x02" + b"x" * 20)
    with pytest.raises(ValueError):
        inputs.read_freesurfer_surface(str(bad))


def test_label_and_annotation_masks(tmp_path):
    n = 10
    lab = tmp_path / "lh.cortex.label"
    lab.write_text(_label_text([0, 2, 3, 9], n))
    assert np.flatnonzero(inputs.read_freesurfer_label(str(lab), n)).tolist() == [0, 2, 3, 9]
    with pytest.raises(ValueError, match="outside"):
        inputs.read_freesurfer_label(str(lab), 5)
    entries = [("unknown", (25, 5, 25)), ("corpuscallosum", (120, 70, 50)), ("precentral", (60, 20, 220))]
    code = {name: r + 256 * g + 65536 * b for name, (r, g, b) in entries}
    labels = [code["precentral"]] * n
    labels[1] = code["unknown"]
    labels[4] = code["corpuscallosum"]
    labels[7] = -1
    ann = tmp_path / "lh.aparc.annot"
    ann.write_bytes(_annot_bytes(labels, entries))
    vals, names = inputs.read_freesurfer_annot(str(ann))
    assert names == code and vals.tolist() == labels
    mask = inputs._annot_cortex_mask(str(ann), n)
    assert np.flatnonzero(~mask).tolist() == [1, 4, 7]


def test_gifti_surface_and_mask(tmp_path, sphere):
    V, F = sphere
    s = tmp_path / "lh.midthickness.surf.gii"
    s.write_text(_gifti_surface_text(V, F))
    m = tmp_path / "lh.mask.shape.gii"
    x = np.zeros(len(V))
    x[::3] = 1
    m.write_text(_gifti_scalar_text(x))
    V2, F2, mask, meta = inputs.load_surface_and_mask(standard="fslr", surface_path=str(s), mask_path=str(m))
    assert np.array_equal(V2, V) and np.array_equal(F2, F)
    assert np.array_equal(mask, x > 0) and meta["output_basename"] == "lh.midthickness"
    with pytest.raises(ValueError, match="requires --mask"):
        inputs.load_surface_and_mask(standard="fslr", surface_path=str(s))


def test_freesurfer_subject_layout_rules(tmp_path, sphere):
    V, F = sphere
    n = len(V)
    subj = tmp_path / "subjects" / "bert"
    (subj / "surf").mkdir(parents=True)
    (subj / "label").mkdir()
    (subj / "mri").mkdir()
    (subj / "surf" / "lh.pial").write_bytes(_fs_surface_bytes(V, F))
    (subj / "label" / "lh.cortex.label").write_text(_label_text(list(range(0, n, 2)), n))
    (subj / "mri" / "orig.mgz").write_bytes(b"x")
    V2, F2, mask, meta = inputs.load_surface_and_mask(standard="freesurfer",
                                                      surface_path=str(subj / "surf" / "lh.pial"))
    assert meta["subject_id"] == "bert" and meta["subject_dir"] == str(tmp_path / "subjects")
    assert meta["mask_source"].startswith("default_label:") and mask.sum() == (n + 1) // 2
    assert meta["output_basename"] == "bert_lh_pial" and meta["mgh_template_path"].endswith("orig.mgz")
    # legacy form: subject_dir + subject_id
    _, _, _, meta2 = inputs.load_surface_and_mask(standard="freesurfer", subject_dir=str(tmp_path / "subjects"),
                                                  subject_id="bert")
    assert meta2["legacy_mode"] is True
    # no cortex label: the aparc annotation fallback, with a warning
    os.remove(subj / "label" / "lh.cortex.label")
    (subj / "label" / "lh.aparc.annot").write_bytes(_annot_bytes([5] * n, [("unknown", (5, 0, 0)), ("x", (1, 1, 1))]))
    with pytest.warns(RuntimeWarning, match="aparc"):
        _, _, mask3, meta3 = inputs.load_surface_and_mask(standard="freesurfer",
                                                          surface_path=str(subj / "surf" / "lh.pial"))
    assert meta3["mask_source"].startswith("aparc_fallback:") and not mask3.any()
    os.remove(subj / "label" / "lh.aparc.annot")
    with pytest.raises(FileNotFoundError):
        inputs.load_surface_and_mask(standard="freesurfer", surface_path=str(subj / "surf" / "lh.pial"))
    _, _, mask4, meta4 = inputs.load_surface_and_mask(standard="freesurfer", surface_path=str(subj / "surf" / "lh.pial"),
                                                      no_mask=True)
    assert mask4.all() and meta4["mask_source"] == "all_vertices_no_mask"
    co = subj / "surf" / "lh.cortexonly.pial"
    co.write_bytes(_fs_surface_bytes(V, F))
    _, _, mask5, meta5 = inputs.load_surface_and_mask(standard="freesurfer", surface_path=str(co))
    assert mask5.all() and meta5["mask_source"] == "cortexonly_surface_all_vertices"
    with pytest.raises(ValueError, match="Unsupported standard"):
        inputs.load_surface_and_mask(standard="mni", surface_path=str(co))
