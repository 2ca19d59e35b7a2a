"""Surface and cortex-mask loaders: FreeSurfer surfaces, labels and annotations, GIFTI surfaces and
masks, and the reference's loading rules (io_utils.py:23-297).

nibabel is not needed: the files are read from their published layouts.

* FreeSurfer triangle surface: 3-byte magic 0xFFFFFE, a comment ended by two newlines, then
  big-endian int32 vertex and face counts, float32 coordinates (n x 3), int32 faces (m x 3).
* FreeSurfer label: text; line 1 holds the count, then one "vertex x y z value" line each.
* FreeSurfer annotation: big-endian int32 count, (vertex, annotation) pairs, then a colour
  table flag and table (version 2 layout: entry index, name, r, g, b, transparency; the
  annotation of an entry is r + 256 g + 65536 b), or the original layout without versions.
* GIFTI: XML with one NIFTI_INTENT_POINTSET and one NIFTI_INTENT_TRIANGLE DataArray (ASCII,
  Base64Binary or GZipBase64Binary data), masks as the first scalar DataArray.
"""

from __future__ import annotations

import base64
import os
import struct
import warnings
import zlib

import numpy as np

__all__ = ["read_freesurfer_surface", "write_freesurfer_surface", "read_freesurfer_label",
           "read_freesurfer_annot", "read_gifti_arrays", "load_surface_and_mask", "infer_output_basename",
           "select_vertex_subset", "random_vertex_sampling", "stratified_vertex_sampling", "farthest_point_sampling"]

FS_TRIANGLE_MAGIC = 0xFFFFFE


def infer_output_basename(surface_path, fallback="surface"):
    """Output basename from a surface filename (io_utils.py:23-32)."""
    if not surface_path:
        return fallback
    name = os.path.basename(surface_path)
    for suffix in (".surf.gii", ".shape.gii", ".func.gii", ".gii"):
        if name.endswith(suffix):
            return name[: -len(suffix)]
    root, _ = os.path.splitext(name)
    return root if root else fallback


# --------------------------------------------------------------------------------------------
# FreeSurfer
# --------------------------------------------------------------------------------------------
def read_freesurfer_surface(path):
    """(vertices float64 n x 3, faces int32 m x 3) of a FreeSurfer triangle surface."""
    if not os.path.exists(path):
        raise FileNotFoundError(f"Surface file not found: {path}")
    with open(path, "rb") as f:
        raw = f.read()
    magic = int.from_bytes(raw[:3], "big")
    if magic != FS_TRIANGLE_MAGIC:
        raise ValueError(f"{path}: not a FreeSurfer triangle surface (magic {magic:#x})")
    end = raw.find(b"\n\n", 3)
    if end < 0:
        raise ValueError(f"{path}: malformed surface header")
    off = end + 2
    nv, nf = struct.unpack(">2i", raw[off:off + 8])
    off += 8
    V = np.frombuffer(raw, dtype=">f4", count=3 * nv, offset=off).reshape(nv, 3).astype(np.float64)
    off += 12 * nv
    F = np.frombuffer(raw, dtype=">i4", count=3 * nf, offset=off).reshape(nf, 3).astype(np.int32)
    return V, F


def write_freesurfer_surface(path, vertices, faces, comment="created by fastcorticalwiring_amd"):
    """Write a FreeSurfer triangle surface (the layout read_freesurfer_surface reads)."""
    V = np.asarray(vertices, dtype=">f4").reshape(-1, 3)
    F = np.asarray(faces, dtype=">i4").reshape(-1, 3)
    with open(path, "wb") as f:
        f.write(FS_TRIANGLE_MAGIC.to_bytes(3, "big"))
        f.write(comment.encode() + b"\n\n")
        f.write(struct.pack(">2i", V.shape[0], F.shape[0]))
        f.write(V.tobytes())
        f.write(F.tobytes())


def read_freesurfer_label(path, n_vertices):
    """Boolean mask of the label's vertices; ValueError when an index lies outside the surface
    (io_utils.py:41-71)."""
    with open(path, "r", encoding="utf-8") as f:
        lines = f.readlines()
    if len(lines) < 2:
        raise ValueError(f"Malformed label file: {path}")
    count = int(lines[1].strip())
    idx = []
    for line in lines[2:2 + count]:
        fields = line.split()
        if fields:
            idx.append(int(fields[0]))
    idx = np.asarray(idx, dtype=np.int64)
    mask = np.zeros(int(n_vertices), dtype=bool)
    valid = (idx >= 0) & (idx < mask.size)
    if not valid.all():
        raise ValueError(f"{path}: {int((~valid).sum())} label indices fall outside this {n_vertices}-vertex "
                         f"surface (max index {int(idx.max())}); the label belongs to a different mesh. "
                         "Use --no-mask for cortex-only surfaces.")
    mask[idx] = True
    return mask


def read_freesurfer_annot(path):
    """(per-vertex annotation values, {structure name: annotation value}) of an annotation file."""
    with open(path, "rb") as f:
        raw = f.read()
    off = 0

    def i32():
        nonlocal off
        (v,) = struct.unpack(">i", raw[off:off + 4])
        off += 4
        return v

    def text():
        nonlocal off
        n = i32()
        s = raw[off:off + n].split(b"\0", 1)[0].decode("utf-8", "ignore")
        off += n
        return s

    nv = i32()
    pairs = np.frombuffer(raw, dtype=">i4", count=2 * nv, offset=off).reshape(nv, 2)
    off += 8 * nv
    labels = np.full(nv, -1, dtype=np.int64)
    labels[pairs[:, 0]] = pairs[:, 1]
    names = {}
    if off < len(raw) and i32() == 1:
        version = i32()
        if version < 0:  # version 2 layout
            if -version != 2:
                raise ValueError(f"{path}: unsupported colour table version {-version}")
            i32()  # max structure index
            text()  # original colour table file
            for _ in range(i32()):
                i32()  # structure index
                name = text()
                r, g, b, _t = (i32() for _ in range(4))
                names[name] = r + 256 * g + 65536 * b
        else:  # original layout: version holds the entry count
            text()
            for _ in range(version):
                name = text()
                r, g, b, _t = (i32() for _ in range(4))
                names[name] = r + 256 * g + 65536 * b
    return labels, names


def _normalize_region_name(name):
    return "".join(ch.lower() for ch in str(name) if ch.isalnum())


def _annot_cortex_mask(path, n_vertices):
    """Cortex = every vertex not in unknown / corpus callosum / medial wall (io_utils.py:74-101)."""
    labels, names = read_freesurfer_annot(path)
    if labels.shape[0] != int(n_vertices):
        raise ValueError(f"Annotation vertex count mismatch: annot has {labels.shape[0]}, mesh has {n_vertices}")
    excl = {code for name, code in names.items()
            if any(tok in _normalize_region_name(name) for tok in ("unknown", "corpuscallosum", "medialwall"))}
    mask = np.ones(int(n_vertices), dtype=bool)
    for code in excl:
        mask[labels == code] = False
    mask[labels == -1] = False
    return mask


# --------------------------------------------------------------------------------------------
# GIFTI
# --------------------------------------------------------------------------------------------
_GIFTI_DTYPES = {"NIFTI_TYPE_FLOAT32": "f4", "NIFTI_TYPE_FLOAT64": "f8", "NIFTI_TYPE_INT32": "i4",
                 "NIFTI_TYPE_UINT8": "u1", "NIFTI_TYPE_INT16": "i2", "NIFTI_TYPE_INT8": "i1"}


def read_gifti_arrays(path):
    """[(intent name, numpy array)] of a GIFTI file's DataArrays."""
    import xml.etree.ElementTree as ET

    if not os.path.exists(path):
        raise FileNotFoundError(f"GIFTI file not found: {path}")
    root = ET.parse(path).getroot()
    if root.tag != "GIFTI":
        raise ValueError(f"Expected GIFTI file, got: {path}")
    out = []
    for da in root.findall("DataArray"):
        dt = _GIFTI_DTYPES.get(da.get("DataType"))
        if dt is None:
            raise ValueError(f"{path}: unsupported GIFTI data type {da.get('DataType')}")
        dims = [int(da.get(f"Dim{i}")) for i in range(int(da.get("Dimensionality", "1")))]
        enc = da.get("Encoding", "ASCII")
        text = (da.find("Data").text or "").strip()
        if enc == "ASCII":
            arr = np.array(text.split(), dtype=dt)
        else:
            raw = base64.b64decode(text)
            if enc == "GZipBase64Binary":
                raw = zlib.decompress(raw)
            elif enc != "Base64Binary":
                raise ValueError(f"{path}: unsupported GIFTI encoding {enc}")
            endian = "<" if da.get("Endian", "LittleEndian") == "LittleEndian" else ">"
            arr = np.frombuffer(raw, dtype=endian + dt)
        order = "F" if da.get("ArrayIndexingOrder") == "ColumnMajorOrder" else "C"
        out.append((da.get("Intent"), arr.reshape(dims, order=order)))
    return out


def _load_gifti_surface(path):
    arrays = read_gifti_arrays(path)
    pts = [a for i, a in arrays if i == "NIFTI_INTENT_POINTSET"]
    tri = [a for i, a in arrays if i == "NIFTI_INTENT_TRIANGLE"]
    if len(pts) != 1 or len(tri) != 1:
        raise ValueError(f"Malformed GIFTI surface '{path}': expected exactly one POINTSET and one TRIANGLE array")
    return np.asarray(pts[0], dtype=np.float64), np.asarray(tri[0], dtype=np.int32)


def _load_gifti_mask(path, n_vertices):
    arrays = read_gifti_arrays(path)
    if not arrays:
        raise ValueError(f"GIFTI mask contains no data arrays: {path}")
    data = np.squeeze(arrays[0][1])
    if data.ndim != 1:
        raise ValueError(f"Mask must be a 1D scalar GIFTI array, got shape {tuple(data.shape)}: {path}")
    if data.shape[0] != int(n_vertices):
        raise ValueError(f"Mask vertex count mismatch: mask has {data.shape[0]}, mesh has {n_vertices} ({path})")
    return data > 0


# --------------------------------------------------------------------------------------------
# The reference's loading rules (io_utils.py:160-297)
# --------------------------------------------------------------------------------------------
def load_surface_and_mask(*, standard, surface_path=None, mask_path=None, subject_dir=None, subject_id=None,
                          hemi="lh", surf_type="pial", custom_label=None, no_mask=False):
    """(vertices, faces, cortex_mask, metadata) for FreeSurfer or fsLR inputs, resolving the
    subject layout, the mask (label, annotation, or the default cortex label / aparc fallback)
    and the MGH template exactly as io_utils.load_surface_and_mask."""
    standard = str(standard).lower()
    meta = {"standard": standard, "subject_dir": subject_dir, "subject_id": subject_id, "hemi": hemi,
            "surf_type": surf_type, "custom_label": custom_label}
    if standard == "freesurfer":
        legacy = False
        if surface_path is not None:
            surf_dir = os.path.dirname(os.path.abspath(surface_path))
            if os.path.basename(surf_dir) == "surf":
                root = os.path.dirname(surf_dir)
                subject_id = subject_id if subject_id is not None else os.path.basename(root)
                subject_dir = subject_dir if subject_dir is not None else os.path.dirname(root)
                meta["subject_dir"], meta["subject_id"] = subject_dir, subject_id
        if surface_path is None:
            if subject_dir is None or subject_id is None:
                raise ValueError("FreeSurfer loading requires either --surface or both subject_dir and subject_id.")
            legacy = True
            surface_path = os.path.join(subject_dir, subject_id, "surf", f"{hemi}.{surf_type}")
        V, F = read_freesurfer_surface(surface_path)
        n = V.shape[0]
        meta["surface_path"] = surface_path
        meta["legacy_mode"] = legacy
        meta["output_basename"] = f"{subject_id}_{hemi}_{surf_type}" if subject_id else infer_output_basename(surface_path)
        if ".cortexonly." in os.path.basename(str(surface_path)) and not no_mask:
            if mask_path or custom_label:
                raise ValueError(f"{surface_path} is a cortex-only surface; explicit masks or labels do not apply. "
                                 "Remove --mask/--custom-label to use all vertices.")
            no_mask = True
            meta["mask_source"] = "cortexonly_surface_all_vertices"
        if no_mask:
            mask = np.ones(n, dtype=bool)
            meta.setdefault("mask_source", "all_vertices_no_mask")
        elif mask_path:
            ext = os.path.splitext(mask_path)[1].lower()
            if ext == ".label":
                mask = read_freesurfer_label(mask_path, n)
                meta["mask_source"] = f"explicit_label:{mask_path}"
            elif ext == ".annot":
                mask = _annot_cortex_mask(mask_path, n)
                meta["mask_source"] = f"explicit_annot:{mask_path}"
            else:
                raise ValueError(f"Unsupported FreeSurfer mask file extension '{ext}'. Use .label or .annot.")
        else:
            label_dir = None
            if subject_dir and subject_id:
                label_dir = os.path.join(subject_dir, subject_id, "label")
            elif surface_path:
                maybe = os.path.join(os.path.dirname(os.path.dirname(surface_path)), "label")
                if os.path.isdir(maybe):
                    label_dir = maybe
            mask = None
            if label_dir:
                if custom_label:
                    cand = os.path.join(label_dir, f"{hemi}.{custom_label}.label")
                    if os.path.exists(cand):
                        mask = read_freesurfer_label(cand, n)
                        meta["mask_source"] = f"custom_label:{cand}"
                if mask is None:
                    cand = os.path.join(label_dir, f"{hemi}.cortex.label")
                    if os.path.exists(cand):
                        mask = read_freesurfer_label(cand, n)
                        meta["mask_source"] = f"default_label:{cand}"
                if mask is None:
                    cand = os.path.join(label_dir, f"{hemi}.aparc.annot")
                    if os.path.exists(cand):
                        warnings.warn("Cortex label not found; using aparc annotation fallback.", RuntimeWarning)
                        mask = _annot_cortex_mask(cand, n)
                        meta["mask_source"] = f"aparc_fallback:{cand}"
            if mask is None:
                raise FileNotFoundError("No cortical mask found for FreeSurfer surface. "
                                        "Provide --mask, use --custom-label if applicable, or pass --no-mask explicitly.")
        if subject_dir and subject_id:
            tpl = os.path.join(subject_dir, subject_id, "mri", "orig.mgz")
            if not os.path.exists(tpl):
                tpl = os.path.join(subject_dir, subject_id, "mri", "T1.mgz")
            if os.path.exists(tpl):
                meta["mgh_template_path"] = tpl
        return V, F, mask.astype(bool), meta
    if standard == "fslr":
        if surface_path is None:
            raise ValueError("fsLR loading requires --surface pointing to a .surf.gii file.")
        V, F = _load_gifti_surface(surface_path)
        n = V.shape[0]
        meta["surface_path"] = surface_path
        meta["legacy_mode"] = False
        meta["output_basename"] = infer_output_basename(surface_path)
        if no_mask:
            mask = np.ones(n, dtype=bool)
            meta["mask_source"] = "all_vertices_no_mask"
        else:
            if not mask_path:
                raise ValueError("fsLR mode requires --mask (GIFTI scalar) or --no-mask.")
            mask = _load_gifti_mask(mask_path, n)
            meta["mask_source"] = f"gifti_mask:{mask_path}"
        return V, F, mask.astype(bool), meta
    raise ValueError(f"Unsupported standard '{standard}'. Expected 'freesurfer' or 'fslr'.")


# --------------------------------------------------------------------------------------------
# Vertex subsets (io_utils.py:496-630): the same selections for the same seed
# --------------------------------------------------------------------------------------------
def random_vertex_sampling(valid_mask, k, random_state=None):
    """k cortical vertices uniformly without replacement (all of them when k covers them)."""
    valid = np.flatnonzero(valid_mask)
    if k >= valid.size:
        return valid.tolist()
    pick = np.random.default_rng(random_state).choice(valid, size=int(k), replace=False)
    return np.asarray(pick, dtype=np.int64).tolist()


def stratified_vertex_sampling(vertices, valid_mask, k, n_bins=None, random_state=None):
    """k cortical vertices spread over the cortical order: contiguous bins get floor shares of k,
    the remainder goes to the bins with the largest fractional shares, each bin sampled without
    replacement (cortical order is a cheap proxy for spatial spread)."""
    del vertices
    valid = np.flatnonzero(valid_mask)
    nv = valid.size
    if k >= nv:
        return valid.tolist()
    k = int(k)
    nb = max(8, min(256, int(np.sqrt(k)))) if n_bins is None else int(n_bins)
    nb = max(1, min(nb, nv))
    bins = np.array_split(valid, nb)
    size = np.array([b.size for b in bins], dtype=np.int64)
    share = size / float(nv) * float(k)
    take = np.minimum(np.floor(share).astype(np.int64), size)
    left = k - int(take.sum())
    if left > 0:
        for i in np.argsort(-(share - take)):
            if left <= 0:
                break
            room = int(size[i] - take[i])
            if room > 0:
                t = min(room, left)
                take[i] += t
                left -= t
    rng = np.random.default_rng(random_state)
    parts = [np.asarray(b, dtype=np.int64) if t >= b.size else np.asarray(rng.choice(b, size=int(t), replace=False), np.int64)
             for b, t in zip(bins, take) if t > 0]
    chosen = np.concatenate(parts) if parts else np.empty(0, np.int64)
    if chosen.size < k:  # top up from what is left (only when rounding left a gap)
        rest = np.setdiff1d(valid, chosen, assume_unique=False)
        extra = rng.choice(rest, size=int(k - chosen.size), replace=False)
        chosen = np.concatenate([chosen, np.asarray(extra, np.int64)])
    return chosen.tolist()


def farthest_point_sampling(vertices, k, valid_mask):
    """Euclidean farthest-point sampling over the cortical vertices, from the first of them."""
    valid = np.flatnonzero(valid_mask)
    if k >= valid.size:
        return valid.tolist()
    P = np.asarray(vertices, dtype=np.float64)[valid]
    out = np.zeros(int(k), dtype=np.int64)
    dmin = np.full(P.shape[0], np.inf)
    cur = 0
    for i in range(int(k)):
        out[i] = cur
        dmin = np.minimum(dmin, np.sum((P - P[cur]) ** 2, axis=1))
        cur = int(np.argmax(dmin))
    return valid[out].tolist()


def select_vertex_subset(vertices, cortex_mask, *, sample_frac=None, sample_count=None, sample_method="stratified",
                         random_state=None):
    """Cortical subset by fraction or count with the named method; None when neither is given."""
    valid = np.flatnonzero(cortex_mask)
    if valid.size == 0:
        return []
    if sample_frac is not None and sample_count is not None:
        raise ValueError("sample_frac and sample_count are mutually exclusive.")
    if sample_count is not None:
        k = int(sample_count)
        if k <= 0:
            raise ValueError("sample_count must be a positive integer.")
    elif sample_frac is not None:
        frac = float(sample_frac)
        if frac <= 0.0 or frac > 1.0:
            raise ValueError("sample_frac must satisfy 0 < sample_frac <= 1.")
        k = max(1, int(round(frac * valid.size)))
    else:
        return None
    k = min(k, valid.size)
    method = str(sample_method or "stratified").lower()
    if method == "stratified":
        return stratified_vertex_sampling(vertices, cortex_mask, k, random_state=random_state)
    if method == "random":
        return random_vertex_sampling(cortex_mask, k, random_state=random_state)
    if method == "fps":
        return farthest_point_sampling(vertices, k, cortex_mask)
    raise ValueError(f"Unknown sample_method: {sample_method}")
