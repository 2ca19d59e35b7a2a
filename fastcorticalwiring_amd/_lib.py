"""ctypes binding of libfcw.so (include/fcw.h).

The library is built in-tree (``fastcorticalwiring_amd/libfcw.so``) by
``fastcorticalwiring_amd/build.py``; importing this module never builds or
falls back to anything else: if the library is missing, ``lib()`` raises.
"""

from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfcw.so")

# HIP serves a process's streams from GPU_MAX_HW_QUEUES hardware queues (4 unless set). A context
# keeps one RHS block in flight per stream (8 by default), and blocks on more streams than queues
# share a queue and run one after another. The HIP runtime reads the variable once, when it
# initialises, so it is raised here, before the library is loaded; a larger value set by the
# caller is kept, and a process that initialised HIP earlier keeps what it has.
HW_QUEUES = 8
try:
    if int(os.environ.get("GPU_MAX_HW_QUEUES", "0")) < HW_QUEUES:
        os.environ["GPU_MAX_HW_QUEUES"] = str(HW_QUEUES)
except ValueError:
    os.environ["GPU_MAX_HW_QUEUES"] = str(HW_QUEUES)

FCW_FP64 = 0
FCW_FP32_IR = 1
FCW_BF16X3_IR = 2

(S_FLAG, S_NEG_FRAC, S_VIOL_FRAC, S_DIST_BOUNDARY, S_MSD_U, S_MSD_W, S_N_EVAL, S_N_SAMPLES,
 S_OVERFLOW, NSCALARS) = range(10)


class Options(ctypes.Structure):
    _fields_ = [
        ("t", ctypes.c_double),
        ("t_coef", ctypes.c_double),
        ("poisson_shift", ctypes.c_double),
        ("precision", ctypes.c_int),
        ("block", ctypes.c_int),
        ("leaf_size", ctypes.c_int),
        ("host_threads", ctypes.c_int),
        ("device", ctypes.c_int),
        ("ir_steps", ctypes.c_int),
        ("streams", ctypes.c_int),
        ("sparse_rhs", ctypes.c_int),
        ("robust", ctypes.c_int),
        ("split_depth", ctypes.c_int),
        ("band_cap", ctypes.c_int),
        ("graphs", ctypes.c_int),
        ("tile_rows", ctypes.c_int),
        ("index_staging", ctypes.c_int),
    ]


class Info(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int),
        ("m", ctypes.c_int),
        ("nnodes", ctypes.c_int),
        ("nlevels", ctypes.c_int),
        ("max_front", ctypes.c_int),
        ("block", ctypes.c_int),
        ("factor_nnz", ctypes.c_int64),
        ("panel_elems", ctypes.c_int64),
        ("factor_flops", ctypes.c_double),
        ("solve_flops_per_rhs", ctypes.c_double),
        ("t", ctypes.c_double),
        ("mean_edge", ctypes.c_double),
        ("setup_seconds", ctypes.c_double * 4),
        ("device_bytes", ctypes.c_int64),
        ("robust", ctypes.c_int),
        ("intrinsic_flips", ctypes.c_int),
        ("intrinsic_skipped", ctypes.c_int),
        ("mollify_delta", ctypes.c_double),
        ("split_levels", ctypes.c_int),
        ("op_faces", ctypes.c_int),
        ("tufted_cover", ctypes.c_int),
        ("nonmanifold_edges", ctypes.c_int),
    ]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "setup_seconds"}
        d["setup_seconds"] = list(self.setup_seconds)
        return d


class MetricsOut(ctypes.Structure):
    _fields_ = [
        ("scalars", ctypes.POINTER(ctypes.c_double)),
        ("radii", ctypes.POINTER(ctypes.c_double)),
        ("perimeters", ctypes.POINTER(ctypes.c_double)),
        ("extra_r", ctypes.POINTER(ctypes.c_double)),
        ("extra_a", ctypes.POINTER(ctypes.c_double)),
        ("sample_r", ctypes.POINTER(ctypes.c_float)),
        ("sample_a", ctypes.POINTER(ctypes.c_float)),
    ]


class MeshPrepOut(ctypes.Structure):
    _fields_ = [
        ("n_sub", ctypes.c_int),
        ("m_sub", ctypes.c_int),
        ("n_boundary", ctypes.c_int),
        ("n_adj", ctypes.c_int),
        ("sub_to_orig", ctypes.POINTER(ctypes.c_int32)),
        ("orig_to_sub", ctypes.POINTER(ctypes.c_int32)),
        ("faces", ctypes.POINTER(ctypes.c_int32)),
        ("vertices", ctypes.POINTER(ctypes.c_double)),
        ("boundary", ctypes.POINTER(ctypes.c_int32)),
        ("order", ctypes.POINTER(ctypes.c_int32)),
        ("face_areas", ctypes.POINTER(ctypes.c_double)),
        ("face_L", ctypes.POINTER(ctypes.c_double)),
        ("vertex_areas", ctypes.POINTER(ctypes.c_double)),
        ("adj_ptr", ctypes.POINTER(ctypes.c_int32)),
        ("adj", ctypes.POINTER(ctypes.c_int32)),
    ]


# Every symbol include/fcw.h declares.
EXPORTED_SYMBOLS = (
    "fcw_default_options", "fcw_create", "fcw_destroy", "fcw_get_info", "fcw_distance_block",
    "fcw_set_metrics", "fcw_wiring_sources", "fcw_geometry_fields", "fcw_stage_times",
    "fcw_set_profiling", "fcw_kernel_times", "fcw_build_info", "fcw_last_error", "fcw_host_factor_solve",
    "fcw_calibration_copy", "fcw_host_intrinsic", "fcw_mesh_prep", "fcw_host_analyse",
    "fcw_wiring_sources_noalloc", "fcw_kernel_log", "fcw_host_intrinsic_cover", "fcw_divergence_block",
)

_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_vp = ctypes.c_void_p


def lib():
    """Load libfcw.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"libfcw.so not found at {LIB_PATH}; build it with "
            "`python -m fastcorticalwiring_amd.build` (hipcc, gfx950)")
    # The host factorisation's OpenMP threads wait passively at their barriers: spinning waiters
    # take the time slices of a team member that shares its core with other work (another
    # context's factorisation, a CPU process), which made the factorisation ten times slower
    # (L7: 1.5 s instead of 0.14 s with two busy cores). libgomp reads this when it is loaded,
    # here with the library; a value the caller has set is kept.
    os.environ.setdefault("OMP_WAIT_POLICY", "PASSIVE")
    L = ctypes.CDLL(LIB_PATH)
    L.fcw_default_options.argtypes = [ctypes.POINTER(Options)]
    L.fcw_default_options.restype = None
    L.fcw_create.argtypes = [_dp, ctypes.c_int, _ip, ctypes.c_int, ctypes.POINTER(Options), ctypes.POINTER(_vp)]
    L.fcw_destroy.argtypes = [_vp]
    L.fcw_get_info.argtypes = [_vp, ctypes.POINTER(Info)]
    L.fcw_distance_block.argtypes = [_vp, _ip, ctypes.c_int, _dp]
    L.fcw_divergence_block.argtypes = [_vp, _dp, ctypes.c_int, _dp]
    L.fcw_set_metrics.argtypes = [_vp, _dp, ctypes.c_int, _dp, ctypes.c_int, _ip, ctypes.c_int, _dp,
                                  ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_double, ctypes.c_int]
    L.fcw_wiring_sources.argtypes = [_vp, _ip, ctypes.c_int, ctypes.POINTER(MetricsOut)]
    L.fcw_wiring_sources_noalloc.argtypes = [_vp, _ip, ctypes.c_int, ctypes.POINTER(MetricsOut)]
    L.fcw_geometry_fields.argtypes = [_vp, _dp, _ip, ctypes.c_int, ctypes.c_int, ctypes.POINTER(MetricsOut)]
    L.fcw_stage_times.argtypes = [_vp, _dp, ctypes.c_int]
    L.fcw_set_profiling.argtypes = [_vp, ctypes.c_int]
    L.fcw_kernel_times.argtypes = [_vp, _dp, ctypes.c_int]
    L.fcw_kernel_log.argtypes = [_vp, _dp, ctypes.c_int]
    L.fcw_kernel_log.restype = ctypes.c_int
    L.fcw_build_info.argtypes = [ctypes.c_char_p, ctypes.c_int]
    L.fcw_last_error.argtypes = []
    L.fcw_last_error.restype = ctypes.c_char_p
    L.fcw_calibration_copy.argtypes = [ctypes.c_int64, ctypes.c_int]
    L.fcw_host_factor_solve.argtypes = [_dp, ctypes.c_int, _ip, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_int, _dp, ctypes.c_int, _dp, _dp]
    L.fcw_host_intrinsic.argtypes = [_dp, ctypes.c_int, _ip, ctypes.c_int, ctypes.c_double, _ip, _dp, _dp, _dp, _dp]
    L.fcw_host_intrinsic_cover.argtypes = [_dp, ctypes.c_int, _ip, ctypes.c_int, ctypes.c_double, _ip, _ip, _dp, _dp,
                                           _dp, _dp, _dp]
    L.fcw_host_intrinsic_cover.restype = ctypes.c_int
    L.fcw_host_analyse.argtypes = [_dp, ctypes.c_int, _ip, ctypes.c_int, ctypes.c_int, ctypes.c_int, _ip, _ip, _ip, _ip,
                                   _dp, ctypes.POINTER(ctypes.c_uint64)]
    L.fcw_mesh_prep.argtypes = [_dp, ctypes.c_int, _ip, ctypes.c_int, ctypes.POINTER(ctypes.c_ubyte), ctypes.c_int,
                                ctypes.c_int, ctypes.POINTER(MeshPrepOut)]
    for name in ("fcw_create", "fcw_destroy", "fcw_get_info", "fcw_distance_block", "fcw_divergence_block", "fcw_set_metrics",
                 "fcw_wiring_sources", "fcw_geometry_fields", "fcw_stage_times", "fcw_set_profiling",
                 "fcw_kernel_times", "fcw_build_info", "fcw_host_factor_solve", "fcw_calibration_copy",
                 "fcw_host_intrinsic", "fcw_mesh_prep", "fcw_host_analyse", "fcw_wiring_sources_noalloc"):
        getattr(L, name).restype = ctypes.c_int
    _lib = L
    return L


class FcwError(RuntimeError):
    pass


def check(rc):
    if rc != 0:
        msg = lib().fcw_last_error().decode("utf-8", "replace")
        raise FcwError(msg or f"libfcw error {rc}")


def build_info() -> str:
    buf = ctypes.create_string_buffer(512)
    lib().fcw_build_info(buf, 512)
    return buf.value.decode()


def dptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def iptr(a):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


def fptr(a):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def host_factor_solve(V, F, l_scale, m_scale, B, leaf_size=64):
    """CPU-only check of the analysis + FP64 factorisation: X = (l*L + m*M)^{-1} B (B is n x k)."""
    V = np.ascontiguousarray(V, dtype=np.float64)
    F = np.ascontiguousarray(F, dtype=np.int32)
    B = np.ascontiguousarray(B, dtype=np.float64)
    if B.ndim == 1:
        B = B[:, None]
    X = np.empty_like(B)
    stats = np.zeros(4)
    check(lib().fcw_host_factor_solve(dptr(V), V.shape[0], iptr(F), F.shape[0], float(l_scale), float(m_scale),
                                      int(leaf_size), dptr(B), B.shape[1], dptr(X), dptr(stats)))
    return X, dict(nnodes=int(stats[0]), nlevels=int(stats[1]), factor_nnz=int(stats[2]),
                   panel_elems=int(stats[3]))


def host_analyse(V, F, leaf_size=64, par_depth=-1, panel_hash=True):
    """CPU-only view of the nested-dissection analysis: dict with perm, node_start, node_s, node_b,
    the counts and (optionally) the FNV-1a hash of the FP64 panels of M + L."""
    V = np.ascontiguousarray(V, dtype=np.float64)
    F = np.ascontiguousarray(F, dtype=np.int32)
    n = V.shape[0]
    perm = np.empty(n, np.int32)
    ns, nsz, nb = (np.empty(n, np.int32) for _ in range(3))
    stats = np.zeros(4)
    h = ctypes.c_uint64(0)
    check(lib().fcw_host_analyse(dptr(V), n, iptr(F), F.shape[0], int(leaf_size), int(par_depth), iptr(perm),
                                 iptr(ns), iptr(nsz), iptr(nb), dptr(stats), ctypes.byref(h) if panel_hash else None))
    nn = int(stats[0])
    return dict(perm=perm, node_start=ns[:nn], node_s=nsz[:nn], node_b=nb[:nn], nnodes=nn, nlevels=int(stats[1]),
                factor_nnz=int(stats[2]), panel_elems=int(stats[3]), panel_hash=int(h.value) if panel_hash else None)


def host_intrinsic(V, F, mollify=1e-6):
    """CPU-only: the intrinsic Delaunay triangulation the robust operators use. Returns (faces m x 3,
    lengths m x 3 opposite each corner, area m, gram m x 6, stats)."""
    V = np.ascontiguousarray(V, dtype=np.float64)
    F = np.ascontiguousarray(F, dtype=np.int32)
    m = F.shape[0]
    faces = np.empty((m, 3), np.int32)
    length = np.empty((m, 3))
    area = np.empty(m)
    gram = np.empty((m, 6))
    stats = np.zeros(3)
    check(lib().fcw_host_intrinsic(dptr(V), V.shape[0], iptr(F), m, float(mollify), iptr(faces), dptr(length),
                                   dptr(area), dptr(gram), dptr(stats)))
    return faces, length, area, gram, dict(flips=int(stats[0]), skipped=int(stats[1]), delta=float(stats[2]))


def host_intrinsic_cover(V, F, mollify=1e-6):
    """CPU-only: the intrinsic Delaunay triangulation for any mesh (tufted cover for non-manifold
    edges). Returns (faces mo x 3, lengths mo x 3, weight mo, area mo (x weight), gram mo x 6, stats)."""
    V = np.ascontiguousarray(V, dtype=np.float64)
    F = np.ascontiguousarray(F, dtype=np.int32)
    m = F.shape[0]
    faces = np.empty((2 * m, 3), np.int32)
    length = np.empty((2 * m, 3))
    weight = np.empty(2 * m)
    area = np.empty(2 * m)
    gram = np.empty((2 * m, 6))
    stats = np.zeros(5)
    mo = ctypes.c_int(0)
    check(lib().fcw_host_intrinsic_cover(dptr(V), V.shape[0], iptr(F), m, float(mollify), ctypes.byref(mo),
                                         iptr(faces), dptr(length), dptr(weight), dptr(area), dptr(gram), dptr(stats)))
    k = mo.value
    return (faces[:k], length[:k], weight[:k], area[:k], gram[:k],
            dict(flips=int(stats[0]), skipped=int(stats[1]), delta=float(stats[2]), cover=bool(stats[3]),
                 nonmanifold_edges=int(stats[4])))


def mesh_prep_device(V, F, mask, start=0, device=0):
    """Mesh preparation on the device (fcw_mesh_prep): a dict with the cortical submesh
    (vertices, faces, sub_to_orig, orig_to_sub), face_areas, face_L, vertex_areas, boundary, the
    adjacency CSR (adj_ptr, adj) and the breadth-first order, each equal to mesh_prep's host result."""
    V = np.ascontiguousarray(V, dtype=np.float64)
    F = np.ascontiguousarray(F, dtype=np.int32)
    M = np.ascontiguousarray(mask, dtype=np.uint8)
    n, m = V.shape[0], F.shape[0]
    buf = dict(sub_to_orig=np.empty(n, np.int32), orig_to_sub=np.empty(n, np.int32), faces=np.empty(3 * m, np.int32),
               vertices=np.empty(3 * n), boundary=np.empty(n, np.int32), order=np.empty(n, np.int32),
               face_areas=np.empty(m), face_L=np.empty(m), vertex_areas=np.empty(n),
               adj_ptr=np.empty(n + 1, np.int32), adj=np.empty(max(1, 6 * m), np.int32))
    out = MeshPrepOut()
    for k, a in buf.items():
        setattr(out, k, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double if a.dtype == np.float64 else ctypes.c_int32)))
    check(lib().fcw_mesh_prep(dptr(V), n, iptr(F), m, M.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)), int(start),
                              int(device), ctypes.byref(out)))
    ns, ms = out.n_sub, out.m_sub
    return dict(vertices=buf["vertices"][:3 * ns].reshape(ns, 3), faces=buf["faces"][:3 * ms].reshape(ms, 3),
                sub_to_orig=buf["sub_to_orig"][:ns], orig_to_sub=buf["orig_to_sub"], face_areas=buf["face_areas"][:ms],
                face_L=buf["face_L"][:ms], vertex_areas=buf["vertex_areas"][:ns],
                boundary=buf["boundary"][:out.n_boundary], order=buf["order"][:ns].astype(np.int64),
                adj_ptr=buf["adj_ptr"][:ns + 1], adj=buf["adj"][:out.n_adj])
