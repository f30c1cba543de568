"""ctypes binding of liboracle.so (oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the
benchmark's CPU-baseline leg as the checker.  The product never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "liboracle.so")

_PF = C.POINTER(C.c_float)
_PD = C.POINTER(C.c_double)
_PU = C.POINTER(C.c_uint32)
_PI = C.POINTER(C.c_int32)
_u32 = C.c_uint32

_SIG = {
    "orc_geometry": (None, [_u32, _PU, _PU, _PU]),
    "orc_dff_top_side": (None, [_u32, _PF, _PF]),
    "orc_dff_sum_test": (C.c_float, [_u32]),
    "orc_dff_cross": (None, [_u32, _PF]),
    "orc_quad_frame": (None, [_PF, _PF]),
    "orc_face_mvp": (None, [_PF, C.c_int, _PF]),
    "orc_face_viewport": (None, [_u32, C.c_int, _PI, _PI]),
    "orc_render_hemicube": (C.c_uint64, [_u32, _PF, _PU, _u32, _PF, _PU, _PF]),
    "orc_render_hemicube_variant": (C.c_uint64, [_u32, _PF, _PU, _u32, _PF, _PU, C.c_int]),
    "orc_render_hemicube_mt": (C.c_uint64, [_u32, _PF, _PU, _u32, _PF, _PU, C.c_int]),
    "orc_plain_pass_mt": (_u32, [_u32, _u32, _u32, _PF, _PU, _u32, _PF, _PF, _PF, _u32, _PF,
                                 _PF, _PF, _PU, _PF, C.c_int]),
    "orc_process_hemicube_runs": (C.c_uint64, [_u32, _u32, _u32, _PU, _PF, _u32, _PF,
                                               C.POINTER(C.c_uint64)]),
    "orc_hemicube_histogram": (None, [_u32, _u32, _PU, _PF, _u32, _PU, _PD]),
    "orc_select_reference": (_u32, [_PF, _u32, _u32, _PU]),
    "orc_select_argmax_power": (_u32, [_PF, _PF, _u32]),
    "orc_shoot": (None, [_u32, _u32, _PU, _u32, _PD, _PF, _PF, _PD, _PD]),
    "orc_residual": (C.c_double, [_PD, _PF, _PF, _u32]),
    "orc_plain_pass": (_u32, [_u32, _u32, _u32, _PF, _PU, _u32, _PF, _PF, _PF, _u32, _PF,
                              _PF, _PF, _PU, _PF]),
    "orc_gather_jacobi": (_u32, [_u32, _u32, _PF, _PF, _PF, _PD, C.c_double, _u32, _PD]),
    "orc_select_topk_power": (_u32, [_PF, _PF, _u32, _u32, _PU]),
    "orc_process_hemicube_triplets": (C.c_uint64, [_u32, _u32, _u32, _PU, _PF, C.c_uint64, _PU,
                                                   _PU, _PF]),
    "orc_pass_k": (_u32, [_u32, _u32, _u32, _u32, _PF, _PU, _u32, _PF, _PF, _PF, _u32, _PF,
                          _PF, _PF, _PU, _PF, _PU, _PF]),
}

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            build()
        L = C.CDLL(LIBPATH)
        for n, (r, a) in _SIG.items():
            f = getattr(L, n)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def _f(a):
    return None if a is None else a.ctypes.data_as(_PF)


def _d(a):
    return None if a is None else a.ctypes.data_as(_PD)


def _u(a):
    return None if a is None else a.ctypes.data_as(_PU)


def geometry(W):
    a, b, c = C.c_uint32(), C.c_uint32(), C.c_uint32()
    lib().orc_geometry(W, C.byref(a), C.byref(b), C.byref(c))
    return a.value, b.value, c.value


def dff_top_side(W):
    top = np.zeros(W * W, np.float32)
    side = np.zeros(W * W // 2, np.float32)
    lib().orc_dff_top_side(W, _f(top), _f(side))
    return top, side


def dff_sum_test(W):
    return lib().orc_dff_sum_test(W)


def dff_cross(W):
    out = np.zeros(geometry(W)[2], np.float32)
    lib().orc_dff_cross(W, _f(out))
    return out


def quad_frame(v4x3):
    v = np.ascontiguousarray(v4x3, np.float32).reshape(12)
    fr = np.zeros(9, np.float32)
    lib().orc_quad_frame(_f(v), _f(fr))
    return fr


def face_mvp(frame, face):
    fr = np.ascontiguousarray(frame, np.float32).reshape(9)
    m = np.zeros(16, np.float32)
    lib().orc_face_mvp(_f(fr), face, _f(m))
    return m


def face_viewport(W, face):
    vp = np.zeros(4, np.int32)
    sc = np.zeros(4, np.int32)
    lib().orc_face_viewport(W, face, vp.ctypes.data_as(_PI), sc.ctypes.data_as(_PI))
    return vp, sc


def render_hemicube(W, tri, tri_patch, frame, depth=False):
    tri = np.ascontiguousarray(tri, np.float32).reshape(-1, 9)
    tp = np.ascontiguousarray(tri_patch, np.uint32)
    fr = np.ascontiguousarray(frame, np.float32).reshape(9)
    res = geometry(W)[2]
    item = np.zeros(res, np.uint32)
    z = np.zeros(res, np.float32) if depth else None
    frags = lib().orc_render_hemicube(W, _f(tri), _u(tp), tri.shape[0], _f(fr), _u(item), _f(z))
    return (item, z, frags) if depth else item


VAR_DEPTH24, VAR_FMA, VAR_DIVIDE = 1, 2, 4


def render_hemicube_variant(W, tri, tri_patch, frame, flags):
    """render_hemicube with the arithmetic of another implementation
    (VAR_* flags): 24-bit depth, FMA transform, divide by w."""
    tri = np.ascontiguousarray(tri, np.float32).reshape(-1, 9)
    tp = np.ascontiguousarray(tri_patch, np.uint32)
    fr = np.ascontiguousarray(frame, np.float32).reshape(9)
    item = np.zeros(geometry(W)[2], np.uint32)
    lib().orc_render_hemicube_variant(W, _f(tri), _u(tp), tri.shape[0], _f(fr), _u(item), flags)
    return item


def render_hemicube_mt(W, tri, tri_patch, frame, nthreads):
    """render_hemicube on nthreads OpenMP threads (same item buffer);
    returns (item, fragments)."""
    tri = np.ascontiguousarray(tri, np.float32).reshape(-1, 9)
    tp = np.ascontiguousarray(tri_patch, np.uint32)
    fr = np.ascontiguousarray(frame, np.float32).reshape(9)
    item = np.zeros(geometry(W)[2], np.uint32)
    frags = lib().orc_render_hemicube_mt(W, _f(tri), _u(tp), tri.shape[0], _f(fr), _u(item),
                                         int(nthreads))
    return item, frags


def process_runs(W, K, item, dff, npatch, workitems_x=4):
    item = np.ascontiguousarray(item, np.uint32)
    ff = np.zeros(K * npatch, np.float32)
    unk = C.c_uint64()
    n = lib().orc_process_hemicube_runs(W, K, workitems_x, _u(item), _f(dff), npatch, _f(ff),
                                        C.byref(unk))
    return ff.reshape(K, npatch), n, unk.value


def histogram(W, K, item, dff, npatch):
    item = np.ascontiguousarray(item, np.uint32)
    hits = np.zeros(K * npatch, np.uint32)
    ff = np.zeros(K * npatch, np.float64)
    lib().orc_hemicube_histogram(W, K, _u(item), _f(dff), npatch, _u(hits), _d(ff))
    return hits.reshape(K, npatch), ff.reshape(K, npatch)


def select_reference(unshot, count):
    u = np.ascontiguousarray(unshot, np.float32)
    ids = np.zeros(count, np.uint32)
    n = lib().orc_select_reference(_f(u), u.shape[0], count, _u(ids))
    return ids, n


def select_argmax_power(unshot, area):
    u = np.ascontiguousarray(unshot, np.float32)
    a = np.ascontiguousarray(area, np.float32)
    return lib().orc_select_argmax_power(_f(u), _f(a), u.shape[0])


def shoot(mode, shooters, ff64, refl, area, unshot64, shot64):
    sh = np.ascontiguousarray(shooters, np.uint32)
    ff = np.ascontiguousarray(ff64, np.float64)
    lib().orc_shoot(mode, unshot64.shape[0], _u(sh), len(sh), _d(ff),
                    _f(np.ascontiguousarray(refl, np.float32)),
                    _f(np.ascontiguousarray(area, np.float32)), _d(unshot64), _d(shot64))


def residual(unshot64, emission, area):
    return lib().orc_residual(_d(np.ascontiguousarray(unshot64, np.float64)),
                              _f(np.ascontiguousarray(emission, np.float32)),
                              _f(np.ascontiguousarray(area, np.float32)), len(area))


def plain_pass(W, mode, rule, sc, dff, unshot, shot, item_scratch=None, ff_scratch=None):
    if item_scratch is None:
        item_scratch = np.zeros(geometry(W)[2], np.uint32)
    if ff_scratch is None:
        ff_scratch = np.zeros(sc.npatch, np.float32)
    return lib().orc_plain_pass(W, mode, rule, _f(sc.tri), _u(sc.tri_patch), sc.ntri,
                                _f(sc.frame), _f(sc.refl), _f(sc.area), sc.npatch, _f(dff),
                                _f(unshot), _f(shot), _u(item_scratch), _f(ff_scratch))


def plain_pass_mt(W, mode, rule, sc, dff, unshot, shot, nthreads, item_scratch=None,
                  ff_scratch=None):
    """plain_pass with the render and the shooting loop on nthreads threads."""
    if item_scratch is None:
        item_scratch = np.zeros(geometry(W)[2], np.uint32)
    if ff_scratch is None:
        ff_scratch = np.zeros(sc.npatch, np.float32)
    return lib().orc_plain_pass_mt(W, mode, rule, _f(sc.tri), _u(sc.tri_patch), sc.ntri,
                                   _f(sc.frame), _f(sc.refl), _f(sc.area), sc.npatch, _f(dff),
                                   _f(unshot), _f(shot), _u(item_scratch), _f(ff_scratch),
                                   int(nthreads))


def gather_jacobi(M, refl, area, E, tol=1e-13, max_it=100000):
    """FP64 Jacobi fixed point of B = E + diag(rho_{r%3}) M B; E, B: R x n."""
    M = np.ascontiguousarray(M, np.float32)
    n = M.shape[0]
    E = np.ascontiguousarray(np.atleast_2d(E), np.float64)
    B = np.zeros_like(E)
    it = lib().orc_gather_jacobi(n, E.shape[0], _f(M), _f(np.ascontiguousarray(refl, np.float32)),
                                 _f(np.ascontiguousarray(area, np.float32)), _d(E), tol, max_it,
                                 _d(B))
    return B, it


def select_topk_power(unshot, area, count):
    u = np.ascontiguousarray(unshot, np.float32)
    a = np.ascontiguousarray(area, np.float32)
    ids = np.zeros(count, np.uint32)
    n = lib().orc_select_topk_power(_f(u), _f(a), u.shape[0], count, _u(ids))
    return ids, n


def pass_k(W, mode, rule, K, sc, dff, unshot, shot):
    """One K-hemicube pass of the OnIdle loop on float state (updated in
    place); returns (shooter ids, lastEnergy as 3 floats)."""
    item = np.zeros(geometry(W)[2], np.uint32)
    ff = np.zeros(sc.npatch, np.float32)
    ids = np.zeros(K, np.uint32)
    le = np.zeros(3, np.float32)
    n = lib().orc_pass_k(W, mode, rule, K, _f(sc.tri), _u(sc.tri_patch), sc.ntri, _f(sc.frame),
                         _f(sc.refl), _f(sc.area), sc.npatch, _f(dff), _f(unshot), _f(shot),
                         _u(item), _f(ff), _u(ids), _f(le))
    return ids[:n], le


def process_triplets(W, K, item, dff, workitems_x=4):
    """The ProcessHemicube kernel's triplets in work-item order, padding
    included: (hemicubes, ids, energies)."""
    item = np.ascontiguousarray(item, np.uint32)
    cap = int(item.size) + 4 * K * geometry(W)[1] * workitems_x
    h = np.zeros(cap, np.uint32)
    i = np.zeros(cap, np.uint32)
    e = np.zeros(cap, np.float32)
    n = lib().orc_process_hemicube_triplets(W, K, workitems_x, _u(item), _f(dff), cap, _u(h),
                                            _u(i), _f(e))
    return h[:n].copy(), i[:n].copy(), e[:n].copy()
