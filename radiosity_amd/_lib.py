"""ctypes binding of libradhip.so (the C ABI in include/radhip.h).

The library is loaded from the package directory only; a missing library is
an error (there is no fallback implementation).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libradhip.so")

RAD_OK = 0
RAD_MODE_REFERENCE = 0
RAD_MODE_PHYSICAL = 1
RAD_SELECT_REFERENCE = 0
RAD_SELECT_POWER = 1
RAD_DENSE_ROWS = 0
RAD_DENSE_SHOOT = 1
RAD_FACES_QUADS = 0
RAD_FACES_TRIANGLES = 1


class RadError(RuntimeError):
    pass


class rad_config(C.Structure):
    _fields_ = [("hemicube", C.c_uint32), ("hemicubes", C.c_uint32), ("shoots", C.c_uint32),
                ("ocl_workitems_x", C.c_uint32), ("mode", C.c_uint32), ("select", C.c_uint32),
                ("small_tri_px", C.c_uint32), ("tile_px", C.c_uint32),
                ("graph_passes", C.c_uint32)]


class rad_stats(C.Structure):
    _fields_ = [("passes", C.c_uint64), ("shooters", C.c_uint64), ("residual", C.c_double),
                ("seconds", C.c_double), ("last_energy", C.c_float), ("reserved", C.c_uint32)]


class rad_cycle_stats(C.Structure):
    _fields_ = [("passes", C.c_uint64), ("shooters", C.c_uint64),
                ("last_energy", C.c_float * 3), ("stopped", C.c_uint32)]


class rad_scene(C.Structure):
    _fields_ = [("ntri", C.c_uint32), ("npatch", C.c_uint32), ("nquad", C.c_uint32),
                ("reserved", C.c_uint32),
                ("tri", C.POINTER(C.c_float)), ("tri_patch", C.POINTER(C.c_uint32)),
                ("frame", C.POINTER(C.c_float)), ("refl", C.POINTER(C.c_float)),
                ("area", C.POINTER(C.c_float)), ("emission", C.POINTER(C.c_float)),
                ("shot0", C.POINTER(C.c_float)), ("patch_quad", C.POINTER(C.c_uint32)),
                ("quad_nbr", C.POINTER(C.c_uint32))]


class rad_gather_opts(C.Structure):
    _fields_ = [("max_outer", C.c_uint32), ("inner", C.c_uint32),
                ("reserved", C.c_uint32 * 2), ("target", C.c_double)]


class rad_gather_stats(C.Structure):
    _fields_ = [("outer", C.c_uint32), ("sweeps", C.c_uint32), ("reserved", C.c_uint32 * 2),
                ("residual", C.c_double), ("seconds", C.c_double),
                ("history", C.c_double * 16), ("sweep_ms", C.c_double),
                ("resid_ms", C.c_double)]


class rad_tune_result(C.Structure):
    _fields_ = [("tile_px", C.c_uint32), ("small_tri_px", C.c_uint32), ("ncand", C.c_uint32),
                ("reserved", C.c_uint32), ("best_ms", C.c_double),
                ("cand_tile", C.c_uint32 * 18), ("cand_small", C.c_uint32 * 18),
                ("cand_ms", C.c_double * 18)]


_lib = None

_P = C.c_void_p
_u32 = C.c_uint32
_PF = C.POINTER(C.c_float)
_PU = C.POINTER(C.c_uint32)

# name -> (restype, argtypes); every symbol here is declared in include/*.h
SIGNATURES = {
    "rad_create": (C.c_int, [C.POINTER(rad_config), C.POINTER(_P)]),
    "rad_destroy": (None, [_P]),
    "rad_last_error": (C.c_char_p, [_P]),
    "rad_geometry": (C.c_int, [_P, _PU, _PU, _PU]),
    "rad_precompute_dff": (C.c_int, [_P, _PF]),
    "rad_scene_upload": (C.c_int, [_P, _PF, _PU, _u32, _u32, _PF, _PF, _PF]),
    "rad_set_state": (C.c_int, [_P, _PF, _PF, _PF]),
    "rad_get_state": (C.c_int, [_P, _PF, _PF]),
    "rad_select": (C.c_int, [_P, _u32, _PU, _PU]),
    "rad_render_itembuf": (C.c_int, [_P, _PU, _u32, _PU, _PF]),
    "rad_process_hemicube": (C.c_int, [_P, _u32, _PU, _PF]),
    "rad_process_itembuf": (C.c_int, [_P, _PU, _u32, _PU, _PF]),
    "rad_process_runs": (C.c_int, [_P, _u32, _u32, _PU, _PU, _PF, _PU]),
    "rad_process_runs_static": (C.c_int, [_P, _P, _u32, _u32, _P, _P, _P, _P]),
    "rad_shoot": (C.c_int, [_P, _PU, _u32]),
    "rad_step": (C.c_int, [_P, _u32]),
    "rad_prepare": (C.c_int, [_P]),
    "rad_run": (C.c_int, [_P, C.c_uint64, C.c_double, _u32, C.POINTER(rad_stats)]),
    "rad_run_cycle": (C.c_int, [_P, _u32, C.c_double, C.POINTER(rad_cycle_stats)]),
    "rad_shot_count": (C.c_int, [_P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "rad_set_deterministic": (C.c_int, [_P, C.c_int]),
    "rad_residual_fp64": (C.c_int, [_P, C.POINTER(C.c_double)]),
    "rad_finish": (C.c_int, [_P]),
    "rad_face_mvps": (C.c_int, [_P, _u32, _PF]),
    "rad_kernel_times": (C.c_int, [_P, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
    "rad_reduce_kernel_ms": (C.c_int, [_P, C.POINTER(C.c_double)]),
    "rad_scene_cornell": (C.c_int, [C.c_double, C.c_int, C.c_int, C.c_uint64,
                                    C.POINTER(C.POINTER(rad_scene))]),
    "rad_scene_boxes": (C.c_int, [_u32, C.c_float, C.c_float, C.c_float, _u32, _u32,
                                  C.c_uint64, C.POINTER(C.POINTER(rad_scene))]),
    "rad_scene_free": (None, [C.POINTER(rad_scene)]),
    "rad_scene_ingest": (C.c_int, [_P, _PF, _u32, _PF, _PF, _PF, C.c_double, C.c_int,
                                   C.POINTER(_u32), C.POINTER(_u32), C.POINTER(_u32)]),
    "rad_scene_download": (C.c_int, [_P, _PF, _PU, _PF, _PF, _PF]),
    "rad_set_quads": (C.c_int, [_P, _PU, _u32, _PU]),
    "rad_get_quads": (C.c_int, [_P, C.POINTER(_u32), _PU, _PU]),
    "rad_vertex_colors": (C.c_int, [_P, _PF]),
    "rad_render_view": (C.c_int, [_P, _PF, _PU]),
    "rad_obj_parse": (C.c_int, [C.c_char_p, C.POINTER(C.POINTER(C.c_float)), C.POINTER(_u32)]),
    "rad_obj_free": (None, [C.POINTER(C.c_float)]),
    "rad_cornell_faces": (C.c_int, [_PF, _PF, _PF, _PF, C.POINTER(_u32)]),
    "rad_scene_faces": (C.c_int, [_PF, _u32, _PF, _PF, _PF, C.c_double, C.c_int,
                                  C.POINTER(C.POINTER(rad_scene))]),
    "rad_debug_bins": (C.c_int, [_P, _PU, _PU, _PU]),
    "rad_debug_counters": (C.c_int, [_P, _PU, _PU]),
    "rad_debug_pool": (C.c_int, [_P, _u32, _PU]),
    "rad_autotune": (C.c_int, [_P, _u32, C.POINTER(rad_tune_result)]),
    "rad_pack_upload": (C.c_int, [_P, _u32, _PU, _PU, _PU, _PF, _PU, _u32, _u32, _PF, _PF, _PF]),
    "rad_pack_begin": (C.c_int, [_P, C.c_double]),
    "rad_pack_hemicubes": (C.c_int, [_P, _PU]),
    "rad_pack_status": (C.c_int, [_P, _PU, C.POINTER(C.c_double), _PU, _PU, _PU]),
    "rad_checkpoint_save": (C.c_int, [_P, C.c_char_p]),
    "rad_checkpoint_load": (C.c_int, [_P, C.c_char_p]),
    "rad_dense_build": (C.c_int, [_P, _u32]),
    "rad_dense_upload": (C.c_int, [_P, _PF]),
    "rad_dense_download": (C.c_int, [_P, _u32, _u32, _PF]),
    "rad_dense_apply": (C.c_int, [_P, _PF, _u32, C.c_int, _PF]),
    "rad_dense_release": (C.c_int, [_P]),
    "rad_gather_check_fp64": (C.c_int, [_P, _PF, _PF, _u32, C.POINTER(C.c_double)]),
    "rad_gather_solve": (C.c_int, [_P, _PF, _u32, C.POINTER(rad_gather_opts), _PF,
                                   C.POINTER(rad_gather_stats)]),
}


def lib():
    """Loads libradhip.so from the package directory (raises if absent)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RadError(f"{LIBPATH} is missing: run radiosity_amd.build.build() first")
        L = C.CDLL(LIBPATH)
        for name, (res, args) in SIGNATURES.items():
            if not hasattr(L, name):
                continue
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def fptr(a):
    return a.ctypes.data_as(_PF) if a is not None else None


def uptr(a):
    return a.ctypes.data_as(_PU) if a is not None else None


def _arr(ptr, n, dtype, width=1):
    if n == 0:
        return np.zeros((0, width) if width > 1 else 0, dtype=dtype)
    a = np.ctypeslib.as_array(ptr, shape=(n * width,)).astype(dtype, copy=True)
    return a.reshape(n, width) if width > 1 else a


class Scene:
    """Host copy of a generated scene (numpy arrays)."""

    def __init__(self, tri, tri_patch, frame, refl, area, emission, shot0, patch_quad=None,
                 quad_nbr=None):
        self.tri = np.ascontiguousarray(tri, dtype=np.float32).reshape(-1, 9)
        self.tri_patch = np.ascontiguousarray(tri_patch, dtype=np.uint32)
        self.frame = np.ascontiguousarray(frame, dtype=np.float32).reshape(-1, 9)
        self.refl = np.ascontiguousarray(refl, dtype=np.float32).reshape(-1, 3)
        self.area = np.ascontiguousarray(area, dtype=np.float32)
        self.emission = np.ascontiguousarray(emission, dtype=np.float32).reshape(-1, 3)
        self.shot0 = np.ascontiguousarray(shot0, dtype=np.float32).reshape(-1, 3)
        self.patch_quad = patch_quad
        self.quad_nbr = quad_nbr

    @property
    def ntri(self):
        return self.tri.shape[0]

    @property
    def npatch(self):
        return self.frame.shape[0]

    @staticmethod
    def _from_c(p):
        s = p.contents
        out = Scene(_arr(s.tri, s.ntri, np.float32, 9), _arr(s.tri_patch, s.ntri, np.uint32),
                    _arr(s.frame, s.npatch, np.float32, 9), _arr(s.refl, s.npatch, np.float32, 3),
                    _arr(s.area, s.npatch, np.float32), _arr(s.emission, s.npatch, np.float32, 3),
                    _arr(s.shot0, s.npatch, np.float32, 3), _arr(s.patch_quad, s.npatch, np.uint32),
                    _arr(s.quad_nbr, s.nquad, np.uint32, 8))
        out.nquad = s.nquad
        lib().rad_scene_free(p)
        return out

    @staticmethod
    def cornell(max_area, triangle_patches=True, physical=True, seed=1):
        p = C.POINTER(rad_scene)()
        rc = lib().rad_scene_cornell(float(max_area), int(triangle_patches), int(physical),
                                     int(seed), C.byref(p))
        if rc != RAD_OK:
            raise RadError(f"rad_scene_cornell failed ({rc})")
        return Scene._from_c(p)

    @staticmethod
    def faces(quads, max_area, mode=0, color=None, emission=None, shot0=None):
        """Scene from source faces (n x 12: A, B, C, D; rad_scene_faces)."""
        q = np.ascontiguousarray(quads, np.float32).reshape(-1, 12)
        opt = [None if a is None else np.ascontiguousarray(a, np.float32).reshape(-1, 3)
               for a in (color, emission, shot0)]
        p = C.POINTER(rad_scene)()
        rc = lib().rad_scene_faces(fptr(q), q.shape[0], *[fptr(a) if a is not None else None
                                                          for a in opt],
                                   float(max_area), int(mode), C.byref(p))
        if rc != RAD_OK:
            raise RadError(f"rad_scene_faces failed ({rc})")
        return Scene._from_c(p)

    @staticmethod
    def boxes(nboxes=400, room=(20.0, 5.0, 20.0), target_tris=200000, nlights=16, seed=3):
        p = C.POINTER(rad_scene)()
        rc = lib().rad_scene_boxes(int(nboxes), float(room[0]), float(room[1]), float(room[2]),
                                   int(target_tris), int(nlights), int(seed), C.byref(p))
        if rc != RAD_OK:
            raise RadError(f"rad_scene_boxes failed ({rc})")
        return Scene._from_c(p)


def obj_parse(path):
    """Source faces of a Wavefront OBJ file (rad_obj_parse): (faces n x 12,
    complete) where complete is False when a bad vertex line ended the parse."""
    q = C.POINTER(C.c_float)()
    n = _u32(0)
    rc = lib().rad_obj_parse(os.fsencode(path), C.byref(q), C.byref(n))
    if rc != RAD_OK and not q:
        raise RadError(f"rad_obj_parse failed ({rc}): {path}")
    out = _arr(q, n.value, np.float32, 12)
    lib().rad_obj_free(q)
    return out, rc == RAD_OK


def cornell_faces():
    """The 17 Cornell source quads with colour, radiosity, illumination."""
    n = _u32(0)
    lib().rad_cornell_faces(None, None, None, None, C.byref(n))
    q = np.zeros((n.value, 12), np.float32)
    c, e, s = (np.zeros((n.value, 3), np.float32) for _ in range(3))
    lib().rad_cornell_faces(fptr(q), fptr(c), fptr(e), fptr(s), C.byref(n))
    return q, c, e, s


class Solver:
    """One rad_ctx: scene, state and the shooting passes on the current GPU."""

    def __init__(self, hemicube=16, hemicubes=1, mode=RAD_MODE_PHYSICAL, select=RAD_SELECT_POWER,
                 shoots=500, small_tri_px=0, tile_px=0, graph_passes=0, workitems_x=4):
        L = lib()
        cfg = rad_config(hemicube, hemicubes, shoots, workitems_x, mode, select, small_tri_px,
                         tile_px, graph_passes)
        self._ctx = C.c_void_p()
        rc = L.rad_create(C.byref(cfg), C.byref(self._ctx))
        if rc != RAD_OK:
            raise RadError(f"rad_create failed ({rc})")
        self.W, self.K = hemicube, hemicubes
        tw, th, res = C.c_uint32(), C.c_uint32(), C.c_uint32()
        L.rad_geometry(self._ctx, C.byref(tw), C.byref(th), C.byref(res))
        self.tex_w, self.tex_h, self.tex_res = tw.value, th.value, res.value
        self.npatch = 0

    def _chk(self, rc, what):
        if rc != RAD_OK:
            msg = lib().rad_last_error(self._ctx)
            raise RadError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def close(self):
        if self._ctx:
            lib().rad_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def dff(self):
        out = np.zeros(self.tex_res, np.float32)
        self._chk(lib().rad_precompute_dff(self._ctx, fptr(out)), "rad_precompute_dff")
        return out

    def upload(self, sc, state=True):
        self._chk(lib().rad_scene_upload(self._ctx, fptr(sc.tri), uptr(sc.tri_patch), sc.ntri,
                                         sc.npatch, fptr(sc.frame), fptr(sc.refl), fptr(sc.area)),
                  "rad_scene_upload")
        self.npatch = sc.npatch
        self.ntri = sc.ntri
        if state:
            self.set_state(sc.emission, sc.shot0, sc.emission)

    def ingest(self, quads, max_area, mode=0, color=None, emission=None, shot0=None):
        """Builds the scene on the device from source faces (rad_scene_ingest);
        returns (ntri, npatch, nquad)."""
        q = np.ascontiguousarray(quads, np.float32).reshape(-1, 12)
        opt = [None if a is None else np.ascontiguousarray(a, np.float32).reshape(-1, 3)
               for a in (color, emission, shot0)]
        nt, np_, nq = _u32(0), _u32(0), _u32(0)
        self._chk(lib().rad_scene_ingest(self._ctx, fptr(q), q.shape[0], fptr(opt[0]),
                                         fptr(opt[1]), fptr(opt[2]), float(max_area), int(mode),
                                         C.byref(nt), C.byref(np_), C.byref(nq)),
                  "rad_scene_ingest")
        self.npatch = np_.value
        self.ntri = nt.value
        return nt.value, np_.value, nq.value

    def download(self):
        """The device scene as a Scene (state arrays from rad_get_state)."""
        nt, np_ = self.ntri, self.npatch
        tri = np.zeros((nt, 9), np.float32)
        tp = np.zeros(nt, np.uint32)
        fr = np.zeros((np_, 9), np.float32)
        rf = np.zeros((np_, 3), np.float32)
        ar = np.zeros(np_, np.float32)
        self._chk(lib().rad_scene_download(self._ctx, fptr(tri), uptr(tp), fptr(fr), fptr(rf),
                                           fptr(ar)), "rad_scene_download")
        u, sh = self.state()
        pq, nb = self.quads()
        return Scene(tri, tp, fr, rf, ar, u, sh, pq, nb)

    def set_quads(self, patch_quad, quad_nbr):
        pq = np.ascontiguousarray(patch_quad, np.uint32)
        nb = np.ascontiguousarray(quad_nbr, np.uint32).reshape(-1, 8)
        self._chk(lib().rad_set_quads(self._ctx, uptr(pq), nb.shape[0], uptr(nb)),
                  "rad_set_quads")

    def quads(self):
        """(patch_quad, quad_nbr) of the device scene, or (None, None)."""
        n = _u32(0)
        if lib().rad_get_quads(self._ctx, C.byref(n), None, None) != RAD_OK:
            return None, None
        pq = np.zeros(self.npatch, np.uint32)
        nb = np.zeros((n.value, 8), np.uint32)
        self._chk(lib().rad_get_quads(self._ctx, C.byref(n), uptr(pq), uptr(nb)), "rad_get_quads")
        return pq, nb

    def vertex_colors(self):
        n = _u32(0)
        self._chk(lib().rad_get_quads(self._ctx, C.byref(n), None, None), "rad_get_quads")
        out = np.zeros((n.value, 4, 3), np.float32)
        self._chk(lib().rad_vertex_colors(self._ctx, fptr(out)), "rad_vertex_colors")
        return out

    def render_view(self, frame):
        fr = np.ascontiguousarray(frame, np.float32).reshape(9)
        ids = np.zeros(self.tex_res, np.uint32)
        self._chk(lib().rad_render_view(self._ctx, fptr(fr), uptr(ids)), "rad_render_view")
        return ids

    def set_state(self, unshot, shot, emission=None):
        u = np.ascontiguousarray(unshot, np.float32)
        s = np.ascontiguousarray(shot, np.float32)
        e = None if emission is None else np.ascontiguousarray(emission, np.float32)
        self._chk(lib().rad_set_state(self._ctx, fptr(u), fptr(s), fptr(e)), "rad_set_state")

    def state(self):
        u = np.zeros((self.npatch, 3), np.float32)
        s = np.zeros((self.npatch, 3), np.float32)
        self._chk(lib().rad_get_state(self._ctx, fptr(u), fptr(s)), "rad_get_state")
        return u, s

    def select(self, k):
        ids = np.zeros(k, np.uint32)
        n = C.c_uint32()
        self._chk(lib().rad_select(self._ctx, k, uptr(ids), C.byref(n)), "rad_select")
        return ids, n.value

    def render(self, shooters, depth=False):
        sh = np.ascontiguousarray(shooters, np.uint32)
        k = len(sh)
        ids = np.zeros(k * self.tex_res, np.uint32)
        z = np.zeros(k * self.tex_res, np.float32) if depth else None
        self._chk(lib().rad_render_itembuf(self._ctx, uptr(sh), k, uptr(ids), fptr(z)),
                  "rad_render_itembuf")
        return (ids, z) if depth else ids

    def process(self, k):
        hits = np.zeros(k * self.npatch, np.uint32)
        ff = np.zeros(k * self.npatch, np.float32)
        self._chk(lib().rad_process_hemicube(self._ctx, k, uptr(hits), fptr(ff)),
                  "rad_process_hemicube")
        return hits.reshape(k, self.npatch), ff.reshape(k, self.npatch)

    def ff_rows(self, shooters):
        """FP32 form-factor rows (K x npatch) of the given shooters: the item
        buffers stay on the device and are reduced there."""
        sh = np.ascontiguousarray(shooters, np.uint32)
        k = len(sh)
        self._chk(lib().rad_render_itembuf(self._ctx, uptr(sh), k, None, None),
                  "rad_render_itembuf")
        ff = np.zeros(k * self.npatch, np.float32)
        self._chk(lib().rad_process_hemicube(self._ctx, k, None, fptr(ff)), "rad_process_hemicube")
        return ff.reshape(k, self.npatch)

    def process_runs(self, k):
        """The run-form triplets (hemicubes, ids, energies) of the last k
        rendered hemicubes, padding included (rad_process_runs)."""
        n = C.c_uint32()
        rc = lib().rad_process_runs(self._ctx, k, 0, None, None, None, C.byref(n))
        if rc not in (RAD_OK, -5):
            self._chk(rc, "rad_process_runs")
        h = np.zeros(n.value, np.uint32)
        i = np.zeros(n.value, np.uint32)
        e = np.zeros(n.value, np.float32)
        self._chk(lib().rad_process_runs(self._ctx, k, n.value, uptr(h), uptr(i), fptr(e),
                                         C.byref(n)), "rad_process_runs")
        return h, i, e

    def process_itembuf(self, item, k):
        it = np.ascontiguousarray(item, np.uint32)
        hits = np.zeros(k * self.npatch, np.uint32)
        ff = np.zeros(k * self.npatch, np.float32)
        self._chk(lib().rad_process_itembuf(self._ctx, uptr(it), k, uptr(hits), fptr(ff)),
                  "rad_process_itembuf")
        return hits.reshape(k, self.npatch), ff.reshape(k, self.npatch)

    def shoot(self, shooters):
        sh = np.ascontiguousarray(shooters, np.uint32)
        self._chk(lib().rad_shoot(self._ctx, uptr(sh), len(sh)), "rad_shoot")

    def step(self, passes=1):
        self._chk(lib().rad_step(self._ctx, passes), "rad_step")

    def prepare(self):
        """Captures the pass graph now (rad_prepare); runs no pass."""
        self._chk(lib().rad_prepare(self._ctx), "rad_prepare")

    def finish(self):
        self._chk(lib().rad_finish(self._ctx), "rad_finish")

    def run(self, max_passes, target, check_every=0):
        st = rad_stats()
        self._chk(lib().rad_run(self._ctx, max_passes, target, check_every, C.byref(st)), "rad_run")
        return st

    def run_cycle(self, passes, stop_energy=0.1):
        """One OnIdle cycle with the reference's per-pass stop test
        (rad_run_cycle); returns the rad_cycle_stats."""
        st = rad_cycle_stats()
        self._chk(lib().rad_run_cycle(self._ctx, passes, stop_energy, C.byref(st)),
                  "rad_run_cycle")
        return st

    def set_deterministic(self, on=True):
        """Fixed-point F accumulation: bit-identical passes run to run."""
        self._chk(lib().rad_set_deterministic(self._ctx, int(bool(on))), "rad_set_deterministic")

    def shot_count(self):
        """(passes that shot, shooters shot) counted on the device."""
        p, n = C.c_uint64(), C.c_uint64()
        self._chk(lib().rad_shot_count(self._ctx, C.byref(p), C.byref(n)), "rad_shot_count")
        return p.value, n.value

    def residual(self):
        r = C.c_double()
        self._chk(lib().rad_residual_fp64(self._ctx, C.byref(r)), "rad_residual_fp64")
        return r.value

    def face_mvps(self, patch):
        out = np.zeros(80, np.float32)
        self._chk(lib().rad_face_mvps(self._ctx, patch, fptr(out)), "rad_face_mvps")
        return out.reshape(5, 16)

    def debug_bins(self):
        nt, tp = C.c_uint32(), C.c_uint32()
        self._chk(lib().rad_debug_bins(self._ctx, None, C.byref(nt), C.byref(tp)), "rad_debug_bins")
        out = np.zeros(self.K * nt.value, np.uint32)
        self._chk(lib().rad_debug_bins(self._ctx, uptr(out), C.byref(nt), C.byref(tp)),
                  "rad_debug_bins")
        return out.reshape(self.K, nt.value), tp.value

    def debug_counters(self):
        """The rasterizer's pass counters of the last pass (dict) and the
        spill, bin-entry and clip-queue capacities."""
        out = np.zeros(8, np.uint32)
        caps = np.zeros(3, np.uint32)
        self._chk(lib().rad_debug_counters(self._ctx, uptr(out), uptr(caps)), "rad_debug_counters")
        names = ["unused0", "unused1", "overflow", "entries", "unknown", "spill", "clip", "sticky"]
        d = {k: int(v) for k, v in zip(names, out)}
        d.update(spill_cap=int(caps[0]), entry_cap=int(caps[1]), clip_cap=int(caps[2]))
        return d

    def debug_pool(self, cap=0):
        """Candidate pool of the POWER selection in captured passes: cap != 0
        sets the most keys a selection takes from it (else it scans every
        patch); returns the selections taken from the pool, the full scans,
        the times its bound was set and whether it is set now."""
        out = np.zeros(4, np.uint32)
        self._chk(lib().rad_debug_pool(self._ctx, int(cap), uptr(out)), "rad_debug_pool")
        return dict(hits=int(out[0]), scans=int(out[1]), arms=int(out[2]), armed=int(out[3]))

    def kernel_times(self, enable=-1):
        out = (C.c_double * 6)()
        n = C.c_uint64()
        self._chk(lib().rad_kernel_times(self._ctx, enable, out, C.byref(n)), "rad_kernel_times")
        return list(out), n.value

    def reduce_kernel_ms(self):
        """Device ms of the last standalone ProcessHemicube launched while
        kernel timing was enabled."""
        ms = C.c_double()
        self._chk(lib().rad_reduce_kernel_ms(self._ctx, C.byref(ms)), "rad_reduce_kernel_ms")
        return ms.value

    def checkpoint_save(self, path):
        self._chk(lib().rad_checkpoint_save(self._ctx, os.fsencode(path)), "rad_checkpoint_save")

    def checkpoint_load(self, path):
        self._chk(lib().rad_checkpoint_load(self._ctx, os.fsencode(path)), "rad_checkpoint_load")

    def autotune(self, passes=3):
        """Picks the fastest (tile_px, small_tri_px); returns (tile, small,
        ms per render pass, {(tile, small): ms})."""
        r = rad_tune_result()
        self._chk(lib().rad_autotune(self._ctx, passes, C.byref(r)), "rad_autotune")
        cands = {(r.cand_tile[k], r.cand_small[k]): r.cand_ms[k] for k in range(r.ncand)}
        return r.tile_px, r.small_tri_px, r.best_ms, cands

    # ------------------------------------------------------------ packs
    def pack_upload(self, scenes, slots):
        """Concatenates `scenes` into one pack; slots[s] shooters per pass for
        scene s.  Sets the state to each scene's emission / initial shot."""
        tb = np.zeros(len(scenes) + 1, np.uint32)
        pb = np.zeros(len(scenes) + 1, np.uint32)
        tb[1:] = np.cumsum([sc.ntri for sc in scenes])
        pb[1:] = np.cumsum([sc.npatch for sc in scenes])
        tri = np.ascontiguousarray(np.concatenate([sc.tri.reshape(-1) for sc in scenes]),
                                   np.float32)
        tp = np.ascontiguousarray(np.concatenate(
            [sc.tri_patch + pb[k] for k, sc in enumerate(scenes)]), np.uint32)
        cat = lambda name: np.ascontiguousarray(  # noqa: E731
            np.concatenate([getattr(sc, name) for sc in scenes]), np.float32)
        frame, refl, area = cat("frame"), cat("refl"), cat("area")
        sl = np.ascontiguousarray(slots, np.uint32)
        self._chk(lib().rad_pack_upload(self._ctx, len(scenes), uptr(tb), uptr(pb), uptr(sl),
                                        fptr(tri), uptr(tp), int(tb[-1]), int(pb[-1]),
                                        fptr(frame), fptr(refl), fptr(area)), "rad_pack_upload")
        self.npatch = int(pb[-1])
        self.pack_offsets = (tb, pb)
        em = cat("emission")
        self.set_state(em, cat("shot0"), em)

    def pack_hemicubes(self, scene_w):
        """Per-scene hemicube sides of the loaded pack (rad_pack_hemicubes)."""
        w = np.ascontiguousarray(scene_w, np.uint32)
        self._chk(lib().rad_pack_hemicubes(self._ctx, uptr(w)), "rad_pack_hemicubes")

    def pack_begin(self, target):
        self._chk(lib().rad_pack_begin(self._ctx, target), "rad_pack_begin")

    def pack_status(self):
        """(done_pass per scene, 0xFFFFFFFF while active; residual per scene;
        passes; scenes remaining).  Shooters per scene: self.pack_shooters."""
        n = len(self.pack_offsets[1]) - 1
        done = np.zeros(n, np.uint32)
        res = np.zeros(n, np.float64)
        shot = np.zeros(n, np.uint32)
        passes, rem = C.c_uint32(), C.c_uint32()
        self._chk(lib().rad_pack_status(self._ctx, uptr(done),
                                        res.ctypes.data_as(C.POINTER(C.c_double)), uptr(shot),
                                        C.byref(passes), C.byref(rem)), "rad_pack_status")
        self.pack_shooters = shot
        return done, res, passes.value, rem.value

    # ------------------------------------------------------ dense / gather
    def dense_build(self, layout=RAD_DENSE_SHOOT):
        self._chk(lib().rad_dense_build(self._ctx, layout), "rad_dense_build")

    def dense_upload(self, m):
        m = np.ascontiguousarray(m, np.float32)
        assert m.shape == (self.npatch, self.npatch)
        self._chk(lib().rad_dense_upload(self._ctx, fptr(m)), "rad_dense_upload")

    def dense_download(self, row0=0, nrows=None):
        if nrows is None:
            nrows = self.npatch - row0
        out = np.zeros((nrows, self.npatch), np.float32)
        self._chk(lib().rad_dense_download(self._ctx, row0, nrows, fptr(out)), "rad_dense_download")
        return out

    def dense_apply(self, x, split=False):
        x = np.ascontiguousarray(np.atleast_2d(x), np.float32)
        y = np.zeros_like(x)
        self._chk(lib().rad_dense_apply(self._ctx, fptr(x), x.shape[0], int(split), fptr(y)),
                  "rad_dense_apply")
        return y

    def gather_check_fp64(self, B, E=None):
        """(FP64 residual, contraction q) of a solution B (R x npatch) against
        freshly rendered FP32 rows in the shooting layout; the relative
        area-weighted error bound is residual / (1 - q)."""
        B = np.ascontiguousarray(np.atleast_2d(B), np.float32)
        Ep = None if E is None else np.ascontiguousarray(np.atleast_2d(E), np.float32)
        out = (C.c_double * 2)()
        self._chk(lib().rad_gather_check_fp64(self._ctx, fptr(B), fptr(Ep), B.shape[0], out),
                  "rad_gather_check_fp64")
        return out[0], out[1]

    def dense_release(self):
        self._chk(lib().rad_dense_release(self._ctx), "rad_dense_release")

    def gather(self, rhs=None, target=1e-3, max_outer=0, inner=0):
        """Returns (B, stats): B is R x npatch (R = 3 and the scene's emission
        when rhs is None)."""
        if rhs is None:
            R, ptr = 3, None
        else:
            rhs = np.ascontiguousarray(np.atleast_2d(rhs), np.float32)
            R, ptr = rhs.shape[0], fptr(rhs)
        out = np.zeros((R, self.npatch), np.float32)
        o = rad_gather_opts(max_outer, inner, (C.c_uint32 * 2)(), target)
        st = rad_gather_stats()
        self._chk(lib().rad_gather_solve(self._ctx, ptr, R, C.byref(o), fptr(out), C.byref(st)),
                  "rad_gather_solve")
        return out, st
