"""Python binding of the C-ABI in include/skity_hip.h (libskity_hip.so).

The rasteriser runs entirely in the HIP library; this module only loads it,
declares the entry points and wraps the opaque handles.  It raises if the
library is missing instead of falling back to anything else.
"""
import ctypes

import numpy as np

from . import abi

_lib = None


def lib():
    global _lib
    if _lib is None:
        L = abi.load("libskity_hip.so")
        c_ctx = ctypes.c_void_p
        c_surf = ctypes.c_void_p
        SP = ctypes.POINTER(abi.Scene)
        L.hipr_context_create.argtypes = [ctypes.c_int32, ctypes.POINTER(c_ctx)]
        L.hipr_context_create.restype = ctypes.c_int
        L.hipr_context_destroy.argtypes = [c_ctx]
        L.hipr_context_destroy.restype = None
        L.hipr_surface_create.argtypes = [c_ctx, ctypes.c_int32, ctypes.c_int32,
                                          ctypes.POINTER(c_surf)]
        L.hipr_surface_create.restype = ctypes.c_int
        L.hipr_surface_create_ex.argtypes = [c_ctx, ctypes.c_int32, ctypes.c_int32,
                                             ctypes.c_int32, ctypes.c_int32,
                                             ctypes.POINTER(c_surf)]
        L.hipr_surface_create_ex.restype = ctypes.c_int
        L.hipr_surface_destroy.argtypes = [c_surf]
        L.hipr_surface_destroy.restype = None
        L.hipr_surface_clear.argtypes = [c_surf, ctypes.c_uint32]
        L.hipr_surface_clear.restype = ctypes.c_int
        L.hipr_surface_read_pixels.argtypes = [c_surf, ctypes.c_void_p, ctypes.c_size_t]
        L.hipr_surface_read_pixels.restype = ctypes.c_int
        L.hipr_surface_write_pixels.argtypes = [c_surf, ctypes.c_void_p, ctypes.c_size_t]
        L.hipr_surface_write_pixels.restype = ctypes.c_int
        L.hipr_surface_draw_scene.argtypes = [c_surf, SP, ctypes.POINTER(abi.Options)]
        L.hipr_surface_draw_scene.restype = ctypes.c_int
        L.hipr_surface_finish.argtypes = [c_surf]
        L.hipr_surface_finish.restype = ctypes.c_int
        L.hipr_surface_get_stats.argtypes = [c_surf, ctypes.POINTER(abi.Stats)]
        L.hipr_surface_get_stats.restype = ctypes.c_int
        L.hipr_surface_get_stage_times.argtypes = [c_surf, ctypes.c_void_p, ctypes.c_void_p,
                                                   ctypes.c_int32]
        L.hipr_surface_get_stage_times.restype = ctypes.c_int
        L.hipr_scene_prepare.argtypes = [c_ctx, SP, ctypes.POINTER(ctypes.c_void_p)]
        L.hipr_scene_prepare.restype = ctypes.c_int
        L.hipr_prepared_destroy.argtypes = [ctypes.c_void_p]
        L.hipr_prepared_destroy.restype = None
        L.hipr_surface_draw_prepared.argtypes = [c_surf, ctypes.c_void_p,
                                                 ctypes.POINTER(abi.Options)]
        L.hipr_surface_draw_prepared.restype = ctypes.c_int
        L.hipr_debug_build_edges.argtypes = [c_ctx, SP, ctypes.c_void_p, ctypes.c_void_p,
                                             ctypes.c_int64]
        L.hipr_debug_build_edges.restype = ctypes.c_int
        L.hipr_stackblur.argtypes = [c_ctx, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                     ctypes.c_int32, ctypes.c_int32]
        L.hipr_stackblur.restype = ctypes.c_int
        L.hipr_stackblur_ex.argtypes = [c_ctx, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                        ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]
        L.hipr_stackblur_ex.restype = ctypes.c_int
        L.hipr_surface_get_jobs.argtypes = [c_surf, ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.c_void_p, ctypes.c_int32,
                                            ctypes.POINTER(ctypes.c_int32)]
        L.hipr_surface_get_jobs.restype = ctypes.c_int
        L.hipr_surface_get_job_edges.argtypes = [c_surf, ctypes.c_int32, ctypes.c_void_p,
                                                 ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)]
        L.hipr_surface_get_job_edges.restype = ctypes.c_int
        L.hipr_surface_get_tile_bins.argtypes = [c_surf, ctypes.c_void_p, ctypes.c_void_p,
                                                 ctypes.c_int64, ctypes.c_int32]
        L.hipr_surface_get_tile_bins.restype = ctypes.c_int
        c_batch = ctypes.c_void_p
        L.hipr_batch_create.argtypes = [c_ctx, ctypes.POINTER(c_batch)]
        L.hipr_batch_create.restype = ctypes.c_int
        L.hipr_batch_destroy.argtypes = [c_batch]
        L.hipr_batch_destroy.restype = None
        L.hipr_batch_add.argtypes = [c_batch, c_surf, ctypes.c_void_p]
        L.hipr_batch_add.restype = ctypes.c_int
        L.hipr_batch_clear.argtypes = [c_batch]
        L.hipr_batch_clear.restype = ctypes.c_int
        L.hipr_batch_draw.argtypes = [c_batch, ctypes.POINTER(abi.Options)]
        L.hipr_batch_draw.restype = ctypes.c_int
        L.hipr_batch_set_split.argtypes = [c_batch, ctypes.c_int32, ctypes.c_int32,
                                           ctypes.c_int32]
        L.hipr_batch_set_split.restype = ctypes.c_int
        L.hipr_batch_autotune.argtypes = [c_batch, ctypes.c_int32, ctypes.c_int32,
                                          ctypes.POINTER(ctypes.c_int32),
                                          ctypes.POINTER(ctypes.c_int32),
                                          ctypes.POINTER(ctypes.c_int32)]
        L.hipr_batch_autotune.restype = ctypes.c_int
        L.hipr_batch_get_stats.argtypes = [c_batch, ctypes.POINTER(abi.Stats)]
        L.hipr_batch_get_stats.restype = ctypes.c_int
        L.hipr_batch_get_stage_times.argtypes = [c_batch, ctypes.c_void_p, ctypes.c_void_p,
                                                 ctypes.c_int32]
        L.hipr_batch_get_stage_times.restype = ctypes.c_int
        _lib = L
    return _lib


class HiprError(RuntimeError):
    def __init__(self, code, what):
        super().__init__(f"{what} failed with hipr_result {code}")
        self.code = code


def check(code, what):
    if code != abi.SUCCESS:
        raise HiprError(code, what)


class Context:
    def __init__(self, device=0):
        self._h = ctypes.c_void_p()
        check(lib().hipr_context_create(device, ctypes.byref(self._h)), "hipr_context_create")

    @property
    def handle(self):
        return self._h

    def build_edges(self, scene_ptr, n_paths):
        offs = np.zeros(n_paths + 1, np.int32)
        rc = lib().hipr_debug_build_edges(self._h, scene_ptr, offs.ctypes.data, None, 0)
        if rc not in (abi.SUCCESS, abi.ERROR_CAPACITY):
            raise HiprError(rc, "hipr_debug_build_edges")
        total = int(offs[-1])
        edges = np.zeros((max(total, 1), 16), np.int32)
        check(lib().hipr_debug_build_edges(self._h, scene_ptr, offs.ctypes.data,
                                           edges.ctypes.data, total), "hipr_debug_build_edges")
        return offs, edges[:total]

    def stackblur(self, img, radius, impl=abi.BLUR_IMPL_LINES):
        img = np.ascontiguousarray(img, np.uint8)
        h, w = img.shape[:2]
        out = np.zeros_like(img)
        check(lib().hipr_stackblur_ex(self._h, img.ctypes.data, out.ctypes.data, w, h, radius,
                                      impl), "hipr_stackblur_ex")
        return out

    def close(self):
        if self._h:
            lib().hipr_context_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _options(mode, edge_estimates, tile_estimates, blur_impl=abi.BLUR_IMPL_MFMA, band_rows=0,
             band_min_edges=0, flags=0, chunk_points=0):
    opt = abi.Options()
    opt.flags = flags
    opt.chunk_points = chunk_points
    opt.mode = mode
    opt.blur_impl = blur_impl
    opt.band_rows = band_rows
    opt.band_min_edges = band_min_edges
    keep = []
    if edge_estimates is not None:
        e = np.ascontiguousarray(edge_estimates, np.int32)
        keep.append(e)
        opt.edge_estimates = e.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    if tile_estimates is not None:
        t = np.ascontiguousarray(tile_estimates, np.int32)
        keep.append(t)
        opt.tile_estimates = t.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    return opt, keep


class Prepared:
    """A scene validated, planned and uploaded once (hipr_scene_prepare)."""

    def __init__(self, ctx, scene_ptr):
        self.ctx = ctx
        self._h = ctypes.c_void_p()
        check(lib().hipr_scene_prepare(ctx.handle, scene_ptr, ctypes.byref(self._h)),
              "hipr_scene_prepare")

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            lib().hipr_prepared_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Surface:
    def __init__(self, ctx, width, height, alpha_type=abi.ALPHA_PREMUL,
                 color_type=abi.COLOR_RGBA):
        self.ctx = ctx
        self.width, self.height = width, height
        self._h = ctypes.c_void_p()
        check(lib().hipr_surface_create_ex(ctx.handle, width, height, alpha_type, color_type,
                                           ctypes.byref(self._h)),
              "hipr_surface_create_ex")

    def clear(self, rgba=0):
        check(lib().hipr_surface_clear(self._h, rgba), "hipr_surface_clear")

    def draw(self, scene, mode=0, edge_estimates=None, tile_estimates=None,
             blur_impl=abi.BLUR_IMPL_MFMA, band_rows=0, band_min_edges=0, flags=0,
             chunk_points=0):
        """Draws a scene (a Scene pointer or a Prepared scene) onto the surface.
        band_rows: rows per walk band (0 = the library default, < 0 = none);
        band_min_edges: in the estimate modes, band only paths whose active
        edge bound reaches this (0 = default, < 0 = all); flags: FLAG_*;
        chunk_points: source points per sub-plan run (0 = default)."""
        opt, keep = _options(mode, edge_estimates, tile_estimates, blur_impl, band_rows,
                             band_min_edges, flags, chunk_points)
        if isinstance(scene, Prepared):
            check(lib().hipr_surface_draw_prepared(self._h, scene.handle, ctypes.byref(opt)),
                  "hipr_surface_draw_prepared")
        else:
            check(lib().hipr_surface_draw_scene(self._h, scene, ctypes.byref(opt)),
                  "hipr_surface_draw_scene")
        del keep

    def finish(self):
        check(lib().hipr_surface_finish(self._h), "hipr_surface_finish")

    def read(self):
        img = np.zeros((self.height, self.width, 4), np.uint8)
        check(lib().hipr_surface_read_pixels(self._h, img.ctypes.data, img.nbytes),
              "hipr_surface_read_pixels")
        return img

    def write(self, img):
        img = np.ascontiguousarray(img, np.uint8)
        check(lib().hipr_surface_write_pixels(self._h, img.ctypes.data, img.nbytes),
              "hipr_surface_write_pixels")

    def stats(self):
        s = abi.Stats()
        check(lib().hipr_surface_get_stats(self._h, ctypes.byref(s)), "hipr_surface_get_stats")
        return {k: getattr(s, k) for k, _ in abi.Stats._fields_}

    def jobs(self):
        """(path, kind, canvas) int32 arrays of the last draw's raster jobs."""
        n = ctypes.c_int32()
        check(lib().hipr_surface_get_jobs(self._h, None, None, None, 0, ctypes.byref(n)),
              "hipr_surface_get_jobs")
        path, kind, canvas = (np.zeros(n.value, np.int32) for _ in range(3))
        check(lib().hipr_surface_get_jobs(self._h, path.ctypes.data, kind.ctypes.data,
                                          canvas.ctypes.data, n.value, ctypes.byref(n)),
              "hipr_surface_get_jobs")
        return path, kind, canvas

    def job_edges(self, job):
        """Edge records (n x 16 int32, emission order) of one job of the last
        draw."""
        n = ctypes.c_int64()
        check(lib().hipr_surface_get_job_edges(self._h, job, None, 0, ctypes.byref(n)),
              "hipr_surface_get_job_edges")
        out = np.zeros((max(n.value, 1), 16), np.int32)
        check(lib().hipr_surface_get_job_edges(self._h, job, out.ctypes.data, n.value,
                                               ctypes.byref(n)),
              "hipr_surface_get_job_edges")
        return out[:n.value]

    def tile_bins(self):
        """Per job of the last draw, the sorted tile-bin codes
        ((tile_y * tiles_x + tile_x) << 1 | plane) as a list of arrays."""
        nj = len(self.jobs()[0])
        offs = np.zeros(nj + 1, np.int64)
        check(lib().hipr_surface_get_tile_bins(self._h, offs.ctypes.data, None, 0, nj),
              "hipr_surface_get_tile_bins")
        bins = np.zeros(max(int(offs[-1]), 1), np.int32)
        check(lib().hipr_surface_get_tile_bins(self._h, offs.ctypes.data, bins.ctypes.data,
                                               int(offs[-1]), nj),
              "hipr_surface_get_tile_bins")
        return [bins[offs[j]:offs[j + 1]] for j in range(nj)]

    def stage_times(self):
        """{stage: (ms, launches)} of the last draw (HIP events on the stream)."""
        n = len(abi.STAGES)
        ms = np.zeros(n, np.float64)
        cnt = np.zeros(n, np.int32)
        check(lib().hipr_surface_get_stage_times(self._h, ms.ctypes.data, cnt.ctypes.data, n),
              "hipr_surface_get_stage_times")
        return {name: (float(ms[i]), int(cnt[i])) for i, name in enumerate(abi.STAGES)}

    def close(self):
        if self._h:
            lib().hipr_surface_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many (surface, prepared scene) pairs drawn by one pipeline run
    (hipr_batch_*)."""

    def __init__(self, ctx):
        self.ctx = ctx
        self._h = ctypes.c_void_p()
        self._pairs = []  # keeps the surfaces and scenes alive
        check(lib().hipr_batch_create(ctx.handle, ctypes.byref(self._h)), "hipr_batch_create")

    def add(self, surface, prepared):
        check(lib().hipr_batch_add(self._h, surface._h, prepared.handle), "hipr_batch_add")
        self._pairs.append((surface, prepared))

    def clear(self):
        check(lib().hipr_batch_clear(self._h), "hipr_batch_clear")
        self._pairs = []

    def set_split(self, band_rows, band_min_edges, flags=0):
        check(lib().hipr_batch_set_split(self._h, band_rows, band_min_edges, flags),
              "hipr_batch_set_split")

    def autotune(self, mode, repeats=2):
        """Picks the fastest work split for `mode` (hipr_batch_autotune);
        returns (band_rows, band_min_edges, flags).  Leaves the targets'
        contents undefined."""
        r, m, f = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        check(lib().hipr_batch_autotune(self._h, mode, repeats, ctypes.byref(r), ctypes.byref(m),
                                        ctypes.byref(f)),
              "hipr_batch_autotune")
        return r.value, m.value, f.value

    def draw(self, mode=0, edge_estimates=None, tile_estimates=None,
             blur_impl=abi.BLUR_IMPL_MFMA, band_rows=0, band_min_edges=0, flags=0,
             chunk_points=0):
        opt, keep = _options(mode, edge_estimates, tile_estimates, blur_impl, band_rows,
                             band_min_edges, flags, chunk_points)
        check(lib().hipr_batch_draw(self._h, ctypes.byref(opt)), "hipr_batch_draw")
        del keep

    def stats(self):
        s = abi.Stats()
        check(lib().hipr_batch_get_stats(self._h, ctypes.byref(s)), "hipr_batch_get_stats")
        return {k: getattr(s, k) for k, _ in abi.Stats._fields_}

    def stage_times(self):
        n = len(abi.STAGES)
        ms = np.zeros(n, np.float64)
        cnt = np.zeros(n, np.int32)
        check(lib().hipr_batch_get_stage_times(self._h, ms.ctypes.data, cnt.ctypes.data, n),
              "hipr_batch_get_stage_times")
        return {name: (float(ms[i]), int(cnt[i])) for i, name in enumerate(abi.STAGES)}

    def close(self):
        if self._h:
            lib().hipr_batch_destroy(self._h)
            self._h = ctypes.c_void_p()
        self._pairs = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
