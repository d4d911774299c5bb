"""ctypes binding of include/skity_hip_picture.h (libskity_hip_canvas.so):
SKP pictures (Picture::MakeFromStream, module/io/include/skity/io/
picture.hpp:57-80) played back into display lists and drawn on a device
surface (DisplayList::Draw, include/skity/recorder/display_list.hpp:76)."""
import ctypes
import os

from . import abi

_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(abi.HERE, "libskity_hip_canvas.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built (python skity_amd/build.py host)")
        L = ctypes.CDLL(path)
        vp = ctypes.c_void_p
        L.skity_hip_picture_make.argtypes = [ctypes.c_char_p, ctypes.c_size_t,
                                             ctypes.POINTER(vp)]
        L.skity_hip_picture_last_error.restype = ctypes.c_char_p
        L.skity_hip_picture_info.argtypes = [vp, ctypes.POINTER(ctypes.c_float),
                                             ctypes.POINTER(ctypes.c_uint32),
                                             ctypes.POINTER(ctypes.c_int32),
                                             ctypes.POINTER(ctypes.c_int32)]
        L.skity_hip_picture_destroy.argtypes = [vp]
        L.skity_hip_picture_record.argtypes = [vp, ctypes.c_int32, ctypes.c_int32,
                                               ctypes.POINTER(ctypes.c_float),
                                               ctypes.POINTER(vp)]
        L.skity_hip_display_list_destroy.argtypes = [vp]
        L.skity_hip_display_list_scene.argtypes = [
            vp, ctypes.POINTER(ctypes.POINTER(abi.Scene))]
        L.skity_hip_display_list_draw.argtypes = [vp, vp, ctypes.c_int32]
        _LIB = L
    return _LIB


class Recording:
    """An SKP read and played back into a display list of a width x height
    canvas under an optional matrix {scale_x, skew_x, trans_x, skew_y,
    scale_y, trans_y}.  .scene is the hipr_scene its flush draws; it is None
    when the file is refused (rc_make, error) or the recording holds a call
    the backend does not draw (rc = ERROR_NOT_SUPPORTED)."""

    def __init__(self, data, width, height, matrix=None):
        L = lib()
        self.width, self.height = width, height
        self.pic, self.dl = ctypes.c_void_p(), ctypes.c_void_p()
        self.rc_make = L.skity_hip_picture_make(data, len(data), ctypes.byref(self.pic))
        self.error = L.skity_hip_picture_last_error().decode()
        self.scene, self.rc = None, self.rc_make
        if self.rc_make != 0:
            return
        m = None if matrix is None else (ctypes.c_float * 6)(*matrix)
        rc = L.skity_hip_picture_record(self.pic, width, height, m, ctypes.byref(self.dl))
        if rc != 0:
            raise RuntimeError(f"skity_hip_picture_record failed ({rc})")
        sc = ctypes.POINTER(abi.Scene)()
        self.rc = L.skity_hip_display_list_scene(self.dl, ctypes.byref(sc))
        if self.rc == 0:
            self.scene = sc

    def info(self):
        """(cull rect, SKP version, paint count, path count)"""
        cull = (ctypes.c_float * 4)()
        v, np_, nq = ctypes.c_uint32(), ctypes.c_int32(), ctypes.c_int32()
        rc = lib().skity_hip_picture_info(self.pic, cull, ctypes.byref(v), ctypes.byref(np_),
                                          ctypes.byref(nq))
        if rc != 0:
            raise RuntimeError(f"skity_hip_picture_info failed ({rc})")
        return tuple(cull), v.value, np_.value, nq.value

    def draw(self, surface, mode=abi.MODE_EXACT):
        """DisplayList::Draw onto the surface's canvas + Flush; the
        hipr_result."""
        return lib().skity_hip_display_list_draw(self.dl, surface._h, mode)

    def close(self):
        L = lib()
        if self.dl:
            L.skity_hip_display_list_destroy(self.dl)
            self.dl = ctypes.c_void_p()
        if self.pic:
            L.skity_hip_picture_destroy(self.pic)
            self.pic = ctypes.c_void_p()
        self.scene = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
