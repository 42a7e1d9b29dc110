"""ctypes view of tests/vshost/libvshost.so, a minimal VapourSynth 4 host
(test infrastructure): load a plugin, inspect its registrations, build
source clips from numpy planes, invoke filters and fetch frames."""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "vshost", "libvshost.so")
CF_YUV, ST_INTEGER, ST_FLOAT = 3, 0, 1
PT_INT, PT_FLOAT = 1, 2


class HostError(RuntimeError):
    pass


def _lib():
    L = C.CDLL(LIB)
    vp = C.c_void_p
    L.vsh_create.restype = vp
    L.vsh_free.argtypes = [vp]
    L.vsh_error.argtypes = [vp]
    L.vsh_error.restype = C.c_char_p
    L.vsh_load.argtypes = [vp, C.c_char_p]
    L.vsh_plugin.argtypes = [vp, C.c_int]
    L.vsh_plugin.restype = C.c_char_p
    L.vsh_num_functions.argtypes = [vp]
    L.vsh_function.argtypes = [vp, C.c_int, C.c_int]
    L.vsh_function.restype = C.c_char_p
    L.vsh_source.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                             C.POINTER(vp), C.POINTER(C.c_ssize_t), C.c_int]
    L.vsh_invoke.argtypes = [vp, C.c_char_p, C.c_int, C.c_int, C.POINTER(C.c_char_p), C.POINTER(C.c_char_p)]
    L.vsh_node_info.argtypes = [vp, C.c_int, C.POINTER(C.c_int)]
    L.vsh_get_frame.argtypes = [vp, C.c_int, C.c_int, C.POINTER(vp), C.POINTER(C.c_ssize_t)]
    L.vsh_shares_plane.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int]
    L.vsh_prop.argtypes = [vp, C.c_char_p, C.POINTER(C.c_int64), C.POINTER(C.c_double)]
    L.vsh_num_props.argtypes = [vp]
    L.vsh_prop_key.argtypes = [vp, C.c_int]
    L.vsh_prop_key.restype = C.c_char_p
    return L


class Host:
    def __init__(self):
        self.L = _lib()
        self.h = self.L.vsh_create()

    def close(self):
        if self.h:
            self.L.vsh_free(self.h)
            self.h = None

    def error(self):
        return self.L.vsh_error(self.h).decode()

    def load(self, path):
        if self.L.vsh_load(self.h, path.encode()) != 0:
            raise HostError(self.error())

    def plugin(self):
        return tuple(self.L.vsh_plugin(self.h, i).decode() for i in range(3))

    def functions(self):
        return {self.L.vsh_function(self.h, i, 0).decode(): (self.L.vsh_function(self.h, i, 1).decode(),
                                                              self.L.vsh_function(self.h, i, 2).decode())
                for i in range(self.L.vsh_num_functions(self.h))}

    def source(self, frames, bits, ss_w, ss_h, chroma_location=-1):
        """frames: list of (Y, U, V) numpy planes (uint8/uint16/float32)."""
        y = frames[0][0]
        st = ST_FLOAT if y.dtype == np.float32 else ST_INTEGER
        arrs = [np.ascontiguousarray(p) for f in frames for p in f]
        ptrs = (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
        strides = (C.c_ssize_t * len(arrs))(*[a.strides[0] for a in arrs])
        return self.L.vsh_source(self.h, CF_YUV, st, bits, ss_w, ss_h, y.shape[1], y.shape[0], len(frames), ptrs,
                                 strides, chroma_location)

    def invoke(self, name, node, **args):
        keys = (C.c_char_p * max(1, len(args)))(*[k.encode() for k in args])
        vals = (C.c_char_p * max(1, len(args)))(*[str(v).encode() for v in args.values()])
        r = self.L.vsh_invoke(self.h, name.encode(), node, len(args), keys, vals)
        if r < 0:
            raise HostError(self.error())
        return r

    def info(self, node):
        v = (C.c_int * 8)()
        self.L.vsh_node_info(self.h, node, v)
        return list(v)

    def frame(self, node, n):
        cf, st, bits, ssw, ssh, w, h, _ = self.info(node)
        dt = np.float32 if st == ST_FLOAT else (np.uint8 if bits <= 8 else np.uint16)
        planes = [np.zeros((h, w), dt), np.zeros((h >> ssh, w >> ssw), dt), np.zeros((h >> ssh, w >> ssw), dt)]
        ptrs = (C.c_void_p * 3)(*[p.ctypes.data for p in planes])
        strides = (C.c_ssize_t * 3)(*[p.strides[0] for p in planes])
        if self.L.vsh_get_frame(self.h, node, n, ptrs, strides) != 0:
            raise HostError(self.error())
        return planes

    def shares_plane(self, plane, src_node, n, src_plane=0):
        return bool(self.L.vsh_shares_plane(self.h, plane, src_node, n, src_plane))

    def props(self):
        out = {}
        for i in range(self.L.vsh_num_props(self.h)):
            k = self.L.vsh_prop_key(self.h, i)
            iv, fv = C.c_int64(), C.c_double()
            t = self.L.vsh_prop(self.h, k, C.byref(iv), C.byref(fv))
            out[k.decode()] = iv.value if t == PT_INT else (fv.value if t == PT_FLOAT else None)
        return out
