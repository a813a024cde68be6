"""ctypes loader for libtp.so (the C ABI of include/tp.h).

The library is built in-tree by `python -m tile_processor_gpu_amd.build`.
There is no fallback: a missing library is an error.
"""
import ctypes
import os

from .geometry import CGeometry

# TP_LIB_PATH selects another build of the same library (for A/B timing)
LIB_PATH = os.environ.get("TP_LIB_PATH") or os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                          "libtp.so")

_P = ctypes.c_void_p
_I = ctypes.c_int
_F = ctypes.c_float
_SZ = ctypes.c_size_t
_G = ctypes.POINTER(CGeometry)

# name -> argument types (return type is int unless listed in _RESTYPES)
PROTOTYPES = {
    "tp_last_error": [],
    "tp_version": [],
    "tp_set_lpf": [_I, _P],
    "tp_convert_correct_tiles": [_G, _P, _P, _P, _P, _P, _SZ, _I, _I, _P],
    "tp_correlate2D": [_G, _P, _I, _F, _F, _F, _F, _SZ, _P, _SZ, _I, _P, _P],
    "tp_imclt_rbg": [_G, _P, _P, _I, _I, _I, _SZ, _P],
    "tp_imclt_rbg_all": [_G, _P, _P, _SZ, _P],
    "tp_textures_accumulate": [_G, _P, _P, _SZ, _P, _P, _I, _I, _F, _F, _F, _F, _F, _F, _F, _F,
                               _I, _I, _SZ, _P, _SZ, _P, _P],
    "tp_prepare_texture_list": [_G, _P, _I, _P, _P, _P, _I, _I, _P],
    "tp_corr_combine": [_G, _SZ, _P, _P, _SZ, _I, _I, _P, _SZ, _P],
    "tp_corr_disparity": [_P, _SZ, _I, _I, _I, _P, _P],
    "tp_generate_RBGA": [_G, _P, _I, _P, _P, _P, _I, _I, _P, _P, _I, _I, _F, _F, _F, _F, _F, _F,
                         _F, _F, _I, _I, _SZ, _P, _P],
    "tp_convert_correlate": [_G, _P, _P, _P, _P, _P, _SZ, _I, _I, _I, _F, _F, _F, _F, _SZ, _I, _P,
                             _P, _SZ, _I, _P, _P],
    "tp_stream_convert_correlate": [_G, _P, _I, _P, _P, _SZ, _I, _I, _F, _F, _F, _F, _SZ, _I, _SZ,
                                    _I, _P, _P],
    "tp_stream_convert_correlate_noalloc": [_G, _P, _I, _P, _SZ, _I, _I, _F, _F, _F, _F, _SZ, _I,
                                            _SZ, _I, _P, _P],
    "tp_stream_graph_create": [_G, _P, _I, _P, _P, _SZ, _I, _I, _F, _F, _F, _F, _SZ, _I, _SZ, _I,
                               _P, _P],
    "tp_graph_launch": [_P, _P],
    "tp_graph_destroy": [_P],
    "tp_wall_clock_hz": [],
    "tp_nn_input": [_G, _P, _I, _P, _SZ, _P, _P],
    "tp_nn_forward": [_P, _P, _I, _P, _P, _P],
    "tp_nn_infer": [_P, _P, _I, _P, _P],
    "tp_nn_infer_corrs": [_G, _P, _P, _I, _P, _SZ, _P, _P],
    "tp_nn_backward": [_P, _P, _P, _P, _I, _F, _F, _F, _P, _P, _P],
    "tp_nn_adam": [_P, _P, _P, _P, _I, _F, _F, _F, _F, _I, _P, _P],
    "tp_nn_pack_bf16": [_P, _I, _P, _P],
    "tp_psf_to_clt": [_G, _P, _I, _I, _I, _P, _P, _P],
    "tp_dtt8x8": [_P, _P, _I, _I, _I, _I, _P],
    "tp_fill": [_P, _SZ, _F, _P],
}
_RESTYPES = {"tp_last_error": ctypes.c_char_p, "tp_version": ctypes.c_char_p,
             "tp_wall_clock_hz": ctypes.c_double}

_lib = None


class TPError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise TPError(f"{LIB_PATH} is not built; run python -m tile_processor_gpu_amd.build")
        L = ctypes.CDLL(LIB_PATH)
        for name, argt in PROTOTYPES.items():
            fn = getattr(L, name)
            fn.argtypes = argt
            fn.restype = _RESTYPES.get(name, ctypes.c_int)
        _lib = L
    return _lib


def call(name, *args):
    rc = getattr(lib(), name)(*args)
    if rc != 0:
        msg = lib().tp_last_error().decode()
        raise TPError(f"{name} failed ({rc}): {msg}")
    return rc
