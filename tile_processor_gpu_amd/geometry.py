"""Geometry and binary layouts shared with the C ABI (include/tp.h).

Layouts follow the reference (src/TileProcessor.cuh):
  tp_task      40 B: int task; int txy (tx | ty << 16); float xy[4][2]   (:179-186)
  CltExtra     8 floats per kernel cell and color                           (:187-196)
  kernels      float[kernels_vert][kernels_hor][3][4][8][8] per camera
  clt          float[tilesy][tilesx][3][4][8][8] per camera
"""
import ctypes
from dataclasses import dataclass

import numpy as np

NUM_CAMS = 4
NUM_COLORS = 3
NUM_PAIRS = 6
DTT_SIZE = 8
KERNELS_LSTEP = 4
CORR_NTILE_SHIFT = 8
TASK_CORR_BITS = 4
TASK_TEXTURE_BITS = 0xF
LIST_TEXTURE_BIT = 7
# camera pairs of the correlation list (src/TileProcessor.cuh:484-490)
PAIRS = ((0, 1), (2, 3), (0, 2), (1, 3), (0, 3), (2, 1))
# camera port offsets used by the reference harness (src/dtt8x8.cu:268-272)
PORT_OFFSETS = ((-0.5, -0.5), (0.5, -0.5), (-0.5, 0.5), (0.5, 0.5))

TASK_DTYPE = np.dtype([("task", "<i4"), ("txy", "<i4"), ("xy", "<f4", (NUM_CAMS, 2))])
assert TASK_DTYPE.itemsize == 40


class CGeometry(ctypes.Structure):
    _fields_ = [("img_width", ctypes.c_int32), ("img_height", ctypes.c_int32),
                ("kernels_hor", ctypes.c_int32), ("kernels_vert", ctypes.c_int32)]


@dataclass(frozen=True)
class Geometry:
    """Run-time replacement for IMG_WIDTH/IMG_HEIGHT/KERNELS_HOR/KERNELS_VERT."""
    img_width: int
    img_height: int
    kernels_hor: int = -1
    kernels_vert: int = -1

    def __post_init__(self):
        if self.kernels_hor < 0:
            object.__setattr__(self, "kernels_hor", self.img_width // 16 + 2)
        if self.kernels_vert < 0:
            object.__setattr__(self, "kernels_vert", self.img_height // 16 + 2)

    @property
    def tilesx(self):
        return self.img_width // DTT_SIZE

    @property
    def tilesy(self):
        return self.img_height // DTT_SIZE

    @property
    def tilesya(self):
        return (self.tilesy + 3) & ~3

    @property
    def num_tiles(self):
        return self.tilesx * self.tilesy

    @property
    def kernel_floats(self):
        """floats of one camera's aberration kernels"""
        return self.kernels_vert * self.kernels_hor * NUM_COLORS * 256

    @property
    def clt_floats(self):
        """floats of one camera's CLT buffer"""
        return self.tilesy * self.tilesx * NUM_COLORS * 256

    def c(self):
        return CGeometry(self.img_width, self.img_height, self.kernels_hor, self.kernels_vert)


def full_tasks(geom, xy, task=TASK_TEXTURE_BITS | (((1 << NUM_PAIRS) - 1) << TASK_CORR_BITS)):
    """All-tiles task list as the reference harness builds it
    (src/dtt8x8.cu:417-428): xy is float[tiles][4][2] (tile centres)."""
    n = geom.num_tiles
    t = np.zeros(n, TASK_DTYPE)
    ty, tx = np.divmod(np.arange(n), geom.tilesx)
    t["task"] = task
    t["txy"] = tx + (ty << 16)
    t["xy"] = xy
    return t


def corr_indices(geom, tasks):
    """Correlation list, tile-major and pair-minor, from the task bits
    (src/dtt8x8.cu:453-472): entry = (tx + ty * tilesx) << 8 | pair."""
    tx = tasks["txy"] & 0xFFFF
    ty = tasks["txy"] >> 16
    nt = (tx + ty * geom.tilesx).astype(np.int64)
    cm = (tasks["task"] >> TASK_CORR_BITS) & ((1 << NUM_PAIRS) - 1)
    bits = (cm[:, None] >> np.arange(NUM_PAIRS)[None, :]) & 1
    ti, pi = np.nonzero(bits)
    return ((nt[ti] << CORR_NTILE_SHIFT) | pi).astype(np.int32)


def texture_indices(geom, tasks):
    """Texture list as the reference harness builds it (src/dtt8x8.cu:474-492):
    entry = tile << 8 | 1 << 7 for every tile with a texture bit."""
    tx = tasks["txy"] & 0xFFFF
    ty = tasks["txy"] >> 16
    nt = (tx + ty * geom.tilesx).astype(np.int64)
    sel = (tasks["task"] & TASK_TEXTURE_BITS) != 0
    return ((nt[sel] << CORR_NTILE_SHIFT) | (1 << LIST_TEXTURE_BIT)).astype(np.int32)
