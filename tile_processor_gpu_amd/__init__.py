"""MI355X (gfx950) tile processor: forward CLT with aberration correction,
pairwise phase correlation, inverse CLT and texture generation for 4-camera
Bayer stereo, as hand-written HIP kernels behind the C ABI of libtp.so
(include/tp.h).

The Python layer is a thin host-side mirror of the reference kernels'
calling convention (mfkiwl/tile_processor_gpu, src/TileProcessor.cuh):
`tp` wraps the C ABI for torch device tensors, `geometry` holds the binary
layouts, `synth` generates seeded synthetic inputs.
"""
from .geometry import (  # noqa: F401
    Geometry, TASK_DTYPE, NUM_CAMS, NUM_COLORS, NUM_PAIRS, PAIRS, PORT_OFFSETS,
    CORR_NTILE_SHIFT, TASK_CORR_BITS, TASK_TEXTURE_BITS, LIST_TEXTURE_BIT,
)

__version__ = "0.1.0"
