"""Host-side checks of the stream form (no GPU): frame descriptors are
validated against the rows the persistent kernel writes, and they follow a
frame whose buffers or task count change after a StreamRun (or its captured
graph) was made."""
import types

import numpy as np
import pytest
import torch

from tile_processor_gpu_amd import tp
from tile_processor_gpu_amd.geometry import Geometry, TASK_DTYPE, full_tasks


def _stub(geom, ntasks, corr_rows):
    """Stand-in for a FrameBuffers with CPU tensors (the descriptor logic only
    reads addresses, shapes and the task count)."""
    fb = types.SimpleNamespace(geom=geom, device=torch.device("cpu"), dstride=geom.img_width,
                               corr_radius=7, corr_stride=225, tasks=None, num_tiles=0)
    for name in ("p_images", "p_kernels", "p_koffs", "p_clt"):
        setattr(fb, name, torch.zeros(4, dtype=torch.int64))
    fb.corrs = torch.zeros((corr_rows, 225))
    fb.combined = None
    fb.combined_buffer = lambda: tp.FrameBuffers.combined_buffer(fb)
    tasks = full_tasks(geom, np.zeros((geom.num_tiles, 4, 2), np.float32))[:ntasks]
    tp.FrameBuffers.upload(fb, tasks=tasks)
    return fb


def test_stream_rejects_short_corrs():
    g = Geometry(64, 64)
    ok = _stub(g, 20, 6 * 20)
    short = _stub(g, 20, 6 * 20 - 1)
    tp.StreamRun([ok], stamps=False)
    with pytest.raises(ValueError, match="6 per task"):
        tp.StreamRun([ok, short], stamps=False)
    # without stored pair rows the row count does not matter
    tp.StreamRun([short], store_pairs=False, combined=True, stamps=False)


def test_task_buffer_keeps_its_address():
    g = Geometry(64, 64)
    fb = _stub(g, 10, 6 * g.num_tiles)
    addr = fb.tasks.data_ptr()
    tasks = full_tasks(g, np.zeros((g.num_tiles, 4, 2), np.float32))
    tp.FrameBuffers.upload(fb, tasks=tasks[:37])
    assert fb.tasks.data_ptr() == addr and fb.num_tiles == 37
    got = fb.tasks[:37 * TASK_DTYPE.itemsize].numpy().view(TASK_DTYPE)
    assert np.array_equal(got, tasks[:37])
    tp.FrameBuffers.upload(fb, tasks=tasks)  # a full list still fits
    assert fb.tasks.data_ptr() == addr and fb.num_tiles == g.num_tiles


def test_descriptors_follow_changed_frames():
    g = Geometry(64, 64)
    a, b = _stub(g, 30, 6 * g.num_tiles), _stub(g, 12, 6 * g.num_tiles)
    run = tp.StreamRun([a, b], combined=True, stamps=False)
    desc = run.desc.numpy().view(tp.FRAME_DTYPE)
    assert list(desc["num_tiles"]) == [30, 12]
    assert desc["tasks"][1] == b.tasks.data_ptr()
    assert desc["combined"][0] == a.combined.data_ptr()
    assert not run.refresh()
    # a new task count and a reallocated task buffer (a list larger than the
    # capacity) are picked up in place: the device array keeps its address
    base = run.desc.data_ptr()
    tasks = full_tasks(g, np.zeros((g.num_tiles, 4, 2), np.float32))
    tp.FrameBuffers.upload(a, tasks=tasks[:5])
    b.tasks = torch.zeros(2 * b.tasks.numel(), dtype=torch.uint8)
    assert run.refresh()
    desc = run.desc.numpy().view(tp.FRAME_DTYPE)
    assert run.desc.data_ptr() == base
    assert list(desc["num_tiles"]) == [5, 12]
    assert desc["tasks"][1] == b.tasks.data_ptr()
    assert not run.refresh()
    # a task count whose rows no longer fit is refused before any launch
    a.num_tiles = g.num_tiles + 1
    with pytest.raises(ValueError, match="6 per task"):
        run.refresh()
