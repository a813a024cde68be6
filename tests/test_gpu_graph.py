"""The drop-in entries enqueue kernels only (no allocation, no memset, no
synchronisation), so a caller may capture them into a HIP graph.  A full
small frame (forward CLT, correlation, six-pair combination, whole-frame
inverse CLT and generate_RBGA, with the task words that generate_RBGA
rewrites restored by a copy first) is captured once and replayed three
times, with new image data and eager work between the replays; every
replay equals an eager run of the same calls on the same data.
"""
import numpy as np
import pytest
import torch

from tile_processor_gpu_amd import synth, tp
from tile_processor_gpu_amd.geometry import corr_indices

pytestmark = pytest.mark.gpu
SCALES = (0.25, 0.25, 0.5)
FZ = 0.05
PORTS = [[-0.5, -0.5], [0.5, -0.5], [-0.5, 0.5], [0.5, 0.5]]


def test_staged_frame_graph_replays():
    cfg = synth.config_small(seed=17)
    g = cfg["geom"]
    print(f"[gpu graph] {g}", flush=True)
    idx = corr_indices(g, cfg["tasks"])
    fb = tp.FrameBuffers(g, "cuda:0", corr_radius=7, max_corr=len(idx))
    fb.upload(cfg["images"], cfg["kernels"], cfg["koffs"], cfg["tasks"], idx)
    dev = torch.device("cuda:0")
    tasks0 = torch.from_numpy(np.ascontiguousarray(cfg["tasks"]).view(np.uint8).copy()).to(dev)
    tasks = tasks0.clone()
    ind = torch.zeros(g.tilesx * g.tilesya, dtype=torch.int32, device=dev)
    num = torch.zeros(8, dtype=torch.int32, device=dev)
    woi = torch.zeros(4, dtype=torch.int32, device=dev)
    stride = (g.tilesx + 1) * 8
    rbga = torch.zeros((4 * (g.tilesy + 1) * 8, stride), device=dev)
    comb = torch.zeros((g.num_tiles, 225), device=dev)
    planes = [torch.zeros((3, g.img_height + 8, g.img_width + 8), device=dev) for _ in range(4)]
    p_planes = tp.ptr_array(planes)
    ports = torch.tensor(PORTS, device=dev)
    kw = tp.TEXTURE_DEFAULTS

    def frame():
        tasks.copy_(tasks0)
        fb.convert()
        fb.correlate(3, SCALES, FZ)
        tp.corr_combine(g, len(idx), fb.corr_idx, fb.corrs, fb.corr_stride, 7, 0x3F, comb, 225)
        tp.imclt_rbg_all(g, fb.p_clt, p_planes, g.img_width + 8)
        tp.generate_RBGA(g, tasks, len(cfg["tasks"]), ind, num, woi, g.tilesx, g.tilesy, fb.p_clt, ports,
                         3, 0, kw["min_shot"], kw["scale_shot"], kw["diff_sigma"], kw["diff_threshold"],
                         kw["min_agree"], kw["weight0"], kw["weight1"], kw["weight2"], kw["dust_remove"], 0,
                         stride, rbga)

    def outputs():
        return [t.clone() for t in fb.clt] + [fb.corrs[:len(idx)].clone(), comb.clone(), rbga.clone(),
                                              woi.clone(), num.clone()] + [p.clone() for p in planes]

    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        frame()  # eager once on the capture stream before capturing
    torch.cuda.current_stream().wait_stream(s)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=s):
        frame()
    torch.cuda.synchronize()

    rng = np.random.default_rng(5)
    for rep in range(3):
        fb.images[rep % 4].mul_(float(rng.uniform(0.5, 2.0)))
        graph.replay()
        torch.cuda.synchronize()
        got = outputs()
        for t in [rbga, comb] + planes:
            t.fill_(-3.0)
        frame()  # eager, same data
        torch.cuda.synchronize()
        ref = outputs()
        for i, (a, b) in enumerate(zip(got, ref)):
            assert torch.equal(a, b), (rep, i)
        assert float(ref[6].abs().max()) > 0  # the RGBA image was written
