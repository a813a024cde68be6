"""The benchmark's HBM probe kernels move the bytes they claim to move."""
import ctypes

import pytest
import torch

from tile_processor_gpu_amd import build as tpb

pytestmark = pytest.mark.gpu
VP = ctypes.c_void_p


def probe():
    P = ctypes.CDLL(tpb.PROBE_LIB)
    P.tp_probe_hbm.restype = ctypes.c_int
    P.tp_probe_hbm.argtypes = [ctypes.c_int, VP, VP, VP, ctypes.c_size_t, ctypes.c_int, VP]
    return P


@pytest.mark.parametrize("nt", [0, 1])
def test_copy_mix_write(nt):
    P = probe()
    n = (1 << 20) + 4  # floats; not a multiple of the workgroup's 4 KB
    g = torch.Generator(device="cuda:0").manual_seed(3)
    src = torch.rand(n, device="cuda:0", generator=g)
    dst = torch.zeros(3 * n, device="cuda:0")
    sink = torch.zeros(1, device="cuda:0")
    assert P.tp_probe_hbm(2, VP(src.data_ptr()), VP(dst.data_ptr()), VP(sink.data_ptr()), 4 * n, nt, None) == 0
    torch.cuda.synchronize()
    assert torch.equal(dst[:n], src)
    assert P.tp_probe_hbm(3, VP(src.data_ptr()), VP(dst.data_ptr()), VP(sink.data_ptr()), 4 * n, nt, None) == 0
    assert P.tp_probe_hbm(0, VP(src.data_ptr()), None, VP(sink.data_ptr()), 4 * n, 0, None) == 0
    torch.cuda.synchronize()
    assert float(sink.item()) == 0.0
    assert torch.equal(dst[:n], src)
    assert torch.equal(dst[n:2 * n], src * 2) and torch.equal(dst[2 * n:], src * 3)
    assert P.tp_probe_hbm(1, None, VP(dst.data_ptr()), None, 4 * n, nt, None) == 0
    torch.cuda.synchronize()
    pat = dst[:n].view(-1, 4)
    assert torch.equal(pat[:, :3], torch.tensor([1.0, 2.0, 3.0], device="cuda:0").expand(n // 4, 3))
    assert torch.equal(dst[n:2 * n], src * 2)  # past the written bytes: untouched
    # bad arguments
    assert P.tp_probe_hbm(1, None, VP(dst.data_ptr()), None, 4 * n, 2, None) == -1
    assert P.tp_probe_hbm(1, None, VP(dst.data_ptr()), None, 4 * n + 4, 0, None) == -1
