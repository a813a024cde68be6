"""GPU parity of the texture path (textures_accumulate, prepare_texture_list,
generate_RBGA) against the oracle.

Every output pixel must be within 1e-5 (max |x - x64| / max |x64| per slice)
of the FP64 oracle, or be a pixel where the 4-camera consensus makes a
choice that the oracle certifies as a tie at FP32 accuracy and within 1e-5
of the FP64 output with that choice overridden (RGBA image: within 1e-5
plus the change such flips of the overlapping tiles can make).  See
tests/texcert.py.
"""
import numpy as np
import pytest
import torch

from tile_processor_gpu_amd import tp
from tile_processor_gpu_amd.geometry import Geometry, texture_indices, TASK_DTYPE
from texcert import scene_cfg, certify_tiles, rbga_slack, certify_rbga

pytestmark = pytest.mark.gpu
PORTS = np.array([[-0.5, -0.5], [0.5, -0.5], [-0.5, 0.5], [0.5, 0.5]], np.float32)


@pytest.fixture(scope="module")
def tex(orc):
    g = Geometry(256, 256)
    imgs, kern, koffs, tasks = scene_cfg(3, g)
    print(f"[gpu textures] {g}", flush=True)
    clt64 = orc.convert_correct_tiles(g, koffs, kern, imgs, tasks)
    clt_dev = [torch.from_numpy(c.astype(np.float32)).cuda() for c in clt64]
    return dict(g=g, tasks=tasks, clt64=clt64, clt_dev=clt_dev, p_clt=tp.ptr_array(clt_dev),
                ports=torch.from_numpy(PORTS).cuda())


@pytest.mark.parametrize("colors", [3, 1])
@pytest.mark.parametrize("keep", [0, 1])
def test_texture_tiles_match_oracle(tex, orc, keep, colors):
    g = tex["g"]
    idx = texture_indices(g, tex["tasks"])
    prm = orc.TexParams(keep_weights=keep, colors=colors)
    nsl = prm.slices()
    stride = nsl * 256 + 64
    d_idx = torch.from_numpy(idx).cuda()
    out = torch.full((len(idx), stride), -9.0, device="cuda:0")
    kw = dict(tp.TEXTURE_DEFAULTS, keep_weights=keep, colors=colors)
    tp.textures_accumulate(g, None, tex["p_clt"], len(idx), d_idx, tex["ports"],
                           kw["colors"], kw["is_lwir"], kw["min_shot"], kw["scale_shot"],
                           kw["diff_sigma"], kw["diff_threshold"], kw["min_agree"], kw["weight0"],
                           kw["weight1"], kw["weight2"], kw["dust_remove"], keep, 0, None, stride, out)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    assert (got[:, nsl * 256:] == -9.0).all()
    n_bad, n_tied = certify_tiles(orc, g, tex["clt64"], idx, prm,
                                  got[:, :nsl * 256].reshape(len(idx), nsl, 256), nsl)
    print(f"{n_bad} pixels beyond 1e-5, all certified ties ({n_tied} tie pixels)")


def test_texture_tiles_odd_list_and_skipped_entries(tex, orc):
    """The kernel takes list entries two at a time: an odd-length list, and
    entries without the texture bit or off the grid (skipped: their output
    rows stay untouched), next to valid ones in the same pair."""
    g = tex["g"]
    idx = texture_indices(g, tex["tasks"])[:-3].copy()
    assert len(idx) % 2 == 1
    idx[5] &= ~(1 << 7)                          # no texture bit
    idx[10] = (g.num_tiles + 3) << 8 | (1 << 7)  # off the grid
    prm = orc.TexParams()
    nsl = prm.slices()
    stride = nsl * 256
    out = torch.full((len(idx), stride), -9.0, device="cuda:0")
    kw = tp.TEXTURE_DEFAULTS
    tp.textures_accumulate(g, None, tex["p_clt"], len(idx), torch.from_numpy(idx).cuda(), tex["ports"],
                           kw["colors"], kw["is_lwir"], kw["min_shot"], kw["scale_shot"],
                           kw["diff_sigma"], kw["diff_threshold"], kw["min_agree"], kw["weight0"],
                           kw["weight1"], kw["weight2"], kw["dust_remove"], 0, 0, None, stride, out)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    assert (got[5] == -9.0).all() and (got[10] == -9.0).all()
    # the oracle skips the same two entries (their rows are ignored here)
    certify_tiles(orc, g, tex["clt64"], idx, prm, got.reshape(len(idx), nsl, 256), nsl)


def _gpu_generate(tex, colors=3, pad=8):
    g = tex["g"]
    t = np.ascontiguousarray(tex["tasks"])
    d_tasks = torch.from_numpy(t.view(np.uint8).copy()).cuda()
    ind = torch.full((g.tilesx * g.tilesya,), -1, dtype=torch.int32, device="cuda:0")
    num = torch.full((8,), -1, dtype=torch.int32, device="cuda:0")
    woi = torch.full((4,), -1, dtype=torch.int32, device="cuda:0")
    stride = (g.tilesx + 1) * 8 + pad
    rbga = torch.full(((colors + 1) * (g.tilesy + 1) * 8, stride), 7.0, device="cuda:0")
    kw = dict(tp.TEXTURE_DEFAULTS, colors=colors)
    tp.generate_RBGA(g, d_tasks, len(t), ind, num, woi, g.tilesx, g.tilesy, tex["p_clt"],
                     tex["ports"], colors, 0, kw["min_shot"], kw["scale_shot"], kw["diff_sigma"],
                     kw["diff_threshold"], kw["min_agree"], kw["weight0"], kw["weight1"],
                     kw["weight2"], kw["dust_remove"], 0, stride, rbga)
    torch.cuda.synchronize()
    tasks_out = d_tasks.cpu().numpy().view(TASK_DTYPE)
    return tasks_out, ind.cpu().numpy(), num.cpu().numpy(), woi.cpu().numpy(), rbga.cpu().numpy(), stride


# pad 3: a row stride that rules out 16-byte stores
@pytest.mark.parametrize("pad,colors", [(8, 3), (3, 3), (8, 1)])
def test_generate_rbga_matches_oracle(tex, orc, pad, colors):
    g = tex["g"]
    tasks_out, ind, num, woi, rbga, stride = _gpu_generate(tex, colors=colors, pad=pad)
    t_ref, woi_ref, num_ref, lists_ref, rbga_ref = orc.generate_rbga(g, tex["tasks"], tex["clt64"],
                                                                     orc.TexParams(colors=colors), stride)
    assert list(woi) == list(woi_ref)
    assert list(num) == list(num_ref)
    assert np.array_equal(tasks_out["task"], t_ref["task"])
    lists = orc.sub_lists(g, ind, num)
    for a, b in zip(lists, lists_ref):
        assert sorted(a.tolist()) == sorted(b.tolist())
    rows, cols = (woi[3] + 1) * 8, (woi[2] + 1) * 8
    got = rbga[:(colors + 1) * rows].reshape(colors + 1, rows, stride)
    assert (got[:, :, cols:] == 7.0).all()  # outside the woi: untouched
    ref = rbga_ref[:, :rows, :cols]
    slack = rbga_slack(orc, g, tex["clt64"], lists_ref, woi_ref, orc.TexParams(colors=colors), ref.shape)
    n = certify_rbga(got[:, :, :cols], ref, slack)
    print(f"{n} RGBA pixels beyond 1e-5, all within the certified-tie slack")


def test_prepare_texture_list_sparse(tex, orc):
    """Sparse, clustered texture mask: list, counts, woi and task bits."""
    g = tex["g"]
    rng = np.random.default_rng(8)
    t = np.ascontiguousarray(tex["tasks"]).copy()
    keep = rng.uniform(size=len(t)) < 0.4
    t["task"][~keep] &= ~0xF
    d_tasks = torch.from_numpy(t.view(np.uint8).copy()).cuda()
    ind = torch.zeros(g.tilesx * g.tilesya, dtype=torch.int32, device="cuda:0")
    num = torch.zeros(8, dtype=torch.int32, device="cuda:0")
    woi = torch.zeros(4, dtype=torch.int32, device="cuda:0")
    tp.prepare_texture_list(g, d_tasks, len(t), ind, num, woi, g.tilesx, g.tilesy)
    torch.cuda.synchronize()
    t_ref, ind_ref, num_ref, woi_ref = orc.prepare_texture_list(g, t)
    assert list(woi.cpu().numpy()) == list(woi_ref)
    n = num.cpu().numpy()
    assert list(n) == list(num_ref)
    assert np.array_equal(d_tasks.cpu().numpy().view(TASK_DTYPE)["task"], t_ref["task"])
    for a, b in zip(orc.sub_lists(g, ind.cpu().numpy(), n), orc.sub_lists(g, ind_ref, num_ref)):
        assert sorted(a.tolist()) == sorted(b.tolist())
