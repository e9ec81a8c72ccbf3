"""GPU draw lists against the oracle: node buffers of every LOD level
(vp_gpu_svl_nodes), the assembled frame list (vp_host_octree_cull +
vp_gpu_draw_gather) and the splat resolve of that list, bit-exact ids."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import refimpl as RI
from tests.gpuutil import dev, host, ids_dev
from voxplat_amd import abi, frustum_planes, octree_cull
from voxplat_amd import workload as W

pytestmark = pytest.mark.gpu


def _remesh(gpu, w, vox_d, ids, smap_d, uniform=None):
    import torch

    recs, offs, counts = gpu.cull_mesh(w, vox_d, ids_dev(ids), shadow_map=smap_d, uniform=uniform)
    torch.cuda.synchronize()
    assert gpu.status() == 0
    return recs, offs, counts


def _nodes(gpu, w, recs, ids, offs, counts):
    import torch

    idd = ids_dev(ids)
    out = [gpu.svl_nodes(w, l, recs, idd, offs, counts) for l in range(5)]
    torch.cuda.synchronize()
    assert gpu.status() == 0
    return out


@pytest.mark.parametrize("rb,mb,seed", [(2, (2, 1, 2), 1), (3, (3, 2, 1), 2), (4, (4, 1, 4), 3),
                                        (6, (2, 1, 2), 4)])
def test_svl_nodes_all_levels_match_oracle(gpu, rb, mb, seed):
    import torch

    w = abi.WorldDesc.make(rb, mb)
    rng = np.random.default_rng(50 + seed)
    vox = RI.random_world(w, rng, 0.3)
    smap = O.shadow_build(w, vox)
    ids = rng.permutation(w.chunks).astype(np.uint32)
    recs, offs, counts = _remesh(gpu, w, dev(vox), ids, dev(smap))
    out = gpu.svl_nodes_all(w, recs, ids_dev(ids), offs, counts)
    torch.cuda.synchronize()
    assert gpu.status() == 0
    rh, oh, ch = host(recs), host(offs, np.uint64).astype(np.int64), host(counts, np.uint32)
    for l in range(5):
        exp, eoffs = O.svl_nodes(w, l, rh, ids, oh, ch.astype(np.int64))
        go = host(out[l][1], np.uint64)
        assert np.array_equal(go, eoffs), l
        assert np.array_equal(host(out[l][0])[: int(go[-1])], exp), l


@pytest.mark.parametrize("rb,mb,seed", [(2, (2, 1, 2), 1), (3, (3, 2, 1), 2), (4, (4, 1, 4), 3),
                                        (6, (2, 1, 2), 4)])
def test_svl_nodes_match_oracle(gpu, rb, mb, seed):
    w = abi.WorldDesc.make(rb, mb)
    rng = np.random.default_rng(seed)
    vox = RI.random_world(w, rng, 0.3)
    smap = O.shadow_build(w, vox)
    ids = rng.permutation(w.chunks)[: max(1, w.chunks - 3)].astype(np.uint32)
    ids = np.concatenate([ids, ids[:2]])
    recs, offs, counts = _remesh(gpu, w, dev(vox), ids, dev(smap))
    nodes = _nodes(gpu, w, recs, ids, offs, counts)
    rh, oh, ch = host(recs), host(offs, np.uint64).astype(np.int64), host(counts, np.uint32)
    for l in range(5):
        exp, eoffs = O.svl_nodes(w, l, rh, ids, oh, ch.astype(np.int64))
        got, goffs = nodes[l]
        go = host(goffs, np.uint64)
        assert np.array_equal(go, eoffs), l
        assert np.array_equal(host(got)[: int(go[-1])], exp), l


def test_svl_nodes_count_only_overflow_empty(gpu):
    import torch

    w = abi.WorldDesc.make(3, (2, 1, 2))
    rng = np.random.default_rng(9)
    vox = RI.random_world(w, rng, 0.4)
    smap = O.shadow_build(w, vox)
    ids = np.arange(w.chunks, dtype=np.uint32)
    recs, offs, counts = _remesh(gpu, w, dev(vox), ids, dev(smap))
    idd = ids_dev(ids)
    noffs = torch.empty(abi.node_count(w, 1) + 1, dtype=torch.int64, device="cuda")
    _, full = gpu.svl_nodes(w, 1, recs, idd, offs, counts)
    torch.cuda.synchronize()
    fo = host(full, np.uint64)
    # capacity for the first node only
    small = torch.empty((int(fo[1]), 4), dtype=torch.int16, device="cuda")
    got, go = gpu.svl_nodes(w, 1, recs, idd, offs, counts, capacity=int(fo[1]), out=(small, noffs))
    torch.cuda.synchronize()
    assert gpu.status() & abi.VP_STATUS_OVERFLOW
    assert np.array_equal(host(go, np.uint64), fo)
    exp, _ = O.svl_nodes(w, 1, host(recs), ids, host(offs, np.uint64).astype(np.int64),
                         host(counts, np.uint32).astype(np.int64))
    assert np.array_equal(host(got), exp[: int(fo[1])])
    # an empty list: every node empty
    _, eo = gpu.svl_nodes(w, 2, recs, ids_dev(np.zeros(0, np.uint32)), offs, counts)
    torch.cuda.synchronize()
    assert gpu.status() == 0 and int(host(eo, np.uint64).max()) == 0


def _frame_world(gpu):
    import torch

    cfg = W.CFG1
    w = cfg.desc()
    vox = torch.empty(w.chunks * w.R ** 3, dtype=torch.uint8, device="cuda")
    uni = torch.empty(w.chunks, dtype=torch.int16, device="cuda")
    gpu.generate(w, cfg.gen(), vox, uni)
    sx, sz = w.shadow_dims()
    smap = torch.empty(abi.shadow_padded_len(sx, sz), dtype=torch.int16, device="cuda")
    gpu.shadow_build(w, vox, smap)
    ids = np.arange(w.chunks, dtype=np.uint32)
    recs, offs, counts = _remesh(gpu, w, vox, ids, smap, uniform=uni)
    return w, ids, recs, offs, counts


# (location, yaw, pitch): above the middle, looking down; outside a corner,
# looking in; low over the terrain
FRAME_POSES = [((1024.0, 1500.0, 1024.0), 30.0, 60.0), ((-900.0, 700.0, -900.0), 135.0, 20.0),
               ((700.0, 300.0, 1500.0), 100.0, 15.0)]


def test_frame_draw_list_and_splat_match_oracle(gpu):
    """The 2048x256x2048 world: node buffers of all five levels against the
    oracle's; for three poses the host LOD selection, the GPU-assembled
    list (records and per-splat level) against the oracle's assembly, and
    the splat resolve of that list (160x90) against the oracle's, ids
    bit-exact and depth within 1e-5 relative."""
    import torch

    w, ids, recs, offs, counts = _frame_world(gpu)
    nodes = _nodes(gpu, w, recs, ids, offs, counts)
    # all five levels in one launch give the same buffers
    both = gpu.svl_nodes_all(w, recs, ids_dev(ids), offs, counts)
    torch.cuda.synchronize()
    assert gpu.status() == 0
    for l in range(5):
        assert torch.equal(both[l][1], nodes[l][1]), l
        n = int(nodes[l][1][-1].item())
        assert torch.equal(both[l][0][:n], nodes[l][0][:n]), l
    del both
    rh, oh = host(recs), host(offs, np.uint64).astype(np.int64)
    ch = host(counts, np.uint32).astype(np.int64)
    enodes = [O.svl_nodes(w, l, rh, ids, oh, ch) for l in range(5)]
    for l in range(5):
        assert np.array_equal(host(nodes[l][1], np.uint64), enodes[l][1]), l
        n = int(enodes[l][1][-1])
        assert np.array_equal(host(nodes[l][0])[:n], enodes[l][0]), l
    node_offs_h = [e[1] for e in enodes]
    levels_seen = set()
    for loc, yaw, pitch in FRAME_POSES:
        vp = W.make_view_projection(loc, yaw, pitch, 160, 90)
        planes = frustum_planes(vp)
        draws, mesh, total = octree_cull(w, planes, loc, node_offs_h)
        items = [np.diff(o.astype(np.int64)) for o in node_offs_h]
        edraws, _, etotal = O.octree_cull(w, planes, loc, items)
        assert [tuple(int(v) for v in d) for d in draws] == [tuple(d) for d in edraws]
        assert total == etotal and total > 0
        levels_seen |= {int(d["level"]) for d in draws}
        lst, lod = gpu.draw_gather(w, [n[0] for n in nodes], [n[1] for n in nodes], draws, total)
        torch.cuda.synchronize()
        assert gpu.status() == 0
        elst = np.concatenate([enodes[l][0][int(enodes[l][1][nd]):int(enodes[l][1][nd + 1])]
                               for l, nd, _ in edraws])
        elod = np.concatenate([np.full(int(enodes[l][1][nd + 1] - enodes[l][1][nd]), l, np.uint8)
                               for l, nd, _ in edraws])
        assert np.array_equal(host(lst)[:total], elst)
        assert np.array_equal(host(lod)[:total], elod)
        cam = W.make_camera(loc, yaw, pitch, 160, 90)
        depth, idb, _, _ = gpu.splat(lst[:total], cam, lod=lod[:total])
        torch.cuda.synchronize()
        edepth, eids, _, _ = O.splat_render(elst, cam, lod=elod)
        gi = host(idb, np.uint32)
        assert np.array_equal(gi, eids)
        hit = eids != 0xFFFFFFFF
        assert hit.mean() > 0.05
        gd = host(depth)
        assert np.allclose(gd[hit], edepth[hit], rtol=1e-5, atol=0)
    assert len(levels_seen) >= 3
