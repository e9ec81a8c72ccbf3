"""The frame's draw queue selected on the device (vp_gpu_octree_cull:
octree_cull, src/game.c:319-395, with gfx_fcull_visible, src/gfx/fcull.c:
38-45) against the host C selection, draw for draw, over the 16 seeded poses
of the 2048x256x2048 world, with and without quad-mesh chunks; then the
device-side frame (selection -> counted gather -> counted bounds and
resolve, no host read in between) against the host-driven frame, ids
bit-exact; and small worlds with the camera on a chunk centre, outside the
world, and looking away."""
import numpy as np
import pytest

from tests import refimpl as RI
from tests.gpuutil import dev, host, ids_dev
from voxplat_amd import abi, frustum_planes, octree_cull
from voxplat_amd import workload as W
from voxplat_amd._lib import _draw_dtype
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _world_nodes(gpu, w, vox_d, smap_d, uni_d=None):
    import torch

    ids = torch.arange(w.chunks, dtype=torch.int32, device="cuda")
    recs, offs, counts = gpu.cull_mesh(w, vox_d, ids, shadow_map=smap_d, uniform=uni_d)
    nodes = gpu.svl_nodes_all(w, recs, ids, offs, counts)
    torch.cuda.synchronize()
    assert gpu.status() == 0
    return nodes


def _same_queue(gpu, w, nodes, loc, planes, chunk_mesh=None):
    import torch

    offs_h = [host(n[1], np.uint64) for n in nodes]
    draws, mesh, total = octree_cull(w, planes, loc, offs_h, chunk_mesh)
    cm = None if chunk_mesh is None else dev(chunk_mesh.astype(np.uint8))
    gd, gm, counts = gpu.octree_cull(w, planes, loc, [n[1] for n in nodes], chunk_mesh=cm)
    torch.cuda.synchronize()
    assert gpu.status() == 0
    c = host(counts, np.uint64)
    assert int(c[0]) == len(draws) and int(c[1]) == len(mesh) and int(c[2]) == total
    got = np.ascontiguousarray(host(gd)).view(_draw_dtype()).reshape(-1)[: len(draws)]
    assert np.array_equal(got["level"], draws["level"])
    assert np.array_equal(got["node"], draws["node"])
    assert np.array_equal(got["first"], draws["first"])
    assert np.array_equal(host(gm, np.uint32)[: len(mesh)], mesh)
    return draws, mesh, total


def test_device_queue_matches_host_16_poses(gpu):
    import torch

    cfg = W.CFG1
    w = cfg.desc()
    vox = torch.empty(w.chunks * w.R ** 3, dtype=torch.uint8, device="cuda")
    uni = torch.empty(w.chunks, dtype=torch.int16, device="cuda")
    gpu.generate(w, cfg.gen(), vox, uni)
    sx, sz = w.shadow_dims()
    smap = torch.empty(abi.shadow_padded_len(sx, sz), dtype=torch.int16, device="cuda")
    gpu.shadow_build(w, vox, smap)
    nodes = _world_nodes(gpu, w, vox, smap, uni)
    cols = vox.view(32, 4, 32, 64, 64, 64)

    def top(x, z):
        col = cols[z >> 6, :, x >> 6, z & 63, :, x & 63].reshape(-1)
        nz = torch.nonzero(col).reshape(-1)
        return int(nz.max().item()) + 1 if nz.numel() else 0
    poses = W.splat_poses(top, n=16)
    # plus poses high above and outside the world, where levels 2-4 are drawn
    far = [((1024.0, 1500.0, 1024.0), 30.0, 60.0), ((-900.0, 700.0, -900.0), 135.0, 20.0),
           ((1024.0, 2500.0, 1024.0), 0.0, 89.0), ((-3000.0, 1500.0, 1024.0), 90.0, 15.0),
           ((4500.0, 1000.0, 1024.0), 270.0, 10.0)]
    rng = np.random.default_rng(3)
    levels = set()
    for k, (loc, yaw, pitch) in enumerate(poses + far):
        planes = frustum_planes(W.make_view_projection(loc, yaw, pitch, 1920, 1080))
        mesh = (rng.random(w.chunks) < 0.1) if k % 4 == 3 else None
        draws, _, total = _same_queue(gpu, w, nodes, loc, planes, mesh)
        assert total > 0
        levels |= {int(d) for d in draws["level"]}
    assert levels == {0, 1, 2, 3, 4}

    # the device frame: selection, counted gather, counted bounds + resolve
    loc, yaw, pitch = poses[5]
    cam = W.make_camera(loc, yaw, pitch, 480, 270)
    planes = frustum_planes(W.make_view_projection(loc, yaw, pitch, 480, 270))
    offs_h = [host(n[1], np.uint64) for n in nodes]
    draws, _, total = octree_cull(w, planes, loc, offs_h)
    lst, lod = gpu.draw_gather(w, [n[0] for n in nodes], [n[1] for n in nodes], draws, total)
    d0, i0, _, _ = gpu.splat(lst[:total], cam, lod=lod[:total])
    cap = 40_000_000
    depth, ids, got_total = gpu.draw_frame(w, [n[0] for n in nodes], [n[1] for n in nodes],
                                           planes, loc, cam, draw_cap=65536,
                                           record_capacity=cap)
    torch.cuda.synchronize()
    assert gpu.status() == 0
    assert got_total == total
    assert torch.equal(ids, i0)
    assert torch.equal(depth, d0)


@pytest.mark.parametrize("rb,mb,seed", [(2, (4, 2, 4), 1), (3, (3, 1, 5), 2), (4, (5, 1, 5), 3)])
def test_device_queue_small_worlds(gpu, rb, mb, seed):
    """A camera on a chunk centre (level-0 leaves), outside the world looking
    in, and looking away from it (an empty queue)."""
    w = abi.WorldDesc.make(rb, mb)
    rng = np.random.default_rng(seed)
    vox = RI.random_world(w, rng, 0.2)
    smap = O.shadow_build(w, vox)
    nodes = _world_nodes(gpu, w, dev(vox), dev(smap))
    R = w.R
    X, Y, Z = w.size
    cams = [((1.5 * R, 0.5 * R, 2.5 * R), 30.0, 20.0),             # on a chunk centre
            ((-3.0 * X, 2.0 * Y, -3.0 * Z), 135.0, 20.0),          # outside, looking in
            ((X / 2, Y / 2, Z / 2), 90.0, -80.0),                   # inside, looking up
            ((-3.0 * X, 2.0 * Y, -3.0 * Z), 315.0, 20.0)]           # outside, looking away
    for loc, yaw, pitch in cams:
        planes = frustum_planes(W.make_view_projection(loc, yaw, pitch, 320, 180))
        _same_queue(gpu, w, nodes, loc, planes)
        _same_queue(gpu, w, nodes, loc, planes, rng.random(w.chunks) < 0.3)
