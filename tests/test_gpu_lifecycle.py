"""The chunk lifecycle on the 2048x256x2048 world: edit, one frame (the frame
graph: remesh of the dirty chunks, RLE encode of the idle ones), then every idle
chunk dropped and opened again from its stream (the reference's
chunk_compress / chunk_open_*, src/chunkset.c:133-230,300-306).

Checked: the reopened world equals the world before the drop, byte for byte,
and so does its uniform metadata; a remesh of the reopened chunks gives the
same records, offsets and counts as before the drop; a second frame over the
reopened world reproduces the first frame's records and streams."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.gpuutil import host
from voxplat_amd import workload as W

pytestmark = pytest.mark.gpu


def test_drop_and_reopen_idle_chunks(gpu):
    import torch
    from voxplat_amd import world_frame as WF

    print("[gpu-case] lifecycle: 2048x256x2048, edits, frame, drop + reopen idle", flush=True)
    world = WF.ResidentWorld(gpu, W.CFG1)
    world.generate()
    ids = torch.arange(world.w.chunks, dtype=torch.int32, device="cuda")
    _, offs, _ = gpu.cull_mesh(world.w, world.voxels, ids, shadow_map=world.smap,
                               uniform=world.uniform, count_only=True)
    _, woffs, _ = gpu.rle_encode(world.w, world.voxels, ids, count_only=True)
    torch.cuda.synchronize()
    fr = WF.Frame(world, int(offs[-1].item()) // 4, 2 * int(woffs[-1].item()))
    spheres = W.edit_burst(world.w, n=200, seed=7)
    world.edit(spheres)
    fr.run()
    fr.sync()
    dirty, idle = fr.dirty.clone(), fr.idle.clone()
    assert dirty.numel() > 0 and idle.numel() > 0
    assert gpu.status() == 0
    vox0, uni0 = world.voxels.clone(), world.uniform.clone()
    recs0, words0 = fr.splats[:fr.records].clone(), fr.words[:fr.words_n].clone()
    # encode's per-entry uniform values are the world's metadata of those chunks
    assert torch.equal(fr.idle_uniform, world.uniform[idle.long()])

    # remesh a list of idle chunks before the drop
    probe = idle[:: max(1, idle.numel() // 512)].contiguous()
    r0, o0, c0 = gpu.cull_mesh(world.w, world.voxels, probe, shadow_map=world.smap,
                               uniform=world.uniform)
    torch.cuda.synchronize()

    # drop: every idle chunk's voxels overwritten, its metadata made wrong
    R3 = world.R3
    v = world.voxels.view(-1, R3)
    v.index_fill_(0, idle.long(), 0x55)
    world.uniform.index_fill_(0, idle.long(), 0x55)
    torch.cuda.synchronize()
    assert not torch.equal(world.voxels, vox0)

    world.open(idle, fr.words, fr.word_offsets, fr.idle_uniform)
    torch.cuda.synchronize()
    assert gpu.status() == 0
    assert torch.equal(world.voxels, vox0)
    assert torch.equal(world.uniform, uni0)

    r1, o1, c1 = gpu.cull_mesh(world.w, world.voxels, probe, shadow_map=world.smap,
                               uniform=world.uniform)
    torch.cuda.synchronize()
    assert torch.equal(o1, o0) and torch.equal(c1, c0)
    n = int(o0[-1].item()) // 4
    assert torch.equal(r1[:n], r0[:n])

    # a second frame over the reopened world, the same chunks marked dirty,
    # reproduces the first
    world.dirty[dirty.long()] = 1
    fr.run()
    fr.sync()
    assert gpu.status() == 0
    assert torch.equal(fr.dirty, dirty) and torch.equal(fr.idle, idle)
    assert torch.equal(fr.splats[:fr.records], recs0)
    assert torch.equal(fr.words[:fr.words_n], words0)

    # a few reopened chunks against the oracle's RLE of the same voxels
    for k in range(0, int(idle.numel()), max(1, int(idle.numel()) // 6)):
        c = int(idle[k])
        a, b = int(fr.word_offsets[k]), int(fr.word_offsets[k + 1])
        exp = O.rle_compress(host(world.chunk(c)))
        assert np.array_equal(host(fr.words[a:b], np.uint32), exp), c
    fr.close()
    del world, fr, vox0
    torch.cuda.empty_cache()
