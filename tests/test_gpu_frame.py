"""The per-frame pass captured once and replayed (vp_host_frame_graph_*,
chunkset_manage's walk, src/chunkset.c:246-313): over several edit bursts
through one graph, the dirty and idle lists are the chunks whose dirty byte
was set and the others (ascending), the bytes are cleared, the remesh of the
dirty list equals vp_gpu_cull_mesh and the oracle, and the streams of the
idle list equal vp_gpu_rle and the oracle.  A frame with no edits remeshes
nothing."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.gpuutil import host
from voxplat_amd import workload as W

pytestmark = pytest.mark.gpu


def test_frame_graph_bursts(gpu):
    import torch

    from voxplat_amd import world_frame as WF

    cfg = W.small_world(5, (3, 2, 3), 0x5EED0F00, hmax=90)
    w = cfg.desc()
    world = WF.ResidentWorld(gpu, cfg)
    world.generate()
    ids = torch.arange(w.chunks, dtype=torch.int32, device="cuda")
    _, offs, _ = gpu.cull_mesh(w, world.voxels, ids, shadow_map=world.smap,
                               uniform=world.uniform, count_only=True)
    _, woffs, _ = gpu.rle_encode(w, world.voxels, ids, count_only=True)
    torch.cuda.synchronize()
    fr = WF.Frame(world, 2 * int(offs[-1].item()) // 4, 2 * int(woffs[-1].item()))
    try:
        for f in range(4):
            print(f"[gpu-case] frame graph burst {f}", flush=True)
            spheres = W.edit_burst(w, n=0 if f == 3 else 40, seed=300 + f, rmin=2, rmax=14)
            if spheres:
                world.edit(spheres)
            exp = host(world.dirty).astype(bool)
            marked_mixed = host(world.uniform) < 0
            fr.run()
            fr.sync()
            assert gpu.status() == 0
            assert not host(world.dirty).any()                       # flags cleared
            assert np.array_equal(host(fr.dirty), np.flatnonzero(exp).astype(np.int32))
            # idle: the chunks the uniform table marks mixed, then the uniform ones
            want_idle = np.concatenate([np.flatnonzero(~exp & marked_mixed),
                                        np.flatnonzero(~exp & ~marked_mixed)]).astype(np.int32)
            assert np.array_equal(host(fr.idle), want_idle)
            nd = int(fr.dirty.numel())
            assert (nd > 0) == (f != 3)
            vh, mh = host(world.voxels), host(world.smap, np.uint16)
            # remesh: the direct call on the same list, and the oracle
            if nd:
                recs, offs2, counts2 = gpu.cull_mesh(w, world.voxels, fr.dirty.clone(),
                                                     shadow_map=world.smap, uniform=world.uniform)
                torch.cuda.synchronize()
                assert torch.equal(offs2, fr.offsets) and torch.equal(counts2, fr.counts)
                got = host(fr.splats[:fr.records])
                assert np.array_equal(got, host(recs[:fr.records]))
                erecs, eoffs, ecounts = O.remesh_list(w, vh, host(fr.dirty).view(np.uint32), mh,
                                                      threads=8)
                assert np.array_equal(got, erecs)
            else:
                assert int(fr.offsets[0]) == 0
            # RLE of the idle list: the direct call, and the oracle on a sample
            words, woffs2, wuni = gpu.rle_encode(w, world.voxels, fr.idle.clone())
            torch.cuda.synchronize()
            assert torch.equal(woffs2, fr.word_offsets)
            assert torch.equal(words[:fr.words_n], fr.words[:fr.words_n])
            assert torch.equal(wuni, fr.idle_uniform)
            R3 = w.R ** 3
            idle_h = host(fr.idle)
            for k in np.linspace(0, len(idle_h) - 1, 12).astype(int):
                c = int(idle_h[k])
                ew = O.rle_compress(vh[c * R3:(c + 1) * R3])
                assert np.array_equal(host(fr.chunk_words(int(k))).view(np.uint32), ew)
    finally:
        fr.close()
