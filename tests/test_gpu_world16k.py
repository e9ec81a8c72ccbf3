"""The 16384x512x16384 resident world (524,288 chunks, 128 GiB in HBM):
generate, then three frames, each a burst of 10k sphere edits (voxels and
shadow map on the device) followed by one replay of the frame graph captured
once: dirty/idle lists on the device, batched remesh of the dirty chunks, RLE
encode of every other chunk.

Checked at full size, on the last frame:
  * the graph's dirty list equals every chunk of every box of the last
    burst's spheres (edit.c:98-112,241), and the dirty bytes are cleared;
  * every idle chunk's RLE stream decodes on the GPU back to its voxels;
  * a direct count pass over the dirty list agrees with every offset;
and against the oracle on sampled chunks (the world does not fit the host):
  * remesh records and level counts of 32 dirty chunks, bit-exact;
  * RLE streams of 32 idle chunks, bit-exact;
  * the voxels of 12 dirty chunks, regenerated on the CPU and edited by the
    spheres of all bursts that reach them, in order, bit-exact."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.gpuutil import host
from voxplat_amd import abi
from voxplat_amd import workload as W

pytestmark = pytest.mark.gpu


def _expected_dirty(w, spheres) -> np.ndarray:
    b = w.root_bitw
    mx = [1 << m for m in w.max_bitw]
    d = np.zeros((mx[2], mx[1], mx[0]), bool)
    for s in spheres:
        c = (s.cx, s.cy, s.cz)
        lo = [max(0, (c[i] - s.radius - 1) >> b) for i in range(3)]
        hi = [min(mx[i] - 1, (c[i] + s.radius + 1) >> b) for i in range(3)]
        if any(lo[i] > hi[i] for i in range(3)):
            continue
        d[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] = True
    return d.reshape(-1)


@pytest.fixture(scope="module")
def world16k(gpu):
    import torch
    from voxplat_amd import world_frame as WF

    print("[gpu-case] world16k: generate 16384x512x16384, 3 frames of 10k edits", flush=True)
    res = WF.run_world16k(frames=3, edits=10000, keep=True, gpu=gpu)
    world, fr, every, last = res.pop("_objects")
    print(f"[gpu-case] world16k: {res}", flush=True)
    yield res, world, fr, (every, last)
    fr.close()
    del world, fr
    torch.cuda.empty_cache()


def test_frame_status_and_dirty_set(world16k):
    res, world, fr, (every, last) = world16k
    assert res["status"] == 0
    assert all(f["status"] == 0 for f in res["frames"])
    ed = _expected_dirty(world.w, last)
    assert np.array_equal(host(fr.dirty), np.flatnonzero(ed).astype(np.int32))
    assert not host(world.dirty).any()
    f = res["frames"][-1]
    assert f["dirty_chunks"] == int(ed.sum())
    assert f["dirty_chunks"] + f["idle_chunks"] == world.w.chunks
    # the bursts differ, and so do their dirty sets
    assert len({p["dirty_chunks"] for p in res["frames"]}) > 1


def test_remesh_offsets_match_count_pass(gpu, world16k):
    import torch

    res, world, fr, (every, last) = world16k
    # the graph's emission rewrote offsets/counts; a fresh count pass must agree
    _, offs, counts = gpu.cull_mesh(world.w, world.voxels, fr.dirty, shadow_map=world.smap,
                                    uniform=world.uniform, count_only=True)
    torch.cuda.synchronize()
    assert torch.equal(offs, fr.offsets) and torch.equal(counts, fr.counts)
    assert int(fr.counts.sum()) == 4 * fr.records


def test_remesh_sampled_chunks_match_oracle(world16k):
    from voxplat_amd import world_frame as WF

    res, world, fr, (every, last) = world16k
    w = world.w
    dirty_h = host(fr.dirty).astype(np.uint32)
    rng = np.random.default_rng(0x16C)
    ks = np.sort(rng.choice(len(dirty_h), 32, replace=False))
    hw = WF.sparse_host_world(w)
    for k in ks:
        WF.copy_chunks_to_host(world, hw, WF.plus_neighbours(w, int(dirty_h[k])))
    smap = host(world.smap).view(np.uint16)
    recs, offs, counts = O.remesh_list(w, hw, dirty_h[ks], smap, lod=True, threads=8)
    gcounts = host(fr.counts).view(np.uint32)
    for i, k in enumerate(ks):
        assert np.array_equal(gcounts[k], counts[i]), f"chunk {dirty_h[k]}"
        a, b = int(offs[i]) // 4, int(offs[i + 1]) // 4
        assert np.array_equal(host(fr.chunk_records(int(k))), recs[a:b]), f"chunk {dirty_h[k]}"


def test_rle_sampled_chunks_match_oracle(world16k):
    res, world, fr, (every, last) = world16k
    idle_h = host(fr.idle).astype(np.uint32)
    rng = np.random.default_rng(0x16D)
    uni = host(fr.idle_uniform)
    for k in np.sort(rng.choice(len(idle_h), 32, replace=False)):
        c = int(idle_h[k])
        vox = host(world.chunk(c))
        ew = O.rle_compress(vox)
        assert np.array_equal(host(fr.chunk_words(int(k))).view(np.uint32), ew), f"chunk {c}"
        eu = -1 if (vox != vox[0]).any() else int(vox[0])
        assert int(uni[k]) == eu


def test_rle_every_idle_chunk_decodes_to_its_voxels(gpu, world16k):
    """All 480k+ streams, decoded 4096 chunks at a time into a 1 GiB scratch
    world and compared with the resident voxels."""
    import torch

    res, world, fr, (every, last) = world16k
    R3 = world.R3
    m = 4096
    scratch_w = abi.WorldDesc.make(6, (5, 2, 5))
    scratch = torch.empty(m * R3, dtype=torch.uint8, device="cuda")
    local = torch.arange(m, dtype=torch.int32, device="cuda")
    vox2 = world.voxels.view(-1, R3)
    n = int(fr.idle.numel())
    for j in range(0, n, m):
        k = min(m, n - j)
        gpu.rle_decode(scratch_w, fr.words, fr.word_offsets[j:j + k + 1], local[:k], scratch)
        want = vox2.index_select(0, fr.idle[j:j + k].long())
        assert torch.equal(scratch[:k * R3].view(k, R3), want), f"idle chunks {j}..{j + k}"
    torch.cuda.synchronize()
    assert gpu.status() == 0


def test_edits_sampled_chunks_match_oracle(world16k):
    from voxplat_amd import world_frame as WF

    res, world, fr, (every, last) = world16k
    w = world.w
    b = w.root_bitw
    dirty_h = host(fr.dirty).astype(np.uint32)
    rng = np.random.default_rng(0x16E)
    hw = WF.sparse_host_world(w)
    params = world.cfg.gen()
    mxb = w.max_bitw
    for k in np.sort(rng.choice(len(dirty_h), 12, replace=False)):
        c = int(dirty_h[k])
        cc = (c & ((1 << mxb[0]) - 1), (c >> mxb[0]) & ((1 << mxb[1]) - 1), c >> (mxb[0] + mxb[1]))
        R3 = world.R3
        hw[c * R3:(c + 1) * R3] = O.gen_chunk(w, params, c)
        for s in every:
            ctr = (s.cx, s.cy, s.cz)
            if all(((ctr[i] - s.radius - 1) >> b) <= cc[i] <= ((ctr[i] + s.radius + 1) >> b)
                   for i in range(3)):
                O.edit_sphere(w, hw, None, ctr, s.radius, s.voxel)
        assert np.array_equal(host(world.chunk(c)), hw[c * R3:(c + 1) * R3]), f"chunk {c}"
