"""vp_gpu_cull_mesh_noalloc (work-queue scratch in a static device array,
no allocation) against vp_gpu_cull_mesh on the same inputs, through the C
ABI: records, offsets and level counts bit-exact for two calls in a row, the
second on a smaller list (the array keeps the first call's contents), plus a
device-side count and the oracle on the smaller list.  A list longer than
the static scratch is refused before anything is launched."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from tests.gpuutil import dev, host, ids_dev
from voxplat_amd import abi
from voxplat_amd import workload as W

pytestmark = pytest.mark.gpu


def _run(gpu, w, vox, ids, smap, uni, noalloc, d_n=None):
    import torch

    recs, offs, counts = gpu.cull_mesh(w, vox, ids, shadow_map=smap, uniform=uni, d_n=d_n,
                                       noalloc=noalloc)
    torch.cuda.synchronize()
    assert gpu.status() == 0
    offs_h = host(offs, np.uint64)
    total = int(offs_h[-1]) // 4
    return host(recs)[:total], offs_h, host(counts, np.uint32)


@pytest.fixture(scope="module")
def world1(gpu):
    import torch

    cfg = W.CFG1
    w = cfg.desc()
    vox = torch.empty(w.chunks * w.R ** 3, dtype=torch.uint8, device="cuda")
    uni = torch.empty(w.chunks, dtype=torch.int16, device="cuda")
    gpu.generate(w, cfg.gen(), vox, uni)
    sx, sz = w.shadow_dims()
    smap = torch.empty(abi.shadow_padded_len(sx, sz), dtype=torch.int16, device="cuda")
    gpu.shadow_build(w, vox, smap)
    torch.cuda.synchronize()
    return w, vox, uni, smap


def test_noalloc_equals_cull_mesh_two_calls(gpu, world1):
    w, vox, uni, smap = world1
    rng = np.random.default_rng(5)
    first = ids_dev(rng.permutation(w.chunks))                       # all 4096 chunks
    small = np.sort(rng.choice(w.chunks, 300, replace=False)).astype(np.uint32)
    for ids in (first, ids_dev(small)):
        print(f"[gpu-case] noalloc: {ids.numel()} chunks", flush=True)
        got = _run(gpu, w, vox, ids, smap, uni, noalloc=True)
        ref = _run(gpu, w, vox, ids, smap, uni, noalloc=False)
        assert len(got[0]) > 1000
        for a, b in zip(got, ref):
            assert np.array_equal(a, b)
    # the smaller list against the oracle as well
    erecs, eoffs, ecounts = O.remesh_list(w, host(vox), small, host(smap, np.uint16), threads=8)
    assert np.array_equal(got[0], erecs)
    assert np.array_equal(got[1], eoffs)
    assert np.array_equal(got[2], ecounts)


def test_noalloc_device_count(gpu, world1):
    """d_n below n: only the first *d_n entries are remeshed."""
    w, vox, uni, smap = world1
    ids = ids_dev(np.arange(w.chunks, dtype=np.uint32)[::-1].copy())
    d_n = dev(np.array([777], np.uint32))
    got = _run(gpu, w, vox, ids, smap, uni, noalloc=True, d_n=d_n)
    ref = _run(gpu, w, vox, ids[:777].contiguous(), smap, uni, noalloc=False)
    n = 777
    assert np.array_equal(got[1][:n + 1], ref[1])
    assert np.array_equal(got[2][:n], ref[2])
    assert np.array_equal(got[0][:len(ref[0])], ref[0])


def test_noalloc_refuses_lists_beyond_its_scratch(gpu, world1):
    import torch

    w, vox, uni, smap = world1
    ids = torch.zeros(4, dtype=torch.int32, device="cuda")
    offs = torch.zeros(8, dtype=torch.int64, device="cuda")
    counts = torch.zeros(20, dtype=torch.int32, device="cuda")
    so = abi.SplatOut()
    so.d_splats, so.capacity_records = None, 0
    so.d_chunk_offsets, so.d_level_counts = offs.data_ptr(), counts.data_ptr()
    sh = abi.ShadowDesc()
    rc = gpu.lib.vp_gpu_cull_mesh_noalloc(
        gpu.ctx, C.byref(w), C.c_void_p(vox.data_ptr()), C.c_void_p(uni.data_ptr()),
        C.byref(sh), C.c_void_p(ids.data_ptr()), abi.VP_NOALLOC_MAX_CHUNKS + 1, None,
        abi.VP_CM_COUNT_ONLY, C.byref(so), None)
    assert rc == abi.VP_ERR_CAPACITY
    with pytest.raises(ValueError):
        gpu.cull_mesh(w, vox, torch.zeros(abi.VP_NOALLOC_MAX_CHUNKS + 1, dtype=torch.int32,
                                           device="cuda"), shadow_map=smap, noalloc=True)
