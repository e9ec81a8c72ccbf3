"""The on-disk RLE world on the GPU (vp_host_world_save / _load over
vp_gpu_rle) against the oracle's file (command_export,
src/deadcode.c:320-346; chunkset_clear_import, src/deadcode.c:159-229):
the written bytes equal the oracle's exactly, and loading gives back the
voxels, the uniform table and the shadow map.  Worlds of more chunks than one
batch, the 2048x256x2048 world, and damaged files (errors, no partial
success)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import refimpl as RI
from tests.gpuutil import dev, host
from voxplat_amd import VPError, abi
from voxplat_amd import workload as W

pytestmark = pytest.mark.gpu


def _world(rb, mb, seed, density):
    w = abi.WorldDesc.make(rb, mb)
    rng = np.random.default_rng(seed)
    vox = RI.random_world(w, rng, density)
    R3 = w.R ** 3
    for c in rng.choice(w.chunks, max(1, w.chunks // 4), replace=False):
        vox[c * R3:(c + 1) * R3] = rng.integers(0, 3) * 20     # uniform chunks (air or solid)
    return w, vox


def _expected_uniform(w, vox):
    R3 = w.R ** 3
    ch = vox.reshape(w.chunks, R3)
    same = (ch == ch[:, :1]).all(1)
    return np.where(same, ch[:, 0].astype(np.int16), np.int16(-1))


@pytest.mark.parametrize("rb,mb,seed,density", [
    (2, (5, 3, 5), 1, 0.5),      # 8192 chunks: two batches
    (3, (2, 2, 1), 2, 0.3),
    (5, (1, 1, 2), 3, 0.6),
    (6, (1, 0, 1), 4, 0.05),
])
def test_save_bytes_and_load_round_trip(gpu, tmp_path, rb, mb, seed, density):
    import torch

    w, vox = _world(rb, mb, seed, density)
    smap = O.shadow_build(w, vox)
    vd, md = dev(vox), dev(smap)
    for with_map in (False, True):
        p = tmp_path / f"w{with_map}.bin"
        n = gpu.world_save(p, w, vd, shadow_map=md if with_map else None)
        data = p.read_bytes()
        assert n == len(data)
        assert data == O.world_file(w, vox, smap if with_map else None)
        out = torch.full_like(vd, 0xA5)
        uni = torch.full((w.chunks,), 7, dtype=torch.int16, device="cuda")
        mout = torch.full_like(md, 0x5A5A) if with_map else None
        gpu.world_load(p, w, out, uniform=uni, shadow_map=mout)
        torch.cuda.synchronize()
        assert np.array_equal(host(out), vox)
        assert np.array_equal(host(uni), _expected_uniform(w, vox))
        if with_map:
            assert np.array_equal(host(mout, np.uint16), smap)      # map and zero padding
    assert gpu.status() == 0


def test_world2k_round_trip(gpu, tmp_path):
    """The 2048x256x2048 world: bytes equal the oracle's, load is exact."""
    import torch

    cfg = W.CFG1
    w = cfg.desc()
    vd = torch.empty(w.chunks * w.R ** 3, dtype=torch.uint8, device="cuda")
    ud = torch.empty(w.chunks, dtype=torch.int16, device="cuda")
    gpu.generate(w, cfg.gen(), vd, ud)
    sx, sz = w.shadow_dims()
    md = torch.empty(abi.shadow_padded_len(sx, sz), dtype=torch.int16, device="cuda")
    gpu.shadow_build(w, vd, md)
    torch.cuda.synchronize()
    p = tmp_path / "world2k.bin"
    n = gpu.world_save(p, w, vd, shadow_map=md)
    vh, mh = host(vd), host(md, np.uint16)
    data = p.read_bytes()
    assert n == len(data)
    assert data == O.world_file(w, vh, mh)
    out = torch.zeros_like(vd)
    uni = torch.zeros_like(ud)
    mout = torch.zeros_like(md)
    gpu.world_load(p, w, out, uniform=uni, shadow_map=mout)
    torch.cuda.synchronize()
    assert torch.equal(out, vd)
    assert torch.equal(mout, md)
    # every generated uniform chunk is found uniform (the stream is one run)
    assert torch.equal(uni, ud)


def test_damaged_files_are_errors(gpu, tmp_path):
    import torch

    w, vox = _world(3, (1, 1, 1), 9, 0.4)
    smap = O.shadow_build(w, vox)
    good = O.world_file(w, vox, smap)
    out = torch.zeros(w.chunks * w.R ** 3, dtype=torch.uint8, device="cuda")
    md = torch.zeros(len(smap), dtype=torch.int16, device="cuda")
    cases = {
        "truncated": good[:len(good) // 3],
        "trailer": good + b"\1\0\0\0",
        "other_world": good[:9] + bytes([2]) + good[10:],
    }
    for name, data in cases.items():
        p = tmp_path / f"{name}.bin"
        p.write_bytes(data)
        with pytest.raises(VPError) as e:
            gpu.world_load(p, w, out, shadow_map=md)
        assert e.value.code == abi.VP_ERR_DATA, name
    # no map in the file, one asked for
    p = tmp_path / "nomap.bin"
    p.write_bytes(O.world_file(w, vox))
    with pytest.raises(VPError) as e:
        gpu.world_load(p, w, out, shadow_map=md)
    assert e.value.code == abi.VP_ERR_DATA
    gpu.world_load(p, w, out)            # without the map it loads
    torch.cuda.synchronize()
    assert np.array_equal(host(out), vox)
    assert gpu.status() == 0
