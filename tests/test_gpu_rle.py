"""GPU RLE codec (vp_gpu_rle) against the oracle's rle_compress /
rle_decompress: words, per-chunk offsets and round trips, bit-exact."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import refimpl as RI
from tests.gpuutil import dev, host, ids_dev
from voxplat_amd import abi
from voxplat_amd import workload as W

pytestmark = pytest.mark.gpu


def _oracle_stream(w, vox, ids):
    R3 = w.R ** 3
    parts = [O.rle_compress(vox[c * R3:(c + 1) * R3]) for c in ids]
    offs = np.zeros(len(ids) + 1, np.uint64)
    offs[1:] = np.cumsum([len(p) for p in parts])
    return np.concatenate(parts), offs


@pytest.mark.parametrize("rb,mb,density,seed", [(2, (2, 1, 1), 0.5, 1), (4, (2, 1, 2), 0.3, 2),
                                                (6, (2, 1, 1), 0.5, 3), (6, (1, 1, 2), 0.999, 4)])
def test_encode_decode_random(gpu, rb, mb, density, seed):
    import torch

    w = abi.WorldDesc.make(rb, mb)
    rng = np.random.default_rng(seed)
    vox = RI.random_world(w, rng, density, colours=(1, 4))
    vox[: w.R ** 3] = 0                                   # chunk 0 all air
    ids = rng.permutation(w.chunks).astype(np.uint32)
    ewords, eoffs = _oracle_stream(w, vox, ids)
    vd = dev(vox)
    words, offs, uni = gpu.rle_encode(w, vd, ids_dev(ids), capacity=len(ewords) + 16)
    torch.cuda.synchronize()
    assert gpu.status() == 0
    assert np.array_equal(host(offs, np.uint64), eoffs)
    assert np.array_equal(host(words, np.uint32)[:len(ewords)], ewords)
    R3 = w.R ** 3
    eu = [(-1 if (vox[c * R3:(c + 1) * R3] != vox[c * R3]).any() else int(vox[c * R3])) for c in ids]
    assert host(uni).tolist() == eu
    # decode into a scrambled buffer -> the original world
    out = torch.full_like(vd, 0xAB)
    gpu.rle_decode(w, words, offs, ids_dev(ids), out)
    torch.cuda.synchronize()
    assert gpu.status() == 0
    assert np.array_equal(host(out), vox)


def test_generated_world_roundtrip(gpu):
    import torch

    cfg = W.small_world(6, (2, 2, 2), 11)
    w = cfg.desc()
    vox = torch.empty(w.chunks * w.R ** 3, dtype=torch.uint8, device="cuda")
    gpu.generate(w, cfg.gen(), vox)
    ids = np.arange(w.chunks, dtype=np.uint32)
    _, coffs, cuni = gpu.rle_encode(w, vox, ids_dev(ids), count_only=True)
    words, offs, uni = gpu.rle_encode(w, vox, ids_dev(ids))     # sized by a counting pass
    torch.cuda.synchronize()
    assert gpu.status() == 0
    assert torch.equal(coffs, offs) and torch.equal(cuni, uni)
    assert words.numel() == int(offs[-1])
    out = torch.zeros_like(vox)
    gpu.rle_decode(w, words, offs, ids_dev(ids), out)
    torch.cuda.synchronize()
    vh = host(vox)
    ewords, eoffs = _oracle_stream(w, vh, ids)
    assert np.array_equal(host(offs, np.uint64), eoffs)
    assert np.array_equal(host(words, np.uint32)[:len(ewords)], ewords)
    assert np.array_equal(host(out), vh)


def test_bad_streams_and_overflow(gpu):
    import torch

    w = abi.WorldDesc.make(3, (1, 1, 1))
    vox = RI.random_world(w, np.random.default_rng(5), 0.5)
    ids = np.arange(w.chunks, dtype=np.uint32)
    ewords, eoffs = _oracle_stream(w, vox, ids)
    # too small a word buffer
    words, offs, _ = gpu.rle_encode(w, dev(vox), ids_dev(ids), capacity=int(eoffs[3]))
    torch.cuda.synchronize()
    assert gpu.status() & abi.VP_STATUS_OVERFLOW
    assert np.array_equal(host(offs, np.uint64), eoffs)
    assert np.array_equal(host(words, np.uint32)[:int(eoffs[3])], ewords[:int(eoffs[3])])
    # a stream one voxel short
    short = np.array([(7 << 24) | 511, 0], np.uint32)
    out = dev(np.zeros(w.chunks * 512, np.uint8))
    gpu.rle_decode(w, dev(short), dev(np.array([0, 2], np.uint64)), ids_dev([2]), out)
    torch.cuda.synchronize()
    assert gpu.status() & abi.VP_STATUS_BAD_RLE
    o = host(out)[2 * 512:3 * 512]
    assert (o[:511] == 7).all()
    # offsets past the buffer are rejected without touching memory
    gpu.rle_decode(w, dev(short), dev(np.array([0, 9], np.uint64)), ids_dev([1]), out)
    torch.cuda.synchronize()
    assert gpu.status() & abi.VP_STATUS_BAD_ID
