"""GPU workload builders against the oracle: generator, classification,
shadow-map build (a different algorithm from the oracle's literal placement
order), sphere edits and their host-side shadow updates."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import refimpl as RI
from tests.gpuutil import dev, host
from voxplat_amd import abi, host_shadow_spheres
from voxplat_amd import workload as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cfg", [W.CFG0, W.small_world(4, (2, 2, 1), 5), W.small_world(6, (2, 1, 2), 9)])
def test_generate_and_classify(gpu, cfg):
    import torch

    w = cfg.desc()
    vox = torch.zeros(w.chunks * w.R ** 3, dtype=torch.uint8, device="cuda")
    uni = torch.empty(w.chunks, dtype=torch.int16, device="cuda")
    gpu.generate(w, cfg.gen(), vox, uni)
    uni2 = torch.empty_like(uni)
    gpu.classify(w, vox, uni2)
    torch.cuda.synchronize()
    exp = O.gen_world(w, cfg.gen())
    assert np.array_equal(host(vox), exp)
    eu = np.array([(-1 if (c != c[0]).any() else int(c[0])) for c in exp.reshape(w.chunks, -1)])
    assert np.array_equal(host(uni), eu) and np.array_equal(host(uni2), eu)


@pytest.mark.parametrize("rb,mb,density,seed", [(2, (2, 2, 2), 0.5, 1), (3, (2, 1, 2), 0.3, 2),
                                                (6, (2, 2, 1), 0.6, 3), (6, (1, 4, 1), 0.4, 4)])
def test_shadow_build_matches_serial_order(gpu, rb, mb, density, seed):
    import torch

    w = abi.WorldDesc.make(rb, mb)
    vox = RI.random_world(w, np.random.default_rng(seed), density)
    sx, sz = w.shadow_dims()
    smap = torch.full((abi.shadow_padded_len(sx, sz),), 77, dtype=torch.int16, device="cuda")
    gpu.shadow_build(w, dev(vox), smap)
    torch.cuda.synchronize()
    assert np.array_equal(host(smap, np.uint16), O.shadow_build(w, vox, serial=True))


def test_shadow_build_generated_cfg1(gpu):
    import torch

    cfg = W.CFG1
    w = cfg.desc()
    vox = torch.empty(w.chunks * w.R ** 3, dtype=torch.uint8, device="cuda")
    gpu.generate(w, cfg.gen(), vox)
    sx, sz = w.shadow_dims()
    smap = torch.empty(abi.shadow_padded_len(sx, sz), dtype=torch.int16, device="cuda")
    gpu.shadow_build(w, vox, smap)
    torch.cuda.synchronize()
    exp = O.shadow_build(w, host(vox), serial=False, threads=16)
    assert np.array_equal(host(smap, np.uint16), exp)


def test_edit_spheres_match_sequential_edits(gpu):
    import torch

    cfg = W.small_world(4, (2, 2, 2), 31)
    w = cfg.desc()
    vox_h = O.gen_world(w, cfg.gen())
    smap_h = O.shadow_build(w, vox_h)
    spheres = W.edit_burst(w, n=60, seed=5, rmin=1, rmax=9)
    vox = dev(vox_h)
    dirty = torch.zeros(w.chunks, dtype=torch.uint8, device="cuda")
    uni = torch.zeros(w.chunks, dtype=torch.int16, device="cuda")
    gpu.edit_spheres(w, vox, spheres, dirty=dirty, uniform=uni)
    torch.cuda.synchronize()
    m_gpu = smap_h.copy()
    host_shadow_spheres(w, m_gpu, spheres)
    exp_dirty = np.zeros(w.chunks, np.uint8)
    for s in spheres:
        O.edit_sphere(w, vox_h, smap_h, (s.cx, s.cy, s.cz), s.radius, s.voxel, exp_dirty)
    assert np.array_equal(host(vox), vox_h)
    assert np.array_equal(host(dirty), exp_dirty)
    assert np.array_equal(host(uni)[exp_dirty == 1], np.full(exp_dirty.sum(), -1))
    assert np.array_equal(m_gpu, smap_h)


@pytest.mark.parametrize("rb,mb,n,rmax,seed", [(4, (2, 2, 2), 60, 9, 5), (3, (3, 1, 3), 200, 12, 6),
                                               (6, (2, 1, 2), 300, 32, 7)])
def test_device_shadow_spheres_match_sequential_order(gpu, rb, mb, n, rmax, seed):
    """vp_gpu_shadow_spheres against the oracle's sequential edits (chunk
    list cx, cy, cz; then x, y, z): the whole map, bit-exact.  Spheres reach
    past the world on every side, and overlap."""
    cfg = W.small_world(rb, mb, 31 + seed)
    w = cfg.desc()
    vox_h = O.gen_world(w, cfg.gen())
    smap_h = O.shadow_build(w, vox_h)
    spheres = W.edit_burst(w, n=n, seed=seed, rmin=1, rmax=rmax)
    md = dev(smap_h)
    gpu.shadow_spheres(w, md, spheres)
    for s in spheres:
        O.edit_sphere(w, vox_h, smap_h, (s.cx, s.cy, s.cz), s.radius, s.voxel)
    assert np.array_equal(host(md, np.uint16), smap_h)
    assert gpu.status() == 0


def test_device_shadow_spheres_cfg1_burst(gpu):
    """2000 spheres (r 1-32) on the 2048x256x2048 world: device map equals the
    host C placement in the same order."""
    import torch

    cfg = W.CFG1
    w = cfg.desc()
    vd = torch.empty(w.chunks * w.R ** 3, dtype=torch.uint8, device="cuda")
    gpu.generate(w, cfg.gen(), vd)
    sx, sz = w.shadow_dims()
    md = torch.empty(abi.shadow_padded_len(sx, sz), dtype=torch.int16, device="cuda")
    gpu.shadow_build(w, vd, md)
    spheres = W.edit_burst(w, n=2000, seed=11)
    m_host = host(md, np.uint16).copy()
    gpu.shadow_spheres(w, md, spheres)
    host_shadow_spheres(w, m_host, spheres)
    assert np.array_equal(host(md, np.uint16), m_host)
