"""The splat workloads at their stated sizes, against the oracle.

Config 2: the level-0 list of the 2048x256x2048 world (130.5 M splats, chunk
index order) resolved at 1920x1080 for two of the 16 seeded poses, by the
plain resolve and by the block-culled one: ids bit-exact, depth within 1e-5
relative (vsplat.c:104-193, splat.vert.glsl, splat.frag.glsl).

Config 4: the colour fit of the 1024x256x1024 world (32 M level-0 splats)
with 64 views at 512x512: the GPU gradient against the oracle's fp64
backward on 4 views, finite differences on 1000 sampled voxel colours over
all 64 views, a fast-mode gradcheck, and the 200-step Adam fit with its
final colour error."""
import os
import warnings

import numpy as np
import pytest

from oracle import oracle as O
from tests.gpuutil import dev, host
from voxplat_amd import workload as W

pytestmark = pytest.mark.gpu

DEPTH_RTOL = 1e-5
THREADS = min(16, os.cpu_count() or 4)


@pytest.fixture(scope="module")
def cfg2(gpu):
    from voxplat_amd import splat_run as S

    w, vox, l0 = S.level0_list(gpu)
    cams = S.poses_for(w, vox, 16, 1920, 1080)
    return l0, host(l0), cams


@pytest.mark.parametrize("pose", [0, 9])
def test_cfg2_1080p_pose_matches_oracle(gpu, cfg2, pose):
    import torch

    l0, l0h, cams = cfg2
    assert l0.shape[0] > 100_000_000
    cam = cams[pose]
    ed, ei, _, _ = O.splat_render(l0h, cam, threads=THREADS)
    hit = ei != 0xFFFFFFFF
    assert hit.mean() > 0.3
    bounds = gpu.splat_bounds(l0)
    for b in (None, bounds):
        depth, ids, _, _ = gpu.splat(l0, cam, bounds=b)
        torch.cuda.synchronize()
        assert gpu.status() == 0
        gi = host(ids, np.uint32)
        assert np.array_equal(gi, ei), f"{(gi != ei).sum()} ids differ (blocks={b is not None})"
        assert np.allclose(host(depth), ed, rtol=DEPTH_RTOL, atol=0)


@pytest.fixture(scope="module")
def cfg4(gpu):
    import torch

    from voxplat_amd import splat_fit as F

    cfg = W.CFG4
    w, vox, splats = F.build_scene(gpu, cfg)
    cams = F.views_of(w, vox, 64, 512, cfg.seed + 1)
    n = splats.shape[0]
    g = torch.Generator(device="cuda").manual_seed(cfg.seed)
    gt = torch.rand((n, 3), generator=g, device="cuda")
    init = torch.rand((n, 3), generator=g, device="cuda")
    bounds = gpu.splat_bounds(splats)
    return splats, cams, gt, init, bounds


def test_cfg4_gradient_matches_oracle_backward(gpu, cfg4):
    """Forward image and dL/drgb of L = sum(G * image) on 4 of the 64 views,
    against the oracle's render and fp64 backward."""
    from voxplat_amd import splat_fit as F

    splats, cams, gt, init, bounds = cfg4
    n = splats.shape[0]
    assert n > 30_000_000
    sh = host(splats)
    rng = np.random.default_rng(41)
    rgb = host(init).astype(np.float32)
    for v in (0, 21, 42, 63):
        cam = cams[v]
        G = rng.standard_normal((cam.height, cam.width, 3)).astype(np.float32)
        x = dev(rgb).requires_grad_(True)
        img = F.render(gpu, x, splats, cam, bounds=bounds)
        img.backward(dev(G))
        _, eids, er, ew = O.splat_render(sh, cam, voxel_rgb=rgb, rgb=True, weight=True,
                                         threads=THREADS)
        assert np.allclose(host(img), er, rtol=1e-5, atol=1e-5)
        eg = O.splat_backward(n, cam, eids, ew, G)
        gg = host(x.grad)
        nz = np.nonzero(np.abs(eg).sum(1))[0]
        # close-up views see a few thousand voxels, far ones tens of thousands
        assert len(nz) > 1000
        assert np.allclose(gg, eg, rtol=1e-5, atol=1e-5)


def test_cfg4_finite_differences_1000_voxels(gpu, cfg4):
    """Central differences of L = sum over the 64 views of sum(G_v * image_v)
    in the colours of 1000 sampled visible voxels.  A pixel shows exactly one
    voxel (its nearest hit, which colours do not move), so perturbing all
    sampled voxels at once and summing each pixel's change into the voxel it
    shows gives every voxel's own difference quotient."""
    import torch

    from voxplat_amd import splat_fit as F

    splats, cams, gt, init, bounds = cfg4
    n = splats.shape[0]
    seen = F.visible(gpu, splats, cams, bounds)
    vis = torch.nonzero(seen).reshape(-1)
    g = torch.Generator(device="cuda").manual_seed(8)
    sample = vis[torch.randperm(vis.numel(), generator=g, device="cuda")[:1000]]
    mask = torch.zeros(n, dtype=torch.bool, device="cuda")
    mask[sample] = True
    rgb = init.double()
    h = 1e-2
    grad = torch.zeros((n, 3), dtype=torch.float64, device="cuda")
    fd = torch.zeros((n, 3), dtype=torch.float64, device="cuda")
    steps = []
    for c in range(3):
        step = torch.zeros((n, 3), dtype=torch.float64, device="cuda")
        step[mask, c] = h
        steps.append(step)
    for v, cam in enumerate(cams):
        G = torch.randn((cam.height, cam.width, 3), generator=g, device="cuda",
                        dtype=torch.float64)
        x = rgb.clone().requires_grad_(True)
        (F.render(gpu, x, splats, cam, bounds=bounds) * G).sum().backward()
        grad += x.grad
        _, ids, _, _ = gpu.splat(splats, cam, bounds=bounds)
        ids = ids.reshape(-1).long()
        on = ids >= 0
        for c, step in enumerate(steps):
            with torch.no_grad():
                ip = F.render(gpu, rgb + step, splats, cam, bounds=bounds)
                im = F.render(gpu, rgb - step, splats, cam, bounds=bounds)
            dpix = ((ip - im) * G).sum(-1).reshape(-1) / (2 * h)
            fd[:, c].index_add_(0, ids[on], dpix[on])
    a, b = grad[sample], fd[sample]
    assert float(a.abs().max()) > 1.0
    assert torch.allclose(a, b, rtol=1e-3, atol=1e-3), float((a - b).abs().max())


def test_cfg4_gradcheck_sampled_voxels(gpu, cfg4):
    """torch.autograd.gradcheck (fast mode) of the image of one view as a
    function of the colours of 1000 sampled voxels, the others held fixed."""
    import torch

    from voxplat_amd import splat_fit as F

    splats, cams, gt, init, bounds = cfg4
    n = splats.shape[0]
    cam = cams[5]
    seen = F.visible(gpu, splats, [cam], bounds)
    vis = torch.nonzero(seen).reshape(-1)
    g = torch.Generator(device="cuda").manual_seed(9)
    sample = vis[torch.randperm(vis.numel(), generator=g, device="cuda")[:1000]]
    base = init.double()

    def image(sub):
        full = base.index_put((sample,), sub)
        return F.render(gpu, full, splats, cam, bounds=bounds)
    sub = base[sample].clone().requires_grad_(True)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        assert torch.autograd.gradcheck(image, (sub,), eps=1e-2, atol=1e-3, rtol=1e-3,
                                        nondet_tol=1e-5, fast_mode=True)


def test_cfg4_adam_200_steps(gpu, cfg4):
    """The stated fit: Adam (lr 0.05, betas 0.9/0.999), 200 steps over the 64
    views; the loss falls by 6 orders of magnitude and the visible colours
    come back to the ground truth."""
    import torch

    from voxplat_amd import splat_fit as F

    splats, cams, gt, init, bounds = cfg4
    with torch.no_grad():
        targets = [F.render(gpu, gt, splats, c, bounds=bounds) for c in cams]
    seen = F.visible(gpu, splats, cams, bounds)
    assert int(seen.sum()) > 500_000
    rgb, hist = F.fit_colours(gpu, splats, cams, targets, init, steps=200, lr=0.05,
                              bounds=bounds)
    assert hist[-1] < 1e-6 * hist[0], (hist[0], hist[-1])
    err0 = float((init[seen] - gt[seen]).abs().mean())
    err = float((rgb[seen] - gt[seen]).abs().mean())
    assert err0 > 0.3 and err < 1e-3, (err0, err)
    assert torch.equal(rgb[~seen], init[~seen].clamp(0, 1))
