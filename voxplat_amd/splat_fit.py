"""Differentiable splat colours: fit per-voxel fp32 RGB to target views.

The forward is vp_gpu_splat with voxel colours overriding the 6-bit palette
(the shading of resources/splat.vert.glsl:24-32 and splat.frag.glsl:93-154);
the backward is vp_gpu_splat_backward, a scatter-add of weight * dL/drgb into
the colour of the nearest splat of each pixel.  For a fixed hit the pixel
colour is affine in that splat's colour:

    rgb(p) = (rgb[id(p)] * s1 * s2) * (1 - fog) + fog_colour * fog
    weight(p) = (s1 * s2) * (1 - fog)

with s1, s2 the two shading factors and fog the depth fog of splat.frag, none
of which depend on colours; the hits (ids, depths) do not either, so
d rgb(p) / d rgb[id(p)] = weight(p) exactly and the gradient is exact up to
fp32 rounding of the product.

    python -m voxplat_amd.splat_fit        # the 1024x256x1024 colour fit
"""
from __future__ import annotations

import json
import sys
import time

import torch

from . import abi
from . import workload as W
from ._lib import GPU


class SplatColour(torch.autograd.Function):
    """image[H, W, 3] = splat(rgb[n, 3]); gradient w.r.t. rgb only."""

    @staticmethod
    def forward(ctx, rgb, gpu: GPU, splats, cam: abi.Camera, lod=None, lod_log2: int = 0,
                bounds=None):
        if rgb.dim() != 2 or rgb.shape[1] != 3 or rgb.shape[0] != splats.shape[0]:
            raise ValueError(f"rgb must be [{splats.shape[0]}, 3], got {tuple(rgb.shape)}")
        rgb_c = rgb.detach().to(torch.float32).contiguous()
        _, ids, img, wt = gpu.splat(splats, cam, lod=lod, lod_log2=lod_log2, voxel_rgb=rgb_c,
                                    rgb=True, weight=True, bounds=bounds)
        ctx.save_for_backward(ids, wt)
        ctx.gpu, ctx.cam, ctx.n, ctx.dtype = gpu, cam, rgb.shape[0], rgb.dtype
        return img.to(rgb.dtype)

    @staticmethod
    def backward(ctx, g_img):
        ids, wt = ctx.saved_tensors
        grad = torch.zeros((ctx.n, 3), dtype=torch.float32, device=g_img.device)
        ctx.gpu.splat_backward(ctx.n, ctx.cam, ids, wt, g_img.to(torch.float32).contiguous(), grad)
        return grad.to(ctx.dtype), None, None, None, None, None, None


class SplatColourViews(torch.autograd.Function):
    """images[V, H, W, 3] = splat(rgb[n, 3]) from V cameras of one size.
    The backward scatters every view's dL/dimage into one [n, 3] gradient,
    so a step over V views holds one dense gradient instead of V."""

    @staticmethod
    def forward(ctx, rgb, gpu: GPU, splats, cams, lod=None, lod_log2: int = 0, bounds=None):
        if rgb.dim() != 2 or rgb.shape[1] != 3 or rgb.shape[0] != splats.shape[0]:
            raise ValueError(f"rgb must be [{splats.shape[0]}, 3], got {tuple(rgb.shape)}")
        W_, H_ = cams[0].width, cams[0].height
        if any(c.width != W_ or c.height != H_ for c in cams):
            raise ValueError("all views must have the same size")
        dev = rgb.device
        V = len(cams)
        rgb_c = rgb.detach().to(torch.float32).contiguous()
        depth = torch.empty((H_, W_), dtype=torch.float32, device=dev)
        ids = torch.empty((V, H_, W_), dtype=torch.int32, device=dev)
        img = torch.empty((V, H_, W_, 3), dtype=torch.float32, device=dev)
        wt = torch.empty((V, H_, W_), dtype=torch.float32, device=dev)
        for v, cam in enumerate(cams):
            gpu.splat(splats, cam, lod=lod, lod_log2=lod_log2, voxel_rgb=rgb_c, rgb=True,
                      weight=True, bounds=bounds, out=(depth, ids[v], img[v], wt[v]))
        ctx.save_for_backward(ids, wt)
        ctx.gpu, ctx.cams, ctx.n, ctx.dtype = gpu, cams, rgb.shape[0], rgb.dtype
        return img.to(rgb.dtype)

    @staticmethod
    def backward(ctx, g_img):
        ids, wt = ctx.saved_tensors
        g = g_img.to(torch.float32).contiguous()
        grad = torch.zeros((ctx.n, 3), dtype=torch.float32, device=g_img.device)
        for v, cam in enumerate(ctx.cams):
            ctx.gpu.splat_backward(ctx.n, cam, ids[v], wt[v], g[v], grad)
        return grad.to(ctx.dtype), None, None, None, None, None, None


def render_views(gpu: GPU, rgb, splats, cams, lod=None, lod_log2: int = 0, bounds=None):
    """All views at once ([V, H, W, 3]); one gradient buffer in the backward."""
    return SplatColourViews.apply(rgb, gpu, splats, tuple(cams), lod, lod_log2, bounds)


def render(gpu: GPU, rgb, splats, cam, lod=None, lod_log2: int = 0, bounds=None):
    """bounds: the list's block bounds (GPU.splat_bounds), computed once while
    the splats stay put; the resolve then culls and orders blocks per view."""
    return SplatColour.apply(rgb, gpu, splats, cam, lod, lod_log2, bounds)


def build_scene(gpu: GPU, cfg: W.WorldConfig = W.CFG4):
    """Generate the world on the GPU and return its level-0 splat list."""
    w = cfg.desc()
    dev = "cuda"
    vox = torch.empty(w.chunks * w.R ** 3, dtype=torch.uint8, device=dev)
    uni = torch.empty(w.chunks, dtype=torch.int16, device=dev)
    gpu.generate(w, cfg.gen(), vox, uni)
    sx, sz = w.shadow_dims()
    smap = torch.empty(abi.shadow_padded_len(sx, sz), dtype=torch.int16, device=dev)
    gpu.shadow_build(w, vox, smap)
    ids = torch.arange(w.chunks, dtype=torch.int32, device=dev)
    recs, offs, counts = gpu.cull_mesh(w, vox, ids, shadow_map=smap, uniform=uni, lod=False)
    torch.cuda.synchronize()
    n = int(offs[-1].item()) // 4
    return w, vox, recs[:n].contiguous()


def views_of(w: abi.WorldDesc, vox, n: int, size: int, seed: int):
    X, Y, Z = w.size
    mx, my, mz = (1 << w.max_bitw[0]), (1 << w.max_bitw[1]), (1 << w.max_bitw[2])
    R = w.R
    cols = vox.view(mz, my, mx, R, R, R)

    def top(x, z):
        col = cols[z // R, :, x // R, z % R, :, x % R].reshape(-1)
        nz = torch.nonzero(col).reshape(-1)
        return int(nz.max().item()) + 1 if nz.numel() else 0
    poses = W.splat_poses(top, n=n, seed=seed, lo=X * 0.125, hi=X * 0.875)
    return [W.make_camera(p, yw, pt, size, size) for p, yw, pt in poses]


def fit_colours(gpu: GPU, splats, cams, targets, init, steps: int = 200, lr: float = 0.05,
                log_every: int = 0, bounds=None):
    """Adam on per-splat RGB from `init`, loss = mean over views of the mean
    squared pixel error; colours are clamped to [0, 1] after each step.
    Returns (rgb, per-step loss list)."""
    rgb = init.detach().clone().to(torch.float32).requires_grad_(True)
    opt = torch.optim.Adam([rgb], lr=lr, betas=(0.9, 0.999))
    hist = torch.zeros(steps, dtype=torch.float64, device=init.device)
    tgt = torch.stack(list(targets))
    for it in range(steps):
        opt.zero_grad(set_to_none=True)
        img = render_views(gpu, rgb, splats, cams, bounds=bounds)
        # the mean over views of each view's mean squared error (equal sizes)
        l = ((img - tgt) ** 2).mean()
        l.backward()
        hist[it] += l.detach()
        opt.step()
        with torch.no_grad():
            rgb.clamp_(0.0, 1.0)
        if log_every and it % log_every == 0:
            print(f"step {it}: loss {float(hist[it]):.6f}", file=sys.stderr, flush=True)
    return rgb.detach(), hist.tolist()


def visible(gpu: GPU, splats, cams, bounds=None):
    """Mask of the splats that are the nearest hit of at least one pixel."""
    seen = torch.zeros(splats.shape[0], dtype=torch.bool, device=splats.device)
    for c in cams:
        _, ids, _, _ = gpu.splat(splats, c, bounds=bounds)
        hit = ids[ids >= 0]
        seen[hit.long()] = True
    return seen


def run_fit(gpu: GPU, views: int = 64, size: int = 512, steps: int = 200, lr: float = 0.05,
            log_every: int = 20) -> dict:
    """The 1024x256x1024 colour fit: seeded ground truth, uniform random
    initial colours, Adam over `views` target views.  Returns its figures."""
    cfg = W.CFG4
    t0 = time.perf_counter()
    w, vox, splats = build_scene(gpu, cfg)
    n = splats.shape[0]
    cams = views_of(w, vox, views, size, cfg.seed + 1)
    g = torch.Generator(device="cuda").manual_seed(cfg.seed)
    gt = torch.rand((n, 3), generator=g, device="cuda")
    init = torch.rand((n, 3), generator=g, device="cuda")
    bounds = gpu.splat_bounds(splats)
    with torch.no_grad():
        targets = [render(gpu, gt, splats, c, bounds=bounds) for c in cams]
    seen = visible(gpu, splats, cams, bounds)
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    rgb, hist = fit_colours(gpu, splats, cams, targets, init, steps=steps, lr=lr,
                            log_every=log_every, bounds=bounds)
    torch.cuda.synchronize()
    t_fit = time.perf_counter() - t0
    out = {
        "workload": f"colour fit, {cfg.name}, {views} views at {size}x{size}, "
                    f"Adam lr {lr}, {steps} steps",
        "splats": n, "visible_splats": int(seen.sum()), "steps": steps,
        "ms_per_step": round(1e3 * t_fit / steps, 3),
        "renders_per_s": round(steps * views / t_fit, 1),
        "loss_first": hist[0], "loss_last": hist[-1],
        "mean_abs_rgb_error_visible_init": round(float((init[seen] - gt[seen]).abs().mean()), 5),
        "mean_abs_rgb_error_visible": round(float((rgb[seen] - gt[seen]).abs().mean()), 5),
        "setup_s": round(t_setup, 2)}
    del splats, vox, targets, bounds
    return out


def main(argv=None) -> int:
    import argparse

    ap = argparse.ArgumentParser(description="per-splat colour fit to rendered target views")
    ap.add_argument("--views", type=int, default=64)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--lr", type=float, default=0.05)
    a = ap.parse_args(argv)
    gpu = GPU(0)
    print(json.dumps(run_fit(gpu, a.views, a.size, a.steps, a.lr)))
    gpu.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
