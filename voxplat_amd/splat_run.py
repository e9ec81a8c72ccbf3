"""Splat resolve of the 2048x256x2048 level-0 list at 1920x1080, for profilers.

    python -m voxplat_amd.splat_run [--poses P] [--rgb] [--plain]
                                    [--path tiles|global] [--check]

Generates the world on the GPU, remeshes it (level 0 only, chunk-index
order: 130.5 M splats), computes the list's block bounds once, then renders
the seeded poses one after another with vp_gpu_splat_blocks (--plain:
vp_gpu_splat), one synchronised frame each, and prints the per-frame times
measured with HIP events.  --path picks the screen-tile resolve or the
global per-fragment path (default); --check then draws every pose again by
the other path and compares the frames bit for bit.
"""
from __future__ import annotations

import argparse
import json
import sys

import torch

from . import abi
from . import workload as W
from ._lib import GPU


def level0_list(gpu: GPU, cfg: W.WorldConfig = W.CFG1):
    """(world, voxels, level-0 records in chunk-index order)."""
    w = cfg.desc()
    vox = torch.empty(w.chunks * w.R ** 3, dtype=torch.uint8, device="cuda")
    uni = torch.empty(w.chunks, dtype=torch.int16, device="cuda")
    gpu.generate(w, cfg.gen(), vox, uni)
    sx, sz = w.shadow_dims()
    smap = torch.empty(abi.shadow_padded_len(sx, sz), dtype=torch.int16, device="cuda")
    gpu.shadow_build(w, vox, smap)
    ids = torch.arange(w.chunks, dtype=torch.int32, device="cuda")
    recs, offs, _ = gpu.cull_mesh(w, vox, ids, shadow_map=smap, uniform=uni, lod=False)
    torch.cuda.synchronize()
    n = int(offs[-1].item()) // 4
    return w, vox, recs[:n].contiguous()


def poses_for(w: abi.WorldDesc, vox, n: int, width: int, height: int):
    R = w.R
    cols = vox.view(1 << w.max_bitw[2], 1 << w.max_bitw[1], 1 << w.max_bitw[0], R, R, R)

    def top(x, z):
        col = cols[z // R, :, x // R, z % R, :, x % R].reshape(-1)
        nz = torch.nonzero(col).reshape(-1)
        return int(nz.max().item()) + 1 if nz.numel() else 0
    return [W.make_camera(p, yw, pt, width, height) for p, yw, pt in W.splat_poses(top, n=n)]


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--poses", type=int, default=16)
    ap.add_argument("--rgb", action="store_true", help="also shade RGB")
    ap.add_argument("--plain", action="store_true",
                    help="vp_gpu_splat without block bounds (no culling or ordering)")
    ap.add_argument("--path", choices=("tiles", "global"), default="global")
    ap.add_argument("--check", action="store_true",
                    help="also draw every pose by the other path and compare the frames")
    a = ap.parse_args(argv)
    gpu = GPU(0)
    gpu.set_splat_tiles(a.path == "tiles")
    w, vox, l0 = level0_list(gpu)
    cams = poses_for(w, vox, a.poses, 1920, 1080)
    gpu.reserve(0, 1920 * 1080)
    depth = torch.empty((1080, 1920), dtype=torch.float32, device="cuda")
    ids = torch.empty((1080, 1920), dtype=torch.int32, device="cuda")
    rgb = torch.empty((1080, 1920, 3), dtype=torch.float32, device="cuda") if a.rgb else None
    out = (depth, ids, rgb, None)
    bounds = None if a.plain else gpu.splat_bounds(l0)
    gpu.splat(l0, cams[0], out=out, rgb=a.rgb, bounds=bounds)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ms, hit = [], []
    for cam in cams:
        e0.record()
        gpu.splat(l0, cam, out=out, rgb=a.rgb, bounds=bounds)
        e1.record()
        e1.synchronize()
        ms.append(e0.elapsed_time(e1))
        hit.append(float((ids != -1).float().mean()))
    same = None
    if a.check:
        same = True
        for cam in cams:
            gpu.set_splat_tiles(a.path == "tiles")
            gpu.splat(l0, cam, out=out, rgb=a.rgb, bounds=bounds)
            d0, i0 = depth.clone(), ids.clone()
            gpu.set_splat_tiles(a.path != "tiles")
            gpu.splat(l0, cam, out=out, rgb=a.rgb, bounds=bounds)
            same = same and bool(torch.equal(i0, ids)) and bool(torch.equal(d0, depth))
        gpu.set_splat_tiles(a.path == "tiles")
    print(json.dumps({"splats": int(l0.shape[0]), "path": a.path,
                      "frame_ms": [round(x, 4) for x in ms],
                      "mean_ms": round(sum(ms) / len(ms), 4),
                      "pixels_hit": round(sum(hit) / len(hit), 4),
                      "paths_identical": same, "status": gpu.status()}))
    gpu.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
