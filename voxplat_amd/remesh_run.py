"""Plain launches of the batched remesh on a generated world, for profilers.

    python -m voxplat_amd.remesh_run [--world 2k|16k] [--launches K] [--copy-gib G]
                                     [--kernel cull|mesh|nodes]

Generates the world on the GPU, puts the chunk list in locality order and
sizes the output with counting passes, then issues K direct (not
graph-captured) launches of vp_gpu_cull_mesh over every chunk, so that a
kernel tracer or counter collector sees one dispatch per launch.
--copy-gib G first runs one device-to-device copy of G GiB: a kernel of
known traffic (G GiB read, G GiB written) against which HBM byte counters
are scaled.  --kernel mesh issues vp_gpu_mesh (the near-field quad mesh)
over every chunk in index order instead, output sized by a counting call;
--kernel nodes remeshes once and then builds the octree node buffers of
the five LOD levels K times one launch per level (vp_gpu_svl_nodes), then K
times in one launch (vp_gpu_svl_nodes_all).
"""
from __future__ import annotations

import argparse
import json
import sys

import torch

from . import abi
from . import workload as W
from ._lib import GPU
from .world_frame import locality_order, mesh_bytes


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--world", choices=("2k", "16k"), default="2k")
    ap.add_argument("--launches", type=int, default=3)
    ap.add_argument("--copy-gib", type=float, default=0.0)
    ap.add_argument("--autotune", action="store_true", help="pick the launch shape first")
    ap.add_argument("--kernel", choices=("cull", "mesh", "nodes"), default="cull")
    a = ap.parse_args(argv)
    gpu = GPU(0)
    if a.copy_gib:
        n = int(a.copy_gib * 2 ** 30)
        src = torch.ones(n, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(f"copy: {n} bytes", file=sys.stderr, flush=True)
        del src, dst
    cfg = W.CFG1 if a.world == "2k" else W.CFG3
    w = cfg.desc()
    vox = torch.empty(w.chunks * w.R ** 3, dtype=torch.uint8, device="cuda")
    uni = torch.empty(w.chunks, dtype=torch.int16, device="cuda")
    gpu.generate(w, cfg.gen(), vox, uni)
    sx, sz = w.shadow_dims()
    smap = torch.empty(abi.shadow_padded_len(sx, sz), dtype=torch.int16, device="cuda")
    gpu.shadow_build(w, vox, smap)
    ids = torch.arange(w.chunks, dtype=torch.int32, device="cuda")
    if a.kernel == "mesh":
        return run_mesh(gpu, cfg, w, vox, uni, smap, ids, a.launches)
    ids, offs, counts = locality_order(gpu, w, vox, ids, smap, uni)
    torch.cuda.synchronize()
    if a.kernel == "nodes":
        return run_nodes(gpu, cfg, w, vox, uni, smap, ids, offs, counts, a.launches)
    cap = int(offs[-1].item()) // 4
    splats = torch.empty((max(cap, 1), 4), dtype=torch.int16, device="cuda")
    if a.autotune:   # the launch shape the benchmark would pick
        t = gpu.autotune_remesh(w, vox, ids, smap, splats, cap, offs, counts, uniform=uni,
                                reps=3)
        print(f"launch shape: {t['block']} threads x {t['wg_per_cu']} per CU, "
              f"{t['records_in_flight']} records in flight ({t['shapes_timed']} shapes timed)",
              file=sys.stderr, flush=True)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ms = []
    for _ in range(a.launches):
        e0.record()
        gpu.cull_mesh(w, vox, ids, shadow_map=smap, uniform=uni, capacity=cap,
                      out=(offs, counts), records=splats)
        e1.record()
        e1.synchronize()
        ms.append(e0.elapsed_time(e1))
    st = gpu.status()
    print(json.dumps({"world": cfg.name, "chunks": w.chunks, "records": cap,
                      "launch_ms": [round(x, 4) for x in ms], "status": st}))
    gpu.close()
    return 0 if st == 0 else 1


def run_nodes(gpu, cfg, w, vox, uni, smap, ids, offs, counts, launches: int) -> int:
    recs, _, _ = gpu.cull_mesh(w, vox, ids, shadow_map=smap, uniform=uni,
                               capacity=int(offs[-1].item()) // 4, out=(offs, counts))
    nodes = [gpu.svl_nodes(w, l, recs, ids, offs, counts) for l in range(5)]
    torch.cuda.synchronize()
    caps = [int(n[1][-1].item()) for n in nodes]
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ms, ms_all = [], []
    for _ in range(launches):         # one launch per level
        e0.record()
        for l in range(5):
            gpu.svl_nodes(w, l, recs, ids, offs, counts, capacity=caps[l], out=nodes[l])
        e1.record()
        e1.synchronize()
        ms.append(e0.elapsed_time(e1))
    for _ in range(launches):         # all five levels in one launch
        e0.record()
        gpu.svl_nodes_all(w, recs, ids, offs, counts, out=nodes)
        e1.record()
        e1.synchronize()
        ms_all.append(e0.elapsed_time(e1))
    st = gpu.status()
    print(json.dumps({"world": cfg.name, "kernel": "k_svl_nodes", "records": sum(caps),
                      "per_level_records": caps, "five_launches_ms": [round(x, 4) for x in ms],
                      "one_launch_ms": [round(x, 4) for x in ms_all], "status": st}))
    gpu.close()
    return 0 if st == 0 else 1


def run_mesh(gpu, cfg, w, vox, uni, smap, ids, launches: int) -> int:
    _, _, offs = gpu.mesh(w, vox, ids, smap, uniform=uni, count_only=True)
    torch.cuda.synchronize()
    faces = int(offs[-1].item())
    geom = torch.empty((max(faces, 1), 4, 4), dtype=torch.int16, device="cuda")
    index = torch.empty((max(faces, 1), 6), dtype=torch.int32, device="cuda")
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ms = []
    for _ in range(launches):
        e0.record()
        gpu.mesh(w, vox, ids, smap, uniform=uni, capacity=faces, out=(geom, index, offs))
        e1.record()
        e1.synchronize()
        ms.append(e0.elapsed_time(e1))
    st = gpu.status()
    codes = ((geom[:faces, 0, 3].to(torch.int32) >> 8) & 7).cpu().numpy()
    algo = mesh_bytes(uni.cpu().numpy(), codes, w.R)
    best = min(ms)
    print(json.dumps({"world": cfg.name, "kernel": "k_mesh", "chunks": w.chunks, "faces": faces,
                      "launch_ms": [round(x, 4) for x in ms], "algorithmic_bytes": algo,
                      "gb_s_best": round(algo / (best * 1e-3) / 1e9, 1), "status": st}))
    gpu.close()
    return 0 if st == 0 else 1


if __name__ == "__main__":
    sys.exit(main())
