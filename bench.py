#!/usr/bin/env python3
"""Benchmark of the voxel pipeline on one MI355X.

Headline (one step): a full remesh -- cull, level-0 splat list and LOD levels
1..4 -- of every chunk of the 16384x512x16384 uint8 world (524,288 chunks of
64^3, 128 GiB resident in HBM, seeded value-noise terrain with 5% caves), as
one batched persistent kernel captured in a hipGraph, the list in locality
order (mixed chunks in descending chunk-index order, the lightest of them
last, then the uniform chunks).  Inputs are resident in HBM before timing starts.

Also measured (extra keys, not the headline value):
  * edit_frame: on the same world, frames of 10k sphere edits each followed by
    one replay of the frame graph (dirty/idle lists, remesh, RLE);
  * world2k: the 2048x256x2048 world -- batched remesh against the plain
    per-chunk path (one launch per chunk from the C host loop), the splat
    resolve of its level-0 list at 1920x1080 over 16 seeded poses, the
    near-field quad mesh, the octree draw lists, one 64^3 chunk alone;
  * fit: the 1024x256x1024 colour fit (64 views at 512x512, 200 Adam steps);
  * roofline of the remesh kernel from HIP events on its launch stream;
  * cpu_baseline: the CPU oracle (4 OpenMP threads, as chunkset_manage's
    team) on 512 seeded chunks of the same world.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import pathlib
import re
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
VOXELS_PER_CHUNK = 64 ** 3


def log(msg: str) -> None:
    sys.stderr.write(f"[bench] {msg}\n")
    sys.stderr.flush()


def algorithmic_bytes(uniform, counts) -> int:
    """Bytes the remesh must move per launch: chunk + 3 halo planes (mixed) or
    the 3 halo planes alone (uniform), 8 B per record written, 4 B per record
    of shadow-map reads, 5 u32 counts per chunk."""
    import numpy as np

    mixed = int((uniform < 0).sum())
    uni = len(uniform) - mixed
    records = int(counts.astype(np.int64).sum()) // 4
    return mixed * (262144 + 3 * 4096) + uni * 3 * 4096 + 12 * records + 20 * len(uniform)


def measured_traffic(world: str):
    """HBM bytes per launch of the remesh kernel on `world` (e.g. "16384x512x16384"),
    measured with rocprofv3 --pmc (committed summary)."""
    p = ROOT / "profiles" / f"cull_mesh_{world}_traffic.md"
    if not p.exists():
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9]+)", p.read_text())
    return int(m.group(1)) if m else None


def leg_world16k(args) -> dict:
    """The headline: `steps` graph launches remeshing all 524,288 chunks of the
    16384x512x16384 world, then the edit-burst frames; the CPU baseline on a
    seeded sample of the same world's chunks."""
    from voxplat_amd import world_frame as WF

    def cpu_leg(world):
        """The oracle on 512 seeded chunks of the world (their + neighbours
        copied into a sparse host image), timed on host cores."""
        if args.no_cpu:
            return None
        import numpy as np

        from oracle import oracle as O

        w = world.w
        rng = np.random.default_rng(0xC0FFEE)
        sample = np.sort(rng.choice(w.chunks, 512, replace=False)).astype(np.uint32)
        hw = WF.sparse_host_world(w)
        for c in sample:
            WF.copy_chunks_to_host(world, hw, WF.plus_neighbours(w, int(c)))
        smap = world.smap.cpu().numpy().view(np.uint16)
        O.remesh_time(w, hw, sample[:16], smap, args.cpu_threads)       # warm
        t, _ = O.remesh_time(w, hw, sample, smap, args.cpu_threads)
        res = {"value": round(len(sample) * VOXELS_PER_CHUNK / t / 1e9, 4), "unit": "Gvoxels/s",
               "cores": args.cpu_threads, "kind": "port",
               "sample": f"512 seeded chunks of the 524,288 (uniform sample, L0..L4 with "
                         f"shadows), {args.cpu_threads} OpenMP threads as chunkset_manage "
                         f"(host has {os.cpu_count()} CPUs); {t:.2f} s"}
        log(f"cpu: {res['sample']}: {res['value']} Gvoxel/s")
        return res

    res = WF.run_world16k(frames=args.frames, edits=10000, log=log, full_steps=args.steps,
                          full_warmup=args.warmup, on_generated=cpu_leg,
                          autotune=not args.no_autotune)
    full = res["full_remesh"]
    gvox = full["gvoxels_s"]
    return {
        "metric": "Gvoxels/s remeshed (batched vs per-chunk), splats/s at 1080p, % HBM peak",
        "value": gvox, "unit": "Gvoxels/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": full["ms"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u8",
        "data": "synthetic (seeded value-noise terrain, 5% caves)",
        "config": {"workload": "remesh L0-L4 of 16384x512x16384 (524,288 chunks of 64^3, "
                               "128 GiB resident), batched persistent kernel in a hipGraph, "
                               "list in locality order",
                   "chunks": full["chunks"], "occupancy": res["occupancy"],
                   "splat_records": full["records"], "launch": full["launch"]},
        "roofline": {"bound": "hbm", "achieved": full["gb_s"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(full["gb_s"] / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic("16384x512x16384"),
                     "algorithmic_bytes": full["algorithmic_bytes"],
                     "kernel_ms": full["kernel_ms"]},
        "cpu_baseline": res["on_generated"],
        "edit_frame": {k: v for k, v in res.items()
                       if k not in ("full_remesh", "on_generated", "occupancy")},
    }


def visible_splats(splats, cam, step: int = 1 << 24) -> int:
    """Splats of a level-0 list whose centre is inside the view volume (the
    vertex stage's clip test, in float64 here: a count for the report, not a
    parity check)."""
    import numpy as np
    import torch

    T = np.array(cam.view_translation, np.float64).reshape(4, 4).T
    PR = np.array(cam.proj_view_rot, np.float64).reshape(4, 4).T
    M = torch.from_numpy(PR @ T).to(splats.device)
    n = 0
    for a in range(0, int(splats.shape[0]), step):
        c = splats[a:a + step, :3].to(torch.float64) + 0.5
        clip = c @ M[:, :3].T + M[:, 3]
        w = clip[:, 3:4]
        n += int(((clip[:, :3] >= -w) & (clip[:, :3] <= w)).all(1).sum().item())
    return n


def leg_world2k(args) -> dict:
    """The 2048x256x2048 legs (see the module docstring)."""
    import numpy as np
    import torch

    from voxplat_amd import GPU, RemeshGraph, abi
    from voxplat_amd import workload as W

    dev = "cuda:0"
    gpu = GPU(0)
    stream = torch.cuda.current_stream()
    cfg = W.CFG1
    w = cfg.desc()
    nch = w.chunks
    log(f"generating {cfg.name}: {nch} chunks")
    vox = torch.empty(nch * VOXELS_PER_CHUNK, dtype=torch.uint8, device=dev)
    uni = torch.empty(nch, dtype=torch.int16, device=dev)
    gpu.generate(w, cfg.gen(), vox, uni)
    sx, sz = w.shadow_dims()
    smap = torch.empty(abi.shadow_padded_len(sx, sz), dtype=torch.int16, device=dev)
    gpu.shadow_build(w, vox, smap)
    # the list is in locality order (the output follows the list; every
    # chunk's records are found through its offset)
    from voxplat_amd.world_frame import locality_order

    ids, offs, counts = locality_order(gpu, w, vox, torch.arange(nch, dtype=torch.int32, device=dev),
                                       smap, uni)
    torch.cuda.synchronize()
    total_items = int(offs[-1].item())
    cap = total_items // 4
    splats = torch.empty((cap, 4), dtype=torch.int16, device=dev)
    uni_h = uni.cpu().numpy()
    counts_h = counts.cpu().numpy().view(np.uint32)
    occupancy = {"uniform_air": int((uni_h == 0).sum()), "uniform_solid": int((uni_h > 0).sum()),
                 "mixed": int((uni_h < 0).sum())}
    log(f"occupancy {occupancy}, {cap} splats (L0..L4)")

    tune = None
    if not args.no_autotune:
        tune = gpu.autotune_remesh(w, vox, ids, smap, splats, cap, offs, counts, uniform=uni,
                                   reps=3)
        tune["ms"] = round(tune["ms"], 4)
        log(f"autotune ({tune['shapes_timed']} resident shapes): block {tune['block']} x "
            f"{tune['wg_per_cu']}/CU, {tune['records_in_flight']} records in flight -> "
            f"{tune['ms']:.3f} ms")
    graph = RemeshGraph(gpu, w, vox, ids, smap, offs, counts, splats, cap, uniform=uni)

    # ---- batched, graph replay: the headline ----
    for _ in range(args.warmup):
        graph.launch()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        graph.launch()
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ev_ms = e0.elapsed_time(e1) / args.steps
    ms_per_step = max(wall * 1e3 / args.steps, ev_ms)
    gvox = nch * VOXELS_PER_CHUNK / (ms_per_step * 1e-3) / 1e9
    assert gpu.status() == 0
    log(f"batched: {ms_per_step:.3f} ms/step ({ev_ms:.3f} ms by events) -> {gvox:.1f} Gvoxel/s")

    # ---- plain per-chunk path (C host loop, one launch per chunk) ----
    offs_h = offs.cpu().numpy().view(np.uint64)
    rec_off = (offs_h[:-1] // 4).astype(np.uint64)
    rec_cap = ((offs_h[1:] - offs_h[:-1]) // 4).astype(np.uint64)
    splats_pc = torch.empty_like(splats)
    offs2 = torch.empty(2 * nch, dtype=torch.int64, device=dev)
    counts_pc = torch.empty_like(counts)
    gpu.remesh_per_chunk(w, vox, ids, smap, splats_pc, rec_off, rec_cap, offs2, counts_pc,
                         uniform=uni)
    torch.cuda.synchronize()
    reps = 2
    t0 = time.perf_counter()
    for _ in range(reps):
        gpu.remesh_per_chunk(w, vox, ids, smap, splats_pc, rec_off, rec_cap, offs2, counts_pc,
                             uniform=uni)
    torch.cuda.synchronize()
    pc_ms = (time.perf_counter() - t0) * 1e3 / reps
    pc_gvox = nch * VOXELS_PER_CHUNK / (pc_ms * 1e-3) / 1e9
    same = bool(torch.equal(splats_pc, splats)) and bool(torch.equal(counts_pc, counts))
    log(f"per-chunk: {pc_ms:.2f} ms/pass -> {pc_gvox:.2f} Gvoxel/s; identical output: {same}")

    # ---- roofline of the remesh kernel ----
    algo = algorithmic_bytes(uni_h, counts_h)
    achieved = algo / (ev_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": measured_traffic("2048x256x2048"), "algorithmic_bytes": algo,
                "kernel_ms": round(ev_ms, 4)}

    # ---- splat resolve at 1080p (level-0 list, chunk order) ----
    splat_res = None
    splat_cpu = None
    if not args.no_splat:
        # the level-0 records in chunk-index order
        ids_h = ids.cpu().numpy()
        starts = (offs_h[:-1] // 4).astype(np.int64)
        lens = (counts_h[:, 0] // 4).astype(np.int64)
        by_chunk = np.argsort(ids_h, kind="stable")
        starts, lens = starts[by_chunk], lens[by_chunk]
        first = np.repeat(starts - np.concatenate([[0], np.cumsum(lens)[:-1]]), lens)
        l0_idx = torch.from_numpy(first + np.arange(int(lens.sum()), dtype=np.int64)).to(dev)
        l0 = splats.index_select(0, l0_idx).contiguous()
        n0 = int(l0.shape[0])
        vh_cols = vox.view(32, 4, 32, 64, 64, 64)          # [cz, cy, cx, z, y, x]

        def top(x, z):
            col = vh_cols[z >> 6, :, x >> 6, z & 63, :, x & 63].reshape(-1)
            nz = torch.nonzero(col).reshape(-1)
            return int(nz.max().item()) + 1 if nz.numel() else 0
        poses = W.splat_poses(top, n=16)
        cams = [W.make_camera(p, yw, pt, 1920, 1080) for p, yw, pt in poses]
        gpu.reserve(nch, 1920 * 1080)
        depth = torch.empty((1080, 1920), dtype=torch.float32, device=dev)
        idb = torch.empty((1080, 1920), dtype=torch.int32, device=dev)
        out = (depth, idb, None, None)
        # block bounds of the list, once per list (the culling data a draw list carries)
        bounds = gpu.splat_bounds(l0)
        gpu.splat(l0, cams[0], out=out, bounds=bounds)
        torch.cuda.synchronize()
        times, hits = [], []
        for cam in cams:
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            gpu.splat(l0, cam, out=out, bounds=bounds)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
            hits.append(float((idb != -1).float().mean().item()))
        fr_ms = 1e3 * sum(times) / len(times)
        vis = sum(visible_splats(l0, cam) for cam in cams) / len(cams)
        splat_res = {"splats": n0, "resolution": "1920x1080", "poses": len(cams),
                     "ms_per_frame": round(fr_ms, 3),
                     "list_splats_per_s": round(n0 / (fr_ms * 1e-3), 1),
                     "visible_splats_mean": int(vis),
                     "visible_splats_per_s": round(vis / (fr_ms * 1e-3), 1),
                     "pixels_hit": round(sum(hits) / len(hits), 3)}
        log(f"splat: {n0} splats in the list, {vis / 1e6:.1f} M in view on average, "
            f"{fr_ms:.2f} ms/frame -> {n0 / (fr_ms * 1e-3) / 1e9:.2f} G list splats/s, "
            f"{vis / (fr_ms * 1e-3) / 1e9:.2f} G visible splats/s")
        splat_cpu = (l0[::32].contiguous().cpu().numpy(), cams[0])   # every 32nd splat: ~4 M

    # ---- draw lists: octree node buffers, then per pose the LOD selection,
    # the assembled list and its splat resolve at 1080p ----
    draw_res = None
    if not args.no_splat:
        from voxplat_amd import frustum_planes, octree_cull

        for _ in range(2):
            nodes = [gpu.svl_nodes(w, l, splats, ids, offs, counts, capacity=None) for l in range(5)]
        torch.cuda.synchronize()
        caps = [int(n[1][-1].item()) for n in nodes]
        lev_ev = [torch.cuda.Event(enable_timing=True) for _ in range(6)]
        lev_ev[0].record(stream)
        for l in range(5):      # one launch per level, for the per-level times
            gpu.svl_nodes(w, l, splats, ids, offs, counts, capacity=caps[l], out=nodes[l])
            lev_ev[l + 1].record(stream)
        torch.cuda.synchronize()
        lev_ms = [round(lev_ev[l].elapsed_time(lev_ev[l + 1]), 4) for l in range(5)]
        gpu.svl_nodes_all(w, splats, ids, offs, counts, out=nodes)
        torch.cuda.synchronize()
        e0.record(stream)       # all five levels in one launch: the figure reported
        gpu.svl_nodes_all(w, splats, ids, offs, counts, out=nodes)
        e1.record(stream)
        torch.cuda.synchronize()
        assert gpu.status() == 0
        n_ms = e0.elapsed_time(e1)
        node_bytes = 16 * sum(caps)                          # every record read once, written once
        node_offs_h = [n[1].cpu().numpy().view(np.uint64) for n in nodes]
        lst = torch.empty((caps[0] + caps[1], 4), dtype=torch.int16, device=dev)
        lodb = torch.empty(caps[0] + caps[1], dtype=torch.uint8, device=dev)
        bnd_buf = torch.empty(((caps[0] + caps[1]) // abi.VP_SPLAT_BLOCK + 1, 8), dtype=torch.float32,
                              device=dev)
        depth_d = torch.empty((1080, 1920), dtype=torch.float32, device=dev)
        id_d = torch.empty((1080, 1920), dtype=torch.int32, device=dev)
        cull_s, gather_ms, frame_ms, drawn = [], [], [], []
        for p, yw, pt in poses:
            t0 = time.perf_counter()
            planes = frustum_planes(W.make_view_projection(p, yw, pt, 1920, 1080))
            draws, _, total = octree_cull(w, planes, p, node_offs_h)
            cull_s.append(time.perf_counter() - t0)
            if total == 0:
                continue
            cam = W.make_camera(p, yw, pt, 1920, 1080)
            torch.cuda.synchronize()
            e0.record(stream)
            gpu.draw_gather(w, [n[0] for n in nodes], [n[1] for n in nodes], draws, total,
                            out=(lst, lodb))
            e2 = torch.cuda.Event(enable_timing=True)
            e2.record(stream)
            # the assembled list changes every frame: its block bounds too
            bnd = gpu.splat_bounds(lst[:total], lod=lodb[:total], out=bnd_buf)
            gpu.splat(lst[:total], cam, lod=lodb[:total], out=(depth_d, id_d, None, None), bounds=bnd)
            e1.record(stream)
            torch.cuda.synchronize()
            gather_ms.append(e0.elapsed_time(e2))   # gather; the frame adds bounds + resolve
            frame_ms.append(e0.elapsed_time(e1))
            drawn.append(total)
        assert gpu.status() == 0
        draw_res = {"node_buffers_ms": round(n_ms, 4), "node_ms_per_level": lev_ms,
                    "node_records": sum(caps),
                    "node_gb_s": round(node_bytes / (n_ms * 1e-3) / 1e9, 1),
                    "poses": len(frame_ms), "resolution": "1920x1080",
                    "mean_splats_drawn": int(np.mean(drawn)) if drawn else 0,
                    "host_lod_selection_ms": round(1e3 * float(np.mean(cull_s)), 3),
                    "gather_ms": round(float(np.mean(gather_ms)), 4) if gather_ms else None,
                    "frame_ms": round(float(np.mean(frame_ms)), 4) if frame_ms else None}
        log(f"draw lists: nodes {n_ms:.3f} ms; per frame {draw_res['mean_splats_drawn']} splats, "
            f"{draw_res['frame_ms']} ms (gather {draw_res['gather_ms']} ms)")
        del nodes, lst, lodb, bnd_buf

    # ---- near-field quad mesh of the same world (every chunk, index order) ----
    mesh_res = None
    if not args.no_mesh:
        from voxplat_amd.world_frame import mesh_bytes

        mids = torch.arange(nch, dtype=torch.int32, device=dev)
        _, _, moffs = gpu.mesh(w, vox, mids, smap, uniform=uni, count_only=True)
        torch.cuda.synchronize()
        nf = int(moffs[-1].item())
        mgeom = torch.empty((nf, 4, 4), dtype=torch.int16, device=dev)
        mindex = torch.empty((nf, 6), dtype=torch.int32, device=dev)
        mout = (mgeom, mindex, moffs)
        gpu.mesh(w, vox, mids, smap, uniform=uni, capacity=nf, out=mout)
        torch.cuda.synchronize()
        reps = 5
        e0.record(stream)
        for _ in range(reps):
            gpu.mesh(w, vox, mids, smap, uniform=uni, capacity=nf, out=mout)
        e1.record(stream)
        torch.cuda.synchronize()
        assert gpu.status() == 0
        m_ms = e0.elapsed_time(e1) / reps
        codes = ((mgeom[:, 0, 3].to(torch.int32) >> 8) & 7).cpu().numpy()
        m_algo = mesh_bytes(uni_h, codes, w.R)
        m_gbs = m_algo / (m_ms * 1e-3) / 1e9
        mesh_res = {"faces": nf, "ms_per_launch": round(m_ms, 4),
                    "gvoxels_s": round(nch * VOXELS_PER_CHUNK / (m_ms * 1e-3) / 1e9, 2),
                    "algorithmic_bytes": m_algo, "achieved_gb_s": round(m_gbs, 1),
                    "frac_hbm": round(m_gbs / HBM_PEAK_GBS, 4)}
        log(f"mesh: {nf} faces, {m_ms:.3f} ms/launch -> {m_gbs:.0f} GB/s algorithmic")
        del mgeom, mindex, mout

    # ---- config 0: one 64^3 chunk, the plain per-chunk path (n = 1) ----
    c0w = W.CFG0.desc()
    c0vox = torch.empty(c0w.R ** 3, dtype=torch.uint8, device=dev)
    gpu.generate(c0w, W.CFG0.gen(), c0vox)
    c0sx, c0sz = c0w.shadow_dims()
    c0map = torch.empty(abi.shadow_padded_len(c0sx, c0sz), dtype=torch.int16, device=dev)
    gpu.shadow_build(c0w, c0vox, c0map)
    c0ids = torch.zeros(1, dtype=torch.int32, device=dev)
    _, c0offs, c0counts = gpu.cull_mesh(c0w, c0vox, c0ids, shadow_map=c0map, count_only=True)
    torch.cuda.synchronize()
    c0cap = int(c0offs[-1].item()) // 4
    c0out = torch.empty((max(c0cap, 1), 4), dtype=torch.int16, device=dev)
    c0reps = 50
    for _ in range(5):
        gpu.cull_mesh(c0w, c0vox, c0ids, shadow_map=c0map, capacity=c0cap, out=(c0offs, c0counts),
                      records=c0out)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(c0reps):
        gpu.cull_mesh(c0w, c0vox, c0ids, shadow_map=c0map, capacity=c0cap, out=(c0offs, c0counts),
                      records=c0out)
    e1.record(stream)
    torch.cuda.synchronize()
    assert gpu.status() == 0
    c0_us = e0.elapsed_time(e1) * 1e3 / c0reps
    cfg0_res = {"workload": "cull + L0..L4 of one 64^3 chunk, one launch", "records": c0cap,
                "gpu_us_per_chunk": round(c0_us, 2)}
    log(f"cfg0: one chunk, {c0cap} records, {c0_us:.1f} us per launch")

    # ---- CPU baseline: the oracle on host cores ----
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O

        vh = vox.cpu().numpy()
        mh = smap.cpu().numpy().view(np.uint16)
        idsh = np.arange(nch, dtype=np.uint32)
        O.remesh_time(w, vh, idsh[:64], mh, args.cpu_threads)      # warm
        t, _ = O.remesh_time(w, vh, idsh, mh, args.cpu_threads)
        cpu = {"value": round(nch * VOXELS_PER_CHUNK / t / 1e9, 4), "unit": "Gvoxels/s",
               "cores": args.cpu_threads, "kind": "port",
               "sample": f"all {nch} chunks of the same world, L0..L4 with shadows, "
                         f"{args.cpu_threads} OpenMP threads as chunkset_manage "
                         f"(host has {os.cpu_count()} CPUs); {t:.2f} s"}
        log(f"cpu: {t:.2f} s -> {cpu['value']} Gvoxel/s")
        ncores = min(16, os.cpu_count() or 1)
        t_all, _ = O.remesh_time(w, vh, idsh, mh, ncores)
        cpu["all_cores"] = {"value": round(nch * VOXELS_PER_CHUNK / t_all / 1e9, 4), "cores": ncores,
                            "seconds": round(t_all, 3)}
        log(f"cpu, {ncores} threads: {t_all:.2f} s")
        c0v, c0m = c0vox.cpu().numpy(), c0map.cpu().numpy().view(np.uint16)
        c0t = sorted(O.remesh_time(c0w, c0v, np.zeros(1, np.uint32), c0m, 1)[0] for _ in range(5))[2]
        cfg0_res["cpu_us_per_chunk_1_thread"] = round(c0t * 1e6, 1)
        # the other legs' oracle on bounded samples of the same inputs
        if mesh_res is not None:
            mixed_ids = np.nonzero(uni_h < 0)[0].astype(np.uint32)
            msample = mixed_ids[:: max(1, len(mixed_ids) // 256)][:256]
            t0 = time.perf_counter()
            O.mesh_list(w, vh, msample, mh, threads=args.cpu_threads)
            tm = time.perf_counter() - t0
            mesh_res["cpu"] = {"gvoxels_s": round(len(msample) * VOXELS_PER_CHUNK / tm / 1e9, 4),
                               "cores": args.cpu_threads, "kind": "port",
                               "sample": f"{len(msample)} mixed chunks, {tm:.2f} s"}
            log(f"cpu mesh: {len(msample)} chunks in {tm:.2f} s")
        if splat_cpu is not None:
            l0h, cam0 = splat_cpu
            t0 = time.perf_counter()
            O.splat_render(l0h, cam0)
            ts = time.perf_counter() - t0
            splat_res["cpu"] = {"splats_s": round(len(l0h) / ts, 1), "cores": 1, "kind": "port",
                                "sample": f"every 32nd splat of the list ({len(l0h)}), first pose, 1920x1080, "
                                          f"{ts:.2f} s"}
            log(f"cpu splat: {len(l0h)} splats in {ts:.2f} s")

    graph.close()
    gpu.close()
    return {
        "workload": "remesh L0-L4 of 2048x256x2048 (4096 chunks of 64^3), batched persistent "
                    "kernel in a hipGraph",
        "chunks": nch, "occupancy": occupancy, "splat_records": cap, "launch": tune,
        "gvoxels_s": round(gvox, 3), "ms_per_step": round(ms_per_step, 4),
        "per_chunk": {"gvoxels_s": round(pc_gvox, 3), "ms_per_pass": round(pc_ms, 3),
                      "identical_output": same},
        "speedup_vs_per_chunk": round(gvox / pc_gvox, 2),
        "roofline": roofline,
        "splat": splat_res,
        "mesh": mesh_res,
        "draw": draw_res,
        "cfg0": cfg0_res,
        "cpu": cpu,
    }


def leg_fit(args) -> dict:
    """The 1024x256x1024 colour fit at its stated size."""
    from voxplat_amd import GPU
    from voxplat_amd.splat_fit import run_fit

    gpu = GPU(0)
    res = run_fit(gpu, views=64, size=512, steps=200, lr=0.05, log_every=0)
    gpu.close()
    log(f"fit: {res['ms_per_step']} ms/step, error {res['mean_abs_rgb_error_visible_init']} -> "
        f"{res['mean_abs_rgb_error_visible']}")
    return res


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline legs")
    ap.add_argument("--no-splat", action="store_true")
    ap.add_argument("--no-mesh", action="store_true", help="skip the quad-mesh leg")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-2k", action="store_true", help="skip the 2048x256x2048 legs")
    ap.add_argument("--no-fit", action="store_true", help="skip the colour fit leg")
    ap.add_argument("--frames", type=int, default=3, help="edit-burst frames on the 16k world")
    ap.add_argument("--cpu-threads", type=int, default=4)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported "
              "(the pipeline runs on a single MI355X; a z-slab split of the chunk grid is how it "
              "would shard)",
              file=sys.stderr)
        return 2
    line = leg_world16k(args)
    if not args.no_2k:
        line["world2k"] = leg_world2k(args)
    if not args.no_fit:
        line["fit"] = leg_fit(args)
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
