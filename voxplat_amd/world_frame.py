"""A resident world on one GPU: generate, edit, and run the per-frame chunk
management pass -- batched remesh of the dirty chunks, then RLE
recompression of every other chunk -- as one hipGraph captured once and
replayed every frame.

This is chunkset_manage's frame (src/chunkset.c:246-507) over a whole world
held uncompressed in HBM: the reference walks the chunks, meshes the dirty
ones one per OpenMP thread (chunkset.c:251,339-487) and compresses the others
with rle_compress (chunkset.c:213-230,300-306).  Here the walk is a device
compaction of the dirty bytes into a dirty and an idle list, and each of the
two passes is one launch that reads its list length from device memory
(vp_host_frame_graph_*), so the same graph serves every edit burst.  The edit
burst is chunkset_edit_sphere (src/chunkset/edit.c:179-244): voxels
(vp_gpu_edit_spheres) and the shadow-map updates of the add spheres in their
sequential order (vp_gpu_shadow_spheres, shadow.h:77-89), both on the device.

    python -m voxplat_amd.world_frame [--frames K]     # 16384x512x16384, 10k edits per frame
"""
from __future__ import annotations

import json
import mmap
import sys
import time

import numpy as np
import torch

from . import abi
from . import workload as W
from ._lib import GPU, FrameGraph, RemeshGraph

MAP_NORESERVE = 0x4000          # Linux mmap flag: no swap reservation for the mapping


class ResidentWorld:
    """Chunk-major u8 voxels, per-chunk uniform values and the padded shadow
    map, all in device memory."""

    def __init__(self, gpu: GPU, cfg: W.WorldConfig, device: str = "cuda"):
        self.gpu, self.cfg = gpu, cfg
        self.w = cfg.desc()
        w = self.w
        self.R3 = w.R ** 3
        self.voxels = torch.empty(w.chunks * self.R3, dtype=torch.uint8, device=device)
        self.uniform = torch.empty(w.chunks, dtype=torch.int16, device=device)
        sx, sz = w.shadow_dims()
        self.smap = torch.empty(abi.shadow_padded_len(sx, sz), dtype=torch.int16, device=device)
        self.dirty = torch.zeros(w.chunks, dtype=torch.uint8, device=device)

    def generate(self) -> None:
        self.gpu.generate(self.w, self.cfg.gen(), self.voxels, self.uniform)
        self.gpu.shadow_build(self.w, self.voxels, self.smap)

    def edit(self, spheres) -> None:
        """Apply the spheres in order; marks every chunk of each sphere's box
        dirty.  The voxels and the shadow map are both updated on the device."""
        self.gpu.edit_spheres(self.w, self.voxels, spheres, self.dirty, self.uniform)
        self.gpu.shadow_spheres(self.w, self.smap, spheres)

    def open(self, ids, words, offsets, uniform) -> None:
        """Decode chunks `ids` from RLE streams back into the world.

        The reference decodes a compressed chunk when it is opened
        (chunk_open_ro / chunk_open_rw, src/chunkset.c:133-200) after
        chunk_compress dropped its voxels (chunkset.c:213-230,300-306).  Here
        the streams are the ones an encode of the idle list wrote (`words`,
        `offsets` in words, one entry per id) and `uniform` the per-entry
        values that encode returned; they become the chunks' uniform
        metadata again."""
        self.gpu.rle_decode(self.w, words, offsets, ids, self.voxels)
        self.uniform.index_copy_(0, ids.long(), uniform)

    def lists(self):
        """(dirty chunk ids, every other chunk id) as int32 device tensors, ascending."""
        d = self.dirty != 0
        ids = torch.arange(self.w.chunks, dtype=torch.int32, device=self.voxels.device)
        return ids[d].contiguous(), ids[~d].contiguous()

    def chunk(self, c: int) -> torch.Tensor:
        return self.voxels[c * self.R3:(c + 1) * self.R3]


def heaviest_first(gpu: GPU, w: abi.WorldDesc, voxels, ids, smap, uniform, lod: bool = True):
    """Reorder a chunk list by record count, heaviest first, and size it.

    The batched remesh hands chunks to workgroups in list order and writes
    its output in list order, so the order is the caller's schedule: with the
    heaviest chunks first, the chunks still running when the list runs dry
    are the light ones.  Returns (ids, offsets, counts) for the new order,
    offsets and counts from a counting pass over it."""
    _, offs, counts = gpu.cull_mesh(w, voxels, ids, shadow_map=smap, uniform=uniform, lod=lod,
                                    count_only=True)
    order = torch.argsort(counts.sum(1), descending=True, stable=True)
    ids = ids[order].contiguous()
    _, offs, counts = gpu.cull_mesh(w, voxels, ids, shadow_map=smap, uniform=uniform, lod=lod,
                                    count_only=True)
    return ids, offs, counts


def locality_order(gpu: GPU, w: abi.WorldDesc, voxels, ids, smap, uniform, lod: bool = True,
                   tail: int | None = None):
    """Order a chunk list for the batched remesh, and size it: the mixed chunks
    in descending chunk-index order, the lightest `tail` of them (default two
    per CU) moved to the end, heaviest of those first, then the uniform
    chunks.

    The batched remesh hands chunks to workgroups in list order, so the order
    is the caller's schedule.  In descending chunk-index order a mixed chunk's
    +x neighbour is the list's previous entry, read by another workgroup just
    before: the 4096 lines the +x column touches are then mostly found in the
    caches instead of memory.  All mixed chunks first keeps the memory system
    busy without gaps, and the light chunks at the end leave no long chunk
    running alone when the list runs dry.  Returns (ids, offsets, counts) for the new
    order, offsets and counts from a counting pass over it."""
    ids = torch.sort(ids, descending=True).values
    _, offs, counts = gpu.cull_mesh(w, voxels, ids, shadow_map=smap, uniform=uniform, lod=lod,
                                    count_only=True)
    mixed = uniform[ids.long()].to(torch.int32) < 0 if uniform is not None else \
        torch.ones_like(ids, dtype=torch.bool)
    mids, weight = ids[mixed], counts.sum(1)[mixed]
    k = min(2 * gpu.num_cus if tail is None else tail, int(mids.numel()))
    light = torch.argsort(weight, descending=True, stable=True)[mids.numel() - k:]
    keep = torch.ones(mids.numel(), dtype=torch.bool, device=ids.device)
    keep[light] = False
    ids = torch.cat([mids[keep], mids[light], ids[~mixed]]).contiguous()
    _, offs, counts = gpu.cull_mesh(w, voxels, ids, shadow_map=smap, uniform=uniform, lod=lod,
                                    count_only=True)
    return ids, offs, counts


class Frame:
    """The per-frame pass of a resident world, captured once (FrameGraph) and
    replayed each frame: every launch rebuilds the dirty and idle lists from
    the world's dirty bytes (clearing them), remeshes the dirty chunks and
    RLE-encodes the others into buffers sized once for `record_capacity`
    records and `word_capacity` words.  After `sync()` the attributes name
    the last frame's lists and outputs."""

    def __init__(self, world: ResidentWorld, record_capacity: int, word_capacity: int,
                 lod: bool = True):
        self.world = world
        self.graph = FrameGraph(world.gpu, world.w, world.voxels, world.uniform, world.smap,
                                world.dirty, record_capacity, word_capacity, lod=lod)
        self.splats, self.words = self.graph.splats, self.graph.words

    def run(self, stream=None) -> None:
        self.graph.launch(stream)

    def sync(self) -> None:
        """Read the last frame's list lengths and expose its outputs."""
        g = self.graph
        torch.cuda.synchronize()
        self.dirty, self.idle = g.lists()
        nd, ni = int(self.dirty.numel()), int(self.idle.numel())
        self.offsets, self.counts = g.offsets[:nd + 1], g.level_counts[:nd]
        self.word_offsets, self.idle_uniform = g.word_offsets[:ni + 1], g.idle_uniform[:ni]
        self.records = int(self.offsets[-1].item()) // 4 if nd else 0
        self.words_n = int(self.word_offsets[-1].item()) if ni else 0

    def chunk_records(self, k: int) -> torch.Tensor:
        """Records of dirty[k], levels 0..4 concatenated."""
        a, b = int(self.offsets[k]) // 4, int(self.offsets[k + 1]) // 4
        return self.splats[a:b]

    def chunk_words(self, k: int) -> torch.Tensor:
        a, b = int(self.word_offsets[k]), int(self.word_offsets[k + 1])
        return self.words[a:b]

    def close(self) -> None:
        if self.graph is not None:
            self.graph.close()
            self.graph = None


def sparse_host_world(w: abi.WorldDesc) -> np.ndarray:
    """A zero-filled chunk-major host buffer of the whole world whose pages
    exist only once written: CPU code indexed by world chunk ids can then run
    on a few chunks copied out of device memory."""
    size = w.chunks * w.R ** 3
    mm = mmap.mmap(-1, size, flags=mmap.MAP_PRIVATE | mmap.MAP_ANONYMOUS | MAP_NORESERVE)
    return np.frombuffer(mm, dtype=np.uint8)


def copy_chunks_to_host(world: ResidentWorld, host: np.ndarray, chunks) -> None:
    R3 = world.R3
    for c in sorted(set(int(c) for c in chunks)):
        host[c * R3:(c + 1) * R3] = world.chunk(c).cpu().numpy()


def plus_neighbours(w: abi.WorldDesc, c: int):
    """The chunk and its +x, +y, +z neighbours inside the world (the voxels a
    remesh of c reads)."""
    mx, my, mz = (1 << w.max_bitw[0]), (1 << w.max_bitw[1]), (1 << w.max_bitw[2])
    cx = c & (mx - 1)
    cy = (c >> w.max_bitw[0]) & (my - 1)
    cz = c >> (w.max_bitw[0] + w.max_bitw[1])
    out = [c]
    if cx + 1 < mx:
        out.append(c + 1)
    if cy + 1 < my:
        out.append(c + mx)
    if cz + 1 < mz:
        out.append(c + mx * my)
    return out


def remesh_bytes(uniform_h: np.ndarray, counts_h: np.ndarray) -> int:
    """Algorithmic bytes of one remesh launch: chunk + 3 halo planes (mixed)
    or the halo planes alone (uniform), 8 B per record written plus 4 B of
    shadow-map reads per record, 5 u32 counts per chunk."""
    mixed = int((uniform_h < 0).sum())
    uni = len(uniform_h) - mixed
    records = int(counts_h.astype(np.int64).sum()) // 4
    return mixed * (262144 + 3 * 4096) + uni * 3 * 4096 + 12 * records + 20 * len(uniform_h)


def mesh_bytes(uniform_h: np.ndarray, face_codes: np.ndarray, R: int = 64) -> int:
    """Algorithmic bytes of one quad-mesh launch: every mixed chunk's voxels
    and every chunk's one-voxel shell ((R+2)^3 - R^3 bytes) read once, 56 B
    written per face (4 vertices of 4 int16, 6 u32 indices), and the shadow
    map entries each face samples: 4 B for x faces towards +x (normal) and y
    faces of solid voxels, 8 B for z faces.  face_codes: per-face
    axis + 3 * normal (attribute bits 8-10)."""
    mixed = int((uniform_h < 0).sum())
    n = len(uniform_h)
    faces = len(face_codes)
    counts = np.bincount(face_codes, minlength=6)
    maps = 4 * (int(counts[3]) + int(counts[1])) + 8 * (int(counts[2]) + int(counts[5]))
    return mixed * R ** 3 + n * ((R + 2) ** 3 - R ** 3) + 56 * faces + maps + 8 * n


def rle_bytes(n_chunks: int, words: int, R3: int = 262144, read_chunks: int | None = None) -> int:
    """Algorithmic bytes of one RLE encode (or decode): the voxels of every
    chunk read (written) once -- of `read_chunks` of them when the others are
    uniform chunks the encoder does not read --, every word (terminators
    included) written (read) once, one u64 offset and i16 value per chunk."""
    nread = n_chunks if read_chunks is None else read_chunks
    return nread * R3 + n_chunks * 10 + 4 * words


def full_remesh(world: ResidentWorld, steps: int = 3, warmup: int = 1, log=None,
                autotune: bool = False):
    """Remesh every chunk of the world (L0-L4) as one graph launch per step,
    the list in locality order (`locality_order`); with `autotune`, the launch shape is
    first tuned on this list and kept as the world's shape (the frames' graph
    then uses it too).  Returns timings; the record buffer is freed before
    returning."""
    gpu, w = world.gpu, world.w
    ids = torch.arange(w.chunks, dtype=torch.int32, device=world.voxels.device)
    ids, offs, counts = locality_order(gpu, w, world.voxels, ids, world.smap, world.uniform)
    torch.cuda.synchronize()
    records = int(offs[-1].item()) // 4
    if log:
        log(f"full remesh: {records} records ({records * 8 / 2**30:.1f} GiB)")
    splats = torch.empty((max(records, 1), 4), dtype=torch.int16, device=world.voxels.device)
    gpu.reserve(w.chunks, 0)
    tune = None
    if autotune:
        tune = gpu.autotune_remesh(w, world.voxels, ids, world.smap, splats, records, offs,
                                   counts, uniform=world.uniform, reps=1)
        tune["ms"] = round(tune["ms"], 3)
        if log:
            log(f"launch shape: {tune}")
    g = RemeshGraph(gpu, w, world.voxels, ids, world.smap, offs, counts, splats, records,
                    uniform=world.uniform)
    st = torch.cuda.current_stream()
    for _ in range(warmup):
        g.launch()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    tw = time.perf_counter()
    e0.record(st)
    for _ in range(steps):
        g.launch()
    e1.record(st)
    e1.synchronize()
    wall = (time.perf_counter() - tw) * 1e3 / steps
    ms = e0.elapsed_time(e1) / steps
    status = gpu.status()
    g.close()
    uni_h = world.uniform.cpu().numpy()
    counts_h = counts.cpu().numpy().view(np.uint32)
    del splats
    torch.cuda.empty_cache()
    algo = remesh_bytes(uni_h, counts_h)
    return {"chunks": w.chunks, "records": records, "ms": round(max(ms, wall), 3),
            "kernel_ms": round(ms, 3),
            "gvoxels_s": round(w.chunks * world.R3 / (max(ms, wall) * 1e-3) / 1e9, 2),
            "algorithmic_bytes": algo, "gb_s": round(algo / (ms * 1e-3) / 1e9, 1),
            "launch": tune if tune else gpu.get_remesh_shape(w), "status": status}


def main(argv=None) -> int:
    import argparse

    ap = argparse.ArgumentParser(description="resident-world frames: edit bursts, remesh, RLE")
    ap.add_argument("--frames", type=int, default=4, help="edit bursts, one frame each")
    ap.add_argument("--edits", type=int, default=10000)
    ap.add_argument("--full-steps", type=int, default=0,
                    help="also time a remesh of every chunk (needs ~100 GB more HBM)")
    ap.add_argument("--warmup", type=int, default=1, help="untimed full remeshes first")
    ap.add_argument("--autotune", action="store_true",
                    help="tune the launch shape on the full list first (with --full-steps)")
    a = ap.parse_args(argv)
    res = run_world16k(a.frames, a.edits, full_steps=a.full_steps, full_warmup=a.warmup,
                       autotune=a.autotune, log=lambda m: print(m, file=sys.stderr, flush=True))
    print(json.dumps(res))
    return 0


def run_world16k(frames: int = 4, edits: int = 10000, log=None, keep: bool = False,
                 gpu: GPU | None = None, full_steps: int = 0, full_warmup: int = 1,
                 on_generated=None, record_capacity: int = 1 << 31, word_margin: float = 1.25,
                 autotune: bool = False):
    """The 16384x512x16384 workload.  Each frame is a new burst of `edits`
    seeded sphere edits (seed EDIT_SEED + frame), then one replay of the frame
    graph: lists, remesh of the dirty chunks, RLE of the others.  Returns a
    dict of timings (and, with keep, (world, frame, spheres of every burst,
    spheres of the last burst) under "_objects").  on_generated(world), if
    given, runs after generation and the full remesh, before the edits; its
    return value is kept under "on_generated"."""
    def say(m):
        if log:
            log(m)
    cfg = W.CFG3
    own = gpu is None
    gpu = GPU(0) if own else gpu
    t0 = time.perf_counter()
    world = ResidentWorld(gpu, cfg)
    world.generate()
    torch.cuda.synchronize()
    t_gen = time.perf_counter() - t0
    uni_h = world.uniform.cpu().numpy()
    occ = {"uniform_air": int((uni_h == 0).sum()), "uniform_solid": int((uni_h > 0).sum()),
           "mixed": int((uni_h < 0).sum())}
    say(f"generated {cfg.name} in {t_gen:.2f} s, occupancy {occ}")
    full = full_remesh(world, steps=full_steps, warmup=full_warmup, log=say,
                       autotune=autotune) if full_steps else None
    if full:
        say(f"full remesh: {full}")
    extra = on_generated(world) if on_generated else None
    # word budget: every chunk's stream now, with a margin for the edits
    ids = torch.arange(world.w.chunks, dtype=torch.int32, device=world.voxels.device)
    _, woffs, _ = gpu.rle_encode(world.w, world.voxels, ids, count_only=True)
    torch.cuda.synchronize()
    word_capacity = int(int(woffs[-1].item()) * word_margin)
    del ids, woffs
    fr = Frame(world, record_capacity, word_capacity)
    say(f"frame graph: {record_capacity} records, {word_capacity} words")
    st = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    per = []
    every, last = [], []
    for f in range(frames):
        spheres = W.edit_burst(world.w, n=edits, seed=W.EDIT_SEED + f)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        world.edit(spheres)
        torch.cuda.synchronize()
        t_edit = time.perf_counter() - t0
        tw = time.perf_counter()
        ev[0].record(st)
        fr.run()
        ev[1].record(st)
        ev[1].synchronize()
        wall = (time.perf_counter() - tw) * 1e3
        ms = ev[0].elapsed_time(ev[1])
        fr.sync()
        status = gpu.status()
        nd, ni = int(fr.dirty.numel()), int(fr.idle.numel())
        dirty_uni = world.uniform[fr.dirty.long()].cpu().numpy()
        rb = remesh_bytes(dirty_uni, fr.counts.cpu().numpy().view(np.uint32))
        # idle chunks the uniform table marks uniform are encoded without a read
        idle_read = int((world.uniform[fr.idle.long()] < 0).sum().item()) if ni else 0
        lb = rle_bytes(ni, fr.words_n, world.R3, read_chunks=idle_read)
        per.append({"edit_s": round(t_edit, 3), "dirty_chunks": nd, "idle_chunks": ni,
                    "idle_chunks_read": idle_read,
                    "records": fr.records, "rle_words": fr.words_n,
                    "ms": round(max(ms, wall), 3), "graph_ms": round(ms, 3),
                    "algorithmic_bytes": rb + lb,
                    "gb_s": round((rb + lb) / (ms * 1e-3) / 1e9, 1), "status": status})
        say(f"frame {f}: {per[-1]}")
        every.extend(spheres)
        last = spheres
    # reopen: every idle chunk of the last frame decoded back in place from
    # the stream just written (the bytes it writes are the bytes already there)
    ev[0].record(st)
    world.open(fr.idle, fr.words, fr.word_offsets, fr.idle_uniform)
    ev[1].record(st)
    ev[1].synchronize()
    ro = ev[0].elapsed_time(ev[1])
    status = gpu.status()
    lb = rle_bytes(int(fr.idle.numel()), fr.words_n, world.R3)
    mean = lambda k: round(float(np.mean([p[k] for p in per])), 3)
    out = {
        "workload": f"{cfg.name} ({world.w.chunks} chunks of 64^3, 128 GiB in HBM); {frames} "
                    f"frames, each: {edits} sphere edits r 1-32 (voxels and shadow map on the "
                    "device), then one replay of a graph captured once: dirty/idle lists on the "
                    "device, batched remesh L0-L4 of the dirty chunks, RLE encode of the others",
        "occupancy": occ, "generate_s": round(t_gen, 3),
        "frames": per, "ms_per_frame": mean("ms"), "edit_s": mean("edit_s"),
        "frame_gb_s": mean("gb_s"),
        "reopen_ms": round(ro, 3),
        "reopen_gb_s": round(lb / (ro * 1e-3) / 1e9, 1) if ro else None,
        "status": status,
        "full_remesh": full,
        "on_generated": extra,
    }
    if keep:
        out["_objects"] = (world, fr, every, last)
    else:
        fr.close()
        del world
        if own:
            gpu.close()
    return out


if __name__ == "__main__":
    sys.exit(main())
