"""A resident world on one GPU: generate, edit, and run the per-frame chunk
management pass -- batched remesh of the dirty chunks in one hipGraph launch,
then RLE recompression of every other chunk.

This is chunkset_manage's frame (src/chunkset.c:246-507) over a whole world
held uncompressed in HBM: the reference meshes dirty chunks one per OpenMP
thread (chunkset.c:251,339-487) and compresses chunks that went idle with
rle_compress (chunkset.c:213-230); here each of the two is one launch over a
list of chunks.  The edit burst is chunkset_edit_sphere (src/chunkset/edit.c:
179-244): voxels on the GPU (vp_gpu_edit_spheres), the shadow-map updates of
the add spheres in their sequential order on the host (vp_host_shadow_spheres,
shadow.h:77-89).

    python -m voxplat_amd.world_frame [--frames K]     # 16384x512x16384, 10k edits
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
from ._lib import GPU, RemeshGraph, host_shadow_spheres

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
        """Apply the spheres in order; marks every chunk of each sphere's box dirty."""
        self.gpu.edit_spheres(self.w, self.voxels, spheres, self.dirty, self.uniform)
        adds = [s for s in spheres if s.voxel]
        if adds:
            m = self.smap.cpu().numpy().view(np.uint16)
            host_shadow_spheres(self.w, m, adds)
            self.smap.copy_(torch.from_numpy(m.view(np.int16)))

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


class Frame:
    """One frame of chunk management with every buffer sized up front: the
    remesh of `dirty` is captured in a hipGraph and the RLE of `idle` reuses
    its word buffer, so a frame allocates nothing."""

    def __init__(self, world: ResidentWorld, dirty, idle, lod: bool = True):
        gpu, w = world.gpu, world.w
        if dirty.numel():
            dirty, self.offsets, self.counts = heaviest_first(gpu, w, world.voxels, dirty,
                                                              world.smap, world.uniform, lod)
        else:
            _, self.offsets, self.counts = gpu.cull_mesh(w, world.voxels, dirty,
                                                         shadow_map=world.smap,
                                                         uniform=world.uniform, lod=lod,
                                                         count_only=True)
        self.world, self.dirty, self.idle = world, dirty, idle
        _, self.word_offsets, self.idle_uniform = gpu.rle_encode(w, world.voxels, idle,
                                                                 count_only=True)
        torch.cuda.synchronize()
        self.records = int(self.offsets[-1].item()) // 4
        self.words_n = int(self.word_offsets[-1].item()) if idle.numel() else 0
        dev = world.voxels.device
        self.splats = torch.empty((max(self.records, 1), 4), dtype=torch.int16, device=dev)
        self.words = torch.empty(max(self.words_n, 1), dtype=torch.int32, device=dev)
        gpu.reserve(max(int(dirty.numel()), int(idle.numel()), 1), 0)
        self.graph = (RemeshGraph(gpu, w, world.voxels, dirty, world.smap, self.offsets,
                                  self.counts, self.splats, self.records,
                                  uniform=world.uniform, lod=lod)
                      if dirty.numel() else None)

    def remesh(self) -> None:
        if self.graph is not None:
            self.graph.launch()

    def recompress(self) -> None:
        if self.idle.numel():
            self.world.gpu.rle_encode(self.world.w, self.world.voxels, self.idle,
                                      capacity=self.words_n,
                                      out=(self.words, self.word_offsets, self.idle_uniform))

    def run(self) -> None:
        self.remesh()
        self.recompress()

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


def rle_bytes(n_chunks: int, words: int, R3: int = 262144) -> int:
    """Algorithmic bytes of one RLE encode: every voxel read once, every word
    (terminators included) written once, one u64 offset and i16 value per chunk."""
    return n_chunks * (R3 + 10) + 4 * words


def full_remesh(world: ResidentWorld, steps: int = 3, warmup: int = 1, log=None):
    """Remesh every chunk of the world (L0-L4) as one graph launch per step,
    the list ordered heaviest first.  Returns timings; the record buffer is
    freed before returning."""
    gpu, w = world.gpu, world.w
    ids = torch.arange(w.chunks, dtype=torch.int32, device=world.voxels.device)
    ids, offs, counts = heaviest_first(gpu, w, world.voxels, ids, world.smap, world.uniform)
    torch.cuda.synchronize()
    records = int(offs[-1].item()) // 4
    if log:
        log(f"full remesh: {records} records ({records * 8 / 2**30:.1f} GiB)")
    splats = torch.empty((max(records, 1), 4), dtype=torch.int16, device=world.voxels.device)
    gpu.reserve(w.chunks, 0)
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
            "status": status}


def main(argv=None) -> int:
    import argparse

    ap = argparse.ArgumentParser(description="resident-world frame: edits, remesh, RLE")
    ap.add_argument("--frames", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--edits", type=int, default=10000)
    ap.add_argument("--full-steps", type=int, default=0,
                    help="also time a remesh of every chunk (needs ~100 GB more HBM)")
    a = ap.parse_args(argv)
    res = run_world16k(a.frames, a.warmup, a.edits, full_steps=a.full_steps,
                       log=lambda m: print(m, file=sys.stderr, flush=True))
    print(json.dumps(res))
    return 0


def run_world16k(frames: int = 5, warmup: int = 1, edits: int = 10000, log=None,
                 keep: bool = False, gpu: GPU | None = None, full_steps: int = 0,
                 on_generated=None):
    """The 16384x512x16384 workload.  Returns a dict of timings (and, with
    keep, the (world, frame, spheres) objects under "_objects").
    on_generated(world), if given, runs after generation and the full
    remesh, before the edits; its return value is kept under "on_generated"."""
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
    full = full_remesh(world, steps=full_steps, log=say) if full_steps else None
    if full:
        say(f"full remesh: {full}")
    extra = on_generated(world) if on_generated else None
    spheres = W.edit_burst(world.w, n=edits)
    t0 = time.perf_counter()
    world.edit(spheres)
    torch.cuda.synchronize()
    t_edit = time.perf_counter() - t0
    dirty, idle = world.lists()
    say(f"{edits} edits in {t_edit:.2f} s: {dirty.numel()} dirty, {idle.numel()} idle chunks")
    fr = Frame(world, dirty, idle)
    say(f"frame buffers: {fr.records} records, {fr.words_n} RLE words")
    st = torch.cuda.current_stream()
    for _ in range(warmup):
        fr.run()
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    rm = rl = 0.0
    tw = time.perf_counter()
    for _ in range(frames):
        ev[0].record(st)
        fr.remesh()
        ev[1].record(st)
        fr.recompress()
        ev[2].record(st)
        ev[2].synchronize()
        rm += ev[0].elapsed_time(ev[1])
        rl += ev[1].elapsed_time(ev[2])
    wall = (time.perf_counter() - tw) * 1e3 / frames
    rm, rl = rm / frames, rl / frames
    # reopen: every idle chunk decoded back in place from the stream just
    # written (the bytes it writes are the bytes already there)
    ev[0].record(st)
    world.open(idle, fr.words, fr.word_offsets, fr.idle_uniform)
    ev[1].record(st)
    ev[1].synchronize()
    ro = ev[0].elapsed_time(ev[1])
    status = gpu.status()
    nd, ni = int(dirty.numel()), int(idle.numel())
    dirty_uni = world.uniform[dirty.long()].cpu().numpy()
    counts_h = fr.counts.cpu().numpy().view(np.uint32)
    rb = remesh_bytes(dirty_uni, counts_h)
    lb = rle_bytes(ni, fr.words_n, world.R3)
    out = {
        "workload": f"{cfg.name} ({world.w.chunks} chunks of 64^3, 128 GiB in HBM), "
                    f"{edits} sphere edits r 1-32, then per frame: batched remesh L0-L4 of the "
                    "dirty chunks (one hipGraph launch) + RLE encode of the other chunks",
        "occupancy": occ, "generate_s": round(t_gen, 3), "edit_s": round(t_edit, 3),
        "dirty_chunks": nd, "idle_chunks": ni, "records": fr.records, "rle_words": fr.words_n,
        "frames": frames, "ms_per_frame": round(max(wall, rm + rl), 3),
        "remesh_ms": round(rm, 3),
        "remesh_gvoxels_s": round(nd * world.R3 / (rm * 1e-3) / 1e9, 2) if rm else None,
        "remesh_gb_s": round(rb / (rm * 1e-3) / 1e9, 1) if rm else None,
        "rle_ms": round(rl, 3),
        "rle_gvoxels_s": round(ni * world.R3 / (rl * 1e-3) / 1e9, 2) if rl else None,
        "rle_gb_s": round(lb / (rl * 1e-3) / 1e9, 1) if rl else None,
        "frame_gb_s": round((rb + lb) / ((rm + rl) * 1e-3) / 1e9, 1),
        "reopen_ms": round(ro, 3),
        "reopen_gb_s": round(lb / (ro * 1e-3) / 1e9, 1) if ro else None,
        "status": status,
        "full_remesh": full,
        "on_generated": extra,
    }
    if keep:
        out["_objects"] = (world, fr, spheres)
    else:
        fr.close()
        del world
        if own:
            gpu.close()
    return out


if __name__ == "__main__":
    sys.exit(main())
