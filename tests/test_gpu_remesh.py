"""GPU remesh (vp_gpu_cull_mesh, vp_gpu_make_mask, vp_gpu_lod_mip) against
the oracle, bit-exact: splat lists, per-level counts, offsets."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import refimpl as RI
from tests.gpuutil import dev, gpu_remesh, host, ids_dev
from voxplat_amd import abi
from voxplat_amd import workload as W

pytestmark = pytest.mark.gpu


def _check(w, vox, ids, smap, got, lod=True):
    recs, offs, counts, st = got
    erecs, eoffs, ecounts = O.remesh_list(w, vox, ids, smap, lod=lod, threads=8)
    assert st == 0
    assert np.array_equal(offs, eoffs)
    assert np.array_equal(counts, ecounts)
    assert np.array_equal(recs, erecs)


def test_cfg0_single_chunk_plain_path(gpu):
    cfg = W.CFG0
    w = cfg.desc()
    vox = O.gen_world(w, cfg.gen())
    smap = O.shadow_build(w, vox)
    got = gpu_remesh(gpu, w, vox, [0], smap)
    _check(w, vox, [0], smap, got)
    assert len(got[0]) > 40000


@pytest.mark.parametrize("rb,mb,density,seed", [
    (2, (2, 1, 2), 0.5, 1), (3, (2, 2, 1), 0.3, 2), (4, (1, 2, 2), 0.6, 3),
    (5, (1, 1, 2), 0.5, 4), (6, (1, 1, 1), 0.5, 5), (6, (2, 1, 1), 0.03, 6),
    (6, (1, 1, 1), 0.97, 7),
])
def test_random_worlds_batched(gpu, rb, mb, density, seed):
    w = abi.WorldDesc.make(rb, mb)
    rng = np.random.default_rng(seed)
    vox = RI.random_world(w, rng, density)
    smap = O.shadow_build(w, vox)
    ids = rng.permutation(w.chunks).astype(np.uint32)
    ids = np.concatenate([ids, ids[:3]])               # repeats are legal
    _check(w, vox, ids, smap, gpu_remesh(gpu, w, vox, ids, smap))
    # the same chunks one launch each (the plain per-chunk path)
    for c in ids[:4]:
        _check(w, vox, [c], smap, gpu_remesh(gpu, w, vox, [c], smap))


@pytest.mark.parametrize("block,wg,u", [(256, 2, 4), (1024, 1, 4), (512, 2, 2), (256, 1, 2),
                                        (1024, 1, 2), (512, 2, 4)])
def test_launch_shapes(gpu, block, wg, u):
    """Every launch shape the autotuner may pick gives the same bytes: set as
    the world's own shape, then as the default (both restored after)."""
    cfg = W.small_world(6, (2, 1, 2), 0x5EED0077, hmax=100)
    w = cfg.desc()
    vox = O.gen_world(w, cfg.gen())
    smap = O.shadow_build(w, vox)
    ids = np.arange(w.chunks, dtype=np.uint32)
    default = gpu.get_remesh_shape()
    w_other = abi.WorldDesc.make(6, (1, 1, 2))
    gpu.set_remesh_shape(w, block, wg, u)
    try:
        assert gpu.get_remesh_shape(w) == {"block": block, "wg_per_cu": wg,
                                           "records_in_flight": u}
        assert gpu.get_remesh_shape(w_other) == default     # other worlds keep the default
        _check(w, vox, ids, smap, gpu_remesh(gpu, w, vox, ids, smap))
    finally:
        gpu.set_remesh_shape(w, 512, 2, 4)
    gpu.set_remesh_shape(None, block, wg, u)
    try:
        vox2 = O.gen_world(w_other, W.small_world(6, (1, 1, 2), 0x5EED0078, hmax=50).gen())
        smap2 = O.shadow_build(w_other, vox2)
        ids2 = np.arange(w_other.chunks, dtype=np.uint32)
        _check(w_other, vox2, ids2, smap2, gpu_remesh(gpu, w_other, vox2, ids2, smap2))
    finally:
        gpu.set_remesh_shape(None, 512, 2, 4)


def test_locality_order(gpu):
    """locality_order: a permutation of the list with the mixed chunks first
    (descending index order, then the `tail` lightest, heaviest of them
    first), the uniform chunks last; the remesh of the reordered list is the
    oracle's for that list."""
    import torch

    from voxplat_amd.world_frame import locality_order

    cfg = W.small_world(6, (2, 2, 2), 0x5EED007A, hmax=70)
    w = cfg.desc()
    vox_d = torch.empty(w.chunks * w.R ** 3, dtype=torch.uint8, device="cuda")
    uni_d = torch.empty(w.chunks, dtype=torch.int16, device="cuda")
    gpu.generate(w, cfg.gen(), vox_d, uni_d)
    sx, sz = w.shadow_dims()
    smap_d = torch.empty(abi.shadow_padded_len(sx, sz), dtype=torch.int16, device="cuda")
    gpu.shadow_build(w, vox_d, smap_d)
    ids0 = torch.arange(w.chunks - 1, -1, -1, dtype=torch.int32, device="cuda")
    ids, offs, counts = locality_order(gpu, w, vox_d, ids0, smap_d, uni_d, tail=2)
    torch.cuda.synchronize()
    ids_h, uni_h = ids.cpu().numpy(), uni_d.cpu().numpy()
    assert sorted(ids_h.tolist()) == list(range(w.chunks))
    mixed = uni_h[ids_h] < 0
    nm = int(mixed.sum())
    assert 3 < nm < w.chunks, "the test world needs mixed and uniform chunks"
    assert mixed[:nm].all() and not mixed[nm:].any()
    assert np.all(np.diff(ids_h[:nm - 2]) < 0) and np.all(np.diff(ids_h[nm:]) < 0)
    tot = counts.cpu().numpy().view(np.uint32).sum(1)
    assert tot[nm - 2] >= tot[nm - 1] and min(tot[nm - 2:nm]) <= min(tot[:nm - 2])
    vox, smap = vox_d.cpu().numpy(), smap_d.cpu().numpy().view(np.uint16)
    _check(w, vox, ids_h.astype(np.uint32), smap, gpu_remesh(gpu, w, vox, ids_h.astype(np.uint32), smap))


def test_residency_and_autotuner(gpu):
    """Resident workgroups per CU follow the kernels' LDS (about 70 KiB at
    512 threads, 156 KiB at 1024); the autotuner times only resident shapes,
    keeps the fastest as the world's own shape, and leaves the default and
    other worlds alone; its pick remeshes to the oracle's bytes."""
    import torch

    cfg = W.small_world(6, (2, 1, 2), 0x5EED0079, hmax=100)
    w = cfg.desc()
    assert gpu.remesh_resident(w, 512) == 2 and gpu.remesh_resident(w, 512, 2) == 2
    assert gpu.remesh_resident(w, 1024) == 1
    res = {(b, u): gpu.remesh_resident(w, b, u) for b in (256, 512, 1024) for u in (2, 4)}
    assert all(1 <= r for r in res.values())
    vox = O.gen_world(w, cfg.gen())
    smap = O.shadow_build(w, vox)
    ids = np.arange(w.chunks, dtype=np.uint32)
    default = gpu.get_remesh_shape()
    vd, sd, idd = dev(vox), dev(smap), dev(ids)
    _, offs, counts = gpu.cull_mesh(w, vd, idd, shadow_map=sd, count_only=True)
    cap = int(offs[-1].item()) // 4
    splats = torch.empty((cap, 4), dtype=torch.int16, device="cuda")
    t = gpu.autotune_remesh(w, vd, idd, sd, splats, cap, offs, counts, reps=1)
    assert t["shapes_timed"] == sum(min(r, 4) for r in res.values())
    assert t["wg_per_cu"] <= res[(t["block"], t["records_in_flight"])]
    best = {k: t[k] for k in ("block", "wg_per_cu", "records_in_flight")}
    assert gpu.get_remesh_shape(w) == best
    assert gpu.get_remesh_shape() == default
    assert gpu.get_remesh_shape(abi.WorldDesc.make(6, (1, 1, 1))) == default
    try:
        _check(w, vox, ids, smap, gpu_remesh(gpu, w, vox, ids, smap))
    finally:
        gpu.set_remesh_shape(w, 512, 2, 4)


@pytest.mark.parametrize("rb", [3, 6])
def test_generated_world_with_uniform_chunks(gpu, rb):
    # heights well inside the world: all-air chunks above, all-solid below
    # (caves make most "solid" chunks mixed; clear them in the bottom layer)
    w = abi.WorldDesc.make(rb, (2, 2, 2))
    R = 1 << rb
    cfg = W.small_world(rb, (2, 2, 2), 77, hmax=3 * R - 1)
    vox = O.gen_world(w, cfg.gen())
    world = RI.world_array(w, vox)
    world[:, :R, :] = 9                                 # cy = 0 layer fully solid
    world[:, 3 * R:, :] = 0                             # cy = 3 layer all air
    world[:, R:R + 2, :] = 9                            # the surface layer stays mixed
    vox = np.ascontiguousarray(world.reshape(4, R, 4, R, 4, R).transpose(0, 2, 4, 1, 3, 5)).reshape(-1)
    smap = O.shadow_build(w, vox)
    uni = np.array([(-1 if (c != c[0]).any() else int(c[0])) for c in vox.reshape(w.chunks, -1)],
                   np.int16)
    assert (uni == 0).any() and (uni == 9).any() and (uni == -1).any()
    ids = np.arange(w.chunks, dtype=np.uint32)
    plain = gpu_remesh(gpu, w, vox, ids, smap)
    with_meta = gpu_remesh(gpu, w, vox, ids, smap, uniform=uni)
    _check(w, vox, ids, smap, plain)
    _check(w, vox, ids, smap, with_meta)


def test_lod_off_and_v2_records(gpu):
    w = abi.WorldDesc.make(4, (2, 1, 2))
    rng = np.random.default_rng(12)
    vox = RI.random_world(w, rng, 0.4, colours=(1, 64))
    smap = O.shadow_build(w, vox)
    ids = np.arange(w.chunks, dtype=np.uint32)
    _check(w, vox, ids, smap, gpu_remesh(gpu, w, vox, ids, smap, lod=False), lod=False)
    comp = gpu_remesh(gpu, w, vox, ids, smap)
    v2 = gpu_remesh(gpu, w, vox, ids, smap, v2=True)
    assert np.array_equal(comp[1], v2[1]) and np.array_equal(comp[2], v2[2])
    raw = v2[0].view(np.uint8).reshape(-1, 8)
    col, flags = raw[:, 6].astype(np.int16), raw[:, 7].astype(np.int16)
    assert np.array_equal(comp[0][:, :3], v2[0][:, :3])
    assert np.array_equal(comp[0][:, 3], col | (flags & 0x40))
    # level-0 face flags against the world: bit 2a = +a open, 2a+1 = -a open
    world = RI.world_array(w, vox)
    P = np.pad(world, 1)
    R = w.R
    start = 0
    for k, c in enumerate(ids):
        n0 = v2[2][k][0] // 4
        recs = v2[0][start:start + n0].astype(np.int64)
        fl = flags[start:start + n0]
        cx, cy, cz = RI.chunk_coords(w, int(c))
        for r, f in zip(recs, fl):
            x, y, z = r[:3]
            lx, ly, lz = x - cx * R, y - cy * R, z - cz * R
            if max(lx, ly, lz) == R:
                exp = 2 if lx == R else 8 if ly == R else 32
            else:
                s = lambda dx, dy, dz: P[z + 1 + dz, y + 1 + dy, x + 1 + dx] != 0
                exp = ((not s(1, 0, 0)) | ((lx > 0 and not s(-1, 0, 0)) << 1) |
                       ((not s(0, 1, 0)) << 2) | ((ly > 0 and not s(0, -1, 0)) << 3) |
                       ((not s(0, 0, 1)) << 4) | ((lz > 0 and not s(0, 0, -1)) << 5))
            assert (f & 0x3F) == exp
        start += sum(v2[2][k]) // 4


def test_capacity_overflow_count_only_and_bad_id(gpu):
    import torch

    w = abi.WorldDesc.make(4, (2, 1, 1))
    vox = RI.random_world(w, np.random.default_rng(3), 0.5)
    smap = O.shadow_build(w, vox)
    ids = np.arange(w.chunks, dtype=np.uint32)
    erecs, eoffs, ecounts = O.remesh_list(w, vox, ids, smap)
    vd, sd, idd = dev(vox), dev(smap), ids_dev(ids)
    # count only
    _, offs, counts = gpu.cull_mesh(w, vd, idd, shadow_map=sd, count_only=True)
    torch.cuda.synchronize()
    assert np.array_equal(host(offs, np.uint64), eoffs)
    assert np.array_equal(host(counts, np.uint32), ecounts)
    # capacity for the first chunk only: it is written, the rest are not
    cap = int(eoffs[1]) // 4
    recs, offs, counts = gpu.cull_mesh(w, vd, idd, shadow_map=sd, capacity=cap)
    torch.cuda.synchronize()
    assert gpu.status() & abi.VP_STATUS_OVERFLOW
    assert np.array_equal(host(offs, np.uint64), eoffs)
    assert np.array_equal(host(recs)[:cap], erecs[:cap])
    # a chunk id beyond the world is flagged and contributes nothing
    bad = np.array([0, w.chunks + 5, 1], np.uint32)
    recs, offs, counts = gpu.cull_mesh(w, vd, ids_dev(bad), shadow_map=sd)
    torch.cuda.synchronize()
    assert gpu.status() & abi.VP_STATUS_BAD_ID
    o = host(offs, np.uint64)
    assert o[2] == o[1] and o[1] == eoffs[1]
    # device-resident count: only the first 3 of the list
    d_n = dev(np.array([3], np.uint32))
    recs, offs, counts = gpu.cull_mesh(w, vd, idd, shadow_map=sd, d_n=d_n,
                                       capacity=int(eoffs[3]) // 4)
    torch.cuda.synchronize()
    assert gpu.status() == 0
    assert np.array_equal(host(offs, np.uint64)[:4], eoffs[:4])
    assert np.array_equal(host(recs)[:int(eoffs[3]) // 4], erecs[:int(eoffs[3]) // 4])


@pytest.mark.parametrize("rb", [3, 6])
def test_make_mask_and_lod_mip(gpu, rb):
    import torch

    w = abi.WorldDesc.make(rb, (1, 1, 2))
    vox = RI.random_world(w, np.random.default_rng(rb), 0.5, colours=(1, 256))
    ids = np.arange(w.chunks, dtype=np.uint32)
    masks = gpu.make_mask(w, dev(vox), ids_dev(ids))
    torch.cuda.synchronize()
    mh = host(masks).reshape(w.chunks, -1)
    for c in range(w.chunks):
        assert np.array_equal(mh[c], O.make_mask(w, vox, c)), c
    cur_g, cur_o = masks, mh
    for level in range(1, 5):
        cur_g = gpu.lod_mip(rb, level, cur_g)
        torch.cuda.synchronize()
        cur_o = np.stack([O.mask_downsample(w.R, level, m) for m in cur_o])
        assert np.array_equal(host(cur_g).reshape(w.chunks, -1), cur_o), level


def test_cfg1_full_world_batched(gpu):
    """The 2048x256x2048 world, all 4096 chunks in one persistent launch."""
    import torch

    cfg = W.CFG1
    w = cfg.desc()
    vox = torch.empty(w.chunks * w.R ** 3, dtype=torch.uint8, device="cuda")
    uni = torch.empty(w.chunks, dtype=torch.int16, device="cuda")
    gpu.generate(w, cfg.gen(), vox, uni)
    sx, sz = w.shadow_dims()
    smap = torch.empty(abi.shadow_padded_len(sx, sz), dtype=torch.int16, device="cuda")
    gpu.shadow_build(w, vox, smap)
    torch.cuda.synchronize()
    vox_h, smap_h = host(vox), host(smap, np.uint16)
    ids = np.arange(w.chunks, dtype=np.uint32)
    got = gpu_remesh(gpu, w, vox, ids, smap, uniform=uni)
    _check(w, vox_h, ids, smap_h, got)
    # the list in descending order: every chunk's records and counts equal
    # those of the list above, at the chunk's place in the new order
    rd, od, cd, st = gpu_remesh(gpu, w, vox, ids[::-1].copy(), smap, uniform=uni)
    ra, oa, ca, _ = got
    assert st == 0 and od[-1] == oa[-1]
    for e, c in enumerate(ids[::-1]):
        assert np.array_equal(cd[e], ca[c])
        assert np.array_equal(rd[od[e] // 4:od[e + 1] // 4], ra[oa[c] // 4:oa[c + 1] // 4]), c


def test_output_past_4gib(gpu):
    """A list whose output passes 2^31 int16 items (4 GiB of records): the
    2048x256x2048 world listed three times over; every copy of a chunk's
    records must be identical and sit at its offset."""
    import torch

    cfg = W.CFG1
    w = cfg.desc()
    vox = torch.empty(w.chunks * w.R ** 3, dtype=torch.uint8, device="cuda")
    uni = torch.empty(w.chunks, dtype=torch.int16, device="cuda")
    gpu.generate(w, cfg.gen(), vox, uni)
    sx, sz = w.shadow_dims()
    smap = torch.empty(abi.shadow_padded_len(sx, sz), dtype=torch.int16, device="cuda")
    gpu.shadow_build(w, vox, smap)
    one = torch.arange(w.chunks, dtype=torch.int32, device="cuda")
    ids = torch.cat([one, one, one])
    recs, offs, counts = gpu.cull_mesh(w, vox, ids, shadow_map=smap, uniform=uni)
    torch.cuda.synchronize()
    assert gpu.status() == 0
    n1 = int(offs[w.chunks]) // 4
    assert int(offs[-1]) == 3 * 4 * n1 and 4 * 3 * n1 > 2 ** 31
    items = counts.sum(1).to(torch.int64)
    assert torch.equal(offs[1:] - offs[:-1], items)
    first = recs[:n1]
    assert bool((first != 0).any(dim=1).all())           # every record has a non-zero field
    assert torch.equal(recs[n1:2 * n1], first)
    assert torch.equal(recs[2 * n1:3 * n1], first)
    del recs
    torch.cuda.empty_cache()


def test_graph_keeps_its_scratch_when_the_context_grows(gpu):
    """A captured remesh graph owns its ticket and look-back words: calls on
    the same context that grow (and so replace) the context's scratch do not
    touch what the graph's replays use.  The graph is replayed after an RLE
    encode and a remesh whose lists need more look-back words than the
    context started with, and after both run beside it, on the same world."""
    import torch

    from voxplat_amd import GPU, RemeshGraph

    cfg = W.small_world(6, (2, 1, 2), 0x5EED0101, hmax=100)
    w = cfg.desc()
    vox = O.gen_world(w, cfg.gen())
    smap = O.shadow_build(w, vox)
    ids = np.random.default_rng(5).permutation(w.chunks).astype(np.uint32)
    erecs, eoffs, ecounts = O.remesh_list(w, vox, ids, smap, threads=8)
    g2 = GPU(0)                  # a fresh context: 4096 look-back words to start with
    try:
        vd, sd, idd = dev(vox), dev(smap), ids_dev(ids)
        dn = dev(np.array([len(ids)], np.uint32))
        total = len(erecs)
        recs = torch.empty(total * 4, dtype=torch.int16, device="cuda")
        offs = torch.empty(len(ids) + 1, dtype=torch.int64, device="cuda")
        counts = torch.empty(len(ids) * 5, dtype=torch.int32, device="cuda")
        graph = RemeshGraph(g2, w, vd, idd, sd, offs, counts, recs, total, d_n=dn)

        def replay_and_check():
            recs.fill_(0)
            graph.launch()
            torch.cuda.synchronize()
            assert g2.status() == 0
            assert np.array_equal(host(offs, np.uint64), eoffs)
            assert np.array_equal(host(counts, np.uint32).reshape(-1, 5), ecounts)
            assert np.array_equal(host(recs).reshape(-1, 4), erecs)

        replay_and_check()
        big = ids_dev(np.tile(ids, 20000 // len(ids) + 1)[:20000])   # 20,000 entries
        g2.rle_encode(w, vd, big)
        torch.cuda.synchronize()
        replay_and_check()
        got = gpu_remesh(g2, w, vd, np.tile(ids, 600)[:9000], sd)
        assert got[3] == 0 and len(got[0]) > total
        replay_and_check()
        graph.close()
    finally:
        g2.close()
