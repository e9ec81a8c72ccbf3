"""Helpers shared by the GPU parity tests."""
from __future__ import annotations

import numpy as np


def dev(a: np.ndarray):
    import torch

    a = np.ascontiguousarray(a)
    if a.dtype == np.uint16:
        a = a.view(np.int16)
    elif a.dtype == np.uint32:
        a = a.view(np.int32)
    elif a.dtype == np.uint64:
        a = a.view(np.int64)
    return torch.from_numpy(a).cuda()


def ids_dev(ids):
    return dev(np.asarray(ids, np.uint32))


def host(t, dtype=None) -> np.ndarray:
    a = t.detach().cpu().numpy()
    return a.view(dtype) if dtype is not None else a


def gpu_remesh(gpu, w, vox, ids, smap, uniform=None, lod=True, v2=False):
    """Returns (records int16[N,4] (v2: raw bytes view), offsets u64, counts u32, status)."""
    import torch

    vd = vox if not isinstance(vox, np.ndarray) else dev(vox)
    sd = smap if not isinstance(smap, np.ndarray) else dev(smap)
    ud = None if uniform is None else (uniform if not isinstance(uniform, np.ndarray) else dev(uniform))
    recs, offs, counts = gpu.cull_mesh(w, vd, ids_dev(ids), shadow_map=sd, uniform=ud, lod=lod,
                                       v2=v2)
    torch.cuda.synchronize()
    st = gpu.status()
    offs_h = host(offs, np.uint64)
    total = int(offs_h[-1]) // 4
    return host(recs)[:total], offs_h, host(counts, np.uint32), st
