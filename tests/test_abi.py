"""The C-ABI library loads without a GPU and exports every symbol the public
headers declare; argument validation fails with error codes, not aborts."""
import ctypes as C
import pathlib
import re

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent


def _declared():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        txt = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(vp_\w+)\s*\(", txt, re.M):
            if "static inline" in txt[max(0, m.start() - 40):m.start() + 60]:
                continue
            names.add(m.group(1))
    return names


def test_library_exports_declared_symbols():
    from voxplat_amd import load_library

    lib = load_library()
    names = _declared()
    assert {"vp_gpu_cull_mesh", "vp_gpu_lod_mip", "vp_gpu_rle", "vp_gpu_splat",
            "vp_gpu_splat_backward", "vp_gpu_make_mask", "vp_ctx_create"} <= names
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing


def test_bindings_cover_declarations():
    from voxplat_amd import abi

    assert _declared() == set(abi.PROTOTYPES)


def test_null_arguments_are_errors_not_crashes():
    from voxplat_amd import abi, load_library

    lib = load_library()
    assert lib.vp_gpu_cull_mesh(None, None, None, None, None, None, 0, None, 0, None, None) == abi.VP_ERR_ARG
    assert lib.vp_gpu_rle(None, 0, None, None, None, 1, None, 0, None, None, None) == abi.VP_ERR_ARG
    assert lib.vp_gpu_splat(None, None, None, 0, 0, None, None, None, None) == abi.VP_ERR_ARG
    assert lib.vp_gpu_lod_mip(None, 6, 1, None, None, 1, None) == abi.VP_ERR_ARG
    assert lib.vp_ctx_create(0, None) == abi.VP_ERR_ARG
    assert lib.vp_ctx_set_splat_tiles(None, 1, 0) == abi.VP_ERR_ARG
    assert lib.vp_strerror(abi.VP_ERR_CAPACITY) == b"output capacity exceeded"
    # host-side shadow updates run without a GPU
    w = abi.WorldDesc.make(2, (1, 1, 1))
    assert lib.vp_host_shadow_spheres(C.byref(w), None, None, 0) == abi.VP_ERR_ARG


def test_host_shadow_spheres_matches_oracle():
    import numpy as np

    from oracle import oracle as O
    from voxplat_amd import abi, host_shadow_spheres

    w = abi.WorldDesc.make(3, (2, 1, 2))
    sx, sz = w.shadow_dims()
    rng = np.random.default_rng(4)
    spheres = [abi.Sphere(int(rng.integers(-4, 36)), int(rng.integers(-2, 18)),
                          int(rng.integers(-4, 36)), int(rng.integers(1, 7)),
                          int(rng.integers(0, 64)) if i % 3 else 0) for i in range(40)]
    a = np.zeros(abi.shadow_padded_len(sx, sz), np.uint16)
    host_shadow_spheres(w, a, spheres)
    b = np.zeros_like(a)
    vox = np.zeros(w.chunks * 512, np.uint8)
    for s in spheres:
        O.edit_sphere(w, vox, b, (s.cx, s.cy, s.cz), s.radius, s.voxel)
    assert a.any() and np.array_equal(a, b)


class _FakeDev:
    """Stands in for a device tensor: only what the size check reads."""

    def __init__(self, nbytes, cuda=True, contiguous=True):
        self.nbytes, self.is_cuda, self._c = nbytes, cuda, contiguous

    def is_contiguous(self):
        return self._c

    def numel(self):
        return self.nbytes

    def element_size(self):
        return 1


def test_wrapper_size_checks_reject_short_or_host_buffers():
    from voxplat_amd import _lib, abi

    _lib._need(_FakeDev(64), 64, "x")
    _lib._need(None, 64, "x")
    with pytest.raises(ValueError, match="needs at least 65"):
        _lib._need(_FakeDev(64), 65, "x")
    with pytest.raises(ValueError, match="device tensor"):
        _lib._need(_FakeDev(64, cuda=False), 8, "x")
    with pytest.raises(ValueError, match="device tensor"):
        _lib._need(_FakeDev(64, contiguous=False), 8, "x")
    w = abi.WorldDesc.make(3, (2, 1, 2))
    sx, sz = w.shadow_dims()
    _lib._need_world(w, _FakeDev(w.chunks * 512), _FakeDev(w.chunks * 2),
                     _FakeDev(abi.shadow_padded_len(sx, sz) * 2))
    with pytest.raises(ValueError, match="voxels"):
        _lib._need_world(w, _FakeDev(w.chunks * 512 - 1))
    with pytest.raises(ValueError, match="shadow_map"):
        _lib._need_world(w, _FakeDev(w.chunks * 512), None, _FakeDev(sx * sz * 2))
