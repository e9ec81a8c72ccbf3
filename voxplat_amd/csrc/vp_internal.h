// vp_internal.h -- shared definitions of the HIP library (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <mutex>

#include "../../include/vp_gpu.h"

#define VP_HIP_TRY(expr)                                   \
    do {                                                   \
        hipError_t e_ = (expr);                            \
        if (e_ != hipSuccess) return VP_ERR_HIP;           \
    } while (0)

// Launch shape of the persistent remesh kernel: threads per workgroup,
// workgroups per CU the grid is sized for, records each lane prepares
// before its stores.
struct vp_remesh_shape_i {
    int block = 512;
    int wg_per_cu = 2;
    int u = 4;
};

struct vp_ctx {
    int device = 0;
    int num_cus = 0;
    int xcds = 8;
    std::mutex mu;
    // device scratch: [0] status word, [1] ticket counter, then look-back
    // state words (u64).  Grown on demand, never shrunk.
    uint32_t *d_status = nullptr;      // 1 word
    uint32_t *d_ticket = nullptr;      // 4 words
    uint64_t *d_state = nullptr;       // state_cap words
    uint64_t  state_cap = 0;
    uint64_t *d_frame = nullptr;       // packed depth|id per pixel for splat
    uint64_t  frame_cap = 0;
    float4   *d_rays = nullptr;        // per-pixel 1/ray direction for splat
    uint32_t *d_blk = nullptr;         // splat block culling: bins, order, counters
    uint64_t  blk_cap = 0;             // blocks d_blk is sized for
    uint64_t  rays_cap = 0;
    uint32_t *d_tiles = nullptr;       // splat screen tiles: histograms, counts, items, entries
    uint64_t  tiles_words = 0;         // u32 words of d_tiles
    int       splat_tiles = 0;         // resolve block-culled frames on screen tiles
    uint64_t  splat_tile_entries = 0;  // entry capacity of a tiled frame (0: default)
    uint32_t *d_cmap = nullptr;        // draw lists: list entry + 1 per chunk index
    uint64_t  cmap_cap = 0;
    uint32_t *d_oct = nullptr;         // device draw-queue selection: key slots
    uint64_t  oct_cap = 0;
    uint64_t *d_oct_part = nullptr;    // and per-workgroup partial counts
    uint64_t  oct_part_cap = 0;
    // device address of the static scratch of vp_gpu_cull_mesh_noalloc on
    // this context's device (noalloc.hip); not owned
    uint64_t *d_noalloc = nullptr;
    // launch shape of the persistent remesh kernel: a default, and tuned
    // shapes per world geometry (key: root_bitw | max_bitw[a] << 8 (a + 1))
    vp_remesh_shape_i cull_shape;
    static constexpr int kMaxShapes = 16;
    uint32_t cull_shape_key[kMaxShapes] = {};
    vp_remesh_shape_i cull_shapes[kMaxShapes];
    int n_cull_shapes = 0;
};

namespace vp {

int ensure_state(vp_ctx *ctx, uint64_t words);
int ensure_frame(vp_ctx *ctx, uint64_t pixels);
int ensure_blocks(vp_ctx *ctx, uint64_t blocks);
int ensure_cmap(vp_ctx *ctx, uint64_t chunks);
int ensure_tiles(vp_ctx *ctx, uint64_t words);
int ensure_oct(vp_ctx *ctx, uint64_t slots, uint64_t blocks);
// fetch the static scratch's device address on the current device
int noalloc_init(vp_ctx *ctx);

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

inline uint32_t world_key(const vp_world_desc *w)
{
    return w->root_bitw | (uint32_t)w->max_bitw[0] << 8 | (uint32_t)w->max_bitw[1] << 16 |
           (uint32_t)w->max_bitw[2] << 24;
}

// The remesh launch shape for a world: its tuned shape, else the default.
// Read without the context lock (launches may hold it); shapes are set
// between calls, not during them.
inline vp_remesh_shape_i remesh_shape(const vp_ctx *ctx, const vp_world_desc *w)
{
    const uint32_t k = world_key(w);
    for (int i = 0; i < ctx->n_cull_shapes; i++)
        if (ctx->cull_shape_key[i] == k) return ctx->cull_shapes[i];
    return ctx->cull_shape;
}

// vp_gpu_cull_mesh's argument checks, and its launch with explicit scratch:
// a ticket word and n look-back words (both unused when n == 1 and d_n is
// NULL).  The context's own scratch, a graph's private scratch and the
// static scratch of the allocation-free entry all go through it.
int cull_mesh_check(const vp_world_desc *world, const uint8_t *d_voxels,
                    const uint32_t *d_chunk_ids, uint32_t flags, const vp_splat_out *out);
int cull_mesh_launch(vp_ctx *ctx, const vp_world_desc *world, const uint8_t *d_voxels,
                     const int16_t *d_uniform, const vp_shadow_desc *shadow,
                     const uint32_t *d_chunk_ids, uint32_t n, const uint32_t *d_n,
                     uint32_t flags, const vp_splat_out *out, uint32_t *ticket,
                     uint64_t *state, hipStream_t st);

// The RLE encode of vp_gpu_rle with explicit scratch (a ticket word and n
// look-back words), an optional device-resident count d_n <= n and an
// optional uniform table (d_uniform_in, per chunk: the value every voxel
// holds, or -1): a chunk marked uniform there is encoded without reading it.
int rle_encode_launch(vp_ctx *ctx, const vp_world_desc *world, const uint8_t *d_voxels,
                      const uint32_t *d_chunk_ids, uint32_t n, const uint32_t *d_n,
                      uint32_t *d_words, uint64_t capacity_words, uint64_t *d_word_offsets,
                      int16_t *d_uniform_out, const int16_t *d_uniform_in, uint32_t *ticket,
                      uint64_t *state, hipStream_t st);

// Stores of output the kernel never reads back: streamed past the caches
// (non-temporal), so they do not evict the inputs still being read.
__device__ __forceinline__ void nt_store16(void *p, uint4 v)
{
    typedef unsigned int v4u __attribute__((ext_vector_type(4)));
    const v4u x = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(x, reinterpret_cast<v4u *>(p));
}
__device__ __forceinline__ void nt_store8(void *p, uint2 v)
{
    typedef unsigned int v2u __attribute__((ext_vector_type(2)));
    const v2u x = {v.x, v.y};
    __builtin_nontemporal_store(x, reinterpret_cast<v2u *>(p));
}

// relaxed device-scope atomics on 64-bit words (coherent across XCDs)
__device__ __forceinline__ uint64_t ld_relaxed(const uint64_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_relaxed(uint64_t *p, uint64_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace vp
