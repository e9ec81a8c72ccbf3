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
    uint32_t *d_cmap = nullptr;        // draw lists: list entry + 1 per chunk index
    uint64_t  cmap_cap = 0;
    uint32_t *d_oct = nullptr;         // device draw-queue selection: key slots
    uint64_t  oct_cap = 0;
    uint64_t *d_oct_part = nullptr;    // and per-workgroup partial counts
    uint64_t  oct_part_cap = 0;
    // device address of the static scratch of vp_gpu_cull_mesh_noalloc on
    // this context's device (noalloc.hip); not owned
    uint64_t *d_noalloc = nullptr;
    // tuned launch shape of the persistent remesh kernel
    int cull_block = 512;              // threads per workgroup
    int cull_wg_per_cu = 2;            // resident workgroups per CU
};

namespace vp {

int ensure_state(vp_ctx *ctx, uint64_t words);
int ensure_frame(vp_ctx *ctx, uint64_t pixels);
int ensure_blocks(vp_ctx *ctx, uint64_t blocks);
int ensure_cmap(vp_ctx *ctx, uint64_t chunks);
int ensure_oct(vp_ctx *ctx, uint64_t slots, uint64_t blocks);
// fetch the static scratch's device address on the current device
int noalloc_init(vp_ctx *ctx);

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

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
// look-back words) and an optional device-resident count d_n <= n.
int rle_encode_launch(vp_ctx *ctx, const vp_world_desc *world, const uint8_t *d_voxels,
                      const uint32_t *d_chunk_ids, uint32_t n, const uint32_t *d_n,
                      uint32_t *d_words, uint64_t capacity_words, uint64_t *d_word_offsets,
                      int16_t *d_uniform_out, uint32_t *ticket, uint64_t *state, hipStream_t st);

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
