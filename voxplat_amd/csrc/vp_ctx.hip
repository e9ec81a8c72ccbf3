// vp_ctx.hip -- context: device selection, scratch arenas, status word.
#include "vp_internal.h"
#include "../../include/vp_host.h"

#include <new>

namespace vp {

int ensure_state(vp_ctx *ctx, uint64_t words)
{
    if (words <= ctx->state_cap) return VP_OK;
    uint64_t cap = ctx->state_cap ? ctx->state_cap : 4096;
    while (cap < words) cap *= 2;
    uint64_t *p = nullptr;
    if (hipMalloc(&p, cap * sizeof(uint64_t)) != hipSuccess) return VP_ERR_NOMEM;
    if (ctx->d_state) {
        // the old buffer may still be in use by queued work
        if (hipDeviceSynchronize() != hipSuccess) return VP_ERR_HIP;
        (void)hipFree(ctx->d_state);
    }
    ctx->d_state = p;
    ctx->state_cap = cap;
    return VP_OK;
}

int ensure_frame(vp_ctx *ctx, uint64_t pixels)
{
    if (pixels <= ctx->frame_cap && pixels <= ctx->rays_cap) return VP_OK;
    // both buffers are allocated before either capacity changes, so a failed
    // allocation leaves the old (smaller) pair and its capacities in place
    uint64_t *p = nullptr;
    float4 *r = nullptr;
    if (hipMalloc(&p, pixels * sizeof(uint64_t)) != hipSuccess) return VP_ERR_NOMEM;
    if (hipMalloc(&r, pixels * sizeof(float4)) != hipSuccess) {
        (void)hipFree(p);
        return VP_ERR_NOMEM;
    }
    if (ctx->d_frame || ctx->d_rays) {
        // the old buffers may still be in use by queued work
        if (hipDeviceSynchronize() != hipSuccess) {
            (void)hipFree(p);
            (void)hipFree(r);
            return VP_ERR_HIP;
        }
        if (ctx->d_frame) (void)hipFree(ctx->d_frame);
        if (ctx->d_rays) (void)hipFree(ctx->d_rays);
    }
    ctx->d_frame = p;
    ctx->d_rays = r;
    ctx->frame_cap = pixels;
    ctx->rays_cap = pixels;
    return VP_OK;
}

int ensure_blocks(vp_ctx *ctx, uint64_t blocks)
{
    if (blocks <= ctx->blk_cap) return VP_OK;
    uint32_t *p = nullptr;
    // a bin and an order slot per block, then 2 * 256 bin counters and the total
    if (hipMalloc(&p, (2 * blocks + 2 * 256 + 1) * sizeof(uint32_t)) != hipSuccess)
        return VP_ERR_NOMEM;
    if (ctx->d_blk) {
        if (hipDeviceSynchronize() != hipSuccess) return VP_ERR_HIP;
        (void)hipFree(ctx->d_blk);
    }
    ctx->d_blk = p;
    ctx->blk_cap = blocks;
    return VP_OK;
}

int ensure_tiles(vp_ctx *ctx, uint64_t words)
{
    if (words <= ctx->tiles_words) return VP_OK;
    uint32_t *p = nullptr;
    if (hipMalloc(&p, words * sizeof(uint32_t)) != hipSuccess) return VP_ERR_NOMEM;
    if (ctx->d_tiles) {
        if (hipDeviceSynchronize() != hipSuccess) {
            (void)hipFree(p);
            return VP_ERR_HIP;
        }
        (void)hipFree(ctx->d_tiles);
    }
    ctx->d_tiles = p;
    ctx->tiles_words = words;
    return VP_OK;
}

int ensure_cmap(vp_ctx *ctx, uint64_t chunks)
{
    if (chunks <= ctx->cmap_cap) return VP_OK;
    uint32_t *p = nullptr;
    if (hipMalloc(&p, chunks * sizeof(uint32_t)) != hipSuccess) return VP_ERR_NOMEM;
    if (ctx->d_cmap) {
        if (hipDeviceSynchronize() != hipSuccess) return VP_ERR_HIP;
        (void)hipFree(ctx->d_cmap);
    }
    ctx->d_cmap = p;
    ctx->cmap_cap = chunks;
    return VP_OK;
}

int ensure_oct(vp_ctx *ctx, uint64_t slots, uint64_t blocks)
{
    if (slots > ctx->oct_cap) {
        uint32_t *p = nullptr;
        if (hipMalloc(&p, slots * sizeof(uint32_t)) != hipSuccess) return VP_ERR_NOMEM;
        if (ctx->d_oct) {
            if (hipDeviceSynchronize() != hipSuccess) {
                (void)hipFree(p);
                return VP_ERR_HIP;
            }
            (void)hipFree(ctx->d_oct);
        }
        ctx->d_oct = p;
        ctx->oct_cap = slots;
    }
    if (3 * blocks > ctx->oct_part_cap) {
        uint64_t *p = nullptr;
        if (hipMalloc(&p, 3 * blocks * sizeof(uint64_t)) != hipSuccess) return VP_ERR_NOMEM;
        if (ctx->d_oct_part) {
            if (hipDeviceSynchronize() != hipSuccess) {
                (void)hipFree(p);
                return VP_ERR_HIP;
            }
            (void)hipFree(ctx->d_oct_part);
        }
        ctx->d_oct_part = p;
        ctx->oct_part_cap = 3 * blocks;
    }
    return VP_OK;
}

}  // namespace vp

extern "C" int vp_ctx_create(int device, vp_ctx **out)
{
    if (!out) return VP_ERR_ARG;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count)
        return VP_ERR_ARG;
    VP_HIP_TRY(hipSetDevice(device));
    vp_ctx *c = new (std::nothrow) vp_ctx();
    if (!c) return VP_ERR_NOMEM;
    c->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
        delete c;
        return VP_ERR_HIP;
    }
    c->num_cus = prop.multiProcessorCount;
    if (hipMalloc(&c->d_status, 64) != hipSuccess || hipMemset(c->d_status, 0, 64) != hipSuccess) {
        delete c;
        return VP_ERR_NOMEM;
    }
    c->d_ticket = c->d_status + 4;
    if (vp::ensure_state(c, 4096) != VP_OK) {
        (void)hipFree(c->d_status);
        delete c;
        return VP_ERR_NOMEM;
    }
    if (vp::noalloc_init(c) != VP_OK) {
        (void)hipFree(c->d_status);
        (void)hipFree(c->d_state);
        delete c;
        return VP_ERR_HIP;
    }
    *out = c;
    return VP_OK;
}

extern "C" int vp_ctx_destroy(vp_ctx *ctx)
{
    if (!ctx) return VP_ERR_ARG;
    (void)hipDeviceSynchronize();
    if (ctx->d_status) (void)hipFree(ctx->d_status);
    if (ctx->d_state) (void)hipFree(ctx->d_state);
    if (ctx->d_frame) (void)hipFree(ctx->d_frame);
    if (ctx->d_rays) (void)hipFree(ctx->d_rays);
    if (ctx->d_blk) (void)hipFree(ctx->d_blk);
    if (ctx->d_tiles) (void)hipFree(ctx->d_tiles);
    if (ctx->d_cmap) (void)hipFree(ctx->d_cmap);
    if (ctx->d_oct) (void)hipFree(ctx->d_oct);
    if (ctx->d_oct_part) (void)hipFree(ctx->d_oct_part);
    delete ctx;
    return VP_OK;
}

extern "C" const char *vp_strerror(int code)
{
    switch (code) {
    case VP_OK: return "ok";
    case VP_ERR_ARG: return "invalid argument";
    case VP_ERR_CAPACITY: return "output capacity exceeded";
    case VP_ERR_HIP: return "HIP runtime error";
    case VP_ERR_UNSUPPORTED: return "unsupported configuration";
    case VP_ERR_DATA: return "malformed input data";
    case VP_ERR_NOMEM: return "out of device memory";
    default: return "unknown error";
    }
}

extern "C" int vp_ctx_status(vp_ctx *ctx, void *stream, uint32_t *status)
{
    if (!ctx || !status) return VP_ERR_ARG;
    hipStream_t st = vp::as_stream(stream);
    uint32_t v = 0;
    VP_HIP_TRY(hipMemcpyAsync(&v, ctx->d_status, sizeof(v), hipMemcpyDeviceToHost, st));
    VP_HIP_TRY(hipStreamSynchronize(st));
    VP_HIP_TRY(hipMemsetAsync(ctx->d_status, 0, sizeof(uint32_t), st));
    VP_HIP_TRY(hipStreamSynchronize(st));
    *status = v;
    return VP_OK;
}

extern "C" int vp_ctx_num_cus(vp_ctx *ctx)
{
    return ctx ? ctx->num_cus : VP_ERR_ARG;
}

extern "C" int vp_ctx_reserve(vp_ctx *ctx, uint64_t max_chunks, uint64_t max_pixels)
{
    if (!ctx) return VP_ERR_ARG;
    std::lock_guard<std::mutex> g(ctx->mu);
    int rc = vp::ensure_state(ctx, max_chunks ? max_chunks : 1);
    if (rc) return rc;
    if (max_pixels) rc = vp::ensure_frame(ctx, max_pixels);
    return rc;
}

extern "C" int vp_ctx_set_launch(vp_ctx *ctx, int block, int wg_per_cu)
{
    if (!ctx || (block != 256 && block != 512 && block != 1024) || wg_per_cu < 1 || wg_per_cu > 4)
        return VP_ERR_ARG;
    std::lock_guard<std::mutex> g(ctx->mu);
    ctx->cull_shape.block = block;
    ctx->cull_shape.wg_per_cu = wg_per_cu;
    return VP_OK;
}

extern "C" int vp_ctx_get_launch(vp_ctx *ctx, int *block, int *wg_per_cu)
{
    if (!ctx || !block || !wg_per_cu) return VP_ERR_ARG;
    std::lock_guard<std::mutex> g(ctx->mu);
    *block = ctx->cull_shape.block;
    *wg_per_cu = ctx->cull_shape.wg_per_cu;
    return VP_OK;
}

extern "C" int vp_ctx_set_remesh_shape(vp_ctx *ctx, const vp_world_desc *world,
                                       const vp_remesh_shape *shape)
{
    if (!ctx || !shape) return VP_ERR_ARG;
    const int b = shape->block, w = shape->wg_per_cu, u = shape->records_in_flight;
    if ((b != 256 && b != 512 && b != 1024) || w < 1 || w > 4 || (u != 2 && u != 4))
        return VP_ERR_ARG;
    vp_remesh_shape_i s;
    s.block = b;
    s.wg_per_cu = w;
    s.u = u;
    std::lock_guard<std::mutex> g(ctx->mu);
    if (!world) {
        ctx->cull_shape = s;
        return VP_OK;
    }
    const uint32_t k = vp::world_key(world);
    for (int i = 0; i < ctx->n_cull_shapes; i++)
        if (ctx->cull_shape_key[i] == k) {
            ctx->cull_shapes[i] = s;
            return VP_OK;
        }
    if (ctx->n_cull_shapes == vp_ctx::kMaxShapes) return VP_ERR_CAPACITY;
    ctx->cull_shape_key[ctx->n_cull_shapes] = k;
    ctx->cull_shapes[ctx->n_cull_shapes++] = s;
    return VP_OK;
}

extern "C" int vp_ctx_get_remesh_shape(vp_ctx *ctx, const vp_world_desc *world,
                                       vp_remesh_shape *shape)
{
    if (!ctx || !shape) return VP_ERR_ARG;
    std::lock_guard<std::mutex> g(ctx->mu);
    const vp_remesh_shape_i s = world ? vp::remesh_shape(ctx, world) : ctx->cull_shape;
    shape->block = s.block;
    shape->wg_per_cu = s.wg_per_cu;
    shape->records_in_flight = s.u;
    return VP_OK;
}
