// noalloc.hip -- vp_gpu_cull_mesh_noalloc: the remesh of vp_gpu_cull_mesh
// with its work-queue ticket and look-back words in a static device array, so
// the call allocates nothing (for callers that may not call hipMalloc: code
// under graph capture, tight serving loops).
//
// The array is a file-scope __device__ symbol.  In host code its name is only
// a stand-in, so its device address is fetched once per device, when a
// context is created on it (noalloc_init, called by vp_ctx_create with that
// device current), and kept in the context; launches pass that address.
// Every call on one GPU shares the array: the launch clears the ticket and
// the look-back words it uses before the remesh kernel runs, and two calls
// that overlap on one GPU (two streams, or two contexts on the same device)
// race on it.
#include "vp_internal.h"

namespace {

// words 0..kTicketWords-1: the ticket word (and padding that keeps the
// look-back words on their own cache lines); then one look-back word per
// list entry
constexpr uint64_t kTicketWords = 32;
static __device__ __align__(256) uint64_t scratch[kTicketWords + VP_NOALLOC_MAX_CHUNKS];

}  // namespace

namespace vp {

int noalloc_init(vp_ctx *ctx)
{
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return VP_ERR_HIP;
    ctx->d_noalloc = static_cast<uint64_t *>(p);
    return VP_OK;
}

}  // namespace vp

extern "C" int vp_gpu_cull_mesh_noalloc(vp_ctx *ctx, const vp_world_desc *world,
                                        const uint8_t *d_voxels, const int16_t *d_uniform,
                                        const vp_shadow_desc *shadow,
                                        const uint32_t *d_chunk_ids, uint32_t n,
                                        const uint32_t *d_n, uint32_t flags,
                                        const vp_splat_out *out, void *stream)
{
    if (!ctx) return VP_ERR_ARG;
    int rc = vp::cull_mesh_check(world, d_voxels, d_chunk_ids, flags, out);
    if (rc) return rc;
    if (n == 0) return VP_OK;
    if ((uint64_t)n > VP_NOALLOC_MAX_CHUNKS) return VP_ERR_CAPACITY;
    if (!ctx->d_noalloc) return VP_ERR_HIP;
    uint32_t *ticket = reinterpret_cast<uint32_t *>(ctx->d_noalloc);
    uint64_t *state = ctx->d_noalloc + kTicketWords;
    // the launch shape is read from the context
    std::lock_guard<std::mutex> g(ctx->mu);
    return vp::cull_mesh_launch(ctx, world, d_voxels, d_uniform, shadow, d_chunk_ids, n, d_n,
                                flags, out, ticket, state, vp::as_stream(stream));
}
