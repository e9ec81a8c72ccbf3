// remesh_graph.hip -- the batched remesh captured once in a hipGraph and
// replayed per frame (vp_host_remesh_graph_*, include/vp_host.h).
//
// This is chunkset_manage's per-frame loop over the dirty chunks
// (src/chunkset.c:246-313) as one replayable launch.  Each graph owns its
// scratch -- the work-queue ticket and the look-back words of the
// persistent kernel -- allocated here and freed with the graph, as each
// chunkset_manage thread owns its mask/work/geom buffers (chunkset.c:258-273).
// So no later call on the context (which may grow and replace the context's
// own scratch) can leave a graph pointing at freed memory, and a replay and a
// direct call on another stream never share a ticket.
#include "vp_internal.h"
#include "../../include/vp_host.h"

#include <new>

struct vp_remesh_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint32_t *d_ticket = nullptr;   // 64 bytes: the ticket word
    uint64_t *d_state = nullptr;    // n look-back words
};

static void graph_free(vp_remesh_graph *g)
{
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    if (g->d_ticket || g->d_state) {
        // a replay still queued may use the scratch
        (void)hipDeviceSynchronize();
        if (g->d_ticket) (void)hipFree(g->d_ticket);
        if (g->d_state) (void)hipFree(g->d_state);
    }
    delete g;
}

extern "C" int vp_host_remesh_graph_create(vp_ctx *ctx, const vp_world_desc *world,
                                           const uint8_t *d_voxels, const int16_t *d_uniform,
                                           const vp_shadow_desc *shadow,
                                           const uint32_t *d_chunk_ids, uint32_t n,
                                           const uint32_t *d_n, uint32_t flags,
                                           const vp_splat_out *out, vp_remesh_graph **graph)
{
    if (!ctx || !graph) return VP_ERR_ARG;
    *graph = nullptr;
    int rc = vp::cull_mesh_check(world, d_voxels, d_chunk_ids, flags, out);
    if (rc) return rc;
    if (n == 0) return VP_ERR_ARG;
    VP_HIP_TRY(hipSetDevice(ctx->device));
    vp_remesh_graph *g = new (std::nothrow) vp_remesh_graph();
    if (!g) return VP_ERR_NOMEM;
    if (hipMalloc(&g->d_ticket, 64) != hipSuccess ||
        hipMalloc(&g->d_state, (size_t)n * sizeof(uint64_t)) != hipSuccess) {
        graph_free(g);
        return VP_ERR_NOMEM;
    }
    hipStream_t s;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
        graph_free(g);
        return VP_ERR_HIP;
    }
    rc = VP_ERR_HIP;
    if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) == hipSuccess) {
        int crc;
        {
            // the launch shape is read from the context
            std::lock_guard<std::mutex> lk(ctx->mu);
            crc = vp::cull_mesh_launch(ctx, world, d_voxels, d_uniform, shadow, d_chunk_ids, n,
                                       d_n, flags, out, g->d_ticket, g->d_state, s);
        }
        hipError_t e = hipStreamEndCapture(s, &g->graph);
        if (crc != VP_OK)
            rc = crc;
        else if (e == hipSuccess &&
                 hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) == hipSuccess)
            rc = VP_OK;
    }
    (void)hipStreamDestroy(s);
    if (rc != VP_OK) {
        graph_free(g);
        return rc;
    }
    *graph = g;
    return VP_OK;
}

extern "C" int vp_host_remesh_graph_launch(vp_remesh_graph *graph, void *stream)
{
    if (!graph) return VP_ERR_ARG;
    return hipGraphLaunch(graph->exec, vp::as_stream(stream)) == hipSuccess ? VP_OK : VP_ERR_HIP;
}

extern "C" int vp_host_remesh_graph_destroy(vp_remesh_graph *graph)
{
    if (!graph) return VP_ERR_ARG;
    graph_free(graph);
    return VP_OK;
}
