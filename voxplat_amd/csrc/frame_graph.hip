// frame_graph.hip -- one frame of chunk management over a world resident in
// HBM, captured once in a hipGraph and replayed every frame
// (vp_host_frame_graph_*, include/vp_host.h).
//
// chunkset_manage (src/chunkset.c:246-313) walks every chunk each frame: a
// dirty chunk has its dirty flag cleared and is remeshed, any other chunk is
// compressed (chunk_compress, chunkset.c:213-230,300-306).  Here the walk is
// two small kernels that turn the per-chunk dirty bytes into a dirty list
// (ascending, as the loop visits them) and an idle list (the chunks the
// uniform table marks mixed, ascending, then the uniform ones, ascending: the
// encoder reads the first and not the second, and so streams its reads
// without gaps) on the device and clear the flags; the batched remesh then reads the dirty count and the RLE
// encode the idle count from device memory, so one captured graph serves
// every frame, whatever the edits marked.  The graph owns its scratch: the
// list kernels' per-workgroup counts, and a ticket word and look-back words
// for each of the two persistent kernels.
#include "vp_internal.h"
#include "../../include/vp_host.h"

#include <new>

namespace vp {

constexpr int kListBS = 256;
constexpr int kListPer = 4;                       // chunks per thread
constexpr int kListChunks = kListBS * kListPer;   // chunks per workgroup

// Kind of chunk i: 1 dirty, 2 idle and marked mixed (or no table), 3 idle
// and marked uniform.
__device__ __forceinline__ uint32_t list_kind(const uint8_t *dirty, const int16_t *uniform,
                                              uint32_t i)
{
    return dirty[i] ? 1u : (!uniform || uniform[i] < 0) ? 2u : 3u;
}

// Per workgroup: (dirty, idle mixed) chunks of its range, packed as
// d | m << 16; its idle uniform chunks are the rest of its range.
__global__ void __launch_bounds__(kListBS) k_frame_count(const uint8_t *dirty,
                                                          const int16_t *uniform, uint32_t nch,
                                                          uint32_t *blk)
{
    __shared__ uint32_t wsum[kListBS / 64];
    const uint32_t base = blockIdx.x * (uint32_t)kListChunks;
    uint32_t v = 0;
    for (int j = 0; j < kListPer; j++) {
        const uint32_t i = base + threadIdx.x * kListPer + j;
        if (i < nch) {
            const uint32_t k = list_kind(dirty, uniform, i);
            v += k == 1u ? 1u : k == 2u ? (1u << 16) : 0u;
        }
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += (uint32_t)__shfl_xor((int)v, d, 64);
    if ((threadIdx.x & 63u) == 0) wsum[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < kListBS / 64; w++) t += wsum[w];
        blk[blockIdx.x] = t;
    }
}

// Ordered scatter: chunk i goes to its place among the dirty chunks, the idle
// mixed ones or (after all those) the idle uniform ones, each ascending, and
// its dirty byte is cleared.  counts[0] = dirty, counts[1] = idle.  The
// workgroups read their counts before any byte is cleared: k_frame_count ran
// first, as its own launch.
__global__ void __launch_bounds__(kListBS) k_frame_scatter(uint8_t *dirty, const int16_t *uniform,
                                                            uint32_t nch, const uint32_t *blk,
                                                            uint32_t nblk, uint32_t *dirty_ids,
                                                            uint32_t *idle_ids, uint32_t *counts)
{
    __shared__ uint32_t wsum[4][kListBS / 64];
    __shared__ uint32_t s_pre[2], s_tot[2];
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    // this workgroup's prefix over the earlier workgroups, and the totals,
    // per kind (the packed per-workgroup counts are unpacked before summing)
    uint32_t pd = 0, pc = 0, td = 0, tc = 0;
    for (uint32_t b = threadIdx.x; b < nblk; b += kListBS) {
        const uint32_t v = blk[b], vd = v & 0xFFFFu, vc = v >> 16;
        td += vd;
        tc += vc;
        if (b < blockIdx.x) {
            pd += vd;
            pc += vc;
        }
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        pd += (uint32_t)__shfl_xor((int)pd, d, 64);
        pc += (uint32_t)__shfl_xor((int)pc, d, 64);
        td += (uint32_t)__shfl_xor((int)td, d, 64);
        tc += (uint32_t)__shfl_xor((int)tc, d, 64);
    }
    if (lane == 0) {
        wsum[0][w] = pd;
        wsum[1][w] = pc;
        wsum[2][w] = td;
        wsum[3][w] = tc;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t a0 = 0, a1 = 0, a2 = 0, a3 = 0;
        for (int k = 0; k < kListBS / 64; k++) {
            a0 += wsum[0][k];
            a1 += wsum[1][k];
            a2 += wsum[2][k];
            a3 += wsum[3][k];
        }
        s_pre[0] = a0;
        s_pre[1] = a1;
        s_tot[0] = a2;
        s_tot[1] = a3;
    }
    __syncthreads();
    // this thread's chunks, counted, then an exclusive scan over the workgroup
    const uint32_t i0 = blockIdx.x * (uint32_t)kListChunks + threadIdx.x * kListPer;
    uint8_t f[kListPer];
    uint32_t mine = 0;
    for (int j = 0; j < kListPer; j++) {
        const uint32_t i = i0 + j;
        f[j] = i < nch ? (uint8_t)list_kind(dirty, uniform, i) : 0;
        mine += f[j] == 1 ? 1u : f[j] == 2 ? (1u << 16) : 0u;
    }
    uint32_t x = mine;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)x, d, 64);
        if (lane >= (uint32_t)d) x += y;
    }
    __syncthreads();
    if (lane == 63) wsum[0][w] = x;
    __syncthreads();
    uint32_t wb = 0;
    for (uint32_t k = 0; k < w; k++) wb += wsum[0][k];
    const uint32_t local = wb + x - mine;          // packed, at most kListChunks per kind
    const uint32_t ld = local & 0xFFFFu, lm = local >> 16;
    // idle uniform chunks before this thread's first: those of the earlier
    // workgroups and threads, which are neither dirty nor idle mixed
    const uint32_t before = i0 < nch ? i0 : nch;
    uint32_t od = s_pre[0] + ld, om = s_pre[1] + lm;
    uint32_t ou = s_tot[1] + (before - s_pre[0] - ld - s_pre[1] - lm);
    for (int j = 0; j < kListPer; j++) {
        const uint32_t i = i0 + j;
        if (f[j] == 1) {
            dirty_ids[od++] = i;
            dirty[i] = 0;
        } else if (f[j] == 2) {
            idle_ids[om++] = i;
        } else if (f[j] == 3) {
            idle_ids[ou++] = i;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        counts[0] = s_tot[0];
        counts[1] = nch - s_tot[0];
    }
}

}  // namespace vp

struct vp_frame_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint32_t *d_blk = nullptr;      // per-workgroup list counts
    uint32_t *d_ticket = nullptr;   // 2 ticket words, 64 bytes apart
    uint64_t *d_state = nullptr;    // 2 * chunks look-back words
};

static void frame_graph_free(vp_frame_graph *g)
{
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    if (g->d_blk || g->d_ticket || g->d_state) {
        (void)hipDeviceSynchronize();     // a queued replay may still use the scratch
        if (g->d_blk) (void)hipFree(g->d_blk);
        if (g->d_ticket) (void)hipFree(g->d_ticket);
        if (g->d_state) (void)hipFree(g->d_state);
    }
    delete g;
}

extern "C" int vp_host_frame_graph_create(vp_ctx *ctx, const vp_world_desc *world,
                                          uint8_t *d_voxels, const int16_t *d_uniform,
                                          const vp_shadow_desc *shadow, uint8_t *d_dirty,
                                          uint32_t flags, const vp_frame_out *out,
                                          vp_frame_graph **graph)
{
    if (!ctx || !graph || !out || !d_dirty || !out->d_dirty_ids || !out->d_idle_ids ||
        !out->d_counts || !out->d_word_offsets)
        return VP_ERR_ARG;
    *graph = nullptr;
    if (flags & VP_CM_COUNT_ONLY) return VP_ERR_ARG;
    if (!out->d_words && out->capacity_words) return VP_ERR_ARG;
    int rc = vp::cull_mesh_check(world, d_voxels, out->d_dirty_ids, flags, &out->remesh);
    if (rc) return rc;
    const uint64_t nch64 = 1ull << (world->max_bitw[0] + world->max_bitw[1] + world->max_bitw[2]);
    // the packed list counts hold 16 bits per kind and workgroup; chunk ids are u32
    if (nch64 > 0xFFFFFFFFull) return VP_ERR_UNSUPPORTED;
    const uint32_t nch = (uint32_t)nch64;
    const uint32_t nblk = (nch + vp::kListChunks - 1) / vp::kListChunks;
    VP_HIP_TRY(hipSetDevice(ctx->device));
    vp_frame_graph *g = new (std::nothrow) vp_frame_graph();
    if (!g) return VP_ERR_NOMEM;
    if (hipMalloc(&g->d_blk, (size_t)nblk * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&g->d_ticket, 128) != hipSuccess ||
        hipMalloc(&g->d_state, 2 * (size_t)nch * sizeof(uint64_t)) != hipSuccess) {
        frame_graph_free(g);
        return VP_ERR_NOMEM;
    }
    hipStream_t s;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
        frame_graph_free(g);
        return VP_ERR_HIP;
    }
    rc = VP_ERR_HIP;
    if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) == hipSuccess) {
        int crc = VP_OK;
        hipLaunchKernelGGL(vp::k_frame_count, dim3(nblk), dim3(vp::kListBS), 0, s, d_dirty, d_uniform, nch,
                           g->d_blk);
        if (hipGetLastError() != hipSuccess) crc = VP_ERR_HIP;
        hipLaunchKernelGGL(vp::k_frame_scatter, dim3(nblk), dim3(vp::kListBS), 0, s, d_dirty, d_uniform, nch,
                           g->d_blk, nblk, out->d_dirty_ids, out->d_idle_ids, out->d_counts);
        if (hipGetLastError() != hipSuccess) crc = VP_ERR_HIP;
        if (crc == VP_OK) {
            // the launch shape is read from the context
            std::lock_guard<std::mutex> lk(ctx->mu);
            crc = vp::cull_mesh_launch(ctx, world, d_voxels, d_uniform, shadow, out->d_dirty_ids,
                                       nch, out->d_counts, flags, &out->remesh, g->d_ticket,
                                       g->d_state, s);
            if (crc == VP_OK)
                crc = vp::rle_encode_launch(ctx, world, d_voxels, out->d_idle_ids, nch,
                                            out->d_counts + 1, out->d_words, out->capacity_words,
                                            out->d_word_offsets, out->d_idle_uniform, d_uniform,
                                            g->d_ticket + 16, g->d_state + nch, s);
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
        frame_graph_free(g);
        return rc;
    }
    *graph = g;
    return VP_OK;
}

extern "C" int vp_host_frame_graph_launch(vp_frame_graph *graph, void *stream)
{
    if (!graph) return VP_ERR_ARG;
    return hipGraphLaunch(graph->exec, vp::as_stream(stream)) == hipSuccess ? VP_OK : VP_ERR_HIP;
}

extern "C" int vp_host_frame_graph_destroy(vp_frame_graph *graph)
{
    if (!graph) return VP_ERR_ARG;
    frame_graph_free(graph);
    return VP_OK;
}
