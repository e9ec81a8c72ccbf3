// drawlist.hip -- octree node lists and frame draw assembly for gfx950.
//
// Replaces the per-node vertex-buffer uploads of gfx_update_svl
// (src/gfx/vsplat.c:197-338) and the per-frame draw queue of gfx_svl_draw
// (src/gfx/vsplat.c:163-183), so that a frame's splats go from the remesh
// output to the splat resolve without a host copy.
//
// k_svl_nodes: one workgroup per node of one level (persistent, ticket order
//   = node index order).  The node's chunks, in the order x outer, y, z inner,
//   get their level-l record counts from a chunk -> list-entry map; a
//   workgroup scan gives each chunk's place in the node, the node total is
//   published and the node's output offset comes from a decoupled look-back;
//   then every thread copies records, finding its chunk by a binary search
//   over the chunk prefixes in LDS, so a node of thousands of small chunks
//   is still one coalesced copy.
// k_draw_gather: sixteen workgroups per draw, each copying a slice of the
//   node's records (and writing their level) to the draw's place in the
//   frame's list.
#include "vp_internal.h"

namespace vp {

constexpr uint64_t kNFlagAgg = 1ull << 62;
constexpr uint64_t kNFlagPre = 2ull << 62;
constexpr uint64_t kNValMask = (1ull << 62) - 1;
constexpr int kNodeBS = 512;
constexpr int kNodeMaxSub = 4096;   // chunks of a level-4 node: 16^3
constexpr int kGatherSlices = 16;   // workgroups per draw in k_draw_gather

// One launch builds the node buffers of up to five levels.  Tickets run
// over the listed levels in table order (the caller lists the high levels,
// with few large nodes, first); a node's look-back stops at its level's
// first ticket.
struct NodeLevel {
    uint32_t level;
    uint32_t nb0, nb1;           // node index bits of x, y
    uint64_t first;              // first ticket of this level
    uint64_t nnodes;
    uint64_t *out;
    uint64_t cap;
    uint64_t *node_off;
};

struct NodeArgs {
    const uint64_t *splats;      // remesh records, 8 bytes each
    const uint32_t *cmap;        // per chunk index: list entry + 1, 0 = not listed
    const uint64_t *chunk_off;   // n+1 exclusive prefix, int16 items
    const uint32_t *lcounts;     // n*5, int16 items
    uint32_t mb0, mb1;           // chunk index bits of x, y
    uint32_t max0, max1, max2;   // chunks per axis
    uint32_t nlev;
    NodeLevel lev[5];
    uint64_t ntickets;
    uint32_t *ticket;
    uint64_t *state;
    uint32_t *status;
};

__global__ void k_cmap_clear(uint32_t *m, uint64_t n)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        m[i] = 0u;
}

// entry + 1 of the last list entry naming each chunk
__global__ void k_cmap_fill(uint32_t *m, const uint32_t *ids, uint32_t n, uint32_t nchunks,
                            uint32_t *status)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t id = ids[k];
    if (id < nchunks) atomicMax(m + id, k + 1u);
    else atomicOr(status, (uint32_t)VP_STATUS_BAD_ID);
}

__global__ void k_node_reset(uint32_t *ticket, uint64_t *state, uint64_t n)
{
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i == 0) ticket[0] = 0u;
    for (uint64_t j = i; j < n; j += (uint64_t)gridDim.x * blockDim.x) state[j] = 0ull;
}

__device__ __forceinline__ uint32_t n_scan_excl(uint32_t v, uint32_t *wsum, uint32_t &total)
{
    constexpr int NW = kNodeBS / 64;
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d, 64);
        if (lane >= (uint32_t)d) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int i = 0; i < NW; i++) {
            const uint32_t t = wsum[i];
            wsum[i] = run;
            run += t;
        }
        wsum[NW] = run;
    }
    __syncthreads();
    total = wsum[NW];
    const uint32_t r = wsum[w] + x - v;
    __syncthreads();
    return r;
}

__global__ void __launch_bounds__(kNodeBS) k_svl_nodes(NodeArgs a)
{
    __shared__ uint32_t pre[kNodeMaxSub];     // records of the node's earlier chunks
    __shared__ uint64_t src[kNodeMaxSub];     // first record of each chunk's level-l segment
    __shared__ uint32_t wsum[kNodeBS / 64 + 1];
    __shared__ uint32_t s_ticket;
    __shared__ uint64_t s_base;
    for (;;) {
        if (threadIdx.x == 0) s_ticket = atomicAdd(a.ticket, 1u);
        __syncthreads();
        const uint64_t t = s_ticket;
        __syncthreads();
        if (t >= a.ntickets) return;
        uint32_t li = 0;
        while (li + 1 < a.nlev && t >= a.lev[li + 1].first) li++;
        const NodeLevel &L = a.lev[li];
        const uint32_t l = L.level;
        const uint64_t k = t - L.first;          // node index within the level
        const uint32_t nx = (uint32_t)(k & ((1ull << L.nb0) - 1ull));
        const uint32_t ny = (uint32_t)((k >> L.nb0) & ((1ull << L.nb1) - 1ull));
        const uint32_t nz = (uint32_t)(k >> (L.nb0 + L.nb1));
        // the node's chunk box (vsplat.c:213-222)
        const uint32_t x0 = nx << l, y0 = ny << l, z0 = nz << l;
        const uint32_t x1 = min((nx + 1u) << l, a.max0), y1 = min((ny + 1u) << l, a.max1),
                       z1 = min((nz + 1u) << l, a.max2);
        const uint32_t dx = x1 > x0 ? x1 - x0 : 0u, dy = y1 > y0 ? y1 - y0 : 0u,
                       dz = z1 > z0 ? z1 - z0 : 0u;
        const uint32_t S = dx * dy * dz;
        // chunk s of the node in scan order: x outer, y, z inner (vsplat.c:263-266)
        auto chunk_of = [&](uint32_t s) -> uint32_t {
            const uint32_t z = s % dz, y = (s / dz) % dy, x = s / (dz * dy);
            return (((z0 + z) << a.mb1 | (y0 + y)) << a.mb0) | (x0 + x);
        };
        const uint32_t K = (S + kNodeBS - 1) / kNodeBS;
        const uint32_t s0 = threadIdx.x * K;
        uint32_t mine = 0;
        for (uint32_t j = 0; j < K; j++) {
            const uint32_t s = s0 + j;
            if (s >= S) break;
            const uint32_t e = a.cmap[chunk_of(s)];
            uint32_t c = 0;
            uint64_t so = 0;
            if (e) {
                const uint32_t *lc = a.lcounts + (uint64_t)(e - 1u) * 5u;
                c = lc[l] / 4u;
                so = a.chunk_off[e - 1u] / 4u;
                for (uint32_t i = 0; i < l; i++) so += lc[i] / 4u;
            }
            pre[s] = c;                   // the count, turned into a prefix below
            src[s] = so;
            mine += c;
        }
        uint32_t total;
        uint32_t run = n_scan_excl(mine, wsum, total);
        for (uint32_t j = 0; j < K; j++) {
            const uint32_t s = s0 + j;
            if (s >= S) break;
            const uint32_t c = pre[s];
            pre[s] = run;
            run += c;
        }
        // ---- publish the node's size, look back for its offset ----
        if (threadIdx.x < 64) {
            uint64_t excl = 0;
            if (k == 0) {
                if (threadIdx.x == 0) st_relaxed(a.state + t, kNFlagPre | total);
            } else {
                if (threadIdx.x == 0) st_relaxed(a.state + t, kNFlagAgg | total);
                int64_t top = (int64_t)t - 1;
                const int64_t lstart = (int64_t)L.first;
                const uint32_t ln = threadIdx.x & 63u;
                for (;;) {
                    const int64_t j = top - (int64_t)ln;
                    const uint64_t sv = (j >= lstart) ? ld_relaxed(a.state + j) : kNFlagPre;
                    const uint64_t f = sv >> 62;
                    const uint64_t pr = __ballot(f == 2);
                    const uint64_t zero = __ballot(f == 0);
                    const uint64_t need = pr ? (pr & (~pr + 1ull)) * 2ull - 1ull : ~0ull;
                    if (zero & need) {
                        __builtin_amdgcn_s_sleep(1);
                        continue;
                    }
                    uint64_t v = ((need >> ln) & 1ull) ? (sv & kNValMask) : 0ull;
#pragma unroll
                    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
                    excl += v;
                    if (pr) break;
                    top -= 64;
                }
                if (threadIdx.x == 0) st_relaxed(a.state + t, kNFlagPre | (excl + total));
            }
            if (threadIdx.x == 0) {
                s_base = excl;
                L.node_off[k] = excl;
                if (k + 1 == L.nnodes) L.node_off[L.nnodes] = excl + total;
                if (L.out && excl + total > L.cap) atomicOr(a.status, (uint32_t)VP_STATUS_OVERFLOW);
            }
        }
        __syncthreads();
        const uint64_t base = s_base;
        if (L.out && total && base + total <= L.cap) {
            // U records per thread per step, all loads issued before any
            // store: a node of a high level is one of few workgroups, so its
            // copy has to keep many loads in flight on its own
            constexpr int U = 4;
            for (uint32_t j0 = threadIdx.x; j0 < total; j0 += U * kNodeBS) {
                uint64_t v[U];
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const uint32_t j = j0 + (uint32_t)u * kNodeBS;
                    v[u] = 0;
                    if (j < total) {
                        // the last chunk whose prefix is <= j holds record j
                        uint32_t lo = 0, hi = S - 1u;
                        while (lo < hi) {
                            const uint32_t mid = (lo + hi + 1u) >> 1;
                            if (pre[mid] <= j) lo = mid;
                            else hi = mid - 1u;
                        }
                        v[u] = a.splats[src[lo] + (j - pre[lo])];
                    }
                }
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const uint32_t j = j0 + (uint32_t)u * kNodeBS;
                    if (j < total) L.out[base + j] = v[u];
                }
            }
        }
        __syncthreads();
    }
}

struct GatherArgs {
    const uint64_t *nodes[5];
    const uint64_t *noff[5];
    uint64_t ncount[5];
    const vp_draw *draws;
    uint32_t n;
    const uint64_t *d_n;        // nullable: the draw count in device memory, n its bound
    uint64_t *out;
    uint8_t *lod;
    uint64_t cap;
    uint32_t *status;
};

// Draw d's records are cut into gridDim.y equal slices; workgroup (x, y)
// copies slice y of draws x, x + gridDim.x, ...: a near chunk's tens of
// thousands of records are not left to a single workgroup.
__global__ void __launch_bounds__(256) k_draw_gather(GatherArgs a)
{
    const uint64_t ns = gridDim.y, sl = blockIdx.y;
    const uint32_t n = a.d_n ? (uint32_t)min(*a.d_n, (uint64_t)a.n) : a.n;
    for (uint32_t d = blockIdx.x; d < n; d += gridDim.x) {
        const uint32_t l = a.draws[d].level, node = a.draws[d].node;
        const uint64_t first = a.draws[d].first;
        if (l >= 5 || node >= a.ncount[l] || !a.nodes[l] || !a.noff[l]) {
            if (threadIdx.x == 0) atomicOr(a.status, (uint32_t)VP_STATUS_BAD_ID);
            continue;
        }
        const uint64_t s0 = a.noff[l][node], cnt = a.noff[l][node + 1] - s0;
        if (first + cnt > a.cap) {
            if (threadIdx.x == 0 && sl == 0) atomicOr(a.status, (uint32_t)VP_STATUS_OVERFLOW);
            continue;
        }
        const uint64_t *src = a.nodes[l] + s0;
        const uint64_t jb = cnt * sl / ns, je = cnt * (sl + 1) / ns;   // this workgroup's slice
        constexpr int U = 4;                    // loads in flight per thread
        for (uint64_t j0 = jb + threadIdx.x; j0 < je; j0 += U * blockDim.x) {
            uint64_t v[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const uint64_t j = j0 + (uint64_t)u * blockDim.x;
                v[u] = j < je ? src[j] : 0ull;
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                const uint64_t j = j0 + (uint64_t)u * blockDim.x;
                if (j < je) {
                    a.out[first + j] = v[u];
                    a.lod[first + j] = (uint8_t)l;
                }
            }
        }
    }
}

}  // namespace vp

using namespace vp;

static int svl_nodes_launch(vp_ctx *ctx, const vp_world_desc *world, uint32_t nlev,
                            const uint32_t *levels, const void *d_splats, const uint32_t *d_chunk_ids,
                            uint32_t n, const uint64_t *d_chunk_offsets, const uint32_t *d_level_counts,
                            const vp_node_out *outs, void *stream)
{
    if (!ctx || !world || !outs) return VP_ERR_ARG;
    if (n && (!d_splats || !d_chunk_ids || !d_chunk_offsets || !d_level_counts)) return VP_ERR_ARG;
    const uint32_t mbs = (uint32_t)world->max_bitw[0] + world->max_bitw[1] + world->max_bitw[2];
    if (mbs > 31u) return VP_ERR_ARG;
    NodeArgs a = {};
    a.nlev = nlev;
    uint64_t tickets = 0;
    for (uint32_t i = 0; i < nlev; i++) {
        const vp_node_out *o = outs + i;
        const uint32_t l = levels[i];
        if (l > 4 || !o->d_node_offsets) return VP_ERR_ARG;
        const bool count_only = !o->d_node_splats && o->capacity_records == 0;
        if (!count_only && !o->d_node_splats) return VP_ERR_ARG;
        NodeLevel &L = a.lev[i];
        L.level = l;
        L.nb0 = vp_node_bits(world, l, 0);
        L.nb1 = vp_node_bits(world, l, 1);
        L.first = tickets;
        L.nnodes = vp_node_count(world, l);
        L.out = count_only ? nullptr : static_cast<uint64_t *>(o->d_node_splats);
        L.cap = o->capacity_records;
        L.node_off = o->d_node_offsets;
        tickets += L.nnodes;
    }
    hipStream_t st = as_stream(stream);
    const uint64_t nchunks = 1ull << mbs;
    std::lock_guard<std::mutex> g(ctx->mu);
    int rc = ensure_state(ctx, tickets);
    if (rc) return rc;
    rc = ensure_cmap(ctx, nchunks);
    if (rc) return rc;
    const int cgrid = (int)((nchunks + 255) / 256 < 4096 ? (nchunks + 255) / 256 : 4096);
    hipLaunchKernelGGL(k_cmap_clear, dim3(cgrid), dim3(256), 0, st, ctx->d_cmap, nchunks);
    VP_HIP_TRY(hipGetLastError());
    if (n) {
        hipLaunchKernelGGL(k_cmap_fill, dim3((n + 255) / 256), dim3(256), 0, st, ctx->d_cmap, d_chunk_ids,
                           n, (uint32_t)nchunks, ctx->d_status);
        VP_HIP_TRY(hipGetLastError());
    }
    const int rgrid = (int)((tickets + 255) / 256 < 1024 ? (tickets + 255) / 256 : 1024);
    hipLaunchKernelGGL(k_node_reset, dim3(rgrid), dim3(256), 0, st, ctx->d_ticket, ctx->d_state, tickets);
    VP_HIP_TRY(hipGetLastError());
    a.splats = static_cast<const uint64_t *>(d_splats);
    a.cmap = ctx->d_cmap;
    a.chunk_off = d_chunk_offsets;
    a.lcounts = d_level_counts;
    a.mb0 = world->max_bitw[0];
    a.mb1 = world->max_bitw[1];
    a.max0 = 1u << world->max_bitw[0];
    a.max1 = 1u << world->max_bitw[1];
    a.max2 = 1u << world->max_bitw[2];
    a.ntickets = tickets;
    a.ticket = ctx->d_ticket;
    a.state = ctx->d_state;
    a.status = ctx->d_status;
    const uint64_t resident = (uint64_t)ctx->num_cus * 3u;
    const int grid = (int)(tickets < resident ? tickets : resident);
    hipLaunchKernelGGL(k_svl_nodes, dim3(grid), dim3(kNodeBS), 0, st, a);
    return hipGetLastError() == hipSuccess ? VP_OK : VP_ERR_HIP;
}

extern "C" int vp_gpu_svl_nodes(vp_ctx *ctx, const vp_world_desc *world, uint32_t level,
                                const void *d_splats, const uint32_t *d_chunk_ids, uint32_t n,
                                const uint64_t *d_chunk_offsets, const uint32_t *d_level_counts,
                                const vp_node_out *out, void *stream)
{
    if (level > 4) return VP_ERR_ARG;
    return svl_nodes_launch(ctx, world, 1, &level, d_splats, d_chunk_ids, n, d_chunk_offsets,
                            d_level_counts, out, stream);
}

extern "C" int vp_gpu_svl_nodes_all(vp_ctx *ctx, const vp_world_desc *world, const void *d_splats,
                                    const uint32_t *d_chunk_ids, uint32_t n,
                                    const uint64_t *d_chunk_offsets, const uint32_t *d_level_counts,
                                    const vp_node_out out[5], void *stream)
{
    // levels 4..0: the few large nodes of the high levels take their tickets first
    const uint32_t levels[5] = {4, 3, 2, 1, 0};
    if (!out) return VP_ERR_ARG;
    const vp_node_out ordered[5] = {out[4], out[3], out[2], out[1], out[0]};
    return svl_nodes_launch(ctx, world, 5, levels, d_splats, d_chunk_ids, n, d_chunk_offsets,
                            d_level_counts, ordered, stream);
}

static int draw_gather(vp_ctx *ctx, const vp_world_desc *world, const void *const d_node_splats[5],
                       const uint64_t *const d_node_offsets[5], const vp_draw *d_draws, uint32_t n,
                       const uint64_t *d_n, void *d_out, uint8_t *d_out_lod,
                       uint64_t capacity_records, void *stream)
{
    if (!ctx || !world || !d_node_splats || !d_node_offsets) return VP_ERR_ARG;
    if (n == 0) return VP_OK;
    if (!d_draws || !d_out || !d_out_lod) return VP_ERR_ARG;
    GatherArgs a = {};
    for (int l = 0; l < 5; l++) {
        a.nodes[l] = static_cast<const uint64_t *>(d_node_splats[l]);
        a.noff[l] = d_node_offsets[l];
        a.ncount[l] = vp_node_count(world, (uint32_t)l);
    }
    a.draws = d_draws;
    a.n = n;
    a.d_n = d_n;
    a.out = static_cast<uint64_t *>(d_out);
    a.lod = d_out_lod;
    a.cap = capacity_records;
    a.status = ctx->d_status;
    const int grid = (int)(n < 65535u ? n : 65535u);
    hipLaunchKernelGGL(k_draw_gather, dim3(grid, kGatherSlices), dim3(256), 0, as_stream(stream), a);
    return hipGetLastError() == hipSuccess ? VP_OK : VP_ERR_HIP;
}

extern "C" int vp_gpu_draw_gather(vp_ctx *ctx, const vp_world_desc *world,
                                  const void *const d_node_splats[5],
                                  const uint64_t *const d_node_offsets[5], const vp_draw *d_draws,
                                  uint32_t n, void *d_out, uint8_t *d_out_lod,
                                  uint64_t capacity_records, void *stream)
{
    return draw_gather(ctx, world, d_node_splats, d_node_offsets, d_draws, n, nullptr, d_out,
                       d_out_lod, capacity_records, stream);
}

extern "C" int vp_gpu_draw_gather_counted(vp_ctx *ctx, const vp_world_desc *world,
                                          const void *const d_node_splats[5],
                                          const uint64_t *const d_node_offsets[5],
                                          const vp_draw *d_draws, uint32_t draw_cap,
                                          const uint64_t *d_counts, void *d_out,
                                          uint8_t *d_out_lod, uint64_t capacity_records,
                                          void *stream)
{
    if (!d_counts) return VP_ERR_ARG;
    return draw_gather(ctx, world, d_node_splats, d_node_offsets, d_draws, draw_cap, d_counts,
                       d_out, d_out_lod, capacity_records, stream);
}
