// octree.hip -- the per-frame draw queue selected on the device
// (vp_gpu_octree_cull, include/vp_gpu.h).
//
// octree_cull (src/game.c:319-395) walks the LOD octree depth first from every
// level-4 node (x outer, y, z inner; game.c:507-512): a node outside the
// world or whose bounding sphere (radius 0.85 of its edge) fails the frustum
// test (gfx_fcull_visible, src/gfx/fcull.c:38-45) is skipped; a node farther
// than level * 1024 from the camera, or at level 0, is drawn (a level-0 chunk
// with a quad mesh is queued as a mesh draw instead); any other node visits
// its eight children (x outer, y, z inner).  A node without records is not
// drawn.  The queue's `first` is the running record total.
//
// Whether a node is drawn depends only on the node and its ancestors, so
// every node of every level is decided by its own thread (k_oct_mark).  The
// drawn nodes cover disjoint parts of the world, and the depth-first order of
// the walk is the order of their first level-0 descendant under the key
//   root index (x outer, y, z inner) * 4096 + the child steps of levels 3..0
//   (three bits each, x, y, z),
// so each drawn node is written to the slot of that key, and an ordered
// compaction of the slots (k_oct_count, k_oct_write) gives the queue, the
// running totals and the counts, all in device memory.
#include "vp_internal.h"

namespace vp {

constexpr int kOctBS = 256;
constexpr int kOctPer = 4;                       // slots per thread
constexpr int kOctSlots = kOctBS * kOctPer;      // slots per workgroup

struct OctArgs {
    float planes[24];
    float origin[3];
    uint32_t rb;                 // root_bitw
    uint32_t mb[3];              // max_bitw
    uint32_t nr[3];              // level-4 roots per axis
    uint64_t lvl_first[6];       // first thread of each level in k_oct_mark
    const uint64_t *noff[5];     // node offsets per level (records), count + 1 each
    const uint8_t *chunk_mesh;   // nullable
    uint32_t *slots;             // nr[0]*nr[1]*nr[2] * 4096 entries, 0 = empty
    uint64_t nslots;
    uint64_t *partial;           // 3 per workgroup of k_oct_count
    vp_draw *draws;
    uint32_t draw_cap;
    uint32_t *mesh;
    uint32_t mesh_cap;
    uint64_t *counts;            // [0] draws, [1] mesh draws, [2] records
    uint32_t *status;
};

// gfx_fcull_visible (fcull.c:38-45), in the host's float order
__device__ __forceinline__ bool oct_visible(const float *p, const float c[3], float radius)
{
    for (int i = 0; i < 6; i++) {
        const float d = (p[4 * i] * c[0] + p[4 * i + 1] * c[1]) + p[4 * i + 2] * c[2];
        if (d + p[4 * i + 3] < -radius) return false;
    }
    return true;
}

// Is node (lod, x, y, z) visible, and its centre's distance to the camera.
__device__ __forceinline__ bool oct_node(const OctArgs &a, uint32_t lod, uint32_t x, uint32_t y,
                                         uint32_t z, float &dist)
{
    const uint32_t ws_root = (1u << a.rb) << lod;
    const float c[3] = {(float)(x * ws_root + ws_root / 2), (float)(y * ws_root + ws_root / 2),
                        (float)(z * ws_root + ws_root / 2)};
    if (!oct_visible(a.planes, c, (float)((double)ws_root * 0.85))) return false;
    const float d0 = a.origin[0] - c[0], d1 = a.origin[1] - c[1], d2 = a.origin[2] - c[2];
    dist = sqrtf((d0 * d0 + d1 * d1) + d2 * d2);
    return true;
}

__device__ __forceinline__ uint32_t oct_nbits(const OctArgs &a, uint32_t lod, int axis)
{
    const uint32_t b = a.mb[axis];
    return b - (lod < b ? lod : b);
}

__global__ void k_oct_clear(uint32_t *slots, uint64_t n)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        slots[i] = 0u;
}

// One thread per node of every level: drawn nodes write their slot.
// Slot value: 1 + (kind << 31 | level << 28 | node), kind 1 = mesh draw.
__global__ void __launch_bounds__(256) k_oct_mark(OctArgs a)
{
    const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (t >= a.lvl_first[5]) return;
    uint32_t l = 0;
    while (t >= a.lvl_first[l + 1]) l++;
    const uint64_t i = t - a.lvl_first[l];
    const uint32_t b0 = oct_nbits(a, l, 0), b1 = oct_nbits(a, l, 1);
    const uint32_t x = (uint32_t)(i & ((1ull << b0) - 1ull));
    const uint32_t y = (uint32_t)((i >> b0) & ((1ull << b1) - 1ull));
    const uint32_t z = (uint32_t)(i >> (b0 + b1));
    // every ancestor must be visible and near enough to be walked into
    for (uint32_t k = 4; k > l; k--) {
        float dist;
        const uint32_t s = k - l;
        if (!oct_node(a, k, x >> s, y >> s, z >> s, dist)) return;
        if (dist > (float)(k * 1024u)) return;          // drawn at the ancestor
    }
    float dist;
    if (!oct_node(a, l, x, y, z, dist)) return;
    // a level-0 node is always a leaf
    if (!(dist > (float)(l * 1024u) || l == 0)) return;
    const uint32_t node = ((z << b1 | y) << b0) | x;
    uint32_t kind = 0;
    if (l == 0 && a.chunk_mesh && a.chunk_mesh[node]) {
        kind = 1;
    } else {
        if (a.noff[l][node + 1] == a.noff[l][node]) return;   // no vertex buffer
    }
    // depth-first key of the node's first level-0 descendant
    const uint32_t cx = x << l, cy = y << l, cz = z << l;
    const uint64_t root = ((uint64_t)(cx >> 4) * a.nr[1] + (cy >> 4)) * a.nr[2] + (cz >> 4);
    uint32_t path = 0;
    for (int k = 3; k >= 0; k--)
        path |= ((((cx >> k) & 1u) << 2) | (((cy >> k) & 1u) << 1) | ((cz >> k) & 1u)) << (3 * k);
    a.slots[root * 4096u + path] = 1u + ((kind << 31) | (l << 28) | node);
}

__device__ __forceinline__ uint64_t oct_items(const OctArgs &a, uint32_t v)
{
    const uint32_t e = v - 1u, l = (e >> 28) & 7u, node = e & 0x0FFFFFFFu;
    return (e >> 31) ? 0ull : a.noff[l][node + 1] - a.noff[l][node];
}

// Per workgroup of slots: draws, mesh draws, records.
__global__ void __launch_bounds__(kOctBS) k_oct_count(OctArgs a)
{
    __shared__ uint64_t ws[3][kOctBS / 64];
    const uint64_t base = blockIdx.x * (uint64_t)kOctSlots;
    uint64_t nd = 0, nm = 0, it = 0;
    for (int j = 0; j < kOctPer; j++) {
        const uint64_t s = base + threadIdx.x * kOctPer + j;
        const uint32_t v = s < a.nslots ? a.slots[s] : 0u;
        if (!v) continue;
        if ((v - 1u) >> 31) nm++;
        else { nd++; it += oct_items(a, v); }
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        nd += __shfl_xor(nd, d, 64);
        nm += __shfl_xor(nm, d, 64);
        it += __shfl_xor(it, d, 64);
    }
    if ((threadIdx.x & 63u) == 0) {
        ws[0][threadIdx.x >> 6] = nd;
        ws[1][threadIdx.x >> 6] = nm;
        ws[2][threadIdx.x >> 6] = it;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        uint64_t s = 0;
        for (int w = 0; w < kOctBS / 64; w++) s += ws[threadIdx.x][w];
        a.partial[3 * (uint64_t)blockIdx.x + threadIdx.x] = s;
    }
}

// Ordered compaction of the slots into the draw queue and the mesh list.
__global__ void __launch_bounds__(kOctBS) k_oct_write(OctArgs a, uint32_t nblk)
{
    __shared__ uint64_t ws[3][kOctBS / 64];
    __shared__ uint64_t s_pre[3], s_tot[3];
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    // prefix over the earlier workgroups, and the totals
    uint64_t pre[3] = {0, 0, 0}, tot[3] = {0, 0, 0};
    for (uint32_t b = threadIdx.x; b < nblk; b += kOctBS)
        for (int k = 0; k < 3; k++) {
            const uint64_t v = a.partial[3 * (uint64_t)b + k];
            tot[k] += v;
            if (b < blockIdx.x) pre[k] += v;
        }
    for (int k = 0; k < 3; k++) {
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) {
            pre[k] += __shfl_xor(pre[k], d, 64);
            tot[k] += __shfl_xor(tot[k], d, 64);
        }
    }
    if (lane == 0)
        for (int k = 0; k < 3; k++) ws[k][w] = pre[k];
    __syncthreads();
    if (threadIdx.x < 3) {
        uint64_t s = 0;
        for (int q = 0; q < kOctBS / 64; q++) s += ws[threadIdx.x][q];
        s_pre[threadIdx.x] = s;
    }
    __syncthreads();
    if (lane == 0)
        for (int k = 0; k < 3; k++) ws[k][w] = tot[k];
    __syncthreads();
    if (threadIdx.x < 3) {
        uint64_t s = 0;
        for (int q = 0; q < kOctBS / 64; q++) s += ws[threadIdx.x][q];
        s_tot[threadIdx.x] = s;
    }
    __syncthreads();
    // this thread's slots, then an exclusive scan of (draws, mesh, records)
    const uint64_t s0 = blockIdx.x * (uint64_t)kOctSlots + threadIdx.x * kOctPer;
    uint32_t v[kOctPer];
    uint64_t mine[3] = {0, 0, 0};
    for (int j = 0; j < kOctPer; j++) {
        v[j] = s0 + j < a.nslots ? a.slots[s0 + j] : 0u;
        if (!v[j]) continue;
        if ((v[j] - 1u) >> 31) mine[1]++;
        else { mine[0]++; mine[2] += oct_items(a, v[j]); }
    }
    uint64_t x[3] = {mine[0], mine[1], mine[2]};
    for (int k = 0; k < 3; k++) {
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint64_t y = __shfl_up(x[k], d, 64);
            if (lane >= (uint32_t)d) x[k] += y;
        }
    }
    __syncthreads();
    if (lane == 63)
        for (int k = 0; k < 3; k++) ws[k][w] = x[k];
    __syncthreads();
    uint64_t off[3];
    for (int k = 0; k < 3; k++) {
        uint64_t wb = 0;
        for (uint32_t q = 0; q < w; q++) wb += ws[k][q];
        off[k] = s_pre[k] + wb + x[k] - mine[k];
    }
    for (int j = 0; j < kOctPer; j++) {
        if (!v[j]) continue;
        const uint32_t e = v[j] - 1u;
        if (e >> 31) {
            if (off[1] < a.mesh_cap) a.mesh[off[1]] = e & 0x0FFFFFFFu;
            off[1]++;
        } else {
            if (off[0] < a.draw_cap) {
                vp_draw d;
                d.level = (e >> 28) & 7u;
                d.node = e & 0x0FFFFFFFu;
                d.first = off[2];
                a.draws[off[0]] = d;
            }
            off[0]++;
            off[2] += oct_items(a, v[j]);
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.counts[0] = s_tot[0];
        a.counts[1] = s_tot[1];
        a.counts[2] = s_tot[2];
        if (s_tot[0] > a.draw_cap || s_tot[1] > a.mesh_cap)
            atomicOr(a.status, (uint32_t)VP_STATUS_OVERFLOW);
    }
}

}  // namespace vp

using namespace vp;

extern "C" int vp_gpu_octree_cull(vp_ctx *ctx, const vp_world_desc *world, const float planes[24],
                                  const float origin[3], const uint64_t *const d_node_offsets[5],
                                  const uint8_t *d_chunk_mesh, vp_draw *d_draws, uint32_t draw_cap,
                                  uint32_t *d_mesh_chunks, uint32_t mesh_cap, uint64_t *d_counts,
                                  void *stream)
{
    if (!ctx || !world || !planes || !origin || !d_node_offsets || !d_counts) return VP_ERR_ARG;
    if ((draw_cap && !d_draws) || (mesh_cap && !d_mesh_chunks)) return VP_ERR_ARG;
    for (int l = 0; l < 5; l++)
        if (!d_node_offsets[l]) return VP_ERR_ARG;
    if (world->root_bitw < 2 || world->root_bitw > 8) return VP_ERR_ARG;
    const uint32_t sb = (uint32_t)world->max_bitw[0] + world->max_bitw[1] + world->max_bitw[2];
    if (world->max_bitw[0] > 16 || world->max_bitw[1] > 16 || world->max_bitw[2] > 16) return VP_ERR_ARG;
    if (sb > 27) return VP_ERR_UNSUPPORTED;     // node indices are packed in 28 bits
    OctArgs a = {};
    for (int i = 0; i < 24; i++) a.planes[i] = planes[i];
    for (int i = 0; i < 3; i++) a.origin[i] = origin[i];
    a.rb = world->root_bitw;
    uint64_t nroots = 1;
    for (int i = 0; i < 3; i++) {
        a.mb[i] = world->max_bitw[i];
        const uint32_t m = 1u << world->max_bitw[i];
        a.nr[i] = (m >> 4) ? (m >> 4) : 1u;
        nroots *= a.nr[i];
    }
    a.lvl_first[0] = 0;
    for (int l = 0; l < 5; l++) {
        a.noff[l] = d_node_offsets[l];
        a.lvl_first[l + 1] = a.lvl_first[l] + vp_node_count(world, (uint32_t)l);
    }
    a.chunk_mesh = d_chunk_mesh;
    a.nslots = nroots * 4096u;
    const uint32_t nblk = (uint32_t)((a.nslots + kOctSlots - 1) / kOctSlots);
    {
        std::lock_guard<std::mutex> g(ctx->mu);
        int rc = ensure_oct(ctx, a.nslots, nblk);
        if (rc) return rc;
        a.slots = ctx->d_oct;
        a.partial = ctx->d_oct_part;
    }
    a.draws = d_draws;
    a.draw_cap = draw_cap;
    a.mesh = d_mesh_chunks;
    a.mesh_cap = mesh_cap;
    a.counts = d_counts;
    a.status = ctx->d_status;
    hipStream_t st = as_stream(stream);
    hipLaunchKernelGGL(k_oct_clear, dim3(1024), dim3(256), 0, st, a.slots, a.nslots);
    const uint64_t nt = a.lvl_first[5];
    hipLaunchKernelGGL(k_oct_mark, dim3((uint32_t)((nt + 255) / 256)), dim3(256), 0, st, a);
    hipLaunchKernelGGL(k_oct_count, dim3(nblk), dim3(kOctBS), 0, st, a);
    hipLaunchKernelGGL(k_oct_write, dim3(nblk), dim3(kOctBS), 0, st, a, nblk);
    return hipGetLastError() == hipSuccess ? VP_OK : VP_ERR_HIP;
}
