// gen.hip -- device-side workload builders: world generator, chunk
// classification, shadow-map build, sphere edits.  None of these is on the
// timed path; they build its inputs in HBM (a 16384x512x16384 world does not
// fit through the host).
#include "vp_internal.h"
#include "../../include/vp_workload.h"

#include <algorithm>
#include <vector>

namespace vp {

__device__ __forceinline__ uint64_t splitmix64(uint64_t x)
{
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ int64_t lattice(uint64_t seed, int o, uint32_t ix, uint32_t iz)
{
    const uint64_t key = seed ^ ((uint64_t)(o + 1) * 0xD1B54A32D192ED03ull) ^
                         (((uint64_t)ix << 32) | (uint64_t)iz);
    return (int64_t)(splitmix64(key) >> 48);
}

__device__ __forceinline__ int64_t smooth_q16(int64_t t)
{
    return (((t * t) >> 16) * (3 * 65536 - 2 * t)) >> 16;
}

__device__ __forceinline__ int64_t lerp_q16(int64_t a, int64_t b, int64_t t)
{
    return a + (((b - a) * t) >> 16);
}

// 4 octaves of Q16 value noise, weights 8,4,2,1, mapped onto [hmin, hmax]
__device__ int32_t height_at(const vp_gen_params &p, uint32_t x, uint32_t z)
{
    int64_t acc = 0;
    for (int o = 0; o < 4; o++) {
        uint32_t per = p.period >> o;
        if (per == 0) per = 1;
        const int64_t w = 8 >> o;
        const uint32_t ix = x / per, iz = z / per;
        const int64_t tx = ((int64_t)(x % per) << 16) / per;
        const int64_t tz = ((int64_t)(z % per) << 16) / per;
        const int64_t sx = smooth_q16(tx), sz = smooth_q16(tz);
        const int64_t a = lerp_q16(lattice(p.seed, o, ix, iz), lattice(p.seed, o, ix + 1, iz), sx);
        const int64_t b =
            lerp_q16(lattice(p.seed, o, ix, iz + 1), lattice(p.seed, o, ix + 1, iz + 1), sx);
        acc += w * lerp_q16(a, b, sz);
    }
    const uint64_t span = (uint64_t)(p.hmax - p.hmin + 1);
    return p.hmin + (int32_t)(((uint64_t)acc * span) / (15ull * 65536ull));
}

template <int B>
__global__ void __launch_bounds__(256) k_generate(vp_gen_params p, uint8_t mb0, uint8_t mb1,
                                                  uint8_t *voxels, uint32_t chunk0,
                                                  uint32_t count, int16_t *uniform)
{
    constexpr int R = 1 << B;
    __shared__ int32_t h[R * R];
    __shared__ uint32_t s_mixed;
    for (uint32_t k = blockIdx.x; k < count; k += gridDim.x) {
        const uint32_t chunk = chunk0 + k;
        const uint32_t cx = chunk & ((1u << mb0) - 1u);
        const uint32_t cy = (chunk >> mb0) & ((1u << mb1) - 1u);
        const uint32_t cz = chunk >> (mb0 + mb1);
        for (int c = threadIdx.x; c < R * R; c += 256)
            h[c] = height_at(p, cx * R + (c % R), cz * R + (c / R));   // h[z*R + x]
        if (threadIdx.x == 0) s_mixed = 0u;
        __syncthreads();
        const uint64_t sc = splitmix64(p.seed ^ (uint64_t)chunk);
        auto vox = [&](uint32_t i) -> uint32_t {
            const uint32_t x = i & (R - 1), y = (i >> B) & (R - 1), z = i >> (2 * B);
            const int32_t wy = (int32_t)(cy * R + y);
            if (wy >= h[z * R + x]) return 0u;
            const int32_t band = wy >> 3;
            if (splitmix64(sc ^ (uint64_t)i) % 20u == 0) return 0u;
            return 1u + (uint32_t)(band < 62 ? band : 62);
        };
        const uint32_t v0 = vox(0);
        uint8_t *dst = voxels + ((size_t)chunk << (3 * B));
        bool mixed = false;
        if constexpr (R * R * R >= 16) {
            for (uint32_t s = threadIdx.x; s < (uint32_t)(R * R * R / 16); s += 256) {
                uint32_t w[4];
                for (int q = 0; q < 4; q++) {
                    uint32_t acc = 0;
                    for (int j = 0; j < 4; j++) {
                        const uint32_t v = vox(s * 16 + q * 4 + j);
                        mixed |= (v != v0);
                        acc |= v << (8 * j);
                    }
                    w[q] = acc;
                }
                *reinterpret_cast<uint4 *>(dst + (size_t)s * 16) = make_uint4(w[0], w[1], w[2], w[3]);
            }
        }
        if (mixed) s_mixed = 1u;
        __syncthreads();
        if (uniform && threadIdx.x == 0) uniform[chunk] = s_mixed ? (int16_t)-1 : (int16_t)v0;
        __syncthreads();
    }
}

template <int B>
__global__ void __launch_bounds__(256) k_classify(const uint8_t *voxels, uint32_t chunk0,
                                                  uint32_t count, int16_t *uniform)
{
    constexpr int R = 1 << B;
    __shared__ uint32_t s_mixed;
    for (uint32_t k = blockIdx.x; k < count; k += gridDim.x) {
        const uint32_t chunk = chunk0 + k;
        const uint8_t *src = voxels + ((size_t)chunk << (3 * B));
        if (threadIdx.x == 0) s_mixed = 0u;
        __syncthreads();
        const uint32_t v0 = src[0];
        const uint32_t rep = v0 * 0x01010101u;
        bool mixed = false;
        for (uint32_t s = threadIdx.x; s < (uint32_t)(R * R * R / 16); s += 256) {
            const uint4 v = *reinterpret_cast<const uint4 *>(src + (size_t)s * 16);
            mixed |= (v.x != rep) | (v.y != rep) | (v.z != rep) | (v.w != rep);
        }
        if (mixed) s_mixed = 1u;
        __syncthreads();
        if (threadIdx.x == 0) uniform[chunk] = s_mixed ? (int16_t)-1 : (int16_t)v0;
        __syncthreads();
    }
}

__global__ void k_fill16(uint16_t *p, uint64_t n)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

// Shadow map rows.  A solid voxel at height wy is placed (shadow.h:77-89)
// unless map[wx+wy] or map[wx+wy+1] is already above wy.  Within one chunk
// band (cy fixed) the serial order visits columns x ascending and y ascending
// inside a column, and an earlier band only ever stores heights below the
// current band, so:
//  * bands do not block each other, and a later band overwrites an earlier
//    band's entries (larger heights);
//  * inside a band, voxel (x, y) is placed iff diagonals x+y and x+y+1 are
//    free of placements from earlier columns -- placements in the same
//    column are on lower diagonals.
// So per band a 64-bit window of taken diagonals is carried across columns:
// placed = solid & ~T & ~(T >> 1); T = (T | placed) >> 1.
template <int B>
__global__ void __launch_bounds__(256) k_shadow_rows(const uint8_t *voxels, uint8_t mb0,
                                                     uint8_t mb1, uint8_t mb2, uint16_t *map,
                                                     uint32_t sx)
{
    constexpr int R = 1 << B, SEG = 1024;
    extern __shared__ uint16_t row[];      // sx entries
    __shared__ uint64_t colbits[SEG];
    const uint32_t mx = 1u << mb0, my = 1u << mb1;
    const uint32_t X = mx * R, Z = (1u << mb2) * R;
    for (uint32_t wz = blockIdx.x; wz < Z; wz += gridDim.x) {
        for (uint32_t i = threadIdx.x; i < sx; i += 256) row[i] = 0;
        __syncthreads();
        const uint32_t cz = wz >> B, zl = wz & (R - 1);
        for (uint32_t cy = 0; cy < my; cy++) {
            uint64_t T = 0;
            for (uint32_t xs = 0; xs < X; xs += SEG) {
                const uint32_t segn = min((uint32_t)SEG, X - xs);
                for (uint32_t c = threadIdx.x; c < segn; c += 256) {
                    const uint32_t x = xs + c, cx = x >> B, xl = x & (R - 1);
                    const uint32_t chunk = ((cz << mb1 | cy) << mb0) | cx;
                    const uint8_t *src = voxels + ((size_t)chunk << (3 * B)) + (zl * R) * R + xl;
                    uint64_t m = 0;
                    for (int y = 0; y < R; y++) m |= (uint64_t)(src[y * R] != 0) << y;
                    colbits[c] = m;
                }
                __syncthreads();
                if (threadIdx.x == 0) {
                    for (uint32_t c = 0; c < segn; c++) {
                        const uint64_t placed = colbits[c] & ~T & ~(T >> 1);
                        colbits[c] = placed;
                        T = (T | placed) >> 1;
                    }
                }
                __syncthreads();
                for (uint32_t c = threadIdx.x; c < segn; c += 256) {
                    uint64_t m = colbits[c];
                    while (m) {
                        const int y = __builtin_ctzll(m);
                        m &= m - 1ull;
                        const uint32_t wy = cy * R + y;
                        row[xs + c + wy] = (uint16_t)wy;   // distinct diagonals per band
                    }
                }
                __syncthreads();
            }
        }
        uint16_t *dst = map + (size_t)sx * wz;
        for (uint32_t i = threadIdx.x; i < sx; i += 256) dst[i] = row[i];
        __syncthreads();
    }
}

// The same map for 64^3 chunks with every band's serial chain in flight at
// once: one workgroup per map row, one wave per band.  Lane y loads voxel row
// y of the band's 64x64 slab in a chunk (16-byte pieces, the wave's 4 KB in
// one go); 64 ballots turn the rows into the slab's 64 column masks (bit y of
// column x: voxel (x, y) solid), which are wave-uniform, so the placement
// recurrence runs in scalar registers.  A placed voxel raises its diagonal's
// entry of an LDS copy of the row with atomicMax: a band places at most once
// per diagonal (the diagonal's bit stays in T until it drops out below y = 0)
// and later bands carry larger heights, so the maximum is the value the
// band-after-band overwrite leaves.
constexpr int kShadowBS = 512;
__global__ void __launch_bounds__(kShadowBS) k_shadow_rows64(const uint8_t *voxels, uint8_t mb0,
                                                            uint8_t mb1, uint8_t mb2,
                                                            uint16_t *map, uint32_t sx)
{
    constexpr int R = 64;
    extern __shared__ uint32_t row32[];    // sx entries
    const uint32_t mx = 1u << mb0, my = 1u << mb1, Z = (1u << mb2) * R;
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    constexpr uint32_t NW = kShadowBS / 64;
    for (uint32_t wz = blockIdx.x; wz < Z; wz += gridDim.x) {
        for (uint32_t i = threadIdx.x; i < sx; i += kShadowBS) row32[i] = 0u;
        __syncthreads();
        const uint32_t cz = wz >> 6, zl = wz & 63u;
        for (uint32_t cy = wv; cy < my; cy += NW) {
            uint64_t T = 0;
            const uint32_t wy = cy * R + lane;
            for (uint32_t cx = 0; cx < mx; cx++) {
                const uint32_t chunk = ((cz << mb1 | cy) << mb0) | cx;
                const uint4 *src = reinterpret_cast<const uint4 *>(
                    voxels + ((size_t)chunk << 18) + ((size_t)zl * R + lane) * R);
                const uint4 v[4] = {src[0], src[1], src[2], src[3]};
                // bit x of this lane's row: voxel (x, lane) != 0
                uint64_t rowbits = 0;
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const uint32_t d[4] = {v[q].x, v[q].y, v[q].z, v[q].w};
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        // high bit of each byte: set iff the byte is not 0
                        const uint32_t nz = (((d[k] & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | d[k]) & 0x80808080u;
                        // bytes 0..3 -> bits 0..3 of this dword's nibble
                        const uint32_t n4 = ((nz >> 7) & 1u) | ((nz >> 14) & 2u) | ((nz >> 21) & 4u) |
                                            ((nz >> 28) & 8u);
                        rowbits |= (uint64_t)n4 << (16 * q + 4 * k);
                    }
                }
                const uint32_t base = cx * R + wy;     // diagonal of (x = 0, y = lane)
#pragma unroll 8
                for (int x = 0; x < R; x++) {
                    const uint64_t col = __ballot((rowbits >> x) & 1ull);
                    const uint64_t placed = col & ~T & ~(T >> 1);
                    T = (T | placed) >> 1;
                    if ((placed >> lane) & 1ull) atomicMax(&row32[base + (uint32_t)x], wy);
                }
            }
        }
        __syncthreads();
        uint16_t *dst = map + (size_t)sx * wz;
        for (uint32_t i = threadIdx.x; i < sx; i += kShadowBS) dst[i] = (uint16_t)row32[i];
        __syncthreads();
    }
}

struct PairList {
    uint32_t chunk;
    uint32_t first, count;   // range in the sphere-index array
};

template <int B>
__global__ void __launch_bounds__(256) k_edit(uint8_t *voxels, uint8_t mb0, uint8_t mb1,
                                              const PairList *lists, uint32_t nlists,
                                              const uint32_t *sidx, const vp_sphere *sph)
{
    constexpr int R = 1 << B;
    for (uint32_t k = blockIdx.x; k < nlists; k += gridDim.x) {
        const PairList pl = lists[k];
        const uint32_t chunk = pl.chunk;
        const int32_t cx = (int32_t)(chunk & ((1u << mb0) - 1u));
        const int32_t cy = (int32_t)((chunk >> mb0) & ((1u << mb1) - 1u));
        const int32_t cz = (int32_t)(chunk >> (mb0 + mb1));
        uint8_t *v = voxels + ((size_t)chunk << (3 * B));
        for (uint32_t j = 0; j < pl.count; j++) {
            const vp_sphere s = sph[sidx[pl.first + j]];
            const int32_t r = (int32_t)s.radius;
            const int64_t r2 = (int64_t)r * r;
            // the sphere's box clipped to the chunk
            int32_t lo[3], hi[3];
            const int32_t c[3] = {s.cx, s.cy, s.cz}, base[3] = {cx * R, cy * R, cz * R};
            for (int a = 0; a < 3; a++) {
                lo[a] = max(c[a] - r - 1, base[a]) - base[a];
                hi[a] = min(c[a] + r + 1, base[a] + R) - base[a];   // exclusive
            }
            const int32_t nx = hi[0] - lo[0], ny = hi[1] - lo[1], nz = hi[2] - lo[2];
            if (nx > 0 && ny > 0 && nz > 0) {
                const int32_t tot = nx * ny * nz;
                for (int32_t t = threadIdx.x; t < tot; t += 256) {
                    const int32_t xl = lo[0] + t % nx, yl = lo[1] + (t / nx) % ny,
                                  zl = lo[2] + t / (nx * ny);
                    const int64_t dx = base[0] + xl - c[0], dy = base[1] + yl - c[1],
                                  dz = base[2] + zl - c[2];
                    if (dx * dx + dy * dy + dz * dz >= r2) continue;
                    if (base[1] + yl < 2) continue;                 // edit.c:151
                    v[(zl * R + yl) * R + xl] = (uint8_t)s.voxel;
                }
            }
            __syncthreads();   // spheres apply in list order
        }
    }
}

__global__ void k_mark(const PairList *lists, uint32_t n, uint8_t *dirty, int16_t *uniform)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (dirty) dirty[lists[i].chunk] = 1;
        if (uniform) uniform[lists[i].chunk] = -1;
    }
}

// shadow_place_update for add-sphere voxels (src/chunkset/shadow.h:77-89,
// called from chunkset_edit_sphere, src/chunkset/edit.c:210-236), on the
// device.  A placement at (wx, wy, wz) reads entries i = wx + wy + sx * wz
// and i + 1 and may write entry i, all in map row wz, so rows are
// independent: one wave per row, the row staged in LDS.  In a row the
// reference's order is: spheres in list order; chunks cx, then cy (cz is the
// row's); then ux; then uy.  One "sweep" (fixed sphere, chunk, ux) places
// uy = y0.. in order at i = ux + uy, and each placement reads i and i + 1,
// which no earlier placement of the same sweep writes (they write below i):
// so the lanes of a wave take the uy of a sweep together, all reads of the
// sweep before its writes, and the sweeps follow each other in order.
__global__ void __launch_bounds__(64) k_shadow_spheres(uint16_t *map, uint32_t sx, int b,
                                                       int32_t mx0, int32_t mx1,
                                                       const uint32_t *row_off,
                                                       const uint32_t *row_sph,
                                                       const vp_sphere *sph)
{
    extern __shared__ uint16_t row[];
    const uint32_t wz = blockIdx.x, lane = threadIdx.x;
    const uint32_t s0 = row_off[wz], s1 = row_off[wz + 1];
    if (s0 == s1) return;
    const int32_t R = 1 << b;
    uint32_t *row32 = reinterpret_cast<uint32_t *>(row);
    const uint32_t *src = reinterpret_cast<const uint32_t *>(map + (size_t)sx * wz);
    for (uint32_t e = lane; e < sx / 2; e += 64) row32[e] = src[e];
    __syncthreads();
    for (uint32_t k = s0; k < s1; k++) {
        const vp_sphere sp = sph[row_sph[k]];
        const int32_t c[3] = {sp.cx, sp.cy, sp.cz}, r = (int32_t)sp.radius;
        const int64_t dz = (int64_t)wz - c[2];
        const int64_t rz2 = (int64_t)r * r - dz * dz;      // > 0 for a listed sphere
        const int32_t lo0 = c[0] - r - 1, hi0 = c[0] + r + 1, lo1 = c[1] - r - 1, hi1 = c[1] + r + 1;
        for (int32_t cx = lo0 >> b; cx <= (hi0 >> b); cx++) {
            if (cx < 0 || cx >= mx0) continue;
            const int32_t x0 = max(lo0, cx * R), x1 = min(hi0, cx * R + R);
            for (int32_t cy = lo1 >> b; cy <= (hi1 >> b); cy++) {
                if (cy < 0 || cy >= mx1) continue;
                const int32_t y0 = max(lo1, cy * R), y1 = min(hi1, cy * R + R);
                for (int32_t ux = x0; ux < x1; ux++) {
                    const int64_t dx = (int64_t)ux - c[0];
                    if (dx * dx >= rz2) continue;
                    const int32_t uy = y0 + (int32_t)lane;
                    const int64_t dy = (int64_t)uy - c[1];
                    if (uy < y1 && dx * dx + dy * dy < rz2) {
                        const uint32_t i = (uint32_t)(ux + uy), wy = (uint32_t)uy;
                        const bool lit = row[i] <= wy && row[i + 1] <= wy;
                        __builtin_amdgcn_wave_barrier();    // every read of the sweep first
                        if (lit) row[i] = (uint16_t)wy;
                    }
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                }
            }
        }
    }
    __syncthreads();
    uint32_t *dst = reinterpret_cast<uint32_t *>(map + (size_t)sx * wz);
    for (uint32_t e = lane; e < sx / 2; e += 64) dst[e] = row32[e];
}

template <typename F>
static hipError_t by_root(int b, F f)
{
    switch (b) {
    case 2: return f(std::integral_constant<int, 2>{});
    case 3: return f(std::integral_constant<int, 3>{});
    case 4: return f(std::integral_constant<int, 4>{});
    case 5: return f(std::integral_constant<int, 5>{});
    default: return f(std::integral_constant<int, 6>{});
    }
}

static int check_world_w(const vp_world_desc *w)
{
    if (!w) return VP_ERR_ARG;
    if (w->root_bitw < 2 || w->root_bitw > 6) return VP_ERR_UNSUPPORTED;
    if ((int)w->max_bitw[0] + w->max_bitw[1] + w->max_bitw[2] > 32) return VP_ERR_ARG;
    return VP_OK;
}

}  // namespace vp

using namespace vp;

extern "C" int vp_gpu_generate(vp_ctx *ctx, const vp_world_desc *world, const vp_gen_params *p,
                               uint8_t *d_voxels, uint32_t chunk0, uint32_t count,
                               int16_t *d_uniform, void *stream)
{
    if (!ctx || !p || !d_voxels || p->period == 0 || p->hmax < p->hmin) return VP_ERR_ARG;
    int rc = check_world_w(world);
    if (rc) return rc;
    if ((uint64_t)chunk0 + count > (1ull << (world->max_bitw[0] + world->max_bitw[1] + world->max_bitw[2])))
        return VP_ERR_ARG;
    if (count == 0) return VP_OK;
    const int grid = (int)std::min<uint32_t>(count, (uint32_t)ctx->num_cus * 8u);
    const vp_gen_params pp = *p;
    hipStream_t st = as_stream(stream);
    VP_HIP_TRY(by_root(world->root_bitw, [&](auto bc) {
        constexpr int B = decltype(bc)::value;
        hipLaunchKernelGGL(k_generate<B>, dim3(grid), dim3(256), 0, st, pp, world->max_bitw[0],
                           world->max_bitw[1], d_voxels, chunk0, count, d_uniform);
        return hipGetLastError();
    }));
    return VP_OK;
}

extern "C" int vp_gpu_classify(vp_ctx *ctx, const vp_world_desc *world, const uint8_t *d_voxels,
                               uint32_t chunk0, uint32_t count, int16_t *d_uniform, void *stream)
{
    if (!ctx || !d_voxels || !d_uniform) return VP_ERR_ARG;
    int rc = check_world_w(world);
    if (rc) return rc;
    if ((uint64_t)chunk0 + count > (1ull << (world->max_bitw[0] + world->max_bitw[1] + world->max_bitw[2])))
        return VP_ERR_ARG;
    if (count == 0) return VP_OK;
    const int grid = (int)std::min<uint32_t>(count, (uint32_t)ctx->num_cus * 8u);
    hipStream_t st = as_stream(stream);
    VP_HIP_TRY(by_root(world->root_bitw, [&](auto bc) {
        constexpr int B = decltype(bc)::value;
        hipLaunchKernelGGL(k_classify<B>, dim3(grid), dim3(256), 0, st, d_voxels, chunk0, count,
                           d_uniform);
        return hipGetLastError();
    }));
    return VP_OK;
}

extern "C" int vp_gpu_shadow_build(vp_ctx *ctx, const vp_world_desc *world,
                                   const uint8_t *d_voxels, uint16_t *d_map, void *stream)
{
    if (!ctx || !d_voxels || !d_map) return VP_ERR_ARG;
    int rc = check_world_w(world);
    if (rc) return rc;
    const uint32_t b = world->root_bitw;
    const uint64_t X = (uint64_t)(1u << world->max_bitw[0]) << b;
    const uint64_t Y = (uint64_t)(1u << world->max_bitw[1]) << b;
    const uint64_t Z = (uint64_t)(1u << world->max_bitw[2]) << b;
    const uint64_t sx = X + Y;
    if (sx * 2 > 60 * 1024) return VP_ERR_UNSUPPORTED;   // one map row must fit in LDS
    hipStream_t st = as_stream(stream);
    const uint64_t len = vp_shadow_padded_len((uint32_t)sx, (uint32_t)Z);
    hipLaunchKernelGGL(k_fill16, dim3(2048), dim3(256), 0, st, d_map, len);
    VP_HIP_TRY(hipGetLastError());
    if (b == 6 && sx * 4 <= 120 * 1024) {
        // 64^3 chunks: every band's chain at once (a u32 row in LDS)
        const int grid = (int)std::min<uint64_t>(Z, (uint64_t)ctx->num_cus * 2u);
        hipLaunchKernelGGL(k_shadow_rows64, dim3(grid), dim3(kShadowBS), sx * sizeof(uint32_t), st,
                           d_voxels, world->max_bitw[0], world->max_bitw[1], world->max_bitw[2],
                           d_map, (uint32_t)sx);
        VP_HIP_TRY(hipGetLastError());
        return VP_OK;
    }
    const int grid = (int)std::min<uint64_t>(Z, (uint64_t)ctx->num_cus * 4u);
    const size_t lds = sx * sizeof(uint16_t);
    VP_HIP_TRY(by_root((int)b, [&](auto bc) {
        constexpr int B = decltype(bc)::value;
        hipLaunchKernelGGL(k_shadow_rows<B>, dim3(grid), dim3(256), lds, st, d_voxels,
                           world->max_bitw[0], world->max_bitw[1], world->max_bitw[2], d_map,
                           (uint32_t)sx);
        return hipGetLastError();
    }));
    return VP_OK;
}

extern "C" int vp_gpu_edit_spheres(vp_ctx *ctx, const vp_world_desc *world, uint8_t *d_voxels,
                                   const vp_sphere *spheres, uint32_t n, uint8_t *d_dirty,
                                   int16_t *d_uniform, void *stream)
{
    if (!ctx || !d_voxels || (n && !spheres)) return VP_ERR_ARG;
    int rc = check_world_w(world);
    if (rc) return rc;
    if (n == 0) return VP_OK;
    const int b = world->root_bitw;
    const int32_t mx[3] = {1 << world->max_bitw[0], 1 << world->max_bitw[1],
                           1 << world->max_bitw[2]};
    // (chunk, sphere) pairs in chunk order, spheres in list order inside a chunk
    std::vector<std::pair<uint32_t, uint32_t>> pairs;
    for (uint32_t i = 0; i < n; i++) {
        const vp_sphere &s = spheres[i];
        const int32_t c[3] = {s.cx, s.cy, s.cz};
        int32_t lo[3], hi[3];
        for (int a = 0; a < 3; a++) {
            lo[a] = std::max((c[a] - (int32_t)s.radius - 1) >> b, 0);   // edit.c:98-101
            hi[a] = std::min((c[a] + (int32_t)s.radius + 1) >> b, mx[a] - 1);
        }
        for (int32_t x = lo[0]; x <= hi[0]; x++)
            for (int32_t y = lo[1]; y <= hi[1]; y++)
                for (int32_t z = lo[2]; z <= hi[2]; z++) {
                    const uint32_t ch = (((uint32_t)z << world->max_bitw[1] | (uint32_t)y)
                                         << world->max_bitw[0]) | (uint32_t)x;
                    pairs.emplace_back(ch, i);
                }
    }
    std::stable_sort(pairs.begin(), pairs.end(),
                     [](const auto &a, const auto &c) { return a.first < c.first; });
    std::vector<PairList> lists;
    std::vector<uint32_t> sidx(pairs.size());
    for (size_t i = 0; i < pairs.size(); i++) {
        sidx[i] = pairs[i].second;
        if (lists.empty() || lists.back().chunk != pairs[i].first)
            lists.push_back({pairs[i].first, (uint32_t)i, 0u});
        lists.back().count++;
    }
    if (lists.empty()) return VP_OK;
    hipStream_t st = as_stream(stream);
    PairList *d_lists = nullptr;
    uint32_t *d_sidx = nullptr;
    vp_sphere *d_sph = nullptr;
    if (hipMalloc(&d_lists, lists.size() * sizeof(PairList)) != hipSuccess ||
        hipMalloc(&d_sidx, sidx.size() * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&d_sph, n * sizeof(vp_sphere)) != hipSuccess) {
        (void)hipFree(d_lists);
        (void)hipFree(d_sidx);
        (void)hipFree(d_sph);
        return VP_ERR_NOMEM;
    }
    hipError_t e = hipMemcpyAsync(d_lists, lists.data(), lists.size() * sizeof(PairList),
                                  hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_sidx, sidx.data(), sidx.size() * sizeof(uint32_t),
                           hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_sph, spheres, n * sizeof(vp_sphere), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
        const int grid = (int)std::min<size_t>(lists.size(), (size_t)ctx->num_cus * 4u);
        e = by_root(b, [&](auto bc) {
            constexpr int BB = decltype(bc)::value;
            hipLaunchKernelGGL(k_edit<BB>, dim3(grid), dim3(256), 0, st, d_voxels,
                               world->max_bitw[0], world->max_bitw[1], d_lists,
                               (uint32_t)lists.size(), d_sidx, d_sph);
            return hipGetLastError();
        });
    }
    if (e == hipSuccess && (d_dirty || d_uniform))
        hipLaunchKernelGGL(k_mark, dim3(256), dim3(256), 0, st, d_lists, (uint32_t)lists.size(),
                           d_dirty, d_uniform);
    if (e == hipSuccess) e = hipGetLastError();
    // the temporary lists are freed below; wait for the kernels that read them
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(d_lists);
    (void)hipFree(d_sidx);
    (void)hipFree(d_sph);
    return e == hipSuccess ? VP_OK : VP_ERR_HIP;
}

extern "C" int vp_gpu_shadow_spheres(vp_ctx *ctx, const vp_world_desc *world, uint16_t *d_map,
                                     const vp_sphere *spheres, uint32_t n, void *stream)
{
    if (!ctx || !d_map || (n && !spheres)) return VP_ERR_ARG;
    int rc = check_world_w(world);
    if (rc) return rc;
    const int b = world->root_bitw;
    const uint32_t X = (1u << world->max_bitw[0]) << b, Y = (1u << world->max_bitw[1]) << b;
    const uint32_t Z = (1u << world->max_bitw[2]) << b;
    const uint32_t sx = X + Y;
    // the row sits in LDS: up to 64 KiB
    if (sx > 32768u) return VP_ERR_UNSUPPORTED;
    // per-row sphere lists, in sphere order: the add-spheres whose voxels reach the row
    std::vector<uint32_t> cnt(Z + 1, 0u);
    for (uint32_t i = 0; i < n; i++) {
        const vp_sphere &s = spheres[i];
        if (!s.voxel || s.radius == 0) continue;
        const int64_t z0 = std::max<int64_t>((int64_t)s.cz - s.radius + 1, 0);
        const int64_t z1 = std::min<int64_t>((int64_t)s.cz + s.radius - 1, (int64_t)Z - 1);
        for (int64_t z = z0; z <= z1; z++) cnt[(size_t)z + 1]++;
    }
    for (uint32_t z = 0; z < Z; z++) cnt[z + 1] += cnt[z];
    if (cnt[Z] == 0) return VP_OK;
    std::vector<uint32_t> lst(cnt[Z]);
    {
        std::vector<uint32_t> fill(cnt.begin(), cnt.end() - 1);
        for (uint32_t i = 0; i < n; i++) {
            const vp_sphere &s = spheres[i];
            if (!s.voxel || s.radius == 0) continue;
            const int64_t z0 = std::max<int64_t>((int64_t)s.cz - s.radius + 1, 0);
            const int64_t z1 = std::min<int64_t>((int64_t)s.cz + s.radius - 1, (int64_t)Z - 1);
            for (int64_t z = z0; z <= z1; z++) lst[fill[(size_t)z]++] = i;
        }
    }
    hipStream_t st = as_stream(stream);
    uint32_t *d_off = nullptr, *d_lst = nullptr;
    vp_sphere *d_sph = nullptr;
    if (hipMalloc(&d_off, cnt.size() * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&d_lst, lst.size() * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&d_sph, n * sizeof(vp_sphere)) != hipSuccess) {
        (void)hipFree(d_off);
        (void)hipFree(d_lst);
        (void)hipFree(d_sph);
        return VP_ERR_NOMEM;
    }
    hipError_t e = hipMemcpyAsync(d_off, cnt.data(), cnt.size() * sizeof(uint32_t),
                                  hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_lst, lst.data(), lst.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                           st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_sph, spheres, n * sizeof(vp_sphere), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
        const size_t lds = (size_t)sx * sizeof(uint16_t);
        if (lds > 65536) e = hipErrorInvalidValue;
        else {
            hipLaunchKernelGGL(k_shadow_spheres, dim3(Z), dim3(64), lds, st, d_map, sx, b,
                               (int32_t)(1u << world->max_bitw[0]),
                               (int32_t)(1u << world->max_bitw[1]), d_off, d_lst, d_sph);
            e = hipGetLastError();
        }
    }
    // the temporary lists are freed below; wait for the kernel that reads them
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(d_off);
    (void)hipFree(d_lst);
    (void)hipFree(d_sph);
    return e == hipSuccess ? VP_OK : VP_ERR_HIP;
}
