// splat.hip -- per-pixel ray/AABB splat resolve and its colour backward, gfx950.
//
// Replaces the splat draw of gfx_svl_draw (src/gfx/vsplat.c:104-193): the
// vertex stage of resources/splat.vert.glsl:34-54 (view translation, clip
// position, point-sprite size) and the fragment stage of
// resources/splat.frag.glsl:60-156 (view ray, slab ray/AABB test, depth
// t0/far, GL_LESS), with GL's point clipping and sprite coverage made
// explicit:
//  * a point survives iff its clip-space vertex is inside the clip volume;
//  * its sprite covers pixel p iff wx - s/2 <= p.x + 0.5 < wx + s/2 (same in y),
//    s = the point size clamped to [1, 2048];
//  * the nearest hit wins, ties to the lower splat index (GL submission
//    order): one 64-bit atomicMin per fragment on (depth bits << 32 | index).
// All float arithmetic is written one IEEE operation at a time, contraction
// off, correctly rounded division and square root, GLSL min/max as selects,
// so the results are those of the stated operation order.
#pragma clang fp contract(off)

#include "vp_internal.h"

namespace vp {

struct SplatCam {
    float T[16], PR[16], IPR[16];
    uint32_t W, H;
    float far_plane;
};

__device__ __forceinline__ void mv4(const float *m, const float v[4], float o[4])
{
#pragma unroll
    for (int r = 0; r < 4; r++) {
        float s = m[0 * 4 + r] * v[0];
        s = s + m[1 * 4 + r] * v[1];
        s = s + m[2 * 4 + r] * v[2];
        s = s + m[3 * 4 + r] * v[3];
        o[r] = s;
    }
}

__device__ __forceinline__ float gmin(float x, float y) { return (y < x) ? y : x; }
__device__ __forceinline__ float gmax(float x, float y) { return (x < y) ? y : x; }
__device__ __forceinline__ float gsign(float x)
{
    return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f);
}

struct VS {
    float vx, vy, vz, half;
    float lox, hix, loy, hiy;
    int x0, x1, y0, y1;
    uint64_t lb;   // (lower bound of the splat's depth t0/far) << 32: no fragment key is below it
};

__device__ __forceinline__ bool vertex_stage(const int16_t *rec, float lodf, const SplatCam &c,
                                             VS &o)
{
    const float half = 0.5f * lodf;
    const float a[4] = {(float)rec[0] + half, (float)rec[1] + half, (float)rec[2] + half, 1.0f};
    float v[4], clip[4];
    mv4(c.T, a, v);
    mv4(c.PR, v, clip);
    const float W = (float)c.W, H = (float)c.H, cw = clip[3];
    if (!(clip[0] >= -cw && clip[0] <= cw && clip[1] >= -cw && clip[1] <= cw &&
          clip[2] >= -cw && clip[2] <= cw))
        return false;
    const float ratio = W / H;
    const float rx = fabsf((clip[0] * ratio) / cw);
    const float ry = fabsf(clip[1] / cw);
    float reduce = gmax(rx, ry);
    reduce = reduce + 0.03f;
    float size = (H * 1.1f) / clip[2];
    size = size * gmax(reduce, 1.0f);
    float ps = size * lodf;
    ps = (ps > 1.0f) ? ps : 1.0f;
    ps = (ps < 2048.0f) ? ps : 2048.0f;
    const float ndx = clip[0] / cw, ndy = clip[1] / cw;
    const float wx = (ndx + 1.0f) * (0.5f * W);
    const float wy = (ndy + 1.0f) * (0.5f * H);
    const float hs = 0.5f * ps;
    o.lox = wx - hs; o.hix = wx + hs;
    o.loy = wy - hs; o.hiy = wy + hs;
    // pixel p is covered iff lo <= p + 0.5 < hi (tested exactly per pixel);
    // every such p lies in [floor(lo), ceil(hi) - 1]
    const int x0 = (int)floorf(o.lox), x1 = (int)ceilf(o.hix) - 1;
    const int y0 = (int)floorf(o.loy), y1 = (int)ceilf(o.hiy) - 1;
    o.x0 = x0 < 0 ? 0 : x0;
    o.y0 = y0 < 0 ? 0 : y0;
    o.x1 = x1 > (int)c.W - 1 ? (int)c.W - 1 : x1;
    o.y1 = y1 > (int)c.H - 1 ? (int)c.H - 1 : y1;
    o.vx = v[0]; o.vy = v[1]; o.vz = v[2];
    o.half = half;
    // Distance from the eye (the origin of the translated frame) to the box
    // v +- half bounds every ray's entry time t0 from below; shrunk by 2^-10,
    // far more than the rounding of the slab test, it stays below the computed
    // t0, and t0/far keeps the order after rounding.
    const float ex = gmax(fabsf(v[0]) - half, 0.0f), ey = gmax(fabsf(v[1]) - half, 0.0f);
    const float ez = gmax(fabsf(v[2]) - half, 0.0f);
    float dd = ex * ex;
    dd = dd + ey * ey;
    dd = dd + ez * ez;
    const float lbd = (sqrtf(dd) * 0.9990234375f) / c.far_plane;
    o.lb = (uint64_t)__float_as_uint(lbd) << 32;
    return true;
}

__device__ __forceinline__ void pixel_dir(const SplatCam &c, int px, int py, float d[3])
{
    const float fx = ((float)px + 0.5f) / (float)c.W;
    const float fy = ((float)py + 0.5f) / (float)c.H;
    const float cl[4] = {fx * 2.0f - 1.0f, fy * 2.0f - 1.0f, -1.0f, 1.0f};
    float r[4];
    mv4(c.IPR, cl, r);
    const float x = r[0] / r[3], y = r[1] / r[3], z = r[2] / r[3];
    float dd = x * x;
    dd = dd + y * y;
    dd = dd + z * z;
    const float len = sqrtf(dd);
    d[0] = x / len; d[1] = y / len; d[2] = z / len;
}

__device__ __forceinline__ void aabb(float vx, float vy, float vz, float half, float ix, float iy,
                                     float iz, float &t0, float &t1)
{
    const float tbx = (vx - half) * ix, ttx = (vx + half) * ix;
    const float tby = (vy - half) * iy, tty = (vy + half) * iy;
    const float tbz = (vz - half) * iz, ttz = (vz + half) * iz;
    const float mnx = gmin(ttx, tbx), mny = gmin(tty, tby), mnz = gmin(ttz, tbz);
    const float mxx = gmax(ttx, tbx), mxy = gmax(tty, tby), mxz = gmax(ttz, tbz);
    t0 = gmax(gmax(mnx, mny), gmax(mnx, mnz));
    t1 = gmin(gmin(mxx, mxy), gmin(mxx, mxz));
}

__global__ void k_rays(SplatCam c, float4 *rays)
{
    const uint32_t px = blockIdx.x * blockDim.x + threadIdx.x, py = blockIdx.y;
    if (px >= c.W) return;
    float d[3];
    pixel_dir(c, (int)px, (int)py, d);
    rays[(size_t)py * c.W + px] = make_float4(1.0f / d[0], 1.0f / d[1], 1.0f / d[2], 0.0f);
}

__global__ void k_clear(uint64_t *frame, uint64_t np)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < np;
         i += (uint64_t)gridDim.x * blockDim.x)
        frame[i] = ~0ull;
}

// The fragment of pixel p whose frame key was read as cur: GL_LESS on
// (depth bits << 32 | id) by one atomicMin.  cur may be older than the
// frame (the loads of a batch go out before any of its tests): a key at or
// above an older value is at or above the current one too, so the tests
// only ever skip fragments that could not win.
__device__ __forceinline__ void fragment_test(const SplatCam &c, const float4 *rays,
                                              uint64_t *frame, float vx, float vy, float vz,
                                              float half, uint32_t id, uint64_t lb, uint32_t p,
                                              uint64_t cur)
{
    if (lb > cur) return;          // every key of this splat is above the stored one
    const float4 inv = rays[p];
    float t0, t1;
    aabb(vx, vy, vz, half, inv.x, inv.y, inv.z, t0, t1);
    if (!(t0 <= t1)) return;
    float d = t0 / c.far_plane;
    if (!(d < 1.0f)) return;
    if (d <= 0.0f) d = 0.0f;
    const uint64_t key = ((uint64_t)__float_as_uint(d) << 32) | id;
    if (key < cur) atomicMin(reinterpret_cast<unsigned long long *>(frame + p), key);
}

// Sprites of up to kSmallArea pixels are rasterised by their own lane,
// larger ones by the whole wave, a few at a time.  Most fragments end at
// the depth lower bound after one frame load, so the loads are issued in
// batches (kLoads per lane) before any test: the lane waits once per batch
// instead of once per pixel.
constexpr int kSmallArea = 64;
constexpr int kLoads = 8;
constexpr int kGroup = 4;                 // wave-shared sprites per round
constexpr int kSteps = kLoads / kGroup;   // pixels of each per lane per round

// Lane src's value (src wave-uniform) through a scalar register: no LDS
// traffic, unlike a general shuffle.
__device__ __forceinline__ int lane_i(int v, int src) { return __builtin_amdgcn_readlane(v, src); }
__device__ __forceinline__ float lane_f(float v, int src)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}

// A wave-shared sprite: the owning lane's vertex stage, broadcast.
struct WaveSprite {
    float vx, vy, vz, half, lox, hix, loy, hiy;
    int x0, y0, w, a, dq, dr;
    uint32_t id;
    uint64_t lb;
};

__device__ __forceinline__ WaveSprite wave_sprite(const VS &vs, int area, uint32_t id, int src)
{
    WaveSprite s;
    s.vx = lane_f(vs.vx, src); s.vy = lane_f(vs.vy, src);
    s.vz = lane_f(vs.vz, src); s.half = lane_f(vs.half, src);
    s.lox = lane_f(vs.lox, src); s.hix = lane_f(vs.hix, src);
    s.loy = lane_f(vs.loy, src); s.hiy = lane_f(vs.hiy, src);
    s.x0 = lane_i(vs.x0, src);
    s.y0 = lane_i(vs.y0, src);
    s.w = lane_i(vs.x1, src) - s.x0 + 1;
    s.a = lane_i(area, src);
    s.id = (uint32_t)lane_i((int)id, src);
    s.lb = (uint64_t)(uint32_t)lane_i((int)(uint32_t)(vs.lb >> 32), src) << 32;
    return s;
}

// Splat i (i >= n: none) of one lane; every lane of the wave calls it.
__device__ __forceinline__ void raster_splat(const int16_t *splats, const uint8_t *lod,
                                             uint32_t lod_log2, uint64_t n, uint64_t i,
                                             const SplatCam &c, const float4 *rays,
                                             uint64_t *frame)
{
    VS vs;
    bool ok = false;
    if (i < n) {
        const float lodf = (float)(1u << (lod ? lod[i] : lod_log2));
        ok = vertex_stage(splats + 4 * i, lodf, c, vs);
    }
    int area = 0;
    if (ok && vs.x0 <= vs.x1 && vs.y0 <= vs.y1) area = (vs.x1 - vs.x0 + 1) * (vs.y1 - vs.y0 + 1);
    const uint32_t id = (uint32_t)i;
    if (area > 0 && area <= kSmallArea) {
        // the box's pixels row by row, kLoads at a time
        const int w = vs.x1 - vs.x0 + 1;
        int col = 0, row = 0;
        for (int k0 = 0; k0 < area; k0 += kLoads) {
            uint32_t pp[kLoads];
            uint64_t cur[kLoads];
            uint32_t live = 0;
#pragma unroll
            for (int u = 0; u < kLoads; u++) {
                const int px = vs.x0 + col, py = vs.y0 + row;
                const float cx = (float)px + 0.5f, cy = (float)py + 0.5f;
                pp[u] = (uint32_t)py * c.W + (uint32_t)px;
                cur[u] = 0;
                if (k0 + u < area && vs.lox <= cx && cx < vs.hix && vs.loy <= cy && cy < vs.hiy) {
                    cur[u] = frame[pp[u]];
                    live |= 1u << u;
                }
                if (++col == w) { col = 0; ++row; }
            }
#pragma unroll
            for (int u = 0; u < kLoads; u++)
                if ((live >> u) & 1u)
                    fragment_test(c, rays, frame, vs.vx, vs.vy, vs.vz, vs.half, id, vs.lb, pp[u],
                                  cur[u]);
        }
    }
    uint64_t big = __ballot(area > kSmallArea);
    const int lane = (int)(threadIdx.x & 63u);
    while (big) {
        // up to kGroup sprites; lane k takes their pixels k, k + 64, ... row
        // by row, the column and row advancing by 64 % w and 64 / w per step.
        // The quotients come from an approximate reciprocal: for a < 2^7 the
        // fraction of (a + 1/2)/w is at least 1/(2w) >= 2^-12 away from an
        // integer, far above the reciprocal's error, so the floor is exact.
        WaveSprite sp[kGroup];
        int row[kGroup], col[kGroup];
        int amax = 0;
#pragma unroll
        for (int g = 0; g < kGroup; g++) {
            sp[g] = WaveSprite{};
            sp[g].w = 1;
            row[g] = col[g] = 0;
            if (big) {
                const int src = __builtin_ctzll(big);
                big &= big - 1ull;
                sp[g] = wave_sprite(vs, area, id, src);
                const float rw = __builtin_amdgcn_rcpf((float)sp[g].w);
                sp[g].dq = (int)(64.5f * rw);
                sp[g].dr = 64 - sp[g].dq * sp[g].w;
                row[g] = (int)(((float)lane + 0.5f) * rw);
                col[g] = lane - row[g] * sp[g].w;
                amax = sp[g].a > amax ? sp[g].a : amax;
            }
        }
        for (int k0 = 0; k0 < amax; k0 += 64 * kSteps) {
            uint32_t pp[kLoads];
            uint64_t cur[kLoads];
            uint32_t live = 0;
#pragma unroll
            for (int g = 0; g < kGroup; g++)
#pragma unroll
                for (int j = 0; j < kSteps; j++) {
                    const int u = g * kSteps + j;
                    const int px = sp[g].x0 + col[g], py = sp[g].y0 + row[g];
                    col[g] += sp[g].dr;
                    row[g] += sp[g].dq;
                    if (col[g] >= sp[g].w) { col[g] -= sp[g].w; row[g] += 1; }
                    const float cx = (float)px + 0.5f, cy = (float)py + 0.5f;
                    pp[u] = (uint32_t)py * c.W + (uint32_t)px;
                    cur[u] = 0;
                    if (k0 + 64 * j + lane < sp[g].a && sp[g].lox <= cx && cx < sp[g].hix &&
                        sp[g].loy <= cy && cy < sp[g].hiy) {
                        cur[u] = frame[pp[u]];
                        live |= 1u << u;
                    }
                }
#pragma unroll
            for (int g = 0; g < kGroup; g++)
#pragma unroll
                for (int j = 0; j < kSteps; j++) {
                    const int u = g * kSteps + j;
                    if ((live >> u) & 1u)
                        fragment_test(c, rays, frame, sp[g].vx, sp[g].vy, sp[g].vz, sp[g].half,
                                      sp[g].id, sp[g].lb, pp[u], cur[u]);
                }
        }
    }
}

// One thread per splat, in list order.
__global__ void __launch_bounds__(256) k_raster(const int16_t *splats, const uint8_t *lod,
                                                uint32_t lod_log2, uint64_t n, SplatCam c,
                                                const float4 *rays, uint64_t *frame)
{
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    raster_splat(splats, lod, lod_log2, n, i, c, rays, frame);
}

// ---- block culling and near-to-far order -------------------------------------
// Block b is splats [256b, 256b+256); its bounds are the box of the splat
// centres and their largest half-size.  Blocks wholly outside the view volume
// are dropped; the rest are bucketed by distance and rasterised nearest bucket
// first, so the early depth rejection in the raster sees the occluders first.
constexpr int kBlock = VP_SPLAT_BLOCK;
constexpr int kBins = 256;
// Rasterised in phases of growing distance: the nearest eighth of the
// visible blocks, up to half, then the rest; a Hi-Z of the frame is built
// between phases.
// counters: [0, kBins) bin counts, [kBins, 2 kBins) bin cursors, the total
constexpr int kTotal = 2 * kBins;
constexpr int kCounters = kTotal + 1;

// One wave per block of kBlock splats (lane l takes splats l, l + 64, ...),
// blocks strided over the grid's waves; a wave-wide min reduction, no LDS.
__global__ void __launch_bounds__(256) k_bounds(const int16_t *splats, const uint8_t *lod,
                                                uint32_t lod_log2, uint64_t n, uint64_t nb,
                                                float *bounds, const uint64_t *d_n)
{
    constexpr int PER = kBlock / 64;
    if (d_n) {          // the list length lives in device memory (n is its bound)
        n = min(*d_n, n);
        nb = (n + kBlock - 1) / kBlock;
    }
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t nwaves = (uint64_t)gridDim.x * (blockDim.x >> 6);
    for (uint64_t b = (uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); b < nb;
         b += nwaves) {
        float v[7];   // min x y z, max half, -max x y z (all reduced with min)
        for (int k = 0; k < 7; k++) v[k] = INFINITY;
        uint2 r[PER];
        uint32_t lv[PER];
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const uint64_t i = b * kBlock + (uint64_t)q * 64 + lane;
            const bool in = i < n;
            r[q] = in ? *reinterpret_cast<const uint2 *>(splats + 4 * i) : make_uint2(0u, 0u);
            lv[q] = (in && lod) ? (uint32_t)lod[i] : lod_log2;
        }
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const uint64_t i = b * kBlock + (uint64_t)q * 64 + lane;
            if (i < n) {
                const float half = 0.5f * (float)(1u << lv[q]);
                const float cx = (float)(int16_t)(r[q].x & 0xFFFFu) + half;
                const float cy = (float)(int16_t)(r[q].x >> 16) + half;
                const float cz = (float)(int16_t)(r[q].y & 0xFFFFu) + half;
                v[0] = fminf(v[0], cx); v[1] = fminf(v[1], cy); v[2] = fminf(v[2], cz);
                v[3] = fminf(v[3], -half);
                v[4] = fminf(v[4], -cx); v[5] = fminf(v[5], -cy); v[6] = fminf(v[6], -cz);
            }
        }
        for (int k = 0; k < 7; k++)
            for (int d = 32; d >= 1; d >>= 1) v[k] = fminf(v[k], __shfl_xor(v[k], d, 64));
        if (lane == 0) {
            float4 *o = reinterpret_cast<float4 *>(bounds + 8 * b);
            o[0] = make_float4(v[0], v[1], v[2], -v[3]);
            o[1] = make_float4(-v[4], -v[5], -v[6], 0.0f);
        }
    }
}

struct CullPlanes {
    float p[6][4];   // inside: p.x*cx + p.y*cy + p.z*cz + p.w >= 0 for a splat centre
    float eye[3];
    float far_plane;
};

__global__ void k_blk_reset(uint32_t *cnt)
{
    for (int i = threadIdx.x; i < kCounters; i += blockDim.x) cnt[i] = 0u;
}

__device__ __forceinline__ uint32_t block_bin(const float *bounds, uint64_t b,
                                              const CullPlanes &cp)
{
    const float *o = bounds + 8 * b;
    const float lo[3] = {o[0], o[1], o[2]}, hi[3] = {o[4], o[5], o[6]};
    if (!(lo[0] <= hi[0])) return 0xFFFFFFFFu;   // empty block
    for (int k = 0; k < 6; k++) {
        // the largest value of the plane function over the box of centres
        float f = cp.p[k][3], mag = fabsf(cp.p[k][3]);
        for (int a = 0; a < 3; a++) {
            const float c = cp.p[k][a];
            f += c * (c > 0.0f ? hi[a] : lo[a]);
            mag += fabsf(c) * fmaxf(fabsf(lo[a]), fabsf(hi[a]));
        }
        // the rounding of the vertex stage's clip test is far below 1e-3 of mag
        if (f < -1e-3f * mag) return 0xFFFFFFFFu;
    }
    float dd = 0.0f;
    for (int a = 0; a < 3; a++) {
        const float e = fmaxf(fmaxf(lo[a] - o[3] - cp.eye[a], cp.eye[a] - hi[a] - o[3]), 0.0f);
        dd += e * e;
    }
    const float t = sqrtf(sqrtf(dd) / cp.far_plane);
    return (uint32_t)fminf(t * (float)kBins, (float)(kBins - 1));
}

// Bins of the blocks, counted per workgroup in LDS and added to the global
// counts once per workgroup and bin.
__global__ void __launch_bounds__(256) k_blk_cull(const float *bounds, uint64_t nb, CullPlanes cp,
                                                  uint32_t *bin, uint32_t *cnt, const uint64_t *d_n)
{
    __shared__ uint32_t hist[kBins];
    if (d_n) nb = min((*d_n + kBlock - 1) / kBlock, nb);
    for (int k = threadIdx.x; k < kBins; k += blockDim.x) hist[k] = 0u;
    __syncthreads();
    for (uint64_t b = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; b < nb;
         b += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t k = block_bin(bounds, b, cp);
        bin[b] = k;
        if (k != 0xFFFFFFFFu) atomicAdd(hist + k, 1u);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < kBins; k += blockDim.x)
        if (hist[k]) atomicAdd(cnt + k, hist[k]);
}

// exclusive scan of the bin counts into the bin cursors; cnt[2*kBins] = total
__global__ void __launch_bounds__(kBins) k_blk_scan(uint32_t *cnt)
{
    __shared__ uint32_t sm[kBins];
    const uint32_t t = threadIdx.x;
    sm[t] = cnt[t];
    __syncthreads();
    for (uint32_t d = 1; d < kBins; d <<= 1) {
        const uint32_t v = t >= d ? sm[t - d] : 0u;
        __syncthreads();
        sm[t] += v;
        __syncthreads();
    }
    cnt[kBins + t] = sm[t] - cnt[t];
    if (t == kBins - 1) cnt[kTotal] = sm[t];
}

// Blocks into their bins' ranges of the order: per 256 blocks, a workgroup
// counts its blocks per bin in LDS, reserves each bin's share with one global
// add, then places every block inside its share.
__global__ void __launch_bounds__(256) k_blk_scatter(const uint32_t *bin, uint64_t nb,
                                                     uint32_t *cnt, uint32_t *order,
                                                     const uint64_t *d_n)
{
    __shared__ uint32_t num[kBins], base[kBins];
    if (d_n) nb = min((*d_n + kBlock - 1) / kBlock, nb);
    for (uint64_t b0 = blockIdx.x * (uint64_t)blockDim.x; b0 < nb;
         b0 += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t b = b0 + threadIdx.x;
        for (int k = threadIdx.x; k < kBins; k += blockDim.x) num[k] = 0u;
        __syncthreads();
        const uint32_t k = b < nb ? bin[b] : 0xFFFFFFFFu;
        if (k != 0xFFFFFFFFu) atomicAdd(num + k, 1u);
        __syncthreads();
        for (int q = threadIdx.x; q < kBins; q += blockDim.x) {
            base[q] = num[q] ? atomicAdd(cnt + kBins + q, num[q]) : 0u;
            num[q] = 0u;
        }
        __syncthreads();
        if (k != 0xFFFFFFFFu) order[base[k] + atomicAdd(num + k, 1u)] = (uint32_t)b;
        __syncthreads();
    }
}

// Workgroup g rasterises the ordered blocks g, g + G, g + 2G, ...: every
// workgroup starts at the near end of the order, and no counter is shared.
__global__ void __launch_bounds__(kBlock) k_raster_blocks(const int16_t *splats,
                                                           const uint8_t *lod, uint32_t lod_log2,
                                                           uint64_t n, SplatCam c,
                                                           const float4 *rays, uint64_t *frame,
                                                           const uint32_t *order,
                                                           const uint32_t *cnt,
                                                           const uint64_t *d_n,
                                                           const uint32_t *tiles_ok)
{
    if (tiles_ok && *tiles_ok) return;    // the screen-tile path resolved this frame
    const uint32_t total = cnt[kTotal];
    if (d_n) n = min(*d_n, n);
    for (uint32_t t = blockIdx.x; t < total; t += gridDim.x) {
        const uint64_t i = (uint64_t)order[t] * kBlock + threadIdx.x;
        raster_splat(splats, lod, lod_log2, n, i, c, rays, frame);
    }
}

// ---- screen-tile resolve ------------------------------------------------------
// The visible blocks' splats are binned into kTile x kTile pixel tiles; each
// tile is then resolved by one workgroup with its rays and frame keys in LDS,
// so a fragment costs LDS reads and one LDS atomicMin instead of global
// gathers and an L2 atomic.  A splat whose pixel box is at most kTileSide
// pixels on a side goes into each tile it touches (at most four entries); a
// larger one -- a near splat, few of them, many pixels each -- goes to the big
// list, which is drawn first against the frame in device memory, spread over
// the whole grid.  The tiles then start from those keys, so the near
// occluders reject the small splats behind them from the first entry on.
// The keys are the same (depth bits << 32 | id) and min is order-free, so the
// frame is identical to the global path's.  If a frame needs more entries
// than the buffers hold, the scan leaves the flag clear and the global path
// (k_raster_blocks) draws the whole frame instead.
constexpr int kTileLog = 5, kTile = 1 << kTileLog, kTilePx = kTile * kTile;
constexpr int kTileSide = 8;                  // larger pixel boxes go to the big list
static_assert(kTilePx % kBlock == 0, "k_tile_raster: whole pixels per thread");
constexpr uint32_t kMaxTiles = 8192;          // LDS bins of the binning kernels: 32 KiB
constexpr uint32_t kItem = 8192;              // tile entries per raster work item
constexpr uint32_t kSingle = 1u << 31;        // item flag: the tile's only item
// tile counters (u32): [0] raster items, [1] entries, [2] big splats counted,
// [3] flag: the tile path draws this frame, [4] big-list cursor
constexpr int kTcItems = 0, kTcPairs = 1, kTcBig = 2, kTcOk = 3, kTcBigCur = 4, kTcWords = 8;

struct TileBufs {
    uint32_t tx, ty, nt;      // tiles across, down, total
    uint32_t *hist;           // [G][nt] per-workgroup entry counts
    uint32_t *cnt;            // [nt] entries per tile
    uint32_t *cur;            // [nt] placement cursors
    uint32_t *tc;             // kTcWords counters
    uint4    *items;          // (tile | kSingle, begin, end, 0)
    uint32_t  item_cap;
    uint32_t *pairs;          // splat index per entry, tile-major
    uint32_t  pair_cap;
    uint32_t *big;            // splat indices of the big list
    uint32_t  big_cap;
};

// Vertex stage of splat i (i < n), and the tiles its pixel box touches:
// 0 none, 1 tile entries (t0 = first tile, span bits: +1 tile in x, +1 in y),
// 2 big.
__device__ __forceinline__ int splat_tiles(const int16_t *splats, const uint8_t *lod,
                                           uint32_t lod_log2, uint64_t i, const SplatCam &c,
                                           uint32_t tx, VS &vs, uint32_t &t0, uint32_t &span)
{
    const float lodf = (float)(1u << (lod ? lod[i] : lod_log2));
    if (!vertex_stage(splats + 4 * i, lodf, c, vs)) return 0;
    if (vs.x0 > vs.x1 || vs.y0 > vs.y1) return 0;
    if (vs.x1 - vs.x0 >= kTileSide || vs.y1 - vs.y0 >= kTileSide) return 2;
    // a box of at most kTileSide <= kTile pixels touches at most 2 x 2 tiles
    const uint32_t a = (uint32_t)vs.x0 >> kTileLog, b = (uint32_t)vs.x1 >> kTileLog;
    const uint32_t d = (uint32_t)vs.y0 >> kTileLog, e = (uint32_t)vs.y1 >> kTileLog;
    t0 = d * tx + a;
    span = (b - a) | ((e - d) << 1);
    return 1;
}

__global__ void k_tile_reset(uint32_t *cnt, uint32_t nt, uint32_t *tc)
{
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < nt; t += gridDim.x * blockDim.x)
        cnt[t] = 0u;
    if (blockIdx.x == 0 && threadIdx.x < kTcWords) tc[threadIdx.x] = 0u;
}

// Workgroup g takes the ordered visible blocks g, g + G, ... (the same
// assignment as k_tile_scatter) and counts its entries per tile in LDS; the
// counts go to its row of the histogram and are added to the tile totals.
__global__ void __launch_bounds__(kBlock) k_tile_count(const int16_t *splats, const uint8_t *lod,
                                                       uint32_t lod_log2, uint64_t n, SplatCam c,
                                                       const uint32_t *order, const uint32_t *cnt,
                                                       const uint64_t *d_n, TileBufs tb)
{
    extern __shared__ uint32_t hist[];
    for (uint32_t t = threadIdx.x; t < tb.nt; t += blockDim.x) hist[t] = 0u;
    __syncthreads();
    const uint32_t total = cnt[kTotal];
    if (d_n) n = min(*d_n, n);
    uint32_t nbig = 0;
    for (uint32_t b = blockIdx.x; b < total; b += gridDim.x) {
        const uint64_t i = (uint64_t)order[b] * kBlock + threadIdx.x;
        if (i >= n) continue;
        VS vs;
        uint32_t t0 = 0, span = 0;
        const int k = splat_tiles(splats, lod, lod_log2, i, c, tb.tx, vs, t0, span);
        if (k == 2) nbig++;
        if (k != 1) continue;
        atomicAdd(hist + t0, 1u);
        if (span & 1u) atomicAdd(hist + t0 + 1, 1u);
        if (span & 2u) atomicAdd(hist + t0 + tb.tx, 1u);
        if (span == 3u) atomicAdd(hist + t0 + tb.tx + 1, 1u);
    }
    __syncthreads();
    uint32_t *row = tb.hist + (size_t)blockIdx.x * tb.nt;
    for (uint32_t t = threadIdx.x; t < tb.nt; t += blockDim.x) {
        const uint32_t v = hist[t];
        row[t] = v;
        if (v) atomicAdd(tb.cnt + t, v);
    }
    for (int d = 32; d >= 1; d >>= 1) nbig += (uint32_t)__shfl_xor((int)nbig, d, 64);
    if ((threadIdx.x & 63u) == 0 && nbig) atomicAdd(tb.tc + kTcBig, nbig);
}

// One workgroup: tile offsets (exclusive scan of the counts), placement
// cursors, the raster items (runs of at most kItem entries of one tile) and
// the flag that hands the frame to the tile path when everything fits.
constexpr int kScanThreads = 1024;
__device__ __forceinline__ uint64_t shfl_up64(uint64_t v, int d)
{
    const uint32_t lo = (uint32_t)__shfl_up((int)(uint32_t)v, d, 64);
    const uint32_t hi = (uint32_t)__shfl_up((int)(uint32_t)(v >> 32), d, 64);
    return ((uint64_t)hi << 32) | lo;
}
__global__ void __launch_bounds__(kScanThreads) k_tile_scan(TileBufs tb)
{
    __shared__ uint64_t wsum[2][kScanThreads / 64];
    __shared__ uint32_t s_ok;
    const uint32_t per = (tb.nt + kScanThreads - 1) / kScanThreads;
    const uint32_t t0 = min(threadIdx.x * per, tb.nt), t1 = min(t0 + per, tb.nt);
    // entry totals in 64 bits: four entries per splat can pass 2^32
    uint64_t sc = 0, si = 0;
    for (uint32_t t = t0; t < t1; t++) {
        const uint32_t v = tb.cnt[t];
        sc += v;
        si += (v + kItem - 1) / kItem;
    }
    // block-wide exclusive scans of (entries, items)
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    uint64_t ic = sc, ii = si;
    for (int d = 1; d < 64; d <<= 1) {
        const uint64_t a = shfl_up64(ic, d), b = shfl_up64(ii, d);
        if ((int)lane >= d) { ic += a; ii += b; }
    }
    if (lane == 63) { wsum[0][w] = ic; wsum[1][w] = ii; }
    __syncthreads();
    uint64_t bc = 0, bi = 0, tc = 0, ti = 0;
    for (uint32_t k = 0; k < kScanThreads / 64; k++) {
        if (k < w) { bc += wsum[0][k]; bi += wsum[1][k]; }
        tc += wsum[0][k];
        ti += wsum[1][k];
    }
    if (threadIdx.x == 0) {
        const uint32_t big = tb.tc[kTcBig];
        const bool ok = tc <= tb.pair_cap && ti <= tb.item_cap && big <= tb.big_cap;
        s_ok = ok ? 1u : 0u;
        tb.tc[kTcPairs] = (uint32_t)min(tc, (uint64_t)0xFFFFFFFFu);
        tb.tc[kTcItems] = ok ? (uint32_t)ti : 0u;
        tb.tc[kTcOk] = s_ok;
    }
    __syncthreads();
    if (!s_ok) return;
    // every total now fits in 32 bits
    uint32_t off = (uint32_t)(bc + ic - sc), io = (uint32_t)(bi + ii - si);
    for (uint32_t t = t0; t < t1; t++) {
        const uint32_t v = tb.cnt[t];
        tb.cur[t] = off;
        const uint32_t k = (v + kItem - 1) / kItem;
        for (uint32_t q = 0; q < k; q++) {
            const uint32_t b = off + q * kItem, e = min(b + kItem, off + v);
            tb.items[io + q] = make_uint4(t | (k == 1 ? kSingle : 0u), b, e, 0u);
        }
        off += v;
        io += k;
    }
}

// The same blocks as k_tile_count: reserve the workgroup's range of every
// tile it touches, then place each entry (LDS cursors); big splats are
// appended to the big list.
__global__ void __launch_bounds__(kBlock) k_tile_scatter(const int16_t *splats, const uint8_t *lod,
                                                         uint32_t lod_log2, uint64_t n, SplatCam c,
                                                         const uint32_t *order, const uint32_t *cnt,
                                                         const uint64_t *d_n, TileBufs tb)
{
    extern __shared__ uint32_t lcur[];
    if (!tb.tc[kTcOk]) return;
    const uint32_t *row = tb.hist + (size_t)blockIdx.x * tb.nt;
    for (uint32_t t = threadIdx.x; t < tb.nt; t += blockDim.x) {
        const uint32_t v = row[t];
        lcur[t] = v ? atomicAdd(tb.cur + t, v) : 0u;
    }
    __syncthreads();
    const uint32_t total = cnt[kTotal];
    if (d_n) n = min(*d_n, n);
    const uint32_t lane = threadIdx.x & 63u;
    for (uint32_t b = blockIdx.x; b < total; b += gridDim.x) {
        const uint64_t i = (uint64_t)order[b] * kBlock + threadIdx.x;
        VS vs;
        uint32_t t0 = 0, span = 0;
        const int k = i < n ? splat_tiles(splats, lod, lod_log2, i, c, tb.tx, vs, t0, span) : 0;
        if (k == 1) {
            uint32_t q = atomicAdd(lcur + t0, 1u);
            if (q < tb.pair_cap) tb.pairs[q] = (uint32_t)i;
            if (span & 1u) {
                q = atomicAdd(lcur + t0 + 1, 1u);
                if (q < tb.pair_cap) tb.pairs[q] = (uint32_t)i;
            }
            if (span & 2u) {
                q = atomicAdd(lcur + t0 + tb.tx, 1u);
                if (q < tb.pair_cap) tb.pairs[q] = (uint32_t)i;
            }
            if (span == 3u) {
                q = atomicAdd(lcur + t0 + tb.tx + 1, 1u);
                if (q < tb.pair_cap) tb.pairs[q] = (uint32_t)i;
            }
        }
        const uint64_t m = __ballot(k == 2);
        if (m) {
            uint32_t base = 0;
            if (lane == (uint32_t)__builtin_ctzll(m))
                base = atomicAdd(tb.tc + kTcBigCur, (uint32_t)__popcll(m));
            base = (uint32_t)lane_i((int)base, __builtin_ctzll(m));
            if (k == 2) {
                const uint32_t q = base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
                if (q < tb.big_cap) tb.big[q] = (uint32_t)i;
            }
        }
    }
}

// One tile's fragment in LDS: lp indexes the tile's pixel.
__device__ __forceinline__ void fragment_lds(const float4 *lray, uint64_t *lkey, float far_plane,
                                             float vx, float vy, float vz, float half, uint32_t id,
                                             uint64_t lb, int lp)
{
    const uint64_t cur = lkey[lp];
    if (lb > cur) return;
    const float4 inv = lray[lp];
    float t0, t1;
    aabb(vx, vy, vz, half, inv.x, inv.y, inv.z, t0, t1);
    if (!(t0 <= t1)) return;
    float d = t0 / far_plane;
    if (!(d < 1.0f)) return;
    if (d <= 0.0f) d = 0.0f;
    const uint64_t key = ((uint64_t)__float_as_uint(d) << 32) | id;
    if (key < cur) atomicMin(reinterpret_cast<unsigned long long *>(lkey + lp), key);
}

// Entry i (i >= n: none) of one lane, clipped to the tile at (px0, py0);
// every lane of the wave calls it.  As raster_splat: boxes of up to
// kSmallArea pixels by their own lane, larger ones by the whole wave.
__device__ __forceinline__ void tile_splat(const int16_t *splats, const uint8_t *lod,
                                           uint32_t lod_log2, uint64_t n, uint64_t i,
                                           const SplatCam &c, int px0, int py0,
                                           const float4 *lray, uint64_t *lkey)
{
    VS vs;
    bool ok = false;
    if (i < n) {
        const float lodf = (float)(1u << (lod ? lod[i] : lod_log2));
        ok = vertex_stage(splats + 4 * i, lodf, c, vs);
    }
    int x0 = 0, x1 = -1, y0 = 0, y1 = -1;
    if (ok) {
        x0 = vs.x0 > px0 ? vs.x0 : px0;
        x1 = vs.x1 < px0 + kTile - 1 ? vs.x1 : px0 + kTile - 1;
        y0 = vs.y0 > py0 ? vs.y0 : py0;
        y1 = vs.y1 < py0 + kTile - 1 ? vs.y1 : py0 + kTile - 1;
    }
    int area = 0;
    if (x0 <= x1 && y0 <= y1) area = (x1 - x0 + 1) * (y1 - y0 + 1);
    const uint32_t id = (uint32_t)i;
    if (area > 0 && area <= kSmallArea) {
        for (int py = y0; py <= y1; py++) {
            const float cy = (float)py + 0.5f;
            if (!(vs.loy <= cy && cy < vs.hiy)) continue;
            for (int px = x0; px <= x1; px++) {
                const float cx = (float)px + 0.5f;
                if (!(vs.lox <= cx && cx < vs.hix)) continue;
                fragment_lds(lray, lkey, c.far_plane, vs.vx, vs.vy, vs.vz, vs.half, id, vs.lb,
                             ((py - py0) << kTileLog) + (px - px0));
            }
        }
    }
    uint64_t big = __ballot(area > kSmallArea);
    const int lane = (int)(threadIdx.x & 63u);
    while (big) {
        const int src = __builtin_ctzll(big);
        big &= big - 1ull;
        const float vx = lane_f(vs.vx, src), vy = lane_f(vs.vy, src);
        const float vz = lane_f(vs.vz, src), half = lane_f(vs.half, src);
        const float lox = lane_f(vs.lox, src), hix = lane_f(vs.hix, src);
        const float loy = lane_f(vs.loy, src), hiy = lane_f(vs.hiy, src);
        const int bx0 = lane_i(x0, src), bx1 = lane_i(x1, src);
        const int by0 = lane_i(y0, src);
        const int a = lane_i(area, src);
        const uint32_t sid = (uint32_t)lane_i((int)id, src);
        const uint64_t lb = (uint64_t)(uint32_t)lane_i((int)(uint32_t)(vs.lb >> 32), src) << 32;
        // lane k takes pixels k, k + 64, ... of the clipped box, row by row
        // (w <= kTile, so 64 / w and the row of k come exactly from the
        // approximate reciprocal, as in raster_splat)
        const int w = bx1 - bx0 + 1;
        const float rw = __builtin_amdgcn_rcpf((float)w);
        const int dq = (int)(64.5f * rw), dr = 64 - dq * w;
        int row = (int)(((float)lane + 0.5f) * rw), col = lane - row * w;
        for (int k = lane; k < a; k += 64) {
            const int px = bx0 + col, py = by0 + row;
            col += dr;
            row += dq;
            if (col >= w) { col -= w; row += 1; }
            const float cx = (float)px + 0.5f, cy = (float)py + 0.5f;
            if (!(lox <= cx && cx < hix && loy <= cy && cy < hiy)) continue;
            fragment_lds(lray, lkey, c.far_plane, vx, vy, vz, half, sid, lb,
                         ((py - py0) << kTileLog) + (px - px0));
        }
    }
}

// Raster items, strided over the grid: a tile's rays in LDS and its keys as
// the big list left them, its entries in kBlock-wide steps, then its keys out
// -- stored when the item is the tile's only one, else merged with a global
// atomicMin where the item lowered a key.
__global__ void __launch_bounds__(kBlock) k_tile_raster(const int16_t *splats, const uint8_t *lod,
                                                        uint32_t lod_log2, uint64_t n, SplatCam c,
                                                        const float4 *rays, uint64_t *frame,
                                                        const uint64_t *d_n, TileBufs tb)
{
    __shared__ float4 lray[kTilePx];
    __shared__ uint64_t lkey[kTilePx];
    if (!tb.tc[kTcOk]) return;
    const uint32_t nitems = tb.tc[kTcItems];
    if (d_n) n = min(*d_n, n);
    for (uint32_t it = blockIdx.x; it < nitems; it += gridDim.x) {
        const uint4 item = tb.items[it];
        const uint32_t t = item.x & ~kSingle;
        const int px0 = (int)(t % tb.tx) << kTileLog, py0 = (int)(t / tb.tx) << kTileLog;
        uint64_t start[kTilePx / kBlock];
#pragma unroll
        for (int k = 0; k < kTilePx / kBlock; k++) {
            const int q = (int)threadIdx.x + k * kBlock;
            const int px = px0 + (q & (kTile - 1)), py = py0 + (q >> kTileLog);
            start[k] = ~0ull;
            if (px < (int)c.W && py < (int)c.H) {
                lray[q] = rays[(size_t)py * c.W + px];
                start[k] = frame[(size_t)py * c.W + px];
            }
            lkey[q] = start[k];
        }
        __syncthreads();
        for (uint32_t j0 = item.y; j0 < item.z; j0 += blockDim.x) {
            const uint32_t j = j0 + threadIdx.x;
            const uint64_t i = j < item.z ? (uint64_t)tb.pairs[j] : n;
            tile_splat(splats, lod, lod_log2, n, i, c, px0, py0, lray, lkey);
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kTilePx / kBlock; k++) {
            const int q = (int)threadIdx.x + k * kBlock;
            const int px = px0 + (q & (kTile - 1)), py = py0 + (q >> kTileLog);
            if (px >= (int)c.W || py >= (int)c.H) continue;
            const uint64_t key = lkey[q];
            if (key >= start[k]) continue;
            uint64_t *f = frame + (size_t)py * c.W + px;
            if (item.x & kSingle)
                *f = key;
            else
                atomicMin(reinterpret_cast<unsigned long long *>(f), key);
        }
        __syncthreads();
    }
}

// The big list against the global frame, before the tiles.
__global__ void __launch_bounds__(kBlock) k_tile_big(const int16_t *splats, const uint8_t *lod,
                                                     uint32_t lod_log2, uint64_t n, SplatCam c,
                                                     const float4 *rays, uint64_t *frame,
                                                     const uint64_t *d_n, TileBufs tb)
{
    if (!tb.tc[kTcOk]) return;
    const uint32_t nbig = min(tb.tc[kTcBigCur], tb.big_cap);
    if (d_n) n = min(*d_n, n);
    for (uint32_t j0 = blockIdx.x * blockDim.x; j0 < nbig; j0 += gridDim.x * blockDim.x) {
        const uint32_t j = j0 + threadIdx.x;
        const uint64_t i = j < nbig ? (uint64_t)tb.big[j] : n;
        raster_splat(splats, lod, lod_log2, n, i, c, rays, frame);
    }
}

__global__ void k_resolve(const int16_t *splats, const uint8_t *lod, uint32_t lod_log2,
                          uint64_t n, SplatCam c, const float *voxel_rgb, const uint64_t *frame,
                          const float4 *rays, float *depth, uint32_t *idbuf, float *rgb,
                          float *weight)
{
    const uint32_t px = blockIdx.x * blockDim.x + threadIdx.x, py = blockIdx.y;
    if (px >= c.W) return;
    const size_t p = (size_t)py * c.W + px;
    const uint64_t key = frame[p];
    if (key == ~0ull) {
        depth[p] = 1.0f;
        idbuf[p] = 0xFFFFFFFFu;
        if (rgb) { rgb[3 * p] = 0.5f; rgb[3 * p + 1] = 0.5f; rgb[3 * p + 2] = 0.7f; }
        if (weight) weight[p] = 0.0f;
        return;
    }
    const uint32_t i = (uint32_t)key;
    const float d = __uint_as_float((uint32_t)(key >> 32));
    depth[p] = d;
    idbuf[p] = i;
    if (!rgb && !weight) return;
    // shading, splat.vert.glsl:24-32 and splat.frag.glsl:93-154
    const float lodf = (float)(1u << (lod ? lod[i] : lod_log2));
    const float half = 0.5f * lodf;
    const int16_t *rec = splats + 4 * (size_t)i;
    const float a[4] = {(float)rec[0] + half, (float)rec[1] + half, (float)rec[2] + half, 1.0f};
    float v[4];
    mv4(c.T, a, v);
    float dir[3];
    pixel_dir(c, (int)px, (int)py, dir);
    const float4 inv = rays[p];
    float t0, t1;
    aabb(v[0], v[1], v[2], half, inv.x, inv.y, inv.z, t0, t1);
    float hit[3], ha[3];
    for (int k = 0; k < 3; k++) {
        hit[k] = v[k] - dir[k] * t0;
        ha[k] = fabsf(hit[k]);
    }
    const float md = gmax(gmax(ha[0], ha[1]), ha[2]);
    float nrm[3];
    for (int k = 0; k < 3; k++) nrm[k] = (ha[k] == md ? 1.0f : 0.0f) * gsign(hit[k]);
    const float lx = -0.70710677f, ly = 0.70710677f, lz = 0.0f;
    float nd = (-nrm[0]) * lx;
    nd = nd + (-nrm[1]) * ly;
    nd = nd + (-nrm[2]) * lz;
    const int attr = rec[3];
    const float sh = (float)((attr >> 6) & 1);
    const float s1 = gmax(0.7f, 1.0f - sh);
    const float s2 = gmax(gmax(0.7f, sh), gsign(nd));
    float base[3];
    if (voxel_rgb) {
        for (int k = 0; k < 3; k++) base[k] = voxel_rgb[3 * (size_t)i + k];
    } else {
        base[0] = (float)((attr >> 4) & 3) / 3.0f;
        base[1] = (float)((attr >> 2) & 3) / 3.0f;
        base[2] = (float)(attr & 3) / 3.0f;
    }
    const float fs = 256.0f / c.far_plane;
    const float fog = sqrtf(gmax(d - fs, 0.0001f));
    const float keep = 1.0f - fog;
    const float fogc[3] = {0.5f, 0.5f, 0.7f};
    if (rgb)
        for (int k = 0; k < 3; k++) {
            const float vcol = base[k] * s1;
            const float col = vcol * s2;
            rgb[3 * p + k] = col * keep + fogc[k] * fog;
        }
    if (weight) weight[p] = (s1 * s2) * keep;
}

__global__ void k_backward(uint64_t np, uint64_t n, const uint32_t *idbuf, const float *weight,
                           const float *dL, float *grad)
{
    for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p < np;
         p += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t i = idbuf[p];
        if (i == 0xFFFFFFFFu || i >= n) continue;
        const float w = weight[p];
        if (w == 0.0f) continue;
        for (int k = 0; k < 3; k++) atomicAdd(grad + 3 * (size_t)i + k, w * dL[3 * p + k]);
    }
}

static SplatCam to_cam(const vp_camera *cam)
{
    SplatCam c;
    for (int i = 0; i < 16; i++) {
        c.T[i] = cam->view_translation[i];
        c.PR[i] = cam->proj_view_rot[i];
        c.IPR[i] = cam->inv_proj_view_rot[i];
    }
    c.W = cam->width;
    c.H = cam->height;
    c.far_plane = cam->far_plane;
    return c;
}

}  // namespace vp

using namespace vp;

static CullPlanes cull_planes(const vp_camera *cam)
{
    // M = proj_view_rot * view_translation (column-major), in double: clip = M (c, 1)
    double M[4][4];
    for (int r = 0; r < 4; r++)
        for (int k = 0; k < 4; k++) {
            double acc = 0.0;
            for (int j = 0; j < 4; j++)
                acc += (double)cam->proj_view_rot[j * 4 + r] * (double)cam->view_translation[k * 4 + j];
            M[r][k] = acc;
        }
    CullPlanes cp;
    // -w <= x, x <= w, -w <= y, y <= w, -w <= z, z <= w
    for (int a = 0; a < 3; a++)
        for (int sgn = 0; sgn < 2; sgn++)
            for (int k = 0; k < 4; k++)
                cp.p[2 * a + sgn][k] = (float)(M[3][k] + (sgn ? -M[a][k] : M[a][k]));
    for (int a = 0; a < 3; a++) cp.eye[a] = -cam->view_translation[12 + a];
    cp.far_plane = cam->far_plane;
    return cp;
}

// The splat frame.  d_n (nullable): the list length in device memory, n its
// bound; with d_n the block bounds are computed into d_bounds here first (the
// list is assembled on the device each frame).
static int splat_frame(vp_ctx *ctx, const int16_t *d_splats, const uint8_t *d_lod,
                       uint32_t lod_log2, uint64_t n, const float *d_bounds,
                       const vp_camera *cam, const float *d_voxel_rgb, const vp_frame *out,
                       void *stream, const uint64_t *d_n = nullptr)
{
    if (!ctx || !cam || !out || !out->d_depth || !out->d_id) return VP_ERR_ARG;
    if (n && !d_splats) return VP_ERR_ARG;
    if (cam->width == 0 || cam->height == 0 || cam->width > 16384 || cam->height > 16384)
        return VP_ERR_ARG;
    if (lod_log2 > 8) return VP_ERR_ARG;
    if (n > 0xFFFFFFFFull) return VP_ERR_UNSUPPORTED;   // ids are 32-bit
    const uint64_t np = (uint64_t)cam->width * cam->height;
    const uint64_t nb = (n + kBlock - 1) / kBlock;
    hipStream_t st = as_stream(stream);
    uint64_t *frame;
    float4 *rays;
    uint32_t *blk = nullptr;
    int ncu;
    // screen tiles (block-culled frames of at most kMaxTiles tiles)
    TileBufs tb = {};
    tb.tx = (cam->width + kTile - 1) / kTile;
    tb.ty = (cam->height + kTile - 1) / kTile;
    tb.nt = tb.tx * tb.ty;
    uint32_t tgrid = 0;
    bool tiles = false;
    {
        std::lock_guard<std::mutex> g(ctx->mu);
        tiles = ctx->splat_tiles && d_bounds && nb && tb.nt <= kMaxTiles;
        int rc = ensure_frame(ctx, np);
        if (rc) return rc;
        if (d_bounds && nb) {
            rc = ensure_blocks(ctx, nb);
            if (rc) return rc;
            blk = ctx->d_blk;
        }
        frame = ctx->d_frame;
        rays = ctx->d_rays;
        ncu = ctx->num_cus;
        if (tiles) {
            // entries: at most four per splat; beyond 2^26 in one frame the
            // global path draws it
            tgrid = (uint32_t)ncu * 4u;
            const uint64_t lim = ctx->splat_tile_entries ? ctx->splat_tile_entries : (1ull << 26);
            tb.pair_cap = (uint32_t)(4 * n < lim ? 4 * n : lim);
            tb.big_cap = (uint32_t)(n < (1ull << 22) ? n : (1ull << 22));
            tb.item_cap = tb.nt + tb.pair_cap / kItem + 1;
            const uint64_t words = (uint64_t)tgrid * tb.nt + 2ull * tb.nt + kTcWords +
                                   4ull * tb.item_cap + tb.pair_cap + tb.big_cap;
            rc = ensure_tiles(ctx, words);
            if (rc) return rc;
            uint32_t *w = ctx->d_tiles;
            tb.items = reinterpret_cast<uint4 *>(w);            // 16-byte aligned
            w += 4ull * tb.item_cap;
            tb.hist = w;
            w += (uint64_t)tgrid * tb.nt;
            tb.cnt = w;
            w += tb.nt;
            tb.cur = w;
            w += tb.nt;
            tb.tc = w;
            w += kTcWords;
            tb.pairs = w;
            w += tb.pair_cap;
            tb.big = w;
        }
    }
    const SplatCam c = to_cam(cam);
    const dim3 pg((cam->width + 255) / 256, cam->height);
    hipLaunchKernelGGL(k_rays, pg, dim3(256), 0, st, c, rays);
    hipLaunchKernelGGL(k_clear, dim3(1024), dim3(256), 0, st, frame, np);
    if (n && blk && d_n) {
        const uint64_t wgs = (nb + 3) / 4, cap = (uint64_t)ncu * 32u;
        hipLaunchKernelGGL(k_bounds, dim3((uint32_t)(wgs < cap ? wgs : cap)), dim3(256), 0, st,
                           d_splats, d_lod, lod_log2, n, nb, const_cast<float *>(d_bounds), d_n);
    }
    if (n && blk) {
        // blk: [bin per block | order | counters]
        uint32_t *bin = blk, *order = blk + nb, *cnt = blk + 2 * nb;
        const uint32_t g = (uint32_t)((nb + 255) / 256 < 4096 ? (nb + 255) / 256 : 4096);
        hipLaunchKernelGGL(k_blk_reset, dim3(1), dim3(256), 0, st, cnt);
        hipLaunchKernelGGL(k_blk_cull, dim3(g), dim3(256), 0, st, d_bounds, nb, cull_planes(cam),
                           bin, cnt, d_n);
        hipLaunchKernelGGL(k_blk_scan, dim3(1), dim3(kBins), 0, st, cnt);
        hipLaunchKernelGGL(k_blk_scatter, dim3(g), dim3(256), 0, st, bin, nb, cnt, order, d_n);
        if (tiles) {
            const size_t lds = (size_t)tb.nt * sizeof(uint32_t);
            hipLaunchKernelGGL(k_tile_reset, dim3((tb.nt + 255) / 256), dim3(256), 0, st, tb.cnt,
                               tb.nt, tb.tc);
            hipLaunchKernelGGL(k_tile_count, dim3(tgrid), dim3(kBlock), lds, st, d_splats, d_lod,
                               lod_log2, n, c, order, cnt, d_n, tb);
            hipLaunchKernelGGL(k_tile_scan, dim3(1), dim3(kScanThreads), 0, st, tb);
            hipLaunchKernelGGL(k_tile_scatter, dim3(tgrid), dim3(kBlock), lds, st, d_splats, d_lod,
                               lod_log2, n, c, order, cnt, d_n, tb);
            hipLaunchKernelGGL(k_tile_big, dim3((uint32_t)ncu * 4u), dim3(kBlock), 0, st, d_splats,
                               d_lod, lod_log2, n, c, rays, frame, d_n, tb);
            hipLaunchKernelGGL(k_tile_raster, dim3((uint32_t)ncu * 6u), dim3(kBlock), 0, st,
                               d_splats, d_lod, lod_log2, n, c, rays, frame, d_n, tb);
        }
        // 4 workgroups per CU: what the raster's registers leave resident
        hipLaunchKernelGGL(k_raster_blocks, dim3((uint32_t)ncu * 4u), dim3(kBlock), 0, st,
                           d_splats, d_lod, lod_log2, n, c, rays, frame, order, cnt, d_n,
                           tiles ? tb.tc + kTcOk : nullptr);
    } else if (n) {
        hipLaunchKernelGGL(k_raster, dim3((uint32_t)nb), dim3(256), 0, st, d_splats, d_lod,
                           lod_log2, n, c, rays, frame);
    }
    hipLaunchKernelGGL(k_resolve, pg, dim3(256), 0, st, d_splats, d_lod, lod_log2, n, c,
                       d_voxel_rgb, frame, rays, out->d_depth, out->d_id, out->d_rgb,
                       out->d_weight);
    VP_HIP_TRY(hipGetLastError());
    return VP_OK;
}

extern "C" int vp_gpu_splat(vp_ctx *ctx, const int16_t *d_splats, const uint8_t *d_lod,
                            uint32_t lod_log2, uint64_t n, const vp_camera *cam,
                            const float *d_voxel_rgb, const vp_frame *out, void *stream)
{
    return splat_frame(ctx, d_splats, d_lod, lod_log2, n, nullptr, cam, d_voxel_rgb, out, stream);
}

extern "C" int vp_gpu_splat_bounds(vp_ctx *ctx, const int16_t *d_splats, const uint8_t *d_lod,
                                   uint32_t lod_log2, uint64_t n, float *d_bounds, void *stream)
{
    if (!ctx || (n && (!d_splats || !d_bounds))) return VP_ERR_ARG;
    if (lod_log2 > 8) return VP_ERR_ARG;
    if (n > 0xFFFFFFFFull) return VP_ERR_UNSUPPORTED;
    if (n == 0) return VP_OK;
    const uint64_t nb = (n + kBlock - 1) / kBlock;
    // four blocks per workgroup (a wave each), at most 32 workgroups per CU
    const uint64_t wgs = (nb + 3) / 4, cap = (uint64_t)ctx->num_cus * 32u;
    hipLaunchKernelGGL(k_bounds, dim3((uint32_t)(wgs < cap ? wgs : cap)), dim3(256), 0,
                       as_stream(stream), d_splats, d_lod, lod_log2, n, nb, d_bounds, nullptr);
    VP_HIP_TRY(hipGetLastError());
    return VP_OK;
}

extern "C" int vp_gpu_splat_blocks(vp_ctx *ctx, const int16_t *d_splats, const uint8_t *d_lod,
                                   uint32_t lod_log2, uint64_t n, const float *d_bounds,
                                   const vp_camera *cam, const float *d_voxel_rgb,
                                   const vp_frame *out, void *stream)
{
    return splat_frame(ctx, d_splats, d_lod, lod_log2, n, d_bounds, cam, d_voxel_rgb, out, stream);
}

extern "C" int vp_ctx_set_splat_tiles(vp_ctx *ctx, int enable, uint64_t max_entries)
{
    if (!ctx || max_entries > 0xFFFFFFFFull) return VP_ERR_ARG;
    std::lock_guard<std::mutex> g(ctx->mu);
    ctx->splat_tiles = enable ? 1 : 0;
    ctx->splat_tile_entries = max_entries;
    return VP_OK;
}

extern "C" int vp_gpu_splat_backward(vp_ctx *ctx, uint64_t n, const vp_camera *cam,
                                     const uint32_t *d_id, const float *d_weight,
                                     const float *d_dL_drgb, float *d_grad, void *stream)
{
    if (!ctx || !cam || !d_id || !d_weight || !d_dL_drgb || !d_grad) return VP_ERR_ARG;
    const uint64_t np = (uint64_t)cam->width * cam->height;
    if (np == 0) return VP_OK;
    uint64_t blocks = (np + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_backward, dim3((uint32_t)blocks), dim3(256), 0, as_stream(stream), np, n,
                       d_id, d_weight, d_dL_drgb, d_grad);
    VP_HIP_TRY(hipGetLastError());
    return VP_OK;
}

extern "C" int vp_gpu_splat_counted(vp_ctx *ctx, const int16_t *d_splats, const uint8_t *d_lod,
                                    uint32_t lod_log2, uint64_t capacity, const uint64_t *d_count,
                                    float *d_bounds, const vp_camera *cam,
                                    const float *d_voxel_rgb, const vp_frame *out, void *stream)
{
    if (!d_count || (capacity && !d_bounds)) return VP_ERR_ARG;
    return splat_frame(ctx, d_splats, d_lod, lod_log2, capacity, d_bounds, cam, d_voxel_rgb, out,
                       stream, d_count);
}
