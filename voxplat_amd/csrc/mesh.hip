// mesh.hip -- near-field quad mesh with ambient occlusion for gfx950.
//
// Replaces chunk_make_mesh (src/chunkset/mesher.c:184-357) with its ambient
// samples (sample_ao / sample_ao_border, mesher.c:72-171) and the shadow
// samples it takes (shadow_sample / shadow_sample_normal,
// src/chunkset/shadow.h:52-75), batched over a chunk list.
//
// One workgroup per chunk (persistent, ticket order = list order):
//  1. the chunk's occupancy and a one-voxel shell around it, (R+2)^3 bits,
//     go to LDS as one u64 row per (z, y) in [-1, R]^2 (bits x = 0..R-1)
//     plus one byte per row for each of x = -1 and x = R.  Rows of uniform
//     chunks are synthesised without a read; rows beyond the world are air.
//  2. face bits of a row are XORs of rows: +x = row ^ (row >> 1 | x=R bit),
//     +y / +z = row ^ the next row; their popcounts give the chunk's face
//     count, which is published and turned into an output offset by a
//     decoupled look-back before any face is written.
//  3. slice by slice (several slices at once for R < 64), each thread takes
//     an 8-voxel segment of a row, a workgroup scan places the segments'
//     faces, and the faces go to an LDS list as 14-bit (cell, axis) entries
//     in (z, y, x, axis) order -- the reference's emission order.
//  4. each thread then builds one face of the list: ambient terms from the
//     LDS bits, the solid voxel's colour, shadow samples from the map, and
//     writes its 32-byte vertex quad and 24-byte index row; consecutive
//     lanes write consecutive faces.
#include "vp_internal.h"

namespace vp {

constexpr uint64_t kMFlagAgg = 1ull << 62;
constexpr uint64_t kMFlagPre = 2ull << 62;
constexpr uint64_t kMValMask = (1ull << 62) - 1;
constexpr int kMeshBS = 512;

struct MeshArgs {
    const uint8_t  *voxels;
    const int16_t  *uniform;
    const uint16_t *map;
    uint32_t        map_sx;
    const uint32_t *ids;
    uint32_t        n;
    uint32_t        nchunks;
    uint32_t        mb0, mb1, mb2;
    uint32_t        single;
    int16_t        *geom;
    uint32_t       *index;
    uint64_t        cap;
    uint64_t       *offsets;
    uint32_t       *ticket;
    uint64_t       *state;
    uint32_t       *status;
};

// corners of the unit cube's +x, +y, +z faces in vertex order, packed as
// 3 bits (x | y << 1 | z << 2) per vertex (mesher.c:19-31)
__device__ __forceinline__ uint32_t m_corner(uint32_t axis, uint32_t v)
{
    // axis 0: (1,0,0) (1,0,1) (1,1,1) (1,1,0) -> 1 5 7 3
    // axis 1: (0,1,0) (1,1,0) (1,1,1) (0,1,1) -> 2 3 7 6
    // axis 2: (1,0,1) (0,0,1) (0,1,1) (1,1,1) -> 5 4 6 7
    constexpr uint32_t lut = (1u | 5u << 3 | 7u << 6 | 3u << 9) |
                             (2u | 3u << 3 | 7u << 6 | 6u << 9) << 12;
    constexpr uint32_t lut2 = 5u | 4u << 3 | 6u << 6 | 7u << 9;
    const uint32_t w = axis == 2 ? lut2 : (lut >> (12 * axis));
    return (w >> (3 * v)) & 7u;
}

// triangle index rows by [normal][rotated] (mesher.c:54-66), 2 bits each
__device__ __forceinline__ uint32_t m_tri(uint32_t normal, uint32_t rot, uint32_t k)
{
    constexpr uint32_t r00 = 0u | 3u << 2 | 1u << 4 | 2u << 6 | 1u << 8 | 3u << 10;
    constexpr uint32_t r01 = 3u | 2u << 2 | 0u << 4 | 1u << 6 | 0u << 8 | 2u << 10;
    constexpr uint32_t r10 = 1u | 3u << 2 | 0u << 4 | 3u << 6 | 1u << 8 | 2u << 10;
    constexpr uint32_t r11 = 0u | 2u << 2 | 3u << 4 | 2u << 6 | 0u << 8 | 1u << 10;
    const uint32_t row = normal ? (rot ? r11 : r10) : (rot ? r01 : r00);
    return (row >> (2 * k)) & 3u;
}

// shadow_sample (side +1) / shadow_sample_normal (side -1): lit iff both map
// entries are below (wy + 1) in u32; at wy = 0xFFFFFFFF that is never true,
// and the map is not read
__device__ __forceinline__ uint32_t m_shadow(const uint16_t *map, uint32_t sx,
                                             uint32_t wx, uint32_t wy, uint32_t wz, int side)
{
    const uint32_t top = wy + 1u;
    if (top == 0u) return 1u;
    const uint32_t i = (wx + wy) + sx * wz;
    const uint32_t a = map[i], b = map[(uint32_t)((int32_t)i + side)];
    return (a < top && b < top) ? 0u : 1u;
}

// 16 bytes -> 16 occupancy bits
__device__ __forceinline__ uint32_t nz4(uint32_t w)
{
    const uint32_t m = (((w & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | w) & 0x80808080u;
    return (((m >> 7) * 0x00204081u) >> 21) & 0xFu;
}

template <int B>
struct MeshShape {
    static constexpr int R = 1 << B;
    static constexpr int P = R + 2;                       // rows per axis with the shell
    static constexpr int SEG = R >= 8 ? R / 8 : 1;        // segments per row
    static constexpr int SW = R >= 8 ? 8 : R;             // voxels per segment
    static constexpr int IS = R * SEG;                    // items per slice
    static constexpr int SPI = (kMeshBS / IS) < R ? (kMeshBS / IS) : R;   // slices per pass
    static constexpr int LIST = 3 * SPI * R * R;          // faces per pass, at most
    static constexpr int QW = R >= 16 ? 16 : R;           // bytes per row load
    static constexpr int Q = R / QW;                      // loads per row
    static constexpr uint64_t FULL = R == 64 ? ~0ull : ((1ull << R) - 1ull);
};

template <int BS>
__device__ __forceinline__ uint32_t m_scan_excl(uint32_t v, uint32_t *wsum, uint32_t &total)
{
    constexpr int NW = BS / 64;
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

template <int B>
__global__ void __launch_bounds__(kMeshBS) k_mesh(MeshArgs a)
{
    using S = MeshShape<B>;
    constexpr int R = S::R, P = S::P;
    __shared__ uint64_t occ[P * P];
    __shared__ uint8_t  elo[P * P];       // x = -1
    __shared__ uint8_t  ehi[P * P];       // x = R
    __shared__ uint16_t list[S::LIST];
    __shared__ uint32_t wsum[kMeshBS / 64 + 1];
    __shared__ uint32_t s_ticket;
    __shared__ uint64_t s_base;

    const uint32_t mx = 1u << a.mb0, my = 1u << a.mb1, mz = 1u << a.mb2;
    for (uint32_t round = 0;; round++) {
        // a single chunk takes no ticket: the one workgroup does list entry 0
        if (threadIdx.x == 0) s_ticket = a.single ? (round == 0 ? 0u : a.n) : atomicAdd(a.ticket, 1u);
        __syncthreads();
        const uint32_t k = s_ticket;
        __syncthreads();
        if (k >= a.n) return;
        const uint32_t id = a.ids[k];
        const bool bad = id >= a.nchunks;
        const uint32_t cx = id & (mx - 1u), cy = (id >> a.mb0) & (my - 1u), cz = id >> (a.mb0 + a.mb1);

        // ---- 1. occupancy with the one-voxel shell ----
        if (!bad) {
            for (int t = threadIdx.x; t < P * P * S::Q; t += kMeshBS) {
                const int row = t / S::Q, q = t % S::Q;
                const int yy = row % P, zz = row / P;
                const int64_t ncy = (int64_t)cy + (yy == 0 ? -1 : (yy == P - 1 ? 1 : 0));
                const int64_t ncz = (int64_t)cz + (zz == 0 ? -1 : (zz == P - 1 ? 1 : 0));
                uint32_t bits = 0;
                if (ncy >= 0 && ncy < my && ncz >= 0 && ncz < mz) {
                    const uint32_t nid = (((uint32_t)ncz << a.mb1 | (uint32_t)ncy) << a.mb0) | cx;
                    const int16_t u = a.uniform ? a.uniform[nid] : (int16_t)-1;
                    if (u >= 0) {
                        bits = u ? (1u << S::QW) - 1u : 0u;
                    } else {
                        const uint32_t ly = (uint32_t)(yy - 1) & (R - 1), lz = (uint32_t)(zz - 1) & (R - 1);
                        const uint8_t *p = a.voxels + ((uint64_t)nid << (3 * B)) +
                                           ((((uint64_t)lz << B) | ly) << B) + q * S::QW;
                        if constexpr (S::QW == 16) {
                            const uint4 v = *reinterpret_cast<const uint4 *>(p);
                            bits = nz4(v.x) | nz4(v.y) << 4 | nz4(v.z) << 8 | nz4(v.w) << 12;
                        } else if constexpr (S::QW == 8) {
                            const uint2 v = *reinterpret_cast<const uint2 *>(p);
                            bits = nz4(v.x) | nz4(v.y) << 4;
                        } else {
                            bits = nz4(*reinterpret_cast<const uint32_t *>(p));
                        }
                    }
                }
                uint64_t r = (uint64_t)bits << (q * S::QW);
                if constexpr (S::Q > 1) {
#pragma unroll
                    for (int d = 1; d < S::Q; d <<= 1) r |= __shfl_xor(r, d, 64);
                }
                if (q == 0) occ[row] = r;
            }
            for (int t = threadIdx.x; t < P * P * 2; t += kMeshBS) {
                const int row = t >> 1, hi = t & 1;
                const int yy = row % P, zz = row / P;
                const int64_t ncx = (int64_t)cx + (hi ? 1 : -1);
                const int64_t ncy = (int64_t)cy + (yy == 0 ? -1 : (yy == P - 1 ? 1 : 0));
                const int64_t ncz = (int64_t)cz + (zz == 0 ? -1 : (zz == P - 1 ? 1 : 0));
                uint8_t bit = 0;
                if (ncx >= 0 && ncx < mx && ncy >= 0 && ncy < my && ncz >= 0 && ncz < mz) {
                    const uint32_t nid = (((uint32_t)ncz << a.mb1 | (uint32_t)ncy) << a.mb0) |
                                         (uint32_t)ncx;
                    const int16_t u = a.uniform ? a.uniform[nid] : (int16_t)-1;
                    if (u >= 0) {
                        bit = u ? 1 : 0;
                    } else {
                        const uint32_t ly = (uint32_t)(yy - 1) & (R - 1), lz = (uint32_t)(zz - 1) & (R - 1);
                        const uint32_t lx = hi ? 0u : (uint32_t)(R - 1);
                        bit = a.voxels[((uint64_t)nid << (3 * B)) + ((((uint64_t)lz << B) | ly) << B) +
                                       lx] != 0;
                    }
                }
                (hi ? ehi : elo)[row] = bit;
            }
        }
        __syncthreads();

        // ---- 2. face count, output offset ----
        uint32_t cnt = 0;
        if (!bad) {
            for (int t = threadIdx.x; t < R * R; t += kMeshBS) {
                const int z = t / R, y = t % R;
                const int r0 = (z + 1) * P + (y + 1);
                const uint64_t o = occ[r0];
                const uint64_t bx = (o >> 1) | ((uint64_t)ehi[r0] << (R - 1));
                cnt += __popcll((o ^ bx) & S::FULL) + __popcll((o ^ occ[r0 + 1]) & S::FULL) +
                       __popcll((o ^ occ[r0 + P]) & S::FULL);
            }
        }
        uint32_t faces;
        (void)m_scan_excl<kMeshBS>(cnt, wsum, faces);
        if (threadIdx.x < 64) {
            uint64_t excl = 0;
            if (!a.single) {
                if (k == 0) {
                    if (threadIdx.x == 0) st_relaxed(a.state, kMFlagPre | faces);
                } else {
                    if (threadIdx.x == 0) st_relaxed(a.state + k, kMFlagAgg | faces);
                    int64_t top = (int64_t)k - 1;
                    const uint32_t ln = threadIdx.x & 63u;
                    for (;;) {
                        const int64_t j = top - (int64_t)ln;
                        const uint64_t s = (j >= 0) ? ld_relaxed(a.state + j) : kMFlagPre;
                        const uint64_t f = s >> 62;
                        const uint64_t pre = __ballot(f == 2);
                        const uint64_t zero = __ballot(f == 0);
                        const uint64_t need = pre ? (pre & (~pre + 1ull)) * 2ull - 1ull : ~0ull;
                        if (zero & need) {
                            __builtin_amdgcn_s_sleep(1);
                            continue;
                        }
                        uint64_t v = ((need >> ln) & 1ull) ? (s & kMValMask) : 0ull;
#pragma unroll
                        for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
                        excl += v;
                        if (pre) break;
                        top -= 64;
                    }
                    if (threadIdx.x == 0) st_relaxed(a.state + k, kMFlagPre | (excl + faces));
                }
            }
            if (threadIdx.x == 0) {
                s_base = excl;
                a.offsets[k] = excl;
                if (k + 1 == a.n) a.offsets[a.n] = excl + faces;
                if (bad) atomicOr(a.status, (uint32_t)VP_STATUS_BAD_ID);
                if (a.geom && excl + faces > a.cap) atomicOr(a.status, (uint32_t)VP_STATUS_OVERFLOW);
            }
        }
        __syncthreads();
        const uint64_t base = s_base;
        if (bad || !a.geom || faces == 0 || base + faces > a.cap) continue;

        // ---- 3./4. faces, SPI slices per pass ----
        const uint32_t ox = cx << B, oy = cy << B, oz = cz << B;
        const uint8_t *cvox = a.voxels + ((uint64_t)id << (3 * B));
        const int16_t cuni = a.uniform ? a.uniform[id] : (int16_t)-1;
        uint32_t done = 0;                      // faces of earlier passes
        for (int z0 = 0; z0 < R; z0 += S::SPI) {
            uint32_t mxb = 0, myb = 0, mzb = 0;
            int zs = 0, y = 0, s = 0;
            if (threadIdx.x < S::SPI * S::IS) {
                zs = threadIdx.x / S::IS;
                const int rem = threadIdx.x % S::IS;
                y = rem / S::SEG;
                s = rem % S::SEG;
                const int r0 = (z0 + zs + 1) * P + (y + 1);
                const uint64_t o = occ[r0];
                const uint64_t bx = (o >> 1) | ((uint64_t)ehi[r0] << (R - 1));
                const uint32_t sh = s * S::SW;
                constexpr uint32_t SM = (S::SW == 32) ? ~0u : ((1u << S::SW) - 1u);
                mxb = (uint32_t)(((o ^ bx) & S::FULL) >> sh) & SM;
                myb = (uint32_t)(((o ^ occ[r0 + 1]) & S::FULL) >> sh) & SM;
                mzb = (uint32_t)(((o ^ occ[r0 + P]) & S::FULL) >> sh) & SM;
            }
            uint32_t pass;
            uint32_t pos = m_scan_excl<kMeshBS>(__popc(mxb) + __popc(myb) + __popc(mzb), wsum, pass);
            uint32_t any = mxb | myb | mzb;
            while (any) {
                const uint32_t j = __builtin_ctz(any);
                any &= any - 1u;
                const uint32_t cell = ((uint32_t)(zs * R + y) * R) + s * S::SW + j;
                if ((mxb >> j) & 1u) list[pos++] = (uint16_t)(cell << 2 | 0u);
                if ((myb >> j) & 1u) list[pos++] = (uint16_t)(cell << 2 | 1u);
                if ((mzb >> j) & 1u) list[pos++] = (uint16_t)(cell << 2 | 2u);
            }
            __syncthreads();
            for (uint32_t f = threadIdx.x; f < pass; f += kMeshBS) {
                const uint32_t e = list[f];
                const uint32_t axis = e & 3u, cell = e >> 2;
                const int x = cell & (R - 1), yl = (cell >> B) & (R - 1), z = z0 + (int)(cell >> (2 * B));
                const int r0 = (z + 1) * P + (yl + 1);
                const uint32_t solid_a = (uint32_t)(occ[r0] >> x) & 1u;
                const uint32_t normal = solid_a ^ 1u;
                // the air cell of the pair (A itself, or B when A is solid)
                const int px = x + (axis == 0 ? (int)solid_a : 0);
                const int py = yl + (axis == 1 ? (int)solid_a : 0);
                const int pz = z + (axis == 2 ? (int)solid_a : 0);
                // occupancy at shell coordinates in [-1, R]
                auto sol = [&](int qx, int qy, int qz) -> uint32_t {
                    const int r = (qz + 1) * P + (qy + 1);
                    if (qx < 0) return elo[r];
                    if (qx >= R) return ehi[r];
                    return (uint32_t)(occ[r] >> qx) & 1u;
                };
                // steps d0 along u0 and d1 along u1, the two axes other than
                // the face's (in ascending order)
                auto at = [&](int d0, int d1) -> uint32_t {
                    return sol(px + (axis != 0 ? d0 : 0),
                               py + (axis == 0 ? d0 : (axis == 2 ? d1 : 0)),
                               pz + (axis != 2 ? d1 : 0));
                };
                const uint32_t n0 = at(-1, 0), n1 = at(0, -1), n2 = at(1, 0), n3 = at(0, 1);
                const uint32_t c0 = (n0 + n1) | at(-1, -1);
                const uint32_t c1 = (n1 + n2) | at(1, -1);
                const uint32_t c2 = (n2 + n3) | at(1, 1);
                const uint32_t c3 = (n3 + n0) | at(-1, 1);
                uint32_t v0, v1, v2, v3;
                if (axis == 0) {
                    v0 = c0; v3 = c1; v2 = c2; v1 = c3;
                } else if (axis == 1) {
                    v0 = c0; v1 = c1; v2 = c2; v3 = c3;
                } else {
                    v1 = c0; v0 = c1; v3 = c2; v2 = c3;
                }
                const uint32_t rot = (v0 + v2 < v1 + v3) ? 1u : 0u;
                const uint32_t wx = ox + x, wy = oy + yl, wz = oz + z;
                // colour of the solid voxel of the pair
                uint32_t colour;
                if (solid_a) {
                    colour = cuni >= 0 ? (uint32_t)cuni
                                       : cvox[(((uint32_t)z << B | (uint32_t)yl) << B) | (uint32_t)x];
                } else {
                    const uint32_t bx = x + (axis == 0), by = yl + (axis == 1), bz = z + (axis == 2);
                    const uint32_t edge = axis == 0 ? bx : (axis == 1 ? by : bz);
                    if (edge < (uint32_t)R) {
                        colour = cuni >= 0 ? (uint32_t)cuni : cvox[((bz << B | by) << B) | bx];
                    } else {
                        // in the +axis neighbour, which exists: the voxel is solid
                        const uint32_t nid = (((cz + (axis == 2)) << a.mb1 | (cy + (axis == 1))) << a.mb0) |
                                             (cx + (axis == 0));
                        const uint32_t li = (((bz & (R - 1)) << B | (by & (R - 1))) << B) | (bx & (R - 1));
                        const int16_t u = a.uniform ? a.uniform[nid] : (int16_t)-1;
                        colour = u >= 0 ? (uint32_t)u : a.voxels[((uint64_t)nid << (3 * B)) + li];
                    }
                }
                uint32_t shadow = 0, diamond = 3;
                if (axis == 0 && normal) {
                    shadow = m_shadow(a.map, a.map_sx, wx + 1u, wy, wz, -1);
                } else if (axis == 1 && !normal) {
                    shadow = m_shadow(a.map, a.map_sx, wx, wy, wz, +1);
                } else if (axis == 2) {
                    const uint32_t sz = wz + (solid_a ? 1u : 0u);   // B air: the cell in front
                    diamond = (m_shadow(a.map, a.map_sx, wx, wy, sz, +1) ^ 1u) << 1;
                    diamond |= m_shadow(a.map, a.map_sx, wx, wy - 1u, sz, +1) ^ 1u;
                }
                const uint32_t common = colour | (axis + 3u * normal) << 8 | shadow << 13 | diamond << 14;
                const uint32_t vao[4] = {v0, v1, v2, v3};
                uint32_t wd[8];
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const uint32_t c = m_corner(axis, v);
                    const uint32_t gx = (wx + (c & 1u)) & 0xFFFFu;
                    const uint32_t gy = (wy + ((c >> 1) & 1u)) & 0xFFFFu;
                    const uint32_t gz = (wz + (c >> 2)) & 0xFFFFu;
                    const uint32_t attr = (common | vao[v] << 6 | (uint32_t)v << 11) & 0xFFFFu;
                    wd[2 * v] = gx | gy << 16;
                    wd[2 * v + 1] = gz | attr << 16;
                }
                const uint64_t gf = base + done + f;
                uint4 *gp = reinterpret_cast<uint4 *>(a.geom + 16 * gf);
                gp[0] = make_uint4(wd[0], wd[1], wd[2], wd[3]);
                gp[1] = make_uint4(wd[4], wd[5], wd[6], wd[7]);
                const uint32_t vb = 4u * (done + f);
                uint2 *ip = reinterpret_cast<uint2 *>(a.index + 6 * gf);
                ip[0] = make_uint2(vb + m_tri(normal, rot, 0), vb + m_tri(normal, rot, 1));
                ip[1] = make_uint2(vb + m_tri(normal, rot, 2), vb + m_tri(normal, rot, 3));
                ip[2] = make_uint2(vb + m_tri(normal, rot, 4), vb + m_tri(normal, rot, 5));
            }
            done += pass;
            __syncthreads();
        }
    }
}

__global__ void k_mesh_reset(uint32_t *ticket, uint64_t *state, uint32_t n, uint64_t *offsets)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        ticket[0] = 0u;
        offsets[0] = 0ull;
    }
    for (uint32_t j = i; j < n; j += gridDim.x * blockDim.x) state[j] = 0ull;
}

template <int B>
static hipError_t mesh_launch(const MeshArgs &a, int grid, hipStream_t st)
{
    hipLaunchKernelGGL((k_mesh<B>), dim3(grid), dim3(kMeshBS), 0, st, a);
    return hipGetLastError();
}

}  // namespace vp

using namespace vp;

extern "C" int vp_gpu_mesh(vp_ctx *ctx, const vp_world_desc *world, const uint8_t *d_voxels,
                           const int16_t *d_uniform, const vp_shadow_desc *shadow,
                           const uint32_t *d_chunk_ids, uint32_t n, const vp_mesh_out *out,
                           void *stream)
{
    if (!ctx || !world || !d_voxels || !shadow || !shadow->d_map || (!d_chunk_ids && n) || !out ||
        !out->d_face_offsets)
        return VP_ERR_ARG;
    const bool count_only = !out->d_geometry && !out->d_index && out->capacity_faces == 0;
    if (!count_only && (!out->d_geometry || !out->d_index)) return VP_ERR_ARG;
    const uint32_t root_bitw = world->root_bitw;
    if (root_bitw < 2 || root_bitw > 6) return VP_ERR_UNSUPPORTED;
    const uint32_t mbs = (uint32_t)world->max_bitw[0] + world->max_bitw[1] + world->max_bitw[2];
    if (mbs > 31u || mbs + 3u * root_bitw > 40u) return VP_ERR_ARG;
    const uint64_t sx = ((uint64_t)1 << (world->max_bitw[0] + root_bitw)) +
                        ((uint64_t)1 << (world->max_bitw[1] + root_bitw));
    if (shadow->size_x != sx) return VP_ERR_ARG;
    hipStream_t st = as_stream(stream);
    MeshArgs a = {};
    a.voxels = d_voxels;
    a.uniform = d_uniform;
    a.map = shadow->d_map;
    a.map_sx = shadow->size_x;
    a.ids = d_chunk_ids;
    a.n = n;
    a.nchunks = (uint32_t)(1ull << mbs);
    a.mb0 = world->max_bitw[0];
    a.mb1 = world->max_bitw[1];
    a.mb2 = world->max_bitw[2];
    a.single = (n == 1) ? 1u : 0u;
    a.geom = count_only ? nullptr : out->d_geometry;
    a.index = count_only ? nullptr : out->d_index;
    a.cap = out->capacity_faces;
    a.offsets = out->d_face_offsets;
    a.status = ctx->d_status;
    std::lock_guard<std::mutex> g(ctx->mu);
    int grid = 1;
    if (!a.single) {
        int rc = ensure_state(ctx, n ? n : 1);
        if (rc) return rc;
        a.ticket = ctx->d_ticket;
        a.state = ctx->d_state;
        const int resident = ctx->num_cus * 2;
        grid = (int)(n < (uint32_t)resident ? n : (uint32_t)resident);
        const int rgrid = (int)((n + 255u) / 256u);
        hipLaunchKernelGGL(k_mesh_reset, dim3(rgrid < 1 ? 1 : (rgrid < 1024 ? rgrid : 1024)),
                           dim3(256), 0, st, a.ticket, a.state, n, a.offsets);
        VP_HIP_TRY(hipGetLastError());
        if (n == 0) return VP_OK;
    }
    hipError_t e;
    switch (root_bitw) {
    case 2: e = mesh_launch<2>(a, grid, st); break;
    case 3: e = mesh_launch<3>(a, grid, st); break;
    case 4: e = mesh_launch<4>(a, grid, st); break;
    case 5: e = mesh_launch<5>(a, grid, st); break;
    default: e = mesh_launch<6>(a, grid, st); break;
    }
    return e == hipSuccess ? VP_OK : VP_ERR_HIP;
}
