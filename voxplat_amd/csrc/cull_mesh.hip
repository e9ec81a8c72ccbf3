// cull_mesh.hip -- face culling + splat-list emission + LOD mip for gfx950.
//
// Replaces, for a list of chunks, the meshing body of chunkset_manage
// (src/chunkset.c:318-505): chunk_make_mask (src/chunkset/mesher.c:377-456),
// chunk_make_splatlist for levels 0..4 (mesher.c:497-536), the four
// chunk_mask_downsample steps (mesher.c:460-493) and shadow_sample
// (src/chunkset/shadow.h:45-63).
//
// One workgroup owns one chunk at a time (R^3 uint8 voxels, R <= 64).
//  1. occupancy: each voxel row (fixed y,z) becomes one 64-bit word in LDS,
//     built from coalesced 16-byte loads; the +x/+y/+z neighbour planes become
//     three small bit arrays (the +x plane through a wave ballot).
//  2. visibility: mask cell q holds a voxel iff that voxel is solid and
//     exposed through a face this chunk tests (mesher.c:421-448); per row that
//     is a handful of shifts, ANDs and NOTs.  Cells at coordinate R hold the
//     +neighbour's voxels exposed towards this chunk.
//  3. LOD: the level-l cell P is non-zero iff one of its children 2P+d is;
//     its colour is that of the greatest-(z,y,x) non-zero child (the last
//     write of the downsample scan), found by descending the bit pyramid.
//  4. counts per row -> workgroup scan -> per-chunk total; the chunk's
//     output offset comes from a decoupled look-back over the work-queue
//     ticket order, so the concatenated output is in list order.
//  5. emission: every z-slice of every level is a contiguous range of the
//     output; waves take slices from a work counter in LDS and write them
//     without workgroup barriers (emit_slices).
#include <type_traits>

#include "vp_internal.h"
#include "../../include/vp_host.h"

namespace vp {

constexpr uint64_t kFlagAgg = 1ull << 62;
constexpr uint64_t kFlagPre = 2ull << 62;
constexpr uint64_t kValMask = (1ull << 62) - 1;

struct CullArgs {
    const uint8_t  *voxels;
    const int16_t  *uniform;
    const uint16_t *map;
    uint32_t        sx;
    const uint32_t *ids;
    uint32_t        n;
    const uint32_t *d_n;
    uint8_t         mb[3];
    uint8_t         lod;
    uint8_t         count_only;
    uint8_t         single;        // n == 1 without a device count: no ticket
    uint64_t       *splats;
    uint64_t        cap;
    uint64_t       *offsets;
    uint32_t       *level_counts;
    uint32_t       *ticket;
    uint64_t       *state;
    uint32_t       *status;
    uint8_t        *masks;         // make_mask mode: (R+1)^3 bytes per chunk
    uint32_t        nchunks;       // chunks in the world (ids are checked against it)
};

template <int B>
struct Geo {
    static constexpr int R = 1 << B;
    static constexpr int N0 = R + 1;
    static constexpr uint64_t ROWMASK = (R == 64) ? ~0ull : ((1ull << R) - 1ull);
    static constexpr int n(int l) { return (R >> l) + 1; }
    static constexpr int RB1 = N0 * N0;
    static constexpr int RB2 = RB1 + n(1) * n(1);
    static constexpr int RB3 = RB2 + n(2) * n(2);
    static constexpr int RB4 = RB3 + n(3) * n(3);
    static constexpr int ROWS = RB4 + n(4) * n(4);
    static __device__ __forceinline__ int row_base(int l)
    {
        return l == 0 ? 0 : l == 1 ? RB1 : l == 2 ? RB2 : l == 3 ? RB3 : l == 4 ? RB4 : ROWS;
    }
};

// LDS layout in 64-bit words.  The level-0 visibility rows V0 are built over
// the occupancy rows S in place (every S row is read before any V0 row is
// written), unless v2 face flags still need S afterwards.  Past the LOD rows
// come the slice offsets, the emission work counter and one area per wave:
// the wave's list of cells to emit and a window of one shadow-map row.
template <int B, bool V2>
struct Rows {
    using G = Geo<B>;
    static constexpr int R = G::R, N0 = G::N0;
    static constexpr int S = 0;
    static constexpr int V0 = V2 ? R * R : 0;
    static constexpr int HX = V0 + N0 * N0;
    static constexpr int HY = HX + R;
    static constexpr int HZ = HY + R;
    static constexpr int XB = HZ + R;
    static constexpr int P1 = XB + R;               // P1..P4 contiguous, row order
    static constexpr int PEND = P1 + (G::ROWS - G::RB1);
    // z-slices of every level, level-major: slice base of level l
    static constexpr int SLB0 = 0, SLB1 = G::n(0), SLB2 = SLB1 + G::n(1), SLB3 = SLB2 + G::n(2),
                         SLB4 = SLB3 + G::n(3), NSLICES = SLB4 + G::n(4);
};

template <int B, int BS, bool V2>
struct Lay : Rows<B, V2> {
    using P = Rows<B, V2>;
    static constexpr int NW = BS / 64;
    // u32 words: slice offsets [0, NSLICES], then the work counter
    static constexpr int SOFF = P::PEND;
    static constexpr int ITEM_U32 = P::NSLICES + 1;
    static constexpr int SOFF_WORDS = (P::NSLICES + 2 + 1) / 2;
    // LOD colours: n(l)^3 bytes per level 1-4, cell (x, y, z) at (z n + y) n + x.
    // Built once per chunk right after the chunk was read (its lines are still
    // in L2), so the emission of levels 1-4 reads colours from LDS.  Only the
    // 1024-thread R = 64 kernel without v2 flags has the LDS room (41.7 KiB).
    static constexpr bool MIP = (B == 6 && BS == 1024 && !V2);
    static constexpr int cube(int v) { return v * v * v; }
    static constexpr int cbase(int l)
    {
        return l <= 1 ? 0 : cbase(l - 1) + cube(Geo<B>::n(l - 1));
    }
    // per wave: LCAP list entries and SWIN shadow entries.  An entry is the
    // cell code y << 7 | x, u16 in the MIP kernel (which keeps no level-1
    // child selection in it, so its lists hold 1920 cells: most slices in
    // one batch, and the byte plane below still fits), else u32 with the
    // selection in bits 16-18.
    static constexpr int mini(int a, int b) { return a < b ? a : b; }
    static constexpr int EB = MIP ? 2 : 4;
    static constexpr int LCAP = mini(MIP ? 1920 : 512, P::N0 * P::N0);
    static constexpr int SWIN = 2 * P::R + 2;
    static constexpr int WAVE_WORDS = (EB * LCAP + 2 * SWIN + 7) / 8;
    static constexpr int WAVES = SOFF + SOFF_WORDS;
    static constexpr int COL = WAVES + NW * WAVE_WORDS;
    // the +x neighbour's x = 0 plane as bytes, voxel (0, y, z) at z R + y:
    // the colours of the cells at x = R, kept from the plane's reads
    static constexpr int XCB = COL + (MIP ? (cbase(5) + 7) / 8 : 0);
    static constexpr int TOTAL = XCB + (P::R * P::R + 7) / 8;
};

__device__ __forceinline__ uint64_t compress_pairs(uint64_t r)
{
    uint64_t x = (r | (r >> 1)) & 0x5555555555555555ull;
    x = (x | (x >> 1)) & 0x3333333333333333ull;
    x = (x | (x >> 2)) & 0x0F0F0F0F0F0F0F0Full;
    x = (x | (x >> 4)) & 0x00FF00FF00FF00FFull;
    x = (x | (x >> 8)) & 0x0000FFFF0000FFFFull;
    x = (x | (x >> 16)) & 0x00000000FFFFFFFFull;
    return x;
}

// 4-bit "byte is non-zero" mask of a little-endian word
__device__ __forceinline__ uint32_t nz4(uint32_t w)
{
    uint32_t m = (((w & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | w) & 0x80808080u;
    return (((m >> 7) * 0x00204081u) >> 21) & 0xFu;
}
__device__ __forceinline__ uint32_t nz16(uint4 v)
{
    return nz4(v.x) | (nz4(v.y) << 4) | (nz4(v.z) << 8) | (nz4(v.w) << 12);
}

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

// wave-uniform values read from LDS: keep them in scalar registers.  The
// builtin returns int: each half goes through uint32_t so that a low word
// >= 2^31 is not sign-extended into the high word.
__device__ __forceinline__ uint32_t uni(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_readfirstlane(v);
}
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint64_t uni(uint64_t v)
{
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)v);
    return ((uint64_t)hi << 32) | (uint64_t)lo;
}

__device__ __forceinline__ uint32_t rank_below(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// position of the k-th set bit (k counted from 0) of v; v has more than k bits
__device__ __forceinline__ int select64(uint64_t v, uint32_t k)
{
    int pos = 0;
    uint32_t c = __popc((uint32_t)v);
    if (k >= c) { k -= c; v >>= 32; pos += 32; }
    c = __popc((uint32_t)v & 0xFFFFu);
    if (k >= c) { k -= c; v >>= 16; pos += 16; }
    c = __popc((uint32_t)v & 0xFFu);
    if (k >= c) { k -= c; v >>= 8; pos += 8; }
    c = __popc((uint32_t)v & 0xFu);
    if (k >= c) { k -= c; v >>= 4; pos += 4; }
    c = __popc((uint32_t)v & 0x3u);
    if (k >= c) { k -= c; v >>= 2; pos += 2; }
    c = (uint32_t)v & 1u;
    if (k >= c) pos += 1;
    return pos;
}

struct ChunkRefs {
    const uint8_t *own, *nb[3];   // nb[a] == nullptr beyond the world
    int own_u, nb_u[3];           // uniform value or -1
    uint32_t org[3];              // world origin of the chunk (offset * R)
};

// Buffer resource over [p, p + bytes): loads share one VGPR offset, take the
// per-load step from a scalar offset, and read 0 past the end.  0x00020000 is
// the gfx9 word-3 format of a raw byte buffer.  p must be wave-uniform: it is
// moved to scalar registers here, so the descriptor is scalar and a load
// needs no per-lane loop over descriptors.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t byte_rsrc(const void *p, uint32_t bytes)
{
    const uint64_t a = (uint64_t)(uintptr_t)p;
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)a);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(a >> 32));
    void *u = (void *)(uintptr_t)(((uint64_t)hi << 32) | lo);
    return __builtin_amdgcn_make_buffer_rsrc(u, (short)0, (int)bytes, 0x00020000);
}

__device__ __forceinline__ uint4 ld16(__amdgpu_buffer_rsrc_t rs, uint32_t voff, uint32_t soff)
{
    typedef unsigned int v4u __attribute__((ext_vector_type(4)));
    const v4u v = __builtin_amdgcn_raw_buffer_load_b128(rs, (int)voff, (int)soff, 0);
    return make_uint4(v.x, v.y, v.z, v.w);
}

// Fill `dst` (row words, rowcount rows of R bits) from R-byte rows at src with
// row stride `stride` (bytes).  Rows are numbered r = 0..rowcount-1 and row r
// starts at src + r*stride.
template <int B, int BS>
__device__ __forceinline__ void load_rows(uint64_t *dst, const uint8_t *src, int rowcount,
                                          int stride)
{
    constexpr int R = 1 << B;
    if constexpr (R == 64) {
        uint16_t *d16 = reinterpret_cast<uint16_t *>(dst);
        const int total = rowcount * 4;
        if (stride == R) {
            // contiguous rows: byte offset = 16 * index, one buffer for all
            constexpr int DEPTH = 8;
            const __amdgpu_buffer_rsrc_t rs = byte_rsrc(src, (uint32_t)total * 16u);
            int i = threadIdx.x;
#pragma unroll 1
            for (; i + (DEPTH - 1) * BS < total; i += DEPTH * BS) {
                uint4 v[DEPTH];
#pragma unroll
                for (int j = 0; j < DEPTH; j++) v[j] = ld16(rs, (uint32_t)i * 16u, (uint32_t)(j * BS * 16));
#pragma unroll
                for (int j = 0; j < DEPTH; j++) d16[i + j * BS] = (uint16_t)nz16(v[j]);
            }
#pragma unroll 1
            for (; i < total; i += BS) d16[i] = (uint16_t)nz16(ld16(rs, (uint32_t)i * 16u, 0u));
        } else {
            for (int i = threadIdx.x; i < total; i += BS) {
                const int r = i >> 2, p = i & 3;
                const uint4 v = *reinterpret_cast<const uint4 *>(src + (size_t)r * stride + p * 16);
                d16[i] = (uint16_t)nz16(v);
            }
        }
    } else if constexpr (R >= 16) {
        constexpr int PARTS = R / 16;
        uint16_t *d16 = reinterpret_cast<uint16_t *>(dst);
        const int total = rowcount * PARTS;
        for (int i = threadIdx.x; i < total; i += BS) {
            const int r = i / PARTS, p = i % PARTS;
            const uint4 v = *reinterpret_cast<const uint4 *>(src + (size_t)r * stride + p * 16);
            d16[r * 4 + p] = (uint16_t)nz16(v);
            if (p == PARTS - 1)
                for (int q = PARTS; q < 4; q++) d16[r * 4 + q] = 0;
        }
    } else {
        for (int r = threadIdx.x; r < rowcount; r += BS) {
            uint64_t m = 0;
            for (int x = 0; x < R; x++) m |= (uint64_t)(src[(size_t)r * stride + x] != 0) << x;
            dst[r] = m;
        }
    }
}

template <int B>
__device__ __forceinline__ uint8_t l0_colour(const ChunkRefs &c, int x, int y, int z)
{
    constexpr int R = 1 << B;
    if (x == R) return c.nb_u[0] >= 0 ? (uint8_t)c.nb_u[0] : c.nb[0][(z * R + y) * R];
    if (y == R) return c.nb_u[1] >= 0 ? (uint8_t)c.nb_u[1] : c.nb[1][(z * R) * R + x];
    if (z == R) return c.nb_u[2] >= 0 ? (uint8_t)c.nb_u[2] : c.nb[2][y * R + x];
    return c.own_u >= 0 ? (uint8_t)c.own_u : c.own[(z * R + y) * R + x];
}

// Bit of level-0 cell (x,y,z), x,y,z in [0,R].
template <int B, bool V2>
__device__ __forceinline__ uint32_t l0_bit(const uint64_t *lds, int x, int y, int z)
{
    using L = Rows<B, V2>;
    constexpr int R = 1 << B, N0 = R + 1;
    if (x < R) return (uint32_t)(lds[L::V0 + z * N0 + y] >> x) & 1u;
    if (y < R && z < R) return (uint32_t)(lds[L::XB + z] >> y) & 1u;
    return 0u;
}

// Descend from level-l cell (x,y,z) to the level-0 cell whose value the
// chained downsample leaves there: at each step the greatest (z,y,x) non-zero
// child among 2P + {0,1}^3 inside the child grid.
template <int B, bool V2>
__device__ __forceinline__ void descend(const uint64_t *lds, int l, int &x, int &y, int &z)
{
    using L = Rows<B, V2>;
    using G = Geo<B>;
    for (int lv = l; lv >= 1; --lv) {
        const int np = G::n(lv - 1);
        bool found = false;
        for (int dz = 1; dz >= 0 && !found; --dz) {
            const int Z = 2 * z + dz;
            if (Z >= np) continue;
            for (int dy = 1; dy >= 0 && !found; --dy) {
                const int Y = 2 * y + dy;
                if (Y >= np) continue;
                const int X1 = 2 * x + 1, X0 = 2 * x;
                uint32_t b1, b0;
                if (lv - 1 == 0) {
                    b1 = (X1 < np) ? l0_bit<B, V2>(lds, X1, Y, Z) : 0u;
                    b0 = l0_bit<B, V2>(lds, X0, Y, Z);
                } else {
                    const uint64_t row =
                        lds[L::P1 + (G::row_base(lv - 1) - G::RB1) + Z * np + Y];
                    b1 = (X1 < np) ? (uint32_t)(row >> X1) & 1u : 0u;
                    b0 = (uint32_t)(row >> X0) & 1u;
                }
                if (b1) { x = X1; y = Y; z = Z; found = true; }
                else if (b0) { x = X0; y = Y; z = Z; found = true; }
            }
        }
    }
}

// The descent of descend() without branches: at each step the four child
// rows (z 2P+1, 2P; y 2Q+1, 2Q) are read together and the greatest (z,y,x)
// non-zero child is picked with bit operations.
template <int B, bool V2>
__device__ __forceinline__ void descend_bf(const uint64_t *lds, int l, int &x, int &y, int &z)
{
    using L = Rows<B, V2>;
    using G = Geo<B>;
    constexpr int R = G::R, N0 = G::N0;
    for (int lv = l; lv >= 1; --lv) {
        const int np = G::n(lv - 1);
        const int X0 = 2 * x, Y0 = 2 * y, Z0 = 2 * z;
        uint32_t b[2][2][2];          // [dz][dy][dx]
#pragma unroll
        for (int dz = 0; dz < 2; dz++)
#pragma unroll
            for (int dy = 0; dy < 2; dy++) {
                const int Z = Z0 + dz, Y = Y0 + dy;
                uint64_t row = 0;
                uint32_t bitR = 0;             // level 0: cell x = R of this row
                if (Z < np && Y < np) {
                    if (lv - 1 == 0) {
                        row = lds[L::V0 + Z * N0 + Y];
                        if (Z < R && Y < R) bitR = (uint32_t)(lds[L::XB + Z] >> Y) & 1u;
                    } else {
                        row = lds[L::P1 + (G::row_base(lv - 1) - G::RB1) + Z * np + Y];
                    }
                }
                // bit X of the child row, X < np; at level 0, X = R is bitR
                auto bit = [&](int X) -> uint32_t {
                    if (X >= np) return 0u;
                    if (lv - 1 == 0 && X == R) return bitR;
                    return (uint32_t)(row >> X) & 1u;
                };
                b[dz][dy][1] = bit(X0 + 1);
                b[dz][dy][0] = bit(X0);
            }
        const uint32_t a11 = b[1][1][1] | b[1][1][0], a10 = b[1][0][1] | b[1][0][0];
        const uint32_t a01 = b[0][1][1] | b[0][1][0];
        const uint32_t dz = a11 | a10;
        const uint32_t dy = dz ? a11 : a01;
        const uint32_t dx = b[dz][dy][1];
        x = X0 + (int)dx;
        y = Y0 + (int)dy;
        z = Z0 + (int)dz;
    }
}

// Face-exposure flags of level-0 cell q (v2 records): bit 2a = +a face open,
// bit 2a+1 = -a face open, as the checks of mesher.c:427-445 found them.
template <int B, bool V2>
__device__ __forceinline__ uint32_t face_flags(const uint64_t *lds, int x, int y, int z)
{
    using L = Rows<B, V2>;
    constexpr int R = 1 << B;
    if (x == R) return 2u;          // seen through the -x face from this chunk
    if (y == R) return 8u;
    if (z == R) return 32u;
    auto solid = [&](int xx, int yy, int zz) -> uint32_t {
        if (xx == R) return (uint32_t)(lds[L::HX + zz] >> yy) & 1u;
        if (yy == R) return (uint32_t)(lds[L::HY + zz] >> xx) & 1u;
        if (zz == R) return (uint32_t)(lds[L::HZ + yy] >> xx) & 1u;
        return (uint32_t)(lds[L::S + zz * R + yy] >> xx) & 1u;
    };
    uint32_t f = 0;
    f |= (!solid(x + 1, y, z)) << 0;
    f |= (x > 0 && !solid(x - 1, y, z)) << 1;
    f |= (!solid(x, y + 1, z)) << 2;
    f |= (y > 0 && !solid(x, y - 1, z)) << 3;
    f |= (!solid(x, y, z + 1)) << 4;
    f |= (z > 0 && !solid(x, y, z - 1)) << 5;
    return f;
}

__device__ __forceinline__ uint32_t shadow_bit(const uint16_t *map, uint32_t sx, uint32_t wx,
                                               uint32_t wy, uint32_t wz)
{
    if (!map) return 0u;
    const uint32_t i = (wx + wy) + sx * wz;        // shadow.h:45-51, u32 wrap
    return !(map[i] <= wy && map[i + 1] <= wy);    // shadow.h:54-63
}

template <bool V2>
__device__ __forceinline__ uint64_t pack_record(uint32_t wx, uint32_t wy, uint32_t wz,
                                                uint32_t colour, uint32_t sh, uint32_t faces)
{
    uint64_t r = (uint64_t)(uint16_t)wx | ((uint64_t)(uint16_t)wy << 16) |
                 ((uint64_t)(uint16_t)wz << 32);
    if constexpr (V2)
        return r | ((uint64_t)(colour & 0xFFu) << 48) | ((uint64_t)(faces | (sh << 6)) << 56);
    else
        return r | ((uint64_t)(uint16_t)(colour | (sh << 6)) << 48);
}

template <int BS>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *wsum, uint32_t &total)
{
    constexpr int NW = BS / 64;
    const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t y = __shfl_up(x, d, 64);
        if (lane >= (uint32_t)d) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    if (w == 0) {
        uint32_t s = lane < (uint32_t)NW ? wsum[lane] : 0u, t = s;
#pragma unroll
        for (int d = 1; d < NW; d <<= 1) {
            uint32_t y = __shfl_up(t, d, 64);
            if (lane >= (uint32_t)d) t += y;
        }
        if (lane < (uint32_t)NW) wsum[lane] = t - s;      // exclusive wave base
        if (lane == (uint32_t)NW - 1) wsum[NW] = t;       // block total
    }
    __syncthreads();
    total = uni(wsum[NW]);
    return wsum[w] + x - v;
}

// ---- emission ----------------------------------------------------------------
// Exclusive prefix sum over the lanes of a wave, and the wave's total.
__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t v, uint32_t &total)
{
    const uint32_t ln = lane_id();
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)x, d, 64);
        if (ln >= (uint32_t)d) x += y;
    }
    total = uni((uint32_t)__shfl((int)x, 63, 64));
    return x - v;
}

__device__ __forceinline__ uint32_t wave_min(uint32_t v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, d, 64));
    return uni(v);
}

// The cell list and the shadow window of a wave are written and read by
// different lanes of that wave: keep the accesses in program order.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Colour of level-0 cell (x, y, z), x, y, z in [0, R]: one select of the
// source and at most one byte load, plus one LDS byte read of the +x plane
// xc, taken when x = R (y, z < R there).  Branch-free: a cell at x = R
// reads the chunk's first byte instead (a line in the caches), so the
// record loop stays one path.
template <int B>
__device__ __forceinline__ uint32_t cell_colour(const ChunkRefs &c, const uint8_t *xc, int x, int y,
                                                int z)
{
    constexpr int R = 1 << B;
    const uint8_t *p = c.own;
    int u = c.own_u;
    size_t i = (size_t)((z * R + y) * R + x);
    if (x == R) {
        i = 0;
    } else if (y == R) {
        p = c.nb[1]; u = c.nb_u[1]; i = (size_t)(z * R * R + x);
    } else if (z == R) {
        p = c.nb[2]; u = c.nb_u[2]; i = (size_t)(y * R + x);
    }
    const uint32_t g = u >= 0 ? (uint32_t)u : (uint32_t)p[i];
    const uint32_t l = xc[(z * R + y) & (R * R - 1)];
    return x == R ? l : g;
}

// Even bits of r (bit 2k -> bit k), 32 of them.
__device__ __forceinline__ uint64_t even_bits(uint64_t r)
{
    uint64_t x = r & 0x5555555555555555ull;
    x = (x | (x >> 1)) & 0x3333333333333333ull;
    x = (x | (x >> 2)) & 0x0F0F0F0F0F0F0F0Full;
    x = (x | (x >> 4)) & 0x00FF00FF00FF00FFull;
    x = (x | (x >> 8)) & 0x0000FFFF0000FFFFull;
    x = (x | (x >> 16)) & 0x00000000FFFFFFFFull;
    return x;
}

// Level-0 row (y, z) as bits x = 0..R: the V0 row and, at bit R, the cell
// (R, y, z) of the +x neighbour (zero outside the mask's (R+1)^3 cells).
template <int B, bool V2>
__device__ __forceinline__ void l0_row65(const uint64_t *lds, int y, int z, uint64_t &lo64,
                                        uint32_t &bitR)
{
    using L = Rows<B, V2>;
    constexpr int R = 1 << B, N0 = R + 1;
    lo64 = 0;
    bitR = 0;
    if (y > R || z > R) return;
    lo64 = lds[L::V0 + z * N0 + y];
    if (y < R && z < R) bitR = (uint32_t)(lds[L::XB + z] >> y) & 1u;
}

// For level-1 row (Y, slice s): the level-0 child each cell X descends to --
// the greatest (z, y, x) non-zero child, the last write of the downsample
// scan (mesher.c:460-493) -- as three bit planes over X: dz (child in z =
// 2s+1), dy (child in y = 2Y+1), dx (child at x = 2X+1).
template <int B, bool V2>
__device__ __forceinline__ void l1_child_planes(const uint64_t *lds, int s, int Y,
                                                uint64_t &pdz, uint64_t &pdy, uint64_t &pdx)
{
    constexpr int R = 1 << B;
    uint64_t hi[2][2], lo[2][2], any[2][2];
#pragma unroll
    for (int dz = 0; dz < 2; dz++)
#pragma unroll
        for (int dy = 0; dy < 2; dy++) {
            uint64_t r;
            uint32_t br;
            l0_row65<B, V2>(lds, 2 * Y + dy, 2 * s + dz, r, br);
            // cell X: children x = 2X (even bit) and 2X + 1 (odd bit); the
            // cell X = R/2 has only x = R, bit R of the 65-bit row
            lo[dz][dy] = even_bits(r) | ((uint64_t)br << (R / 2));
            hi[dz][dy] = even_bits(r >> 1);
            any[dz][dy] = lo[dz][dy] | hi[dz][dy];
        }
    pdz = any[1][1] | any[1][0];
    pdy = (pdz & any[1][1]) | (~pdz & any[0][1]);
    pdx = (pdz & pdy & hi[1][1]) | (pdz & ~pdy & hi[1][0]) | (~pdz & pdy & hi[0][1]) |
          (~pdz & ~pdy & hi[0][0]);
}

// Child planes of a level-l row (l >= 2) from its four child rows c[dz][dy]
// of level l - 1 (zero past the child grid): for every cell X the greatest
// (z, y, x) non-zero child, as bit planes dz, dy, dx over X.
__device__ __forceinline__ void lod_child_planes(const uint64_t ch[2][2], uint64_t &pdz,
                                                 uint64_t &pdy, uint64_t &pdx)
{
    uint64_t hi[2][2], an[2][2];
#pragma unroll
    for (int dz = 0; dz < 2; dz++)
#pragma unroll
        for (int dy = 0; dy < 2; dy++) {
            hi[dz][dy] = even_bits(ch[dz][dy] >> 1);
            an[dz][dy] = even_bits(ch[dz][dy]) | hi[dz][dy];
        }
    pdz = an[1][1] | an[1][0];
    pdy = (pdz & an[1][1]) | (~pdz & an[0][1]);
    pdx = (pdz & pdy & hi[1][1]) | (pdz & ~pdy & hi[1][0]) | (~pdz & pdy & hi[0][1]) |
          (~pdz & ~pdy & hi[0][0]);
}

// Level-1 colours of the level-1 row (Y, Z), Y, Z <= R/2, R = 64: for every
// cell X the colour of its greatest (z, y, x) visible level-0 child, the value
// the downsample's last write leaves there (mesher.c:460-493).  The four child
// rows (y = 2Y + dy, z = 2Z + dz) are read in 16-byte loads, from the chunk
// or, at y = R or z = R, from the first plane of the +y / +z neighbour; the
// cell X = R/2 takes its voxel from the +x neighbour.  m: the row's bits.
template <int B>
__device__ __forceinline__ void col1_row(const ChunkRefs &c, const uint64_t *lds, const uint8_t *xc,
                                         uint8_t *dst, int Y, int Z, uint64_t m)
{
    constexpr int R = 1 << B, H = R / 2;
    static_assert(R == 64, "16-byte row pieces of a 64-voxel row");
    uint64_t pdz, pdy, pdx;
    l1_child_planes<B, false>(lds, Z, Y, pdz, pdy, pdx);
    const uint8_t *src[4];      // child row q = 2 dz + dy, or its uniform value
    uint32_t fill[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int y = 2 * Y + (q & 1), z = 2 * Z + (q >> 1);
        const uint8_t *p = nullptr;
        int u = 0;
        if (y < R && z < R) {
            p = c.own + (size_t)(z * R + y) * R;
            u = c.own_u;
        } else if (y == R && z < R) {
            u = c.nb_u[1];
            if (u < 0) p = c.nb[1] + (size_t)z * R * R;
        } else if (z == R && y < R) {
            u = c.nb_u[2];
            if (u < 0) p = c.nb[2] + (size_t)y * R;
        }
        src[q] = u < 0 ? p : nullptr;
        fill[q] = (uint32_t)(u < 0 ? 0 : u) * 0x01010101u;
    }
#pragma unroll
    for (int h = 0; h < 2; h++) {
        if (!((m >> (16 * h)) & 0xFFFFull)) continue;
        uint32_t w[4][8];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            if (src[q]) {
                const uint4 v0 = *reinterpret_cast<const uint4 *>(src[q] + 32 * h);
                const uint4 v1 = *reinterpret_cast<const uint4 *>(src[q] + 32 * h + 16);
                w[q][0] = v0.x; w[q][1] = v0.y; w[q][2] = v0.z; w[q][3] = v0.w;
                w[q][4] = v1.x; w[q][5] = v1.y; w[q][6] = v1.z; w[q][7] = v1.w;
            } else {
#pragma unroll
                for (int k = 0; k < 8; k++) w[q][k] = fill[q];
            }
        }
        // cell X = 16 h + i: children x = 2X + dx are bytes 2i + dx of the half
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const int X = 16 * h + i;
            const bool bz = (pdz >> X) & 1ull, by = (pdy >> X) & 1ull;
            const uint32_t bx = (uint32_t)(pdx >> X) & 1u;
            const uint32_t t1 = by ? w[3][i >> 1] : w[2][i >> 1];
            const uint32_t t0 = by ? w[1][i >> 1] : w[0][i >> 1];
            const uint32_t wd = bz ? t1 : t0;
            dst[X] = (uint8_t)(wd >> (16 * (i & 1) + 8 * bx));
        }
    }
    if ((m >> H) & 1ull) {
        const int dz = (int)((pdz >> H) & 1ull), dy = (int)((pdy >> H) & 1ull);
        dst[H] = xc[(2 * Z + dz) * R + 2 * Y + dy];
    }
}

// Emission of one chunk whose slice offsets are in LDS.  Records are ordered
// level, z, y, x, so each z-slice of each level is a contiguous range of the
// output starting at its slice offset.  Waves take slices from the work
// counter, with no workgroup barrier: lane y counts row y of the slice (plus
// the cell x = R of level 0), a wave scan gives each row its place, each lane
// writes its row's cells (y << 7 | x) to the wave's list, and then the wave
// writes the list's records with consecutive lanes on consecutive records,
// U records per lane in flight.  A slice with more than LCAP cells goes in
// batches of whole rows.  Colours are byte loads from the chunk (or from the
// neighbour planes for cells at coordinate R); the shadow bits read a window
// of the one map row the slice samples, staged in the wave's LDS area.
template <int B, int BS, bool V2, int U>
__device__ void emit_slices(const CullArgs &a, const ChunkRefs &c, uint64_t *lds, uint64_t *out)
{
    using L = Lay<B, BS, V2>;
    using G = Geo<B>;
    constexpr int R = G::R, N0 = G::N0;
    constexpr int LCAP = L::LCAP, SWIN = L::SWIN;
    const uint32_t ln = lane_id(), w = threadIdx.x >> 6;
    uint32_t *soff = reinterpret_cast<uint32_t *>(lds + L::SOFF);
    using LE = typename std::conditional<L::MIP, uint16_t, uint32_t>::type;
    const uint8_t *xc = reinterpret_cast<const uint8_t *>(lds + L::XCB);
    LE *list = reinterpret_cast<LE *>(lds + L::WAVES + (int)w * L::WAVE_WORDS);
    uint16_t *win = reinterpret_cast<uint16_t *>(list + LCAP);
    const int nitems = a.lod ? L::NSLICES : L::SLB1;
    auto take = [&]() -> uint32_t {
        uint32_t t = 0;
        if (ln == 0)
            t = __hip_atomic_fetch_add(&soff[L::ITEM_U32], 1u, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
        return uni((uint32_t)__shfl((int)t, 0, 64));
    };
    auto level_of = [](uint32_t t) {
        return (int)t < L::SLB1 ? 0 : (int)t < L::SLB2 ? 1 : (int)t < L::SLB3 ? 2
             : (int)t < L::SLB4 ? 3 : 4;
    };
    auto slice_of = [](uint32_t t, int lv) {
        return (int)t - (lv == 0 ? L::SLB0 : lv == 1 ? L::SLB1 : lv == 2 ? L::SLB2
                        : lv == 3 ? L::SLB3 : L::SLB4);
    };
    // The shadow-map row a slice samples (2R + 2 entries, lane e holds e,
    // e + 64, e + 128) is loaded into registers one item ahead, so its wait
    // overlaps the previous slice's work.
    constexpr int WQ = (SWIN + 63) / 64;
    auto fetch_win = [&](uint32_t t, uint32_t (&wv)[WQ]) {
#pragma unroll
        for (int q = 0; q < WQ; q++) wv[q] = 0;
        if (!a.map || (int)t >= nitems || soff[t + 1] == soff[t]) return;
        const int lv = level_of(t), sl = slice_of(t, lv);
        const uint32_t dd = lv ? (1u << lv) : 0u;
        const size_t wb = (size_t)a.sx * (c.org[2] + ((uint32_t)sl << lv) + dd) +
                          c.org[0] + c.org[1] + 2u * dd;
#pragma unroll
        for (int q = 0; q < WQ; q++)
            if ((int)ln + 64 * q < SWIN) wv[q] = a.map[wb + ln + 64 * q];
    };
    uint32_t nxt = take();
    uint32_t wv[WQ];
    fetch_win(nxt, wv);
#pragma unroll 1
    for (;;) {
        const uint32_t it = nxt;
        if ((int)it >= nitems) break;
        const int l = level_of(it), s = slice_of(it, l);
        const int nl = G::n(l);
        const uint32_t sbeg = soff[it], T = soff[it + 1] - sbeg;
        const uint32_t d = l ? (1u << l) : 0u;        // LOD samples at +(1 << l) (mesher.c:523-531)
        if (T != 0) {
            wave_sync();                               // the previous slice's window reads are done
            if (a.map) {
#pragma unroll
                for (int q = 0; q < WQ; q++)
                    if ((int)ln + 64 * q < SWIN) win[ln + 64 * q] = (uint16_t)wv[q];
            }
        }
        nxt = take();
        fetch_win(nxt, wv);
        if (T == 0) continue;
        const uint64_t *rows = l == 0 ? lds + L::V0 + s * N0
                                      : lds + L::P1 + (G::row_base(l) - G::RB1) + s * nl;
        uint64_t rb = 0;
        uint32_t xb = 0;
        if ((int)ln < nl) rb = rows[ln];
        if (l == 0 && s < R && (int)ln < R) xb = (uint32_t)(lds[L::XB + s] >> ln) & 1u;
        const uint32_t cnt = (uint32_t)__popcll(rb) + xb;
        uint32_t tot64;
        const uint32_t off = wave_excl_scan(cnt, tot64);
        const uint64_t rbx = nl > 64 ? rows[64] : 0ull;    // row y = 64 of level 0 (R = 64)
        uint64_t pdz = 0, pdy = 0, pdx = 0;     // level-1 child planes of lane's row
        uint64_t *o = out + sbeg;
        const uint32_t wz = c.org[2] + ((uint32_t)s << l);
#pragma unroll 1
        for (uint32_t b0 = 0; b0 < T;) {
            const uint32_t lim = b0 + LCAP;
            const bool pending = cnt && off >= b0;
            uint32_t b1 = wave_min(pending && off + cnt > lim ? off : 0xFFFFFFFFu);
            if (b1 == 0xFFFFFFFFu) b1 = T <= lim ? T : tot64;
            wave_sync();                               // the previous batch's list reads are done
            if (pending && off < b1) {
                uint32_t pos = off - b0;
                uint64_t m = rb;
                const uint32_t code = ln << 7;
                if (!L::MIP && l == 1) {
                    // level 1: the descent to the level-0 child for the whole row
                    // at once, as bit planes over X (the child's dz, dy, dx)
                    l1_child_planes<B, V2>(lds, s, (int)ln, pdz, pdy, pdx);
                    while (m) {
                        const uint32_t x = (uint32_t)__builtin_ctzll(m);
                        m &= m - 1ull;
                        const uint32_t sel = (uint32_t)((pdz >> x) & 1ull) |
                                             ((uint32_t)((pdy >> x) & 1ull) << 1) |
                                             ((uint32_t)((pdx >> x) & 1ull) << 2);
                        list[pos++] = (LE)(code | x | (sel << 16));
                    }
                } else {
                    while (m) {
                        const uint32_t x = (uint32_t)__builtin_ctzll(m);
                        m &= m - 1ull;
                        list[pos++] = (LE)(code | x);
                    }
                    if (xb) list[pos] = (LE)(code | (uint32_t)R);
                }
            }
            if (rbx && b1 == T && tot64 >= b0 && ((rbx >> ln) & 1ull))
                list[tot64 - b0 + rank_below(rbx)] = (LE)((64u << 7) | ln);
            wave_sync();
            const uint32_t nb = b1 - b0;
            uint64_t *ob = o + b0;
            if (l == 0) {
                // The colour loads of a round go out together, with no branch
                // around them: lanes past the end of the batch take its last
                // entry and store that entry's record again, to the same
                // place.  (With a branch per record the compiler waited for
                // each load in turn: loads and stores share one completion
                // counter here, so it cannot wait for a subset.)
                const uint32_t jmax = nb - 1u;
#pragma unroll 1
                for (uint32_t j0 = 0; j0 < nb; j0 += 64 * U) {
                    uint32_t col[U];
                    uint64_t rec[U];
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        const uint32_t j = j0 + (uint32_t)u * 64u + ln;
                        const uint32_t e = list[j < jmax ? j : jmax];
                        const int x = (int)(e & 127u), y = (int)((e >> 7) & 127u);
                        col[u] = cell_colour<B>(c, xc, x, y, s);
                        const uint32_t wy = c.org[1] + (uint32_t)y;
                        uint32_t sh = 0;
                        if (a.map) sh = !(win[x + y] <= wy && win[x + y + 1] <= wy);
                        uint32_t fl = 0;
                        if constexpr (V2) fl = face_flags<B, V2>(lds, x, y, s);
                        rec[u] = pack_record<V2>(c.org[0] + (uint32_t)x, wy, wz, 0u, sh, fl);
                    }
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        const uint32_t j = j0 + (uint32_t)u * 64u + ln;
                        __builtin_nontemporal_store(rec[u] | ((uint64_t)(col[u] & 0xFFu) << 48),
                                                    &ob[j < jmax ? j : jmax]);
                    }
                }
            } else {
                const uint8_t *colm = reinterpret_cast<const uint8_t *>(lds + L::COL) + L::cbase(l);
#pragma unroll 1
                for (uint32_t j0 = 0; j0 < nb; j0 += 64 * U) {
                    uint32_t col[U];
                    uint64_t rec[U];
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        const uint32_t j = j0 + (uint32_t)u * 64u + ln;
                        col[u] = 0;
                        rec[u] = 0;
                        if (j < nb) {
                            const uint32_t e = list[j];
                            const int X = (int)(e & 127u), Y = (int)((e >> 7) & 127u);
                            int qx = X, qy = Y, qz = s;
                            if constexpr (L::MIP) {
                                col[u] = colm[(s * nl + Y) * nl + X];
                            } else {
                                if (l == 1) {
                                    qz = 2 * s + (int)((e >> 16) & 1u);
                                    qy = 2 * Y + (int)((e >> 17) & 1u);
                                    qx = 2 * X + (int)((e >> 18) & 1u);
                                } else {
                                    descend_bf<B, V2>(lds, l, qx, qy, qz);
                                }
                                col[u] = cell_colour<B>(c, xc, qx, qy, qz);
                            }
                            const uint32_t wy = c.org[1] + ((uint32_t)Y << l);
                            const uint32_t ex = (uint32_t)(X + Y) << l;
                            uint32_t sh = 0;
                            if (a.map) sh = !(win[ex] <= wy + d && win[ex + 1] <= wy + d);
                            uint32_t fl = 0;
                            if constexpr (V2) fl = face_flags<B, V2>(lds, qx, qy, qz);
                            rec[u] = pack_record<V2>(c.org[0] + ((uint32_t)X << l), wy, wz, 0u,
                                                     sh, fl);
                        }
                    }
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        const uint32_t j = j0 + (uint32_t)u * 64u + ln;
                        if (j < nb) __builtin_nontemporal_store(rec[u] | ((uint64_t)(col[u] & 0xFFu) << 48), &ob[j]);
                    }
                }
            }
            b0 = b1;
        }
    }
}

// Work for one chunk; all threads of the workgroup call it.
template <int B, int BS, bool V2, bool MASK_OUT, int U>
__device__ void cull_chunk(const CullArgs &a, uint32_t k, uint32_t n, uint64_t *lds,
                           uint32_t *wsum, uint64_t *s_base)
{
    using G = Geo<B>;
    using L = Lay<B, BS, V2>;
    constexpr int R = G::R, N0 = G::N0;
    constexpr int NW = BS / 64;
    constexpr uint64_t RM = G::ROWMASK;

    uint32_t chunk = uni(a.ids[k]);
    const bool bad_id = chunk >= a.nchunks;
    if (bad_id) {
        if (threadIdx.x == 0) atomicOr(a.status, (uint32_t)VP_STATUS_BAD_ID);
        chunk = 0;   // read nothing of it: forced trivial below
    }
    const uint32_t mx = 1u << a.mb[0], my = 1u << a.mb[1], mz = 1u << a.mb[2];
    const uint32_t cc[3] = {chunk & (mx - 1u), (chunk >> a.mb[0]) & (my - 1u),
                            chunk >> (a.mb[0] + a.mb[1])};
    const uint32_t step[3] = {1u, mx, mx * my};
    const uint32_t lim[3] = {mx, my, mz};
    ChunkRefs c;
    c.own = a.voxels + ((size_t)chunk << (3 * B));
    c.own_u = a.uniform ? a.uniform[chunk] : -1;
    int nb_state[3];   // 0 air, 1 solid, -1 mixed
    for (int ax = 0; ax < 3; ax++) {
        c.org[ax] = cc[ax] << B;
        if (cc[ax] + 1u < lim[ax]) {
            const uint32_t nc = chunk + step[ax];
            c.nb[ax] = a.voxels + ((size_t)nc << (3 * B));
            c.nb_u[ax] = a.uniform ? a.uniform[nc] : -1;
        } else {
            c.nb[ax] = nullptr;
            c.nb_u[ax] = 0;             // the null chunk: all air (chunkset.c:92-98)
        }
        nb_state[ax] = c.nb_u[ax] < 0 ? -1 : (c.nb_u[ax] != 0);
    }

    // A uniform chunk whose three + neighbours are uniform in the same
    // air/solid sense has no exposed face in any cell it owns.
    const bool trivial = bad_id ||
                         (c.own_u >= 0 && nb_state[0] == (c.own_u != 0) &&
                          nb_state[1] == (c.own_u != 0) && nb_state[2] == (c.own_u != 0));

    uint32_t total = 0, lvl[5] = {0, 0, 0, 0, 0};
    if (!trivial) {
        // ---- 1. occupancy rows and neighbour planes ----
        // (and the slice counters, which phases 2-3 add into)
        for (int i = threadIdx.x; i < L::NSLICES; i += BS)
            reinterpret_cast<uint32_t *>(lds + L::SOFF)[i] = 0u;
        if (c.own_u >= 0) {
            const uint64_t v = c.own_u ? RM : 0ull;
            for (int r = threadIdx.x; r < R * R; r += BS) lds[L::S + r] = v;
        } else {
            load_rows<B, BS>(lds + L::S, c.own, R * R, R);
        }
        // +y plane: row (y=0, z) of the +y neighbour, one row per z
        if (c.nb_u[1] >= 0) {
            for (int z = threadIdx.x; z < R; z += BS) lds[L::HY + z] = c.nb_u[1] ? RM : 0ull;
        } else {
            load_rows<B, BS>(lds + L::HY, c.nb[1], R, R * R);
        }
        // +z plane: rows (y, z=0) of the +z neighbour, contiguous
        if (c.nb_u[2] >= 0) {
            for (int y = threadIdx.x; y < R; y += BS) lds[L::HZ + y] = c.nb_u[2] ? RM : 0ull;
        } else {
            load_rows<B, BS>(lds + L::HZ, c.nb[2], R, R);
        }
        // +x plane: voxel (0, y, z) of the +x neighbour -> HX[z] bit y, and
        // the byte itself at XCB[z R + y] (the colour of cell (R, y, z))
        uint8_t *xcb = reinterpret_cast<uint8_t *>(lds + L::XCB);
        if (c.nb_u[0] >= 0) {
            for (int z = threadIdx.x; z < R; z += BS) lds[L::HX + z] = c.nb_u[0] ? RM : 0ull;
            for (int i = threadIdx.x; i < R * R; i += BS) xcb[i] = (uint8_t)c.nb_u[0];
        } else {
            // every byte load first (no branch around them: a lane past the
            // plane reads its first byte again), then the ballots
            constexpr int NX = (R * R + BS - 1) / BS;
            uint32_t v[NX];
#pragma unroll
            for (int k = 0; k < NX; k++) {
                const int i = k * BS + (int)threadIdx.x;
                v[k] = c.nb[0][(size_t)(i < R * R ? i : 0) * R];
            }
            const int ln = (int)lane_id();
#pragma unroll
            for (int k = 0; k < NX; k++) {
                const int i = k * BS + (int)threadIdx.x;
                const bool act = i < R * R;
                if (act) xcb[i] = (uint8_t)v[k];
                const uint64_t bal = __ballot(act && v[k] != 0);
                if (act && (ln % R) == 0) lds[L::HX + i / R] = (bal >> ln) & RM;
            }
        }
        __syncthreads();

        // ---- 2. level-0 visibility (mesher.c:421-448 as row algebra) ----
        // Each thread computes its rows into registers; V0 overwrites S only
        // after the barrier, when every S row has been read.  The rows'
        // cells are added to their slice's counter.
        uint32_t *soff = reinterpret_cast<uint32_t *>(lds + L::SOFF);
        {
            constexpr int RPT = (R * R + BS - 1) / BS;
            uint64_t vrow[RPT];
#pragma unroll
            for (int i = 0; i < RPT; i++) {
                const int r = i * BS + (int)threadIdx.x;
                const bool act = r < R * R;
                uint32_t bx = 0;
                uint64_t v = 0;
                if (act) {
                    const int z = r >> B, y = r & (R - 1);
                    const uint64_t s = lds[L::S + r];
                    const uint64_t hx = (lds[L::HX + z] >> y) & 1ull;
                    const uint64_t px = (s >> 1) | (hx << (R - 1));
                    uint64_t open = ~px | (~(s << 1) & ~1ull);
                    open |= ~((y + 1 < R) ? lds[L::S + r + 1] : lds[L::HY + z]);
                    if (y > 0) open |= ~lds[L::S + r - 1];
                    open |= ~((z + 1 < R) ? lds[L::S + r + R] : lds[L::HZ + y]);
                    if (z > 0) open |= ~lds[L::S + r - R];
                    v = s & open & RM;
                    bx = (uint32_t)(hx & ~(s >> (R - 1)) & 1ull);
                }
                vrow[i] = v;
                const uint64_t bal = __ballot(bx);
                const int ln = (int)lane_id();
                if (act && (ln % R) == 0) lds[L::XB + (r >> B)] = (bal >> ln) & RM;
                uint32_t cv = act ? (uint32_t)__popcll(v) + bx : 0u;
                if constexpr (R == 64) {   // a wave's rows are one slice's
#pragma unroll
                    for (int d = 32; d >= 1; d >>= 1) cv += __shfl_xor(cv, d, 64);
                    if (act && ln == 0 && cv) atomicAdd(&soff[r >> B], cv);
                } else {
                    if (cv) atomicAdd(&soff[r >> B], cv);
                }
            }
            // cells (x, R, z): +y neighbour voxel exposed to this chunk's (x, R-1, z);
            // cells (x, y, R): +z neighbour voxel exposed to (x, y, R-1)
            uint64_t ey = 0, ez = 0;
            const int t = (int)threadIdx.x;
            if (t < R) {
                ey = lds[L::HY + t] & ~lds[L::S + t * R + (R - 1)] & RM;
                ez = lds[L::HZ + t] & ~lds[L::S + (R - 1) * R + t] & RM;
                if (ey) atomicAdd(&soff[t], (uint32_t)__popcll(ey));
                if (ez) atomicAdd(&soff[R], (uint32_t)__popcll(ez));
            }
            __syncthreads();
#pragma unroll
            for (int i = 0; i < RPT; i++) {
                const int r = i * BS + (int)threadIdx.x;
                if (r < R * R) lds[L::V0 + (r >> B) * N0 + (r & (R - 1))] = vrow[i];
            }
            if (t < R) {
                lds[L::V0 + t * N0 + R] = ey;
                lds[L::V0 + R * N0 + t] = ez;
            }
            if (t == 0) lds[L::V0 + R * N0 + R] = 0ull;
        }
        __syncthreads();

        // ---- 3. LOD bit pyramid ----
        if (a.lod) {
            {   // level 1 from level 0 (65-bit rows: V0 + the XB boundary bit)
                constexpr int n1 = G::n(1);
                for (int r = threadIdx.x; r < n1 * n1; r += BS) {
                    const int pz = r / n1, py = r % n1;
                    uint64_t acc = 0;
                    uint32_t bxa = 0;
                    for (int dz = 0; dz < 2; dz++) {
                        const int Z = 2 * pz + dz;
                        if (Z > R) continue;
                        for (int dy = 0; dy < 2; dy++) {
                            const int Y = 2 * py + dy;
                            if (Y > R) continue;
                            acc |= lds[L::V0 + Z * N0 + Y];
                            if (Z < R && Y < R) bxa |= (uint32_t)(lds[L::XB + Z] >> Y) & 1u;
                        }
                    }
                    const uint64_t row = compress_pairs(acc) | ((uint64_t)bxa << (R / 2));
                    lds[L::P1 + r] = row;
                    if (row) atomicAdd(&soff[L::SLB1 + pz], (uint32_t)__popcll(row));
                    if constexpr (L::MIP)
                        if (row)
                            col1_row<B>(c, lds, reinterpret_cast<const uint8_t *>(lds + L::XCB),
                                        reinterpret_cast<uint8_t *>(lds + L::COL) + r * n1, py, pz,
                                        row);
                }
            }
            __syncthreads();
#pragma unroll 1
            for (int l = 2; l <= 4; l++) {
                const int nc = G::n(l), np = G::n(l - 1);
                const int dst = L::P1 + (G::row_base(l) - G::RB1);
                const int src = L::P1 + (G::row_base(l - 1) - G::RB1);
                for (int r = threadIdx.x; r < nc * nc; r += BS) {
                    const int pz = r / nc, py = r % nc;
                    uint64_t ch[2][2];
#pragma unroll
                    for (int dz = 0; dz < 2; dz++)
#pragma unroll
                        for (int dy = 0; dy < 2; dy++) {
                            const int Z = 2 * pz + dz, Y = 2 * py + dy;
                            ch[dz][dy] = (Z < np && Y < np) ? lds[src + Z * np + Y] : 0ull;
                        }
                    const uint64_t row = compress_pairs(ch[0][0] | ch[0][1] | ch[1][0] | ch[1][1]);
                    lds[dst + r] = row;
                    if (row)
                        atomicAdd(&soff[(l == 2 ? L::SLB2 : l == 3 ? L::SLB3 : L::SLB4) + pz],
                                  (uint32_t)__popcll(row));
                    if constexpr (L::MIP) {
                        // colour of cell X: the level-(l-1) colour of its greatest
                        // non-zero child, which is where descending from X ends
                        if (row) {
                            uint64_t pdz, pdy, pdx;
                            lod_child_planes(ch, pdz, pdy, pdx);
                            const uint8_t *cs = reinterpret_cast<const uint8_t *>(lds + L::COL) +
                                                L::cbase(l - 1);
                            uint8_t *cd = reinterpret_cast<uint8_t *>(lds + L::COL) + L::cbase(l) +
                                          r * nc;
                            uint64_t mm = row;
                            while (mm) {
                                const int X = __builtin_ctzll(mm);
                                mm &= mm - 1ull;
                                const int dz = (int)((pdz >> X) & 1ull), dy = (int)((pdy >> X) & 1ull),
                                          dx = (int)((pdx >> X) & 1ull);
                                cd[X] = cs[((2 * pz + dz) * np + 2 * py + dy) * np + 2 * X + dx];
                            }
                        }
                    }
                }
                __syncthreads();
            }
        } else {
            for (int r = threadIdx.x; r < G::ROWS - G::RB1; r += BS) lds[L::P1 + r] = 0ull;
            __syncthreads();
        }

        // ---- 4. slice counts (added up in phases 2-3) -> scan -> slice offsets ----
        // Records are ordered by level, z, y, x, so every z-slice of every
        // level owns a contiguous range: one offset per slice is enough.
        {
            const int w = (int)(threadIdx.x >> 6);
            const uint32_t ln = lane_id();
            if (w == 0) {   // exclusive scan of the slice counts in one wave
                constexpr int KS = (L::NSLICES + 63) / 64;
                uint32_t part = 0;
                for (int j = 0; j < KS; j++) {
                    const int sl = (int)ln * KS + j;
                    if (sl < L::NSLICES) part += soff[sl];
                }
                uint32_t x = part;
#pragma unroll
                for (int d = 1; d < 64; d <<= 1) {
                    const uint32_t y = __shfl_up(x, d, 64);
                    if (ln >= (uint32_t)d) x += y;
                }
                uint32_t run = x - part;
                for (int j = 0; j < KS; j++) {
                    const int sl = (int)ln * KS + j;
                    if (sl < L::NSLICES) {
                        const uint32_t cv = soff[sl];
                        soff[sl] = run;
                        run += cv;
                    }
                }
                if (ln == 63) soff[L::NSLICES] = x;
                if (ln == 0) soff[L::ITEM_U32] = 0u;   // emission work counter
            }
            __syncthreads();
        }
        total = uni(soff[L::NSLICES]);
        {
            const int slb[6] = {L::SLB0, L::SLB1, L::SLB2, L::SLB3, L::SLB4, L::NSLICES};
            for (int l = 0; l < 5; l++) lvl[l] = soff[slb[l + 1]] - soff[slb[l]];
        }
    }

    if constexpr (MASK_OUT) {
        // dense level-0 mask, (R+1)^3 bytes, zero where no face
        uint8_t *m = a.masks + (size_t)k * N0 * N0 * N0;
        for (int idx = threadIdx.x; idx < N0 * N0 * N0; idx += BS) {
            const int x = idx % N0, y = (idx / N0) % N0, z = idx / (N0 * N0);
            uint8_t v = 0;
            if (!trivial && l0_bit<B, V2>(lds, x, y, z)) v = l0_colour<B>(c, x, y, z);
            m[idx] = v;
        }
        __syncthreads();
        return;
    }

    // ---- publish the chunk's size, look back for its offset ----
    const uint64_t items = 4ull * total;
    if (threadIdx.x < 64) {
        uint64_t excl = 0;
        if (!a.single) {
            if (k == 0) {
                if (threadIdx.x == 0) st_relaxed(a.state, kFlagPre | items);
            } else {
                if (threadIdx.x == 0) st_relaxed(a.state + k, kFlagAgg | items);
                int64_t top = (int64_t)k - 1;       // nearest predecessor not yet summed
                const uint32_t ln = lane_id();
                for (;;) {
                    const int64_t j = top - (int64_t)ln;
                    uint64_t s = (j >= 0) ? ld_relaxed(a.state + j) : kFlagPre;
                    const uint64_t f = s >> 62;
                    const uint64_t pre = __ballot(f == 2);
                    const uint64_t zero = __ballot(f == 0);
                    // lanes up to and including the first prefix must all be ready
                    const uint64_t need = pre ? (pre & (~pre + 1ull)) * 2ull - 1ull : ~0ull;
                    if (zero & need) {
                        __builtin_amdgcn_s_sleep(1);
                        continue;
                    }
                    uint64_t v = (need >> ln) & 1ull ? (s & kValMask) : 0ull;
#pragma unroll
                    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
                    excl += v;
                    if (pre) break;
                    top -= 64;
                }
                if (threadIdx.x == 0) st_relaxed(a.state + k, kFlagPre | (excl + items));
            }
        }
        if (threadIdx.x == 0) {
            *s_base = excl;
            a.offsets[k] = excl;
            if (k + 1 == n) a.offsets[n] = excl + items;
            for (int l = 0; l < 5; l++) a.level_counts[(size_t)k * 5 + l] = 4u * lvl[l];
            if (!a.count_only && total && excl / 4 + total > a.cap)
                atomicOr(a.status, (uint32_t)VP_STATUS_OVERFLOW);
        }
    }
    __syncthreads();
    if (trivial || a.count_only || total == 0) return;
    const uint64_t base = uni(*s_base) / 4;
    if (base + total > a.cap) return;

    // ---- 5. emission: waves take z-slices from the work counter ----
    emit_slices<B, BS, V2, U>(a, c, lds, a.splats + base);
}

// LDS (about 67 KiB at 512 threads for R = 64) allows two workgroups per CU;
// ask the register allocator for the matching occupancy (512 threads: 4 waves
// per SIMD, at most 128 VGPRs; 1024 threads: one workgroup of 4 waves per SIMD).
// U: records each lane prepares before its stores in the emission.
template <int B, int BS, bool V2, bool MASK_OUT, int U>
__global__ void __launch_bounds__(BS)
__attribute__((amdgpu_waves_per_eu(BS == 256 ? 2 : 4))) k_cull(CullArgs a)
{
    __shared__ uint64_t lds[Lay<B, BS, V2>::TOTAL];
    __shared__ uint32_t wsum[BS / 64 + 1];
    __shared__ uint32_t s_ticket;
    __shared__ uint64_t s_base;
    if (a.single) {
        cull_chunk<B, BS, V2, MASK_OUT, U>(a, 0u, 1u, lds, wsum, &s_base);
        return;
    }
    if (MASK_OUT) {   // make_mask: one workgroup per listed chunk, no offsets
        for (uint32_t k = blockIdx.x; k < a.n; k += gridDim.x) {
            cull_chunk<B, BS, V2, MASK_OUT, U>(a, k, a.n, lds, wsum, &s_base);
            __syncthreads();
        }
        return;
    }
    const uint32_t n = a.d_n ? min(*a.d_n, a.n) : a.n;
    for (;;) {
        if (threadIdx.x == 0) s_ticket = atomicAdd(a.ticket, 1u);
        __syncthreads();
        const uint32_t k = uni(s_ticket);
        if (k >= n) break;
        cull_chunk<B, BS, V2, MASK_OUT, U>(a, k, n, lds, wsum, &s_base);
        __syncthreads();
    }
}

__global__ void k_cull_reset(uint32_t *ticket, uint64_t *state, uint32_t n, uint64_t *offsets)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        ticket[0] = 0u;
        offsets[0] = 0ull;
    }
    for (uint32_t j = i; j < n; j += gridDim.x * blockDim.x) state[j] = 0ull;
}

// A kernel instantiation: launch it, or ask how many of its workgroups fit
// on one CU (registers and LDS).
struct CullLaunch {
    const CullArgs *a;      // launch when set, else query residency
    int grid;
    hipStream_t st;
    int *per_cu;
};

template <int B, int BS, bool V2, bool MASK_OUT, int U>
static hipError_t launch_cull(const CullLaunch &L)
{
    if (!L.a) {
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(
            L.per_cu, reinterpret_cast<const void *>(&k_cull<B, BS, V2, MASK_OUT, U>), BS, 0);
    }
    hipLaunchKernelGGL((k_cull<B, BS, V2, MASK_OUT, U>), dim3(L.grid), dim3(BS), 0, L.st, *L.a);
    return hipGetLastError();
}

// Records in flight per lane: 2 or 4 for R = 64 (a tuning choice), 4 below.
template <int B, int BS, bool V2, bool MASK_OUT>
static hipError_t dispatch_u(const CullLaunch &L, int u)
{
    if constexpr (B == 6 && !MASK_OUT)
        if (u == 2) return launch_cull<B, BS, V2, MASK_OUT, 2>(L);
    return launch_cull<B, BS, V2, MASK_OUT, 4>(L);
}

template <int B, bool MASK_OUT>
static hipError_t dispatch_bs(const CullLaunch &L, bool v2, int bs, int u)
{
    if (v2) {
        switch (bs) {
        case 256: return dispatch_u<B, 256, true, MASK_OUT>(L, u);
        case 1024: return dispatch_u<B, 1024, true, MASK_OUT>(L, u);
        default: return dispatch_u<B, 512, true, MASK_OUT>(L, u);
        }
    }
    switch (bs) {
    case 256: return dispatch_u<B, 256, false, MASK_OUT>(L, u);
    case 1024: return dispatch_u<B, 1024, false, MASK_OUT>(L, u);
    default: return dispatch_u<B, 512, false, MASK_OUT>(L, u);
    }
}

template <bool MASK_OUT>
static hipError_t dispatch(int b, const CullLaunch &L, bool v2, int bs, int u)
{
    switch (b) {
    case 2: return dispatch_bs<2, MASK_OUT>(L, v2, bs, u);
    case 3: return dispatch_bs<3, MASK_OUT>(L, v2, bs, u);
    case 4: return dispatch_bs<4, MASK_OUT>(L, v2, bs, u);
    case 5: return dispatch_bs<5, MASK_OUT>(L, v2, bs, u);
    default: return dispatch_bs<6, MASK_OUT>(L, v2, bs, u);
    }
}

static int check_world(const vp_world_desc *w)
{
    if (!w) return VP_ERR_ARG;
    if (w->root_bitw < 2 || w->root_bitw > 6) return VP_ERR_UNSUPPORTED;
    if (w->max_bitw[0] > 16 || w->max_bitw[1] > 16 || w->max_bitw[2] > 16) return VP_ERR_ARG;
    if ((int)w->max_bitw[0] + w->max_bitw[1] + w->max_bitw[2] > 32) return VP_ERR_ARG;
    // world coordinates must fit the int16 records (README: 16k^3 cap)
    for (int a = 0; a < 3; a++)
        if (((1u << w->max_bitw[a]) << w->root_bitw) > 32768u) return VP_ERR_UNSUPPORTED;
    return VP_OK;
}

}  // namespace vp

using namespace vp;

namespace vp {

int cull_mesh_check(const vp_world_desc *world, const uint8_t *d_voxels,
                    const uint32_t *d_chunk_ids, uint32_t flags, const vp_splat_out *out)
{
    if (!d_voxels || !d_chunk_ids || !out || !out->d_chunk_offsets || !out->d_level_counts)
        return VP_ERR_ARG;
    int rc = check_world(world);
    if (rc) return rc;
    if (!(flags & VP_CM_COUNT_ONLY) && !out->d_splats && out->capacity_records) return VP_ERR_ARG;
    return VP_OK;
}

int cull_mesh_launch(vp_ctx *ctx, const vp_world_desc *world, const uint8_t *d_voxels,
                     const int16_t *d_uniform, const vp_shadow_desc *shadow,
                     const uint32_t *d_chunk_ids, uint32_t n, const uint32_t *d_n,
                     uint32_t flags, const vp_splat_out *out, uint32_t *ticket,
                     uint64_t *state, hipStream_t st)
{
    CullArgs a = {};
    a.voxels = d_voxels;
    a.uniform = d_uniform;
    a.map = shadow ? shadow->d_map : nullptr;
    a.sx = shadow ? shadow->size_x : 0u;
    a.ids = d_chunk_ids;
    a.n = n;
    a.d_n = d_n;
    for (int i = 0; i < 3; i++) a.mb[i] = world->max_bitw[i];
    a.lod = (flags & VP_CM_LOD) ? 1 : 0;
    a.count_only = (flags & VP_CM_COUNT_ONLY) ? 1 : 0;
    a.single = (n == 1 && !d_n) ? 1 : 0;
    a.splats = static_cast<uint64_t *>(out->d_splats);
    a.cap = out->d_splats ? out->capacity_records : 0;
    a.offsets = out->d_chunk_offsets;
    a.level_counts = out->d_level_counts;
    a.status = ctx->d_status;
    a.nchunks = (uint32_t)(1ull << (world->max_bitw[0] + world->max_bitw[1] + world->max_bitw[2]));
    const vp_remesh_shape_i sh = remesh_shape(ctx, world);
    int grid = 1, bs = sh.block;
    // one 64^3 chunk alone: a 1024-thread workgroup (16 waves on one CU)
    // finishes it in 0.70x the time of the batched path's 512-thread shape
    if (a.single && world->root_bitw == 6) bs = 1024;
    if (!a.single) {
        if (!ticket || !state) return VP_ERR_ARG;
        a.ticket = ticket;
        a.state = state;
        const int resident = ctx->num_cus * sh.wg_per_cu;
        grid = (int)((uint64_t)n < (uint64_t)resident ? n : (uint32_t)resident);
        const int rgrid = (int)((n + 255u) / 256u);
        hipLaunchKernelGGL(k_cull_reset, dim3(rgrid < 1024 ? rgrid : 1024), dim3(256), 0, st,
                           a.ticket, a.state, n, a.offsets);
        VP_HIP_TRY(hipGetLastError());
    }
    const CullLaunch L = {&a, grid, st, nullptr};
    VP_HIP_TRY(dispatch<false>(world->root_bitw, L, (flags & VP_CM_RECORD_V2) != 0, bs, sh.u));
    return VP_OK;
}

}  // namespace vp

extern "C" int vp_gpu_cull_mesh(vp_ctx *ctx, const vp_world_desc *world,
                                const uint8_t *d_voxels, const int16_t *d_uniform,
                                const vp_shadow_desc *shadow,
                                const uint32_t *d_chunk_ids, uint32_t n, const uint32_t *d_n,
                                uint32_t flags, const vp_splat_out *out, void *stream)
{
    if (!ctx) return VP_ERR_ARG;
    int rc = cull_mesh_check(world, d_voxels, d_chunk_ids, flags, out);
    if (rc) return rc;
    if (n == 0) return VP_OK;
    hipStream_t st = as_stream(stream);
    if (n == 1 && !d_n)   // the plain per-chunk path needs no look-back state
        return cull_mesh_launch(ctx, world, d_voxels, d_uniform, shadow, d_chunk_ids, n, d_n,
                                flags, out, nullptr, nullptr, st);
    // the context's look-back state is grown to n words under the lock and
    // used by the launch; calls on one context run one stream at a time
    std::lock_guard<std::mutex> g(ctx->mu);
    rc = ensure_state(ctx, n);
    if (rc) return rc;
    return cull_mesh_launch(ctx, world, d_voxels, d_uniform, shadow, d_chunk_ids, n, d_n, flags,
                            out, ctx->d_ticket, ctx->d_state, st);
}

extern "C" int vp_gpu_make_mask(vp_ctx *ctx, const vp_world_desc *world,
                                const uint8_t *d_voxels, const int16_t *d_uniform,
                                const uint32_t *d_chunk_ids, uint32_t n,
                                uint8_t *d_masks, void *stream)
{
    if (!ctx || !d_voxels || !d_chunk_ids || !d_masks) return VP_ERR_ARG;
    int rc = check_world(world);
    if (rc) return rc;
    if (n == 0) return VP_OK;
    CullArgs a = {};
    a.voxels = d_voxels;
    a.uniform = d_uniform;
    a.ids = d_chunk_ids;
    a.n = n;
    for (int i = 0; i < 3; i++) a.mb[i] = world->max_bitw[i];
    a.lod = 0;
    a.masks = d_masks;
    a.status = ctx->d_status;
    a.nchunks = (uint32_t)(1ull << (world->max_bitw[0] + world->max_bitw[1] + world->max_bitw[2]));
    const int resident = ctx->num_cus * 2;
    const int grid = (int)(n < (uint32_t)resident ? n : (uint32_t)resident);
    const CullLaunch L = {&a, grid, as_stream(stream), nullptr};
    VP_HIP_TRY(dispatch<true>(world->root_bitw, L, false, 512, 4));
    return VP_OK;
}

extern "C" int vp_ctx_remesh_resident(vp_ctx *ctx, const vp_world_desc *world, uint32_t flags,
                                      const vp_remesh_shape *shape, int *per_cu)
{
    if (!ctx || !shape || !per_cu) return VP_ERR_ARG;
    int rc = check_world(world);
    if (rc) return rc;
    const int b = shape->block, u = shape->records_in_flight;
    if ((b != 256 && b != 512 && b != 1024) || (u != 2 && u != 4)) return VP_ERR_ARG;
    *per_cu = 0;
    const CullLaunch L = {nullptr, 0, nullptr, per_cu};
    VP_HIP_TRY(dispatch<false>(world->root_bitw, L, (flags & VP_CM_RECORD_V2) != 0, b, u));
    return VP_OK;
}

// ---- standalone LOD step on dense masks (chunk_mask_downsample) ----------
namespace vp {
__global__ void k_lod_mip(int R, int level, const uint8_t *src, uint8_t *dst, uint32_t n)
{
    const int N0 = R + 1, cells = N0 * N0 * N0;
    const int nc = (R >> level) + 1, np = (R >> (level - 1)) + 1;
    for (uint32_t k = blockIdx.y; k < n; k += gridDim.y) {
        const uint8_t *s = src + (size_t)k * cells;
        uint8_t *d = dst + (size_t)k * cells;
        for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < cells;
             idx += gridDim.x * blockDim.x) {
            const int x = idx % N0, y = (idx / N0) % N0, z = idx / (N0 * N0);
            uint8_t v = 0;
            if (x < nc && y < nc && z < nc) {
                // last write of the z,y,x scan = greatest non-zero child
                for (int dd = 7; dd >= 0 && !v; --dd) {
                    const int X = 2 * x + (dd & 1), Y = 2 * y + ((dd >> 1) & 1),
                              Z = 2 * z + (dd >> 2);
                    if (X < np && Y < np && Z < np) v = s[(Z * N0 + Y) * N0 + X];
                }
            }
            d[idx] = v;
        }
    }
}
}  // namespace vp

extern "C" int vp_gpu_lod_mip(vp_ctx *ctx, uint32_t root_bitw, int level,
                              const uint8_t *d_src, uint8_t *d_dst, uint32_t n, void *stream)
{
    if (!ctx || !d_src || !d_dst || level < 1 || level > 4) return VP_ERR_ARG;
    if (root_bitw < 1 || root_bitw > 8) return VP_ERR_ARG;
    if (d_src == d_dst) return VP_ERR_ARG;
    if (n == 0) return VP_OK;
    const int R = 1 << root_bitw;
    const int cells = (R + 1) * (R + 1) * (R + 1);
    int gx = (cells + 255) / 256;
    if (gx > 64) gx = 64;
    const uint32_t gy = n < 65535u ? n : 65535u;
    hipLaunchKernelGGL(k_lod_mip, dim3(gx, gy), dim3(256), 0, as_stream(stream), R, level, d_src,
                       d_dst, n);
    VP_HIP_TRY(hipGetLastError());
    return VP_OK;
}
