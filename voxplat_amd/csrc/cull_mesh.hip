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
//  5. emission: one wave per row, lane = x; the row word is the wave's ballot,
//     so mbcnt gives each lane its slot and the stores are contiguous.
#include "vp_internal.h"

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

// LDS layout in 64-bit words.  Once visibility is built the occupancy rows
// are dead (unless v2 face flags still read them) and their space holds the
// emission scratch: slice offsets, the staged slice pair, the row offsets of
// the rows being expanded and the list of their cells.
template <int B, bool V2>
struct Lay {
    using G = Geo<B>;
    static constexpr int R = G::R, N0 = G::N0;
    static constexpr int S = 0;
    static constexpr int HX = S + R * R;
    static constexpr int HY = HX + R;
    static constexpr int HZ = HY + R;
    static constexpr int V0 = HZ + R;
    static constexpr int XB = V0 + N0 * N0;
    static constexpr int P1 = XB + R;               // P1..P4 contiguous, row order
    static constexpr int PEND = P1 + (G::ROWS - G::RB1);
    // z-slices of every level, level-major: slice base of level l
    static constexpr int SLB0 = 0, SLB1 = G::n(0), SLB2 = SLB1 + G::n(1), SLB3 = SLB2 + G::n(2),
                         SLB4 = SLB3 + G::n(3), NSLICES = SLB4 + G::n(4);
    // staged slice pair: two z-slices, +x column and +y row bytes, 3 shadow rows
    static constexpr int SW = 2 * R + 8;              // shadow entries per staged row
    static constexpr int ST_BYTES = 2 * R * R + 4 * R + 3 * SW * 2;
    static constexpr int maxi(int a, int b) { return a > b ? a : b; }
    static constexpr int al16(int v) { return (v + 15) & ~15; }
    static constexpr int PR_ROWS = maxi(2 * N0 + G::n(1), G::n(2) * G::n(2));
    static constexpr int LIST_N = maxi(2 * N0 * N0 + G::n(1) * G::n(1),
                                       G::n(2) * G::n(2) * G::n(2));
    static constexpr int O_STG = 0;
    static constexpr int O_SOFF = al16(ST_BYTES);
    static constexpr int O_ROWP = O_SOFF + al16((NSLICES + 1) * 4);
    static constexpr int O_LIST = O_ROWP + al16(PR_ROWS * 2);
    static constexpr int SCR_BYTES = O_LIST + al16(LIST_N * 2);
    static constexpr bool ALIAS = !V2 && SCR_BYTES <= R * R * 8;
    static constexpr int SCR = ALIAS ? S : PEND;
    static constexpr int TOTAL = ALIAS ? PEND : PEND + SCR_BYTES / 8;
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
    using L = Lay<B, V2>;
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
    using L = Lay<B, V2>;
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

// Face-exposure flags of level-0 cell q (v2 records): bit 2a = +a face open,
// bit 2a+1 = -a face open, as the checks of mesher.c:427-445 found them.
template <int B, bool V2>
__device__ __forceinline__ uint32_t face_flags(const uint64_t *lds, int x, int y, int z)
{
    using L = Lay<B, V2>;
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

// ---- emission helpers --------------------------------------------------------
// Registers carrying one slice pair from global memory to the LDS staging.
template <int B, int BS>
struct PairRegs {
    static constexpr int R = 1 << B;
    static constexpr int NSL = (2 * R * R + 15) / 16;   // 16-byte pieces of two slices
    static constexpr int NXY = 4 * R;                   // +x column and +y row bytes
    static constexpr int NSH = 3 * (2 * R + 8);         // shadow entries
    static constexpr int KSL = (NSL + BS - 1) / BS;
    static constexpr int KXY = (NXY + BS - 1) / BS;
    static constexpr int KSH = (NSH + BS - 1) / BS;
    uint4 sl[KSL];
    uint32_t xy[KXY];
    uint32_t sh[KSH];
};

__device__ __forceinline__ uint4 splat4(uint32_t v)
{
    const uint32_t w = (v & 0xFFu) * 0x01010101u;
    return make_uint4(w, w, w, w);
}

// Issue the loads of pair (z0, z0+1); slice z = R is the +z neighbour's z=0 plane.
template <int B, int BS>
__device__ __forceinline__ void pair_load(const CullArgs &a, const ChunkRefs &c, int z0,
                                          PairRegs<B, BS> &pr)
{
    constexpr int R = 1 << B, SL = R * R;
    using P = PairRegs<B, BS>;
#pragma unroll
    for (int j = 0; j < P::KSL; j++) {
        const int i = (int)threadIdx.x + j * BS;
        uint4 v = make_uint4(0, 0, 0, 0);
        if (i < P::NSL) {
            const int byte = i * 16, sidx = byte / SL, off = byte % SL, z = z0 + sidx;
            if (SL >= 16) {
                if (z < R) {
                    v = c.own_u >= 0 ? splat4((uint32_t)c.own_u)
                                     : *reinterpret_cast<const uint4 *>(c.own + (size_t)z * SL + off);
                } else if (z == R) {
                    v = c.nb_u[2] >= 0 ? splat4((uint32_t)c.nb_u[2])
                                       : *reinterpret_cast<const uint4 *>(c.nb[2] + off);
                }
            } else {   // slices smaller than 16 bytes (R = 2): byte by byte
                uint32_t wv[4] = {0, 0, 0, 0};
                for (int q = 0; q < 16; q++) {
                    const int bb = byte + q, zz = z0 + bb / SL, oo = bb % SL;
                    uint32_t x = 0;
                    if (zz < R) x = c.own_u >= 0 ? (uint32_t)c.own_u : c.own[(size_t)zz * SL + oo];
                    else if (zz == R) x = c.nb_u[2] >= 0 ? (uint32_t)c.nb_u[2] : c.nb[2][oo];
                    wv[q >> 2] |= (x & 0xFFu) << (8 * (q & 3));
                }
                v = make_uint4(wv[0], wv[1], wv[2], wv[3]);
            }
        }
        pr.sl[j] = v;
    }
#pragma unroll
    for (int j = 0; j < P::KXY; j++) {
        const int i = (int)threadIdx.x + j * BS;
        uint32_t v = 0;
        if (i < P::NXY) {
            const int kind = i / (2 * R), sidx = (i % (2 * R)) / R, e = i % R, z = z0 + sidx;
            if (z < R) {
                if (kind == 0)   // voxel (0, e, z) of the +x neighbour
                    v = c.nb_u[0] >= 0 ? (uint32_t)c.nb_u[0] : c.nb[0][(size_t)(z * R + e) * R];
                else             // voxel (e, 0, z) of the +y neighbour
                    v = c.nb_u[1] >= 0 ? (uint32_t)c.nb_u[1] : c.nb[1][(size_t)(z * R) * R + e];
            }
        }
        pr.xy[j] = v;
    }
#pragma unroll
    for (int j = 0; j < P::KSH; j++) {
        const int i = (int)threadIdx.x + j * BS;
        uint32_t v = 0;
        if (i < P::NSH && a.map) {
            const int SWn = 2 * R + 8, row = i / SWn, e = i % SWn;
            v = a.map[(size_t)a.sx * (c.org[2] + z0 + row) + (c.org[0] + c.org[1]) + e];
        }
        pr.sh[j] = v;
    }
}

template <int B, int BS>
__device__ __forceinline__ void pair_store(const PairRegs<B, BS> &pr, uint8_t *stg)
{
    constexpr int R = 1 << B;
    using P = PairRegs<B, BS>;
    constexpr int XC = 2 * R * R, SH = XC + 4 * R;
#pragma unroll
    for (int j = 0; j < P::KSL; j++) {
        const int i = (int)threadIdx.x + j * BS;
        if (i < P::NSL) *reinterpret_cast<uint4 *>(stg + i * 16) = pr.sl[j];
    }
#pragma unroll
    for (int j = 0; j < P::KXY; j++) {
        const int i = (int)threadIdx.x + j * BS;
        if (i < P::NXY) stg[XC + i] = (uint8_t)pr.xy[j];
    }
#pragma unroll
    for (int j = 0; j < P::KSH; j++) {
        const int i = (int)threadIdx.x + j * BS;
        if (i < P::NSH) reinterpret_cast<uint16_t *>(stg + SH)[i] = (uint16_t)pr.sh[j];
    }
}

// Colour of level-0 cell (x, y, z) of the staged pair z0, z0+1.
template <int B>
__device__ __forceinline__ uint32_t staged_colour(const uint8_t *stg, int x, int y, int z, int z0)
{
    constexpr int R = 1 << B;
    const int sidx = z - z0;
    if (x == R) return stg[2 * R * R + sidx * R + y];
    if (y == R) return stg[2 * R * R + 2 * R + sidx * R + x];
    return stg[sidx * R * R + y * R + x];
}

// shadow_sample on a staged row: entry e is map[sx*(oz+z0+row) + ox+oy + e]
template <int B>
__device__ __forceinline__ uint32_t staged_shadow(const uint8_t *stg, int row, int e, uint32_t wy)
{
    constexpr int R = 1 << B;
    const uint16_t *sh = reinterpret_cast<const uint16_t *>(stg + 2 * R * R + 4 * R);
    const uint16_t *rw = sh + row * (2 * R + 8);
    return !(rw[e] <= wy && rw[e + 1] <= wy);
}

// Exclusive scan of the counts of rows [0, nrows) into rowp (u16, relative
// to row 0); each thread sums a contiguous block of rows.  Returns the total.
template <int BS, typename F>
__device__ __forceinline__ uint32_t scan_rows(int nrows, F count, uint16_t *rowp, uint32_t *wsum)
{
    const int K = (nrows + BS - 1) / BS;
    const int r0 = (int)threadIdx.x * K;
    uint32_t mine = 0;
    for (int j = 0; j < K; j++)
        if (r0 + j < nrows) mine += count(r0 + j);
    uint32_t tot = 0;
    uint32_t run = block_excl_scan<BS>(mine, wsum, tot);
    for (int j = 0; j < K; j++)
        if (r0 + j < nrows) {
            rowp[r0 + j] = (uint16_t)run;
            run += count(r0 + j);
        }
    __syncthreads();
    return tot;
}

// Work for one chunk; all threads of the workgroup call it.
template <int B, int BS, bool V2, bool MASK_OUT>
__device__ void cull_chunk(const CullArgs &a, uint32_t k, uint32_t n, uint64_t *lds,
                           uint32_t *wsum, uint64_t *s_base)
{
    using G = Geo<B>;
    using L = Lay<B, V2>;
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
        // +x plane: voxel (0, y, z) of the +x neighbour -> HX[z] bit y
        if (c.nb_u[0] >= 0) {
            for (int z = threadIdx.x; z < R; z += BS) lds[L::HX + z] = c.nb_u[0] ? RM : 0ull;
        } else {
#pragma unroll 1
            for (int base = 0; base < R * R; base += BS) {
                const int i = base + (int)threadIdx.x;
                const bool act = i < R * R;
                const uint32_t bit = act ? (c.nb[0][(size_t)i * R] != 0) : 0u;
                const uint64_t bal = __ballot(bit);
                const int ln = (int)lane_id();
                if (act && (ln % R) == 0) lds[L::HX + i / R] = (bal >> ln) & RM;
            }
        }
        __syncthreads();

        // ---- 2. level-0 visibility (mesher.c:421-448 as row algebra) ----
#pragma unroll 1
        for (int base = 0; base < R * R; base += BS) {
            const int r = base + (int)threadIdx.x;
            const bool act = r < R * R;
            uint32_t bx = 0;
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
                lds[L::V0 + z * N0 + y] = s & open & RM;
                bx = (uint32_t)(hx & ~(s >> (R - 1)) & 1ull);
            }
            const uint64_t bal = __ballot(bx);
            const int ln = (int)lane_id();
            if (act && (ln % R) == 0) lds[L::XB + (r >> B)] = (bal >> ln) & RM;
        }
        for (int i = threadIdx.x; i < R; i += BS) {
            // cells (x, R, z): +y neighbour voxel exposed to this chunk's (x, R-1, z)
            lds[L::V0 + i * N0 + R] = lds[L::HY + i] & ~lds[L::S + i * R + (R - 1)] & RM;
            // cells (x, y, R): +z neighbour voxel exposed to (x, y, R-1)
            lds[L::V0 + R * N0 + i] = lds[L::HZ + i] & ~lds[L::S + (R - 1) * R + i] & RM;
        }
        if (threadIdx.x == 0) lds[L::V0 + R * N0 + R] = 0ull;
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
                    lds[L::P1 + r] = compress_pairs(acc) | ((uint64_t)bxa << (R / 2));
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
                    uint64_t acc = 0;
                    for (int dz = 0; dz < 2; dz++) {
                        const int Z = 2 * pz + dz;
                        if (Z >= np) continue;
                        for (int dy = 0; dy < 2; dy++) {
                            const int Y = 2 * py + dy;
                            if (Y >= np) continue;
                            acc |= lds[src + Z * np + Y];
                        }
                    }
                    lds[dst + r] = compress_pairs(acc);
                }
                __syncthreads();
            }
        } else {
            for (int r = threadIdx.x; r < G::ROWS - G::RB1; r += BS) lds[L::P1 + r] = 0ull;
            __syncthreads();
        }

        // ---- 4. per-slice counts -> scan -> slice offsets ----
        // Records are ordered by level, z, y, x, so every z-slice of every
        // level owns a contiguous range: one offset per slice is enough.
        uint32_t *soff = reinterpret_cast<uint32_t *>(
            reinterpret_cast<uint8_t *>(lds + L::SCR) + L::O_SOFF);
        {
            const int w = (int)(threadIdx.x >> 6);
            const uint32_t ln = lane_id();
            for (int sl = w; sl < L::NSLICES; sl += NW) {
                const int l = sl < L::SLB1 ? 0 : sl < L::SLB2 ? 1 : sl < L::SLB3 ? 2
                            : sl < L::SLB4 ? 3 : 4;
                const int nl = G::n(l);
                const int z = sl - (l == 0 ? L::SLB0 : l == 1 ? L::SLB1 : l == 2 ? L::SLB2
                                    : l == 3 ? L::SLB3 : L::SLB4);
                uint32_t cv = 0;
                for (int y = (int)ln; y < nl; y += 64) {
                    if (l == 0) {
                        cv += __popcll(lds[L::V0 + z * N0 + y]);
                        if (z < R && y < R) cv += (uint32_t)(lds[L::XB + z] >> y) & 1u;
                    } else {
                        cv += __popcll(lds[L::P1 + (G::row_base(l) - G::RB1) + z * nl + y]);
                    }
                }
#pragma unroll
                for (int d = 32; d >= 1; d >>= 1) cv += __shfl_xor(cv, d, 64);
                if (ln == 0) soff[sl] = cv;
            }
            __syncthreads();
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

    // ---- 5. emission ----
    // The chunk is walked in z-slice pairs (2p, 2p+1).  For each pair the
    // workgroup stages in LDS the two slices, the +x/+y neighbour voxels next
    // to them and the three shadow-map rows the pair's records sample; scans
    // the counts of the pair's level-0 rows and of the level-1 rows of L1
    // slice p (whose colours come from these two slices); expands their set
    // bits into a list of cells; and writes one record per thread, in order,
    // with no global access but the stores.  The next pair's loads are in
    // flight meanwhile.  Levels 2..4 follow level by level, reading colours
    // and shadows from global memory (a few percent of the records).
    uint8_t *scr = reinterpret_cast<uint8_t *>(lds + L::SCR);
    const uint32_t *soff = reinterpret_cast<const uint32_t *>(scr + L::O_SOFF);
    uint8_t *stg = scr + L::O_STG;
    uint16_t *rowp = reinterpret_cast<uint16_t *>(scr + L::O_ROWP);
    uint16_t *list = reinterpret_cast<uint16_t *>(scr + L::O_LIST);
    uint64_t *out = a.splats + base;
    constexpr int NP = R / 2 + 1, n1 = G::n(1);
    constexpr int SEGW = R < 16 ? R : 16;             // bits per expansion unit
    constexpr int NSEG0 = R / SEGW + 1;              // level-0 row: segments + the x = R cell
    constexpr int NSEG1 = (n1 + 15) / 16;
    PairRegs<B, BS> pr;
    pair_load<B, BS>(a, c, 0, pr);
    for (int p = 0; p < NP; p++) {
        const int z0 = 2 * p;
        const int nrow0 = (z0 + 1 <= R ? 2 : 1) * N0;
        const int nrows = nrow0 + (a.lod ? n1 : 0);
        __syncthreads();                      // the previous pair is done with the scratch
        pair_store<B, BS>(pr, stg);
        auto row_bits = [&](int i) -> uint64_t {
            return i < nrow0 ? lds[L::V0 + z0 * N0 + i] : lds[L::P1 + p * n1 + (i - nrow0)];
        };
        auto row_count = [&](int i) -> uint32_t {
            if (i < nrow0) {
                const int z = z0 + i / N0, y = i % N0;
                uint32_t cv = __popcll(lds[L::V0 + z * N0 + y]);
                if (z < R && y < R) cv += (uint32_t)(lds[L::XB + z] >> y) & 1u;
                return cv;
            }
            return (uint32_t)__popcll(lds[L::P1 + p * n1 + (i - nrow0)]);
        };
        const uint32_t npair = scan_rows<BS>(nrows, row_count, rowp, wsum);
        if (p + 1 < NP) pair_load<B, BS>(a, c, z0 + 2, pr);
        const uint32_t nl0 = nrow0 < nrows ? (uint32_t)rowp[nrow0] : npair;
        // expansion: one unit per (row, 16-bit segment)
        const int nunits = nrow0 * NSEG0 + (nrows - nrow0) * NSEG1;
        for (int u = threadIdx.x; u < nunits; u += BS) {
            if (u < nrow0 * NSEG0) {
                const int i = u / NSEG0, sg = u % NSEG0;
                const int sidx = i / N0, y = i % N0, z = z0 + sidx;
                const uint64_t rb = row_bits(i);
                const uint32_t code0 = ((uint32_t)sidx << 14) | ((uint32_t)y << 7);
                if (sg == NSEG0 - 1) {             // the cell x = R of this row
                    if (z < R && y < R && ((lds[L::XB + z] >> y) & 1ull))
                        list[rowp[i] + __popcll(rb)] = (uint16_t)(code0 | (uint32_t)R);
                } else {
                    uint32_t pos = rowp[i] + (sg ? (uint32_t)__popcll(rb & ((1ull << (sg * SEGW)) - 1ull)) : 0u);
                    uint32_t seg = (uint32_t)(rb >> (sg * SEGW)) & ((SEGW == 32 ? 0u : (1u << SEGW)) - 1u);
                    while (seg) {
                        const uint32_t x = (uint32_t)(sg * SEGW) + (uint32_t)__builtin_ctz(seg);
                        seg &= seg - 1u;
                        list[pos++] = (uint16_t)(code0 | x);
                    }
                }
            } else {
                const int v = u - nrow0 * NSEG0, i = nrow0 + v / NSEG1, sg = v % NSEG1;
                const int y = i - nrow0;
                const uint64_t rb = row_bits(i);
                uint32_t pos = rowp[i] + (sg ? (uint32_t)__popcll(rb & ((1ull << (sg * 16)) - 1ull)) : 0u);
                uint32_t seg = (uint32_t)(rb >> (sg * 16)) & 0xFFFFu;
                while (seg) {
                    const uint32_t x = (uint32_t)(sg * 16) + (uint32_t)__builtin_ctz(seg);
                    seg &= seg - 1u;
                    list[pos++] = (uint16_t)(0x8000u | ((uint32_t)y << 7) | x);
                }
            }
        }
        // Wait for the next pair's prefetch here (vmcnt 0; the encoding leaves
        // the other counters free), before this pair's stores are issued:
        // vmcnt retires loads and stores in issue order, so a wait for the
        // prefetch after the stores would also wait for the stores.
        __builtin_amdgcn_s_waitcnt(0x0F70);
        __syncthreads();
        const uint32_t o0 = soff[L::SLB0 + z0], o1 = soff[L::SLB1 + p];
        for (uint32_t j = threadIdx.x; j < npair; j += BS) {
            const uint32_t e = list[j];
            const int x = (int)(e & 0x7Fu), y = (int)((e >> 7) & 0x7Fu);
            if (j < nl0) {
                const int z = z0 + (int)((e >> 14) & 1u);
                const uint32_t col = staged_colour<B>(stg, x, y, z, z0);
                const uint32_t wy = c.org[1] + y;
                const uint32_t sh = staged_shadow<B>(stg, z - z0, x + y, wy);
                uint32_t fl = 0;
                if constexpr (V2) fl = face_flags<B, V2>(lds, x, y, z);
                out[o0 + j] = pack_record<V2>(c.org[0] + x, wy, c.org[2] + z, col, sh, fl);
            } else {
                const int yy = (int)((e >> 7) & 0x3Fu);
                int qx = x, qy = yy, qz = p;
                descend<B, V2>(lds, 1, qx, qy, qz);
                const uint32_t col = staged_colour<B>(stg, qx, qy, qz, z0);
                const uint32_t wy = c.org[1] + 2u * yy;
                const uint32_t sh = staged_shadow<B>(stg, 2, 2 * (x + yy) + 4, wy + 2u);
                uint32_t fl = 0;
                if constexpr (V2) fl = face_flags<B, V2>(lds, qx, qy, qz);
                out[o1 + (j - nl0)] = pack_record<V2>(c.org[0] + 2u * x, wy, c.org[2] + 2u * p, col,
                                                      sh, fl);
            }
        }
    }
    if (!a.lod) return;
    // levels 2..4: one level at a time, (z, y, x) of each cell in the list
    for (int l = 2; l <= 4; l++) {
        const int nl = G::n(l), nrows = nl * nl;
        const uint64_t *prow = lds + L::P1 + (G::row_base(l) - G::RB1);
        __syncthreads();
        auto row_count = [&](int i) -> uint32_t { return (uint32_t)__popcll(prow[i]); };
        const uint32_t nlev = scan_rows<BS>(nrows, row_count, rowp, wsum);
        for (int i = threadIdx.x; i < nrows; i += BS) {
            uint64_t rb = prow[i];
            uint32_t pos = rowp[i];
            const uint32_t zy = ((uint32_t)(i / nl) << 10) | ((uint32_t)(i % nl) << 5);
            while (rb) {
                const uint32_t x = (uint32_t)__builtin_ctzll(rb);
                rb &= rb - 1ull;
                list[pos++] = (uint16_t)(zy | x);
            }
        }
        __syncthreads();
        const uint32_t ol = soff[l == 2 ? L::SLB2 : l == 3 ? L::SLB3 : L::SLB4];
        const uint32_t d = 1u << l;
        for (uint32_t j = threadIdx.x; j < nlev; j += BS) {
            const uint32_t e = list[j];
            const int x = (int)(e & 31u), y = (int)((e >> 5) & 31u), z = (int)(e >> 10);
            int qx = x, qy = y, qz = z;
            descend<B, V2>(lds, l, qx, qy, qz);
            const uint32_t col = l0_colour<B>(c, qx, qy, qz);
            const uint32_t wx = c.org[0] + ((uint32_t)x << l);
            const uint32_t wy = c.org[1] + ((uint32_t)y << l);
            const uint32_t wz = c.org[2] + ((uint32_t)z << l);
            const uint32_t sh = shadow_bit(a.map, a.sx, wx + d, wy + d, wz + d);
            uint32_t fl = 0;
            if constexpr (V2) fl = face_flags<B, V2>(lds, qx, qy, qz);
            out[ol + j] = pack_record<V2>(wx, wy, wz, col, sh, fl);
        }
    }
}

// LDS (~79 KiB) allows two workgroups per CU; ask the register allocator for
// the matching occupancy (512 threads: 4 waves per SIMD, at most 128 VGPRs;
// 1024 threads: one workgroup of 4 waves per SIMD).
template <int B, int BS, bool V2, bool MASK_OUT>
__global__ void __launch_bounds__(BS)
__attribute__((amdgpu_waves_per_eu(BS == 256 ? 2 : 4))) k_cull(CullArgs a)
{
    __shared__ uint64_t lds[Lay<B, V2>::TOTAL];
    __shared__ uint32_t wsum[BS / 64 + 1];
    __shared__ uint32_t s_ticket;
    __shared__ uint64_t s_base;
    if (a.single) {
        cull_chunk<B, BS, V2, MASK_OUT>(a, 0u, 1u, lds, wsum, &s_base);
        return;
    }
    if (MASK_OUT) {   // make_mask: one workgroup per listed chunk, no offsets
        for (uint32_t k = blockIdx.x; k < a.n; k += gridDim.x) {
            cull_chunk<B, BS, V2, MASK_OUT>(a, k, a.n, lds, wsum, &s_base);
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
        cull_chunk<B, BS, V2, MASK_OUT>(a, k, n, lds, wsum, &s_base);
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

template <int B, int BS, bool V2, bool MASK_OUT>
static hipError_t launch_cull(const CullArgs &a, int grid, hipStream_t st)
{
    hipLaunchKernelGGL((k_cull<B, BS, V2, MASK_OUT>), dim3(grid), dim3(BS), 0, st, a);
    return hipGetLastError();
}

template <int B, bool MASK_OUT>
static hipError_t dispatch_bs(const CullArgs &a, bool v2, int bs, int grid, hipStream_t st)
{
    if (v2) {
        switch (bs) {
        case 256: return launch_cull<B, 256, true, MASK_OUT>(a, grid, st);
        case 1024: return launch_cull<B, 1024, true, MASK_OUT>(a, grid, st);
        default: return launch_cull<B, 512, true, MASK_OUT>(a, grid, st);
        }
    }
    switch (bs) {
    case 256: return launch_cull<B, 256, false, MASK_OUT>(a, grid, st);
    case 1024: return launch_cull<B, 1024, false, MASK_OUT>(a, grid, st);
    default: return launch_cull<B, 512, false, MASK_OUT>(a, grid, st);
    }
}

template <bool MASK_OUT>
static hipError_t dispatch(int b, const CullArgs &a, bool v2, int bs, int grid, hipStream_t st)
{
    switch (b) {
    case 2: return dispatch_bs<2, MASK_OUT>(a, v2, bs, grid, st);
    case 3: return dispatch_bs<3, MASK_OUT>(a, v2, bs, grid, st);
    case 4: return dispatch_bs<4, MASK_OUT>(a, v2, bs, grid, st);
    case 5: return dispatch_bs<5, MASK_OUT>(a, v2, bs, grid, st);
    default: return dispatch_bs<6, MASK_OUT>(a, v2, bs, grid, st);
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
    int grid = 1, bs = ctx->cull_block;
    // one 64^3 chunk alone: a 1024-thread workgroup (16 waves on one CU)
    // finishes it in 0.70x the time of the batched path's 512-thread shape
    if (a.single && world->root_bitw == 6) bs = 1024;
    if (!a.single) {
        if (!ticket || !state) return VP_ERR_ARG;
        a.ticket = ticket;
        a.state = state;
        const int resident = ctx->num_cus * ctx->cull_wg_per_cu;
        grid = (int)((uint64_t)n < (uint64_t)resident ? n : (uint32_t)resident);
        const int rgrid = (int)((n + 255u) / 256u);
        hipLaunchKernelGGL(k_cull_reset, dim3(rgrid < 1024 ? rgrid : 1024), dim3(256), 0, st,
                           a.ticket, a.state, n, a.offsets);
        VP_HIP_TRY(hipGetLastError());
    }
    VP_HIP_TRY(dispatch<false>(world->root_bitw, a, (flags & VP_CM_RECORD_V2) != 0, bs, grid, st));
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
    VP_HIP_TRY(dispatch<true>(world->root_bitw, a, false, 512, grid, as_stream(stream)));
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
