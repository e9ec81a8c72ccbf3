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
//     plus one byte per row for x = -1 and x = R.  Rows of uniform chunks
//     (and of chunks beyond the world: air) are synthesised without a read.
//  2. face bits of a row are XORs of rows: +x = row ^ (row >> 1 | x=R bit),
//     +y / +z = row ^ the next row.  Their popcounts give per-row counts,
//     per-group (8 rows at R = 64) offsets by a workgroup scan, and the
//     chunk's face count, which is published and turned into an output
//     offset by a decoupled look-back before any face is written.
//  3. faces are written in passes of whole row groups that fit an LDS list
//     of (row, x, axis) entries, in (z, y, x, axis) order -- the reference's
//     emission order.  Each (row, 8-voxel segment) item finds its slots from
//     its row's offset and the popcounts of the row's earlier bits.
//  4. per pass, each thread first does everything but the stores for its
//     faces (ambient terms from the LDS bits, the solid voxel's colour and
//     the shadow-map entries from memory), packed into one word per face,
//     and then only the stores: the vertex quad (32 B) and index row (24 B)
//     of consecutive faces come from consecutive lanes.  A load issued after
//     a store completes after it (the vector memory counter retires in
//     order), so keeping the loads ahead of the stores keeps the stores from
//     stalling the next face's loads.
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

// 16 bytes -> 16 occupancy bits
__device__ __forceinline__ uint32_t nz4(uint32_t w)
{
    const uint32_t m = (((w & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | w) & 0x80808080u;
    return (((m >> 7) * 0x00204081u) >> 21) & 0xFu;
}

template <int B>
struct MeshShape {
    static constexpr int R = 1 << B;
    static constexpr int P = R + 2;                        // rows per axis with the shell
    static constexpr int ROWS = R * R;                     // (z, y) rows of the chunk
    static constexpr int GR = ROWS >= kMeshBS ? ROWS / kMeshBS : 1;   // rows per group
    static constexpr int NG = ROWS / GR;                   // groups (one per thread)
    static constexpr int LIST = 3 * R * R * R < 4096 ? 3 * R * R * R : 4096;   // faces per pass
    static constexpr int KF = (LIST + kMeshBS - 1) / kMeshBS;   // faces per thread per pass
    static constexpr int QW = R >= 16 ? 16 : R;            // bytes per row load
    static constexpr int Q = R / QW;                       // loads per row
    static constexpr uint64_t FULL = R == 64 ? ~0ull : ((1ull << R) - 1ull);
    static_assert(GR * 3 * R <= LIST, "a group's faces must fit one pass");
};

template <int QW> struct RowLoad;
template <> struct RowLoad<16> {
    using T = uint4;
    static __device__ __forceinline__ uint32_t bits(T v)
    {
        return nz4(v.x) | nz4(v.y) << 4 | nz4(v.z) << 8 | nz4(v.w) << 12;
    }
};
template <> struct RowLoad<8> {
    using T = uint2;
    static __device__ __forceinline__ uint32_t bits(T v) { return nz4(v.x) | nz4(v.y) << 4; }
};
template <> struct RowLoad<4> {
    using T = uint32_t;
    static __device__ __forceinline__ uint32_t bits(T v) { return nz4(v); }
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

// per-face state between the LDS work and the global loads: the colour
// byte's address (null: the colour is in meta), the shadow-map entry i the
// face's samples are centred on, the voxel's y, and
// meta = colour | axis << 8 | vao << 10 | rot << 18 | valid << 19 |
//        (reads i - 1) << 20 | (reads i + 1) << 21 | (reads i) << 22
struct FaceWork {
    const uint8_t *cptr;
    uint32_t mi;
    uint32_t wy;
    uint32_t meta;
};

// LDS half of one face of the list: the air cell's in-plane neighbours give
// the vertex ambient terms (sample_ao, mesher.c:72-171, and the vertex order
// of mesher.c:254-272), and the addresses of the colour byte and the shadow
// map entries the face needs (mesher.c:282-320, shadow.h:52-75)
template <int B>
__device__ __forceinline__ void prep_face(FaceWork &w, uint32_t j, uint32_t pf, const uint32_t *list,
                                          const uint64_t *occ, const uint8_t *edge, const int16_t *s_uni,
                                          const uint32_t *s_nid, const uint8_t *voxels, uint32_t map_sx,
                                          uint32_t ox, uint32_t oy, uint32_t oz)
{
    constexpr int R = 1 << B, P = R + 2;
    const uint32_t e = list[j < pf ? j : 0u];
    const uint32_t axis = e & 3u, x = (e >> 2) & 63u, r = e >> 8;
    const int z = (int)(r >> B), yl = (int)(r & (R - 1));
    const int r0 = (z + 1) * P + (yl + 1);
    const uint32_t solid_a = (uint32_t)(occ[r0] >> x) & 1u;
    // the air cell of the pair (A itself, or B when A is solid)
    const int px = (int)x + (axis == 0 ? (int)solid_a : 0);
    const int py = yl + (axis == 1 ? (int)solid_a : 0);
    const int pz = z + (axis == 2 ? (int)solid_a : 0);
    // occupancy at shell coordinates in [-1, R]; both words are read and the
    // bit selected without a branch, so the eight samples' reads issue together
    auto sol = [&](int qx, int qy, int qz) -> uint32_t {
        const int rr = (qz + 1) * P + (qy + 1);
        const uint64_t rw = occ[rr];
        const uint32_t eg = edge[rr];
        const uint32_t in = (uint32_t)(rw >> (qx & 63)) & 1u;
        const uint32_t m_lo = 0u - (uint32_t)(qx < 0), m_hi = 0u - (uint32_t)(qx >= R);
        return (in & ~(m_lo | m_hi)) | (eg & m_lo & 1u) | ((eg >> 1) & m_hi);
    };
    // steps d0 along u0 and d1 along u1, the two axes other than the
    // face's, in ascending order
    auto at = [&](int d0, int d1) -> uint32_t {
        return sol(px + (axis != 0 ? d0 : 0), py + (axis == 0 ? d0 : (axis == 2 ? d1 : 0)),
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
    const uint32_t vao = v0 | v1 << 2 | v2 << 4 | v3 << 6;
    const uint32_t valid = j < pf ? 1u : 0u;
    const uint32_t wx = ox + x, wy = oy + (uint32_t)yl, wz = oz + (uint32_t)z;
    // the solid voxel of the pair: A, or B at +axis (maybe in the next chunk)
    const uint32_t bx = x + (axis == 0), by = (uint32_t)yl + (axis == 1), bz = (uint32_t)z + (axis == 2);
    const uint32_t sx = solid_a ? x : bx, sy = solid_a ? (uint32_t)yl : by, sz = solid_a ? (uint32_t)z : bz;
    const int nb = 13 + (sx >= (uint32_t)R) + 3 * (sy >= (uint32_t)R) + 9 * (sz >= (uint32_t)R);
    const int16_t un = s_uni[nb];
    const uint32_t cval = un >= 0 ? (uint32_t)un : 0u;
    w.cptr = un >= 0 ? nullptr
                     : voxels + ((uint64_t)s_nid[nb] << (3 * B)) +
                           ((((sz & (R - 1)) << B | (sy & (R - 1))) << B) | (sx & (R - 1)));
    // shadow samples: entries (i, i - 1) of the solid +x cell for x faces of
    // air; (i, i + 1) of the voxel for y faces of solid voxels; for z faces
    // (i, i + 1) of the cell in front and (i - 1, i) one below it
    uint32_t need = 0;
    w.mi = 0;
    if (axis == 0 && !solid_a) {
        w.mi = (wx + 1u + wy) + map_sx * wz;
        need = 1u << 20 | 1u << 22;
    } else if (axis == 1 && solid_a) {
        w.mi = (wx + wy) + map_sx * wz;
        need = 1u << 21 | 1u << 22;
    } else if (axis == 2) {
        w.mi = (wx + wy) + map_sx * (wz + solid_a);
        need = (wy != 0u ? 1u << 20 : 0u) | 1u << 21 | 1u << 22;
    }
    w.wy = wy;
    w.meta = cval | axis << 8 | vao << 10 | rot << 18 | valid << 19 | (valid ? need : 0u);
}

template <int B>
__global__ void __launch_bounds__(kMeshBS, 4) k_mesh(MeshArgs a)
{
    using S = MeshShape<B>;
    constexpr int R = S::R, P = S::P, BS = kMeshBS, KF = S::KF;
    using RL = RowLoad<S::QW>;
    __shared__ uint64_t occ[P * P];          // bits x = 0..R-1 of row (z, y), shell coordinates + 1
    __shared__ uint8_t  edge[P * P];         // bit 0: x = -1, bit 1: x = R
    __shared__ uint8_t  rowcnt[S::ROWS];     // faces of each chunk row
    __shared__ uint32_t goff[S::NG + 1];     // face offset of each row group
    __shared__ uint32_t list[S::LIST];       // (row << 8 | x << 2 | axis) of a pass's faces
    __shared__ uint32_t wsum[BS / 64 + 1];
    __shared__ int16_t  s_uni[27];           // 3x3x3 neighbourhood: value if uniform (0 = air/outside), -1 mixed
    __shared__ uint32_t s_nid[27];
    __shared__ uint32_t s_ticket;
    __shared__ uint64_t s_base;

    const uint32_t mx = 1u << a.mb0, my = 1u << a.mb1, mz = 1u << a.mb2;
    for (uint32_t visit = 0;; visit++) {
        // a single chunk takes no ticket: the one workgroup does list entry 0
        if (threadIdx.x == 0) s_ticket = a.single ? (visit == 0 ? 0u : a.n) : atomicAdd(a.ticket, 1u);
        __syncthreads();
        const uint32_t k = s_ticket;
        if (k >= a.n) return;
        const uint32_t id = a.ids[k];
        const bool bad = id >= a.nchunks;
        const uint32_t cx = id & (mx - 1u), cy = (id >> a.mb0) & (my - 1u), cz = id >> (a.mb0 + a.mb1);
        if (threadIdx.x < 27 && !bad) {
            const int dx = (int)(threadIdx.x % 3) - 1, dy = (int)(threadIdx.x / 3 % 3) - 1,
                      dz = (int)(threadIdx.x / 9) - 1;
            const int64_t nx = (int64_t)cx + dx, ny = (int64_t)cy + dy, nz = (int64_t)cz + dz;
            int16_t u = 0;
            uint32_t nid = 0;
            if (nx >= 0 && nx < mx && ny >= 0 && ny < my && nz >= 0 && nz < mz) {
                nid = (((uint32_t)nz << a.mb1 | (uint32_t)ny) << a.mb0) | (uint32_t)nx;
                u = a.uniform ? a.uniform[nid] : (int16_t)-1;
            }
            s_uni[threadIdx.x] = u;
            s_nid[threadIdx.x] = nid;
        }
        __syncthreads();

        // ---- 1. occupancy with the one-voxel shell, 4 loads in flight ----
        if (!bad) {
            constexpr int NI = P * P * S::Q, U = 4;
            for (int t0 = threadIdx.x; t0 < NI; t0 += U * BS) {
                typename RL::T v[U];
                int mode[U];                     // 0 air, 1 solid, 2 loaded
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const int t = t0 + u * BS;
                    mode[u] = 0;
                    v[u] = typename RL::T{};
                    if (t < NI) {
                        const int row = t / S::Q, q = t % S::Q;
                        const int yy = row % P, zz = row / P;
                        const int nb = 13 + 3 * (yy == 0 ? -1 : (yy == P - 1 ? 1 : 0)) +
                                       9 * (zz == 0 ? -1 : (zz == P - 1 ? 1 : 0));
                        const int16_t un = s_uni[nb];
                        if (un >= 0) {
                            mode[u] = un ? 1 : 0;
                        } else {
                            const uint32_t ly = (uint32_t)(yy - 1) & (R - 1), lz = (uint32_t)(zz - 1) & (R - 1);
                            const uint8_t *p = a.voxels + ((uint64_t)s_nid[nb] << (3 * B)) +
                                               ((((uint64_t)lz << B) | ly) << B) + q * S::QW;
                            v[u] = *reinterpret_cast<const typename RL::T *>(p);
                            mode[u] = 2;
                        }
                    }
                }
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const int t = t0 + u * BS;
                    const uint32_t bits = mode[u] == 2 ? RL::bits(v[u])
                                                       : (mode[u] == 1 ? (1u << S::QW) - 1u : 0u);
                    uint64_t r = (uint64_t)bits << ((t % S::Q) * S::QW);
                    if constexpr (S::Q > 1) {
#pragma unroll
                        for (int d = 1; d < S::Q; d <<= 1) r |= __shfl_xor(r, d, 64);
                    }
                    if (t < NI && t % S::Q == 0) occ[t / S::Q] = r;
                }
            }
            constexpr int NE = P * P, UE = 3;
            for (int t0 = threadIdx.x; t0 < NE; t0 += UE * BS) {
                uint8_t lo[UE], hi[UE];
#pragma unroll
                for (int u = 0; u < UE; u++) {
                    const int row = t0 + u * BS;
                    lo[u] = hi[u] = 0;
                    if (row < NE) {
                        const int yy = row % P, zz = row / P;
                        const int nb = 13 + 3 * (yy == 0 ? -1 : (yy == P - 1 ? 1 : 0)) +
                                       9 * (zz == 0 ? -1 : (zz == P - 1 ? 1 : 0));
                        const uint64_t ro = ((((uint64_t)((uint32_t)(zz - 1) & (R - 1))) << B) |
                                             ((uint32_t)(yy - 1) & (R - 1))) << B;
                        const int16_t ul = s_uni[nb - 1], uh = s_uni[nb + 1];
                        lo[u] = ul >= 0 ? (ul != 0)
                                        : (a.voxels[((uint64_t)s_nid[nb - 1] << (3 * B)) + ro + (R - 1)] != 0);
                        hi[u] = uh >= 0 ? (uh != 0)
                                        : (a.voxels[((uint64_t)s_nid[nb + 1] << (3 * B)) + ro] != 0);
                    }
                }
#pragma unroll
                for (int u = 0; u < UE; u++) {
                    const int row = t0 + u * BS;
                    if (row < NE) edge[row] = (uint8_t)(lo[u] | hi[u] << 1);
                }
            }
        }
        __syncthreads();

        // ---- 2. per-row face counts, group offsets, output offset ----
        uint32_t gcnt = 0;
        if (!bad && threadIdx.x < S::NG) {
#pragma unroll
            for (int i = 0; i < S::GR; i++) {
                const int r = threadIdx.x * S::GR + i;
                const int z = r / R, y = r % R;
                const int r0 = (z + 1) * P + (y + 1);
                const uint64_t o = occ[r0];
                const uint64_t bx = (o >> 1) | ((uint64_t)(edge[r0] >> 1) << (R - 1));
                const uint32_t c = __popcll((o ^ bx) & S::FULL) + __popcll((o ^ occ[r0 + 1]) & S::FULL) +
                                   __popcll((o ^ occ[r0 + P]) & S::FULL);
                rowcnt[r] = (uint8_t)c;
                gcnt += c;
            }
        }
        uint32_t faces;
        const uint32_t gex = m_scan_excl<BS>(gcnt, wsum, faces);
        if (threadIdx.x < S::NG) goff[threadIdx.x] = gex;
        if (threadIdx.x == 0) goff[S::NG] = faces;
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

        // ---- 3./4. faces, in passes of whole row groups that fit the list ----
        const uint32_t ox = cx << B, oy = cy << B, oz = cz << B;
        uint32_t g0 = 0;
        while (g0 < (uint32_t)S::NG) {
            const uint32_t lo = goff[g0];
            const bool in = threadIdx.x >= g0 && threadIdx.x < (uint32_t)S::NG &&
                            goff[threadIdx.x + 1] - lo <= (uint32_t)S::LIST;
            const uint32_t g1 = g0 + (uint32_t)__syncthreads_count(in);
            const uint32_t pf = goff[g1] - lo;
            // list entries, one (row, 8-voxel segment) item per thread at a
            // time: a segment's first slot is its row's offset plus the faces
            // of the row's earlier segments, so no scan is needed
            {
                constexpr int SEG = R >= 8 ? R / 8 : 1, SW = R >= 8 ? 8 : R;
                const uint32_t rfirst = g0 * S::GR, nitems = (g1 - g0) * S::GR * SEG;
                for (uint32_t it = threadIdx.x; it < nitems; it += BS) {
                    const uint32_t r = rfirst + it / SEG, sg = it % SEG;
                    if (!rowcnt[r]) continue;
                    const uint32_t g = r / S::GR;
                    uint32_t pos = goff[g] - lo;
                    for (uint32_t q = g * S::GR; q < r; q++) pos += rowcnt[q];
                    const int z = (int)(r / R), y = (int)(r % R);
                    const int r0 = (z + 1) * P + (y + 1);
                    const uint64_t o = occ[r0];
                    const uint64_t bx = (o >> 1) | ((uint64_t)(edge[r0] >> 1) << (R - 1));
                    const uint64_t fx = (o ^ bx) & S::FULL, fy = (o ^ occ[r0 + 1]) & S::FULL,
                                   fz = (o ^ occ[r0 + P]) & S::FULL;
                    const uint32_t sh = sg * SW;
                    const uint64_t below = (1ull << sh) - 1ull;
                    pos += __popcll(fx & below) + __popcll(fy & below) + __popcll(fz & below);
                    constexpr uint32_t SM = (1u << SW) - 1u;
                    const uint32_t mxb = (uint32_t)(fx >> sh) & SM, myb = (uint32_t)(fy >> sh) & SM,
                                   mzb = (uint32_t)(fz >> sh) & SM;
                    uint32_t any = mxb | myb | mzb;
                    while (any) {
                        const uint32_t j = (uint32_t)__builtin_ctz(any);
                        any &= any - 1u;
                        const uint32_t e = r << 8 | (sh + j) << 2;
                        if ((mxb >> j) & 1u) list[pos++] = e | 0u;
                        if ((myb >> j) & 1u) list[pos++] = e | 1u;
                        if ((mzb >> j) & 1u) list[pos++] = e | 2u;
                    }
                }
            }
            __syncthreads();
            // E1: everything but the stores -- ambient terms from LDS, the
            // colour byte and the shadow samples from memory -- packed into
            // one word per face; E2: the stores only, so no load waits behind
            // them (vector memory counters retire loads and stores in order)
            uint32_t packed[KF];
#pragma unroll
            for (int k0 = 0; k0 < KF; k0 += 2) {
                if (threadIdx.x + (uint32_t)k0 * BS >= pf) break;
                FaceWork fw[2];
#pragma unroll
                for (int u = 0; u < 2; u++) {
                    if (k0 + u < KF)
                        prep_face<B>(fw[u], threadIdx.x + (uint32_t)(k0 + u) * BS, pf, list, occ, edge,
                                     s_uni, s_nid, a.voxels, a.map_sx, ox, oy, oz);
                    __builtin_amdgcn_sched_barrier(0);
                }
                uint32_t col[2], elo[2], emid[2], ehi[2];
#pragma unroll
                for (int u = 0; u < 2; u++) {
                    if (k0 + u >= KF) break;
                    const FaceWork &w = fw[u];
                    col[u] = ((w.meta >> 19) & 1u) && w.cptr ? (uint32_t)*w.cptr : (w.meta & 0xFFu);
                    elo[u] = ((w.meta >> 20) & 1u) ? a.map[w.mi - 1u] : 0u;
                    emid[u] = ((w.meta >> 22) & 1u) ? a.map[w.mi] : 0u;
                    ehi[u] = ((w.meta >> 21) & 1u) ? a.map[w.mi + 1u] : 0u;
                }
#pragma unroll
                for (int u = 0; u < 2; u++) {
                    if (k0 + u >= KF) break;
                    const FaceWork &w = fw[u];
                    const uint32_t axis = (w.meta >> 8) & 3u, top = w.wy + 1u;
                    uint32_t shadow = 0, diamond = 3;
                    if ((w.meta >> 22) & 1u) {
                        if (axis == 0) {
                            shadow = (emid[u] < top && elo[u] < top) ? 0u : 1u;
                        } else if (axis == 1) {
                            shadow = (emid[u] < top && ehi[u] < top) ? 0u : 1u;
                        } else {
                            const uint32_t lit0 = (emid[u] < top && ehi[u] < top) ? 1u : 0u;
                            const uint32_t lit1 = (w.wy != 0u && elo[u] < w.wy && emid[u] < w.wy) ? 1u : 0u;
                            diamond = lit0 << 1 | lit1;
                        }
                    }
                    packed[k0 + u] = col[u] | shadow << 8 | diamond << 9 | ((w.meta >> 10) & 0x1FFu) << 11;
                }
                __builtin_amdgcn_sched_barrier(0);
            }
#pragma unroll
            for (int kf = 0; kf < KF; kf++) {
                const uint32_t j = threadIdx.x + (uint32_t)kf * BS;
                if (j >= pf) break;
                const uint32_t e = list[j];
                const uint32_t axis = e & 3u, x = (e >> 2) & 63u, r = e >> 8;
                const uint32_t z = r >> B, yl = r & (R - 1);
                const uint32_t normal = ((uint32_t)(occ[(z + 1) * P + (yl + 1)] >> x) & 1u) ^ 1u;
                const uint32_t pk = packed[kf];
                const uint32_t common = (pk & 0xFFu) | (axis + 3u * normal) << 8 | ((pk >> 8) & 1u) << 13 |
                                        ((pk >> 9) & 3u) << 14;
                const uint32_t vao = (pk >> 11) & 0xFFu, rot = (pk >> 19) & 1u;
                const uint32_t wx = ox + x, wy = oy + yl, wz = oz + z;
                uint32_t wd[8];
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const uint32_t c = m_corner(axis, v);
                    const uint32_t gx = (wx + (c & 1u)) & 0xFFFFu;
                    const uint32_t gy = (wy + ((c >> 1) & 1u)) & 0xFFFFu;
                    const uint32_t gz = (wz + (c >> 2)) & 0xFFFFu;
                    const uint32_t attr = (common | ((vao >> (2 * v)) & 3u) << 6 | (uint32_t)v << 11) & 0xFFFFu;
                    wd[2 * v] = gx | gy << 16;
                    wd[2 * v + 1] = gz | attr << 16;
                }
                const uint32_t fl = lo + j;                     // chunk-local face index
                const uint64_t gf = base + fl;
                uint4 *gp = reinterpret_cast<uint4 *>(a.geom + 16 * gf);
                nt_store16(gp, make_uint4(wd[0], wd[1], wd[2], wd[3]));
                nt_store16(gp + 1, make_uint4(wd[4], wd[5], wd[6], wd[7]));
                const uint32_t vb = 4u * fl;
                uint2 *ip = reinterpret_cast<uint2 *>(a.index + 6 * gf);
                nt_store8(ip, make_uint2(vb + m_tri(normal, rot, 0), vb + m_tri(normal, rot, 1)));
                nt_store8(ip + 1, make_uint2(vb + m_tri(normal, rot, 2), vb + m_tri(normal, rot, 3)));
                nt_store8(ip + 2, make_uint2(vb + m_tri(normal, rot, 4), vb + m_tri(normal, rot, 5)));
                __builtin_amdgcn_sched_barrier(0);   // one face at a time: bounded live ranges
            }
            __syncthreads();
            g0 = g1;
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
