// rle.hip -- chunk RLE encode/decode for gfx950.
//
// Replaces rle_compress / rle_decompress (src/chunkset/rle.c:44-116) and the
// per-chunk (de)compression done by chunk_compress / chunk_open_*
// (src/chunkset.c:133-230).  Word format: (value << 24) | run, runs maximal
// in voxel order, one 0 word after each chunk's runs.
//
// Encode, one workgroup per chunk (persistent, ticket order = list order):
//  1. 16-byte loads; a run starts where a byte differs from the byte before it
//     -> one start bit per voxel in an LDS bitmap (R^3 bits = 32 KiB at R=64).
//  2. popcounts per 64-bit bitmap word -> workgroup scan -> word offsets; a
//     reverse scan gives, per bitmap word, the next start after it, so every
//     run's length is local arithmetic.
//  3. the chunk's word count is published and its output offset taken by a
//     decoupled look-back (same scheme as the remesh kernel).
//  4. each thread writes the words of its bitmap words.
// Decode, one workgroup per chunk: a stream of one full-chunk run is a fill;
// otherwise the words are read eight per thread at a time, their values kept
// in LDS and their lengths scanned into start positions, set as bits of an
// LDS bitmap; an output byte's run is the rank of the last start at or before
// it, so every 16-byte store is independent.
#include "vp_internal.h"

namespace vp {

constexpr uint64_t kRFlagAgg = 1ull << 62;
constexpr uint64_t kRFlagPre = 2ull << 62;
constexpr uint64_t kRValMask = (1ull << 62) - 1;

struct RleArgs {
    uint8_t        *voxels;
    const uint32_t *ids;
    uint32_t        n;
    const uint32_t *d_n;           // nullable: device-resident count <= n (encode)
    uint32_t       *words;
    uint64_t        cap;
    uint64_t       *offsets;
    int16_t        *uniform_out;
    const int16_t  *uniform_in;    // nullable: a chunk marked uniform here is not read (encode)
    uint32_t       *ticket;
    uint64_t       *state;
    uint32_t       *status;
    uint8_t         single;
    uint32_t        nchunks;
};

__device__ __forceinline__ uint32_t r_lane() { return threadIdx.x & 63u; }

// start bits of the 4 bytes of w given the byte before it
__device__ __forceinline__ uint32_t diff4(uint32_t w, uint32_t prev)
{
    const uint32_t shifted = (w << 8) | (prev & 0xFFu);
    const uint32_t d = w ^ shifted;
    const uint32_t m = (((d & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | d) & 0x80808080u;
    return (((m >> 7) * 0x00204081u) >> 21) & 0xFu;
}

template <int BS>
__device__ __forceinline__ uint32_t r_scan_excl(uint32_t v, uint32_t *wsum, uint32_t &total)
{
    constexpr int NW = BS / 64;
    const uint32_t lane = r_lane(), w = threadIdx.x >> 6;
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
        if (lane < (uint32_t)NW) wsum[lane] = t - s;
        if (lane == (uint32_t)NW - 1) wsum[NW] = t;
    }
    __syncthreads();
    total = wsum[NW];
    const uint32_t r = wsum[w] + x - v;
    __syncthreads();
    return r;
}

// suffix-min over the block (thread order): returns min of v over threads > t
template <int BS>
__device__ __forceinline__ uint32_t r_suffix_min_excl(uint32_t v, uint32_t *wtmp)
{
    constexpr int NW = BS / 64;
    const uint32_t lane = r_lane(), w = threadIdx.x >> 6;
    uint32_t x = v;   // inclusive suffix min inside the wave
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t y = __shfl_down(x, d, 64);
        if (lane + d < 64u) x = min(x, y);
    }
    if (lane == 0) wtmp[w] = x;
    __syncthreads();
    uint32_t after = 0xFFFFFFFFu;   // min over later waves
    for (int j = (int)w + 1; j < NW; j++) after = min(after, wtmp[j]);
    const uint32_t nxt = __shfl_down(x, 1, 64);
    const uint32_t excl = (lane < 63u) ? min(nxt, after) : after;
    __syncthreads();
    return excl;
}

template <int B, int BS>
__device__ void rle_encode_chunk(const RleArgs &a, uint32_t k, uint32_t n, uint64_t *bm,
                                 uint32_t *woff, uint32_t *wnext, uint32_t *wsum,
                                 uint64_t *s_base)
{
    constexpr int R = 1 << B, VOX = R * R * R, WORDS = VOX / 64, WPT = (WORDS + BS - 1) / BS;
    uint32_t chunk = a.ids[k];
    if (chunk >= a.nchunks) {       // encode the all-air null chunk instead
        if (threadIdx.x == 0) atomicOr(a.status, (uint32_t)VP_STATUS_BAD_ID);
        chunk = 0xFFFFFFFFu;
    }
    const uint8_t *src = chunk == 0xFFFFFFFFu ? nullptr : a.voxels + ((size_t)chunk << (3 * B));
    // A chunk the caller's uniform table marks uniform (every voxel holds that
    // value; edits mark their chunks mixed) is one run and the terminator:
    // none of its voxels is read.
    const int uval = (src && a.uniform_in) ? (int)a.uniform_in[chunk] : -1;
    const bool fast = uval >= 0;
    uint32_t runs = 1;
    if (!fast) {

    // ---- 1. start bits ----
    // DEPTH 16-byte loads per thread in flight; the byte before a piece is
    // the previous lane's last byte, except at lane 0, which loads it.
    if constexpr (VOX >= 16) {
        constexpr int PIECES = VOX / 16;
        constexpr int PER = PIECES / BS;
        constexpr int DEPTH = PER >= 8 ? 8 : (PER >= 1 ? PER : 1);
        uint16_t *b16 = reinterpret_cast<uint16_t *>(bm);
        const uint32_t ln = r_lane();
        if (!src) {
            for (int s = threadIdx.x; s < PIECES; s += BS) b16[s] = s ? 0 : 1;
        } else {
#pragma unroll 1
            for (int s0 = 0; s0 < PIECES; s0 += BS * DEPTH) {
                uint4 v[DEPTH];
                uint32_t pb[DEPTH];
#pragma unroll
                for (int d = 0; d < DEPTH; d++) {
                    const int sp = s0 + d * BS + (int)threadIdx.x;
                    const int sc = sp < PIECES ? sp : 0;
                    v[d] = *reinterpret_cast<const uint4 *>(src + (size_t)sc * 16);
                    pb[d] = (ln == 0u && sp > 0 && sp < PIECES) ? src[(size_t)sp * 16 - 1] : 0u;
                }
#pragma unroll
                for (int d = 0; d < DEPTH; d++) {
                    const int sp = s0 + d * BS + (int)threadIdx.x;
                    uint32_t prev = (uint32_t)__shfl_up((int)(v[d].w >> 24), 1, 64);
                    if (ln == 0u) prev = sp ? pb[d] : (~v[d].x & 0xFFu);
                    uint32_t m = diff4(v[d].x, prev);
                    m |= diff4(v[d].y, v[d].x >> 24) << 4;
                    m |= diff4(v[d].z, v[d].y >> 24) << 8;
                    m |= diff4(v[d].w, v[d].z >> 24) << 12;
                    if (sp < PIECES) b16[sp] = (uint16_t)m;
                }
            }
        }
    }
    __syncthreads();

    // ---- 2. per-thread bitmap words: counts, offsets, next start ----
    uint32_t mine = 0, first = 0xFFFFFFFFu;
    const int w0 = (int)threadIdx.x * WPT;
    for (int j = 0; j < WPT; j++) {
        const int w = w0 + j;
        if (w >= WORDS) break;
        const uint64_t b = bm[w];
        mine += (uint32_t)__popcll(b);
        if (b && first == 0xFFFFFFFFu) first = (uint32_t)(w * 64 + __builtin_ctzll(b));
    }
    uint32_t run = r_scan_excl<BS>(mine, wsum, runs);
    uint32_t nxt = r_suffix_min_excl<BS>(first, wsum);
    if (nxt == 0xFFFFFFFFu) nxt = VOX;
    // walk own words backwards to give each word the first start after it
    for (int j = WPT - 1; j >= 0; j--) {
        const int w = w0 + j;
        if (w >= WORDS) continue;
        wnext[w] = nxt;
        const uint64_t b = bm[w];
        if (b) nxt = (uint32_t)(w * 64 + __builtin_ctzll(b));
    }
    for (int j = 0; j < WPT; j++) {
        const int w = w0 + j;
        if (w >= WORDS) break;
        woff[w] = run;
        run += (uint32_t)__popcll(bm[w]);
    }
    __syncthreads();
    }   // !fast

    // ---- 3. publish the word count (runs + terminator), look back ----
    const uint64_t nwords = (uint64_t)runs + 1u;
    if (threadIdx.x < 64) {
        uint64_t excl = 0;
        if (!a.single) {
            if (k == 0) {
                if (threadIdx.x == 0) st_relaxed(a.state, kRFlagPre | nwords);
            } else {
                if (threadIdx.x == 0) st_relaxed(a.state + k, kRFlagAgg | nwords);
                int64_t top = (int64_t)k - 1;
                const uint32_t ln = r_lane();
                for (;;) {
                    const int64_t j = top - (int64_t)ln;
                    const uint64_t s = (j >= 0) ? ld_relaxed(a.state + j) : kRFlagPre;
                    const uint64_t f = s >> 62;
                    const uint64_t pre = __ballot(f == 2);
                    const uint64_t zero = __ballot(f == 0);
                    const uint64_t need = pre ? (pre & (~pre + 1ull)) * 2ull - 1ull : ~0ull;
                    if (zero & need) {
                        __builtin_amdgcn_s_sleep(1);
                        continue;
                    }
                    uint64_t v = ((need >> ln) & 1ull) ? (s & kRValMask) : 0ull;
#pragma unroll
                    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
                    excl += v;
                    if (pre) break;
                    top -= 64;
                }
                if (threadIdx.x == 0) st_relaxed(a.state + k, kRFlagPre | (excl + nwords));
            }
        }
        if (threadIdx.x == 0) {
            *s_base = excl;
            a.offsets[k] = excl;
            if (k + 1 == n) a.offsets[n] = excl + nwords;
            if (a.uniform_out)
                a.uniform_out[k] = fast ? (int16_t)uval
                                 : (runs == 1) ? (int16_t)(src ? src[0] : 0) : (int16_t)-1;
            if (a.words && excl + nwords > a.cap) atomicOr(a.status, (uint32_t)VP_STATUS_OVERFLOW);
        }
    }
    __syncthreads();
    const uint64_t base = *s_base;
    if (!a.words || base + nwords > a.cap) return;

    // ---- 4. words ----
    // adjacent lanes take adjacent bitmap words, so a store instruction's
    // addresses stay within a few cache lines.  The starts of a word are taken
    // G at a time: G value loads issued together (the chunk left the L2 since
    // step 1, so each is a memory round trip), then their G words stored.
    constexpr int G = 8;
    uint32_t *out = a.words + base;
    if (fast && threadIdx.x == 0) out[0] = ((uint32_t)uval << 24) | (uint32_t)VOX;
    for (int w = threadIdx.x; w < (fast ? 0 : WORDS); w += BS) {
        uint64_t b = bm[w];
        uint32_t o = woff[w];
        const uint32_t wn = wnext[w];
        while (b) {
            uint32_t pos[G], val[G];
            int cnt = 0;
#pragma unroll
            for (int j = 0; j < G; j++) {
                const bool v = b != 0ull;    // true for j = 0
                pos[j] = v ? (uint32_t)(w * 64 + __builtin_ctzll(b)) : (j ? pos[0] : 0u);
                b = v ? (b & (b - 1ull)) : b;
                cnt += v ? 1 : 0;
            }
            const uint32_t after = b ? (uint32_t)(w * 64 + __builtin_ctzll(b)) : wn;
            if (src) {
#pragma unroll
                for (int j = 0; j < G; j++) val[j] = src[pos[j]];
            } else {
#pragma unroll
                for (int j = 0; j < G; j++) val[j] = 0u;
            }
#pragma unroll
            for (int j = 0; j < G; j++)
                if (j < cnt) {
                    const uint32_t end = (j + 1 < cnt) ? pos[j + 1] : after;
                    out[o + j] = (val[j] << 24) | (end - pos[j]);
                }
            o += (uint32_t)cnt;
        }
    }
    if (threadIdx.x == 0) out[runs] = 0u;
}

template <int B, int BS>
__global__ void __launch_bounds__(BS) k_rle_encode(RleArgs a)
{
    constexpr int R = 1 << B, WORDS = R * R * R / 64;
    __shared__ uint64_t bm[WORDS];
    __shared__ uint32_t woff[WORDS];
    __shared__ uint32_t wnext[WORDS];
    __shared__ uint32_t wsum[BS / 64 + 1];
    __shared__ uint32_t s_ticket;
    __shared__ uint64_t s_base;
    if (a.single) {
        rle_encode_chunk<B, BS>(a, 0u, 1u, bm, woff, wnext, wsum, &s_base);
        return;
    }
    const uint32_t n = a.d_n ? min(*a.d_n, a.n) : a.n;
    for (;;) {
        if (threadIdx.x == 0) s_ticket = atomicAdd(a.ticket, 1u);
        __syncthreads();
        const uint32_t k = s_ticket;
        if (k >= n) break;
        rle_encode_chunk<B, BS>(a, k, n, bm, woff, wnext, wsum, &s_base);
        __syncthreads();
    }
}

template <int B, int BS>
__global__ void __launch_bounds__(BS) k_rle_decode(RleArgs a)
{
    constexpr int R = 1 << B, VOX = R * R * R, WORDS = VOX / 64, WPT = (WORDS + BS - 1) / BS;
    constexpr int K = 8;             // stream words per thread per batch, loads in flight
    // run values kept in LDS for streams of up to VCAP runs: at R = 64 what
    // leaves two workgroups per CU (2 x 81.8 KB); a chunk with 5% of its
    // voxels as single-voxel holes has about 26,000 runs
    constexpr int VCAP = VOX < 32640 ? VOX : 32640;
    constexpr uint32_t LMAX = (uint32_t)VOX + 1u;   // longer runs are malformed; clamped
    __shared__ uint64_t bm[WORDS];
    __shared__ uint32_t rank[WORDS];
    __shared__ uint8_t val8[VCAP];
    __shared__ uint32_t wsum[BS / 64 + 1];
    __shared__ uint32_t s_bad, s_nz;
    for (uint32_t k = blockIdx.x; k < a.n; k += gridDim.x) {
        const uint32_t chunk = a.ids[k];
        const uint64_t o0 = a.offsets[k], o1 = a.offsets[k + 1];
        if (chunk >= a.nchunks || o0 > o1 || o1 > a.cap) {
            if (threadIdx.x == 0) atomicOr(a.status, (uint32_t)VP_STATUS_BAD_ID);
            continue;   // uniform across the workgroup
        }
        uint8_t *dst = a.voxels + ((size_t)chunk << (3 * B));
        const uint32_t *wd = a.words + o0;
        const uint64_t avail = o1 - o0;
        // one run over the whole chunk, then the terminator: a fill
        if (avail >= 2) {
            const uint32_t w0 = wd[0], w1 = wd[1];   // the same address in every lane
            if ((w0 & 0xFFFFFFu) == (uint32_t)VOX && w1 == 0u) {
                const uint32_t b = w0 >> 24;
                const uint32_t v4 = b * 0x01010101u;
                const uint4 q = make_uint4(v4, v4, v4, v4);
                for (int s = threadIdx.x; s < VOX / 16; s += BS)
                    nt_store16(dst + (size_t)s * 16, q);
                continue;   // uniform across the workgroup
            }
        }
        for (int w = threadIdx.x; w < WORDS; w += BS) bm[w] = 0ull;
        // the runs end at the first 0 word after word 0 (rle.c:100-108: the
        // first word is always expanded)
        if (threadIdx.x == 0) {
            s_bad = 0u;
            s_nz = (uint32_t)(avail < 0xFFFFFFFFull ? avail : 0xFFFFFFFFull);
        }
        __syncthreads();
        uint64_t nz = s_nz;
        // run starts: batches of K words per thread; the terminator search,
        // the run values (into LDS) and a scan of the lengths into start bits
        uint64_t pos = 0;
        for (uint64_t b0 = 0; b0 < nz; b0 += (uint64_t)BS * K) {
            const uint64_t i0 = b0 + (uint64_t)threadIdx.x * K;
            uint32_t wv[K];
#pragma unroll
            for (int j = 0; j < K; j++) wv[j] = (i0 + j < avail) ? wd[i0 + j] : 1u;
#pragma unroll
            for (int j = 0; j < K; j++)
                if (i0 + j >= 1 && i0 + j < avail && i0 + j <= 0xFFFFFFFFull && wv[j] == 0u)
                    atomicMin(&s_nz, (uint32_t)(i0 + j));
            __syncthreads();
            nz = s_nz;
            uint32_t len[K], lsum = 0;
            bool bad = false;
#pragma unroll
            for (int j = 0; j < K; j++) {
                const uint64_t i = i0 + j;
                uint32_t l = (i < nz) ? (wv[j] & 0xFFFFFFu) : 0u;
                bad |= (i < nz) && (l == 0u || l > LMAX);   // zero-length or oversized runs
                l = l < LMAX ? l : LMAX;
                len[j] = l;
                lsum += l;
                if (i < nz && i < (uint64_t)VCAP) val8[i] = (uint8_t)(wv[j] >> 24);
            }
            if (bad) s_bad = 1u;
            uint32_t tot = 0;
            uint64_t st = pos + r_scan_excl<BS>(lsum, wsum, tot);
#pragma unroll
            for (int j = 0; j < K; j++) {
                if (len[j] && st < (uint64_t)VOX)
                    atomicOr(reinterpret_cast<unsigned long long *>(&bm[st >> 6]), 1ull << (st & 63u));
                st += len[j];
            }
            pos += tot;
        }
        if (threadIdx.x == 0 && pos != (uint64_t)VOX) s_bad = 1u;
        __syncthreads();
        const uint64_t nw = nz;
        // rank of each bitmap word = starts before it
        uint32_t mine = 0;
        const int w0 = (int)threadIdx.x * WPT;
        for (int j = 0; j < WPT; j++)
            if (w0 + j < WORDS) mine += (uint32_t)__popcll(bm[w0 + j]);
        uint32_t tot = 0;
        uint32_t run = r_scan_excl<BS>(mine, wsum, tot);
        for (int j = 0; j < WPT; j++)
            if (w0 + j < WORDS) {
                rank[w0 + j] = run;
                run += (uint32_t)__popcll(bm[w0 + j]);
            }
        __syncthreads();
        // bytes: run index of byte p = starts at or before p, minus one; each
        // start belongs to a run index below nw
        const bool in_lds = nw <= (uint64_t)VCAP;
        for (int s = threadIdx.x; s < VOX / 16; s += BS) {
            const int w = s >> 2;
            const uint64_t b = bm[w];
            const int sh = (s & 3) * 16;
            const uint64_t below = sh ? (b & ((1ull << sh) - 1ull)) : 0ull;
            const int r0 = (int)(rank[w] + (uint32_t)__popcll(below)) - 1;
            const uint32_t bits16 = (uint32_t)(b >> sh) & 0xFFFFu;
            uint32_t out[4] = {0, 0, 0, 0};
            const uint64_t p0 = (uint64_t)s * 16;
#pragma unroll
            for (int j = 0; j < 16; j++) {
                const int ri = r0 + __popc(bits16 & ((2u << j) - 1u));
                uint32_t v = 0;
                if (ri >= 0 && p0 + j < pos) v = in_lds ? (uint32_t)val8[ri] : (wd[ri] >> 24);
                out[j >> 2] |= v << ((j & 3) * 8);
            }
            nt_store16(dst + p0, make_uint4(out[0], out[1], out[2], out[3]));
        }
        if (threadIdx.x == 0 && s_bad) atomicOr(a.status, (uint32_t)VP_STATUS_BAD_RLE);
        __syncthreads();
    }
}

__global__ void k_rle_reset(uint32_t *ticket, uint64_t *state, uint32_t n, uint64_t *offsets)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        ticket[0] = 0u;
        offsets[0] = 0ull;
    }
    for (uint32_t j = i; j < n; j += gridDim.x * blockDim.x) state[j] = 0ull;
}

template <int B>
static hipError_t rle_launch(int dir, const RleArgs &a, int grid, hipStream_t st)
{
    if (dir == VP_RLE_ENCODE)
        hipLaunchKernelGGL((k_rle_encode<B, 512>), dim3(grid), dim3(512), 0, st, a);
    else
        hipLaunchKernelGGL((k_rle_decode<B, 512>), dim3(grid), dim3(512), 0, st, a);
    return hipGetLastError();
}

static int rle_check(const vp_world_desc *world)
{
    if (!world) return VP_ERR_ARG;
    if (world->root_bitw < 2 || world->root_bitw > 6) return VP_ERR_UNSUPPORTED;
    if ((int)world->max_bitw[0] + world->max_bitw[1] + world->max_bitw[2] > 32) return VP_ERR_ARG;
    return VP_OK;
}

static hipError_t rle_dispatch(uint32_t root_bitw, int dir, const RleArgs &a, int grid,
                               hipStream_t st)
{
    switch (root_bitw) {
    case 2: return rle_launch<2>(dir, a, grid, st);
    case 3: return rle_launch<3>(dir, a, grid, st);
    case 4: return rle_launch<4>(dir, a, grid, st);
    case 5: return rle_launch<5>(dir, a, grid, st);
    default: return rle_launch<6>(dir, a, grid, st);
    }
}

int rle_encode_launch(vp_ctx *ctx, const vp_world_desc *world, const uint8_t *d_voxels,
                      const uint32_t *d_chunk_ids, uint32_t n, const uint32_t *d_n,
                      uint32_t *d_words, uint64_t capacity_words, uint64_t *d_word_offsets,
                      int16_t *d_uniform_out, const int16_t *d_uniform_in, uint32_t *ticket,
                      uint64_t *state, hipStream_t st)
{
    int rc = rle_check(world);
    if (rc) return rc;
    if (!d_voxels || !d_chunk_ids || !d_word_offsets || !ticket || !state) return VP_ERR_ARG;
    if (!d_words && capacity_words != 0) return VP_ERR_ARG;
    if (n == 0) return VP_OK;
    RleArgs a = {};
    a.voxels = const_cast<uint8_t *>(d_voxels);
    a.ids = d_chunk_ids;
    a.n = n;
    a.d_n = d_n;
    a.words = d_words;
    a.cap = capacity_words;
    a.offsets = d_word_offsets;
    a.uniform_out = d_uniform_out;
    a.uniform_in = d_uniform_in;
    a.status = ctx->d_status;
    a.single = 0;
    a.nchunks = (uint32_t)(1ull << (world->max_bitw[0] + world->max_bitw[1] + world->max_bitw[2]));
    a.ticket = ticket;
    a.state = state;
    const int resident = ctx->num_cus * 2;
    const int grid = (int)(n < (uint32_t)resident ? n : (uint32_t)resident);
    const int rgrid = (int)((n + 255u) / 256u);
    hipLaunchKernelGGL(k_rle_reset, dim3(rgrid < 1024 ? rgrid : 1024), dim3(256), 0, st, a.ticket,
                       a.state, n, a.offsets);
    VP_HIP_TRY(hipGetLastError());
    VP_HIP_TRY(rle_dispatch(world->root_bitw, VP_RLE_ENCODE, a, grid, st));
    return VP_OK;
}

}  // namespace vp

using namespace vp;

extern "C" int vp_gpu_rle(vp_ctx *ctx, int dir, const vp_world_desc *world,
                          uint8_t *d_voxels, const uint32_t *d_chunk_ids, uint32_t n,
                          uint32_t *d_words, uint64_t capacity_words,
                          uint64_t *d_word_offsets, int16_t *d_uniform_out, void *stream)
{
    if (!ctx || !world || !d_voxels || !d_chunk_ids || !d_word_offsets) return VP_ERR_ARG;
    if (dir != VP_RLE_ENCODE && dir != VP_RLE_DECODE) return VP_ERR_ARG;
    // d_words == NULL with capacity 0: an encode that only sizes the streams
    if (!d_words && (dir != VP_RLE_ENCODE || capacity_words != 0)) return VP_ERR_ARG;
    const uint32_t root_bitw = world->root_bitw;
    if (root_bitw < 2 || root_bitw > 6) return VP_ERR_UNSUPPORTED;
    if ((int)world->max_bitw[0] + world->max_bitw[1] + world->max_bitw[2] > 32) return VP_ERR_ARG;
    if (n == 0) return VP_OK;
    hipStream_t st = as_stream(stream);
    RleArgs a = {};
    a.voxels = d_voxels;
    a.ids = d_chunk_ids;
    a.n = n;
    a.words = d_words;
    a.cap = capacity_words;
    a.offsets = d_word_offsets;
    a.uniform_out = d_uniform_out;
    a.status = ctx->d_status;
    a.single = (n == 1) ? 1 : 0;
    a.nchunks = (uint32_t)(1ull << (world->max_bitw[0] + world->max_bitw[1] + world->max_bitw[2]));
    int grid = 1;
    if (dir == VP_RLE_ENCODE) {
        std::lock_guard<std::mutex> g(ctx->mu);
        if (!a.single) {
            int rc = ensure_state(ctx, n);
            if (rc) return rc;
            a.ticket = ctx->d_ticket;
            a.state = ctx->d_state;
            const int resident = ctx->num_cus * 2;
            grid = (int)(n < (uint32_t)resident ? n : (uint32_t)resident);
            const int rgrid = (int)((n + 255u) / 256u);
            hipLaunchKernelGGL(k_rle_reset, dim3(rgrid < 1024 ? rgrid : 1024), dim3(256), 0, st,
                               a.ticket, a.state, n, a.offsets);
            VP_HIP_TRY(hipGetLastError());
        }
    } else {
        const int resident = ctx->num_cus * 2;
        grid = (int)(n < (uint32_t)resident ? n : (uint32_t)resident);
    }
    return rle_dispatch(root_bitw, dir, a, grid, st) == hipSuccess ? VP_OK : VP_ERR_HIP;
}
