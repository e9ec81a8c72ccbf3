/*
 * vp_oracle.c -- CPU restatement of voxplat's voxel pipeline.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path and the CPU leg of
 * the benchmark.  The product (voxplat_amd/) never links or calls this file.
 *
 * Build: oracle/build_oracle.py (gcc -O2 -fopenmp -ffp-contract=off).
 * The floating-point code of the splat port is written with one operation per
 * expression and no contraction so that its results are the IEEE results of
 * the stated operation order.
 */
#include "vp_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ========================================================================
 * Generator: 2D value noise in Q16 fixed point, 4 octaves, caves at 5%.
 * The reference generator (src/chunkset/gen.c:89-142,187-341) uses FastNoise
 * simplex seeded from time(0); only its shape is mirrored: columns solid below
 * a height, colour by height band.
 * ====================================================================== */
uint64_t vpo_splitmix64(uint64_t x)
{
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static int64_t lattice_q16(uint64_t seed, int octave, uint32_t ix, uint32_t iz)
{
    uint64_t key = seed ^ ((uint64_t)(octave + 1) * 0xD1B54A32D192ED03ull)
                 ^ (((uint64_t)ix << 32) | (uint64_t)iz);
    return (int64_t)(vpo_splitmix64(key) >> 48);
}

static int64_t smooth_q16(int64_t t)      /* t^2 (3 - 2t), t in [0, 65536) */
{
    return (((t * t) >> 16) * (3 * 65536 - 2 * t)) >> 16;
}

static int64_t lerp_q16(int64_t a, int64_t b, int64_t t)
{
    return a + (((b - a) * t) >> 16);     /* arithmetic shift */
}

int32_t vpo_gen_height(const vpo_gen_params *p, uint32_t x, uint32_t z)
{
    int64_t acc = 0;
    for (int o = 0; o < 4; o++) {
        uint32_t per = p->period >> o;
        if (per == 0) per = 1;
        int64_t  w  = 8 >> o;
        uint32_t ix = x / per, iz = z / per;
        int64_t  tx = ((int64_t)(x % per) << 16) / per;
        int64_t  tz = ((int64_t)(z % per) << 16) / per;
        int64_t  sx = smooth_q16(tx), sz = smooth_q16(tz);
        int64_t  v00 = lattice_q16(p->seed, o, ix, iz);
        int64_t  v10 = lattice_q16(p->seed, o, ix + 1, iz);
        int64_t  v01 = lattice_q16(p->seed, o, ix, iz + 1);
        int64_t  v11 = lattice_q16(p->seed, o, ix + 1, iz + 1);
        int64_t  a = lerp_q16(v00, v10, sx);
        int64_t  b = lerp_q16(v01, v11, sx);
        acc += w * lerp_q16(a, b, sz);
    }
    uint64_t span = (uint64_t)(p->hmax - p->hmin + 1);
    return p->hmin + (int32_t)(((uint64_t)acc * span) / (15ull * 65536ull));
}

void vpo_chunk_coords(const vp_world_desc *w, uint32_t chunk, uint32_t c[3])
{
    c[0] = chunk & ((1u << w->max_bitw[0]) - 1u);
    c[1] = (chunk >> w->max_bitw[0]) & ((1u << w->max_bitw[1]) - 1u);
    c[2] = chunk >> (w->max_bitw[0] + w->max_bitw[1]);
}

/* flatten3, src/chunkset.c:124-126 */
uint32_t vpo_chunk_index(const vp_world_desc *w, const uint32_t c[3])
{
    return ((c[2] << w->max_bitw[1] | c[1]) << w->max_bitw[0]) | c[0];
}

void vpo_gen_chunk(const vp_world_desc *w, const vpo_gen_params *p,
                   uint32_t chunk, uint8_t *out)
{
    uint32_t b = w->root_bitw, R = 1u << b, c[3];
    vpo_chunk_coords(w, chunk, c);
    uint64_t seed_chunk = vpo_splitmix64(p->seed ^ (uint64_t)chunk);
    int32_t *h = (int32_t *)malloc(sizeof(int32_t) * R * R);
    for (uint32_t z = 0; z < R; z++)
        for (uint32_t x = 0; x < R; x++)
            h[z * R + x] = vpo_gen_height(p, c[0] * R + x, c[2] * R + z);
    for (uint32_t z = 0; z < R; z++)
        for (uint32_t y = 0; y < R; y++)
            for (uint32_t x = 0; x < R; x++) {
                uint32_t i  = ((z << b | y) << b) | x;   /* flatten1, chunkset.c:128-130 */
                int32_t  wy = (int32_t)(c[1] * R + y);
                uint8_t  v  = 0;
                if (wy < h[z * R + x]) {
                    int32_t band = wy >> 3;
                    v = (uint8_t)(1 + (band < 62 ? band : 62));
                    if (vpo_splitmix64(seed_chunk ^ (uint64_t)i) % 20u == 0) v = 0;
                }
                out[i] = v;
            }
    free(h);
}

void vpo_gen_world(const vp_world_desc *w, const vpo_gen_params *p, uint8_t *out, int threads)
{
    uint32_t R3 = 1u << (3 * w->root_bitw);
    int64_t  nc = 1ll << (w->max_bitw[0] + w->max_bitw[1] + w->max_bitw[2]);
#pragma omp parallel for schedule(dynamic, 4) num_threads(threads > 0 ? threads : 1)
    for (int64_t c = 0; c < nc; c++)
        vpo_gen_chunk(w, p, (uint32_t)c, out + (uint64_t)c * R3);
}

/* ========================================================================
 * Shadow map, src/chunkset/shadow.h:26-89
 * ====================================================================== */
void vpo_shadow_dims(const vp_world_desc *w, uint32_t *size_x, uint32_t *size_z)
{
    /* shadow_init, shadow.h:30-33: SX = X + Y, SZ = Z */
    uint32_t b = w->root_bitw;
    *size_x = ((1u << w->max_bitw[0]) << b) + ((1u << w->max_bitw[1]) << b);
    *size_z = (1u << w->max_bitw[2]) << b;
}

/* shadow_index, shadow.h:45-51 (u32 arithmetic, x+y overflow wraps into the
 * next z-row exactly as the reference does) */
uint32_t vpo_shadow_index(uint32_t size_x, const uint32_t ws[3])
{
    return (ws[0] + ws[1]) + size_x * ws[2];
}

/* shadow_sample, shadow.h:54-63: shadowed unless both the entry and its +1
 * neighbour are at or below the voxel's height */
int vpo_shadow_sample(const uint16_t *map, uint32_t size_x, const uint32_t ws[3])
{
    uint32_t i = vpo_shadow_index(size_x, ws);
    int lit = (map[i] <= ws[1]) && (map[i + 1] <= ws[1]);
    return !lit;
}

/* shadow_place_update, shadow.h:77-89 */
void vpo_shadow_place(uint16_t *map, uint32_t size_x, const uint32_t ws[3])
{
    if (vpo_shadow_sample(map, size_x, ws)) return;
    map[vpo_shadow_index(size_x, ws)] = (uint16_t)ws[1];
}

static inline uint8_t world_voxel(const vp_world_desc *w, const uint8_t *voxels,
                                  uint32_t wx, uint32_t wy, uint32_t wz)
{
    uint32_t b = w->root_bitw, R = 1u << b;
    uint32_t c[3] = {wx >> b, wy >> b, wz >> b};
    uint32_t l[3] = {wx & (R - 1), wy & (R - 1), wz & (R - 1)};
    uint64_t ci = vpo_chunk_index(w, c);
    return voxels[(ci << (3 * b)) + (((l[2] << b | l[1]) << b) | l[0])];
}

/* Placement order of chunkset_gen (gen.c:236-306): chunks by index, then x,
 * z, y ascending.  The reference runs chunks in an OpenMP loop (a race); the
 * build fixes the serial order. */
void vpo_shadow_build(const vp_world_desc *w, const uint8_t *voxels,
                      uint16_t *map, int serial, int threads)
{
    uint32_t b = w->root_bitw, R = 1u << b, sx, sz;
    vpo_shadow_dims(w, &sx, &sz);
    memset(map, 0, sizeof(uint16_t) * vp_shadow_padded_len(sx, sz));
    uint32_t mx = 1u << w->max_bitw[0], my = 1u << w->max_bitw[1], mz = 1u << w->max_bitw[2];
    if (serial) {
        uint64_t nc = (uint64_t)mx * my * mz;
        for (uint64_t ci = 0; ci < nc; ci++) {
            uint32_t c[3];
            vpo_chunk_coords(w, (uint32_t)ci, c);
            const uint8_t *v = voxels + (ci << (3 * b));
            for (uint32_t x = 0; x < R; x++)
                for (uint32_t z = 0; z < R; z++)
                    for (uint32_t y = 0; y < R; y++) {
                        if (!v[((z << b | y) << b) | x]) continue;
                        uint32_t ws[3] = {c[0] * R + x, c[1] * R + y, c[2] * R + z};
                        vpo_shadow_place(map, sx, ws);
                    }
        }
        return;
    }
    /* A voxel at wz only touches row wz (indices SX*wz + [0, X+Y-1]), so the
     * rows are independent; inside a row the serial order restricts to
     * cy, cx, x, y. */
#pragma omp parallel for schedule(dynamic, 8) num_threads(threads > 0 ? threads : 1)
    for (int64_t wz = 0; wz < (int64_t)sz; wz++) {
        for (uint32_t cy = 0; cy < my; cy++)
            for (uint32_t cx = 0; cx < mx; cx++)
                for (uint32_t x = 0; x < R; x++)
                    for (uint32_t y = 0; y < R; y++) {
                        uint32_t ws[3] = {cx * R + x, cy * R + y, (uint32_t)wz};
                        if (!world_voxel(w, voxels, ws[0], ws[1], ws[2])) continue;
                        vpo_shadow_place(map, sx, ws);
                    }
    }
}

/* ========================================================================
 * Mesher, src/chunkset/mesher.c
 * ====================================================================== */

/* flatten1_no_po2, mesher.c:365-374 */
static inline uint32_t flat_np2(uint32_t x, uint32_t y, uint32_t z, uint32_t n)
{
    return (z * n + y) * n + x;
}

/* chunk_make_mask, mesher.c:377-456.  For every voxel A of the chunk and each
 * axis, B is the next voxel along +axis (from the +axis neighbour chunk at the
 * chunk's far face, the all-air null chunk beyond the world).  When exactly one
 * of A, B is air, the solid one's value is stored at the cell of the solid
 * voxel: A's own cell, or A's cell + axis when A is air.  The early-out for
 * all-air chunks (mesher.c:404-409) only skips work whose result is empty, so
 * the full computation is the defined behaviour. */
void vpo_make_mask(const vp_world_desc *w, const uint8_t *voxels,
                   uint32_t chunk, uint8_t *mask)
{
    uint32_t b = w->root_bitw, R = 1u << b, n = R + 1, c[3];
    vpo_chunk_coords(w, chunk, c);
    const uint8_t *own = voxels + ((uint64_t)chunk << (3 * b));
    const uint8_t *nb[3] = {NULL, NULL, NULL};
    for (int a = 0; a < 3; a++) {
        uint32_t cc[3] = {c[0], c[1], c[2]};
        cc[a]++;
        if (cc[a] < (1u << w->max_bitw[a]))
            nb[a] = voxels + ((uint64_t)vpo_chunk_index(w, cc) << (3 * b));
    }
    for (uint32_t z = 0; z < R; z++)
        for (uint32_t y = 0; y < R; y++)
            for (uint32_t x = 0; x < R; x++) {
                uint32_t p[3] = {x, y, z};
                uint32_t ai = ((z << b | y) << b) | x;
                uint8_t  A = own[ai];
                for (int a = 0; a < 3; a++) {
                    uint8_t B;
                    if (p[a] == R - 1) {
                        uint32_t bi = ai - ((R - 1) << (a * b));
                        B = nb[a] ? nb[a][bi] : 0;
                    } else {
                        B = own[ai + (1u << (a * b))];
                    }
                    if ((A == 0) == (B == 0)) continue;
                    uint32_t q[3] = {x, y, z};
                    q[a] += (A == 0);
                    mask[flat_np2(q[0], q[1], q[2], n)] = (uint8_t)(A | B);
                }
            }
}

/* chunk_mask_downsample, mesher.c:460-493: scan the level-(l-1) cells in
 * z, y, x order and copy each non-zero cell to its parent; the last write
 * wins.  Both buffers have stride R+1; dst must be zero on entry. */
void vpo_mask_downsample(uint32_t root, int level, const uint8_t *src, uint8_t *dst)
{
    uint32_t n = root + 1, range = (root >> (level - 1)) + 1;
    for (uint32_t z = 0; z < range; z++)
        for (uint32_t y = 0; y < range; y++)
            for (uint32_t x = 0; x < range; x++) {
                uint8_t v = src[flat_np2(x, y, z, n)];
                if (!v) continue;
                dst[flat_np2(x >> 1, y >> 1, z >> 1, n)] = v;
            }
}

/* chunk_make_splatlist, mesher.c:497-536: one record per non-zero cell of
 * [0, (R>>l)]^3 in z, y, x order: world position offset*R + (q << l) as int16,
 * then colour | shadow << 6 with the shadow sampled at the position, moved by
 * +(1 << l) on all three axes for l > 0.  Returns the int16 items written. */
uint32_t vpo_make_splatlist(const vp_world_desc *w, uint32_t chunk, int level,
                            const uint8_t *mask, const uint16_t *map, uint32_t size_x,
                            int16_t *geom)
{
    uint32_t b = w->root_bitw, R = 1u << b, n = R + 1, range = (R >> level) + 1, c[3];
    vpo_chunk_coords(w, chunk, c);
    uint32_t v = 0;
    for (uint32_t z = 0; z < range; z++)
        for (uint32_t y = 0; y < range; y++)
            for (uint32_t x = 0; x < range; x++) {
                uint8_t m = mask[flat_np2(x, y, z, n)];
                if (!m) continue;
                uint32_t q[3] = {x, y, z}, ws[3];
                for (int j = 0; j < 3; j++) {
                    ws[j] = (c[j] << b) + (q[j] << level);
                    geom[v++] = (int16_t)ws[j];
                }
                if (level)
                    for (int j = 0; j < 3; j++) ws[j] += 1u << level;
                geom[v++] = (int16_t)(m | (vpo_shadow_sample(map, size_x, ws) << 6));
            }
    return v;
}

/* chunkset_manage's meshing sequence for one chunk, chunkset.c:321-458:
 * mask -> L0 list; downsample mask->work (L1), work->mask (L2), mask->work
 * (L3), work->mask (L4), each target zeroed first, each followed by its list. */
uint32_t vpo_chunk_remesh(const vp_world_desc *w, const uint8_t *voxels, uint32_t chunk,
                          const uint16_t *map, uint32_t size_x, int lod,
                          int16_t *geom, uint32_t items[5],
                          uint8_t *mask, uint8_t *work)
{
    uint32_t R = 1u << w->root_bitw, msz = (R + 1) * (R + 1) * (R + 1);
    memset(mask, 0, msz);
    vpo_make_mask(w, voxels, chunk, mask);
    uint32_t total = vpo_make_splatlist(w, chunk, 0, mask, map, size_x, geom);
    items[0] = total;
    for (int l = 1; l < 5; l++) items[l] = 0;
    if (!lod) return total;
    uint8_t *src = mask, *dst = work;
    for (int l = 1; l < 5; l++) {
        memset(dst, 0, msz);
        vpo_mask_downsample(R, l, src, dst);
        items[l] = vpo_make_splatlist(w, chunk, l, dst, map, size_x, geom + total);
        total += items[l];
        uint8_t *t = src; src = dst; dst = t;
    }
    return total;
}

static uint64_t max_items_per_chunk(uint32_t R)
{
    uint64_t n = 0;
    for (int l = 0; l < 5; l++) {
        uint64_t r = (R >> l) + 1;
        n += r * r * r;
    }
    return 4 * n;
}

int64_t vpo_remesh_list(const vp_world_desc *w, const uint8_t *voxels,
                        const uint32_t *ids, uint32_t n,
                        const uint16_t *map, uint32_t size_x, int lod, int threads,
                        int16_t *out, uint64_t cap_items,
                        uint64_t *offsets, uint32_t *level_counts)
{
    uint32_t R = 1u << w->root_bitw, msz = (R + 1) * (R + 1) * (R + 1);
    uint64_t gmax = max_items_per_chunk(R);
    int16_t **lists = (int16_t **)calloc(n ? n : 1, sizeof(int16_t *));
    if (threads < 1) threads = 1;
#pragma omp parallel num_threads(threads)
    {
        uint8_t *mask = (uint8_t *)malloc(msz), *work = (uint8_t *)malloc(msz);
        int16_t *geom = (int16_t *)malloc(gmax * sizeof(int16_t));
#pragma omp for schedule(dynamic, 1)
        for (int64_t k = 0; k < (int64_t)n; k++) {
            uint32_t *it = level_counts + 5 * k;
            uint32_t tot = vpo_chunk_remesh(w, voxels, ids[k], map, size_x, lod,
                                            geom, it, mask, work);
            lists[k] = (int16_t *)malloc((tot ? tot : 1) * sizeof(int16_t));
            memcpy(lists[k], geom, tot * sizeof(int16_t));
        }
        free(mask); free(work); free(geom);
    }
    uint64_t off = 0;
    int ok = 1;
    for (uint32_t k = 0; k < n; k++) {
        offsets[k] = off;
        uint64_t tot = 0;
        for (int l = 0; l < 5; l++) tot += level_counts[5 * k + l];
        if (out && off + tot <= cap_items)
            memcpy(out + off, lists[k], tot * sizeof(int16_t));
        else if (tot)
            ok = 0;
        off += tot;
        free(lists[k]);
    }
    offsets[n] = off;
    free(lists);
    return ok ? (int64_t)off : -1;
}

double vpo_remesh_time(const vp_world_desc *w, const uint8_t *voxels,
                       const uint32_t *ids, uint32_t n,
                       const uint16_t *map, uint32_t size_x, int threads,
                       uint64_t *total_items)
{
    uint32_t R = 1u << w->root_bitw, msz = (R + 1) * (R + 1) * (R + 1);
    uint64_t gmax = max_items_per_chunk(R), total = 0;
    struct timespec t0, t1;
    if (threads < 1) threads = 1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
#pragma omp parallel num_threads(threads) reduction(+ : total)
    {
        uint8_t *mask = (uint8_t *)malloc(msz), *work = (uint8_t *)malloc(msz);
        int16_t *geom = (int16_t *)malloc(gmax * sizeof(int16_t));
        uint32_t items[5];
#pragma omp for schedule(dynamic, 1)
        for (int64_t k = 0; k < (int64_t)n; k++) {
            uint32_t tot = vpo_chunk_remesh(w, voxels, ids[k], map, size_x, 1,
                                            geom, items, mask, work);
            /* publish a private copy, as chunkset.c:475-480 does */
            int16_t *svl = (int16_t *)malloc((tot ? tot : 1) * sizeof(int16_t));
            memcpy(svl, geom, tot * sizeof(int16_t));
            total += tot;
            free(svl);
        }
        free(mask); free(work); free(geom);
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    if (total_items) *total_items = total;
    return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}

/* ========================================================================
 * Near-field quad mesh: chunk_make_mesh (src/chunkset/mesher.c:184-357),
 * with its ambient-occlusion sampling (sample_ao / sample_ao_border,
 * mesher.c:73-170) and shadow_sample / shadow_sample_normal
 * (src/chunkset/shadow.h:52-75).  A face is emitted for every voxel A of
 * the chunk and axis i whose +i neighbour B differs from A in occupancy
 * (B from the +i chunk at the edge, air beyond the world), in voxel order
 * z, y, x and then axis 0, 1, 2.  Per face: 4 vertices of int16
 * {x, y, z, attribute} and 6 u32 indices counted from the chunk's first
 * vertex.
 * ====================================================================== */

/* solid (nonzero) voxel at world coordinates, air beyond the world; the
 * coordinates are u32, so -1 wraps past the end (edit.c:10-35) */
static inline int world_solid(const vp_world_desc *w, const uint8_t *voxels,
                              uint32_t x, uint32_t y, uint32_t z)
{
    uint32_t b = w->root_bitw;
    if (x >= (1u << w->max_bitw[0]) << b || y >= (1u << w->max_bitw[1]) << b ||
        z >= (1u << w->max_bitw[2]) << b)
        return 0;
    return world_voxel(w, voxels, x, y, z) != 0;
}

static inline uint8_t world_value(const vp_world_desc *w, const uint8_t *voxels,
                                  uint32_t x, uint32_t y, uint32_t z)
{
    uint32_t b = w->root_bitw;
    if (x >= (1u << w->max_bitw[0]) << b || y >= (1u << w->max_bitw[1]) << b ||
        z >= (1u << w->max_bitw[2]) << b)
        return 0;
    return world_voxel(w, voxels, x, y, z);
}

/* shadow_sample with the reference's comparison M < y+1 in u32: at
 * y = 0xFFFFFFFF (one below the world) y+1 wraps to 0 and the sample is
 * shadowed whatever the map holds, so the map is not read there */
static int mesh_shadow(const uint16_t *map, uint32_t size_x, const uint32_t ws[3])
{
    if (ws[1] == 0xFFFFFFFFu) return 1;
    return vpo_shadow_sample(map, size_x, ws);
}

/* shadow_sample_normal, shadow.h:65-75: the entry and its -1 neighbour */
static int mesh_shadow_normal(const uint16_t *map, uint32_t size_x, const uint32_t ws[3])
{
    uint32_t i = vpo_shadow_index(size_x, ws);
    int lit = (map[i] <= ws[1]) && (map[i - 1] <= ws[1]);
    return !lit;
}

/* corners of the unit cube's +x, +y, +z faces in vertex order (mesher.c:19-31) */
static const uint8_t mesh_corner[3][4][3] = {
    {{1, 0, 0}, {1, 0, 1}, {1, 1, 1}, {1, 1, 0}},
    {{0, 1, 0}, {1, 1, 0}, {1, 1, 1}, {0, 1, 1}},
    {{1, 0, 1}, {0, 0, 1}, {0, 1, 1}, {1, 1, 1}},
};
/* triangle indices by [normal][rotated] (mesher.c:54-66) */
static const uint8_t mesh_tri[2][2][6] = {
    {{0, 3, 1, 2, 1, 3}, {3, 2, 0, 1, 0, 2}},
    {{1, 3, 0, 3, 1, 2}, {0, 2, 3, 2, 0, 1}},
};

uint32_t vpo_chunk_make_mesh(const vp_world_desc *w, const uint8_t *voxels, uint32_t chunk,
                             const uint16_t *map, uint32_t size_x,
                             int16_t *geom, uint32_t *index)
{
    uint32_t b = w->root_bitw, R = 1u << b, cc[3];
    vpo_chunk_coords(w, chunk, cc);
    const uint32_t org[3] = {cc[0] << b, cc[1] << b, cc[2] << b};
    uint32_t nf = 0;
    for (uint32_t z = 0; z < R; z++)
        for (uint32_t y = 0; y < R; y++)
            for (uint32_t x = 0; x < R; x++) {
                const uint32_t wa[3] = {org[0] + x, org[1] + y, org[2] + z};
                const uint8_t A = world_value(w, voxels, wa[0], wa[1], wa[2]);
                for (int i = 0; i < 3; i++) {
                    uint32_t wb[3] = {wa[0], wa[1], wa[2]};
                    wb[i]++;
                    const uint8_t B = world_value(w, voxels, wb[0], wb[1], wb[2]);
                    if ((A == 0) == (B == 0)) continue;
                    const int normal = (A == 0);
                    /* the air cell of the pair and the 8 cells around it in
                     * the plane normal to axis i */
                    uint32_t air[3] = {wa[0], wa[1], wa[2]};
                    air[i] += (B == 0);
                    const int u0 = i == 0 ? 1 : 0, u1 = i == 2 ? 1 : 2;
                    int aon[4], aoc[4];
                    {
                        uint32_t q[3];
#define SOL(d0, d1) (q[0] = air[0], q[1] = air[1], q[2] = air[2], q[u0] += (uint32_t)(d0), \
                     q[u1] += (uint32_t)(d1), world_solid(w, voxels, q[0], q[1], q[2]))
                        aon[0] = SOL(-1, 0);
                        aon[2] = SOL(+1, 0);
                        aon[1] = SOL(0, -1);
                        aon[3] = SOL(0, +1);
                        aoc[0] = SOL(-1, -1);
                        aoc[1] = SOL(+1, -1);
                        aoc[2] = SOL(+1, +1);
                        aoc[3] = SOL(-1, +1);
#undef SOL
                    }
                    int cv[4], vao[4];
                    for (int k = 0; k < 4; k++) cv[k] = (aon[k] + aon[(k + 1) & 3]) | aoc[k];
                    if (i == 0) {
                        vao[0] = cv[0]; vao[3] = cv[1]; vao[2] = cv[2]; vao[1] = cv[3];
                    } else if (i == 1) {
                        vao[0] = cv[0]; vao[1] = cv[1]; vao[2] = cv[2]; vao[3] = cv[3];
                    } else {
                        vao[1] = cv[0]; vao[0] = cv[1]; vao[3] = cv[2]; vao[2] = cv[3];
                    }
                    const int rot = (vao[0] + vao[2] < vao[1] + vao[3]);
                    int shadow = 0, diamond = 3;
                    if (i == 0 && normal) {
                        shadow = mesh_shadow_normal(map, size_x, wb);   /* the solid cell */
                    } else if (i == 1 && !normal) {
                        shadow = mesh_shadow(map, size_x, wa);
                    } else if (i == 2) {
                        uint32_t s[3] = {wa[0], wa[1], wa[2]};
                        if (!B) s[2] += 1;
                        diamond = (!mesh_shadow(map, size_x, s)) << 1;
                        s[1] -= 1;
                        diamond |= !mesh_shadow(map, size_x, s);
                    }
                    for (int v = 0; v < 4; v++) {
                        int16_t *g = geom + 16 * (size_t)nf + 4 * v;
                        for (int a = 0; a < 3; a++)
                            g[a] = (int16_t)(wa[a] + mesh_corner[i][v][a]);
                        g[3] = (int16_t)((A | B) | (vao[v] << 6) | ((i + 3 * normal) << 8) |
                                         (v << 11) | (shadow << 13) | (diamond << 14));
                    }
                    for (int k = 0; k < 6; k++)
                        index[6 * (size_t)nf + k] = 4 * nf + mesh_tri[normal][rot][k];
                    nf++;
                }
            }
    return nf;
}

int64_t vpo_mesh_list(const vp_world_desc *w, const uint8_t *voxels, const uint32_t *ids,
                      uint32_t n, const uint16_t *map, uint32_t size_x, int threads,
                      int16_t *geom_out, uint32_t *index_out, uint64_t cap_faces,
                      uint64_t *face_offsets)
{
    uint32_t R = 1u << w->root_bitw;
    uint64_t fmax = 3ull * R * R * R;
    int16_t **gl = (int16_t **)calloc(n ? n : 1, sizeof(int16_t *));
    uint32_t **il = (uint32_t **)calloc(n ? n : 1, sizeof(uint32_t *));
    uint32_t *nfs = (uint32_t *)calloc(n ? n : 1, sizeof(uint32_t));
    if (threads < 1) threads = 1;
#pragma omp parallel num_threads(threads)
    {
        int16_t *g = (int16_t *)malloc(fmax * 16 * sizeof(int16_t));
        uint32_t *ix = (uint32_t *)malloc(fmax * 6 * sizeof(uint32_t));
#pragma omp for schedule(dynamic, 1)
        for (int64_t k = 0; k < (int64_t)n; k++) {
            uint32_t nf = vpo_chunk_make_mesh(w, voxels, ids[k], map, size_x, g, ix);
            nfs[k] = nf;
            gl[k] = (int16_t *)malloc((nf ? nf : 1) * 16 * sizeof(int16_t));
            il[k] = (uint32_t *)malloc((nf ? nf : 1) * 6 * sizeof(uint32_t));
            memcpy(gl[k], g, (size_t)nf * 16 * sizeof(int16_t));
            memcpy(il[k], ix, (size_t)nf * 6 * sizeof(uint32_t));
        }
        free(g); free(ix);
    }
    uint64_t off = 0;
    int ok = 1;
    for (uint32_t k = 0; k < n; k++) {
        face_offsets[k] = off;
        if (geom_out && off + nfs[k] <= cap_faces) {
            memcpy(geom_out + 16 * off, gl[k], (size_t)nfs[k] * 16 * sizeof(int16_t));
            memcpy(index_out + 6 * off, il[k], (size_t)nfs[k] * 6 * sizeof(uint32_t));
        } else if (nfs[k]) {
            ok = 0;
        }
        off += nfs[k];
        free(gl[k]); free(il[k]);
    }
    face_offsets[n] = off;
    free(gl); free(il); free(nfs);
    return ok ? (int64_t)off : -1;
}

/* ========================================================================
 * RLE, src/chunkset/rle.c:44-116
 * ====================================================================== */
uint32_t vpo_rle_compress(const uint8_t *data, uint32_t len, uint32_t *words, uint32_t cap)
{
    uint32_t b = 0, count = 1;
    uint8_t  last = data[0];
    for (uint32_t i = 1; i < len; i++) {
        uint8_t cur = data[i];
        if (cur == last && count < 0xFFFFFFu) {
            count++;
            continue;
        }
        if (b + 2 > cap) return 0;
        words[b++] = count | ((uint32_t)last << 24);
        count = 1;
        last = cur;
    }
    if (b + 2 > cap) return 0;
    words[b++] = count | ((uint32_t)last << 24);
    words[b++] = 0;
    return b;
}

/* rle_decompress, rle.c:90-116: expand words until the first 0 word (the
 * first word is always expanded).  The reference bounds nothing; here the
 * output stops at cap and the full expanded length is returned. */
int64_t vpo_rle_decompress(const uint32_t *words, uint32_t nwords, uint8_t *out, uint32_t cap)
{
    if (nwords == 0) return -1;
    uint64_t b = 0;
    uint32_t i = 0;
    do {
        uint32_t count = words[i] & 0xFFFFFFu;
        uint8_t  item = (uint8_t)(words[i] >> 24);
        for (uint32_t k = 0; k < count; k++)
            if (b + k < cap) out[b + k] = item;
        b += count;
        i++;
    } while (i < nwords && words[i]);
    return (int64_t)b;
}

/* ========================================================================
 * Sphere edit, src/chunkset/edit.c:89-244 with chunk_ws_write's y<2 guard
 * (edit.c:151) and shadow_place_update for added voxels.
 * ====================================================================== */
void vpo_edit_sphere(const vp_world_desc *w, uint8_t *voxels, uint16_t *map,
                     uint32_t size_x, const int32_t centre[3], uint32_t radius,
                     uint8_t voxel, uint8_t *dirty)
{
    int32_t b = w->root_bitw, R = 1 << b;
    int32_t lo[3], hi[3], cmin[3], cmax[3], mx[3];
    for (int i = 0; i < 3; i++) {
        lo[i] = centre[i] - (int32_t)radius - 1;   /* edit.c:190-193 */
        hi[i] = centre[i] + (int32_t)radius + 1;
        cmin[i] = lo[i] >> b;                      /* edit.c:98-101 */
        cmax[i] = hi[i] >> b;
        mx[i] = 1 << w->max_bitw[i];
    }
    int64_t r2 = (int64_t)radius * radius;
    /* chunk list order: cx outer, cy, cz inner (edit.c:110-112) */
    for (int32_t cx = cmin[0]; cx <= cmax[0]; cx++)
        for (int32_t cy = cmin[1]; cy <= cmax[1]; cy++)
            for (int32_t cz = cmin[2]; cz <= cmax[2]; cz++) {
                if (cx < 0 || cy < 0 || cz < 0 || cx >= mx[0] || cy >= mx[1] || cz >= mx[2])
                    continue;
                uint32_t cc[3] = {(uint32_t)cx, (uint32_t)cy, (uint32_t)cz};
                uint32_t ci = vpo_chunk_index(w, cc);
                uint8_t *v = voxels + ((uint64_t)ci << (3 * b));
                int32_t base[3] = {cx * R, cy * R, cz * R};
                /* u order x, y, z (edit.c:213-215) */
                for (int32_t ux = lo[0]; ux < hi[0]; ux++)
                    for (int32_t uy = lo[1]; uy < hi[1]; uy++)
                        for (int32_t uz = lo[2]; uz < hi[2]; uz++) {
                            int64_t dx = ux - centre[0], dy = uy - centre[1], dz = uz - centre[2];
                            /* float distance < radius  <=>  d^2 < r^2 for integers */
                            if (dx * dx + dy * dy + dz * dz >= r2) continue;
                            if (ux < base[0] || ux >= base[0] + R || uy < base[1] ||
                                uy >= base[1] + R || uz < base[2] || uz >= base[2] + R)
                                continue;                         /* chunk_ws_inside */
                            if (uy >= 2) {                        /* chunk_ws_write */
                                uint32_t l[3] = {(uint32_t)(ux - base[0]), (uint32_t)(uy - base[1]),
                                                 (uint32_t)(uz - base[2])};
                                v[((l[2] << b | l[1]) << b) | l[0]] = voxel;
                            }
                            if (voxel && map) {
                                uint32_t ws[3] = {(uint32_t)ux, (uint32_t)uy, (uint32_t)uz};
                                vpo_shadow_place(map, size_x, ws);
                            }
                        }
                if (dirty) dirty[ci] = 1;   /* edit.c:241-242 */
            }
}

/* ========================================================================
 * Splat resolve: a CPU restatement of splat.vert.glsl:34-54 and
 * splat.frag.glsl:25-156 with GL's point clipping, point-sprite coverage
 * and GL_LESS depth test (game.c:537-538) made explicit.
 * ====================================================================== */

/* GLSL mat4 * vec4 with a column-major matrix, summed left to right */
static void mat4_mul_vec4(const float m[16], const float v[4], float o[4])
{
    for (int r = 0; r < 4; r++) {
        float s = m[0 * 4 + r] * v[0];
        s = s + m[1 * 4 + r] * v[1];
        s = s + m[2 * 4 + r] * v[2];
        s = s + m[3 * 4 + r] * v[3];
        o[r] = s;
    }
}

static inline float glsl_min(float x, float y) { return (y < x) ? y : x; }
static inline float glsl_max(float x, float y) { return (x < y) ? y : x; }

typedef struct {
    float vxl[3];     /* view-translated centre (vVert.xyz) */
    float half_lod;   /* 0.5 * u_lod */
    float lodf;
    int   x0, x1, y0, y1;   /* covered pixel range, inclusive; empty if x0 > x1 */
    float lo_x, hi_x, lo_y, hi_y;
} splat_vs;

/* Vertex stage: returns 0 when GL would clip the point. */
static int splat_vertex(const int16_t *rec, float lodf, const vp_camera *cam, splat_vs *o)
{
    float half = 0.5f * lodf;
    float a[4] = {(float)rec[0] + half, (float)rec[1] + half, (float)rec[2] + half, 1.0f};
    float v[4], clip[4];
    mat4_mul_vec4(cam->view_translation, a, v);
    mat4_mul_vec4(cam->proj_view_rot, v, clip);
    float W = (float)cam->width, H = (float)cam->height;
    /* GL point clipping: the point survives iff its vertex is in the volume */
    float cw = clip[3];
    if (!(clip[0] >= -cw && clip[0] <= cw && clip[1] >= -cw && clip[1] <= cw &&
          clip[2] >= -cw && clip[2] <= cw))
        return 0;
    float ratio = W / H;
    float rx = fabsf((clip[0] * ratio) / cw);
    float ry = fabsf(clip[1] / cw);
    float reduce = glsl_max(rx, ry);
    reduce = reduce + 0.03f;
    float size = (H * 1.1f) / clip[2];
    size = size * glsl_max(reduce, 1.0f);
    float ps = size * lodf;
    ps = (ps > 1.0f) ? ps : 1.0f;          /* also maps NaN to 1 */
    ps = (ps < 2048.0f) ? ps : 2048.0f;
    float ndx = clip[0] / cw, ndy = clip[1] / cw;
    float wx = (ndx + 1.0f) * (0.5f * W);
    float wy = (ndy + 1.0f) * (0.5f * H);
    float hs = 0.5f * ps;
    o->lo_x = wx - hs; o->hi_x = wx + hs;
    o->lo_y = wy - hs; o->hi_y = wy + hs;
    /* pixel p is covered iff lo <= p + 0.5 < hi; start one pixel early and
     * let the exact test decide */
    int x0 = (int)floorf(o->lo_x) - 1, x1 = (int)ceilf(o->hi_x) + 1;
    int y0 = (int)floorf(o->lo_y) - 1, y1 = (int)ceilf(o->hi_y) + 1;
    o->x0 = x0 < 0 ? 0 : x0;
    o->y0 = y0 < 0 ? 0 : y0;
    o->x1 = x1 > (int)cam->width - 1 ? (int)cam->width - 1 : x1;
    o->y1 = y1 > (int)cam->height - 1 ? (int)cam->height - 1 : y1;
    o->vxl[0] = v[0]; o->vxl[1] = v[1]; o->vxl[2] = v[2];
    o->half_lod = half;
    o->lodf = lodf;
    return 1;
}

/* ss2wsVec, splat.frag.glsl:48-58 */
static void pixel_ray(const vp_camera *cam, int px, int py, float dir[3])
{
    float fx = ((float)px + 0.5f) / (float)cam->width;
    float fy = ((float)py + 0.5f) / (float)cam->height;
    float c[4] = {fx * 2.0f - 1.0f, fy * 2.0f - 1.0f, -1.0f, 1.0f}, r[4];
    mat4_mul_vec4(cam->inv_proj_view_rot, c, r);
    float x = r[0] / r[3], y = r[1] / r[3], z = r[2] / r[3];
    float d = x * x;
    d = d + y * y;
    d = d + z * z;
    float len = sqrtf(d);
    dir[0] = x / len; dir[1] = y / len; dir[2] = z / len;
}

/* AABBIntersect from the origin, splat.frag.glsl:25-46 */
static void aabb_hit(const float vxl[3], float half, const float inv[3], float *t0, float *t1)
{
    float tmin[3], tmax[3];
    for (int k = 0; k < 3; k++) {
        float mn = vxl[k] - half, mxv = vxl[k] + half;
        float tb = mn * inv[k], tt = mxv * inv[k];
        tmin[k] = glsl_min(tt, tb);
        tmax[k] = glsl_max(tt, tb);
    }
    float a = glsl_max(tmin[0], tmin[1]), b = glsl_max(tmin[0], tmin[2]);
    *t0 = glsl_max(a, b);
    a = glsl_min(tmax[0], tmax[1]); b = glsl_min(tmax[0], tmax[2]);
    *t1 = glsl_min(a, b);
}

static inline uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

static float glsl_sign(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }

/* Rasterise splats [i0, i1) into best[] (packed depth|id keys, min wins). */
static void splat_range(const int16_t *splats, const uint8_t *lod, uint32_t lod_log2,
                        uint64_t i0, uint64_t i1, const vp_camera *cam, const float *rays,
                        uint64_t *best)
{
    uint32_t W = cam->width;
    for (uint64_t i = i0; i < i1; i++) {
        float lodf = (float)(1u << (lod ? lod[i] : lod_log2));
        splat_vs vs;
        if (!splat_vertex(splats + 4 * i, lodf, cam, &vs)) continue;
        for (int py = vs.y0; py <= vs.y1; py++) {
            float cy = (float)py + 0.5f;
            if (!(vs.lo_y <= cy && cy < vs.hi_y)) continue;
            for (int px = vs.x0; px <= vs.x1; px++) {
                float cx = (float)px + 0.5f;
                if (!(vs.lo_x <= cx && cx < vs.hi_x)) continue;
                uint64_t pi = (uint64_t)py * W + (uint32_t)px;
                const float *r = rays + 6 * pi;
                float t0, t1;
                aabb_hit(vs.vxl, vs.half_lod, r + 3, &t0, &t1);
                if (!(t0 <= t1)) continue;
                float d = t0 / cam->far_plane;
                if (!(d < 1.0f)) continue;          /* GL_LESS against the 1.0 clear */
                if (d <= 0.0f) d = 0.0f;            /* depth clamp; also folds -0 */
                uint64_t key = ((uint64_t)f2u(d) << 32) | (uint32_t)i;
                if (key < best[pi]) best[pi] = key;
            }
        }
    }
}

/* threads > 1: each thread rasterises a contiguous range of splats into a
 * private frame and the frames are merged by the minimum key.  The key
 * (depth bits, index) orders hits exactly as the serial loop's depth test
 * with ties to the lower index does, so the result does not depend on the
 * thread count. */
void vpo_splat_render(const int16_t *splats, const uint8_t *lod, uint32_t lod_log2,
                      uint64_t n, const vp_camera *cam, const float *voxel_rgb,
                      float *depth, uint32_t *id, float *rgb, float *weight, int threads)
{
    uint32_t W = cam->width, H = cam->height;
    uint64_t np = (uint64_t)W * H;
    float *rays = (float *)malloc(sizeof(float) * 6 * np);
    if (threads < 1) threads = 1;
    if ((uint64_t)threads > n / 4096 + 1) threads = (int)(n / 4096 + 1);
    uint64_t *frames = (uint64_t *)malloc(sizeof(uint64_t) * np * (uint64_t)threads);
    uint64_t *best = frames;
#pragma omp parallel for num_threads(threads) schedule(static)
    for (uint32_t py = 0; py < H; py++)
        for (uint32_t px = 0; px < W; px++) {
            float *r = rays + 6 * ((uint64_t)py * W + px);
            pixel_ray(cam, (int)px, (int)py, r);
            r[3] = 1.0f / r[0]; r[4] = 1.0f / r[1]; r[5] = 1.0f / r[2];
        }
#pragma omp parallel num_threads(threads)
    {
        int t = omp_get_thread_num(), nt = omp_get_num_threads();
        uint64_t *mine = frames + np * (uint64_t)t;
        for (uint64_t p = 0; p < np; p++) mine[p] = ~0ull;
        uint64_t i0 = n * (uint64_t)t / (uint64_t)nt, i1 = n * (uint64_t)(t + 1) / (uint64_t)nt;
        splat_range(splats, lod, lod_log2, i0, i1, cam, rays, mine);
#pragma omp barrier
#pragma omp for schedule(static)
        for (uint64_t p = 0; p < np; p++) {
            uint64_t k = frames[p];
            for (int u = 1; u < nt; u++) {
                uint64_t v = frames[np * (uint64_t)u + p];
                if (v < k) k = v;
            }
            best[p] = k;
        }
    }
    const float light[3] = {-0.70710677f, 0.70710677f, 0.0f};  /* normalize(-1,1,0) */
#pragma omp parallel for num_threads(threads) schedule(static)
    for (uint64_t p = 0; p < np; p++) {
        uint64_t k = best[p];
        if (k == ~0ull) {
            depth[p] = 1.0f;
            id[p] = 0xFFFFFFFFu;
            if (rgb) { rgb[3 * p] = 0.5f; rgb[3 * p + 1] = 0.5f; rgb[3 * p + 2] = 0.7f; }
            if (weight) weight[p] = 0.0f;
            continue;
        }
        uint32_t i = (uint32_t)k;
        float d = u2f((uint32_t)(k >> 32));
        depth[p] = d;
        id[p] = i;
        if (!rgb && !weight) continue;
        /* shading, splat.vert.glsl:24-32 and splat.frag.glsl:93-154 */
        float lodf = (float)(1u << (lod ? lod[i] : lod_log2));
        splat_vs vs;
        splat_vertex(splats + 4 * i, lodf, cam, &vs);
        const float *r = rays + 6 * p;
        float t0, t1;
        aabb_hit(vs.vxl, vs.half_lod, r + 3, &t0, &t1);
        float hit[3], ha[3];
        for (int c = 0; c < 3; c++) {
            hit[c] = vs.vxl[c] - r[c] * t0;
            ha[c] = fabsf(hit[c]);
        }
        float md = glsl_max(glsl_max(ha[0], ha[1]), ha[2]);
        float nrm[3];
        for (int c = 0; c < 3; c++) nrm[c] = (ha[c] == md ? 1.0f : 0.0f) * glsl_sign(hit[c]);
        float nd = (-nrm[0]) * light[0];
        nd = nd + (-nrm[1]) * light[1];
        nd = nd + (-nrm[2]) * light[2];
        int attr = splats[4 * i + 3];
        float sh = (float)((attr >> 6) & 1);
        float s1 = glsl_max(0.7f, 1.0f - sh);
        float s2 = glsl_max(glsl_max(0.7f, sh), glsl_sign(nd));
        float base[3];
        if (voxel_rgb) {
            for (int c = 0; c < 3; c++) base[c] = voxel_rgb[3 * (uint64_t)i + c];
        } else {
            base[0] = (float)((attr >> 4) & 3) / 3.0f;
            base[1] = (float)((attr >> 2) & 3) / 3.0f;
            base[2] = (float)(attr & 3) / 3.0f;
        }
        /* pow(x, 0.5) of splat.frag.glsl:144 is restated as sqrt */
        float fs = 256.0f / cam->far_plane;
        float fog = sqrtf(glsl_max(d - fs, 0.0001f));
        float keep = 1.0f - fog;
        const float fogc[3] = {0.5f, 0.5f, 0.7f};
        if (rgb)
            for (int c = 0; c < 3; c++) {
                float vcol = base[c] * s1;          /* vColor, splat.vert.glsl:31 */
                float col = vcol * s2;              /* splat.frag.glsl:130 */
                rgb[3 * p + c] = col * keep + fogc[c] * fog;   /* mix() */
            }
        if (weight) weight[p] = (s1 * s2) * keep;
    }
    free(rays);
    free(frames);
}

void vpo_splat_backward(uint64_t n, const vp_camera *cam, const uint32_t *id,
                        const float *weight, const float *dL_drgb, double *grad)
{
    uint64_t np = (uint64_t)cam->width * cam->height;
    for (uint64_t p = 0; p < np; p++) {
        uint32_t i = id[p];
        if (i == 0xFFFFFFFFu || i >= n) continue;
        for (int c = 0; c < 3; c++)
            grad[3 * (uint64_t)i + c] += (double)weight[p] * (double)dL_drgb[3 * p + c];
    }
}
