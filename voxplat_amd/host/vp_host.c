/*
 * vp_host.c -- host-side C of the pipeline: the order-dependent shadow-map
 * updates of sphere edits, the remesh drivers (plain per-chunk loop,
 * launch-shape autotuner) and the per-frame draw
 * queue (frustum planes, octree LOD selection) declared in vp_host.h.
 *
 * shadow_place_update (src/chunkset/shadow.h:77-89) reads map[i] and map[i+1]
 * before writing map[i], so the result depends on the order of the placed
 * voxels; chunkset_edit_sphere (src/chunkset/edit.c:179-244) places every
 * voxel of an add-sphere that lies inside an in-world chunk of the sphere's
 * box, chunk by chunk (cx outer, then cy, cz) and inside a chunk in u order
 * (x outer, then y, z).  The updates are cheap and strictly sequential, so
 * they run here, on a host copy of the map.
 */
#include <math.h>
#include <stdint.h>
#include <stddef.h>

#include "vp_gpu.h"
#include "vp_workload.h"

static void place(uint16_t *map, uint32_t sx, uint32_t wx, uint32_t wy, uint32_t wz)
{
    const uint32_t i = (wx + wy) + sx * wz;
    if (!(map[i] <= wy && map[i + 1] <= wy)) return;   /* shadowed: no update */
    map[i] = (uint16_t)wy;
}

int vp_host_shadow_spheres(const vp_world_desc *world, uint16_t *map,
                           const vp_sphere *spheres, uint32_t n)
{
    if (!world || !map || (n && !spheres)) return VP_ERR_ARG;
    const int32_t b = world->root_bitw, R = 1 << b;
    int32_t mx[3];
    for (int a = 0; a < 3; a++) mx[a] = 1 << world->max_bitw[a];
    const uint32_t sx = (uint32_t)((mx[0] << b) + (mx[1] << b));
    for (uint32_t s = 0; s < n; s++) {
        const vp_sphere *sp = &spheres[s];
        if (!sp->voxel) continue;                        /* shadow_break_update: stub */
        const int32_t c[3] = {sp->cx, sp->cy, sp->cz}, r = (int32_t)sp->radius;
        const int64_t r2 = (int64_t)r * r;
        int32_t lo[3], hi[3], cl[3], ch[3];
        for (int a = 0; a < 3; a++) {
            lo[a] = c[a] - r - 1;
            hi[a] = c[a] + r + 1;                        /* exclusive */
            cl[a] = lo[a] >> b;
            ch[a] = hi[a] >> b;
        }
        for (int32_t cx = cl[0]; cx <= ch[0]; cx++)
            for (int32_t cy = cl[1]; cy <= ch[1]; cy++)
                for (int32_t cz = cl[2]; cz <= ch[2]; cz++) {
                    if (cx < 0 || cy < 0 || cz < 0 || cx >= mx[0] || cy >= mx[1] || cz >= mx[2])
                        continue;
                    const int32_t bx = cx * R, by = cy * R, bz = cz * R;
                    /* only u inside this chunk can pass chunk_ws_inside */
                    const int32_t x0 = lo[0] > bx ? lo[0] : bx, x1 = hi[0] < bx + R ? hi[0] : bx + R;
                    const int32_t y0 = lo[1] > by ? lo[1] : by, y1 = hi[1] < by + R ? hi[1] : by + R;
                    const int32_t z0 = lo[2] > bz ? lo[2] : bz, z1 = hi[2] < bz + R ? hi[2] : bz + R;
                    for (int32_t ux = x0; ux < x1; ux++)
                        for (int32_t uy = y0; uy < y1; uy++)
                            for (int32_t uz = z0; uz < z1; uz++) {
                                const int64_t dx = ux - c[0], dy = uy - c[1], dz = uz - c[2];
                                if (dx * dx + dy * dy + dz * dz >= r2) continue;
                                place(map, sx, (uint32_t)ux, (uint32_t)uy, (uint32_t)uz);
                            }
                }
    }
    return VP_OK;
}

/* ------------------------------------------------------------------------
 * Remesh drivers: the plain per-chunk loop and the launch-shape autotuner
 * (the captured batched frame is in csrc/remesh_graph.hip).
 * ---------------------------------------------------------------------- */
#include <stdlib.h>
#include <string.h>
#include <hip/hip_runtime_api.h>

#include "vp_host.h"

int vp_host_remesh_per_chunk(vp_ctx *ctx, const vp_world_desc *world,
                             const uint8_t *d_voxels, const int16_t *d_uniform,
                             const vp_shadow_desc *shadow, const uint32_t *d_chunk_ids,
                             uint32_t n, uint32_t flags, void *d_splats,
                             const uint64_t *rec_offsets, const uint64_t *rec_caps,
                             uint64_t *d_offsets2, uint32_t *d_level_counts, void *stream)
{
    if (!ctx || !world || !d_chunk_ids || !d_offsets2 || !d_level_counts) return VP_ERR_ARG;
    if (!(flags & VP_CM_COUNT_ONLY) && (!d_splats || !rec_offsets || !rec_caps)) return VP_ERR_ARG;
    for (uint32_t k = 0; k < n; k++) {
        vp_splat_out o;
        o.d_splats = (flags & VP_CM_COUNT_ONLY) ? NULL
                                                : (void *)((uint64_t *)d_splats + rec_offsets[k]);
        o.capacity_records = (flags & VP_CM_COUNT_ONLY) ? 0 : rec_caps[k];
        o.d_chunk_offsets = d_offsets2 + 2 * (uint64_t)k;
        o.d_level_counts = d_level_counts + 5 * (uint64_t)k;
        int rc = vp_gpu_cull_mesh(ctx, world, d_voxels, d_uniform, shadow, d_chunk_ids + k, 1,
                                  NULL, flags, &o, stream);
        if (rc) return rc;
    }
    return VP_OK;
}

static int cmp_float(const void *a, const void *b)
{
    float x = *(const float *)a, y = *(const float *)b;
    return (x > y) - (x < y);
}

int vp_host_autotune_remesh(vp_ctx *ctx, const vp_world_desc *world,
                            const uint8_t *d_voxels, const int16_t *d_uniform,
                            const vp_shadow_desc *shadow, const uint32_t *d_chunk_ids,
                            uint32_t n, uint32_t flags, const vp_splat_out *out, int reps,
                            void *stream, vp_remesh_shape *best, float *best_ms,
                            int *shapes_timed)
{
    static const int blocks[3] = {256, 512, 1024};
    if (!ctx || !world || !best || reps < 1 || reps > 64) return VP_ERR_ARG;
    /* the shape the calls below run with is the world's own: remember the
     * one in force, to restore it if a call fails */
    vp_remesh_shape keep;
    int rc = vp_ctx_get_remesh_shape(ctx, world, &keep);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return VP_ERR_HIP;
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return VP_ERR_HIP;
    }
    const int nu = world->root_bitw == 6 ? 2 : 1;
    float bestt = 1e30f, t[64];
    vp_remesh_shape bs = keep;
    int timed = 0;
    for (int bi = 0; bi < 3 && rc == VP_OK; bi++)
        for (int ui = 0; ui < nu && rc == VP_OK; ui++) {
            vp_remesh_shape s = {blocks[bi], 1, ui ? 2 : 4};
            int resident = 0;
            rc = vp_ctx_remesh_resident(ctx, world, flags, &s, &resident);
            if (rc) break;
            if (resident > 4) resident = 4;
            for (int wg = 1; wg <= resident && rc == VP_OK; wg++) {
                s.wg_per_cu = wg;
                rc = vp_ctx_set_remesh_shape(ctx, world, &s);
                if (rc) break;
                rc = vp_gpu_cull_mesh(ctx, world, d_voxels, d_uniform, shadow, d_chunk_ids, n,
                                      NULL, flags, out, stream);   /* warm-up */
                for (int r = 0; r < reps && rc == VP_OK; r++) {
                    if (hipEventRecord(e0, st) != hipSuccess) { rc = VP_ERR_HIP; break; }
                    rc = vp_gpu_cull_mesh(ctx, world, d_voxels, d_uniform, shadow, d_chunk_ids,
                                          n, NULL, flags, out, stream);
                    if (hipEventRecord(e1, st) != hipSuccess ||
                        hipEventSynchronize(e1) != hipSuccess ||
                        hipEventElapsedTime(&t[r], e0, e1) != hipSuccess)
                        rc = VP_ERR_HIP;
                }
                if (rc) break;
                qsort(t, (size_t)reps, sizeof(float), cmp_float);
                timed++;
                if (t[reps / 2] < bestt) {
                    bestt = t[reps / 2];
                    bs = s;
                }
            }
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) {
        (void)vp_ctx_set_remesh_shape(ctx, world, &keep);
        return rc;
    }
    if (timed == 0) return VP_ERR_UNSUPPORTED;
    *best = bs;
    if (best_ms) *best_ms = bestt;
    if (shapes_timed) *shapes_timed = timed;
    return vp_ctx_set_remesh_shape(ctx, world, &bs);
}

/* ---- frame draw queue -------------------------------------------------------
 * gfx_fcull_extract_planes (src/gfx/fcull.c:18-36): rows w - x, w + x,
 * w - y, w + y, w - z, w + z of the column-major view-projection, each
 * divided by the length of its xyz part (a float sum of squares, square root
 * in double, as vec3length computes it). */
void vp_host_frustum_planes(const float m[16], float planes[24])
{
    int n = 0;
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 4; j++) planes[n++] = m[3 + 4 * j] - m[4 * j + i];
        for (int j = 0; j < 4; j++) planes[n++] = m[3 + 4 * j] + m[4 * j + i];
    }
    for (int i = 0; i < 6; i++) {
        float *p = planes + 4 * i;
        const float sq = p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
        const float len = (float)sqrt((double)sq);
        p[0] /= len;
        p[1] /= len;
        p[2] /= len;
        p[3] /= len;
    }
}

typedef struct {
    const vp_world_desc *w;
    const float *planes;
    const float *origin;
    const uint64_t *const *node_off;
    const uint8_t *chunk_mesh;
    vp_draw *draws;
    uint32_t draw_cap, n_draws;
    uint32_t *mesh;
    uint32_t mesh_cap, n_mesh;
    uint64_t total;
} cull_state;

/* gfx_fcull_visible (fcull.c:38-45): outside iff some plane's signed
 * distance is below -radius */
static int sphere_visible(const float *planes, const float c[3], float radius)
{
    for (int i = 0; i < 6; i++) {
        const float *p = planes + 4 * i;
        const float d = (p[0] * c[0] + p[1] * c[1]) + p[2] * c[2];
        if (d + p[3] < -radius) return 0;
    }
    return 1;
}

/* octree_cull (src/game.c:319-395) */
static void cull_node(cull_state *s, uint32_t lod, uint32_t x, uint32_t y, uint32_t z)
{
    const vp_world_desc *w = s->w;
    const uint32_t ws_root = (uint32_t)(1u << w->root_bitw) << lod;
    if ((x << lod) >= (1u << w->max_bitw[0]) || (y << lod) >= (1u << w->max_bitw[1]) ||
        (z << lod) >= (1u << w->max_bitw[2]))
        return;
    const float c[3] = {(float)(x * ws_root + ws_root / 2), (float)(y * ws_root + ws_root / 2),
                        (float)(z * ws_root + ws_root / 2)};
    if (!sphere_visible(s->planes, c, (float)((double)ws_root * 0.85))) return;
    const float d0 = s->origin[0] - c[0], d1 = s->origin[1] - c[1], d2 = s->origin[2] - c[2];
    const float dist = sqrtf((d0 * d0 + d1 * d1) + d2 * d2);
    /* A level-0 node is always a leaf.  The reference would visit level
     * 255 (uint8 lod - 1, game.c:386) when the camera sits on a chunk
     * centre, where dist > 0 * 1024 fails. */
    if (dist > (float)(lod * 1024u) || lod == 0) {
        if (lod == 0 && s->chunk_mesh) {
            const uint32_t chunk = ((z << w->max_bitw[1] | y) << w->max_bitw[0]) | x;
            if (s->chunk_mesh[chunk]) {
                if (s->n_mesh < s->mesh_cap) s->mesh[s->n_mesh] = chunk;
                s->n_mesh++;
                return;
            }
        }
        const uint32_t b0 = vp_node_bits(w, lod, 0), b1 = vp_node_bits(w, lod, 1);
        const uint32_t node = ((z << b1 | y) << b0) | x;
        const uint64_t items = s->node_off[lod][node + 1] - s->node_off[lod][node];
        if (!items) return;                    /* no vertex buffer */
        if (s->n_draws < s->draw_cap) {
            vp_draw *d = s->draws + s->n_draws;
            d->level = lod;
            d->node = node;
            d->first = s->total;
        }
        s->n_draws++;
        s->total += items;
        return;
    }
    for (uint32_t nx = 0; nx < 2; nx++)
        for (uint32_t ny = 0; ny < 2; ny++)
            for (uint32_t nz = 0; nz < 2; nz++)
                cull_node(s, lod - 1, 2 * x + nx, 2 * y + ny, 2 * z + nz);
}

int vp_host_octree_cull(const vp_world_desc *world, const float planes[24], const float origin[3],
                        const uint64_t *const node_offsets[5], const uint8_t *chunk_mesh,
                        vp_draw *draws, uint32_t draw_cap, uint32_t *n_draws,
                        uint32_t *mesh_chunks, uint32_t mesh_cap, uint32_t *n_mesh,
                        uint64_t *total_records)
{
    if (!world || !planes || !origin || !node_offsets || !n_draws || !n_mesh) return VP_ERR_ARG;
    if ((draw_cap && !draws) || (mesh_cap && !mesh_chunks)) return VP_ERR_ARG;
    for (int l = 0; l < 5; l++)
        if (!node_offsets[l]) return VP_ERR_ARG;
    cull_state s = {world, planes, origin, node_offsets, chunk_mesh, draws, draw_cap, 0,
                    mesh_chunks, mesh_cap, 0, 0};
    /* every level-4 node, x outer, y, z inner (game.c:507-512) */
    const uint32_t mx = 1u << world->max_bitw[0], my = 1u << world->max_bitw[1],
                   mz = 1u << world->max_bitw[2];
    for (uint32_t x = 0; x < (mx >> 4) + 1; x++)
        for (uint32_t y = 0; y < (my >> 4) + 1; y++)
            for (uint32_t z = 0; z < (mz >> 4) + 1; z++) cull_node(&s, 4, x, y, z);
    *n_draws = s.n_draws;
    *n_mesh = s.n_mesh;
    if (total_records) *total_records = s.total;
    return (s.n_draws > draw_cap || s.n_mesh > mesh_cap) ? VP_ERR_CAPACITY : VP_OK;
}
