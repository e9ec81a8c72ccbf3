/*
 * vp_host.c -- host-side C of the pipeline: the order-dependent shadow-map
 * updates of sphere edits, and the remesh drivers (plain per-chunk loop,
 * captured batched frame, launch-shape autotuner) declared in vp_host.h.
 *
 * shadow_place_update (src/chunkset/shadow.h:77-89) reads map[i] and map[i+1]
 * before writing map[i], so the result depends on the order of the placed
 * voxels; chunkset_edit_sphere (src/chunkset/edit.c:179-244) places every
 * voxel of an add-sphere that lies inside an in-world chunk of the sphere's
 * box, chunk by chunk (cx outer, then cy, cz) and inside a chunk in u order
 * (x outer, then y, z).  The updates are cheap and strictly sequential, so
 * they run here, on a host copy of the map.
 */
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
 * Remesh drivers: the plain per-chunk loop, the captured batched frame and
 * the launch-shape autotuner.
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

struct vp_remesh_graph {
    hipGraph_t graph;
    hipGraphExec_t exec;
};

int vp_host_remesh_graph_create(vp_ctx *ctx, const vp_world_desc *world,
                                const uint8_t *d_voxels, const int16_t *d_uniform,
                                const vp_shadow_desc *shadow, const uint32_t *d_chunk_ids,
                                uint32_t n, const uint32_t *d_n, uint32_t flags,
                                const vp_splat_out *out, vp_remesh_graph **graph)
{
    if (!ctx || !graph) return VP_ERR_ARG;
    *graph = NULL;
    int rc = vp_ctx_reserve(ctx, n, 0);
    if (rc) return rc;
    hipStream_t s;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return VP_ERR_HIP;
    vp_remesh_graph *g = (vp_remesh_graph *)calloc(1, sizeof(*g));
    if (!g) {
        (void)hipStreamDestroy(s);
        return VP_ERR_NOMEM;
    }
    rc = VP_ERR_HIP;
    if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) == hipSuccess) {
        int crc = vp_gpu_cull_mesh(ctx, world, d_voxels, d_uniform, shadow, d_chunk_ids, n, d_n,
                                   flags, out, s);
        hipError_t e = hipStreamEndCapture(s, &g->graph);
        if (crc != VP_OK)
            rc = crc;
        else if (e == hipSuccess &&
                 hipGraphInstantiate(&g->exec, g->graph, NULL, NULL, 0) == hipSuccess)
            rc = VP_OK;
    }
    (void)hipStreamDestroy(s);
    if (rc != VP_OK) {
        if (g->graph) (void)hipGraphDestroy(g->graph);
        free(g);
        return rc;
    }
    *graph = g;
    return VP_OK;
}

int vp_host_remesh_graph_launch(vp_remesh_graph *graph, void *stream)
{
    if (!graph) return VP_ERR_ARG;
    return hipGraphLaunch(graph->exec, (hipStream_t)stream) == hipSuccess ? VP_OK : VP_ERR_HIP;
}

int vp_host_remesh_graph_destroy(vp_remesh_graph *graph)
{
    if (!graph) return VP_ERR_ARG;
    (void)hipGraphExecDestroy(graph->exec);
    (void)hipGraphDestroy(graph->graph);
    free(graph);
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
                            void *stream, int *best_block, int *best_wg_per_cu,
                            float *best_ms)
{
    static const int blocks[3] = {256, 512, 1024};
    if (!ctx || !best_block || !best_wg_per_cu || reps < 1 || reps > 64) return VP_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return VP_ERR_HIP;
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return VP_ERR_HIP;
    }
    float best = 1e30f, t[64];
    int bb = 512, bw = 2, rc = VP_OK;
    for (int bi = 0; bi < 3 && rc == VP_OK; bi++)
        for (int wg = 1; wg <= 4 && rc == VP_OK; wg++) {
            if (blocks[bi] * wg > 2048) continue;          /* 32 waves per CU at most */
            rc = vp_ctx_set_launch(ctx, blocks[bi], wg);
            if (rc) break;
            rc = vp_gpu_cull_mesh(ctx, world, d_voxels, d_uniform, shadow, d_chunk_ids, n, NULL,
                                  flags, out, stream);   /* warm-up */
            for (int r = 0; r < reps && rc == VP_OK; r++) {
                if (hipEventRecord(e0, st) != hipSuccess) { rc = VP_ERR_HIP; break; }
                rc = vp_gpu_cull_mesh(ctx, world, d_voxels, d_uniform, shadow, d_chunk_ids, n,
                                      NULL, flags, out, stream);
                if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                    hipEventElapsedTime(&t[r], e0, e1) != hipSuccess)
                    rc = VP_ERR_HIP;
            }
            if (rc) break;
            qsort(t, (size_t)reps, sizeof(float), cmp_float);
            if (t[reps / 2] < best) {
                best = t[reps / 2];
                bb = blocks[bi];
                bw = wg;
            }
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    *best_block = bb;
    *best_wg_per_cu = bw;
    if (best_ms) *best_ms = best;
    return vp_ctx_set_launch(ctx, bb, bw);
}
