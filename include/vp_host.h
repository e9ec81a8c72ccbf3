/*
 * vp_host.h -- host-side drivers above the C ABI, in C.
 *
 * These are the parts of chunkset_manage (src/chunkset.c:246-507) that stay
 * on the host when the per-chunk meshing moves to the GPU: walking the dirty
 * list, one launch per chunk (the plain path), or one captured graph per frame
 * (the batched path), and choosing the launch shape.  No counterpart in the
 * reference's own interface beyond that loop.
 */
#ifndef VP_HOST_H
#define VP_HOST_H

#include "vp_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Pre-size the context's scratch so later calls allocate nothing (required
 * before capturing a call in a graph): look-back state for lists of up to
 * max_chunks entries, frame scratch for max_pixels pixels. */
int vp_ctx_reserve(vp_ctx *ctx, uint64_t max_chunks, uint64_t max_pixels);

/* Launch shape of the persistent remesh kernel: threads per workgroup (256,
 * 512 or 1024) and resident workgroups per CU (1..4). */
int vp_ctx_set_launch(vp_ctx *ctx, int block, int wg_per_cu);
int vp_ctx_get_launch(vp_ctx *ctx, int *block, int *wg_per_cu);

/* The plain per-chunk path: one vp_gpu_cull_mesh(n = 1) launch per listed
 * chunk, as chunkset_manage meshes one chunk at a time.  Chunk k's records go
 * to d_splats + 8 * rec_offsets[k] (host array, records) with room for
 * rec_caps[k] records; its local offsets to d_offsets2[2k..2k+1] and counts
 * to d_level_counts[5k..5k+4]. */
int vp_host_remesh_per_chunk(vp_ctx *ctx, const vp_world_desc *world,
                             const uint8_t *d_voxels, const int16_t *d_uniform,
                             const vp_shadow_desc *shadow, const uint32_t *d_chunk_ids,
                             uint32_t n, uint32_t flags, void *d_splats,
                             const uint64_t *rec_offsets, const uint64_t *rec_caps,
                             uint64_t *d_offsets2, uint32_t *d_level_counts, void *stream);

/* The batched path captured once in a hipGraph and replayed per frame.  The
 * arguments are those of vp_gpu_cull_mesh; with d_n set, each replay remeshes
 * the first *d_n entries of d_chunk_ids (the dirty count lives on the GPU). */
typedef struct vp_remesh_graph vp_remesh_graph;
int vp_host_remesh_graph_create(vp_ctx *ctx, const vp_world_desc *world,
                                const uint8_t *d_voxels, const int16_t *d_uniform,
                                const vp_shadow_desc *shadow, const uint32_t *d_chunk_ids,
                                uint32_t n, const uint32_t *d_n, uint32_t flags,
                                const vp_splat_out *out, vp_remesh_graph **graph);
int vp_host_remesh_graph_launch(vp_remesh_graph *graph, void *stream);
int vp_host_remesh_graph_destroy(vp_remesh_graph *graph);

/* Autotuner: time every launch shape on the given workload (reps runs each,
 * median), keep the fastest in the context.  Returns it and its time. */
int vp_host_autotune_remesh(vp_ctx *ctx, const vp_world_desc *world,
                            const uint8_t *d_voxels, const int16_t *d_uniform,
                            const vp_shadow_desc *shadow, const uint32_t *d_chunk_ids,
                            uint32_t n, uint32_t flags, const vp_splat_out *out, int reps,
                            void *stream, int *best_block, int *best_wg_per_cu,
                            float *best_ms);

#ifdef __cplusplus
}
#endif
#endif
