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
 * 512 or 1024), workgroups per CU the grid is sized for (1..4), and the
 * records each lane prepares before it stores them (2 or 4; chunks of 64^3
 * only, smaller chunks always use 4).  No counterpart in the reference, which
 * meshes on the CPU.  A world without a shape of its own (set for its
 * root_bitw and max_bitw) uses the default (world = NULL): 512, 2, 4. */
typedef struct {
    int block;
    int wg_per_cu;
    int records_in_flight;
} vp_remesh_shape;
int vp_ctx_set_remesh_shape(vp_ctx *ctx, const vp_world_desc *world,
                            const vp_remesh_shape *shape);
int vp_ctx_get_remesh_shape(vp_ctx *ctx, const vp_world_desc *world, vp_remesh_shape *shape);
/* Workgroups of the shape's kernel (for the world's chunk size and the
 * record flags) that fit on one CU at once, from its registers and LDS;
 * wg_per_cu above it only queues workgroups. */
int vp_ctx_remesh_resident(vp_ctx *ctx, const vp_world_desc *world, uint32_t flags,
                           const vp_remesh_shape *shape, int *per_cu);
/* The default shape's block and wg_per_cu. */
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
 * the first *d_n entries of d_chunk_ids (the dirty count lives on the GPU).
 * Ownership: the graph allocates its own work-queue ticket and n look-back
 * words at creation and frees them in _destroy, so later calls on the
 * context (which may grow and replace the context's scratch) cannot leave it
 * pointing at freed memory.  The buffers the caller passes (voxels, ids,
 * d_n, outputs) must outlive the graph.  Replays of one graph run one at a
 * time (they share its scratch). */
typedef struct vp_remesh_graph vp_remesh_graph;
int vp_host_remesh_graph_create(vp_ctx *ctx, const vp_world_desc *world,
                                const uint8_t *d_voxels, const int16_t *d_uniform,
                                const vp_shadow_desc *shadow, const uint32_t *d_chunk_ids,
                                uint32_t n, const uint32_t *d_n, uint32_t flags,
                                const vp_splat_out *out, vp_remesh_graph **graph);
int vp_host_remesh_graph_launch(vp_remesh_graph *graph, void *stream);
int vp_host_remesh_graph_destroy(vp_remesh_graph *graph);

/* One frame of chunk management over a resident world, captured once and
 * replayed every frame: chunkset_manage's walk over the chunks
 * (src/chunkset.c:246-313), which remeshes each dirty chunk (clearing its
 * dirty flag) and compresses every other one (chunk_compress,
 * chunkset.c:213-230,300-306).  Each replay
 *   1. turns the dirty bytes d_dirty[chunks] (set by vp_gpu_edit_spheres) into
 *      d_dirty_ids (the dirty chunks, ascending) and d_idle_ids (every other
 *      chunk: those d_uniform marks mixed, ascending, then those it marks
 *      uniform, ascending), with their counts in d_counts[0] and d_counts[1],
 *      and clears the dirty bytes;
 *   2. remeshes the dirty list as vp_gpu_cull_mesh does with d_n = d_counts:
 *      records to remesh.d_splats, offsets and level counts per entry;
 *   3. RLE-encodes the idle list as vp_gpu_rle does with n = d_counts[1]:
 *      the look-back that places each stream also sizes it, so one pass
 *      writes d_word_offsets (entries + 1) and the words that fit
 *      capacity_words (VP_STATUS_OVERFLOW otherwise), and d_idle_uniform.
 *      A chunk that d_uniform marks uniform (the value every voxel holds;
 *      the table the remesh takes, which edits mark mixed) gets its one-run
 *      stream without a read of its voxels, as the reference keeps all-air
 *      chunks on the shared null-chunk stream.
 * Arrays indexed by list entry must hold `chunks` entries (chunks + 1 for
 * offsets, 5 * chunks for level counts).  The graph owns its scratch (list
 * counts, tickets and look-back words of both persistent kernels) and frees
 * it on destroy; replays of one graph run one at a time. */
typedef struct {
    uint32_t    *d_dirty_ids;
    uint32_t    *d_idle_ids;
    uint32_t    *d_counts;          /* [0] dirty entries, [1] idle entries      */
    vp_splat_out remesh;
    uint32_t    *d_words;
    uint64_t     capacity_words;
    uint64_t    *d_word_offsets;
    int16_t     *d_idle_uniform;    /* nullable */
} vp_frame_out;
typedef struct vp_frame_graph vp_frame_graph;
int vp_host_frame_graph_create(vp_ctx *ctx, const vp_world_desc *world, uint8_t *d_voxels,
                               const int16_t *d_uniform, const vp_shadow_desc *shadow,
                               uint8_t *d_dirty, uint32_t flags, const vp_frame_out *out,
                               vp_frame_graph **graph);
int vp_host_frame_graph_launch(vp_frame_graph *graph, void *stream);
int vp_host_frame_graph_destroy(vp_frame_graph *graph);

/* Autotuner: time every launch shape whose workgroups are all resident
 * (vp_ctx_remesh_resident) on the given workload -- block 256, 512, 1024,
 * wg_per_cu 1 up to the resident count, 2 and 4 records in flight for 64^3
 * chunks -- reps runs each, median; keep the fastest as the world's shape in
 * the context.  Returns it, its time, and (nullable) the number of shapes
 * timed. */
int vp_host_autotune_remesh(vp_ctx *ctx, const vp_world_desc *world,
                            const uint8_t *d_voxels, const int16_t *d_uniform,
                            const vp_shadow_desc *shadow, const uint32_t *d_chunk_ids,
                            uint32_t n, uint32_t flags, const vp_splat_out *out, int reps,
                            void *stream, vp_remesh_shape *best, float *best_ms,
                            int *shapes_timed);

/* ---- on-disk RLE world -------------------------------------------------------
 * The file of command_export (src/deadcode.c:320-346), read back as
 * chunkset_clear_import does (src/deadcode.c:159-229).  Little endian:
 *   0x89 "VOXPLAT", root_bitw, max_bitw[3]          (12 bytes)
 *   every chunk's RLE stream in chunk-index order, u32 words
 *   (value << 24 | run) with the 0 terminator      (rle.c:44-87)
 *   optionally the shadow map, size_x * size_z u16  (shadow.h:26-42)
 * Unlike the dead export, each stream is written whole (its two-byte length
 * scan stops inside the first 32-bit word) and the map entirely (it writes
 * shadow_map_length bytes of a u16 map).  The streams are made and expanded
 * by vp_gpu_rle in batches of chunks; the calls are synchronous on `stream`
 * and read and clear the context's status word.  Errors: VP_ERR_ARG (bad
 * argument, file not openable or not writable), VP_ERR_DATA (bad header, a
 * world other than `world`, a truncated or malformed stream, a trailer that
 * is neither empty nor a whole map, a map asked for that the file lacks). */
int vp_host_world_save(vp_ctx *ctx, const char *path, const vp_world_desc *world,
                       const uint8_t *d_voxels, const uint16_t *d_shadow_map, void *stream,
                       uint64_t *bytes_written);
/* The world geometry in a file's header. */
int vp_host_world_read_header(const char *path, vp_world_desc *world);
/* Decode every chunk into d_voxels (chunk-major, as vp_gpu_cull_mesh reads
 * it); d_uniform (nullable) receives -1 or the value of each chunk whose
 * stream is one full-chunk run; d_shadow_map (nullable) receives the map and
 * zero padding up to vp_shadow_padded_len entries. */
int vp_host_world_load(vp_ctx *ctx, const char *path, const vp_world_desc *world,
                       uint8_t *d_voxels, int16_t *d_uniform, uint16_t *d_shadow_map,
                       void *stream);

/* ---- frame draw queue ------------------------------------------------------
 * Frustum planes of a view-projection matrix (column-major), normalised, in
 * the order -x, +x, -y, +y, -z, +z of the clip-space bounds
 * (gfx_fcull_extract_planes, src/gfx/fcull.c:18-36). */
void vp_host_frustum_planes(const float view_projection[16], float planes[24]);

/* The reference's per-frame LOD selection (octree_cull, src/game.c:319-395,
 * driven from src/game.c:507-515): from every level-4 node, a node whose
 * bounding sphere (radius 0.85 of its edge) passes the frustum test is drawn
 * at its level if its centre is farther than level * 1024 from `origin`,
 * else its eight children are visited (x outer, y, z inner).  At level 0 a
 * chunk with chunk_mesh[chunk] != 0 (nullable) is queued as a mesh draw
 * instead.  A node with no records is skipped, as a node without a vertex
 * buffer is.  node_offsets[l]: host copy of level l's node offsets (from
 * vp_gpu_svl_nodes).  Draws are written in queue order with `first` set to
 * the running record total; returns VP_ERR_CAPACITY if draw_cap or mesh_cap
 * is too small (the counts are still the full ones). */
int vp_host_octree_cull(const vp_world_desc *world, const float planes[24], const float origin[3],
                        const uint64_t *const node_offsets[5], const uint8_t *chunk_mesh,
                        vp_draw *draws, uint32_t draw_cap, uint32_t *n_draws,
                        uint32_t *mesh_chunks, uint32_t mesh_cap, uint32_t *n_mesh,
                        uint64_t *total_records);

#ifdef __cplusplus
}
#endif
#endif
