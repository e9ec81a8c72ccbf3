/*
 * vp_gpu.h -- C ABI of the MI355X (gfx950) voxel pipeline.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, returns an
 * int (VP_OK = 0, negative = error, never aborts) and is asynchronous on the
 * caller's HIP stream (passed as void*; NULL = the null stream).  Device
 * buffers belong to the caller; a vp_ctx owns only its small scratch arenas.
 * Output order is deterministic: it never depends on scheduling.
 *
 * Streams: the calls made on one vp_ctx share its scratch (the work-queue
 * ticket and look-back words, the splat frame and ray buffers, the block
 * lists, the status word), so a context must be used from one stream at a
 * time: calls queued on two streams at once may overwrite each other's
 * scratch.  Use one context per stream (as the reference uses one set of
 * buffers per meshing thread, src/chunkset.c:258-273).  A remesh graph
 * (vp_host_remesh_graph_create) owns its own ticket and look-back words and
 * may replay beside calls on the context; it shares only the status word,
 * whose bits are sticky flags.
 *
 * Reference interfaces replaced (paths are relative to kosshi-net/voxplat):
 *   vp_gpu_cull_mesh  <- chunkset_manage per-chunk body   src/chunkset.c:318-505
 *                        = chunk_make_mask                src/chunkset/mesher.c:377-456
 *                        + chunk_make_splatlist (L0..L4)  src/chunkset/mesher.c:497-536
 *                        + chunk_mask_downsample (L1..L4) src/chunkset/mesher.c:460-493
 *                        + shadow_sample                  src/chunkset/shadow.h:45-63
 *   vp_gpu_mesh       <- chunk_make_mesh                  src/chunkset/mesher.c:184-357
 *                        + sample_ao / sample_ao_border   src/chunkset/mesher.c:72-171
 *                        + shadow_sample(_normal)         src/chunkset/shadow.h:52-75
 *   vp_gpu_make_mask  <- chunk_make_mask                  src/chunkset/mesher.c:377-456
 *   vp_gpu_lod_mip    <- chunk_mask_downsample            src/chunkset/mesher.c:460-493
 *   vp_gpu_rle        <- rle_compress / rle_decompress    src/chunkset/rle.c:44-116
 *                        (callers chunk_compress / chunk_open_*, src/chunkset.c:133-230)
 *   vp_gpu_splat      <- gfx_svl_draw + splat.vert/.frag  src/gfx/vsplat.c:104-193,
 *                        resources/splat.vert.glsl:34-54, resources/splat.frag.glsl:60-156
 *   vp_gpu_splat_backward -- no counterpart in the reference (colour fitting).
 *   vp_gpu_svl_nodes  <- gfx_update_svl (node uploads)   src/gfx/vsplat.c:197-338
 *   vp_gpu_draw_gather <- gfx_svl_draw's draw queue       src/gfx/vsplat.c:163-183
 *                        (queue built by octree_cull,     src/game.c:319-395, in
 *                        vp_host_octree_cull, include/vp_host.h)
 *   vp_gpu_splat_bounds / vp_gpu_splat_blocks -- per-block culling and
 *                        near-to-far order (the reference culls octree nodes,
 *                        src/game.c:319-395)
 */
#ifndef VP_GPU_H
#define VP_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define VP_OK               0
#define VP_ERR_ARG         -1  /* bad argument (NULL pointer, size out of range)   */
#define VP_ERR_CAPACITY    -2  /* an output buffer is too small (sync queries)     */
#define VP_ERR_HIP         -3  /* a HIP runtime call failed                        */
#define VP_ERR_UNSUPPORTED -4  /* valid in the reference but not on this path      */
#define VP_ERR_DATA        -5  /* malformed input stream (RLE length != chunk size)*/
#define VP_ERR_NOMEM       -6  /* scratch allocation failed                        */

/* device status bits, read with vp_ctx_status() */
#define VP_STATUS_OVERFLOW  1u  /* a splat / word list did not fit its capacity */
#define VP_STATUS_BAD_RLE   2u  /* an RLE stream did not decode to R^3 voxels    */
#define VP_STATUS_BAD_ID    4u  /* a listed chunk id / stream range was invalid  */

typedef struct vp_ctx vp_ctx;

/* World geometry, as ChunkSet::root_bitw / max_bitw (src/chunkset.h:36-40).
 * Chunk edge R = 1 << root_bitw; chunks per axis = 1 << max_bitw[a].
 * Voxel index inside a chunk: ((z << b | y) << b) | x   (src/chunkset.c:128-130)
 * Chunk index:               ((cz << by | cy) << bx) | cx (src/chunkset.c:124-126)
 * World voxels are chunk-major: chunk c occupies bytes [c*R^3, (c+1)*R^3). */
typedef struct {
    uint8_t root_bitw;
    uint8_t max_bitw[3];
} vp_world_desc;

/* Shadow map (src/chunkset/shadow.h:26-63): u16 occluder heights indexed by
 * (wx + wy) + size_x * wz, size_x = X + Y, size_z = Z.  The buffer must hold
 * vp_shadow_padded_len() entries; entries past size_x*size_z are zero so the
 * LOD samples that land beyond the world read "not shadowed". */
typedef struct {
    const uint16_t *d_map;
    uint32_t size_x;
    uint32_t size_z;
} vp_shadow_desc;

static inline uint64_t vp_shadow_padded_len(uint32_t size_x, uint32_t size_z)
{
    return (uint64_t)size_x * (uint64_t)(size_z + 17u) + 64u;
}

/* ---- context ------------------------------------------------------------ */
int         vp_ctx_create(int device, vp_ctx **out);
int         vp_ctx_destroy(vp_ctx *ctx);
const char *vp_strerror(int code);
/* Synchronises `stream`, returns and clears the device status bits. */
int         vp_ctx_status(vp_ctx *ctx, void *stream, uint32_t *status);
/* Number of compute units of the context's device. */
int         vp_ctx_num_cus(vp_ctx *ctx);

/* ---- cull + splat list (+ LOD) ------------------------------------------ */
#define VP_CM_LOD        1u  /* also emit LOD levels 1..4 (chunkset.c:390-458)  */
#define VP_CM_RECORD_V2  2u  /* v2 records instead of the reference bytes       */
#define VP_CM_COUNT_ONLY 4u  /* write counts/offsets only, no records           */

/* Splat records are 8 bytes.
 * compat (reference bytes, mesher.c:521-531): int16 {wx, wy, wz, attr},
 *         attr = colour | shadow << 6.
 * v2:     uint16 {wx, wy, wz}; uint8 colour; uint8 flags
 *         (bit a*2 = +a face exposed, bit a*2+1 = -a face exposed, bit 6 = shadow).
 *         LOD records carry the face bits of the level-0 cell their colour
 *         comes from.  compat == {wx, wy, wz, colour | (flags & 0x40)}. */
typedef struct {
    void     *d_splats;           /* capacity_records * 8 bytes                     */
    uint64_t  capacity_records;
    uint64_t *d_chunk_offsets;    /* n+1 exclusive prefix, in int16 items (4/splat) */
    uint32_t *d_level_counts;     /* n*5 per-level counts, int16 items (svl_items)  */
} vp_splat_out;

/* Remesh the chunks d_chunk_ids[0..n) of a world resident in HBM.
 * Output: the concatenation, in list order, of each chunk's L0||L1||L2||L3||L4
 * records, each level in (z,y,x) scan order -- exactly the bytes
 * chunkset_manage stores in ChunkMD::svl, with ChunkMD::svl_items in
 * d_level_counts.  Chunks whose list does not fit capacity_records are not
 * written (their counts are) and VP_STATUS_OVERFLOW is raised.
 * d_uniform (nullable): per chunk, -1 = mixed, else the value every voxel has;
 *   uniform chunks are not read (only their halo planes are synthesised).
 * d_n (nullable): device-resident count <= n (for replayed graphs).
 * n == 1 is the plain per-chunk path; n > 1 runs one persistent work-queue
 * kernel whose output offsets come from a decoupled look-back.
 * Supported roots on the GPU: 4..64 (root_bitw 2..6). */
int vp_gpu_cull_mesh(vp_ctx *ctx, const vp_world_desc *world,
                     const uint8_t *d_voxels, const int16_t *d_uniform,
                     const vp_shadow_desc *shadow,
                     const uint32_t *d_chunk_ids, uint32_t n, const uint32_t *d_n,
                     uint32_t flags, const vp_splat_out *out, void *stream);

/* vp_gpu_cull_mesh that allocates nothing: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under graph capture,
 * tight serving loops).  Same arguments, output and errors as
 * vp_gpu_cull_mesh, but the work-queue ticket and the n look-back words live
 * in a static device array of each GPU (4 MiB + 256 B, held until the process
 * ends) instead of the context's scratch, so n is limited to
 * VP_NOALLOC_MAX_CHUNKS (every chunk of a 16384x512x16384 world of 64^3
 * chunks); a larger n returns VP_ERR_CAPACITY and launches nothing.  All calls
 * on one GPU share the array: each call clears what it uses before its
 * kernel runs, and calls that overlap on one GPU (two streams, or two
 * contexts of the same device) race on it. */
#define VP_NOALLOC_MAX_CHUNKS 524288u
int vp_gpu_cull_mesh_noalloc(vp_ctx *ctx, const vp_world_desc *world,
                             const uint8_t *d_voxels, const int16_t *d_uniform,
                             const vp_shadow_desc *shadow,
                             const uint32_t *d_chunk_ids, uint32_t n, const uint32_t *d_n,
                             uint32_t flags, const vp_splat_out *out, void *stream);

/* ---- near-field quad mesh ---------------------------------------------------
 * Replaces chunk_make_mesh (src/chunkset/mesher.c:184-357), one call per
 * chunk in the reference, with one launch over a chunk list.
 * Per chunk, faces in the reference's order (voxels z, y, x; axis 0, 1, 2):
 * one face wherever a voxel A of the chunk and its +axis neighbour B differ
 * in occupancy (B from the next chunk at the edge, air beyond the world).
 * Face f of the concatenation: d_geometry[16 f ..] = 4 vertices
 * {x, y, z, attribute} as int16, attribute bits __SUVNNN AORRGGBB
 * (mesher.c:332-342); d_index[6 f ..] = two triangles, counted from the
 * chunk's first vertex as the reference's per-chunk index buffer.
 * The reference's geometry_items / index_items of list entry k are
 * 16 * faces_k and 6 * faces_k, faces_k = d_face_offsets[k+1] - d_face_offsets[k].
 * Shadow samples read the map as vp_shadow_desc describes (zero padding
 * past the world: a z face on the world's far z plane samples a lit row). */
typedef struct {
    int16_t  *d_geometry;         /* capacity_faces * 16 int16                     */
    uint32_t *d_index;            /* capacity_faces * 6 u32                        */
    uint64_t  capacity_faces;
    uint64_t *d_face_offsets;     /* n+1 exclusive prefix, in faces                */
} vp_mesh_out;

/* Mesh the chunks d_chunk_ids[0..n), concatenated in list order.
 * d_geometry = d_index = NULL with capacity_faces = 0 counts only: the
 * offsets are written, nothing else.  Chunks whose faces do not fit are not
 * written (their offsets are) and VP_STATUS_OVERFLOW is raised; ids outside
 * the world get no faces and raise VP_STATUS_BAD_ID.
 * d_uniform (nullable): per chunk, -1 = mixed, else every voxel's value;
 * uniform chunks and uniform neighbours are not read.
 * Supported roots: 4..64 (root_bitw 2..6). */
int vp_gpu_mesh(vp_ctx *ctx, const vp_world_desc *world, const uint8_t *d_voxels,
                const int16_t *d_uniform, const vp_shadow_desc *shadow,
                const uint32_t *d_chunk_ids, uint32_t n, const vp_mesh_out *out,
                void *stream);

/* Level-0 visibility masks ((R+1)^3 bytes per chunk, index (z*(R+1)+y)*(R+1)+x)
 * exactly as chunk_make_mask leaves them in a zeroed buffer. */
int vp_gpu_make_mask(vp_ctx *ctx, const vp_world_desc *world,
                     const uint8_t *d_voxels, const int16_t *d_uniform,
                     const uint32_t *d_chunk_ids, uint32_t n,
                     uint8_t *d_masks, void *stream);

/* One LOD step on n masks of stride (R+1)^3: d_dst = downsample(level, d_src)
 * into a zeroed target (chunk_mask_downsample, mesher.c:460-493).  Each dst
 * mask is fully overwritten. level in 1..4. */
int vp_gpu_lod_mip(vp_ctx *ctx, uint32_t root_bitw, int level,
                   const uint8_t *d_src, uint8_t *d_dst, uint32_t n, void *stream);

/* ---- RLE codec ------------------------------------------------------------
 * Word = (value << 24) | run (1..0xFFFFFF), runs maximal in voxel order,
 * one 0 word terminates each chunk's stream (rle.c:44-87).
 * ENCODE: chunks d_chunk_ids[0..n) of d_voxels -> d_words, concatenated in
 *   list order with d_word_offsets[n+1] (in words, terminators included).
 *   d_uniform_out (nullable) receives -1 / the uniform value per listed chunk.
 *   d_words = NULL with capacity_words = 0 sizes the streams only: offsets
 *   (and d_uniform_out) are written, no words, no overflow is raised.
 * DECODE: d_words/d_word_offsets -> chunks d_chunk_ids[k] of d_voxels.
 *   A stream that does not expand to exactly R^3 voxels raises
 *   VP_STATUS_BAD_RLE (the chunk is written up to R^3 bytes); offsets outside
 *   [0, capacity_words] raise VP_STATUS_BAD_ID and the chunk is skipped.
 * Chunk ids outside the world raise VP_STATUS_BAD_ID and are skipped. */
#define VP_RLE_ENCODE 0
#define VP_RLE_DECODE 1
int vp_gpu_rle(vp_ctx *ctx, int dir, const vp_world_desc *world,
               uint8_t *d_voxels, const uint32_t *d_chunk_ids, uint32_t n,
               uint32_t *d_words, uint64_t capacity_words,
               uint64_t *d_word_offsets, int16_t *d_uniform_out, void *stream);

/* ---- splat resolve --------------------------------------------------------
 * Camera matrices are column-major 4x4 (m[col*4+row], cglm layout), built on
 * the host as in src/game.c:458-491. */
typedef struct {
    float    view_translation[16];   /* u_view_translation                   */
    float    proj_view_rot[16];      /* u_projection * u_view_rotation       */
    float    inv_proj_view_rot[16];  /* u_inv_proj_view_rot                  */
    uint32_t width, height;          /* uViewport                            */
    float    far_plane;              /* u_far                                */
} vp_camera;

typedef struct {
    float    *d_depth;   /* W*H, t0/far of the nearest hit, 1.0 = none          */
    uint32_t *d_id;      /* W*H, splat index of the nearest hit, ~0u = none     */
    float    *d_rgb;     /* nullable W*H*3: shaded + fogged colour              */
    float    *d_weight;  /* nullable W*H: d rgb / d voxel_rgb (shade*(1-fog))    */
} vp_frame;

/* d_splats: n compat records; lod per splat = 1 << d_lod[i] (nullable: lod_log2
 * for all). d_voxel_rgb (nullable): n*3 fp32 colours that replace the 6-bit
 * palette (the fitting path). Pixel (px,py) has gl_FragCoord (px+.5, py+.5),
 * row 0 at the bottom, as GL. Ties go to the lower splat index. */
int vp_gpu_splat(vp_ctx *ctx, const int16_t *d_splats, const uint8_t *d_lod,
                 uint32_t lod_log2, uint64_t n, const vp_camera *cam,
                 const float *d_voxel_rgb, const vp_frame *out, void *stream);

/* Block bounds of a splat list, for vp_gpu_splat_blocks.  Block b holds the
 * splats [VP_SPLAT_BLOCK*b, VP_SPLAT_BLOCK*(b+1)); d_bounds receives
 * ceil(n / VP_SPLAT_BLOCK) entries of 8 floats: min x, y, z of the splat
 * centres (position + lod/2), the largest lod/2, max x, y, z of the centres, 0.
 * Computed once per list (one read of it).  No counterpart in the reference,
 * which culls per octree node on the CPU (src/game.c:319-395); this is that
 * culling at block granularity on the GPU. */
#define VP_SPLAT_BLOCK 256
int vp_gpu_splat_bounds(vp_ctx *ctx, const int16_t *d_splats, const uint8_t *d_lod,
                        uint32_t lod_log2, uint64_t n, float *d_bounds, void *stream);

/* vp_gpu_splat over a list with block bounds: blocks whose splat centres all
 * lie outside the view volume (with a margin far above rounding) are skipped,
 * and the others are rasterised nearest first, so the early depth rejection
 * meets occluders early.  The result equals vp_gpu_splat's bit for bit. */
int vp_gpu_splat_blocks(vp_ctx *ctx, const int16_t *d_splats, const uint8_t *d_lod,
                        uint32_t lod_log2, uint64_t n, const float *d_bounds,
                        const vp_camera *cam, const float *d_voxel_rgb,
                        const vp_frame *out, void *stream);

/* How vp_gpu_splat_blocks and vp_gpu_splat_counted resolve a frame of at
 * most 8192 screen tiles of 32x32 pixels (up to 3840x2160 and a little more).
 * With enable = 1 the visible splats of at most 8x8 pixels are binned into
 * the tiles they cover and each tile is resolved in on-chip memory, after the
 * larger ones are drawn against the frame in device memory; frames with more
 * than max_entries (tile, splat) entries (0: 2^26, about 256 MiB of context
 * memory) fall back to rasterising every splat against the frame in device
 * memory, as do all frames with enable = 0 (the default).  Both give the same
 * frame bit for bit; this sets memory and speed only.  No counterpart in the
 * reference (the GL rasteriser bins fragments in hardware). */
int vp_ctx_set_splat_tiles(vp_ctx *ctx, int enable, uint64_t max_entries);

/* dL/d voxel_rgb += sum over pixels p with id(p)=i of weight(p) * dL/drgb(p).
 * d_id/d_weight are the forward's outputs; d_grad (n*3) is accumulated. */
int vp_gpu_splat_backward(vp_ctx *ctx, uint64_t n, const vp_camera *cam,
                          const uint32_t *d_id, const float *d_weight,
                          const float *d_dL_drgb, float *d_grad, void *stream);

/* ---- draw lists -----------------------------------------------------------
 * The reference draws splats per octree node: at LOD level l a node covers
 * the chunks whose coordinates shifted right by l equal the node's, and its
 * vertex buffer is the concatenation, over those chunks in the order x
 * outer, y, z inner, of each chunk's level-l records (gfx_update_svl,
 * src/gfx/vsplat.c:197-338).  Nodes per axis at level l:
 * 1 << (max_bitw[a] - min(l, max_bitw[a])); node index = flatten3 of the
 * node coordinates with those bits (vsplat.c:213-232). */
static inline uint32_t vp_node_bits(const vp_world_desc *w, uint32_t level, uint32_t axis)
{
    const uint32_t b = w->max_bitw[axis];
    return b - (level < b ? level : b);
}
static inline uint64_t vp_node_count(const vp_world_desc *w, uint32_t level)
{
    return (uint64_t)1 << (vp_node_bits(w, level, 0) + vp_node_bits(w, level, 1) +
                           vp_node_bits(w, level, 2));
}

typedef struct {
    void     *d_node_splats;      /* capacity_records * 8 bytes                      */
    uint64_t  capacity_records;
    uint64_t *d_node_offsets;     /* vp_node_count + 1 exclusive prefix, in records  */
} vp_node_out;

/* Node lists of one level from a remesh output: d_splats, d_chunk_offsets
 * (n+1, int16 items) and d_level_counts (n*5, int16 items) as
 * vp_gpu_cull_mesh wrote them for the list d_chunk_ids[0..n).  Chunks not in
 * the list contribute nothing; a chunk listed twice uses its last entry.
 * d_node_splats = NULL with capacity 0 writes the offsets only.  Nodes that
 * do not fit are not written (their offsets are) and VP_STATUS_OVERFLOW is
 * raised.  Replaces the per-node uploads of gfx_update_svl. */
int vp_gpu_svl_nodes(vp_ctx *ctx, const vp_world_desc *world, uint32_t level,
                     const void *d_splats, const uint32_t *d_chunk_ids, uint32_t n,
                     const uint64_t *d_chunk_offsets, const uint32_t *d_level_counts,
                     const vp_node_out *out, void *stream);

/* All five levels in one launch (out[l] for level l), high levels first, so
 * that their few large nodes run beside the many small ones of levels 0-1. */
int vp_gpu_svl_nodes_all(vp_ctx *ctx, const vp_world_desc *world, const void *d_splats,
                         const uint32_t *d_chunk_ids, uint32_t n,
                         const uint64_t *d_chunk_offsets, const uint32_t *d_level_counts,
                         const vp_node_out out[5], void *stream);

/* One draw of a frame: node `node` of level `level`, whose records go to
 * the assembled list from record `first` on (vp_host_octree_cull fills it). */
typedef struct {
    uint32_t level;
    uint32_t node;
    uint64_t first;
} vp_draw;

/* Assemble a frame's splat list from its draws (device array d_draws[0..n)):
 * each draw's node records are copied to d_out + 8 * first and its level to
 * d_out_lod (log2 splat size per splat, the per-draw u_lod of
 * vsplat.c:163-183).  d_node_splats / d_node_offsets: the five levels' node
 * lists (host arrays of device pointers).  A draw naming a level or node
 * outside the world raises VP_STATUS_BAD_ID, one reaching past
 * capacity_records VP_STATUS_OVERFLOW; either is skipped.  The list is then one
 * vp_gpu_splat(d_lod = d_out_lod) call: splat ids follow the draw order, so
 * equal depths resolve to the earlier draw, as GL_LESS does for the
 * reference's draw queue. */
int vp_gpu_draw_gather(vp_ctx *ctx, const vp_world_desc *world,
                       const void *const d_node_splats[5],
                       const uint64_t *const d_node_offsets[5], const vp_draw *d_draws,
                       uint32_t n, void *d_out, uint8_t *d_out_lod,
                       uint64_t capacity_records, void *stream);

/* The frame's draw queue selected on the device: octree_cull
 * (src/game.c:319-395) from every level-4 node (game.c:507-512), with
 * gfx_fcull_visible's bounding-sphere test (src/gfx/fcull.c:38-45), exactly as
 * vp_host_octree_cull (include/vp_host.h) selects it, draw for draw: queue
 * order, levels, nodes and running `first` record offsets.  d_node_offsets[l]:
 * device node offsets of level l (vp_gpu_svl_nodes); d_chunk_mesh (nullable,
 * one byte per chunk): level-0 chunks with a quad mesh go to d_mesh_chunks.
 * d_counts[0..2] receives (draws, mesh draws, records) in device memory; the
 * host reads nothing, so gather and splat can follow on the stream with
 * vp_gpu_draw_gather_counted and vp_gpu_splat_counted.  Draws or mesh chunks
 * past their capacity are not written (their counts are) and
 * VP_STATUS_OVERFLOW is raised.  Needs sum(max_bitw) <= 27. */
int vp_gpu_octree_cull(vp_ctx *ctx, const vp_world_desc *world, const float planes[24],
                       const float origin[3], const uint64_t *const d_node_offsets[5],
                       const uint8_t *d_chunk_mesh, vp_draw *d_draws, uint32_t draw_cap,
                       uint32_t *d_mesh_chunks, uint32_t mesh_cap, uint64_t *d_counts,
                       void *stream);

/* vp_gpu_draw_gather with the draw count read from device memory:
 * min(d_counts[0], draw_cap) draws of d_draws (d_counts as vp_gpu_octree_cull
 * writes it). */
int vp_gpu_draw_gather_counted(vp_ctx *ctx, const vp_world_desc *world,
                               const void *const d_node_splats[5],
                               const uint64_t *const d_node_offsets[5], const vp_draw *d_draws,
                               uint32_t draw_cap, const uint64_t *d_counts, void *d_out,
                               uint8_t *d_out_lod, uint64_t capacity_records, void *stream);

/* vp_gpu_splat_bounds + vp_gpu_splat_blocks for a list whose length is in
 * device memory: the first min(*d_count, capacity) splats.  d_bounds holds
 * ceil(capacity / VP_SPLAT_BLOCK) * 8 floats of workspace; the bounds are
 * computed into it, then the frame is resolved.  The result equals
 * vp_gpu_splat's on that many splats. */
int vp_gpu_splat_counted(vp_ctx *ctx, const int16_t *d_splats, const uint8_t *d_lod,
                         uint32_t lod_log2, uint64_t capacity, const uint64_t *d_count,
                         float *d_bounds, const vp_camera *cam, const float *d_voxel_rgb,
                         const vp_frame *out, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* VP_GPU_H */
