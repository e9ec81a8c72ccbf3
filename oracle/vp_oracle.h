/*
 * vp_oracle.h -- CPU restatement of voxplat's voxel pipeline.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the HIP path is
 * compared against (tests/, __graft_entry__.smoke) and the CPU leg timed
 * beside it (bench.py's cpu_baseline).  Nothing in voxplat_amd/ links, loads
 * or calls it.
 *
 * Each function cites the reference lines (kosshi-net/voxplat) whose
 * behaviour it restates.  It is written from the behaviour, not copied.
 * World layout and record formats are those of include/vp_gpu.h.
 */
#ifndef VP_ORACLE_H
#define VP_ORACLE_H

#include <stdint.h>
#include "../include/vp_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- seeded integer generator (the workload's own; reference gen.c uses
 * FastNoise, which is not vendored) ------------------------------------- */
typedef struct {
    uint64_t seed;
    uint32_t period;   /* base lattice period of octave 0 */
    int32_t  hmin;     /* height range, inclusive */
    int32_t  hmax;
} vpo_gen_params;

uint64_t vpo_splitmix64(uint64_t x);
int32_t  vpo_gen_height(const vpo_gen_params *p, uint32_t x, uint32_t z);
void     vpo_gen_chunk(const vp_world_desc *w, const vpo_gen_params *p,
                       uint32_t chunk, uint8_t *out);
void     vpo_gen_world(const vp_world_desc *w, const vpo_gen_params *p,
                       uint8_t *out, int threads);

/* ---- shadow map (shadow.h:26-89) --------------------------------------- */
void     vpo_shadow_dims(const vp_world_desc *w, uint32_t *size_x, uint32_t *size_z);
uint32_t vpo_shadow_index(uint32_t size_x, const uint32_t ws[3]);
int      vpo_shadow_sample(const uint16_t *map, uint32_t size_x, const uint32_t ws[3]);
void     vpo_shadow_place(uint16_t *map, uint32_t size_x, const uint32_t ws[3]);
/* Build from a world: zero the padded map, then place every solid voxel in
 * the serial order chunks-by-index, x, z, y.  `serial`=1 runs that order
 * literally; 0 runs the z-rows in parallel (same result: rows never mix). */
void     vpo_shadow_build(const vp_world_desc *w, const uint8_t *voxels,
                          uint16_t *map, int serial, int threads);

/* ---- mesher (mesher.c:365-536) ------------------------------------------ */
uint32_t vpo_chunk_index(const vp_world_desc *w, const uint32_t c[3]);
void     vpo_chunk_coords(const vp_world_desc *w, uint32_t chunk, uint32_t c[3]);
void     vpo_make_mask(const vp_world_desc *w, const uint8_t *voxels,
                       uint32_t chunk, uint8_t *mask);
void     vpo_mask_downsample(uint32_t root, int level, const uint8_t *src, uint8_t *dst);
uint32_t vpo_make_splatlist(const vp_world_desc *w, uint32_t chunk, int level,
                            const uint8_t *mask, const uint16_t *map, uint32_t size_x,
                            int16_t *geom);
/* One chunk through the chunkset_manage sequence (chunkset.c:318-505).
 * Writes L0..L4 records into geom, items[l] = int16 items of level l, returns
 * the total.  mask/work are caller scratch of (R+1)^3 bytes each. */
uint32_t vpo_chunk_remesh(const vp_world_desc *w, const uint8_t *voxels, uint32_t chunk,
                          const uint16_t *map, uint32_t size_x, int lod,
                          int16_t *geom, uint32_t items[5],
                          uint8_t *mask, uint8_t *work);
/* A list of chunks on `threads` OpenMP threads (chunkset.c:251), then the
 * lists concatenated in list order.  Returns total items, or -1 if out_items
 * capacity is too small (offsets/counts are still filled). */
int64_t  vpo_remesh_list(const vp_world_desc *w, const uint8_t *voxels,
                         const uint32_t *ids, uint32_t n,
                         const uint16_t *map, uint32_t size_x, int lod, int threads,
                         int16_t *out, uint64_t cap_items,
                         uint64_t *offsets, uint32_t *level_counts);
/* Same work without the concatenation: timing leg for the CPU baseline. */
double   vpo_remesh_time(const vp_world_desc *w, const uint8_t *voxels,
                         const uint32_t *ids, uint32_t n,
                         const uint16_t *map, uint32_t size_x, int threads,
                         uint64_t *total_items);

/* ---- near-field quad mesh (mesher.c:184-357) ------------------------------ */
/* One chunk: returns the face count f; geom gets 16 int16 per face (4
 * vertices of x, y, z, attribute), index 6 u32 per face (chunk-local). */
uint32_t vpo_chunk_make_mesh(const vp_world_desc *w, const uint8_t *voxels, uint32_t chunk,
                             const uint16_t *map, uint32_t size_x,
                             int16_t *geom, uint32_t *index);
/* A list of chunks, concatenated in list order; face_offsets[n+1]. Returns
 * the face total, -1 if it exceeds cap_faces (offsets are still written). */
int64_t  vpo_mesh_list(const vp_world_desc *w, const uint8_t *voxels, const uint32_t *ids,
                       uint32_t n, const uint16_t *map, uint32_t size_x, int threads,
                       int16_t *geom_out, uint32_t *index_out, uint64_t cap_faces,
                       uint64_t *face_offsets);

/* ---- RLE (rle.c:44-116) -------------------------------------------------- */
/* returns words written (terminator included), 0 if cap is too small */
uint32_t vpo_rle_compress(const uint8_t *data, uint32_t len, uint32_t *words, uint32_t cap);
/* returns voxels the stream expands to (writes at most cap), -1 on empty input */
int64_t  vpo_rle_decompress(const uint32_t *words, uint32_t nwords, uint8_t *out, uint32_t cap);

/* ---- edits (edit.c:89-244) ---------------------------------------------- */
void     vpo_edit_sphere(const vp_world_desc *w, uint8_t *voxels, uint16_t *map,
                         uint32_t size_x, const int32_t centre[3], uint32_t radius,
                         uint8_t voxel, uint8_t *dirty);

/* ---- splat resolve (splat.vert.glsl / splat.frag.glsl) ------------------- */
void     vpo_splat_render(const int16_t *splats, const uint8_t *lod, uint32_t lod_log2,
                          uint64_t n, const vp_camera *cam, const float *voxel_rgb,
                          float *depth, uint32_t *id, float *rgb, float *weight,
                          int threads);
void     vpo_splat_backward(uint64_t n, const vp_camera *cam, const uint32_t *id,
                            const float *weight, const float *dL_drgb, double *grad);

#ifdef __cplusplus
}
#endif
#endif
