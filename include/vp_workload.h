/*
 * vp_workload.h -- device-side workload builders of the benchmark.
 *
 * These have no counterpart in the reference's remesh/splat boundary: they
 * build the inputs of the configurations on the GPU (a 128 GiB world does not
 * go through the host).  Same conventions as vp_gpu.h.
 *
 *  vp_gpu_generate      seeded value-noise world (the reference's generator,
 *                       src/chunkset/gen.c:187-341, uses FastNoise; only its
 *                       shape is kept: solid below a height, colour by band)
 *  vp_gpu_classify      per-chunk uniform value (-1 = mixed), the metadata the
 *                       reference keeps as "rle == null_chunk->rle"
 *                       (src/chunkset.c:143-145, 225-228)
 *  vp_gpu_shadow_build  shadow map of a world in the serial placement order
 *                       chunks-by-index, x, z, y (src/chunkset/gen.c:236-306,
 *                       src/chunkset/shadow.h:77-89)
 *  vp_gpu_edit_spheres  sphere add/remove edits applied in list order
 *                       (src/chunkset/edit.c:179-244); marks touched chunks
 *  vp_gpu_shadow_spheres the edits' shadow_place_update calls, in their order
 *                       (src/chunkset/shadow.h:77-89, edit.c:210-236)
 */
#ifndef VP_WORKLOAD_H
#define VP_WORKLOAD_H

#include "vp_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    uint64_t seed;
    uint32_t period;   /* lattice period of octave 0 (halves per octave)   */
    int32_t  hmin;     /* column height range, inclusive                    */
    int32_t  hmax;
} vp_gen_params;

/* Generate chunks [chunk0, chunk0 + count) of d_voxels; d_uniform (nullable,
 * indexed by chunk) receives each chunk's uniform value or -1. */
int vp_gpu_generate(vp_ctx *ctx, const vp_world_desc *world, const vp_gen_params *p,
                    uint8_t *d_voxels, uint32_t chunk0, uint32_t count,
                    int16_t *d_uniform, void *stream);

int vp_gpu_classify(vp_ctx *ctx, const vp_world_desc *world, const uint8_t *d_voxels,
                    uint32_t chunk0, uint32_t count, int16_t *d_uniform, void *stream);

/* d_map: vp_shadow_padded_len(X+Y, Z) entries, fully overwritten. */
int vp_gpu_shadow_build(vp_ctx *ctx, const vp_world_desc *world, const uint8_t *d_voxels,
                        uint16_t *d_map, void *stream);

typedef struct {
    int32_t  cx, cy, cz;   /* centre (world voxels)          */
    uint32_t radius;       /* 1..                            */
    uint32_t voxel;        /* 0 = remove, else the colour    */
} vp_sphere;

/* Apply spheres[0..n) (host array) in order: a voxel ends with the value of
 * the last sphere that covers it (y >= 2 only, edit.c:151).  d_dirty
 * (nullable, one byte per chunk) is set to 1 for every in-world chunk of a
 * sphere's bounding box (edit.c:241-242); d_uniform (nullable) becomes -1 for
 * those chunks.  The shadow map is not touched: see vp_host_shadow_spheres. */
int vp_gpu_edit_spheres(vp_ctx *ctx, const vp_world_desc *world, uint8_t *d_voxels,
                        const vp_sphere *spheres, uint32_t n, uint8_t *d_dirty,
                        int16_t *d_uniform, void *stream);

/* Device twin of vp_host_shadow_spheres: the same placements in the same
 * order on the device map, so edits need no host copy of it.  Map rows (fixed
 * wz) are independent and each is updated by one wave with the row in LDS;
 * spheres[0..n) is a host array.  Needs size_x = X + Y <= 32768
 * (VP_ERR_UNSUPPORTED otherwise).  Synchronous on `stream`. */
int vp_gpu_shadow_spheres(vp_ctx *ctx, const vp_world_desc *world, uint16_t *d_map,
                          const vp_sphere *spheres, uint32_t n, void *stream);

/* Host: shadow_place_update for every voxel the add-spheres cover, in the
 * order of edit.c:210-236 (chunk list cx, cy, cz; then x, y, z), on a host
 * copy of the map (remove-spheres leave the map alone, shadow.h:91-97). */
int vp_host_shadow_spheres(const vp_world_desc *world, uint16_t *map,
                           const vp_sphere *spheres, uint32_t n);

#ifdef __cplusplus
}
#endif
#endif
