/*
 * vp_world_io.c -- the on-disk RLE world: the file command_export writes
 * (src/deadcode.c:320-346) and chunkset_clear_import reads back
 * (src/deadcode.c:159-229).  The chunks' RLE streams are made and expanded on
 * the GPU by vp_gpu_rle; this file moves them between device memory and the
 * file in batches of chunks, so a world of any size needs only a bounded
 * amount of host and device memory.
 *
 * File layout (little endian, as the reference writes it on x86):
 *   byte 0        0x89
 *   bytes 1..7    "VOXPLAT"
 *   byte 8        root_bitw
 *   bytes 9..11   max_bitw[0..2]
 *   then, for every chunk in chunk-index order, its RLE stream: u32 words
 *   (value << 24 | run), the 0 terminator included (rle.c:44-87);
 *   then, optionally, the shadow map: size_x * size_z u16 entries.
 * Two byte counts of the dead export are not kept, because they no longer
 * match the data they write: its stream length scan steps two bytes at a time
 * and stops at the first zero byte at an even offset (the two-byte units of
 * an older stream format; on the 32-bit words it stops inside the first
 * word), and it writes shadow_map_length bytes of a map of shadow_map_length
 * u16 entries (shadow.h:35-42).  Here each stream is written whole and the
 * map entirely.
 */
#define _POSIX_C_SOURCE 200809L
#define _FILE_OFFSET_BITS 64
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <hip/hip_runtime_api.h>

#include "vp_gpu.h"
#include "vp_host.h"

static const uint8_t kMagic[8] = {0x89, 'V', 'O', 'X', 'P', 'L', 'A', 'T'};

/* chunks per batch, and the word budget of one load batch */
#define IO_BATCH_CHUNKS 4096u
#define IO_BLOCK_WORDS (16u << 20)

static int world_ok(const vp_world_desc *w)
{
    if (!w || w->root_bitw < 2 || w->root_bitw > 8) return 0;
    unsigned s = 0;
    for (int a = 0; a < 3; a++) {
        if (w->max_bitw[a] > 16) return 0;
        s += w->max_bitw[a];
    }
    return s <= 32;
}

static uint64_t world_chunks(const vp_world_desc *w)
{
    return (uint64_t)1 << (w->max_bitw[0] + w->max_bitw[1] + w->max_bitw[2]);
}

static void shadow_dims(const vp_world_desc *w, uint32_t *sx, uint32_t *sz)
{
    const uint32_t b = w->root_bitw;
    *sx = ((1u << w->max_bitw[0]) << b) + ((1u << w->max_bitw[1]) << b);
    *sz = (1u << w->max_bitw[2]) << b;
}

/* device buffer that grows on demand */
typedef struct {
    void *p;
    size_t cap;
} dbuf;

static int dbuf_need(dbuf *b, size_t bytes)
{
    if (bytes <= b->cap) return VP_OK;
    if (b->p && hipFree(b->p) != hipSuccess) return VP_ERR_HIP;
    b->p = NULL;
    b->cap = 0;
    if (hipMalloc(&b->p, bytes) != hipSuccess) return VP_ERR_NOMEM;
    b->cap = bytes;
    return VP_OK;
}

static void dbuf_free(dbuf *b)
{
    if (b->p) (void)hipFree(b->p);
    b->p = NULL;
    b->cap = 0;
}

/* Status bits raised by the batch's kernels: any of them makes the file bad. */
static int check_status(vp_ctx *ctx, void *stream)
{
    uint32_t st = 0;
    int rc = vp_ctx_status(ctx, stream, &st);
    if (rc) return rc;
    return st ? VP_ERR_DATA : VP_OK;
}

int vp_host_world_save(vp_ctx *ctx, const char *path, const vp_world_desc *world,
                       const uint8_t *d_voxels, const uint16_t *d_shadow_map, void *stream,
                       uint64_t *bytes_written)
{
    if (!ctx || !path || !world_ok(world) || !d_voxels) return VP_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    FILE *f = fopen(path, "wb");
    if (!f) return VP_ERR_ARG;
    int rc = VP_OK;
    uint64_t written = 0;
    dbuf ids = {0}, offs = {0}, words = {0};
    uint32_t *h_ids = NULL, *h_words = NULL;
    uint64_t *h_offs = NULL;
    size_t h_words_cap = 0;
    const uint64_t nch = world_chunks(world);
    const uint32_t nb_max = nch < IO_BATCH_CHUNKS ? (uint32_t)nch : IO_BATCH_CHUNKS;
    uint8_t head[12];
    memcpy(head, kMagic, 8);
    head[8] = world->root_bitw;
    memcpy(head + 9, world->max_bitw, 3);
    if (fwrite(head, 1, 12, f) != 12) { rc = VP_ERR_ARG; goto done; }
    written = 12;
    h_ids = (uint32_t *)malloc(nb_max * sizeof(uint32_t));
    h_offs = (uint64_t *)malloc((nb_max + 1) * sizeof(uint64_t));
    if (!h_ids || !h_offs) { rc = VP_ERR_NOMEM; goto done; }
    if ((rc = dbuf_need(&ids, nb_max * sizeof(uint32_t))) ||
        (rc = dbuf_need(&offs, (nb_max + 1) * sizeof(uint64_t))))
        goto done;
    for (uint64_t c0 = 0; c0 < nch; c0 += nb_max) {
        const uint32_t nb = (uint32_t)(nch - c0 < nb_max ? nch - c0 : nb_max);
        for (uint32_t i = 0; i < nb; i++) h_ids[i] = (uint32_t)(c0 + i);
        if (hipMemcpyAsync(ids.p, h_ids, nb * sizeof(uint32_t), hipMemcpyHostToDevice, st) !=
            hipSuccess) { rc = VP_ERR_HIP; goto done; }
        /* sizing pass: offsets only */
        rc = vp_gpu_rle(ctx, VP_RLE_ENCODE, world, (uint8_t *)d_voxels, (const uint32_t *)ids.p,
                        nb, NULL, 0, (uint64_t *)offs.p, NULL, stream);
        if (rc) goto done;
        if (hipMemcpyAsync(h_offs, offs.p, (nb + 1) * sizeof(uint64_t), hipMemcpyDeviceToHost,
                           st) != hipSuccess || hipStreamSynchronize(st) != hipSuccess) {
            rc = VP_ERR_HIP;
            goto done;
        }
        const uint64_t nw = h_offs[nb];
        if ((rc = dbuf_need(&words, nw * sizeof(uint32_t)))) goto done;
        if (nw > h_words_cap) {
            free(h_words);
            h_words = (uint32_t *)malloc(nw * sizeof(uint32_t));
            h_words_cap = h_words ? nw : 0;
            if (!h_words) { rc = VP_ERR_NOMEM; goto done; }
        }
        rc = vp_gpu_rle(ctx, VP_RLE_ENCODE, world, (uint8_t *)d_voxels, (const uint32_t *)ids.p,
                        nb, (uint32_t *)words.p, nw, (uint64_t *)offs.p, NULL, stream);
        if (rc) goto done;
        if (hipMemcpyAsync(h_words, words.p, nw * sizeof(uint32_t), hipMemcpyDeviceToHost, st) !=
            hipSuccess) { rc = VP_ERR_HIP; goto done; }
        if ((rc = check_status(ctx, stream))) goto done;
        if (fwrite(h_words, sizeof(uint32_t), nw, f) != nw) { rc = VP_ERR_ARG; goto done; }
        written += nw * sizeof(uint32_t);
    }
    if (d_shadow_map) {
        uint32_t sx, sz;
        shadow_dims(world, &sx, &sz);
        const size_t n = (size_t)sx * sz;
        uint16_t *h_map = (uint16_t *)malloc(n * sizeof(uint16_t));
        if (!h_map) { rc = VP_ERR_NOMEM; goto done; }
        if (hipMemcpyAsync(h_map, d_shadow_map, n * sizeof(uint16_t), hipMemcpyDeviceToHost, st) !=
                hipSuccess || hipStreamSynchronize(st) != hipSuccess) {
            free(h_map);
            rc = VP_ERR_HIP;
            goto done;
        }
        const size_t put = fwrite(h_map, sizeof(uint16_t), n, f);
        free(h_map);
        if (put != n) { rc = VP_ERR_ARG; goto done; }
        written += n * sizeof(uint16_t);
    }
done:
    if (fclose(f) != 0 && rc == VP_OK) rc = VP_ERR_ARG;
    dbuf_free(&ids);
    dbuf_free(&offs);
    dbuf_free(&words);
    free(h_ids);
    free(h_offs);
    free(h_words);
    if (bytes_written) *bytes_written = rc == VP_OK ? written : 0;
    return rc;
}

int vp_host_world_read_header(const char *path, vp_world_desc *world)
{
    if (!path || !world) return VP_ERR_ARG;
    FILE *f = fopen(path, "rb");
    if (!f) return VP_ERR_ARG;
    uint8_t head[12];
    const size_t got = fread(head, 1, 12, f);
    fclose(f);
    if (got != 12 || memcmp(head, kMagic, 8) != 0) return VP_ERR_DATA;
    vp_world_desc w;
    w.root_bitw = head[8];
    memcpy(w.max_bitw, head + 9, 3);
    if (!world_ok(&w)) return VP_ERR_DATA;
    *world = w;
    return VP_OK;
}

/* Upload and decode the complete streams of one load batch. */
static int decode_batch(vp_ctx *ctx, const vp_world_desc *world, uint8_t *d_voxels,
                        const uint32_t *h_words, uint64_t nw, const uint64_t *h_offs,
                        const uint32_t *h_ids, uint32_t nb, dbuf *words, dbuf *offs, dbuf *ids,
                        void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    int rc;
    if ((rc = dbuf_need(words, nw * sizeof(uint32_t))) ||
        (rc = dbuf_need(offs, (nb + 1) * sizeof(uint64_t))) ||
        (rc = dbuf_need(ids, nb * sizeof(uint32_t))))
        return rc;
    if (hipMemcpyAsync(words->p, h_words, nw * sizeof(uint32_t), hipMemcpyHostToDevice, st) !=
            hipSuccess ||
        hipMemcpyAsync(offs->p, h_offs, (nb + 1) * sizeof(uint64_t), hipMemcpyHostToDevice, st) !=
            hipSuccess ||
        hipMemcpyAsync(ids->p, h_ids, nb * sizeof(uint32_t), hipMemcpyHostToDevice, st) !=
            hipSuccess)
        return VP_ERR_HIP;
    rc = vp_gpu_rle(ctx, VP_RLE_DECODE, world, d_voxels, (const uint32_t *)ids->p, nb,
                    (uint32_t *)words->p, nw, (uint64_t *)offs->p, NULL, stream);
    if (rc) return rc;
    /* the host buffers are reused by the next batch */
    return check_status(ctx, stream);
}

int vp_host_world_load(vp_ctx *ctx, const char *path, const vp_world_desc *world,
                       uint8_t *d_voxels, int16_t *d_uniform, uint16_t *d_shadow_map,
                       void *stream)
{
    if (!ctx || !path || !world_ok(world) || !d_voxels) return VP_ERR_ARG;
    vp_world_desc fw;
    int rc = vp_host_world_read_header(path, &fw);
    if (rc) return rc;
    if (fw.root_bitw != world->root_bitw || memcmp(fw.max_bitw, world->max_bitw, 3) != 0)
        return VP_ERR_DATA;
    FILE *f = fopen(path, "rb");
    if (!f) return VP_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    const uint64_t nch = world_chunks(world);
    const uint64_t cells = (uint64_t)1 << (3 * world->root_bitw);
    /* the longest valid stream: one word per voxel and the terminator */
    const uint64_t max_stream = cells + 1;
    const uint64_t cap = max_stream > IO_BLOCK_WORDS ? 2 * max_stream : IO_BLOCK_WORDS;
    uint32_t *h_words = (uint32_t *)malloc(cap * sizeof(uint32_t));
    uint64_t *h_offs = (uint64_t *)malloc((IO_BATCH_CHUNKS + 1) * sizeof(uint64_t));
    uint32_t *h_ids = (uint32_t *)malloc(IO_BATCH_CHUNKS * sizeof(uint32_t));
    int16_t *h_uni = d_uniform ? (int16_t *)malloc(nch * sizeof(int16_t)) : NULL;
    dbuf words = {0}, offs = {0}, ids = {0};
    if (!h_words || !h_offs || !h_ids || (d_uniform && !h_uni)) { rc = VP_ERR_NOMEM; goto done; }
    if (fseeko(f, 12, SEEK_SET) != 0) { rc = VP_ERR_DATA; goto done; }
    uint64_t have = 0, chunk = 0;   /* words in the buffer; next chunk to find */
    int eof = 0;
    while (chunk < nch) {
        if (!eof && have < cap) {
            const size_t want = (size_t)(cap - have);
            const size_t got = fread(h_words + have, sizeof(uint32_t), want, f);
            have += got;
            if (got < want) eof = 1;
        }
        /* cut complete streams out of the buffer */
        uint64_t pos = 0;
        uint32_t nb = 0;
        h_offs[0] = 0;
        while (chunk + nb < nch && nb < IO_BATCH_CHUNKS) {
            uint64_t e = pos;
            while (e < have && h_words[e] != 0u) e++;
            if (e >= have) break;                     /* stream continues past the buffer */
            const uint64_t len = e + 1 - pos;
            if (len > max_stream) { rc = VP_ERR_DATA; goto done; }
            const uint64_t c = chunk + nb;
            if (h_uni) {
                /* a single full-chunk run: the uniform value; anything else mixed */
                const uint32_t w0 = h_words[pos];
                h_uni[c] = (len == 2 && (w0 & 0xFFFFFFu) == cells) ? (int16_t)(w0 >> 24)
                                                                    : (int16_t)-1;
            }
            h_ids[nb] = (uint32_t)c;
            pos = e + 1;
            h_offs[++nb] = pos;
        }
        if (nb == 0) {
            /* no complete stream in a buffer that holds more than the longest one */
            rc = VP_ERR_DATA;
            goto done;
        }
        rc = decode_batch(ctx, world, d_voxels, h_words, pos, h_offs, h_ids, nb, &words, &offs,
                          &ids, stream);
        if (rc) goto done;
        chunk += nb;
        memmove(h_words, h_words + pos, (have - pos) * sizeof(uint32_t));
        have -= pos;
    }
    /* what follows the streams: nothing, or the whole shadow map */
    {
        uint32_t sx, sz;
        shadow_dims(world, &sx, &sz);
        /* u16 entries read as u32 words: sx and sz are multiples of the chunk edge */
        const uint64_t map_words = (uint64_t)sx * sz / 2;
        uint64_t rest = have;
        if (!eof) {
            const off_t here = ftello(f);
            if (fseeko(f, 0, SEEK_END) != 0) { rc = VP_ERR_DATA; goto done; }
            const off_t end = ftello(f);
            rest += (uint64_t)(end - here) / sizeof(uint32_t);
            if ((end - here) % sizeof(uint32_t)) { rc = VP_ERR_DATA; goto done; }
            if (fseeko(f, here, SEEK_SET) != 0) { rc = VP_ERR_DATA; goto done; }
        }
        const int has_map = rest != 0;
        if (has_map && rest != map_words) { rc = VP_ERR_DATA; goto done; }
        if (d_shadow_map) {
            if (!has_map) { rc = VP_ERR_DATA; goto done; }
            const uint64_t padded = vp_shadow_padded_len(sx, sz);
            uint16_t *h_map = (uint16_t *)calloc(padded, sizeof(uint16_t));
            if (!h_map) { rc = VP_ERR_NOMEM; goto done; }
            memcpy(h_map, h_words, have * sizeof(uint32_t));
            const size_t need = (size_t)(map_words - have);
            if (need && fread((uint32_t *)h_map + have, sizeof(uint32_t), need, f) != need) {
                free(h_map);
                rc = VP_ERR_DATA;
                goto done;
            }
            const int bad = hipMemcpyAsync(d_shadow_map, h_map, padded * sizeof(uint16_t),
                                           hipMemcpyHostToDevice, st) != hipSuccess ||
                            hipStreamSynchronize(st) != hipSuccess;
            free(h_map);
            if (bad) { rc = VP_ERR_HIP; goto done; }
        }
    }
    if (h_uni) {
        if (hipMemcpyAsync(d_uniform, h_uni, nch * sizeof(int16_t), hipMemcpyHostToDevice, st) !=
                hipSuccess || hipStreamSynchronize(st) != hipSuccess)
            rc = VP_ERR_HIP;
    }
done:
    fclose(f);
    (void)hipStreamSynchronize(st);
    dbuf_free(&words);
    dbuf_free(&offs);
    dbuf_free(&ids);
    free(h_words);
    free(h_offs);
    free(h_ids);
    free(h_uni);
    return rc;
}
