/*
 * cpu_zlib.c -- stand-in for the reference's CPU backend, timed beside the GPU
 * in bench.py.  BENCH INFRASTRUCTURE ONLY; the product never uses it.
 *
 * Same structure as gpzip-codec-cpu (codec-cpu/src/backend.rs:14-35,
 * codec-cpu/src/parallel.rs:21-229, codec-cpu/src/compress.rs:52-59): chunks
 * compressed independently on `threads` workers, each chunk one gzip member
 * (raw deflate at the given level, windowBits -15, wrapped with the 10-byte
 * header 1f 8b 08 00 00000000 00 ff, CRC-32 and ISIZE), members written in
 * chunk order.  It runs on system zlib, not the reference's zlib-rs 0.6.3, so
 * member bytes are not expected to match the reference CPU path.
 */
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <zlib.h>

typedef struct {
    const uint8_t *in;
    const uint64_t *off, *len;
    uint8_t **out;
    uint64_t *out_len;
    size_t n;
    int level;
    size_t next;
    pthread_mutex_t mu;
    int err;
} job_t;

static int gzip_member(const uint8_t *in, uint64_t n, int level, uint8_t **out, uint64_t *out_len) {
    z_stream zs;
    memset(&zs, 0, sizeof(zs));
    if (deflateInit2(&zs, level, Z_DEFLATED, -15, 8, Z_DEFAULT_STRATEGY) != Z_OK) return -1;
    uLong bound = deflateBound(&zs, (uLong)n) + 18;
    uint8_t *buf = (uint8_t *)malloc(bound);
    if (!buf) {
        deflateEnd(&zs);
        return -1;
    }
    static const uint8_t hdr[10] = {0x1f, 0x8b, 0x08, 0, 0, 0, 0, 0, 0, 0xff};
    memcpy(buf, hdr, 10);
    zs.next_in = (Bytef *)in;
    zs.avail_in = (uInt)n;
    zs.next_out = buf + 10;
    zs.avail_out = (uInt)(bound - 18);
    int rc = deflate(&zs, Z_FINISH);
    uint64_t dl = zs.total_out;
    deflateEnd(&zs);
    if (rc != Z_STREAM_END) {
        free(buf);
        return -1;
    }
    uint32_t crc = (uint32_t)crc32(0L, in, (uInt)n), isz = (uint32_t)n;
    for (int i = 0; i < 4; i++) buf[10 + dl + i] = (uint8_t)(crc >> (8 * i));
    for (int i = 0; i < 4; i++) buf[14 + dl + i] = (uint8_t)(isz >> (8 * i));
    *out = buf;
    *out_len = dl + 18;
    return 0;
}

static void *worker(void *arg) {
    job_t *j = (job_t *)arg;
    for (;;) {
        pthread_mutex_lock(&j->mu);
        size_t i = j->next++;
        pthread_mutex_unlock(&j->mu);
        if (i >= j->n) break;
        if (gzip_member(j->in + j->off[i], j->len[i], j->level, &j->out[i], &j->out_len[i])) j->err = 1;
    }
    return NULL;
}

/* Compress n chunks (in + off[i], len[i]); writes the concatenated members to
 * out (capacity cap) and each member's length to out_lens.  Returns total
 * bytes, or -1. */
int64_t cpuz_compress_chunks(const uint8_t *in, const uint64_t *off, const uint64_t *len, size_t n, int level,
                             int threads, uint8_t *out, uint64_t cap, uint64_t *out_lens) {
    job_t j;
    memset(&j, 0, sizeof(j));
    j.in = in;
    j.off = off;
    j.len = len;
    j.n = n;
    j.level = level;
    j.out = (uint8_t **)calloc(n ? n : 1, sizeof(uint8_t *));
    j.out_len = out_lens;
    pthread_mutex_init(&j.mu, NULL);
    if (threads < 1) threads = 1;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * threads);
    for (int t = 0; t < threads; t++) pthread_create(&th[t], NULL, worker, &j);
    for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
    free(th);
    pthread_mutex_destroy(&j.mu);
    int64_t total = 0;
    for (size_t i = 0; i < n; i++) {
        if (!j.err && out && (uint64_t)total + out_lens[i] <= cap) memcpy(out + total, j.out[i], out_lens[i]);
        else if (out) j.err = 1;
        total += (int64_t)out_lens[i];
        free(j.out[i]);
    }
    free(j.out);
    return j.err ? -1 : total;
}

const char *cpuz_zlib_version(void) { return zlibVersion(); }
