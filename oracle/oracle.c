/*
 * oracle.c -- CPU restatement of the reference TopoSZp OpenMP path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load this file's library; the product library never
 * links it and never falls back to it.
 *
 * Written from a study of the reference's behaviour; each function cites the
 * reference file:line it restates.  Serial code: a chunk count T only changes the
 * chunk-offset header, exactly as the reference's OpenMP writers do.
 *
 * Pinning (see tests/test_oracle.py):
 *   - golden vectors produced by the reference's own CPU sources (tests/golden/),
 *   - brute-force sort positions, exhaustive bit_count table, encode/decode round trips.
 */
#include "oracle.h"

#include <float.h>
#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* scalar rules                                                              */
/* ------------------------------------------------------------------------ */

/* bit_count = floorf(log2f((float)max)) + 1   (SZp/src/szp_float.cc:1243; the
 * `(int)(log2f(..)) + 1` spelling at :822 and :2241 is the same for max >= 1). */
unsigned or_bitcount(unsigned maxv)
{
    if (maxv == 0) return 0;
    return (unsigned)floorf(log2f((float)maxv)) + 1u;
}

uint64_t or_check_bitcount_table(const uint32_t *thr, uint64_t lo, uint64_t hi)
{
    uint64_t bad = 0;
    for (uint64_t x = lo; x < hi; x++) {
        unsigned v = (unsigned)x;
        if (v == 0) continue;
        unsigned k = 31u - (unsigned)__builtin_clz(v);
        unsigned want = k + 1u + (v >= thr[k] ? 1u : 0u);
        if (or_bitcount(v) != want) bad++;
    }
    return bad;
}

/* The reference converts with a C cast, which on x86-64 is cvttsd2si: truncation
 * toward zero and 0x80000000 for NaN or out-of-range input. */
int or_cvt_i32(double d)
{
    if (!(d > -2147483649.0 && d < 2147483648.0)) return INT_MIN;
    return (int)d;
}

/* SZp/include/szp_errbound.h:34-51: serial float min/max, float range, float product. */
int or_rel_to_abs(const float *data, size_t n, float rel, float *eb, float *range)
{
    if (!data || n == 0 || rel <= 0) return -1;
    float lo = data[0], hi = data[0];
    for (size_t i = 1; i < n; i++) {
        if (data[i] < lo) lo = data[i];
        if (data[i] > hi) hi = data[i];
    }
    float r = hi - lo;
    if (r <= 0) r = FLT_EPSILON;
    *eb = rel * r;
    if (range) *range = r;
    return 0;
}

/* quantized bin of a critical point: float add, then double multiply
 * (SZp/src/szp_float.cc:46,76; SZp/src/szp_topology_3d.cc:101). */
static int cp_bin(float v, float eb, double inv)
{
    float s = v + eb;
    return or_cvt_i32((double)s * inv);
}

/* ------------------------------------------------------------------------ */
/* critical points                                                          */
/* ------------------------------------------------------------------------ */

static int type_4nbr(float c, float up, float dn, float lf, float rt)
{
    /* SZp/src/szp_float.cc:73-90 */
    if (c > up && c > dn && c > lf && c > rt) return 1;
    if (c < up && c < dn && c < lf && c < rt) return 2;
    if ((c < up && c < dn && c > lf && c > rt) || (c > up && c > dn && c < lf && c < rt)) return 3;
    return 0;
}

static int type_6nbr(float c, float xm, float xp, float ym, float yp, float zm, float zp)
{
    /* SZp/src/szp_topology_3d.cc:79-98 */
    if (c > xm && c > xp && c > ym && c > yp && c > zm && c > zp) return 1;
    if (c < xm && c < xp && c < ym && c < yp && c < zm && c < zp) return 2;
    int hi = (c > xm && c > xp) + (c > ym && c > yp) + (c > zm && c > zp);
    int lo = (c < xm && c < xp) + (c < ym && c < yp) + (c < zm && c < zp);
    return (hi >= 1 && lo >= 1) ? 3 : 0;
}

/* SZp/src/szp_float.cc:28-120 with one thread: interior points in row-major order. */
size_t or_find_cps_2d(const float *data, int rows, int cols, float eb, or_cp2 *out)
{
    if (!data || rows <= 2 || cols <= 2) return 0;
    double inv = 1.0 / eb;
    size_t cnt = 0;
    for (int i = 1; i < rows - 1; i++) {
        for (int j = 1; j < cols - 1; j++) {
            size_t f = (size_t)i * cols + j;
            float c = data[f];
            int t = type_4nbr(c, data[f - cols], data[f + cols], data[f - 1], data[f + 1]);
            if (!t) continue;
            or_cp2 *p = &out[cnt++];
            p->x = i; p->y = j; p->type = t;
            p->quantized_bin = cp_bin(c, eb, inv);
            p->sort_position = 0;
        }
    }
    return cnt;
}

/* SZp/src/szp_topology_3d.cc:30-129 with one thread. */
size_t or_find_cps_3d(const float *data, int d1, int d2, int d3, float eb, or_cp3 *out)
{
    if (!data || d1 <= 2 || d2 <= 2 || d3 <= 2) return 0;
    double inv = 1.0 / eb;
    size_t plane = (size_t)d2 * d3, cnt = 0;
    for (int x = 1; x < d1 - 1; x++)
        for (int y = 1; y < d2 - 1; y++)
            for (int z = 1; z < d3 - 1; z++) {
                size_t f = (size_t)x * plane + (size_t)y * d3 + z;
                float c = data[f];
                int t = type_6nbr(c, data[f - plane], data[f + plane], data[f - d3], data[f + d3],
                                  data[f - 1], data[f + 1]);
                if (!t) continue;
                or_cp3 *p = &out[cnt++];
                p->x = x; p->y = y; p->z = z; p->type = t;
                p->quantized_bin = cp_bin(c, eb, inv);
                p->sort_position = 0;
            }
    return cnt;
}

void or_classify_2d(const float *data, int rows, int cols, unsigned char *types)
{
    memset(types, 0, (size_t)rows * cols);
    for (int i = 1; i < rows - 1; i++)
        for (int j = 1; j < cols - 1; j++) {
            size_t f = (size_t)i * cols + j;
            types[f] = (unsigned char)type_4nbr(data[f], data[f - cols], data[f + cols],
                                                data[f - 1], data[f + 1]);
        }
}

void or_classify_3d(const float *data, int d1, int d2, int d3, unsigned char *types)
{
    size_t plane = (size_t)d2 * d3;
    memset(types, 0, (size_t)d1 * plane);
    for (int x = 1; x < d1 - 1; x++)
        for (int y = 1; y < d2 - 1; y++)
            for (int z = 1; z < d3 - 1; z++) {
                size_t f = (size_t)x * plane + (size_t)y * d3 + z;
                types[f] = (unsigned char)type_6nbr(data[f], data[f - plane], data[f + plane],
                                                    data[f - d3], data[f + d3], data[f - 1], data[f + 1]);
            }
}

/* ------------------------------------------------------------------------ */
/* sort positions: dense rank of distinct values within a quantized bin      */
/* SZp/src/szp_float.cc:137-509, SZp/src/szp_topology_3d.cc:135-247          */
/* ------------------------------------------------------------------------ */

typedef struct { int bin; float v; size_t idx; } rank_item;

static int rank_cmp(const void *a, const void *b)
{
    const rank_item *p = (const rank_item *)a, *q = (const rank_item *)b;
    if (p->bin != q->bin) return p->bin < q->bin ? -1 : 1;
    if (p->v < q->v) return -1;
    if (p->v > q->v) return 1;
    return p->idx < q->idx ? -1 : (p->idx > q->idx);
}

/* Ranks count distinct float values (so -0.0 and +0.0 share one rank, as the
 * reference's `!=` comparison does); the CP array order is left unchanged. */
static void dense_rank(rank_item *it, size_t n, int *rank_out /* indexed by idx */)
{
    qsort(it, n, sizeof(rank_item), rank_cmp);
    int r = 0;
    for (size_t k = 0; k < n; k++) {
        if (k == 0 || it[k].bin != it[k - 1].bin) r = 0;
        else if (it[k].v != it[k - 1].v) r++;
        rank_out[it[k].idx] = r;
    }
}

void or_sort_positions_2d(or_cp2 *cps, size_t n, const float *data, int cols)
{
    if (!cps || !n || !data) return;
    rank_item *it = (rank_item *)malloc(n * sizeof(rank_item));
    int *r = (int *)malloc(n * sizeof(int));
    for (size_t k = 0; k < n; k++) {
        it[k].bin = cps[k].quantized_bin;
        it[k].v = data[(size_t)cps[k].x * cols + cps[k].y];
        it[k].idx = k;
    }
    dense_rank(it, n, r);
    for (size_t k = 0; k < n; k++) cps[k].sort_position = r[k];
    free(it); free(r);
}

void or_sort_positions_3d(or_cp3 *cps, size_t n, const float *data, int d2, int d3)
{
    if (!cps || !n || !data) return;
    rank_item *it = (rank_item *)malloc(n * sizeof(rank_item));
    int *r = (int *)malloc(n * sizeof(int));
    for (size_t k = 0; k < n; k++) {
        it[k].bin = cps[k].quantized_bin;
        it[k].v = data[(size_t)cps[k].x * d2 * d3 + (size_t)cps[k].y * d3 + cps[k].z];
        it[k].idx = k;
    }
    dense_rank(it, n, r);
    for (size_t k = 0; k < n; k++) cps[k].sort_position = r[k];
    free(it); free(r);
}

size_t or_extrema_sortpos_2d(const or_cp2 *cps, size_t n, int *sp)
{
    size_t e = 0;
    for (size_t k = 0; k < n; k++)
        if (cps[k].type == 1 || cps[k].type == 2) sp[e++] = cps[k].sort_position;
    return e;
}

size_t or_extrema_sortpos_3d(const or_cp3 *cps, size_t n, int *sp)
{
    size_t e = 0;
    for (size_t k = 0; k < n; k++)
        if (cps[k].type == 1 || cps[k].type == 2) sp[e++] = cps[k].sort_position;
    return e;
}

void or_mark_types_2d(const or_cp2 *cps, size_t n, int rows, int cols, unsigned char *types)
{
    /* SZp/src/szp_float.cc:1040-1064 */
    memset(types, 0, (size_t)rows * cols);
    for (size_t k = 0; k < n; k++) {
        int x = cps[k].x, y = cps[k].y;
        if (x < 0 || x >= rows || y < 0 || y >= cols) continue;
        if (cps[k].type >= 1 && cps[k].type <= 3)
            types[(size_t)x * cols + y] = (unsigned char)cps[k].type;
    }
}

void or_mark_types_3d(const or_cp3 *cps, size_t n, int d1, int d2, int d3, unsigned char *types)
{
    /* SZp/src/szp_topology_3d.cc:470-480 */
    memset(types, 0, (size_t)d1 * d2 * d3);
    for (size_t k = 0; k < n; k++) {
        int x = cps[k].x, y = cps[k].y, z = cps[k].z;
        if (x < 0 || x >= d1 || y < 0 || y >= d2 || z < 0 || z >= d3) continue;
        if (cps[k].type >= 1 && cps[k].type <= 3)
            types[(size_t)x * d2 * d3 + (size_t)y * d3 + z] = (unsigned char)cps[k].type;
    }
}

/* ------------------------------------------------------------------------ */
/* block codec (SZp/src/szp_TypeManager.cc)                                  */
/* ------------------------------------------------------------------------ */

/* write `w` low bits of v at bit position *pos of an MSB-first bit stream */
static void put_bits_msb(unsigned char *dst, size_t *pos, unsigned v, unsigned w)
{
    for (unsigned b = 0; b < w; b++) {
        unsigned bit = (v >> (w - 1 - b)) & 1u;
        size_t p = *pos + b;
        if (bit) dst[p >> 3] |= (unsigned char)(0x80u >> (p & 7));
    }
    *pos += w;
}

static unsigned get_bits_msb(const unsigned char *src, size_t pos, unsigned w)
{
    unsigned v = 0;
    for (unsigned b = 0; b < w; b++) {
        size_t p = pos + b;
        v = (v << 1) | ((src[p >> 3] >> (7 - (p & 7))) & 1u);
    }
    return v;
}

/* bytes of the magnitude section for n values of width bc
 * (Jiajun_save_fixed_length_bits, SZp/src/szp_TypeManager.cc:13-103) */
static size_t mag_bytes(size_t n, unsigned bc)
{
    size_t full = (size_t)(bc / 8) * n;
    unsigned rem = bc % 8;
    return rem ? full + (rem * n + 7) / 8 : full;
}

/* One block: [i32 anchor][u8 bc][signs][magnitudes][2-bit types].  The bc byte is
 * always written, also for a one-element block (SZp/src/szp_float.cc:1239-1242). */
static size_t put_block(unsigned char *p, const int *q, size_t cur, const unsigned char *types)
{
    unsigned char *s = p;
    memcpy(p, &q[0], 4);
    p += 4;
    size_t n = cur - 1;
    unsigned mags[1024];
    unsigned char neg[1024];
    unsigned maxv = 0;
    for (size_t j = 0; j < n; j++) {
        unsigned d = (unsigned)q[j + 1] - (unsigned)q[j];  /* two's-complement difference */
        int sd = (int)d;
        neg[j] = sd < 0;
        mags[j] = sd < 0 ? 0u - d : d;
        if (mags[j] > maxv) maxv = mags[j];
    }
    unsigned bc = or_bitcount(maxv);
    *p++ = (unsigned char)bc;
    if (bc) {
        size_t sb = (n + 7) / 8;
        memset(p, 0, sb);
        for (size_t j = 0; j < n; j++)
            if (neg[j]) p[j >> 3] |= (unsigned char)(0x80u >> (j & 7));
        p += sb;
        /* full-byte part: (v >> rem) little-endian over bc/8 bytes per value */
        unsigned nfull = bc / 8, rem = bc % 8;
        for (size_t j = 0; j < n; j++) {
            unsigned v = mags[j] >> rem;
            for (unsigned k = 0; k < nfull; k++) { *p++ = (unsigned char)v; v >>= 8; }
        }
        /* remainder part: low `rem` bits, contiguous MSB-first */
        if (rem) {
            size_t rb = (rem * n + 7) / 8, pos = 0;
            memset(p, 0, rb);
            for (size_t j = 0; j < n; j++) put_bits_msb(p, &pos, mags[j] & ((1u << rem) - 1u), rem);
            p += rb;
        }
    }
    if (types) {
        size_t tb = (2 * cur + 7) / 8, pos = 0;
        memset(p, 0, tb);
        for (size_t j = 0; j < cur; j++) put_bits_msb(p, &pos, types[j] & 3u, 2);
        p += tb;
    }
    return (size_t)(p - s);
}

/* Decode one block; returns bytes consumed.  `read_bc_always` mirrors the topology
 * decoder (SZp/src/szpd_float.cc:735), the others read bc only when cur > 1
 * (SZp/src/szpd_float.cc:487, :958). */
static size_t get_block(const unsigned char *p, int *q, size_t cur, unsigned char *types,
                        int read_bc_always)
{
    const unsigned char *s = p;
    int prior;
    memcpy(&prior, p, 4);
    p += 4;
    q[0] = prior;
    size_t n = cur - 1;
    if (cur > 1 || read_bc_always) {
        unsigned bc = *p++;
        if (bc == 0) {
            for (size_t j = 0; j < n; j++) q[j + 1] = prior;
        } else {
            const unsigned char *sg = p;
            p += (n + 7) / 8;
            unsigned nfull = bc / 8, rem = bc % 8;
            const unsigned char *full = p, *rb = p + (size_t)nfull * n;
            for (size_t j = 0; j < n; j++) {
                unsigned v = 0;
                for (unsigned k = 0; k < nfull; k++) v |= (unsigned)full[j * nfull + k] << (8 * k);
                v <<= rem;
                if (rem) v |= get_bits_msb(rb, j * rem, rem);
                int neg = (sg[j >> 3] >> (7 - (j & 7))) & 1;
                int d = neg ? -(int)v : (int)v;
                prior = (int)((unsigned)prior + (unsigned)d);
                q[j + 1] = prior;
            }
            p += mag_bytes(n, bc);
        }
    }
    if (types) {
        for (size_t j = 0; j < cur; j++) types[j] = (unsigned char)get_bits_msb(p, 2 * j, 2);
        p += (2 * cur + 7) / 8;
    }
    return (size_t)(p - s);
}

static size_t block_worst(int bs, int topo)
{
    size_t n = (size_t)bs - 1;
    return 4 + 1 + (n + 7) / 8 + 4 * n + 1 + (topo ? (2 * (size_t)bs + 7) / 8 : 0);
}

size_t or_topo_bound(size_t n, int bs, unsigned T)
{
    return 8 * (size_t)T + ((n + bs - 1) / bs) * block_worst(bs, 1);
}
size_t or_plain_bound(size_t n, int bs, unsigned T)
{
    return 4 + 8 * (size_t)T + ((n + bs - 1) / bs) * block_worst(bs, 0);
}
size_t or_sortpos_bound(size_t e, int bs, unsigned T)
{
    return 8 + 8 * (size_t)T + ((e + bs - 1) / bs) * block_worst(bs, 0);
}

/* OpenMP block partition of the writers (SZp/src/szp_float.cc:1148-1176):
 * static ceil(nb/T) blocks per chunk, or, when `adaptive` and T >= 16 and
 * nb < 8T, ceil(nb/ceil(nb/8)) blocks for the first ceil(nb/8) chunks only. */
void or_chunk_range(size_t nb, unsigned T, int adaptive, unsigned t, size_t *b0, size_t *b1)
{
    size_t bpt, s, e;
    if (adaptive && T >= 16 && nb < (size_t)T * 8) {
        size_t opt = (nb + 7) / 8;
        if (opt < 1) opt = 1;
        bpt = (nb + opt - 1) / opt;
        size_t et = t % opt;
        s = et * bpt;
        e = (et + 1) * bpt;
        if (e > nb) e = nb;
        if (t >= opt) s = e;
    } else {
        bpt = (nb + T - 1) / T;
        s = (size_t)t * bpt;
        e = ((size_t)t + 1) * bpt;
        if (e > nb) e = nb;
    }
    if (s > e) s = e;
    *b0 = s; *b1 = e;
}

/* Shared writer: header of T chunk offsets, then every chunk's blocks in order. */
static size_t write_chunks(unsigned char *hdr, const int *q_all, size_t n, int bs,
                           const unsigned char *types, unsigned T, int adaptive)
{
    size_t nb = (n + bs - 1) / bs;
    unsigned char *payload = hdr + 8 * (size_t)T;
    size_t pos = 0;
    for (unsigned t = 0; t < T; t++) {
        size_t b0, b1;
        or_chunk_range(nb, T, adaptive, t, &b0, &b1);
        uint64_t off = pos;
        memcpy(hdr + 8 * (size_t)t, &off, 8);
        for (size_t b = b0; b < b1; b++) {
            size_t i = b * (size_t)bs;
            size_t cur = (i + bs > n) ? n - i : (size_t)bs;
            pos += put_block(payload + pos, q_all + i, cur, types ? types + i : NULL);
        }
    }
    return 8 * (size_t)T + pos;
}

/* SZp/src/szp_float.cc:998-1318 (2D), SZp/src/szp_topology_3d.cc:433-632 (3D, adaptive=0) */
size_t or_topo_compress(const float *data, size_t n, float eb, int bs,
                        const unsigned char *types, unsigned T, int adaptive, unsigned char *out)
{
    if (!(eb > 0) || n == 0 || bs <= 0 || bs > 1025 || T == 0) return 0;
    double inv = 1.0 / eb;
    int *q = (int *)malloc(n * sizeof(int));
    for (size_t i = 0; i < n; i++) q[i] = or_cvt_i32((double)data[i] * inv);
    size_t r = write_chunks(out, q, n, bs, types, T, adaptive);
    free(q);
    return r;
}

/* SZp/src/szp_float.cc:2061-2292: reciprocal taken in float, then widened
 * (:2125), header is the big-endian float error bound (:2071). */
size_t or_plain_compress(const float *data, size_t n, float eb, int bs, unsigned T,
                         int adaptive, unsigned char *out)
{
    if (!(eb > 0) || n == 0 || bs <= 1 || bs > 1025 || T == 0) return 0;
    float finv = 1 / eb;
    double inv = (double)finv;
    uint32_t bits;
    memcpy(&bits, &eb, 4);
    out[0] = (unsigned char)(bits >> 24); out[1] = (unsigned char)(bits >> 16);
    out[2] = (unsigned char)(bits >> 8);  out[3] = (unsigned char)bits;
    int *q = (int *)malloc(n * sizeof(int));
    for (size_t i = 0; i < n; i++) q[i] = or_cvt_i32((double)data[i] * inv);
    size_t r = 4 + write_chunks(out + 4, q, n, bs, NULL, T, adaptive);
    free(q);
    return r;
}

/* SZp/src/szp_float.cc:521-996 (2D, adaptive) and SZp/src/szp_topology_3d.cc:253-426
 * (3D, adaptive=0): [u64 E][u64 off[T]][blocks]. */
size_t or_compress_sortpos(const int *sp, size_t e, int bs, unsigned T, int adaptive,
                           unsigned char *out)
{
    if (!sp || e == 0 || bs <= 0 || bs > 1025 || T == 0) return 0;
    uint64_t ec = e;
    memcpy(out, &ec, 8);
    return 8 + write_chunks(out + 8, sp, e, bs, NULL, T, adaptive);
}

/* ------------------------------------------------------------------------ */
/* chunk-count detection (SZp/include/szp_detect_threads.h)                  */
/* ------------------------------------------------------------------------ */

static uint64_t rd64(const unsigned char *p) { uint64_t v; memcpy(&v, p, 8); return v; }

static size_t fixed_len(unsigned n, unsigned bc)
{
    unsigned nfull = bc / 8, rem = bc % 8;
    unsigned full = nfull * n;
    return rem ? full + (rem * n - 1) / 8 + 1 : full;
}

/* szp_validate_nbThreads_randomaccess[_topo]  (:54-103, :229-279) */
static int validate_T(const unsigned char *cmp, size_t n, int bs, unsigned T, size_t lim, int topo)
{
    const unsigned char *rcp = cmp + 8 * (size_t)T, *end = cmp + lim;
    size_t nb = (n + bs - 1) / bs, bpt = (nb + T - 1) / T;
    for (unsigned t = 0; t < T; t++) {
        size_t s = (size_t)t * bpt, e = ((size_t)t + 1) * bpt;
        if (e > nb) e = nb;
        const unsigned char *p = rcp + rd64(cmp + 8 * (size_t)t);
        if (p >= end || p < rcp) return 0;
        for (size_t b = s; b < e; b++) {
            size_t i = b * (size_t)bs;
            if (i >= n) break;
            size_t cur = (i + bs > n) ? n - i : (size_t)bs;
            if (p + 4 > end) return 0;
            p += 4;
            if (cur > 1) {
                if (p >= end) return 0;
                unsigned bc = *p++;
                if (bc > 32) return 0;
                if (bc) {
                    unsigned m = (unsigned)(cur - 1);
                    p += (m + 7) / 8 + fixed_len(m, bc);
                    if (p > end) return 0;
                }
            }
            if (topo) {
                p += (2 * (unsigned)cur + 7) / 8;
                if (p > end) return 0;
            }
        }
    }
    return 1;
}

/* szp_detect_nbThreads_randomaccess_topo (:284-309): first valid T in 1..128 */
unsigned or_detect_T_topo(const unsigned char *cmp, size_t n, int bs)
{
    size_t lim = 8 * n + 1024;
    if (rd64(cmp) != 0) return 1;
    for (unsigned t = 1; t <= 128; t++) {
        for (unsigned k = 1; k < t; k++) {
            uint64_t a = rd64(cmp + 8 * (size_t)k), b = rd64(cmp + 8 * (size_t)(k - 1));
            if (a > lim || a < b) return 1;
        }
        if (validate_T(cmp, n, bs, t, lim, 1)) return t;
    }
    return 1;
}

/* szp_detect_nbThreads_randomaccess (:161-190): T in 2..128, else 1 */
unsigned or_detect_T_ra(const unsigned char *cmp, size_t n, int bs, size_t elem)
{
    size_t lim = elem * n + elem;
    if (rd64(cmp) != 0) return 1;
    for (unsigned t = 2; t <= 128; t++) {
        for (unsigned k = 1; k < t; k++) {
            uint64_t a = rd64(cmp + 8 * (size_t)k), b = rd64(cmp + 8 * (size_t)(k - 1));
            if (a > lim || a < b) return 1;
        }
        if (validate_T(cmp, n, bs, t, lim, 0)) return t;
    }
    return 1;
}

/* ------------------------------------------------------------------------ */
/* decoders (SZp/src/szpd_float.cc)                                           */
/* ------------------------------------------------------------------------ */

/* :616-781 -- always reads bc; types become int 0..3, boundary points 0. */
void or_topo_decompress(const unsigned char *cmp, size_t n, float eb, int bs, float *out, int *types)
{
    unsigned T = or_detect_T_topo(cmp, n, bs);
    const unsigned char *rcp = cmp + 8 * (size_t)T;
    size_t nb = (n + bs - 1) / bs;
    const float scale = 1 * eb;
    int q[1024];
    unsigned char ty[1024];
    for (unsigned t = 0; t < T; t++) {
        size_t b0, b1;
        or_chunk_range(nb, T, 1, t, &b0, &b1);
        const unsigned char *p = rcp + rd64(cmp + 8 * (size_t)t);
        for (size_t b = b0; b < b1; b++) {
            size_t i = b * (size_t)bs;
            size_t cur = (i + bs > n) ? n - i : (size_t)bs;
            p += get_block(p, q, cur, ty, 1);
            for (size_t j = 0; j < cur; j++) {
                out[i + j] = (float)q[j] * scale;
                if (types) types[i + j] = ty[j];
            }
        }
    }
}

/* :792-1016 -- `cmp` is the stream after its u64 extrema count. */
void or_sortpos_decompress(const unsigned char *cmp, size_t e, int bs, int *out)
{
    unsigned T = or_detect_T_ra(cmp, e, bs, sizeof(float));
    const unsigned char *rcp = cmp + 8 * (size_t)T;
    size_t nb = (e + bs - 1) / bs;
    int q[1024];
    for (unsigned t = 0; t < T; t++) {
        size_t b0, b1;
        or_chunk_range(nb, T, 1, t, &b0, &b1);
        const unsigned char *p = rcp + rd64(cmp + 8 * (size_t)t);
        for (size_t b = b0; b < b1; b++) {
            size_t i = b * (size_t)bs;
            size_t cur = (i + bs > e) ? e - i : (size_t)bs;
            p += get_block(p, q, cur, NULL, 0);
            memcpy(out + i, q, cur * sizeof(int));
        }
    }
}

/* :431-524 -- `cmp` is the stream after its 4-byte error-bound header; the chunk count
 * comes from the reference's detection (szp_detect_nbThreads_randomaccess), which can
 * misread some single-chunk streams, so or_plain_decompress_T takes it from the caller. */
void or_plain_decompress(const unsigned char *cmp, size_t n, float eb, int bs, float *out)
{
    or_plain_decompress_T(cmp, n, eb, bs, or_detect_T_ra(cmp, n, bs, sizeof(float)), out);
}

void or_plain_decompress_T(const unsigned char *cmp, size_t n, float eb, int bs, unsigned T, float *out)
{
    const unsigned char *rcp = cmp + 8 * (size_t)T;
    size_t nb = (n + bs - 1) / bs;
    int q[1024];
    for (unsigned t = 0; t < T; t++) {
        size_t b0, b1;
        or_chunk_range(nb, T, 0, t, &b0, &b1);
        const unsigned char *p = rcp + rd64(cmp + 8 * (size_t)t);
        for (size_t b = b0; b < b1; b++) {
            size_t i = b * (size_t)bs;
            size_t cur = (i + bs > n) ? n - i : (size_t)bs;
            p += get_block(p, q, cur, NULL, 0);
            for (size_t j = 0; j < cur; j++) out[i + j] = (float)q[j] * eb;
        }
    }
}

/* Serial offset walk of a single-chunk payload (the host loop of
 * SZp/src/szp_cuda_decompress.cu:391-477).  `bc_always` selects whether a
 * one-element block carries its bc byte (true for every stream the CPU writers emit). */
int or_block_offsets(const unsigned char *payload, size_t n, int bs, int topo, int bc_always,
                     uint64_t *offs)
{
    size_t nb = (n + bs - 1) / bs;
    uint64_t pos = 0;
    for (size_t b = 0; b < nb; b++) {
        size_t i = b * (size_t)bs;
        size_t cur = (i + bs > n) ? n - i : (size_t)bs;
        offs[b] = pos;
        pos += 4;
        if (cur > 1 || bc_always) {
            unsigned bc = payload[pos++];
            if (bc > 33) return -1;
            if (bc) pos += (cur - 1 + 7) / 8 + mag_bytes(cur - 1, bc);
        }
        if (topo) pos += (2 * cur + 7) / 8;
    }
    offs[nb] = pos;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* element-partitioned writers: the threadblock (chained) codec and the     */
/* double random-access codec                                                */
/* ------------------------------------------------------------------------ */

/* Chunk t of T covers elements [t * (n / T), (t + 1) * (n / T)), the last chunk up
 * to n (SZp/src/szp_float.cc:1489-1504, SZp/src/szp_double.cc:188-199, :790-800). */
static void elem_chunk(size_t n, unsigned T, unsigned t, size_t *lo, size_t *hi)
{
    size_t ce = n / T;
    *lo = (size_t)t * ce;
    *hi = (t == T - 1) ? n : (size_t)(t + 1) * ce;
}

/* [bc][signs][magnitudes] of the m deltas q[j] - q[j - 1], j = 0 .. m - 1 (q[-1] is
 * read); bc only when m == 0 and `bc_always` is false. */
static size_t put_run(unsigned char *p, const int *q, size_t m, int bc_always)
{
    unsigned char *s = p;
    if (m == 0 && !bc_always) return 0;
    unsigned *mags = (unsigned *)malloc((m + 1) * sizeof(unsigned));
    unsigned char *neg = (unsigned char *)malloc(m + 1);
    unsigned maxv = 0;
    for (size_t j = 0; j < m; j++) {
        unsigned d = (unsigned)q[j] - (unsigned)q[(ptrdiff_t)j - 1];
        int sd = (int)d;
        neg[j] = sd < 0;
        mags[j] = sd < 0 ? 0u - d : d;
        if (mags[j] > maxv) maxv = mags[j];
    }
    unsigned bc = or_bitcount(maxv);
    *p++ = (unsigned char)bc;
    if (bc) {
        size_t sb = (m + 7) / 8;
        memset(p, 0, sb);
        for (size_t j = 0; j < m; j++)
            if (neg[j]) p[j >> 3] |= (unsigned char)(0x80u >> (j & 7));
        p += sb;
        unsigned nfull = bc / 8, rem = bc % 8;
        for (size_t j = 0; j < m; j++) {
            unsigned v = mags[j] >> rem;
            for (unsigned k = 0; k < nfull; k++) { *p++ = (unsigned char)v; v >>= 8; }
        }
        if (rem) {
            size_t rb = (rem * m + 7) / 8, pos = 0;
            memset(p, 0, rb);
            for (size_t j = 0; j < m; j++) put_bits_msb(p, &pos, mags[j] & ((1u << rem) - 1u), rem);
            p += rb;
        }
    }
    free(mags);
    free(neg);
    return (size_t)(p - s);
}

/* inverse of put_run: q[j] = q[j - 1] + delta_j; returns bytes consumed */
static size_t get_run(const unsigned char *p, int *q, size_t m, int bc_always)
{
    const unsigned char *s = p;
    if (m == 0 && !bc_always) return 0;
    unsigned bc = *p++;
    int prior = q[-1];
    if (bc == 0) {
        for (size_t j = 0; j < m; j++) q[j] = prior;
        return (size_t)(p - s);
    }
    const unsigned char *sg = p;
    p += (m + 7) / 8;
    unsigned nfull = bc / 8, rem = bc % 8;
    const unsigned char *full = p, *rb = p + (size_t)nfull * m;
    for (size_t j = 0; j < m; j++) {
        unsigned v = 0;
        for (unsigned k = 0; k < nfull; k++) v |= (unsigned)full[j * nfull + k] << (8 * k);
        v <<= rem;
        if (rem) v |= get_bits_msb(rb, j * rem, rem);
        int ng = (sg[j >> 3] >> (7 - (j & 7))) & 1;
        int d = ng ? -(int)v : (int)v;
        prior = (int)((unsigned)prior + (unsigned)d);
        q[j] = prior;
    }
    p += mag_bytes(m, bc);
    return (size_t)(p - s);
}

static void put_be(unsigned char *out, const void *v, int size)
{
    const unsigned char *b = (const unsigned char *)v;
    for (int k = 0; k < size; k++) out[k] = b[size - 1 - k];
}

/* One chunk of the element-partitioned stream.
 *   ra = 0: threadblock -- [i32 q[lo]] then blocks of bs deltas from lo + 1, chained
 *           across blocks (SZp/src/szp_float.cc:1505-1570, szp_double.cc:200-278);
 *   ra = 1: double random access -- blocks of bs elements from lo, each
 *           [i32 anchor][bc][signs][magnitudes] over its cur - 1 deltas, the bc byte
 *           also for a one-element block (SZp/src/szp_double.cc:810-880). */
static size_t put_elem_chunk(unsigned char *p, const int *q, size_t lo, size_t hi, int bs, int ra)
{
    unsigned char *s = p;
    if (lo >= hi) return 0;
    if (ra) {
        for (size_t i = lo; i < hi; i += (size_t)bs) {
            size_t cur = (i + bs > hi) ? hi - i : (size_t)bs;
            memcpy(p, &q[i], 4);
            p += 4;
            p += put_run(p, q + i + 1, cur - 1, 1);
        }
    } else {
        memcpy(p, &q[lo], 4);
        p += 4;
        for (size_t i = lo + 1; i < hi; i += (size_t)bs) {
            size_t cur = (i + bs > hi) ? hi - i : (size_t)bs;
            p += put_run(p, q + i, cur, 0);
        }
    }
    return (size_t)(p - s);
}

static size_t get_elem_chunk(const unsigned char *p, int *q, size_t lo, size_t hi, int bs, int ra)
{
    const unsigned char *s = p;
    if (lo >= hi) return 0;
    if (ra) {
        for (size_t i = lo; i < hi; i += (size_t)bs) {
            size_t cur = (i + bs > hi) ? hi - i : (size_t)bs;
            memcpy(&q[i], p, 4);
            p += 4;
            p += get_run(p, q + i + 1, cur - 1, 1);
        }
    } else {
        memcpy(&q[lo], p, 4);
        p += 4;
        for (size_t i = lo + 1; i < hi; i += (size_t)bs) {
            size_t cur = (i + bs > hi) ? hi - i : (size_t)bs;
            p += get_run(p, q + i, cur, 0);
        }
    }
    return (size_t)(p - s);
}

size_t or_elem_bound(size_t n, int bs, unsigned T, int dsize, int ra)
{
    size_t nb = (n + bs - 1) / bs + T;
    return (size_t)dsize + 8 * (size_t)T + 4 * (size_t)T + nb * (4 + 1 + ((size_t)bs + 7) / 8 + 4 * (size_t)bs) +
           (ra ? 0 : 4 * n);
}

/* Writers: [eb big-endian (4 or 8 bytes)][u64 off[T]][chunks].  Float: the reciprocal
 * is an int / float division, widened (SZp/src/szp_float.cc:1489); double: 1 / eb
 * (SZp/src/szp_double.cc:188, :787).  q = C cast of (double)x * inv. */
size_t or_elem_compress(const void *data, int dsize, size_t n, double eb, int bs, unsigned T, int ra,
                        unsigned char *out)
{
    if (n == 0 || bs < 1 || T == 0 || (dsize != 4 && dsize != 8) || !(eb > 0)) return 0;
    double inv;
    if (dsize == 4) {
        float ebf = (float)eb;
        float finv = 1 / ebf;
        inv = (double)finv;
        put_be(out, &ebf, 4);
    } else {
        inv = 1 / eb;
        put_be(out, &eb, 8);
    }
    int *q = (int *)malloc(n * sizeof(int));
    for (size_t i = 0; i < n; i++) {
        double x = dsize == 4 ? (double)((const float *)data)[i] : ((const double *)data)[i];
        q[i] = or_cvt_i32(x * inv);
    }
    unsigned char *hdr = out + dsize, *payload = hdr + 8 * (size_t)T;
    uint64_t pos = 0;
    for (unsigned t = 0; t < T; t++) {
        size_t lo, hi;
        elem_chunk(n, T, t, &lo, &hi);
        memcpy(hdr + 8 * (size_t)t, &pos, 8);
        pos += put_elem_chunk(payload + pos, q, lo, hi, bs, ra);
    }
    free(q);
    return (size_t)dsize + 8 * (size_t)T + pos;
}

/* Decoders of the writers above (SZp/src/szpd_float.cc:28-135, szpd_double.cc:6-110):
 * `cmp` starts at the offset table, the chunk count is given.  Float values are
 * (float)q * eb in float, double values (double)q * eb. */
int or_elem_decompress(const unsigned char *cmp, size_t n, double eb, int bs, unsigned T, int ra, int dsize,
                       void *out)
{
    if (n == 0 || bs < 1 || T == 0 || (dsize != 4 && dsize != 8)) return -1;
    int *q = (int *)malloc(n * sizeof(int));
    const unsigned char *rcp = cmp + 8 * (size_t)T;
    for (unsigned t = 0; t < T; t++) {
        size_t lo, hi;
        elem_chunk(n, T, t, &lo, &hi);
        get_elem_chunk(rcp + rd64(cmp + 8 * (size_t)t), q, lo, hi, bs, ra);
    }
    if (dsize == 4) {
        float ebf = (float)eb;
        for (size_t i = 0; i < n; i++) ((float *)out)[i] = (float)q[i] * ebf;
    } else {
        for (size_t i = 0; i < n; i++) ((double *)out)[i] = (double)q[i] * eb;
    }
    free(q);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* double-precision 3D topology stream (SZp/src/szp_cuda_topology_3d.cu)      */
/* ------------------------------------------------------------------------ */

/* One Jacobi pass of the saddle fix (:736-826): every interior point typed 3 whose
 * quantized 6-neighbourhood (read from the pass's snapshot) is no longer a saddle
 * takes the first of c - 1, c + 1 that is high along one axis and low along another
 * without reaching a neighbouring maximum or minimum. Returns the points changed. */
size_t or_fix_saddles_3d_pass(const int *src, int *dst, const unsigned char *types, int d1, int d2, int d3)
{
    size_t changed = 0;
    const size_t pl = (size_t)d2 * d3;
    for (int x = 1; x + 1 < d1; x++)
        for (int y = 1; y + 1 < d2; y++)
            for (int z = 1; z + 1 < d3; z++) {
                size_t f = (size_t)x * pl + (size_t)y * d3 + z;
                if (types[f] != 3) continue;
                const size_t o[6] = {f - pl, f + pl, f - d3, f + d3, f - 1, f + 1};
                int c = src[f], v[6];
                for (int k = 0; k < 6; k++) v[k] = src[o[k]];
                int hi = 0, lo = 0;
                for (int a = 0; a < 3; a++) {
                    hi += c > v[2 * a] && c > v[2 * a + 1];
                    lo += c < v[2 * a] && c < v[2 * a + 1];
                }
                if (hi >= 1 && lo >= 1) continue;
                const int trials[2] = {(int)((unsigned)c - 1u), (int)((unsigned)c + 1u)};
                for (int k = 0; k < 2; k++) {
                    int t = trials[k], th = 0, tl = 0, ok = 1;
                    for (int a = 0; a < 3; a++) {
                        int mx = v[2 * a] > v[2 * a + 1] ? v[2 * a] : v[2 * a + 1];
                        int mn = v[2 * a] < v[2 * a + 1] ? v[2 * a] : v[2 * a + 1];
                        th += t > mx;
                        tl += t < mn;
                    }
                    if (!(th >= 1 && tl >= 1)) continue;
                    for (int j = 0; j < 6; j++) {
                        if (types[o[j]] == 1 && v[j] <= t) ok = 0;
                        if (types[o[j]] == 2 && v[j] >= t) ok = 0;
                    }
                    if (ok) {
                        dst[f] = t;
                        changed++;
                        break;
                    }
                }
            }
    return changed;
}

/* szp_cuda_double_compress_topology_3d (:1065-1188): q = C cast of x * (1 / eb) in
 * double, three saddle-fix passes, then the topology writer with one chunk:
 * [u64 0][blocks]. `types` is the type map marked from the critical-point list.
 * q_out (nullable) receives the fixed quantized field. */
size_t or_topo3d_f64_compress(const double *data, int d1, int d2, int d3, double eb, int bs,
                              const unsigned char *types, unsigned char *out, int *q_out)
{
    size_t n = (size_t)d1 * d2 * d3;
    if (n == 0 || !(eb > 0) || bs < 1 || bs > 1025) return 0;
    double inv = 1.0 / eb;
    int *q = (int *)malloc(n * sizeof(int)), *snap = (int *)malloc(n * sizeof(int));
    for (size_t i = 0; i < n; i++) q[i] = or_cvt_i32(data[i] * inv);
    for (int pass = 0; pass < 3; pass++) {
        memcpy(snap, q, n * sizeof(int));
        or_fix_saddles_3d_pass(snap, q, types, d1, d2, d3);
    }
    size_t r = write_chunks(out, q, n, bs, types, 1, 0);
    if (q_out) memcpy(q_out, q, n * sizeof(int));
    free(q);
    free(snap);
    return r;
}
