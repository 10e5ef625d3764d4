/*
 * artea_oracle.c -- CPU restatement of the ARTEA search and refine hot path.
 *
 * TEST INFRASTRUCTURE ONLY (parity oracle + timed CPU baseline).  Nothing in
 * the product path links, loads or calls this file.  Built with
 * -ffp-contract=off so every FP32 operation is rounded exactly where the C
 * source says; the distance kernels use fmaf() where the reference uses
 * _mm512_fmadd_ps.  Citations are relative to include/artea/cpu/ of the
 * reference checkout.
 */
#define _GNU_SOURCE
#include "artea_oracle.h"

#include <float.h>
#include <math.h>
#include <omp.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static void set_err(char* err, size_t errlen, const char* fmt, ...) {
    if (!err || errlen == 0) return;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(err, errlen, fmt, ap);
    va_end(ap);
}

static int pick_threads(int nthreads) {
    int t = nthreads > 0 ? nthreads : omp_get_max_threads();
    return t > 256 ? 256 : t; /* the refine log buffers are per thread, 256 at most */
}

/* ======================================================================
 * Distance: SIMDDistance<..., 1> (utils/simd_distance.hpp:64-79).
 * Lane l of a 16-float chunk accumulates acc[l] = fma(x, y, acc[l]) over
 * chunks 0..D/16-1 in order (:108-115); the 16 lanes are then summed by GCC's
 * _mm512_reduce_add_ps tree (avx512fintrin.h __MM512_REDUCE_OP): upper 8 +
 * lower 8, upper 4 + lower 4, then (t[0]+t[2]) + (t[1]+t[3]).
 * ====================================================================== */
static inline float reduce16(const float* a) {
    float t8[8], t4[4];
    for (int i = 0; i < 8; ++i) t8[i] = a[i + 8] + a[i];
    for (int i = 0; i < 4; ++i) t4[i] = t8[i + 4] + t8[i];
    const float x0 = t4[0] + t4[2];
    const float x1 = t4[1] + t4[3];
    return x0 + x1;
}

static inline float l2sqr_f32(const float* a, const float* b, uint32_t D) {
    float acc[16] = {0};
    for (uint32_t c = 0; c < D / 16; ++c) {
        for (int l = 0; l < 16; ++l) {
            const float d = a[c * 16 + l] - b[c * 16 + l];
            acc[l] = fmaf(d, d, acc[l]);
        }
    }
    return reduce16(acc);
}

static inline float dot_f32(const float* a, const float* b, uint32_t D) {
    float acc[16] = {0};
    for (uint32_t c = 0; c < D / 16; ++c)
        for (int l = 0; l < 16; ++l) acc[l] = fmaf(a[c * 16 + l], b[c * 16 + l], acc[l]);
    return reduce16(acc);
}

/* _impl_cosine (simd_distance.hpp:241-275): three chains, one reduce each,
 * denom = sqrtf(sq1 * sq2), 1 - dot/denom, or 1 when denom is not > 0. */
static inline float cosine_f32(const float* a, const float* b, uint32_t D) {
    float cr[16] = {0}, n1[16] = {0}, n2[16] = {0};
    for (uint32_t c = 0; c < D / 16; ++c) {
        for (int l = 0; l < 16; ++l) {
            const float x = a[c * 16 + l], y = b[c * 16 + l];
            cr[l] = fmaf(x, y, cr[l]);
            n1[l] = fmaf(x, x, n1[l]);
            n2[l] = fmaf(y, y, n2[l]);
        }
    }
    const float dot = reduce16(cr);
    const float sq1 = reduce16(n1);
    const float sq2 = reduce16(n2);
    const float denom = sqrtf(sq1 * sq2);
    return denom > 0.0f ? 1.0f - dot / denom : 1.0f;
}

float orc_distance(int metric, const float* a, const float* b, uint32_t D) {
    switch (metric) {
    case ORC_L2SQR: return l2sqr_f32(a, b, D);
    case ORC_L2: return sqrtf(l2sqr_f32(a, b, D));
    case ORC_DOT: return -dot_f32(a, b, D);
    case ORC_COSINE: return cosine_f32(a, b, D);
    default: return NAN;
    }
}

double orc_distance_f64(int metric, const float* a, const float* b, uint32_t D) {
    double s = 0, na = 0, nb = 0, dt = 0;
    for (uint32_t i = 0; i < D; ++i) {
        const double x = a[i], y = b[i];
        s += (x - y) * (x - y);
        dt += x * y;
        na += x * x;
        nb += y * y;
    }
    switch (metric) {
    case ORC_L2SQR: return s;
    case ORC_L2: return sqrt(s);
    case ORC_DOT: return -dt;
    case ORC_COSINE: {
        const double den = sqrt(na * nb);
        return den > 0 ? 1.0 - dt / den : 1.0;
    }
    default: return NAN;
    }
}

void orc_distance_matrix(int metric, const float* a, uint32_t na, const float* b, uint32_t nb,
                         uint32_t D, float* out, int nthreads) {
#pragma omp parallel for schedule(static) num_threads(pick_threads(nthreads))
    for (int64_t i = 0; i < (int64_t)na; ++i)
        for (uint32_t j = 0; j < nb; ++j)
            out[(uint64_t)i * nb + j] = orc_distance(metric, a + (uint64_t)i * D, b + (uint64_t)j * D, D);
}

/* ======================================================================
 * Compact hierarchical graph (compact_structure/hierarchical_graph.hpp).
 * Group h holds the vertices whose final highest level is h; each has h+1
 * rows ordered h..0; row index = local*(h+1) + (h-level) (:99-108).
 * ====================================================================== */
struct orc_graph {
    uint32_t n, top, ul_max, bl_max, entry;
    uint32_t* highest; /* n */
    uint32_t* local;   /* n */
    uint64_t* count;   /* top+1 */
    uint64_t** offs;   /* top+1 arrays of count[h]*(h+1)+1 */
    uint32_t** bucket; /* top+1 arrays of count[h] */
    uint32_t* nbrs;
    uint64_t n_nbrs;
};

static orc_graph* graph_alloc(uint32_t n, uint32_t top, uint32_t ul, uint32_t bl, const uint64_t* counts) {
    orc_graph* g = (orc_graph*)calloc(1, sizeof(orc_graph));
    g->n = n;
    g->top = top;
    g->ul_max = ul;
    g->bl_max = bl;
    g->entry = 0xFFFFFFFFu;
    g->highest = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    g->local = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    for (uint32_t v = 0; v < n; ++v) {
        g->highest[v] = ORC_INVALID_LEVEL;
        g->local[v] = 0xFFFFFFFFu;
    }
    g->count = (uint64_t*)calloc(top + 1, sizeof(uint64_t));
    g->offs = (uint64_t**)calloc(top + 1, sizeof(uint64_t*));
    g->bucket = (uint32_t**)calloc(top + 1, sizeof(uint32_t*));
    for (uint32_t h = 0; h <= top; ++h) {
        g->count[h] = counts ? counts[h] : 0;
        g->offs[h] = (uint64_t*)calloc(g->count[h] * (h + 1) + 1, sizeof(uint64_t));
        g->bucket[h] = (uint32_t*)malloc(sizeof(uint32_t) * (g->count[h] ? g->count[h] : 1));
    }
    return g;
}

void orc_graph_free(orc_graph* g) {
    if (!g) return;
    for (uint32_t h = 0; h <= g->top; ++h) {
        free(g->offs[h]);
        free(g->bucket[h]);
    }
    free(g->offs);
    free(g->bucket);
    free(g->count);
    free(g->highest);
    free(g->local);
    free(g->nbrs);
    free(g);
}

/* allocate_neighbors_from_row_counts (compact hierarchical_graph.hpp:143-158):
 * row-start entries hold counts; convert to a prefix sum across groups. */
static void graph_finalize_offsets(orc_graph* g) {
    uint64_t total = 0;
    for (uint32_t h = 0; h <= g->top; ++h) {
        uint64_t rows = g->count[h] * (h + 1);
        for (uint64_t r = 0; r < rows; ++r) {
            uint64_t len = g->offs[h][r];
            g->offs[h][r] = total;
            total += len;
        }
        g->offs[h][rows] = total;
    }
    g->nbrs = (uint32_t*)malloc(sizeof(uint32_t) * (total ? total : 1));
    g->n_nbrs = total;
}

static inline uint64_t row_index(uint32_t local, uint32_t h, uint32_t level) {
    return (uint64_t)local * (h + 1) + (h - level);
}

uint32_t orc_graph_num_vertices(const orc_graph* g) { return g->n; }
uint32_t orc_graph_top(const orc_graph* g) { return g->n == 0 ? ORC_INVALID_LEVEL : g->top; }
uint32_t orc_graph_entry(const orc_graph* g) { return g->entry; }
uint32_t orc_graph_ul_max(const orc_graph* g) { return g->ul_max; }
uint32_t orc_graph_bl_max(const orc_graph* g) { return g->bl_max; }
uint32_t orc_graph_highest(const orc_graph* g, uint32_t v) { return g->highest[v]; }
uint32_t orc_graph_local(const orc_graph* g, uint32_t v) { return g->local[v]; }
uint64_t orc_graph_group_count(const orc_graph* g, uint32_t h) { return h <= g->top ? g->count[h] : 0; }
const uint64_t* orc_graph_group_offsets(const orc_graph* g, uint32_t h) { return g->offs[h]; }
const uint32_t* orc_graph_nbrs(const orc_graph* g, uint64_t* n) {
    *n = g->n_nbrs;
    return g->nbrs;
}

const uint32_t* orc_graph_level_nbrs(const orc_graph* g, uint32_t vid, uint32_t level, uint32_t* count) {
    const uint32_t h = g->highest[vid];
    if (h == ORC_INVALID_LEVEL || level > h) { /* precondition violated: empty span */
        *count = 0;
        return g->nbrs;
    }
    const uint64_t r = row_index(g->local[vid], h, level);
    *count = (uint32_t)(g->offs[h][r + 1] - g->offs[h][r]);
    return g->nbrs + g->offs[h][r];
}

orc_graph* orc_graph_from_csr(uint32_t n, uint32_t top, uint32_t ul, uint32_t bl, uint32_t entry,
                              const uint32_t* vi, const uint64_t* const* go, const uint32_t* nbrs, uint64_t n_nbrs,
                              char* err, size_t errlen) {
    uint64_t* counts = (uint64_t*)calloc(top + 1, sizeof(uint64_t));
    for (uint32_t v = 0; v < n; ++v) {
        const uint32_t h = vi[2 * v];
        if (h == ORC_INVALID_LEVEL) continue;
        if (h > top) {
            set_err(err, errlen, "vid %u highest %u > top %u", v, h, top);
            free(counts);
            return NULL;
        }
        counts[h]++;
    }
    orc_graph* g = graph_alloc(n, top, ul, bl, counts);
    for (uint32_t v = 0; v < n; ++v) {
        const uint32_t h = vi[2 * v], loc = vi[2 * v + 1];
        if (h == ORC_INVALID_LEVEL) continue;
        if (loc >= counts[h]) {
            set_err(err, errlen, "vid %u local id %u out of range", v, loc);
            orc_graph_free(g);
            free(counts);
            return NULL;
        }
        g->highest[v] = h;
        g->local[v] = loc;
        g->bucket[h][loc] = v;
    }
    for (uint32_t h = 0; h <= top; ++h) {
        const uint64_t rows = counts[h] * (h + 1);
        for (uint64_t r = 0; r <= rows; ++r) {
            g->offs[h][r] = go[h][r];
            if (go[h][r] > n_nbrs || (r > 0 && go[h][r] < go[h][r - 1])) {
                set_err(err, errlen, "group %u offsets not monotone within the neighbour array", h);
                orc_graph_free(g);
                free(counts);
                return NULL;
            }
        }
    }
    g->nbrs = (uint32_t*)malloc(sizeof(uint32_t) * (n_nbrs ? n_nbrs : 1));
    memcpy(g->nbrs, nbrs, sizeof(uint32_t) * n_nbrs);
    g->n_nbrs = n_nbrs;
    g->entry = entry;
    free(counts);
    return g;
}

/* _compute_entry_point_vid (compactor/hierarchical_graph_compactor.hpp:261-327):
 * FP64 coordinate sums over the top bucket, times 1/N, cast to FP32; argmin
 * of dist(centroid, v) over the bucket.  Ties keep the earliest bucket
 * position (the sequential outcome; the reference's TBB combine order is
 * unspecified). */
static uint32_t compute_entry(const uint32_t* bucket, uint64_t N, const float* vecs, uint32_t D, int metric) {
    if (N == 0) return 0xFFFFFFFFu;
    double* sums = (double*)calloc(D, sizeof(double));
    for (uint64_t i = 0; i < N; ++i) {
        const float* v = vecs + (uint64_t)bucket[i] * D;
        for (uint32_t d = 0; d < D; ++d) sums[d] += (double)v[d];
    }
    const double inv = 1.0 / (double)N;
    float* c = (float*)malloc(sizeof(float) * D);
    for (uint32_t d = 0; d < D; ++d) c[d] = (float)(sums[d] * inv);
    float best = FLT_MAX;
    uint32_t best_vid = 0xFFFFFFFFu;
    for (uint64_t i = 0; i < N; ++i) {
        const float dd = orc_distance(metric, c, vecs + (uint64_t)bucket[i] * D, D);
        if (dd < best) {
            best = dd;
            best_vid = bucket[i];
        }
    }
    free(sums);
    free(c);
    return best_vid;
}

/* HierarchicalGraphCompactor::compact_graph (:118-252). Source bucket order
 * for this API is ascending vid. */
orc_graph* orc_compact_graph(uint32_t n, uint32_t max_level, uint32_t ul_max, uint32_t bl_max,
                             const uint32_t* highest, const uint64_t* const* row_off,
                             const uint32_t* const* row_ids, const float* vecs, uint32_t D, int metric,
                             char* err, size_t errlen) {
    return orc_compact_graph_ordered(n, max_level, ul_max, bl_max, highest, row_off, row_ids, NULL, vecs, D, metric,
                                     err, errlen);
}

/* The dynamic graph's buckets list vertices in the order they were assigned
 * (assign_layer appends, dynamic_structure/hierarchical_graph.hpp); `order`
 * gives that order (NULL = ascending vid) and the compactor keeps it within
 * every bucket (compactor/hierarchical_graph_compactor.hpp:146-179). */
orc_graph* orc_compact_graph_ordered(uint32_t n, uint32_t max_level, uint32_t ul_max, uint32_t bl_max,
                                     const uint32_t* highest, const uint64_t* const* row_off,
                                     const uint32_t* const* row_ids, const uint32_t* order, const float* vecs,
                                     uint32_t D, int metric, char* err, size_t errlen) {
    uint32_t src_top = ORC_INVALID_LEVEL;
    for (uint32_t v = 0; v < n; ++v) {
        if (highest[v] == ORC_INVALID_LEVEL) continue;
        if (highest[v] > max_level) {
            set_err(err, errlen, "vid %u highest level %u > max level %u", v, highest[v], max_level);
            return NULL;
        }
        if (src_top == ORC_INVALID_LEVEL || highest[v] > src_top) src_top = highest[v];
    }
    if (n == 0 || src_top == ORC_INVALID_LEVEL) {
        uint64_t zero = 0;
        orc_graph* g = graph_alloc(n, 0, ul_max, bl_max, &zero);
        graph_finalize_offsets(g);
        return g;
    }
    uint64_t* src_count = (uint64_t*)calloc(src_top + 1, sizeof(uint64_t));
    for (uint32_t v = 0; v < n; ++v)
        if (highest[v] != ORC_INVALID_LEVEL) src_count[highest[v]]++;
    /* Step 1: the first bucket from the top with >= min_layer_cap (32,
     * framework/type_traits/index_traits.hpp:79) vertices is the new top. */
    uint32_t new_top = 0;
    for (uint32_t h = src_top;; --h) {
        if (src_count[h] >= 32) {
            new_top = h;
            break;
        }
        if (h == 0) break;
    }
    uint64_t* counts = (uint64_t*)calloc(new_top + 1, sizeof(uint64_t));
    for (uint32_t h = 0; h <= src_top; ++h) counts[h < new_top ? h : new_top] += src_count[h];
    orc_graph* g = graph_alloc(n, new_top, ul_max, bl_max, counts);
    /* Step 2: groups below new_top keep source order; new_top gets its own
     * bucket then demoted buckets in ascending source level. */
    uint64_t* fill = (uint64_t*)calloc(new_top + 1, sizeof(uint64_t));
    for (uint32_t src_h = 0; src_h <= src_top; ++src_h) {
        const uint32_t grp = src_h < new_top ? src_h : new_top;
        for (uint32_t i = 0; i < n; ++i) {
            const uint32_t v = order ? order[i] : i;
            if (highest[v] != src_h) continue;
            g->bucket[grp][fill[grp]] = v;
            g->highest[v] = grp;
            g->local[v] = (uint32_t)fill[grp];
            fill[grp]++;
        }
    }
    /* Row counts, then offsets, then ids in row order (trimmed levels dropped). */
    for (uint32_t h = 0; h <= new_top; ++h) {
        for (uint64_t loc = 0; loc < g->count[h]; ++loc) {
            const uint32_t v = g->bucket[h][loc];
            for (uint32_t l = 0; l <= h; ++l)
                g->offs[h][row_index((uint32_t)loc, h, l)] = row_off[l][v + 1] - row_off[l][v];
        }
    }
    graph_finalize_offsets(g);
    for (uint32_t v = 0; v < n; ++v) {
        const uint32_t h = g->highest[v];
        if (h == ORC_INVALID_LEVEL) continue;
        for (uint32_t l = 0; l <= h; ++l) {
            const uint64_t r = row_index(g->local[v], h, l);
            memcpy(g->nbrs + g->offs[h][r], row_ids[l] + row_off[l][v],
                   sizeof(uint32_t) * (g->offs[h][r + 1] - g->offs[h][r]));
        }
    }
    for (uint64_t i = 0; i < g->n_nbrs; ++i) {
        if (g->nbrs[i] >= n) {
            set_err(err, errlen, "neighbour id %u out of range", g->nbrs[i]);
            orc_graph_free(g);
            g = NULL;
            break;
        }
    }
    if (g) g->entry = compute_entry(g->bucket[new_top], g->count[new_top], vecs, D, metric);
    free(src_count);
    free(counts);
    free(fill);
    return g;
}

/* ======================================================================
 * HGRA v1 (persistence/hierarchical_graph_file_manager.hpp:20-40).
 * ====================================================================== */
#define HGRA_MAGIC 0x48475241u
#define HGRA_VERSION 1u

int orc_graph_snapshot(const orc_graph* g, const char* path, char* err, size_t errlen) {
    FILE* f = fopen(path, "wb");
    if (!f) {
        set_err(err, errlen, "snapshot: failed to open %s", path);
        return -1;
    }
    const uint32_t hdr[7] = {HGRA_MAGIC, HGRA_VERSION, g->n, g->n == 0 ? 0u : g->top, g->ul_max,
                             g->bl_max, g->entry};
    fwrite(hdr, 4, 7, f);
    for (uint32_t v = 0; v < g->n; ++v) {
        const uint32_t h = g->highest[v];
        fwrite(&h, 4, 1, f);
        if (h == ORC_INVALID_LEVEL) continue;
        for (uint32_t l = h;; --l) {
            uint32_t cnt;
            const uint32_t* ids = orc_graph_level_nbrs(g, v, l, &cnt);
            fwrite(&cnt, 4, 1, f);
            if (cnt) fwrite(ids, 4, cnt, f);
            if (l == 0) break;
        }
    }
    int bad = ferror(f);
    fclose(f);
    if (bad) {
        set_err(err, errlen, "snapshot: write failure on %s", path);
        return -1;
    }
    return 0;
}

static int rd32(FILE* f, uint32_t* x) { return fread(x, 4, 1, f) == 1; }

/* restore (:152-324): validates magic, version, header, highest <= top,
 * count <= level capacity, neighbour ids < N; buckets in ascending vid. */
orc_graph* orc_graph_restore(const char* path, char* err, size_t errlen) {
    FILE* f = fopen(path, "rb");
    if (!f) {
        set_err(err, errlen, "restore: failed to open %s", path);
        return NULL;
    }
    uint32_t hdr[7];
    if (fread(hdr, 4, 7, f) != 7) {
        set_err(err, errlen, "restore: header read failed: %s", path);
        fclose(f);
        return NULL;
    }
    if (hdr[0] != HGRA_MAGIC) {
        set_err(err, errlen, "restore: bad magic %#x in %s", hdr[0], path);
        fclose(f);
        return NULL;
    }
    if (hdr[1] != HGRA_VERSION) {
        set_err(err, errlen, "restore: unsupported version %u in %s", hdr[1], path);
        fclose(f);
        return NULL;
    }
    const uint32_t n = hdr[2], top = hdr[3], ul = hdr[4], bl = hdr[5], entry = hdr[6];
    if (top == ORC_INVALID_LEVEL || (entry != 0xFFFFFFFFu && entry >= n)) {
        set_err(err, errlen, "restore: invalid header in %s", path);
        fclose(f);
        return NULL;
    }
    if (n == 0) {
        uint64_t* z = (uint64_t*)calloc(top + 1, sizeof(uint64_t));
        orc_graph* g = graph_alloc(0, top, ul, bl, z);
        free(z);
        graph_finalize_offsets(g);
        g->entry = entry;
        fclose(f);
        return g;
    }
    uint32_t* hl = (uint32_t*)malloc(sizeof(uint32_t) * n);
    uint64_t* cnt_total = (uint64_t*)calloc(n + 1, sizeof(uint64_t)); /* ids per vid */
    /* First pass: payload into one growable array (per vid levels h..0). */
    uint64_t cap_ids = 1u << 20, n_ids = 0, cap_lens = 1u << 16, n_lens = 0;
    uint32_t* ids = (uint32_t*)malloc(sizeof(uint32_t) * cap_ids);
    uint32_t* lens = (uint32_t*)malloc(sizeof(uint32_t) * cap_lens);
    uint64_t* hist = (uint64_t*)calloc(top + 1, sizeof(uint64_t));
    orc_graph* g = NULL;
    for (uint32_t v = 0; v < n; ++v) {
        if (!rd32(f, &hl[v])) {
            set_err(err, errlen, "restore: truncated at vid %u in %s", v, path);
            goto fail;
        }
        if (hl[v] == ORC_INVALID_LEVEL) continue;
        if (hl[v] > top) {
            set_err(err, errlen, "restore: vid %u has highest_level_id %u > top_level_id %u in %s", v,
                    hl[v], top, path);
            goto fail;
        }
        hist[hl[v]]++;
        for (uint32_t l = hl[v];; --l) {
            uint32_t c;
            if (!rd32(f, &c)) {
                set_err(err, errlen, "restore: truncated at vid %u level %u in %s", v, l, path);
                goto fail;
            }
            const uint32_t capl = l == 0 ? bl : ul;
            if (c > capl) {
                set_err(err, errlen, "restore: vid %u level %u count %u > capacity %u in %s", v, l, c, capl,
                        path);
                goto fail;
            }
            if (n_lens == cap_lens) lens = (uint32_t*)realloc(lens, sizeof(uint32_t) * (cap_lens *= 2));
            lens[n_lens++] = c;
            while (n_ids + c > cap_ids) ids = (uint32_t*)realloc(ids, sizeof(uint32_t) * (cap_ids *= 2));
            if (c && fread(ids + n_ids, 4, c, f) != c) {
                set_err(err, errlen, "restore: truncated reading nbrs of vid %u level %u in %s", v, l, path);
                goto fail;
            }
            for (uint32_t i = 0; i < c; ++i) {
                if (ids[n_ids + i] >= n) {
                    set_err(err, errlen, "restore: invalid neighbor %u in %s", ids[n_ids + i], path);
                    goto fail;
                }
            }
            n_ids += c;
            cnt_total[v] += c;
            if (l == 0) break;
        }
    }
    g = graph_alloc(n, top, ul, bl, hist);
    {
        uint64_t* fill = (uint64_t*)calloc(top + 1, sizeof(uint64_t));
        uint64_t li = 0;
        for (uint32_t v = 0; v < n; ++v) {
            const uint32_t h = hl[v];
            if (h == ORC_INVALID_LEVEL) continue;
            const uint64_t loc = fill[h]++;
            g->highest[v] = h;
            g->local[v] = (uint32_t)loc;
            g->bucket[h][loc] = v;
            for (uint32_t l = h;; --l) {
                g->offs[h][row_index((uint32_t)loc, h, l)] = lens[li++];
                if (l == 0) break;
            }
        }
        free(fill);
        graph_finalize_offsets(g);
        uint64_t pos = 0;
        for (uint32_t v = 0; v < n; ++v) {
            const uint32_t h = hl[v];
            if (h == ORC_INVALID_LEVEL) continue;
            for (uint32_t l = h;; --l) {
                uint32_t c;
                orc_graph_level_nbrs(g, v, l, &c);
                const uint64_t r = row_index(g->local[v], h, l);
                memcpy(g->nbrs + g->offs[h][r], ids + pos, sizeof(uint32_t) * c);
                pos += c;
                if (l == 0) break;
            }
        }
        g->entry = entry;
    }
    free(hl);
    free(cnt_total);
    free(ids);
    free(lens);
    free(hist);
    fclose(f);
    return g;
fail:
    free(hl);
    free(cnt_total);
    free(ids);
    free(lens);
    free(hist);
    fclose(f);
    return NULL;
}

/* ======================================================================
 * StdCandidateQueue (router/data_structures/std_candidate_queue.hpp).
 * U = unexplored min-heap, T = top max-heap, both ordered by (dist, vid)
 * (candidate_entry.hpp:170-180).  Evicted entries stay in U.
 * ====================================================================== */
typedef struct {
    uint32_t vid;
    float d;
} centry;

static inline int ce_lt(centry a, centry b) { return a.d < b.d || (a.d == b.d && a.vid < b.vid); }

typedef struct {
    centry* a;
    uint32_t n, cap;
    int is_max; /* 1 = max-heap (T), 0 = min-heap (U) */
} heap_t;

static inline int heap_before(const heap_t* h, centry x, centry y) { /* x closer to top than y */
    return h->is_max ? ce_lt(y, x) : ce_lt(x, y);
}
static void heap_push(heap_t* h, centry e) {
    if (h->n == h->cap) {
        h->cap = h->cap ? h->cap * 2 : 64;
        h->a = (centry*)realloc(h->a, sizeof(centry) * h->cap);
    }
    uint32_t i = h->n++;
    h->a[i] = e;
    while (i > 0) {
        uint32_t p = (i - 1) / 2;
        if (!heap_before(h, h->a[i], h->a[p])) break;
        centry t = h->a[i];
        h->a[i] = h->a[p];
        h->a[p] = t;
        i = p;
    }
}
static centry heap_pop(heap_t* h) {
    centry top = h->a[0];
    h->a[0] = h->a[--h->n];
    uint32_t i = 0;
    for (;;) {
        uint32_t l = 2 * i + 1, r = l + 1, b = i;
        if (l < h->n && heap_before(h, h->a[l], h->a[b])) b = l;
        if (r < h->n && heap_before(h, h->a[r], h->a[b])) b = r;
        if (b == i) break;
        centry t = h->a[i];
        h->a[i] = h->a[b];
        h->a[b] = t;
        i = b;
    }
    return top;
}

struct orc_queue {
    heap_t U, T;
    uint32_t capacity;
    float lb;
};

static void queue_init(orc_queue* q, uint32_t capacity) {
    memset(q, 0, sizeof(*q));
    q->U.is_max = 0;
    q->T.is_max = 1;
    q->capacity = capacity;
    q->lb = FLT_MAX;
}
static void queue_release(orc_queue* q) {
    free(q->U.a);
    free(q->T.a);
}
static void queue_clear(orc_queue* q) {
    q->U.n = q->T.n = 0;
    q->lb = FLT_MAX;
}

orc_queue* orc_queue_create(uint32_t capacity) {
    orc_queue* q = (orc_queue*)malloc(sizeof(orc_queue));
    queue_init(q, capacity);
    return q;
}
void orc_queue_free(orc_queue* q) {
    if (!q) return;
    queue_release(q);
    free(q);
}

/* try_push (:358-388): reject iff |T| >= L && d >= lb (distance only). */
static inline int queue_try_push(orc_queue* q, uint32_t vid, float d) {
    if (q->T.n >= q->capacity && d >= q->lb) return 0;
    centry e = {vid & ORC_INVALID_VID, d};
    heap_push(&q->U, e);
    heap_push(&q->T, e);
    if (q->T.n > q->capacity) heap_pop(&q->T);
    q->lb = q->T.n ? q->T.a[0].d : FLT_MAX;
    return 1;
}
int orc_queue_try_push(orc_queue* q, uint32_t vid, float dist) { return queue_try_push(q, vid, dist); }

/* pop_best_unexplored_entry (:398-406) */
int orc_queue_pop_best(orc_queue* q, uint32_t* vid, float* dist) {
    if (q->U.n == 0) return 0;
    centry e = heap_pop(&q->U);
    *vid = e.vid;
    *dist = e.d;
    return 1;
}
/* should_terminate (:503-513) */
static inline int queue_should_terminate(const orc_queue* q) {
    if (q->U.n == 0) return 1;
    if (q->T.n < q->capacity) return 0;
    return q->U.a[0].d > q->lb;
}
int orc_queue_should_terminate(const orc_queue* q) { return queue_should_terminate(q); }
float orc_queue_lower_bound(const orc_queue* q) { return q->lb; }
uint32_t orc_queue_result_size(const orc_queue* q) { return q->T.n; }
uint32_t orc_queue_unexplored_size(const orc_queue* q) { return q->U.n; }
/* reset_exploration (:421-426) */
void orc_queue_reset_exploration(orc_queue* q) {
    q->U.n = 0;
    for (uint32_t i = 0; i < q->T.n; ++i) heap_push(&q->U, q->T.a[i]);
}
/* extract_results (:542-561): all of T in descending order, reversed, first k. */
uint32_t orc_queue_extract(orc_queue* q, uint32_t k, uint32_t* vids, float* dists) {
    const uint32_t n = q->T.n;
    centry* tmp = (centry*)malloc(sizeof(centry) * (n ? n : 1));
    for (uint32_t i = 0; i < n; ++i) tmp[n - 1 - i] = heap_pop(&q->T);
    const uint32_t m = k < n ? k : n;
    for (uint32_t i = 0; i < m; ++i) {
        vids[i] = tmp[i].vid;
        dists[i] = tmp[i].d;
    }
    free(tmp);
    return m;
}

/* ======================================================================
 * VersionTagTable (containers/version_tag_table.hpp:47-164).
 * ====================================================================== */
typedef struct {
    uint16_t* tags;
    uint32_t n;
    uint16_t version;
} vtt_t;

static void vtt_init(vtt_t* t, uint32_t n) {
    t->tags = (uint16_t*)calloc(n ? n : 1, sizeof(uint16_t));
    t->n = n;
    t->version = 1;
}
static void vtt_clear(vtt_t* t) {
    if (++t->version == 0) {
        memset(t->tags, 0, sizeof(uint16_t) * t->n);
        t->version = 1;
    }
}
static inline void vtt_set(vtt_t* t, uint32_t v) { t->tags[v] = t->version; }
static inline int vtt_test_and_set(vtt_t* t, uint32_t v) {
    const int prev = t->tags[v] == t->version;
    t->tags[v] = t->version;
    return prev;
}

/* ======================================================================
 * Router (router/hierarchical_graph_router.hpp, single_layer_router.hpp,
 * router/detail/beam_loop.hpp).
 * ====================================================================== */
typedef struct {
    const orc_graph* g;
    const float* base;
    uint32_t D;
    int metric;
} router_ctx;

/* greedy_loop_body (beam_loop.hpp:129-166); the caller marks the seed. */
static void greedy(const router_ctx* c, const float* q, uint32_t level, uint32_t* best_vid, float* best_d,
                   vtt_t* vis, int early, float thr, uint64_t* ndc) {
    for (;;) {
        if (early && *best_d <= thr) break;
        uint32_t next = *best_vid;
        float nd = *best_d;
        uint32_t cnt;
        const uint32_t* row = orc_graph_level_nbrs(c->g, *best_vid, level, &cnt);
        for (uint32_t i = 0; i < cnt; ++i) {
            const uint32_t u = row[i];
            if (vtt_test_and_set(vis, u)) continue;
            const float d = orc_distance(c->metric, q, c->base + (uint64_t)u * c->D, c->D);
            (*ndc)++;
            if (d < nd) {
                nd = d;
                next = u;
            }
        }
        if (next == *best_vid) break;
        *best_vid = next;
        *best_d = nd;
    }
}

/* SingleLayerRouter::beam_search (single_layer_router.hpp:145-163) +
 * beam_loop_body (beam_loop.hpp:90-119). */
static void beam(const router_ctx* c, const float* q, orc_queue* cq, vtt_t* vis, uint64_t* ndc) {
    orc_queue_reset_exploration(cq);
    if (cq->U.n == 0) return;
    for (uint32_t i = 0; i < cq->T.n; ++i) vtt_set(vis, cq->T.a[i].vid);
    while (cq->U.n != 0) {
        if (queue_should_terminate(cq)) break;
        centry cur = heap_pop(&cq->U);
        if (cur.vid == ORC_INVALID_VID) break;
        uint32_t cnt;
        const uint32_t* row = orc_graph_level_nbrs(c->g, cur.vid, 0, &cnt);
        for (uint32_t i = 0; i < cnt; ++i) {
            const uint32_t u = row[i];
            if (vtt_test_and_set(vis, u)) continue;
            const float d = orc_distance(c->metric, q, c->base + (uint64_t)u * c->D, c->D);
            (*ndc)++;
            queue_try_push(cq, u, d);
        }
    }
}

/* _query (:195-248) and query_l0_only (:258-286). */
static uint32_t run_query(const router_ctx* c, const float* q, uint32_t topk, uint32_t queue, uint32_t flags,
                          const float* thr, orc_queue* cq, vtt_t* vis, uint32_t* out_v, float* out_d,
                          uint32_t* cursors, uint64_t* ndc) {
    const orc_graph* g = c->g;
    if (g->n == 0) return 0;
    vtt_clear(vis);
    queue_clear(cq);
    cq->capacity = queue;
    const uint32_t entry = g->entry;
    const float ed = orc_distance(c->metric, q, c->base + (uint64_t)entry * c->D, c->D);
    (*ndc)++;
    if (flags & ORC_L0_ONLY) {
        queue_try_push(cq, entry, ed);
    } else {
        uint32_t cur = entry;
        float cd = ed;
        for (uint32_t level = g->top; level >= 1; --level) {
            vtt_set(vis, cur);
            greedy(c, q, level, &cur, &cd, vis, (flags & ORC_EARLY_STOP) != 0,
                   (flags & ORC_EARLY_STOP) ? thr[level] : 0.0f, ndc);
            if (cursors) cursors[level] = cur;
            vtt_clear(vis);
        }
        queue_try_push(cq, cur, cd);
    }
    beam(c, q, cq, vis, ndc);
    const uint32_t k = topk < cq->T.n ? topk : cq->T.n;
    if (k == 0) return 0;
    return orc_queue_extract(cq, k, out_v, out_d);
}

static int check_router_args(const orc_graph* g, uint32_t topk, uint32_t queue, char* err, size_t errlen) {
    if (queue < topk) {
        set_err(err, errlen, "candidate_queue_size (%u) must be >= topk (%u)", queue, topk);
        return -1;
    }
    if (g->n != 0 && (g->entry == 0xFFFFFFFFu || g->entry >= g->n)) {
        set_err(err, errlen, "graph has no valid entry point");
        return -1;
    }
    return 0;
}

int orc_batch_query(const orc_graph* g, const float* base, uint32_t D, int metric, const float* queries,
                    uint32_t nq, uint32_t topk, uint32_t queue, uint32_t flags, const float* thresholds,
                    uint32_t* out_vids, float* out_dists, uint64_t* ndc, int nthreads, char* err,
                    size_t errlen) {
    if (check_router_args(g, topk, queue, err, errlen)) return -1;
    router_ctx c = {g, base, D, metric};
#pragma omp parallel num_threads(pick_threads(nthreads))
    {
        vtt_t vis;
        vtt_init(&vis, g->n);
        orc_queue cq;
        queue_init(&cq, queue);
#pragma omp for schedule(dynamic, 4)
        for (int64_t i = 0; i < (int64_t)nq; ++i) {
            uint32_t* ov = out_vids + (uint64_t)i * topk;
            float* od = out_dists + (uint64_t)i * topk;
            uint64_t cnt = 0;
            uint32_t m = run_query(&c, queries + (uint64_t)i * D, topk, queue, flags, thresholds, &cq, &vis, ov,
                                   od, NULL, &cnt);
            for (uint32_t j = m; j < topk; ++j) {
                ov[j] = ORC_INVALID_VID;
                od[j] = FLT_MAX;
            }
            if (ndc) ndc[i] = cnt;
        }
        queue_release(&cq);
        free(vis.tags);
    }
    return 0;
}

int orc_query_trace(const orc_graph* g, const float* base, uint32_t D, int metric, const float* query,
                    uint32_t topk, uint32_t queue, uint32_t flags, const float* thresholds, uint32_t* out_vids,
                    float* out_dists, uint32_t* out_count, uint32_t* cursor_vids, uint64_t* ndc) {
    if (check_router_args(g, topk, queue, NULL, 0)) return -1;
    router_ctx c = {g, base, D, metric};
    vtt_t vis;
    vtt_init(&vis, g->n);
    orc_queue cq;
    queue_init(&cq, queue);
    uint64_t cnt = 0;
    *out_count = run_query(&c, query, topk, queue, flags, thresholds, &cq, &vis, out_vids, out_dists,
                           cursor_vids, &cnt);
    if (ndc) *ndc = cnt;
    queue_release(&cq);
    free(vis.tags);
    return 0;
}

/* ======================================================================
 * Refine: SingleLayerGraph rows + NbrLogTable + PropagateEngine + updaters.
 * ====================================================================== */
typedef struct {
    orc_nbr* a;
    uint32_t n, cap;
} nvec;

static inline void nvec_push(nvec* v, orc_nbr e) {
    if (v->n == v->cap) {
        v->cap = v->cap ? v->cap * 2 : 16;
        v->a = (orc_nbr*)realloc(v->a, sizeof(orc_nbr) * v->cap);
    }
    v->a[v->n++] = e;
}

typedef struct {
    uint32_t row, vid;
    float d;
} logrec;

typedef struct {
    logrec* a;
    uint64_t n, cap;
} logbuf;

static inline void logbuf_push(logbuf* b, uint32_t row, uint32_t vid, float d) {
    if (b->n == b->cap) {
        b->cap = b->cap ? b->cap * 2 : 1024;
        b->a = (logrec*)realloc(b->a, sizeof(logrec) * b->cap);
    }
    b->a[b->n].row = row;
    b->a[b->n].vid = vid;
    b->a[b->n].d = d;
    b->n++;
}

struct orc_layer {
    const float* vecs;
    uint32_t n_global, D, cap, n_local;
    int metric;
    uint32_t* l2g; /* NULL = identity */
    uint32_t* g2l; /* NULL = identity; else n_global, 0xFFFFFFFF = absent */
    nvec* rows;
    logbuf* tlogs; /* per thread */
    int n_tlogs;
    uint64_t last_logs, pair_dists;
};

static inline uint32_t storage_vid(const orc_layer* L, uint32_t local) { return L->l2g ? L->l2g[local] : local; }
static inline uint32_t local_of(const orc_layer* L, uint32_t global) { return L->g2l ? L->g2l[global] : global; }
static inline const float* vec_of(const orc_layer* L, uint32_t global) { return L->vecs + (uint64_t)global * L->D; }

orc_layer* orc_layer_create(const float* vecs, uint32_t n_global, uint32_t D, int metric, uint32_t cap,
                            const uint32_t* l2g, uint32_t n_local) {
    orc_layer* L = (orc_layer*)calloc(1, sizeof(orc_layer));
    L->vecs = vecs;
    L->n_global = n_global;
    L->D = D;
    L->metric = metric;
    L->cap = cap;
    L->n_local = l2g ? n_local : n_global;
    if (l2g) {
        L->l2g = (uint32_t*)malloc(sizeof(uint32_t) * (n_local ? n_local : 1));
        memcpy(L->l2g, l2g, sizeof(uint32_t) * n_local);
        L->g2l = (uint32_t*)malloc(sizeof(uint32_t) * (n_global ? n_global : 1));
        for (uint32_t i = 0; i < n_global; ++i) L->g2l[i] = 0xFFFFFFFFu;
        for (uint32_t i = 0; i < n_local; ++i) L->g2l[l2g[i]] = i;
    }
    L->rows = (nvec*)calloc(L->n_local ? L->n_local : 1, sizeof(nvec));
    L->n_tlogs = 256;
    L->tlogs = (logbuf*)calloc(L->n_tlogs, sizeof(logbuf));
    return L;
}

void orc_layer_free(orc_layer* L) {
    if (!L) return;
    for (uint32_t i = 0; i < L->n_local; ++i) free(L->rows[i].a);
    for (int t = 0; t < L->n_tlogs; ++t) free(L->tlogs[t].a);
    free(L->rows);
    free(L->tlogs);
    free(L->l2g);
    free(L->g2l);
    free(L);
}

uint32_t orc_layer_num_rows(const orc_layer* L) { return L->n_local; }
uint64_t orc_layer_last_logs(const orc_layer* L) { return L->last_logs; }
uint64_t orc_layer_pair_distances(const orc_layer* L) { return L->pair_dists; }

/* StrictNbrComparator (index/neighbor.hpp:247-258): (dist, vid). */
static inline int nbr_lt(orc_nbr a, orc_nbr b) {
    if (a.dist != b.dist) return a.dist < b.dist;
    return (a.vid_status & ORC_INVALID_VID) < (b.vid_status & ORC_INVALID_VID);
}

static void sort_nbrs(orc_nbr* a, uint32_t n) { /* insertion/merge sort, stable */
    if (n < 2) return;
    if (n <= 32) {
        for (uint32_t i = 1; i < n; ++i) {
            orc_nbr x = a[i];
            uint32_t j = i;
            while (j > 0 && nbr_lt(x, a[j - 1])) {
                a[j] = a[j - 1];
                --j;
            }
            a[j] = x;
        }
        return;
    }
    const uint32_t h = n / 2;
    sort_nbrs(a, h);
    sort_nbrs(a + h, n - h);
    orc_nbr* t = (orc_nbr*)malloc(sizeof(orc_nbr) * n);
    uint32_t i = 0, j = h, k = 0;
    while (i < h && j < n) t[k++] = nbr_lt(a[j], a[i]) ? a[j++] : a[i++];
    while (i < h) t[k++] = a[i++];
    while (j < n) t[k++] = a[j++];
    memcpy(a, t, sizeof(orc_nbr) * n);
    free(t);
}

void orc_layer_fill_row(orc_layer* L, uint32_t local, const orc_nbr* e, uint32_t count) {
    nvec* r = &L->rows[local];
    r->n = 0;
    for (uint32_t i = 0; i < count; ++i) {
        if ((e[i].vid_status & ORC_INVALID_VID) == ORC_INVALID_VID) break;
        nvec_push(r, e[i]);
    }
    sort_nbrs(r->a, r->n);
    if (r->n > L->cap) r->n = L->cap;
}

void orc_layer_set_row_raw(orc_layer* L, uint32_t local, const orc_nbr* e, uint32_t count) {
    nvec* r = &L->rows[local];
    r->n = 0;
    for (uint32_t i = 0; i < count; ++i) nvec_push(r, e[i]);
}

/* Every row at once: counts[i] entries of `entries` (row-major), as given. */
void orc_layer_set_rows_raw(orc_layer* L, const uint32_t* counts, const orc_nbr* entries) {
    uint64_t pos = 0;
    for (uint32_t i = 0; i < L->n_local; ++i) {
        orc_layer_set_row_raw(L, i, entries + pos, counts[i]);
        pos += counts[i];
    }
}

uint32_t orc_layer_row_size(const orc_layer* L, uint32_t local) { return L->rows[local].n; }

uint32_t orc_layer_get_row(const orc_layer* L, uint32_t local, orc_nbr* out, uint32_t max_out) {
    const nvec* r = &L->rows[local];
    const uint32_t m = r->n < max_out ? r->n : max_out;
    memcpy(out, r->a, sizeof(orc_nbr) * m);
    return r->n;
}

uint64_t orc_layer_export(const orc_layer* L, uint32_t* counts, orc_nbr* entries, uint64_t max_entries) {
    uint64_t pos = 0;
    for (uint32_t i = 0; i < L->n_local; ++i) {
        const nvec* r = &L->rows[i];
        counts[i] = r->n;
        if (entries && pos + r->n <= max_entries) memcpy(entries + pos, r->a, sizeof(orc_nbr) * r->n);
        pos += r->n;
    }
    return pos;
}

/* NbrLogTable::write_log (nbr_log_table.hpp:75-87): logs carry is_new=true. */
static inline void write_log(orc_layer* L, uint32_t row, uint32_t vid, float d) {
    logbuf_push(&L->tlogs[omp_get_thread_num()], row, vid, d);
}

static int cmp_log_in_row(const void* pa, const void* pb) {
    const logrec* a = (const logrec*)pa;
    const logrec* b = (const logrec*)pb;
    if (a->d != b->d) return a->d < b->d ? -1 : 1;
    if (a->vid != b->vid) return a->vid < b->vid ? -1 : 1;
    return 0;
}

/* PropagateEngine::merge_logs (propagate_engine.hpp:175-198) with
 * NbrLogTable::apply_logs (nbr_log_table.hpp:127-182): sort logs by (dist,
 * vid), stable merge after the row (row entries win ties), adjacent unique by
 * vid keeping the first. */
static void merge_logs(orc_layer* L, int nthreads) {
    uint64_t total = 0;
    for (int t = 0; t < L->n_tlogs; ++t) total += L->tlogs[t].n;
    L->last_logs = total;
    if (total == 0) return;
    uint64_t* off = (uint64_t*)calloc((uint64_t)L->n_local + 1, sizeof(uint64_t));
    for (int t = 0; t < L->n_tlogs; ++t)
        for (uint64_t i = 0; i < L->tlogs[t].n; ++i) off[L->tlogs[t].a[i].row + 1]++;
    for (uint32_t i = 0; i < L->n_local; ++i) off[i + 1] += off[i];
    logrec* all = (logrec*)malloc(sizeof(logrec) * total);
    uint64_t* fill = (uint64_t*)malloc(sizeof(uint64_t) * ((uint64_t)L->n_local + 1));
    memcpy(fill, off, sizeof(uint64_t) * ((uint64_t)L->n_local + 1));
    for (int t = 0; t < L->n_tlogs; ++t) {
        for (uint64_t i = 0; i < L->tlogs[t].n; ++i) all[fill[L->tlogs[t].a[i].row]++] = L->tlogs[t].a[i];
        L->tlogs[t].n = 0;
    }
    free(fill);
#pragma omp parallel for schedule(dynamic, 256) num_threads(pick_threads(nthreads))
    for (int64_t ri = 0; ri < (int64_t)L->n_local; ++ri) {
        const uint64_t b = off[ri], e = off[ri + 1];
        if (b == e) continue;
        qsort(all + b, e - b, sizeof(logrec), cmp_log_in_row);
        nvec* r = &L->rows[ri];
        const uint32_t m = (uint32_t)(e - b);
        orc_nbr* merged = (orc_nbr*)malloc(sizeof(orc_nbr) * (r->n + m));
        uint32_t i = 0, k = 0;
        uint64_t j = b;
        while (i < r->n && j < e) {
            orc_nbr lg = {all[j].vid | ORC_NBR_NEW, all[j].d};
            if (nbr_lt(lg, r->a[i])) {
                merged[k++] = lg;
                ++j;
            } else {
                merged[k++] = r->a[i++];
            }
        }
        while (i < r->n) merged[k++] = r->a[i++];
        while (j < e) {
            orc_nbr lg = {all[j].vid | ORC_NBR_NEW, all[j].d};
            merged[k++] = lg;
            ++j;
        }
        /* std::unique by vid */
        uint32_t w = 0;
        for (uint32_t x = 0; x < k; ++x) {
            if (w > 0 && (merged[w - 1].vid_status & ORC_INVALID_VID) == (merged[x].vid_status & ORC_INVALID_VID))
                continue;
            merged[w++] = merged[x];
        }
        free(r->a);
        r->a = merged;
        r->n = w;
        r->cap = k;
    }
    free(all);
    free(off);
}

/* Counter-based uniform draws shared with the GPU prefill.  SplitMix64
 * finaliser (Steele, Lea, Flood 2014) chained over (seed, stream, row, j),
 * mapped to [lo, hi) by the 64x32 multiply-high (Lemire 2019). */
static inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
uint32_t orc_rand_draw(uint64_t seed, uint64_t stream, uint32_t row, uint32_t j, uint32_t lo, uint32_t hi) {
    const uint64_t x = mix64(mix64(mix64(seed ^ mix64(stream)) ^ (uint64_t)row) + (uint64_t)j);
    const uint64_t span = (uint64_t)(hi - lo);
    return lo + (uint32_t)(((x >> 32) * span) >> 32);
}

/* RandomUpdater::update_impl (refiner/updaters/random_updater.hpp:97-142). */
void orc_next_prefill(orc_layer* L, uint32_t gen_size, uint32_t start, uint32_t end, uint32_t row_begin,
                      uint32_t row_end, uint64_t seed, uint64_t stream, int nthreads) {
    if (row_end > L->n_local) row_end = L->n_local;
#pragma omp parallel for schedule(dynamic, 256) num_threads(pick_threads(nthreads))
    for (int64_t ri = row_begin; ri < (int64_t)row_end; ++ri) {
        const nvec* r = &L->rows[ri];
        if (r->n >= gen_size) continue;
        const uint32_t need = gen_size - r->n;
        const uint32_t sv = storage_vid(L, (uint32_t)ri);
        const float* pv = vec_of(L, sv);
        for (uint32_t j = 0; j < need; ++j) {
            const uint32_t rl = orc_rand_draw(seed, stream, (uint32_t)ri, j, start, end);
            const uint32_t rg = storage_vid(L, rl);
            if (rg == sv) continue;
            const float d = orc_distance(L->metric, pv, vec_of(L, rg), L->D);
            if (r->n >= L->cap && r->a[L->cap - 1].dist <= d) continue;
            write_log(L, (uint32_t)ri, rg, d);
        }
    }
    merge_logs(L, nthreads);
}

/* TriangleUpdater::update_impl / _internal_check
 * (refiner/updaters/triangle_updater.hpp:114-157, 200-237). */
/* next_range(TriangleUpdater, row_begin, row_end)
 * (refiner/propagate_engine.hpp:217-225): the updater runs on the rows of
 * the window, the logs are merged into every row. */
void orc_next_triangle_range(orc_layer* L, float scale_coeffs, float shifted_coeffs, uint32_t row_begin,
                             uint32_t row_end, int nthreads) {
    const float inv = 1.0f / scale_coeffs;
    const float shift_term = (1.0f + inv) * shifted_coeffs;
    uint64_t pairs = 0;
    const int64_t re = row_end < L->n_local ? (int64_t)row_end : (int64_t)L->n_local;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : pairs) num_threads(pick_threads(nthreads))
    for (int64_t ri = (int64_t)row_begin; ri < re; ++ri) {
        nvec* row = &L->rows[ri];
        if (row->n == 0) continue;
        orc_nbr* kept = (orc_nbr*)malloc(sizeof(orc_nbr) * row->n);
        uint32_t nk = 0;
        kept[nk++] = row->a[0];
        for (uint32_t i = 1; i < row->n; ++i) {
            const orc_nbr c = row->a[i];
            const uint32_t cvid = c.vid_status & ORC_INVALID_VID;
            const int c_old = (c.vid_status & ORC_NBR_NEW) == 0;
            const float rec_thr = c.dist;
            const float prune_thr = c.dist * inv - shift_term;
            uint32_t rec_vid = ORC_INVALID_VID;
            float rec_d = FLT_MAX;
            int prune = 0;
            const float* cv = vec_of(L, cvid);
            for (uint32_t k = nk; k-- > 0;) {
                const orc_nbr r = kept[k];
                if (c_old && (r.vid_status & ORC_NBR_NEW) == 0) continue;
                const uint32_t rvid = r.vid_status & ORC_INVALID_VID;
                const float t = orc_distance(L->metric, cv, vec_of(L, rvid), L->D);
                pairs++;
                if (t < rec_thr && t < rec_d) {
                    rec_d = t;
                    rec_vid = rvid;
                }
                if (t < prune_thr) {
                    prune = 1;
                    break;
                }
            }
            if (rec_vid != ORC_INVALID_VID) write_log(L, local_of(L, rec_vid), cvid, rec_d);
            if (!prune) kept[nk++] = c;
        }
        for (uint32_t k = 0; k < nk; ++k) kept[k].vid_status &= ORC_INVALID_VID; /* mark_as_old */
        memcpy(row->a, kept, sizeof(orc_nbr) * nk);
        row->n = nk;
        free(kept);
    }
    L->pair_dists = pairs;
    merge_logs(L, nthreads);
}

void orc_next_triangle(orc_layer* L, float scale_coeffs, float shifted_coeffs, int nthreads) {
    orc_next_triangle_range(L, scale_coeffs, shifted_coeffs, 0, L->n_local, nthreads);
}

/* TruncateUpdater (refiner/updaters/truncate_updater.hpp:58-76) with the
 * default max_distance cutoff, i.e. resize(min(size, cap)). */
void orc_next_truncate(orc_layer* L, int nthreads) {
    (void)nthreads;
    for (uint32_t i = 0; i < L->n_local; ++i)
        if (L->rows[i].n > L->cap) L->rows[i].n = L->cap;
    merge_logs(L, nthreads);
}

/* ReverseUpdater::update_impl (refiner/updaters/reverse_updater.hpp:72-100). */
void orc_next_reverse(orc_layer* L, int nthreads) { orc_next_reverse_range(L, 0, L->n_local, nthreads); }

/* next_range(ReverseUpdater, row_begin, row_end). */
void orc_next_reverse_range(orc_layer* L, uint32_t row_begin, uint32_t row_end, int nthreads) {
    const int64_t re = row_end < L->n_local ? (int64_t)row_end : (int64_t)L->n_local;
#pragma omp parallel for schedule(dynamic, 256) num_threads(pick_threads(nthreads))
    for (int64_t ri = (int64_t)row_begin; ri < re; ++ri) {
        const nvec* row = &L->rows[ri];
        const uint32_t sv = storage_vid(L, (uint32_t)ri);
        for (uint32_t i = 0; i < row->n; ++i) {
            const uint32_t tl = local_of(L, row->a[i].vid_status & ORC_INVALID_VID);
            if (tl == 0xFFFFFFFFu) continue; /* neighbour outside this layer: no row to log into */
            write_log(L, tl, sv, row->a[i].dist);
        }
    }
    merge_logs(L, nthreads);
}

void orc_refine_schedule(orc_layer* L, uint32_t build_loops, uint32_t triu_iters, float scale_coeffs,
                         float effective_shift, int nthreads) {
    for (uint32_t b = 0; b < build_loops; ++b) {
        for (uint32_t t = 0; t < triu_iters; ++t) {
            orc_next_triangle(L, scale_coeffs, effective_shift, nthreads);
            orc_next_truncate(L, nthreads);
        }
        if (b + 1 < build_loops) orc_next_reverse(L, nthreads);
        orc_next_truncate(L, nthreads);
    }
}

/* ======================================================================
 * Routing loops: RoutingUpdater then PruningUpdater, with reverse + ARC
 * between loops (index/artea_graph/index_factory.hpp:378-403).
 * ====================================================================== */

/* SingleLayerRouter::_query_impl(seed_nbrs) (router/single_layer_router.hpp
 * :259-282): the pivot's own row seeds the queue (row order, distances from
 * the row, each seed marked visited), then beam_loop_body
 * (router/detail/beam_loop.hpp:90-119) over the layer's current rows
 * (DynamicFlatRange: a row's entries until the first invalid one), then
 * extract_results(topk).  The pivot itself is not marked visited. */
static uint32_t route_warm(const orc_layer* L, uint32_t ri, uint32_t topk, uint32_t queue, orc_queue* q, vtt_t* vis,
                           uint32_t* ov, float* od) {
    vtt_clear(vis);
    queue_clear(q);
    q->capacity = topk > queue ? topk : queue;
    const nvec* row = &L->rows[ri];
    const float* pv = vec_of(L, storage_vid(L, ri));
    for (uint32_t i = 0; i < row->n; ++i) {
        const uint32_t sv = row->a[i].vid_status & ORC_INVALID_VID;
        if (sv == ORC_INVALID_VID) break;
        if (vtt_test_and_set(vis, sv)) continue;
        queue_try_push(q, sv, row->a[i].dist);
    }
    while (q->U.n != 0) {
        if (queue_should_terminate(q)) break;
        const centry cur = heap_pop(&q->U);
        if (cur.vid == ORC_INVALID_VID) break;
        const uint32_t lr = local_of(L, cur.vid);
        if (lr == ORC_INVALID_VID || lr >= L->n_local) continue;
        const nvec* nb = &L->rows[lr];
        for (uint32_t j = 0; j < nb->n; ++j) {
            const uint32_t u = nb->a[j].vid_status & ORC_INVALID_VID;
            if (u == ORC_INVALID_VID) break;
            if (vtt_test_and_set(vis, u)) continue;
            queue_try_push(q, u, orc_distance(L->metric, pv, vec_of(L, u), L->D));
        }
    }
    const uint32_t k = topk < q->T.n ? topk : q->T.n;
    return k ? orc_queue_extract(q, k, ov, od) : 0;
}

/* RoutingUpdater::update_impl (refiner/updaters/routing_updater.hpp:84-114):
 * every result except the pivot itself is logged into the pivot's row,
 * unless the row is full and its last entry (position max_nbr_size - 1) is
 * not farther than the result; then merge_logs (no degree cap). */
void orc_next_routing(orc_layer* L, uint32_t topk, uint32_t queue, int nthreads) {
    orc_next_routing_range(L, topk, queue, 0, L->n_local, nthreads);
}

/* The same updater on the pivot rows [row_begin, row_end) only
 * (next_range, refiner/propagate_engine.hpp:217-225). */
void orc_next_routing_range(orc_layer* L, uint32_t topk, uint32_t queue, uint32_t row_begin, uint32_t row_end,
                            int nthreads) {
    const int T = pick_threads(nthreads);
    if (row_end > L->n_local) row_end = L->n_local;
    if (row_begin > row_end) row_begin = row_end;
#pragma omp parallel num_threads(T)
    {
        orc_queue q;
        vtt_t vis;
        queue_init(&q, topk > queue ? topk : queue);
        vtt_init(&vis, L->n_global);
        const uint32_t kk = topk ? topk : 1;
        uint32_t* ov = (uint32_t*)malloc(sizeof(uint32_t) * kk);
        float* od = (float*)malloc(sizeof(float) * kk);
#pragma omp for schedule(dynamic, 64)
        for (int64_t ri = row_begin; ri < (int64_t)row_end; ++ri) {
            const uint32_t sv = storage_vid(L, (uint32_t)ri);
            const uint32_t m = route_warm(L, (uint32_t)ri, topk, queue, &q, &vis, ov, od);
            const nvec* row = &L->rows[ri];
            for (uint32_t i = 0; i < m; ++i) {
                if (ov[i] == ORC_INVALID_VID || ov[i] == sv) continue;
                if (row->n >= L->cap && row->a[L->cap - 1].dist <= od[i]) continue;
                write_log(L, (uint32_t)ri, ov[i], od[i]);
            }
        }
        free(ov);
        free(od);
        queue_release(&q);
        free(vis.tags);
    }
    merge_logs(L, nthreads);
}

/* PruningUpdater::update_impl (refiner/updaters/pruning_updater.hpp:170-177)
 * = CandidatePruningUpdater<scaled_ineq, SkipOldPairs = true> (:63-102):
 * entries in row order; a candidate is dropped when some retained entry r
 * (pairs where both are old skipped) has dist(candidate, r) <
 * candidate.dist * (1 / scale); survivors marked old.  No logs. */
void orc_next_pruning(orc_layer* L, float scale_coeffs, int nthreads) {
    const float inv = 1.0f / scale_coeffs;
#pragma omp parallel for schedule(dynamic, 64) num_threads(pick_threads(nthreads))
    for (int64_t ri = 0; ri < (int64_t)L->n_local; ++ri) {
        nvec* row = &L->rows[ri];
        if (row->n == 0) continue;
        orc_nbr* kept = (orc_nbr*)malloc(sizeof(orc_nbr) * row->n);
        uint32_t nk = 0;
        for (uint32_t i = 0; i < row->n; ++i) {
            const orc_nbr c = row->a[i];
            const int c_old = (c.vid_status & ORC_NBR_NEW) == 0;
            const float thr = c.dist * inv;
            const float* cv = vec_of(L, c.vid_status & ORC_INVALID_VID);
            int keep = 1;
            for (uint32_t k = 0; k < nk; ++k) {
                if (c_old && (kept[k].vid_status & ORC_NBR_NEW) == 0) continue;
                const float t = orc_distance(L->metric, cv, vec_of(L, kept[k].vid_status & ORC_INVALID_VID), L->D);
                if (t < thr) {
                    keep = 0;
                    break;
                }
            }
            if (keep) kept[nk++] = c;
        }
        for (uint32_t k = 0; k < nk; ++k) kept[k].vid_status &= ORC_INVALID_VID;
        memcpy(row->a, kept, sizeof(orc_nbr) * nk);
        row->n = nk;
        free(kept);
    }
    merge_logs(L, nthreads);
}

/* The routing loops of refine_layer (artea_graph/index_factory.hpp:400-404):
 * next(routing).next(pruning), then next(reverse).next(arc) between loops.
 * Rows may end longer than the cap; the writeback keeps the first slots. */
void orc_routing_loops(orc_layer* L, uint32_t loops, uint32_t topk, uint32_t queue, float scale_coeffs,
                       int nthreads) {
    for (uint32_t l = 0; l < loops; ++l) {
        orc_next_routing(L, topk, queue, nthreads);
        orc_next_pruning(L, scale_coeffs, nthreads);
        if (l + 1 != loops) {
            orc_next_reverse(L, nthreads);
            orc_next_truncate(L, nthreads);
        }
    }
}

/* compute_max_allowed_level_id (index/stacked_rgraph/configs.hpp:241-249) */
uint32_t orc_max_level(uint32_t n) {
    if (n == 0) return 1;
    const double ratio = (double)n / 1000.0;
    if (ratio <= 1.0) return 1;
    const uint32_t c = (uint32_t)ceil(log(ratio));
    return c > 1 ? c : 1;
}

/* Source level of vertex v: level_fraction f = 1 puts every vertex on every
 * level 0..max_level; otherwise the highest h <= max_level with u(v) < f^h,
 * u(v) a seeded uniform draw in [0, 1).  This geometric assignment stands in
 * for the r-net insertion (index/stacked_rgraph/index_factory.hpp:281-480),
 * which is not restated here. */
uint32_t orc_build_level_of(uint64_t seed, uint32_t v, uint32_t max_level, float f) {
    if (f >= 1.0f) return max_level;
    uint64_t z = seed ^ 0x6C6576656C73ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z += (uint64_t)v * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    const double u = (double)(z >> 11) * 0x1.0p-53;
    uint32_t h = 0;
    double p = (double)f;
    while (h < max_level && u < p) {
        ++h;
        p *= (double)f;
    }
    return h;
}

/* refine_layer for every level of the source hierarchy (level l holds the
 * vertices with source level >= l, as a sparse layer when not all), rows
 * written back per vid, then compact_graph (trim, demotion, entry point). */
orc_graph* orc_build_full(const float* vecs, uint32_t n, uint32_t D, int build_metric, uint32_t ul_max,
                          uint32_t bl_max, uint32_t max_level, uint32_t build_loops, uint32_t triu_iters,
                          float prefill_ratio, float scale_coeffs, float shifted_coeffs, float l0_min_distance,
                          uint64_t seed, float level_fraction, int nthreads, double* secs) {
    return orc_build_full_routing(vecs, n, D, build_metric, ul_max, bl_max, max_level, build_loops, triu_iters,
                                  prefill_ratio, scale_coeffs, shifted_coeffs, l0_min_distance, seed, level_fraction,
                                  0, 64, 96, nthreads, secs);
}

/* The same with routing loops after each level's build loops. */
orc_graph* orc_build_full_routing(const float* vecs, uint32_t n, uint32_t D, int build_metric, uint32_t ul_max,
                                  uint32_t bl_max, uint32_t max_level, uint32_t build_loops, uint32_t triu_iters,
                                  float prefill_ratio, float scale_coeffs, float shifted_coeffs,
                                  float l0_min_distance, uint64_t seed, float level_fraction, uint32_t routing_loops,
                                  uint32_t routing_topk, uint32_t routing_queue, int nthreads, double* secs) {
    uint32_t* highest = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint32_t src_top = 0;
    for (uint32_t v = 0; v < n; ++v) {
        highest[v] = orc_build_level_of(seed, v, max_level, level_fraction);
        if (highest[v] > src_top) src_top = highest[v];
    }
    uint32_t* members = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint64_t** offs = (uint64_t**)calloc(max_level + 1, sizeof(uint64_t*));
    uint32_t** ids = (uint32_t**)calloc(max_level + 1, sizeof(uint32_t*));
    for (uint32_t l = 0; l <= max_level; ++l) {
        if (secs) secs[l] = 0.0;
        offs[l] = (uint64_t*)calloc((uint64_t)n + 1, sizeof(uint64_t));
        if (l > src_top) {
            ids[l] = (uint32_t*)malloc(sizeof(uint32_t));
            continue;
        }
        const double t0 = omp_get_wtime();
        uint32_t nl = 0;
        for (uint32_t v = 0; v < n; ++v)
            if (highest[v] >= l) members[nl++] = v;
        const uint32_t slot_cap = l == 0 ? bl_max : ul_max;
        /* IndexStructure refine caps: floor(1.5 * max) (artea_graph/index_structure.hpp:52-57) */
        const uint32_t cap = (uint32_t)((float)slot_cap * 1.5f);
        orc_layer* L = orc_layer_create(vecs, n, D, build_metric, cap, nl == n ? NULL : members, nl);
        const uint32_t g = (uint32_t)((float)cap * prefill_ratio);
        if (g > 0 && nl > 0) orc_next_prefill(L, g, 0, nl, 0, nl, seed, l, nthreads);
        const float shift = l == 0 ? shifted_coeffs * l0_min_distance : 0.0f;
        orc_refine_schedule(L, build_loops, triu_iters, scale_coeffs, shift, nthreads);
        orc_routing_loops(L, routing_loops, routing_topk, routing_queue, scale_coeffs, nthreads);
        for (uint32_t i = 0; i < nl; ++i) {
            const uint32_t c = L->rows[i].n < slot_cap ? L->rows[i].n : slot_cap;
            offs[l][members[i] + 1] = c;
        }
        for (uint32_t v = 0; v < n; ++v) offs[l][v + 1] += offs[l][v];
        ids[l] = (uint32_t*)malloc(sizeof(uint32_t) * (offs[l][n] ? offs[l][n] : 1));
        for (uint32_t i = 0; i < nl; ++i) {
            const uint32_t v = members[i];
            for (uint64_t k = 0; k < offs[l][v + 1] - offs[l][v]; ++k)
                ids[l][offs[l][v] + k] = L->rows[i].a[k].vid_status & ORC_INVALID_VID;
        }
        orc_layer_free(L);
        if (secs) secs[l] = omp_get_wtime() - t0;
    }
    orc_graph* gr = orc_compact_graph(n, max_level, ul_max, bl_max, highest, (const uint64_t* const*)offs,
                                      (const uint32_t* const*)ids, vecs, D, build_metric, NULL, 0);
    for (uint32_t l = 0; l <= max_level; ++l) {
        free(offs[l]);
        free(ids[l]);
    }
    free(offs);
    free(ids);
    free(highest);
    free(members);
    return gr;
}

/* ======================================================================
 * FP64 brute force (the recall oracle's ground truth).
 * ====================================================================== */
typedef struct {
    double d;
    uint32_t v;
} dpair;

static inline int dp_lt(dpair a, dpair b) { return a.d < b.d || (a.d == b.d && a.v < b.v); }

void orc_bruteforce_knn(int metric, const float* base, uint32_t n, const float* queries, uint32_t nq,
                        uint32_t D, uint32_t k, uint32_t* out_vids, double* out_dists, int nthreads) {
#pragma omp parallel for schedule(dynamic, 1) num_threads(pick_threads(nthreads))
    for (int64_t qi = 0; qi < (int64_t)nq; ++qi) {
        const float* q = queries + (uint64_t)qi * D;
        dpair* best = (dpair*)malloc(sizeof(dpair) * (k + 1));
        uint32_t nb = 0;
        for (uint32_t v = 0; v < n; ++v) {
            dpair x = {orc_distance_f64(metric, q, base + (uint64_t)v * D, D), v};
            if (nb == k && !dp_lt(x, best[k - 1])) continue;
            uint32_t j = nb < k ? nb++ : k - 1;
            while (j > 0 && dp_lt(x, best[j - 1])) {
                best[j] = best[j - 1];
                --j;
            }
            best[j] = x;
        }
        for (uint32_t i = 0; i < k; ++i) {
            out_vids[(uint64_t)qi * k + i] = i < nb ? best[i].v : ORC_INVALID_VID;
            out_dists[(uint64_t)qi * k + i] = i < nb ? best[i].d : INFINITY;
        }
        free(best);
    }
}

/* ======================================================================
 * Scoring and dataset probes.
 * ====================================================================== */
static int cmp_u32(const void* a, const void* b) {
    const uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

/* RecallEstimator::calculate_recall_at_k (utils/recall_estimator.hpp:54-137):
 * per query, sort the first k ground-truth ids and the first k predicted
 * ids, drop repeated predictions (std::unique), and count the sorted
 * intersection (std::set_intersection); the sum over queries. */
uint64_t orc_recall_matches(const uint32_t* pred, uint32_t pstride, const uint32_t* gt, uint32_t gstride, uint32_t nq,
                            uint32_t k) {
    uint64_t total = 0;
    uint32_t* sg = (uint32_t*)malloc(sizeof(uint32_t) * (k ? k : 1));
    uint32_t* sp = (uint32_t*)malloc(sizeof(uint32_t) * (k ? k : 1));
    for (uint32_t q = 0; q < nq; ++q) {
        memcpy(sg, gt + (size_t)q * gstride, sizeof(uint32_t) * k);
        memcpy(sp, pred + (size_t)q * pstride, sizeof(uint32_t) * k);
        qsort(sg, k, sizeof(uint32_t), cmp_u32);
        qsort(sp, k, sizeof(uint32_t), cmp_u32);
        uint32_t u = 0; /* std::unique */
        for (uint32_t i = 0; i < k; ++i)
            if (u == 0 || sp[u - 1] != sp[i]) sp[u++] = sp[i];
        uint32_t i = 0, j = 0; /* std::set_intersection */
        while (i < k && j < u) {
            if (sg[i] < sp[j]) ++i;
            else if (sp[j] < sg[i]) ++j;
            else {
                ++total;
                ++i;
                ++j;
            }
        }
    }
    free(sg);
    free(sp);
    return total;
}

/* ADREstimator::calculate_adr_at_1 (utils/adr_estimator.hpp): mean over
 * queries of double(dist(q, retrieved)) / double(dist(q, true NN)), queries
 * with a non-positive ground-truth distance skipped; summed in query order.
 * A top-1 slot without a vertex (id >= n) is skipped too (the reference
 * would read past the base there). */
double orc_adr_at_1(int metric, const float* base, uint32_t n, const float* queries, uint32_t nq, uint32_t D,
                    const uint32_t* pred, uint32_t pstride, const uint32_t* gt, uint32_t gstride, uint64_t* counted) {
    double sum = 0.0;
    uint64_t c = 0;
    for (uint32_t q = 0; q < nq; ++q) {
        const uint32_t r = pred[(size_t)q * pstride], t = gt[(size_t)q * gstride];
        if (r >= n || t >= n) continue;
        const float* x = queries + (size_t)q * D;
        const float dr = orc_distance(metric, x, base + (size_t)r * D, D);
        const float dt = orc_distance(metric, x, base + (size_t)t * D, D);
        if (dt <= 0.0f) continue;
        sum += (double)dr / (double)dt;
        ++c;
    }
    if (counted) *counted = c;
    return c ? sum / (double)c : 0.0;
}

/* Max-heap of floats (std::priority_queue<distance_t>). */
static void fheap_push(float* h, uint32_t* n, float v) {
    uint32_t i = (*n)++;
    h[i] = v;
    while (i > 0) {
        const uint32_t p = (i - 1) / 2;
        if (!(h[p] < h[i])) break;
        const float t = h[p];
        h[p] = h[i];
        h[i] = t;
        i = p;
    }
}
static float fheap_pop(float* h, uint32_t* n) {
    const float top = h[0];
    h[0] = h[--(*n)];
    uint32_t i = 0;
    for (;;) {
        const uint32_t l = 2 * i + 1, r = l + 1;
        uint32_t m = i;
        if (l < *n && h[m] < h[l]) m = l;
        if (r < *n && h[m] < h[r]) m = r;
        if (m == i) break;
        const float t = h[m];
        h[m] = h[i];
        h[i] = t;
        i = m;
    }
    return top;
}

/* DatasetProber::_compute_knn_table (utils/dataset_prober.hpp:471-540) for
 * given samples: per sample, a max-heap of the 128 smallest dist(sample, j)
 * over j != sample, extracted ascending into out[s * 128 ..]. */
void orc_probe_knn(int metric, const float* base, uint32_t n, uint32_t D, const uint32_t* samples, uint32_t ns,
                   float* out, int nthreads) {
    enum { K = 128 };
#pragma omp parallel for schedule(dynamic, 1) num_threads(pick_threads(nthreads))
    for (int64_t s = 0; s < (int64_t)ns; ++s) {
        float heap[K];
        uint32_t hn = 0;
        const uint32_t vid = samples[s];
        const float* q = base + (size_t)vid * D;
        for (uint32_t j = 0; j < n; ++j) {
            if (j == vid) continue;
            const float d = orc_distance(metric, q, base + (size_t)j * D, D);
            if (hn < K) {
                fheap_push(heap, &hn, d);
            } else if (d < heap[0]) {
                fheap_pop(heap, &hn);
                fheap_push(heap, &hn, d);
            }
        }
        for (uint32_t k = hn; k > 0; --k) out[(size_t)s * K + k - 1] = fheap_pop(heap, &hn);
    }
}

/* DatasetProber::_compute_farthest (:542-585) for given samples. */
void orc_probe_farthest(int metric, const float* base, uint32_t n, uint32_t D, const uint32_t* samples, uint32_t ns,
                        float* out, int nthreads) {
#pragma omp parallel for schedule(dynamic, 1) num_threads(pick_threads(nthreads))
    for (int64_t s = 0; s < (int64_t)ns; ++s) {
        const uint32_t vid = samples[s];
        float m = -FLT_MAX;
        for (uint32_t j = 0; j < n; ++j) {
            if (j == vid) continue;
            const float d = orc_distance(metric, base + (size_t)vid * D, base + (size_t)j * D, D);
            if (d > m) m = d;
        }
        out[s] = m;
    }
}

/* ======================================================================
 * R-net insertion (index/stacked_rgraph/index_factory.hpp:192-583) in
 * waves: the vertices of a wave all search the graph as it stood at the
 * wave's start and are then applied in insertion order (level, forward
 * rows, reverse edges).  A wave ends after the first vertex that grows the
 * hierarchy.  Waves of one vertex are the reference's serial insertion.
 * The apex entry of a descent is drawn from the counter-based stream (the
 * reference draws it from a thread-local std::mt19937_64 seeded by
 * std::random_device, candidate_sample_utils.hpp:108-148).
 * ====================================================================== */
#define RNET_STREAM 0x726E6574ull /* entry draws */

typedef struct {
    const float* vecs;
    uint32_t n, D, max_level;
    int metric;
    orc_rnet_params p;
    nvec* rows;         /* (max_level + 1) * n rows, level-major */
    uint32_t* highest;  /* ORC_INVALID_LEVEL until assigned */
    uint32_t** bucket;  /* vids per highest level, in assignment order */
    uint32_t* bucket_n;
    uint32_t top;       /* ORC_INVALID_LEVEL: empty */
    uint32_t waves;
} rnet_t;

static inline const float* vec_of_g(const float* vecs, uint32_t D, uint32_t v) { return vecs + (uint64_t)v * D; }

static inline nvec* rnet_row(const rnet_t* R, uint32_t l, uint32_t v) { return &R->rows[(size_t)l * R->n + v]; }
static inline uint32_t rnet_cap(const rnet_t* R, uint32_t l) { return l == 0 ? R->p.bl_max : R->p.ul_max; }

/* RGraphConfig::radius_at (stacked_rgraph/configs.hpp:130-138). */
float orc_rnet_radius(float l0_min_distance, float tau_k, float beta, uint32_t h) {
    if (h == 0) return l0_min_distance;
    return (float)((double)l0_min_distance * (1.0 + (double)tau_k) * pow((double)beta, (double)(h - 1)));
}

/* Wave length after `inserted` vertices: the reference's serial startup
 * (startup_points = 10, index_factory.hpp:104-105, 242-254), then
 * inserted / wave_div vertices, at most wave_max. */
uint32_t orc_rnet_wave(uint32_t inserted, uint32_t wave_div, uint32_t wave_max) {
    if (inserted < 10) return 1;
    uint32_t w = wave_div ? inserted / wave_div : wave_max;
    if (w < 1) w = 1;
    return w > wave_max ? wave_max : w;
}

/* HierarchicalPruningUpdater::update_impl (refiner/updaters/
 * hierarchical_pruning_updater.hpp:96-138): the first entry is kept; each
 * later one while fewer than max_nbr are kept is kept iff every kept entry
 * (pairs of two old entries skipped) is at least dist * (1 / scale) away
 * from it; kept entries marked old. */
static uint32_t rnet_prune(const rnet_t* R, orc_nbr* a, uint32_t m, uint32_t max_nbr) {
    if (m == 0) return 0;
    const float inv = 1.0f / R->p.scale;
    uint32_t nk = 1;
    for (uint32_t i = 1; i < m && nk < max_nbr; ++i) {
        const orc_nbr c = a[i];
        const float thr = c.dist * inv;
        const int c_old = (c.vid_status & ORC_NBR_NEW) == 0;
        const float* cv = vec_of_g(R->vecs, R->D, c.vid_status & ORC_INVALID_VID);
        int keep = 1;
        for (uint32_t k = 0; k < nk; ++k) {
            if (c_old && (a[k].vid_status & ORC_NBR_NEW) == 0) continue;
            if (orc_distance(R->metric, cv, vec_of_g(R->vecs, R->D, a[k].vid_status & ORC_INVALID_VID), R->D) < thr) {
                keep = 0;
                break;
            }
        }
        if (keep) a[nk++] = c;
    }
    for (uint32_t k = 0; k < nk; ++k) a[k].vid_status &= ORC_INVALID_VID;
    return nk;
}

/* SingleLayerRouter::beam_search (single_layer_router.hpp:145-163) over
 * level l's rows (DynamicLayerRange: entries until the first invalid one). */
static void rnet_beam(const rnet_t* R, const float* q, uint32_t l, orc_queue* cq, vtt_t* vis) {
    orc_queue_reset_exploration(cq);
    if (cq->U.n == 0) return;
    for (uint32_t i = 0; i < cq->T.n; ++i) vtt_set(vis, cq->T.a[i].vid);
    while (cq->U.n != 0) {
        if (queue_should_terminate(cq)) break;
        const centry cur = heap_pop(&cq->U);
        if (cur.vid == ORC_INVALID_VID) break;
        const nvec* row = rnet_row(R, l, cur.vid);
        for (uint32_t j = 0; j < row->n; ++j) {
            const uint32_t u = row->a[j].vid_status & ORC_INVALID_VID;
            if (vtt_test_and_set(vis, u)) continue;
            queue_try_push(cq, u, orc_distance(R->metric, q, vec_of_g(R->vecs, R->D, u), R->D));
        }
    }
}

typedef struct {
    uint32_t h;
    int grows;
    uint32_t* nf;   /* forward entries per level */
    orc_nbr** f;
} rnet_plan;

static void queue_copy_top(orc_queue* dst, const orc_queue* src, uint32_t capacity) {
    /* seed_from_queue (std_candidate_queue.hpp:456-472) */
    queue_clear(dst);
    dst->capacity = capacity;
    for (uint32_t i = 0; i < src->T.n; ++i) queue_try_push(dst, src->T.a[i].vid, src->T.a[i].d);
}

/* Steps A, B and the searches of D of _insert_one (index_factory.hpp
 * :281-478) against the current graph; nothing is written. */
static void rnet_plan_vertex(const rnet_t* R, uint32_t qv, orc_queue* stash, orc_queue* cq, vtt_t* vis,
                             rnet_plan* out) {
    const orc_rnet_params* P = &R->p;
    const uint32_t L = R->max_level;
    const float* q = vec_of_g(R->vecs, R->D, qv);
    float* min_dist = (float*)malloc(sizeof(float) * (L + 1));
    for (uint32_t l = 0; l <= L; ++l) {
        min_dist[l] = FLT_MAX;
        queue_clear(&stash[l]);
        stash[l].capacity = P->search_qs;
    }
    const uint32_t top = R->top;
    if (top != ORC_INVALID_LEVEL && top >= 1) {
        queue_clear(cq);
        cq->capacity = P->search_qs;
        const uint32_t sz = R->bucket_n[top];
        const uint32_t e = R->bucket[top][orc_rand_draw(P->seed, RNET_STREAM, qv, 0, 0, sz)];
        queue_try_push(cq, e, orc_distance(R->metric, q, vec_of_g(R->vecs, R->D, e), R->D));
        vtt_clear(vis);
        for (uint32_t l = top; l >= 1; --l) {
            rnet_beam(R, q, l, cq, vis);
            if (cq->T.n > 0) {
                float b = FLT_MAX;
                for (uint32_t i = 0; i < cq->T.n; ++i)
                    if (cq->T.a[i].d < b) b = cq->T.a[i].d;
                min_dist[l] = b;
            }
            queue_copy_top(&stash[l], cq, P->search_qs);
            if (l == 1) break;
            orc_queue nx;
            queue_init(&nx, P->search_qs);
            queue_copy_top(&nx, cq, P->search_qs);
            queue_release(cq);
            *cq = nx;
        }
    }
    if (P->insert_on_l0 && stash[1].T.n > 0) queue_copy_top(&stash[0], &stash[1], P->search_qs);
    uint32_t h;
    if (top == ORC_INVALID_LEVEL) {
        h = 1;
    } else {
        h = top + 1 < L ? top + 1 : L;
        for (uint32_t k = 1; k <= top; ++k)
            if (min_dist[k] <= orc_rnet_radius(P->l0_min_distance, P->tau_k, P->beta, k)) {
                h = k - 1;
                break;
            }
    }
    out->h = h;
    out->grows = top == ORC_INVALID_LEVEL || h > top;
    out->nf = (uint32_t*)calloc(L + 1, sizeof(uint32_t));
    out->f = (orc_nbr**)calloc(L + 1, sizeof(orc_nbr*));
    for (uint32_t l = P->insert_on_l0 ? 0 : 1; l <= h; ++l) {
        if (stash[l].T.n == 0) continue;
        orc_queue sq;
        queue_init(&sq, l == 0 ? P->bl_select_qs : P->ul_select_qs);
        for (uint32_t i = 0; i < stash[l].T.n; ++i) queue_try_push(&sq, stash[l].T.a[i].vid, stash[l].T.a[i].d);
        vtt_clear(vis);
        rnet_beam(R, q, l, &sq, vis);
        const uint32_t m = sq.T.n;
        uint32_t* vv = (uint32_t*)malloc(sizeof(uint32_t) * (m ? m : 1));
        float* dd = (float*)malloc(sizeof(float) * (m ? m : 1));
        orc_queue_extract(&sq, m, vv, dd);
        orc_nbr* c = (orc_nbr*)malloc(sizeof(orc_nbr) * (m ? m : 1));
        uint32_t nc = 0;
        for (uint32_t i = 0; i < m; ++i) {
            if (vv[i] == qv) continue;
            c[nc].vid_status = vv[i] | ORC_NBR_NEW;
            c[nc].dist = dd[i];
            ++nc;
        }
        out->nf[l] = rnet_prune(R, c, nc, rnet_cap(R, l));
        out->f[l] = c;
        free(vv);
        free(dd);
        queue_release(&sq);
    }
    free(min_dist);
}

/* Stable order by distance for the re-pruned full slot (the reference's
 * std::sort by distance leaves equal distances in an unspecified order). */
static void stable_by_dist(orc_nbr* a, uint32_t n) {
    for (uint32_t i = 1; i < n; ++i) {
        const orc_nbr x = a[i];
        uint32_t j = i;
        while (j > 0 && a[j - 1].dist > x.dist) {
            a[j] = a[j - 1];
            --j;
        }
        a[j] = x;
    }
}

/* Step C and the writes of D: assign_layer, _write_forward_edges and
 * _write_reverse_edges (index_factory.hpp:497-580). */
static void rnet_apply(rnet_t* R, uint32_t qv, const rnet_plan* pl) {
    const uint32_t h = pl->h;
    R->highest[qv] = h;
    if (R->bucket_n[h] % 1024 == 0) R->bucket[h] = (uint32_t*)realloc(R->bucket[h], sizeof(uint32_t) * (R->bucket_n[h] + 1024));
    R->bucket[h][R->bucket_n[h]++] = qv;
    if (pl->grows) R->top = h;
    for (uint32_t l = R->p.insert_on_l0 ? 0 : 1; l <= h; ++l) {
        const uint32_t m = pl->nf[l];
        if (m == 0) continue;
        const uint32_t cap = rnet_cap(R, l);
        nvec* own = rnet_row(R, l, qv);
        own->n = 0;
        for (uint32_t i = 0; i < m && i < cap; ++i) nvec_push(own, pl->f[l][i]);
        for (uint32_t i = 0; i < m; ++i) {
            const uint32_t nb = pl->f[l][i].vid_status & ORC_INVALID_VID;
            const orc_nbr back = {qv | ORC_NBR_NEW, pl->f[l][i].dist};
            nvec* row = rnet_row(R, l, nb);
            if (row->n < cap) {
                nvec_push(row, back);
                continue;
            }
            orc_nbr* merged = (orc_nbr*)malloc(sizeof(orc_nbr) * (cap + 1));
            memcpy(merged, row->a, sizeof(orc_nbr) * cap);
            merged[cap] = back;
            stable_by_dist(merged, cap + 1);
            const uint32_t k = rnet_prune(R, merged, cap + 1, cap);
            memcpy(row->a, merged, sizeof(orc_nbr) * k);
            row->n = k;
            free(merged);
        }
    }
}

static void plan_free(rnet_plan* pl, uint32_t L) {
    for (uint32_t l = 0; l <= L; ++l) free(pl->f[l]);
    free(pl->f);
    free(pl->nf);
}

orc_rnet* orc_rnet_build(const float* vecs, uint32_t n, uint32_t D, int metric, const orc_rnet_params* p,
                         int nthreads) {
    rnet_t* R = (rnet_t*)calloc(1, sizeof(rnet_t));
    R->vecs = vecs;
    R->n = n;
    R->D = D;
    R->metric = metric;
    R->p = *p;
    R->max_level = orc_max_level(n);
    const uint32_t L = R->max_level;
    R->rows = (nvec*)calloc((size_t)(L + 1) * (n ? n : 1), sizeof(nvec));
    R->highest = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    for (uint32_t v = 0; v < n; ++v) R->highest[v] = ORC_INVALID_LEVEL;
    R->bucket = (uint32_t**)calloc(L + 1, sizeof(uint32_t*));
    R->bucket_n = (uint32_t*)calloc(L + 1, sizeof(uint32_t));
    R->top = ORC_INVALID_LEVEL;
    const int T = pick_threads(nthreads);
    rnet_plan* plans = (rnet_plan*)calloc(p->wave_max ? p->wave_max : 1, sizeof(rnet_plan));
    uint32_t pos = 0;
    R->waves = 0;
    while (pos < n) {
        uint32_t w = orc_rnet_wave(pos, p->wave_div, p->wave_max);
        if (w > n - pos) w = n - pos;
#pragma omp parallel num_threads(T)
        {
            orc_queue cq;
            queue_init(&cq, p->search_qs);
            orc_queue* stash = (orc_queue*)malloc(sizeof(orc_queue) * (L + 1));
            for (uint32_t l = 0; l <= L; ++l) queue_init(&stash[l], p->search_qs);
            vtt_t vis;
            vtt_init(&vis, n);
#pragma omp for schedule(dynamic, 1)
            for (int64_t i = 0; i < (int64_t)w; ++i) rnet_plan_vertex(R, pos + (uint32_t)i, stash, &cq, &vis, &plans[i]);
            for (uint32_t l = 0; l <= L; ++l) queue_release(&stash[l]);
            free(stash);
            queue_release(&cq);
            free(vis.tags);
        }
        uint32_t end = w;
        for (uint32_t i = 0; i < w; ++i)
            if (plans[i].grows) {
                end = i + 1;
                break;
            }
        for (uint32_t i = 0; i < end; ++i) rnet_apply(R, pos + i, &plans[i]);
        for (uint32_t i = 0; i < w; ++i) plan_free(&plans[i], L);
        pos += end;
        R->waves++;
    }
    free(plans);
    return (orc_rnet*)R;
}

void orc_rnet_free(orc_rnet* h) {
    rnet_t* R = (rnet_t*)h;
    if (!R) return;
    for (size_t i = 0; i < (size_t)(R->max_level + 1) * R->n; ++i) free(R->rows[i].a);
    free(R->rows);
    for (uint32_t l = 0; l <= R->max_level; ++l) free(R->bucket[l]);
    free(R->bucket);
    free(R->bucket_n);
    free(R->highest);
    free(R);
}
uint32_t orc_rnet_max_level(const orc_rnet* h) { return ((const rnet_t*)h)->max_level; }
uint32_t orc_rnet_top(const orc_rnet* h) { return ((const rnet_t*)h)->top; }
uint32_t orc_rnet_waves(const orc_rnet* h) { return ((const rnet_t*)h)->waves; }
const uint32_t* orc_rnet_highest(const orc_rnet* h) { return ((const rnet_t*)h)->highest; }
/* Rows of level l, by vid: counts[v] and the entries (vid_status, dist)
 * back to back; returns the total. */
uint64_t orc_rnet_export_level(const orc_rnet* h, uint32_t l, uint32_t* counts, orc_nbr* entries, uint64_t max) {
    const rnet_t* R = (const rnet_t*)h;
    uint64_t pos = 0;
    for (uint32_t v = 0; v < R->n; ++v) {
        const nvec* r = rnet_row(R, l, v);
        if (counts) counts[v] = r->n;
        for (uint32_t k = 0; k < r->n; ++k, ++pos)
            if (entries && pos < max) entries[pos] = r->a[k];
    }
    return pos;
}

/* The artea_graph build (index/artea_graph/index_factory.hpp:135-410): r-net
 * insertion, then refine_layer on every occupied level with at least
 * min_layer_cap (32) vertices (rows imported sorted by (distance, vid),
 * L0 topped up by the random updater, upper levels from their r-net rows),
 * the build loops and the routing loops, the writeback of the first
 * max_nbr_size entries, and the compactor with the insertion order. */
orc_graph* orc_build_full_rnet(const float* vecs, uint32_t n, uint32_t D, int build_metric, const orc_rnet_params* rp,
                               uint32_t build_loops, uint32_t triu_iters, float prefill_ratio, float shifted_coeffs,
                               uint32_t routing_loops, uint32_t routing_topk, uint32_t routing_queue, int nthreads,
                               double* secs) {
    const double t0 = omp_get_wtime();
    rnet_t* R = (rnet_t*)orc_rnet_build(vecs, n, D, build_metric, rp, nthreads);
    if (secs) secs[0] = omp_get_wtime() - t0;
    const uint32_t L = R->max_level;
    uint32_t* members = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint64_t** offs = (uint64_t**)calloc(L + 1, sizeof(uint64_t*));
    uint32_t** ids = (uint32_t**)calloc(L + 1, sizeof(uint32_t*));
    const uint32_t top = R->top == ORC_INVALID_LEVEL ? 0 : R->top;
    for (uint32_t l = 0; l <= L; ++l) {
        offs[l] = (uint64_t*)calloc((uint64_t)n + 1, sizeof(uint64_t));
        uint32_t nl = 0;
        for (uint32_t v = 0; v < n; ++v)
            if (R->highest[v] != ORC_INVALID_LEVEL && R->highest[v] >= l) members[nl++] = v;
        const uint32_t slot_cap = rnet_cap(R, l);
        uint32_t* cnt = (uint32_t*)calloc(nl ? nl : 1, sizeof(uint32_t));
        orc_nbr** rowp = (orc_nbr**)calloc(nl ? nl : 1, sizeof(orc_nbr*));
        orc_layer* Ly = NULL;
        if (l <= top && nl >= 32) {
            const uint32_t cap = (uint32_t)((float)slot_cap * 1.5f);
            Ly = orc_layer_create(vecs, n, D, build_metric, cap, nl == n ? NULL : members, nl);
            for (uint32_t i = 0; i < nl; ++i) {
                const nvec* r = rnet_row(R, l, members[i]);
                orc_layer_fill_row(Ly, i, r->a, r->n);
            }
            if (l == 0) {
                const uint32_t g = (uint32_t)((float)cap * prefill_ratio);
                if (g > 0) orc_next_prefill(Ly, g, 0, nl, 0, nl, rp->seed, 0, nthreads);
            }
            const float shift = l == 0 ? shifted_coeffs * rp->l0_min_distance : 0.0f;
            orc_refine_schedule(Ly, build_loops, triu_iters, rp->scale, shift, nthreads);
            orc_routing_loops(Ly, routing_loops, routing_topk, routing_queue, rp->scale, nthreads);
            for (uint32_t i = 0; i < nl; ++i) {
                cnt[i] = Ly->rows[i].n < slot_cap ? Ly->rows[i].n : slot_cap;
                rowp[i] = Ly->rows[i].a;
            }
        } else {
            for (uint32_t i = 0; i < nl; ++i) {
                const nvec* r = rnet_row(R, l, members[i]);
                cnt[i] = r->n < slot_cap ? r->n : slot_cap;
                rowp[i] = r->a;
            }
        }
        for (uint32_t i = 0; i < nl; ++i) offs[l][members[i] + 1] = cnt[i];
        for (uint32_t v = 0; v < n; ++v) offs[l][v + 1] += offs[l][v];
        ids[l] = (uint32_t*)malloc(sizeof(uint32_t) * (offs[l][n] ? offs[l][n] : 1));
        for (uint32_t i = 0; i < nl; ++i) {
            const uint32_t v = members[i];
            for (uint32_t k = 0; k < cnt[i]; ++k) ids[l][offs[l][v] + k] = rowp[i][k].vid_status & ORC_INVALID_VID;
        }
        if (Ly) orc_layer_free(Ly);
        free(cnt);
        free(rowp);
    }
    if (secs) secs[1] = omp_get_wtime() - t0;
    uint32_t* order = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    for (uint32_t v = 0; v < n; ++v) order[v] = v;
    orc_graph* gr = orc_compact_graph_ordered(n, L, rp->ul_max, rp->bl_max, R->highest, (const uint64_t* const*)offs,
                                              (const uint32_t* const*)ids, order, vecs, D, build_metric, NULL, 0);
    for (uint32_t l = 0; l <= L; ++l) {
        free(offs[l]);
        free(ids[l]);
    }
    free(offs);
    free(ids);
    free(order);
    free(members);
    orc_rnet_free((orc_rnet*)R);
    return gr;
}

/* HierarchicalPruningUpdater::update_impl on a caller's list (sorted by
 * distance); returns the kept count, the list rewritten in place. */
uint32_t orc_hier_prune(const float* vecs, uint32_t D, int metric, orc_nbr* a, uint32_t m, uint32_t max_nbr,
                        float scale) {
    rnet_t R;
    memset(&R, 0, sizeof R);
    R.vecs = vecs;
    R.D = D;
    R.metric = metric;
    R.p.scale = scale;
    return rnet_prune(&R, a, m, max_nbr);
}
