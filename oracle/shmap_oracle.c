/*
 * shmap oracle -- TEST INFRASTRUCTURE ONLY (see shmap_oracle.h).
 *
 * A line-by-line restatement, in plain C, of shmap-rs's per-read mapping path.
 * It is the parity checker for the HIP library and the CPU baseline timed by
 * the benchmark; the shipped library never links it.
 *
 * Reference files are cited as path:line relative to the shmap-rs crate root.
 */
#define _GNU_SOURCE
#include "shmap_oracle.h"

#include <limits.h>
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define MIN_HALFLEN 5 /* src/buckets.rs:15 */

static void *xmalloc(size_t n) {
    void *p = malloc(n ? n : 1);
    if (!p) {
        fprintf(stderr, "oracle: out of memory (%zu bytes)\n", n);
        abort();
    }
    return p;
}
static void *xrealloc(void *p, size_t n) {
    void *q = realloc(p, n ? n : 1);
    if (!q) {
        fprintf(stderr, "oracle: out of memory (%zu bytes)\n", n);
        abort();
    }
    return q;
}
#define GROW(ptr, cap, need)                                                   \
    do {                                                                       \
        if ((int64_t)(need) > (int64_t)(cap)) {                                \
            int64_t nc_ = (int64_t)(cap) * 2;                                  \
            if (nc_ < (int64_t)(need)) nc_ = (int64_t)(need);                  \
            if (nc_ < 16) nc_ = 16;                                            \
            (ptr) = xrealloc((ptr), (size_t)nc_ * sizeof(*(ptr)));             \
            (cap) = nc_;                                                       \
        }                                                                      \
    } while (0)

/* ------------------------------------------------------------------------- */
/* Sketching: src/sketch.rs                                                  */
/* ------------------------------------------------------------------------- */

static inline uint64_t rotl64(uint64_t x, uint32_t n) {
    n &= 63u;
    return n ? (x << n) | (x >> (64u - n)) : x;
}
static inline uint64_t rotr64(uint64_t x, uint32_t n) {
    n &= 63u;
    return n ? (x >> n) | (x << (64u - n)) : x;
}

typedef struct {
    uint64_t fw[256], rc[256], fw_k[256], rc_r1[256], rc_k1[256];
    int32_t k;
    double h_frac;
    uint64_t thres;
} sketcher_t;

/* src/sketch.rs:126-132 -- `(h_frac * u64::MAX as f64) as u64`; u64::MAX as
 * f64 rounds to 2^64. */
uint64_t ok_threshold(double h_frac) {
    if (h_frac < 1.0) {
        double v = h_frac * 18446744073709551616.0;
        if (!(v > 0.0)) return 0; /* `as u64` saturates NaN/negatives to 0 */
        if (v >= 18446744073709551616.0) return UINT64_MAX;
        return (uint64_t)v;
    }
    return UINT64_MAX;
}

/* src/sketch.rs:58-109 -- ntHash LUTs; any non-ACGT byte contributes 0. */
static void sketcher_init(sketcher_t *sk, int32_t k, double h_frac) {
    const uint64_t A = 0x3c8bfbb395c60474ull, C = 0x3193c18562a02b4cull,
                   G = 0x20323ed082572324ull, T = 0x295549f54be24456ull;
    memset(sk, 0, sizeof(*sk));
    sk->fw['a'] = sk->fw['A'] = A;
    sk->fw['c'] = sk->fw['C'] = C;
    sk->fw['g'] = sk->fw['G'] = G;
    sk->fw['t'] = sk->fw['T'] = T;
    sk->rc['a'] = sk->rc['A'] = sk->fw['T'];
    sk->rc['c'] = sk->rc['C'] = sk->fw['G'];
    sk->rc['g'] = sk->rc['G'] = sk->fw['C'];
    sk->rc['t'] = sk->rc['T'] = sk->fw['A'];
    for (int c = 0; c < 256; c++) {
        sk->fw_k[c] = rotl64(sk->fw[c], (uint32_t)k);
        sk->rc_r1[c] = rotr64(sk->rc[c], 1);
        sk->rc_k1[c] = rotl64(sk->rc[c], (uint32_t)(k - 1));
    }
    sk->k = k;
    sk->h_frac = h_frac;
    sk->thres = ok_threshold(h_frac);
}

typedef struct {
    ok_kmer *v;
    int64_t n, cap;
} kvec_t;

static inline void kvec_push(kvec_t *kv, int32_t r, uint64_t h, uint8_t strand) {
    GROW(kv->v, kv->cap, kv->n + 1);
    kv->v[kv->n].r = r;
    kv->v[kv->n].h = h;
    kv->v[kv->n].strand = strand;
    kv->n++;
}

/* src/sketch.rs:167-219 -- sketch_slice_into + emit.  Appends to kv. */
static void sketch_slice(const sketcher_t *sk, const uint8_t *s, int64_t len, int32_t offset,
                         kvec_t *kv) {
    const int32_t k = sk->k;
    if ((int32_t)len < k || k <= 0) return; /* :170-173 (`s.len() as RPos`) */
    uint64_t hf = 0, hr = 0;
    for (int32_t i = 0; i < k; i++) { /* :180-184 */
        const uint8_t c = s[i];
        hf ^= rotl64(sk->fw[c], (uint32_t)(k - i - 1));
        hr ^= rotl64(sk->rc[c], (uint32_t)i);
    }
    int32_t r = k - 1 + offset; /* :196 */
    {
        const uint64_t h = hr ^ hf; /* emit :214-219 */
        if (h <= sk->thres) kvec_push(kv, r, h, hf > hr);
    }
    for (int64_t j = 0; j + k < len; j++) { /* :197-206 */
        const uint8_t in = s[j + k], out = s[j];
        hf = rotl64(hf, 1) ^ sk->fw_k[out] ^ sk->fw[in];
        hr = rotr64(hr, 1) ^ sk->rc_r1[out] ^ sk->rc_k1[in];
        r++;
        const uint64_t h = hr ^ hf;
        if (h <= sk->thres) kvec_push(kv, r, h, hf > hr);
    }
}

int64_t ok_sketch(int32_t k, double h_frac, const uint8_t *s, int64_t len, int32_t offset,
                  ok_kmer *out, int64_t cap) {
    sketcher_t sk;
    sketcher_init(&sk, k, h_frac);
    kvec_t kv = {0};
    sketch_slice(&sk, s, len, offset, &kv);
    const int64_t n = kv.n;
    if (out) memcpy(out, kv.v, (size_t)(n < cap ? n : cap) * sizeof(ok_kmer));
    free(kv.v);
    return n;
}

/* ------------------------------------------------------------------------- */
/* Reference index: src/index.rs                                              */
/* ------------------------------------------------------------------------- */

typedef struct {
    int32_t r, tpos, segm_id;
    uint8_t strand;
} hit_t; /* src/types.rs:44-51 */

typedef struct {
    uint64_t key;
    uint32_t start;
    uint32_t count; /* 0 = empty slot */
} hslot_t;

struct ok_index {
    int32_t k;
    double h_frac;
    int32_t n_segments;
    char **names;
    int32_t *sz;        /* segment length in bases */
    int64_t *kstart;    /* first global k-mer index of each segment */
    int64_t n_kmers;    /* N_T */
    uint64_t *t_h;      /* reference sketch, concatenated over segments */
    int32_t *t_r;
    uint8_t *t_strand;
    hit_t *hits;        /* every indexed hit, grouped by hash, (segm_id, r) order */
    hslot_t *table;     /* hash -> (start, count); keys over -M are absent */
    uint64_t mask;
    int64_t indexed_kmers, indexed_hits, blacklisted_kmers;
};

static inline uint64_t slot_of(uint64_t h, uint64_t mask) {
    return ((h ^ (h >> 29)) * 0x9E3779B97F4A7C15ull >> 17) & mask;
}

static const hslot_t *index_find(const ok_index *idx, uint64_t h) {
    uint64_t pos = slot_of(h, idx->mask);
    for (;;) {
        const hslot_t *s = &idx->table[pos];
        if (s->count == 0) return NULL;
        if (s->key == h) return s;
        pos = (pos + 1) & idx->mask;
    }
}

/* src/index.rs:204-213 */
int32_t ok_index_count(const ok_index *idx, uint64_t h) {
    const hslot_t *s = index_find(idx, h);
    return s ? (int32_t)s->count : 0;
}

int32_t ok_index_hits(const ok_index *idx, uint64_t h, int32_t *out4, int32_t cap) {
    const hslot_t *s = index_find(idx, h);
    if (!s) return 0;
    for (uint32_t i = 0; i < s->count && (int32_t)i < cap; i++) {
        const hit_t *x = &idx->hits[s->start + i];
        out4[4 * i + 0] = x->r;
        out4[4 * i + 1] = x->tpos;
        out4[4 * i + 2] = x->strand;
        out4[4 * i + 3] = x->segm_id;
    }
    return (int32_t)s->count;
}

int32_t ok_index_n_segments(const ok_index *idx) { return idx->n_segments; }
int64_t ok_index_segment_kmers(const ok_index *idx, int32_t s) {
    return idx->kstart[s + 1] - idx->kstart[s];
}
int64_t ok_index_indexed_kmers(const ok_index *idx) { return idx->indexed_kmers; }
int64_t ok_index_indexed_hits(const ok_index *idx) { return idx->indexed_hits; }
int64_t ok_index_blacklisted_kmers(const ok_index *idx) { return idx->blacklisted_kmers; }

typedef struct {
    const sketcher_t *sk;
    const uint8_t *seq;
    int64_t w0, w1;
    kvec_t out;
} chunk_job_t;

typedef struct {
    chunk_job_t *jobs;
    int64_t n_jobs;
    int64_t next;
    pthread_mutex_t mu;
} chunk_pool_t;

static void *chunk_worker(void *arg) {
    chunk_pool_t *pool = (chunk_pool_t *)arg;
    for (;;) {
        pthread_mutex_lock(&pool->mu);
        const int64_t j = pool->next++;
        pthread_mutex_unlock(&pool->mu);
        if (j >= pool->n_jobs) break;
        chunk_job_t *job = &pool->jobs[j];
        const int32_t k = job->sk->k;
        /* src/index.rs:477-483 -- window chunk [w0,w1) reads bases
         * seq[w0 .. w1+k-1], sketched at offset w0. */
        sketch_slice(job->sk, job->seq + job->w0, job->w1 - job->w0 + (k - 1), (int32_t)job->w0,
                     &job->out);
    }
    return NULL;
}

/* A parallel loop: fn(ctx, lo, hi, t) over n_threads contiguous slices of
 * [0, n).  Index construction only; the mapping path is not threaded here. */
typedef void (*range_fn)(void *ctx, int64_t lo, int64_t hi, int t);
typedef struct {
    range_fn fn;
    void *ctx;
    int64_t lo, hi;
    int t;
} range_job_t;

static void *range_worker(void *arg) {
    range_job_t *j = (range_job_t *)arg;
    j->fn(j->ctx, j->lo, j->hi, j->t);
    return NULL;
}

static void par_for(int n_threads, int64_t n, range_fn fn, void *ctx) {
    if (n_threads < 1) n_threads = 1;
    if (n_threads > 256) n_threads = 256;
    pthread_t th[256];
    range_job_t jobs[256];
    for (int t = 0; t < n_threads; t++) {
        jobs[t] = (range_job_t){fn, ctx, n * t / n_threads, n * (t + 1) / n_threads, t};
        pthread_create(&th[t], NULL, range_worker, &jobs[t]);
    }
    for (int t = 0; t < n_threads; t++) pthread_join(th[t], NULL);
}

/* Stable LSD radix sort of (key, val) pairs by key, 16-bit digits: each
 * thread histograms its slice, the counts are scanned in (digit, thread)
 * order, and each thread scatters its slice in order -- so equal keys keep
 * their input order. */
typedef struct {
    uint64_t *key, *k2;
    uint32_t *val, *v2;
    int shift, nt;
    int64_t *cnt; /* nt x 65536 */
} radix_ctx;

static void radix_count(void *c, int64_t lo, int64_t hi, int t) {
    radix_ctx *x = (radix_ctx *)c;
    int64_t *cnt = x->cnt + (size_t)t * 65536;
    memset(cnt, 0, 65536 * sizeof(int64_t));
    for (int64_t i = lo; i < hi; i++) cnt[(x->key[i] >> x->shift) & 0xffff]++;
}

static void radix_scatter(void *c, int64_t lo, int64_t hi, int t) {
    radix_ctx *x = (radix_ctx *)c;
    int64_t *cnt = x->cnt + (size_t)t * 65536;
    for (int64_t i = lo; i < hi; i++) {
        const int64_t o = cnt[(x->key[i] >> x->shift) & 0xffff]++;
        x->k2[o] = x->key[i];
        x->v2[o] = x->val[i];
    }
}

/* Sorts (*key, *val) in place of the pointers: the sorted arrays may end up
 * in the scratch buffers, which the caller then owns (the others are freed). */
static void radix_sort_kv(uint64_t **keyp, uint32_t **valp, int64_t n, int n_threads) {
    if (n <= 1) return;
    if (n_threads < 1) n_threads = 1;
    if (n < (1 << 20)) n_threads = 1;
    radix_ctx x = {*keyp, xmalloc((size_t)n * 8), *valp, xmalloc((size_t)n * 4), 0, n_threads,
                   xmalloc(sizeof(int64_t) * 65536 * (size_t)n_threads)};
    uint64_t all_or = 0;
    for (int64_t i = 0; i < n; i++) all_or |= x.key[i];
    for (int shift = 0; shift < 64; shift += 16) {
        if ((all_or >> shift) == 0) break;
        x.shift = shift;
        par_for(n_threads, n, radix_count, &x);
        int64_t sum = 0;
        for (int d = 0; d < 65536; d++)
            for (int t = 0; t < n_threads; t++) {
                int64_t *c = &x.cnt[(size_t)t * 65536 + d];
                const int64_t v = *c;
                *c = sum;
                sum += v;
            }
        par_for(n_threads, n, radix_scatter, &x);
        uint64_t *tk = x.key;
        x.key = x.k2;
        x.k2 = tk;
        uint32_t *tv = x.val;
        x.val = x.v2;
        x.v2 = tv;
    }
    free(x.k2);
    free(x.v2);
    free(x.cnt);
    *keyp = x.key;
    *valp = x.val;
}

typedef struct {
    ok_index *idx;
    const uint32_t *gid;
} hits_ctx;

static void fill_hits(void *c, int64_t lo, int64_t hi, int t) {
    hits_ctx *x = (hits_ctx *)c;
    ok_index *idx = x->idx;
    const int32_t n_segments = idx->n_segments;
    for (int64_t i = lo; i < hi; i++) {
        const int64_t g = x->gid[i];
        /* segment of global k-mer g */
        int32_t s = 0, top = n_segments - 1;
        while (s < top) {
            const int32_t mid = (s + top + 1) / 2;
            if (idx->kstart[mid] <= g) s = mid; else top = mid - 1;
        }
        idx->hits[i].r = idx->t_r[g];
        idx->hits[i].tpos = (int32_t)(g - idx->kstart[s]);
        idx->hits[i].segm_id = s;
        idx->hits[i].strand = idx->t_strand[g];
    }
}

/* Inserts the runs of equal keys that start in [lo, hi) (a run belongs to the
 * slice holding its first element); slots are claimed by compare-and-swap on
 * the count, so a key's slot depends on timing but every lookup's answer
 * does not (keys are distinct). */
typedef struct {
    ok_index *idx;
    const uint64_t *key;
    int64_t total;
    int32_t max_matches;
    int64_t kmers[256], hits[256], black[256];
} insert_ctx;

static void insert_runs(void *c, int64_t lo, int64_t hi, int t) {
    insert_ctx *x = (insert_ctx *)c;
    ok_index *idx = x->idx;
    const uint64_t *key = x->key;
    int64_t i = lo;
    while (i > 0 && i < hi && key[i] == key[i - 1]) i++;
    int64_t nk = 0, nh = 0, nb = 0;
    while (i < hi) {
        int64_t e = i + 1;
        while (e < x->total && key[e] == key[i]) e++;
        const int64_t cnt = e - i;
        nk++;
        nh += cnt;
        if (x->max_matches > 0 && cnt > x->max_matches && cnt > 1) {
            nb++;
        } else {
            uint64_t pos = slot_of(key[i], idx->mask);
            for (;;) {
                uint32_t zero = 0;
                if (__atomic_compare_exchange_n(&idx->table[pos].count, &zero, (uint32_t)cnt, 0, __ATOMIC_RELAXED,
                                                __ATOMIC_RELAXED))
                    break;
                pos = (pos + 1) & idx->mask;
            }
            idx->table[pos].key = key[i];
            idx->table[pos].start = (uint32_t)i;
        }
        i = e;
    }
    x->kmers[t] = nk;
    x->hits[t] = nh;
    x->black[t] = nb;
}

/* src/index.rs:384-663 -- build_index.  Segments get ids in file order; the
 * per-hash hit lists are sorted by (segm_id, r) (:295-315).  With -M, a hash
 * whose total count c exceeds M is dropped: insert_hit keeps min(c-1, M+1)
 * hits in h2multi (:286-288), finalize_shard adds the single hit back
 * (:297-299), and erase_frequent_kmers removes lists longer than M (:360-376),
 * so a key survives iff c <= M (c = 1 is never in h2multi and always stays). */
ok_index *ok_index_build(int32_t k, double h_frac, int32_t max_matches, int32_t n_segments,
                         const char *const *names, const uint8_t *const *seqs,
                         const int64_t *lens, int32_t n_threads) {
    sketcher_t sk;
    sketcher_init(&sk, k, h_frac);
    ok_index *idx = calloc(1, sizeof(ok_index));
    idx->k = k;
    idx->h_frac = h_frac;
    idx->n_segments = n_segments;
    idx->names = xmalloc(sizeof(char *) * (size_t)(n_segments > 0 ? n_segments : 1));
    idx->sz = xmalloc(sizeof(int32_t) * (size_t)(n_segments > 0 ? n_segments : 1));
    idx->kstart = xmalloc(sizeof(int64_t) * (size_t)(n_segments + 1));
    if (n_threads < 1) n_threads = 1;

    /* Chunk every segment into window runs (src/index.rs:145-153). */
    const int64_t CHUNK = 1 << 21;
    int64_t n_jobs = 0;
    for (int32_t s = 0; s < n_segments; s++) {
        const int64_t nw = lens[s] >= k ? lens[s] - k + 1 : 0;
        n_jobs += nw > 0 ? (nw + CHUNK - 1) / CHUNK : 0;
    }
    chunk_job_t *jobs = calloc((size_t)(n_jobs > 0 ? n_jobs : 1), sizeof(chunk_job_t));
    int64_t *seg_first_job = xmalloc(sizeof(int64_t) * (size_t)(n_segments + 1));
    n_jobs = 0;
    for (int32_t s = 0; s < n_segments; s++) {
        seg_first_job[s] = n_jobs;
        const int64_t nw = lens[s] >= k ? lens[s] - k + 1 : 0;
        for (int64_t w0 = 0; w0 < nw; w0 += CHUNK) {
            jobs[n_jobs].sk = &sk;
            jobs[n_jobs].seq = seqs[s];
            jobs[n_jobs].w0 = w0;
            jobs[n_jobs].w1 = w0 + CHUNK < nw ? w0 + CHUNK : nw;
            n_jobs++;
        }
    }
    seg_first_job[n_segments] = n_jobs;
    chunk_pool_t pool = {jobs, n_jobs, 0, PTHREAD_MUTEX_INITIALIZER};
    {
        pthread_t th[256];
        const int nt = n_threads > 256 ? 256 : n_threads;
        for (int t = 0; t < nt; t++) pthread_create(&th[t], NULL, chunk_worker, &pool);
        for (int t = 0; t < nt; t++) pthread_join(th[t], NULL);
    }
    int64_t total = 0;
    for (int64_t j = 0; j < n_jobs; j++) total += jobs[j].out.n;
    idx->n_kmers = total;
    idx->t_h = xmalloc(sizeof(uint64_t) * (size_t)(total > 0 ? total : 1));
    idx->t_r = xmalloc(sizeof(int32_t) * (size_t)(total > 0 ? total : 1));
    idx->t_strand = xmalloc((size_t)(total > 0 ? total : 1));
    int64_t at = 0;
    for (int32_t s = 0; s < n_segments; s++) {
        idx->names[s] = strdup(names[s]);
        idx->sz[s] = (int32_t)lens[s];
        idx->kstart[s] = at;
        for (int64_t j = seg_first_job[s]; j < seg_first_job[s + 1]; j++) {
            for (int64_t i = 0; i < jobs[j].out.n; i++) {
                idx->t_h[at] = jobs[j].out.v[i].h;
                idx->t_r[at] = jobs[j].out.v[i].r;
                idx->t_strand[at] = jobs[j].out.v[i].strand;
                at++;
            }
            free(jobs[j].out.v);
        }
    }
    idx->kstart[n_segments] = at;
    free(jobs);
    free(seg_first_job);

    /* Group hits by hash; a stable sort keeps (segm_id, tpos) order, which is
     * (segm_id, r) order, inside each group. */
    uint64_t *key = xmalloc(sizeof(uint64_t) * (size_t)(total > 0 ? total : 1));
    uint32_t *gid = xmalloc(sizeof(uint32_t) * (size_t)(total > 0 ? total : 1));
    for (int64_t i = 0; i < total; i++) {
        key[i] = idx->t_h[i];
        gid[i] = (uint32_t)i;
    }
    radix_sort_kv(&key, &gid, total, n_threads);
    idx->hits = xmalloc(sizeof(hit_t) * (size_t)(total > 0 ? total : 1));
    {
        hits_ctx hc = {idx, gid};
        par_for(total < (1 << 20) ? 1 : n_threads, total, fill_hits, &hc);
    }
    int64_t n_keys = 0;
    for (int64_t i = 0; i < total; i++)
        if (i == 0 || key[i] != key[i - 1]) n_keys++;
    uint64_t cap = 16;
    while (cap < (uint64_t)n_keys * 2) cap <<= 1;
    idx->table = calloc(cap, sizeof(hslot_t));
    idx->mask = cap - 1;
    {
        const int nt = total < (1 << 20) ? 1 : (n_threads > 256 ? 256 : n_threads);
        insert_ctx *ic = calloc(1, sizeof(insert_ctx));
        ic->idx = idx;
        ic->key = key;
        ic->total = total;
        ic->max_matches = max_matches;
        par_for(nt, total, insert_runs, ic);
        for (int t = 0; t < nt; t++) {
            idx->indexed_kmers += ic->kmers[t];
            idx->indexed_hits += ic->hits[t];
            idx->blacklisted_kmers += ic->black[t];
        }
        free(ic);
    }
    free(key);
    free(gid);
    return idx;
}

void ok_index_free(ok_index *idx) {
    if (!idx) return;
    for (int32_t s = 0; s < idx->n_segments; s++) free(idx->names[s]);
    free(idx->names);
    free(idx->sz);
    free(idx->kstart);
    free(idx->t_h);
    free(idx->t_r);
    free(idx->t_strand);
    free(idx->hits);
    free(idx->table);
    free(idx);
}

/* ------------------------------------------------------------------------- */
/* Per-read mapping: src/shmap/{mod,seeding,pruning,scoring}.rs, buckets.rs,  */
/* mapping.rs                                                                */
/* ------------------------------------------------------------------------- */

typedef struct {
    uint64_t h;
    int32_t r;        /* kmer.r of the run's last element (smallest r) */
    uint8_t strand;   /* its strand */
    int32_t hits_in_t;
    int32_t occs_in_p;
    int32_t seed_num;
    int32_t pm_off;   /* pmatches = sorted_p[pm_off .. pm_off+occs).r, decreasing */
} seed_t;             /* src/types.rs:132-139 */

typedef struct {
    int32_t segm_id, b;
    int32_t matches, codirection, r_min, r_max;
} bentry_t; /* src/buckets.rs:107-114 */

typedef struct {
    int32_t i, seeds, matches, codirection, r_min, r_max;
} bcontent_t; /* src/types.rs:230-250 */

typedef struct {
    int32_t segm_id, b;
    bcontent_t c;
} sbucket_t;

typedef struct {
    int32_t p_start, p_end;
    int32_t strand;
    int32_t segm_id;
    int32_t t_l, t_r;
    int32_t s_sz, intersection, same_strand;
    double j, j2, sh;
    int32_t b_segm, b_b;
} mapping_t; /* src/mapping.rs:12-25, 110-138 (fields on the mapping path) */

static void mapping_default(mapping_t *m) { /* src/mapping.rs:28-43, 140-160 */
    m->p_start = m->p_end = -1;
    m->strand = '?';
    m->segm_id = -1;
    m->t_l = m->t_r = -1;
    m->s_sz = -1;
    m->intersection = -1;
    m->same_strand = -1;
    m->j = -1.0;
    m->j2 = -1.0;
    m->sh = -1.0;
    m->b_segm = m->b_b = -1;
}

/* src/mapping.rs:299-327 */
static void mapping_update(mapping_t *m, int32_t p_start, int32_t p_end, int32_t t_l, int32_t t_r,
                           int32_t segm_id, int32_t intersection, double score, int32_t same,
                           int32_t sz) {
    if (score > m->j) {
        m->p_start = p_start;
        m->p_end = p_end;
        m->strand = same > 0 ? '+' : '-';
        m->segm_id = segm_id;
        m->t_l = t_l;
        m->t_r = t_r;
        m->s_sz = sz;
        m->same_strand = same;
        m->intersection = intersection;
        m->j = score;
    }
}

/* src/mapping.rs:412-420 */
double ok_overlap(int32_t sa, int32_t la, int32_t ra, int32_t sb, int32_t lb, int32_t rb) {
    if (sa != sb) return -0.0;
    int32_t cap = (ra < rb ? ra : rb) - (la > lb ? la : lb);
    if (cap < 0) cap = 0;
    const int32_t cup = (ra > rb ? ra : rb) - (la < lb ? la : lb);
    return (double)cap / (double)cup;
}

/* src/shmap/pruning.rs:10-13 */
double ok_hseed(int32_t p, int32_t seeds, int32_t matches) {
    return 1.0 - (double)(seeds - matches) / (double)p;
}

/* src/shmap/scoring.rs:133-154 -- patience LIS length (strict increase). */
int32_t ok_lis_len(const int32_t *x, int32_t n) {
    int32_t *lis = xmalloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    int64_t len = 0;
    for (int32_t i = 0; i < n; i++) {
        int64_t l = -1, r = len;
        while (l + 1 < r) {
            const int64_t mid = (l + r) / 2;
            if (lis[mid] < x[i]) l = mid; else r = mid;
        }
        if (r == len) lis[len++] = x[i]; else lis[r] = x[i];
    }
    free(lis);
    return (int32_t)len;
}

/* src/mapping.rs:341-354 */
int32_t ok_calc_mapq(double j, double j2, int32_t same_strand, int32_t intersection,
                     double theta2, double min_diff, double *j2_out) {
    if (j2 < theta2) j2 = theta2;
    if (j2_out) *j2_out = j2;
    const double frac = 1.0 - j2 / j;
    const int32_t strand5 = abs(same_strand) < intersection / 2 ? 5 : 60;
    if (frac > min_diff) return 60 < strand5 ? 60 : strand5;
    return 0;
}

typedef struct {
    sketcher_t sk;
    kvec_t p;
    seed_t *seeds;    int64_t seeds_cap;   /* p_unique, rarest first */
    int32_t *ht;      int64_t ht_cap;      /* p_ht: hash -> seed index in hash order + 1 */
    seed_t *by_hash;  int64_t by_hash_cap; /* seeds indexed by seed_num */
    int32_t *diff;    int64_t diff_cap;    /* diff_hist */
    bentry_t *ent;    int64_t ent_cap;  int64_t n_ent;
    sbucket_t *sb;    int64_t sb_cap;
    int32_t *ppos;    int64_t ppos_cap;
    struct { int32_t idx; mapping_t m; } *memo; int64_t memo_cap; int64_t n_memo;
    uint64_t ht_mask;
    int64_t prune_hits;  /* hit records the prune pass read (a measurement, not a result) */
} ws_t;

static int cmp_kmer_h_rdesc(const void *a, const void *b) {
    const ok_kmer *x = a, *y = b;
    if (x->h != y->h) return x->h < y->h ? -1 : 1;
    return (x->r > y->r) ? -1 : (x->r < y->r); /* r descending */
}
static int cmp_seed_hits(const void *a, const void *b) {
    const seed_t *x = a, *y = b; /* stable sort by hits_in_t == (hits, seed_num) */
    if (x->hits_in_t != y->hits_in_t) return x->hits_in_t < y->hits_in_t ? -1 : 1;
    return (x->seed_num > y->seed_num) - (x->seed_num < y->seed_num);
}
static int cmp_bentry_loc(const void *a, const void *b) {
    const bentry_t *x = a, *y = b;
    if (x->segm_id != y->segm_id) return x->segm_id < y->segm_id ? -1 : 1;
    return (x->b > y->b) - (x->b < y->b);
}
static int cmp_sbucket_order(const void *a, const void *b) {
    /* src/buckets.rs:685-724: key (u32::MAX - matches) << 32 | index, i.e.
     * matches descending, then location-sorted index ascending.  The
     * location-sorted index is held in c.i during the sort. */
    const sbucket_t *x = a, *y = b;
    if (x->c.matches != y->c.matches) return x->c.matches > y->c.matches ? -1 : 1;
    return (x->c.i > y->c.i) - (x->c.i < y->c.i);
}

static const seed_t *pht_get(const ws_t *w, uint64_t h) {
    uint64_t pos = slot_of(h, w->ht_mask);
    for (;;) {
        const int32_t v = w->ht[pos];
        if (v == 0) return NULL;
        if (w->by_hash[v - 1].h == h) return &w->by_hash[v - 1];
        pos = (pos + 1) & w->ht_mask;
    }
}

typedef struct {
    const ok_index *idx;
    const ok_params *prm;
    int32_t halflen;
    int64_t *prune_hits;  /* the workspace's count of hit records the prune pass reads */
} ctx_t;

static inline int32_t bucket_begin(const ctx_t *c, int32_t b) { return b * c->halflen; }
static inline int32_t bucket_end(const ctx_t *c, int32_t b) { return (b + 2) * c->halflen; }

/* src/buckets.rs:642-662 (add_to_pos) */
static void add_to_pos(ws_t *w, const ctx_t *c, const hit_t *hit, int32_t matches, int32_t codir,
                       int32_t r_min, int32_t r_max) {
    const int32_t pos = c->prm->abs_pos ? hit->r : hit->tpos;
    const int32_t b = pos / c->halflen;
    GROW(w->ent, w->ent_cap, w->n_ent + 2);
    w->ent[w->n_ent++] = (bentry_t){hit->segm_id, b, matches, codir, r_min, r_max};
    if (b > 0) w->ent[w->n_ent++] = (bentry_t){hit->segm_id, b - 1, matches, codir, r_min, r_max};
}

/* src/buckets.rs:664-674 (add_to_bucket) */
static void add_to_bucket(ws_t *w, int32_t segm_id, int32_t b, int32_t matches, int32_t codir,
                          int32_t r_min, int32_t r_max) {
    GROW(w->ent, w->ent_cap, w->n_ent + 1);
    w->ent[w->n_ent++] = (bentry_t){segm_id, b, matches, codir, r_min, r_max};
}

static inline void bc_add(bcontent_t *a, int32_t matches, int32_t codir, int32_t r_min,
                          int32_t r_max) { /* src/types.rs:265-273 */
    a->matches += matches;
    a->codirection += codir;
    if (r_min < a->r_min) a->r_min = r_min;
    if (r_max > a->r_max) a->r_max = r_max;
}
static inline bcontent_t bc_default(void) {
    bcontent_t b = {-1, 0, 0, 0, INT32_MAX, -1};
    return b;
}

/* src/shmap/seeding.rs:13-21 */
static void flush_slot(ws_t *w, int32_t segm_id, int32_t bb, const bcontent_t *acc, int32_t occs) {
    add_to_bucket(w, segm_id, bb, acc->matches < occs ? acc->matches : occs, acc->codirection,
                  acc->r_min, acc->r_max);
}

/* src/shmap/seeding.rs:98-212 -- match_seeds.  Returns via the out params the
 * consumed seed index i and the consumed occurrence total. */
static void match_seeds(ws_t *w, const ctx_t *c, const seed_t *pu, int32_t u, int32_t S,
                        int32_t *bi, int32_t *bseeds, int64_t *seed_matches_out,
                        int64_t *max_seed_matches) {
    const ok_index *idx = c->idx;
    int32_t seed_matches = 0;
    const int32_t halflen = c->halflen;
    const int ap = c->prm->abs_pos;
    while (*bi < u && *bseeds < S) {
        const seed_t *seed = &pu[*bi];
        if (seed->hits_in_t > 0) {
            seed_matches += seed->hits_in_t;
            if (seed->hits_in_t > *max_seed_matches) *max_seed_matches = seed->hits_in_t;
            const hslot_t *slot = index_find(idx, seed->h);
            if (seed->hits_in_t == 1) {
                const hit_t *hit = &idx->hits[slot->start];
                add_to_pos(w, c, hit, 1, hit->strand == seed->strand ? 1 : -1, hit->r, hit->r);
            } else {
                const int32_t occs = seed->occs_in_p;
                int32_t cur_sid = -1, b = 0, b_hi = 0, base = 0;
                bcontent_t acc[2] = {bc_default(), bc_default()};
                for (uint32_t t = 0; t < slot->count; t++) {
                    const hit_t *hit = &idx->hits[slot->start + t];
                    const int32_t pos = ap ? hit->r : hit->tpos;
                    const int32_t codir = hit->strand == seed->strand ? 1 : -1;
                    if (hit->segm_id != cur_sid) {
                        for (int j = 0; j < 2; j++) {
                            if (acc[j].matches > 0) {
                                flush_slot(w, cur_sid, base + j, &acc[j], occs);
                                acc[j] = bc_default();
                            }
                        }
                        cur_sid = hit->segm_id;
                        b = pos / halflen;
                        b_hi = (b + 1) * halflen;
                        base = b - 1 > 0 ? b - 1 : 0;
                    } else {
                        while (pos >= b_hi) {
                            b += 1;
                            b_hi += halflen;
                        }
                        const int32_t new_base = b - 1 > 0 ? b - 1 : 0;
                        if (new_base > base) {
                            if (acc[0].matches > 0) flush_slot(w, cur_sid, base, &acc[0], occs);
                            if (new_base == base + 1) {
                                acc[0] = acc[1];
                            } else {
                                if (acc[1].matches > 0)
                                    flush_slot(w, cur_sid, base + 1, &acc[1], occs);
                                acc[0] = bc_default();
                            }
                            acc[1] = bc_default();
                            base = new_base;
                        }
                    }
                    bc_add(&acc[b - base], 1, codir, hit->r, hit->r);
                    if (b > 0) bc_add(&acc[b - 1 - base], 1, codir, hit->r, hit->r);
                }
                for (int j = 0; j < 2; j++)
                    if (acc[j].matches > 0) flush_slot(w, cur_sid, base + j, &acc[j], occs);
            }
        }
        *bseeds += seed->occs_in_p;
        *bi += 1;
    }
    *seed_matches_out = seed_matches;
}

/* src/buckets.rs:603-629, 685-724 -- merge_entries (sparse path) and
 * get_sorted_buckets.  Returns the number of sorted buckets in w->sb. */
static int64_t sorted_buckets(ws_t *w, int32_t prop_i, int32_t prop_seeds) {
    int64_t n = 0;
    if (w->n_ent > 0) {
        qsort(w->ent, (size_t)w->n_ent, sizeof(bentry_t), cmp_bentry_loc);
        int64_t wr = 0;
        for (int64_t rd = 1; rd < w->n_ent; rd++) {
            if (w->ent[rd].segm_id == w->ent[wr].segm_id && w->ent[rd].b == w->ent[wr].b) {
                bentry_t *e = &w->ent[wr];
                const bentry_t *o = &w->ent[rd];
                e->matches += o->matches;
                e->codirection += o->codirection;
                if (o->r_min < e->r_min) e->r_min = o->r_min;
                if (o->r_max > e->r_max) e->r_max = o->r_max;
            } else {
                w->ent[++wr] = w->ent[rd];
            }
        }
        n = wr + 1;
    }
    GROW(w->sb, w->sb_cap, n);
    for (int64_t i = 0; i < n; i++) {
        const bentry_t *e = &w->ent[i];
        w->sb[i].segm_id = e->segm_id;
        w->sb[i].b = e->b;
        w->sb[i].c = (bcontent_t){(int32_t)i, prop_seeds, e->matches, e->codirection, e->r_min,
                                  e->r_max};
    }
    qsort(w->sb, (size_t)n, sizeof(sbucket_t), cmp_sbucket_order);
    for (int64_t i = 0; i < n; i++) w->sb[i].c.i = prop_i;
    return n;
}

/* src/shmap/pruning.rs:18-80 -- matches_in_bucket */
static void matches_in_bucket(const ctx_t *c, int32_t bsegm, int32_t bb, bcontent_t *bucket,
                              const seed_t *s) {
    const ok_index *idx = c->idx;
    const int ap = c->prm->abs_pos;
    bucket->seeds += s->occs_in_p;
    if (s->hits_in_t == 0) return;
    const hslot_t *slot = index_find(idx, s->h);
    const int32_t begin = bucket_begin(c, bb), end = bucket_end(c, bb);
    if (s->hits_in_t == 1) {
        *c->prune_hits += 1;
        const hit_t *hit = &idx->hits[slot->start];
        const int32_t pos = ap ? hit->r : hit->tpos;
        if (hit->segm_id == bsegm && begin <= pos && pos < end) {
            bucket->matches += 1;
            bucket->codirection += hit->strand == s->strand ? 1 : -1;
            if (hit->r < bucket->r_min) bucket->r_min = hit->r;
            if (hit->r > bucket->r_max) bucket->r_max = hit->r;
        }
        return;
    }
    const hit_t *hits = &idx->hits[slot->start];
    const int64_t n = slot->count;
    /* partition_point: first hit not (segm < bsegm || (segm == bsegm && pos < begin)) */
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = (lo + hi) / 2;
        const hit_t *h = &hits[mid];
        *c->prune_hits += 1;
        const int32_t pos = ap ? h->r : h->tpos;
        const int before = h->segm_id != bsegm ? h->segm_id < bsegm : pos < begin;
        if (before) lo = mid + 1; else hi = mid;
    }
    int32_t matches = 0;
    for (int64_t t = lo; t < n; t++) {
        const hit_t *h = &hits[t];
        *c->prune_hits += 1;
        const int32_t pos = ap ? h->r : h->tpos;
        if (!(h->segm_id == bsegm && pos < end)) break;
        matches++;
        bucket->codirection += h->strand == s->strand ? 1 : -1;
        if (h->r < bucket->r_min) bucket->r_min = h->r;
        if (h->r > bucket->r_max) bucket->r_max = h->r;
    }
    bucket->matches += matches < s->occs_in_p ? matches : s->occs_in_p;
}

/* src/shmap/pruning.rs:95-119 -- seed_heuristic_pass */
static int seed_heuristic_pass(const ctx_t *c, const seed_t *pu, int32_t u, int32_t m,
                               int32_t bsegm, int32_t bb, bcontent_t *bucket, double *sh,
                               double thr) {
    if (!c->prm->no_bucket_pruning) {
        for (;;) {
            *sh = ok_hseed(m, bucket->seeds, bucket->matches);
            if (*sh < thr) return 0;
            if (bucket->i >= u) break;
            matches_in_bucket(c, bsegm, bb, bucket, &pu[bucket->i]);
            bucket->i += 1;
        }
    }
    return 1;
}

/* src/shmap/scoring.rs:189-291 -- best_fixed_length (Containment/Jaccard) */
static void best_fixed_length(ws_t *w, const ctx_t *c, int32_t segm_id, int32_t from, int32_t to,
                              int32_t p_sz, int32_t m, int32_t metric, mapping_t *best) {
    const ok_index *idx = c->idx;
    const int64_t ks = idx->kstart[segm_id];
    const int32_t tlen = (int32_t)(idx->kstart[segm_id + 1] - ks);
    const uint64_t *th = idx->t_h + ks;
    const int32_t *tr_ = idx->t_r + ks;
    const uint8_t *ts = idx->t_strand + ks;
    const int ap = c->prm->abs_pos;
    int32_t l = from > 0 ? from : 0;
    int32_t r = l;
    const int32_t end = tlen < to ? tlen : to;
    int32_t intersection = 0, same = 0;
    mapping_default(best);
    while (l < end) {
        const int32_t tl_r = tr_[l];
        while (r < end) {
            const int in_window = ap ? tr_[r] < tl_r + m : r < l + m;
            if (!in_window) break;
            const seed_t *seed = pht_get(w, th[r]);
            if (seed) {
                same += seed->strand == ts[r] ? 1 : -1;
                int32_t *cnt = &w->diff[seed->seed_num];
                *cnt -= 1;
                if (*cnt >= 0) intersection += 1;
            }
            r++;
        }
        const int32_t s_kmers = r - l;
        double score;
        if (metric == OK_JACCARD)
            score = (double)intersection / (double)(m + s_kmers - intersection);
        else
            score = (double)intersection / (double)m;
        if (l < r && score > best->j) {
            mapping_update(best, 0, p_sz - 1, tr_[l], tr_[r - 1], segm_id, intersection, score,
                           same, tr_[r - 1] - tr_[l] + 1);
        }
        const seed_t *seed = pht_get(w, th[l]);
        if (seed) {
            same -= seed->strand == ts[l] ? 1 : -1;
            int32_t *cnt = &w->diff[seed->seed_num];
            *cnt += 1;
            if (*cnt >= 1) intersection -= 1;
        }
        l++;
    }
}

/* src/shmap/scoring.rs:114-170 -- lcs */
static int32_t lcs(ws_t *w, const ctx_t *c, int32_t segm_id, int32_t b) {
    const ok_index *idx = c->idx;
    const int64_t ks = idx->kstart[segm_id];
    const int32_t tlen = (int32_t)(idx->kstart[segm_id + 1] - ks);
    const int32_t begin = bucket_begin(c, b);
    int32_t end = bucket_end(c, b);
    if (tlen < end) end = tlen;
    int64_t n = 0;
    for (int32_t l = begin; l < end; l++) {
        const seed_t *seed = pht_get(w, idx->t_h[ks + l]);
        if (seed) {
            GROW(w->ppos, w->ppos_cap, n + seed->occs_in_p);
            for (int32_t q = 0; q < seed->occs_in_p; q++) w->ppos[n++] = w->p.v[seed->pm_off + q].r;
        }
    }
    const int32_t fwd = ok_lis_len(w->ppos, (int32_t)n);
    for (int64_t i = 0, j = n - 1; i < j; i++, j--) {
        const int32_t t = w->ppos[i];
        w->ppos[i] = w->ppos[j];
        w->ppos[j] = t;
    }
    const int32_t rev = ok_lis_len(w->ppos, (int32_t)n);
    return fwd > rev ? fwd : rev;
}

/* src/shmap/scoring.rs:296-359 -- find_best_mapping */
static void find_best_mapping(ws_t *w, const ctx_t *c, int32_t segm_id, int32_t b,
                              const bcontent_t *content, int32_t p_sz, int32_t m, int32_t lmax,
                              double sh, mapping_t *out) {
    const int32_t metric = c->prm->metric;
    if (metric == OK_BUCKET_SH) {
        mapping_default(out);
        mapping_update(out, 0, p_sz - 1, content->r_min, content->r_max, segm_id,
                       content->matches, sh, content->codirection,
                       content->r_max - content->r_min);
    } else if (metric == OK_BUCKET_LCS) {
        const int32_t lcs_cnt = lcs(w, c, segm_id, b);
        const double score = (double)lcs_cnt / (double)m;
        mapping_default(out);
        mapping_update(out, 0, p_sz - 1, content->r_min, content->r_max, segm_id,
                       content->matches, score, content->codirection,
                       content->r_max - content->r_min);
    } else {
        best_fixed_length(w, c, segm_id, bucket_begin(c, b), bucket_end(c, b), p_sz - c->prm->k,
                          lmax, metric, out);
    }
    out->b_segm = segm_id;
    out->b_b = b;
    out->sh = sh;
}

/* src/shmap/scoring.rs:367-489 -- match_rest (rarity tie-break off), with the
 * RefineCache of :38-109 (record in pass 1, replay in pass 2). */
static int match_rest(ws_t *w, const ctx_t *c, int32_t p_sz, int32_t m, int32_t lmax,
                      const seed_t *pu, int32_t u, sbucket_t *sb, int64_t nsb, double thr,
                      const mapping_t *forbidden, int replay, mapping_t *best,
                      int64_t *refined, int64_t *memo_hits, int64_t *final_buckets) {
    int have = 0;
    const int memoizable = c->prm->metric == OK_CONTAINMENT || c->prm->metric == OK_JACCARD;
    int64_t cursor = 0;
    if (!replay) w->n_memo = 0;
    for (int64_t idx = 0; idx < nsb; idx++) {
        double sh = 1.0;
        if (!seed_heuristic_pass(c, pu, u, m, sb[idx].segm_id, sb[idx].b, &sb[idx].c, &sh, thr))
            continue;
        mapping_t cand;
        int have_cand = 0;
        int memo = -1;
        if (memoizable && replay) {
            while (cursor < w->n_memo && w->memo[cursor].idx < idx) cursor++;
            if (cursor < w->n_memo && w->memo[cursor].idx == idx) memo = (int)cursor;
        }
        if (memo >= 0) {
            *memo_hits += 1;
            if (w->memo[memo].m.j > thr) {
                cand = w->memo[memo].m;
                cand.sh = sh;
                have_cand = 1;
            }
        } else {
            *refined += 1;
            mapping_t scored;
            find_best_mapping(w, c, sb[idx].segm_id, sb[idx].b, &sb[idx].c, p_sz, m, lmax, sh,
                              &scored);
            if (memoizable && !replay) {
                GROW(w->memo, w->memo_cap, w->n_memo + 1);
                w->memo[w->n_memo].idx = (int32_t)idx;
                w->memo[w->n_memo].m = scored;
                w->n_memo++;
            }
            if (scored.j > thr) {
                cand = scored;
                have_cand = 1;
            }
        }
        if (have_cand) {
            *final_buckets += 1;
            const int ok = !forbidden ||
                           ok_overlap(cand.segm_id, cand.t_l, cand.t_r, forbidden->segm_id,
                                      forbidden->t_l, forbidden->t_r) < c->prm->max_overlap;
            if (ok && (!have || cand.j > best->j)) {
                if (cand.j > thr) thr = cand.j;
                *best = cand;
                have = 1;
            }
        }
    }
    return have;
}

static void ws_init(ws_t *w, const ok_params *p) {
    memset(w, 0, sizeof(*w));
    sketcher_init(&w->sk, p->k, p->h_frac);
}
static void ws_free(ws_t *w) {
    free(w->p.v);
    free(w->seeds);
    free(w->ht);
    free(w->by_hash);
    free(w->diff);
    free(w->ent);
    free(w->sb);
    free(w->ppos);
    free(w->memo);
}

/* src/shmap/mod.rs:653-971 -- map_read */
static int map_read_ws(ws_t *w, const ok_index *idx, const ok_params *prm, const uint8_t *seq,
                       int32_t len, ok_result *out) {
    memset(out, 0, sizeof(*out));
    out->read_len = len;
    out->b2_segm = out->b2_b = -1;
    w->p.n = 0;
    sketch_slice(&w->sk, seq, len, 0, &w->p); /* :669-672 */
    const int32_t m = (int32_t)w->p.n;
    out->m = m;

    /* unique_elements_with_info -- src/shmap/seeding.rs:26-77 */
    qsort(w->p.v, (size_t)m, sizeof(ok_kmer), cmp_kmer_h_rdesc);
    int32_t u = 0;
    GROW(w->seeds, w->seeds_cap, m);
    GROW(w->by_hash, w->by_hash_cap, m);
    for (int32_t ppos = 0, gs = 0; ppos < m; ppos++) {
        if (ppos + 1 == m || w->p.v[ppos].h != w->p.v[ppos + 1].h) {
            seed_t *s = &w->by_hash[u];
            s->h = w->p.v[ppos].h;
            s->r = w->p.v[ppos].r;
            s->strand = w->p.v[ppos].strand;
            s->hits_in_t = ok_index_count(idx, s->h);
            s->occs_in_p = ppos - gs + 1;
            s->seed_num = u;
            s->pm_off = gs;
            u++;
            gs = ppos + 1;
        }
    }
    out->u = u;
    memcpy(w->seeds, w->by_hash, sizeof(seed_t) * (size_t)u);
    qsort(w->seeds, (size_t)u, sizeof(seed_t), cmp_seed_hits); /* :71 stable by hits_in_t */

    /* p_ht and diff_hist -- src/shmap/mod.rs:682-692 */
    uint64_t cap = 16;
    while (cap < (uint64_t)u * 2) cap <<= 1;
    GROW(w->ht, w->ht_cap, (int64_t)cap);
    memset(w->ht, 0, sizeof(int32_t) * cap);
    w->ht_mask = cap - 1;
    GROW(w->diff, w->diff_cap, u);
    int64_t possible = 0;
    for (int32_t i = 0; i < u; i++) {
        uint64_t pos = slot_of(w->by_hash[i].h, w->ht_mask);
        while (w->ht[pos] != 0) pos = (pos + 1) & w->ht_mask;
        w->ht[pos] = i + 1;
        w->diff[i] = w->by_hash[i].occs_in_p;
        possible += w->by_hash[i].hits_in_t;
    }
    out->possible_matches = (int32_t)possible; /* summed as i32 (:689) */

    const int32_t lmax = m;                                       /* :726 */
    const double theta = prm->theta;
    const double theta2 = theta - prm->min_diff;                  /* :728 */
    const int32_t S = prm->one_sweep ? u : (int32_t)((1.0 - theta2) * (double)m) + 1; /* :732-736 */
    out->S = S;
    ctx_t c = {idx, prm, prm->abs_pos ? len : m, &w->prune_hits};  /* :738-742 */
    if (c.halflen < MIN_HALFLEN) {                                /* :772-783 */
        out->unmappable_short = 1;
        return 0;
    }

    /* match_seeds -- :785-787 */
    w->n_ent = 0;
    int32_t bi = 0, bseeds = 0;
    int64_t seed_matches = 0, max_seed_matches = 0;
    match_seeds(w, &c, w->seeds, u, S, &bi, &bseeds, &seed_matches, &max_seed_matches);
    out->seed_matches = seed_matches;
    out->total_matches = seed_matches;
    out->max_seed_matches = max_seed_matches;

    /* propagate_seeds_to_buckets + get_sorted_buckets -- :789-793 */
    const int64_t nsb = sorted_buckets(w, bi, bseeds);
    out->seeded_buckets = nsb;

    /* match_rest, two passes -- :809-861 */
    mapping_t best, best2;
    const int have = match_rest(w, &c, len, m, lmax, w->seeds, u, w->sb, nsb, theta, NULL, 0,
                                &best, &out->refined_buckets, &out->refine_memo_hits,
                                &out->final_buckets);
    int have2 = 0;
    if (have) {
        const double thr2 = best.j * (1.0 - prm->min_diff);
        have2 = match_rest(w, &c, len, m, lmax, w->seeds, u, w->sb, nsb, thr2, &best, 1, &best2,
                           &out->refined_buckets, &out->refine_memo_hits, &out->final_buckets);
    }
    if (!have) return 0;

    /* set_second_best + set_global_stats -- :873-898, src/mapping.rs:356-408 */
    int32_t i2 = -1;
    double j2 = -1.0, sig = -1.0;
    int32_t b2s = -1, b2b = -1;
    if (have2) {
        i2 = best2.intersection;
        j2 = best2.j;
        b2s = best2.b_segm;
        b2b = best2.b_b;
        const double x = i2 == -1 ? 0.0 : (double)i2;
        const double y = best.intersection == -1 ? 0.0 : (double)best.intersection;
        sig = fabs(x - y) / sqrt(x + y);
    }
    double j2p;
    out->mapq = ok_calc_mapq(best.j, j2, best.same_strand, best.intersection, theta2,
                             prm->min_diff, &j2p);
    out->mapped = 1;
    out->p_start = best.p_start;
    out->p_end = best.p_end;
    out->strand = best.strand;
    out->segm_id = best.segm_id;
    out->t_l = best.t_l;
    out->t_r = best.t_r;
    out->s_sz = best.s_sz;
    out->intersection = best.intersection;
    out->intersection2 = i2;
    out->same_strand = best.same_strand;
    out->b_segm = best.b_segm;
    out->b_b = best.b_b;
    out->b2_segm = b2s;
    out->b2_b = b2b;
    out->J = best.j;
    out->J2 = j2p;
    out->sh = best.sh;
    out->sigmas = sig;
    out->match_inefficiency = (double)(int32_t)seed_matches / (double)best.intersection;
    return 0;
}

/* unique_elements_with_info's grouping (src/shmap/seeding.rs:26-68), exposed
 * for tests: seeds in hash order, each with its occurrence count and pmatches
 * (decreasing r) concatenated into pm.  Returns the number of seeds. */
int32_t ok_unique_pmatches(const ok_kmer *p, int32_t n, uint64_t *hash, int32_t *occs, int32_t *pm) {
    ok_kmer *q = xmalloc(sizeof(ok_kmer) * (size_t)(n > 0 ? n : 1));
    memcpy(q, p, sizeof(ok_kmer) * (size_t)n);
    qsort(q, (size_t)n, sizeof(ok_kmer), cmp_kmer_h_rdesc);
    int32_t u = 0;
    for (int32_t i = 0, gs = 0; i < n; i++) {
        if (i + 1 == n || q[i].h != q[i + 1].h) {
            hash[u] = q[i].h;
            occs[u] = i - gs + 1;
            u++;
            gs = i + 1;
        }
        pm[i] = q[i].r;
    }
    free(q);
    return u;
}

int ok_map_read(const ok_index *idx, const ok_params *p, const uint8_t *seq, int32_t len,
                ok_result *out) {
    ws_t w;
    ws_init(&w, p);
    const int rc = map_read_ws(&w, idx, p, seq, len, out);
    ws_free(&w);
    return rc;
}

typedef struct {
    const ok_index *idx;
    const ok_params *p;
    int32_t n;
    const uint8_t *const *seqs;
    const int32_t *lens;
    ok_result *out;
    int32_t next;
    pthread_mutex_t mu;
    int64_t prune_hits;
} pool_t;

static void *map_worker(void *arg) {
    pool_t *pool = (pool_t *)arg;
    ws_t w;
    ws_init(&w, pool->p);
    for (;;) {
        pthread_mutex_lock(&pool->mu);
        const int32_t i = pool->next;
        pool->next += 8;
        pthread_mutex_unlock(&pool->mu);
        if (i >= pool->n) break;
        const int32_t e = i + 8 < pool->n ? i + 8 : pool->n;
        for (int32_t j = i; j < e; j++)
            map_read_ws(&w, pool->idx, pool->p, pool->seqs[j], pool->lens[j], &pool->out[j]);
    }
    pthread_mutex_lock(&pool->mu);
    pool->prune_hits += w.prune_hits;
    pthread_mutex_unlock(&pool->mu);
    ws_free(&w);
    return NULL;
}

int ok_map_reads_counted(const ok_index *idx, const ok_params *p, int32_t n, const uint8_t *const *seqs,
                         const int32_t *lens, ok_result *out, int32_t n_threads, int64_t *prune_hits) {
    if (n_threads < 1) n_threads = 1;
    if (n_threads > 256) n_threads = 256;
    pool_t pool = {idx, p, n, seqs, lens, out, 0, PTHREAD_MUTEX_INITIALIZER, 0};
    if (n_threads == 1) {
        map_worker(&pool);
    } else {
        pthread_t th[256];
        for (int t = 0; t < n_threads; t++) pthread_create(&th[t], NULL, map_worker, &pool);
        for (int t = 0; t < n_threads; t++) pthread_join(th[t], NULL);
    }
    if (prune_hits) *prune_hits = pool.prune_hits;
    return 0;
}

int ok_map_reads(const ok_index *idx, const ok_params *p, int32_t n, const uint8_t *const *seqs,
                 const int32_t *lens, ok_result *out, int32_t n_threads) {
    return ok_map_reads_counted(idx, p, n, seqs, lens, out, n_threads, NULL);
}

/* Rust `{:.N}` on f64: exact value rounded half-to-even (glibc printf agrees in
 * the default rounding mode) except that NaN and infinities are spelled
 * NaN / inf / -inf. */
static int fmt_f(char *buf, size_t cap, double v, int prec) {
    if (isnan(v)) return snprintf(buf, cap, "NaN");
    if (isinf(v)) return snprintf(buf, cap, v > 0 ? "inf" : "-inf");
    return snprintf(buf, cap, "%.*f", prec, v);
}

/* src/mapping.rs:87-107, 161-182, 228-247 */
int ok_format_paf(const ok_index *idx, const ok_params *p, const char *qid, const ok_result *r,
                  double map_time, char *buf, size_t cap) {
    char f_isd[64], f_j[64], f_j2[64], f_jd[64], f_sh[64], f_t[64], f_mi[64];
    fmt_f(f_isd, sizeof f_isd, r->sigmas, 5);
    fmt_f(f_j, sizeof f_j, r->J, 5);
    fmt_f(f_j2, sizeof f_j2, r->J2, 5);
    fmt_f(f_jd, sizeof f_jd, r->J - r->J2, 5);
    fmt_f(f_sh, sizeof f_sh, r->sh, 5);
    fmt_f(f_t, sizeof f_t, map_time, 5);
    fmt_f(f_mi, sizeof f_mi, r->match_inefficiency, 3);
    return snprintf(
        buf, cap,
        "%s\t%d\t%d\t%d\t%c\t%s\t%d\t%d\t%d\t%d\t%d\t%d"
        "\tp:i:%d\ts:i:%d\tI:i:%d\tI2:i:%d\tIdiff:i:%d\tIsdiff:f:%s\tJ:f:%s\tJ2:f:%s"
        "\tJdiff:f:%s\tsh:f:%s\tstrand:i:%d\tt:f:%s\tb:s:(%d,%d)\tb2:s:(%d,%d)"
        "\tk:i:%d\tgt_J:f:-1.000\tgt_C:f:-1.000\tgt_C_bucket:f:-1.000\tseeds:i:%d"
        "\tmax_seed_matches:i:%d\tseed_matches:i:%d\ttotal_matches:i:%d"
        "\tmatch_inefficiency:f:%s\tseeded_buckets:i:%d\tfinal_buckets:i:%d\tFPTP:f:-1.000",
        qid, r->read_len, r->p_start, r->p_end, (char)r->strand, idx->names[r->segm_id],
        idx->sz[r->segm_id], r->t_l, r->t_r, r->read_len, r->read_len, r->mapq, r->m, r->s_sz,
        r->intersection, r->intersection2, r->intersection - r->intersection2, f_isd, f_j, f_j2,
        f_jd, f_sh, r->same_strand, f_t, r->b_segm, r->b_b, r->b2_segm, r->b2_b, p->k, r->S,
        (int32_t)r->max_seed_matches, (int32_t)r->seed_matches, (int32_t)r->total_matches, f_mi,
        (int32_t)r->seeded_buckets, (int32_t)r->final_buckets);
}
