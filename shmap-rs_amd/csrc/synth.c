/*
 * Seeded synthetic references and reads for the mapping workloads.
 *
 * Every base and every read is a pure function of (seed, position) or
 * (seed, read index) through a counter-based splitmix64 stream, so output does
 * not depend on the thread count.  Reads follow the error model of the
 * reference's simulator (independent substitution / insertion / deletion per
 * base; simulate/simulate_reads.py:87-125 of shmap-rs) and carry truth
 * coordinates (contig, start, end, strand) for the placement check.
 */
#define _GNU_SOURCE
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static inline uint64_t mix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
static inline uint64_t rng_at(uint64_t seed, uint64_t stream, uint64_t i) {
    return mix64(mix64(seed ^ (stream * 0xD1B54A32D192ED03ull)) + i * 0x9E3779B97F4A7C15ull);
}
static inline double unif(uint64_t r) { return (double)(r >> 11) * (1.0 / 9007199254740992.0); }

static const uint8_t ACGT[4] = {'A', 'C', 'G', 'T'};

typedef struct {
    int64_t start, len; /* global offsets */
    int32_t unit;
    double div;
} array_t;

typedef struct {
    uint64_t seed;
    uint8_t *out;
    int64_t total;
    const int64_t *coff;
    int32_t n_contigs;
    const array_t *arr;
    int64_t n_arr;
    const uint8_t *units;
    int32_t unit_len;
    int64_t next;
    int64_t block;
    pthread_mutex_t mu;
} ref_job_t;

static void *ref_worker(void *arg) {
    ref_job_t *J = arg;
    for (;;) {
        pthread_mutex_lock(&J->mu);
        const int64_t b0 = J->next;
        J->next += J->block;
        pthread_mutex_unlock(&J->mu);
        if (b0 >= J->total) break;
        const int64_t b1 = b0 + J->block < J->total ? b0 + J->block : J->total;
        for (int64_t p = b0; p < b1; p++) J->out[p] = ACGT[rng_at(J->seed, 1, (uint64_t)p) & 3];
        /* overlay tandem arrays intersecting [b0, b1) */
        int64_t lo = 0, hi = J->n_arr;
        while (lo < hi) {
            const int64_t mid = (lo + hi) / 2;
            if (J->arr[mid].start + J->arr[mid].len <= b0) lo = mid + 1; else hi = mid;
        }
        for (int64_t a = lo; a < J->n_arr && J->arr[a].start < b1; a++) {
            const array_t *A = &J->arr[a];
            const int64_t s = A->start > b0 ? A->start : b0;
            const int64_t e = A->start + A->len < b1 ? A->start + A->len : b1;
            const uint8_t *unit = J->units + (size_t)A->unit * J->unit_len;
            for (int64_t p = s; p < e; p++) {
                uint8_t base = unit[(p - A->start) % J->unit_len];
                const uint64_t r = rng_at(J->seed, 2, (uint64_t)p);
                if (unif(r) < A->div) {
                    int code = base == 'A' ? 0 : base == 'C' ? 1 : base == 'G' ? 2 : 3;
                    base = ACGT[(code + 1 + (int)((r >> 3) % 3)) & 3];
                }
                J->out[p] = base;
            }
        }
    }
    return NULL;
}

/* Contigs of the given lengths, concatenated into out (sum(lens) bytes):
 * uniform ACGT background with a fraction repeat_frac laid out as tandem arrays
 * of array_len bases, each copying one of n_units random units of unit_len
 * bases, every copy base substituted with probability div in [div_min, div_max]
 * (drawn per array). */
int synth_reference(uint64_t seed, int32_t n_contigs, const int64_t *lens, double repeat_frac,
                    int32_t unit_len, double div_min, double div_max, int32_t n_units, int64_t array_len,
                    uint8_t *out, int32_t n_threads) {
    int64_t *coff = malloc(sizeof(int64_t) * (size_t)(n_contigs + 1));
    coff[0] = 0;
    for (int32_t c = 0; c < n_contigs; c++) coff[c + 1] = coff[c] + lens[c];
    uint8_t *units = malloc((size_t)(n_units > 0 ? n_units : 1) * (size_t)(unit_len > 0 ? unit_len : 1));
    for (int32_t u = 0; u < n_units; u++)
        for (int32_t i = 0; i < unit_len; i++)
            units[(size_t)u * unit_len + i] = ACGT[rng_at(seed, 3, (uint64_t)u * 1000003ull + i) & 3];
    int64_t n_arr = 0, cap = 16;
    array_t *arr = malloc(sizeof(array_t) * (size_t)cap);
    if (repeat_frac > 0 && n_units > 0 && unit_len > 0 && array_len > 0) {
        for (int32_t c = 0; c < n_contigs; c++) {
            const int64_t na = (int64_t)llround((double)lens[c] * repeat_frac / (double)array_len);
            if (na <= 0) continue;
            const int64_t slot = lens[c] / na;
            if (slot < array_len) continue;
            for (int64_t j = 0; j < na; j++) {
                const uint64_t r = rng_at(seed, 4, (uint64_t)c * 1000003ull + (uint64_t)j);
                if (n_arr == cap) {
                    cap *= 2;
                    arr = realloc(arr, sizeof(array_t) * (size_t)cap);
                }
                arr[n_arr].len = array_len;
                arr[n_arr].start = coff[c] + j * slot + (int64_t)(r % (uint64_t)(slot - array_len + 1));
                arr[n_arr].unit = (int32_t)((r >> 20) % (uint64_t)n_units);
                arr[n_arr].div = div_min + (div_max - div_min) * unif(mix64(r));
                n_arr++;
            }
        }
    }
    ref_job_t J = {seed, out, coff[n_contigs], coff, n_contigs, arr, n_arr, units, unit_len, 0, 1 << 20,
                   PTHREAD_MUTEX_INITIALIZER};
    if (n_threads < 1) n_threads = 1;
    if (n_threads > 64) n_threads = 64;
    pthread_t th[64];
    for (int t = 0; t < n_threads; t++) pthread_create(&th[t], NULL, ref_worker, &J);
    for (int t = 0; t < n_threads; t++) pthread_join(th[t], NULL);
    free(arr);
    free(units);
    free(coff);
    return 0;
}

typedef struct {
    uint64_t seed;
    const uint8_t *ref;
    const int64_t *coff;
    int32_t n_contigs;
    int32_t n_reads;
    int32_t len_mode;
    double mu, sigma, len_mean;
    int32_t len_min, len_max;
    const double *err;
    int32_t n_classes;
    int32_t both;
    uint8_t *bases; /* NULL: lengths only */
    int64_t *read_off;
    int32_t *lens;
    int32_t *truth;
    int32_t next;
    pthread_mutex_t mu_lock;
} reads_job_t;

static int32_t contig_of(const int64_t *coff, int32_t n, int64_t p) {
    int32_t lo = 0, hi = n - 1;
    while (lo < hi) {
        const int32_t mid = (lo + hi + 1) / 2;
        if (coff[mid] <= p) lo = mid; else hi = mid - 1;
    }
    return lo;
}

/* Generates read i; writes into dst when non-NULL; returns its length. */
static int32_t make_read(const reads_job_t *J, int32_t i, uint8_t *dst, int32_t *truth) {
    uint64_t ctr = 0;
#define NEXT() rng_at(J->seed, 100 + (uint64_t)i, ctr++)
    int64_t T;
    if (J->len_mode == 0) {
        T = (int64_t)J->len_mean;
    } else {
        const double u1 = unif(NEXT()) + 1e-300, u2 = unif(NEXT());
        const double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
        T = (int64_t)llround(exp(J->mu + J->sigma * z));
        if (T < J->len_min) T = J->len_min;
        if (T > J->len_max) T = J->len_max;
    }
    const int64_t total = J->coff[J->n_contigs];
    if (T > total) T = total;
    int64_t s = (int64_t)(NEXT() % (uint64_t)(total - T + 1));
    int32_t c = contig_of(J->coff, J->n_contigs, s);
    const int64_t clen = J->coff[c + 1] - J->coff[c];
    if (T > clen) T = clen;
    if (s + T > J->coff[c + 1]) s = J->coff[c + 1] - T;
    const int strand = J->both ? (int)(NEXT() & 1) : 0;
    const int cls = J->n_classes > 1 ? (int)(NEXT() % (uint64_t)J->n_classes) : 0;
    const double rate = J->err[4 * cls], fs = J->err[4 * cls + 1], fi = J->err[4 * cls + 2];
    int32_t n = 0;
    for (int64_t p = s; p < s + T; p++) {
        const uint8_t b = J->ref[p];
        const uint64_t r = NEXT();
        const double u = unif(r);
        if (u < rate) {
            const double t = u / rate;
            if (t < fs) {
                int code = b == 'A' ? 0 : b == 'C' ? 1 : b == 'G' ? 2 : 3;
                if (dst) dst[n] = ACGT[(code + 1 + (int)((r >> 7) % 3)) & 3];
                n++;
            } else if (t < fs + fi) {
                if (dst) {
                    dst[n] = b;
                    dst[n + 1] = ACGT[(r >> 9) & 3];
                }
                n += 2;
            } /* else deletion */
        } else {
            if (dst) dst[n] = b;
            n++;
        }
    }
    if (dst && strand) {
        for (int32_t a = 0, z = n - 1; a <= z; a++, z--) {
            const uint8_t x = dst[a], y = dst[z];
            dst[a] = y == 'A' ? 'T' : y == 'C' ? 'G' : y == 'G' ? 'C' : y == 'T' ? 'A' : y;
            dst[z] = x == 'A' ? 'T' : x == 'C' ? 'G' : x == 'G' ? 'C' : x == 'T' ? 'A' : x;
        }
    }
    if (truth) {
        truth[0] = c;
        truth[1] = (int32_t)(s - J->coff[c]);
        truth[2] = (int32_t)(s + T - J->coff[c]);
        truth[3] = strand;
    }
#undef NEXT
    return n;
}

static void *reads_worker(void *arg) {
    reads_job_t *J = arg;
    for (;;) {
        pthread_mutex_lock(&J->mu_lock);
        const int32_t i0 = J->next;
        J->next += 64;
        pthread_mutex_unlock(&J->mu_lock);
        if (i0 >= J->n_reads) break;
        const int32_t i1 = i0 + 64 < J->n_reads ? i0 + 64 : J->n_reads;
        for (int32_t i = i0; i < i1; i++) {
            if (J->bases)
                make_read(J, i, J->bases + J->read_off[i], J->truth ? J->truth + 4 * (size_t)i : NULL);
            else
                J->lens[i] = make_read(J, i, NULL, NULL);
        }
    }
    return NULL;
}

/* Reads sampled from the concatenated reference (contig offsets coff, n+1).
 * len_mode 0: fixed length len_mean; 1: lognormal with mean len_mean and
 * standard deviation len_sd, clipped to [len_min, len_max].  err holds
 * n_classes rows of (rate, f_sub, f_ins, f_del); each read draws one class.
 * With bases == NULL only read_off (n+1 offsets) is filled; call again with a
 * buffer of read_off[n] bytes to generate the bases and truth (n*4). */
int64_t synth_reads(uint64_t seed, const uint8_t *ref, const int64_t *coff, int32_t n_contigs,
                    int32_t n_reads, int32_t len_mode, double len_mean, double len_sd, int32_t len_min,
                    int32_t len_max, const double *err, int32_t n_classes, int32_t both_strands,
                    uint8_t *bases, int64_t *read_off, int32_t *truth, int32_t n_threads) {
    reads_job_t J;
    memset(&J, 0, sizeof J);
    J.seed = seed;
    J.ref = ref;
    J.coff = coff;
    J.n_contigs = n_contigs;
    J.n_reads = n_reads;
    J.len_mode = len_mode;
    J.len_mean = len_mean;
    const double v = len_sd * len_sd / (len_mean * len_mean);
    J.sigma = sqrt(log(1.0 + v));
    J.mu = log(len_mean) - 0.5 * J.sigma * J.sigma;
    J.len_min = len_min;
    J.len_max = len_max;
    J.err = err;
    J.n_classes = n_classes;
    J.both = both_strands;
    J.bases = bases;
    J.read_off = read_off;
    J.truth = truth;
    pthread_mutex_init(&J.mu_lock, NULL);
    if (n_threads < 1) n_threads = 1;
    if (n_threads > 64) n_threads = 64;
    pthread_t th[64];
    if (!bases) {
        J.lens = malloc(sizeof(int32_t) * (size_t)(n_reads > 0 ? n_reads : 1));
        for (int t = 0; t < n_threads; t++) pthread_create(&th[t], NULL, reads_worker, &J);
        for (int t = 0; t < n_threads; t++) pthread_join(th[t], NULL);
        read_off[0] = 0;
        for (int32_t i = 0; i < n_reads; i++) read_off[i + 1] = read_off[i] + J.lens[i];
        free(J.lens);
        return read_off[n_reads];
    }
    for (int t = 0; t < n_threads; t++) pthread_create(&th[t], NULL, reads_worker, &J);
    for (int t = 0; t < n_threads; t++) pthread_join(th[t], NULL);
    return read_off[n_reads];
}
