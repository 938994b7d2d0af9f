/*
 * shmap oracle command line -- TEST INFRASTRUCTURE ONLY.
 *
 * Restates the reference CLI (src/main.rs:16-70, src/params.rs:14-263) over the
 * oracle library so parity tests can run the reference's own CLI-level test
 * cases (golden PAF, short-read floor, thread parity) against it.  Output
 * ordering follows the in-order collector of src/shmap/mod.rs:568-622.
 */
#define _GNU_SOURCE
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

#include "shmap_oracle.h"

typedef struct {
    char **names;
    uint8_t **seqs;
    int64_t *lens;
    int32_t n, cap;
} records_t;

static void rec_push(records_t *r, const char *name, size_t nlen, uint8_t *seq, int64_t len) {
    if (r->n == r->cap) {
        r->cap = r->cap ? r->cap * 2 : 64;
        r->names = realloc(r->names, sizeof(char *) * (size_t)r->cap);
        r->seqs = realloc(r->seqs, sizeof(uint8_t *) * (size_t)r->cap);
        r->lens = realloc(r->lens, sizeof(int64_t) * (size_t)r->cap);
    }
    r->names[r->n] = strndup(name, nlen);
    r->seqs[r->n] = seq;
    r->lens[r->n] = len;
    r->n++;
}

/* FASTA/FASTQ records as needletail delivers them (src/io/mod.rs:383-427):
 * id truncated at the first space or tab, line breaks and '\r' removed. */
static int read_fastx(const char *path, records_t *out) {
    FILE *f = fopen(path, "rb");
    if (!f) return -1;
    fseek(f, 0, SEEK_END);
    const long sz = ftell(f);
    fseek(f, 0, SEEK_SET);
    char *buf = malloc((size_t)sz + 1);
    if (fread(buf, 1, (size_t)sz, f) != (size_t)sz) {
        fclose(f);
        free(buf);
        return -1;
    }
    fclose(f);
    buf[sz] = 0;
    long pos = 0;
    while (pos < sz) {
        const char c0 = buf[pos];
        long eol = pos;
        while (eol < sz && buf[eol] != '\n') eol++;
        long hl = eol;
        if (hl > pos && buf[hl - 1] == '\r') hl--;
        if (c0 == '>' || c0 == '@') {
            const char *h = buf + pos + 1;
            size_t nlen = 0;
            while (h + nlen < buf + hl && h[nlen] != ' ' && h[nlen] != '\t') nlen++;
            const char *name = h;
            pos = eol + 1;
            int64_t cap = 1024, len = 0;
            uint8_t *seq = malloc((size_t)cap);
            if (c0 == '>') {
                while (pos < sz && buf[pos] != '>') {
                    long e = pos;
                    while (e < sz && buf[e] != '\n') e++;
                    long le = e;
                    if (le > pos && buf[le - 1] == '\r') le--;
                    if (len + (le - pos) > cap) {
                        while (len + (le - pos) > cap) cap *= 2;
                        seq = realloc(seq, (size_t)cap);
                    }
                    memcpy(seq + len, buf + pos, (size_t)(le - pos));
                    len += le - pos;
                    pos = e + 1;
                }
            } else {
                long e = pos;
                while (e < sz && buf[e] != '\n') e++;
                long le = e;
                if (le > pos && buf[le - 1] == '\r') le--;
                len = le - pos;
                seq = realloc(seq, (size_t)(len > 0 ? len : 1));
                memcpy(seq, buf + pos, (size_t)len);
                pos = e + 1;
                for (int skip = 0; skip < 2 && pos < sz; skip++) { /* '+' and quality lines */
                    while (pos < sz && buf[pos] != '\n') pos++;
                    pos++;
                }
            }
            rec_push(out, name, nlen, seq, len);
        } else {
            pos = eol + 1;
        }
    }
    free(buf);
    return 0;
}

static double now_s(void) {
    struct timeval tv;
    gettimeofday(&tv, NULL);
    return (double)tv.tv_sec + 1e-6 * (double)tv.tv_usec;
}

static void usage(void) {
    fprintf(stderr, "usage: shmap_oracle -s ref.fa -p reads.fa [-k K] [-r R] [-t T] [-d D] [-o O]"
                    " [-m METRIC] [-M M] [-P] [-B] [-F] [-@ N] [-z PREFIX]"
                    " [--per-read-stats FILE] [--per-read-stats-sample N]\n");
}

int main(int argc, char **argv) {
    const char *p_file = NULL, *t_file = NULL, *params_file = "", *stats_file = NULL;
    ok_params prm = {15, 0.05, 0.9, 0.02, 0.5, OK_CONTAINMENT, 0, 0, 0, -1};
    int threads = 1;
    long stats_sample = 1;
    int have_M = 0;
    for (int i = 1; i < argc; i++) {
        const char *a = argv[i];
        const char *v = i + 1 < argc ? argv[i + 1] : NULL;
#define NEEDV                                                                  \
    do {                                                                       \
        if (!v) {                                                              \
            usage();                                                           \
            return 2;                                                          \
        }                                                                      \
        i++;                                                                   \
    } while (0)
        if (!strcmp(a, "-p") || !strcmp(a, "--pattern")) { NEEDV; p_file = v; }
        else if (!strcmp(a, "-s") || !strcmp(a, "--text")) { NEEDV; t_file = v; }
        else if (!strcmp(a, "-k") || !strcmp(a, "--ksize")) { NEEDV; prm.k = atoi(v); }
        else if (!strcmp(a, "-r") || !strcmp(a, "--hashratio")) { NEEDV; prm.h_frac = strtod(v, NULL); }
        else if (!strcmp(a, "-t") || !strcmp(a, "--threshold")) { NEEDV; prm.theta = strtod(v, NULL); }
        else if (!strcmp(a, "-d") || !strcmp(a, "--min_diff")) { NEEDV; prm.min_diff = strtod(v, NULL); }
        else if (!strcmp(a, "-o") || !strcmp(a, "--max_overlap")) { NEEDV; prm.max_overlap = strtod(v, NULL); }
        else if (!strcmp(a, "-M") || !strcmp(a, "--max_matches")) { NEEDV; prm.max_matches = atoi(v); have_M = 1; }
        else if (!strcmp(a, "-z") || !strcmp(a, "--params")) { NEEDV; params_file = v; }
        else if (!strcmp(a, "-@") || !strcmp(a, "--threads")) { NEEDV; threads = atoi(v); }
        else if (!strcmp(a, "--per-read-stats")) { NEEDV; stats_file = v; }
        else if (!strcmp(a, "--per-read-stats-sample")) { NEEDV; stats_sample = atol(v); }
        else if (!strcmp(a, "-P") || !strcmp(a, "--no_bucket_pruning")) prm.no_bucket_pruning = 1;
        else if (!strcmp(a, "-B") || !strcmp(a, "--one_sweep")) prm.one_sweep = 1;
        else if (!strcmp(a, "-F") || !strcmp(a, "--abs_pos")) prm.abs_pos = 1;
        else if (!strcmp(a, "-m") || !strcmp(a, "--metric")) {
            NEEDV;
            if (!strcmp(v, "Containment")) prm.metric = OK_CONTAINMENT;
            else if (!strcmp(v, "Jaccard")) prm.metric = OK_JACCARD;
            else if (!strcmp(v, "bucket_SH")) prm.metric = OK_BUCKET_SH;
            else if (!strcmp(v, "bucket_LCS")) prm.metric = OK_BUCKET_LCS;
            else {
                fprintf(stderr, "error: invalid value '%s' for '--metric <METRIC>'\n", v);
                return 2;
            }
        } else {
            usage();
            return 2;
        }
    }
    if (!p_file || !t_file) {
        usage();
        return 2;
    }
    /* src/params.rs:191-263 */
    if (prm.k <= 0) {
        fprintf(stderr, "ERROR: K-mer length (-k) should be positive. You provided %d.\n", prm.k);
        return 1;
    }
    if (prm.h_frac <= 0.0 || prm.h_frac > 1.0) {
        fprintf(stderr, "ERROR: Given hash ratio (-r) should be between 0 and 1. You provided %g.\n", prm.h_frac);
        return 1;
    }
    if (have_M && prm.max_matches <= 0) {
        fprintf(stderr, "ERROR: The number of seed matches (-M) should be positive. You provided %d.\n", prm.max_matches);
        return 1;
    }
    if (prm.theta < 0.0 || prm.theta > 1.0) {
        fprintf(stderr, "ERROR: The threshold (-t) should be between 0 and 1. You provided %g.\n", prm.theta);
        return 1;
    }
    if (prm.min_diff < 0.0) {
        fprintf(stderr, "ERROR: The minimum difference (-d) should be >= 0. You provided %g.\n", prm.min_diff);
        return 1;
    }
    if (prm.max_overlap < 0.0 || prm.max_overlap > 1.0) {
        fprintf(stderr, "ERROR: The maximum overlap (-o) should be between 0 and 1. You provided %g.\n", prm.max_overlap);
        return 1;
    }
    if (threads <= 0) {
        fprintf(stderr, "ERROR: The number of threads (-@) should be positive. You provided %d.\n", threads);
        return 1;
    }
    if (stats_sample <= 0) {
        fprintf(stderr, "ERROR: --per-read-stats-sample should be >= 1 (1 keeps every read). You provided 0.\n");
        return 1;
    }

    records_t ref = {0}, reads = {0};
    if (read_fastx(t_file, &ref) != 0) {
        fprintf(stderr, "Error: failed to build index from %s\n", t_file);
        return 1;
    }
    ok_index *idx = ok_index_build(prm.k, prm.h_frac, prm.max_matches, ref.n,
                                   (const char *const *)ref.names, (const uint8_t *const *)ref.seqs,
                                   ref.lens, threads);
    if (read_fastx(p_file, &reads) != 0) {
        fprintf(stderr, "Error: failed while mapping reads from %s\n", p_file);
        return 1;
    }
    int32_t *lens = malloc(sizeof(int32_t) * (size_t)(reads.n + 1));
    for (int32_t i = 0; i < reads.n; i++) lens[i] = (int32_t)reads.lens[i];
    ok_result *res = calloc((size_t)(reads.n + 1), sizeof(ok_result));
    const double t0 = now_s();
    ok_map_reads(idx, &prm, reads.n, (const uint8_t *const *)reads.seqs, lens, res, threads);
    const double per_read = reads.n ? (now_s() - t0) / reads.n : 0.0;

    FILE *unm = NULL, *st = NULL;
    if (params_file[0]) {
        char fn[4096];
        snprintf(fn, sizeof fn, "%s.unmapped.paf", params_file);
        unm = fopen(fn, "w");
    }
    if (stats_file) {
        st = fopen(stats_file, "w");
        fprintf(st, "query_id\tread_len\tkmers_sketched\tkmers_seeds\tpossible_matches\t"
                    "examined_matches\tseed_matches\tseeded_buckets\trefined_buckets\t"
                    "final_buckets\tmatches_in_mapping\tmapped\tmapq\tquery_mapping_s\n");
    }
    size_t cap = 1 << 16;
    char *line = malloc(cap);
    for (int32_t i = 0; i < reads.n; i++) {
        const ok_result *r = &res[i];
        if (st && i % stats_sample == 0)
            fprintf(st, "%s\t%d\t%d\t%d\t%d\t%d\t%d\t%d\t%d\t%d\t%d\t%d\t%d\t%.9f\n",
                    reads.names[i], r->read_len, r->m, r->S, (int32_t)r->possible_matches,
                    (int32_t)r->total_matches, (int32_t)r->seed_matches,
                    (int32_t)r->seeded_buckets, (int32_t)r->refined_buckets,
                    (int32_t)r->final_buckets, r->mapped ? r->intersection : 0, r->mapped,
                    r->mapped ? (r->mapq == 60 ? 60 : 0) : -1, per_read);
        if (r->mapped) {
            int n = ok_format_paf(idx, &prm, reads.names[i], r, per_read, line, cap);
            if ((size_t)n >= cap) {
                cap = (size_t)n + 1;
                line = realloc(line, cap);
                ok_format_paf(idx, &prm, reads.names[i], r, per_read, line, cap);
            }
            fputs(line, stdout);
            fputc('\n', stdout);
        } else if (unm) {
            fprintf(unm, "%s\t%d\t0\t0\t*\t*\t0\t0\t0\t0\t0\t0\n", reads.names[i], r->read_len);
        }
    }
    if (unm) fclose(unm);
    if (st) fclose(st);
    ok_index_free(idx);
    return 0;
}
