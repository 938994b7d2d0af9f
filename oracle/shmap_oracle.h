/*
 * shmap oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, single-source restatement of shmap-rs's CPU mapping path
 * (sketch -> index -> seeding -> buckets -> pruning -> scoring -> MAPQ ->
 * PAF text).  It exists to check the HIP product path and to be timed as the
 * CPU baseline.  Only tests, the smoke check and the benchmark's CPU-baseline
 * leg may load it; nothing in the shipped library calls into it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the shmap-rs crate root).
 *
 * Parity pins (what checks each part; see tests/test_oracle_*.py):
 *   sketch      -- direct-definition ntHash, xor canonical form, RC symmetry,
 *                  chunk invariance, threshold constants
 *                  (tests/nthash_equivalence.rs, src/sketch.rs tests)
 *   index       -- count fixtures (src/index.rs:716-769), -M closed form
 *   seeding     -- pmatches grouping (src/shmap/seeding.rs:222-245)
 *   buckets     -- geometry tests (src/buckets.rs:748-849), brute-force
 *                  per-(seed,bucket) clamp on random inputs
 *   pruning     -- worked hseed example (docs/sections/09_pruning.tex)
 *   scoring     -- LCS = 3 fixture (src/shmap/scoring.rs:499-535), brute-force
 *                  window enumeration for Containment/Jaccard
 *   mapping     -- update/overlap/calc_mapq fixtures (src/mapping.rs:445-495)
 *   end to end  -- golden PAF (tests/fixtures/tiny.golden.paf), short-read
 *                  floor (tests/short_read_sketch_floor.rs)
 *   match_rest  -- best / second-best buckets, scores, sigma, MAPQ,
 *                  final_buckets, refined_buckets and refine_memo_hits
 *                  against an independent brute-force mapper
 *                  (tests/brute_map.py: every bucket scored in closed form,
 *                  the two passes' threshold sequences replayed), four
 *                  metrics x {default, -P, -B}; -P and the default agree on
 *                  the mappings and sh >= J
 * Parity unpinned (no executable reference vector; they follow the source
 * text line by line): -F (abs_pos) bucket geometry; needletail parsing of
 * FASTQ / compressed input; Rust float spelling on exact decimal ties.
 */
#ifndef SHMAP_ORACLE_H
#define SHMAP_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* src/types.rs:24-28 */
typedef struct {
    int32_t r;
    uint64_t h;
    uint8_t strand;
} ok_kmer;

enum { OK_CONTAINMENT = 0, OK_JACCARD = 1, OK_BUCKET_SH = 2, OK_BUCKET_LCS = 3 };

/* src/params.rs:14-186 (only the fields the mapping path reads) */
typedef struct {
    int32_t k;
    double h_frac;
    double theta;
    double min_diff;
    double max_overlap;
    int32_t metric;
    int32_t no_bucket_pruning; /* -P */
    int32_t one_sweep;         /* -B */
    int32_t abs_pos;           /* -F */
    int32_t max_matches;       /* -M, <=0 means None */
} ok_params;

/* One read's outcome: every PAF field plus every per-read counter. */
typedef struct {
    int32_t mapped;
    int32_t unmappable_short;
    int32_t read_len;
    int32_t m;          /* sketch size (p:i:) */
    int32_t u;          /* distinct k-mers */
    int32_t S;          /* seeds:i: */
    int64_t possible_matches;
    int64_t seed_matches;
    int64_t total_matches;
    int64_t max_seed_matches;
    int64_t seeded_buckets;
    int64_t refined_buckets;
    int64_t refine_memo_hits;
    int64_t final_buckets;
    /* mapping (valid when mapped) */
    int32_t p_start, p_end;
    int32_t strand; /* '+' or '-' */
    int32_t segm_id;
    int32_t t_l, t_r;
    int32_t mapq;
    int32_t s_sz;
    int32_t intersection, intersection2;
    int32_t same_strand;
    int32_t b_segm, b_b, b2_segm, b2_b;
    double J, J2, sh, sigmas, match_inefficiency;
} ok_result;

typedef struct ok_index ok_index;

/* src/sketch.rs:58-109, 126-132, 167-219.  Returns the number of k-mers
 * written; writes at most cap (returns the full count regardless). */
int64_t ok_sketch(int32_t k, double h_frac, const uint8_t *s, int64_t len, int32_t offset,
                  ok_kmer *out, int64_t cap);
uint64_t ok_threshold(double h_frac);

/* src/index.rs:384-663 (+ -M closed form of :264-291, :295-315, :360-376). */
ok_index *ok_index_build(int32_t k, double h_frac, int32_t max_matches, int32_t n_segments,
                         const char *const *names, const uint8_t *const *seqs,
                         const int64_t *lens, int32_t n_threads);
void ok_index_free(ok_index *idx);
int32_t ok_index_count(const ok_index *idx, uint64_t h);
/* Writes up to cap hits of h as (r, tpos, strand, segm_id) quadruples. */
int32_t ok_index_hits(const ok_index *idx, uint64_t h, int32_t *out4, int32_t cap);
int32_t ok_index_n_segments(const ok_index *idx);
int64_t ok_index_segment_kmers(const ok_index *idx, int32_t segm_id);
int64_t ok_index_indexed_kmers(const ok_index *idx);
int64_t ok_index_indexed_hits(const ok_index *idx);
int64_t ok_index_blacklisted_kmers(const ok_index *idx);

/* src/shmap/mod.rs:653-971 (map_read) for one read. */
int ok_map_read(const ok_index *idx, const ok_params *p, const uint8_t *seq, int32_t len,
                ok_result *out);
/* The -@ worker pool of src/shmap/mod.rs:374-650 restated: maps n reads on
 * n_threads threads; results land at out[i] in input order. */
int ok_map_reads(const ok_index *idx, const ok_params *p, int32_t n, const uint8_t *const *seqs,
                 const int32_t *lens, ok_result *out, int32_t n_threads);
/* The same, also returning the hit records the seed-heuristic prune pass
 * (matches_in_bucket, src/shmap/pruning.rs:18-80) read in total: the P term of
 * the path's algorithmic bytes.  A measurement, not part of any result. */
int ok_map_reads_counted(const ok_index *idx, const ok_params *p, int32_t n, const uint8_t *const *seqs,
                         const int32_t *lens, ok_result *out, int32_t n_threads, int64_t *prune_hits);

/* src/mapping.rs:87-107, 161-182, 228-247: PAF line (no newline), t:f: printed
 * from map_time.  Returns bytes written (snprintf semantics). */
int ok_format_paf(const ok_index *idx, const ok_params *p, const char *qid,
                  const ok_result *r, double map_time, char *buf, size_t cap);

/* Helpers restated for unit tests. */
double ok_hseed(int32_t p, int32_t seeds, int32_t matches);          /* pruning.rs:10-13 */
int32_t ok_lis_len(const int32_t *x, int32_t n);                        /* scoring.rs:133-154 */
int32_t ok_calc_mapq(double j, double j2, int32_t same_strand, int32_t intersection,
                     double theta2, double min_diff, double *j2_out); /* mapping.rs:341-354 */
int32_t ok_unique_pmatches(const ok_kmer *p, int32_t n, uint64_t *hash, int32_t *occs, int32_t *pm);
double ok_overlap(int32_t sa, int32_t la, int32_t ra, int32_t sb, int32_t lb, int32_t rb);

#ifdef __cplusplus
}
#endif
#endif
