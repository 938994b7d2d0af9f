/*
 * shmap_hip -- C ABI of the MI355X (gfx950) per-read mapping library.
 *
 * The entry points mirror shmap-rs's library surface for the mapping path:
 * the FracMinHash sketcher, the reference SketchIndex and the Mapper plugin
 * (reference paths are relative to the shmap-rs crate root).  All calls
 * return 0 on success and a nonzero shm_status otherwise; shm_last_error()
 * returns a thread-local message for the last failure.  Callers own every
 * host buffer they pass; handles own their device memory.
 */
#ifndef SHMAP_HIP_H
#define SHMAP_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum shm_status {
    SHM_OK = 0,
    SHM_ERR_ARG = 1,     /* invalid argument (the reference's Params::validate) */
    SHM_ERR_HIP = 2,     /* HIP runtime failure */
    SHM_ERR_NOMEM = 3,   /* device or host allocation failed */
    SHM_ERR_CAPACITY = 4, /* caller buffer too small; the needed size is returned */
    SHM_ERR_IO = 5        /* a file could not be opened, read or written */
};

/* Metric -- src/types.rs:300-322 */
enum shm_metric { SHM_CONTAINMENT = 0, SHM_JACCARD = 1, SHM_BUCKET_SH = 2, SHM_BUCKET_LCS = 3 };

/* Kmer{r, h, strand} -- src/types.rs:24-28.  strand: 0 forward, 1 reverse. */
typedef struct {
    uint64_t h;
    int32_t r;
    uint8_t strand;
    uint8_t pad[3];
} shm_kmer;

/* Hit{r, tpos, strand, segm_id} -- src/types.rs:44-51 */
typedef struct {
    int32_t r;
    int32_t tpos;
    int32_t strand;
    int32_t segm_id;
} shm_hit;

/* One reference segment (FASTA record) handed to the index build. */
typedef struct {
    const char *name;
    const uint8_t *seq;
    int64_t len;
} shm_segment;

/* One read handed to the mapper (host memory). */
typedef struct {
    const uint8_t *seq;
    int32_t len;
} shm_read;

/* The parameters map_read consumes -- src/params.rs:14-186.  max_matches is
 * the index's -M and is given to shm_index_build instead. */
typedef struct {
    int32_t k;
    double h_frac;
    double theta;            /* -t */
    double min_diff;         /* -d */
    double max_overlap;      /* -o */
    int32_t metric;          /* -m, enum shm_metric */
    int32_t no_bucket_pruning; /* -P */
    int32_t one_sweep;       /* -B */
    int32_t abs_pos;         /* -F */
} shm_map_params;

/* Everything map_read produces for one read (src/shmap/mod.rs:653-971): the
 * fields of the PAF record (src/mapping.rs:12-247) and the per-read counters
 * printed in it or in --per-read-stats (src/shmap/mod.rs:280-327). */
typedef struct {
    int32_t read_len;
    int32_t m;                /* sketch size, p:i: */
    int32_t u;                /* distinct k-mers */
    int32_t S;                /* seeds:i: */
    int32_t mapped;
    int32_t unmappable_short; /* sketch below MIN_HALFLEN */
    int32_t possible_matches;
    int32_t seed_matches;
    int32_t total_matches;
    int32_t max_seed_matches;
    int32_t seeded_buckets;
    int32_t refined_buckets;
    int32_t refine_memo_hits;
    int32_t final_buckets;
    int32_t p_start, p_end;
    int32_t strand;           /* '+' or '-' */
    int32_t segm_id;
    int32_t t_l, t_r;
    int32_t mapq;
    int32_t s_sz;
    int32_t intersection, intersection2;
    int32_t same_strand;
    int32_t b_segm, b_b, b2_segm, b2_b;
    int32_t status;           /* 0 = ok; nonzero (a shm_status) = this read failed (see shm_map_batch) */
    int32_t has_second;       /* a second-best mapping was found */
    double J, J2, sh, sigmas, match_inefficiency;
} shm_read_result;

typedef struct shm_sketcher shm_sketcher;
typedef struct shm_index shm_index;
typedef struct shm_mapper shm_mapper;

const char *shm_last_error(void);

/* FracMinHash::new -- src/sketch.rs:58-109 */
int shm_sketcher_create(int32_t k, double h_frac, shm_sketcher **out);
void shm_sketcher_free(shm_sketcher *sk);

/* FracMinHash::sketch_slice_into -- src/sketch.rs:167-209 (sketch/sketch_into
 * are offset 0).  Runs on the GPU.  *n receives the number of k-mers; when it
 * exceeds cap nothing is written and SHM_ERR_CAPACITY is returned. */
int shm_sketch(const shm_sketcher *sk, const uint8_t *seq, int64_t len, int32_t offset,
               shm_kmer *out, int64_t cap, int64_t *n);

/* No reference counterpart: shm_sketch for callers that cannot allocate (code
 * under stream capture, tight serving loops, harnesses that forbid hipMalloc).
 * The same k-mers as shm_sketch, with every device intermediate carved from
 * one fixed array per GPU of SHM_NOALLOC_SCRATCH_BYTES (held until the process
 * ends) instead of allocated: no hipMalloc, hipFree, stream creation or other
 * allocating call.  seq and out are the caller's host memory; the work runs on
 * `stream` (a hipStream_t of the current device; NULL for the null stream)
 * and the call returns when out is written.  A sequence whose intermediates do
 * not fit the array is refused with SHM_ERR_CAPACITY before any
 * device work; *n and SHM_ERR_CAPACITY as shm_sketch when out is too small.
 * Calls on one GPU share its array and would race: this library serialises
 * them (one host mutex per device), so threads of one process may call it
 * freely, each call waiting for the one before. */
#define SHM_NOALLOC_SCRATCH_BYTES (64ull << 20)
int shm_sketch_noalloc(const shm_sketcher *sk, const uint8_t *seq, int64_t len, int32_t offset,
                       shm_kmer *out, int64_t cap, int64_t *n, void *stream);
/* Bytes of the fixed array shm_sketch_noalloc needs for len bases and room for
 * cap k-mers (-1 on a bad argument): 29 B per k-mer of room plus about 2 B per
 * base, so 64 MiB holds a 23 Mbp sequence at k=25 r=0.01 with cap = 1.2 len r,
 * and 2.1 Mbp with cap = len. */
int64_t shm_sketch_noalloc_bytes(const shm_sketcher *sk, int64_t len, int64_t cap);

/* SketchIndex::build_index -- src/index.rs:384-663.  Builds the whole index on
 * the GPU and keeps it resident.  max_matches <= 0 means None (-M unset). */
int shm_index_build(const shm_sketcher *sk, const shm_segment *segs, int32_t n_segments,
                    int32_t max_matches, int32_t device_id, shm_index **out);
void shm_index_free(shm_index *idx);

/* SketchIndex::count -- src/index.rs:204-213 */
int shm_index_count(const shm_index *idx, uint64_t h, int32_t *count);
/* SketchIndex::single_hit / multi_hits -- src/index.rs:192-201: every hit of h
 * sorted by (segm_id, r).  *n receives the count; SHM_ERR_CAPACITY if > cap. */
int shm_index_hits(const shm_index *idx, uint64_t h, shm_hit *out, int32_t cap, int32_t *n);
/* SketchIndex::get_segment -- src/index.rs:184-186 (name, sz, kmers.len()) */
int shm_index_n_segments(const shm_index *idx, int32_t *n);
int shm_index_segment(const shm_index *idx, int32_t segm_id, const char **name, int32_t *sz,
                      int64_t *n_kmers);
/* get_kmer_stats counters -- src/index.rs:343-376 */
int shm_index_stats(const shm_index *idx, int64_t *indexed_kmers, int64_t *indexed_hits,
                    int64_t *highest_freq, int64_t *blacklisted_kmers, int64_t *blacklisted_hits);
/* The parameters the index was built with: k and h_frac of its sketcher and
 * max_matches (0: -M unset).  No reference counterpart (the caller's own
 * Params hold them there). */
int shm_index_params(const shm_index *idx, int32_t *k, double *h_frac, int32_t *max_matches);

/* Persistent index.  shm_index_save writes the whole device index (parameters,
 * segment table, sketch, hit arrays and hash table) to a file;
 * shm_index_load reads it onto device_id, and the loaded index maps exactly as
 * the built one.  No reference counterpart: shmap-rs rebuilds its index from
 * the FASTA on every run, and names an on-disk index as its next step
 * (RESULTS.md:1646-1650).  shm_index_load checks the file's magic, version,
 * header and size before it touches the device (SHM_ERR_ARG with a message
 * for a file that is not an index, of another version, or truncated;
 * SHM_ERR_IO when it cannot be opened). */
int shm_index_save(const shm_index *idx, const char *path);
int shm_index_load(const char *path, int32_t device_id, shm_index **out);

/* create_mapper -- src/mapper.rs:25-39: a mapper over a built index, holding
 * the device scratch for batches of reads. */
int shm_mapper_create(const shm_index *idx, const shm_map_params *params, shm_mapper **out);
void shm_mapper_free(shm_mapper *mp);

/* Mapper::map_reads for a batch of host reads -- src/shmap/mod.rs:374-650,
 * with map_read (:653-971) for each.  out[i] is read i's result, in input
 * order.  As in the reference, where a read that panics becomes one Err while
 * every other read is still written (src/shmap/mod.rs:219-238, 597-612), a
 * read can fail alone: a malformed read (negative length, no bases:
 * SHM_ERR_ARG), a read over the mapper's length limit (SHM_ERR_CAPACITY, see
 * shm_mapper_set_read_limit), or a read that does not fit device memory even
 * in a batch of its own (SHM_ERR_NOMEM; a batch that runs out of memory is
 * split and mapped again).  Such a read has out[i].status != 0 and
 * mapped = 0; every other record is complete.  The call then returns the
 * first failed read's status, with shm_last_error() naming it.  A failure of
 * the whole batch (SHM_ERR_HIP) sets every read's status and is returned. */
int shm_map_batch(shm_mapper *mp, const shm_read *reads, int32_t n, shm_read_result *out);

/* Reads longer than max_read_len bases fail alone in shm_map_batch (status
 * SHM_ERR_CAPACITY) before any device work; 0 (the default) means no limit.
 * No reference counterpart: an admission guard for callers (the serving loop)
 * that must keep one oversized request from taking a batch's device memory. */
int shm_mapper_set_read_limit(shm_mapper *mp, int32_t max_read_len);

/* The same over reads already resident in device memory: bases d_seq with
 * n+1 offsets d_off (device pointers), results copied to host out. */
int shm_map_batch_device(shm_mapper *mp, const uint8_t *d_seq, const int64_t *d_off, int32_t n,
                         shm_read_result *out);

/* 2-bit packed reads: a quarter of the bytes to move to the device.  No
 * reference counterpart (the reference maps reads from memory it parsed
 * itself); for callers whose reads start in host memory, where the copy over
 * the link, not the mapping, sets the rate.
 *
 * shm_pack_reads packs a batch on the host (threads > 1 splits it): base i of
 * the batch (reads back to back) goes to bits 2*(i%4) of packed[i/4], A C G T
 * (either case) as 0 1 2 3; off[0..n] receives the reads' base offsets, and
 * exc[0..2*n_exc) the runs of every other byte as sorted [start, end) base
 * positions (any non-ACGT byte has the ntHash seed 0, src/sketch.rs:73-88, so
 * the mapping does not depend on which byte it was).  packed needs
 * (bases + 3) / 4 bytes; SHM_ERR_CAPACITY when packed or exc is too small
 * (*n_exc is still set).
 *
 * shm_map_batch_packed maps such a batch: the same records as shm_map_batch on
 * the reads it was packed from.  packed, off and exc are host memory (from
 * shm_host_alloc for full link speed); the bases are unpacked on the device. */
int shm_pack_reads(const shm_read *reads, int32_t n, uint8_t *packed, int64_t packed_cap, int64_t *off,
                   int64_t *exc, int64_t exc_cap, int64_t *n_exc, int32_t threads);
int shm_map_batch_packed(shm_mapper *mp, const uint8_t *packed, const int64_t *off, int32_t n,
                         const int64_t *exc, int64_t n_exc, shm_read_result *out);

/* Per-read record formatting shared by every caller: the PAF line of a mapped
 * read (src/mapping.rs:87-107, 161-182, 228-247; no newline) with map_time as
 * t:f:, or the unmapped line (src/mapping.rs:82-84).  Returns the length
 * written (snprintf semantics). */
int shm_format_paf(const shm_index *idx, const shm_map_params *params, const char *query_id,
                   const shm_read_result *r, double map_time, char *buf, size_t cap);
int shm_format_unmapped(const char *query_id, int32_t read_len, char *buf, size_t cap);

/* Page-locked host memory for read and result buffers: copies between it and
 * the device run at full link speed (shm_map_batch_device writes its results
 * straight into `out`).  No reference counterpart: host allocation is the
 * caller's in the Rust crate. */
int shm_host_alloc(size_t bytes, void **out);
void shm_host_free(void *p);

/* The number of compute units of device `device` (for occupancy and issue
 * figures); no reference counterpart. */
int shm_device_compute_units(int32_t device, int32_t *cus);

/* Device kernel timing of the last batch on this mapper, in milliseconds:
 * total GPU time and the time of each stage (HIP events on the mapper's own
 * stream).  stage_ms receives up to n_stages values; names via
 * shm_mapper_stage_name. */
int shm_mapper_last_timing(const shm_mapper *mp, double *total_ms, double *stage_ms,
                           int32_t n_stages, int32_t *n_written);
const char *shm_mapper_stage_name(int32_t stage);
/* Work counts of the last batch on this mapper: [0] k-mers sketched, [1]
 * distinct k-mers (seeds), [2] buckets counted, [3] buckets refined (a superset
 * of what the match_rest passes read), [4] read groups (table-budget batches),
 * [5] first-pass candidates (bound at least theta after the full-extension
 * prune), [6] buckets only a second pass can reach, [7] examined hits (the
 * index hits of every consumed seed), [8] reads whose buckets were counted in
 * device-memory hash tables rather than on chip (many hits or buckets), [9]
 * host synchronisations of the batch's stream, [10] 1 when a speculative
 * batch outgrew its buffers and was mapped again (see
 * shm_mapper_set_speculation), [11] device and page-locked buffer
 * allocations the batch made (0 once the mapper's buffers fit its batches),
 * [12] 1 when the batch ran as a graph launch, [13] 1 when that graph was
 * captured by this batch (see shm_mapper_set_graphs).
 * Writes up to n values and their number to *n_written. */
int shm_mapper_last_counts(const shm_mapper *mp, int64_t *counts, int32_t n, int32_t *n_written);
/* Turns per-stage event timing on or off (off by default: it adds a stream
 * synchronisation per stage group). */
int shm_mapper_set_timing(shm_mapper *mp, int32_t on);
/* Turns speculative batches of shm_map_batch and shm_map_batch_packed on (the
 * default) or off.  A speculative batch sizes its device buffers from the
 * batches the mapper has seen, runs every kernel without reading a size back,
 * and synchronises with the host once; device-side checks stop it when a size
 * outgrows its buffer, and it is then mapped again with a read-back of each
 * size before what depends on it (three to four synchronisations, as every
 * batch takes with speculation off, and shm_map_batch_device always: it has
 * no host copy of the offsets).  The records are the same either way.  No
 * reference counterpart. */
int shm_mapper_set_speculation(shm_mapper *mp, int32_t on);
/* Turns graph capture of speculative batches on (the default) or off.  The
 * stream work of a speculative batch whose reads are in the mapper's own
 * upload buffers (shm_map_batch, shm_map_batch_packed) depends only on its
 * read count and the mapper's buffer capacities; the batch is padded with
 * empty reads to its read class (1, 2, 3, 4, 6, 8, 12, ...: powers of two and
 * their halfway points), captured into a HIP graph once per class and
 * replayed for every later batch of the class, until a buffer grows.  Records
 * are the same either way; per-stage timing (shm_mapper_set_timing) runs
 * without graphs.  No reference counterpart. */
int shm_mapper_set_graphs(shm_mapper *mp, int32_t on);

/* Batching choices of a mapper: where per-read work moves from one execution
 * shape to another.  No reference counterpart.  Any values give the same
 * records, only the speed changes; 0 in a field means the built-in rule.
 *   contrib_wave_hits   examined hits up to which one wave counts a read's
 *                       buckets whole; larger reads go to a workgroup
 *                       (built-in: 1024 for batches under 2^24 examined
 *                       hits, else 4096)
 *   contrib_chunk_hits  chunk length of the reads past a workgroup, counted
 *                       through global tables, a power of two from 64 to
 *                       4096 (built-in: 512-4096, about 8192 chunks)
 *   prune_thread_seeds  seeds one thread extends a bucket with before a wave
 *                       takes it over (built-in: 16)
 *   contrib_wg_hits     examined hits up to which one workgroup counts a
 *                       read's buckets whole in a shared table of 1,024
 *                       slots (built-in: 8192)
 *   seed_key_bits       leading bits of each k-mer hash (from the keep
 *                       threshold's highest bit down) in the one-word keys a
 *                       wave sorts a read's k-mers by, 1 to 55; k-mers whose
 *                       distinct hashes share them are put in order after
 *                       the sort (built-in: 55) */
typedef struct {
    int32_t contrib_wave_hits;
    int32_t contrib_chunk_hits;
    int32_t prune_thread_seeds;
    int32_t contrib_wg_hits;
    int32_t seed_key_bits;
} shm_tuning;
int shm_mapper_set_tuning(shm_mapper *mp, const shm_tuning *t);
int shm_mapper_get_tuning(const shm_mapper *mp, shm_tuning *t);
/* Autotuner: maps the sample batch repeatedly under candidate tunings (a
 * coordinate search over the four batching fields, workgroup bound first,
 * each candidate timed twice by the batch's GPU time and taken only when 2%
 * faster than the best so far; seed_key_bits is left as set), keeps the
 * fastest on the mapper and writes it to *out (optional).  The sample should
 * look like the batches to come; a few thousand reads suffice. */
int shm_mapper_autotune(shm_mapper *mp, const shm_read *reads, int32_t n_reads, shm_tuning *out);

/* ---- Serving loop: continuous batching over one resident index ----------
 *
 * No reference counterpart: shmap-rs maps one read file per run
 * (src/shmap/mod.rs:374-650).  The server keeps the index on the device and
 * maps reads as callers submit them.  `workers` host threads each own one
 * mapper (one HIP stream); a free worker takes every queued read up to the
 * admission budget (max_batch_reads reads, max_batch_bases bases, which bound
 * the batch's device memory) and maps them as one batch, so reads that arrive
 * while other batches are on the GPU are admitted as soon as a slot frees, and
 * the batches of different workers overlap on the device.  A read's result
 * does not depend on the batch it lands in. */
typedef struct shm_server shm_server;

typedef struct {
    int32_t workers;         /* concurrent batches (mappers/streams); 0 = 3 */
    int32_t max_batch_reads; /* admission budget per batch; 0 = 65536 */
    int64_t max_batch_bases; /* admission budget per batch; 0 = 2^31 */
    int32_t max_read_len;    /* longer reads fail alone (their status); 0 = no limit */
    int32_t flags;           /* SHM_SERVER_*: the workers' mapper settings; 0 = defaults */
} shm_server_config;

#define SHM_SERVER_NO_SPECULATION 1 /* workers' batches synchronised (shm_mapper_set_speculation) */
#define SHM_SERVER_NO_GRAPHS 2      /* workers' batches not captured (shm_mapper_set_graphs) */

typedef struct {
    uint64_t tag;          /* the caller's id from shm_server_submit */
    double submit_s;       /* seconds since shm_server_create: submitted, */
    double start_s;        /* its batch started, */
    double done_s;         /* its result was ready */
    int32_t batch_reads;   /* reads in the batch it was mapped with */
    int32_t worker;
    shm_read_result r;     /* r.status != 0: this read failed (see shm_map_batch, shm_server_error) */
} shm_served;

int shm_server_create(const shm_index *idx, const shm_map_params *params, const shm_server_config *cfg,
                      shm_server **out);
/* Queues reads (the bases are copied); thread-safe. */
int shm_server_submit(shm_server *s, uint64_t tag, const uint8_t *seq, int32_t len);
int shm_server_submit_many(shm_server *s, const uint64_t *tags, const shm_read *reads, int32_t n);
/* Queues reads without copying: the caller keeps each read's bytes alive and
 * unchanged until its result has been polled (or the server freed).  Reads
 * laid out back to back in one buffer reach the device in one copy per batch
 * (at full link speed when the buffer comes from shm_host_alloc). */
int shm_server_submit_borrowed(shm_server *s, const uint64_t *tags, const shm_read *reads, int32_t n);
/* Takes up to cap finished results, waiting up to timeout_ms (-1: no limit)
 * for at least min_results of them (fewer if nothing more is queued or running). */
int shm_server_poll(shm_server *s, shm_served *out, int32_t cap, int32_t min_results, int32_t timeout_ms,
                    int32_t *n);
/* Reads queued, being mapped, and finished but not yet polled. */
int shm_server_counts(shm_server *s, int64_t *queued, int64_t *running, int64_t *finished);
/* Batches the workers have mapped so far: [0] batches, [1] speculative
 * batches mapped again, [2] batches that allocated device or page-locked
 * memory, [3] graph captures (shm_mapper_last_counts [10], [11], [13]).
 * Writes up to n values. */
int shm_server_batch_stats(shm_server *s, int64_t *stats, int32_t n);
double shm_server_now(const shm_server *s);
/* The message of the first failed batch ("" if none). */
const char *shm_server_error(shm_server *s);
/* Waits for running batches, drops queued reads, frees everything. */
void shm_server_free(shm_server *s);

#ifdef __cplusplus
}
#endif
#endif
