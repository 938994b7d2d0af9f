// Kernel argument structures and launch constants shared by the gfx950 kernels
// and the host driver.
#pragma once

#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdint.h>

#include "../../include/shmap_hip.h"

// FracMinHash parameters (src/sketch.rs:58-132): k, the hash ratio and its
// keep threshold.
struct shm_sketcher {
    int32_t k;
    double h_frac;
    uint64_t thres;
};

namespace shm {

constexpr int LANE_W = 128;               // consecutive windows rolled by one lane
constexpr int RS_ITEMS = 16;              // radix sort keys per thread per tile
constexpr int RS_TILE = 256 * RS_ITEMS;   // radix sort tile
constexpr uint32_t SORT_CAP = 4096;       // (u64,u32) pairs sorted in LDS
constexpr int32_t MIN_HALFLEN = 5;        // src/buckets.rs:15
constexpr uint32_t SCAN_PARTS = 1024;     // workgroups of the device-wide scan
constexpr uint64_t SCAN_SINGLE_MAX = 8192; // arrays up to this length scan in one workgroup
constexpr uint32_t WAVE_SEED_CAP = 512;   // k-mers of a read seeded by one wave
constexpr uint32_t WAVE_SEED_LDS = WAVE_SEED_CAP * 14;  // per-wave LDS of k_seed_wave
constexpr int SEED_KEY_BITS_MAX = 55;     // hash bits in k_seed_wave's sort keys (with 9 bits of k-mer index)
static_assert(WAVE_SEED_CAP <= 512, "k_seed_wave's sort keys hold a k-mer index in 9 bits");
constexpr uint32_t TASK_COUNTERS = 8;       // per group: first-pass, second band, 3 refine rounds, prune list,
                                            // reads on the global-table bucket path, reads past one wave
constexpr uint32_t EXTRACT_WAVE_CAP = 512;  // buckets of a read sorted by one wave (in registers)

enum { METRIC_CONTAINMENT = 0, METRIC_JACCARD = 1, METRIC_BUCKET_SH = 2, METRIC_BUCKET_LCS = 3 };

// Halo units per sequence: the lane rolling a unit's LANE_W windows also reads
// the bases of the next ceil((k + 7) / LANE_W) units of its sequence (the k - 1
// bases of its last window plus a dword of read-ahead), so every sequence with
// windows gets that many extra units whose lanes only hold bases.
__host__ __device__ inline uint32_t sketch_halo(int32_t k) {
    return (uint32_t)((k + 7 + LANE_W - 1) / LANE_W);
}
// LDS bytes per wave: WAVE + halo unit slots of LANE_W / 8 code dwords, with one
// dword of bank skew per 16, then one 8-byte base offset per slot.
__host__ __device__ inline uint32_t sketch_codes_bytes(int32_t k) {
    return 4u * ((uint32_t)WAVE + sketch_halo(k)) * (uint32_t)(LANE_W / 8 + LANE_W / 128);
}
__host__ __device__ inline uint32_t sketch_span(int32_t k) {
    return (((sketch_codes_bytes(k) + 7u) & ~7u) + 8u * ((uint32_t)WAVE + sketch_halo(k)) + 15u) & ~15u;
}

// A set of sequences: sequence i is bases[off[i], off[i+1]); bytes up to
// bases + off[n] are readable.
struct SeqSet {
    const uint8_t *bases;
    const int64_t *off;
    int32_t n;
};

struct ShardGeom {
    uint64_t base[8];
    uint32_t bits[8];
};

// Everything the per-read stages read and write.  Arrays indexed by k-mer
// ("koff") belong to the whole batch; arrays indexed by "tab"/"bk" offsets
// belong to the current group of reads [r0, r0 + RG).
struct MapArgs {
    int32_t R;          // reads in the batch
    int32_t r0, RG;     // current group
    int32_t k;
    double theta, min_diff, max_overlap, thr_min;
    int32_t metric, no_bucket_pruning, one_sweep, abs_pos;
    const int64_t *roff;   // R+1 read offsets into the bases
    const int64_t *kstart; // reference segment k-mer starts (nseg+1)
    const int64_t *seg_cum;// reference segment base starts (nseg+1)
    int32_t nseg;
    // sketch of every read, at koff[rd]
    const uint64_t *koff;
    uint64_t *k_h;
    int32_t *k_r;
    uint8_t *k_s;
    uint64_t *ps_h;
    uint32_t *ps_v;
    // seeds, hash order, at koff[rd]
    uint64_t *sd_h;
    uint32_t *sd_first, *sd_occs, *sd_cnt, *sd_hoff, *sd_ord;
    uint8_t *sd_strand;
    uint64_t *sd_key;
    // seeds in rarest-first order, at koff[rd]: {hit offset, hit count,
    // occurrences | strand << 31, first examined position (consumed seeds)}
    uint4 *sd_row;
    // per read (whole batch)
    int32_t *rd_u, *rd_possible, *rd_S, *rd_hl, *rd_ok, *rd_c, *rd_bseeds, *rd_max_seed,
        *rd_seed_matches;
    int32_t *rd_z;  // consumed seeds without hits (they sort first): rd_z is the first with hits
    uint64_t *rd_E, *rd_cap;
    // per group read (local index), RG+1 entries
    const uint64_t *hit_base, *tab_base, *bk_base;
    uint64_t *grp_nb;  // each group read's bucket count (counted by k_contrib's inserts), scanned into bk_base
    // per-read bucket tables (group); tb_list holds each read's used slots
    uint32_t *tb_key, *tb_list;
    uint64_t *tb_mc;
    int32_t *tb_min, *tb_max;
    // buckets (group), at bk_base
    uint64_t *bk_key;
    uint32_t *bk_val;
    uint32_t *bs_g;
    int32_t *bs_m, *bs_c, *bs_min, *bs_max;
    uint8_t *bs_alive;
    double *bs_sh;
    double *sc_j;
    int32_t *sc_i, *sc_same, *sc_tl, *sc_tr, *sc_ssz, *sc_pend, *sc_seg;
    // refine tasks: (group read, sorted bucket) of every bucket that can survive
    // a match_rest pass.  Buckets whose bound reaches theta (first-pass
    // candidates) fill tasks[0, ..) (task_count[0]); those only the second pass
    // can reach (bound in [theta*(1-min_diff), theta)) fill tasks from the end,
    // tasks[cap-1-i] (task_count[1]).  rd_first holds each read's first
    // first-pass candidate in sorted order, rd_flag the first sorted bucket
    // scoring above theta (INT32_MAX if none).  Refinement runs in three rounds
    // (refine_pass 0-2) that refine a superset of what the two match_rest
    // passes read; each round's buckets are selected into sel, counted in
    // task_count[2 + round].  Before that, sel lists the buckets k_prune_ext
    // extends (task_count[5]), with k_prune's next seed and seeds so far in
    // sc_i and sc_same (refinement writes those later for what it refines).
    uint2 *tasks, *sel;
    uint32_t *task_count;
    uint32_t task_cap;
    int32_t refine_pass;
    // dynamic LDS bytes of a refinement workgroup: what the batch's largest
    // span and sketch need (refine_lds_bytes), so that more workgroups share a
    // CU when the reads are short; arrays past it go to global scratch
    uint32_t refine_lds;
    int32_t *rd_flag, *rd_first;
    // reads too large for one wave's seeding slice, for k_seed_wg
    uint32_t *big_list, *big_count;
    // group reads whose bucket tables exceed one wave's slice, for k_extract
    uint32_t *ext_list, *ext_count;
    // group reads past one wave's bucket count (listed by k_contrib_wave for
    // k_contrib_wg), and those whose buckets go through their global tables in
    // chunks (k_contrib_chunks, listed by k_contrib_wg); ov_kpre holds the
    // latter's chunk prefix and then the chunk length
    uint32_t *ov_list, *ov_count;
    uint32_t *gl_list, *gl_count;
    uint64_t *ov_kpre;
    uint32_t contrib_wave_e;  // examined hits one k_contrib_wave wave takes whole
    uint32_t seed_sh;         // k_seed_wave's sort keys: hash >> seed_sh (< 2^SEED_KEY_BITS_MAX) above the k-mer index
    uint32_t contrib_chunk;   // chunk length of the listed reads (0: chosen by k_contrib_over_scan)
    uint32_t contrib_wg_e;    // examined hits one k_contrib_wg workgroup takes whole
    uint32_t prune_seq;       // seeds one k_prune thread extends a bucket with
    // k_final's batch totals: the mapped reads (for the host's finalization)
    // and the sum of distinct k-mers
    uint32_t *mapped_list, *mapped_count;
    uint64_t *u_total;
    // per-workgroup scratch for spans/reads that overflow LDS
    int32_t *g_sid, *g_tr;     // sid_stride entries each
    uint64_t sid_stride;
    int32_t *g_diff, *g_tails; // diff_stride entries each
    uint64_t diff_stride;
    // set (nonzero) when a size computed on the device exceeds what the
    // batch's buffers were sized for: every later stage then returns at once
    // and the host maps the batch again with exact sizes (null: never set)
    const uint32_t *abort;
};

__device__ __forceinline__ bool aborted(const uint32_t *abort) {
    return abort && *(const volatile uint32_t *)abort;
}

}  // namespace shm

namespace shm {

struct DevIndex;

enum Stage {
    ST_SKETCH = 0,
    ST_SKETCH_COMPACT,
    ST_SEED_WAVE,
    ST_SEED_WG,
    ST_TAB_INIT,
    ST_CONTRIB,
    ST_EXTRACT,
    ST_PRUNE,
    ST_REFINE,
    ST_FINAL,
    ST_COUNT
};

// Sets the calling thread's shm_last_error() message (shm_lib.hip); allocates nothing.
void set_error(const char *fmt, va_list ap);

// host launch wrappers (shm_kernels.hip)
// p[0, n) = v, as a kernel (no memset node when a caller captures the stream).
void launch_fill_u32(uint32_t *p, uint64_t n, uint32_t v, hipStream_t st);
void launch_fill_u64(uint64_t *p, uint64_t n, uint64_t v, hipStream_t st);
// *abort = 1 when any *val[i] > cap[i] (val[i] null: unused), for the batch's
// capacity checks.
void launch_gate(const uint64_t *v0, uint64_t c0, const uint64_t *v1, uint64_t c1, const uint64_t *v2, uint64_t c2,
                 const uint64_t *v3, uint64_t c3, uint32_t *abort, hipStream_t st);
// Runs of 32-bit words to clear in one launch.
struct ZeroList {
    uint32_t *p[12];
    uint32_t n[12];
    int32_t count;
    void add(void *ptr, uint32_t words) {
        p[count] = (uint32_t *)ptr;
        n[count++] = words;
    }
};
void launch_zero_list(const ZeroList &z, hipStream_t st);
// Scalars (8 bytes when wide, else 4) gathered into one array in one launch.
struct GatherList {
    const void *src[32];
    uint8_t wide[32];
    int32_t count;
    void add(const void *ptr, bool w) {
        src[count] = ptr;
        wide[count++] = w;
    }
};
void launch_gather(const GatherList &g, uint64_t *dst, hipStream_t st);
void launch_scan(const uint64_t *in, uint64_t *out, const uint64_t *n_dev, uint64_t n_hint, uint64_t *part,
                 hipStream_t st);
// Units of LANE_W windows per sequence (plus its halo units; none without windows).
void launch_unit_counts(SeqSet S, int32_t k, uint64_t *nunit, hipStream_t st);
size_t sketch_lds(int32_t k);
// unit_seq[u] = the sequence of unit u, from the scanned unit counts.
void launch_unit_seq(int32_t n, const uint64_t *unit_start, uint32_t *unit_seq, hipStream_t st,
                     const uint32_t *abort = nullptr);
// A chunk is WAVE consecutive units (one per lane).  pass 0: roll every chunk
// into lane-strided staging (lcap 16-byte records per lane), listing chunks with
// a lane over lcap, each lane's count and its exclusive prefix in the chunk;
// pass 1: re-roll the listed chunks straight to their final positions
// (chunk_val = scanned chunk offsets).  With n_units_dev the unit count is
// read from the device and n_units, n_chunks are capacities (grid sizes).
void launch_sketch(int pass, SeqSet S, int32_t k, uint64_t thres, const uint64_t *unit_start,
                   const uint32_t *unit_seq, uint64_t n_units, uint64_t *chunk_val, uint32_t *lane_cnt,
                   uint32_t *lane_pre, uint4 *stage, uint64_t *out_h, int32_t *out_r, uint8_t *out_s, uint32_t lcap,
                   uint32_t *ovf_list, uint32_t *ovf_count, uint64_t n_chunks, uint32_t max_grid, hipStream_t st,
                   const uint32_t *abort = nullptr, const uint64_t *n_units_dev = nullptr);
void launch_sketch_compact(uint64_t n_chunks, const uint32_t *lane_cnt, const uint64_t *chunk_koff,
                           uint32_t lcap, const uint4 *stage, uint64_t *out_h, int32_t *out_r, uint8_t *out_s,
                           uint32_t max_grid, hipStream_t st, const uint32_t *abort = nullptr,
                           const uint64_t *n_units_dev = nullptr);
// Staging slots per lane for a keep threshold: the expected kept count of
// LANE_W windows plus five standard deviations and a margin, at most LANE_W.
uint32_t sketch_lane_cap(uint64_t thres);
void launch_seq_koff(int32_t n, const uint64_t *unit_start, const uint64_t *chunk_koff, const uint32_t *lane_pre,
                     uint64_t *koff, uint64_t *max_m, hipStream_t st, const uint32_t *abort = nullptr);
void launch_max_i32(const int32_t *a, int32_t n, uint64_t *out, hipStream_t st);
void launch_fill_pairs(const uint64_t *t_h, uint64_t n, uint64_t *keys, uint32_t *vals, hipStream_t st);
void launch_radix_pass(const uint64_t *kin, const uint32_t *vin, uint64_t *kout, uint32_t *vout, uint64_t n,
                       uint32_t shift, uint64_t *hist, uint64_t *offs, uint64_t *part, hipStream_t st);
void launch_mark_heads(const uint64_t *keys, uint64_t n, uint64_t *flag, hipStream_t st);
void launch_emit_keys(const uint64_t *keys, uint64_t n, const uint64_t *flag, const uint64_t *kid,
                      uint64_t *uk_h, uint64_t *uk_off, hipStream_t st);
void launch_fill_hits(const uint32_t *vals, uint64_t n, const int32_t *t_r, const uint8_t *t_s,
                      const int64_t *kstart, int32_t nseg, uint32_t *hit_g, uint32_t *hit_r, uint32_t *hit_seg,
                      uint64_t *sbits, hipStream_t st);
void launch_narrow_prefix(const uint64_t *in, uint64_t n, uint32_t *out, hipStream_t st);
void launch_shard_count(const uint64_t *uk_h, const uint64_t *uk_off, uint64_t nk, int32_t max_matches,
                        unsigned long long *shard_n, unsigned long long *stats, hipStream_t st);
void launch_insert(const uint64_t *uk_h, const uint64_t *uk_off, uint64_t nk, int32_t max_matches,
                   uint4 *table, ShardGeom geom, hipStream_t st);
// Dynamic LDS of a refinement workgroup for spans of at most 2 * max_hl
// positions and sketches of at most max_m k-mers (bounded by the largest a
// workgroup may take).
uint32_t refine_lds_bytes(uint64_t max_hl, uint64_t max_m, bool lcs);
void launch_stage(Stage stage, const MapArgs &A, const DevIndex &I, shm_read_result *d_out, uint64_t flat_n,
                  uint32_t max_grid, hipStream_t st);
void launch_lookup(const DevIndex &I, const uint64_t *hs, uint32_t n, uint32_t *cnt, uint32_t *off,
                   hipStream_t st);
// 2-bit packed bases to ASCII letters in out[0, nbases), then the exception runs
// (n_exc sorted [start, end) pairs of device int64) written as 'N'.  packed is
// 4-byte aligned and readable up to a whole dword past the last base.
void launch_unpack_bases(const uint8_t *packed, uint64_t nbases, const int64_t *exc, int64_t n_exc, uint8_t *out,
                         hipStream_t st);
// Counts into *bad (zeroed by the caller) the hits whose position or strand is
// not their k-mer's and the slot runs not strictly increasing in (segment,
// k-mer id); for an index read from a file, after the host's bounds checks.
void launch_validate_index(const DevIndex &I, uint64_t slots, unsigned long long *bad, hipStream_t st);

}  // namespace shm
