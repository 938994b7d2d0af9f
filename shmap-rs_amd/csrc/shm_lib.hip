// Host side of the gfx950 mapping library: the C ABI of include/shmap_hip.h.
//
// Mirrors shmap-rs's library surface for the mapping path -- FracMinHash
// (src/sketch.rs), SketchIndex (src/index.rs), create_mapper/Mapper
// (src/mapper.rs) -- with the work done by the kernels of shm_kernels.hip.
// PAF text follows src/mapping.rs.
#include <hip/hip_runtime.h>

#include <errno.h>
#include <unistd.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <memory>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/shmap_hip.h"
#include "shm_device.hpp"
#include "shm_kernels.hpp"

using namespace shm;

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------

// The calling thread's last failure message: a fixed buffer, so that setting
// it allocates nothing (shm_sketch_noalloc reports through it too).
static thread_local char g_err[1024];

void shm::set_error(const char *fmt, va_list ap) { vsnprintf(g_err, sizeof g_err, fmt, ap); }

static int fail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    shm::set_error(fmt, ap);
    va_end(ap);
    return code;
}

#define HIPCHK(expr)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail(e_ == hipErrorOutOfMemory ? SHM_ERR_NOMEM : SHM_ERR_HIP, "%s: %s (%s:%d)", \
                        #expr, hipGetErrorString(e_), __FILE__, __LINE__);                  \
    } while (0)

#define RCHK(expr)                \
    do {                          \
        int rc_ = (expr);         \
        if (rc_ != SHM_OK) return rc_; \
    } while (0)

extern "C" const char *shm_last_error(void) { return g_err; }

// ---------------------------------------------------------------------------
// device buffers
// ---------------------------------------------------------------------------

// Buffer allocations made on this thread (a batch reports its own).
static thread_local int64_t g_allocs = 0;

template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t cap = 0;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    int ensure(size_t n) {
        if (n <= cap && p) return SHM_OK;
        release();
        size_t c = n < 16 ? 16 : n + n / 4;
        g_allocs++;
        hipError_t e = hipMalloc((void **)&p, c * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            return fail(SHM_ERR_NOMEM, "hipMalloc of %zu bytes failed: %s", c * sizeof(T),
                        hipGetErrorString(e));
        }
        cap = c;
        return SHM_OK;
    }
};

// Page-locked host memory: device copies to and from it run at full link speed
// without the runtime's staging.
template <typename T>
struct PinnedBuf {
    T *p = nullptr;
    size_t cap = 0;
    ~PinnedBuf() { release(); }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
    }
    int ensure(size_t n) {
        if (n <= cap && p) return SHM_OK;
        release();
        size_t c = n < 16 ? 16 : n + n / 4;
        g_allocs++;
        hipError_t e = hipHostMalloc((void **)&p, c * sizeof(T), hipHostMallocDefault);
        if (e != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            return fail(SHM_ERR_NOMEM, "hipHostMalloc of %zu bytes failed: %s", c * sizeof(T),
                        hipGetErrorString(e));
        }
        cap = c;
        return SHM_OK;
    }
};

extern "C" int shm_host_alloc(size_t bytes, void **out) {
    if (!out) return fail(SHM_ERR_ARG, "null output");
    *out = nullptr;
    if (bytes == 0) bytes = 1;
    hipError_t e = hipHostMalloc(out, bytes, hipHostMallocDefault);
    if (e != hipSuccess) {
        *out = nullptr;
        (void)hipGetLastError();
        return fail(SHM_ERR_NOMEM, "hipHostMalloc of %zu bytes failed: %s", bytes, hipGetErrorString(e));
    }
    return SHM_OK;
}

extern "C" void shm_host_free(void *p) {
    if (p) (void)hipHostFree(p);
}

// FracMinHash threshold (src/sketch.rs:126-132): (h_frac * u64::MAX as f64) as
// u64 below 1, where u64::MAX as f64 is 2^64; u64::MAX otherwise.
static uint64_t hash_threshold(double h_frac) {
    if (h_frac < 1.0) {
        const double v = h_frac * 18446744073709551616.0;
        if (!(v > 0.0)) return 0;
        if (v >= 18446744073709551616.0) return UINT64_MAX;
        return (uint64_t)v;
    }
    return UINT64_MAX;
}

static uint32_t device_grid(int device) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return 1024;
    const uint32_t g = (uint32_t)prop.multiProcessorCount * 8;
    return g < 64 ? 64 : g;
}

// ---------------------------------------------------------------------------
// sketcher
// ---------------------------------------------------------------------------


extern "C" int shm_sketcher_create(int32_t k, double h_frac, shm_sketcher **out) {
    if (!out) return fail(SHM_ERR_ARG, "null output");
    if (k <= 0) return fail(SHM_ERR_ARG, "K-mer length (-k) should be positive. You provided %d.", k);
    if (k > 4096) return fail(SHM_ERR_ARG, "K-mer length (-k) above 4096 is not supported.");
    if (!(h_frac > 0.0) || h_frac > 1.0)
        return fail(SHM_ERR_ARG, "Given hash ratio (-r) should be between 0 and 1. You provided %g.", h_frac);
    shm_sketcher *s = new shm_sketcher;
    s->k = k;
    s->h_frac = h_frac;
    s->thres = hash_threshold(h_frac);
    *out = s;
    return SHM_OK;
}

extern "C" void shm_sketcher_free(shm_sketcher *sk) { delete sk; }

// Sketches a set of sequences already on the device: chunk table on the host,
// count pass, scan, write pass.  koff (n+1) receives per-sequence offsets.
struct SketchOut {
    DevBuf<uint64_t> nunit, unit_start, chunk_cnt, chunk_koff, koff, part, stat;  // stat: [0] largest sketch
    DevBuf<uint32_t> lane_cnt, lane_pre, unit_seq, ovf, ovf_n;  // overflow chunk list and its count
    DevBuf<uint4> stage;                        // lane-strided staging records of the rolling pass
    std::vector<uint64_t> h_koff;  // per-sequence offsets, when asked for
    PinnedBuf<uint64_t> h_sum;     // total, largest sketch, overflow chunks
    int syncs = 0;                 // host synchronisations so far
    uint64_t M = 0, max_m = 0;
    uint32_t n_overflow = 0;
};

// Units of a set of sequences, as k_unit_counts counts them (its halo units
// included), from host offsets.
static uint64_t count_units(int32_t k, const int64_t *h_off, int32_t n) {
    const uint64_t halo = sketch_halo(k);
    uint64_t U = 0;
    for (int32_t i = 0; i < n; i++) {
        const int64_t len = h_off[i + 1] - h_off[i];
        const int64_t nw = (len >= k && len >= 1) ? len - k + 1 : 0;
        if (nw > 0) U += (uint64_t)((nw + LANE_W - 1) / LANE_W) + halo;
    }
    return U;
}

// Sketches a set of sequences already on the device: count pass, scan, write
// pass.  koff (n+1) receives per-sequence offsets.  Sizes come one of three
// ways.  With an abort flag (the caller then clears so.ovf_n and so.stat
// beforehand) nothing is read back: the unit table holds u_cap
// units and the outputs m_cap + 1 k-mers, and more units, more k-mers or a
// sequence of more than mm_cap k-mers set *abort and skip every later write
// (the caller reads so.koff[n] and so.stat afterwards); the launches then
// depend on the capacities alone, so they can be captured into a graph.
// Otherwise the unit table is sized from the host's offsets (h_off) or from a
// read-back, the k-mer count is read back, and so.M, so.max_m and
// so.n_overflow are set.  so.syncs counts the host synchronisations.
static int sketch_device(const shm_sketcher *sk, const uint8_t *d_bases, const int64_t *d_off, int32_t n,
                         SketchOut &so, DevBuf<uint64_t> &out_h, DevBuf<int32_t> &out_r, DevBuf<uint8_t> &out_s,
                         uint32_t max_grid, hipStream_t st, hipEvent_t *ev_count, hipEvent_t *ev_write,
                         bool full_koff, const int64_t *h_off = nullptr, uint32_t *abort = nullptr,
                         uint64_t u_cap = 0, uint64_t m_cap = 0, uint64_t mm_cap = 0) {
    // unit table: units of LANE_W windows per sequence, scanned on the device;
    // a chunk is WAVE consecutive units
    SeqSet S{d_bases, d_off, n};
    RCHK(so.unit_start.ensure(n + 1));
    RCHK(so.nunit.ensure(n + 1));
    RCHK(so.koff.ensure(n + 1));
    RCHK(so.part.ensure(SCAN_PARTS + 1));
    RCHK(so.h_sum.ensure(3));
    RCHK(so.ovf_n.ensure(1));
    RCHK(so.stat.ensure(1));
    launch_unit_counts(S, sk->k, so.nunit.p, st);
    launch_scan(so.nunit.p, so.unit_start.p, nullptr, n, so.part.p, st);
    if (!abort) {  // (with an abort flag the caller clears them with its own counters)
        launch_fill_u32(so.ovf_n.p, 1, 0, st);
        launch_fill_u64(so.stat.p, 1, 0, st);
    }
    uint64_t U;                         // units, or their capacity
    const uint64_t *u_dev = nullptr;    // the unit count on the device, when U is a capacity
    if (abort) {
        U = u_cap;
        u_dev = so.unit_start.p + n;
        launch_gate(u_dev, u_cap, nullptr, 0, nullptr, 0, nullptr, 0, abort, st);
    } else if (h_off) {
        U = count_units(sk->k, h_off, n);
    } else {
        HIPCHK(hipMemcpyAsync(so.h_sum.p, so.unit_start.p + n, 8, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        so.syncs++;
        U = so.h_sum.p[0];
    }
    const uint64_t C = (U + WAVE - 1) / WAVE;
    RCHK(so.chunk_cnt.ensure(C + 1));
    RCHK(so.chunk_koff.ensure(C + 1));
    RCHK(so.lane_cnt.ensure(C * WAVE + 1));
    RCHK(so.lane_pre.ensure(C * WAVE + 1));
    RCHK(so.unit_seq.ensure(U + 1));
    if (C > 0) launch_unit_seq(n, so.unit_start.p, so.unit_seq.p, st, abort);
    const uint32_t lcap = sketch_lane_cap(sk->thres);
    const uint64_t slots = C * WAVE * (uint64_t)lcap;
    RCHK(so.stage.ensure(slots + 1));
    RCHK(so.ovf.ensure(C + 1));
    uint32_t *ovf_count = so.ovf_n.p, *ovf_list = so.ovf.p;
    if (ev_count) HIPCHK(hipEventRecord(ev_count[0], st));
    if (C > 0)
        launch_sketch(0, S, sk->k, sk->thres, so.unit_start.p, so.unit_seq.p, U, so.chunk_cnt.p, so.lane_cnt.p,
                      so.lane_pre.p, so.stage.p, nullptr, nullptr, nullptr, lcap, ovf_list, ovf_count, C, max_grid, st,
                      abort, u_dev);
    if (ev_count) HIPCHK(hipEventRecord(ev_count[1], st));
    // (with U a capacity, the chunk offsets past the last chunk are stale;
    // koff reads none of them)
    launch_scan(so.chunk_cnt.p, so.chunk_koff.p, nullptr, C, so.part.p, st);
    launch_seq_koff(n, so.unit_start.p, so.chunk_koff.p, so.lane_pre.p, so.koff.p, so.stat.p, st, abort);
    if (full_koff) {
        so.h_koff.resize(n + 1);
        HIPCHK(hipMemcpyAsync(so.h_koff.data(), so.koff.p, (n + 1) * 8, hipMemcpyDeviceToHost, st));
    }
    uint64_t M = m_cap;
    if (abort) {
        RCHK(out_h.ensure(m_cap + 1));
        RCHK(out_r.ensure(m_cap + 1));
        RCHK(out_s.ensure(m_cap + 1));
        launch_gate(so.koff.p + n, m_cap, so.stat.p, mm_cap, nullptr, 0, nullptr, 0, abort, st);
    } else {
        so.h_sum.p[2] = 0;
        HIPCHK(hipMemcpyAsync(so.h_sum.p, so.koff.p + n, 8, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(so.h_sum.p + 1, so.stat.p, 8, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(so.h_sum.p + 2, ovf_count, 4, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        so.syncs++;
        so.M = M = so.h_sum.p[0];
        so.max_m = so.h_sum.p[1];
        so.n_overflow = (uint32_t)so.h_sum.p[2];
        RCHK(out_h.ensure(M + 1));
        RCHK(out_r.ensure(M + 1));
        RCHK(out_s.ensure(M + 1));
    }
    if (ev_write) HIPCHK(hipEventRecord(ev_write[0], st));
    if (C > 0)
        launch_sketch_compact(C, so.lane_cnt.p, so.chunk_koff.p, lcap, so.stage.p, out_h.p, out_r.p, out_s.p,
                              max_grid, st, abort, u_dev);
    // the re-roll of overflowing chunks: their count is on the device when
    // the host has not read it back (the kernel takes it from there)
    if (C > 0 && (abort || so.n_overflow > 0))
        launch_sketch(1, S, sk->k, sk->thres, so.unit_start.p, so.unit_seq.p, U, so.chunk_koff.p, so.lane_cnt.p,
                      so.lane_pre.p, nullptr, out_h.p, out_r.p, out_s.p, lcap, ovf_list, ovf_count,
                      abort ? C : so.n_overflow, max_grid, st, abort, u_dev);
    if (ev_write) HIPCHK(hipEventRecord(ev_write[1], st));
    HIPCHK(hipGetLastError());
    return SHM_OK;
}

// A stream owned by one call, destroyed on every return path.
struct OwnedStream {
    hipStream_t st = nullptr;
    ~OwnedStream() {
        if (st) (void)hipStreamDestroy(st);
    }
};

extern "C" int shm_sketch(const shm_sketcher *sk, const uint8_t *seq, int64_t len, int32_t offset,
                          shm_kmer *out, int64_t cap, int64_t *n) {
    if (!sk || (!seq && len > 0) || !n || len < 0) return fail(SHM_ERR_ARG, "invalid argument");
    OwnedStream os;
    HIPCHK(hipStreamCreate(&os.st));
    hipStream_t st = os.st;
    DevBuf<uint8_t> bases;
    DevBuf<int64_t> off;
    RCHK(bases.ensure(len + 64));
    RCHK(off.ensure(2));
    std::vector<int64_t> h_off = {0, len};
    if (len > 0) HIPCHK(hipMemcpyAsync(bases.p, seq, len, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(off.p, h_off.data(), 16, hipMemcpyHostToDevice, st));
    SketchOut so;
    DevBuf<uint64_t> kh;
    DevBuf<int32_t> kr;
    DevBuf<uint8_t> ks;
    // (s.len() as RPos) < k yields nothing (src/sketch.rs:170-173)
    const bool empty = (int32_t)len < sk->k;
    if (!empty)
        RCHK(sketch_device(sk, bases.p, off.p, 1, so, kh, kr, ks, 1024, st, nullptr, nullptr, false));
    const int64_t M = empty ? 0 : (int64_t)so.M;
    *n = M;
    if (M > cap)
        return fail(SHM_ERR_CAPACITY, "sketch has %lld k-mers, buffer holds %lld", (long long)M, (long long)cap);
    if (M > 0) {
        std::vector<uint64_t> h(M);
        std::vector<int32_t> r(M);
        std::vector<uint8_t> s(M);
        HIPCHK(hipMemcpyAsync(h.data(), kh.p, M * 8, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(r.data(), kr.p, M * 4, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(s.data(), ks.p, M, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        for (int64_t i = 0; i < M; i++) {
            out[i].h = h[i];
            out[i].r = r[i] + offset;
            out[i].strand = s[i];
            out[i].pad[0] = out[i].pad[1] = out[i].pad[2] = 0;
        }
    }
    return SHM_OK;
}

// ---------------------------------------------------------------------------
// index
// ---------------------------------------------------------------------------

struct shm_index {
    int device = 0;
    int32_t k = 0;
    double h_frac = 0;
    int32_t max_matches = 0;  // -M (<= 0: none)
    int32_t nseg = 0;
    std::vector<std::string> names;
    std::vector<int32_t> sz;
    std::vector<int64_t> h_kstart, h_segcum;
    int64_t n_kmers = 0;
    int64_t indexed_kmers = 0, indexed_hits = 0, highest = 0, black_kmers = 0, black_hits = 0;
    DevBuf<uint64_t> t_h;
    DevBuf<int32_t> t_r;
    DevBuf<uint8_t> t_s;
    DevBuf<int64_t> kstart, segcum;
    DevBuf<uint32_t> hit_g, hit_r, hit_seg, strand_pre;
    DevBuf<uint4> table;
    ShardGeom geom;
    DevIndex view() const {
        DevIndex I;
        I.t_h = t_h.p;
        I.t_r = t_r.p;
        I.t_s = t_s.p;
        I.kstart = kstart.p;
        I.nseg = nseg;
        I.n_kmers = n_kmers;
        I.hit_g = hit_g.p;
        I.hit_r = hit_r.p;
        I.hit_seg = hit_seg.p;
        I.strand_pre = strand_pre.p;
        I.table = table.p;
        for (int i = 0; i < 8; i++) {
            I.shard_base[i] = geom.base[i];
            I.shard_bits[i] = geom.bits[i];
        }
        return I;
    }
};

// SketchIndex::build_index (src/index.rs:384-663), on the device: sketch every
// segment, radix-sort (hash, k-mer id), run-length encode, and insert each kept
// hash into its shard's open-addressing table.
extern "C" int shm_index_build(const shm_sketcher *sk, const shm_segment *segs, int32_t n_segments,
                               int32_t max_matches, int32_t device_id, shm_index **out) {
    if (!sk || !out || n_segments < 0 || (n_segments > 0 && !segs)) return fail(SHM_ERR_ARG, "invalid argument");
    // positions are i32 in the reference (RPos, src/types.rs) and 31 bits in
    // the hit records here; the PAF's target length is an i32 too
    for (int32_t s = 0; s < n_segments; s++)
        if (segs[s].len < 0 || segs[s].len > INT32_MAX || (segs[s].len > 0 && !segs[s].seq))
            return fail(SHM_ERR_ARG, "segment %d: length %lld is outside [0, 2^31-1] or its bases are missing", s,
                        (long long)segs[s].len);
    HIPCHK(hipSetDevice(device_id));
    shm_index *idx = new shm_index;
    idx->device = device_id;
    idx->k = sk->k;
    idx->h_frac = sk->h_frac;
    idx->max_matches = max_matches > 0 ? max_matches : 0;
    idx->nseg = n_segments;
    hipStream_t st;
    HIPCHK(hipStreamCreate(&st));
    const uint32_t max_grid = device_grid(device_id);

    std::vector<int64_t> h_off(n_segments + 1, 0);
    for (int32_t s = 0; s < n_segments; s++) {
        idx->names.push_back(segs[s].name ? segs[s].name : "");
        idx->sz.push_back((int32_t)segs[s].len);
        h_off[s + 1] = h_off[s] + segs[s].len;
    }
    idx->h_segcum = h_off;
    const int64_t total_bases = h_off[n_segments];
    DevBuf<uint8_t> bases;
    DevBuf<int64_t> d_off;
    int rc = SHM_OK;
    auto cleanup = [&](int code) {
        (void)hipStreamDestroy(st);
        if (code != SHM_OK) delete idx;
        return code;
    };
    if ((rc = bases.ensure(total_bases + 64)) != SHM_OK) return cleanup(rc);
    if ((rc = d_off.ensure(n_segments + 1)) != SHM_OK) return cleanup(rc);
    for (int32_t s = 0; s < n_segments; s++)
        if (segs[s].len > 0 &&
            hipMemcpyAsync(bases.p + h_off[s], segs[s].seq, segs[s].len, hipMemcpyHostToDevice, st) != hipSuccess)
            return cleanup(fail(SHM_ERR_HIP, "copying segment %d to the device failed", s));
    if (hipMemcpyAsync(d_off.p, h_off.data(), (n_segments + 1) * 8, hipMemcpyHostToDevice, st) != hipSuccess)
        return cleanup(fail(SHM_ERR_HIP, "copy failed"));

    // 1. sketch every segment (segments shorter than k contribute nothing)
    SketchOut so;
    if ((rc = sketch_device(sk, bases.p, d_off.p, n_segments, so, idx->t_h, idx->t_r, idx->t_s, max_grid, st,
                            nullptr, nullptr, true)) != SHM_OK)
        return cleanup(rc);
    const uint64_t N = so.h_koff[n_segments];
    idx->n_kmers = (int64_t)N;
    idx->h_kstart.assign(so.h_koff.begin(), so.h_koff.end());
    if ((rc = idx->kstart.ensure(n_segments + 1)) != SHM_OK) return cleanup(rc);
    if ((rc = idx->segcum.ensure(n_segments + 1)) != SHM_OK) return cleanup(rc);
    if (hipMemcpyAsync(idx->kstart.p, idx->h_kstart.data(), (n_segments + 1) * 8, hipMemcpyHostToDevice, st) !=
            hipSuccess ||
        hipMemcpyAsync(idx->segcum.p, idx->h_segcum.data(), (n_segments + 1) * 8, hipMemcpyHostToDevice, st) !=
            hipSuccess)
        return cleanup(fail(SHM_ERR_HIP, "copy failed"));
    bases.release();
    if (N >= 0xffffffffull) return cleanup(fail(SHM_ERR_CAPACITY, "more than 2^32-1 indexed k-mers"));

    // 2. stable radix sort of (hash, k-mer id) by hash
    {
        DevBuf<uint64_t> k0, k1, hist, offs, part;
        DevBuf<uint32_t> v0, v1;
        if ((rc = k0.ensure(N + 1)) || (rc = k1.ensure(N + 1)) || (rc = v0.ensure(N + 1)) ||
            (rc = v1.ensure(N + 1)) || (rc = part.ensure(SCAN_PARTS + 1)))
            return cleanup(rc);
        const uint64_t ntiles = (N + RS_TILE - 1) / RS_TILE;
        if ((rc = hist.ensure(ntiles * 256 + 1)) || (rc = offs.ensure(ntiles * 256 + 1))) return cleanup(rc);
        if (N > 0) launch_fill_pairs(idx->t_h.p, N, k0.p, v0.p, st);
        const uint64_t thres = sk->thres;
        const int key_bits = thres == 0 ? 1 : 64 - __builtin_clzll(thres);
        uint64_t *ka = k0.p, *kb = k1.p;
        uint32_t *va = v0.p, *vb = v1.p;
        for (int shift = 0; shift < key_bits && N > 1; shift += 8) {
            launch_radix_pass(ka, va, kb, vb, N, (uint32_t)shift, hist.p, offs.p, part.p, st);
            std::swap(ka, kb);
            std::swap(va, vb);
        }
        // 3. run-length encode: key ids, per-key offsets
        DevBuf<uint64_t> flag, kid, uk_h, uk_off;
        if ((rc = flag.ensure(N + 1)) || (rc = kid.ensure(N + 1))) return cleanup(rc);
        if (N > 0) launch_mark_heads(ka, N, flag.p, st);
        launch_scan(flag.p, kid.p, nullptr, N, part.p, st);
        uint64_t U = 0;
        if (hipMemcpyAsync(&U, kid.p + N, 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            return cleanup(fail(SHM_ERR_HIP, "sync failed"));
        if ((rc = uk_h.ensure(U + 1)) || (rc = uk_off.ensure(U + 1))) return cleanup(rc);
        if (N > 0) launch_emit_keys(ka, N, flag.p, kid.p, uk_h.p, uk_off.p, st);
        else if (hipMemsetAsync(uk_off.p, 0, 8, st) != hipSuccess) return cleanup(fail(SHM_ERR_HIP, "memset"));
        // 4. hit arrays in (hash, segm_id, r) order and the strand prefix
        if ((rc = idx->hit_g.ensure(N + 1)) || (rc = idx->hit_r.ensure(N + 1)) || (rc = idx->hit_seg.ensure(N + 1)) ||
            (rc = idx->strand_pre.ensure(N + 1)))
            return cleanup(rc);
        DevBuf<uint64_t> sbits, spre;
        if ((rc = sbits.ensure(N + 1)) || (rc = spre.ensure(N + 1))) return cleanup(rc);
        if (N > 0)
            launch_fill_hits(va, N, idx->t_r.p, idx->t_s.p, idx->kstart.p, n_segments, idx->hit_g.p, idx->hit_r.p,
                             idx->hit_seg.p, sbits.p, st);
        launch_scan(sbits.p, spre.p, nullptr, N, part.p, st);
        launch_narrow_prefix(spre.p, N, idx->strand_pre.p, st);
        // 5. shards: count kept keys per shard (and -M statistics)
        DevBuf<unsigned long long> shard_n, stats;
        if ((rc = shard_n.ensure(8)) || (rc = stats.ensure(4))) return cleanup(rc);
        if (hipMemsetAsync(shard_n.p, 0, 8 * 8, st) != hipSuccess || hipMemsetAsync(stats.p, 0, 4 * 8, st) != hipSuccess)
            return cleanup(fail(SHM_ERR_HIP, "memset"));
        if (U > 0) launch_shard_count(uk_h.p, uk_off.p, U, max_matches, shard_n.p, stats.p, st);
        unsigned long long hn[8], hs[4];
        if (hipMemcpyAsync(hn, shard_n.p, 64, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipMemcpyAsync(hs, stats.p, 32, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            return cleanup(fail(SHM_ERR_HIP, "sync failed"));
        uint64_t tot = 0;
        for (int s = 0; s < 8; s++) {
            uint32_t bits = 6;
            while ((1ull << bits) < 4ull * hn[s] + 1) bits++;  // load <= 1/4: short probe chains
            idx->geom.bits[s] = bits;
            idx->geom.base[s] = tot;
            tot += 1ull << bits;
        }
        if ((rc = idx->table.ensure(tot)) != SHM_OK) return cleanup(rc);
        if (hipMemsetAsync(idx->table.p, 0, tot * sizeof(uint4), st) != hipSuccess)
            return cleanup(fail(SHM_ERR_HIP, "memset"));
        if (U > 0) launch_insert(uk_h.p, uk_off.p, U, max_matches, idx->table.p, idx->geom, st);
        if (hipStreamSynchronize(st) != hipSuccess || hipGetLastError() != hipSuccess)
            return cleanup(fail(SHM_ERR_HIP, "index build kernels failed"));
        // get_kmer_stats + erase_frequent_kmers counters (src/index.rs:343-376)
        idx->indexed_kmers = (int64_t)U;
        idx->indexed_hits = (int64_t)N;
        idx->highest = (int64_t)hs[2];
        idx->black_kmers = (int64_t)hs[0];
        idx->black_hits = (int64_t)hs[1];
    }
    *out = idx;
    (void)hipStreamDestroy(st);
    return SHM_OK;
}

extern "C" void shm_index_free(shm_index *idx) { delete idx; }

static int lookup_one(const shm_index *idx, uint64_t h, uint32_t *cnt, uint32_t *off) {
    HIPCHK(hipSetDevice(idx->device));
    DevBuf<uint64_t> dh;
    DevBuf<uint32_t> dc, dofs;
    RCHK(dh.ensure(1));
    RCHK(dc.ensure(1));
    RCHK(dofs.ensure(1));
    HIPCHK(hipMemcpy(dh.p, &h, 8, hipMemcpyHostToDevice));
    launch_lookup(idx->view(), dh.p, 1, dc.p, dofs.p, 0);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpy(cnt, dc.p, 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(off, dofs.p, 4, hipMemcpyDeviceToHost));
    return SHM_OK;
}

extern "C" int shm_index_count(const shm_index *idx, uint64_t h, int32_t *count) {
    if (!idx || !count) return fail(SHM_ERR_ARG, "invalid argument");
    uint32_t c, o;
    RCHK(lookup_one(idx, h, &c, &o));
    *count = (int32_t)c;
    return SHM_OK;
}

extern "C" int shm_index_hits(const shm_index *idx, uint64_t h, shm_hit *out, int32_t cap, int32_t *n) {
    if (!idx || !n) return fail(SHM_ERR_ARG, "invalid argument");
    uint32_t c, o;
    RCHK(lookup_one(idx, h, &c, &o));
    *n = (int32_t)c;
    if ((int32_t)c > cap) return fail(SHM_ERR_CAPACITY, "%u hits, buffer holds %d", c, cap);
    if (c == 0) return SHM_OK;
    std::vector<uint32_t> g(c), r(c), s(c);
    HIPCHK(hipMemcpy(g.data(), idx->hit_g.p + o, c * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(r.data(), idx->hit_r.p + o, c * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(s.data(), idx->hit_seg.p + o, c * 4, hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < c; i++) {
        out[i].r = (int32_t)(r[i] & 0x7fffffffu);
        out[i].strand = (int32_t)(r[i] >> 31);
        out[i].segm_id = (int32_t)s[i];
        out[i].tpos = (int32_t)((int64_t)g[i] - idx->h_kstart[s[i]]);
    }
    return SHM_OK;
}

extern "C" int shm_index_n_segments(const shm_index *idx, int32_t *n) {
    if (!idx || !n) return fail(SHM_ERR_ARG, "invalid argument");
    *n = idx->nseg;
    return SHM_OK;
}

extern "C" int shm_index_segment(const shm_index *idx, int32_t s, const char **name, int32_t *sz,
                                 int64_t *n_kmers) {
    if (!idx || s < 0 || s >= idx->nseg) return fail(SHM_ERR_ARG, "segment id out of range");
    if (name) *name = idx->names[s].c_str();
    if (sz) *sz = idx->sz[s];
    if (n_kmers) *n_kmers = idx->h_kstart[s + 1] - idx->h_kstart[s];
    return SHM_OK;
}

extern "C" int shm_index_stats(const shm_index *idx, int64_t *ik, int64_t *ih, int64_t *hf, int64_t *bk,
                               int64_t *bh) {
    if (!idx) return fail(SHM_ERR_ARG, "invalid argument");
    if (ik) *ik = idx->indexed_kmers;
    if (ih) *ih = idx->indexed_hits;
    if (hf) *hf = idx->highest;
    if (bk) *bk = idx->black_kmers;
    if (bh) *bh = idx->black_hits;
    return SHM_OK;
}

extern "C" int shm_index_params(const shm_index *idx, int32_t *k, double *h_frac, int32_t *max_matches) {
    if (!idx) return fail(SHM_ERR_ARG, "invalid argument");
    if (k) *k = idx->k;
    if (h_frac) *h_frac = idx->h_frac;
    if (max_matches) *max_matches = idx->max_matches;
    return SHM_OK;
}

// ---------------------------------------------------------------------------
// index file: the build parameters, the segment table and every device array,
// little-endian, between a magic and an end marker
// ---------------------------------------------------------------------------

static const char IDX_MAGIC[8] = {'S', 'H', 'M', 'I', 'D', 'X', '0', '1'};
static const char IDX_END[8] = {'S', 'H', 'M', 'I', 'D', 'X', 'E', 'N'};
static const uint32_t IDX_VERSION = 1;
static const size_t IDX_STAGE = 64ull << 20;  // bytes per staged copy

struct IdxHeader {
    char magic[8];
    uint32_t version, nseg;
    int32_t k, max_matches;
    double h_frac;
    int64_t n_kmers, indexed_kmers, indexed_hits, highest, black_kmers, black_hits;
    uint64_t shard_base[8];
    uint32_t shard_bits[8];
    uint64_t table_slots, name_bytes;
};
static_assert(sizeof(IdxHeader) == 192, "index file header layout");

static int idx_write(FILE *f, const void *p, size_t n) {
    if (n && fwrite(p, 1, n, f) != n) return fail(SHM_ERR_IO, "write failed: %s", strerror(errno));
    return SHM_OK;
}

static int idx_read(FILE *f, void *p, size_t n, const char *path) {
    if (n && fread(p, 1, n, f) != n) return fail(SHM_ERR_ARG, "\"%s\" is truncated", path);
    return SHM_OK;
}

// device array -> file, through a page-locked stage
template <typename T>
static int idx_put(FILE *f, const T *d, size_t n, PinnedBuf<uint8_t> &stage) {
    size_t bytes = n * sizeof(T), done = 0;
    while (done < bytes) {
        const size_t c = bytes - done < IDX_STAGE ? bytes - done : IDX_STAGE;
        HIPCHK(hipMemcpy(stage.p, (const uint8_t *)d + done, c, hipMemcpyDeviceToHost));
        RCHK(idx_write(f, stage.p, c));
        done += c;
    }
    return SHM_OK;
}

// file -> device array (allocated with n + extra entries)
template <typename T>
static int idx_get(FILE *f, DevBuf<T> &d, size_t n, size_t extra, PinnedBuf<uint8_t> &stage, const char *path) {
    RCHK(d.ensure(n + extra));
    size_t bytes = n * sizeof(T), done = 0;
    while (done < bytes) {
        const size_t c = bytes - done < IDX_STAGE ? bytes - done : IDX_STAGE;
        RCHK(idx_read(f, stage.p, c, path));
        HIPCHK(hipMemcpy((uint8_t *)d.p + done, stage.p, c, hipMemcpyHostToDevice));
        done += c;
    }
    return SHM_OK;
}

extern "C" int shm_index_save(const shm_index *idx, const char *path) {
    if (!idx || !path) return fail(SHM_ERR_ARG, "invalid argument");
    HIPCHK(hipSetDevice(idx->device));
    IdxHeader h{};
    memcpy(h.magic, IDX_MAGIC, 8);
    h.version = IDX_VERSION;
    h.nseg = (uint32_t)idx->nseg;
    h.k = idx->k;
    h.max_matches = idx->max_matches;
    h.h_frac = idx->h_frac;
    h.n_kmers = idx->n_kmers;
    h.indexed_kmers = idx->indexed_kmers;
    h.indexed_hits = idx->indexed_hits;
    h.highest = idx->highest;
    h.black_kmers = idx->black_kmers;
    h.black_hits = idx->black_hits;
    for (int i = 0; i < 8; i++) {
        h.shard_base[i] = idx->geom.base[i];
        h.shard_bits[i] = idx->geom.bits[i];
        h.table_slots += 1ull << idx->geom.bits[i];
    }
    for (const std::string &nm : idx->names) h.name_bytes += 4 + nm.size();
    FILE *f = fopen(path, "wb");
    if (!f) return fail(SHM_ERR_IO, "cannot create \"%s\": %s", path, strerror(errno));
    PinnedBuf<uint8_t> stage;
    int rc = stage.ensure(IDX_STAGE);
    const size_t N = (size_t)idx->n_kmers;
    if (rc == SHM_OK) rc = idx_write(f, &h, sizeof h);
    for (size_t i = 0; rc == SHM_OK && i < idx->names.size(); i++) {
        const uint32_t len = (uint32_t)idx->names[i].size();
        rc = idx_write(f, &len, 4);
        if (rc == SHM_OK) rc = idx_write(f, idx->names[i].data(), len);
    }
    if (rc == SHM_OK) rc = idx_write(f, idx->sz.data(), idx->sz.size() * 4);
    if (rc == SHM_OK) rc = idx_write(f, idx->h_kstart.data(), idx->h_kstart.size() * 8);
    if (rc == SHM_OK) rc = idx_write(f, idx->h_segcum.data(), idx->h_segcum.size() * 8);
    if (rc == SHM_OK) rc = idx_put(f, idx->t_h.p, N, stage);
    if (rc == SHM_OK) rc = idx_put(f, idx->t_r.p, N, stage);
    if (rc == SHM_OK) rc = idx_put(f, idx->t_s.p, N, stage);
    if (rc == SHM_OK) rc = idx_put(f, idx->hit_g.p, N, stage);
    if (rc == SHM_OK) rc = idx_put(f, idx->hit_r.p, N, stage);
    if (rc == SHM_OK) rc = idx_put(f, idx->hit_seg.p, N, stage);
    if (rc == SHM_OK) rc = idx_put(f, idx->strand_pre.p, N + 1, stage);
    if (rc == SHM_OK) rc = idx_put(f, idx->table.p, (size_t)h.table_slots, stage);
    if (rc == SHM_OK) rc = idx_write(f, IDX_END, 8);
    if (fclose(f) != 0 && rc == SHM_OK) rc = fail(SHM_ERR_IO, "write failed: %s", strerror(errno));
    if (rc != SHM_OK) remove(path);
    return rc;
}

// Reads n elements at file offset `at` in pieces of at most `step`, handing
// each piece to fn(first element index, pointer, count); false from fn stops.
template <typename T, typename F>
static int idx_scan(int fd, uint64_t at, uint64_t n, std::vector<T> &buf, const char *path, F fn) {
    const uint64_t step = buf.size();
    for (uint64_t i = 0; i < n; i += step) {
        const uint64_t c = n - i < step ? n - i : step;
        const size_t bytes = (size_t)c * sizeof(T);
        size_t got = 0;
        while (got < bytes) {
            const ssize_t r = pread(fd, (uint8_t *)buf.data() + got, bytes - got, (off_t)(at + i * sizeof(T) + got));
            if (r <= 0) return fail(SHM_ERR_IO, "cannot read \"%s\"", path);
            got += (size_t)r;
        }
        if (!fn(i, buf.data(), c)) return fail(SHM_ERR_ARG, "\"%s\" is inconsistent", path);
    }
    return SHM_OK;
}

// The host's checks of an index file's arrays (at file offset `at`: t_h, t_r,
// t_s, hit_g, hit_r, hit_seg, strand_pre, table), before any of it reaches the
// device: sketch positions strictly increasing inside each segment and within
// its length, strands 0/1; every hit's segment in range and its k-mer id inside
// that segment; the strand prefix counting exactly the reverse-strand hits;
// every slot's run inside the hit arrays, the runs' lengths summing to at most
// the hit count, and an empty slot in every shard, so that each probe ends.
static int check_index_arrays(int fd, uint64_t at, const IdxHeader &h, const shm_index &ix, const char *path) {
    const uint64_t N = (uint64_t)h.n_kmers, S = h.nseg;
    const uint64_t o_tr = at + N * 8, o_ts = o_tr + N * 4, o_hg = o_ts + N, o_hr = o_hg + N * 4,
                   o_hs = o_hr + N * 4, o_sp = o_hs + N * 4, o_tab = o_sp + (N + 1) * 4;
    const uint64_t STEP = 1u << 20;
    {
        std::vector<int32_t> tr(STEP);
        uint64_t seg = 0;
        int32_t prev = -1;
        RCHK(idx_scan(fd, o_tr, N, tr, path, [&](uint64_t i0, const int32_t *v, uint64_t c) {
            for (uint64_t j = 0; j < c; j++) {
                const uint64_t g = i0 + j;
                while (seg < S && (uint64_t)ix.h_kstart[seg + 1] <= g) {
                    seg++;
                    prev = -1;
                }
                if (seg >= S || v[j] <= prev || v[j] >= ix.sz[seg]) return false;
                prev = v[j];
            }
            return true;
        }));
        std::vector<uint8_t> ts(STEP);
        RCHK(idx_scan(fd, o_ts, N, ts, path, [&](uint64_t, const uint8_t *v, uint64_t c) {
            for (uint64_t j = 0; j < c; j++)
                if (v[j] > 1) return false;
            return true;
        }));
    }
    {
        std::vector<uint32_t> hg(STEP), hr(STEP), hs(STEP), sp(STEP);
        uint32_t sp0 = 0;
        if (pread(fd, &sp0, 4, (off_t)o_sp) != 4) return fail(SHM_ERR_IO, "cannot read \"%s\"", path);
        if (sp0 != 0) return fail(SHM_ERR_ARG, "\"%s\" is inconsistent", path);
        // hit_g, hit_r and hit_seg in step; strand_pre one ahead
        RCHK(idx_scan(fd, o_hg, N, hg, path, [&](uint64_t i0, const uint32_t *g, uint64_t c) {
            if (idx_scan(fd, o_hr + i0 * 4, c, hr, path, [](uint64_t, const uint32_t *, uint64_t) { return true; }) ||
                idx_scan(fd, o_hs + i0 * 4, c, hs, path, [](uint64_t, const uint32_t *, uint64_t) { return true; }) ||
                idx_scan(fd, o_sp + (i0 + 1) * 4, c, sp, path, [](uint64_t, const uint32_t *, uint64_t) { return true; }))
                return false;
            for (uint64_t j = 0; j < c; j++) {
                const uint32_t sg = hs[j];
                if (sg >= S || (int64_t)g[j] < ix.h_kstart[sg] || (int64_t)g[j] >= ix.h_kstart[sg + 1]) return false;
                const uint32_t before = j ? sp[j - 1] : sp0;
                if (sp[j] != before + (hr[j] >> 31)) return false;
            }
            sp0 = sp[c - 1];
            return true;
        }));
    }
    {
        std::vector<uint4> tab(STEP);
        uint64_t total = 0;
        int shard = 0;
        bool empty_seen = false;
        RCHK(idx_scan(fd, o_tab, h.table_slots, tab, path, [&](uint64_t i0, const uint4 *t, uint64_t c) {
            for (uint64_t j = 0; j < c; j++) {
                const uint64_t slot = i0 + j;
                while (shard < 7 && slot >= h.shard_base[shard + 1]) {
                    if (!empty_seen) return false;
                    shard++;
                    empty_seen = false;
                }
                if (t[j].w == 0) {
                    empty_seen = true;
                    continue;
                }
                total += t[j].w;
                if ((uint64_t)t[j].z + t[j].w > N || total > N) return false;
            }
            return true;
        }));
        if (!empty_seen) return fail(SHM_ERR_ARG, "\"%s\" is inconsistent", path);
    }
    return SHM_OK;
}

extern "C" int shm_index_load(const char *path, int32_t device_id, shm_index **out) {
    if (!path || !out) return fail(SHM_ERR_ARG, "invalid argument");
    FILE *f = fopen(path, "rb");
    if (!f) return fail(SHM_ERR_IO, "cannot open \"%s\": %s", path, strerror(errno));
    struct Closer {
        FILE *f;
        ~Closer() { fclose(f); }
    } closer{f};
    // the header is checked before the device is touched
    IdxHeader h{};
    if (fread(&h, 1, sizeof h, f) != sizeof h || memcmp(h.magic, IDX_MAGIC, 8) != 0)
        return fail(SHM_ERR_ARG, "\"%s\" is not a shmap index file", path);
    if (h.version != IDX_VERSION)
        return fail(SHM_ERR_ARG, "\"%s\" is an index file of version %u; this build reads version %u", path,
                    h.version, IDX_VERSION);
    if (h.n_kmers < 0 || (uint64_t)h.n_kmers >= 0xffffffffull || h.nseg > (1u << 30))
        return fail(SHM_ERR_ARG, "\"%s\" has an invalid header", path);
    // the parameters shm_sketcher_create accepts (the mapper's kernels take k
    // and the threshold from here)
    if (h.k <= 0 || h.k > 4096 || !(h.h_frac > 0.0) || h.h_frac > 1.0 || h.max_matches < 0)
        return fail(SHM_ERR_ARG, "\"%s\" has invalid index parameters (k=%d r=%g M=%d)", path, h.k, h.h_frac,
                    h.max_matches);
    // the device probes use 32-bit slot positions and masks: at most 2^31 slots
    // a shard (the build makes far fewer)
    uint64_t slots = 0;
    for (int i = 0; i < 8; i++) {
        if (h.shard_bits[i] < 6 || h.shard_bits[i] > 31 || h.shard_base[i] != slots)
            return fail(SHM_ERR_ARG, "\"%s\" has an invalid header", path);
        slots += 1ull << h.shard_bits[i];
    }
    if (slots != h.table_slots || h.name_bytes > (1ull << 40))
        return fail(SHM_ERR_ARG, "\"%s\" has an invalid header", path);
    // the whole file must be there: header, names, tables, arrays, end marker
    const uint64_t N = (uint64_t)h.n_kmers, S = h.nseg;
    const uint64_t want = sizeof h + h.name_bytes + S * 4 + 2 * (S + 1) * 8 + N * (8 + 4 + 1 + 4 * 3) + (N + 1) * 4 +
                          h.table_slots * 16 + 8;
    if (fseeko(f, 0, SEEK_END) != 0) return fail(SHM_ERR_IO, "cannot read \"%s\"", path);
    const uint64_t have = (uint64_t)ftello(f);
    if (have != want)
        return fail(SHM_ERR_ARG, "\"%s\" is %s (%llu bytes, the header describes %llu)", path,
                    have < want ? "truncated" : "longer than its header describes", (unsigned long long)have,
                    (unsigned long long)want);
    char end[8];
    if (fseeko(f, (off_t)(want - 8), SEEK_SET) != 0 || fread(end, 1, 8, f) != 8 || memcmp(end, IDX_END, 8) != 0)
        return fail(SHM_ERR_ARG, "\"%s\" has no end marker", path);
    if (fseeko(f, (off_t)sizeof h, SEEK_SET) != 0) return fail(SHM_ERR_IO, "cannot read \"%s\"", path);
    std::unique_ptr<shm_index> idx(new shm_index);
    idx->device = device_id;
    idx->k = h.k;
    idx->h_frac = h.h_frac;
    idx->max_matches = h.max_matches;
    idx->nseg = (int32_t)S;
    idx->n_kmers = h.n_kmers;
    idx->indexed_kmers = h.indexed_kmers;
    idx->indexed_hits = h.indexed_hits;
    idx->highest = h.highest;
    idx->black_kmers = h.black_kmers;
    idx->black_hits = h.black_hits;
    for (int i = 0; i < 8; i++) {
        idx->geom.base[i] = h.shard_base[i];
        idx->geom.bits[i] = h.shard_bits[i];
    }
    uint64_t name_left = h.name_bytes;
    for (uint64_t i = 0; i < S; i++) {
        uint32_t len = 0;
        RCHK(idx_read(f, &len, 4, path));
        if ((uint64_t)len + 4 > name_left) return fail(SHM_ERR_ARG, "\"%s\" has an invalid segment table", path);
        name_left -= 4 + (uint64_t)len;
        std::string nm(len, '\0');
        RCHK(idx_read(f, &nm[0], len, path));
        idx->names.push_back(std::move(nm));
    }
    if (name_left != 0) return fail(SHM_ERR_ARG, "\"%s\" has an invalid segment table", path);
    idx->sz.resize(S);
    idx->h_kstart.resize(S + 1);
    idx->h_segcum.resize(S + 1);
    RCHK(idx_read(f, idx->sz.data(), S * 4, path));
    RCHK(idx_read(f, idx->h_kstart.data(), (S + 1) * 8, path));
    RCHK(idx_read(f, idx->h_segcum.data(), (S + 1) * 8, path));
    if (idx->h_kstart[S] != h.n_kmers) return fail(SHM_ERR_ARG, "\"%s\" has an invalid segment table", path);
    // segment lengths and base offsets agree, k-mer starts never decrease
    if (idx->h_kstart[0] != 0 || idx->h_segcum[0] != 0)
        return fail(SHM_ERR_ARG, "\"%s\" has an invalid segment table", path);
    for (uint64_t s = 0; s < S; s++)
        if (idx->sz[s] < 0 || idx->h_segcum[s + 1] - idx->h_segcum[s] != idx->sz[s] ||
            idx->h_kstart[s + 1] < idx->h_kstart[s])
            return fail(SHM_ERR_ARG, "\"%s\" has an invalid segment table", path);
    // every array the device will index through, checked on the host before
    // the device is touched
    RCHK(check_index_arrays(fileno(f), (uint64_t)ftello(f), h, *idx, path));
    HIPCHK(hipSetDevice(device_id));
    PinnedBuf<uint8_t> stage;
    RCHK(stage.ensure(IDX_STAGE));
    RCHK(idx->kstart.ensure(S + 1));
    RCHK(idx->segcum.ensure(S + 1));
    HIPCHK(hipMemcpy(idx->kstart.p, idx->h_kstart.data(), (S + 1) * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(idx->segcum.p, idx->h_segcum.data(), (S + 1) * 8, hipMemcpyHostToDevice));
    RCHK(idx_get(f, idx->t_h, N, 1, stage, path));
    RCHK(idx_get(f, idx->t_r, N, 1, stage, path));
    RCHK(idx_get(f, idx->t_s, N, 1, stage, path));
    RCHK(idx_get(f, idx->hit_g, N, 1, stage, path));
    RCHK(idx_get(f, idx->hit_r, N, 1, stage, path));
    RCHK(idx_get(f, idx->hit_seg, N, 1, stage, path));
    RCHK(idx_get(f, idx->strand_pre, N + 1, 0, stage, path));
    RCHK(idx_get(f, idx->table, (size_t)h.table_slots, 0, stage, path));
    // what the host cannot check in one pass (random access): each hit carries
    // its k-mer's position and strand, and every slot's run is ordered
    {
        DevBuf<unsigned long long> bad;
        RCHK(bad.ensure(1));
        HIPCHK(hipMemset(bad.p, 0, 8));
        launch_validate_index(idx->view(), h.table_slots, bad.p, 0);
        HIPCHK(hipGetLastError());
        unsigned long long nbad = 0;
        HIPCHK(hipMemcpy(&nbad, bad.p, 8, hipMemcpyDeviceToHost));
        if (nbad) return fail(SHM_ERR_ARG, "\"%s\" is inconsistent (%llu hit records disagree with the sketch "
                                           "or are out of order)", path, nbad);
    }
    *out = idx.release();
    return SHM_OK;
}

// ---------------------------------------------------------------------------
// mapper
// ---------------------------------------------------------------------------

static const char *kStageNames[ST_COUNT] = {"sketch", "sketch_compact", "seed_wave", "seed_wg", "tab_init",
                                            "contrib", "extract",        "prune",     "refine",  "final"};

constexpr int SHM_N_COUNTS = 14;  // see shm_mapper_last_counts

// The sizes batches have had so far, from which a batch's buffers are sized
// before its own sizes are known (a speculative batch): the largest k-mers,
// hits and buckets per base, k-mers of the largest sketch per base of the
// longest read, and the capacities (in items) the buffers of speculative
// batches were last sized to.  Ratios per base come from batches large enough
// to show them (or the first batch), so that a batch of a few reads does not
// inflate every later batch's buffers.
struct BatchShape {
    bool seen = false;
    double m_per_base = 0, e_per_base = 0, b_per_base = 0, mm_per_base = 0;
    uint64_t U = 0, M = 0, T = 0, B = 0, max_m = 0, max_len = 0;
    void learn(uint64_t bases, uint64_t M_, uint64_t max_m_, int64_t max_len, uint64_t E_, uint64_t B_) {
        if (bases == 0 || M_ == 0) return;
        if (!seen || M_ >= 4096) {
            const double b = (double)bases;
            m_per_base = fmax(m_per_base, (double)M_ / b);
            e_per_base = fmax(e_per_base, (double)E_ / b);
            b_per_base = fmax(b_per_base, (double)B_ / b);
        }
        if (max_len > 0) mm_per_base = fmax(mm_per_base, (double)max_m_ / (double)max_len);
        seen = true;
    }
};

// A batch's stream work captured for one read class, and what it was
// captured with.
struct MapGraph {
    hipGraphExec_t exec = nullptr;
    uint64_t key = 0;
    uint64_t seen_key = 0;  // the key of the class's latest eager batches,
    int32_t seen = 0;       // and how many ran in a row
};

// A class is captured once GRAPH_AFTER batches in a row had the same key: a
// capture and instantiation cost several milliseconds, about what a hundred
// replays save, so classes that come rarely, or while capacities still grow,
// run eagerly.
constexpr int32_t GRAPH_AFTER = 8;

struct shm_mapper {
    const shm_index *idx = nullptr;
    shm_map_params prm{};
    shm_sketcher sk{};
    hipStream_t st = nullptr;
    uint32_t max_grid = 1024;
    shm_tuning tune{};  // batching choices (0: the built-in rule)
    uint64_t tab_budget = 1ull << 30;  // bucket-table slots per group (24 B each, ~26 GB)
    bool timing = false;
    hipEvent_t ev[ST_COUNT][2];
    hipEvent_t t0 = nullptr, t1 = nullptr;  // a batch's first and last GPU work
    hipEvent_t ev_up = nullptr;              // the end of a batch's upload
    int32_t max_read_len = 0;  // reads longer than this fail alone (0: no limit)
    bool speculate = true;     // size batches from earlier ones and synchronise once
    bool graphs_on = true;     // capture speculative batches per read class
    std::map<int32_t, MapGraph> graphs;
    BatchShape shape;
    double stage_ms[ST_COUNT] = {};
    double total_ms = 0;
    int64_t counts[SHM_N_COUNTS] = {};
    // batch buffers
    DevBuf<uint8_t> bases;
    DevBuf<int64_t> roff;
    PinnedBuf<uint8_t> h_bases;  // upload staging of shm_map_batch
    PinnedBuf<int64_t> h_roff;
    DevBuf<uint32_t> packed;  // shm_map_batch_packed: the 2-bit bases,
    DevBuf<int64_t> exc;      // and their runs of other bytes
    SketchOut so;
    DevBuf<uint64_t> k_h, ps_h, sd_h, sd_key;
    DevBuf<int32_t> k_r;
    DevBuf<uint8_t> k_s, sd_strand;
    DevBuf<uint32_t> ps_v, sd_first, sd_occs, sd_cnt, sd_hoff, sd_ord;
    DevBuf<uint4> sd_row;  // seeds in rarest-first order: offset, count, occs | strand, first position
    DevBuf<int32_t> rd_i32;  // 12 per-read int arrays
    DevBuf<uint64_t> rd_u64; // 2 per-read u64 arrays
    DevBuf<uint64_t> grp_base; // 3 x (RG+1): hit, table and bucket offsets of a group's reads
    DevBuf<uint64_t> grp_nb, part, stat;
    PinnedBuf<uint64_t> h_sum;  // totals read back between stages
    DevBuf<uint32_t> tb_key, tb_list;
    DevBuf<uint64_t> tb_mc;
    DevBuf<int32_t> tb_min, tb_max;
    DevBuf<uint64_t> bk_key;
    DevBuf<uint32_t> bk_val, bs_g;
    DevBuf<int32_t> bk_i32;  // bs_m,c,min,max + sc_* (7)
    DevBuf<uint8_t> bs_alive;
    DevBuf<double> bs_sh, sc_j;
    DevBuf<int32_t> g_sid, g_tr, g_diff, g_tails;
    DevBuf<uint2> tasks, sel;
    DevBuf<uint32_t> big;  // [0] count, then the reads seeded by a workgroup
    DevBuf<uint32_t> ext;  // [0] count, then the group reads extracted by a workgroup
    DevBuf<uint32_t> ov;       // the group reads bucketed through their global tables
    DevBuf<uint32_t> gl;       // the group reads counted through global tables
    DevBuf<uint64_t> ov_kpre;  // their chunk prefix, then the chunk length
    DevBuf<uint32_t> mapped;   // [0] count, then the mapped reads
    DevBuf<uint64_t> u_total;  // sum of distinct k-mers
    PinnedBuf<uint32_t> h_mapped;
    DevBuf<uint32_t> task_count;
    DevBuf<shm_read_result> d_out;
    DevBuf<uint32_t> abort_flag;  // set by the capacity checks of a speculative batch
    PinnedBuf<uint64_t> h_spec;   // its sizes, flag and task counters, gathered there by the device

    void drop_graphs() {
        for (auto &g : graphs)
            if (g.second.exec) (void)hipGraphExecDestroy(g.second.exec);
        graphs.clear();
    }

    // Frees every batch buffer (after a failed allocation, so that a retry
    // with fewer reads starts from the device memory a smaller batch needs).
    void release_batch_buffers() {
        bases.release(), roff.release(), h_bases.release(), packed.release(), exc.release();
        so.nunit.release(), so.unit_start.release(), so.chunk_cnt.release(), so.chunk_koff.release();
        so.koff.release(), so.lane_cnt.release(), so.lane_pre.release(), so.unit_seq.release();
        so.ovf.release(), so.stage.release();
        k_h.release(), ps_h.release(), sd_h.release(), sd_key.release(), k_r.release(), k_s.release();
        sd_strand.release(), ps_v.release(), sd_first.release(), sd_occs.release(), sd_cnt.release();
        sd_hoff.release(), sd_ord.release(), sd_row.release(), rd_i32.release(), rd_u64.release();
        grp_base.release(), grp_nb.release(), tb_key.release(), tb_list.release(), tb_mc.release();
        tb_min.release(), tb_max.release(), bk_key.release(), bk_val.release(), bs_g.release();
        bk_i32.release(), bs_alive.release(), bs_sh.release(), sc_j.release(), g_sid.release();
        g_tr.release(), g_diff.release(), g_tails.release(), tasks.release(), sel.release(), big.release();
        ext.release(), ov.release(), gl.release(), ov_kpre.release(), mapped.release(), h_mapped.release();
        task_count.release(), d_out.release(), abort_flag.release();
        drop_graphs();
        shape.U = shape.M = shape.T = shape.B = shape.max_m = shape.max_len = 0;  // the capacities went with them
    }
};

extern "C" int shm_mapper_create(const shm_index *idx, const shm_map_params *p, shm_mapper **out) {
    if (!idx || !p || !out) return fail(SHM_ERR_ARG, "invalid argument");
    if (p->k != idx->k || p->h_frac != idx->h_frac)
        return fail(SHM_ERR_ARG, "mapper k/h_frac must match the index (k=%d r=%g)", idx->k, idx->h_frac);
    if (p->theta < 0.0 || p->theta > 1.0)
        return fail(SHM_ERR_ARG, "The threshold (-t) should be between 0 and 1. You provided %g.", p->theta);
    if (p->min_diff < 0.0)
        return fail(SHM_ERR_ARG, "The minimum difference (-d) should be >= 0. You provided %g.", p->min_diff);
    if (p->max_overlap < 0.0 || p->max_overlap > 1.0)
        return fail(SHM_ERR_ARG, "The maximum overlap (-o) should be between 0 and 1. You provided %g.",
                    p->max_overlap);
    if (p->metric < 0 || p->metric > 3) return fail(SHM_ERR_ARG, "invalid metric %d", p->metric);
    HIPCHK(hipSetDevice(idx->device));
    shm_mapper *mp = new shm_mapper;
    mp->idx = idx;
    mp->prm = *p;
    mp->sk.k = p->k;
    mp->sk.h_frac = p->h_frac;
    mp->sk.thres = hash_threshold(p->h_frac);
    mp->max_grid = device_grid(idx->device);

    if (const char *e = getenv("SHM_TAB_BUDGET")) mp->tab_budget = strtoull(e, nullptr, 10);
    if (hipStreamCreate(&mp->st) != hipSuccess) {
        delete mp;
        return fail(SHM_ERR_HIP, "hipStreamCreate failed");
    }
    for (int s = 0; s < ST_COUNT; s++)
        for (int j = 0; j < 2; j++) (void)hipEventCreate(&mp->ev[s][j]);
    if (hipEventCreate(&mp->t0) != hipSuccess || hipEventCreate(&mp->t1) != hipSuccess ||
        hipEventCreateWithFlags(&mp->ev_up, hipEventDisableTiming) != hipSuccess) {
        shm_mapper_free(mp);
        return fail(SHM_ERR_HIP, "hipEventCreate failed");
    }
    *out = mp;
    return SHM_OK;
}

extern "C" void shm_mapper_free(shm_mapper *mp) {
    if (!mp) return;
    mp->drop_graphs();
    for (int s = 0; s < ST_COUNT; s++)
        for (int j = 0; j < 2; j++) (void)hipEventDestroy(mp->ev[s][j]);
    if (mp->t0) (void)hipEventDestroy(mp->t0);
    if (mp->t1) (void)hipEventDestroy(mp->t1);
    if (mp->ev_up) (void)hipEventDestroy(mp->ev_up);
    (void)hipStreamDestroy(mp->st);
    delete mp;
}

extern "C" int shm_device_compute_units(int32_t device, int32_t *cus) {
    if (!cus) return fail(SHM_ERR_ARG, "invalid argument");
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device));
    *cus = prop.multiProcessorCount;
    return SHM_OK;
}

extern "C" int shm_mapper_set_timing(shm_mapper *mp, int32_t on) {
    if (!mp) return fail(SHM_ERR_ARG, "invalid argument");
    mp->timing = on != 0;
    return SHM_OK;
}

extern "C" int shm_mapper_set_speculation(shm_mapper *mp, int32_t on) {
    if (!mp) return fail(SHM_ERR_ARG, "invalid argument");
    mp->speculate = on != 0;
    return SHM_OK;
}

extern "C" int shm_mapper_set_graphs(shm_mapper *mp, int32_t on) {
    if (!mp) return fail(SHM_ERR_ARG, "invalid argument");
    mp->graphs_on = on != 0;
    if (!mp->graphs_on) mp->drop_graphs();
    return SHM_OK;
}

extern "C" int shm_mapper_set_tuning(shm_mapper *mp, const shm_tuning *t) {
    if (!mp || !t) return fail(SHM_ERR_ARG, "invalid argument");
    if (t->contrib_wave_hits < 0 || t->prune_thread_seeds < 0 || t->contrib_chunk_hits < 0 || t->contrib_wg_hits < 0 ||
        (t->contrib_chunk_hits &&
         (t->contrib_chunk_hits < 64 || t->contrib_chunk_hits > 4096 || (t->contrib_chunk_hits & (t->contrib_chunk_hits - 1)))) ||
        t->seed_key_bits < 0 || t->seed_key_bits > SEED_KEY_BITS_MAX)
        return fail(SHM_ERR_ARG, "tuning out of range (contrib_chunk_hits: 0 or a power of two from 64 to 4096; "
                                 "seed_key_bits: 0 to 55; the others >= 0)");
    mp->tune = *t;
    return SHM_OK;
}

extern "C" int shm_mapper_get_tuning(const shm_mapper *mp, shm_tuning *t) {
    if (!mp || !t) return fail(SHM_ERR_ARG, "invalid argument");
    *t = mp->tune;
    return SHM_OK;
}

extern "C" int shm_map_batch(shm_mapper *mp, const shm_read *reads, int32_t n, shm_read_result *out);

// Coordinate search: each field in turn over its candidates, the others held
// at the best so far, the workgroup bound first (it moves the most work, and
// the one-wave bound and chunk length act on what it leaves); a candidate's
// cost is the better of two timed batches (GPU time from the first kernel to
// the last record), and it replaces the best only when 2% faster, so timing
// noise does not move the choice.
extern "C" int shm_mapper_autotune(shm_mapper *mp, const shm_read *reads, int32_t n, shm_tuning *out) {
    if (!mp || n < 0 || (n > 0 && !reads)) return fail(SHM_ERR_ARG, "invalid argument");
    std::vector<shm_read_result> res((size_t)(n > 0 ? n : 1));
    shm_tuning best = mp->tune;
    auto cost = [&](const shm_tuning &t, float *ms) -> int {
        mp->tune = t;
        float b = 1e30f;
        for (int rep = 0; rep < 2; rep++) {
            RCHK(shm_map_batch(mp, reads, n, res.data()));
            b = mp->total_ms < b ? mp->total_ms : b;
        }
        *ms = b;
        return SHM_OK;
    };
    RCHK(shm_map_batch(mp, reads, n, res.data()));  // buffers sized before timing
    float best_ms;
    RCHK(cost(best, &best_ms));
    static const int32_t cand_prune[] = {4, 8, 16, 32, 64, 128};
    static const int32_t cand_wave[] = {512, 1024, 2048, 4096, 8192};
    static const int32_t cand_chunk[] = {0, 512, 1024, 2048, 4096};
    static const int32_t cand_wg[] = {1, 8192, 32768, 1 << 20};  // 1: no read takes a workgroup
    for (int field = 0; field < 4; field++) {
        const int32_t *cand = field == 0 ? cand_wg : field == 1 ? cand_wave : field == 2 ? cand_chunk : cand_prune;
        const int nc = field == 0 ? 4 : field == 3 ? 6 : 5;
        for (int i = 0; i < nc; i++) {
            shm_tuning t = best;
            int32_t &f = field == 0   ? t.contrib_wg_hits
                         : field == 1 ? t.contrib_wave_hits
                         : field == 2 ? t.contrib_chunk_hits
                                      : t.prune_thread_seeds;
            if (f == cand[i]) continue;
            f = cand[i];
            float ms;
            const int rc = cost(t, &ms);
            if (rc != SHM_OK) {
                mp->tune = best;
                return rc;
            }
            if (ms < 0.98f * best_ms) {
                best_ms = ms;
                best = t;
            }
        }
    }
    mp->tune = best;
    if (out) *out = best;
    return SHM_OK;
}

extern "C" const char *shm_mapper_stage_name(int32_t s) {
    return (s >= 0 && s < ST_COUNT) ? kStageNames[s] : "";
}

extern "C" int shm_mapper_last_counts(const shm_mapper *mp, int64_t *counts, int32_t n, int32_t *n_written) {
    if (!mp || (n > 0 && !counts)) return fail(SHM_ERR_ARG, "invalid argument");
    int32_t w = 0;
    for (; w < n && w < SHM_N_COUNTS; w++) counts[w] = mp->counts[w];
    if (n_written) *n_written = w;
    return SHM_OK;
}

extern "C" int shm_mapper_last_timing(const shm_mapper *mp, double *total_ms, double *stage_ms, int32_t n,
                                      int32_t *n_written) {
    if (!mp) return fail(SHM_ERR_ARG, "invalid argument");
    if (total_ms) *total_ms = mp->total_ms;
    int32_t w = 0;
    for (; w < n && w < ST_COUNT; w++) stage_ms[w] = mp->stage_ms[w];
    if (n_written) *n_written = w;
    return SHM_OK;
}

// src/mapping.rs:341-361, 370-408 -- second best, MAPQ, printed J2 and the
// match_inefficiency of the reported mapping.
static void finalize_result(const shm_map_params &p, shm_read_result &r) {
    if (!r.mapped) return;
    const double theta2 = p.theta - p.min_diff;
    double j2 = -1.0, sig = -1.0;
    int32_t i2 = -1;
    if (r.has_second) {
        i2 = r.intersection2;
        j2 = r.J2;
        const double x = i2 == -1 ? 0.0 : (double)i2;
        const double y = r.intersection == -1 ? 0.0 : (double)r.intersection;
        sig = fabs(x - y) / sqrt(x + y);
    } else {
        r.b2_segm = r.b2_b = -1;
    }
    if (j2 < theta2) j2 = theta2;
    const double frac = 1.0 - j2 / r.J;
    const int32_t strand5 = abs(r.same_strand) < r.intersection / 2 ? 5 : 60;
    r.mapq = frac > p.min_diff ? (60 < strand5 ? 60 : strand5) : 0;
    r.intersection2 = i2;
    r.J2 = j2;
    r.sigmas = sig;
    r.match_inefficiency = (double)r.total_matches / (double)r.intersection;
}

#define TIME_BEGIN(s) \
    if (mp->timing) HIPCHK(hipEventRecord(mp->ev[s][0], st))
#define TIME_END(s)                                   \
    if (mp->timing) {                                 \
        HIPCHK(hipEventRecord(mp->ev[s][1], st));     \
        pending.push_back(s);                         \
    }

static int flush_timing(shm_mapper *mp, std::vector<int> &pending) {
    if (!mp->timing) return SHM_OK;
    HIPCHK(hipStreamSynchronize(mp->st));
    for (int s : pending) {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, mp->ev[s][0], mp->ev[s][1]));
        mp->stage_ms[s] += ms;
    }
    pending.clear();
    return SHM_OK;
}

// Results go straight into the caller's buffer when it is page-locked memory
// the device can address (k_final's stores cross the link while it runs),
// else into d_out and one copy at the end.
static int result_target(shm_mapper *mp, shm_read_result *out, int32_t R, shm_read_result **res_dev, bool *direct) {
    hipPointerAttribute_t pa;
    if (hipPointerGetAttributes(&pa, out) == hipSuccess && pa.type == hipMemoryTypeHost && pa.devicePointer) {
        *res_dev = (shm_read_result *)pa.devicePointer;
        *direct = true;
        return SHM_OK;
    }
    (void)hipGetLastError();  // pageable memory: the query's error is not ours
    RCHK(mp->d_out.ensure(R));
    *res_dev = mp->d_out.p;
    *direct = false;
    return SHM_OK;
}

// Per-read and per-k-mer buffers of a batch of R reads and up to M k-mers.
static int ensure_read_buffers(shm_mapper *mp, int32_t R, uint64_t M) {
    RCHK(mp->ps_h.ensure(M + 1));
    RCHK(mp->ps_v.ensure(M + 1));
    RCHK(mp->sd_h.ensure(M + 1));
    RCHK(mp->sd_key.ensure(M + 1));
    RCHK(mp->sd_strand.ensure(M + 1));
    RCHK(mp->sd_first.ensure(M + 1));
    RCHK(mp->sd_occs.ensure(M + 1));
    RCHK(mp->sd_cnt.ensure(M + 1));
    RCHK(mp->sd_hoff.ensure(M + 1));
    RCHK(mp->sd_ord.ensure(M + 1));
    RCHK(mp->sd_row.ensure(M + 1));
    RCHK(mp->rd_i32.ensure((size_t)R * 12 + 1));
    RCHK(mp->rd_u64.ensure((size_t)R * 2 + 1));
    RCHK(mp->big.ensure((size_t)R + 2));
    RCHK(mp->grp_base.ensure(3 * (size_t)(R + 1)));
    RCHK(mp->grp_nb.ensure((size_t)R + 1));
    RCHK(mp->part.ensure(SCAN_PARTS + 1));
    RCHK(mp->stat.ensure(1));
    RCHK(mp->mapped.ensure((size_t)R + 2));
    RCHK(mp->u_total.ensure(1));
    RCHK(mp->h_mapped.ensure((size_t)R + 1));
    RCHK(mp->h_sum.ensure(6));
    RCHK(mp->ov.ensure((size_t)R + 1));
    RCHK(mp->gl.ensure((size_t)R + 1));
    RCHK(mp->ov_kpre.ensure((size_t)R + 2));
    RCHK(mp->ext.ensure((size_t)R + 2));
    return SHM_OK;
}

// The shift that leaves a hash's leading key bits (from the keep threshold's
// highest bit down) for the one-word keys of k_seed_wave's sort: every kept
// hash is at most the threshold, so hash >> shift < 2^key_bits.
static uint32_t seed_key_shift(const shm_mapper *mp) {
    const int kb = mp->tune.seed_key_bits > 0 ? mp->tune.seed_key_bits : SEED_KEY_BITS_MAX;
    const int b = mp->sk.thres ? 64 - __builtin_clzll(mp->sk.thres) : 0;
    return b > kb ? (uint32_t)(b - kb) : 0u;
}

// The read-level arguments of the mapping kernels.
static MapArgs read_args(shm_mapper *mp, int32_t R, const int64_t *d_off) {
    const shm_index *ix = mp->idx;
    MapArgs A{};
    A.R = R;
    A.k = mp->prm.k;
    A.theta = mp->prm.theta;
    A.min_diff = mp->prm.min_diff;
    A.max_overlap = mp->prm.max_overlap;
    A.thr_min = mp->prm.theta * (1.0 - mp->prm.min_diff);
    A.metric = mp->prm.metric;
    A.no_bucket_pruning = mp->prm.no_bucket_pruning;
    A.one_sweep = mp->prm.one_sweep;
    A.seed_sh = seed_key_shift(mp);
    A.abs_pos = mp->prm.abs_pos;
    A.roff = d_off;
    A.kstart = ix->kstart.p;
    A.seg_cum = ix->segcum.p;
    A.nseg = ix->nseg;
    A.koff = mp->so.koff.p;
    A.k_h = mp->k_h.p;
    A.k_r = mp->k_r.p;
    A.k_s = mp->k_s.p;
    A.ps_h = mp->ps_h.p;
    A.ps_v = mp->ps_v.p;
    A.sd_h = mp->sd_h.p;
    A.sd_first = mp->sd_first.p;
    A.sd_occs = mp->sd_occs.p;
    A.sd_cnt = mp->sd_cnt.p;
    A.sd_hoff = mp->sd_hoff.p;
    A.sd_ord = mp->sd_ord.p;
    A.sd_strand = mp->sd_strand.p;
    A.sd_key = mp->sd_key.p;
    A.sd_row = mp->sd_row.p;
    int32_t *ri = mp->rd_i32.p;
    A.rd_u = ri + 0 * (size_t)R;
    A.rd_possible = ri + 1 * (size_t)R;
    A.rd_S = ri + 2 * (size_t)R;
    A.rd_hl = ri + 3 * (size_t)R;
    A.rd_ok = ri + 4 * (size_t)R;
    A.rd_c = ri + 5 * (size_t)R;
    A.rd_bseeds = ri + 6 * (size_t)R;
    A.rd_max_seed = ri + 7 * (size_t)R;
    A.rd_seed_matches = ri + 8 * (size_t)R;
    A.rd_z = ri + 9 * (size_t)R;
    A.rd_flag = ri + 10 * (size_t)R;
    A.rd_first = ri + 11 * (size_t)R;
    A.rd_E = mp->rd_u64.p;
    A.rd_cap = mp->rd_u64.p + R;
    A.r0 = 0;
    A.RG = R;
    A.big_count = mp->big.p;
    A.big_list = mp->big.p + 1;
    A.mapped_count = mp->mapped.p;
    A.mapped_list = mp->mapped.p + 1;
    A.u_total = mp->u_total.p;
    return A;
}

// Bucket and refinement buffers of a group of RG reads with up to B buckets
// and T table slots, whose reads have halflen at most max_hl and at most max_m
// k-mers.
static int ensure_group_buffers(shm_mapper *mp, uint64_t T, uint64_t B, int32_t RG, int32_t max_hl, uint64_t max_m) {
    RCHK(mp->tb_key.ensure(T + 1));
    RCHK(mp->tb_list.ensure(T + 1));
    RCHK(mp->tb_mc.ensure(T + 1));
    RCHK(mp->tb_min.ensure(T + 1));
    RCHK(mp->tb_max.ensure(T + 1));
    RCHK(mp->bk_key.ensure(B + 1));
    RCHK(mp->bk_val.ensure(B + 1));
    RCHK(mp->bs_g.ensure(B + 1));
    RCHK(mp->bk_i32.ensure(11 * (B + 1)));
    RCHK(mp->bs_alive.ensure(B + 1));
    RCHK(mp->bs_sh.ensure(B + 1));
    RCHK(mp->sc_j.ensure(B + 1));
    RCHK(mp->tasks.ensure(B + 1));
    RCHK(mp->sel.ensure(B + (uint64_t)RG + 1));
    // per-workgroup scratch of refinement, for the arrays of a span or read
    // that do not fit the workgroup's LDS (a span has at most 2 * hl
    // positions; the grid is max_grid workgroups)
    const uint64_t stride = 2ull * (uint64_t)max_hl + 2;
    const uint64_t dstride = max_m + 1;
    RCHK(mp->g_sid.ensure(stride * mp->max_grid + 1));
    RCHK(mp->g_tr.ensure(stride * mp->max_grid + 1));
    RCHK(mp->g_diff.ensure(dstride * mp->max_grid + 1));
    RCHK(mp->g_tails.ensure(mp->prm.metric == SHM_BUCKET_LCS ? dstride * mp->max_grid + 1 : 1));
    return SHM_OK;
}

// The table arguments of group [r0, r0 + RG), its offsets at grp_base.
static void table_args(shm_mapper *mp, MapArgs &A, int32_t r0, int32_t RG, uint32_t *task_count, uint64_t E_grp) {
    A.r0 = r0;
    A.RG = RG;
    A.hit_base = mp->grp_base.p;
    A.tab_base = mp->grp_base.p + (RG + 1);
    A.bk_base = mp->grp_base.p + 2 * (size_t)(RG + 1);
    A.grp_nb = mp->grp_nb.p;
    A.tb_key = mp->tb_key.p;
    A.tb_list = mp->tb_list.p;
    A.tb_mc = mp->tb_mc.p;
    A.tb_min = mp->tb_min.p;
    A.tb_max = mp->tb_max.p;
    // the group's task counters ([6] counts the reads k_contrib_wave hands
    // to the global tables)
    A.task_count = task_count;
    A.ov_count = task_count + 7;
    A.ov_list = mp->ov.p;
    A.gl_count = task_count + 6;
    A.gl_list = mp->gl.p;
    A.ov_kpre = mp->ov_kpre.p;
    // a big batch fills the device with whole reads; a small one splits
    // long reads so no read's serial walk sets the batch's time
    A.contrib_wave_e = mp->tune.contrib_wave_hits > 0 ? (uint32_t)mp->tune.contrib_wave_hits
                                                      : (E_grp >= (1ull << 24) ? 4096u : 1024u);
    A.contrib_chunk = (uint32_t)mp->tune.contrib_chunk_hits;
    A.contrib_wg_e = mp->tune.contrib_wg_hits > 0 ? (uint32_t)mp->tune.contrib_wg_hits : 8192u;
    A.prune_seq = mp->tune.prune_thread_seeds > 0 ? (uint32_t)mp->tune.prune_thread_seeds : 16u;
}


// The bucket arguments of a group whose buffers hold B buckets.
static void bucket_args(shm_mapper *mp, MapArgs &A, uint64_t B, int32_t max_hl, uint64_t max_m) {
    A.ext_count = mp->ext.p;
    A.ext_list = mp->ext.p + 1;
    A.bk_key = mp->bk_key.p;
    A.bk_val = mp->bk_val.p;
    A.bs_g = mp->bs_g.p;
    int32_t *bi = mp->bk_i32.p;
    const size_t Bs = B + 1;
    A.bs_m = bi + 0 * Bs;
    A.bs_c = bi + 1 * Bs;
    A.bs_min = bi + 2 * Bs;
    A.bs_max = bi + 3 * Bs;
    A.sc_i = bi + 4 * Bs;
    A.sc_same = bi + 5 * Bs;
    A.sc_tl = bi + 6 * Bs;
    A.sc_tr = bi + 7 * Bs;
    A.sc_ssz = bi + 8 * Bs;
    A.sc_pend = bi + 9 * Bs;
    A.sc_seg = bi + 10 * Bs;
    A.bs_alive = mp->bs_alive.p;
    A.bs_sh = mp->bs_sh.p;
    A.sc_j = mp->sc_j.p;
    A.g_sid = mp->g_sid.p;
    A.g_tr = mp->g_tr.p;
    A.sid_stride = 2ull * (uint64_t)max_hl + 2;
    A.g_diff = mp->g_diff.p;
    A.g_tails = mp->g_tails.p;
    A.diff_stride = max_m + 1;
    A.tasks = mp->tasks.p;
    A.sel = mp->sel.p;
    A.task_cap = (uint32_t)(B + 1);
    A.refine_lds = refine_lds_bytes((uint64_t)max_hl, max_m, mp->prm.metric == METRIC_BUCKET_LCS);
}

// What every batch ends with once its records and counters are on the host:
// the host's double-precision finalisation of the mapped reads and the work
// counts.
static void finish_batch(shm_mapper *mp, shm_read_result *out, const std::vector<uint32_t> &ntask, int32_t ngroups,
                         uint64_t M, uint64_t seeds, int64_t nbuckets, uint64_t E_all) {
    const uint32_t n_mapped = mp->h_mapped.p[0];
    for (uint32_t i = 0; i < n_mapped; i++) finalize_result(mp->prm, out[mp->h_mapped.p[1 + i]]);
    int64_t cand = 0, deferred = 0, refined = 0, global_reads = 0;
    for (int32_t g = 0; g < ngroups; g++) {
        const uint32_t *tc = &ntask[TASK_COUNTERS * (size_t)g];
        cand += tc[0];
        deferred += tc[1];
        refined += tc[2] + tc[3] + tc[4];
        global_reads += tc[6];
    }
    mp->counts[0] = (int64_t)M;
    mp->counts[1] = (int64_t)seeds;
    mp->counts[2] = nbuckets;
    mp->counts[3] = refined;
    mp->counts[4] = ngroups;
    mp->counts[5] = cand;
    mp->counts[6] = deferred;
    mp->counts[7] = (int64_t)E_all;
    mp->counts[8] = global_reads;
}

// Bases and longest read of a batch, from host offsets.
static void batch_extent(const int64_t *h_off, int32_t R, uint64_t *bases, int64_t *max_len) {
    int64_t ml = 0;
    for (int32_t i = 0; i < R; i++) ml = std::max(ml, h_off[i + 1] - h_off[i]);
    *bases = (uint64_t)(h_off[R] - h_off[0]);
    *max_len = ml;
}

// Slots of a speculative batch's gathered sizes and counters.
enum { SPEC_M, SPEC_MAX_M, SPEC_E, SPEC_B, SPEC_SEEDS, SPEC_ABORT, SPEC_TASKS, SPEC_SLOTS = SPEC_TASKS + TASK_COUNTERS };

// The capacities of a speculative batch of R reads, `bases` bases and longest
// read max_len: its estimates from the shapes seen so far, with a margin, and
// never below what the buffers were last sized to (buffers only grow, to
// twice their size at least, so that a stream of slowly growing batches
// reallocates rarely: a reallocation waits for the whole device).
// A sequence of nw windows has ceil(nw / LANE_W) + halo sketch units, so
// U <= bases / LANE_W + R (1 + halo).  A read's table has the smallest power
// of two of at least twice its hits slots, 64 at least, so T <= 64 R + 4 E;
// its hits give at most two buckets each.  False when the batch's tables
// could exceed the table budget.
static bool speculative_caps(shm_mapper *mp, int32_t R, uint64_t bases, int64_t max_len, double *E_est) {
    BatchShape &sh = mp->shape;
    // margins: k-mers per base vary little between batches, hits per k-mer
    // and buckets per hit with the reads' repeat content
    auto grow = [](uint64_t have, double want, double margin) {
        const uint64_t w = (uint64_t)(want * margin) + 1024;
        if (w <= have) return have;
        return w > 2 * have ? w : 2 * have;
    };
    const double M_est = (double)bases * sh.m_per_base;
    *E_est = (double)bases * sh.e_per_base;
    const double T_est = 64.0 * R + 4.0 * *E_est;
    if (T_est * 1.3 + 1024 > (double)mp->tab_budget) return false;
    sh.U = grow(sh.U, (double)bases / LANE_W + (double)R * (1 + sketch_halo(mp->sk.k)), 1.0);
    sh.M = grow(sh.M, M_est, 1.15);
    sh.max_m = grow(sh.max_m, (double)max_len * 2.0 * fmax(sh.mm_per_base, sh.m_per_base), 1.15);
    if (sh.max_m > sh.M) sh.max_m = sh.M;
    sh.max_len = grow(sh.max_len, (double)max_len, 1.15);
    sh.T = std::min<uint64_t>(grow(sh.T, T_est, 1.3), mp->tab_budget);
    sh.B = std::min<uint64_t>(grow(sh.B, fmin(2.0 * *E_est, (double)bases * sh.b_per_base), 1.5),
                              (uint64_t)UINT32_MAX - 1);
    return true;
}

// The largest halflen a speculative batch's buffers hold: a read's halflen is
// its k-mer count, or its length with absolute positions.
static int32_t speculative_hl(const shm_mapper *mp) {
    const uint64_t hl = mp->prm.abs_pos ? mp->shape.max_len : mp->shape.max_m;
    return hl < (uint64_t)INT32_MAX ? (int32_t)hl : INT32_MAX;
}

// Every buffer a speculative batch of R reads uses, at the capacities of
// speculative_caps.
static int reserve_speculative(shm_mapper *mp, int32_t R) {
    const BatchShape &sh = mp->shape;
    SketchOut &so = mp->so;
    const uint64_t C = (sh.U + WAVE - 1) / WAVE;
    RCHK(so.unit_start.ensure((size_t)R + 1));
    RCHK(so.nunit.ensure((size_t)R + 1));
    RCHK(so.koff.ensure((size_t)R + 1));
    RCHK(so.part.ensure(SCAN_PARTS + 1));
    RCHK(so.h_sum.ensure(3));
    RCHK(so.ovf_n.ensure(1));
    RCHK(so.stat.ensure(1));
    RCHK(so.chunk_cnt.ensure(C + 1));
    RCHK(so.chunk_koff.ensure(C + 1));
    RCHK(so.lane_cnt.ensure(C * WAVE + 1));
    RCHK(so.lane_pre.ensure(C * WAVE + 1));
    RCHK(so.unit_seq.ensure(sh.U + 1));
    RCHK(so.stage.ensure(C * WAVE * (uint64_t)sketch_lane_cap(mp->sk.thres) + 1));
    RCHK(so.ovf.ensure(C + 1));
    RCHK(mp->k_h.ensure(sh.M + 1));
    RCHK(mp->k_r.ensure(sh.M + 1));
    RCHK(mp->k_s.ensure(sh.M + 1));
    RCHK(ensure_read_buffers(mp, R, sh.M));
    RCHK(ensure_group_buffers(mp, sh.T, sh.B, R, speculative_hl(mp), sh.max_m));
    RCHK(mp->task_count.ensure(TASK_COUNTERS + 1));
    RCHK(mp->abort_flag.ensure(1));
    RCHK(mp->h_spec.ensure(SPEC_SLOTS));
    RCHK(mp->d_out.ensure((size_t)R));  // the records of a captured batch
    return SHM_OK;
}

// Read counts of captured batches: powers of two and their halfway points
// (1, 2, 3, 4, 6, 8, 12, ...); a batch is padded with empty reads up to its
// class, so that one graph serves every batch of the class.
static int32_t read_class(int32_t R) {
    if (R <= 2) return R;
    int32_t c = 2;
    while (true) {
        if (c >= R) return c;
        if (c + c / 2 >= R) return c + c / 2;
        if (c > INT32_MAX / 2) return R;
        c *= 2;
    }
}

// The stream work of a speculative batch of R reads (arguments A): every
// kernel, then the copies of its sizes, counters and mapped list to the
// mapper's page-locked buffers.  It depends on the capacities and buffers
// alone, so it can be captured.
static int enqueue_speculative(shm_mapper *mp, const uint8_t *d_seq, const int64_t *d_off, int32_t R, MapArgs &A,
                               shm_read_result *res_dev, std::vector<int> &pending) {
    hipStream_t st = mp->st;
    const DevIndex I = mp->idx->view();
    const BatchShape &sh = mp->shape;
    uint32_t *abort = mp->abort_flag.p;
    // every counter of the batch cleared in one launch
    ZeroList z{};
    z.add(abort, 1);
    z.add(mp->so.ovf_n.p, 1);
    z.add(mp->so.stat.p, 2);
    z.add(A.big_count, 1);
    z.add(mp->stat.p, 2);
    z.add(A.mapped_count, 1);
    z.add(A.u_total, 2);
    z.add(A.task_count, TASK_COUNTERS);
    z.add(A.ext_count, 1);
    launch_zero_list(z, st);
    RCHK(sketch_device(&mp->sk, d_seq, d_off, R, mp->so, mp->k_h, mp->k_r, mp->k_s, mp->max_grid, st,
                       mp->timing ? mp->ev[ST_SKETCH] : nullptr, mp->timing ? mp->ev[ST_SKETCH_COMPACT] : nullptr,
                       false, nullptr, abort, sh.U, sh.M, sh.max_m));
    if (mp->timing) {
        pending.push_back(ST_SKETCH);
        pending.push_back(ST_SKETCH_COMPACT);
    }
    for (Stage s : {ST_SEED_WAVE, ST_SEED_WG}) {
        TIME_BEGIN(s);
        launch_stage(s, A, I, nullptr, 0, mp->max_grid, st);
        TIME_END(s);
    }
    launch_max_i32(A.rd_hl, R, mp->stat.p, st);
    launch_scan(A.rd_E, mp->grp_base.p, nullptr, R, mp->part.p, st);
    launch_scan(A.rd_cap, mp->grp_base.p + (R + 1), nullptr, R, mp->part.p, st);
    launch_gate(mp->grp_base.p + (R + 1) + R, sh.T, nullptr, 0, nullptr, 0, nullptr, 0, abort, st);
    TIME_BEGIN(ST_TAB_INIT);
    launch_stage(ST_TAB_INIT, A, I, nullptr, R, mp->max_grid, st);
    TIME_END(ST_TAB_INIT);
    TIME_BEGIN(ST_CONTRIB);
    launch_stage(ST_CONTRIB, A, I, nullptr, 0, mp->max_grid, st);
    TIME_END(ST_CONTRIB);
    launch_scan(mp->grp_nb.p, mp->grp_base.p + 2 * (size_t)(R + 1), nullptr, R, mp->part.p, st);
    launch_gate(mp->grp_base.p + 2 * (size_t)(R + 1) + R, sh.B, nullptr, 0, nullptr, 0, nullptr, 0, abort, st);
    for (Stage s : {ST_EXTRACT, ST_PRUNE, ST_REFINE, ST_FINAL}) {
        TIME_BEGIN(s);
        launch_stage(s, A, I, res_dev, sh.B, mp->max_grid, st);
        TIME_END(s);
    }
    HIPCHK(hipGetLastError());
    // the batch's sizes and counters, gathered straight into page-locked
    // memory (SPEC_* slots), then its mapped list
    uint64_t *hs_dev = nullptr;
    HIPCHK(hipHostGetDevicePointer((void **)&hs_dev, mp->h_spec.p, 0));
    GatherList g{};
    g.add(mp->so.koff.p + R, true);                            // SPEC_M
    g.add(mp->so.stat.p, true);                                // SPEC_MAX_M
    g.add(mp->grp_base.p + R, true);                           // SPEC_E
    g.add(mp->grp_base.p + 2 * (size_t)(R + 1) + R, true);     // SPEC_B
    g.add(A.u_total, true);                                    // SPEC_SEEDS
    g.add(abort, false);                                       // SPEC_ABORT
    for (uint32_t t = 0; t < TASK_COUNTERS; t++) g.add(A.task_count + t, false);  // SPEC_TASKS
    launch_gather(g, hs_dev, st);
    HIPCHK(hipMemcpyAsync(mp->h_mapped.p, mp->mapped.p, ((size_t)R + 1) * 4, hipMemcpyDeviceToHost, st));
    return SHM_OK;
}

// 64-bit FNV-1a over a byte range, chained.
static uint64_t fnv1a(uint64_t h, const void *p, size_t n) {
    const uint8_t *b = (const uint8_t *)p;
    for (size_t i = 0; i < n; i++) h = (h ^ b[i]) * 1099511628211ull;
    return h;
}

// What a captured batch depends on: its kernel arguments (every buffer
// address and capacity among them) and the sketch's buffers and capacities.
static uint64_t graph_key(const shm_mapper *mp, const MapArgs &A, const uint8_t *d_seq, const int64_t *d_off,
                          const shm_read_result *res_dev) {
    const SketchOut &so = mp->so;
    const void *ptrs[] = {d_seq, d_off, res_dev, so.unit_start.p, so.nunit.p, so.koff.p, so.part.p, so.ovf_n.p,
                          so.stat.p, so.chunk_cnt.p, so.chunk_koff.p, so.lane_cnt.p, so.lane_pre.p, so.unit_seq.p,
                          so.stage.p, so.ovf.p, mp->abort_flag.p, mp->h_spec.p, mp->h_mapped.p,
                          mp->stat.p, mp->part.p, mp->grp_base.p, mp->grp_nb.p};
    const BatchShape &sh = mp->shape;
    const uint64_t caps[] = {sh.U, sh.M, sh.max_m, sh.T, sh.B, sh.max_len, (uint64_t)mp->max_grid};
    uint64_t h = 1469598103934665603ull;
    h = fnv1a(h, &A, sizeof A);
    h = fnv1a(h, ptrs, sizeof ptrs);
    return fnv1a(h, caps, sizeof caps);
}

// Maps R reads with one host synchronisation, at the end.  Every buffer is
// sized before the first launch from the shapes of earlier batches, with a
// margin; device checks after the sketch's unit scan (units), after the sketch
// (k-mers, the largest sketch), after seeding (table slots) and after
// counting (buckets) set an abort flag when a size exceeds its buffer, and
// every later kernel then returns at once.  For reads in the mapper's own
// upload buffers the batch is padded with empty reads to its read class and
// its stream work captured into a graph once per class, then replayed.
// *done is false when the batch stopped (or would need table groups): the
// caller maps it again by map_synced.
static int map_speculative(shm_mapper *mp, const uint8_t *d_seq, const int64_t *d_off, int32_t R,
                           shm_read_result *out, const int64_t *h_off, bool *done) {
    *done = false;
    hipStream_t st = mp->st;
    BatchShape &sh = mp->shape;
    uint64_t bases;
    int64_t max_len;
    batch_extent(h_off, R, &bases, &max_len);
    // a graph needs the mapper's upload buffers, with room for the padding
    const int32_t Rc = read_class(R);
    const bool graph = mp->graphs_on && !mp->timing && d_seq == mp->bases.p && d_off == mp->roff.p &&
                       mp->roff.cap >= (size_t)Rc + 1;
    const int32_t RL = graph ? Rc : R;  // reads launched
    double E_est;
    if (!speculative_caps(mp, RL, bases, max_len, &E_est)) return SHM_OK;  // table groups

    shm_read_result *res_dev;
    bool direct = false;
    if (graph) {
        RCHK(mp->d_out.ensure(RL));  // a fixed address for the graph; one copy to `out` after it
        res_dev = mp->d_out.p;
    } else {
        RCHK(result_target(mp, out, R, &res_dev, &direct));
    }
    RCHK(reserve_speculative(mp, RL));
    const int32_t max_hl = speculative_hl(mp);

    MapArgs A = read_args(mp, RL, d_off);
    A.abort = mp->abort_flag.p;
    // the hit count only picks the contribution split, which records never
    // depend on: the batch's estimate stands in for it
    table_args(mp, A, 0, RL, mp->task_count.p, (uint64_t)E_est);
    bucket_args(mp, A, sh.B, max_hl, sh.max_m);

    std::vector<int> pending;
    HIPCHK(hipEventRecord(mp->t0, st));
    if (RL > R) launch_fill_u64((uint64_t *)mp->roff.p + R + 1, (uint64_t)(RL - R), (uint64_t)h_off[R], st);
    bool replay = false;
    MapGraph *gp = nullptr;
    if (graph) {
        MapGraph &g = mp->graphs[RL];
        gp = &g;
        const uint64_t key = graph_key(mp, A, d_seq, d_off, res_dev);
        if (g.exec && g.key == key) {
            replay = true;
        } else {
            g.seen = g.seen_key == key ? g.seen + 1 : 1;
            g.seen_key = key;
            replay = g.seen >= GRAPH_AFTER;
        }
        if (replay && (!g.exec || g.key != key)) {
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
            g.exec = nullptr;
            const int64_t a0 = g_allocs;
            HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
            const int rc = enqueue_speculative(mp, d_seq, d_off, RL, A, res_dev, pending);
            hipGraph_t gr = nullptr;
            const hipError_t e = hipStreamEndCapture(st, &gr);
            if (rc != SHM_OK || e != hipSuccess) {
                if (gr) (void)hipGraphDestroy(gr);
                if (rc != SHM_OK) return rc;
                HIPCHK(e);
            }
            const hipError_t ei = hipGraphInstantiate(&g.exec, gr, nullptr, nullptr, 0);
            (void)hipGraphDestroy(gr);
            if (ei != hipSuccess) g.exec = nullptr;
            HIPCHK(ei);
            // an allocation during the capture changed an address the key read
            g.key = g_allocs == a0 ? key : 0;
            mp->counts[13] = 1;
        }
    }
    if (replay) {
        HIPCHK(hipGraphLaunch(gp->exec, st));
        mp->counts[12] = 1;
    } else {
        RCHK(enqueue_speculative(mp, d_seq, d_off, RL, A, res_dev, pending));
    }
    if (!direct)  // (a batch that may be captured writes its records to d_out)
        HIPCHK(hipMemcpyAsync(out, res_dev, (size_t)R * sizeof(shm_read_result), hipMemcpyDeviceToHost, st));
    HIPCHK(hipEventRecord(mp->t1, st));
    HIPCHK(hipStreamSynchronize(st));
    RCHK(flush_timing(mp, pending));
    const uint64_t *hs = mp->h_spec.p;
    if (hs[SPEC_ABORT] != 0) return SHM_OK;  // a size outgrew its buffer: mapped again, synchronised
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, mp->t0, mp->t1));
    mp->total_ms = ms;
    const uint64_t M = hs[SPEC_M], E_all = hs[SPEC_E], B = hs[SPEC_B];
    std::vector<uint32_t> ntask(TASK_COUNTERS);
    for (uint32_t t = 0; t < TASK_COUNTERS; t++) ntask[t] = (uint32_t)hs[SPEC_TASKS + t];
    finish_batch(mp, out, ntask, 1, M, hs[SPEC_SEEDS], (int64_t)B, E_all);
    mp->counts[9] = 1;
    sh.learn(bases, M, hs[SPEC_MAX_M], max_len, E_all, B);
    *done = true;
    return SHM_OK;
}

// Maps R reads whose bases and offsets are on the device, reading each size
// back before sizing what depends on it: the sketch's k-mer count (and, with
// no host offsets, its unit count), the table and hit totals after seeding,
// and each table group's bucket count.  Tables of more slots than the table
// budget are split into groups of reads here.
static int map_synced(shm_mapper *mp, const uint8_t *d_seq, const int64_t *d_off, int32_t R, shm_read_result *out,
                      const int64_t *h_off) {
    hipStream_t st = mp->st;
    const DevIndex I = mp->idx->view();
    std::vector<int> pending;
    HIPCHK(hipEventRecord(mp->t0, st));
    mp->so.syncs = 0;
    RCHK(sketch_device(&mp->sk, d_seq, d_off, R, mp->so, mp->k_h, mp->k_r, mp->k_s, mp->max_grid, st,
                       mp->timing ? mp->ev[ST_SKETCH] : nullptr, mp->timing ? mp->ev[ST_SKETCH_COMPACT] : nullptr,
                       false, h_off));
    int syncs = mp->so.syncs;
    if (mp->timing) {
        pending.push_back(ST_SKETCH);
        pending.push_back(ST_SKETCH_COMPACT);
    }
    const uint64_t M = mp->so.M;
    shm_read_result *res_dev;
    bool direct;
    RCHK(result_target(mp, out, R, &res_dev, &direct));
    RCHK(ensure_read_buffers(mp, R, M));
    MapArgs A = read_args(mp, R, d_off);

    const bool any_large = mp->so.max_m > WAVE_SEED_CAP;  // a read whose sketch exceeds one wave's LDS slice
    launch_fill_u32(A.big_count, 1, 0, st);
    for (Stage s : {ST_SEED_WAVE, ST_SEED_WG}) {
        if (s == ST_SEED_WG && !any_large) continue;
        TIME_BEGIN(s);
        launch_stage(s, A, I, nullptr, 0, mp->max_grid, st);
        TIME_END(s);
    }
    HIPCHK(hipGetLastError());
    // hit and table offsets of every read by device scans; only the totals and
    // the largest halflen come back
    launch_fill_u64(mp->stat.p, 1, 0, st);
    launch_max_i32(A.rd_hl, R, mp->stat.p, st);
    launch_scan(A.rd_E, mp->grp_base.p, nullptr, R, mp->part.p, st);
    launch_scan(A.rd_cap, mp->grp_base.p + (R + 1), nullptr, R, mp->part.p, st);
    HIPCHK(hipMemcpyAsync(mp->h_sum.p, mp->grp_base.p + R, 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(mp->h_sum.p + 1, mp->grp_base.p + (R + 1) + R, 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(mp->h_sum.p + 2, mp->stat.p, 8, hipMemcpyDeviceToHost, st));
    // batch counters of k_final, cleared while the host waits
    launch_fill_u32(mp->mapped.p, 1, 0, st);
    launch_fill_u64(mp->u_total.p, 1, 0, st);
    HIPCHK(hipStreamSynchronize(st));
    syncs++;
    RCHK(flush_timing(mp, pending));
    const uint64_t E_all = mp->h_sum.p[0], T_all = mp->h_sum.p[1];
    const int32_t hl_all = (int32_t)mp->h_sum.p[2];
    // Tables of all reads over the budget: split into groups on the host.
    const bool one_group = T_all <= mp->tab_budget;
    std::vector<uint64_t> E, cap;
    if (!one_group) {
        E.resize(R);
        cap.resize(R);
        HIPCHK(hipMemcpyAsync(E.data(), A.rd_E, R * 8, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(cap.data(), A.rd_cap, R * 8, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        syncs++;
    }

    // bucketing, in groups of reads whose tables fit the budget
    RCHK(mp->task_count.ensure(TASK_COUNTERS * ((size_t)R + 1)));  // refine-task counters per group
    int32_t ngroups = 0;
    int64_t nbuckets = 0;
    int32_t r0 = 0;
    while (r0 < R) {
        int32_t r1 = R;
        uint64_t T = T_all, E_grp = E_all;
        if (!one_group) {
            r1 = r0;
            uint64_t tsum = 0;
            while (r1 < R && (r1 == r0 || tsum + cap[r1] <= mp->tab_budget)) tsum += cap[r1++];
            const int32_t RG = r1 - r0;
            std::vector<uint64_t> hb(RG + 1), tb(RG + 1);
            hb[0] = tb[0] = 0;
            for (int32_t i = 0; i < RG; i++) {
                hb[i + 1] = hb[i] + E[r0 + i];
                tb[i + 1] = tb[i] + cap[r0 + i];
            }
            HIPCHK(hipMemcpyAsync(mp->grp_base.p, hb.data(), (RG + 1) * 8, hipMemcpyHostToDevice, st));
            HIPCHK(hipMemcpyAsync(mp->grp_base.p + (RG + 1), tb.data(), (RG + 1) * 8, hipMemcpyHostToDevice, st));
            T = tb[RG];
            E_grp = hb[RG];
            HIPCHK(hipStreamSynchronize(st));  // hb and tb leave scope
            syncs++;
        }
        const int32_t RG = r1 - r0;
        RCHK(mp->tb_key.ensure(T + 1));
        RCHK(mp->tb_list.ensure(T + 1));
        RCHK(mp->tb_mc.ensure(T + 1));
        RCHK(mp->tb_min.ensure(T + 1));
        RCHK(mp->tb_max.ensure(T + 1));
        uint32_t *task_count = mp->task_count.p + TASK_COUNTERS * (size_t)ngroups;
        launch_fill_u32(task_count, TASK_COUNTERS, 0, st);
        table_args(mp, A, r0, RG, task_count, E_grp);
        TIME_BEGIN(ST_TAB_INIT);
        launch_stage(ST_TAB_INIT, A, I, nullptr, RG, mp->max_grid, st);
        TIME_END(ST_TAB_INIT);
        TIME_BEGIN(ST_CONTRIB);
        launch_stage(ST_CONTRIB, A, I, nullptr, E_grp, mp->max_grid, st);
        TIME_END(ST_CONTRIB);
        launch_scan(mp->grp_nb.p, mp->grp_base.p + 2 * (size_t)(RG + 1), nullptr, RG, mp->part.p, st);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(mp->h_sum.p + 3, mp->grp_base.p + 2 * (size_t)(RG + 1) + RG, 8, hipMemcpyDeviceToHost,
                              st));
        // the group's extract list, cleared while the host waits
        launch_fill_u32(mp->ext.p, 1, 0, st);
        HIPCHK(hipStreamSynchronize(st));
        syncs++;
        RCHK(flush_timing(mp, pending));
        const uint64_t B = mp->h_sum.p[3];
        RCHK(ensure_group_buffers(mp, T, B, RG, hl_all, mp->so.max_m));
        bucket_args(mp, A, B, hl_all, mp->so.max_m);
        for (Stage s : {ST_EXTRACT, ST_PRUNE, ST_REFINE, ST_FINAL}) {
            TIME_BEGIN(s);
            launch_stage(s, A, I, res_dev, B, mp->max_grid, st);
            TIME_END(s);
        }
        HIPCHK(hipGetLastError());
        nbuckets += (int64_t)B;
        ngroups++;
        r0 = r1;
    }
    std::vector<uint32_t> ntask(TASK_COUNTERS * (size_t)ngroups);
    HIPCHK(hipMemcpyAsync(ntask.data(), mp->task_count.p, ntask.size() * 4, hipMemcpyDeviceToHost, st));
    if (!direct)
        HIPCHK(hipMemcpyAsync(out, res_dev, (size_t)R * sizeof(shm_read_result), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(mp->h_mapped.p, mp->mapped.p, ((size_t)R + 1) * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(mp->h_sum.p + 4, A.u_total, 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipEventRecord(mp->t1, st));
    HIPCHK(hipStreamSynchronize(st));
    syncs++;
    RCHK(flush_timing(mp, pending));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, mp->t0, mp->t1));
    mp->total_ms = ms;
    finish_batch(mp, out, ntask, ngroups, M, mp->h_sum.p[4], nbuckets, E_all);
    mp->counts[9] = syncs;
    if (h_off && one_group) {
        uint64_t bases;
        int64_t max_len;
        batch_extent(h_off, R, &bases, &max_len);
        mp->shape.learn(bases, M, mp->so.max_m, max_len, E_all, (uint64_t)nbuckets);
        // the buffers of the next batch, if it is like this one, while the
        // stream is idle
        double E_est;
        const int32_t RL = mp->graphs_on && d_seq == mp->bases.p && d_off == mp->roff.p ? read_class(R) : R;
        if (mp->speculate && mp->shape.seen && speculative_caps(mp, RL, bases, max_len, &E_est))
            RCHK(reserve_speculative(mp, RL));
    }
    return SHM_OK;
}

// Maps R reads whose bases and offsets are on the device.  With the host's
// copy of the offsets (h_off) a batch is speculative (map_speculative), once
// an earlier batch has shown the mapper its shape; otherwise, and when a
// speculative batch stops, it is mapped by map_synced.
static int map_device(shm_mapper *mp, const uint8_t *d_seq, const int64_t *d_off, int32_t R, shm_read_result *out,
                      const int64_t *h_off = nullptr) {
    for (int s = 0; s < ST_COUNT; s++) mp->stage_ms[s] = 0;
    for (int i = 0; i < SHM_N_COUNTS; i++) mp->counts[i] = 0;
    if (R == 0) {
        mp->total_ms = 0;
        return SHM_OK;
    }
    const int64_t allocs0 = g_allocs;
    struct Allocs {  // the batch's allocations, on every return path
        shm_mapper *mp;
        int64_t a0;
        ~Allocs() { mp->counts[11] = g_allocs - a0; }
    } allocs{mp, allocs0};
    int64_t redone = 0;
    if (h_off && mp->speculate && mp->shape.seen) {
        bool done = false;
        RCHK(map_speculative(mp, d_seq, d_off, R, out, h_off, &done));
        if (done) return SHM_OK;
        for (int s = 0; s < ST_COUNT; s++) mp->stage_ms[s] = 0;
        redone = 1;
    }
    RCHK(map_synced(mp, d_seq, d_off, R, out, h_off));
    mp->counts[9] += redone;  // the speculative attempt's synchronisation
    mp->counts[10] = redone;
    return SHM_OK;
}

// Large host-to-device uploads of concurrent mappers on one device go one at a
// time, in turn: copies that share the link at once all finish late together
// and leave it idle while their mappers compute, whereas in turn one mapper's
// kernels run while the next one's batch goes up.  The caller holds the turn
// from its first copy until they are done (small uploads skip it).
constexpr uint64_t UPLOAD_TURN_BYTES = 8ull << 20;
static std::mutex g_upload_turn[64];

struct UploadTurn {
    std::unique_lock<std::mutex> lk;
    shm_mapper *mp = nullptr;
    UploadTurn(shm_mapper *m, uint64_t bytes);
    int done();  // waits for the copies queued so far on the mapper's stream
};

// Uploads n valid host reads and maps them as one batch.
static int map_host_batch(shm_mapper *mp, const shm_read *reads, int32_t n, shm_read_result *out) {
    std::vector<int64_t> off(n + 1, 0);
    for (int32_t i = 0; i < n; i++) off[i + 1] = off[i] + reads[i].len;
    RCHK(mp->bases.ensure(off[n] + 64));
    RCHK(mp->roff.ensure((size_t)read_class(n) + 1));  // room to pad a captured batch
    // Reads back to back in one caller buffer go up in one copy; otherwise they
    // are gathered into one page-locked staging copy first.  (The previous call
    // ended with a stream synchronisation, so no copy still reads the staging
    // buffers.)
    bool contiguous = n > 0;
    for (int32_t i = 1; i < n && contiguous; i++) contiguous = reads[i].seq == reads[i - 1].seq + reads[i - 1].len;
    RCHK(mp->h_roff.ensure(n + 1));
    memcpy(mp->h_roff.p, off.data(), (n + 1) * 8);
    const uint8_t *src = contiguous ? reads[0].seq : nullptr;
    if (!contiguous && off[n] > 0) {
        RCHK(mp->h_bases.ensure(off[n] + 1));
        for (int32_t i = 0; i < n; i++)
            if (reads[i].len) memcpy(mp->h_bases.p + off[i], reads[i].seq, reads[i].len);
        src = mp->h_bases.p;
    }
    UploadTurn turn(mp, (uint64_t)off[n]);
    if (off[n] > 0) HIPCHK(hipMemcpyAsync(mp->bases.p, src, off[n], hipMemcpyHostToDevice, mp->st));
    HIPCHK(hipMemcpyAsync(mp->roff.p, mp->h_roff.p, (n + 1) * 8, hipMemcpyHostToDevice, mp->st));
    RCHK(turn.done());
    return map_device(mp, mp->bases.p, mp->roff.p, n, out, mp->h_roff.p);
}

UploadTurn::UploadTurn(shm_mapper *m, uint64_t bytes) {
    if (bytes < UPLOAD_TURN_BYTES) return;
    const int dev = m->idx->device;
    lk = std::unique_lock<std::mutex>(g_upload_turn[(dev >= 0 && dev < 64) ? dev : 0]);
    mp = m;
}

int UploadTurn::done() {
    if (!mp) return SHM_OK;
    HIPCHK(hipEventRecord(mp->ev_up, mp->st));
    const hipError_t e = hipEventSynchronize(mp->ev_up);
    lk.unlock();
    mp = nullptr;
    HIPCHK(e);
    return SHM_OK;
}

// The record of a read that failed alone: its length and status, nothing mapped.
static void failed_record(shm_read_result &r, int32_t len, int code) {
    memset(&r, 0, sizeof r);
    r.read_len = len > 0 ? len : 0;
    r.b_segm = r.b_b = r.b2_segm = r.b2_b = -1;
    r.status = code;
}

// The first read (in input order) that failed alone, with its message.
struct ReadFailure {
    int32_t read = -1;
    int code = SHM_OK;
    std::string msg;
    void note(int32_t i, int c, const std::string &m) {
        if (read < 0 || i < read) {
            read = i;
            code = c;
            msg = m;
        }
    }
};

// Maps the reads ids[0, n) (ascending) into their records.  A batch that runs
// out of device memory is halved and each half mapped again, down to single
// reads: a read that cannot be mapped even alone fails alone (its status), as
// a read that panics does in the reference (src/shmap/mod.rs:219-238).  Any
// other failure is the whole batch's (every read's status) and is returned.
static int map_isolating(shm_mapper *mp, const shm_read *reads, const int32_t *ids, int32_t n,
                         shm_read_result *out, ReadFailure &rf) {
    if (n == 0) return SHM_OK;
    const bool dense = ids[n - 1] - ids[0] == n - 1;  // a contiguous run: the caller's arrays as they are
    std::vector<shm_read> sub;
    std::vector<shm_read_result> tmp;
    if (!dense) {
        sub.resize(n);
        tmp.resize(n);
        for (int32_t j = 0; j < n; j++) sub[j] = reads[ids[j]];
    }
    const int rc = dense ? map_host_batch(mp, reads + ids[0], n, out + ids[0]) : map_host_batch(mp, sub.data(), n, tmp.data());
    if (rc == SHM_OK) {
        if (!dense)
            for (int32_t j = 0; j < n; j++) out[ids[j]] = tmp[j];
        return SHM_OK;
    }
    const std::string msg = g_err;  // a failed batch: this path may allocate
    if (rc == SHM_ERR_NOMEM) {
        (void)hipStreamSynchronize(mp->st);
        (void)hipGetLastError();
        mp->release_batch_buffers();
        if (n == 1) {
            failed_record(out[ids[0]], reads[ids[0]].len, rc);
            rf.note(ids[0], rc, msg);
            return SHM_OK;
        }
        const int32_t h = n / 2;
        const int r1 = map_isolating(mp, reads, ids, h, out, rf);
        if (r1 != SHM_OK) {
            for (int32_t j = h; j < n; j++) failed_record(out[ids[j]], reads[ids[j]].len, r1);
            return r1;
        }
        return map_isolating(mp, reads, ids + h, n - h, out, rf);
    }
    for (int32_t j = 0; j < n; j++) failed_record(out[ids[j]], reads[ids[j]].len, rc);
    snprintf(g_err, sizeof g_err, "%s", msg.c_str());
    return rc;
}

extern "C" int shm_map_batch(shm_mapper *mp, const shm_read *reads, int32_t n, shm_read_result *out) {
    if (!mp || n < 0 || (n > 0 && (!reads || !out))) return fail(SHM_ERR_ARG, "invalid argument");
    HIPCHK(hipSetDevice(mp->idx->device));
    // reads that fail alone before the device sees them: malformed, or over
    // the mapper's length limit
    ReadFailure rf;
    std::vector<int32_t> ids;
    ids.reserve(n);
    for (int32_t i = 0; i < n; i++) {
        const int32_t len = reads[i].len;
        if (len < 0 || (len > 0 && !reads[i].seq)) {
            failed_record(out[i], len, SHM_ERR_ARG);
            rf.note(i, SHM_ERR_ARG, "invalid read (negative length or no bases)");
        } else if (mp->max_read_len > 0 && len > mp->max_read_len) {
            failed_record(out[i], len, SHM_ERR_CAPACITY);
            char m[128];
            snprintf(m, sizeof m, "read of %d bases is over the mapper's limit of %d", len, mp->max_read_len);
            rf.note(i, SHM_ERR_CAPACITY, m);
        } else {
            ids.push_back(i);
        }
    }
    RCHK(map_isolating(mp, reads, ids.data(), (int32_t)ids.size(), out, rf));
    if (rf.read >= 0) return fail(rf.code, "read %d failed: %s", rf.read, rf.msg.c_str());
    return SHM_OK;
}

extern "C" int shm_mapper_set_read_limit(shm_mapper *mp, int32_t max_read_len) {
    if (!mp || max_read_len < 0) return fail(SHM_ERR_ARG, "invalid argument");
    mp->max_read_len = max_read_len;
    return SHM_OK;
}

// ---------------------------------------------------------------------------
// 2-bit packed reads
// ---------------------------------------------------------------------------

// Packs reads [r0, r1) of a batch (bases at offsets off[]) into `packed`, and
// lists their runs of non-ACGT bytes.  Bytes shared with the neighbouring
// ranges (a read boundary inside a byte) are or-ed atomically; `packed` starts
// zeroed.
static int64_t pack_range(const shm_read *reads, const int64_t *off, int32_t r0, int32_t r1, uint8_t *packed,
                          std::vector<int64_t> &exc) {
    static const int8_t code[256] = {
#define X4(a) a, a, a, a
#define X16(a) X4(a), X4(a), X4(a), X4(a)
        X16(-1), X16(-1), X16(-1), X16(-1),
        -1, 0, -1, 1, -1, -1, -1, 2, -1, -1, -1, -1, -1, -1, -1, -1,  // @ A B C D E F G
        -1, -1, -1, -1, 3, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1,  // T
        -1, 0, -1, 1, -1, -1, -1, 2, -1, -1, -1, -1, -1, -1, -1, -1,  // ` a b c d e f g
        -1, -1, -1, -1, 3, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1,  // t
        X16(-1), X16(-1), X16(-1), X16(-1), X16(-1), X16(-1), X16(-1), X16(-1)
#undef X16
#undef X4
    };
    const int64_t lo = off[r0], hi = off[r1];
    int64_t run = -1;  // start of the open exception run
    for (int32_t r = r0; r < r1; r++) {
        const uint8_t *s = reads[r].seq;
        const int64_t at = off[r];
        for (int64_t i = 0; i < reads[r].len; i++) {
            const int64_t q = at + i;
            const int c = code[s[i]];
            if (c < 0) {
                if (run < 0) run = q;
            } else {
                if (run >= 0) {
                    exc.push_back(run);
                    exc.push_back(q);
                    run = -1;
                }
                const uint8_t bits = (uint8_t)(c << (2 * (q & 3)));
                if (bits == 0) continue;
                uint8_t *b = packed + (q >> 2);
                const bool shared = (q >> 2) == (lo >> 2) || (q >> 2) == ((hi - 1) >> 2);
                if (shared) __atomic_fetch_or(b, bits, __ATOMIC_RELAXED);
                else *b |= bits;
            }
        }
        if (run >= 0) {  // runs end at read boundaries
            exc.push_back(run);
            exc.push_back(at + reads[r].len);
            run = -1;
        }
    }
    return (int64_t)exc.size() / 2;
}

extern "C" int shm_pack_reads(const shm_read *reads, int32_t n, uint8_t *packed, int64_t packed_cap, int64_t *off,
                              int64_t *exc, int64_t exc_cap, int64_t *n_exc, int32_t threads) {
    if (n < 0 || (n > 0 && (!reads || !off)) || !n_exc || exc_cap < 0 || (exc_cap > 0 && !exc))
        return fail(SHM_ERR_ARG, "invalid argument");
    off[0] = 0;
    for (int32_t i = 0; i < n; i++) {
        if (reads[i].len < 0 || (reads[i].len > 0 && !reads[i].seq)) return fail(SHM_ERR_ARG, "read %d invalid", i);
        off[i + 1] = off[i] + reads[i].len;
    }
    const int64_t nbytes = (off[n] + 3) / 4;
    if (packed_cap < nbytes)
        return fail(SHM_ERR_CAPACITY, "%lld bases need %lld packed bytes, the buffer holds %lld", (long long)off[n],
                    (long long)nbytes, (long long)packed_cap);
    if (nbytes > 0 && !packed) return fail(SHM_ERR_ARG, "invalid argument");
    memset(packed, 0, (size_t)nbytes);
    int32_t nt = threads > 0 ? threads : 1;
    if (nt > 64) nt = 64;
    if (nt > n) nt = n > 0 ? n : 1;
    std::vector<std::vector<int64_t>> part(nt);
    std::vector<std::thread> th;
    for (int32_t t = 0; t < nt; t++) {
        // read ranges of about equal bases
        auto bound = [&](int32_t tt) {
            if (tt == 0) return (int32_t)0;
            if (tt == nt) return n;
            return (int32_t)(std::upper_bound(off, off + n + 1, off[n] * tt / nt) - off - 1);
        };
        const int32_t a = bound(t), b = bound(t + 1);
        th.emplace_back([&, a, b, t] { if (a < b) pack_range(reads, off, a, b, packed, part[t]); });
    }
    for (std::thread &x : th) x.join();
    int64_t total = 0;
    for (const auto &p : part) total += (int64_t)p.size() / 2;
    *n_exc = total;
    if (total > exc_cap)
        return fail(SHM_ERR_CAPACITY, "%lld runs of non-ACGT bytes, the list holds %lld", (long long)total,
                    (long long)exc_cap);
    int64_t k = 0;
    for (const auto &p : part)
        for (int64_t v : p) exc[k++] = v;
    return SHM_OK;
}

extern "C" int shm_map_batch_packed(shm_mapper *mp, const uint8_t *packed, const int64_t *off, int32_t n,
                                    const int64_t *exc, int64_t n_exc, shm_read_result *out) {
    if (!mp || n < 0 || (n > 0 && (!off || !out)) || n_exc < 0 || (n_exc > 0 && !exc))
        return fail(SHM_ERR_ARG, "invalid argument");
    HIPCHK(hipSetDevice(mp->idx->device));
    const int64_t nb = n > 0 ? off[n] : 0;
    if (nb < 0 || (nb > 0 && !packed)) return fail(SHM_ERR_ARG, "invalid argument");
    for (int32_t i = 0; i < n; i++)
        if (off[i + 1] < off[i] || off[i + 1] - off[i] > INT32_MAX) return fail(SHM_ERR_ARG, "read %d invalid", i);
    for (int64_t r = 0; r < n_exc; r++)
        if (exc[2 * r] < (r ? exc[2 * r - 1] : 0) || exc[2 * r + 1] < exc[2 * r] || exc[2 * r + 1] > nb)
            return fail(SHM_ERR_ARG, "exception runs must be sorted, disjoint and inside the bases");
    // reads over the length limit fail alone: the unpacked bases go through
    // shm_map_batch, which isolates them
    bool over = false;
    for (int32_t i = 0; i < n && !over; i++) over = mp->max_read_len > 0 && off[i + 1] - off[i] > mp->max_read_len;
    const int64_t nbytes = (nb + 3) / 4;
    RCHK(mp->packed.ensure((uint64_t)(nbytes + 3) / 4 + 1));
    RCHK(mp->exc.ensure(2 * (uint64_t)n_exc + 2));
    RCHK(mp->bases.ensure(nb + 64));
    RCHK(mp->roff.ensure((size_t)read_class(n) + 1));  // room to pad a captured batch
    UploadTurn turn(mp, (uint64_t)nbytes);
    if (nbytes > 0) HIPCHK(hipMemcpyAsync(mp->packed.p, packed, (size_t)nbytes, hipMemcpyHostToDevice, mp->st));
    if (n_exc > 0) HIPCHK(hipMemcpyAsync(mp->exc.p, exc, (size_t)n_exc * 16, hipMemcpyHostToDevice, mp->st));
    HIPCHK(hipMemcpyAsync(mp->roff.p, off, (size_t)(n + 1) * 8, hipMemcpyHostToDevice, mp->st));
    RCHK(turn.done());
    launch_unpack_bases((const uint8_t *)mp->packed.p, (uint64_t)nb, mp->exc.p, n_exc, mp->bases.p, mp->st);
    HIPCHK(hipGetLastError());
    if (over) {
        std::vector<uint8_t> h((size_t)nb + 1);
        if (nb > 0) HIPCHK(hipMemcpyAsync(h.data(), mp->bases.p, (size_t)nb, hipMemcpyDeviceToHost, mp->st));
        HIPCHK(hipStreamSynchronize(mp->st));
        std::vector<shm_read> rs(n);
        for (int32_t i = 0; i < n; i++) rs[i] = shm_read{h.data() + off[i], (int32_t)(off[i + 1] - off[i])};
        return shm_map_batch(mp, rs.data(), n, out);
    }
    return map_device(mp, mp->bases.p, mp->roff.p, n, out, off);
}

extern "C" int shm_map_batch_device(shm_mapper *mp, const uint8_t *d_seq, const int64_t *d_off, int32_t n,
                                    shm_read_result *out) {
    if (!mp || n < 0 || (n > 0 && (!d_seq || !d_off || !out))) return fail(SHM_ERR_ARG, "invalid argument");
    HIPCHK(hipSetDevice(mp->idx->device));
    // no host copy of the offsets: the synchronised path (a read-back of the
    // offsets first costs more than its four synchronisations on large batches)
    return map_device(mp, d_seq, d_off, n, out);
}

// ---------------------------------------------------------------------------
// PAF text (src/mapping.rs:87-107, 161-182, 228-247)
// ---------------------------------------------------------------------------

// Rust's {:.N} on f64 prints the exact binary value rounded half to even, as
// glibc printf does in the default rounding mode, but spells NaN and the
// infinities "NaN", "inf" and "-inf".
static void rust_fixed(char *buf, size_t cap, double v, int prec) {
    if (isnan(v)) snprintf(buf, cap, "NaN");
    else if (isinf(v)) snprintf(buf, cap, v > 0 ? "inf" : "-inf");
    else snprintf(buf, cap, "%.*f", prec, v);
}

extern "C" int shm_format_paf(const shm_index *idx, const shm_map_params *p, const char *qid,
                              const shm_read_result *r, double map_time, char *buf, size_t cap) {
    if (!idx || !p || !qid || !r || !r->mapped || r->segm_id < 0 || r->segm_id >= idx->nseg) {
        fail(SHM_ERR_ARG, "not a mapped record");
        return -1;
    }
    char isd[48], j[48], j2[48], jd[48], sh[48], t[48], mi[48];
    rust_fixed(isd, sizeof isd, r->sigmas, 5);
    rust_fixed(j, sizeof j, r->J, 5);
    rust_fixed(j2, sizeof j2, r->J2, 5);
    rust_fixed(jd, sizeof jd, r->J - r->J2, 5);
    rust_fixed(sh, sizeof sh, r->sh, 5);
    rust_fixed(t, sizeof t, map_time, 5);
    rust_fixed(mi, sizeof mi, r->match_inefficiency, 3);
    return snprintf(buf, cap,
                    // MappingPaf: 12 mandatory columns; 10 and 11 repeat the read length
                    "%s\t%d\t%d\t%d\t%c\t%s\t%d\t%d\t%d\t%d\t%d\t%d"
                    // LocalMappingStats
                    "\tp:i:%d\ts:i:%d\tI:i:%d\tI2:i:%d\tIdiff:i:%d\tIsdiff:f:%s\tJ:f:%s\tJ2:f:%s"
                    "\tJdiff:f:%s\tsh:f:%s\tstrand:i:%d\tt:f:%s\tb:s:(%d,%d)\tb2:s:(%d,%d)"
                    // GlobalMappingStats
                    "\tk:i:%d\tgt_J:f:-1.000\tgt_C:f:-1.000\tgt_C_bucket:f:-1.000\tseeds:i:%d"
                    "\tmax_seed_matches:i:%d\tseed_matches:i:%d\ttotal_matches:i:%d"
                    "\tmatch_inefficiency:f:%s\tseeded_buckets:i:%d\tfinal_buckets:i:%d\tFPTP:f:-1.000",
                    qid, r->read_len, r->p_start, r->p_end, (char)r->strand, idx->names[r->segm_id].c_str(),
                    idx->sz[r->segm_id], r->t_l, r->t_r, r->read_len, r->read_len, r->mapq, r->m, r->s_sz,
                    r->intersection, r->intersection2, r->intersection - r->intersection2, isd, j, j2, jd, sh,
                    r->same_strand, t, r->b_segm, r->b_b, r->b2_segm, r->b2_b, p->k, r->S, r->max_seed_matches,
                    r->seed_matches, r->total_matches, mi, r->seeded_buckets, r->final_buckets);
}

extern "C" int shm_format_unmapped(const char *qid, int32_t read_len, char *buf, size_t cap) {
    return snprintf(buf, cap, "%s\t%d\t0\t0\t*\t*\t0\t0\t0\t0\t0\t0", qid ? qid : "", read_len);
}
