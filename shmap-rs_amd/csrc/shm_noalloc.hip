// shm_sketch_noalloc: FracMinHash::sketch_slice_into (src/sketch.rs:167-209)
// on the GPU with no device (or host) allocation, for callers that cannot
// allocate -- code under stream capture, tight serving loops, harnesses that
// forbid hipMalloc.  Every intermediate of the sketch pipeline (the unit
// table, lane-strided staging, per-chunk counts, the compacted k-mers and
// their packed records) is carved from one fixed device array per GPU,
// `scratch`, of SHM_NOALLOC_SCRATCH_BYTES.  Each GPU the library runs on holds
// that array until the process ends.  Calls on one GPU share it: this file
// serialises them with a host mutex per device, so calls from the threads of
// one process never overlap; a call holds it until its results are back.
#include <hip/hip_runtime.h>

#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <mutex>

#include "../../include/shmap_hip.h"
#include "shm_device.hpp"
#include "shm_kernels.hpp"

using namespace shm;

static constexpr uint64_t SCRATCH_BYTES = SHM_NOALLOC_SCRATCH_BYTES;
static_assert(SCRATCH_BYTES % 256 == 0, "scratch regions are 256-byte aligned");
static constexpr int MAX_DEVICES = 64;

__device__ __align__(256) uint8_t scratch[SCRATCH_BYTES];

namespace {

struct DeviceScratch {
    std::mutex mu;           // one call at a time on this device
    bool ready = false;
    uint8_t *ptr = nullptr;  // this device's address of `scratch`
    uint32_t max_grid = 0;
};
DeviceScratch g_dev[MAX_DEVICES];
std::mutex g_setup;

int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    shm::set_error(fmt, ap);
    va_end(ap);
    return code;
}

// Byte offsets of every region of one call's scratch (each 256-aligned).
struct Layout {
    uint64_t bases, off, nunit, unit_start, koff, part, stat, ovf_n, chunk_cnt, chunk_koff, lane_cnt, lane_pre,
        unit_seq, stage, ovf, out_h, out_r, out_s, packed, total;
};

uint64_t carve(uint64_t &at, uint64_t bytes) {
    const uint64_t o = at;
    at = (at + bytes + 255) & ~255ull;
    return o;
}

// The regions for a sequence of len bases with units U (chunks C) and room
// for mcap k-mers.
Layout layout(int64_t len, uint64_t U, uint64_t C, uint32_t lcap, uint64_t mcap) {
    Layout L{};
    uint64_t at = 0;
    L.bases = carve(at, (uint64_t)len + 64);
    L.off = carve(at, 2 * 8);
    L.nunit = carve(at, 2 * 8);
    L.unit_start = carve(at, 2 * 8);
    L.koff = carve(at, 2 * 8);
    L.part = carve(at, (SCAN_PARTS + 1) * 8);
    L.stat = carve(at, 8);
    L.ovf_n = carve(at, 4);
    L.chunk_cnt = carve(at, (C + 1) * 8);
    L.chunk_koff = carve(at, (C + 1) * 8);
    L.lane_cnt = carve(at, (C * WAVE + 1) * 4);
    L.lane_pre = carve(at, (C * WAVE + 1) * 4);
    L.unit_seq = carve(at, (U + 1) * 4);
    L.stage = carve(at, (C * WAVE * lcap + 1) * 16);
    L.ovf = carve(at, (C + 1) * 4);
    L.out_h = carve(at, (mcap + 1) * 8);
    L.out_r = carve(at, (mcap + 1) * 4);
    L.out_s = carve(at, mcap + 1);
    L.packed = carve(at, (mcap + 1) * sizeof(shm_kmer));
    L.total = at;
    return L;
}

__global__ void k_zero_u32(uint32_t *p, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

// The compacted (hash, position, strand) arrays as shm_kmer records, the
// caller's offset added to every position (src/sketch.rs:196-199).
__global__ void k_pack_kmers(const uint64_t *h, const int32_t *r, const uint8_t *s, uint64_t n, int32_t offset,
                             shm_kmer *out) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        shm_kmer k;
        k.h = h[i];
        k.r = r[i] + offset;
        k.strand = s[i];
        k.pad[0] = k.pad[1] = k.pad[2] = 0;
        out[i] = k;
    }
}

// Once per device, with that device current: the device address of its copy
// of `scratch` (the host-side name is not one), and the grid bound.
int setup(int dev, DeviceScratch *&ds) {
    if (dev < 0 || dev >= MAX_DEVICES) return fail(SHM_ERR_ARG, "device %d is beyond this library's %d", dev, MAX_DEVICES);
    ds = &g_dev[dev];
    std::lock_guard<std::mutex> lk(g_setup);
    if (ds->ready) return SHM_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(SHM_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d failed: %s", dev, hipGetErrorString(e));
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 8;
    ds->ptr = (uint8_t *)p;
    ds->max_grid = (uint32_t)cus * 8;
    ds->ready = true;
    return SHM_OK;
}

}  // namespace

#define NCHK(expr)                                                                                       \
    do {                                                                                                 \
        const hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess) return fail(SHM_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_));          \
    } while (0)

extern "C" int64_t shm_sketch_noalloc_bytes(const shm_sketcher *sk, int64_t len, int64_t cap) {
    if (!sk || len < 0 || cap < 0) return -1;
    const int64_t nw = len >= sk->k ? len - sk->k + 1 : 0;
    if (nw == 0) return 0;
    const uint64_t U = (uint64_t)((nw + LANE_W - 1) / LANE_W) + sketch_halo(sk->k), C = (U + WAVE - 1) / WAVE;
    const uint64_t mcap = (uint64_t)(cap < nw ? cap : nw);
    return (int64_t)layout(len, U, C, sketch_lane_cap(sk->thres), mcap).total;
}

extern "C" int shm_sketch_noalloc(const shm_sketcher *sk, const uint8_t *seq, int64_t len, int32_t offset,
                                  shm_kmer *out, int64_t cap, int64_t *n, void *stream) {
    if (!sk || (!seq && len > 0) || !n || len < 0 || cap < 0 || (cap > 0 && !out))
        return fail(SHM_ERR_ARG, "invalid argument");
    *n = 0;
    // (s.len() as RPos) < k yields nothing (src/sketch.rs:170-173)
    const int64_t nw = len >= sk->k ? len - sk->k + 1 : 0;
    if (nw == 0) return SHM_OK;
    if (len > INT32_MAX) return fail(SHM_ERR_ARG, "sequence of %lld bases is over 2^31-1", (long long)len);
    // the unit table of one sequence, as k_unit_counts builds it
    const uint64_t U = (uint64_t)((nw + LANE_W - 1) / LANE_W) + sketch_halo(sk->k), C = (U + WAVE - 1) / WAVE;
    const uint32_t lcap = sketch_lane_cap(sk->thres);
    const uint64_t mcap = (uint64_t)(cap < nw ? cap : nw);
    const Layout L = layout(len, U, C, lcap, mcap);
    if (L.total > SCRATCH_BYTES)
        return fail(SHM_ERR_CAPACITY,
                    "a %lld-base sequence needs %llu bytes of scratch at k=%d r=%g; the fixed array holds %llu "
                    "(use shm_sketch)",
                    (long long)len, (unsigned long long)L.total, sk->k, sk->h_frac,
                    (unsigned long long)SCRATCH_BYTES);
    int dev = 0;
    NCHK(hipGetDevice(&dev));
    DeviceScratch *ds = nullptr;
    if (const int rc = setup(dev, ds)) return rc;
    std::lock_guard<std::mutex> lk(ds->mu);
    hipStream_t st = (hipStream_t)stream;
    uint8_t *S0 = ds->ptr;
    uint8_t *bases = S0 + L.bases;
    int64_t *d_off = (int64_t *)(S0 + L.off);
    uint64_t *nunit = (uint64_t *)(S0 + L.nunit), *unit_start = (uint64_t *)(S0 + L.unit_start);
    uint64_t *koff = (uint64_t *)(S0 + L.koff), *part = (uint64_t *)(S0 + L.part), *stat = (uint64_t *)(S0 + L.stat);
    uint32_t *ovf_n = (uint32_t *)(S0 + L.ovf_n), *ovf = (uint32_t *)(S0 + L.ovf);
    uint64_t *chunk_cnt = (uint64_t *)(S0 + L.chunk_cnt), *chunk_koff = (uint64_t *)(S0 + L.chunk_koff);
    uint32_t *lane_cnt = (uint32_t *)(S0 + L.lane_cnt), *lane_pre = (uint32_t *)(S0 + L.lane_pre);
    uint32_t *unit_seq = (uint32_t *)(S0 + L.unit_seq);
    uint4 *stage = (uint4 *)(S0 + L.stage);
    uint64_t *out_h = (uint64_t *)(S0 + L.out_h);
    int32_t *out_r = (int32_t *)(S0 + L.out_r);
    uint8_t *out_s = S0 + L.out_s;
    shm_kmer *packed = (shm_kmer *)(S0 + L.packed);

    // inputs; the array keeps the previous call's contents, so the counters
    // start from zero here (a fill kernel: no memset node in a captured stream)
    const int64_t h_off[2] = {0, len};
    NCHK(hipMemcpyAsync(bases, seq, (size_t)len, hipMemcpyHostToDevice, st));
    NCHK(hipMemcpyAsync(d_off, h_off, sizeof h_off, hipMemcpyHostToDevice, st));
    k_zero_u32<<<1, 64, 0, st>>>(ovf_n, 1);
    k_zero_u32<<<1, 64, 0, st>>>((uint32_t *)stat, 2);
    const SeqSet S{bases, d_off, 1};
    launch_unit_counts(S, sk->k, nunit, st);
    launch_scan(nunit, unit_start, nullptr, 1, part, st);
    launch_unit_seq(1, unit_start, unit_seq, st);
    launch_sketch(0, S, sk->k, sk->thres, unit_start, unit_seq, U, chunk_cnt, lane_cnt, lane_pre, stage, nullptr,
                  nullptr, nullptr, lcap, ovf, ovf_n, C, ds->max_grid, st);
    launch_scan(chunk_cnt, chunk_koff, nullptr, C, part, st);
    launch_seq_koff(1, unit_start, chunk_koff, lane_pre, koff, stat, st);
    NCHK(hipGetLastError());
    uint64_t M = 0;
    uint32_t n_overflow = 0;
    NCHK(hipMemcpyAsync(&M, koff + 1, 8, hipMemcpyDeviceToHost, st));
    NCHK(hipMemcpyAsync(&n_overflow, ovf_n, 4, hipMemcpyDeviceToHost, st));
    NCHK(hipStreamSynchronize(st));
    *n = (int64_t)M;
    if ((int64_t)M > cap)
        return fail(SHM_ERR_CAPACITY, "sketch has %llu k-mers, buffer holds %lld", (unsigned long long)M,
                    (long long)cap);
    launch_sketch_compact(C, lane_cnt, chunk_koff, lcap, stage, out_h, out_r, out_s, ds->max_grid, st);
    if (n_overflow > 0)
        launch_sketch(1, S, sk->k, sk->thres, unit_start, unit_seq, U, chunk_koff, lane_cnt, lane_pre, nullptr,
                      out_h, out_r, out_s, lcap, ovf, ovf_n, n_overflow, ds->max_grid, st);
    if (M > 0) {
        const uint32_t grid = (uint32_t)((M + 255) / 256 < 4096 ? (M + 255) / 256 : 4096);
        k_pack_kmers<<<grid, 256, 0, st>>>(out_h, out_r, out_s, M, offset, packed);
        NCHK(hipGetLastError());
        NCHK(hipMemcpyAsync(out, packed, M * sizeof(shm_kmer), hipMemcpyDeviceToHost, st));
    }
    NCHK(hipStreamSynchronize(st));
    return SHM_OK;
}
