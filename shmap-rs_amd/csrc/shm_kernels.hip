// gfx950 kernels of the mapping path.
//
// Sketching (src/sketch.rs:167-219) runs one wave64 per 4096-window chunk of a
// sequence with a k-1 base halo staged in LDS; the index build
// (src/index.rs:384-663) is a device radix sort of (hash, k-mer id) followed by
// run-length encoding into an 8-shard open-addressing table; per-read mapping
// (src/shmap/mod.rs:653-971) is a sequence of persistent kernels over a batch of
// reads, each read's intermediate data living in flat arrays at per-read
// offsets.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <mutex>
#include <vector>

#include "shm_device.hpp"
#include "shm_kernels.hpp"

namespace shm {

// ============================================================================
// Device-wide exclusive scan (three phases over a fixed grid).  n is read from
// device memory so no host round trip is needed to launch it.
// ============================================================================

__global__ void __launch_bounds__(WG) k_scan_reduce(const uint64_t *in, const uint64_t *n_ptr,
                                                    uint64_t n_static, uint64_t *part) {
    const uint64_t n = n_ptr ? *n_ptr : n_static;
    const uint64_t per = (n + gridDim.x - 1) / gridDim.x;
    const uint64_t b = (uint64_t)blockIdx.x * per, e = b + per < n ? b + per : n;
    uint64_t s = 0;
    for (uint64_t i = b + threadIdx.x; i < e; i += WG) s += in[i];
    __shared__ uint64_t tmp[NWAVES + 1];
    uint64_t tot;
    wg_excl_scan64(s, tmp, &tot);
    if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(WG) k_scan_top(uint64_t *part, uint32_t nparts) {
    __shared__ uint64_t tmp[NWAVES + 1];
    uint64_t carry = 0;
    for (uint32_t base = 0; base < nparts; base += WG) {
        const uint32_t i = base + threadIdx.x;
        const uint64_t v = i < nparts ? part[i] : 0;
        uint64_t tot;
        const uint64_t ex = wg_excl_scan64(v, tmp, &tot);
        if (i < nparts) part[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) part[nparts] = carry;
}

__global__ void __launch_bounds__(WG) k_scan_down(const uint64_t *in, uint64_t *out,
                                                  const uint64_t *n_ptr, uint64_t n_static,
                                                  const uint64_t *part) {
    const uint64_t n = n_ptr ? *n_ptr : n_static;
    const uint64_t per = (n + gridDim.x - 1) / gridDim.x;
    const uint64_t b = (uint64_t)blockIdx.x * per, e = b + per < n ? b + per : n;
    __shared__ uint64_t tmp[NWAVES + 1];
    uint64_t carry = part[blockIdx.x];
    for (uint64_t base = b; base < e; base += WG) {
        const uint64_t i = base + threadIdx.x;
        const uint64_t v = i < e ? in[i] : 0;
        uint64_t tot;
        const uint64_t ex = wg_excl_scan64(v, tmp, &tot);
        if (i < e) out[i] = carry + ex;
        carry += tot;
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = part[gridDim.x];
}

// ============================================================================
// Sketching -- FracMinHash::sketch_slice_into (src/sketch.rs:167-219)
// ============================================================================

__device__ __forceinline__ uint32_t upper_bound_u64(const uint64_t *a, uint32_t n, uint64_t x) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a[mid] <= x) lo = mid + 1; else hi = mid;
    }
    return lo;
}

__global__ void __launch_bounds__(WG) k_unit_counts(SeqSet S, int32_t k, uint64_t *nunit) {
    const uint32_t halo = sketch_halo(k);
    for (int32_t i = blockIdx.x * WG + threadIdx.x; i < S.n; i += gridDim.x * WG) {
        const int64_t len = S.off[i + 1] - S.off[i];
        const int64_t nw = (len >= k && len >= 1) ? len - k + 1 : 0;
        nunit[i] = nw > 0 ? (uint64_t)((nw + LANE_W - 1) / LANE_W) + halo : 0;
    }
}

// Sequence of every unit: one wave per sequence writes its run of unit_seq.
__global__ void __launch_bounds__(WG) k_unit_seq(int32_t n, const uint64_t *unit_start, uint32_t *unit_seq,
                                                 const uint32_t *abort) {
    if (aborted(abort)) return;
    const int lane = threadIdx.x & (WAVE - 1);
    for (int64_t s = ((int64_t)blockIdx.x * WG + threadIdx.x) / WAVE; s < n; s += (int64_t)gridDim.x * NWAVES) {
        const uint64_t a = unit_start[s], e = unit_start[s + 1];
        for (uint64_t u = a + (uint64_t)lane; u < e; u += WAVE) unit_seq[u] = (uint32_t)s;
    }
}

// Codes of four packed bytes: ((c >> 1) ^ (c >> 2)) & 3 sends a/c/g/t in
// either case to 0/1/2/3; any other byte (whose ntHash seed is 0,
// src/sketch.rs:73-88) gets bit 2 set (codes 4-7).  Validity is checked by
// rebuilding each letter from its code with a byte permute and testing the
// difference for a nonzero byte.
__device__ __forceinline__ uint32_t base_codes4(uint32_t v) {
    const uint32_t x = v | 0x20202020u;                               // lower case
    const uint32_t code = ((x >> 1) ^ (x >> 2)) & 0x03030303u;
    const uint32_t back = __builtin_amdgcn_perm(0x74676361u, 0x74676361u, code); // "acgt"[code]
    const uint32_t d = x ^ back;                                      // nonzero byte = not ACGT
    const uint32_t nz = (((d & 0x7f7f7f7fu) + 0x7f7f7f7fu) | d) & 0x80808080u;
    return code | (nz >> 5);
}

// Four code bytes (each < 16) to four nibbles, first byte lowest.
__device__ __forceinline__ uint32_t pack_nibbles4(uint32_t c) {
    c = (c | (c >> 4)) & 0x00ff00ffu;
    return (c | (c >> 8)) & 0x0000ffffu;
}

// Eight bases (v0 then v1, first base lowest) to eight 4-bit codes as
// base_codes4 gives them.  A, C, G, T (either case) differ in their low three
// bits (1, 3, 7, 4), so byte permutes over 8-entry tables indexed by those bits
// give each byte's code, its code times 16, and the lower-case letter the code
// stands for; when every byte is that letter (in either case) the codes pack
// with one shift-or per dword and one byte gather.  Any other byte takes the
// general path.
__device__ __forceinline__ uint32_t stage_codes8(uint32_t v0, uint32_t v1) {
    constexpr uint32_t LET_LO = 0x63006100u, LET_HI = 0x67000074u;  // low 3 bits -> letter: 1 a, 3 c, 4 t, 7 g
    constexpr uint32_t COD_LO = 0x01000000u, COD_HI = 0x02000003u;  // -> code
    constexpr uint32_t C16_LO = 0x10000000u, C16_HI = 0x20000030u;  // -> code << 4
    const uint32_t s0 = v0 & 0x07070707u, s1 = v1 & 0x07070707u;
    const uint32_t d = ((v0 | 0x20202020u) ^ __builtin_amdgcn_perm(LET_HI, LET_LO, s0)) |
                       ((v1 | 0x20202020u) ^ __builtin_amdgcn_perm(LET_HI, LET_LO, s1));
    if (__builtin_expect(d != 0u, 0))
        return pack_nibbles4(base_codes4(v0)) | (pack_nibbles4(base_codes4(v1)) << 16);
    const uint32_t t0 = (__builtin_amdgcn_perm(COD_HI, COD_LO, s0) << 8) | __builtin_amdgcn_perm(C16_HI, C16_LO, s0);
    const uint32_t t1 = (__builtin_amdgcn_perm(COD_HI, COD_LO, s1) << 8) | __builtin_amdgcn_perm(C16_HI, C16_LO, s1);
    return __builtin_amdgcn_perm(t1, t0, 0x07050301u);  // bytes 1 and 3 of t0, then of t1
}

// Byte b of v times 16 (a 16-byte table entry's offset) as one byte-select
// shift.  Written out because the compiler, knowing the bytes' high bits are
// clear, otherwise emits a shift and a mask.  b is a constant after unrolling.
template <int B> __device__ __forceinline__ uint32_t byte_x16_c(uint32_t v) {
    uint32_t r;
    if constexpr (B == 0)
        asm("v_lshlrev_b32_sdwa %0, 4, %1 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:BYTE_0" : "=v"(r) : "v"(v));
    else if constexpr (B == 1)
        asm("v_lshlrev_b32_sdwa %0, 4, %1 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:BYTE_1" : "=v"(r) : "v"(v));
    else if constexpr (B == 2)
        asm("v_lshlrev_b32_sdwa %0, 4, %1 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:BYTE_2" : "=v"(r) : "v"(v));
    else
        asm("v_lshlrev_b32_sdwa %0, 4, %1 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:BYTE_3" : "=v"(r) : "v"(v));
    return r;
}
__device__ __forceinline__ uint32_t byte_x16(uint32_t v, int b) {
    switch (b) {
        case 0: return byte_x16_c<0>(v);
        case 1: return byte_x16_c<1>(v);
        case 2: return byte_x16_c<2>(v);
        default: return byte_x16_c<3>(v);
    }
}

// One wave per chunk of WAVE units, one unit per lane: unit u of sequence s is
// its windows [LANE_W * (u - unit_start[s]), + LANE_W), and every sequence ends
// with sketch_halo(k) halo units that hold bases only, so the lanes of a wave
// may roll windows of several sequences.  The chunk's WAVE + halo unit slots
// (LANE_W bases each, from the unit's first base on) are staged in LDS as 4-bit
// codes, base p of slot q at nibble p & 7 of code dword 16q + p / 8, with one
// dword of bank skew per 16 dwords (lanes 16 dwords apart hit different
// banks); a lane reads its own slot and the next halo slots, which always hold
// the continuation of its sequence.  The
// rolling update (src/sketch.rs:90-97, 202-203) xors, per strand, the outgoing
// base's rotated seed and the incoming base's; both come from one 64-entry LDS
// table indexed by (outgoing code << 3 | incoming code), entry
// (rotl(fw_out, k) ^ fw_in, rotr(rc_out, 1) ^ rotl(rc_in, k-1)), one 16-byte
// load per window (codes 4-7 are non-ACGT bytes, seed 0).  Each lane rolls
// up to LANE_W consecutive windows; its first window is rolled in from a zero hash
// two bases per step (a second 64-entry table of base pairs) and, for odd k,
// one more through the outgoing-code-4 row, which equals the direct definition
// of src/sketch.rs:180-184.
//
// PASS 0 visits every chunk and writes each lane's kept k-mers, in position
// order, to lane-strided staging slots (slot i of lane l of chunk c at
// (c*lcap + i)*WAVE + l, for i < lcap); it stores every lane's full kept count
// and the chunk's total, and lists the chunks in which a lane kept more than
// lcap.  PASS 1 re-rolls only the listed chunks and writes straight to the final
// positions: the chunk's scanned offset plus the lane's exclusive prefix.
template <int PASS>
__global__ void __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(7))) k_sketch(SeqSet S, int32_t k, uint64_t thres,
                                               const uint64_t *unit_start, const uint32_t *unit_seq,
                                               uint64_t n_units, uint64_t *chunk_val, uint32_t *lane_cnt,
                                               uint32_t *lane_pre, uint4 *stage,
                                               uint64_t *out_h, int32_t *out_r, uint8_t *out_s, uint32_t lcap,
                                               uint32_t *ovf_list, uint32_t *ovf_count, const uint32_t *abort,
                                               const uint64_t *n_units_dev) {
    if (aborted(abort)) return;
    if (n_units_dev) n_units = *n_units_dev;  // sized on the device (the grid covers a capacity)
    extern __shared__ uint8_t lds[];
    // T[(out code << 3) | in code]: the xor terms of both strands for one step
    // (row 4, an N rolling out, adds nothing: the lane starts use it).
    // T2[(a << 3) | b]: two bases a, b rolled in from nothing, for lane starts.
    __shared__ ulonglong2 T[64], T2[64];
    const int lane = threadIdx.x & (WAVE - 1), w = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
    const uint32_t span = sketch_span(k);
    uint8_t *buf = lds + (size_t)w * span;
    if (threadIdx.x < 64) {
        const uint8_t co = "ACGTNNNN"[threadIdx.x >> 3], ci = "ACGTNNNN"[threadIdx.x & 7];  // code order of base_codes4
        const uint64_t ri = rotl64(nt_rc(ci), (uint32_t)(k - 1)), rl = rotl64(nt_rc(co), (uint32_t)(k - 1));
        T[threadIdx.x] = make_ulonglong2(rotl64(nt_fw(co), (uint32_t)k) ^ nt_fw(ci), rotr64(nt_rc(co), 1) ^ ri);
        T2[threadIdx.x] = make_ulonglong2(rotl64(nt_fw(co), 1) ^ nt_fw(ci), rotr64(rl, 1) ^ ri);
    }
    __syncthreads();
    const uint32_t thres_hi = (uint32_t)(thres >> 32);
    const uint32_t halo = sketch_halo(k), nslot = (uint32_t)WAVE + halo;
    uint64_t *sbase = (uint64_t *)(buf + ((sketch_codes_bytes(k) + 7u) & ~7u));  // slot q's first base, or ~0
    const uint64_t total = PASS == 0 ? (n_units + WAVE - 1) / WAVE : (uint64_t)*ovf_count;
    uint32_t *b32 = (uint32_t *)buf;
#define SKW(q) b32[(uint32_t)(q) + ((uint32_t)(q) >> 4)]
#define SKN(p) ((SKW((uint32_t)(p) >> 3) >> (((uint32_t)(p) & 7u) * 4u)) & 0xfu)
    static_assert(LANE_W == 128, "a unit slot is 16 code dwords, one skew step");
    // Each wave stages and reads only its own slice, so the waves of a
    // workgroup run independently (wave-level ordering only).
    for (uint64_t ci = (uint64_t)blockIdx.x * NWAVES + w; ci < total; ci += (uint64_t)gridDim.x * NWAVES) {
        const uint64_t c = PASS == 0 ? ci : (uint64_t)ovf_list[ci];
        int64_t ws = 0;    // the lane's first window in its sequence
        int32_t nwl = 0;   // its window count (0 for a halo unit or past the last unit)
        {
            // slot table: the first base of each slot's unit, as an offset into
            // S.bases; bit 63 marks a slot to stage byte by byte: one past the
            // last unit (all ~0) or one whose aligned dwords may reach past
            // the end of the set
            const uint64_t set_end = (uint64_t)S.off[S.n];
            auto slot_entry = [&](uint64_t u) -> uint64_t {
                if (u >= n_units) return ~0ull;
                const uint32_t sq = unit_seq[u];
                const uint64_t b = (uint64_t)(S.off[sq] + (int64_t)(u - unit_start[sq]) * LANE_W);
                return (b & ~3ull) + 4u * (2u * (LANE_W / 8) + 1u) <= set_end ? b : b | (1ull << 63);
            };
            const uint64_t u = c * WAVE + (uint64_t)lane;
            if (u < n_units) {
                const uint32_t sq = unit_seq[u];
                const int64_t nw = S.off[sq + 1] - S.off[sq] - k + 1;
                ws = (int64_t)(u - unit_start[sq]) * LANE_W;
                const int64_t left = nw - ws;
                nwl = left <= 0 ? 0 : (left < LANE_W ? (int32_t)left : LANE_W);
            }
            sbase[lane] = slot_entry(u);
            if ((uint32_t)lane < halo) sbase[WAVE + lane] = slot_entry(c * WAVE + WAVE + (uint64_t)lane);
        }
        wave_sync();
        {
            // Lane l stages code dword l % 16 (eight bases from the slot's first
            // base plus 8 * (l % 16)) of slots l / 16, l / 16 + 4, ...  Only the
            // last chunks of a set have marked slots; every other chunk takes
            // the fast path: the WAVE main slots in four rounds of four, each
            // round's loads first so four are in flight, then the halo slots.
            // Pointer arithmetic (not integer casts) keeps the loads global, so
            // they do not wait on the LDS writes.
            const uint32_t qi = (uint32_t)lane >> 4, di = (uint32_t)lane & 15u;
            const uint8_t *pl = S.bases + 8u * di;
            const bool mk = (sbase[lane] >> 63) || ((uint32_t)lane < halo && (sbase[WAVE + lane] >> 63));
            if (__ballot(mk) == 0) {
                auto code8 = [&](const uint32_t (&d)[3], uint32_t sh) {
                    return stage_codes8(__builtin_amdgcn_alignbit(d[1], d[0], sh),
                                        __builtin_amdgcn_alignbit(d[2], d[1], sh));
                };
#pragma unroll 1
                for (uint32_t q0 = 0; q0 < (uint32_t)WAVE; q0 += 16) {
                    uint32_t d[4][3], sh[4];
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        const uint64_t bq = sbase[q0 + 4 * i + qi];
                        const uint32_t *a32 = (const uint32_t *)(pl + (bq & ~3ull));
                        d[i][0] = a32[0];
                        d[i][1] = a32[1];
                        d[i][2] = a32[2];
                        sh[i] = ((uint32_t)bq & 3u) * 8u;  // alignbit by 0 returns the low dword
                    }
#pragma unroll
                    for (int i = 0; i < 4; i++) b32[(q0 + 4 * i + qi) * 17u + di] = code8(d[i], sh[i]);
                }
                for (uint32_t q = WAVE + qi; q < nslot; q += 4) {
                    const uint64_t bq = sbase[q];
                    const uint32_t *a32 = (const uint32_t *)(pl + (bq & ~3ull));
                    const uint32_t d[3] = {a32[0], a32[1], a32[2]};
                    b32[q * 17u + di] = code8(d, ((uint32_t)bq & 3u) * 8u);
                }
            } else {
                // byte by byte; bytes past the set read as 0 (no window uses
                // them), and a slot past the last unit is never read (a unit's
                // halo slots belong to its own sequence)
                const uintptr_t lim = (uintptr_t)(S.bases + S.off[S.n]);
                for (uint32_t q = qi; q < nslot; q += 4) {
                    const uint64_t bq = sbase[q];
                    if (bq == ~0ull) continue;
                    const uint8_t *p = pl + (bq & ~(1ull << 63));
                    uint32_t v0 = 0, v1 = 0;
                    for (uint32_t j = 0; j < 4; j++) {
                        if ((uintptr_t)(p + j) < lim) v0 |= (uint32_t)p[j] << (8 * j);
                        if ((uintptr_t)(p + 4 + j) < lim) v1 |= (uint32_t)p[4 + j] << (8 * j);
                    }
                    b32[q * 17u + di] = stage_codes8(v0, v1);
                }
            }
        }
        wave_sync();
        {
            const int64_t lw0 = (int64_t)lane * LANE_W;  // the lane's slot, in LDS base positions
            const int64_t lw1 = lw0 + nwl;
            uint32_t cnt = 0;                        // PASS 1: k-mers written so far
            uint32_t so = (uint32_t)lane * 16u;      // PASS 0: byte offset of the next staging slot
            const uint32_t so_lim = lcap * (WAVE * 16u);
            uint64_t base_out = 0;
            if (PASS == 1) {
                const uint32_t mine = lane_cnt[c * WAVE + lane];
                base_out = chunk_val[c] + (wave_incl_scan(mine) - mine);
            }
            if (lw0 < lw1) {
                // the lane's start: bases lw0 .. lw0+k-2 rolled in from nothing,
                // two at a time (pairs are the bytes of the code stream; the
                // lane start is dword aligned since LANE_W is a multiple of 8)
                uint32_t fl = 0, fh = 0, rl = 0, rh = 0;  // the strands as 32-bit halves
                const uint32_t q0 = (uint32_t)lw0 >> 3;
                const int32_t npair = (k - 1) / 2;
                for (int32_t p = 0; p < npair; p += 4) {
                    const uint32_t d = SKW(q0 + ((uint32_t)p >> 2));
                    const uint32_t e = ((d & 0x07070707u) << 3) | ((d >> 4) & 0x07070707u);
#pragma unroll
                    for (int t = 0; t < 4; t++) {
                        if (p + t < npair) {
                            const uint4 tt = *(const uint4 *)((const uint8_t *)T2 + byte_x16(e, t));
                            const uint32_t nfh = __builtin_amdgcn_alignbit(fh, fl, 30) ^ tt.y;  // rotl2
                            const uint32_t nfl = __builtin_amdgcn_alignbit(fl, fh, 30) ^ tt.x;
                            const uint32_t nrh = __builtin_amdgcn_alignbit(rl, rh, 2) ^ tt.w;   // rotr2
                            const uint32_t nrl = __builtin_amdgcn_alignbit(rh, rl, 2) ^ tt.z;
                            fh = nfh; fl = nfl; rh = nrh; rl = nrl;
                        }
                    }
                }
                if ((k - 1) & 1) {
                    const uint4 tt = ((const uint4 *)T)[32 + SKN(lw0 + k - 2)];  // row 4: nothing rolls out
                    const uint32_t nfh = __builtin_amdgcn_alignbit(fh, fl, 31) ^ tt.y;
                    const uint32_t nfl = __builtin_amdgcn_alignbit(fl, fh, 31) ^ tt.x;
                    const uint32_t nrh = __builtin_amdgcn_alignbit(rl, rh, 1) ^ tt.w;
                    const uint32_t nrl = __builtin_amdgcn_alignbit(rh, rl, 1) ^ tt.z;
                    fh = nfh; fl = nfl; rh = nrh; rl = nrl;
                }
                // A kept k-mer goes to the lane's next staging slot (PASS 0):
                // one 16-byte record (hash, position, strand) at chunk c's
                // uniform base plus a running 32-bit byte offset (slots of one
                // lane are WAVE * 16 bytes apart), so one add and one store;
                // the offset keeps counting past lcap slots without storing.
                uint4 *st_c = stage + (PASS == 0 ? (uint64_t)c * lcap * WAVE : 0);
                // window lw0 + j of the slot is window ws + j of the sequence,
                // which ends at base ws + j + k - 1
                const int32_t r0 = (int32_t)(ws - lw0 + k - 1);
                auto put = [&](int64_t wi, uint64_t h, bool strand) {
                    if (PASS == 1) {
                        const uint64_t o = base_out + cnt;
                        out_h[o] = h;
                        out_r[o] = r0 + (int32_t)wi;
                        out_s[o] = strand;
                        cnt++;
                    } else {
                        if (so < so_lim)
                            *(uint4 *)((uint8_t *)st_c + so) =
                                make_uint4((uint32_t)h, (uint32_t)(h >> 32), (uint32_t)(r0 + (int32_t)wi), strand ? 1u : 0u);
                        so += WAVE * 16u;
                    }
                };
                // windows lw0 .. lw1-1; the step into window lw0+j rolls out
                // base lw0+j-1 (nothing for j = 0: code 4) and in base lw0+j+k-1.
                // Eight steps at a time: both code runs are funnel shifts of two
                // dwords (lw0 is a multiple of LANE_W, so a lane's outgoing
                // dwords share one skew); even and odd nibbles spread to bytes
                // give the eight (out << 3 | in) table indices.  A full lane is
                // sixteen such groups.  The strands roll as 32-bit halves.
                // Kept windows are rare, so the steps only test the high word of
                // the hash against the threshold's, and the exact test and the
                // writes come after.
                const int32_t nwin_l = (int32_t)(lw1 - lw0);
                const uint32_t ish = (uint32_t)((k - 1) & 7) * 4;
                // The lane's first code dword q0 = lw0 / 8 is a multiple of 16,
                // so SKW(q0 + m) = lb[m + m / 16] with lb = &SKW(q0): a per-lane
                // base plus an offset that is the same for every lane.
                static_assert(LANE_W % 128 == 0, "a lane's code dwords must start on a skew step");
                const uint32_t *lb = &SKW((uint32_t)lw0 >> 3);
                const uint4 *T4 = (const uint4 *)T;  // (fw lo, fw hi, rc lo, rc hi)
                int32_t j = 0;
                for (; j + 8 <= nwin_l; j += 8) {
                    // outgoing dwords: m = j / 8 < LANE_W / 8 <= 16, no skew step
                    const uint32_t oc = lb[j >> 3], op = j ? lb[(j >> 3) - 1] : 0x44444444u;
                    const uint32_t od = __builtin_amdgcn_alignbit(oc, op, 28);  // bases lw0+j-1 .. lw0+j+6
                    const uint32_t mi = (uint32_t)(j + k - 1) >> 3;
                    const uint32_t i0 = lb[mi + (mi >> 4)], i1 = lb[mi + 1 + ((mi + 1) >> 4)];
                    const uint32_t id = ish ? __builtin_amdgcn_alignbit(i1, i0, ish) : i0;
                    const uint32_t pe = ((od & 0x0f0f0f0fu) << 3) | (id & 0x0f0f0f0fu);            // windows 0, 2, 4, 6
                    const uint32_t po = (((od >> 4) & 0x0f0f0f0fu) << 3) | ((id >> 4) & 0x0f0f0f0fu); // windows 1, 3, 5, 7
#pragma unroll
                    for (int half = 0; half < 2; half++) {
                        uint32_t sfl[4], sfh[4], srl[4], srh[4];
                        uint32_t cand = 0;
#pragma unroll
                        for (int t4 = 0; t4 < 4; t4++) {
                            const int t = half * 4 + t4;
                            const uint32_t toff = byte_x16((t & 1) ? po : pe, t >> 1);
                            const uint4 tt = *(const uint4 *)((const uint8_t *)T4 + toff);
                            const uint32_t nfh = __builtin_amdgcn_alignbit(fh, fl, 31) ^ tt.y;  // rotl1
                            const uint32_t nfl = __builtin_amdgcn_alignbit(fl, fh, 31) ^ tt.x;
                            const uint32_t nrh = __builtin_amdgcn_alignbit(rl, rh, 1) ^ tt.w;   // rotr1
                            const uint32_t nrl = __builtin_amdgcn_alignbit(rh, rl, 1) ^ tt.z;
                            fh = nfh; fl = nfl; rh = nrh; rl = nrl;
                            sfl[t4] = fl; sfh[t4] = fh; srl[t4] = rl; srh[t4] = rh;
                            cand |= (uint32_t)((fh ^ rh) <= thres_hi) << t4;
                        }
                        if (cand) {
#pragma unroll
                            for (int t4 = 0; t4 < 4; t4++)
                                if (cand & (1u << t4)) {
                                    const uint64_t f = ((uint64_t)sfh[t4] << 32) | sfl[t4];
                                    const uint64_t r = ((uint64_t)srh[t4] << 32) | srl[t4];
                                    const uint64_t h = f ^ r;
                                    if (h <= thres) put(lw0 + j + half * 4 + t4, h, f > r);
                                }
                        }
                    }
                }
                // a partial lane's last windows (the read's end)
                uint64_t hf = ((uint64_t)fh << 32) | fl, hr = ((uint64_t)rh << 32) | rl;
                for (; j < nwin_l; j++) {
                    const uint32_t oc = j ? SKN(lw0 + j - 1) : 4u;
                    const ulonglong2 tt = T[(oc << 3) | SKN(lw0 + j + k - 1)];
                    hf = rotl1(hf) ^ tt.x;
                    hr = rotr1(hr) ^ tt.y;
                    const uint64_t h = hr ^ hf;
                    if (h <= thres) put(lw0 + j, h, hf > hr);
                }
            }
            if (PASS == 0) {
                cnt = (so - (uint32_t)lane * 16u) / (WAVE * 16u);
                lane_cnt[c * WAVE + lane] = cnt;
                lane_pre[c * WAVE + lane] = wave_incl_scan(cnt) - cnt;
                uint32_t s2 = cnt;
                for (int d = WAVE / 2; d > 0; d >>= 1) s2 += __shfl_xor(s2, d);
                if (lane == 0) chunk_val[c] = s2;
                if (__ballot(cnt > lcap) != 0 && lane == 0) ovf_list[atomicAdd(ovf_count, 1u)] = (uint32_t)c;
            }
        }
        wave_sync();  // every lane is done with the slice before it is restaged
    }
#undef SKW
#undef SKN
}

// Moves the staged k-mers of every chunk without a lane overflow to their
// final positions (the chunk's scanned offset plus the lane's exclusive
// prefix); the listed overflow chunks are written by k_sketch<1>.  A chunk of
// at most COMPACT_TILE k-mers goes row by row: row i holds each lane's i-th
// record side by side, so the wave reads the rows whole (lanes past their
// count stay idle) and places every record at its output index in LDS, then
// writes the chunk's k-mers out contiguously.  A larger chunk (sketches at
// high ratios) gathers each output element from its slot instead, the lane
// found by a search of the lane prefix.
constexpr uint32_t COMPACT_TILE = 512;
__global__ void __launch_bounds__(WG) k_sketch_compact(uint64_t n_chunks, const uint32_t *lane_cnt,
                                                       const uint64_t *chunk_koff, uint32_t lcap,
                                                       const uint4 *stage, uint64_t *out_h, int32_t *out_r,
                                                       uint8_t *out_s, const uint32_t *abort,
                                                       const uint64_t *n_units_dev) {
    if (aborted(abort)) return;
    if (n_units_dev) n_chunks = (*n_units_dev + WAVE - 1) / WAVE;
    __shared__ uint32_t s_pre[NWAVES][WAVE];
    __shared__ uint64_t s_h[NWAVES][COMPACT_TILE];
    __shared__ int32_t s_r[NWAVES][COMPACT_TILE];
    __shared__ uint8_t s_s[NWAVES][COMPACT_TILE];
    const int lane = threadIdx.x & (WAVE - 1), w = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
    uint32_t *pre_w = s_pre[w];
    for (uint64_t c = (uint64_t)blockIdx.x * NWAVES + w; c < n_chunks; c += (uint64_t)gridDim.x * NWAVES) {
        const uint32_t cnt = lane_cnt[c * WAVE + lane];
        if (__ballot(cnt > lcap) != 0) continue;
        const uint32_t incl = wave_incl_scan(cnt);
        const uint32_t total = __shfl(incl, WAVE - 1);
        const uint32_t pre = incl - cnt;
        const uint64_t ob = chunk_koff[c];
        const uint64_t sb = (uint64_t)c * lcap * WAVE;
        if (total <= COMPACT_TILE) {
            uint32_t rows = cnt;
            for (int d = WAVE / 2; d > 0; d >>= 1) {
                const uint32_t o = __shfl_xor(rows, d);
                rows = o > rows ? o : rows;
            }
            for (uint32_t i0 = 0; i0 < rows; i0 += 4) {
                uint4 rec[4];
#pragma unroll
                for (uint32_t j = 0; j < 4; j++)
                    if (i0 + j < cnt) rec[j] = stage[sb + (uint64_t)(i0 + j) * WAVE + lane];
#pragma unroll
                for (uint32_t j = 0; j < 4; j++)
                    if (i0 + j < cnt) {
                        const uint32_t p = pre + i0 + j;  // < pre + cnt <= total
                        s_h[w][p] = ((uint64_t)rec[j].y << 32) | rec[j].x;
                        s_r[w][p] = (int32_t)rec[j].z;
                        s_s[w][p] = (uint8_t)rec[j].w;
                    }
            }
            wave_sync();
            for (uint32_t t = lane; t < total; t += WAVE) {
                out_h[ob + t] = s_h[w][t];
                out_r[ob + t] = s_r[w][t];
                out_s[ob + t] = s_s[w][t];
            }
        } else {
            pre_w[lane] = pre;
            wave_sync();
            for (uint32_t t = lane; t < total; t += WAVE) {
                uint32_t lo = 0, hi = WAVE;  // the last lane whose prefix is <= t
                while (hi - lo > 1) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (pre_w[mid] <= t) lo = mid; else hi = mid;
                }
                const uint64_t si = sb + (uint64_t)(t - pre_w[lo]) * WAVE + lo;
                const uint4 rec = stage[si];
                out_h[ob + t] = ((uint64_t)rec.y << 32) | rec.x;
                out_r[ob + t] = (int32_t)rec.z;
                out_s[ob + t] = (uint8_t)rec.w;
            }
        }
        wave_sync();  // the wave's LDS slices are rewritten by its next chunk
    }
}

// Per-sequence k-mer offsets: sequence i's first unit u = unit_start[i] is lane
// u % WAVE of chunk u / WAVE, so koff[i] = chunk_koff[u / WAVE] plus that lane's
// exclusive prefix of kept counts in its chunk (lane_pre; none for lane 0, which
// also covers u = n_units at a chunk boundary).  The largest per-sequence
// sketch goes into *max_m (zeroed by the caller).
__device__ __forceinline__ uint64_t seq_koff_at(uint64_t u, const uint64_t *chunk_koff, const uint32_t *lane_pre) {
    return chunk_koff[u / WAVE] + ((u & (WAVE - 1)) ? lane_pre[u] : 0u);
}

__global__ void __launch_bounds__(WG) k_seq_koff(int32_t n, const uint64_t *unit_start,
                                                 const uint64_t *chunk_koff, const uint32_t *lane_pre,
                                                 uint64_t *koff, unsigned long long *max_m,
                                                 const uint32_t *abort) {
    if (aborted(abort)) return;
    for (int32_t i0 = blockIdx.x * WG; i0 <= n; i0 += gridDim.x * WG) {
        const int32_t i = i0 + (int32_t)threadIdx.x;
        uint64_t m = 0;
        if (i <= n) {
            const uint64_t a = seq_koff_at(unit_start[i], chunk_koff, lane_pre);
            koff[i] = a;
            if (i < n) m = seq_koff_at(unit_start[i + 1], chunk_koff, lane_pre) - a;
        }
        for (int d = WAVE / 2; d > 0; d >>= 1) {
            const uint64_t o = __shfl_xor(m, d);
            m = o > m ? o : m;
        }
        if ((threadIdx.x & (WAVE - 1)) == 0 && m) atomicMax(max_m, (unsigned long long)m);
    }
}

// Largest value of a[0, n) into *out (zeroed by the caller; values >= 0).
__global__ void __launch_bounds__(WG) k_max_i32(const int32_t *a, int32_t n, unsigned long long *out) {
    for (int32_t i0 = blockIdx.x * WG; i0 < n; i0 += gridDim.x * WG) {
        const int32_t i = i0 + (int32_t)threadIdx.x;
        int32_t v = i < n ? a[i] : 0;
        for (int d = WAVE / 2; d > 0; d >>= 1) {
            const int32_t o = __shfl_xor(v, d);
            v = o > v ? o : v;
        }
        if ((threadIdx.x & (WAVE - 1)) == 0 && v > 0) atomicMax(out, (unsigned long long)v);
    }
}

// ============================================================================
// Index build (src/index.rs:233-376)
// ============================================================================

__global__ void __launch_bounds__(WG) k_fill_pairs(const uint64_t *t_h, uint64_t n, uint64_t *keys,
                                                   uint32_t *vals) {
    for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < n; i += (uint64_t)gridDim.x * WG) {
        keys[i] = t_h[i];
        vals[i] = (uint32_t)i;
    }
}

// Stable LSD radix sort pass over 8-bit digits: per-tile digit histograms ...
__global__ void __launch_bounds__(WG) k_rs_hist(const uint64_t *keys, uint64_t n, uint32_t shift,
                                                uint64_t *hist, uint32_t ntiles) {
    __shared__ uint32_t cnt[256];
    for (uint32_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        cnt[threadIdx.x] = 0;
        __syncthreads();
        const uint64_t t0 = (uint64_t)tile * RS_TILE;
        for (int it = 0; it < RS_ITEMS; it++) {
            const uint64_t e = t0 + (uint64_t)it * WG + threadIdx.x;
            if (e < n) atomicAdd(&cnt[(keys[e] >> shift) & 0xff], 1u);
        }
        __syncthreads();
        hist[(uint64_t)threadIdx.x * ntiles + tile] = cnt[threadIdx.x];
        __syncthreads();
    }
}

// ... then a stable scatter: within a tile, keys are visited in index order in
// rounds of WG, ranked inside each wave by ballot matching on the digit, and
// across waves through LDS counters.
__global__ void __launch_bounds__(WG) k_rs_scatter(const uint64_t *kin, const uint32_t *vin,
                                                   uint64_t *kout, uint32_t *vout, uint64_t n,
                                                   uint32_t shift, const uint64_t *offs,
                                                   uint32_t ntiles) {
    __shared__ uint32_t run[256];
    __shared__ uint32_t wcnt[NWAVES][256];
    __shared__ uint64_t gbase[256];
    const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
    const uint64_t lt = (1ull << lane) - 1ull;
    for (uint32_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        run[threadIdx.x] = 0;
        for (int i = 0; i < NWAVES; i++) wcnt[i][threadIdx.x] = 0;
        gbase[threadIdx.x] = offs[(uint64_t)threadIdx.x * ntiles + tile];
        __syncthreads();
        const uint64_t t0 = (uint64_t)tile * RS_TILE;
        for (int it = 0; it < RS_ITEMS; it++) {
            const uint64_t e = t0 + (uint64_t)it * WG + threadIdx.x;
            const bool valid = e < n;
            uint64_t key = 0;
            uint32_t val = 0;
            uint32_t d = 256;
            if (valid) {
                key = kin[e];
                val = vin[e];
                d = (uint32_t)((key >> shift) & 0xff);
            }
            uint64_t peers = ~0ull;
#pragma unroll
            for (int b = 0; b < 9; b++) {
                const uint64_t bal = __ballot((d >> b) & 1u);
                peers &= ((d >> b) & 1u) ? bal : ~bal;
            }
            const uint32_t rank = __popcll(peers & lt);
            const uint32_t cntw = __popcll(peers);
            if (valid && rank == 0) wcnt[w][d] = cntw;
            __syncthreads();
            if (valid) {
                uint32_t pos = run[d] + rank;
                for (int i = 0; i < w; i++) pos += wcnt[i][d];
                const uint64_t o = gbase[d] + pos;
                kout[o] = key;
                vout[o] = val;
            }
            __syncthreads();
            if (valid && rank == 0) {
                atomicAdd(&run[d], cntw);
                wcnt[w][d] = 0;
            }
            __syncthreads();
        }
    }
}

__global__ void __launch_bounds__(WG) k_mark_heads(const uint64_t *keys, uint64_t n, uint64_t *flag) {
    for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < n; i += (uint64_t)gridDim.x * WG)
        flag[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1 : 0;
}

__global__ void __launch_bounds__(WG) k_emit_keys(const uint64_t *keys, uint64_t n, const uint64_t *flag,
                                                  const uint64_t *kid, uint64_t *uk_h, uint64_t *uk_off) {
    for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < n; i += (uint64_t)gridDim.x * WG) {
        if (flag[i]) {
            uk_h[kid[i]] = keys[i];
            uk_off[kid[i]] = i;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) uk_off[kid[n]] = n;
}

__device__ __forceinline__ uint32_t seg_of(const int64_t *kstart, int32_t nseg, uint64_t g) {
    int32_t lo = 0, hi = nseg - 1;
    while (lo < hi) {
        const int32_t mid = (lo + hi + 1) >> 1;
        if ((uint64_t)kstart[mid] <= g) lo = mid; else hi = mid - 1;
    }
    return (uint32_t)lo;
}

__global__ void __launch_bounds__(WG) k_fill_hits(const uint32_t *vals, uint64_t n, const int32_t *t_r,
                                                  const uint8_t *t_s, const int64_t *kstart,
                                                  int32_t nseg, uint32_t *hit_g, uint32_t *hit_r,
                                                  uint32_t *hit_seg, uint64_t *sbits) {
    for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < n; i += (uint64_t)gridDim.x * WG) {
        const uint32_t g = vals[i];
        hit_g[i] = g;
        hit_r[i] = (uint32_t)t_r[g] | ((uint32_t)t_s[g] << 31);
        hit_seg[i] = seg_of(kstart, nseg, g);
        sbits[i] = t_s[g];
    }
}

__global__ void __launch_bounds__(WG) k_narrow_prefix(const uint64_t *in, uint64_t n, uint32_t *out) {
    for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i <= n; i += (uint64_t)gridDim.x * WG)
        out[i] = (uint32_t)in[i];
}

// The -M closed form (src/index.rs:264-291, :295-315, :360-376): a hash whose
// total count exceeds max_matches is erased; a single-hit hash never is.
__device__ __forceinline__ bool key_kept(uint64_t cnt, int32_t max_matches) {
    return !(max_matches > 0 && cnt > 1 && cnt > (uint64_t)max_matches);
}

__global__ void __launch_bounds__(WG) k_shard_count(const uint64_t *uk_h, const uint64_t *uk_off,
                                                    uint64_t nk, int32_t max_matches,
                                                    unsigned long long *shard_n, unsigned long long *stats) {
    __shared__ unsigned long long sc[8 + 4];
    if (threadIdx.x < 12) sc[threadIdx.x] = 0;
    __syncthreads();
    for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < nk; i += (uint64_t)gridDim.x * WG) {
        const uint64_t cnt = uk_off[i + 1] - uk_off[i];
        if (key_kept(cnt, max_matches)) {
            atomicAdd(&sc[uk_h[i] & 7], 1ull);
        } else {
            atomicAdd(&sc[8], 1ull);
            atomicAdd(&sc[9], (unsigned long long)cnt);
        }
        if (cnt > 1) atomicMax(&sc[10], (unsigned long long)cnt);
    }
    __syncthreads();
    if (threadIdx.x < 8) atomicAdd(&shard_n[threadIdx.x], sc[threadIdx.x]);
    if (threadIdx.x == 8) atomicAdd(&stats[0], sc[8]);
    if (threadIdx.x == 9) atomicAdd(&stats[1], sc[9]);
    if (threadIdx.x == 10) atomicMax(&stats[2], sc[10]);
}

__global__ void __launch_bounds__(WG) k_insert(const uint64_t *uk_h, const uint64_t *uk_off, uint64_t nk,
                                               int32_t max_matches, uint4 *table, ShardGeom geom) {
    for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < nk; i += (uint64_t)gridDim.x * WG) {
        const uint64_t cnt = uk_off[i + 1] - uk_off[i];
        if (!key_kept(cnt, max_matches)) continue;
        const uint64_t h = uk_h[i];
        const uint32_t sh = (uint32_t)(h & 7u);
        const uint32_t bits = geom.bits[sh];
        const uint32_t mask = (1u << bits) - 1u;
        uint4 *T = table + geom.base[sh];
        uint32_t pos = table_pos(h, bits);
        for (;;) {
            const uint32_t old = atomicCAS(&T[pos].w, 0u, (uint32_t)cnt);
            if (old == 0u) {
                T[pos].x = (uint32_t)h;
                T[pos].y = (uint32_t)(h >> 32);
                T[pos].z = (uint32_t)uk_off[i];
                break;
            }
            pos = (pos + 1u) & mask;
        }
    }
}

// ============================================================================
// Per-read mapping stages (src/shmap/mod.rs:653-971).  Seeding kernels cover
// the whole batch; bucketing kernels cover one group of reads [r0, r0+RG)
// whose tables fit the device budget.
// ============================================================================

// A read's k-mers (m <= E * WAVE) in (hash, position) order, as hashes in kh
// and k-mer indices in kv (the wave's LDS slice).  The wave sorts one-word keys
// (hash >> A.seed_sh) << 9 | index: the shifted hash keeps the hashes' order
// and the index breaks ties, so two keys are out of order only if their
// distinct hashes agree in the kept bits.  Such neighbours are looked for
// after the sort, and if there are any, one lane puts the array in order by
// insertion (few and short runs are out of place, so this is close to one
// pass).
template <int E>
__device__ __forceinline__ void seed_sort_e(const MapArgs &A, uint64_t o, uint32_t m, uint64_t *kh, uint32_t *kv,
                                            uint32_t lane) {
    const uint32_t ssh = A.seed_sh;
    uint64_t k[E];
    uint32_t v[E];  // (unused: the keys carry the index)
#pragma unroll
    for (int e = 0; e < E; e++) {
        const uint32_t x = lane * E + (uint32_t)e;
        k[e] = ~0ull;
        v[e] = 0;
        if (x < m) {
            const uint64_t h = A.k_h[o + x];
            kh[x] = h;
            k[e] = ((h >> ssh) << 9) | x;
        }
    }
    wave_sync();
    wave_sort_regs<E, false>(k, v);
    uint64_t hv[E];
#pragma unroll
    for (int e = 0; e < E; e++) {
        const uint32_t x = lane * E + (uint32_t)e;
        hv[e] = x < m ? kh[(uint32_t)k[e] & 511u] : 0ull;
    }
    wave_sync();  // every hash read before any is overwritten
#pragma unroll
    for (int e = 0; e < E; e++) {
        const uint32_t x = lane * E + (uint32_t)e;
        if (x < m) {
            kh[x] = hv[e];
            kv[x] = (uint32_t)k[e] & 511u;
        }
    }
    wave_sync();
    bool mixed = false;
    for (uint32_t x = lane + 1; x < m; x += WAVE)
        mixed |= kh[x] != kh[x - 1] && (kh[x] >> ssh) == (kh[x - 1] >> ssh);
    if (__ballot(mixed) == 0) return;
    if (lane == 0) {
        for (uint32_t i = 1; i < m; i++) {
            const uint64_t h = kh[i];
            const uint32_t p = kv[i];
            uint32_t j = i;
            while (j > 0 && (kh[j - 1] > h || (kh[j - 1] == h && kv[j - 1] > p))) {
                kh[j] = kh[j - 1];
                kv[j] = kv[j - 1];
                j--;
            }
            kh[j] = h;
            kv[j] = p;
        }
    }
    wave_sync();
}

// The seeding stages for one read per wave (reads with at most WAVE_SEED_CAP
// sketched k-mers; larger reads are listed for k_seed_wg):
//   unique_elements_with_info (src/shmap/seeding.rs:26-77): sort by (hash,
//     position), one seed per hash run keeping the smallest-r strand, index
//     probe, stable rarest-first order as a sort by (hits_in_t, seed_num);
//   the seed budget S, halflen and the consumed prefix of match_seeds
//     (src/shmap/mod.rs:726-742, src/shmap/seeding.rs:101-106, 207-209).
// Everything stays in the wave's LDS slice; seeds are written out once.
// (five waves per SIMD: 96 VGPRs without spills since the one-word sort keys;
// its waves mostly wait, so the fifth pays: 1.0 -> 0.97 ms on genome)
__global__ void __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(5))) k_seed_wave(MapArgs A, DevIndex I) {
    if (aborted(A.abort)) return;
    extern __shared__ uint8_t lds[];
    const uint32_t lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
    // Per-wave LDS, 14 B per k-mer slot:
    //   kh  u64: hashes for the (hash, position) sort and the run scan; then the
    //            (hits << 32 | seed id) keys of the rarest-first sort
    //   kv  u32: positions; then, at index seed id, each seed's hit count
    //   so  u16: occurrences of each seed id
    uint8_t *my = lds + (size_t)w * WAVE_SEED_LDS;
    uint64_t *kh = (uint64_t *)my;
    uint32_t *kv = (uint32_t *)(my + WAVE_SEED_CAP * 8);
    uint16_t *so = (uint16_t *)(my + WAVE_SEED_CAP * 12);
    const uint64_t lt = (1ull << lane) - 1ull;
    const double theta2 = A.theta - A.min_diff;
    for (int32_t rd = blockIdx.x * NWAVES + w; rd < A.R; rd += gridDim.x * NWAVES) {
        const uint64_t o = A.koff[rd];
        const uint32_t m = (uint32_t)(A.koff[rd + 1] - o);
        if (m > WAVE_SEED_CAP) {  // a workgroup's job (k_seed_wg)
            if (lane == 0) A.big_list[atomicAdd(A.big_count, 1u)] = (uint32_t)rd;
            continue;
        }
        // (hash, position) order, in registers (m <= WAVE_SEED_CAP = 8 * WAVE)
        if (m <= WAVE) seed_sort_e<1>(A, o, m, kh, kv, lane);
        else if (m <= 2 * WAVE) seed_sort_e<2>(A, o, m, kh, kv, lane);
        else if (m <= 4 * WAVE) seed_sort_e<4>(A, o, m, kh, kv, lane);
        else seed_sort_e<8>(A, o, m, kh, kv, lane);
        wave_sync();
        // runs of equal hash -> seeds in hash order.  A seed id never exceeds
        // the position of its run head (and equals it only while every earlier
        // element heads its own run, when kh[sid] already holds the hash), so
        // kh[sid] can take the seed's hash once this chunk has read its own
        // hashes: later chunks read only kh from the chunk's last element on.
        uint32_t u = 0;
        for (uint32_t b0 = 0; b0 < m; b0 += WAVE) {
            const uint32_t i = b0 + lane;
            const uint32_t pos = i < m ? kv[i] : 0u;
            const bool head = i < m && (i == 0 || kh[i] != kh[i - 1]);
            const uint64_t bal = __ballot(head);
            if (i < m) A.ps_v[o + i] = pos;
            uint32_t sid = 0, e = 0;
            uint64_t h = 0;
            if (head) {
                sid = u + __popcll(bal & lt);
                h = kh[i];
                e = i + 1;
                while (e < m && kh[e] == h) e++;
                A.sd_h[o + sid] = h;
                A.sd_first[o + sid] = i;
                A.sd_occs[o + sid] = e - i;
                A.sd_strand[o + sid] = A.k_s[o + pos];
            }
            wave_sync();  // every lane has read kv[b0, b0+64) and kh
            if (head) {
                kh[sid] = h;
                so[sid] = (uint16_t)(e - i);
            }
            u += __popcll(bal);
        }
        wave_sync();
        // index probes of every seed, four per lane in flight together; kv[sid]
        // takes the seed's hit count (kv is no longer read as positions)
        uint32_t pm = 0;
        for (uint32_t q0 = 0; q0 < u; q0 += 4 * WAVE) {
            uint64_t hq[4];
            bool act[4];
            uint32_t cnt[4], hoff[4];
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const uint32_t q = q0 + (uint32_t)t * WAVE + lane;
                act[t] = q < u;
                hq[t] = act[t] ? kh[q] : 0ull;
            }
            index_lookup4(I, hq, act, cnt, hoff);
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const uint32_t q = q0 + (uint32_t)t * WAVE + lane;
                if (act[t]) {
                    A.sd_cnt[o + q] = cnt[t];
                    A.sd_hoff[o + q] = hoff[t];
                    kv[q] = cnt[t];
                    pm += cnt[t];
                }
            }
        }
        for (int d = WAVE / 2; d > 0; d >>= 1) pm += __shfl_xor(pm, d);
        wave_sync();
        // rarest-first order: keys (hits << 32 | seed id), distinct.  Seeds
        // without hits lead, in seed-id order, so they are placed by ballot
        // and only the others are sorted (after them, in kh[zs, u)).
        uint32_t zs = 0;
        for (uint32_t b0 = 0; b0 < u; b0 += WAVE) zs += __popcll(__ballot(b0 + lane < u && kv[b0 + lane] == 0));
        {
            uint32_t zi = 0, ni = zs;
            for (uint32_t b0 = 0; b0 < u; b0 += WAVE) {
                const uint32_t q = b0 + lane;
                const uint32_t cq = q < u ? kv[q] : 0u;
                const uint64_t bz = __ballot(q < u && cq == 0), bn = __ballot(q < u && cq != 0);
                if (q < u) {
                    const uint32_t at = cq == 0 ? zi + __popcll(bz & lt) : ni + __popcll(bn & lt);
                    kh[at] = ((uint64_t)cq << 32) | q;
                }
                zi += __popcll(bz);
                ni += __popcll(bn);
            }
        }
        wave_sync();
        if (u - zs > 1 &&
            !wave_sort_n<false>(
                u - zs, [&](uint32_t x, uint64_t &key, uint32_t &) { key = kh[zs + x]; },
                [&](uint32_t x, uint64_t key, uint32_t) { kh[zs + x] = key; })) {
            wave_bitonic_keys(kh + zs, u - zs);
        }
        wave_sync();
        // budget: consume seeds while fewer than S occurrences are consumed
        // (the rows read back the hit offsets and strands this wave's lanes
        // stored above)
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        const int32_t len = (int32_t)(A.roff[rd + 1] - A.roff[rd]);
        const int32_t S = A.one_sweep ? (int32_t)u : (int32_t)((1.0 - theta2) * (double)m) + 1;
        const int32_t hl = A.abs_pos ? len : (int32_t)m;
        const bool ok = hl >= MIN_HALFLEN;
        uint32_t occ_carry = 0, cnt_carry = 0, c = 0, z = 0, seeds = 0, mx = 0;
        for (uint32_t b0 = 0; ok && b0 < u; b0 += WAVE) {
            const uint32_t q = b0 + lane;
            uint32_t occ = 0, cnt = 0, sid = 0;
            if (q < u) {
                const uint64_t key = kh[q];
                sid = (uint32_t)key;
                occ = so[sid];
                cnt = (uint32_t)(key >> 32);
            }
            const uint32_t oinc = wave_incl_scan(occ);
            const bool consumed = q < u && (int64_t)(occ_carry + oinc - occ) < (int64_t)S;
            const uint32_t cv = consumed ? cnt : 0u;
            const uint32_t cinc = wave_incl_scan(cv);
            const uint32_t epre = cnt_carry + cinc - cv;
            if (q < u) {
                // the seed's row (a seed without hits needs no offset or strand)
                uint32_t hoff = 0, str = 0;
                if (cnt) {
                    hoff = A.sd_hoff[o + sid];
                    str = A.sd_strand[o + sid];
                }
                A.sd_row[o + q] = make_uint4(hoff, cnt, occ | (str << 31), consumed ? epre : 0u);
            }
            const uint64_t cb = __ballot(consumed);
            c += __popcll(cb);
            z += __popcll(__ballot(consumed && cnt == 0));  // hitless seeds sort first
            uint32_t so_ = consumed ? occ : 0u, mm = consumed ? cnt : 0u;
            for (int d = WAVE / 2; d > 0; d >>= 1) {
                so_ += __shfl_xor(so_, d);
                const uint32_t om = __shfl_xor(mm, d);
                mm = om > mm ? om : mm;
            }
            seeds += so_;
            mx = mm > mx ? mm : mx;
            occ_carry += __shfl(oinc, WAVE - 1);
            cnt_carry += __shfl(cinc, WAVE - 1);
        }
        if (lane == 0) {
            const uint32_t E = cnt_carry;
            A.rd_u[rd] = (int32_t)u;
            A.rd_possible[rd] = (int32_t)pm;  // an i32 sum (src/shmap/mod.rs:689)
            A.rd_S[rd] = S;
            A.rd_hl[rd] = hl;
            A.rd_ok[rd] = ok ? 1 : 0;
            A.rd_c[rd] = (int32_t)c;
            A.rd_z[rd] = (int32_t)z;
            A.rd_bseeds[rd] = (int32_t)seeds;
            A.rd_max_seed[rd] = (int32_t)mx;
            A.rd_seed_matches[rd] = (int32_t)E;
            A.rd_E[rd] = E;
            uint64_t cap = 0;
            if (E > 0) {
                cap = 64;
                while (cap < 2ull * E) cap <<= 1;
            }
            A.rd_cap[rd] = cap;
        }
        wave_sync();
    }
}

// The seeding stages for reads with more than WAVE_SEED_CAP sketched k-mers
// (listed by k_seed_wave), one workgroup per read, through global memory with
// the sorts staged in LDS when they fit: the same steps as k_seed_wave --
// sort by (hash, position) (the order of unique_elements_with_info,
// src/shmap/seeding.rs:28-34, with every run in increasing r, so a run's first
// element is the smallest-r occurrence whose strand the seed keeps and
// pmatches is the run read backwards), one seed per run with the index probe
// (SketchIndex::count), the rarest-first order as a sort by (hits_in_t,
// seed_num) (src/shmap/seeding.rs:26-77), then the seed budget S, halflen and
// the consumed prefix of match_seeds (src/shmap/mod.rs:726-742,
// src/shmap/seeding.rs:101-106, 207-209).
__global__ void __launch_bounds__(WG) k_seed_wg(MapArgs A, DevIndex I) {
    if (aborted(A.abort)) return;
    extern __shared__ uint8_t lds[];
    uint64_t *sk = (uint64_t *)lds;
    uint32_t *sv = (uint32_t *)(lds + (size_t)SORT_CAP * 8);
    __shared__ uint32_t tmp[NWAVES + 1];
    __shared__ uint32_t s_u, s_pm, s_c, s_seeds, s_max, s_occ_carry, s_cnt_carry;
    const double theta2 = A.theta - A.min_diff;
    const uint32_t nbig = *A.big_count;
    for (uint32_t t = blockIdx.x; t < nbig; t += gridDim.x) {
        const int32_t rd = (int32_t)A.big_list[t];
        const uint64_t o = A.koff[rd];
        const uint32_t m = (uint32_t)(A.koff[rd + 1] - o);
        // 1. (hash, position) order
        if (threadIdx.x == 0) {
            s_u = 0;
            s_pm = 0;
            s_c = 0;
            s_seeds = 0;
            s_max = 0;
            s_occ_carry = 0;
            s_cnt_carry = 0;
        }
        __syncthreads();
        if (!wg_sort_n<true>(
                m, [&](uint32_t x, uint64_t &key, uint32_t &val) { key = A.k_h[o + x], val = x; },
                [&](uint32_t x, uint64_t key, uint32_t val) { A.ps_h[o + x] = key, A.ps_v[o + x] = val; }, sk, sv)) {
            for (uint32_t i = threadIdx.x; i < m; i += WG) {
                A.ps_h[o + i] = A.k_h[o + i];
                A.ps_v[o + i] = i;
            }
            __syncthreads();
            wg_sort_pairs(A.ps_h + o, A.ps_v + o, m, sk, sv, SORT_CAP);
        }
        // 2. runs -> seeds in hash order, one probe each
        for (uint32_t base = 0; base < m; base += WG) {
            const uint32_t i = base + threadIdx.x;
            const bool head = i < m && (i == 0 || A.ps_h[o + i] != A.ps_h[o + i - 1]);
            uint32_t tot;
            const uint32_t ex = wg_excl_scan(head ? 1u : 0u, tmp, &tot);
            const uint32_t u0 = s_u;
            if (head) A.sd_first[o + u0 + ex] = i;
            __syncthreads();
            if (threadIdx.x == 0) s_u = u0 + tot;
            __syncthreads();
        }
        const uint32_t u = s_u;
        uint32_t pm = 0;
        // one probe per seed, four per thread in flight together
        for (uint32_t s0 = 0; s0 < u; s0 += 4 * WG) {
            uint64_t hq[4];
            bool act[4];
            uint32_t cnt[4], hoff[4];
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const uint32_t sid = s0 + (uint32_t)t * WG + threadIdx.x;
                act[t] = sid < u;
                hq[t] = act[t] ? A.ps_h[o + A.sd_first[o + sid]] : 0ull;
            }
            index_lookup4(I, hq, act, cnt, hoff);
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const uint32_t sid = s0 + (uint32_t)t * WG + threadIdx.x;
                if (!act[t]) continue;
                const uint32_t first = A.sd_first[o + sid];
                const uint32_t next = sid + 1 < u ? A.sd_first[o + sid + 1] : m;
                A.sd_h[o + sid] = hq[t];
                A.sd_occs[o + sid] = next - first;
                A.sd_strand[o + sid] = A.k_s[o + A.ps_v[o + first]];
                A.sd_cnt[o + sid] = cnt[t];
                A.sd_hoff[o + sid] = hoff[t];
                pm += cnt[t];
            }
        }
        for (int d = WAVE / 2; d > 0; d >>= 1) pm += __shfl_xor(pm, d);
        if ((threadIdx.x & (WAVE - 1)) == 0) atomicAdd(&s_pm, pm);
        __syncthreads();
        // 3. rarest-first order: seeds without hits lead, in seed-id order,
        //    placed by a scan; only the others are sorted, after them
        uint32_t zs = 0;
        for (uint32_t base = 0; base < u; base += WG) {
            const uint32_t q = base + threadIdx.x;
            uint32_t tot;
            (void)wg_excl_scan(q < u && A.sd_cnt[o + q] == 0 ? 1u : 0u, tmp, &tot);
            zs += tot;
        }
        {
            uint32_t zi = 0, ni = zs;
            for (uint32_t base = 0; base < u; base += WG) {
                const uint32_t q = base + threadIdx.x;
                const uint32_t cq = q < u ? A.sd_cnt[o + q] : 1u;
                uint32_t tz, tn;
                const uint32_t ez = wg_excl_scan(q < u && cq == 0 ? 1u : 0u, tmp, &tz);
                const uint32_t en = wg_excl_scan(q < u && cq != 0 ? 1u : 0u, tmp, &tn);
                if (q < u) {
                    const uint32_t at = cq == 0 ? zi + ez : ni + en;
                    A.sd_key[o + at] = ((uint64_t)cq << 32) | q;
                    A.sd_ord[o + at] = q;
                }
                zi += tz;
                ni += tn;
            }
        }
        __syncthreads();
        if (u - zs > 1 &&
            !wg_sort_n<true>(
                u - zs, [&](uint32_t x, uint64_t &key, uint32_t &val) { key = A.sd_key[o + zs + x], val = A.sd_ord[o + zs + x]; },
                [&](uint32_t x, uint64_t key, uint32_t val) { A.sd_key[o + zs + x] = key, A.sd_ord[o + zs + x] = val; },
                sk, sv))
            wg_sort_pairs(A.sd_key + o + zs, A.sd_ord + o + zs, u - zs, sk, sv, SORT_CAP);
        // 4. budget: consume seeds while fewer than S occurrences are consumed
        const int32_t len = (int32_t)(A.roff[rd + 1] - A.roff[rd]);
        const int32_t S = A.one_sweep ? (int32_t)u : (int32_t)((1.0 - theta2) * (double)m) + 1;
        const int32_t hl = A.abs_pos ? len : (int32_t)m;
        const bool ok = hl >= MIN_HALFLEN;
        if (ok) {
            for (int32_t base = 0; base < (int32_t)u; base += WG) {
                const int32_t q = base + threadIdx.x;
                uint32_t occs = 0, cnt = 0, sid = 0;
                if (q < (int32_t)u) {
                    sid = A.sd_ord[o + q];
                    occs = A.sd_occs[o + sid];
                    cnt = A.sd_cnt[o + sid];
                }
                uint32_t tot;
                const uint32_t ex = wg_excl_scan(occs, tmp, &tot);
                const uint32_t pre = s_occ_carry + ex;
                const bool consumed = q < (int32_t)u && (int64_t)pre < (int64_t)S;
                uint32_t ctot;
                const uint32_t cex = wg_excl_scan(consumed ? cnt : 0u, tmp, &ctot);
                if (q < (int32_t)u) {
                    uint32_t hoff = 0, str = 0;
                    if (cnt) {
                        hoff = A.sd_hoff[o + sid];
                        str = A.sd_strand[o + sid];
                    }
                    A.sd_row[o + q] = make_uint4(hoff, cnt, occs | (str << 31), consumed ? s_cnt_carry + cex : 0u);
                }
                // the consumed seeds' count, occurrences and largest hit count:
                // one LDS atomic per wave
                const uint64_t cb = __ballot(consumed);
                uint32_t so_ = consumed ? occs : 0u, mm = consumed ? cnt : 0u;
                for (int d = WAVE / 2; d > 0; d >>= 1) {
                    so_ += __shfl_xor(so_, d);
                    const uint32_t om = __shfl_xor(mm, d);
                    mm = om > mm ? om : mm;
                }
                if ((threadIdx.x & (WAVE - 1)) == 0 && cb) {
                    atomicAdd(&s_c, (uint32_t)__popcll(cb));
                    atomicAdd(&s_seeds, so_);
                    atomicMax(&s_max, mm);
                }
                __syncthreads();
                if (threadIdx.x == 0) {
                    s_occ_carry += tot;
                    s_cnt_carry += ctot;
                }
                __syncthreads();
            }
        }
        if (threadIdx.x == 0) {
            const uint32_t E = s_cnt_carry;
            A.rd_u[rd] = (int32_t)u;
            A.rd_possible[rd] = (int32_t)s_pm;  // an i32 sum (src/shmap/mod.rs:689)
            A.rd_S[rd] = S;
            A.rd_hl[rd] = hl;
            A.rd_ok[rd] = ok ? 1 : 0;
            A.rd_c[rd] = (int32_t)s_c;
            A.rd_z[rd] = (int32_t)(zs < s_c ? zs : s_c);  // hitless seeds lead the order
            A.rd_bseeds[rd] = (int32_t)s_seeds;
            A.rd_max_seed[rd] = (int32_t)s_max;
            A.rd_seed_matches[rd] = (int32_t)E;
            A.rd_E[rd] = E;
            uint64_t cap = 0;
            if (E > 0) {
                cap = 64;
                while (cap < 2ull * E) cap <<= 1;
            }
            A.rd_cap[rd] = cap;
        }
        __syncthreads();
    }
}

__global__ void __launch_bounds__(WG) k_tab_init(MapArgs A) {
    if (aborted(A.abort)) return;
    for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < (uint64_t)A.RG; i += (uint64_t)gridDim.x * WG) {
        A.grp_nb[i] = 0;
        A.rd_first[A.r0 + i] = INT32_MAX;
        A.rd_flag[A.r0 + i] = INT32_MAX;
    }
}

// Bucket geometry.  A hit at pos (tpos, or r with abs_pos) lies in half-bucket
// pos / hl; bucket b spans half-buckets b and b+1 (src/buckets.rs:512-518).
// g = ext_start[seg] / hl + 2*seg + b numbers buckets in (segm_id, b) order
// without gaps shared between segments.
__device__ __forceinline__ int32_t hit_pos(const DevIndex &I, uint32_t idx, uint32_t seg, int ap) {
    return ap ? (int32_t)(I.hit_r[idx] & 0x7fffffffu)
              : (int32_t)((int64_t)I.hit_g[idx] - I.kstart[seg]);
}

__device__ __forceinline__ uint64_t bucket_id(const MapArgs &A, uint32_t seg, int32_t b, int32_t hl) {
    const int64_t ext = A.abs_pos ? A.seg_cum[seg] : A.kstart[seg];
    return (uint64_t)(ext / hl) + 2ull * seg + (uint64_t)b;
}

__device__ __forceinline__ void decode_bucket(const MapArgs &A, uint32_t g, int32_t hl, uint32_t &seg,
                                              int32_t &b) {
    int32_t lo = 0, hi = A.nseg - 1;
    while (lo < hi) {
        const int32_t mid = (lo + hi + 1) >> 1;
        if (bucket_id(A, (uint32_t)mid, 0, hl) <= g) lo = mid; else hi = mid - 1;
    }
    seg = (uint32_t)lo;
    b = (int32_t)(g - bucket_id(A, seg, 0, hl));
}

// First hit index in [lo, hi) at or after (seg, pos) -- the partition_point of
// src/shmap/pruning.rs:52-61.  With sketch positions (not abs_pos) a hit's
// global k-mer id is kstart[seg] + pos, below kstart[seg + 1]: ids increase in
// (segment, position) order, so the search compares ids alone against
// kstart[seg] + pos clamped to the segment (one array and no kstart load per
// step).  With abs_pos (positions in bases) it compares segment, then r.
__device__ __forceinline__ uint32_t hits_lower_bound(const DevIndex &I, uint32_t lo, uint32_t hi,
                                                     uint32_t seg, int64_t pos, int ap) {
    if (!ap) {
        const int64_t ks = I.kstart[seg], len = I.kstart[seg + 1] - ks;
        const uint32_t g = (uint32_t)(ks + (pos < 0 ? 0 : (pos < len ? pos : len)));
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (I.hit_g[mid] < g) lo = mid + 1; else hi = mid;
        }
        return lo;
    }
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        const uint32_t s = I.hit_seg[mid];
        const bool before = s != seg ? s < seg : (int64_t)hit_pos(I, mid, s, ap) < pos;
        if (before) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// Adds a bucket's partial sums (mc = sum of (matches << 32) + codirection,
// and the r range) to read lr's global bucket table; the insert that claims an
// empty slot counts the read's new bucket and lists its slot.
__device__ __forceinline__ void tab_add(const MapArgs &A, uint32_t lr, uint64_t base, uint64_t cap, uint32_t g,
                                        uint64_t mcv, int32_t rmin, int32_t rmax) {
    const uint32_t lg = 63 - __clzll(cap);
    const uint32_t mask = (uint32_t)(cap - 1);
    uint32_t pos = ((g * 0x9E3779B1u) >> (32 - lg)) & mask;
    for (;;) {
        const uint32_t old = atomicCAS(&A.tb_key[base + pos], EMPTY_KEY, g);
        if (old == EMPTY_KEY) {
            const uint64_t at = atomicAdd((unsigned long long *)&A.grp_nb[lr], 1ull);
            A.tb_list[base + at] = pos;  // the read's used slots, for k_extract
            break;
        }
        if (old == g) break;
        pos = (pos + 1u) & mask;
    }
    atomicAdd((unsigned long long *)&A.tb_mc[base + pos], (unsigned long long)mcv);
    atomicMin(&A.tb_min[base + pos], rmin);
    atomicMax(&A.tb_max[base + pos], rmax);
}

// A wave's bucket table in LDS: LTAB slots, of which at most LTAB_MAX are
// filled before the table is emptied into the read's global table (or the
// read is handed to that path).
constexpr uint32_t LTAB_BITS = 8;
constexpr uint32_t LTAB = 1u << LTAB_BITS;
constexpr uint32_t LTAB_MAX = LTAB * 7 / 8;
// One wave takes a whole read up to A.contrib_wave_e examined hits (the host
// picks it per batch: long for big batches, which fill the device anyway,
// short for small ones, whose latency a long serial read would set).  Larger
// reads, and reads with more buckets than one LDS table holds, are split into
// chunks, one wave per chunk, whose LDS tables are merged in the read's global
// table (k_contrib_chunks).  The chunk length is A.contrib_chunk when set,
// else chosen per group: the largest power of two from CONTRIB_CHUNK_MIN to
// CONTRIB_CHUNK_MAX hits that still gives CONTRIB_CHUNKS_WANT chunks, so a few
// listed reads are spread over the device and many are taken in long chunks
// (fewer merges).
constexpr uint32_t CONTRIB_CHUNK_MIN = 512, CONTRIB_CHUNK_MAX = 4096;
constexpr uint64_t CONTRIB_CHUNKS_WANT = 8192;

// A workgroup takes a read past one wave whole, its waves splitting the read's
// examined hits into quarters, with a table of WTAB slots in LDS shared by its
// waves, when the read has at most A.contrib_wg_e examined hits and its
// buckets fit (waves reserve a window's worst case of new slots before adding,
// so the table never passes WTAB_MAX); other reads go through global tables.
constexpr uint32_t WTAB_BITS = 10;
constexpr uint32_t WTAB = 1u << WTAB_BITS;
constexpr uint32_t WTAB_MAX = WTAB * 7 / 8;

enum ContribMode { CM_LDS, CM_LDS_FLUSH, CM_WG };

// Rows of 64 positions a wave takes per step of the window walk, and rows
// whose run-end loads are issued together.
constexpr int CROWS = 4;
constexpr int CPREP = 1;
// Reads past one wave are counted by a workgroup of NWAVES waves, those with
// more than CWG_SPLIT examined hits by one of CWG waves.
constexpr int CWG = 16;
constexpr uint32_t CWG_SPLIT = 16384;
static_assert(CROWS % CPREP == 0, "rows are prepared in whole groups");

struct LTab {
    uint32_t *key;
    uint64_t *mark;  // seed starts of the current block, CROWS words
    uint32_t *used;
    uint64_t *mc;
    int32_t *mn, *mx;
    uint32_t *over;  // CM_WG: set when the shared table would pass WTAB_MAX
};

__device__ __forceinline__ void ltab_clear(const LTab &T, uint32_t lane) {
    for (uint32_t i = lane; i < LTAB; i += WAVE) {
        T.key[i] = EMPTY_KEY;
        T.mc[i] = 0;
        T.mn[i] = INT32_MAX;
        T.mx[i] = -1;
    }
}

// Adds a contribution to an LDS table of 2^BITS slots; returns 1 when it
// claimed a new slot (the caller counts claims into *T.used).
// BOUNDED (a table shared by a workgroup's waves, filled without reserving
// room): a probe that finds no slot in the whole table sets *T.over and adds
// nothing.
template <uint32_t BITS, bool BOUNDED = false>
__device__ __forceinline__ uint32_t ltab_add(const LTab &T, uint32_t g, uint64_t mcv, int32_t rmin, int32_t rmax) {
    uint32_t pos = (g * 0x9E3779B1u) >> (32 - BITS), claimed = 0;
    for (uint32_t tries = 0;; tries++) {
        if (BOUNDED && tries == (1u << BITS)) {
            atomicExch(T.over, 1u);
            return 0;
        }
        const uint32_t old = atomicCAS(&T.key[pos], EMPTY_KEY, g);
        if (old == EMPTY_KEY) {
            claimed = 1;
            break;
        }
        if (old == g) break;
        pos = (pos + 1u) & ((1u << BITS) - 1);
    }
    atomicAdd((unsigned long long *)&T.mc[pos], (unsigned long long)mcv);
    atomicMin(&T.mn[pos], rmin);
    atomicMax(&T.mx[pos], rmax);
    return claimed;
}

// Empties the LDS table into read lr's global table.
__device__ __forceinline__ void ltab_flush(const MapArgs &A, const LTab &T, uint32_t lr, uint64_t tb, uint64_t cap,
                                           uint32_t lane) {
    for (uint32_t i = lane; i < LTAB; i += WAVE) {
        const uint32_t k = T.key[i];
        if (k != EMPTY_KEY) {
            tab_add(A, lr, tb, cap, k, T.mc[i], T.mn[i], T.mx[i]);
            T.key[i] = EMPTY_KEY;
            T.mc[i] = 0;
            T.mn[i] = INT32_MAX;
            T.mx[i] = -1;
        }
    }
    if (lane == 0) *T.used = 0;
    wave_sync();
}

// match_seeds (src/shmap/seeding.rs:98-212) over positions [pa, pb) of read
// lr's examined hits, by one wave into its LDS table.  The examined hits are
// each consumed seed's hit run in seed order, flattened: seed q covers
// positions [w, w + y) of its row sd_row[q].  A step takes a block of up to
// CROWS rows of 64 positions (lane j of row i at position P0 + 64 i + j) and
// issues the hit loads of all its rows together, so one chain of dependent
// loads (seed rows, hits, run ends) serves up to 64 * CROWS positions.  From
// the first seed with hits on (rd_z) the seeds' first positions strictly
// increase; a block holds positions of the 64 seeds whose rows are loaded (the
// whole first row, then the positions before the last loaded seed's start),
// and a bitmap of the seed starts in the block gives each position its seed.
// A seed's hits are in (segm_id, pos) order; the first hit of each (segment,
// half-bucket) run of a seed -- a run head -- emits the clamped per-(seed,
// bucket) contribution of bucket b (the hits up to the first at half-bucket
// b+2 or in another segment or seed) and -- unless the previous run, at b-1,
// already did -- of bucket b-1 (up to the first at half-bucket b+1).  Run ends
// come from the block's head masks: the next head ends bucket b-1, and ends
// bucket b too unless it starts the run at b+1 (its previous hit is at b, which
// the second mask marks), when the head after it does.  An end past the block
// leaves the contribution pending (at most three: the last run's two and the
// run before it), settled by ballots in the block where its run ends, or by
// binary search at the end of the range.  A wave's own table is given room
// per row, one new slot per contribution the row can add: CM_LDS returns
// false, the table partly filled, when the table would pass LTAB_MAX;
// CM_LDS_FLUSH then empties it into the global table and goes on.  CM_WG (a
// table of WTAB slots shared by a workgroup's waves) adds without reserving:
// a row that takes the table past WTAB_MAX slots, or a probe that finds it
// full, sets *T.over, and every wave returns false at its next row.
template <int MODE>
__device__ bool contrib_windows(const MapArgs &A, const DevIndex &I, const LTab &T, int32_t lr, uint32_t pa,
                                uint32_t pb, uint32_t E, uint32_t lane) {
    constexpr uint32_t BITS = MODE == CM_WG ? WTAB_BITS : LTAB_BITS;
    constexpr int J = CROWS;
    const int ap = A.abs_pos;
    const uint64_t upto = (((1ull << lane) - 1) << 1) | 1ull;  // lanes 0..lane
    const int32_t rd = A.r0 + lr;
    const uint64_t o = A.koff[rd];
    const uint32_t c = (uint32_t)A.rd_c[rd], z = (uint32_t)A.rd_z[rd];
    const int32_t hl = A.rd_hl[rd];
    uint32_t qc = z;  // the seed holding P0
    if (pa > 0) {
        uint32_t lo = z, hi = c;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (A.sd_row[o + mid].w <= pa) lo = mid + 1; else hi = mid;
        }
        qc = lo - 1;
    }
    // one contribution: hits [idx0, e) of a seed (occs | strand << 31) to bucket bb
    uint32_t claims = 0;  // this lane's new slots in the row (at most three)
    auto add = [&](uint32_t e, uint32_t idx0, uint32_t ocv, int32_t rmin0, uint32_t sg, int32_t bb) {
        const uint32_t n = e - idx0, occs = ocv & 0x7fffffffu;
        const uint32_t ones = I.strand_pre[e] - I.strand_pre[idx0];
        const uint32_t same = (ocv >> 31) ? ones : n - ones;
        const uint32_t matches = n < occs ? n : occs;
        const int32_t codir = (int32_t)(2 * same) - (int32_t)n;
        claims += ltab_add<BITS, MODE == CM_WG>(T, (uint32_t)bucket_id(A, sg, bb, hl), ((uint64_t)matches << 32) + (uint64_t)(int64_t)codir,
                           rmin0, (int32_t)(I.hit_r[e - 1] & 0x7fffffffu));
    };
    // the same from values already loaded: n hits, `ones` of them on the
    // reverse strand, the last hit's r | strand << 31
    auto put = [&](uint32_t n, uint32_t ones, uint32_t ocv, int32_t rmin0, uint32_t rlast, uint32_t sg, int32_t bb) {
        const uint32_t occs = ocv & 0x7fffffffu;
        const uint32_t same = (ocv >> 31) ? ones : n - ones;
        const uint32_t matches = n < occs ? n : occs;
        const int32_t codir = (int32_t)(2 * same) - (int32_t)n;
        claims += ltab_add<BITS, MODE == CM_WG>(T, (uint32_t)bucket_id(A, sg, bb, hl), ((uint64_t)matches << 32) + (uint64_t)(int64_t)codir,
                           rmin0, (int32_t)(rlast & 0x7fffffffu));
    };
    // room for `need` new slots before a row's adds: false when the caller
    // must return false (CM_WG: the table has overflowed)
    auto reserve = [&](uint32_t need) -> bool {
        if (MODE == CM_WG) return *(volatile uint32_t *)T.over == 0;
        if (*T.used + need > LTAB_MAX) {
            if (MODE == CM_LDS) return false;
            ltab_flush(A, T, (uint32_t)lr, A.tab_base[lr], A.rd_cap[rd], lane);
        }
        return true;
    };
    // after the adds: the new slots counted (for CM_WG, *T.over set past
    // WTAB_MAX)
    auto account = [&](uint32_t need) {
        const uint32_t nclaim = (uint32_t)__popcll(__ballot(claims & 1u)) + 2u * (uint32_t)__popcll(__ballot(claims & 2u)) +
                                4u * (uint32_t)__popcll(__ballot(claims & 4u));
        claims = 0;
        wave_sync();
        if (lane == 0) {
            if (MODE == CM_WG) {
                if (nclaim && atomicAdd(T.used, nclaim) + nclaim > WTAB_MAX) atomicExch(T.over, 1u);
            } else {
                *T.used += nclaim;
            }
        }
        wave_sync();
    };
    // pending contributions: lane i < 3 holds item i (pmask, uniform): its
    // seed, segment and first position/hit, the half-bucket that ends it, the
    // bucket, the seed's end, occs | strand and the first hit's r
    uint32_t pmask = 0;
    uint32_t pq = 0, psg = 0, ppos = 0, pidx = 0, pee = 0, poc = 0;
    int32_t ptb = 0, pbb = 0, prmin = 0;
    for (uint32_t P0 = pa; P0 < pb;) {
        if (MODE == CM_WG && *(volatile uint32_t *)T.over) return false;
        const uint32_t qi = qc + lane;
        const bool qv = qi < c;
        const uint4 row = qv ? A.sd_row[o + qi] : make_uint4(0, 0, 0, 0xffffffffu);
        const uint32_t st = row.w;
        // the block [P0, P1): the whole first row (at most 64 seeds start in
        // it), beyond it only positions before the last loaded seed's start
        uint32_t P1 = P0 + (uint32_t)J * WAVE;
        if (J > 1) {
            const uint32_t st_last = __builtin_amdgcn_readlane(st, WAVE - 1);  // ~0 when not loaded
            if (st_last < P1) P1 = st_last > P0 + WAVE ? st_last : P0 + WAVE;
        }
        // the read's end lies in the block: positions from E on end runs
        const bool tail = P1 >= E;
        const uint32_t vis_end = tail ? E + 1 : P1;  // rows holding a position below this are visited
        if (lane > 0 && st > P0 && st < P1)
            atomicOr((unsigned long long *)&T.mark[(st - P0) >> 6], 1ull << ((st - P0) & 63u));
        wave_sync();
        uint64_t M[J];
#pragma unroll
        for (int j = 0; j < J; j++) {
            // (readfirstlane returns int: widen each half as unsigned)
            const uint64_t v = T.mark[j];
            M[j] = ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(v >> 32)) << 32) |
                   (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)v);
        }
        wave_sync();
        if (lane < (uint32_t)J) T.mark[lane] = 0;
        const bool s0start = __builtin_amdgcn_readlane(st, 0) == P0;  // seed qc starts at P0
        const uint32_t sa = row.x - st, send = row.x + row.y, socc = row.z;
        // each row's seed (qc + dq), hit, segment and half-bucket
        uint32_t dq[J], idx[J], seg[J];
        int32_t bk[J];
        uint32_t pre = 0;
#pragma unroll
        for (int j = 0; j < J; j++) {
            const uint32_t p = P0 + (uint32_t)j * WAVE + lane;
            dq[j] = pre + (uint32_t)__popcll(M[j] & upto);
            pre += (uint32_t)__popcll(M[j]);
            idx[j] = __shfl(sa, (int)dq[j]) + p;
            seg[j] = 0xffffffffu;
            if (P0 + (uint32_t)j * WAVE < vis_end && p < P1 && p < E) seg[j] = I.hit_seg[idx[j]];
        }
        uint32_t sg_prev = 0xffffffffu;  // the hit before P0 (same seed)
        int32_t b_prev = -1;
        if (lane == 0 && !s0start) {
            sg_prev = I.hit_seg[idx[0] - 1];
            b_prev = hit_pos(I, idx[0] - 1, sg_prev, ap) / hl;
        }
#pragma unroll
        for (int j = 0; j < J; j++) {
            const uint32_t p = P0 + (uint32_t)j * WAVE + lane;
            bk[j] = -1;
            if (P0 + (uint32_t)j * WAVE < vis_end && p < P1 && p < E) bk[j] = hit_pos(I, idx[j], seg[j], ap) / hl;
        }
        // head masks: H (heads, and positions from E on in the last block),
        // B1 (heads whose previous hit is in the same seed and segment at the
        // previous half-bucket)
        uint64_t H[J], B1[J];
#pragma unroll
        for (int j = 0; j < J; j++) {
            H[j] = B1[j] = 0;
            if (P0 + (uint32_t)j * WAVE >= vis_end) continue;
            const uint32_t p = P0 + (uint32_t)j * WAVE + lane;
            const bool vis = p < P1 || tail, live = p < P1 && p < E;
            uint32_t ps = __shfl_up(seg[j], 1);
            int32_t pbk = __shfl_up(bk[j], 1);
            if (j == 0) {
                if (lane == 0) {
                    ps = sg_prev;
                    pbk = b_prev;
                }
            } else {
                const uint32_t ls = __builtin_amdgcn_readlane(seg[j - 1], WAVE - 1);
                const int32_t lb = __builtin_amdgcn_readlane(bk[j - 1], WAVE - 1);
                if (lane == 0) {
                    ps = ls;
                    pbk = lb;
                }
            }
            const bool sstart = (j == 0 && lane == 0) ? s0start : ((M[j] >> lane) & 1ull) != 0;
            const bool head = live && (sstart || ps != seg[j] || pbk != bk[j]);
            const bool pbm1 = live && !sstart && ps == seg[j] && pbk == bk[j] - 1;
            H[j] = __ballot(head || (vis && !live));
            B1[j] = __ballot(head && pbm1);
        }
        // pending contributions whose run ends in this block
        bool settle = false;
        uint32_t e_set = 0, nset = 0;
        for (uint32_t i = 0; i < 3; i++) {
            if (!((pmask >> i) & 1u)) continue;
            const uint32_t fq = __builtin_amdgcn_readlane(pq, i), fsg = __builtin_amdgcn_readlane(psg, i);
            const int32_t ftb = __builtin_amdgcn_readlane(ptb, i);
            bool found = false;
#pragma unroll
            for (int j = 0; j < J; j++) {
                if (found || P0 + (uint32_t)j * WAVE >= vis_end) continue;
                const uint32_t p = P0 + (uint32_t)j * WAVE + lane;
                const bool vis = p < P1 || tail, live = p < P1 && p < E;
                const uint64_t m = __ballot(vis && (!live || qc + dq[j] != fq || seg[j] != fsg || bk[j] >= ftb));
                if (m) {
                    const uint32_t l = (uint32_t)__ffsll((unsigned long long)m) - 1;
                    if (lane == i) {
                        settle = true;
                        e_set = pidx + (P0 + (uint32_t)j * WAVE + l - ppos);
                    }
                    found = true;
                }
            }
            if (found) {
                pmask &= ~(1u << i);
                nset++;
            }
        }
        // the first row after each row whose head mask is not empty (J: none)
        uint32_t nzr[J];
        {
            uint32_t nx = J;
#pragma unroll
            for (int j = J - 1; j >= 0; j--) {
                nzr[j] = nx;
                if (H[j]) nx = (uint32_t)j;
            }
        }
        auto pick = [&](const uint64_t (&a)[J], uint32_t r) -> uint64_t {  // a[r] for a uniform r < J, else 0
            uint64_t v = 0;
#pragma unroll
            for (int t = 0; t < J; t++)
                if (r == (uint32_t)t) v = a[t];
            return v;
        };
        auto pick_nz = [&](uint32_t r) -> uint32_t {
            uint32_t v = J;
#pragma unroll
            for (int t = 0; t < J; t++)
                if (r == (uint32_t)t) v = nzr[t];
            return v;
        };
        // Rows are taken CPREP at a time: each row's run ends found and the
        // hits they need (strand prefix at both ends, r of the first and last
        // hit) loaded for all CPREP rows, then the table updates row by row,
        // so the loads of the rows overlap.
#pragma unroll
        for (int j0 = 0; j0 < J; j0 += CPREP) {
            uint32_t nA[CPREP], nB[CPREP], sA[CPREP], sB[CPREP], s0[CPREP], xA[CPREP], xB[CPREP], rmn[CPREP];
            uint32_t pf[CPREP];  // bit 0: bucket b's contribution pending, bit 1: bucket b-1's
            uint32_t nh[CPREP];
#pragma unroll
            for (int t = 0; t < CPREP; t++) {
                const int j = j0 + t;
                nA[t] = nB[t] = sA[t] = sB[t] = s0[t] = xA[t] = xB[t] = rmn[t] = pf[t] = nh[t] = 0;
                if (P0 + (uint32_t)j * WAVE >= vis_end) continue;
                const uint32_t p = P0 + (uint32_t)j * WAVE + lane;
                const bool emit = ((H[j] >> lane) & 1ull) && p < P1 && p < E && p < pb;
                nh[t] = (uint32_t)__popcll(__ballot(emit));
                if (emit) {
                    // h1: the next head; h2: the head after it (block positions)
                    const uint32_t ra = nzr[j];  // uniform
                    const uint64_t ha = pick(H, ra), ba = pick(B1, ra);
                    const uint64_t mj = H[j] & ~upto;
                    const bool f1 = mj != 0 || ra < (uint32_t)J;
                    const uint64_t h1m = mj ? H[j] : ha;
                    const uint32_t r1 = mj ? (uint32_t)j : ra;
                    const uint32_t b1 = f1 ? (uint32_t)__ffsll((unsigned long long)(mj ? mj : ha)) - 1 : 0;
                    const bool h1b1 = f1 && (((mj ? B1[j] : ba) >> b1) & 1ull);
                    const uint64_t m2 = b1 < 63 ? (h1m & (~0ull << (b1 + 1))) : 0ull;
                    const uint32_t rb = mj ? ra : pick_nz(ra);  // row after r1 holding a head
                    const uint64_t hb = mj ? ha : pick(H, pick_nz(ra));
                    const bool f2 = f1 && (m2 != 0 || rb < (uint32_t)J);
                    const uint32_t r2 = m2 ? r1 : rb;
                    const uint32_t b2 = m2 ? (uint32_t)__ffsll((unsigned long long)m2) - 1
                                           : (f2 ? (uint32_t)__ffsll((unsigned long long)hb) - 1 : 0);
                    const uint32_t here = (uint32_t)j * WAVE + lane;
                    const uint32_t e1 = r1 * WAVE + b1, e2 = r2 * WAVE + b2;
                    const uint32_t ix = idx[j];
                    s0[t] = I.strand_pre[ix];
                    rmn[t] = I.hit_r[ix] & 0x7fffffffu;
                    if (h1b1 ? f2 : f1) {
                        nA[t] = (h1b1 ? e2 : e1) - here;
                        sA[t] = I.strand_pre[ix + nA[t]];
                        xA[t] = I.hit_r[ix + nA[t] - 1];
                    } else {
                        pf[t] |= 1u;
                    }
                    if (bk[j] > 0 && !((B1[j] >> lane) & 1ull)) {
                        if (f1) {
                            nB[t] = e1 - here;
                            sB[t] = I.strand_pre[ix + nB[t]];
                            xB[t] = I.hit_r[ix + nB[t] - 1];
                        } else {
                            pf[t] |= 2u;
                        }
                    }
                }
            }
#pragma unroll
            for (int t = 0; t < CPREP; t++) {
                const int j = j0 + t;
                if (P0 + (uint32_t)j * WAVE >= vis_end) continue;
                const uint32_t need = 2 * nh[t] + (j == 0 ? nset : 0);
                if (need == 0) continue;
                if (!reserve(need)) return false;
                if (j == 0 && settle) add(e_set, pidx, poc, prmin, psg, pbb);
                const uint32_t oc = __shfl(socc, (int)dq[j]);
                if (nA[t]) put(nA[t], sA[t] - s0[t], oc, (int32_t)rmn[t], xA[t], seg[j], bk[j]);
                if (nB[t]) put(nB[t], sB[t] - s0[t], oc, (int32_t)rmn[t], xB[t], seg[j], bk[j] - 1);
                // this row's open contributions join the pending ones
                uint64_t m0 = __ballot(pf[t] & 1u), m1 = __ballot(pf[t] & 2u);
                while (m0 | m1) {
                    const bool s1 = m0 == 0;
                    const uint32_t L = (uint32_t)__ffsll((unsigned long long)(s1 ? m1 : m0)) - 1;
                    if (s1) m1 &= m1 - 1; else m0 &= m0 - 1;
                    const uint32_t slot = (uint32_t)__ffs(~pmask & 7u) - 1;
                    pmask |= 1u << slot;
                    const uint32_t vd = __builtin_amdgcn_readlane(dq[j], L);
                    const uint32_t vsg = __builtin_amdgcn_readlane(seg[j], L), vidx = __builtin_amdgcn_readlane(idx[j], L);
                    const uint32_t vee = __builtin_amdgcn_readlane(send, vd), voc = __builtin_amdgcn_readlane(socc, vd);
                    const int32_t vbk = __builtin_amdgcn_readlane(bk[j], L);
                    const int32_t vrmin = (int32_t)__builtin_amdgcn_readlane(rmn[t], L);
                    if (lane == slot) {
                        pq = qc + vd;
                        psg = vsg;
                        ppos = P0 + (uint32_t)j * WAVE + L;
                        pidx = vidx;
                        pee = vee;
                        poc = voc;
                        ptb = vbk + (s1 ? 1 : 2);
                        pbb = vbk - (s1 ? 1 : 0);
                        prmin = vrmin;
                    }
                }
                account(need);
            }
        }
        // the range ends: what is still open ends past the block
        if (P1 >= pb && pmask) {
            const uint32_t need = (uint32_t)__popc(pmask);
            if (!reserve(need)) return false;
            if (lane < 3 && ((pmask >> lane) & 1u)) {
                const uint32_t from = pidx + (P1 - ppos);
                add(hits_lower_bound(I, from, pee, psg, (int64_t)ptb * hl, ap), pidx, poc, prmin, psg, pbb);
            }
            pmask = 0;
            account(need);
        }
        // the seed holding P1: every seed starting at or before it
        uint32_t adv = (uint32_t)__popcll(__ballot(lane > 0 && st <= P1));
        if (adv == WAVE - 1 && qc + WAVE < c && A.sd_row[o + qc + WAVE].w <= P1) adv++;
        qc += adv;
        P0 = P1;
        wave_sync();
    }
    return true;
}

#define LTAB_DECLARE()                                                                         \
    __shared__ uint32_t s_key[NWAVES][LTAB];                                                   \
    __shared__ uint64_t s_mc[NWAVES][LTAB];                                                    \
    __shared__ int32_t s_mn[NWAVES][LTAB], s_mx[NWAVES][LTAB];                                 \
    __shared__ uint64_t s_mark[NWAVES][CROWS];                                                 \
    __shared__ uint32_t s_used[NWAVES];                                                        \
    const uint32_t lane = threadIdx.x & (WAVE - 1), w = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE); \
    const LTab T{s_key[w], s_mark[w], &s_used[w], s_mc[w], s_mn[w], s_mx[w], nullptr};         \
    ltab_clear(T, lane);                                                                       \
    if (lane < (uint32_t)CROWS) T.mark[lane] = 0;                                              \
    if (lane == 0) *T.used = 0;                                                                \
    wave_sync()

// One wave per read with at most A.contrib_wave_e examined hits: its buckets end
// as the first grp_nb rows of its table region, listed in order.  Larger
// reads, and reads with more than LTAB_MAX buckets, are listed for
// k_contrib_wg.
__global__ void __launch_bounds__(WG) k_contrib_wave(MapArgs A, DevIndex I) {
    if (aborted(A.abort)) return;
    LTAB_DECLARE();
    const uint64_t below = (1ull << lane) - 1;
    for (int32_t lr = blockIdx.x * NWAVES + w; lr < A.RG; lr += gridDim.x * NWAVES) {
        const uint32_t E = (uint32_t)(A.hit_base[lr + 1] - A.hit_base[lr]);
        if (E == 0) continue;
        if (E > A.contrib_wave_e) {
            if (lane == 0) A.ov_list[atomicAdd(A.ov_count, 1u)] = (uint32_t)lr;
            continue;
        }
        if (!contrib_windows<CM_LDS>(A, I, T, lr, 0, E, E, lane)) {  // too many buckets
            ltab_clear(T, lane);
            if (lane == 0) {
                *T.used = 0;
                A.ov_list[atomicAdd(A.ov_count, 1u)] = (uint32_t)lr;
            }
            wave_sync();
            continue;
        }
        // the used slots become rows 0.. of the read's table region
        const uint64_t tb = A.tab_base[lr];
        uint32_t nb = 0;
        for (uint32_t s0 = 0; s0 < LTAB; s0 += WAVE) {
            const uint32_t sl = s0 + lane;
            const uint32_t k = T.key[sl];
            const bool used = k != EMPTY_KEY;
            const uint64_t bm = __ballot(used);
            if (used) {
                const uint32_t at = nb + (uint32_t)__popcll(bm & below);
                A.tb_key[tb + at] = k;
                A.tb_mc[tb + at] = T.mc[sl];
                A.tb_min[tb + at] = T.mn[sl];
                A.tb_max[tb + at] = T.mx[sl];
                A.tb_list[tb + at] = at;
                T.key[sl] = EMPTY_KEY;
                T.mc[sl] = 0;
                T.mn[sl] = INT32_MAX;
                T.mx[sl] = -1;
            }
            nb += (uint32_t)__popcll(bm);
        }
        if (lane == 0) {
            A.grp_nb[lr] = nb;
            *T.used = 0;
        }
        wave_sync();
    }
}

// One workgroup of CW waves per read listed by k_contrib_wave whose examined
// hits E satisfy e_lo < E <= e_hi: the hits split into CW parts, one per wave,
// counted into one LDS table of WTAB slots; its buckets end as the first
// grp_nb rows of its table region, as k_contrib_wave leaves them.  Reads with
// more than A.contrib_wg_e examined hits, or whose buckets overfill the table,
// are listed for the global tables (k_contrib_chunks).  Satellite reads with
// hundreds of thousands of hits take 16 waves, so a few of them do not
// outlast the rest of the launch.
template <int CW>
__global__ void __launch_bounds__(CW * WAVE) k_contrib_wg(MapArgs A, DevIndex I, uint32_t e_lo, uint32_t e_hi) {
    if (aborted(A.abort)) return;
    constexpr uint32_t NT = (uint32_t)CW * WAVE;
    __shared__ uint32_t s_key[WTAB];
    __shared__ uint64_t s_mc[WTAB];
    __shared__ int32_t s_mn[WTAB], s_mx[WTAB];
    __shared__ uint64_t s_mark[CW][CROWS];
    __shared__ uint32_t s_used, s_over;
    const uint32_t lane = threadIdx.x & (WAVE - 1), w = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
    const LTab T{s_key, s_mark[w], &s_used, s_mc, s_mn, s_mx, &s_over};
    if (lane < (uint32_t)CROWS) s_mark[w][lane] = 0;
    const uint32_t n = *A.ov_count;
    for (uint32_t t = blockIdx.x; t < n; t += gridDim.x) {
        const int32_t lr = (int32_t)A.ov_list[t];
        const uint32_t E = (uint32_t)(A.hit_base[lr + 1] - A.hit_base[lr]);
        if (E <= e_lo || E > e_hi) continue;  // the other width's read
        if (E > A.contrib_wg_e) {
            if (threadIdx.x == 0) A.gl_list[atomicAdd(A.gl_count, 1u)] = (uint32_t)lr;
            continue;
        }
        for (uint32_t i = threadIdx.x; i < WTAB; i += NT) {
            s_key[i] = EMPTY_KEY;
            s_mc[i] = 0;
            s_mn[i] = INT32_MAX;
            s_mx[i] = -1;
        }
        if (threadIdx.x == 0) s_used = s_over = 0;
        __syncthreads();
        const uint32_t pa = (uint32_t)((uint64_t)E * w / CW), pb = (uint32_t)((uint64_t)E * (w + 1) / CW);
        if (pa < pb) (void)contrib_windows<CM_WG>(A, I, T, lr, pa, pb, E, lane);
        __syncthreads();
        if (s_over) {
            if (threadIdx.x == 0) A.gl_list[atomicAdd(A.gl_count, 1u)] = (uint32_t)lr;
            __syncthreads();
            continue;
        }
        // the used slots become rows 0.. of the read's table region, in slot
        // order (the first wave)
        if (w == 0) {
            const uint64_t tb = A.tab_base[lr], below = (1ull << lane) - 1;
            uint32_t nb = 0;
            for (uint32_t s0 = 0; s0 < WTAB; s0 += WAVE) {
                const uint32_t sl = s0 + lane;
                const uint32_t k = s_key[sl];
                const bool used = k != EMPTY_KEY;
                const uint64_t bm = __ballot(used);
                if (used) {
                    const uint32_t at = nb + (uint32_t)__popcll(bm & below);
                    A.tb_key[tb + at] = k;
                    A.tb_mc[tb + at] = s_mc[sl];
                    A.tb_min[tb + at] = s_mn[sl];
                    A.tb_max[tb + at] = s_mx[sl];
                    A.tb_list[tb + at] = at;
                }
                nb += (uint32_t)__popcll(bm);
            }
            if (lane == 0) A.grp_nb[lr] = nb;
        }
        __syncthreads();
    }
}

// The chunk length for the listed reads (kept at ov_kpre[n + 1]) and their
// chunk prefix (one workgroup).
__global__ void __launch_bounds__(WG) k_contrib_over_scan(MapArgs A) {
    if (aborted(A.abort)) return;
    __shared__ uint64_t tmp[NWAVES + 1];
    const uint32_t n = *A.gl_count;
    uint64_t part = 0;
    for (uint32_t i = threadIdx.x; i < n; i += WG) {
        const uint32_t lr = A.gl_list[i];
        part += A.hit_base[lr + 1] - A.hit_base[lr];
    }
    uint64_t tot_e;
    (void)wg_excl_scan64(part, tmp, &tot_e);
    uint32_t cs = A.contrib_chunk;
    if (cs == 0) {
        cs = CONTRIB_CHUNK_MAX;
        while (cs > CONTRIB_CHUNK_MIN && tot_e / cs < CONTRIB_CHUNKS_WANT) cs >>= 1;
    }
    uint64_t carry = 0;
    for (uint32_t b0 = 0; b0 < n; b0 += WG) {
        const uint32_t i = b0 + threadIdx.x;
        uint64_t k = 0;
        if (i < n) {
            const uint32_t lr = A.gl_list[i];
            k = (A.hit_base[lr + 1] - A.hit_base[lr] + cs - 1) / cs;
        }
        uint64_t tot;
        const uint64_t ex = wg_excl_scan64(k, tmp, &tot);
        if (i < n) A.ov_kpre[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) {
        A.ov_kpre[n] = carry;
        A.ov_kpre[n + 1] = cs;
    }
}

// Empty global tables for the listed reads, one workgroup per read.
__global__ void __launch_bounds__(WG) k_contrib_over_clear(MapArgs A) {
    if (aborted(A.abort)) return;
    const uint32_t n = *A.gl_count;
    for (uint32_t t = blockIdx.x; t < n; t += gridDim.x) {
        const uint32_t lr = A.gl_list[t];
        const uint64_t tb = A.tab_base[lr], cap = A.rd_cap[A.r0 + (int32_t)lr];
        for (uint64_t i = threadIdx.x; i < cap; i += WG) {
            A.tb_key[tb + i] = EMPTY_KEY;
            A.tb_mc[tb + i] = 0;
            A.tb_min[tb + i] = INT32_MAX;
            A.tb_max[tb + i] = -1;
        }
    }
}

// The listed reads in chunks (the length at ov_kpre[n + 1]), one wave per
// chunk: each chunk is counted in LDS and merged into the read's global table
// (when its table fills, and at the end).
__global__ void __launch_bounds__(WG) k_contrib_chunks(MapArgs A, DevIndex I) {
    if (aborted(A.abort)) return;
    LTAB_DECLARE();
    const uint32_t nl = *A.gl_count;
    const uint64_t nk = A.ov_kpre[nl];
    const uint32_t cs = (uint32_t)A.ov_kpre[nl + 1];
    for (uint64_t k = (uint64_t)blockIdx.x * NWAVES + w; k < nk; k += (uint64_t)gridDim.x * NWAVES) {
        const uint32_t i = upper_bound_u64(A.ov_kpre, nl + 1, k) - 1;
        const int32_t lr = (int32_t)A.gl_list[i];
        const uint32_t E = (uint32_t)(A.hit_base[lr + 1] - A.hit_base[lr]);
        const uint32_t pa = (uint32_t)(k - A.ov_kpre[i]) * cs;
        const uint32_t pb = pa + cs < E ? pa + cs : E;
        contrib_windows<CM_LDS_FLUSH>(A, I, T, lr, pa, pb, E, lane);
        ltab_flush(A, T, (uint32_t)lr, A.tab_base[lr], A.rd_cap[A.r0 + lr], lane);
    }
}

// propagate_seeds_to_buckets + get_sorted_buckets (src/buckets.rs:603-640,
// 685-724): a read's buckets ordered by (matches descending, bucket id
// ascending).  Bucket ids are unique within a read, so one sort on the key
// ((~matches) << 32 | id) gives that order directly.
__device__ __forceinline__ uint32_t mc_matches(uint64_t mc, int32_t &codir) {
    codir = (int32_t)(uint32_t)mc;  // the low word is the (signed) codirection sum
    return (uint32_t)((mc - (uint64_t)(int64_t)codir) >> 32);
}

__device__ __forceinline__ void emit_bucket(const MapArgs &A, uint64_t base, uint32_t p, uint64_t key,
                                            uint64_t tb, uint32_t slot) {
    int32_t codir;
    (void)mc_matches(A.tb_mc[tb + slot], codir);
    A.bs_g[base + p] = (uint32_t)key;
    A.bs_m[base + p] = (int32_t)(0xffffffffu - (uint32_t)(key >> 32));
    A.bs_c[base + p] = codir;
    A.bs_min[base + p] = A.tb_min[tb + slot];
    A.bs_max[base + p] = A.tb_max[tb + slot];
}

// One wave per read with at most EXTRACT_WAVE_CAP buckets: the listed slots are
// keyed, sorted in registers across the wave and emitted straight from them.
// Reads with more buckets are listed for k_extract.
__global__ void __launch_bounds__(WG) k_extract_wave(MapArgs A) {
    if (aborted(A.abort)) return;
    const uint32_t lane = threadIdx.x & (WAVE - 1), w = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
    for (int32_t lr = blockIdx.x * NWAVES + w; lr < A.RG; lr += gridDim.x * NWAVES) {
        const uint32_t n = (uint32_t)A.grp_nb[lr];
        if (n > EXTRACT_WAVE_CAP) {
            if (lane == 0) A.ext_list[atomicAdd(A.ext_count, 1u)] = (uint32_t)lr;
            continue;
        }
        const uint64_t tb = A.tab_base[lr], base = A.bk_base[lr];
        wave_sort_n<true>(
            n,
            [&](uint32_t i, uint64_t &key, uint32_t &slot) {
                slot = A.tb_list[tb + i];
                int32_t codir;
                const uint32_t matches = mc_matches(A.tb_mc[tb + slot], codir);
                key = ((uint64_t)(0xffffffffu - matches) << 32) | A.tb_key[tb + slot];
            },
            [&](uint32_t p, uint64_t key, uint32_t slot) { emit_bucket(A, base, p, key, tb, slot); });
    }
}

// The same for the listed reads with more buckets, one workgroup each, through
// global memory with the sort staged in LDS when it fits.
__global__ void __launch_bounds__(WG) k_extract(MapArgs A) {
    if (aborted(A.abort)) return;
    extern __shared__ uint8_t lds[];
    uint64_t *sk = (uint64_t *)lds;
    uint32_t *sv = (uint32_t *)(lds + (size_t)SORT_CAP * 8);
    const uint32_t nlist = *A.ext_count;
    for (uint32_t t = blockIdx.x; t < nlist; t += gridDim.x) {
        const int32_t lr = (int32_t)A.ext_list[t];
        const uint64_t tb = A.tab_base[lr], base = A.bk_base[lr];
        const uint32_t nb = (uint32_t)A.grp_nb[lr];
        // key ((~matches) << 32 | bucket id), value the slot; sorted in
        // registers when the read's buckets fit, else through global memory
        auto key_of = [&](uint32_t i, uint64_t &key, uint32_t &slot) {
            slot = A.tb_list[tb + i];
            int32_t codir;
            const uint32_t matches = mc_matches(A.tb_mc[tb + slot], codir);
            key = ((uint64_t)(0xffffffffu - matches) << 32) | A.tb_key[tb + slot];
        };
        if (!wg_sort_n<true>(
                nb, key_of, [&](uint32_t p, uint64_t key, uint32_t slot) { emit_bucket(A, base, p, key, tb, slot); },
                sk, sv)) {
            for (uint32_t i = threadIdx.x; i < nb; i += WG) key_of(i, A.bk_key[base + i], A.bk_val[base + i]);
            __syncthreads();
            wg_sort_pairs(A.bk_key + base, A.bk_val + base, nb, sk, sv, SORT_CAP);
            for (uint32_t p = threadIdx.x; p < nb; p += WG)
                emit_bucket(A, base, p, A.bk_key[base + p], tb, A.bk_val[base + p]);
            __syncthreads();
        }
    }
}

// seed_heuristic_pass + matches_in_bucket (src/shmap/pruning.rs:10-119).  Each
// extension adds occs to seeds and at most occs to matches, so sh never
// increases: survival under any threshold is decided by the fully extended
// value, and a bucket the reference refines is always fully extended.  Every
// bucket of the group is tested against theta*(1-min_diff) (the lowest
// threshold either match_rest pass can use).  k_prune takes one thread per
// bucket: the thread extends its bucket seed by seed (the two binary searches
// of matches_in_bucket for the bucket's span in the seed's hit run), stopping
// as soon as sh falls below the threshold, for up to A.prune_seq seeds
// (most buckets die within a few); a bucket still alive with seeds left is
// listed, with its partial sums, for k_prune_ext, where one wave extends it 64
// seeds at a time, stopping at the first chunk that kills it.  Survivors are listed for refinement, and
// each read's first first-pass candidate is kept in rd_first.

// A surviving bucket's refine task: first-pass list when sh reaches theta,
// else the second-pass band at the end of the task array.
__device__ __forceinline__ void prune_keep(const MapArgs &A, int32_t lr, int32_t p, double sh) {
    if (!(sh < A.theta)) {
        const uint32_t t = atomicAdd(&A.task_count[0], 1u);
        A.tasks[t] = make_uint2((uint32_t)lr, (uint32_t)p);
        atomicMin(&A.rd_first[A.r0 + lr], p);
    } else {
        const uint32_t t = atomicAdd(&A.task_count[1], 1u);
        A.tasks[A.task_cap - 1 - t] = make_uint2((uint32_t)lr, (uint32_t)p);
    }
}

__global__ void __launch_bounds__(WG) k_prune(MapArgs A, DevIndex I) {
    if (aborted(A.abort)) return;
    __shared__ uint64_t s_bb[WG + 1];
    __shared__ uint32_t s_r0, s_nr;
    const uint64_t B = A.bk_base[A.RG];
    const int ap = A.abs_pos;
    for (uint64_t f0 = (uint64_t)blockIdx.x * WG; f0 < B; f0 += (uint64_t)gridDim.x * WG) {
        // the block's buckets belong to group reads [r0, r0 + nr)
        if (threadIdx.x == 0) {
            const uint64_t fl = f0 + WG - 1 < B ? f0 + WG - 1 : B - 1;
            const uint32_t a = upper_bound_u64(A.bk_base, (uint32_t)A.RG + 1, f0) - 1;
            const uint32_t b = upper_bound_u64(A.bk_base + a, (uint32_t)A.RG + 1 - a, fl) - 1 + a;
            s_r0 = a;
            s_nr = b - a + 1;
        }
        __syncthreads();
        const uint32_t r0 = s_r0, nr = s_nr;
        const bool staged = nr + 1 <= (uint32_t)WG + 1;
        if (staged)
            for (uint32_t i = threadIdx.x; i <= nr; i += WG) s_bb[i] = A.bk_base[r0 + i];
        __syncthreads();
        const uint64_t f = f0 + threadIdx.x;
        if (f < B) {
            const int32_t lr = (int32_t)(r0 - 1 + (staged ? upper_bound_u64(s_bb, nr + 1, f)
                                                          : upper_bound_u64(A.bk_base + r0, nr + 1, f)));
            const int32_t rd = A.r0 + lr;
            const uint64_t base = A.bk_base[lr], o = A.koff[rd];
            const int32_t p = (int32_t)(f - base);
            const int32_t m = (int32_t)(A.koff[rd + 1] - o);
            const int32_t u = A.rd_u[rd], c = A.rd_c[rd], hl = A.rd_hl[rd];
            const int32_t seeds0 = A.rd_bseeds[rd];
            if (A.no_bucket_pruning) {
                A.bs_alive[base + p] = 1;
                A.bs_sh[base + p] = 1.0;
                const uint32_t t = atomicAdd(&A.task_count[0], 1u);
                A.tasks[t] = make_uint2((uint32_t)lr, (uint32_t)p);
                atomicMin(&A.rd_first[rd], p);
            } else if (const int32_t m0 = A.bs_m[base + p];
                       1.0 - (double)(seeds0 - m0) / (double)m < A.thr_min) {
                // dead before its first seed (the loop below would stop at
                // once and leave its sums as they are): only the verdict
                A.bs_alive[base + p] = 0;
                A.bs_sh[base + p] = 1.0 - (double)(seeds0 - m0) / (double)m;
            } else {
                // this thread extends its bucket seed by seed, stopping as soon
                // as it dies, for up to A.prune_seq seeds; a bucket still
                // alive with seeds left goes to k_prune_ext with its partial
                // sums, its next seed (sc_i) and its seeds so far (sc_same)
                int32_t matches = m0, codir = A.bs_c[base + p];
                int32_t rmin = A.bs_min[base + p], rmax = A.bs_max[base + p];
                int32_t seeds = seeds0;
                uint32_t seg;
                int32_t b;
                decode_bucket(A, A.bs_g[base + p], hl, seg, b);
                const int64_t begin = (int64_t)b * hl, end = (int64_t)(b + 2) * hl;
                uint8_t alive = 1;
                double sh;
                const int32_t qend = u - c <= (int32_t)A.prune_seq ? u : c + (int32_t)A.prune_seq;
                int32_t q = c;
                for (;; q++) {
                    sh = 1.0 - (double)(seeds - matches) / (double)m;
                    if (sh < A.thr_min) {
                        alive = 0;
                        break;
                    }
                    if (q >= qend) break;
                    const uint4 row = A.sd_row[o + q];
                    const uint32_t occs = row.z & 0x7fffffffu, cnt = row.y, hoff = row.x;
                    seeds += (int32_t)occs;
                    if (cnt == 0) continue;
                    const uint32_t lo = hits_lower_bound(I, hoff, hoff + cnt, seg, begin, ap);
                    const uint32_t e = hits_lower_bound(I, lo, hoff + cnt, seg, end, ap);
                    const uint32_t n = e - lo;
                    if (n == 0) continue;
                    const uint32_t ones = I.strand_pre[e] - I.strand_pre[lo];
                    const uint32_t same = (row.z >> 31) ? ones : n - ones;
                    codir += (int32_t)(2 * same) - (int32_t)n;
                    const int32_t x0 = (int32_t)(I.hit_r[lo] & 0x7fffffffu);
                    const int32_t x1 = (int32_t)(I.hit_r[e - 1] & 0x7fffffffu);
                    if (x0 < rmin) rmin = x0;
                    if (x1 > rmax) rmax = x1;
                    matches += (int32_t)(n < occs ? n : occs);
                }
                A.bs_m[base + p] = matches;
                A.bs_c[base + p] = codir;
                A.bs_min[base + p] = rmin;
                A.bs_max[base + p] = rmax;
                if (alive && q < u) {
                    A.sc_i[base + p] = q;
                    A.sc_same[base + p] = seeds;
                    A.sel[atomicAdd(&A.task_count[5], 1u)] = make_uint2((uint32_t)lr, (uint32_t)p);
                } else {
                    A.bs_alive[base + p] = alive;
                    A.bs_sh[base + p] = sh;
                    if (alive) prune_keep(A, lr, p, sh);
                }
            }
        }
        __syncthreads();
    }
}

__global__ void __launch_bounds__(WG) k_prune_ext(MapArgs A, DevIndex I) {
    if (aborted(A.abort)) return;
    const uint32_t lane = threadIdx.x & (WAVE - 1), w = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
    const int ap = A.abs_pos;
    const uint32_t nl = A.task_count[5];
    for (uint32_t k = blockIdx.x * NWAVES + w; k < nl; k += gridDim.x * NWAVES) {
        const uint2 e2 = A.sel[k];
        const int32_t lr = (int32_t)e2.x, pb = (int32_t)e2.y;
        const int32_t rd = A.r0 + lr;
        const uint64_t base = A.bk_base[lr], o = A.koff[rd];
        const int32_t m = (int32_t)(A.koff[rd + 1] - o);
        const int32_t u = A.rd_u[rd], hl = A.rd_hl[rd];
        int32_t matches = A.bs_m[base + pb], codir = A.bs_c[base + pb];
        int32_t rmin = A.bs_min[base + pb], rmax = A.bs_max[base + pb];
        int32_t seeds = A.sc_same[base + pb];  // k_prune's hand-over
        uint32_t seg;
        int32_t b;
        decode_bucket(A, A.bs_g[base + pb], hl, seg, b);
        const int64_t begin = (int64_t)b * hl, end = (int64_t)(b + 2) * hl;
        double sh = 1.0 - (double)(seeds - matches) / (double)m;
        bool alive = true;
        for (int32_t q0 = A.sc_i[base + pb]; alive && q0 < u; q0 += WAVE) {
            const int32_t q = q0 + (int32_t)lane;
            int32_t d_seeds = 0, d_matches = 0, d_codir = 0, d_min = INT32_MAX, d_max = -1;
            if (q < u) {
                const uint4 row = A.sd_row[o + q];
                const uint32_t occs = row.z & 0x7fffffffu, cnt = row.y;
                d_seeds = (int32_t)occs;
                if (cnt) {
                    const uint32_t hoff = row.x;
                    const uint32_t lo = hits_lower_bound(I, hoff, hoff + cnt, seg, begin, ap);
                    const uint32_t e = hits_lower_bound(I, lo, hoff + cnt, seg, end, ap);
                    const uint32_t n = e - lo;
                    if (n) {
                        const uint32_t ones = I.strand_pre[e] - I.strand_pre[lo];
                        const uint32_t same = (row.z >> 31) ? ones : n - ones;
                        d_codir = (int32_t)(2 * same) - (int32_t)n;
                        d_min = (int32_t)(I.hit_r[lo] & 0x7fffffffu);
                        d_max = (int32_t)(I.hit_r[e - 1] & 0x7fffffffu);
                        d_matches = (int32_t)(n < occs ? n : occs);
                    }
                }
            }
            for (int d = WAVE / 2; d > 0; d >>= 1) {
                d_seeds += __shfl_xor(d_seeds, d);
                d_matches += __shfl_xor(d_matches, d);
                d_codir += __shfl_xor(d_codir, d);
                d_min = min(d_min, (int32_t)__shfl_xor(d_min, d));
                d_max = max(d_max, (int32_t)__shfl_xor(d_max, d));
            }
            seeds += d_seeds;
            matches += d_matches;
            codir += d_codir;
            rmin = min(rmin, d_min);
            rmax = max(rmax, d_max);
            sh = 1.0 - (double)(seeds - matches) / (double)m;
            alive = !(sh < A.thr_min);
        }
        if (lane == 0) {
            A.bs_m[base + pb] = matches;
            A.bs_c[base + pb] = codir;
            A.bs_min[base + pb] = rmin;
            A.bs_max[base + pb] = rmax;
            A.bs_alive[base + pb] = alive ? 1 : 0;
            A.bs_sh[base + pb] = sh;
            if (alive) prune_keep(A, lr, pb, sh);
        }
    }
}

__device__ __forceinline__ int32_t read_seed_of(const MapArgs &A, uint64_t o, int32_t u, uint64_t h) {
    int32_t lo = 0, hi = u;
    while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (A.sd_h[o + mid] < h) lo = mid + 1; else hi = mid;
    }
    return (lo < u && A.sd_h[o + lo] == h) ? lo : -1;
}

// Inclusive prefix sum of a[0, n) in place, over the workgroup.
__device__ void wg_scan_inplace(int32_t *a, int64_t n, uint32_t *tmp) {
    __shared__ int32_t s_carry;
    if (threadIdx.x == 0) s_carry = 0;
    __syncthreads();
    for (int64_t b0 = 0; b0 < n; b0 += WG) {
        const int64_t i = b0 + threadIdx.x;
        const int32_t v = i < n ? a[i] : 0;
        uint32_t tot;
        const uint32_t ex = wg_excl_scan((uint32_t)v, tmp, &tot);
        const int32_t c = s_carry;
        if (i < n) a[i] = c + (int32_t)ex + v;
        __syncthreads();
        if (threadIdx.x == 0) s_carry = c + (int32_t)tot;
        __syncthreads();
    }
}

// Selects one refine round's buckets into A.sel (count task_count[2 + round]).
// A superset of what the two match_rest passes read is refined, in rounds:
//   0  each read's first first-pass candidate in sorted order -- the first pass
//      always refines it (its threshold is still theta there);
//   1  the other first-pass candidates whose bound reaches
//      L1 = max(theta, J(first)): after the first candidate the first pass's
//      threshold is at least L1, so lower bounds are never refined;
//   2  for reads the first pass maps, what the second pass can reach and rounds
//      0-1 did not refine: bounds of at least
//      L2 = J(first bucket above theta) * (1 - min_diff), which is at most the
//      second pass's starting threshold best * (1 - min_diff).
__global__ void __launch_bounds__(WG) k_refine_select(MapArgs A) {
    if (aborted(A.abort)) return;
    const int32_t round = A.refine_pass;
    const uint32_t c0 = A.task_count[0], c1 = A.task_count[1];
    const uint64_t n = round == 0 ? (uint64_t)A.RG : round == 1 ? c0 : (uint64_t)c0 + c1;
    const uint32_t lane = threadIdx.x & (WAVE - 1);
    for (uint64_t b0 = (uint64_t)blockIdx.x * WG; b0 < n; b0 += (uint64_t)gridDim.x * WG) {
        const uint64_t t = b0 + threadIdx.x;
        bool keep = false;
        uint2 task = make_uint2(0, 0);
        if (t < n) {
            if (round == 0) {
                const int32_t p = A.rd_first[A.r0 + (int32_t)t];
                keep = p != INT32_MAX;
                task = make_uint2((uint32_t)t, (uint32_t)p);
            } else {
                task = A.tasks[t < c0 ? t : A.task_cap - 1 - (t - c0)];
                const int32_t lr = (int32_t)task.x, p = (int32_t)task.y;
                const int32_t rd = A.r0 + lr;
                const uint64_t bx = A.bk_base[lr];
                const int32_t first = A.rd_first[rd];
                const double shp = A.bs_sh[bx + p];
                bool in_r1 = false;  // refined in round 1 (first-pass candidates only)
                if (t < c0 && p != first) {
                    const double jf = A.sc_j[bx + first];
                    in_r1 = !(shp < (jf > A.theta ? jf : A.theta));
                }
                if (round == 1) {
                    keep = in_r1;
                } else {
                    const int32_t succ = A.rd_flag[rd];
                    keep = succ != INT32_MAX && p != first && !in_r1 &&
                           !(shp < A.sc_j[bx + succ] * (1.0 - A.min_diff));
                }
            }
        }
        const uint64_t bal = __ballot(keep);
        if (bal) {
            uint32_t at = 0;
            if (lane == 0) at = atomicAdd(&A.task_count[2 + round], (uint32_t)__popcll(bal));
            at = __shfl(at, 0);
            if (keep) A.sel[at + __popcll(bal & ((1ull << lane) - 1ull))] = task;
        }
    }
}

// bucket_LCS (src/shmap/scoring.rs:114-170) for one wave: the longest strictly
// increasing subsequence of the pmatches values of span positions [0, n) --
// for each matched position in order, the read positions of its seed's
// occurrences, latest first -- and of the reversed sequence, by patience
// sorting (each value replaces the first tail at or above it, or is appended).
// The wave walks 64 positions at a time: a scan of their occurrence counts
// places each of the next 64 values on a lane, which loads it.  Up to
// LCS_TAILS tails live across the wave's registers (tail i in register i / 64
// of lane i % 64, and each chunk's first tail in one more register, lane =
// chunk); a value's place is the number of smaller tails: one ballot over the
// chunk heads picks the chunk, one more counts inside it.
//
// A position's values form a run, strictly decreasing (forward) or increasing
// (reversed).  Single values (and runs shorter than LCS_RUN_MIN) are taken
// one by one.  A run of LCS_RUN_MIN or more values in a round is applied at
// once from the tails in LDS (Ts):
// every value's lower bound p among the tails before the run, by binary
// search.  Decreasing: processing them one by one would give each value that
// same p (earlier, larger values only lower tails at or after it), so tail p
// becomes the smallest value with that p, and the run appends at most once.
// Increasing: value j's place is max over i <= j of p_i + (j - i), a running
// maximum, and the places are distinct.  When the tails outnumber the
// registers (a sketch above LCS_TAILS k-mers can), they stay in Ts for good and
// a single value's place comes from ballots over 64 evenly spaced tails, a
// few rounds each; Ts holds max(m, LCS_TAILS) + 2 entries (the tails are
// distinct read positions, so at most m).
constexpr int32_t LCS_TAILS = 32 * WAVE;
constexpr uint32_t LCS_RUN_MIN = 4;
__device__ int32_t wave_lcs(const MapArgs &A, const int32_t *sid, int64_t n, uint64_t o, int32_t m, int32_t *Ts) {
    const uint32_t lane = threadIdx.x & (WAVE - 1);
    int32_t best = 0;
    for (int dir = 0; dir < 2; dir++) {
        int32_t t[LCS_TAILS / WAVE];
#pragma unroll
        for (int r = 0; r < LCS_TAILS / WAVE; r++) t[r] = 0;
        int32_t head = INT32_MAX;  // lane r: the first tail of chunk r (tails 64r ..), INT32_MAX if none
        int32_t nt = 0;
        bool in_lds = false;    // the tails are in Ts rather than the registers
        bool lds_only = false;  // more tails than the registers hold: they stay in Ts
        for (int64_t g0 = 0; g0 < n; g0 += WAVE) {
            const int64_t gi = g0 + (int64_t)lane;
            const int64_t x = dir == 0 ? gi : n - 1 - gi;
            uint32_t c = 0, first = 0;
            if (gi < n) {
                const int32_t v = sid[x];
                if (v >= 0) {
                    const int32_t sd = v & 0x3fffffff;
                    c = A.sd_occs[o + sd];
                    first = A.sd_first[o + sd];
                }
            }
            const uint32_t incl = wave_incl_scan(c);
            const uint32_t T = __shfl(incl, WAVE - 1);
            for (uint32_t v0 = 0; v0 < T; v0 += WAVE) {
                // value vi belongs to the first lane a with incl > vi (found by
                // binary lifting, every lane shuffling in step; idle lanes
                // search for the last value)
                const uint32_t vi = v0 + lane;
                const uint32_t vs = vi < T ? vi : T - 1;
                uint32_t a = 0;
#pragma unroll
                for (uint32_t step = WAVE / 2; step > 0; step >>= 1)
                    if ((uint32_t)__shfl(incl, (int)(a + step - 1)) <= vs) a += step;
                const uint32_t ca = __shfl(c, (int)a), fa = __shfl(first, (int)a);
                const uint32_t q = vs - ((uint32_t)__shfl(incl, (int)a) - ca);
                int32_t val = 0;
                if (vi < T) val = A.k_r[o + A.ps_v[o + (dir == 0 ? fa + ca - 1 - q : fa + q)]];
                const uint32_t cnt = T - v0 < (uint32_t)WAVE ? T - v0 : (uint32_t)WAVE;
                // the round's runs: lanes whose value starts a position's run
                const uint32_t ap = __shfl_up(a, 1);
                uint64_t rest = __ballot(lane < cnt && (lane == 0 || a != ap));
                while (rest) {
                    const uint32_t s0 = (uint32_t)__ffsll((unsigned long long)rest) - 1;
                    rest &= rest - 1;
                    const uint32_t e0 = rest ? (uint32_t)__ffsll((unsigned long long)rest) - 1 : cnt;
                    if (e0 - s0 < LCS_RUN_MIN) {
                        if (in_lds && !lds_only) {  // back to the registers
                            wave_sync();
#pragma unroll
                            for (int r = 0; r < LCS_TAILS / WAVE; r++)
                                t[r] = r * WAVE + (int32_t)lane < nt ? Ts[r * WAVE + lane] : 0;
                            head = (int32_t)lane < LCS_TAILS / WAVE && (int32_t)lane * WAVE < nt ? Ts[lane * WAVE]
                                                                                                : INT32_MAX;
                            in_lds = false;
                        }
                        for (uint32_t j = s0; j < e0; j++) {
                            const int32_t v = __builtin_amdgcn_readlane(val, (int)j);
                            if (!in_lds) {
                                // chunks whose first tail is below v, then the tails
                                // below v in the last of them
                                const uint32_t k = (uint32_t)__popcll(__ballot(head < v));
                                uint32_t pos = 0;
                                if (k > 0) {
                                    const int32_t ck = (int32_t)k - 1;
                                    pos = (uint32_t)ck * WAVE +
                                          (uint32_t)__popcll(__ballot(ck * WAVE + (int32_t)lane < nt && t[ck] < v));
                                }
                                if (pos < (uint32_t)LCS_TAILS) {
                                    if (pos == (uint32_t)nt) nt++;
                                    if (lane == pos % WAVE) t[pos / WAVE] = v;
                                    if (pos % WAVE == 0 && lane == pos / WAVE) head = v;
                                    continue;
                                }
                                // the registers are full: the tails move to Ts for good
#pragma unroll
                                for (int r = 0; r < LCS_TAILS / WAVE; r++) Ts[r * WAVE + lane] = t[r];
                                in_lds = lds_only = true;
                                wave_sync();
                            }
                            // the tails in Ts: v's lower bound among them, by
                            // ballots over 64 evenly spaced tails at a time
                            uint32_t lo = 0, hi = (uint32_t)nt;
                            while (hi - lo > (uint32_t)WAVE) {
                                const uint32_t step = (hi - lo + WAVE - 1) / WAVE;
                                const uint32_t at = lo + lane * step;
                                const uint32_t c = (uint32_t)__popcll(__ballot(at < hi && Ts[at] < v));
                                if (c == 0) {
                                    hi = lo;
                                    break;
                                }
                                const uint32_t nh = lo + c * step;
                                lo += (c - 1) * step + 1;
                                if (nh < hi) hi = nh;
                            }
                            const uint32_t pos = lo + (uint32_t)__popcll(__ballot(lo + lane < hi && Ts[lo + lane] < v));
                            wave_sync();  // every lane has read before the tail changes
                            if (lane == 0) Ts[pos] = v;
                            if (pos == (uint32_t)nt) nt++;
                            wave_sync();
                        }
                    } else {
                        if (!in_lds) {
#pragma unroll
                            for (int r = 0; r < LCS_TAILS / WAVE; r++) Ts[r * WAVE + lane] = t[r];
                            in_lds = true;
                            wave_sync();
                        }
                        if (lane == 0) Ts[nt] = INT32_MAX;  // an append slot (Ts holds LCS_TAILS + 1)
                        wave_sync();
                        const bool mine = lane >= s0 && lane < e0;
                        uint32_t lo = 0, hi = (uint32_t)nt;
                        if (mine)
                            while (lo < hi) {
                                const uint32_t mid = (lo + hi) >> 1;
                                if (Ts[mid] < val) lo = mid + 1; else hi = mid;
                            }
                        if (dir == 0) {
                            const bool grew = __ballot(mine && lo == (uint32_t)nt) != 0;
                            wave_sync();  // every search is done before a tail changes
                            if (mine) atomicMin(&Ts[lo], val);
                            if (grew) nt++;
                        } else {
                            int32_t pm = mine ? (int32_t)lo - (int32_t)(lane - s0) : INT32_MIN;
#pragma unroll
                            for (int d = 1; d < WAVE; d <<= 1) {
                                const int32_t o2 = __shfl_up(pm, d);
                                if ((int32_t)lane >= d && o2 > pm) pm = o2;
                            }
                            const int32_t place = mine ? pm + (int32_t)(lane - s0) : 0;
                            const int32_t last = __builtin_amdgcn_readlane(place, (int)(e0 - 1));
                            wave_sync();  // every search is done before a tail changes
                            if (mine) Ts[place] = val;
                            if (last + 1 > nt) nt = last + 1;
                        }
                        wave_sync();
                    }
                }
            }
        }
        best = nt > best ? nt : best;
    }
    return best;
}

// find_best_mapping (src/shmap/scoring.rs:296-359) for every bucket a
// refinement round selected, one workgroup per bucket.
//
// The span [l0, end) of the segment's sketch is read in order, as
// best_fixed_length and lcs_get_ppos_in_t do (:114-135, :189-291): each
// reference k-mer is matched by looking its hash up among the read's seeds
// (hash-ordered; binary search, the hashes staged in LDS when they fit), so a
// task reads its span once, coalesced, whatever the read's seed count, and a
// hash the index dropped under -M still matches, as p_ht.get does.
//
// Containment/Jaccard (best_fixed_length) run in parallel: window l covers
// [l, min(l+m, n)); a matched position x of seed s counts in window l exactly
// when fewer than occs(s) occurrences of s lie in [l, x), i.e. when the
// occs(s)-th previous occurrence P(x) is before l.  So x adds 1 to every l in
// [max(P(x)+1, x-m+1, 0), x] and its strand agreement to every l in
// [max(x-m+1, 0), x].  Both difference arrays share one int32 per position
// (intersection * 2^16 + strand agreement, |agreement| <= m < 2^15), so one
// scan gives both, and the first l of the greatest score wins, as the strict
// `score > best` of the sequential sweep picks.  P(x) comes from the index: the
// seed's hits are a position-ordered run, so P(x) is the hit occs(s) before x's
// own (found by binary search, and only for seeds with more hits than occs).
// A seed the index dropped (-M) has no run: such a task, and abs_pos (whose
// window edge depends on positions), LCS (whose patience LIS is sequential by
// nature) and reads with m >= 2^15, take the sequential sweep.  Arrays live in
// LDS while they fit, else in the workgroup's global scratch.
// LCS: the bucket_LCS instance (its patience sweep needs registers the other
// metrics' instance is better off without).
constexpr uint32_t REFINE_LDS_BYTES = 60 * 1024;  // the most a workgroup takes

// Takes LDS bytes (8-aligned) while the launch's `cap` bytes last; nullptr
// once they do not fit.
struct LdsCarve {
    uint8_t *base;
    uint32_t used, cap;
    __device__ void *take(uint64_t bytes) {
        const uint64_t b = (bytes + 7) & ~7ull;
        if ((uint64_t)used + b > cap) return nullptr;
        void *p = base + used;
        used += (uint32_t)b;
        return p;
    }
};

// Index of hash h among the u hash-ordered seed hashes sh, or -1.
__device__ __forceinline__ int32_t seed_find(const uint64_t *sh, int32_t u, uint64_t h) {
    int32_t lo = 0, hi = u;
    while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (sh[mid] < h) lo = mid + 1; else hi = mid;
    }
    return (lo < u && sh[lo] == h) ? lo : -1;
}

template <bool LCS>
__global__ void __launch_bounds__(WG) k_refine(MapArgs A, DevIndex I) {
    if (aborted(A.abort)) return;
    extern __shared__ uint8_t lds[];
    __shared__ uint32_t tmp[NWAVES + 1];
    __shared__ double s_bj[NWAVES];
    __shared__ int32_t s_bl[NWAVES];
    __shared__ int32_t s_lcs, s_dropped;
    __shared__ uint32_t s_pre[WG], s_lo[WG], s_occ[WG];
    __shared__ uint8_t s_str[WG];
    const int ap = A.abs_pos;
    // Rounds (see the task-list comment of MapArgs):
    //   0  each read's first first-pass candidate in sorted order -- the first
    //      pass always refines it (its threshold is still theta there);
    //   1  the other first-pass candidates whose bound reaches
    //      L1 = max(theta, J(first)): after the first candidate the first
    //      pass's threshold is at least L1, so lower bounds are never refined;
    //   2  for reads the first pass maps, what the second pass can reach and
    //      rounds 0-1 did not refine: bounds of at least
    //      L2 = J(first bucket above theta) * (1 - min_diff), which is at most
    //      the second pass's starting threshold best * (1 - min_diff).
    const uint32_t ntask = A.task_count[2 + A.refine_pass];
    for (uint32_t t = blockIdx.x; t < ntask; t += gridDim.x) {
        const uint2 task = A.sel[t];
        const int32_t lr = (int32_t)task.x, p = (int32_t)task.y;
        const int32_t rd = A.r0 + lr;
        const uint64_t base = A.bk_base[lr], o = A.koff[rd];
        const int32_t m = (int32_t)(A.koff[rd + 1] - o);
        const int32_t u = A.rd_u[rd], hl = A.rd_hl[rd];
        const int32_t len = (int32_t)(A.roff[rd + 1] - A.roff[rd]);
        uint32_t seg;
        int32_t b;
        decode_bucket(A, A.bs_g[base + p], hl, seg, b);
        const double sh = A.no_bucket_pruning ? 1.0 : A.bs_sh[base + p];
        if (A.metric == METRIC_BUCKET_SH) {
            if (threadIdx.x == 0) {
                const int32_t tl = A.bs_min[base + p], tr = A.bs_max[base + p];
                if (sh > A.theta) atomicMin(&A.rd_flag[rd], p);
                A.sc_j[base + p] = sh;
                A.sc_i[base + p] = A.bs_m[base + p];
                A.sc_same[base + p] = A.bs_c[base + p];
                A.sc_tl[base + p] = tl;
                A.sc_tr[base + p] = tr;
                A.sc_ssz[base + p] = tr - tl;
                A.sc_pend[base + p] = len - 1;
                A.sc_seg[base + p] = (int32_t)seg;
            }
            continue;
        }
        const int64_t ks = I.kstart[seg];
        const int64_t tlen = I.kstart[seg + 1] - ks;
        const int64_t bb = (int64_t)b * hl, be = (int64_t)(b + 2) * hl;
        const int64_t l0 = bb > 0 ? bb : 0;
        const int64_t end = be < tlen ? be : tlen;
        const int64_t n = end > l0 ? end - l0 : 0;
        int32_t *const wsid = A.g_sid + (uint64_t)blockIdx.x * A.sid_stride;
        int32_t *const wtr = A.g_tr + (uint64_t)blockIdx.x * A.sid_stride;
        bool parallel = n > 0 && !ap && !LCS && m < 32768;
        if (parallel) {
            LdsCarve C{lds, 0, A.refine_lds};
            int32_t *pk = (int32_t *)C.take((uint64_t)(n + 1) * 4);
            if (!pk) pk = wsid;  // sid_stride >= 2 * hl + 2 > n
            // seeds whose P(x) needs their hit run (more hits than occs), then
            // the list of them; in global scratch when LDS is short
            uint32_t *need = (uint32_t *)C.take((uint64_t)u * 4);
            if (!need) need = (uint32_t *)(A.g_diff + (uint64_t)blockIdx.x * A.diff_stride);
            // each seed's hash, and its kind: strand (bit 0), more hits than
            // occs (bit 1), dropped from the index (bit 2)
            uint64_t *s_h = (uint64_t *)C.take((uint64_t)u * 8);
            uint8_t *s_kind = (uint8_t *)C.take((uint64_t)u);
            for (int64_t x = threadIdx.x; x <= n; x += WG) pk[x] = 0;
            for (int32_t i = threadIdx.x; i < u; i += WG) {
                need[i] = 0;
                if (s_h) s_h[i] = A.sd_h[o + i];
                if (s_kind) {
                    const uint32_t cnt = A.sd_cnt[o + i], occs = A.sd_occs[o + i];
                    s_kind[i] = (uint8_t)(A.sd_strand[o + i] | (cnt > occs ? 2u : 0u) | (cnt == 0 ? 4u : 0u));
                }
            }
            if (threadIdx.x == 0) s_dropped = 0;
            __syncthreads();
            const uint64_t *seeds = s_h ? s_h : A.sd_h + o;
            // x counts in windows [max(P(x)+1, x-m+1, 0), x]; its strand
            // agreement in [max(x-m+1, 0), x]
            auto put = [&](int64_t x, int64_t P, int32_t cd) {
                int64_t lw = x - m + 1;
                if (lw < 0) lw = 0;
                const int64_t lc = P + 1 > lw ? P + 1 : lw;
                if (lc == lw) {
                    atomicAdd(&pk[lw], 65536 + cd);
                } else {
                    atomicAdd(&pk[lw], cd);
                    atomicAdd(&pk[lc], 65536);
                }
                atomicAdd(&pk[x + 1], -65536 - cd);
            };
            // a. the span in order: positions of seeds with at most occs hits
            //    (P(x) = -1) are placed at once, the others' seeds marked
            // (four positions a thread at a time, their loads in flight together)
            constexpr int PF = 4;
            for (int64_t x0 = 0; x0 < n; x0 += (int64_t)PF * WG) {
                uint64_t h[PF];
                uint32_t ts[PF];
#pragma unroll
                for (int k = 0; k < PF; k++) {
                    const int64_t x = x0 + k * WG + threadIdx.x;
                    h[k] = x < n ? I.t_h[ks + l0 + x] : 0ull;
                    ts[k] = x < n ? (uint32_t)I.t_s[ks + l0 + x] : 0u;
                }
#pragma unroll
                for (int k = 0; k < PF; k++) {
                    const int64_t x = x0 + k * WG + threadIdx.x;
                    if (x >= n) continue;
                    const int32_t s = seed_find(seeds, u, h[k]);
                    if (s < 0) continue;
                    uint32_t kind;
                    if (s_kind) {
                        kind = s_kind[s];
                    } else {
                        const uint32_t cnt = A.sd_cnt[o + s], occs = A.sd_occs[o + s];
                        kind = A.sd_strand[o + s] | (cnt > occs ? 2u : 0u) | (cnt == 0 ? 4u : 0u);
                    }
                    if (kind & 4u) {  // dropped from the index (-M): no hit run
                        s_dropped = 1;
                    } else if (kind & 2u) {
                        atomicAdd(&need[s], 1u);  // its positions in the span
                    } else {
                        put(x, -1, ts[k] == (kind & 1u) ? 1 : -1);
                    }
                }
            }
            __syncthreads();
            parallel = s_dropped == 0;
            if (parallel) {
                // b. the marked seeds, WG at a time: each thread finds where its
                //    seed's hits enter the span (it knows how many lie there),
                //    a scan lays the ranges end to end and the workgroup visits
                //    the hits, so a repeat k-mer with many hits in the span is
                //    not left to one thread.  A seed's hits are a
                //    position-ordered run of the index, so P(x) of hit j is hit
                //    j - occs.  The list holds (seed << 16 | positions): m and
                //    so u are below 2^15, a span below 2^16.
                uint32_t nneed = 0;
                for (int32_t s0 = 0; s0 < u; s0 += WG) {
                    const int32_t sd = s0 + (int32_t)threadIdx.x;
                    const uint32_t c = sd < u ? need[sd] : 0u;
                    uint32_t tot;
                    const uint32_t ex = wg_excl_scan(c ? 1u : 0u, tmp, &tot);
                    if (c) need[nneed + ex] = ((uint32_t)sd << 16) | c;
                    nneed += tot;
                    __syncthreads();
                }
                for (uint32_t s0 = 0; s0 < nneed; s0 += WG) {
                    const uint32_t li = s0 + threadIdx.x;
                    uint32_t lo = 0, nin = 0, sstr = 0, occ = 0;
                    if (li < nneed) {
                        const uint32_t e = need[li], sd = e >> 16;
                        const uint32_t cnt = A.sd_cnt[o + sd], hoff = A.sd_hoff[o + sd];
                        lo = hits_lower_bound(I, hoff, hoff + cnt, seg, l0, 0);
                        nin = e & 0xffffu;
                        sstr = A.sd_strand[o + sd];
                        occ = A.sd_occs[o + sd];
                    }
                    uint32_t tot;
                    const uint32_t ex = wg_excl_scan(nin, tmp, &tot);  // ends with a barrier
                    s_pre[threadIdx.x] = ex;
                    s_lo[threadIdx.x] = lo;
                    s_occ[threadIdx.x] = occ;
                    s_str[threadIdx.x] = (uint8_t)sstr;
                    __syncthreads();
                    for (uint32_t i = threadIdx.x; i < tot; i += WG) {
                        // the last list slot whose range starts at or before i
                        uint32_t a = 0, b2 = WG;
                        while (b2 - a > 1) {
                            const uint32_t mid = (a + b2) >> 1;
                            if (s_pre[mid] <= i) a = mid; else b2 = mid;
                        }
                        const uint32_t kk = i - s_pre[a];
                        const uint32_t j = s_lo[a] + kk;
                        const int64_t x = (int64_t)I.hit_g[j] - ks - l0;
                        const uint32_t oc = s_occ[a];
                        const int64_t P = kk >= oc ? (int64_t)I.hit_g[j - oc] - ks - l0 : -1;
                        put(x, P, s_str[a] == (I.hit_r[j] >> 31) ? 1 : -1);
                    }
                    __syncthreads();
                }
            }
        }
        if (parallel) {
            LdsCarve C{lds, 0, A.refine_lds};
            int32_t *pk = (int32_t *)C.take((uint64_t)(n + 1) * 4);
            if (!pk) pk = wsid;
            wg_scan_inplace(pk, n, tmp);
            // first window with the greatest score; pk[l] = I(l) * 2^16 + S(l)
            const bool jac = A.metric == METRIC_JACCARD;
            double bj = -1.0;
            int32_t bl = 0x7fffffff;
            for (int64_t l = threadIdx.x; l < n; l += WG) {
                const int32_t v = pk[l];
                const int32_t I2 = (v - (int32_t)(int16_t)(v & 0xffff)) >> 16;
                const int64_t r = l + m < n ? l + m : n;
                const double score = jac ? (double)I2 / (double)(m + (int32_t)(r - l) - I2) : (double)I2 / (double)m;
                if (score > bj) {
                    bj = score;
                    bl = (int32_t)l;
                }
            }
            for (int d = WAVE / 2; d > 0; d >>= 1) {
                const double oj = __shfl_xor(bj, d);
                const int32_t ol = __shfl_xor(bl, d);
                if (oj > bj || (oj == bj && ol < bl)) {
                    bj = oj;
                    bl = ol;
                }
            }
            if ((threadIdx.x & (WAVE - 1)) == 0) {
                s_bj[threadIdx.x / WAVE] = bj;
                s_bl[threadIdx.x / WAVE] = bl;
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                double J = s_bj[0];
                int32_t L = s_bl[0];
                for (int w = 1; w < NWAVES; w++)
                    if (s_bj[w] > J || (s_bj[w] == J && s_bl[w] < L)) {
                        J = s_bj[w];
                        L = s_bl[w];
                    }
                const int64_t r = L + m < n ? L + m : n;
                const int32_t tl = I.t_r[ks + l0 + L], tr = I.t_r[ks + l0 + r - 1];
                const int32_t v = pk[L];
                const int32_t S = (int32_t)(int16_t)(v & 0xffff);
                if (J > A.theta) atomicMin(&A.rd_flag[rd], p);
                A.sc_j[base + p] = J;
                A.sc_i[base + p] = (v - S) >> 16;
                A.sc_same[base + p] = S;
                A.sc_tl[base + p] = tl;
                A.sc_tr[base + p] = tr;
                A.sc_ssz[base + p] = tr - tl + 1;
                A.sc_pend[base + p] = len - A.k - 1;
                A.sc_seg[base + p] = (int32_t)seg;
            }
            __syncthreads();
            continue;
        }
        // sequential forms: LCS, abs_pos windows, seeds dropped by -M, m >= 2^15.
        // sid[x]: the seed matching span position x (| 0x40000000 when its
        // strand disagrees), -1 if none.
        LdsCarve C{lds, 0, A.refine_lds};
        // LCS: the tails (at most m, the read's distinct positions) and the
        // append slot; a sketch too large for them takes thread 0's sweep
        int32_t *Ts = LCS ? (int32_t *)C.take((uint64_t)((m > LCS_TAILS ? m : LCS_TAILS) + 2) * 4) : nullptr;
        int32_t *sid = (int32_t *)C.take((uint64_t)n * 4);
        if (!sid) sid = wsid;
        int32_t *tr_arr = nullptr, *diff = nullptr;
        if (!LCS) {
            tr_arr = (int32_t *)C.take((uint64_t)n * 4);
            if (!tr_arr) tr_arr = wtr;
            diff = (int32_t *)C.take((uint64_t)u * 4);
            if (!diff) diff = A.g_diff + (uint64_t)blockIdx.x * A.diff_stride;
        }
        uint64_t *s_h = (uint64_t *)C.take((uint64_t)u * 8);
        if (s_h)
            for (int32_t i = threadIdx.x; i < u; i += WG) s_h[i] = A.sd_h[o + i];
        if (!LCS)
            for (int32_t x = threadIdx.x; x < u; x += WG) diff[x] = (int32_t)A.sd_occs[o + x];
        __syncthreads();
        const uint64_t *seeds = s_h ? s_h : A.sd_h + o;
        for (int64_t x = threadIdx.x; x < n; x += WG) {
            const int64_t q = ks + l0 + x;
            const int32_t s = seed_find(seeds, u, I.t_h[q]);
            sid[x] = s < 0 ? -1 : (s | ((uint32_t)I.t_s[q] == (uint32_t)A.sd_strand[o + s] ? 0 : 0x40000000));
            if (!LCS) tr_arr[x] = I.t_r[q];
        }
        __syncthreads();
        if constexpr (LCS) {
            // one wave streams the LIS (s_lcs < 0: thread 0 sweeps it below)
            if (threadIdx.x < WAVE) {
                const int32_t best = Ts ? wave_lcs(A, sid, n, o, m, Ts) : -1;
                if (threadIdx.x == 0) s_lcs = best;
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double J = -1.0;
            int32_t inter = -1, same = -1, tl = -1, tr = -1, ssz = -1, pend = -1, mseg = -1;
            if (LCS) {
                // Patience LIS of the pmatches sequence and of its reverse; the
                // tails hold distinct read positions, so at most m of them.
                int32_t *tails = A.g_tails + (uint64_t)blockIdx.x * A.diff_stride;
                int32_t best = s_lcs > 0 ? s_lcs : 0;
                for (int dir = 0; s_lcs < 0 && dir < 2; dir++) {
                    int32_t nt = 0;
                    for (int64_t ii = 0; ii < n; ii++) {
                        const int64_t x = dir == 0 ? ii : n - 1 - ii;
                        const int32_t v = sid[x];
                        if (v < 0) continue;
                        const int32_t sd = v & 0x3fffffff;
                        const uint32_t first = A.sd_first[o + sd], occs = A.sd_occs[o + sd];
                        for (uint32_t q = 0; q < occs; q++) {
                            const uint32_t at = dir == 0 ? first + occs - 1 - q : first + q;
                            const int32_t val = A.k_r[o + A.ps_v[o + at]];
                            int32_t lo2 = -1, hi2 = nt;
                            while (lo2 + 1 < hi2) {
                                const int32_t mid = (lo2 + hi2) / 2;
                                if (tails[mid] < val) lo2 = mid; else hi2 = mid;
                            }
                            if (hi2 == nt) tails[nt++] = val; else tails[hi2] = val;
                        }
                    }
                    if (nt > best) best = nt;
                }
                J = (double)best / (double)m;
                inter = A.bs_m[base + p];
                same = A.bs_c[base + p];
                tl = A.bs_min[base + p];
                tr = A.bs_max[base + p];
                ssz = tr - tl;
                pend = len - 1;
                mseg = (int32_t)seg;
            } else {
                const bool jac = A.metric == METRIC_JACCARD;
                int64_t l = 0, r = 0, bl = -1, br = -1;
                int32_t I2 = 0, ss = 0;
                while (l < n) {
                    const int32_t tlr = tr_arr[l];
                    while (r < n) {
                        const bool in_w = ap ? tr_arr[r] < tlr + m : r < l + m;
                        if (!in_w) break;
                        const int32_t v = sid[r];
                        if (v >= 0) {
                            ss += (v & 0x40000000) ? -1 : 1;
                            if (--diff[v & 0x3fffffff] >= 0) I2++;
                        }
                        r++;
                    }
                    const int32_t s_kmers = (int32_t)(r - l);
                    const double score = jac ? (double)I2 / (double)(m + s_kmers - I2) : (double)I2 / (double)m;
                    if (l < r && score > J) {
                        J = score;
                        inter = I2;
                        same = ss;
                        bl = l;
                        br = r - 1;
                    }
                    const int32_t v = sid[l];
                    if (v >= 0) {
                        ss -= (v & 0x40000000) ? -1 : 1;
                        if (++diff[v & 0x3fffffff] >= 1) I2--;
                    }
                    l++;
                }
                if (bl >= 0) {
                    tl = tr_arr[bl];
                    tr = tr_arr[br];
                    ssz = tr - tl + 1;
                    pend = len - A.k - 1;
                    mseg = (int32_t)seg;
                }
            }
            if (J > A.theta) atomicMin(&A.rd_flag[rd], p);
            A.sc_j[base + p] = J;
            A.sc_i[base + p] = inter;
            A.sc_same[base + p] = same;
            A.sc_tl[base + p] = tl;
            A.sc_tr[base + p] = tr;
            A.sc_ssz[base + p] = ssz;
            A.sc_pend[base + p] = pend;
            A.sc_seg[base + p] = mseg;
        }
        __syncthreads();
    }
}

__device__ __forceinline__ double overlap_of(int32_t sa, int32_t la, int32_t ra, int32_t sb, int32_t lb,
                                             int32_t rb) {
    if (sa != sb) return -0.0;
    int32_t cap = (ra < rb ? ra : rb) - (la > lb ? la : lb);
    if (cap < 0) cap = 0;
    const int32_t cup = (ra > rb ? ra : rb) - (la < lb ? la : lb);
    return (double)cap / (double)cup;
}

// match_rest, both passes (src/shmap/scoring.rs:367-489 called from
// src/shmap/mod.rs:809-861), over the precomputed survival and scores; the
// refine memo's counters follow :38-109 (a Containment/Jaccard bucket refined in
// pass 1 is a memo hit in pass 2).  One wave per read.  Each pass is a scan in
// sorted order whose threshold rises only when a bucket is taken (its score
// exceeds the threshold -- and, in pass 2, it overlaps the best mapping by less
// than max_overlap -- so the taken score is above every earlier one): a chunk
// of 64 buckets is tested against the current threshold, the first taken lane
// settles every lane up to it, and the lanes after it are tested again with
// the raised threshold.  The result is staged in LDS and written by the wave.
__global__ void __launch_bounds__(WG) k_final(MapArgs A, shm_read_result *out) {
    if (aborted(A.abort)) return;
    static_assert(sizeof(shm_read_result) % 4 == 0 && sizeof(shm_read_result) / 4 <= WAVE, "result layout");
    constexpr uint32_t RES_DW = sizeof(shm_read_result) / 4;
    __shared__ shm_read_result s_res[NWAVES];
    const uint32_t lane = threadIdx.x & (WAVE - 1), w = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
    const bool nbp = A.no_bucket_pruning;
    const bool memoizable = A.metric == METRIC_CONTAINMENT || A.metric == METRIC_JACCARD;
    uint64_t u_sum = 0;  // distinct k-mers of this wave's reads (lane 0)
    for (int32_t lr = blockIdx.x * NWAVES + w; lr < A.RG; lr += gridDim.x * NWAVES) {
        const int32_t rd = A.r0 + lr;
        shm_read_result r = {};
        const uint64_t base = A.bk_base[lr], o = A.koff[rd];
        const int32_t m = (int32_t)(A.koff[rd + 1] - o);
        r.read_len = (int32_t)(A.roff[rd + 1] - A.roff[rd]);
        r.m = m;
        r.u = A.rd_u[rd];
        u_sum += (uint64_t)(uint32_t)r.u;
        r.S = A.rd_S[rd];
        r.possible_matches = A.rd_possible[rd];
        r.b_segm = r.b_b = r.b2_segm = r.b2_b = -1;
        r.intersection2 = -1;
        r.J2 = -1.0;
        if (!A.rd_ok[rd]) {
            r.unmappable_short = 1;
        } else {
            r.seed_matches = A.rd_seed_matches[rd];
            r.total_matches = A.rd_seed_matches[rd];
            r.max_seed_matches = A.rd_max_seed[rd];
            const int32_t nb = (int32_t)A.grp_nb[lr];
            const int32_t hl = A.rd_hl[rd];
            r.seeded_buckets = nb;
            int32_t refined = 0, memo = 0, fin = 0;
            // pass 1: threshold theta, then the best score so far
            double thr = A.theta;
            int32_t best = -1;
            for (int32_t c0 = 0; c0 < nb; c0 += WAVE) {
                const int32_t p = c0 + (int32_t)lane;
                uint8_t al = 0;
                double sh = 0.0, sc = 0.0;
                if (p < nb) {
                    al = A.bs_alive[base + p];
                    if (al) {
                        sh = A.bs_sh[base + p];
                        sc = A.sc_j[base + p];
                    }
                }
                uint32_t start = 0;
                for (;;) {
                    const bool ref = al && lane >= start && (nbp || !(sh < thr));
                    const uint64_t am = __ballot(ref && sc > thr), rm = __ballot(ref);
                    const uint32_t f = am ? (uint32_t)__builtin_ctzll(am) : (uint32_t)WAVE;
                    const uint64_t upto = f >= WAVE - 1 ? ~0ull : ((2ull << f) - 1ull);
                    refined += __popcll(rm & upto);
                    if (ref && lane <= f) A.bs_alive[base + p] = 2;  // refined in pass 1: a memo entry
                    if (!am) break;
                    fin++;
                    thr = __shfl(sc, (int)f);
                    best = c0 + (int32_t)f;
                    start = f + 1;
                    if (start >= WAVE) break;
                }
            }
            int32_t best2 = -1;
            if (best >= 0) {
                const uint64_t bb = base + best;
                double thr2 = A.sc_j[bb] * (1.0 - A.min_diff);
                const double thr2_0 = thr2;
                const int32_t fs = A.sc_seg[bb], fl = A.sc_tl[bb], fr = A.sc_tr[bb];
                for (int32_t c0 = 0; c0 < nb; c0 += WAVE) {
                    const int32_t p = c0 + (int32_t)lane;
                    uint8_t al = 0;
                    double sh = 0.0, sc = 0.0, ov = 1.0;
                    if (p < nb) {
                        al = A.bs_alive[base + p];
                        if (al) {
                            sh = A.bs_sh[base + p];
                            sc = A.sc_j[base + p];
                            if (sc > thr2_0)  // the threshold only rises: others never score above it
                                ov = overlap_of(A.sc_seg[base + p], A.sc_tl[base + p], A.sc_tr[base + p], fs, fl,
                                                fr);
                        }
                    }
                    const bool mh = memoizable && al == 2;
                    uint32_t start = 0;
                    for (;;) {
                        const bool ref = al && lane >= start && (nbp || !(sh < thr2));
                        const bool gt = ref && sc > thr2;
                        const uint64_t am = __ballot(gt && ov < A.max_overlap);
                        const uint32_t f = am ? (uint32_t)__builtin_ctzll(am) : (uint32_t)WAVE;
                        const uint64_t upto = f >= WAVE - 1 ? ~0ull : ((2ull << f) - 1ull);
                        memo += __popcll(__ballot(ref && mh) & upto);
                        refined += __popcll(__ballot(ref && !mh) & upto);
                        fin += __popcll(__ballot(gt) & upto);
                        if (!am) break;
                        thr2 = __shfl(sc, (int)f);
                        best2 = c0 + (int32_t)f;
                        start = f + 1;
                        if (start >= WAVE) break;
                    }
                }
            }
            r.refined_buckets = refined;
            r.refine_memo_hits = memo;
            r.final_buckets = fin;
            if (best >= 0) {
                const uint64_t bp = base + best;
                uint32_t seg;
                int32_t b;
                r.mapped = 1;
                r.p_start = 0;
                r.p_end = A.sc_pend[bp];
                r.strand = A.sc_same[bp] > 0 ? '+' : '-';
                r.segm_id = A.sc_seg[bp];
                r.t_l = A.sc_tl[bp];
                r.t_r = A.sc_tr[bp];
                r.s_sz = A.sc_ssz[bp];
                r.intersection = A.sc_i[bp];
                r.same_strand = A.sc_same[bp];
                r.J = A.sc_j[bp];
                r.sh = nbp ? 1.0 : A.bs_sh[bp];
                decode_bucket(A, A.bs_g[bp], hl, seg, b);
                r.b_segm = (int32_t)seg;
                r.b_b = b;
                if (best2 >= 0) {
                    const uint64_t bq = base + best2;
                    r.has_second = 1;
                    r.intersection2 = A.sc_i[bq];
                    r.J2 = A.sc_j[bq];
                    decode_bucket(A, A.bs_g[bq], hl, seg, b);
                    r.b2_segm = (int32_t)seg;
                    r.b2_b = b;
                }
            }
        }
        if (lane == 0) {
            s_res[w] = r;
            if (r.mapped) A.mapped_list[atomicAdd(A.mapped_count, 1u)] = (uint32_t)rd;
        }
        wave_sync();
        if (lane < RES_DW) ((uint32_t *)(out + rd))[lane] = ((const uint32_t *)&s_res[w])[lane];
        wave_sync();  // the slot is rewritten for the wave's next read
    }
    if (lane == 0 && u_sum) atomicAdd((unsigned long long *)A.u_total, (unsigned long long)u_sum);
}

}  // namespace shm

// ============================================================================
// Host launch wrappers
// ============================================================================
namespace shm {

// Workgroups of kernel fn (with `lds` bytes of dynamic LDS) that are resident
// on the current device at once.  A persistent grid-stride loop launched with
// more would leave the extra workgroups waiting for a second round.
static uint32_t resident_wgs(const void *fn, size_t lds) {
    struct Entry {
        const void *fn;
        size_t lds;
        int dev;
        uint32_t n;
    };
    // a fixed table (no allocation on any launch path); past its end the
    // count is recomputed each time
    constexpr int CACHE = 256;
    static std::mutex mu;
    static Entry cache[CACHE];
    static int used = 0;
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::lock_guard<std::mutex> lk(mu);
    for (int i = 0; i < used; i++)
        if (cache[i].fn == fn && cache[i].lds == lds && cache[i].dev == dev) return cache[i].n;
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, WG, lds) != hipSuccess || per_cu < 1) per_cu = 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 1;
    const uint32_t n = (uint32_t)(per_cu * cus);
    if (used < CACHE) cache[used++] = Entry{fn, lds, dev, n};
    return n;
}

// Grid of a persistent kernel: enough workgroups for `want` units of work, at
// most the resident count and max_grid.
static uint32_t persistent_grid(const void *fn, size_t lds, uint64_t want, uint32_t max_grid) {
    uint64_t g = want < 1 ? 1 : want;
    const uint32_t r = resident_wgs(fn, lds);
    if (g > r) g = r;
    if (g > max_grid) g = max_grid;
    return (uint32_t)g;
}

static inline uint32_t flat_grid(uint64_t n, uint32_t cap) {
    uint64_t g = (n + WG - 1) / WG;
    if (g < 1) g = 1;
    return (uint32_t)(g < cap ? g : cap);
}

// One workgroup's exclusive scan of a short array, out[n] = the total.
__global__ void __launch_bounds__(WG) k_scan_single(const uint64_t *in, uint64_t *out, const uint64_t *n_ptr,
                                                    uint64_t n_static) {
    const uint64_t n = n_ptr ? *n_ptr : n_static;
    __shared__ uint64_t tmp[NWAVES + 1];
    uint64_t carry = 0;
    for (uint64_t base = 0; base < n; base += WG) {
        const uint64_t i = base + threadIdx.x;
        const uint64_t v = i < n ? in[i] : 0;
        uint64_t tot;
        const uint64_t ex = wg_excl_scan64(v, tmp, &tot);
        if (i < n) out[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) out[n] = carry;
}

void launch_scan(const uint64_t *in, uint64_t *out, const uint64_t *n_dev, uint64_t n_hint,
                 uint64_t *part, hipStream_t st) {
    // a short array in one workgroup and one launch (a few WG-wide steps
    // cost less than two more launches)
    if (n_hint <= SCAN_SINGLE_MAX) {
        k_scan_single<<<1, WG, 0, st>>>(in, out, n_dev, n_hint);
        return;
    }
    // n_hint bounds the grid; the kernels read the true length from n_dev.
    const uint32_t G = flat_grid(n_hint, SCAN_PARTS);
    k_scan_reduce<<<G, WG, 0, st>>>(in, n_dev, n_hint, part);
    k_scan_top<<<1, WG, 0, st>>>(part, G);
    k_scan_down<<<G, WG, 0, st>>>(in, out, n_dev, n_hint, part);
}

void launch_unit_counts(SeqSet S, int32_t k, uint64_t *nunit, hipStream_t st) {
    k_unit_counts<<<flat_grid((uint64_t)S.n, 1024), WG, 0, st>>>(S, k, nunit);
}

size_t sketch_lds(int32_t k) { return (size_t)NWAVES * sketch_span(k); }

uint32_t sketch_lane_cap(uint64_t thres) {
    const double frac = ((double)thres + 1.0) / 18446744073709551616.0;  // keep probability
    const double lam = frac * LANE_W;
    const double cap = ceil(lam + 5.0 * sqrt(lam) + 4.0);
    return cap >= (double)LANE_W ? (uint32_t)LANE_W : (uint32_t)cap;
}

void launch_unit_seq(int32_t n, const uint64_t *unit_start, uint32_t *unit_seq, hipStream_t st,
                     const uint32_t *abort) {
    if (n > 0) k_unit_seq<<<flat_grid((uint64_t)n * WAVE, 4096), WG, 0, st>>>(n, unit_start, unit_seq, abort);
}

void launch_sketch(int pass, SeqSet S, int32_t k, uint64_t thres, const uint64_t *unit_start,
                   const uint32_t *unit_seq, uint64_t n_units, uint64_t *chunk_val, uint32_t *lane_cnt,
                   uint32_t *lane_pre, uint4 *stage, uint64_t *out_h, int32_t *out_r, uint8_t *out_s, uint32_t lcap,
                   uint32_t *ovf_list, uint32_t *ovf_count, uint64_t n_chunks, uint32_t max_grid, hipStream_t st,
                   const uint32_t *abort, const uint64_t *n_units_dev) {
    const size_t lds = sketch_lds(k);
    const uint64_t want = (n_chunks + NWAVES - 1) / NWAVES;
    if (pass == 0)
        k_sketch<0><<<persistent_grid((const void *)k_sketch<0>, lds, want, max_grid), WG, lds, st>>>(
            S, k, thres, unit_start, unit_seq, n_units, chunk_val, lane_cnt, lane_pre, stage, out_h, out_r, out_s,
            lcap, ovf_list, ovf_count, abort, n_units_dev);
    else
        k_sketch<1><<<persistent_grid((const void *)k_sketch<1>, lds, want, max_grid), WG, lds, st>>>(
            S, k, thres, unit_start, unit_seq, n_units, chunk_val, lane_cnt, lane_pre, stage, out_h, out_r, out_s,
            lcap, ovf_list, ovf_count, abort, n_units_dev);
}

void launch_sketch_compact(uint64_t n_chunks, const uint32_t *lane_cnt, const uint64_t *chunk_koff,
                           uint32_t lcap, const uint4 *stage, uint64_t *out_h, int32_t *out_r, uint8_t *out_s,
                           uint32_t max_grid, hipStream_t st, const uint32_t *abort, const uint64_t *n_units_dev) {
    const uint32_t g =
        persistent_grid((const void *)k_sketch_compact, 0, (n_chunks + NWAVES - 1) / NWAVES, max_grid);
    k_sketch_compact<<<g, WG, 0, st>>>(n_chunks, lane_cnt, chunk_koff, lcap, stage, out_h, out_r, out_s, abort,
                                       n_units_dev);
}

__global__ void __launch_bounds__(WG) k_fill_u32(uint32_t *p, uint64_t n, uint32_t v) {
    for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < n; i += (uint64_t)gridDim.x * WG) p[i] = v;
}
__global__ void __launch_bounds__(WG) k_fill_u64(uint64_t *p, uint64_t n, uint64_t v) {
    for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < n; i += (uint64_t)gridDim.x * WG) p[i] = v;
}

void launch_fill_u32(uint32_t *p, uint64_t n, uint32_t v, hipStream_t st) {
    if (n) k_fill_u32<<<flat_grid(n, 1024), WG, 0, st>>>(p, n, v);
}
void launch_fill_u64(uint64_t *p, uint64_t n, uint64_t v, hipStream_t st) {
    if (n) k_fill_u64<<<flat_grid(n, 1024), WG, 0, st>>>(p, n, v);
}

__global__ void k_gate(const uint64_t *v0, uint64_t c0, const uint64_t *v1, uint64_t c1, const uint64_t *v2,
                       uint64_t c2, const uint64_t *v3, uint64_t c3, uint32_t *abort) {
    if (threadIdx.x != 0) return;
    const bool over = (v0 && *v0 > c0) || (v1 && *v1 > c1) || (v2 && *v2 > c2) || (v3 && *v3 > c3);
    if (over) *abort = 1u;
}

void launch_gate(const uint64_t *v0, uint64_t c0, const uint64_t *v1, uint64_t c1, const uint64_t *v2, uint64_t c2,
                 const uint64_t *v3, uint64_t c3, uint32_t *abort, hipStream_t st) {
    k_gate<<<1, 64, 0, st>>>(v0, c0, v1, c1, v2, c2, v3, c3, abort);
}

// Every listed run of 32-bit words set to 0, in one launch (a batch's
// counters live in several buffers).
__global__ void __launch_bounds__(WG) k_zero_list(ZeroList z) {
    for (int32_t j = 0; j < z.count; j++)
        for (uint32_t i = threadIdx.x; i < z.n[j]; i += WG) z.p[j][i] = 0u;
}

void launch_zero_list(const ZeroList &z, hipStream_t st) { k_zero_list<<<1, WG, 0, st>>>(z); }

// dst[j] = the 8- or 4-byte value at src[j], one thread each: a batch's
// scattered sizes and counters gathered into one (page-locked) block.
__global__ void k_gather(GatherList g, uint64_t *dst) {
    const int32_t j = threadIdx.x;
    if (j < g.count) dst[j] = g.wide[j] ? *(const uint64_t *)g.src[j] : (uint64_t)*(const uint32_t *)g.src[j];
}

void launch_gather(const GatherList &g, uint64_t *dst, hipStream_t st) { k_gather<<<1, 64, 0, st>>>(g, dst); }

void launch_seq_koff(int32_t n, const uint64_t *unit_start, const uint64_t *chunk_koff, const uint32_t *lane_pre,
                     uint64_t *koff, uint64_t *max_m, hipStream_t st, const uint32_t *abort) {
    k_seq_koff<<<flat_grid((uint64_t)n + 1, 1024), WG, 0, st>>>(n, unit_start, chunk_koff, lane_pre, koff,
                                                                (unsigned long long *)max_m, abort);
}

void launch_max_i32(const int32_t *a, int32_t n, uint64_t *out, hipStream_t st) {
    k_max_i32<<<flat_grid((uint64_t)n, 1024), WG, 0, st>>>(a, n, (unsigned long long *)out);
}

void launch_fill_pairs(const uint64_t *t_h, uint64_t n, uint64_t *keys, uint32_t *vals, hipStream_t st) {
    k_fill_pairs<<<flat_grid(n, 8192), WG, 0, st>>>(t_h, n, keys, vals);
}

void launch_radix_pass(const uint64_t *kin, const uint32_t *vin, uint64_t *kout, uint32_t *vout, uint64_t n,
                       uint32_t shift, uint64_t *hist, uint64_t *offs, uint64_t *part, hipStream_t st) {
    const uint32_t ntiles = (uint32_t)((n + RS_TILE - 1) / RS_TILE);
    const uint32_t g = ntiles < 4096 ? ntiles : 4096;
    k_rs_hist<<<g, WG, 0, st>>>(kin, n, shift, hist, ntiles);
    launch_scan(hist, offs, nullptr, (uint64_t)ntiles * 256, part, st);
    k_rs_scatter<<<g, WG, 0, st>>>(kin, vin, kout, vout, n, shift, offs, ntiles);
}

void launch_mark_heads(const uint64_t *keys, uint64_t n, uint64_t *flag, hipStream_t st) {
    k_mark_heads<<<flat_grid(n, 8192), WG, 0, st>>>(keys, n, flag);
}

void launch_emit_keys(const uint64_t *keys, uint64_t n, const uint64_t *flag, const uint64_t *kid,
                      uint64_t *uk_h, uint64_t *uk_off, hipStream_t st) {
    k_emit_keys<<<flat_grid(n, 8192), WG, 0, st>>>(keys, n, flag, kid, uk_h, uk_off);
}

void launch_fill_hits(const uint32_t *vals, uint64_t n, const int32_t *t_r, const uint8_t *t_s,
                      const int64_t *kstart, int32_t nseg, uint32_t *hit_g, uint32_t *hit_r,
                      uint32_t *hit_seg, uint64_t *sbits, hipStream_t st) {
    k_fill_hits<<<flat_grid(n, 8192), WG, 0, st>>>(vals, n, t_r, t_s, kstart, nseg, hit_g, hit_r, hit_seg,
                                                    sbits);
}

void launch_narrow_prefix(const uint64_t *in, uint64_t n, uint32_t *out, hipStream_t st) {
    k_narrow_prefix<<<flat_grid(n + 1, 8192), WG, 0, st>>>(in, n, out);
}

void launch_shard_count(const uint64_t *uk_h, const uint64_t *uk_off, uint64_t nk, int32_t max_matches,
                        unsigned long long *shard_n, unsigned long long *stats, hipStream_t st) {
    k_shard_count<<<flat_grid(nk, 4096), WG, 0, st>>>(uk_h, uk_off, nk, max_matches, shard_n, stats);
}

void launch_insert(const uint64_t *uk_h, const uint64_t *uk_off, uint64_t nk, int32_t max_matches,
                   uint4 *table, ShardGeom geom, hipStream_t st) {
    k_insert<<<flat_grid(nk, 8192), WG, 0, st>>>(uk_h, uk_off, nk, max_matches, table, geom);
}

static inline uint32_t read_grid(int32_t n, uint32_t max_grid) {
    return (uint32_t)(n < 1 ? 1 : ((uint32_t)n < max_grid ? (uint32_t)n : max_grid));
}

static const size_t SORT_LDS = (size_t)SORT_CAP * 12;

uint32_t refine_lds_bytes(uint64_t max_hl, uint64_t max_m, bool lcs) {
    // the carves of k_refine, each rounded up to 8 bytes: span n <= 2 max_hl,
    // seeds u <= max_m.  Parallel form: pk (n+1) int32, need u uint32, s_h u
    // uint64, s_kind u bytes; sequential form: sid and tr n int32 each, diff u
    // int32, s_h; LCS: tails max(m, LCS_TAILS) + 2 int32, sid, s_h.
    const uint64_t n = 2 * max_hl, u = max_m;
    uint64_t need;
    if (lcs) {
        need = 4 * ((max_m > (uint64_t)LCS_TAILS ? max_m : (uint64_t)LCS_TAILS) + 2) + 4 * n + 8 * u;
    } else {
        const uint64_t par = 4 * (n + 1) + 4 * u + 8 * u + u;
        const uint64_t seq = 8 * n + 4 * u + 8 * u;
        need = par > seq ? par : seq;
    }
    need += 64;                          // alignment of the carves
    need = (need + 1023) & ~1023ull;     // whole KiB
    if (need < 4096) need = 4096;
    return (uint32_t)(need < REFINE_LDS_BYTES ? need : REFINE_LDS_BYTES);
}
static const size_t SEED_WAVE_LDS = (size_t)NWAVES * WAVE_SEED_LDS;

// One wrapper per per-read stage kernel so the driver can time each launch.
void launch_stage(Stage stage, const MapArgs &A, const DevIndex &I, shm_read_result *d_out,
                  uint64_t flat_n, uint32_t max_grid, hipStream_t st) {
#define PGRID(fn, lds, want) persistent_grid((const void *)(fn), (lds), (uint64_t)(want), max_grid)
    switch (stage) {
    case ST_SEED_WAVE:
        k_seed_wave<<<PGRID(k_seed_wave, SEED_WAVE_LDS, (A.R + NWAVES - 1) / NWAVES), WG, SEED_WAVE_LDS, st>>>(A, I);
        break;
    case ST_SEED_WG: k_seed_wg<<<read_grid(A.R, max_grid), WG, SORT_LDS, st>>>(A, I); break;
    case ST_TAB_INIT: k_tab_init<<<flat_grid((uint64_t)A.RG, 1024), WG, 0, st>>>(A); break;
    case ST_CONTRIB:
        // one wave per read (the grid covers every read: reads differ widely
        // in cost), a workgroup per read past one wave, then the reads past
        // that in chunks through global tables
        k_contrib_wave<<<(uint32_t)((A.RG + NWAVES - 1) / NWAVES), WG, 0, st>>>(A, I);
        k_contrib_wg<NWAVES><<<max_grid, WG, 0, st>>>(A, I, 0u, CWG_SPLIT);
        k_contrib_wg<CWG><<<(max_grid * NWAVES + CWG - 1) / CWG, CWG * WAVE, 0, st>>>(A, I, CWG_SPLIT, UINT32_MAX);
        k_contrib_over_scan<<<1, WG, 0, st>>>(A);
        k_contrib_over_clear<<<max_grid, WG, 0, st>>>(A);
        k_contrib_chunks<<<PGRID(k_contrib_chunks, 0, max_grid), WG, 0, st>>>(A, I);
        break;
    // per-read workgroup kernels with data-dependent cost: a grid above the
    // resident count lets the hardware backfill as reads finish
    case ST_EXTRACT:
        k_extract_wave<<<PGRID(k_extract_wave, 0, (A.RG + NWAVES - 1) / NWAVES), WG, 0, st>>>(A);
        k_extract<<<read_grid(A.RG, max_grid), WG, SORT_LDS, st>>>(A);
        break;
    case ST_PRUNE:
        k_prune<<<flat_grid(flat_n, 16384), WG, 0, st>>>(A, I);
        k_prune_ext<<<max_grid, WG, 0, st>>>(A, I);
        break;
    case ST_REFINE: {
        // three rounds: first candidates, other first-pass candidates, second pass
        MapArgs P = A;
        for (int32_t round = 0; round < 3; round++) {
            P.refine_pass = round;
            k_refine_select<<<max_grid, WG, 0, st>>>(P);
            if (A.metric == METRIC_BUCKET_LCS)
                k_refine<true><<<max_grid, WG, A.refine_lds, st>>>(P, I);
            else
                k_refine<false><<<max_grid, WG, A.refine_lds, st>>>(P, I);
        }
        break;
    }
    case ST_FINAL: k_final<<<read_grid((A.RG + NWAVES - 1) / NWAVES, max_grid), WG, 0, st>>>(A, d_out); break;
    default: break;
    }
#undef PGRID
}

__global__ void __launch_bounds__(WG) k_lookup(DevIndex I, const uint64_t *hs, uint32_t n, uint32_t *cnt,
                                               uint32_t *off) {
    for (uint32_t i = blockIdx.x * WG + threadIdx.x; i < n; i += gridDim.x * WG) cnt[i] = index_lookup(I, hs[i], off[i]);
}

// 2-bit packed bases (base i at bits 2*(i%4) of byte i/4; A C G T = 0 1 2 3)
// to ASCII letters, sixteen a thread, then the runs of other bytes (sorted,
// disjoint [start, end) pairs) written as 'N': every non-ACGT byte has the
// ntHash seed 0 (src/sketch.rs:73-88), so 'N' stands for any of them.
__global__ void __launch_bounds__(WG) k_unpack_bases(const uint32_t *packed, uint64_t nbases, uint8_t *out) {
    const uint64_t nq = (nbases + 15) / 16;
    for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < nq; i += (uint64_t)gridDim.x * WG) {
        const uint32_t v = packed[i];
        uint32_t w[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            // four 2-bit codes to four bytes, then the letter of each by a byte permute
            const uint32_t c = (v >> (8 * j)) & 0xffu;
            const uint32_t b = (c & 3u) | ((c & 0x0cu) << 6) | ((c & 0x30u) << 12) | ((c & 0xc0u) << 18);
            w[j] = __builtin_amdgcn_perm(0u, 0x54474341u, b);  // "ACGT"[code]
        }
        if ((i + 1) * 16 <= nbases) {
            *(uint4 *)(out + i * 16) = make_uint4(w[0], w[1], w[2], w[3]);
        } else {
            for (uint64_t q = i * 16; q < nbases; q++) out[q] = (uint8_t)(w[(q - i * 16) / 4] >> (8 * (q % 4)));
        }
    }
}

__global__ void __launch_bounds__(WG) k_apply_exceptions(const int64_t *exc, int64_t n_exc, uint8_t *out) {
    for (int64_t r = blockIdx.x; r < n_exc; r += gridDim.x)
        for (int64_t q = exc[2 * r] + threadIdx.x; q < exc[2 * r + 1]; q += WG) out[q] = 'N';
}

void launch_unpack_bases(const uint8_t *packed, uint64_t nbases, const int64_t *exc, int64_t n_exc, uint8_t *out,
                         hipStream_t st) {
    if (nbases > 0)
        k_unpack_bases<<<flat_grid((nbases + 15) / 16, 8192), WG, 0, st>>>((const uint32_t *)packed, nbases, out);
    if (n_exc > 0) k_apply_exceptions<<<(uint32_t)(n_exc < 4096 ? n_exc : 4096), WG, 0, st>>>(exc, n_exc, out);
}

void launch_lookup(const DevIndex &I, const uint64_t *hs, uint32_t n, uint32_t *cnt, uint32_t *off,
                   hipStream_t st) {
    k_lookup<<<flat_grid(n, 1024), WG, 0, st>>>(I, hs, n, cnt, off);
}

// Consistency of an index read from a file, after the host has checked that
// every hit's k-mer id lies inside its segment and every slot's run inside the
// hit arrays (run lengths summing to at most the hit count): each hit carries
// its k-mer's position and strand, and each slot's run is strictly increasing
// in (segment, k-mer id), which the lower-bound searches over runs rely on.
// Violations are counted into *bad.
__global__ void __launch_bounds__(WG) k_validate_hits(DevIndex I, unsigned long long *bad) {
    uint32_t nb = 0;
    const uint64_t N = (uint64_t)I.n_kmers;
    for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < N; i += (uint64_t)gridDim.x * WG) {
        const uint32_t g = I.hit_g[i];
        nb += I.hit_r[i] != ((uint32_t)I.t_r[g] | ((uint32_t)I.t_s[g] << 31));
    }
    if (nb) atomicAdd(bad, (unsigned long long)nb);
}

__global__ void __launch_bounds__(WG) k_validate_runs(DevIndex I, uint64_t slots, unsigned long long *bad) {
    uint32_t nb = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < slots; i += (uint64_t)gridDim.x * WG) {
        const uint4 s = I.table[i];
        for (uint64_t j = (uint64_t)s.z + 1; s.w != 0 && j < (uint64_t)s.z + s.w; j++) {
            const uint32_t sa = I.hit_seg[j - 1], sb = I.hit_seg[j];
            nb += sa > sb || (sa == sb && I.hit_g[j - 1] >= I.hit_g[j]);
        }
    }
    if (nb) atomicAdd(bad, (unsigned long long)nb);
}

void launch_validate_index(const DevIndex &I, uint64_t slots, unsigned long long *bad, hipStream_t st) {
    if (I.n_kmers > 0) k_validate_hits<<<flat_grid((uint64_t)I.n_kmers, 4096), WG, 0, st>>>(I, bad);
    if (slots > 0) k_validate_runs<<<flat_grid(slots, 4096), WG, 0, st>>>(I, slots, bad);
}

}  // namespace shm
