// Device-side building blocks for the gfx950 mapping path: the resident index
// view, ntHash rolling, wave/workgroup scans and a workgroup bitonic sort.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace shm {

constexpr int WG = 256;      // threads per workgroup (4 wave64s)
constexpr int WAVE = 64;
constexpr int NWAVES = WG / WAVE;
constexpr uint32_t EMPTY_KEY = 0xffffffffu;

// Resident reference index (built once on the device).
//   t_h/t_r/t_s   reference sketch of every segment, concatenated; segment s
//                 owns global k-mer ids [kstart[s], kstart[s+1])
//   hit_*         every indexed k-mer occurrence grouped by hash, each group in
//                 (segm_id, r) order: global k-mer id, r | strand << 31, segment
//   strand_pre    exclusive prefix count of reverse-strand hits over hit_*
//   table         8 shards (by h & 7) of open-addressing slots
//                 {key lo, key hi, hit offset, count}; count 0 = empty slot
struct DevIndex {
    const uint64_t *t_h;
    const int32_t *t_r;
    const uint8_t *t_s;
    const int64_t *kstart;
    int32_t nseg;
    int64_t n_kmers;
    const uint32_t *hit_g;
    const uint32_t *hit_r;
    const uint32_t *hit_seg;
    const uint32_t *strand_pre;
    const uint4 *table;
    uint64_t shard_base[8];
    uint32_t shard_bits[8];
};

__device__ __forceinline__ uint64_t rotl64(uint64_t x, uint32_t n) {
    n &= 63u;
    return n ? (x << n) | (x >> (64u - n)) : x;
}
__device__ __forceinline__ uint64_t rotr64(uint64_t x, uint32_t n) {
    n &= 63u;
    return n ? (x >> n) | (x << (64u - n)) : x;
}

// Rotations by one bit as two funnel shifts (v_alignbit_b32) on the halves.
__device__ __forceinline__ uint64_t rotl1(uint64_t x) {
    const uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
    const uint32_t nhi = __builtin_amdgcn_alignbit(hi, lo, 31);  // hi << 1 | lo >> 31
    const uint32_t nlo = __builtin_amdgcn_alignbit(lo, hi, 31);  // lo << 1 | hi >> 31
    return ((uint64_t)nhi << 32) | nlo;
}
__device__ __forceinline__ uint64_t rotr1(uint64_t x) {
    const uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
    const uint32_t nhi = __builtin_amdgcn_alignbit(lo, hi, 1);  // hi >> 1 | lo << 31
    const uint32_t nlo = __builtin_amdgcn_alignbit(hi, lo, 1);  // lo >> 1 | hi << 31
    return ((uint64_t)nhi << 32) | nlo;
}

// ntHash seed of a base, forward strand; 0 for anything but ACGTacgt.
__device__ __forceinline__ uint64_t nt_fw(uint8_t c) {
    const uint8_t l = c | 0x20;
    return l == 'a' ? 0x3c8bfbb395c60474ull
         : l == 'c' ? 0x3193c18562a02b4cull
         : l == 'g' ? 0x20323ed082572324ull
         : l == 't' ? 0x295549f54be24456ull : 0ull;
}
// Reverse-complement strand seed: the forward seed of the complement.
__device__ __forceinline__ uint64_t nt_rc(uint8_t c) {
    const uint8_t l = c | 0x20;
    return l == 'a' ? 0x295549f54be24456ull
         : l == 'c' ? 0x20323ed082572324ull
         : l == 'g' ? 0x3193c18562a02b4cull
         : l == 't' ? 0x3c8bfbb395c60474ull : 0ull;
}

__device__ __forceinline__ uint32_t table_pos(uint64_t h, uint32_t bits) {
    return (uint32_t)(((h >> 3) * 0x9E3779B97F4A7C15ull) >> (64u - bits));
}

// SketchIndex::count + hit offset (src/index.rs:192-213).
__device__ __forceinline__ uint32_t index_lookup(const DevIndex &I, uint64_t h, uint32_t &off) {
    const uint32_t sh = (uint32_t)(h & 7u);
    const uint32_t bits = I.shard_bits[sh];
    const uint32_t mask = (1u << bits) - 1u;
    const uint4 *T = I.table + I.shard_base[sh];
    uint32_t pos = table_pos(h, bits);
    const uint32_t lo = (uint32_t)h, hi = (uint32_t)(h >> 32);
    for (;;) {
        const uint4 s = T[pos];
        if (s.w == 0) {
            off = 0;
            return 0;
        }
        if (s.x == lo && s.y == hi) {
            off = s.z;
            return s.w;
        }
        pos = (pos + 1u) & mask;
    }
}

// index_lookup of up to four hashes per lane (act[k]) probed in lockstep, so
// the table loads of all four are in flight together.
__device__ __forceinline__ void index_lookup4(const DevIndex &I, const uint64_t (&h)[4], const bool (&act)[4],
                                              uint32_t (&cnt)[4], uint32_t (&off)[4]) {
    const uint4 *T[4];
    uint32_t pos[4], mask[4];
    bool live[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const uint32_t sh = (uint32_t)(h[k] & 7u);
        const uint32_t bits = I.shard_bits[sh];
        mask[k] = (1u << bits) - 1u;
        T[k] = I.table + I.shard_base[sh];
        pos[k] = table_pos(h[k], bits);
        live[k] = act[k];
        cnt[k] = 0;
        off[k] = 0;
    }
    for (;;) {
        uint4 s[4];
#pragma unroll
        for (int k = 0; k < 4; k++) s[k] = live[k] ? T[k][pos[k]] : make_uint4(0, 0, 0, 0);
        bool any = false;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (!live[k]) continue;
            if (s[k].w == 0) {
                live[k] = false;
            } else if (s[k].x == (uint32_t)h[k] && s[k].y == (uint32_t)(h[k] >> 32)) {
                cnt[k] = s[k].w;
                off[k] = s[k].z;
                live[k] = false;
            } else {
                pos[k] = (pos[k] + 1u) & mask[k];
                any = true;
            }
        }
        if (!any) break;
    }
}

// ---- wave / workgroup scans --------------------------------------------------

// Orders the wave's LDS accesses around a point every lane reaches.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const int lane = threadIdx.x & (WAVE - 1);
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        const uint32_t t = __shfl_up(v, d, WAVE);
        if (lane >= d) v += t;
    }
    return v;
}

__device__ __forceinline__ uint64_t wave_incl_scan64(uint64_t v) {
    const int lane = threadIdx.x & (WAVE - 1);
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        const uint64_t t = __shfl_up(v, d, WAVE);
        if (lane >= d) v += t;
    }
    return v;
}

// Exclusive scan over the workgroup; returns this thread's exclusive prefix and
// writes the total to *total.  `tmp` is NWAVES+1 words of LDS.
__device__ __forceinline__ uint32_t wg_excl_scan(uint32_t v, uint32_t *tmp, uint32_t *total) {
    const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
    const uint32_t inc = wave_incl_scan(v);
    if (lane == WAVE - 1) tmp[w] = inc;
    __syncthreads();
    uint32_t base = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < NWAVES; i++) {
        const uint32_t t = tmp[i];
        if (i < w) base += t;
        tot += t;
    }
    __syncthreads();
    *total = tot;
    return base + inc - v;
}

__device__ __forceinline__ uint64_t wg_excl_scan64(uint64_t v, uint64_t *tmp, uint64_t *total) {
    const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
    const uint64_t inc = wave_incl_scan64(v);
    if (lane == WAVE - 1) tmp[w] = inc;
    __syncthreads();
    uint64_t base = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < NWAVES; i++) {
        const uint64_t t = tmp[i];
        if (i < w) base += t;
        tot += t;
    }
    __syncthreads();
    *total = tot;
    return base + inc - v;
}

// ---- workgroup bitonic sort of (key, val) pairs ------------------------------
//
// Sorts n pairs ascending by (key, val).  All merge steps run ascending (the
// first step of each merge compares mirrored positions), so positions >= n act
// as +infinity and are simply skipped: no padding storage is needed, and the
// same loop runs on LDS copies (n <= cap) or directly on the global arrays.

template <typename K, typename V>
__device__ __forceinline__ void cswap(K *k, V *v, uint32_t a, uint32_t b) {
    const K ka = k[a], kb = k[b];
    const V va = v[a], vb = v[b];
    if (kb < ka || (kb == ka && vb < va)) {
        k[a] = kb;
        k[b] = ka;
        v[a] = vb;
        v[b] = va;
    }
}

template <typename K, typename V>
__device__ void bitonic_core(K *k, V *v, uint32_t n) {
    uint32_t N = 1;
    while (N < n) N <<= 1;
    for (uint32_t size = 2; size <= N; size <<= 1) {
        const uint32_t half = size >> 1;
        for (uint32_t i = threadIdx.x; i < (N >> 1); i += WG) {
            const uint32_t blk = i >> (31 - __builtin_clz(half)), p = i & (half - 1);  // half: a power of two
            const uint32_t a = blk * size + p, b = blk * size + size - 1 - p;
            if (b < n) cswap(k, v, a, b);
        }
        __syncthreads();
        for (uint32_t stride = half >> 1; stride > 0; stride >>= 1) {
            for (uint32_t i = threadIdx.x; i < (N >> 1); i += WG) {
                const uint32_t a = ((i & ~(stride - 1)) << 1) | (i & (stride - 1)), b = a + stride;  // stride: a power of two
                if (b < n) cswap(k, v, a, b);
            }
            __syncthreads();
        }
    }
}

// Key-only variant (distinct keys), in place; caller synchronises before use.
template <typename K>
__device__ __forceinline__ void cswap_k(K *k, uint32_t a, uint32_t b) {
    const K ka = k[a], kb = k[b];
    if (kb < ka) {
        k[a] = kb;
        k[b] = ka;
    }
}

template <typename K>
__device__ void bitonic_keys(K *k, uint32_t n) {
    uint32_t N = 1;
    while (N < n) N <<= 1;
    for (uint32_t size = 2; size <= N; size <<= 1) {
        const uint32_t half = size >> 1;
        for (uint32_t i = threadIdx.x; i < (N >> 1); i += WG) {
            const uint32_t blk = i >> (31 - __builtin_clz(half)), p = i & (half - 1);  // half: a power of two
            const uint32_t a = blk * size + p, b = blk * size + size - 1 - p;
            if (b < n) cswap_k(k, a, b);
        }
        __syncthreads();
        for (uint32_t stride = half >> 1; stride > 0; stride >>= 1) {
            for (uint32_t i = threadIdx.x; i < (N >> 1); i += WG) {
                const uint32_t a = ((i & ~(stride - 1)) << 1) | (i & (stride - 1)), b = a + stride;  // stride: a power of two
                if (b < n) cswap_k(k, a, b);
            }
            __syncthreads();
        }
    }
}

// Register-resident bitonic sort of one wave's 64*E elements, ascending by
// (key, val) (by key alone when !PAIRS); element x = lane*E + e is k[e], v[e].
// Every step pairs x with x ^ m (m = size-1 for the first step of a merge,
// whose partner mirrors x inside its block, then the strides size/4 .. 1);
// the lower index keeps the smaller element.  Masks below E pair registers of
// one lane; larger ones pair lanes through shuffles.
template <bool PAIRS>
__device__ __forceinline__ bool sort_less(uint64_t ka, uint32_t va, uint64_t kb, uint32_t vb) {
    return ka < kb || (PAIRS && ka == kb && va < vb);
}

template <int E, bool PAIRS, int M>
__device__ __forceinline__ void sort_reg_step(uint64_t (&k)[E], uint32_t (&v)[E]) {
    // in-register pairs (e, e ^ M), M < E; the lower register keeps the min
    constexpr int HB = (M & (M + 1)) == 0 ? (M + 1) / 2 : M;  // highest bit of M (2^j - 1 or 2^j)
#pragma unroll
    for (int e = 0; e < E; e++) {
        if (e & HB) continue;
        const int p = e ^ M;
        const bool t = sort_less<PAIRS>(k[p], v[p], k[e], v[e]);
        const uint64_t ke = k[e], kp = k[p];
        const uint32_t ve = v[e], vp = v[p];
        k[e] = t ? kp : ke;
        k[p] = t ? ke : kp;
        if (PAIRS) {
            v[e] = t ? vp : ve;
            v[p] = t ? ve : vp;
        }
    }
}

template <int E, bool PAIRS, int ST>
__device__ __forceinline__ void sort_reg_strides(uint64_t (&k)[E], uint32_t (&v)[E]) {
    if constexpr (ST > 0) {
        sort_reg_step<E, PAIRS, ST>(k, v);
        sort_reg_strides<E, PAIRS, ST / 2>(k, v);
    }
}

template <int E, bool PAIRS, int SIZE>
__device__ __forceinline__ void sort_reg_merges(uint64_t (&k)[E], uint32_t (&v)[E]) {
    if constexpr (SIZE <= E) {
        sort_reg_step<E, PAIRS, SIZE - 1>(k, v);
        sort_reg_strides<E, PAIRS, SIZE / 4>(k, v);
        sort_reg_merges<E, PAIRS, SIZE * 2>(k, v);
    }
}

// Cross-lane step: this lane's register e meets register (MIRROR ? E-1-e : e)
// of the lane whose byte address is pa; lower says whether this lane holds the
// lower indices.  Elements are distinct or identical, so the partner is taken
// exactly when (partner < mine) == lower.  Registers go in groups of at most
// four (a mirror group holds e and E-1-e together, so no group reads a
// register an earlier group has replaced).
template <int E, bool PAIRS, bool MIRROR>
__device__ __forceinline__ void sort_lane_step(uint64_t (&k)[E], uint32_t (&v)[E], int pa, bool lower) {
    constexpr int G = E < 4 ? E : 4;
#pragma unroll
    for (int j = 0; j < E / G; j++) {
        int es[G];
#pragma unroll
        for (int i = 0; i < G; i++)
            es[i] = !MIRROR ? j * G + i : (i < G / 2 ? j * (G / 2) + i : E - 1 - (j * (G / 2) + i - G / 2));
        if (MIRROR && G == 1) es[0] = 0;
        uint32_t pl[G], ph[G], pv[G];
#pragma unroll
        for (int i = 0; i < G; i++) {
            const int q = MIRROR ? E - 1 - es[i] : es[i];
            pl[i] = (uint32_t)__builtin_amdgcn_ds_bpermute(pa, (int)(uint32_t)k[q]);
            ph[i] = (uint32_t)__builtin_amdgcn_ds_bpermute(pa, (int)(uint32_t)(k[q] >> 32));
            pv[i] = PAIRS ? (uint32_t)__builtin_amdgcn_ds_bpermute(pa, (int)v[q]) : 0u;
        }
#pragma unroll
        for (int i = 0; i < G; i++) {
            const int e = es[i];
            const uint64_t pk = ((uint64_t)ph[i] << 32) | pl[i];
            const bool take = sort_less<PAIRS>(pk, pv[i], k[e], v[e]) == lower;
            k[e] = take ? pk : k[e];
            if (PAIRS) v[e] = take ? pv[i] : v[e];
        }
    }
}

template <int E, bool PAIRS>
__device__ void wave_sort_regs(uint64_t (&k)[E], uint32_t (&v)[E]) {
    const uint32_t lane = threadIdx.x & (WAVE - 1);
    sort_reg_merges<E, PAIRS, 2>(k, v);  // blocks of up to E elements: within each lane
#pragma unroll 1
    for (uint32_t lsz = 2; lsz <= WAVE; lsz <<= 1) {  // blocks of lsz lanes
        sort_lane_step<E, PAIRS, true>(k, v, (int)((lane ^ (lsz - 1)) << 2), (lane & (lsz >> 1)) == 0);
#pragma unroll 1
        for (uint32_t ls = lsz >> 2; ls > 0; ls >>= 1)
            sort_lane_step<E, PAIRS, false>(k, v, (int)((lane ^ ls) << 2), (lane & ls) == 0);
        sort_reg_strides<E, PAIRS, E / 2>(k, v);
    }
}

// One wave sorts n <= 64*E elements held in registers between load(x, key, val)
// and store(x, key, val) over x < n (lane-major: lane l holds x = l*E .. l*E+E-1).
template <int E, bool PAIRS, typename LoadF, typename StoreF>
__device__ __forceinline__ void wave_sort_e(uint32_t n, LoadF load, StoreF store) {
    const uint32_t lane = threadIdx.x & (WAVE - 1);
    uint64_t k[E];
    uint32_t v[E];
#pragma unroll
    for (int e = 0; e < E; e++) {
        const uint32_t x = lane * E + (uint32_t)e;
        k[e] = ~0ull;
        v[e] = ~0u;
        if (x < n) load(x, k[e], v[e]);
    }
    wave_sort_regs<E, PAIRS>(k, v);
#pragma unroll
    for (int e = 0; e < E; e++) {
        const uint32_t x = lane * E + (uint32_t)e;
        if (x < n) store(x, k[e], v[e]);
    }
}

// n <= 8*WAVE: in registers; returns false (nothing done) for larger n.
template <bool PAIRS, typename LoadF, typename StoreF>
__device__ bool wave_sort_n(uint32_t n, LoadF load, StoreF store) {
    if (n <= WAVE) wave_sort_e<1, PAIRS>(n, load, store);
    else if (n <= 2 * WAVE) wave_sort_e<2, PAIRS>(n, load, store);
    else if (n <= 4 * WAVE) wave_sort_e<4, PAIRS>(n, load, store);
    else if (n <= 8 * WAVE) wave_sort_e<8, PAIRS>(n, load, store);
    else return false;
    return true;
}

// One wave sorts n (key, val) pairs in LDS, ascending by (key, val).
template <typename K, typename V>
__device__ void wave_bitonic(K *k, V *v, uint32_t n) {
    const uint32_t lane = threadIdx.x & (WAVE - 1);
    uint32_t N = 1;
    while (N < n) N <<= 1;
    for (uint32_t size = 2; size <= N; size <<= 1) {
        const uint32_t half = size >> 1;
        for (uint32_t i = lane; i < (N >> 1); i += WAVE) {
            const uint32_t blk = i >> (31 - __builtin_clz(half)), p = i & (half - 1);  // half: a power of two
            const uint32_t a = blk * size + p, b = blk * size + size - 1 - p;
            if (b < n) cswap(k, v, a, b);
        }
        wave_sync();
        for (uint32_t stride = half >> 1; stride > 0; stride >>= 1) {
            for (uint32_t i = lane; i < (N >> 1); i += WAVE) {
                const uint32_t a = ((i & ~(stride - 1)) << 1) | (i & (stride - 1)), b = a + stride;  // stride: a power of two
                if (b < n) cswap(k, v, a, b);
            }
            wave_sync();
        }
    }
}

// One wave sorts n distinct keys in LDS, ascending.
template <typename K>
__device__ void wave_bitonic_keys(K *k, uint32_t n) {
    const uint32_t lane = threadIdx.x & (WAVE - 1);
    uint32_t N = 1;
    while (N < n) N <<= 1;
    for (uint32_t size = 2; size <= N; size <<= 1) {
        const uint32_t half = size >> 1;
        for (uint32_t i = lane; i < (N >> 1); i += WAVE) {
            const uint32_t blk = i >> (31 - __builtin_clz(half)), p = i & (half - 1);  // half: a power of two
            const uint32_t a = blk * size + p, b = blk * size + size - 1 - p;
            if (b < n) cswap_k(k, a, b);
        }
        wave_sync();
        for (uint32_t stride = half >> 1; stride > 0; stride >>= 1) {
            for (uint32_t i = lane; i < (N >> 1); i += WAVE) {
                const uint32_t a = ((i & ~(stride - 1)) << 1) | (i & (stride - 1)), b = a + stride;  // stride: a power of two
                if (b < n) cswap_k(k, a, b);
            }
            wave_sync();
        }
    }
}

// Cross-wave step of a workgroup's register sort (x = threadIdx.x * E + e):
// every element meets x ^ mask through LDS (element-major, so a step's reads
// and writes are conflict-free); HB is mask's highest bit, and the lower
// index keeps the smaller element.  mask >= WAVE * E.
template <int E, bool PAIRS>
__device__ __forceinline__ void wg_cross_step(uint64_t (&k)[E], uint32_t (&v)[E], uint64_t *sk, uint32_t *sv,
                                              uint32_t mask, uint32_t hb) {
    const uint32_t t = threadIdx.x;
#pragma unroll
    for (int e = 0; e < E; e++) {
        sk[(uint32_t)e * WG + t] = k[e];
        if (PAIRS) sv[(uint32_t)e * WG + t] = v[e];
    }
    __syncthreads();
    const uint32_t pt = t ^ (mask / (uint32_t)E);
    const bool lower = (t & (hb / (uint32_t)E)) == 0;
#pragma unroll
    for (int e = 0; e < E; e++) {
        const uint32_t pe = (uint32_t)e ^ (mask & (uint32_t)(E - 1));
        const uint64_t pk = sk[pe * WG + pt];
        const uint32_t pv = PAIRS ? sv[pe * WG + pt] : 0u;
        const bool take = sort_less<PAIRS>(pk, pv, k[e], v[e]) == lower;
        k[e] = take ? pk : k[e];
        if (PAIRS) v[e] = take ? pv : v[e];
    }
    __syncthreads();
}

// The strides below one wave's block (WAVE * E) of a bitonic merge, within the
// wave: lane steps for strides of E and more, then register steps.
template <int E, bool PAIRS>
__device__ __forceinline__ void wave_merge_tail(uint64_t (&k)[E], uint32_t (&v)[E]) {
    const uint32_t lane = threadIdx.x & (WAVE - 1);
#pragma unroll 1
    for (uint32_t ls = WAVE >> 1; ls > 0; ls >>= 1)
        sort_lane_step<E, PAIRS, false>(k, v, (int)((lane ^ ls) << 2), (lane & ls) == 0);
    sort_reg_strides<E, PAIRS, E / 2>(k, v);
}

// A workgroup sorts n <= WG * E elements held in registers between load(x,
// key, val) and store(x, key, val) over x < n (thread-major: thread t holds x
// = t*E .. t*E+E-1): each wave sorts its block in registers, then bitonic
// merges of 2 and 4 waves' blocks take their cross-wave steps through LDS
// (sk, sv: WG * E entries) and the rest within each wave.  Ascending by (key,
// val); ends with a barrier.
template <int E, bool PAIRS, typename LoadF, typename StoreF>
__device__ __forceinline__ void wg_sort_e(uint32_t n, LoadF load, StoreF store, uint64_t *sk, uint32_t *sv) {
    uint64_t k[E];
    uint32_t v[E];
#pragma unroll
    for (int e = 0; e < E; e++) {
        const uint32_t x = threadIdx.x * E + (uint32_t)e;
        k[e] = ~0ull;
        v[e] = ~0u;
        if (x < n) load(x, k[e], v[e]);
    }
    wave_sort_regs<E, PAIRS>(k, v);
    constexpr uint32_t WB = WAVE * E;  // one wave's block
#pragma unroll 1
    for (uint32_t size = 2 * WB; size <= (uint32_t)WG * E; size <<= 1) {
        wg_cross_step<E, PAIRS>(k, v, sk, sv, size - 1, size >> 1);
#pragma unroll 1
        for (uint32_t st = size >> 2; st >= WB; st >>= 1) wg_cross_step<E, PAIRS>(k, v, sk, sv, st, st);
        wave_merge_tail<E, PAIRS>(k, v);
    }
#pragma unroll
    for (int e = 0; e < E; e++) {
        const uint32_t x = threadIdx.x * E + (uint32_t)e;
        if (x < n) store(x, k[e], v[e]);
    }
    __syncthreads();
}

// n <= WG * 16: a workgroup's register sort; returns false (nothing done) for
// larger n.
template <bool PAIRS, typename LoadF, typename StoreF>
__device__ bool wg_sort_n(uint32_t n, LoadF load, StoreF store, uint64_t *sk, uint32_t *sv) {
    if (n <= (uint32_t)WG * 4) wg_sort_e<4, PAIRS>(n, load, store, sk, sv);
    else if (n <= (uint32_t)WG * 8) wg_sort_e<8, PAIRS>(n, load, store, sk, sv);
    else if (n <= (uint32_t)WG * 16) wg_sort_e<16, PAIRS>(n, load, store, sk, sv);
    else return false;
    return true;
}

// Sorts global (gk, gv)[0, n) using LDS scratch (sk, sv) of `cap` entries when
// it fits.  Ends with a barrier; results are visible to the whole workgroup.
template <typename K, typename V>
__device__ void wg_sort_pairs(K *gk, V *gv, uint32_t n, K *sk, V *sv, uint32_t cap) {
    if (n <= 1) {
        __syncthreads();
        return;
    }
    if (n <= cap) {
        for (uint32_t i = threadIdx.x; i < n; i += WG) {
            sk[i] = gk[i];
            sv[i] = gv[i];
        }
        __syncthreads();
        bitonic_core(sk, sv, n);
        for (uint32_t i = threadIdx.x; i < n; i += WG) {
            gk[i] = sk[i];
            gv[i] = sv[i];
        }
        __syncthreads();
    } else {
        bitonic_core(gk, gv, n);
    }
}

}  // namespace shm
