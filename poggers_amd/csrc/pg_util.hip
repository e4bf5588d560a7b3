// pg_util.hip -- fill counting, zero fill, key streams and the random 64 B
// sector microbenchmark (gfx950).
#include "pg_internal.hpp"
#include "pg_tile.hpp"

namespace pg {

constexpr int kBlock = 256;

__device__ inline uint64_t wave_sum(uint64_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// count_fill_kernel (tables/bucketed_table.cuh:36-52) counts per bucket with
// one tile each; here each thread counts slots grid-stride and a wave adds
// its total once.  Filled = neither empty nor tombstone.
__global__ __launch_bounds__(kBlock) void k_fill_packed(const uint32_t *w, uint64_t n, uint64_t *counter) {
    uint64_t c = 0;
    for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += uint64_t(gridDim.x) * kBlock) {
        const uint32_t f = w[i] & 0xFFFFu;
        c += (f != 0u && f != 0xFFFFu);
    }
    c = wave_sum(c);
    if (__lane_id() == 0 && c) atomicAdd((unsigned long long *)counter, (unsigned long long)c);
}

// 16-bit words of the 8/8 and 10/6 formats: the key bits are the low
// popcount(kmask) bits.
__global__ __launch_bounds__(kBlock) void k_fill_packed16(const uint16_t *w, uint64_t n, uint32_t kmask,
                                                           uint64_t *counter) {
    uint64_t c = 0;
    for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += uint64_t(gridDim.x) * kBlock) {
        const uint32_t f = w[i] & kmask;
        c += (f != 0u && f != kmask);
    }
    c = wave_sum(c);
    if (__lane_id() == 0 && c) atomicAdd((unsigned long long *)counter, (unsigned long long)c);
}

// 12-bit buckets: a slot is filled when its composite is nonzero
// (12_bit_bucket.cuh:51-68 get_fill); one thread per bucket.
__global__ __launch_bounds__(kBlock) void k_fill_twelve(LevelDev L, uint64_t *counter) {
    uint64_t c = 0;
    for (uint64_t b = uint64_t(blockIdx.x) * kBlock + threadIdx.x; b < L.nb.d; b += uint64_t(gridDim.x) * kBlock) {
        const uint16_t *w = (const uint16_t *)bucket_ptr(L, b);
        for (uint32_t i = 0; i < L.B; i++) {
            const uint16_t *g = w + 3 * (i >> 2);
            uint32_t v;
            switch (i & 3u) {
                case 0: v = g[0] & 0x0FFFu; break;
                case 1: v = (g[0] & 0xF000u) | (g[1] & 0x00FFu); break;
                case 2: v = (g[1] & 0xFF00u) | (g[2] & 0x000Fu); break;
                default: v = g[2] & 0xFFF0u; break;
            }
            c += (v != 0u);
        }
    }
    c = wave_sum(c);
    if (__lane_id() == 0 && c) atomicAdd((unsigned long long *)counter, (unsigned long long)c);
}

__global__ __launch_bounds__(kBlock) void k_fill_key64(const uint64_t *s, uint64_t n, uint64_t *counter) {
    uint64_t c = 0;
    for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += uint64_t(gridDim.x) * kBlock) {
        const uint64_t k = s[i];
        c += (k != 0ull && k != ~0ull);
    }
    c = wave_sum(c);
    if (__lane_id() == 0 && c) atomicAdd((unsigned long long *)counter, (unsigned long long)c);
}

__global__ __launch_bounds__(kBlock) void k_fill_kv(const ulonglong2 *s, uint64_t n, uint64_t *counter) {
    uint64_t c = 0;
    for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += uint64_t(gridDim.x) * kBlock) {
        const uint64_t k = s[i].x;
        c += (k != 0ull && k != ~0ull);
    }
    c = wave_sum(c);
    if (__lane_id() == 0 && c) atomicAdd((unsigned long long *)counter, (unsigned long long)c);
}

__global__ __launch_bounds__(64) void k_zero_u64(uint64_t *p) { p[threadIdx.x] = 0; }

static unsigned stride_grid(uint64_t n) {
    const uint64_t want = (n + kBlock - 1) / kBlock;
    return (unsigned)(want < 8192 ? (want ? want : 1) : 8192);
}

int launch_fill(const TableDev &T, uint64_t *d_counter, hipStream_t s) {
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, s, d_counter);
    for (uint32_t l = 0; l < T.nlevels; l++) {
        const LevelDev &L = T.lv[l];
        const uint64_t nslots = L.nb.d * (uint64_t)L.B;
        // packed words of the whole array: a grouped level's padding is zero,
        // i.e. empty, so it is scanned like slots
        const uint64_t bytes = L.group ? ((L.nb.d - 1) / L.group + 1) * 128 : L.nb.d * (uint64_t)L.bucket_bytes;
        if (T.rep == PG_REP_FP12)
            hipLaunchKernelGGL(k_fill_twelve, dim3(stride_grid(L.nb.d)), dim3(kBlock), 0, s, L, d_counter);
        else if (T.rep == PG_REP_KEY64)
            hipLaunchKernelGGL(k_fill_key64, dim3(stride_grid(nslots)), dim3(kBlock), 0, s,
                               (const uint64_t *)L.slots, nslots, d_counter);
        else if (T.rep == PG_REP_KV64)
            hipLaunchKernelGGL(k_fill_kv, dim3(stride_grid(nslots)), dim3(kBlock), 0, s,
                               (const ulonglong2 *)L.slots, nslots, d_counter);
        else if (T.rep == PG_REP_FP8_V8 || T.rep == PG_REP_FP10_V6)
            hipLaunchKernelGGL(k_fill_packed16, dim3(stride_grid(bytes / 2)), dim3(kBlock), 0, s,
                               (const uint16_t *)L.slots, bytes / 2, T.rep == PG_REP_FP8_V8 ? 0xFFu : 0x3FFu,
                               d_counter);
        else
            hipLaunchKernelGGL(k_fill_packed, dim3(stride_grid(bytes / 4)), dim3(kBlock), 0, s,
                               (const uint32_t *)L.slots, bytes / 4, d_counter);
    }
    return hipGetLastError() == hipSuccess ? PG_OK : PG_ERR_HIP;
}

// Zero fill with 16 B stores (a kernel, not hipMemsetAsync, so it can be
// recorded in a hipGraph).  `bytes` is a multiple of 16.
__global__ __launch_bounds__(kBlock) void k_zero16(uint4 *p, uint64_t n16) {
    const uint4 z = make_uint4(0, 0, 0, 0);
    for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < n16; i += uint64_t(gridDim.x) * kBlock)
        p[i] = z;
}

int launch_zero(void *p, uint64_t bytes, hipStream_t s) {
    const uint64_t n16 = bytes / 16;
    if (n16) hipLaunchKernelGGL(k_zero16, dim3(stride_grid(n16)), dim3(kBlock), 0, s, (uint4 *)p, n16);
    return hipGetLastError() == hipSuccess ? PG_OK : PG_ERR_HIP;
}

// Occupancy index of a linear filter level from its slots: one thread per
// index word sets bit j when bucket 32w + j holds any non-empty word (16-bit
// words for the 8/8 slots, 32-bit for 16/16; key bits under kmask).
template <class W>
__global__ __launch_bounds__(kBlock) void k_occ_rebuild(const W *slots, uint64_t nb, uint32_t B, uint32_t kmask,
                                                        uint32_t *occ) {
    const uint64_t words = (nb + 31) / 32;
    for (uint64_t w = uint64_t(blockIdx.x) * kBlock + threadIdx.x; w < words; w += uint64_t(gridDim.x) * kBlock) {
        uint32_t bits = 0;
        for (uint32_t j = 0; j < 32 && 32 * w + j < nb; j++) {
            const W *bk = slots + (32 * w + j) * B;
            bool any = false;
            for (uint32_t i = 0; i < B && !any; i++) any = (bk[i] & kmask) != 0;
            bits |= (any ? 1u : 0u) << j;
        }
        occ[w] = bits;
    }
}

int launch_occ_rebuild(const LevelDev &L, uint32_t rep, hipStream_t s) {
    if (!L.occ || (rep != PG_REP_FP16_V16 && rep != PG_REP_FP8_V8)) return PG_ERR_INVALID;
    const uint64_t words = (L.nb.d + 31) / 32;
    if (rep == PG_REP_FP16_V16)
        hipLaunchKernelGGL(k_occ_rebuild<uint32_t>, dim3(stride_grid(words)), dim3(kBlock), 0, s,
                           (const uint32_t *)L.slots, L.nb.d, L.B, Fmt16::kmask, L.occ);
    else
        hipLaunchKernelGGL(k_occ_rebuild<uint16_t>, dim3(stride_grid(words)), dim3(kBlock), 0, s,
                           (const uint16_t *)L.slots, L.nb.d, L.B, Fmt8::kmask, L.occ);
    return hipGetLastError() == hipSuccess ? PG_OK : PG_ERR_HIP;
}

__global__ __launch_bounds__(kBlock) void k_gen_keys(uint64_t seed, uint64_t start, uint64_t n, uint64_t *out) {
    for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += uint64_t(gridDim.x) * kBlock)
        out[i] = key_at(seed, start + i);
}

int launch_gen_keys(uint64_t seed, uint64_t start, uint64_t n, uint64_t *out, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_gen_keys, dim3(stride_grid(n)), dim3(kBlock), 0, s, seed, start, n, out);
    return hipGetLastError() == hipSuccess ? PG_OK : PG_ERR_HIP;
}

// Random 64 B-sector reads, shaped like a TCF bucket probe.  Sector indices
// come from fmix64 of a counter (no index array is read).  Variants (packed in
// `op`, see pg_random_sector_bench):
//   bits 0-1  0 = read only; 1 = add the CAS an insert performs on one word of
//             the sector (it rewrites the value it read, so the line is
//             dirtied); 2 = add a plain store of that word instead
//   bits 4-7  v: sectors per tile kept in flight, 1 << (v - 1) for v in 1..4
//             (1, 2, 4, 8); 0 = 4
//   bit 8     nontemporal loads
//   bits 9-10 tile shape: 0 = 4 lanes x 16 B, 1 = 1 lane x 4 x 16 B,
//             2 = 16 lanes x 4 B
//   bits 11-14 w: 0 = one tile per group of sectors (the grid covers the
//             work, as the table kernels do); 1..15 = w blocks of 256 threads
//             per CU (w waves per SIMD) walking the work grid-stride, which
//             caps the requests in flight
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

template <bool NT>
__device__ inline u32x4 ld16(const u32x4 *p) {
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}
template <bool NT>
__device__ inline uint32_t ld4(const uint32_t *p) {
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}

template <int LANES, int I, bool NT>
__global__ __launch_bounds__(kBlock) void k_random_sector(uint32_t *buf, FastMod nsec, uint64_t n,
                                                          uint32_t wop, uint64_t seed, uint64_t *sink) {
    const uint32_t lane = __lane_id();
    const uint32_t rank = lane & (LANES - 1);
    const uint64_t tiles_in_grid = (uint64_t(gridDim.x) * kBlock) / LANES;
    const uint64_t ntiles = (n + I - 1) / I;
    uint32_t x = 0;
    for (uint64_t tile = (uint64_t(blockIdx.x) * kBlock + threadIdx.x) / LANES; tile < ntiles;
         tile += tiles_in_grid) {
        uint64_t sec[I];
        u32x4 v[I];
#pragma unroll
        for (int i = 0; i < I; i++) {
            const uint64_t r = tile * I + i;
            sec[i] = fastmod(fmix64(r ^ seed), nsec);
            const uint32_t *base = buf + sec[i] * 16;
            if (r >= n) {
                v[i] = u32x4{0, 0, 0, 0};
            } else if constexpr (LANES == 4) {
                v[i] = ld16<NT>((const u32x4 *)(base + rank * 4));
            } else if constexpr (LANES == 1) {
                const u32x4 a = ld16<NT>((const u32x4 *)base), b = ld16<NT>((const u32x4 *)(base + 4));
                const u32x4 c = ld16<NT>((const u32x4 *)(base + 8)), d = ld16<NT>((const u32x4 *)(base + 12));
                v[i] = a ^ b ^ c ^ d;
            } else {
                const uint32_t w = ld4<NT>(base + rank);
                v[i] = u32x4{w, 0, 0, 0};
            }
        }
#pragma unroll
        for (int i = 0; i < I; i++) {
            x ^= v[i].x ^ v[i].y ^ v[i].z ^ v[i].w;
            if (wop && rank == 0 && tile * I + i < n) {
                uint32_t *w = buf + sec[i] * 16;
                if (wop == 1) x ^= atomicCAS(w, v[i].x, v[i].x);
                else *w = v[i].x ^ v[i].y ^ v[i].z ^ v[i].w;  // rewrites the sector's first word
            }
        }
    }
    if (x == 0x9E3779B9u) sink[0] = x;
}

template <int LANES, int I, bool NT>
static void launch_rs(uint32_t *buf, FastMod f, uint64_t n, uint32_t wop, uint32_t per_cu, uint64_t seed,
                      uint64_t *sink, hipStream_t s) {
    const uint64_t tiles = (n + I - 1) / I;
    unsigned grid = (unsigned)((tiles * LANES + kBlock - 1) / kBlock);
    if (per_cu) {
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
        const unsigned cap = (unsigned)cus * per_cu;
        if (cap < grid) grid = cap;
    }
    hipLaunchKernelGGL((k_random_sector<LANES, I, NT>), dim3(grid), dim3(kBlock), 0, s, buf, f, n, wop, seed,
                       sink);
}

template <int LANES, bool NT>
static int launch_rs_i(uint32_t iv, uint32_t *buf, FastMod f, uint64_t n, uint32_t wop, uint32_t per_cu,
                       uint64_t seed, uint64_t *sink, hipStream_t s) {
    switch (iv) {
        case 1: launch_rs<LANES, 1, NT>(buf, f, n, wop, per_cu, seed, sink, s); break;
        case 2: launch_rs<LANES, 2, NT>(buf, f, n, wop, per_cu, seed, sink, s); break;
        case 0:
        case 3: launch_rs<LANES, 4, NT>(buf, f, n, wop, per_cu, seed, sink, s); break;
        case 4: launch_rs<LANES, 8, NT>(buf, f, n, wop, per_cu, seed, sink, s); break;
        default: return PG_ERR_INVALID;
    }
    return PG_OK;
}

int launch_random_sector(void *buf, uint64_t bytes, uint64_t n, uint32_t op, uint64_t seed,
                         uint64_t *sink, hipStream_t s) {
    const uint64_t nsec = bytes / 64;
    if (nsec == 0) return PG_ERR_INVALID;
    const FastMod f = make_fastmod(nsec);
    const uint32_t wop = op & 3u;
    const uint32_t iv = (op >> 4) & 0xFu;
    const bool nt = (op >> 8) & 1u;
    const uint32_t shape = (op >> 9) & 3u;
    const uint32_t per_cu = (op >> 11) & 0xFu;
    if (wop == 3) return PG_ERR_INVALID;
    if (n == 0) return PG_OK;
    uint32_t *b = (uint32_t *)buf;
    int rc;
    if (shape == 0) rc = nt ? launch_rs_i<4, true>(iv, b, f, n, wop, per_cu, seed, sink, s)
                            : launch_rs_i<4, false>(iv, b, f, n, wop, per_cu, seed, sink, s);
    else if (shape == 1) rc = nt ? launch_rs_i<1, true>(iv, b, f, n, wop, per_cu, seed, sink, s)
                                 : launch_rs_i<1, false>(iv, b, f, n, wop, per_cu, seed, sink, s);
    else if (shape == 2) rc = nt ? launch_rs_i<16, true>(iv, b, f, n, wop, per_cu, seed, sink, s)
                                 : launch_rs_i<16, false>(iv, b, f, n, wop, per_cu, seed, sink, s);
    else return PG_ERR_INVALID;
    if (rc != PG_OK) return rc;
    return hipGetLastError() == hipSuccess ? PG_OK : PG_ERR_HIP;
}

}  // namespace pg
