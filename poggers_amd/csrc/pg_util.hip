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
        if (T.rep == PG_REP_KV64)
            hipLaunchKernelGGL(k_fill_kv, dim3(stride_grid(nslots)), dim3(kBlock), 0, s,
                               (const ulonglong2 *)L.slots, nslots, d_counter);
        else
            hipLaunchKernelGGL(k_fill_packed, dim3(stride_grid(nslots)), dim3(kBlock), 0, s,
                               (const uint32_t *)L.slots, nslots, d_counter);
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

__global__ __launch_bounds__(kBlock) void k_gen_keys(uint64_t seed, uint64_t start, uint64_t n, uint64_t *out) {
    for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += uint64_t(gridDim.x) * kBlock)
        out[i] = key_at(seed, start + i);
}

int launch_gen_keys(uint64_t seed, uint64_t start, uint64_t n, uint64_t *out, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_gen_keys, dim3(stride_grid(n)), dim3(kBlock), 0, s, seed, start, n, out);
    return hipGetLastError() == hipSuccess ? PG_OK : PG_ERR_HIP;
}

// Random 64 B-sector reads, shaped like a TCF bucket probe: 4-lane tiles,
// 16 B per lane, each tile reading kItems independent sectors whose indices
// come from fmix64 of a counter (no index array is read).  op 1 adds the CAS
// an insert performs on one word of the sector (it rewrites the same value).
constexpr int kItems = 4;

__global__ __launch_bounds__(kBlock) void k_random_sector(const uint4 *buf, FastMod nsec, uint64_t n,
                                                          uint32_t op, uint64_t seed, uint64_t *sink) {
    Tile<4> t;
    const uint64_t tile = (uint64_t(blockIdx.x) * kBlock + threadIdx.x) / 4;
    uint64_t sec[kItems];
    uint4 v[kItems];
#pragma unroll
    for (int i = 0; i < kItems; i++) {
        const uint64_t r = tile * kItems + i;
        sec[i] = fastmod(fmix64(r ^ seed), nsec);
        if (r < n) v[i] = buf[sec[i] * 4 + t.rank];
        else v[i] = make_uint4(0, 0, 0, 0);
    }
    uint32_t x = 0;
#pragma unroll
    for (int i = 0; i < kItems; i++) {
        x ^= v[i].x ^ v[i].y ^ v[i].z ^ v[i].w;
        if (op == 1 && t.rank == 0 && tile * kItems + i < n) {
            unsigned int *w = (unsigned int *)(buf + sec[i] * 4);
            x ^= atomicCAS(w, v[i].x, v[i].x);
        }
    }
    if (x == 0x9E3779B9u) sink[0] = x;
}

int launch_random_sector(void *buf, uint64_t bytes, uint64_t n, uint32_t op, uint64_t seed,
                         uint64_t *sink, hipStream_t s) {
    const uint64_t nsec = bytes / 64;
    if (nsec == 0) return PG_ERR_INVALID;
    const FastMod f = make_fastmod(nsec);
    const uint64_t tiles = (n + kItems - 1) / kItems;
    const unsigned grid = (unsigned)((tiles * 4 + kBlock - 1) / kBlock);
    if (n) hipLaunchKernelGGL(k_random_sector, dim3(grid), dim3(kBlock), 0, s, (const uint4 *)buf, f, n, op,
                              seed, sink);
    return hipGetLastError() == hipSuccess ? PG_OK : PG_ERR_HIP;
}

}  // namespace pg
