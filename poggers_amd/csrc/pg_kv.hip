// pg_kv.hip -- bulk kernels for the key-value slot tables: key_val_pair<u64,
// u64>, key-only and 16/16 key_val_pair slots (device ops in pg_kv_ops.hpp).
#include "pg_kv_ops.hpp"

namespace pg {

constexpr int kBlock = 256;

template <int P, bool ACC, class SL, bool RP = false>
__global__ __launch_bounds__(kBlock) void k_kv_insert(TableDev T, OpArgs a) {
    Tile<P> t;
    t.concurrent = true;
    const uint64_t tid = (uint64_t(blockIdx.x) * kBlock + threadIdx.x) / P;
    const bool act = tid < a.n;
    int st = 1;
    if (act) {
        const uint64_t val = a.vals ? (uint64_t)((const typename SL::V *)a.vals)[tid] : 0ull;
        st = kv_insert_key<P, SL, RP>(T, a.keys[tid], val, t);
        if (a.flags && t.leader()) a.flags[tid] = (uint8_t)st;
    }
    wave_count(a.fail_count, act && t.leader() && st == 0);
    if constexpr (ACC) flush_acc(a.stats, t, act);
}

template <int P, bool ACC, class SL>
__global__ __launch_bounds__(kBlock) void k_kv_query(TableDev T, OpArgs a) {
    Tile<P> t;
    const uint64_t tid = (uint64_t(blockIdx.x) * kBlock + threadIdx.x) / P;
    if (tid < a.n) {
        uint64_t val = 0;
        bool own = false;
        const bool found = kv_query_key<P, SL>(T, a.keys[tid], t, val, own);
        const bool writer = found ? own : t.leader();
        if (writer) {
            if (a.flags) a.flags[tid] = found ? 1 : 0;
            if (a.vals_out) ((typename SL::V *)a.vals_out)[tid] = found ? (typename SL::V)val : 0;
        }
    }
    if constexpr (ACC) flush_acc(a.stats, t, tid < a.n);
}

template <int P, bool ACC, class SL>
__global__ __launch_bounds__(kBlock) void k_kv_remove(TableDev T, OpArgs a) {
    Tile<P> t;
    const uint64_t tid = (uint64_t(blockIdx.x) * kBlock + threadIdx.x) / P;
    if (tid < a.n) {
        const bool ok = kv_remove_key<P, SL>(T, a.keys[tid], t);
        if (a.flags && t.leader()) a.flags[tid] = ok ? 1 : 0;
    }
    if constexpr (ACC) flush_acc(a.stats, t, tid < a.n);
}

template <int P, class SL, bool RP = false>
__global__ __launch_bounds__(64) void k_kv_insert_seq(TableDev T, OpArgs a) {
    Tile<P> t;
    uint64_t fails = 0;
    for (uint64_t i = 0; i < a.n; i++) {
        const uint64_t val = a.vals ? (uint64_t)((const typename SL::V *)a.vals)[i] : 0ull;
        const int st = kv_insert_key<P, SL, RP>(T, a.keys[i], val, t);
        if (a.flags && t.leader()) a.flags[i] = (uint8_t)st;
        fails += (st == 0);
        fence_agent();
    }
    if (a.fail_count && t.leader() && fails) atomicAdd((unsigned long long *)a.fail_count, fails);
    flush_acc(a.stats, t, true);
}

template <int P, class SL>
__global__ __launch_bounds__(64) void k_kv_remove_seq(TableDev T, OpArgs a) {
    Tile<P> t;
    for (uint64_t i = 0; i < a.n; i++) {
        const bool ok = kv_remove_key<P, SL>(T, a.keys[i], t);
        if (a.flags && t.leader()) a.flags[i] = ok ? 1 : 0;
        fence_agent();
    }
    flush_acc(a.stats, t, true);
}

template <int P, bool ACC, class SL>
static void launch_kv_conc(Op op, unsigned grid, const TableDev &T, const OpArgs &a, bool rp, hipStream_t s) {
    if (op == Op::Query) hipLaunchKernelGGL((k_kv_query<P, ACC, SL>), dim3(grid), dim3(kBlock), 0, s, T, a);
    else if (op == Op::Insert && rp)
        hipLaunchKernelGGL((k_kv_insert<P, ACC, SL, true>), dim3(grid), dim3(kBlock), 0, s, T, a);
    else if (op == Op::Insert) hipLaunchKernelGGL((k_kv_insert<P, ACC, SL>), dim3(grid), dim3(kBlock), 0, s, T, a);
    else hipLaunchKernelGGL((k_kv_remove<P, ACC, SL>), dim3(grid), dim3(kBlock), 0, s, T, a);
}

// Whether a level of the table runs the replace scheme.
static bool has_replace_level(const TableDev &T) {
    for (uint32_t l = 0; l < T.nlevels; l++)
        if (T.lv[l].scheme == PG_SCHEME_POWER_OF_N_SHORTCUT_REPLACE) return true;
    return false;
}

template <int P, class SL>
static int launch_kv_ps(const TableDev &T, Op op, const OpArgs &a, hipStream_t s) {
    if (a.n == 0) return PG_OK;
    const unsigned grid = (unsigned)((a.n * (uint64_t)P + kBlock - 1) / kBlock);
    const bool rp = has_replace_level(T);
    if (op != Op::Query && a.deterministic) {
        if (op == Op::Insert && rp) hipLaunchKernelGGL((k_kv_insert_seq<P, SL, true>), dim3(1), dim3(P), 0, s, T, a);
        else if (op == Op::Insert) hipLaunchKernelGGL((k_kv_insert_seq<P, SL>), dim3(1), dim3(P), 0, s, T, a);
        else hipLaunchKernelGGL((k_kv_remove_seq<P, SL>), dim3(1), dim3(P), 0, s, T, a);
    } else if (a.stats) {
        launch_kv_conc<P, true, SL>(op, grid, T, a, rp, s);
    } else {
        launch_kv_conc<P, false, SL>(op, grid, T, a, rp, s);
    }
    return hipGetLastError() == hipSuccess ? PG_OK : PG_ERR_HIP;
}

template <int P>
static int launch_kv_p(const TableDev &T, Op op, const OpArgs &a, hipStream_t s) {
    if (T.rep == PG_REP_KEY64) return launch_kv_ps<P, SlotKey>(T, op, a, s);
    if (T.rep == PG_REP_KV16) return launch_kv_ps<P, SlotKV16>(T, op, a, s);
    return launch_kv_ps<P, SlotKV>(T, op, a, s);
}

int launch_kv(const TableDev &T, uint32_t tile, Op op, const OpArgs &a, hipStream_t s) {
    switch (tile) {
        case 1: return launch_kv_p<1>(T, op, a, s);
        case 2: return launch_kv_p<2>(T, op, a, s);
        case 4: return launch_kv_p<4>(T, op, a, s);
        case 8: return launch_kv_p<8>(T, op, a, s);
        case 16: return launch_kv_p<16>(T, op, a, s);
        case 32: return launch_kv_p<32>(T, op, a, s);
        case 64: return launch_kv_p<64>(T, op, a, s);
        default: return PG_ERR_INVALID;
    }
}

}  // namespace pg
