// pg_tcf.hip -- tile-width dispatch of the filter kernels (pg_tcf.hpp).
#include "pg_internal.hpp"

namespace pg {

template <int P>
int launch_tcf_p(const TableDev &T, Op op, const OpArgs &a, hipStream_t s);

int launch_tcf(const TableDev &T, uint32_t tile, Op op, const OpArgs &a, hipStream_t s) {
    switch (tile) {
        case 1: return launch_tcf_p<1>(T, op, a, s);
        case 2: return launch_tcf_p<2>(T, op, a, s);
        case 4: return launch_tcf_p<4>(T, op, a, s);
        case 8: return launch_tcf_p<8>(T, op, a, s);
        case 16: return launch_tcf_p<16>(T, op, a, s);
        case 32: return launch_tcf_p<32>(T, op, a, s);
        case 64: return launch_tcf_p<64>(T, op, a, s);
        default: return PG_ERR_INVALID;
    }
}

}  // namespace pg
