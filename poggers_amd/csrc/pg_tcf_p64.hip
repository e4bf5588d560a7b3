// Filter kernels for tile width 64 (one translation unit per width keeps
// the build parallel).
#include "pg_tcf.hpp"

namespace pg {
template int launch_tcf_p<64>(const TableDev &, Op, const OpArgs &, hipStream_t);
}  // namespace pg
