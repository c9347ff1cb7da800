// Launchers for the FP64 (complex128) mixed-radix kernels.
#include "wos_launch_mixed.inc"

namespace wos {
int launch_rows_mixed_f64(int N, int dir, const RowArgs& a, hipStream_t s) {
  return launch_rows_mixed_T<double>(N, dir, a, s);
}
int launch_cols_mixed_f64(int N, int mode, int dir, int W, const ColArgs& a, hipStream_t s) {
  return launch_cols_mixed_T<double>(N, mode, dir, W, a, s);
}
int mixed_col_w_f64(int N) { return mixed_col_w_T<double>(N); }
bool mixed_col_valid_f64(int N, int W) { return mixed_col_valid_T<double>(N, W); }
int mixed_rows_wg_f64(int N) { return mixed_rows_wg_T<double>(N); }
}  // namespace wos
