// Launchers for the FP32 (complex64) mixed-radix kernels.
#include "wos_launch_mixed.inc"

namespace wos {
int launch_rows_mixed_f32(int N, int dir, const RowArgs& a, hipStream_t s) {
  return launch_rows_mixed_T<float>(N, dir, a, s);
}
int launch_cols_mixed_f32(int N, int mode, int dir, int W, const ColArgs& a, hipStream_t s) {
  return launch_cols_mixed_T<float>(N, mode, dir, W, a, s);
}
int mixed_col_w_f32(int N) { return mixed_col_w_T<float>(N); }
bool mixed_col_valid_f32(int N, int W) { return mixed_col_valid_T<float>(N, W); }
int mixed_rows_wg_f32(int N) { return mixed_rows_wg_T<float>(N); }
}  // namespace wos
