// Launchers for the FP32 (complex64) kernels.
#include "wos_launch.inc"

namespace wos {
int launch_rows_f32(int N, int dir, const RowArgs& a, hipStream_t s) { return launch_rows_T<float>(N, dir, a, s); }
int launch_cols_f32(int N, int mode, int dir, int W, const ColArgs& a, hipStream_t s) {
  return launch_cols_T<float>(N, mode, dir, W, a, s);
}
int default_col_w_f32(int N, int cols) { return default_col_w_T<float>(N, cols); }
}  // namespace wos
