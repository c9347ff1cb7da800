// Launchers for the FP64 (complex128) kernels under the radix-8 stage schedule.
#include "wos_launch.inc"

namespace wos {
int launch_rows8_f64(int N, int dir, int rows, const RowArgs& a, hipStream_t s) {
  return launch_rows8_T<double>(N, dir, rows, a, s);
}
bool rows_valid8_f64(int N, int rows) { return rows_valid8_T<double>(N, rows); }
int launch_cols8_f64(int N, int mode, int W, const ColArgs& a, hipStream_t s) {
  return launch_cols8_T<double>(N, mode, W, a, s);
}
bool col_w_valid8_f64(int N, int W) { return col_w_valid8_T<double>(N, W); }
int launch_cols32_f64(int N, int mode, int W, const ColArgs& a, hipStream_t s) {
  return launch_cols32_T<double>(N, mode, W, a, s);
}
bool col_w_valid32_f64(int N, int W) { return col_w_valid32_T<double>(N, W); }
}  // namespace wos
