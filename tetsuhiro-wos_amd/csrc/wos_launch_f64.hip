// Launchers for the FP64 (complex128) kernels.
#include "wos_launch.inc"

namespace wos {
int launch_rows_f64(int N, int dir, int rows, const RowArgs& a, hipStream_t s) {
  return launch_rows_T<double>(N, dir, rows, a, s);
}
bool rows_valid_f64(int N, int rows) { return rows_valid_T<double>(N, rows); }
bool rows_pipe_lds_valid_f64(int N, int rows) { return rows_pipe_lds_valid_T<double>(N, rows); }
int launch_cols_f64(int N, int mode, int dir, int W, const ColArgs& a, hipStream_t s) {
  return launch_cols_T<double>(N, mode, dir, W, a, s);
}
int default_col_w_f64(int N, int cols) { return default_col_w_T<double>(N, cols); }
bool col_w_valid_f64(int N, int W) { return col_w_valid_T<double>(N, W); }
int launch_rows_bs_f64(int M, const BsRowArgs& A, hipStream_t s) { return launch_rows_bs_T<double>(M, A, s); }
int launch_cols_bs_f64(int M, int mode, int W, const BsColArgs& A, hipStream_t s) {
  return launch_cols_bs_T<double>(M, mode, W, A, s);
}
int bs_default_w_f64(int M, int cols) { return bs_default_w_T<double>(M, cols); }
}  // namespace wos
