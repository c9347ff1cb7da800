// wos_engine.hpp — host-side engine of libwos_hip: device context, buffer
// pools, twiddle/chirp caches and typed launchers over the kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "wos_kernels.hpp"

namespace wos {

// Thread-local error text, surfaced by wos_hip_last_error().
void set_error(const std::string& msg);

struct Status {
  int code = 0;
  std::string msg;
  static Status ok() { return {}; }
};

#define WOS_HIP_TRY(expr)                                                                      \
  do {                                                                                         \
    hipError_t e_ = (expr);                                                                    \
    if (e_ != hipSuccess) {                                                                    \
      ::wos::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));                    \
      return 3;                                                                                \
    }                                                                                          \
  } while (0)

// A grow-only device buffer.
struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  int ensure(size_t bytes);
  void release();
};

// Per-device state. All work of the C ABI runs on `stream` under `mu`.
struct Device {
  bool ready = false;
  int id = 0;
  char name[256] = {0};
  hipStream_t stream = nullptr;
  std::recursive_mutex mu;
  std::map<long long, void*> twiddles;     // key: n*2 + prec
  std::map<long long, void*> chirps;       // Bluestein tables, key: (n*2+inverse)*2+prec
  // field: resident fields; stage: complex128 host staging for FP32;
  // work: per-unit working copies; aux: spectral-diagnostic copies;
  // scratch: padded intermediates; ir: fresnel_ir spectrum; screens: BPM tables
  DevBuf field, stage, work, aux, scratch, ir, screens, params, partial_in, partial_out, hist, out;
  // fresnel_ir: the impulse-response spectrum of the last (precision, grid,
  // dx, wl, z), reused while they repeat
  DevBuf fir;
  bool fir_ok = false;
  int fir_prec = 0, fir_m = 0;
  double fir_dx = 0, fir_wl = 0, fir_z = 0;
  void* noalloc = nullptr;                 // this device's static scratch (wos_hip_propagate_device_noalloc)
  size_t noalloc_bytes = 0;
  void* host_pin = nullptr;                // pinned staging for parameter uploads
  size_t host_pin_cap = 0;
  hipEvent_t pin_ev = nullptr;             // recorded after the last copy out of host_pin
  int ensure_host_pin(size_t bytes);
  // Copy `bytes` from host memory to `dst` (device) in stream order, through
  // host_pin; waits for the previous copy out of host_pin first.
  int upload_params(void* dst, const void* src, size_t bytes);
};

Device& device();
int ensure_init();

// Transfers between pageable host memory (a Go slice, a numpy array) and the
// device, in the order of stream s. They go through a ring of pinned staging
// chunks: host threads copy a chunk into (or out of) the ring while the DMA
// engine moves the previous one, instead of the runtime's own pageable
// staging. host_to_dev returns once the source has been read (the DMA may
// still be in flight); dev_to_host returns when the data is in dst.
int host_to_dev(void* dst, const void* src, size_t bytes, hipStream_t s);
int dev_to_host(void* dst, const void* src, size_t bytes, hipStream_t s);

// Stage twiddle tables of an n-point transform in precision prec (0 = FP64);
// acc = 1: compensated [hi, lo] pairs; radix: stage schedule (16 or 8).
int get_twiddles(int n, int prec, const void** out, int acc = 0, int radix = 16);

// ---- launchers (return 0 or a WOS_ERR_* code) ----
// Row pass on `units` units of rows x N; rows = rows per workgroup (0: default).
// a.radix = 8 selects the radix-8 schedule's kernels (tables to match).
int launch_rows(int prec, int N, int dir, const RowArgs& a, hipStream_t s, int rows = 0);
bool rows_valid(int prec, int N, int rows, int radix = 16);
bool col_w_valid(int prec, int N, int W, int radix = 16);
// Column pass; mode 0 plain (dir), 1 fused forward/TF/inverse. W = 0: default.
int launch_cols(int prec, int N, int mode, int dir, int W, const ColArgs& a, hipStream_t s);
int default_col_w(int prec, int N, int cols);
bool pow2(int n);
int max_pow2_len();

}  // namespace wos
