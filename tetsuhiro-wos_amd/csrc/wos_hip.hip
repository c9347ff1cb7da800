// wos_hip.hip — C ABI of libwos_hip and the host orchestration behind it.
//
// The ABI mirrors what a `-tags hip` cgo file binds in place of the cuFFT
// backend (optics/gpu.go) and of the FFT-based operators of optics/
// (Propagate, PropagateSplitStep, PropagatePolychromatic, Propagate3D,
// Propagate1D). Device work runs on one non-blocking stream of device 0 under
// a recursive mutex (the gpuExecMu contract, optics/gpu.go:97-100).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/wos_hip.h"
#include "wos_engine.hpp"
#include "wos_bluestein.hpp"

// Device memory of wos_hip_propagate_device_noalloc (callers that cannot
// allocate): the TF records, twiddle tables and column-tiled intermediate of
// one call live here. Each GPU holds its own copy for the life of the
// process; host code uses the address hipGetSymbolAddress returns for the
// current device (Device::noalloc), never the array's name.
constexpr size_t kNoAllocScratchBytes = (size_t)72 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace wos {

int launch_rows_f64(int N, int dir, int rows, const RowArgs& a, hipStream_t s);
int launch_rows_f32(int N, int dir, int rows, const RowArgs& a, hipStream_t s);
bool rows_valid_f64(int N, int rows);
bool rows_valid_f32(int N, int rows);
bool rows_pipe_lds_valid_f64(int N, int rows);
bool rows_pipe_lds_valid_f32(int N, int rows);
int launch_cols_f64(int N, int mode, int dir, int W, const ColArgs& a, hipStream_t s);
int launch_cols_f32(int N, int mode, int dir, int W, const ColArgs& a, hipStream_t s);
int default_col_w_f64(int N, int cols);
int default_col_w_f32(int N, int cols);
bool col_w_valid_f64(int N, int W);
bool col_w_valid_f32(int N, int W);
int launch_rows_bs_f64(int M, const BsRowArgs& A, hipStream_t s);
int launch_rows_bs_f32(int M, const BsRowArgs& A, hipStream_t s);
int launch_rows8_f64(int N, int dir, int rows, const RowArgs& a, hipStream_t s);
int launch_rows8_f32(int N, int dir, int rows, const RowArgs& a, hipStream_t s);
bool rows_valid8_f64(int N, int rows);
bool rows_valid8_f32(int N, int rows);
int launch_cols8_f64(int N, int mode, int W, const ColArgs& a, hipStream_t s);
int launch_cols8_f32(int N, int mode, int W, const ColArgs& a, hipStream_t s);
bool col_w_valid8_f64(int N, int W);
bool col_w_valid8_f32(int N, int W);
int launch_cols32_f64(int N, int mode, int W, const ColArgs& a, hipStream_t s);
int launch_cols32_f32(int N, int mode, int W, const ColArgs& a, hipStream_t s);
bool col_w_valid32_f64(int N, int W);
bool col_w_valid32_f32(int N, int W);
int launch_rows_mixed_f64(int N, int dir, const RowArgs& a, hipStream_t s);
int launch_rows_mixed_f32(int N, int dir, const RowArgs& a, hipStream_t s);
int launch_cols_mixed_f64(int N, int mode, int dir, int W, const ColArgs& a, hipStream_t s);
int launch_cols_mixed_f32(int N, int mode, int dir, int W, const ColArgs& a, hipStream_t s);
int mixed_col_w_f64(int N);
int mixed_col_w_f32(int N);
int mixed_rows_wg_f64(int N);
int mixed_rows_wg_f32(int N);
bool mixed_col_valid_f64(int N, int W);
bool mixed_col_valid_f32(int N, int W);
// persistent transfer (wos_persist.hip)
bool persist_built(int prec, int N);
int persist_rows(int prec);
int persist_grid(int prec, int N, int* grid);
int launch_transfer_persist(int prec, int N, int grid, const PersistArgs& a, hipStream_t s);
int persist_xcc_map_ok(hipStream_t s, bool* ok);
int launch_cols_bs_f64(int M, int mode, int W, const BsColArgs& A, hipStream_t s);
int launch_cols_bs_f32(int M, int mode, int W, const BsColArgs& A, hipStream_t s);
int bs_default_w_f64(int M, int cols);
int bs_default_w_f32(int M, int cols);
static int launch_rows_bs(int prec, int M, const BsRowArgs& A, hipStream_t s);

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }

bool pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }
int max_pow2_len() { return 8192; }
// internal power-of-two length of a Bluestein transform of length L
static int bs_len(int L) {
  int m = 1;
  while (m < 2 * L - 1) m <<= 1;
  return m;
}
// lengths the engine transforms: powers of two up to 8192, others through
// Bluestein with an internal length up to 8192
static bool len_ok(int L) { return L >= 1 && (pow2(L) ? L <= max_pow2_len() : bs_len(L) <= max_pow2_len()); }
// the drop-in transforms on the other lengths (wos_long.inc); the working
// buffers of one chunk of rows stay under kLongChunkBytes
constexpr size_t kLongChunkBytes = (size_t)2 << 30;
static bool long_len_ok(long long L);
static int long_pow2_dev(int prec, void* d, void* tmp, long long N, int units, int inverse, double scale,
                         hipStream_t s);
static int long_fft_dev(int prec, void* d, long long L, long long units, int inverse, double scale, hipStream_t s);
static int long_fft2d_dev(int prec, void* d, int n, int inverse);

static size_t csize(int prec) { return prec == WOS_FP64 ? 16 : 8; }

int DevBuf::ensure(size_t bytes) {
  if (bytes <= cap) return 0;
  if (p) {
    (void)hipStreamSynchronize(device().stream);  // no pending work may touch the old block
    (void)hipFree(p);
  }
  p = nullptr;
  cap = 0;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) {
    p = nullptr;
    set_error(std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e));
    return WOS_ERR_NOMEM;
  }
  cap = bytes;
  return 0;
}
void DevBuf::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  cap = 0;
}

int Device::ensure_host_pin(size_t bytes) {
  if (bytes <= host_pin_cap) return 0;
  if (host_pin) {
    (void)hipStreamSynchronize(stream);
    (void)hipHostFree(host_pin);
  }
  host_pin = nullptr;
  host_pin_cap = 0;
  hipError_t e = hipHostMalloc(&host_pin, bytes, hipHostMallocDefault);
  if (e != hipSuccess) {
    host_pin = nullptr;
    set_error(std::string("hipHostMalloc: ") + hipGetErrorString(e));
    return WOS_ERR_NOMEM;
  }
  host_pin_cap = bytes;
  return 0;
}

int Device::upload_params(void* dst, const void* src, size_t bytes) {
  int rc;
  if (pin_ev) WOS_HIP_TRY(hipEventSynchronize(pin_ev));
  if ((rc = ensure_host_pin(bytes))) return rc;
  memcpy(host_pin, src, bytes);
  WOS_HIP_TRY(hipMemcpyAsync(dst, host_pin, bytes, hipMemcpyHostToDevice, stream));
  if (!pin_ev) WOS_HIP_TRY(hipEventCreateWithFlags(&pin_ev, hipEventDisableTiming));
  WOS_HIP_TRY(hipEventRecord(pin_ev, stream));
  return 0;
}

Device& device() {
  static Device d;
  return d;
}

// ---- host copies for staged transfers ----
// A parallel memcpy over a few persistent threads: one pageable <-> pinned
// copy runs at ~27 GB/s on one host thread and ~100 GB/s on eight, while the
// DMA engine moves ~57 GB/s from pinned memory. The pool is created on first
// use and never destroyed (its threads sleep between jobs); one job runs at a
// time, and the calling thread copies a share of it too.
class CopyPool {
 public:
  static CopyPool& get() {
    static CopyPool* pool = new CopyPool();  // outlives static destructors: threads may still wait on it
    return *pool;
  }
  void copy(void* dst, const void* src, size_t n) {
    if (workers_ == 0 || n < ((size_t)1 << 20)) {
      memcpy(dst, src, n);
      return;
    }
    std::lock_guard<std::mutex> call(call_mu_);
    std::unique_lock<std::mutex> lk(mu_);
    dst_ = (char*)dst;
    src_ = (const char*)src;
    n_ = n;
    const size_t want = (size_t)workers_ + 1;
    piece_ = ((n + want - 1) / want + 63) & ~(size_t)63;
    pieces_ = (int)((n + piece_ - 1) / piece_);
    next_ = 0;
    left_ = pieces_;
    cv_work_.notify_all();
    while (next_ < pieces_) take(lk);
    cv_done_.wait(lk, [&] { return left_ == 0; });
    pieces_ = 0;
    next_ = 0;
  }

 private:
  CopyPool() {
    const unsigned hw = std::thread::hardware_concurrency();
    workers_ = (int)std::min(7u, hw > 2 ? hw / 2 - 1 : 0u);
    for (int i = 0; i < workers_; ++i) std::thread([this] { run(); }).detach();
  }
  // claim one piece under the lock, copy it without the lock
  void take(std::unique_lock<std::mutex>& lk) {
    const size_t off = (size_t)next_++ * piece_;
    char* d = dst_ + off;
    const char* s = src_ + off;
    const size_t len = std::min(piece_, n_ - off);
    lk.unlock();
    memcpy(d, s, len);
    lk.lock();
    if (--left_ == 0) cv_done_.notify_one();
  }
  void run() {
    std::unique_lock<std::mutex> lk(mu_);
    for (;;) {
      cv_work_.wait(lk, [&] { return next_ < pieces_; });
      take(lk);
    }
  }
  std::mutex call_mu_, mu_;
  std::condition_variable cv_work_, cv_done_;
  char* dst_ = nullptr;
  const char* src_ = nullptr;
  size_t n_ = 0, piece_ = 0;
  int pieces_ = 0, next_ = 0, left_ = 0, workers_ = 0;
};

// Pinned staging ring: kSlots chunks of kChunk bytes, each with the event of
// the last DMA that used it.
constexpr int kSlots = 8;
constexpr size_t kChunk = (size_t)8 << 20;
struct XferRing {
  char* pin = nullptr;
  hipEvent_t ev[kSlots] = {};
  bool used[kSlots] = {};  // ev[i] has been recorded
  int wait(int slot) {
    if (used[slot]) WOS_HIP_TRY(hipEventSynchronize(ev[slot]));
    return 0;
  }
  int mark(int slot, hipStream_t s) {
    WOS_HIP_TRY(hipEventRecord(ev[slot], s));
    used[slot] = true;
    return 0;
  }
  int ensure() {
    if (pin) return 0;
    hipError_t e = hipHostMalloc((void**)&pin, kChunk * kSlots, hipHostMallocDefault);
    if (e != hipSuccess) {
      pin = nullptr;
      set_error(std::string("hipHostMalloc(staging ring): ") + hipGetErrorString(e));
      return WOS_ERR_NOMEM;
    }
    for (auto& v : ev) WOS_HIP_TRY(hipEventCreateWithFlags(&v, hipEventDisableTiming));
    return 0;
  }
};
static XferRing& xfer_ring() {
  static XferRing r;
  return r;
}

int host_to_dev(void* dst, const void* src, size_t bytes, hipStream_t s) {
  if (bytes < ((size_t)2 << 20)) {
    WOS_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    return 0;
  }
  XferRing& r = xfer_ring();
  int rc;
  if ((rc = r.ensure())) return rc;
  const size_t nch = (bytes + kChunk - 1) / kChunk;
  for (size_t i = 0; i < nch; ++i) {
    const int slot = (int)(i % kSlots);
    const size_t off = i * kChunk, len = std::min(kChunk, bytes - off);
    if ((rc = r.wait(slot))) return rc;  // the slot's previous DMA has finished
    CopyPool::get().copy(r.pin + slot * kChunk, (const char*)src + off, len);
    WOS_HIP_TRY(hipMemcpyAsync((char*)dst + off, r.pin + slot * kChunk, len, hipMemcpyHostToDevice, s));
    if ((rc = r.mark(slot, s))) return rc;
  }
  return 0;
}

int dev_to_host(void* dst, const void* src, size_t bytes, hipStream_t s) {
  if (bytes < ((size_t)2 << 20)) {
    WOS_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
    WOS_HIP_TRY(hipStreamSynchronize(s));
    return 0;
  }
  XferRing& r = xfer_ring();
  int rc;
  if ((rc = r.ensure())) return rc;
  const size_t nch = (bytes + kChunk - 1) / kChunk;
  auto issue = [&](size_t i) -> int {
    const int slot = (int)(i % kSlots);
    const size_t off = i * kChunk, len = std::min(kChunk, bytes - off);
    int rc2;
    if ((rc2 = r.wait(slot))) return rc2;  // a host-to-device DMA out of the slot may still run
    WOS_HIP_TRY(hipMemcpyAsync(r.pin + slot * kChunk, (const char*)src + off, len, hipMemcpyDeviceToHost, s));
    return r.mark(slot, s);
  };
  for (size_t i = 0; i < nch && i < (size_t)kSlots; ++i)
    if ((rc = issue(i))) return rc;
  for (size_t i = 0; i < nch; ++i) {
    const int slot = (int)(i % kSlots);
    const size_t off = i * kChunk, len = std::min(kChunk, bytes - off);
    if ((rc = r.wait(slot))) return rc;
    CopyPool::get().copy((char*)dst + off, r.pin + slot * kChunk, len);
    if (i + kSlots < nch && (rc = issue(i + kSlots))) return rc;
  }
  return 0;
}

int ensure_init() {
  Device& d = device();
  std::lock_guard<std::recursive_mutex> lk(d.mu);
  if (d.ready) return 0;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
    set_error("no HIP device");
    return WOS_ERR_NO_DEVICE;
  }
  WOS_HIP_TRY(hipSetDevice(0));
  hipDeviceProp_t prop;
  WOS_HIP_TRY(hipGetDeviceProperties(&prop, 0));
  snprintf(d.name, sizeof(d.name), "%s (%s)", prop.name, prop.gcnArchName);
  WOS_HIP_TRY(hipStreamCreateWithFlags(&d.stream, hipStreamNonBlocking));
  // this device's copy of the no-allocation scratch array
  void* na = nullptr;
  hipError_t e = hipGetSymbolAddress(&na, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !na) {
    set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    return WOS_ERR_HIP;
  }
  d.noalloc = na;
  d.noalloc_bytes = kNoAllocScratchBytes;
  d.ready = true;
  return 0;
}

// Host mirror of Plan<N, radix> (wos_fft.hpp): stage radices of the forward
// and inverse plans under the radix-16 or radix-8 schedule.
static void plan_stages(int n, int radix, std::vector<int>& fwd, std::vector<int>& inv) {
  fwd.clear();
  inv.clear();
  if (n < radix) {
    fwd.push_back(n);
    inv.push_back(n);
    return;
  }
  int lg = 0, lv = 0;
  while ((1 << lg) < n) ++lg;
  while ((1 << lv) < radix) ++lv;
  const int nf = lg / lv, rem = 1 << (lg % lv);
  for (int i = 0; i < nf; ++i) fwd.push_back(radix);
  if (rem > 1) {
    fwd.push_back(rem);
    inv.push_back(rem);
  }
  for (int i = 0; i < nf; ++i) inv.push_back(radix);
}

// Stages of the twiddle table of an n-point transform under the radix-16 or
// radix-8 schedule: forward stages 1..S-1, then inverse stages 1..S-1, the
// layout Plan<N>::tw_off_* addresses. Returns the stage count (at most
// kMaxTwStages) and the table's entry count in *entries.
struct TwStage {
  int off, ns, rad;
};
constexpr int kMaxTwStages = 32;
static int twiddle_stages(int n, int radix, TwStage* st, int* entries) {
  std::vector<int> fwd, inv;
  plan_stages(n, radix == 8 ? 8 : 16, fwd, inv);
  int off = 0, cnt = 0;
  for (const std::vector<int>* plan : {&fwd, &inv}) {
    int ns = (*plan)[0];
    for (size_t i = 1; i < plan->size() && cnt < kMaxTwStages; ++i) {
      const int rad = (*plan)[i];
      st[cnt++] = {off, ns, rad};
      off += ns * (rad - 1);
      ns *= rad;
    }
  }
  *entries = off;
  return cnt;
}

// Generate the twiddle table of an n-point transform into p (entries of
// csize(prec) bytes, twice that for acc) in the order of stream s.
static int launch_twiddles(void* p, int n, int prec, int acc, int radix, hipStream_t s) {
  TwStage st[kMaxTwStages];
  int entries;
  const int ns = twiddle_stages(n, radix, st, &entries);
  for (int i = 0; i < ns; ++i) {
    const TwStage& q = st[i];
    const int cnt = q.ns * (q.rad - 1), bs = 256, gs = (cnt + bs - 1) / bs;
    if (prec == WOS_FP64)
      hipLaunchKernelGGL(k_stage_twiddles<double>, dim3(gs), dim3(bs), 0, s, (cpx<double>*)p, q.off, q.ns, q.rad, acc);
    else
      hipLaunchKernelGGL(k_stage_twiddles<float>, dim3(gs), dim3(bs), 0, s, (cpx<float>*)p, q.off, q.ns, q.rad, acc);
    WOS_HIP_TRY(hipGetLastError());
  }
  return 0;
}

// Stage twiddle tables of an n-point transform, cached per (n, precision,
// acc, radix). acc: [hi, lo] pairs for the compensated split-step kernels;
// radix: the stage schedule (16 or 8) whose stages the table holds.
int get_twiddles(int n, int prec, const void** out, int acc, int radix) {
  Device& d = device();
  if (radix != 8) radix = 16;
  long long key = (((long long)n * 2 + prec) * 2 + (acc ? 1 : 0)) * 2 + (radix == 8 ? 1 : 0);
  auto it = d.twiddles.find(key);
  if (it != d.twiddles.end()) {
    *out = it->second;
    return 0;
  }
  TwStage st[kMaxTwStages];
  int entries;
  twiddle_stages(n, radix, st, &entries);
  void* p = nullptr;
  WOS_HIP_TRY(hipMalloc(&p, (size_t)std::max(entries, 1) * csize(prec) * (acc ? 2 : 1)));
  int rc = launch_twiddles(p, n, prec, acc, radix, d.stream);
  if (rc) {
    (void)hipFree(p);
    return rc;
  }
  // consumers run on other streams (batches, pipelines): the table is
  // complete before its first use
  WOS_HIP_TRY(hipStreamSynchronize(d.stream));
  d.twiddles[key] = p;
  *out = p;
  return 0;
}

// Host mirror of MPlan<N> and mixed_rt_plan (wos_mixed.hpp,
// wos_launch_mixed.inc): stage radices of the mixed-radix plan of
// n = 2^a 3^b 5^c 7^d points.
static std::vector<int> mixed_plan_host(int n) {
  int a = 0, b = 0, c = 0, d = 0, m = n;
  while (m % 2 == 0) m /= 2, ++a;
  while (m % 3 == 0) m /= 3, ++b;
  while (m % 5 == 0) m /= 5, ++c;
  while (m % 7 == 0) m /= 7, ++d;
  std::vector<int> rad;
  for (int i = 0; i < a / 4; ++i) rad.push_back(16);
  if (a % 4) rad.push_back(1 << (a % 4));
  for (int i = 0; i < b; ++i) rad.push_back(3);
  for (int i = 0; i < c; ++i) rad.push_back(5);
  for (int i = 0; i < d; ++i) rad.push_back(7);
  return rad;
}

// Host mirror of MRDef<N> (wos_mixreg.hpp): forward stage radices of the
// register-resident column pass where one is built, else empty.
static std::vector<int> mixed_reg_plan_host(int n) {
  if (n == 2000) return {10, 10, 10, 2};
  if (n == 1536) return {8, 8, 8, 3};
  if (n == 1000) return {10, 10, 10};
  return {};
}

// Stage twiddle tables of the mixed-radix plan of n points (stages 1..S-1,
// the layout MPlan<N>::tw_off addresses), followed, where a register-resident
// column pass is built, by its forward stages 1..S-1 and inverse stages
// 1..S-1 (MRPlan<N>::tw_off): entry [k][r-1] of stage (NS, RAD) is
// exp(-2 pi i r k/(NS RAD)), formed in long double from the reduced fraction
// (r k mod NS RAD)/(NS RAD) and rounded once; cached.
static int get_mixed_twiddles(int n, int prec, const void** out) {
  Device& d = device();
  const long long key = (1LL << 40) + (long long)n * 2 + prec;
  auto it = d.twiddles.find(key);
  if (it != d.twiddles.end()) {
    *out = it->second;
    return 0;
  }
  std::vector<double> tab;  // (re, im) pairs
  auto add_stages = [&tab](const std::vector<int>& rad) {
    long long ns = rad[0];
    for (size_t s = 1; s < rad.size(); ++s) {
      const long long L = ns * rad[s];
      for (long long k = 0; k < ns; ++k)
        for (int r = 1; r < rad[s]; ++r) {
          const long double ang =
              -2.0L * 3.141592653589793238462643383279502884L * (long double)((r * k) % L) / (long double)L;
          tab.push_back((double)cosl(ang));
          tab.push_back((double)sinl(ang));
        }
      ns = L;
    }
  };
  add_stages(mixed_plan_host(n));
  const std::vector<int> reg = mixed_reg_plan_host(n);
  if (!reg.empty()) {
    add_stages(reg);
    add_stages(std::vector<int>(reg.rbegin(), reg.rend()));
  }
  const size_t cnt = std::max<size_t>(tab.size() / 2, 1);
  void* p = nullptr;
  WOS_HIP_TRY(hipMalloc(&p, cnt * csize(prec)));
  if (prec == WOS_FP64) {
    if (!tab.empty()) WOS_HIP_TRY(hipMemcpy(p, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));
  } else {
    std::vector<float> tf(tab.begin(), tab.end());
    if (!tf.empty()) WOS_HIP_TRY(hipMemcpy(p, tf.data(), tf.size() * sizeof(float), hipMemcpyHostToDevice));
  }
  d.twiddles[key] = p;
  *out = p;
  return 0;
}

// widest column tile of the mixed-radix kernels for n (0: n not built)
static int mixed_col_w(int prec, int n) {
  if (n <= 1 || pow2(n)) return 0;
  return prec == WOS_FP64 ? mixed_col_w_f64(n) : mixed_col_w_f32(n);
}
// whether the mixed-radix column pass of n is built with tile width W
static bool mixed_col_valid(int prec, int n, int W) {
  if (mixed_col_w(prec, n) <= 0) return false;
  return prec == WOS_FP64 ? mixed_col_valid_f64(n, W) : mixed_col_valid_f32(n, W);
}
// a mixed-radix shape: layout tiles of tw columns (a power of two up to 16
// dividing n) and column workgroups of cw columns (a built width dividing tw)
static bool mixed_shape_valid(int prec, int n, int tw, int cw) {
  return tw > 0 && tw <= 16 && pow2(tw) && n % tw == 0 && cw > 0 && tw % cw == 0 && mixed_col_valid(prec, n, cw);
}
static int launch_rows_mixed(int prec, int N, int dir, const RowArgs& a, hipStream_t s) {
  return prec == WOS_FP64 ? launch_rows_mixed_f64(N, dir, a, s) : launch_rows_mixed_f32(N, dir, a, s);
}
// mode 0: one column transform (dir); mode 2: forward -> TF -> inverse
static int launch_cols_mixed(int prec, int N, int mode, int dir, int W, const ColArgs& a, hipStream_t s) {
  return prec == WOS_FP64 ? launch_cols_mixed_f64(N, mode, dir, W, a, s) : launch_cols_mixed_f32(N, mode, dir, W, a, s);
}

__global__ void k_convert_scale_f64(const cpx<double>* __restrict__ in, cpx<double>* __restrict__ out, int n, double sc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = {in[i].x * sc, in[i].y * sc};
}
__global__ void k_convert_scale_f32(const cpx<double>* __restrict__ in, cpx<float>* __restrict__ out, int n, double sc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = {(float)(in[i].x * sc), (float)(in[i].y * sc)};
}

int launch_rows(int prec, int N, int dir, const RowArgs& a, hipStream_t s, int rows);

// b[k] of the Bluestein convolution (optics/fft.go:182-187) on M points:
// b[0] = 1, b[k] = b[M-k] = conj(c[k]) for 0 < k < L, else 0 (FP64).
__global__ void k_bs_b0(const cpx<double>* __restrict__ chirp, cpx<double>* __restrict__ b, int L, int M) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  cpx<double> v = {0.0, 0.0};
  if (i == 0) v = {1.0, 0.0};
  else if (i < L) v = {chirp[i].x, -chirp[i].y};
  else if (i > M - L) v = {chirp[M - i].x, -chirp[M - i].y};
  b[i] = v;
}

// chirp[k] = cexpI(sign*math.Pi*float64(k*k)/float64(L)) (optics/fft.go:174-176)
__global__ void k_bs_chirp(cpx<double>* __restrict__ chirp, int L, double sign) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= L) return;
  double ph;
  {
#pragma clang fp contract(off)
    ph = sign * M_PI * (double)((long long)k * k) / (double)L;
  }
  double s, c;
  sincos(ph, &s, &c);
  chirp[k] = {c, s};
}

// Bluestein tables of (L, direction) in precision prec: chirp[L] and
// FFT_M(b)/M, generated in FP64 on the device (the spectrum by the engine's
// own M-point FFT) and rounded to prec. Cached.
static int get_bs_tables(int L, int inverse, int prec, BsTab* out) {
  Device& d = device();
  const long long key = ((long long)L * 2 + inverse) * 2 + prec;
  auto it = d.chirps.find(key);
  if (it != d.chirps.end()) {
    out->chirp = it->second;
    out->bspec = (const char*)it->second + (size_t)L * csize(prec);
    return 0;
  }
  const int M = bs_len(L);
  if (!long_len_ok(L)) {
    set_error("Bluestein length not supported: " + std::to_string(L));
    return WOS_ERR_UNSUPPORTED;
  }
  int rc;
  void* tmp = nullptr;  // FP64 chirp[L] then b[M] (then M of four-step workspace for M > 8192)
  WOS_HIP_TRY(hipMalloc(&tmp, (size_t)(L + M + (M > max_pow2_len() ? M : 0)) * 16));
  cpx<double>* c64 = (cpx<double>*)tmp;
  cpx<double>* b64 = c64 + L;
  hipLaunchKernelGGL(k_bs_chirp, dim3((L + 255) / 256), dim3(256), 0, d.stream, c64, L, inverse ? 1.0 : -1.0);
  hipLaunchKernelGGL(k_bs_b0, dim3((M + 255) / 256), dim3(256), 0, d.stream, (const cpx<double>*)c64, b64, L, M);
  WOS_HIP_TRY(hipGetLastError());
  if (M > max_pow2_len()) {  // the four-step M-point transform
    if ((rc = long_pow2_dev(WOS_FP64, b64, b64 + M, M, 1, 0, 1.0, d.stream))) {
      (void)hipFree(tmp);
      return rc;
    }
  } else {
    const void* tw;
    if ((rc = get_twiddles(M, WOS_FP64, &tw))) return rc;
    RowArgs ra = row_args_init();
    ra.in = b64;
    ra.out = b64;
    ra.in_unit = ra.out_unit = M;
    ra.in_ld = ra.out_ld = M;
    ra.rows = 1;
    ra.units = 1;
    ra.in_cols = M;
    ra.out_cols = M;
    ra.tw = tw;
    if ((rc = launch_rows(WOS_FP64, M, -1, ra, d.stream, 0))) return rc;
  }
  void* p = nullptr;
  WOS_HIP_TRY(hipMalloc(&p, (size_t)(L + M) * csize(prec)));
  if (prec == WOS_FP64) {
    hipLaunchKernelGGL(k_convert_scale_f64, dim3((L + 255) / 256), dim3(256), 0, d.stream, (const cpx<double>*)c64, (cpx<double>*)p, L, 1.0);
    hipLaunchKernelGGL(k_convert_scale_f64, dim3((M + 255) / 256), dim3(256), 0, d.stream, (const cpx<double>*)b64, (cpx<double>*)p + L, M, 1.0 / M);
  } else {
    hipLaunchKernelGGL(k_convert_scale_f32, dim3((L + 255) / 256), dim3(256), 0, d.stream, (const cpx<double>*)c64, (cpx<float>*)p, L, 1.0);
    hipLaunchKernelGGL(k_convert_scale_f32, dim3((M + 255) / 256), dim3(256), 0, d.stream, (const cpx<double>*)b64, (cpx<float>*)p + L, M, 1.0 / M);
  }
  WOS_HIP_TRY(hipGetLastError());
  WOS_HIP_TRY(hipStreamSynchronize(d.stream));
  (void)hipFree(tmp);
  d.chirps[key] = p;
  out->chirp = p;
  out->bspec = (const char*)p + (size_t)L * csize(prec);
  return 0;
}

// In-place row transforms of `units` x rows rows of length L (row-major, ld)
// in either engine path.
static int rows_any(int prec, void* d, int L, int rows, int ld, long long unit, int units, int inverse, double scale,
                    hipStream_t s, double* pw_in = nullptr, double* pw_out = nullptr) {
  int rc;
  RowArgs ra = row_args_init();
  ra.pw_in = pw_in;
  ra.pw_out = pw_out;
  ra.in = d;
  ra.out = d;
  ra.in_unit = ra.out_unit = unit;
  ra.in_ld = ra.out_ld = ld;
  ra.rows = rows;
  ra.units = units;
  ra.in_cols = L;
  ra.out_cols = L;
  if (mixed_col_w(prec, L) > 0) {  // 2^a 3^b 5^c lengths with a built mixed-radix plan
    const void* tw;
    if ((rc = get_mixed_twiddles(L, prec, &tw))) return rc;
    ra.tw = tw;
    ra.scale_re = scale;
    return launch_rows_mixed(prec, L, inverse ? 1 : -1, ra, s);
  }
  if (pow2(L)) {
    const void* tw;
    if ((rc = get_twiddles(L, prec, &tw))) return rc;
    ra.tw = tw;
    ra.scale_re = scale;
    return launch_rows(prec, L, inverse ? 1 : -1, ra, s, 0);
  }
  const int M = bs_len(L);
  BsRowArgs A;
  A.io = ra;
  A.len = L;
  A.post = scale;
  if ((rc = get_bs_tables(L, inverse, prec, &A.tab))) return rc;
  const void* tw;
  if ((rc = get_twiddles(M, prec, &tw))) return rc;
  A.io.tw = tw;
  return launch_rows_bs(prec, M, A, s);
}

int launch_rows(int prec, int N, int dir, const RowArgs& a, hipStream_t s, int rows) {
  if (a.radix == 8)
    return prec == WOS_FP64 ? launch_rows8_f64(N, dir, rows, a, s) : launch_rows8_f32(N, dir, rows, a, s);
  return prec == WOS_FP64 ? launch_rows_f64(N, dir, rows, a, s) : launch_rows_f32(N, dir, rows, a, s);
}
bool rows_valid(int prec, int N, int rows, int radix) {
  if (radix == 8) return prec == WOS_FP64 ? rows_valid8_f64(N, rows) : rows_valid8_f32(N, rows);
  return prec == WOS_FP64 ? rows_valid_f64(N, rows) : rows_valid_f32(N, rows);
}
// the pipelined row pass with split exchanges (twiddles in LDS), radix 16
bool rows_pipe_lds_valid(int prec, int N, int rows) {
  return prec == WOS_FP64 ? rows_pipe_lds_valid_f64(N, rows) : rows_pipe_lds_valid_f32(N, rows);
}
bool col_w_valid(int prec, int N, int W, int radix) {
  if (radix == 8) return prec == WOS_FP64 ? col_w_valid8_f64(N, W) : col_w_valid8_f32(N, W);
  if (radix == 32) return prec == WOS_FP64 ? col_w_valid32_f64(N, W) : col_w_valid32_f32(N, W);
  return prec == WOS_FP64 ? col_w_valid_f64(N, W) : col_w_valid_f32(N, W);
}
int launch_cols(int prec, int N, int mode, int dir, int W, const ColArgs& a, hipStream_t s) {
  if (W <= 0) W = default_col_w(prec, N, a.cols);
  if (a.radix == 8) {
    if (mode != 1 && mode != 3) {
      set_error("the radix-8 schedule builds the asm and split-step column passes only");
      return WOS_ERR_UNSUPPORTED;
    }
    return prec == WOS_FP64 ? launch_cols8_f64(N, mode, W, a, s) : launch_cols8_f32(N, mode, W, a, s);
  }
  if (a.radix == 32) return prec == WOS_FP64 ? launch_cols32_f64(N, mode, W, a, s) : launch_cols32_f32(N, mode, W, a, s);
  return prec == WOS_FP64 ? launch_cols_f64(N, mode, dir, W, a, s) : launch_cols_f32(N, mode, dir, W, a, s);
}
int default_col_w(int prec, int N, int cols) {
  return prec == WOS_FP64 ? default_col_w_f64(N, cols) : default_col_w_f32(N, cols);
}
static int launch_rows_bs(int prec, int M, const BsRowArgs& A, hipStream_t s) {
  return prec == WOS_FP64 ? launch_rows_bs_f64(M, A, s) : launch_rows_bs_f32(M, A, s);
}
static int launch_cols_bs(int prec, int M, int mode, int W, const BsColArgs& A, hipStream_t s) {
  return prec == WOS_FP64 ? launch_cols_bs_f64(M, mode, W, A, s) : launch_cols_bs_f32(M, mode, W, A, s);
}
static int bs_default_w(int prec, int M, int cols) {
  return prec == WOS_FP64 ? bs_default_w_f64(M, cols) : bs_default_w_f32(M, cols);
}

// rows handled by one row-pass workgroup (mirrors rows_per_block<N, radix>())
static int rows_per_block_host(int N, int radix = 16) {
  int V = N >= radix ? radix : N;
  int T = N / V;
  return T >= 256 ? 1 : 256 / T;
}

// ---------------------------------------------------------------------------
// Auxiliary kernels (pointwise, reductions).
// ---------------------------------------------------------------------------
__global__ void k_z2c(const cpx<double>* __restrict__ in, cpx<float>* __restrict__ out, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    out[i] = {(float)in[i].x, (float)in[i].y};
}
__global__ void k_c2z(const cpx<float>* __restrict__ in, cpx<double>* __restrict__ out, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    out[i] = {(double)in[i].x, (double)in[i].y};
}

// Radial spectral histogram of spectralRadius (optics/propagate.go:282-304):
// bin = int(hypot(fx, fy)/fnyq*96) clamped to 96, weights |F|^2; one
// 97-bin partial per workgroup.
template <typename R>
__global__ void k_spec_hist(const cpx<R>* __restrict__ spec, int n, size_t total, double den, double fnyq,
                            double* __restrict__ partial) {
  __shared__ double h[97];
  for (int i = threadIdx.x; i < 97; i += blockDim.x) h[i] = 0;
  __syncthreads();
  const size_t nn = (size_t)n * n;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
    cpx<R> v = spec[idx];
    double w = (double)v.x * (double)v.x + (double)v.y * (double)v.y;
    if (w == 0) continue;
    size_t rem = idx % nn;
    int j = (int)(rem / n), i = (int)(rem % n);
    double fr;
    {
#pragma clang fp contract(off)
      fr = go_hypot(freq_of(i, n, den), freq_of(j, n, den)) / fnyq * 96;
    }
    int b = (int)fr;
    if (b > 96) b = 96;
    atomicAdd(&h[b], w);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 97; i += blockDim.x) partial[(size_t)blockIdx.x * 97 + i] = h[i];
}

// Spectral centroid sums of the FP32 path (spectralCentroid,
// optics/propagate.go:391-416): per-workgroup FP64 partial sums of |U|^2,
// fx |U|^2 and fy |U|^2 over a fixed grid, summed by the host in workgroup
// order, so the result does not depend on scheduling. The FP32 field's own
// rounding (~1e-7) dominates any change of summation order.
constexpr int kCentroidBlocks = 1024;
template <typename R>
__global__ __launch_bounds__(256) void k_spec_centroid_partial(const cpx<R>* __restrict__ spec, int n, int ncomp,
                                                               double den, double* __restrict__ partial) {
  __shared__ double red[3][4];
  const size_t nn = (size_t)n * n, total = (size_t)ncomp * nn;
  double sw = 0, sx = 0, sy = 0;
  for (size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x; idx < total; idx += (size_t)gridDim.x * 256) {
    const size_t rem = idx % nn;
    const int j = (int)(rem / n), i = (int)(rem % n);
    const double re = (double)spec[idx].x, im = (double)spec[idx].y;
    const double w = re * re + im * im;
    sw += w;
    sx += w * freq_of(i, n, den);
    sy += w * freq_of(j, n, den);
  }
  double v[3] = {sw, sx, sy};
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v[q] += __shfl_xor(v[q], o, 64);
    if (lane == 0) red[q][wave] = v[q];
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const int q = threadIdx.x;
    partial[(size_t)blockIdx.x * 3 + q] = ((red[q][0] + red[q][1]) + red[q][2]) + red[q][3];
  }
}

// Spectral centroid sums of spectralCentroid (optics/propagate.go:391-416)
// for the FP64 path, in the reference's serial order (component, row, column),
// one lane accumulating while the other waves prefetch the next chunk into
// LDS. The carrier feeds a phase ramp whose result moves by ~4e-12 rel-L2
// per 1e-15 relative change of the centroid, so the reference's naive serial
// rounding is reproduced rather than a more accurate tree sum.
template <typename R>
__global__ __launch_bounds__(256) void k_spec_centroid_serial(const cpx<R>* __restrict__ spec, int n, int ncomp,
                                                              double den, double* __restrict__ out3) {
  constexpr int CH = 2048;
  __shared__ double fq[8192];
  __shared__ cpx<R> buf[2][CH];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  for (int i = tid; i < n; i += 256) fq[i] = freq_of(i, n, den);
  const size_t total = (size_t)ncomp * n * n;
  const size_t nch = (total + CH - 1) / CH;
  for (int i = tid; i < CH && (size_t)i < total; i += 256) buf[0][i] = spec[i];
  __syncthreads();
  double sw = 0, sx = 0, sy = 0;
  int ci = 0, cj = 0;  // column / row of the next element
  for (size_t c = 0; c < nch; ++c) {
    const size_t base = c * CH;
    const int len = (int)((total - base) < (size_t)CH ? (total - base) : (size_t)CH);
    if (wave != 0) {
      const size_t nb = base + CH;
      if (nb < total) {
        const int nlen = (int)((total - nb) < (size_t)CH ? (total - nb) : (size_t)CH);
        for (int i = tid - 64; i < nlen; i += 192) buf[(c + 1) & 1][i] = spec[nb + i];
      }
    } else if (lane == 0) {
      const cpx<R>* b = buf[c & 1];
      for (int e = 0; e < len; ++e) {
#pragma clang fp contract(off)
        const double re = (double)b[e].x, im = (double)b[e].y;
        const double w = re * re + im * im;
        sw += w;
        sx += w * fq[ci];
        sy += w * fq[cj];
        if (++ci == n) {
          ci = 0;
          if (++cj == n) cj = 0;
        }
      }
    }
    __syncthreads();
  }
  if (tid == 0) {
    out3[0] = sw;
    out3[1] = sx;
    out3[2] = sy;
  }
}

// Carrier ramp of propASMShift (optics/propagate.go:437-444):
// a *= cexpI(sgn*2*pi*(fxc*X(i) + fyc*y)), phase formed in FP64 in the
// reference's order.
template <typename R>
__global__ void k_ramp(cpx<R>* __restrict__ a, int n, int units, double dx, double fxc, double fyc, double sgn) {
  const size_t nn = (size_t)n * n, total = nn * units;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
    size_t rem = idx % nn;
    int j = (int)(rem / n), i = (int)(rem % n);
    double ph;
    {
#pragma clang fp contract(off)
      double y = ((double)j - (double)n / 2) * dx;
      double x = ((double)i - (double)n / 2) * dx;
      ph = sgn * 2 * M_PI * (fxc * x + fyc * y);
    }
    double s, c;
    sincos(ph, &s, &c);
    cpx<double> v = {(double)a[idx].x, (double)a[idx].y};
    cpx<double> t = {c, s};
    cpx<double> r;
    {
#pragma clang fp contract(off)
      r = cmul(v, t);
    }
    a[idx] = {(R)r.x, (R)r.y};
  }
}

// BPM screen table of applyMediumPhase (optics/medium.go:50-65):
// t = complex(exp(-k*Im n*dz), 0) * cexpI(k*(Re n - 1)*dz), built in FP64.
__global__ void k_screen_table(const cpx<double>* __restrict__ index, cpx<double>* __restrict__ t, size_t nn,
                               double k, double dz) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += (size_t)gridDim.x * blockDim.x) {
#pragma clang fp contract(off)
    cpx<double> nv = index[i];
    double ph = k * (nv.x - 1) * dz;
    double e = exp(-k * nv.y * dz);
    double s, c;
    sincos(ph, &s, &c);
    cpx<double> ee = {e, 0.0};
    cpx<double> cc = {c, s};
    t[i] = cmul(ee, cc);
  }
}

// a[u] *= t (t shared by all units); FP64 multiply, stored in R.
template <typename R>
__global__ void k_mul_table(cpx<R>* __restrict__ a, const cpx<double>* __restrict__ t, size_t nn, int units) {
  const size_t total = nn * units;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
    cpx<double> v = {(double)a[idx].x, (double)a[idx].y};
    cpx<double> r;
    {
#pragma clang fp contract(off)
      r = cmul(v, t[idx % nn]);
    }
    a[idx] = {(R)r.x, (R)r.y};
  }
}

// a[u] *= t for complex64 fields and table
__global__ void k_mul_table32(cpx<float>* __restrict__ a, const cpx<float>* __restrict__ t, size_t nn, int units) {
  const size_t total = nn * units;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x)
    a[idx] = cmul(a[idx], t[idx % nn]);
}

// out[i] += w * Intensity(i) over the components of one field
// (optics/polychromatic.go:36-38, optics/field.go:65-74).
template <typename R>
__global__ void k_accum_intensity(const cpx<R>* __restrict__ a, int ncomp, size_t nn, double w, double* __restrict__ out) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += (size_t)gridDim.x * blockDim.x) {
#pragma clang fp contract(off)
    double p = 0;
    for (int c = 0; c < ncomp; ++c) {
      cpx<R> v = a[(size_t)c * nn + i];
      double re = v.x, im = v.y;
      if (c == 0) p = re * re + im * im;
      else p += re * re + im * im;
    }
    out[i] += w * p;
  }
}

// Fresnel impulse response of propFresnelIR (optics/propagate.go:543-558) on
// the 2N grid in wrap-around layout: cexpI(pi*(x^2 + y^2)/(wl*z)).
template <typename R>
__global__ void k_fresnel_ir_kernel(cpx<R>* __restrict__ h, int m, double dx, double wl, double z) {
  const size_t total = (size_t)m * m;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
    int j = (int)(idx / m), i = (int)(idx % m);
    double ph;
    {
#pragma clang fp contract(off)
      double dy = (double)j;
      if (j > m / 2) dy -= (double)m;
      double y = dy * dx;
      double dxi = (double)i;
      if (i > m / 2) dxi -= (double)m;
      double x = dxi * dx;
      ph = M_PI * (x * x + y * y) / (wl * z);
    }
    double s, c;
    sincos(ph, &s, &c);
    h[idx] = {(R)c, (R)s};
  }
}

// asmTF1D (optics/propagate.go:679-701) applied to a 1-D spectrum.
template <typename R>
__global__ void k_tf1d(cpx<R>* __restrict__ a, int n, double dx, TFUnit u) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    cpx<double> h;
    {
#pragma clang fp contract(off)
      double fx = i <= n / 2 ? (double)i / ((double)n * dx) : (double)(i - n) / ((double)n * dx);
      double rho = fabs(u.wl * fx);
      if (rho <= 1) {
        double s, c;
        sincos(u.k * u.z * sqrt((1 - rho) * (1 + rho)), &s, &c);
        h = {c, s};
      } else {
        double decay = u.k * u.az * sqrt((rho - 1) * (rho + 1));
        if (u.zero_ev || (u.limit > 0 && decay > u.limit)) h = {0.0, 0.0};
        else if (u.backward) {
          if (u.reg) {
            double e = exp(decay);
            h = {e / (1 + u.alpha * u.alpha * e * e), 0.0};
          } else h = {exp(decay), 0.0};
        } else h = {exp(-decay), 0.0};
      }
    }
    cpx<double> v = {(double)a[i].x, (double)a[i].y};
    cpx<double> r;
    {
#pragma clang fp contract(off)
      r = cmul(v, h);
    }
    a[i] = {(R)r.x, (R)r.y};
  }
}

// Counter-based standard normal (SplitMix64 pair, Box-Muller), the same
// generator the oracle uses in place of Go's math/rand stream.
__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ULL;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
  return x ^ (x >> 31);
}
__device__ __forceinline__ double normal_at(uint64_t seed, uint64_t counter) {
  uint64_t h1 = splitmix64(seed ^ splitmix64(2 * counter));
  uint64_t h2 = splitmix64(seed ^ splitmix64(2 * counter + 1));
  double u1 = ((double)(h1 >> 11) + 0.5) / 9007199254740992.0;
  double u2 = (double)(h2 >> 11) / 9007199254740992.0;
  return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

// Degree of coherence mu on the wrapped difference grid
// (optics/coherence.go:29-45), FP64.
__global__ void k_schell_mu(cpx<double>* __restrict__ s, int n, double dx, double sigma_c) {
  const size_t nn = (size_t)n * n;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < nn; idx += (size_t)gridDim.x * blockDim.x) {
#pragma clang fp contract(off)
    int j = (int)(idx / n), i = (int)(idx % n);
    double dy = (j <= n / 2) ? (double)j * dx : (double)(j - n) * dx;
    double d = (i <= n / 2) ? (double)i * dx : (double)(i - n) * dx;
    s[idx] = {exp(-(d * d + dy * dy) / (2 * sigma_c * sigma_c)), 0.0};
  }
}
__global__ void k_schell_sqrt(cpx<double>* __restrict__ s, size_t nn) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += (size_t)gridDim.x * blockDim.x) {
    double v = s[i].x;
    if (v < 0) v = 0;
    s[i] = {sqrt(v), 0.0};
  }
}
// complex Gaussian noise / sqrt(2) times sqrt(S) (optics/coherence.go:63-73)
__global__ void k_schell_noise(cpx<double>* __restrict__ out, const cpx<double>* __restrict__ sq, size_t nn,
                               uint64_t seed, uint64_t realization) {
  const double inv_sqrt2 = 0.70710678118654752440;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += (size_t)gridDim.x * blockDim.x) {
#pragma clang fp contract(off)
    uint64_t c = (realization * (uint64_t)nn + i) * 2;
    cpx<double> v = {normal_at(seed, c) * inv_sqrt2, normal_at(seed, c + 1) * inv_sqrt2};
    out[i] = cmul(v, sq[i]);
  }
}
// times sqrt(I), I = exp(-2 r^2 / w0^2) (optics/coherence.go:54-62, 74-77), stored in R
template <typename R>
__global__ void k_schell_envelope(const cpx<double>* __restrict__ in, cpx<R>* __restrict__ out, int n, double dx,
                                  double w0) {
  const size_t nn = (size_t)n * n;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < nn; idx += (size_t)gridDim.x * blockDim.x) {
#pragma clang fp contract(off)
    int j = (int)(idx / n), i = (int)(idx % n);
    double y = ((double)j - (double)n / 2) * dx;
    double x = ((double)i - (double)n / 2) * dx;
    double a = sqrt(exp(-2 * (x * x + y * y) / (w0 * w0)));
    cpx<double> v = cmul(in[idx], cpx<double>{a, 0.0});
    out[idx] = {(R)v.x, (R)v.y};
  }
}

static unsigned grid_for(size_t total, int bs = 256) {
  size_t g = (total + bs - 1) / bs;
  if (g > 65536) g = 65536;
  if (g == 0) g = 1;
  return (unsigned)g;
}

// ---------------------------------------------------------------------------
// Host helpers.
// ---------------------------------------------------------------------------
struct Neumaier {
  double s = 0, c = 0;
  void add(double x) {
    double t = s + x;
    if (std::fabs(s) >= std::fabs(x)) c += (s - t) + x;
    else c += (x - t) + s;
    s = t;
  }
  double value() const { return s + c; }
};

static void warn_add(wos_warnings* w, int code, double value) {
  if (!w) return;
  w->count[code]++;
  if (value != 0) w->value[code] = value;
}

// evanescentPolicy (optics/propagate.go:186-199)
static void evanescent_policy(const wos_ctx& ctx, double z, int* zero, int* reg, double* alpha, double* limit) {
  *limit = ctx.evanescent_limit;
  if (z < 0) {
    if (ctx.backward_regularize) {
      *alpha = ctx.tikhonov_alpha;
      if (*alpha <= 0) *alpha = 1e-3;
      *zero = 0;
      *reg = 1;
      return;
    }
    *zero = 1;
    *reg = 0;
    *alpha = 0;
    return;
  }
  *zero = ctx.evanescent_zero != 0;
  *reg = 0;
  *alpha = 0;
}

static TFUnit make_asm_tf(double wl, double z, const wos_ctx& ctx, int M, double dx, double fxo, double fyo) {
  TFUnit u;
  memset(&u, 0, sizeof(u));
  int zero, reg;
  double alpha, limit;
  evanescent_policy(ctx, z, &zero, &reg, &alpha, &limit);
  u.wl = wl;
  u.z = z;
  u.k = 2.0 * M_PI / wl;
  u.az = std::fabs(z);
  u.alpha = alpha;
  u.limit = limit;
  u.fx_off = fxo;
  u.fy_off = fyo;
  u.den = (double)M * dx;
  u.inv_den = 1.0 / u.den;
  u.kz = u.k * z;
  u.M = M;
  u.kind = TF_ASM;
  u.zero_ev = zero;
  u.reg = reg;
  u.backward = z < 0;
  return u;
}

// asmWarn (optics/propagate.go:203-218)
static void asm_warn(wos_warnings* w, double z, double p_in, double p_out, int zero_ev, int reg) {
  if (p_in > 0 && p_out < p_in * (1 - 1e-8) && z > 0) warn_add(w, WOS_WARN_EVANESCENT_FILTERED, (p_in - p_out) / p_in);
  if (z < 0) {
    if (reg) warn_add(w, WOS_WARN_BACKWARD_REGULARIZED, 0);
    else if (zero_ev) warn_add(w, WOS_WARN_BACKWARD_EVANESCENT, 0);
  }
}

// ---------------------------------------------------------------------------
// Device-resident building blocks.
// ---------------------------------------------------------------------------
// In-place 2-D transform of `units` n x n fields (power-of-two n).
static int fft2d_dev(int prec, void* d, int n, int units, int inverse) {
  Device& dv = device();
  if (!len_ok(n)) {
    set_error("2-D FFT length not supported: " + std::to_string(n));
    return WOS_ERR_UNSUPPORTED;
  }
  const double sc = inverse ? 1.0 / (double)n : 1.0;
  int rc;
  // rows (fft2DCPU order: rows, then columns; inverse scales 1/n per pass)
  if ((rc = rows_any(prec, d, n, n, n, (long long)n * n, units, inverse, sc, dv.stream))) return rc;
  ColArgs ca;
  memset(&ca, 0, sizeof(ca));
  ca.data = d;
  ca.unit = (long long)n * n;
  ca.ld = n;
  ca.cols = n;
  ca.units = units;
  ca.in_row0 = 0;
  ca.rows = n;
  ca.out_row0 = 0;
  ca.out_rows = n;
  if (pow2(n)) {
    const void* tw;
    if ((rc = get_twiddles(n, prec, &tw))) return rc;
    ca.scale = sc;
    ca.tw = tw;
    return launch_cols(prec, n, 0, inverse ? 1 : -1, 0, ca, dv.stream);
  }
  if (mixed_col_w(prec, n) > 0) {
    const void* tw;
    if ((rc = get_mixed_twiddles(n, prec, &tw))) return rc;
    ca.scale = sc;
    ca.tw = tw;
    return launch_cols_mixed(prec, n, 0, inverse ? 1 : -1, mixed_col_w(prec, n), ca, dv.stream);
  }
  const int M = bs_len(n);
  BsColArgs A;
  A.io = ca;
  A.io.scale = 1.0;
  A.len = n;
  A.post0 = sc;
  A.post1 = 1.0;
  if ((rc = get_bs_tables(n, inverse, prec, &A.tab0))) return rc;
  A.tab1 = A.tab0;
  const void* tw;
  if ((rc = get_twiddles(M, prec, &tw))) return rc;
  A.io.tw = tw;
  const int W = bs_default_w(prec, M, n);
  A.io.cols = n;
  return launch_cols_bs(prec, M, 0, W, A, dv.stream);
}

struct TransferOut {
  std::vector<double> p_in, p_out;  // per field group (sum of |.|^2, no dx^2)
};

// Launch shape of one transfer step: tw = column-tile width of the
// intermediate (also the column pass's W), rows = rows per row-pass
// workgroup, xcd_remap = XCD-aware order of the column workgroups.
struct Tuning {
  int tw = 0;
  int rows = 0;
  int xcd_remap = 0;
  int groups = 0;  // row groups per workgroup of the pipelined row kernels (<= 1: plain)
  int col_w = 0;   // columns per column-pass workgroup, dividing tw (0: tw)
  int split = 0;   // one-group row kernels exchange real and imaginary parts separately
  int radix = 0;      // stage schedule of the row passes: 8 radix-8, else radix-16
  int col_radix = 0;  // stage schedule of the column pass: 8 radix-8, 32 radix-16 with 32 values per thread, else radix-16
  int col_split = 0;  // radix-16 asm column pass: exchanges through half the LDS
  // the inverse row pass's rows per workgroup, row groups and split
  // exchanges (it reads the tiled layout the forward pass writes, and the
  // two passes can prefer different shapes); inv_rows < 0: the forward
  // pass's three values
  int inv_rows = -1;
  int inv_groups = 0;
  int inv_split = 0;
};

// The inverse row pass takes the forward pass's shape unless it has its own.
static Tuning inv_from_fwd(Tuning t) {
  if (t.inv_rows < 0) {
    t.inv_rows = t.rows;
    t.inv_groups = t.groups;
    t.inv_split = t.split;
  }
  return t;
}

// The radix-8 schedule is built for the asm transfer kernels from 256 points.
static bool radix8_ok(int M) { return pow2(M) && M >= 256 && M <= 8192; }
// The 32-value column schedule (radix-16 stages, split asm column pass).
static bool col32_ok(int M) { return M == 1024 || M == 2048 || M == 4096; }

// Fill unset fields: the tile width and rows per workgroup that make one
// tile chunk of ROWS x TW elements a whole 128-byte line.
static Tuning resolve_tuning(int prec, int M, Tuning t) {
  const int S = (int)csize(prec);
  t.radix = t.radix == 8 && radix8_ok(M) ? 8 : 16;
  t.col_radix = t.col_radix == 8 && radix8_ok(M) ? 8 : (t.col_radix == 32 && col32_ok(M) ? 32 : 16);
  if (t.tw <= 0) {
    int cand[] = {prec == WOS_FP32 ? 4 : 2, 4, 8, 16, 2, 1};
    for (int w : cand)
      if (w <= M && (M % w) == 0 && col_w_valid(prec, M, w)) {
        t.tw = w;
        break;
      }
  }
  if (t.col_w <= 0 || t.col_w > t.tw || (t.tw % t.col_w) != 0 || !col_w_valid(prec, M, t.col_w, t.col_radix)) {
    t.col_w = t.tw;  // the widest column workgroup the schedule builds
    while (t.col_w > 1 && !col_w_valid(prec, M, t.col_w, t.col_radix)) t.col_w /= 2;
  }
  auto default_rows = [&]() {
    const int rd = rows_per_block_host(M, t.radix);
    int want = std::max(rd, 128 / (S * std::max(1, t.tw)));
    while (want > rd && !rows_valid(prec, M, want, t.radix)) want /= 2;
    return rows_valid(prec, M, want, t.radix) ? want : rd;
  };
  if (t.rows <= 0) t.rows = default_rows();
  t = inv_from_fwd(t);
  if (t.inv_rows == 0) t.inv_rows = default_rows();
  return t;
}

// The three launches of one transfer step (forward rows, fused columns,
// inverse rows) on `units` fields of n x n: power-of-two grid side M = n, or
// M = 2n zero-padded with the centre crop. inter is the column-tiled
// intermediate of units * inter_unit_elems(n, M) elements and must not alias d.
struct TransferLaunch {
  int prec = 0, M = 0, mode = 1;
  Tuning tu;
  RowArgs fwd, inv;
  ColArgs col;
  // non-power-of-two M: the mixed-radix kernels (mr) where built, else
  // Bluestein kernels with internal length Mb
  bool mr = false;
  bool bs = false;
  int Mb = 0;
  BsRowArgs bfwd, binv;
  BsColArgs bcol;
};

// Column-tiled intermediate: tiles of tw columns, each n*tw elements plus
// kTilePad, so the tiles a row pass touches together do not start at
// power-of-two distances (which concentrate on the same HBM channels).
constexpr int kTilePad = 32;
static inline long long inter_unit_elems(int n, int M) { return (long long)n * M + (long long)M * kTilePad; }

static int build_transfer(TransferLaunch& L, int prec, void* d, void* inter, int n, int units, bool pad,
                          const TFUnit* d_tf, int tf_kind, double post_re, double post_im, double* pw_in,
                          double* pw_out, Tuning tu, const void* const* tabs = nullptr) {
  const int M = pad ? 2 * n : n;
  const int off = pad ? n / 2 : 0;
  if (!len_ok(M)) {
    set_error("transform length not supported: " + std::to_string(M));
    return WOS_ERR_UNSUPPORTED;
  }
  const bool mr = !pow2(M) && mixed_col_w(prec, M) > 0;
  const bool bs = !pow2(M) && !mr;
  const int Mk = bs ? bs_len(M) : M;  // kernel FFT length
  if (bs) {
    tu.tw = bs_default_w(prec, Mk, M);
    tu.rows = rows_per_block_host(Mk);
    tu.radix = tu.col_radix = 16;
    tu.inv_rows = tu.rows;
  } else if (mr) {  // one row per row workgroup; the requested shape if built, else the widest column tile for both
    const int cw = tu.col_w > 0 ? tu.col_w : tu.tw;
    const bool ok = mixed_shape_valid(prec, M, tu.tw, cw);
    const int w = mixed_col_w(prec, M);
    const int t_w = ok ? tu.tw : w, c_w = ok ? cw : w;
    tu = Tuning();
    tu.tw = t_w;
    tu.col_w = c_w;
    tu.rows = prec == WOS_FP64 ? mixed_rows_wg_f64(M) : mixed_rows_wg_f32(M);
    tu.radix = tu.col_radix = 16;
    tu = inv_from_fwd(tu);
  } else {
    const bool lean = tf_kind == TF_ASM || tf_kind == TF_CONJ_ASM;
    if (!lean) {  // generic TFs use the default tile and schedule
      tu.tw = default_col_w(prec, M, M);
      tu.radix = tu.col_radix = 16;
    }
    tu = resolve_tuning(prec, M, tu);
  }
  const void *tw, *tw64, *twc;
  int rc;
  if (mr) {
    if ((rc = get_mixed_twiddles(M, prec, &tw))) return rc;
    tw64 = twc = tw;
  } else if (tabs) {  // tables the caller generated (row schedule, its FP64 twin, column schedule)
    tw = tabs[0];
    tw64 = tabs[1];
    twc = tabs[2];
  } else {
    if ((rc = get_twiddles(Mk, prec, &tw, 0, tu.radix))) return rc;
    if ((rc = get_twiddles(Mk, WOS_FP64, &tw64, 0, tu.radix))) return rc;
    if ((rc = get_twiddles(Mk, prec, &twc, 0, tu.col_radix == 32 ? 16 : tu.col_radix))) return rc;
  }
  L.prec = prec;
  L.M = M;
  L.mode = (tf_kind == TF_ASM || tf_kind == TF_CONJ_ASM) ? 1 : 2;
  L.tu = tu;
  L.mr = mr;
  L.bs = bs;
  L.Mb = Mk;
  const long long iunit = inter_unit_elems(n, M);
  const int tstride = n * tu.tw + kTilePad;
  const int groups = bs ? 0 : tu.groups;
  const int split = bs ? 0 : tu.split;
  const int lg = ilog2(tu.tw);
  RowArgs& ra = L.fwd;
  ra = row_args_init();
  ra.in = d;
  ra.out = inter;
  ra.in_unit = (long long)n * n;
  ra.out_unit = iunit;
  ra.in_ld = n;
  ra.out_tw = lg;
  ra.out_tstride = tstride;
  ra.rows = n;
  ra.units = units;
  ra.in_col0 = off;
  ra.in_cols = n;
  ra.out_col0 = 0;
  ra.out_cols = M;
  ra.scale_re = 1.0;
  ra.pw_in = pw_in;
  ra.tw = tw;
  ra.tw64 = tw64;
  ra.groups = groups;
  ra.split = split;
  ra.radix = tu.radix;
  ColArgs& ca = L.col;
  memset(&ca, 0, sizeof(ca));
  ca.data = inter;
  ca.unit = iunit;
  ca.tiled = 1;
  ca.tlg = lg;
  ca.tstride = tstride;
  ca.cols = M;
  ca.units = units;
  ca.in_row0 = off;
  ca.rows = n;
  ca.out_row0 = off;
  ca.out_rows = n;
  ca.scale = 1.0;
  // column workgroups narrower than the layout tile share its rows' lines:
  // the XCD-aware order puts the workgroups of one tile on one XCD (one L2)
  ca.xcd_remap = tu.xcd_remap || (!bs && tu.col_w < tu.tw);
  ca.tf = d_tf;
  ca.tw = twc;
  ca.radix = tu.col_radix;
  ca.split = !bs && !mr && L.mode == 1 && tu.col_radix != 8 && M >= 256 ? tu.col_split : 0;
  RowArgs& rb = L.inv;
  rb = row_args_init();
  rb.in = inter;
  rb.out = d;
  rb.in_unit = iunit;
  rb.out_unit = (long long)n * n;
  rb.in_tw = lg;
  rb.in_tstride = tstride;
  rb.out_ld = n;
  rb.rows = n;
  rb.units = units;
  rb.in_col0 = 0;
  rb.in_cols = M;
  rb.out_col0 = off;
  rb.out_cols = n;
  const double inv = 1.0 / ((double)M * (double)M);
  rb.scale_re = post_re * inv;
  rb.scale_im = post_im * inv;
  rb.pw_out = pw_out;
  rb.tw = tw;
  rb.tw64 = tw64;
  rb.groups = bs ? 0 : tu.inv_groups;
  rb.split = bs ? 0 : tu.inv_split;
  rb.radix = tu.radix;
  if (bs) {
    // fft2DCPU's per-pass 1/M of the inverse is applied by the Bluestein posts
    BsTab fwd_t, inv_t;
    if ((rc = get_bs_tables(M, 0, prec, &fwd_t))) return rc;
    if ((rc = get_bs_tables(M, 1, prec, &inv_t))) return rc;
    L.bfwd.io = ra;
    L.bfwd.len = M;
    L.bfwd.tab = fwd_t;
    L.bfwd.post = 1.0;
    L.bcol.io = ca;
    L.bcol.len = M;
    L.bcol.tab0 = fwd_t;
    L.bcol.tab1 = inv_t;
    L.bcol.post0 = 1.0;
    L.bcol.post1 = 1.0 / (double)M;
    L.binv.io = rb;
    L.binv.io.scale_re = post_re;
    L.binv.io.scale_im = post_im;
    L.binv.len = M;
    L.binv.tab = inv_t;
    L.binv.post = 1.0 / (double)M;
  }
  return 0;
}

// Set the power-partial outputs of a built transfer.
static void set_transfer_power(TransferLaunch& L, double* pw_in, double* pw_out) {
  L.fwd.pw_in = pw_in;
  L.inv.pw_out = pw_out;
  L.bfwd.io.pw_in = pw_in;
  L.binv.io.pw_out = pw_out;
}

// kernel k of a transfer step: 0 forward rows, 1 columns, 2 inverse rows
static int launch_transfer_kernel(const TransferLaunch& L, int k, hipStream_t s) {
  if (L.mr) {
    if (k == 0) return launch_rows_mixed(L.prec, L.M, -1, L.fwd, s);
    if (k == 1) return launch_cols_mixed(L.prec, L.M, 2, 1, L.tu.col_w, L.col, s);
    return launch_rows_mixed(L.prec, L.M, 1, L.inv, s);
  }
  if (L.bs) {
    if (k == 0) return launch_rows_bs(L.prec, L.Mb, L.bfwd, s);
    if (k == 1) return launch_cols_bs(L.prec, L.Mb, 2, L.tu.tw, L.bcol, s);
    return launch_rows_bs(L.prec, L.Mb, L.binv, s);
  }
  if (k == 0) return launch_rows(L.prec, L.M, -1, L.fwd, s, L.tu.rows);
  if (k == 1) return launch_cols(L.prec, L.M, L.mode, 0, L.bs ? L.tu.tw : L.tu.col_w, L.col, s);
  return launch_rows(L.prec, L.M, 1, L.inv, s, L.tu.inv_rows);
}

// Synchronous wrapper on the device stream with per-field power sums.
static int transfer_dev(int prec, void* d, int n, int units, bool pad, const std::vector<TFUnit>& tf, int group,
                        double post_re, double post_im, TransferOut* pw) {
  Device& dv = device();
  const int M = pad ? 2 * n : n;
  int rc;
  if (!len_ok(M)) {
    set_error("transform length not supported: " + std::to_string(M));
    return WOS_ERR_UNSUPPORTED;
  }
  if ((rc = dv.scratch.ensure((size_t)units * inter_unit_elems(n, M) * csize(prec)))) return rc;
  if ((rc = dv.params.ensure(sizeof(TFUnit) * units))) return rc;
  if ((rc = dv.upload_params(dv.params.p, tf.data(), sizeof(TFUnit) * units))) return rc;
  // the asm column modes take no spectral shift (their transfer function
  // is mirrored in fy): shifted transfers (asm_shift*) run the generic
  // column mode, which evaluates every unit's own TF kind per element
  bool shifted = false;
  for (const TFUnit& u : tf) shifted = shifted || u.fy_off != 0.0;
  const int mode_kind = shifted ? (int)TF_NONE : tf[0].kind;
  TransferLaunch L;
  if ((rc = build_transfer(L, prec, d, dv.scratch.p, n, units, pad, (const TFUnit*)dv.params.p, mode_kind, post_re,
                           post_im, nullptr, nullptr, Tuning())))
    return rc;
  const int bpu = (n + L.tu.rows - 1) / L.tu.rows;           // forward row workgroups per unit
  const int bpo = (n + L.tu.inv_rows - 1) / L.tu.inv_rows;   // inverse row workgroups per unit
  if (pw) {
    if ((rc = dv.partial_in.ensure(sizeof(double) * bpu * units))) return rc;
    if ((rc = dv.partial_out.ensure(sizeof(double) * bpo * units))) return rc;
    set_transfer_power(L, (double*)dv.partial_in.p, (double*)dv.partial_out.p);
  }
  for (int k = 0; k < 3; ++k)
    if ((rc = launch_transfer_kernel(L, k, dv.stream))) return rc;
  if (pw) {
    std::vector<double> hin((size_t)bpu * units), hout((size_t)bpo * units);
    WOS_HIP_TRY(hipMemcpyAsync(hin.data(), dv.partial_in.p, sizeof(double) * hin.size(), hipMemcpyDeviceToHost,
                               dv.stream));
    WOS_HIP_TRY(hipMemcpyAsync(hout.data(), dv.partial_out.p, sizeof(double) * hout.size(), hipMemcpyDeviceToHost,
                               dv.stream));
    WOS_HIP_TRY(hipStreamSynchronize(dv.stream));
    const int groups = units / group;
    pw->p_in.assign(groups, 0);
    pw->p_out.assign(groups, 0);
    for (int g = 0; g < groups; ++g) {
      Neumaier si, so;
      for (int b = g * group * bpu; b < (g + 1) * group * bpu; ++b) si.add(hin[b]);
      for (int b = g * group * bpo; b < (g + 1) * group * bpo; ++b) so.add(hout[b]);
      pw->p_in[g] = si.value();
      pw->p_out[g] = so.value();
    }
  }
  return 0;
}

// spectralRadius (optics/propagate.go:270-320) of one field (ncomp
// components at d), computed from a device FFT of a copy.
static int spectral_radius_dev(int prec, const void* d, int n, int ncomp, double dx, double frac, double* out) {
  Device& dv = device();
  if (frac <= 0 || frac >= 1) frac = 0.99;
  const double fnyq = 1 / (2 * dx);
  *out = 0;
  if (n == 0 || dx <= 0) return 0;
  int rc;
  const size_t bytes = (size_t)ncomp * n * n * csize(prec);
  if ((rc = dv.aux.ensure(bytes))) return rc;
  WOS_HIP_TRY(hipMemcpyAsync(dv.aux.p, d, bytes, hipMemcpyDeviceToDevice, dv.stream));
  if ((rc = fft2d_dev(prec, dv.aux.p, n, ncomp, 0))) return rc;
  const size_t total = (size_t)ncomp * n * n;
  const unsigned gs = std::min<unsigned>(grid_for(total), 1024);
  if ((rc = dv.hist.ensure(sizeof(double) * 97 * gs))) return rc;
  const double den = (double)n * dx;
  if (prec == WOS_FP64)
    hipLaunchKernelGGL(k_spec_hist<double>, dim3(gs), dim3(256), 0, dv.stream, (const cpx<double>*)dv.aux.p, n, total, den, fnyq, (double*)dv.hist.p);
  else
    hipLaunchKernelGGL(k_spec_hist<float>, dim3(gs), dim3(256), 0, dv.stream, (const cpx<float>*)dv.aux.p, n, total, den, fnyq, (double*)dv.hist.p);
  WOS_HIP_TRY(hipGetLastError());
  std::vector<double> h((size_t)97 * gs);
  WOS_HIP_TRY(hipMemcpyAsync(h.data(), dv.hist.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, dv.stream));
  WOS_HIP_TRY(hipStreamSynchronize(dv.stream));
  double hist[97];
  double tot = 0;
  for (int b = 0; b < 97; ++b) {
    Neumaier s;
    for (unsigned g = 0; g < gs; ++g) s.add(h[(size_t)g * 97 + b]);
    hist[b] = s.value();
    tot += hist[b];
  }
  if (tot <= 0) return 0;
  double run = 0.0;
  for (int b = 0; b <= 96; ++b) {
    run += hist[b];
    if (run >= frac * tot) {
      *out = (double)b / 96 * fnyq;
      return 0;
    }
  }
  *out = fnyq;
  return 0;
}

// asmWrapWarn (optics/propagate.go:247-264)
static int asm_wrap_warn_dev(int prec, const void* d, int n, int ncomp, double dx, double z, double wl,
                             wos_warnings* w, double carry) {
  if (!w || z == 0 || carry <= 0) return 0;
  if (std::fabs(z) * wl / (2 * dx) <= carry) return 0;
  double fsig = 0;
  int rc = spectral_radius_dev(prec, d, n, ncomp, dx, 0.90, &fsig);
  if (rc) return rc;
  if (fsig <= 0) return 0;
  double walk = std::fabs(z) * wl * fsig;
  if (walk > carry) warn_add(w, WOS_WARN_ASM_ALIAS_WRAP, walk / carry);
  return 0;
}

// spectralCentroid (optics/propagate.go:391-416), summed in the
// reference's serial order.
static int spectral_centroid_dev(int prec, const void* d, int n, int ncomp, double dx, double* fx, double* fy) {
  Device& dv = device();
  int rc;
  if (n > 8192) {
    set_error("spectral centroid: grid too large");
    return WOS_ERR_UNSUPPORTED;
  }
  const size_t bytes = (size_t)ncomp * n * n * csize(prec);
  if ((rc = dv.aux.ensure(bytes))) return rc;
  WOS_HIP_TRY(hipMemcpyAsync(dv.aux.p, d, bytes, hipMemcpyDeviceToDevice, dv.stream));
  if ((rc = fft2d_dev(prec, dv.aux.p, n, ncomp, 0))) return rc;
  if ((rc = dv.hist.ensure(sizeof(double) * 3 * kCentroidBlocks))) return rc;
  const double den = (double)n * dx;
  double h[3] = {0, 0, 0};
  if (prec == WOS_FP64) {
    hipLaunchKernelGGL(k_spec_centroid_serial<double>, dim3(1), dim3(256), 0, dv.stream, (const cpx<double>*)dv.aux.p, n,
                       ncomp, den, (double*)dv.hist.p);
    WOS_HIP_TRY(hipGetLastError());
    WOS_HIP_TRY(hipMemcpyAsync(h, dv.hist.p, sizeof(h), hipMemcpyDeviceToHost, dv.stream));
    WOS_HIP_TRY(hipStreamSynchronize(dv.stream));
  } else {
    hipLaunchKernelGGL(k_spec_centroid_partial<float>, dim3(kCentroidBlocks), dim3(256), 0, dv.stream,
                       (const cpx<float>*)dv.aux.p, n, ncomp, den, (double*)dv.hist.p);
    WOS_HIP_TRY(hipGetLastError());
    std::vector<double> part(3 * kCentroidBlocks);
    WOS_HIP_TRY(hipMemcpyAsync(part.data(), dv.hist.p, sizeof(double) * part.size(), hipMemcpyDeviceToHost, dv.stream));
    WOS_HIP_TRY(hipStreamSynchronize(dv.stream));
    for (int q = 0; q < 3; ++q) {
      Neumaier acc;
      for (int b = 0; b < kCentroidBlocks; ++b) acc.add(part[(size_t)b * 3 + q]);
      h[q] = acc.value();
    }
  }
  if (h[0] <= 0) {
    *fx = 0;
    *fy = 0;
    return 0;
  }
  *fx = h[1] / h[0];
  *fy = h[2] / h[0];
  return 0;
}

static int ramp_dev(int prec, void* d, int n, int units, double dx, double fxc, double fyc, double sgn) {
  Device& dv = device();
  const size_t total = (size_t)units * n * n;
  if (prec == WOS_FP64)
    hipLaunchKernelGGL(k_ramp<double>, dim3(grid_for(total)), dim3(256), 0, dv.stream, (cpx<double>*)d, n, units, dx, fxc, fyc, sgn);
  else
    hipLaunchKernelGGL(k_ramp<float>, dim3(grid_for(total)), dim3(256), 0, dv.stream, (cpx<float>*)d, n, units, dx, fxc, fyc, sgn);
  WOS_HIP_TRY(hipGetLastError());
  return 0;
}

// propFraunhofer's spectrum post-processing (optics/propagate.go:603-650) for
// `comps` spectra of n x n (n even): output (i, j) takes bin
// ((i + n/2) % n, (j + n/2) % n) (the fftshift), times dx^2, negated for odd
// bin parity (the centred-layout checkerboard), times
// c0 * cexpI(k (x^2 + y^2) / (2 z)) on the output grid of spacing dxo, with
// c0 = cexpI(k z) * complex(0, -1/(wl z)) formed on the host.
template <typename R>
__global__ void k_fraunhofer(const cpx<R>* __restrict__ spec, cpx<R>* __restrict__ out, int n, int comps, double dx2,
                             double dxo, double k, double z, double c0r, double c0i) {
#pragma clang fp contract(off)
  const size_t nn = (size_t)n * n;
  const int half = n / 2;
  for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < nn * comps; g += (size_t)gridDim.x * blockDim.x) {
    const size_t c = g / nn, idx = g % nn;
    const int j = (int)(idx / n), i = (int)(idx % n);
    const int sj = (j + half) % n, si = (i + half) % n;
    const cpx<R> a = spec[c * nn + (size_t)sj * n + si];
    double ar = (double)a.x * dx2 - (double)a.y * 0.0, ai = (double)a.x * 0.0 + (double)a.y * dx2;
    if ((si + sj) & 1) {
      ar = -ar;
      ai = -ai;
    }
    const double y = ((double)j - (double)n / 2) * dxo;
    const double x = ((double)i - (double)n / 2) * dxo;
    double sq, cq;
    sincos(k * (x * x + y * y) / (2 * z), &sq, &cq);
    const double tr = c0r * cq - c0i * sq, ti = c0r * sq + c0i * cq;
    out[g] = {(R)(ar * tr - ai * ti), (R)(ar * ti + ai * tr)};
  }
}

// PropagateVectorial's spectral step (optics/vectorial.go:34-75) on the
// spectra of Ex, Ey: Ez = -(kx Ex + ky Ey)/kz (kz = i kzi on evanescent bins,
// Go's complex division of a real or an imaginary divisor), then all three
// times H = cexpI(kz z) or the evanescent policy's real factor.
template <typename R>
__global__ void k_vectorial(cpx<R>* __restrict__ ex, cpx<R>* __restrict__ ey, cpx<R>* __restrict__ ez, int n,
                            double dx, double wl, double k, double z, int zero, int reg, double alpha, double limit) {
#pragma clang fp contract(off)
  const size_t nn = (size_t)n * n;
  const double den = (double)n * dx;
  const double az = fabs(z);
  const bool backward = z < 0;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < nn; idx += (size_t)gridDim.x * blockDim.x) {
    const int j = (int)(idx / n), i = (int)(idx % n);
    const double fy = (double)(j <= n / 2 ? j : j - n) / den;
    const double fx = (double)(i <= n / 2 ? i : i - n) / den;
    const double ky = 2 * M_PI * fy, kx = 2 * M_PI * fx;
    const double rho = wl * go_hypot(fx, fy);
    const double exr = (double)ex[idx].x, exi = (double)ex[idx].y, eyr = (double)ey[idx].x, eyi = (double)ey[idx].y;
    // -(complex(kx,0)*Ex + complex(ky,0)*Ey)
    const double nr = -((kx * exr - 0.0 * exi) + (ky * eyr - 0.0 * eyi));
    const double ni = -((kx * exi + 0.0 * exr) + (ky * eyi + 0.0 * eyr));
    double hr, hi, zr = 0, zi = 0;
    if (rho <= 1) {
      const double kz = k * sqrt((1 - rho) * (1 + rho));
      sincos(kz * z, &hi, &hr);
      if (kz > 0) {  // n / complex(kz, 0)
        const double ratio = 0.0 / kz, denom = kz + ratio * 0.0;
        zr = (nr + ni * ratio) / denom;
        zi = (ni - nr * ratio) / denom;
      }
    } else {
      const double decay = k * az * sqrt((rho - 1) * (rho + 1));
      if (zero || (limit > 0 && decay > limit)) hr = 0;
      else if (backward && reg) {
        const double a = exp(decay);
        hr = a / (1 + alpha * alpha * a * a);
      } else if (backward) hr = exp(decay);
      else hr = exp(-decay);
      hi = 0;
      const double kzi = k * sqrt((rho - 1) * (rho + 1));
      if (kzi > 0) {  // n / complex(0, kzi)
        const double ratio = 0.0 / kzi, denom = kzi + ratio * 0.0;
        zr = (nr * ratio + ni) / denom;
        zi = (ni * ratio - nr) / denom;
      }
    }
    ex[idx] = {(R)(exr * hr - exi * hi), (R)(exr * hi + exi * hr)};
    ey[idx] = {(R)(eyr * hr - eyi * hi), (R)(eyr * hi + eyi * hr)};
    ez[idx] = {(R)(zr * hr - zi * hi), (R)(zr * hi + zi * hr)};
  }
}

// PropagateUniaxial's transfer functions (optics/birefringent.go:38-57) on
// `comps` spectra: component 0 (Ex) extraordinary, component 1 (Ey) ordinary.
template <typename R>
__global__ void k_uniaxial(cpx<R>* __restrict__ a, int n, int comps, double dx, double z, double k0, double no,
                           double ne) {
#pragma clang fp contract(off)
  const size_t nn = (size_t)n * n;
  const double den = (double)n * dx;
  const double az = fabs(z);
  for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < nn * comps; g += (size_t)gridDim.x * blockDim.x) {
    const size_t idx = g % nn;
    const int j = (int)(idx / n), i = (int)(idx % n);
    const double fy = (double)(j <= n / 2 ? j : j - n) / den;
    const double fx = (double)(i <= n / 2 ? i : i - n) / den;
    double arg;
    if (g < nn) {
      const double kx = 2 * M_PI * fx, ky = 2 * M_PI * fy;
      arg = k0 * k0 * ne * ne - (ne * ne / (no * no)) * kx * kx - ky * ky;
    } else {
      arg = k0 * k0 * no * no - (2 * M_PI * fx) * (2 * M_PI * fx) - (2 * M_PI * fy) * (2 * M_PI * fy);
    }
    double hr, hi;
    if (arg < 0) {
      hr = exp(-az * sqrt(-arg));
      hi = 0;
    } else {
      sincos(z * sqrt(arg), &hi, &hr);
    }
    const double ar = (double)a[g].x, ai = (double)a[g].y;
    a[g] = {(R)(ar * hr - ai * hi), (R)(ar * hi + ai * hr)};
  }
}

// PropagateUniaxial (optics/birefringent.go:19-61) of a resident field:
// ncomp 1 (scalar: Ex only) or 2.
static int uniaxial_dev(int prec, void* d, int n, int ncomp, double dx, double z, double no, double ne, double wl) {
  int rc;
  if ((rc = fft2d_dev(prec, d, n, ncomp, 0))) return rc;
  Device& dv = device();
  const size_t nn = (size_t)n * n;
  const double k0 = 2 * M_PI / wl;
  const unsigned gs = std::min<unsigned>(grid_for(nn * ncomp), 8192);
  if (prec == WOS_FP64)
    hipLaunchKernelGGL(k_uniaxial<double>, dim3(gs), dim3(256), 0, dv.stream, (cpx<double>*)d, n, ncomp, dx, z, k0, no,
                       ne);
  else
    hipLaunchKernelGGL(k_uniaxial<float>, dim3(gs), dim3(256), 0, dv.stream, (cpx<float>*)d, n, ncomp, dx, z, k0, no,
                       ne);
  WOS_HIP_TRY(hipGetLastError());
  return fft2d_dev(prec, d, n, ncomp, 1);
}

}  // namespace wos
// ComputeStats on device buffers (wos_stats.inc)
static int stats_dev(int prec, const void* ex, const void* ey, int n, double dx, double wl, double ap, double dist,
                     wos_plane_stats* out);
namespace wos {

// Propagate (optics/propagate.go:65-102) of one field (ncomp components,
// device resident) by one FFT method.
static int propagate_dev(int prec, void* d, int n, int ncomp, double dx, double z, int method, const wos_ctx& ctx,
                         wos_warnings* w, wos_prop_stats* st) {
  if (st) {
    memset(st, 0, sizeof(*st));
    st->method_used = method;
  }
  if (z == 0) return 0;
  const double wl = ctx.wavelength;
  int rc;
  int zero, reg;
  double alpha, limit;
  evanescent_policy(ctx, z, &zero, &reg, &alpha, &limit);
  const double carry = (double)n * dx;
  switch (method) {
    case WOS_METHOD_ASM:
    case WOS_METHOD_ASM_PAD: {
      const bool pad = method == WOS_METHOD_ASM_PAD;
      const int M = pad ? 2 * n : n;
      if ((rc = asm_wrap_warn_dev(prec, d, n, ncomp, dx, z, wl, w, carry))) return rc;
      std::vector<TFUnit> tf(ncomp, make_asm_tf(wl, z, ctx, M, dx, 0, 0));
      TransferOut pw;
      if ((rc = transfer_dev(prec, d, n, ncomp, pad, tf, ncomp, 1.0, 0.0, &pw))) return rc;
      const double p_in = pw.p_in[0] * dx * dx, p_out = pw.p_out[0] * dx * dx;
      asm_warn(w, z, p_in, p_out, zero, reg);
      if (st) {
        st->p_in = p_in;
        st->p_out = p_out;
      }
      return 0;
    }
    case WOS_METHOD_ASM_SHIFT:
    case WOS_METHOD_ASM_SHIFT_PAD: {
      const bool pad = method == WOS_METHOD_ASM_SHIFT_PAD;
      double fxc, fyc;
      if ((rc = spectral_centroid_dev(prec, d, n, ncomp, dx, &fxc, &fyc))) return rc;
      if (st) {
        st->fx_c = fxc;
        st->fy_c = fyc;
      }
      const double fbin = 1 / ((double)n * dx);
      if (std::fabs(fxc) < 0.5 * fbin && std::fabs(fyc) < 0.5 * fbin) {
        const int fallback = pad ? WOS_METHOD_ASM_PAD : WOS_METHOD_ASM;
        wos_prop_stats inner;
        rc = propagate_dev(prec, d, n, ncomp, dx, z, fallback, ctx, w, &inner);
        if (st) {
          st->p_in = inner.p_in;
          st->p_out = inner.p_out;
          st->method_used = fallback;
        }
        return rc;
      }
      const int M = pad ? 2 * n : n;
      // propASMShift checks wrap on the unshifted field; propASMShiftPad on the shifted one
      if (!pad && (rc = asm_wrap_warn_dev(prec, d, n, ncomp, dx, z, wl, w, carry))) return rc;
      if ((rc = ramp_dev(prec, d, n, ncomp, dx, fxc, fyc, -1))) return rc;
      if (pad && (rc = asm_wrap_warn_dev(prec, d, n, ncomp, dx, z, wl, w, carry))) return rc;
      std::vector<TFUnit> tf(ncomp, make_asm_tf(wl, z, ctx, M, dx, fxc, fyc));
      TransferOut pw;
      if ((rc = transfer_dev(prec, d, n, ncomp, pad, tf, ncomp, 1.0, 0.0, &pw))) return rc;
      if ((rc = ramp_dev(prec, d, n, ncomp, dx, fxc, fyc, +1))) return rc;
      const double p_in = pw.p_in[0] * dx * dx, p_out = pw.p_out[0] * dx * dx;
      asm_warn(w, z, p_in, p_out, zero, reg);
      if (st) {
        st->p_in = p_in;
        st->p_out = p_out;
      }
      return 0;
    }
    case WOS_METHOD_FRESNEL_TF: {
      const double k = 2.0 * M_PI / wl;
      TFUnit u = make_asm_tf(wl, z, ctx, n, dx, 0, 0);
      u.kind = TF_FRESNEL;
      u.fr_c = -M_PI * wl * z;
      double s, c;
      sincos(k * z, &s, &c);
      u.pre_re = c;
      u.pre_im = s;
      std::vector<TFUnit> tf(ncomp, u);
      if ((rc = transfer_dev(prec, d, n, ncomp, false, tf, ncomp, 1.0, 0.0, nullptr))) return rc;
      if (std::fabs(z) > (double)n * dx * dx / wl) warn_add(w, WOS_WARN_FRESNEL_TF_ALIAS, 0);
      return 0;
    }
    case WOS_METHOD_FRESNEL_IR: {
      Device& dv = device();
      const double k = 2.0 * M_PI / wl;
      const int m = 2 * n;
      if (!len_ok(m)) {
        set_error("fresnel_ir transform length not supported");
        return WOS_ERR_UNSUPPORTED;
      }
      // impulse-response spectrum (optics/propagate.go:543-559) in its own
      // buffer, computed once per (precision, grid, dx, wl, z)
      const size_t hb = (size_t)m * m * csize(prec);
      DevBuf& ir = dv.fir;
      const bool cached = dv.fir_ok && dv.fir_prec == prec && dv.fir_m == m && dv.fir_dx == dx && dv.fir_wl == wl &&
                          dv.fir_z == z && ir.cap >= hb;
      if (!cached) {
        dv.fir_ok = false;
        if ((rc = ir.ensure(hb))) return rc;
        const size_t total = (size_t)m * m;
        if (prec == WOS_FP64)
          hipLaunchKernelGGL(k_fresnel_ir_kernel<double>, dim3(grid_for(total)), dim3(256), 0, dv.stream,
                             (cpx<double>*)ir.p, m, dx, wl, z);
        else
          hipLaunchKernelGGL(k_fresnel_ir_kernel<float>, dim3(grid_for(total)), dim3(256), 0, dv.stream,
                             (cpx<float>*)ir.p, m, dx, wl, z);
        WOS_HIP_TRY(hipGetLastError());
        if ((rc = fft2d_dev(prec, ir.p, m, 1, 0))) return rc;
        dv.fir_ok = true;
        dv.fir_prec = prec;
        dv.fir_m = m;
        dv.fir_dx = dx;
        dv.fir_wl = wl;
        dv.fir_z = z;
      }
      TFUnit u = make_asm_tf(wl, z, ctx, m, dx, 0, 0);
      u.kind = TF_TABLE;
      u.table = ir.p;
      // pre = complex(0, -1/(wl*z)) * cexpI(k*z) * complex(dx*dx, 0)
      double s, c;
      sincos(k * z, &s, &c);
      const double b = -1 / (wl * z);
      double p1r = 0 * c - b * s, p1i = 0 * s + b * c;
      double d2 = dx * dx;
      double pr = p1r * d2 - p1i * 0, pi = p1r * 0 + p1i * d2;
      std::vector<TFUnit> tf(ncomp, u);
      if ((rc = transfer_dev(prec, d, n, ncomp, true, tf, ncomp, pr, pi, nullptr))) return rc;
      if (std::fabs(z) < (double)n * dx * dx / wl) warn_add(w, WOS_WARN_FRESNEL_IR_ALIAS, 0);
      return 0;
    }
    case WOS_METHOD_FRAUNHOFER: {
      // propFraunhofer (optics/propagate.go:603-662); the output spacing
      // wl |z| / (n dx) is the caller's new Field.DX
      if (n & 1) {
        set_error("fraunhofer: odd grid sizes are not supported");
        return WOS_ERR_UNSUPPORTED;
      }
      const size_t nn = (size_t)n * n;
      const double k = 2 * M_PI / wl;
      const double dxo = wl * std::fabs(z) / ((double)n * dx);
      // rmsRadius of the input (optics/propagate.go:105-130) = 2 sqrt(rms_x^2 + rms_y^2)
      wos_plane_stats ps;
      if ((rc = stats_dev(prec, d, ncomp > 1 ? (const char*)d + nn * csize(prec) : nullptr, n, dx, wl, 0, 0, &ps)))
        return rc;
      const double d_in = ps.power > 0 ? 2 * std::sqrt(ps.rms_x * ps.rms_x + ps.rms_y * ps.rms_y) : 0.0;
      if ((rc = fft2d_dev(prec, d, n, ncomp, 0))) return rc;
      Device& dv = device();
      if ((rc = dv.work.ensure(nn * ncomp * csize(prec)))) return rc;
      // c0 = cexpI(k z) * complex(0, -1/(wl z)), Go complex product order
      const double m = -1 / (wl * z);
      const double ckz = std::cos(k * z), skz = std::sin(k * z);
      const double c0r = ckz * 0.0 - skz * m, c0i = ckz * m + skz * 0.0;
      const unsigned gs = std::min<unsigned>(grid_for(nn * ncomp), 8192);
      if (prec == WOS_FP64)
        hipLaunchKernelGGL(k_fraunhofer<double>, dim3(gs), dim3(256), 0, dv.stream, (const cpx<double>*)d,
                           (cpx<double>*)dv.work.p, n, ncomp, dx * dx, dxo, k, z, c0r, c0i);
      else
        hipLaunchKernelGGL(k_fraunhofer<float>, dim3(gs), dim3(256), 0, dv.stream, (const cpx<float>*)d,
                           (cpx<float>*)dv.work.p, n, ncomp, dx * dx, dxo, k, z, c0r, c0i);
      WOS_HIP_TRY(hipGetLastError());
      WOS_HIP_TRY(hipMemcpyAsync(d, dv.work.p, nn * ncomp * csize(prec), hipMemcpyDeviceToDevice, dv.stream));
      if (d_in > 0) {
        const double nf = d_in * d_in / (wl * std::fabs(z));
        if (nf > 0.5) warn_add(w, WOS_WARN_FRAUNHOFER_NEARFIELD, nf);
      }
      return 0;
    }
    default:
      set_error("unknown propagation method");
      return WOS_ERR_INVALID;
  }
}

// evanescent_filtered checks of merged split-step propagations from the
// spectral power sums of the column passes (Parseval: the ratio of the
// spatial powers before and after a propagation equals the ratio of these
// sums): |U|^2 -> |HU|^2 for one step, -> |H^2 U|^2 for two merged steps.
// Spectral power totals of every slice of a run: one workgroup per slice
// sums that slice's per-workgroup partials (pw3 + s*3*nblocks) into
// tot[3*s .. 3*s+2] (Neumaier per thread, fixed-order block sum).
__global__ void k_split_totals(const double* __restrict__ pw3, int nblocks, double* __restrict__ tot) {
  __shared__ double red[3][256];
  const double* p = pw3 + (size_t)blockIdx.x * 3 * nblocks;
  double s[3] = {0, 0, 0}, c[3] = {0, 0, 0};
  for (int b = threadIdx.x; b < nblocks; b += blockDim.x)
    for (int q = 0; q < 3; ++q) {
      const double x = p[(size_t)b * 3 + q];
      const double t = s[q] + x;
      c[q] += fabs(s[q]) >= fabs(x) ? (s[q] - t) + x : (x - t) + s[q];
      s[q] = t;
    }
  for (int q = 0; q < 3; ++q) red[q][threadIdx.x] = s[q] + c[q];
  __syncthreads();
  if (threadIdx.x < 3) {
    double a = 0;
    for (int i = 0; i < (int)blockDim.x; ++i) a += red[threadIdx.x][i];
    tot[(size_t)blockIdx.x * 3 + threadIdx.x] = a;
  }
}

// The evanescent_filtered checks of every half step, in order (one thread):
// slice s has 1 half step if it is the first or the last, else 2.
__global__ void k_split_warn_all(const double* __restrict__ tot, int slices, double z, wos_warnings* dw) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  for (int sl = 0; sl < slices; ++sl) {
    const double* t = tot + (size_t)sl * 3;
    const int halves = (sl == 0 || sl == slices - 1) ? 1 : 2;
    for (int h = 0; h < halves; ++h) {
      const double pin = t[h], pout = t[h + 1];
      if (pin > 0 && pout < pin * (1 - 1e-8) && z > 0) {
        dw->count[WOS_WARN_EVANESCENT_FILTERED]++;
        dw->value[WOS_WARN_EVANESCENT_FILTERED] = (pin - pout) / pin;
      }
    }
  }
}

// H(dz/2) of the split-step column pass, FP64, evaluated once per run by the
// column kernel's own function and stored in its W-column tile order:
// element (row p, column c) at (c / W) * n * W + p * W + c % W.
__global__ void k_h_table(const TFUnit* __restrict__ u, int n, int W, cpx<double>* __restrict__ h) {
  const size_t nn = (size_t)n * n;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < nn; i += (size_t)gridDim.x * blockDim.x) {
    const int tile = (int)(i / ((size_t)n * W));
    const int rem = (int)(i % ((size_t)n * W));
    const int row = rem / W, col = tile * W + rem % W;
    h[i] = asm_tf_fp64_at(*u, col, row);
  }
}

__global__ void k_zero_warn(wos_warnings* dw) {
  if (threadIdx.x < WOS_WARN_SLOTS) {
    dw->count[threadIdx.x] = 0;
    dw->value[threadIdx.x] = 0;
  }
}

static double g_last_device_ms = 0;

// Launch shape of the split-step slices: column tile width and rows per row
// workgroup. One field of n^2 is small for 256 CUs, so narrower shapes give
// more workgroups; wos_hip_split_step_set_shape overrides the choice.
static std::map<long long, Tuning> g_split_shapes;

static Tuning split_step_shape(int prec, int n) {
  auto it = g_split_shapes.find((long long)n * 2 + prec);
  if (it != g_split_shapes.end()) return it->second;
  Tuning t;
  if (prec == WOS_FP64 && n == 1024) {
    // measured best at 1024^2 FP64 (400 slices): radix-8 rows and columns on
    // 2-column tiles, 35.3k slices/s against 33.8k for the radix-16 shape
    t.tw = t.col_w = 2;
    t.radix = t.col_radix = 8;
    return t;
  }
  t.tw = default_col_w(prec, n, n);
  if (n >= 512 && t.tw > 4) t.tw = 4;  // 1024^2: 4-column tiles with 4 (FP32) or 2 (FP64) rows measured best
  if (n >= 512) {
    const int rows = prec == WOS_FP32 ? 4 : 2;
    if (rows_valid(prec, n, rows)) t.rows = rows;
  }
  return t;
}

// Split-step BPM (optics/medium.go:22-45) on the device for a power-of-two
// grid: the two ASM half steps between screens are merged into one column
// pass applying H(dz/2)^2, each screen runs inside the row pass between the
// inverse and the forward row transform, and every half step's
// evanescent_filtered check is evaluated on the device from spectral power
// sums. `screens` holds n_planes complex128 tables t(x, y). Backward (z < 0)
// warnings are added by the caller.
static int split_step_fast(int prec, void* d, int n, int ncomp, double dx, double z, int steps, const void* screens,
                           int n_planes, const wos_ctx& ctx, wos_warnings* w) {
  Device& dv = device();
  int rc;
  const size_t nn = (size_t)n * n;
  const double dz = z / (double)steps;
  std::vector<TFUnit> tf(ncomp, make_asm_tf(ctx.wavelength, dz / 2, ctx, n, dx, 0, 0));
  if ((rc = dv.params.ensure(sizeof(TFUnit) * ncomp))) return rc;
  if ((rc = dv.upload_params(dv.params.p, tf.data(), sizeof(TFUnit) * ncomp))) return rc;
  if ((rc = dv.scratch.ensure(ncomp * inter_unit_elems(n, n) * csize(prec)))) return rc;
  TransferLaunch L;
  Tuning tu = split_step_shape(prec, n);
  if ((rc = build_transfer(L, prec, d, dv.scratch.p, n, ncomp, false, (const TFUnit*)dv.params.p, TF_ASM, 1.0, 0.0,
                           nullptr, nullptr, tu)))
    return rc;
  const int ncolblk = (n / L.tu.tw) * ncomp;
  // power partials of every slice (read by one warning pass after the loop)
  const size_t pw_bytes = sizeof(double) * 3 * (size_t)ncolblk * (steps + 1);
  const size_t tot_bytes = sizeof(double) * 3 * (size_t)(steps + 1);
  if ((rc = dv.hist.ensure(pw_bytes + tot_bytes + 256))) return rc;
  double* pw3 = (double*)dv.hist.p;
  double* tot = (double*)((char*)dv.hist.p + pw_bytes);
  wos_warnings* dw = (wos_warnings*)((char*)dv.hist.p + pw_bytes + tot_bytes);
  dw = (wos_warnings*)(((uintptr_t)dw + 15) & ~(uintptr_t)15);
  if ((rc = dv.hist.ensure((char*)(dw + 1) - (char*)dv.hist.p))) return rc;
  hipLaunchKernelGGL(k_zero_warn, dim3(1), dim3(64), 0, dv.stream, dw);
  // H(dz/2) once per run, in the column kernel's tile order
  if ((rc = dv.ir.ensure(nn * 16))) return rc;
  hipLaunchKernelGGL(k_h_table, dim3(std::min<unsigned>(grid_for(nn), 4096)), dim3(256), 0, dv.stream,
                     (const TFUnit*)dv.params.p, n, L.tu.tw, (cpx<double>*)dv.ir.p);
  WOS_HIP_TRY(hipGetLastError());
  // the screen row pass: tiled intermediate in place, inverse -> screen -> forward
  RowArgs scr = L.inv;
  scr.out = const_cast<void*>(L.inv.in);
  scr.out_unit = L.inv.in_unit;
  scr.out_tw = L.inv.in_tw;
  scr.out_tstride = L.inv.in_tstride;
  scr.out_ld = 0;
  scr.out_col0 = 0;
  scr.out_cols = n;
  scr.pw_out = nullptr;
  ColArgs col = L.col;
  col.pw3 = w ? pw3 : nullptr;
  col.htab = dv.ir.p;
  if (prec == WOS_FP32) {  // the slice kernels use compensated twiddles
    const void *twr, *twc;
    if ((rc = get_twiddles(n, prec, &twr, 1, L.tu.radix))) return rc;
    if ((rc = get_twiddles(n, prec, &twc, 1, L.tu.col_radix))) return rc;
    scr.tw = twr;
    col.tw = twc;
  }
  const bool want = w != nullptr;
  hipEvent_t e0, e1;
  WOS_HIP_TRY(hipEventCreate(&e0));
  WOS_HIP_TRY(hipEventCreate(&e1));
  WOS_HIP_TRY(hipEventRecord(e0, dv.stream));
  if ((rc = launch_rows(prec, n, -1, L.fwd, dv.stream, L.tu.rows))) return rc;
  for (int s = 0; s <= steps; ++s) {
    col.tf_square = (s == 0 || s == steps) ? 0 : 1;
    col.pw3 = w ? pw3 + (size_t)s * 3 * ncolblk : nullptr;
    if ((rc = launch_cols(prec, n, 3, 0, L.tu.tw, col, dv.stream))) return rc;
    if (s == steps) break;
    scr.screen = (const char*)screens + (n_planes == 1 ? 0 : (size_t)s * nn * 16);
    if ((rc = launch_rows(prec, n, 1, scr, dv.stream, L.tu.rows))) return rc;
  }
  if ((rc = launch_rows(prec, n, 1, L.inv, dv.stream, L.tu.inv_rows))) return rc;
  if (want) {
    hipLaunchKernelGGL(k_split_totals, dim3(steps + 1), dim3(256), 0, dv.stream, (const double*)pw3, ncolblk, tot);
    hipLaunchKernelGGL(k_split_warn_all, dim3(1), dim3(64), 0, dv.stream, (const double*)tot, steps + 1, dz / 2, dw);
    WOS_HIP_TRY(hipGetLastError());
  }
  WOS_HIP_TRY(hipEventRecord(e1, dv.stream));
  if (want) {
    wos_warnings hw;
    WOS_HIP_TRY(hipMemcpyAsync(&hw, dw, sizeof(hw), hipMemcpyDeviceToHost, dv.stream));
    WOS_HIP_TRY(hipStreamSynchronize(dv.stream));
    for (int c = 0; c < WOS_WARN_SLOTS; ++c) {
      w->count[c] += hw.count[c];
      if (hw.value[c] != 0) w->value[c] = hw.value[c];
    }
  }
  WOS_HIP_TRY(hipEventSynchronize(e1));
  float ms = 0;
  WOS_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
  g_last_device_ms = ms;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return 0;
}

// Upload host complex128 arrays into the field buffer in `prec`.
static int upload_fields(int prec, const double* const* src, int ncomp, size_t nn, void** dptr) {
  Device& dv = device();
  int rc;
  if ((rc = dv.field.ensure((size_t)ncomp * nn * csize(prec)))) return rc;
  if (prec == WOS_FP64) {
    for (int c = 0; c < ncomp; ++c)
      if ((rc = host_to_dev((char*)dv.field.p + (size_t)c * nn * 16, src[c], nn * 16, dv.stream))) return rc;
  } else {
    if ((rc = dv.stage.ensure((size_t)ncomp * nn * 16))) return rc;
    for (int c = 0; c < ncomp; ++c)
      if ((rc = host_to_dev((char*)dv.stage.p + (size_t)c * nn * 16, src[c], nn * 16, dv.stream))) return rc;
    hipLaunchKernelGGL(k_z2c, dim3(grid_for(ncomp * nn)), dim3(256), 0, dv.stream, (const cpx<double>*)dv.stage.p,
                       (cpx<float>*)dv.field.p, ncomp * nn);
    WOS_HIP_TRY(hipGetLastError());
  }
  *dptr = dv.field.p;
  return 0;
}

static int download_fields(int prec, const void* d, double* const* dst, int ncomp, size_t nn) {
  Device& dv = device();
  int rc;
  if (prec == WOS_FP64) {
    for (int c = 0; c < ncomp; ++c)
      if ((rc = dev_to_host(dst[c], (const char*)d + (size_t)c * nn * 16, nn * 16, dv.stream))) return rc;
  } else {
    if ((rc = dv.stage.ensure((size_t)ncomp * nn * 16))) return rc;
    hipLaunchKernelGGL(k_c2z, dim3(grid_for(ncomp * nn)), dim3(256), 0, dv.stream, (const cpx<float>*)d,
                       (cpx<double>*)dv.stage.p, ncomp * nn);
    WOS_HIP_TRY(hipGetLastError());
    for (int c = 0; c < ncomp; ++c)
      if ((rc = dev_to_host(dst[c], (const char*)dv.stage.p + (size_t)c * nn * 16, nn * 16, dv.stream))) return rc;
  }
  WOS_HIP_TRY(hipStreamSynchronize(dv.stream));
  return 0;
}

__global__ void k_fill_c128(cpx<double>* __restrict__ a, size_t count, double re, double im) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
    a[i] = {re, im};
}

// Screen tables t(x, y) of PropagateSplitStep's slices (optics/medium.go:
// 47-65) in dv.screens: n_planes index planes from host complex128 arrays,
// or one uniform index nre + i nim when index is NULL; FP64 tables, then a
// rounded FP32 copy for FP32. *tables gets the precision's copy.
static int split_step_screens(int prec, int n, int n_planes, double k, double dz, const double* index, double nre,
                              double nim, const void** tables) {
  Device& dv = device();
  const size_t nn = (size_t)n * n;
  const size_t tb = (size_t)n_planes * nn * 16;
  int rc;
  if ((rc = dv.screens.ensure(tb + (prec == WOS_FP32 ? (size_t)n_planes * nn * 8 : 0)))) return rc;
  if ((rc = dv.work.ensure(nn * 16))) return rc;
  for (int p = 0; p < n_planes; ++p) {
    if (index) {
      if ((rc = host_to_dev(dv.work.p, index + (size_t)p * nn * 2, nn * 16, dv.stream))) return rc;
    } else {
      hipLaunchKernelGGL(k_fill_c128, dim3(grid_for(nn)), dim3(256), 0, dv.stream, (cpx<double>*)dv.work.p, nn, nre,
                         nim);
    }
    hipLaunchKernelGGL(k_screen_table, dim3(grid_for(nn)), dim3(256), 0, dv.stream, (const cpx<double>*)dv.work.p,
                       (cpx<double>*)dv.screens.p + (size_t)p * nn, nn, k, dz);
    WOS_HIP_TRY(hipGetLastError());
    WOS_HIP_TRY(hipStreamSynchronize(dv.stream));
  }
  *tables = dv.screens.p;
  if (prec == WOS_FP32) {
    *tables = (char*)dv.screens.p + tb;
    const size_t tot = (size_t)n_planes * nn;
    hipLaunchKernelGGL(k_z2c, dim3(grid_for(tot)), dim3(256), 0, dv.stream, (const cpx<double>*)dv.screens.p,
                       (cpx<float>*)*tables, tot);
    WOS_HIP_TRY(hipGetLastError());
  }
  return 0;
}

// PropagateSplitStep (optics/medium.go:22-45) on ncomp device components at
// d with the screens of split_step_screens: the merged device pipeline on
// power-of-two grids, else the half steps one by one.
static int split_step_run(int prec, void* d, int n, int ncomp, double dx, double z, int steps, int n_planes,
                          const void* tables, const wos_ctx& ctx, wos_warnings* w) {
  Device& dv = device();
  const size_t nn = (size_t)n * n;
  const double dz = z / (double)steps;
  int rc;
  // merged device pipeline unless a half step needs the wrap diagnostic
  // (its quick check, optics/propagate.go:251, passes for BPM slices)
  const bool quick_ok = std::fabs(dz / 2) * ctx.wavelength / (2 * dx) <= (double)n * dx;
  if (pow2(n) && quick_ok && !getenv("WOS_SPLIT_STEP_REFERENCE_ORDER")) {
    if ((rc = split_step_fast(prec, d, n, ncomp, dx, z, steps, dv.screens.p, n_planes, ctx, w))) return rc;
    if (w && dz / 2 < 0) {  // asmWarn's backward diagnostics, one per half step
      int zero, reg;
      double alpha, limit;
      evanescent_policy(ctx, dz / 2, &zero, &reg, &alpha, &limit);
      for (int i = 0; i < 2 * steps; ++i) asm_warn(w, dz / 2, 0, 0, zero, reg);
    }
    return 0;
  }
  for (int s = 0; s < steps; ++s) {
    if ((rc = propagate_dev(prec, d, n, ncomp, dx, dz / 2, WOS_METHOD_ASM, ctx, w, nullptr))) return rc;
    const void* t = (const char*)tables + (n_planes == 1 ? 0 : (size_t)s * nn * csize(prec));
    if (prec == WOS_FP64)
      hipLaunchKernelGGL(k_mul_table<double>, dim3(grid_for(nn * ncomp)), dim3(256), 0, dv.stream, (cpx<double>*)d,
                         (const cpx<double>*)t, nn, ncomp);
    else
      hipLaunchKernelGGL(k_mul_table32, dim3(grid_for(nn * ncomp)), dim3(256), 0, dv.stream, (cpx<float>*)d,
                         (const cpx<float>*)t, nn, ncomp);
    WOS_HIP_TRY(hipGetLastError());
    if ((rc = propagate_dev(prec, d, n, ncomp, dx, dz / 2, WOS_METHOD_ASM, ctx, w, nullptr))) return rc;
  }
  return 0;
}

static bool valid_prec(int p) { return p == WOS_FP64 || p == WOS_FP32; }

}  // namespace wos

using namespace wos;

// Failure injection for the callers' fallback paths: WOS_HIP_FAIL holds a
// comma-separated list of entry-point names, or "all"; a listed entry point
// returns WOS_ERR_HIP at once, before it reads its arguments or touches the
// device, so the caller's buffers are exactly as it passed them (the Go
// caller then runs its CPU path, optics/fft.go:136-141, 210-215).
static bool fail_injected(const char* fn) {
  const char* e = getenv("WOS_HIP_FAIL");
  if (!e || !*e) return false;
  const std::string spec(e), name(fn);
  if (spec == "all") return true;
  size_t pos = 0;
  while (true) {
    const size_t q = spec.find(',', pos);
    if (spec.compare(pos, q == std::string::npos ? std::string::npos : q - pos, name) == 0) return true;
    if (q == std::string::npos) return false;
    pos = q + 1;
  }
}

#define WOS_ENTER()                                                                   \
  if (fail_injected(__func__)) {                                                      \
    set_error(std::string(__func__) + ": injected failure (WOS_HIP_FAIL)");           \
    return WOS_ERR_HIP;                                                               \
  }                                                                                   \
  int rc_ = ensure_init();                                                            \
  if (rc_) return rc_;                                                                \
  std::lock_guard<std::recursive_mutex> lk_(device().mu);                             \
  (void)hipSetDevice(device().id)

namespace wos {
// humanBytes (optics/memory.go:57-70)
static std::string human_bytes(uint64_t b) {
  char buf[64];
  if (b >= (1ull << 40)) snprintf(buf, sizeof(buf), "%.1f TiB", (double)b / (double)(1ull << 40));
  else if (b >= (1ull << 30)) snprintf(buf, sizeof(buf), "%.1f GiB", (double)b / (double)(1ull << 30));
  else if (b >= (1ull << 20)) snprintf(buf, sizeof(buf), "%.1f MiB", (double)b / (double)(1ull << 20));
  else if (b >= (1ull << 10)) snprintf(buf, sizeof(buf), "%.1f KiB", (double)b / (double)(1ull << 10));
  else snprintf(buf, sizeof(buf), "%llu B", (unsigned long long)b);
  return buf;
}

// Device bytes of `units` resident propagations of n x n (fields, the
// column-tiled intermediate of the M-point transforms, the complex128
// staging an FP32 upload passes through).
static uint64_t grid_bytes(int n, int method, int prec, int units) {
  const int M = method == WOS_METHOD_ASM_PAD || method == WOS_METHOD_ASM_SHIFT_PAD ? 2 * n : n;
  const uint64_t S = csize(prec), nn = (uint64_t)n * n;
  return (uint64_t)units * (nn * S + (uint64_t)inter_unit_elems(n, M) * S) + (prec == WOS_FP32 ? nn * 16 : 0);
}

// CheckGridMemory (optics/memory.go:74-81) for device memory: the budget is
// 75% of the device's memory (memoryBudgetBytes' share on machines above
// 8 GiB); need is the caller's estimate for the whole job.
static int check_grid_budget(const char* what, int n, uint64_t need, uint64_t* budget_out = nullptr) {
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess || total_b == 0) return WOS_OK;  // unknown: hipMalloc decides
  const uint64_t budget = (uint64_t)total_b - (uint64_t)total_b / 4;
  if (budget_out) *budget_out = budget;
  if (need > budget) {
    set_error(std::string(what) + ": grid " + std::to_string(n) + "x" + std::to_string(n) + " would need " +
              human_bytes(need) + " of device memory but only " + human_bytes(budget) +
              " is available; reduce the grid size or the unit count");
    return WOS_ERR_NOMEM;
  }
  return WOS_OK;
}
}  // namespace wos

extern "C" {

// CheckGridMemory (optics/memory.go:74-81) on the device: WOS_OK when `units`
// resident propagations of n x n by `method` in `precision` fit the budget,
// else WOS_ERR_NOMEM with the reference's kind of message; need / budget
// (nullable) receive the estimate and the budget in bytes.
int wos_hip_check_grid_memory(int n, int method, int precision, int units, uint64_t* need, uint64_t* budget) {
  WOS_ENTER();
  if (n <= 0 || units <= 0 || !valid_prec(precision)) {
    set_error("wos_hip_check_grid_memory: bad arguments");
    return WOS_ERR_INVALID;
  }
  const uint64_t b = grid_bytes(n, method, precision, units);
  if (need) *need = b;
  return check_grid_budget("check_grid_memory", n, b, budget);
}

int wos_hip_device_count(void) {
  if (fail_injected(__func__)) return 0;   // GPUAvailable() then reports no GPU
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int wos_hip_device_name(char* buf, int len) {
  WOS_ENTER();
  if (!buf || len <= 0) {
    set_error("wos_hip_device_name: empty buffer");
    return WOS_ERR_INVALID;
  }
  snprintf(buf, (size_t)len, "%s", device().name);
  return WOS_OK;
}

int wos_hip_init(void) { return ensure_init(); }

const char* wos_hip_last_error(void) { return g_err.c_str(); }

double wos_hip_last_device_ms(void) { return g_last_device_ms; }

const char* wos_hip_version(void) { return "libwos_hip 0.1 (gfx950, Stockham radix-16/8/4/2, LDS-staged)"; }

int wos_hip_release_cache(void) {
  WOS_ENTER();
  Device& d = device();
  (void)hipStreamSynchronize(d.stream);
  for (auto& kv : d.twiddles) (void)hipFree(kv.second);
  d.twiddles.clear();
  for (auto& kv : d.chirps) (void)hipFree(kv.second);
  d.chirps.clear();
  for (DevBuf* b : {&d.field, &d.stage, &d.work, &d.aux, &d.scratch, &d.ir, &d.screens, &d.params, &d.partial_in,
                    &d.partial_out, &d.hist, &d.out, &d.fir})
    b->release();
  d.fir_ok = false;
  return WOS_OK;
}

static int fft2d_host(int prec, void* a, int n, int inverse) {
  if (!a || n <= 0) {
    set_error("fft2d: bad arguments");
    return WOS_ERR_INVALID;
  }
  if (n == 1) return WOS_OK;  // the 1-point transform is the identity in both directions
  const bool lng = !len_ok(n);
  if (lng && !long_len_ok(n)) {
    set_error("2-D FFT size not supported: " + std::to_string(n));
    return WOS_ERR_UNSUPPORTED;
  }
  Device& dv = device();
  const size_t nn = (size_t)n * n;
  const size_t bytes = nn * csize(prec);
  int rc;
  if (lng && (rc = check_grid_budget("fft2d", n, 2 * bytes + (prec == WOS_FP32 ? nn * 16 : 0) + kLongChunkBytes * 2)))
    return rc;
  if ((rc = dv.field.ensure(bytes))) return rc;
  if ((rc = host_to_dev(dv.field.p, a, bytes, dv.stream))) return rc;
  if ((rc = lng ? long_fft2d_dev(prec, dv.field.p, n, inverse) : fft2d_dev(prec, dv.field.p, n, 1, inverse))) return rc;
  if ((rc = dev_to_host(a, dv.field.p, bytes, dv.stream))) return rc;
  WOS_HIP_TRY(hipStreamSynchronize(dv.stream));
  return WOS_OK;
}

static int fft1d_host(int prec, void* a, int n, int inverse) {
  if (!a || n <= 0) {
    set_error("fft1d: bad arguments");
    return WOS_ERR_INVALID;
  }
  if (n == 1) return WOS_OK;
  Device& dv = device();
  const bool lng = !len_ok(n);
  if (lng && !long_len_ok(n)) {
    set_error("1-D FFT length not supported: " + std::to_string(n));
    return WOS_ERR_UNSUPPORTED;
  }
  const size_t bytes = (size_t)n * csize(prec);
  int rc;
  if ((rc = dv.field.ensure(bytes))) return rc;
  if ((rc = host_to_dev(dv.field.p, a, bytes, dv.stream))) return rc;
  const double sc = inverse ? 1.0 / (double)n : 1.0;
  if ((rc = lng ? long_fft_dev(prec, dv.field.p, n, 1, inverse, sc, dv.stream)
                : rows_any(prec, dv.field.p, n, 1, n, n, 1, inverse, sc, dv.stream)))
    return rc;
  if ((rc = dev_to_host(a, dv.field.p, bytes, dv.stream))) return rc;
  WOS_HIP_TRY(hipStreamSynchronize(dv.stream));
  return WOS_OK;
}

int wos_hip_fft2d_z(double* a, int n, int inverse) {
  WOS_ENTER();
  return fft2d_host(WOS_FP64, a, n, inverse);
}
int wos_hip_fft2d_c(float* a, int n, int inverse) {
  WOS_ENTER();
  return fft2d_host(WOS_FP32, a, n, inverse);
}
int wos_hip_fft1d_z(double* a, int n, int inverse) {
  WOS_ENTER();
  return fft1d_host(WOS_FP64, a, n, inverse);
}
int wos_hip_fft1d_c(float* a, int n, int inverse) {
  WOS_ENTER();
  return fft1d_host(WOS_FP32, a, n, inverse);
}

int wos_hip_propagate(double* ex, double* ey, int n, double dx, double z, int method, const wos_ctx* ctx,
                      int precision, wos_warnings* w, wos_prop_stats* st) {
  WOS_ENTER();
  if (!ex || n <= 0 || !ctx || !valid_prec(precision) || !(dx > 0)) {
    set_error("wos_hip_propagate: bad arguments");
    return WOS_ERR_INVALID;
  }
  if (st) {
    memset(st, 0, sizeof(*st));
    st->method_used = method;
  }
  if (z == 0) return WOS_OK;
  const int ncomp = ey ? 2 : 1;
  const size_t nn = (size_t)n * n;
  const double* src[2] = {ex, ey};
  double* dst[2] = {ex, ey};
  void* d;
  int rc;
  if ((rc = upload_fields(precision, src, ncomp, nn, &d))) return rc;
  if ((rc = propagate_dev(precision, d, n, ncomp, dx, z, method, *ctx, w, st))) return rc;
  return download_fields(precision, d, dst, ncomp, nn);
}


// PropagateUniaxial (optics/birefringent.go:19-61)
int wos_hip_propagate_uniaxial(double* ex, double* ey, int n, double dx, double z, double no, double ne,
                               const wos_ctx* ctx, int precision) {
  WOS_ENTER();
  if (!ex || n <= 0 || !ctx || !valid_prec(precision) || !(dx > 0) || !(ctx->wavelength > 0)) {
    set_error("wos_hip_propagate_uniaxial: bad arguments");
    return WOS_ERR_INVALID;
  }
  if (no <= 0 || ne <= 0) {
    set_error("uniaxial: n_o and n_e must be > 0");
    return WOS_ERR_INVALID;
  }
  if (!len_ok(n)) {
    set_error("wos_hip_propagate_uniaxial: grid size not supported: " + std::to_string(n));
    return WOS_ERR_UNSUPPORTED;
  }
  const int ncomp = ey ? 2 : 1;
  const size_t nn = (size_t)n * n;
  const double* src[2] = {ex, ey};
  double* dst[2] = {ex, ey};
  void* d;
  int rc;
  if ((rc = upload_fields(precision, src, ncomp, nn, &d))) return rc;
  if ((rc = uniaxial_dev(precision, d, n, ncomp, dx, z, no, ne, ctx->wavelength))) return rc;
  return download_fields(precision, d, dst, ncomp, nn);
}

// PropagateVectorial (optics/vectorial.go:14-80)
int wos_hip_propagate_vectorial(double* ex, double* ey, double* ez, int n, double dx, double z, const wos_ctx* ctx,
                                int precision) {
  WOS_ENTER();
  if (!ex || !ey || !ez || n <= 0 || !ctx || !valid_prec(precision) || !(dx > 0) || !(ctx->wavelength > 0)) {
    set_error("wos_hip_propagate_vectorial: bad arguments");
    return WOS_ERR_INVALID;
  }
  if (!len_ok(n)) {
    set_error("wos_hip_propagate_vectorial: grid size not supported: " + std::to_string(n));
    return WOS_ERR_UNSUPPORTED;
  }
  const size_t nn = (size_t)n * n;
  const double* src[2] = {ex, ey};
  double* dst[2] = {ex, ey};
  void* d;
  int rc;
  if ((rc = upload_fields(precision, src, 2, nn, &d))) return rc;
  Device& dv = device();
  if ((rc = dv.work.ensure(nn * csize(precision)))) return rc;
  void* dz = dv.work.p;
  if ((rc = fft2d_dev(precision, d, n, 2, 0))) return rc;
  int zero, reg;
  double alpha, limit;
  evanescent_policy(*ctx, z, &zero, &reg, &alpha, &limit);
  const double wl = ctx->wavelength, k = 2 * M_PI / wl;
  const unsigned gs = std::min<unsigned>(grid_for(nn), 8192);
  if (precision == WOS_FP64)
    hipLaunchKernelGGL(k_vectorial<double>, dim3(gs), dim3(256), 0, dv.stream, (cpx<double>*)d,
                       (cpx<double>*)d + nn, (cpx<double>*)dz, n, dx, wl, k, z, zero, reg, alpha, limit);
  else
    hipLaunchKernelGGL(k_vectorial<float>, dim3(gs), dim3(256), 0, dv.stream, (cpx<float>*)d, (cpx<float>*)d + nn,
                       (cpx<float>*)dz, n, dx, wl, k, z, zero, reg, alpha, limit);
  WOS_HIP_TRY(hipGetLastError());
  if ((rc = fft2d_dev(precision, d, n, 2, 1))) return rc;
  if ((rc = fft2d_dev(precision, dz, n, 1, 1))) return rc;
  if ((rc = download_fields(precision, d, dst, 2, nn))) return rc;
  double* dzh[1] = {ez};
  return download_fields(precision, dz, dzh, 1, nn);
}

// Propagate1D (optics/propagate.go:718-738)
int wos_hip_propagate1d(double* a, int n, double dx, double z, double wl, const wos_ctx* ctx, int precision,
                        wos_warnings* w) {
  WOS_ENTER();
  if (!a || n <= 0 || !ctx || !valid_prec(precision)) {
    set_error("wos_hip_propagate1d: bad arguments");
    return WOS_ERR_INVALID;
  }
  if (!len_ok(n)) {
    set_error("1-D propagation length not supported: " + std::to_string(n));
    return WOS_ERR_UNSUPPORTED;
  }
  Device& dv = device();
  const double* src[1] = {a};
  double* dst[1] = {a};
  void* d;
  int rc;
  if ((rc = upload_fields(precision, src, 1, (size_t)n, &d))) return rc;
  if ((rc = dv.partial_in.ensure(sizeof(double)))) return rc;
  if ((rc = dv.partial_out.ensure(sizeof(double)))) return rc;
  if ((rc = rows_any(precision, d, n, 1, n, n, 1, 0, 1.0, dv.stream, (double*)dv.partial_in.p, nullptr))) return rc;
  TFUnit u = make_asm_tf(wl, z, *ctx, n, dx, 0, 0);
  if (precision == WOS_FP64)
    hipLaunchKernelGGL(k_tf1d<double>, dim3(grid_for(n)), dim3(256), 0, dv.stream, (cpx<double>*)d, n, dx, u);
  else
    hipLaunchKernelGGL(k_tf1d<float>, dim3(grid_for(n)), dim3(256), 0, dv.stream, (cpx<float>*)d, n, dx, u);
  WOS_HIP_TRY(hipGetLastError());
  if ((rc = rows_any(precision, d, n, 1, n, n, 1, 1, 1.0 / (double)n, dv.stream, nullptr, (double*)dv.partial_out.p)))
    return rc;
  double p_in = 0, p_out = 0;
  WOS_HIP_TRY(hipMemcpyAsync(&p_in, dv.partial_in.p, sizeof(double), hipMemcpyDeviceToHost, dv.stream));
  WOS_HIP_TRY(hipMemcpyAsync(&p_out, dv.partial_out.p, sizeof(double), hipMemcpyDeviceToHost, dv.stream));
  if ((rc = download_fields(precision, d, dst, 1, (size_t)n))) return rc;
  if (p_in > 0 && p_out < p_in * (1 - 1e-8) && z > 0) warn_add(w, WOS_WARN_EVANESCENT_FILTERED, (p_in - p_out) / p_in);
  return WOS_OK;
}

// PropagateSplitStep (optics/medium.go:22-45): Strang slices
// [ASM dz/2, medium phase dz, ASM dz/2], each ASM half step a full Propagate
// with its own diagnostics, as the reference executes them.
int wos_hip_split_step_set_shape(int precision, int n, int tile_w, int rows, int radix, int col_radix) {
  WOS_ENTER();
  if (!valid_prec(precision) || n <= 0 || !pow2(n)) {
    set_error("wos_hip_split_step_set_shape: bad arguments");
    return WOS_ERR_INVALID;
  }
  const long long key = (long long)n * 2 + precision;
  if (tile_w == 0) {
    g_split_shapes.erase(key);
    return WOS_OK;
  }
  if (radix == 0) radix = 16;
  if (col_radix == 0) col_radix = 16;
  const bool r8 = radix == 8 || col_radix == 8;
  if (!((radix == 16 || radix == 8) && (col_radix == 16 || col_radix == 8)) || (r8 && !radix8_ok(n))) {
    set_error("wos_hip_split_step_set_shape: radix-16 or radix-8 (256..8192 points) schedules only");
    return WOS_ERR_INVALID;
  }
  const int dw = default_col_w(precision, n, n);
  const bool tile_built = col_radix == 8 ? tile_w >= 2 : (tile_w == dw || (n >= 512 && tile_w >= 2 && tile_w < dw));
  if (!tile_built || (n % tile_w) != 0 || !col_w_valid(precision, n, tile_w, col_radix) ||
      (rows > 0 && !rows_valid(precision, n, rows, radix))) {
    set_error("wos_hip_split_step_set_shape: column tile or rows not built for the split-step kernels");
    return WOS_ERR_INVALID;
  }
  Tuning t;
  t.tw = tile_w;
  t.col_w = tile_w;
  t.rows = rows;
  t.radix = radix;
  t.col_radix = col_radix;
  g_split_shapes[key] = t;
  return WOS_OK;
}

int wos_hip_split_step(double* ex, double* ey, int n, double dx, double z, int steps, const double* index,
                       int n_planes, const wos_ctx* ctx, int precision, wos_warnings* w) {
  WOS_ENTER();
  if (!ex || n <= 0 || !ctx || !index || !valid_prec(precision) || !(dx > 0)) {
    set_error("wos_hip_split_step: bad arguments");
    return WOS_ERR_INVALID;
  }
  if (z == 0) return WOS_OK;
  if (steps <= 0) steps = 1;
  if (n_planes != 1 && n_planes != steps) {
    set_error("wos_hip_split_step: n_planes must be 1 or steps");
    return WOS_ERR_INVALID;
  }
  const int ncomp = ey ? 2 : 1;
  const size_t nn = (size_t)n * n;
  // a z-invariant medium sampled once per slice: every plane repeats the
  // first bit for bit, so one screen table serves every slice
  if (n_planes > 1) {
    int p = 1;
    while (p < n_planes && !memcmp(index, index + (size_t)p * nn * 2, nn * 16)) ++p;
    if (p == n_planes) n_planes = 1;
  }
  const double* src[2] = {ex, ey};
  double* dst[2] = {ex, ey};
  void* d;
  int rc;
  if ((rc = upload_fields(precision, src, ncomp, nn, &d))) return rc;
  const void* tables;
  if ((rc = split_step_screens(precision, n, n_planes, 2.0 * M_PI / ctx->wavelength, z / (double)steps, index, 0, 0,
                               &tables)))
    return rc;
  if ((rc = split_step_run(precision, d, n, ncomp, dx, z, steps, n_planes, tables, *ctx, w))) return rc;
  return download_fields(precision, d, dst, ncomp, nn);
}

// PropagatePolychromatic (optics/polychromatic.go:18-43)
int wos_hip_polychromatic(const double* ex, const double* ey, int n, double dx, double z, const double* wavelengths,
                          const double* weights, int n_samples, int method, const wos_ctx* ctx, int precision,
                          double* out, wos_warnings* w) {
  WOS_ENTER();
  if (!ex || n <= 0 || !ctx || !out || !wavelengths || !weights || !valid_prec(precision) || !(dx > 0)) {
    set_error("wos_hip_polychromatic: bad arguments");
    return WOS_ERR_INVALID;
  }
  if (n_samples <= 0) {
    set_error("polychromatic: no wavelength samples");
    return WOS_ERR_INVALID;
  }
  Device& dv = device();
  const int ncomp = ey ? 2 : 1;
  const size_t nn = (size_t)n * n;
  const double* src[2] = {ex, ey};
  void* d;
  int rc;
  if ((rc = upload_fields(precision, src, ncomp, nn, &d))) return rc;
  if ((rc = dv.out.ensure(nn * sizeof(double)))) return rc;
  WOS_HIP_TRY(hipMemsetAsync(dv.out.p, 0, nn * sizeof(double), dv.stream));
  const size_t fb = (size_t)ncomp * nn * csize(precision);
  if ((rc = dv.work.ensure(fb))) return rc;
  for (int s = 0; s < n_samples; ++s) {
    if (wavelengths[s] <= 0) {
      set_error("polychromatic: wavelength must be > 0");
      return WOS_ERR_INVALID;
    }
    WOS_HIP_TRY(hipMemcpyAsync(dv.work.p, d, fb, hipMemcpyDeviceToDevice, dv.stream));
    wos_ctx c = *ctx;
    c.wavelength = wavelengths[s];
    if ((rc = propagate_dev(precision, dv.work.p, n, ncomp, dx, z, method, c, w, nullptr))) return rc;
    if (precision == WOS_FP64)
      hipLaunchKernelGGL(k_accum_intensity<double>, dim3(grid_for(nn)), dim3(256), 0, dv.stream,
                         (const cpx<double>*)dv.work.p, ncomp, nn, weights[s], (double*)dv.out.p);
    else
      hipLaunchKernelGGL(k_accum_intensity<float>, dim3(grid_for(nn)), dim3(256), 0, dv.stream,
                         (const cpx<float>*)dv.work.p, ncomp, nn, weights[s], (double*)dv.out.p);
    WOS_HIP_TRY(hipGetLastError());
  }
  WOS_HIP_TRY(hipMemcpyAsync(out, dv.out.p, nn * sizeof(double), hipMemcpyDeviceToHost, dv.stream));
  WOS_HIP_TRY(hipStreamSynchronize(dv.stream));
  return WOS_OK;
}

// Propagate3D (optics/volume.go:8-18)
int wos_hip_volume(const double* ex, const double* ey, int n, double dx, const double* zs, int nz, int method,
                   const wos_ctx* ctx, int precision, double* out_ex, double* out_ey, wos_warnings* w) {
  WOS_ENTER();
  if (!ex || n <= 0 || !ctx || !out_ex || (nz > 0 && !zs) || !valid_prec(precision) || !(dx > 0) ||
      (ey && !out_ey)) {
    set_error("wos_hip_volume: bad arguments");
    return WOS_ERR_INVALID;
  }
  Device& dv = device();
  const int ncomp = ey ? 2 : 1;
  const size_t nn = (size_t)n * n;
  const double* src[2] = {ex, ey};
  void* d;
  int rc;
  if ((rc = upload_fields(precision, src, ncomp, nn, &d))) return rc;
  const size_t fb = (size_t)ncomp * nn * csize(precision);
  if ((rc = dv.work.ensure(fb))) return rc;
  for (int i = 0; i < nz; ++i) {
    WOS_HIP_TRY(hipMemcpyAsync(dv.work.p, d, fb, hipMemcpyDeviceToDevice, dv.stream));
    if ((rc = propagate_dev(precision, dv.work.p, n, ncomp, dx, zs[i], method, *ctx, w, nullptr))) return rc;
    double* dst[2] = {out_ex + (size_t)i * nn * 2, ey ? out_ey + (size_t)i * nn * 2 : nullptr};
    if ((rc = download_fields(precision, dv.work.p, dst, ncomp, nn))) return rc;
  }
  return WOS_OK;
}


// ---------------------------------------------------------------------------
// Device-resident batches.
// ---------------------------------------------------------------------------
}  // extern "C"

// sqrt(max(0, Re FFT(mu))) of GenerateSchellRealizations (optics/coherence.go:29-53)
// into sq (n x n complex128), in FP64 on the device stream.
static int schell_spectrum(int n, double dx, double sigma_c, void* sq) {
  Device& dv = device();
  const size_t nn = (size_t)n * n;
  hipLaunchKernelGGL(k_schell_mu, dim3(grid_for(nn)), dim3(256), 0, dv.stream, (cpx<double>*)sq, n, dx, sigma_c);
  WOS_HIP_TRY(hipGetLastError());
  int rc;
  if ((rc = fft2d_dev(WOS_FP64, sq, n, 1, 0))) return rc;
  hipLaunchKernelGGL(k_schell_sqrt, dim3(grid_for(nn)), dim3(256), 0, dv.stream, (cpx<double>*)sq, nn);
  WOS_HIP_TRY(hipGetLastError());
  return WOS_OK;
}

// Realisation r of `seed` (optics/coherence.go:54-78 with the counter-based
// normal generator): noise times sq, inverse FFT, envelope, generated in FP64
// in `work` (n x n complex128) and stored in precision prec at dst; stream
// ordered.
static int schell_unit(int prec, int n, double dx, double w0, uint64_t seed, uint64_t r, const void* sq, void* work,
                       void* dst) {
  Device& dv = device();
  const size_t nn = (size_t)n * n;
  hipLaunchKernelGGL(k_schell_noise, dim3(grid_for(nn)), dim3(256), 0, dv.stream, (cpx<double>*)work,
                     (const cpx<double>*)sq, nn, seed, r);
  WOS_HIP_TRY(hipGetLastError());
  int rc;
  if ((rc = fft2d_dev(WOS_FP64, work, n, 1, 1))) return rc;
  if (prec == WOS_FP64)
    hipLaunchKernelGGL(k_schell_envelope<double>, dim3(grid_for(nn)), dim3(256), 0, dv.stream,
                       (const cpx<double>*)work, (cpx<double>*)dst, n, dx, w0);
  else
    hipLaunchKernelGGL(k_schell_envelope<float>, dim3(grid_for(nn)), dim3(256), 0, dv.stream,
                       (const cpx<double>*)work, (cpx<float>*)dst, n, dx, w0);
  WOS_HIP_TRY(hipGetLastError());
  return WOS_OK;
}

// Gaussian-Schell realisations u0 .. u0+count-1 of `seed` (optics/coherence.go:22-80
// with the counter-based normal generator) into dst (count fields of n x n,
// precision prec), generated in FP64 on the device stream; synchronous.
static int schell_fill(int prec, int n, double dx, double w0, double sigma_c, uint64_t seed, int u0, int count,
                       void* dst) {
  Device& dv = device();
  const size_t nn = (size_t)n * n;
  int rc;
  if ((rc = dv.aux.ensure(nn * 16))) return rc;    // sqrt(S)
  if ((rc = dv.work.ensure(nn * 16))) return rc;   // one realisation
  if ((rc = schell_spectrum(n, dx, sigma_c, dv.aux.p))) return rc;
  for (int u = 0; u < count; ++u)
    if ((rc = schell_unit(prec, n, dx, w0, seed, (uint64_t)(u0 + u), dv.aux.p,
                          dv.work.p, (char*)dst + (size_t)u * nn * csize(prec))))
      return rc;
  WOS_HIP_TRY(hipStreamSynchronize(dv.stream));
  return WOS_OK;
}

// Device memory owned by one call, freed when the call returns.
struct CallBuf {
  void* p = nullptr;
  ~CallBuf() {
    if (p) (void)hipFree(p);
  }
  int alloc(size_t bytes) {
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
      p = nullptr;
      set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
      return e == hipErrorOutOfMemory ? WOS_ERR_NOMEM : WOS_ERR_HIP;
    }
    return WOS_OK;
  }
};

__global__ void k_scale_f64(double* __restrict__ a, size_t count, double s) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
    a[i] *= s;
}

extern "C" {

// GenerateSchellRealizations (optics/coherence.go:22-81)
int wos_hip_schell_realizations(int n, double dx, double w0, double sigma_c, int64_t seed, int m, double* out) {
  WOS_ENTER();
  if (n <= 0 || !(dx > 0) || !out) {
    set_error("wos_hip_schell_realizations: bad arguments");
    return WOS_ERR_INVALID;
  }
  if (!(w0 > 0) || !(sigma_c > 0)) {
    set_error("schell: width and coherence must be > 0");
    return WOS_ERR_INVALID;
  }
  if (m <= 0) m = 1;
  Device& dv = device();
  const size_t nn = (size_t)n * n;
  CallBuf sq, work, f;
  int rc;
  if ((rc = sq.alloc(nn * 16)) || (rc = work.alloc(nn * 16)) || (rc = f.alloc(nn * 16))) return rc;
  if ((rc = schell_spectrum(n, dx, sigma_c, sq.p))) return rc;
  for (int k = 0; k < m; ++k) {
    if ((rc = schell_unit(WOS_FP64, n, dx, w0, (uint64_t)seed, (uint64_t)k, sq.p, work.p, f.p))) return rc;
    if ((rc = dev_to_host(out + (size_t)k * nn * 2, f.p, nn * 16, dv.stream))) return rc;
  }
  return WOS_OK;
}

// AverageIntensity (optics/coherence.go:85-100)
int wos_hip_average_intensity(const double* ex, const double* ey, int m, int n, double* out) {
  WOS_ENTER();
  if (n <= 0 || m < 0 || (m > 0 && (!ex || !out))) {
    set_error("wos_hip_average_intensity: bad arguments");
    return WOS_ERR_INVALID;
  }
  if (m == 0) return WOS_OK;   // the reference returns nil: nothing is written
  Device& dv = device();
  const size_t nn = (size_t)n * n;
  const int ncomp = ey ? 2 : 1;
  CallBuf f, acc;
  int rc;
  if ((rc = f.alloc((size_t)ncomp * nn * 16)) || (rc = acc.alloc(nn * sizeof(double)))) return rc;
  WOS_HIP_TRY(hipMemsetAsync(acc.p, 0, nn * sizeof(double), dv.stream));
  for (int k = 0; k < m; ++k) {
    if ((rc = host_to_dev(f.p, ex + (size_t)k * nn * 2, nn * 16, dv.stream))) return rc;
    if (ey && (rc = host_to_dev((char*)f.p + nn * 16, ey + (size_t)k * nn * 2, nn * 16, dv.stream))) return rc;
    hipLaunchKernelGGL(k_accum_intensity<double>, dim3(grid_for(nn)), dim3(256), 0, dv.stream,
                       (const cpx<double>*)f.p, ncomp, nn, 1.0, (double*)acc.p);
    WOS_HIP_TRY(hipGetLastError());
  }
  hipLaunchKernelGGL(k_scale_f64, dim3(grid_for(nn)), dim3(256), 0, dv.stream, (double*)acc.p, nn,
                     1 / (double)m);
  WOS_HIP_TRY(hipGetLastError());
  return dev_to_host(out, acc.p, nn * sizeof(double), dv.stream);
}

// PropagatePartiallyCoherent (optics/coherence.go:105-116)
int wos_hip_propagate_partially_coherent(int n, double dx, double w0, double sigma_c, int64_t seed, int m, double z,
                                         int method, const wos_ctx* ctx, int precision, double* out,
                                         wos_warnings* w) {
  WOS_ENTER();
  if (n <= 0 || !(dx > 0) || !ctx || !out || !valid_prec(precision) || method < WOS_METHOD_ASM ||
      method > WOS_METHOD_FRAUNHOFER) {
    set_error("wos_hip_propagate_partially_coherent: bad arguments");
    return WOS_ERR_INVALID;
  }
  if (!(w0 > 0) || !(sigma_c > 0)) {
    set_error("schell: width and coherence must be > 0");
    return WOS_ERR_INVALID;
  }
  if (m <= 0) m = 1;
  Device& dv = device();
  const size_t nn = (size_t)n * n;
  CallBuf sq, work, f, acc;
  int rc;
  if ((rc = sq.alloc(nn * 16)) || (rc = work.alloc(nn * 16)) || (rc = f.alloc(nn * csize(precision))) ||
      (rc = acc.alloc(nn * sizeof(double))))
    return rc;
  if ((rc = schell_spectrum(n, dx, sigma_c, sq.p))) return rc;
  WOS_HIP_TRY(hipMemsetAsync(acc.p, 0, nn * sizeof(double), dv.stream));
  for (int k = 0; k < m; ++k) {
    if ((rc = schell_unit(precision, n, dx, w0, (uint64_t)seed, (uint64_t)k, sq.p, work.p, f.p))) return rc;
    if ((rc = propagate_dev(precision, f.p, n, 1, dx, z, method, *ctx, w, nullptr))) return rc;
    if (precision == WOS_FP64)
      hipLaunchKernelGGL(k_accum_intensity<double>, dim3(grid_for(nn)), dim3(256), 0, dv.stream,
                         (const cpx<double>*)f.p, 1, nn, 1.0, (double*)acc.p);
    else
      hipLaunchKernelGGL(k_accum_intensity<float>, dim3(grid_for(nn)), dim3(256), 0, dv.stream,
                         (const cpx<float>*)f.p, 1, nn, 1.0, (double*)acc.p);
    WOS_HIP_TRY(hipGetLastError());
  }
  hipLaunchKernelGGL(k_scale_f64, dim3(grid_for(nn)), dim3(256), 0, dv.stream, (double*)acc.p, nn,
                     1 / (double)m);
  WOS_HIP_TRY(hipGetLastError());
  return dev_to_host(out, acc.p, nn * sizeof(double), dv.stream);
}

}  // extern "C"

struct wos_batch_s {
  int n = 0, M = 0, method = 0, prec = 0, units = 0;
  bool pad = false;
  double dx = 0;
  hipStream_t stream = nullptr;
  void* fields = nullptr;   // units * n * n, row-major
  void* inter = nullptr;    // units * inter_unit_elems(n, M), column-tiled intermediate
  TFUnit* d_tf = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  Tuning tune;              // zero fields: defaults
  // persistent transfer: one launch per step (wos_hip_batch_set_persistent)
  bool persist = false;
  int persist_grid = 0;
  int persist_interleave = 0;
  unsigned* persist_ctr = nullptr;  // persist_ctr_words(units) words
};

// Sleeps (about 60 ns each, plus a counter read) before a persistent task's
// dependency wait gives up: far beyond any legitimate wait, which lasts
// about one task (tens of microseconds).
constexpr int kPersistSpinLimit = 1 << 22;

// The persistent kernel's shape: row and column workgroups of one size
// (ROWS rows, ROWS-column tiles), split LDS exchanges, radix-16 schedules.
static Tuning persist_tuning(int prec) {
  Tuning t;
  t.tw = t.col_w = t.rows = persist_rows(prec);
  t.groups = 0;
  t.split = 1;
  t.radix = t.col_radix = 16;
  t.col_split = 1;
  return t;
}

static int batch_build(wos_batch_s* b, TransferLaunch& L) {
  return build_transfer(L, b->prec, b->fields, b->inter, b->n, b->units, b->pad, b->d_tf, TF_ASM, 1.0, 0.0,
                        nullptr, nullptr, b->persist ? persist_tuning(b->prec) : b->tune);
}

static PersistArgs batch_persist_args(const wos_batch_s* b, const TransferLaunch& L) {
  PersistArgs a;
  memset(&a, 0, sizeof(a));
  a.fwd = L.fwd;
  a.col = L.col;
  a.col.xcd_remap = 0;  // the queues place the work on XCDs themselves
  a.inv = L.inv;
  a.units = b->units;
  a.rb = (b->n + L.tu.rows - 1) / L.tu.rows;
  a.cb = L.M / L.tu.col_w;
  a.interleave = b->persist_interleave && a.rb == a.cb;
  a.ctr = b->persist_ctr;
  a.spin_limit = kPersistSpinLimit;
  return a;
}

static int batch_enqueue(wos_batch_s* b) {
  TransferLaunch L;
  int rc = batch_build(b, L);
  if (!rc && b->persist) return launch_transfer_persist(b->prec, b->M, b->persist_grid, batch_persist_args(b, L), b->stream);
  for (int k = 0; k < 3 && !rc; ++k) rc = launch_transfer_kernel(L, k, b->stream);
  return rc;
}

// The sticky status of the persistent launches since the last check
// (after the stream is idle): a wait that gave up or a queue left with
// unclaimed tasks means some unit was not propagated.
static int batch_persist_status(wos_batch_s* b) {
  if (!b->persist_ctr) return 0;
  unsigned st = 0;
  WOS_HIP_TRY(hipMemcpy(&st, b->persist_ctr + kPersistStatus, sizeof(st), hipMemcpyDeviceToHost));
  if (st) {
    set_error(std::string("persistent transfer: ") +
              ((st & 1) ? "a dependency wait gave up" : "a work queue was not drained") +
              "; the batch's fields are not valid");
    return WOS_ERR_HIP;
  }
  return 0;
}

static void batch_drop_graph(wos_batch_s* b) {
  if (b->exec) {
    (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
  }
  if (b->graph) {
    (void)hipGraphDestroy(b->graph);
    b->graph = nullptr;
  }
}

extern "C" {

int wos_hip_batch_create(wos_batch* out, int n, double dx, int method, int precision, int n_units,
                         const double* wavelengths, const double* zs, const wos_ctx* ctx) {
  WOS_ENTER();
  if (!out || n <= 0 || n_units <= 0 || !wavelengths || !zs || !ctx || !valid_prec(precision) || !(dx > 0)) {
    set_error("wos_hip_batch_create: bad arguments");
    return WOS_ERR_INVALID;
  }
  if (method != WOS_METHOD_ASM && method != WOS_METHOD_ASM_PAD) {
    set_error("wos_hip_batch_create: batches run asm or asm_pad");
    return WOS_ERR_UNSUPPORTED;
  }
  const bool pad = method == WOS_METHOD_ASM_PAD;
  const int M = pad ? 2 * n : n;
  if (!len_ok(M)) {
    set_error("wos_hip_batch_create: transform length not supported");
    return WOS_ERR_UNSUPPORTED;
  }
  {
    const int brc = check_grid_budget("wos_hip_batch_create", n, grid_bytes(n, method, precision, n_units));
    if (brc) return brc;
  }
  wos_batch_s* b = new wos_batch_s();
  b->n = n;
  b->M = M;
  b->method = method;
  b->prec = precision;
  b->units = n_units;
  b->pad = pad;
  b->dx = dx;
  const size_t nn = (size_t)n * n;
  auto fail = [&](int code) {
    if (b->fields) (void)hipFree(b->fields);
    if (b->inter) (void)hipFree(b->inter);
    if (b->d_tf) (void)hipFree(b->d_tf);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return code;
  };
  hipError_t e;
  if ((e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)) != hipSuccess) {
    set_error(std::string("hipStreamCreate: ") + hipGetErrorString(e));
    return fail(WOS_ERR_HIP);
  }
  if ((e = hipMalloc(&b->fields, nn * n_units * csize(precision))) != hipSuccess) {
    b->fields = nullptr;
    set_error(std::string("hipMalloc(batch fields): ") + hipGetErrorString(e));
    return fail(WOS_ERR_NOMEM);
  }
  (void)hipMemset(b->fields, 0, nn * n_units * csize(precision));
  if ((e = hipMalloc(&b->inter, (size_t)inter_unit_elems(n, M) * n_units * csize(precision))) != hipSuccess) {
    b->inter = nullptr;
    set_error(std::string("hipMalloc(batch intermediate): ") + hipGetErrorString(e));
    return fail(WOS_ERR_NOMEM);
  }
  std::vector<TFUnit> tf(n_units);
  for (int u = 0; u < n_units; ++u) tf[u] = make_asm_tf(wavelengths[u], zs[u], *ctx, M, dx, 0, 0);
  if ((e = hipMalloc((void**)&b->d_tf, sizeof(TFUnit) * n_units)) != hipSuccess) {
    b->d_tf = nullptr;
    set_error(std::string("hipMalloc(batch params): ") + hipGetErrorString(e));
    return fail(WOS_ERR_NOMEM);
  }
  if ((e = hipMemcpy(b->d_tf, tf.data(), sizeof(TFUnit) * n_units, hipMemcpyHostToDevice)) != hipSuccess) {
    set_error(std::string("hipMemcpy(batch params): ") + hipGetErrorString(e));
    return fail(WOS_ERR_HIP);
  }
  const void* tw;
  int rc = get_twiddles(M, precision, &tw);
  if (rc) return fail(rc);
  if ((e = hipStreamSynchronize(device().stream)) != hipSuccess) {
    set_error(std::string("twiddles: ") + hipGetErrorString(e));
    return fail(WOS_ERR_HIP);
  }
  *out = b;
  return WOS_OK;
}

int wos_hip_batch_destroy(wos_batch b) {
  WOS_ENTER();
  if (!b) return WOS_OK;
  (void)hipStreamSynchronize(b->stream);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->persist_ctr) (void)hipFree(b->persist_ctr);
  if (b->inter) (void)hipFree(b->inter);
  if (b->fields) (void)hipFree(b->fields);
  if (b->d_tf) (void)hipFree(b->d_tf);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return WOS_OK;
}

void* wos_hip_batch_device_ptr(wos_batch b) { return b ? b->fields : nullptr; }

int wos_hip_batch_upload(wos_batch b, int unit, const double* ex) {
  WOS_ENTER();
  if (!b || !ex || unit < 0 || unit >= b->units) {
    set_error("wos_hip_batch_upload: bad arguments");
    return WOS_ERR_INVALID;
  }
  Device& dv = device();
  const size_t nn = (size_t)b->n * b->n;
  int rc;
  WOS_HIP_TRY(hipStreamSynchronize(b->stream));
  if (b->prec == WOS_FP64) {
    if ((rc = host_to_dev((char*)b->fields + unit * nn * 16, ex, nn * 16, dv.stream))) return rc;
    WOS_HIP_TRY(hipStreamSynchronize(dv.stream));
    return WOS_OK;
  }
  if ((rc = dv.stage.ensure(nn * 16))) return rc;
  if ((rc = host_to_dev(dv.stage.p, ex, nn * 16, dv.stream))) return rc;
  hipLaunchKernelGGL(k_z2c, dim3(grid_for(nn)), dim3(256), 0, dv.stream, (const cpx<double>*)dv.stage.p,
                     (cpx<float>*)b->fields + unit * nn, nn);
  WOS_HIP_TRY(hipGetLastError());
  WOS_HIP_TRY(hipStreamSynchronize(dv.stream));
  return WOS_OK;
}

int wos_hip_batch_download(wos_batch b, int unit, double* ex) {
  WOS_ENTER();
  if (!b || !ex || unit < 0 || unit >= b->units) {
    set_error("wos_hip_batch_download: bad arguments");
    return WOS_ERR_INVALID;
  }
  Device& dv = device();
  const size_t nn = (size_t)b->n * b->n;
  int rc;
  WOS_HIP_TRY(hipStreamSynchronize(b->stream));
  if (b->prec == WOS_FP64) {
    if ((rc = dev_to_host(ex, (char*)b->fields + unit * nn * 16, nn * 16, dv.stream))) return rc;
    return WOS_OK;
  }
  if ((rc = dv.stage.ensure(nn * 16))) return rc;
  hipLaunchKernelGGL(k_c2z, dim3(grid_for(nn)), dim3(256), 0, dv.stream, (const cpx<float>*)b->fields + unit * nn,
                     (cpx<double>*)dv.stage.p, nn);
  WOS_HIP_TRY(hipGetLastError());
  if ((rc = dev_to_host(ex, dv.stage.p, nn * 16, dv.stream))) return rc;
  WOS_HIP_TRY(hipStreamSynchronize(dv.stream));
  return WOS_OK;
}

int wos_hip_batch_fill_schell(wos_batch b, double w0, double sigma_c, uint64_t seed) {
  WOS_ENTER();
  if (!b || !(w0 > 0) || !(sigma_c > 0)) {
    set_error("wos_hip_batch_fill_schell: width and coherence must be > 0");
    return WOS_ERR_INVALID;
  }
  WOS_HIP_TRY(hipStreamSynchronize(b->stream));
  return schell_fill(b->prec, b->n, b->dx, w0, sigma_c, seed, 0, b->units, b->fields);
}

// One run of the batch timed with HIP events (mean over iters after a warm-up).
static int batch_time_runs(wos_batch_s* b, int iters, hipEvent_t e0, hipEvent_t e1, double* ms_out) {
  int rc;
  if ((rc = batch_enqueue(b))) return rc;
  WOS_HIP_TRY(hipEventRecord(e0, b->stream));
  for (int i = 0; i < iters; ++i)
    if ((rc = batch_enqueue(b))) return rc;
  WOS_HIP_TRY(hipEventRecord(e1, b->stream));
  WOS_HIP_TRY(hipEventSynchronize(e1));
  float ms = 0;
  WOS_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
  *ms_out = ms / iters;
  return 0;
}

}  // extern "C"

// Time kernel k (0 forward rows, 1 columns, 2 inverse rows) of a built
// transfer: one warm-up launch, then `iters` launches back to back between
// two events on stream s.
static int time_transfer_kernel(const TransferLaunch& L, int k, int iters, hipStream_t s, hipEvent_t e0,
                                hipEvent_t e1, double* ms_out) {
  int rc;
  if ((rc = launch_transfer_kernel(L, k, s))) return rc;
  WOS_HIP_TRY(hipEventRecord(e0, s));
  for (int i = 0; i < iters; ++i)
    if ((rc = launch_transfer_kernel(L, k, s))) return rc;
  WOS_HIP_TRY(hipEventRecord(e1, s));
  WOS_HIP_TRY(hipEventSynchronize(e1));
  float ms = 0;
  WOS_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
  *ms_out = ms / iters;
  return 0;
}

// Time the whole step (the three kernels in order) of a built transfer: one
// warm-up step, then `iters` steps back to back between two events.
static int time_transfer_step(const TransferLaunch& L, int iters, hipStream_t s, hipEvent_t e0, hipEvent_t e1,
                              double* ms_out) {
  int rc;
  for (int k = 0; k < 3; ++k)
    if ((rc = launch_transfer_kernel(L, k, s))) return rc;
  WOS_HIP_TRY(hipEventRecord(e0, s));
  for (int i = 0; i < iters; ++i)
    for (int k = 0; k < 3; ++k)
      if ((rc = launch_transfer_kernel(L, k, s))) return rc;
  WOS_HIP_TRY(hipEventRecord(e1, s));
  WOS_HIP_TRY(hipEventSynchronize(e1));
  float ms = 0;
  WOS_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
  *ms_out = ms / iters;
  return 0;
}

// Launch-shape search of one transfer step on a power-of-two grid of side M
// (n rows per unit). The row passes and the column pass share only the
// intermediate layout's tile width tw, so for each tw they are searched
// apart, each kernel timed on its own:
//   row passes: stage schedule (radix-16, or radix-8 from 256 points)
//     shared by both, and for the forward and the inverse pass each its
//     fastest rows per workgroup, row groups of the pipelined kernels and
//     split exchanges (radix-16);
//   column pass: stage schedule (radix-16, radix-8, or radix-16 stages with
//     32 values per thread), columns per workgroup (tw, tw/2 or tw/4), split
//     exchanges (radix-16).
// The tw with the smallest sum of its best row and column times wins. A
// pass timed alone runs on caches and a memory system the other passes do
// not leave it in the step, so the three fastest forward and the three
// fastest inverse row shapes of the winning tw (and its row schedule) are
// then timed in whole steps, nine pairings, and the fastest step is kept.
// build(t, L) builds the three launches of shape t.
template <typename Build>
static int tune_transfer(int prec, int M, int n, int iters, hipStream_t s, const Build& build, Tuning* out) {
  const int S = (int)csize(prec);
  hipEvent_t e0, e1;
  WOS_HIP_TRY(hipEventCreate(&e0));
  WOS_HIP_TRY(hipEventCreate(&e1));
  double best = 1e300;
  Tuning bt;
  // the fastest row shapes (time, shape) of one tw and row schedule
  typedef std::vector<std::pair<double, Tuning>> Ranked;
  auto keep3 = [](Ranked& r, double ms, const Tuning& t) {
    r.emplace_back(ms, t);
    std::stable_sort(r.begin(), r.end(),
                     [](const std::pair<double, Tuning>& a, const std::pair<double, Tuning>& b) { return a.first < b.first; });
    if (r.size() > 3) r.resize(3);
  };
  Ranked best_f, best_i;  // of the winning tw
  auto search = [&]() -> int {
    int rc;
    for (int tw = 1; tw <= 16; tw *= 2) {
      if ((M % tw) != 0) continue;
      double brow = 1e300;
      Tuning tr;
      Ranked rf_tw, ri_tw;
      for (int radix : {16, 8}) {
        if (radix == 8 && !radix8_ok(M)) continue;
        // the forward and the inverse pass each take their fastest shape
        double bf = 1e300, bi = 1e300;
        Tuning tf, ti;
        Ranked rf, ri;
        for (int rows = 1; rows <= 16; rows *= 2) {
          if (!rows_valid(prec, M, rows, radix)) continue;
          if (rows * tw * S < 64) continue;  // chunks under half a line: never competitive
          for (int groups : {0, 2, 4}) {
            if (groups && (M < 256 || n % (rows * groups) != 0)) continue;
            // split exchanges: the one-group kernels, and the pipelined
            // kernels of 2 or 4 rows (twiddles in LDS)
            const int max_split =
                M >= 256 && radix == 16 && (groups == 0 || rows_pipe_lds_valid(prec, M, rows)) ? 1 : 0;
            for (int split = 0; split <= max_split; ++split) {
              Tuning t;
              t.tw = tw;
              t.rows = rows;
              t.groups = groups;
              t.split = split;
              t.radix = radix;
              TransferLaunch L;
              double f, i;
              if ((rc = build(t, L))) return rc;
              if ((rc = time_transfer_kernel(L, 0, iters, s, e0, e1, &f))) return rc;
              if ((rc = time_transfer_kernel(L, 2, iters, s, e0, e1, &i))) return rc;
              keep3(rf, f, t);
              keep3(ri, i, t);
              if (f < bf) {
                bf = f;
                tf = t;
              }
              if (i < bi) {
                bi = i;
                ti = t;
              }
            }
          }
        }
        if (bf + bi < brow) {
          brow = bf + bi;
          tr = tf;
          tr.inv_rows = ti.rows;
          tr.inv_groups = ti.groups;
          tr.inv_split = ti.split;
          rf_tw = rf;
          ri_tw = ri;
        }
      }
      if (brow >= 1e300) continue;
      double bcol = 1e300;
      Tuning tc = tr;
      for (int radix : {16, 8, 32}) {
        if ((radix == 8 && !radix8_ok(M)) || (radix == 32 && !col32_ok(M))) continue;
        for (int cw = tw; cw >= 1 && cw * 4 >= tw; cw /= 2) {
          if (!col_w_valid(prec, M, cw, radix)) continue;
          // split exchanges: radix-16; the 32-value schedule builds them only
          for (int cs = radix == 32 ? 1 : 0; cs <= (radix != 8 && M >= 256 ? 1 : 0); ++cs) {
            Tuning t = tr;
            t.col_w = cw;
            t.col_radix = radix;
            t.col_split = cs;
            TransferLaunch L;
            double c;
            if ((rc = build(t, L))) return rc;
            if ((rc = time_transfer_kernel(L, 1, iters, s, e0, e1, &c))) return rc;
            if (c < bcol) {
              bcol = c;
              tc = t;
            }
          }
        }
      }
      if (bcol < 1e300 && brow + bcol < best) {
        best = brow + bcol;
        bt = tc;
        best_f = rf_tw;
        best_i = ri_tw;
      }
    }
    // whole steps: the winning tw's fastest forward x inverse row shapes
    double bstep = 1e300;
    Tuning bs = bt;
    for (const auto& f : best_f)
      for (const auto& i : best_i) {
        Tuning t = bt;
        t.rows = f.second.rows;
        t.groups = f.second.groups;
        t.split = f.second.split;
        t.inv_rows = i.second.rows;
        t.inv_groups = i.second.groups;
        t.inv_split = i.second.split;
        TransferLaunch L;
        double ms;
        if ((rc = build(t, L))) return rc;
        if ((rc = time_transfer_step(L, iters, s, e0, e1, &ms))) return rc;
        if (ms < bstep) {
          bstep = ms;
          bs = t;
        }
      }
    bt = bs;
    return 0;
  };
  const int rc = search();
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (!rc) *out = bt;  // no shape timed (tiny grids): the defaults
  return rc;
}

extern "C" {

static void tuning_out(const Tuning& t, int* shape) {
  if (!shape) return;
  shape[0] = t.tw;
  shape[1] = t.col_w;
  shape[2] = t.rows;
  shape[3] = t.groups;
  shape[4] = t.split;
  shape[5] = t.radix == 8 ? 8 : 16;
  shape[6] = t.col_radix == 8 ? 8 : t.col_radix == 32 ? 32 : 16;
  shape[7] = t.col_split;
  const Tuning ti = inv_from_fwd(t);
  shape[8] = ti.inv_rows;
  shape[9] = ti.inv_groups;
  shape[10] = ti.inv_split;
}

int wos_hip_batch_autotune(wos_batch b, int iters, int* shape, double* best_ms) {
  WOS_ENTER();
  if (!b || iters <= 0) return WOS_ERR_INVALID;
  int rc;
  WOS_HIP_TRY(hipStreamSynchronize(b->stream));
  batch_drop_graph(b);
  hipEvent_t e0, e1;
  WOS_HIP_TRY(hipEventCreate(&e0));
  WOS_HIP_TRY(hipEventCreate(&e1));
  double best = 1e300;
  Tuning bt = b->tune;
  if (!pow2(b->M) && mixed_col_w(b->prec, b->M) > 0) {  // mixed radix: layout tile and column tile widths
    rc = 0;
    for (int tw : {16, 8, 4, 2, 1}) {
      for (int cw = tw; cw >= 1 && !rc; cw /= 2) {
        if (!mixed_shape_valid(b->prec, b->M, tw, cw)) continue;
        Tuning t;
        t.tw = tw;
        t.col_w = cw;
        b->tune = t;
        double ms;
        if ((rc = batch_time_runs(b, iters, e0, e1, &ms))) break;
        if (ms < best) {
          best = ms;
          bt = t;
        }
      }
      if (rc) break;
    }
    b->tune = bt;
  } else if (!pow2(b->M)) {  // Bluestein grids run their fixed launch shape
    rc = batch_time_runs(b, iters, e0, e1, &best);
    bt = Tuning();
  } else {
    auto build = [b](const Tuning& t, TransferLaunch& L) {
      return build_transfer(L, b->prec, b->fields, b->inter, b->n, b->units, b->pad, b->d_tf, TF_ASM, 1.0, 0.0,
                            nullptr, nullptr, t);
    };
    rc = tune_transfer(b->prec, b->M, b->n, iters, b->stream, build, &bt);
    if (!rc) {
      b->tune = bt;
      rc = batch_time_runs(b, iters, e0, e1, &best);  // the whole step at the chosen shape
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  tuning_out(pow2(b->M) ? resolve_tuning(b->prec, b->M, bt) : bt, shape);
  if (best_ms) *best_ms = best;
  return WOS_OK;
}

}  // extern "C"

// A launch shape {tile_w, col_w, rows, groups, split, radix, col_radix,
// col_split, inv_rows, inv_groups, inv_split} checked against what the
// kernels build for an M-point grid in precision prec (mixed-radix grids
// take the two tile widths only).
static int shape_to_tuning(int prec, int M, const int* sh, Tuning* out, const char* what) {
  int tile_w = sh[0], col_w = sh[1], rows = sh[2], groups = sh[3], split = sh[4], radix = sh[5], col_radix = sh[6],
      col_split = sh[7], inv_rows = sh[8], inv_groups = sh[9], inv_split = sh[10];
  if (radix == 0) radix = 16;
  if (col_radix == 0) col_radix = 16;
  Tuning t;
  if (!pow2(M) && mixed_col_w(prec, M) > 0) {  // mixed radix: the tile widths only
    if (!mixed_shape_valid(prec, M, tile_w, col_w)) {
      set_error(std::string(what) + ": tile width " + std::to_string(tile_w) + " / column width " +
                std::to_string(col_w) + " not built for the mixed-radix grid of " + std::to_string(M));
      return WOS_ERR_INVALID;
    }
    t.tw = tile_w;
    t.col_w = col_w;
    *out = t;
    return WOS_OK;
  }
  if (col_split && col_radix == 8) {
    set_error(std::string(what) + ": split column exchanges are radix-16 column passes only");
    return WOS_ERR_INVALID;
  }
  if (col_radix == 32 && (!col_split || !col32_ok(M))) {
    set_error(std::string(what) + ": the 32-value column schedule is the split radix-16 column pass of 1024, 2048 "
              "and 4096 points");
    return WOS_ERR_INVALID;
  }
  if (!(radix == 16 || (radix == 8 && radix8_ok(M) && !split)) ||
      !(col_radix == 16 || col_radix == 32 || (col_radix == 8 && radix8_ok(M)))) {
    set_error(std::string(what) + ": row radix " + std::to_string(radix) + (split ? " with split exchanges" : "") +
              " or column radix " + std::to_string(col_radix) + " not built for a grid of " + std::to_string(M));
    return WOS_ERR_INVALID;
  }
  if (!pow2(M) || tile_w <= 0 || (M % tile_w) != 0 || col_w <= 0 || (tile_w % col_w) != 0 ||
      !col_w_valid(prec, M, col_w, col_radix) || (rows > 0 && !rows_valid(prec, M, rows, radix)) ||
      !(groups == 0 || groups == 1 || groups == 2 || groups == 4) ||
      (groups >= 2 && split && !rows_pipe_lds_valid(prec, M, rows)) ||
      (inv_rows >= 0 && ((inv_rows > 0 && !rows_valid(prec, M, inv_rows, radix)) ||
                         !(inv_groups == 0 || inv_groups == 1 || inv_groups == 2 || inv_groups == 4) ||
                         (inv_split && radix != 16) ||
                         (inv_groups >= 2 && inv_split && !rows_pipe_lds_valid(prec, M, inv_rows))))) {
    set_error(std::string(what) + ": tile width " + std::to_string(tile_w) + ", column width " +
              std::to_string(col_w) + ", rows " + std::to_string(rows) + " or row groups " + std::to_string(groups) +
              " not built for a grid of " + std::to_string(M) + " under the radix-" + std::to_string(radix) +
              " row / radix-" + std::to_string(col_radix) + " column schedules");
    return WOS_ERR_INVALID;
  }
  t.tw = tile_w;
  t.col_w = col_w;
  t.rows = rows;
  t.groups = groups;
  t.split = split ? 1 : 0;
  t.radix = radix;
  t.col_radix = col_radix;
  t.col_split = col_split ? 1 : 0;
  if (inv_rows >= 0) {
    t.inv_rows = inv_rows;
    t.inv_groups = inv_groups;
    t.inv_split = inv_split ? 1 : 0;
  }
  *out = t;
  return WOS_OK;
}

extern "C" {

int wos_hip_batch_set_tuning(wos_batch b, int tile_w, int col_w, int rows, int groups, int split, int radix,
                             int col_radix, int col_split, int inv_rows, int inv_groups, int inv_split) {
  WOS_ENTER();
  if (!b) return WOS_ERR_INVALID;
  const int sh[WOS_SHAPE_LEN] = {tile_w, col_w, rows, groups, split, radix, col_radix, col_split,
                                 inv_rows, inv_groups, inv_split};
  Tuning t;
  int rc = shape_to_tuning(b->prec, b->M, sh, &t, "wos_hip_batch_set_tuning");
  if (rc) return rc;
  WOS_HIP_TRY(hipStreamSynchronize(b->stream));
  batch_drop_graph(b);
  b->tune = t;
  return WOS_OK;
}

int wos_hip_batch_capture(wos_batch b) {
  WOS_ENTER();
  if (!b) return WOS_ERR_INVALID;
  WOS_HIP_TRY(hipStreamSynchronize(b->stream));
  batch_drop_graph(b);
  {  // twiddle and chirp tables are created (allocated) outside the capture
    TransferLaunch L;
    int rc0 = batch_build(b, L);
    if (rc0) return rc0;
  }
  WOS_HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  int rc = batch_enqueue(b);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) {
    set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    return WOS_ERR_HIP;
  }
  b->graph = g;
  WOS_HIP_TRY(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
  return WOS_OK;
}

int wos_hip_batch_run(wos_batch b, int iters) {
  WOS_ENTER();
  if (!b || iters < 0) return WOS_ERR_INVALID;
  for (int i = 0; i < iters; ++i) {
    if (b->exec) {
      WOS_HIP_TRY(hipGraphLaunch(b->exec, b->stream));
    } else {
      int rc = batch_enqueue(b);
      if (rc) return rc;
    }
  }
  return WOS_OK;
}

int wos_hip_batch_sync(wos_batch b) {
  if (!b) return WOS_ERR_INVALID;
  WOS_HIP_TRY(hipStreamSynchronize(b->stream));
  return batch_persist_status(b);
}

int wos_hip_batch_set_persistent(wos_batch b, int on) {
  WOS_ENTER();
  if (!b) return WOS_ERR_INVALID;
  WOS_HIP_TRY(hipStreamSynchronize(b->stream));
  if (!on) {
    if (b->persist) batch_drop_graph(b);
    b->persist = false;
    if (b->persist_ctr) (void)hipFree(b->persist_ctr);
    b->persist_ctr = nullptr;
    return WOS_OK;
  }
  if (b->pad || !persist_built(b->prec, b->M)) {
    set_error("wos_hip_batch_set_persistent: built for asm batches of side 512, 1024, 2048 (FP32) and 1024, 2048 (FP64)");
    return WOS_ERR_UNSUPPORTED;
  }
  static int xcc_checked = 0;  // 1: workgroups go to the 8 XCDs in turn; -1: they do not
  static std::string xcc_map;
  if (!xcc_checked) {
    bool ok = false;
    int rc = persist_xcc_map_ok(b->stream, &ok);
    if (rc) return rc;
    xcc_checked = ok ? 1 : -1;
    if (!ok) xcc_map = g_err;
  }
  if (xcc_checked < 0) {
    set_error("wos_hip_batch_set_persistent: workgroups are not placed on the XCDs in turn on this GPU (" + xcc_map +
              ")");
    return WOS_ERR_UNSUPPORTED;
  }
  int grid = 0;
  int rc = persist_grid(b->prec, b->M, &grid);
  if (rc) return rc;
  const size_t bytes = sizeof(unsigned) * persist_ctr_words(b->units);
  if (!b->persist_ctr) {
    hipError_t e = hipMalloc((void**)&b->persist_ctr, bytes);
    if (e != hipSuccess) {
      b->persist_ctr = nullptr;
      set_error(std::string("hipMalloc(persistent counters): ") + hipGetErrorString(e));
      return WOS_ERR_NOMEM;
    }
  }
  WOS_HIP_TRY(hipMemset(b->persist_ctr, 0, bytes));
  WOS_HIP_TRY(hipDeviceSynchronize());
  batch_drop_graph(b);
  b->persist = true;
  b->persist_grid = grid;
  b->persist_interleave = on == 2;
  return WOS_OK;
}

int wos_hip_batch_time(wos_batch b, int warmup, int iters, double* ms_run, double* kernel_ms) {
  WOS_ENTER();
  if (!b || iters <= 0 || !ms_run) return WOS_ERR_INVALID;
  int rc;
  hipEvent_t e0, e1;
  WOS_HIP_TRY(hipEventCreate(&e0));
  WOS_HIP_TRY(hipEventCreate(&e1));
  if ((rc = wos_hip_batch_run(b, warmup))) return rc;
  WOS_HIP_TRY(hipEventRecord(e0, b->stream));
  if ((rc = wos_hip_batch_run(b, iters))) return rc;
  WOS_HIP_TRY(hipEventRecord(e1, b->stream));
  WOS_HIP_TRY(hipEventSynchronize(e1));
  float ms = 0;
  WOS_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
  *ms_run = ms / iters;
  if ((rc = batch_persist_status(b))) return rc;
  if (kernel_ms && b->persist) {
    // one launch per run: its time, and no separate passes
    WOS_HIP_TRY(hipEventRecord(e0, b->stream));
    for (int i = 0; i < iters; ++i)
      if ((rc = batch_enqueue(b))) return rc;
    WOS_HIP_TRY(hipEventRecord(e1, b->stream));
    WOS_HIP_TRY(hipEventSynchronize(e1));
    WOS_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    kernel_ms[0] = ms / iters;
    kernel_ms[1] = kernel_ms[2] = 0;
    if ((rc = batch_persist_status(b))) return rc;
  } else if (kernel_ms) {
    // each kernel of the pipeline back to back on the batch stream
    TransferLaunch L;
    if ((rc = batch_build(b, L))) return rc;
    for (int k = 0; k < 3; ++k) {
      if ((rc = launch_transfer_kernel(L, k, b->stream))) return rc;
      WOS_HIP_TRY(hipEventRecord(e0, b->stream));
      for (int i = 0; i < iters; ++i)
        if ((rc = launch_transfer_kernel(L, k, b->stream))) return rc;
      WOS_HIP_TRY(hipEventRecord(e1, b->stream));
      WOS_HIP_TRY(hipEventSynchronize(e1));
      WOS_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
      kernel_ms[k] = ms / iters;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return WOS_OK;
}

}  // extern "C"

#include "wos_adjoint.inc"
#include "wos_pipeline.inc"
#include "wos_stats.inc"
#include "wos_field.inc"
#include "wos_aniso.inc"
#include "wos_noalloc.inc"
#include "wos_tuning.inc"
#include "wos_long.inc"
