// Host-side check of libwos_hip's C ABI, built with the host code under a
// sanitizer (tests/native/Makefile: asan, tsan). It drives the host
// machinery -- the pinned transfer ring and its copy threads, the device
// mutex under concurrent callers, batches, pipelines, resident fields, the
// error paths and failure injection -- and checks results against a direct
// DFT on small grids and round trips on large ones. Exit status 0: every
// check passed and the sanitizer reported nothing.
#include <complex>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include <unistd.h>

#include "wos_hip.h"

using cd = std::complex<double>;
static int g_fail = 0;
#define CHECK(cond, ...)                                               \
  do {                                                                 \
    if (!(cond)) {                                                     \
      fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__);             \
      fprintf(stderr, __VA_ARGS__);                                    \
      fprintf(stderr, " (last error: %s)\n", wos_hip_last_error());    \
      ++g_fail;                                                        \
    }                                                                  \
  } while (0)

static std::vector<cd> field(int n, double seed) {
  std::vector<cd> a((size_t)n * n);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i)
      a[(size_t)j * n + i] = cd(std::sin(seed * (i + 1) * 0.7) + 0.5 * std::cos(seed * (j + 1)),
                                std::cos(seed * (i - j) * 0.3) - 0.25 * std::sin(seed * (i * j + 1)));
  return a;
}

static std::vector<cd> dft2(const std::vector<cd>& a, int n) {
  std::vector<cd> out((size_t)n * n);
  const double tp = -2 * M_PI / n;
  for (int v = 0; v < n; ++v)
    for (int u = 0; u < n; ++u) {
      cd s = 0;
      for (int y = 0; y < n; ++y)
        for (int x = 0; x < n; ++x) s += a[(size_t)y * n + x] * std::polar(1.0, tp * ((double)u * x + (double)v * y));
      out[(size_t)v * n + u] = s;
    }
  return out;
}

static double rel(const std::vector<cd>& a, const std::vector<cd>& b) {
  double num = 0, den = 0;
  for (size_t i = 0; i < a.size(); ++i) {
    num += std::norm(a[i] - b[i]);
    den += std::norm(b[i]);
  }
  return den > 0 ? std::sqrt(num / den) : 0;
}

static double* d(std::vector<cd>& a) { return reinterpret_cast<double*>(a.data()); }

int main() {
  setvbuf(stdout, nullptr, _IONBF, 0);
  if (wos_hip_device_count() <= 0) {
    printf("no device\n");
    return 2;
  }
  CHECK(wos_hip_init() == WOS_OK, "init");
  // 1. drop-in 2-D FFTs against a direct DFT: power of two, mixed radix, Bluestein
  for (int n : {16, 12, 11}) {
    std::vector<cd> a = field(n, 1.3), want = dft2(a, n);
    CHECK(wos_hip_fft2d_z(d(a), n, 0) == WOS_OK, "fft2d_z n=%d", n);
    CHECK(rel(a, want) < 1e-12, "fft2d_z n=%d rel %g", n, rel(a, want));
  }
  // 2. large buffers through the pinned ring (64 MiB each way): round trip
  {
    const int n = 2048;
    std::vector<cd> a = field(n, 0.9), keep = a;
    CHECK(wos_hip_fft2d_z(d(a), n, 0) == WOS_OK && wos_hip_fft2d_z(d(a), n, 1) == WOS_OK, "2048 round trip");
    CHECK(rel(a, keep) < 1e-13, "2048 round trip rel %g", rel(a, keep));
  }
  // 3. concurrent callers: 8 host threads, 2 MiB and larger buffers (the ring) and small ones
  {
    std::vector<std::thread> th;
    std::vector<int> bad(8, 0);
    for (int k = 0; k < 8; ++k)
      th.emplace_back([k, &bad] {
        const int n = (k & 1) ? 1024 : 128;
        std::vector<cd> a = field(n, 0.3 + k), keep = a;
        for (int it = 0; it < 6; ++it) {
          if (wos_hip_fft2d_z(d(a), n, it & 1) != WOS_OK) bad[k] = 1;
        }
        if (rel(a, keep) > 1e-12) bad[k] = 2;
      });
    for (auto& t : th) t.join();
    for (int k = 0; k < 8; ++k) CHECK(bad[k] == 0, "thread %d: %d", k, bad[k]);
  }
  // 4. Propagate with warnings and stats; FP32
  {
    const int n = 256;
    std::vector<cd> ex = field(n, 0.5), ey = field(n, 0.6);
    wos_ctx ctx = {632.8e-9, 0, 0, 0.0, 0.0};
    wos_warnings w;
    wos_prop_stats st;
    memset(&w, 0, sizeof(w));
    CHECK(wos_hip_propagate(d(ex), d(ey), n, 2e-6, 5e-3, WOS_METHOD_ASM_PAD, &ctx, WOS_FP64, &w, &st) == WOS_OK,
          "propagate");
    CHECK(st.p_in > 0 && std::fabs(st.p_out / st.p_in - 1) < 0.5, "powers %g %g", st.p_in, st.p_out);
    CHECK(wos_hip_propagate(d(ex), nullptr, n, 2e-6, 5e-3, WOS_METHOD_ASM_SHIFT, &ctx, WOS_FP32, nullptr, nullptr) ==
              WOS_OK,
          "propagate fp32");
  }
  // 5. a batch: create, fill, autotune, capture, replay, download, destroy
  {
    const int n = 512, u = 4;
    const double wl[u] = {450e-9, 500e-9, 600e-9, 700e-9}, zs[u] = {1e-3, 2e-3, 3e-3, 4e-3};
    wos_ctx ctx = {550e-9, 0, 0, 0.0, 0.0};
    wos_batch b = nullptr;
    CHECK(wos_hip_batch_create(&b, n, 2e-6, WOS_METHOD_ASM, WOS_FP32, u, wl, zs, &ctx) == WOS_OK, "batch_create");
    if (b) {
      int shape[8];
      double ms = 0;
      CHECK(wos_hip_batch_fill_schell(b, 100e-6, 20e-6, 7) == WOS_OK, "fill_schell");
      CHECK(wos_hip_batch_autotune(b, 1, shape, &ms) == WOS_OK && ms > 0, "autotune");
      CHECK(wos_hip_batch_fill_schell(b, 100e-6, 20e-6, 7) == WOS_OK, "fill_schell");
      CHECK(wos_hip_batch_capture(b) == WOS_OK && wos_hip_batch_run(b, 3) == WOS_OK && wos_hip_batch_sync(b) == WOS_OK,
            "capture/run");
      std::vector<cd> out((size_t)n * n);
      CHECK(wos_hip_batch_download(b, 3, d(out)) == WOS_OK, "download");
      CHECK(wos_hip_batch_destroy(b) == WOS_OK, "destroy");
    }
  }
  // 6. a pipeline of two stages (one shared source), captured and replayed
  {
    wos_pipeline p = nullptr;
    wos_ctx ctx = {550e-9, 0, 0, 0.0, 0.0};
    CHECK(wos_hip_pipeline_create(&p, WOS_FP64) == WOS_OK, "pipeline_create");
    if (p) {
      const double wl[3] = {450e-9, 550e-9, 650e-9}, zs[3] = {1e-3, 1e-3, 1e-3};
      int s0 = -1, s1 = -1;
      CHECK(wos_hip_pipeline_add(p, 256, 2e-6, WOS_METHOD_ASM, 3, 1, wl, zs, &ctx, &s0) == WOS_OK, "add");
      CHECK(wos_hip_pipeline_add(p, 100, 2e-6, WOS_METHOD_ASM_PAD, 1, 1, wl, zs, &ctx, &s1) == WOS_OK, "add");
      const double wts[3] = {1, 1, 1};
      CHECK(wos_hip_pipeline_set_intensity(p, s0, wts) == WOS_OK, "set_intensity");
      std::vector<cd> src = field(256, 0.2);
      CHECK(wos_hip_pipeline_upload(p, s0, 0, d(src)) == WOS_OK, "upload");
      CHECK(wos_hip_pipeline_capture(p) == WOS_OK && wos_hip_pipeline_run(p, 2) == WOS_OK &&
                wos_hip_pipeline_sync(p) == WOS_OK,
            "pipeline run");
      std::vector<double> inten(256 * 256);
      CHECK(wos_hip_pipeline_intensity(p, s0, inten.data()) == WOS_OK, "intensity");
      CHECK(wos_hip_pipeline_destroy(p) == WOS_OK, "pipeline_destroy");
    }
  }
  // 7. a resident field through a few elements, a view and its stats
  {
    wos_field f = nullptr;
    CHECK(wos_hip_field_create(&f, 256, 2e-6, 1, WOS_FP64) == WOS_OK, "field_create");
    if (f) {
      std::vector<cd> ex = field(256, 0.4);
      wos_ctx ctx = {633e-9, 0, 0, 0.0, 0.0};
      CHECK(wos_hip_field_upload(f, d(ex), nullptr) == WOS_OK, "field_upload");
      CHECK(wos_hip_field_lens(f, 633e-9, 0.1, 0.0, 0.0, 0.0) == WOS_OK, "lens");
      CHECK(wos_hip_field_propagate(f, 5e-3, WOS_METHOD_ASM, &ctx, nullptr, nullptr) == WOS_OK, "field_propagate");
      std::vector<float> view(256 * 256);
      CHECK(wos_hip_field_view(f, 0, 0.0, view.data()) == WOS_OK, "view");
      wos_plane_stats ps;
      CHECK(wos_hip_field_stats(f, 633e-9, 0.0, 0.0, &ps) == WOS_OK && ps.power > 0, "stats");
      CHECK(wos_hip_field_destroy(f) == WOS_OK, "field_destroy");
    }
  }
  // 8. error paths and failure injection; the library stays usable
  {
    std::vector<cd> a = field(16, 2.0), keep = a;
    CHECK(wos_hip_fft2d_z(nullptr, 16, 0) == WOS_ERR_INVALID && strlen(wos_hip_last_error()) > 0, "bad pointer");
    CHECK(wos_hip_fft2d_z(d(a), 1 << 15, 0) != WOS_OK, "unsupported length");
    setenv("WOS_HIP_FAIL", "wos_hip_fft2d_z", 1);
    CHECK(wos_hip_fft2d_z(d(a), 16, 0) == WOS_ERR_HIP, "injected failure");
    CHECK(rel(a, keep) == 0, "injected failure left the buffer");
    unsetenv("WOS_HIP_FAIL");
    uint64_t need = 0, budget = 0;
    CHECK(wos_hip_check_grid_memory(2048, WOS_METHOD_ASM, WOS_FP64, 100000, &need, &budget) == WOS_ERR_NOMEM,
          "budget refusal");
    CHECK(wos_hip_release_cache() == WOS_OK, "release_cache");
    std::vector<cd> want = dft2(keep, 16);
    CHECK(wos_hip_fft2d_z(d(a), 16, 0) == WOS_OK && rel(a, want) < 1e-12, "fft after release");
  }
  printf("host check: %d failure(s)\n", g_fail);
  fflush(stdout);
  fflush(stderr);
  // leave without running the runtimes' exit-time teardown, where the HIP
  // runtime's release of its allocations meets the sanitizer's allocator
  _exit(g_fail ? 1 : 0);
}
