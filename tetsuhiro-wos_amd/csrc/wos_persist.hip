// wos_persist.hip — the persistent transfer: one launch runs the forward
// rows, the column pass and the inverse rows of every unit of a batch (asm
// transfer function, power-of-two side, radix-16 schedules, split LDS
// exchanges) from per-XCD work queues.
//
// Placement. A workgroup reads the XCD it runs on from the XCC_ID hardware
// register and serves only that XCD's queue. Unit u belongs to XCD u % 8,
// so every task reads data that earlier tasks wrote through the same L2: a
// producer waits for its stores (s_waitcnt vmcnt(0)) and then counts the
// task done; a consumer waits for the count and then invalidates its CU's
// L1 (buffer_inv sc0) before it loads. These are the release and acquire of
// the workgroup scope in threadgroup-split mode, where the waves of one
// workgroup may sit on different CUs of one XCD. Every counter access is an
// atomic read-modify-write (a poll adds an opaque zero), so reads and
// updates meet at one point of coherence wherever the hardware performs
// them; each XCD's ticket and abort flag and each unit's counts sit on
// their own 128-byte line.
//
// Order. Queue x holds rounds r = 0 .. nu + 1 (nu units on the XCD): the
// inverse rows of its unit r - 2, the columns of unit r - 1 and the forward
// rows of unit r, either pass after pass (a task's inputs were finished a
// whole pass of tickets earlier, so waits are rare) or interleaved task by
// task (a CU mixes VALU-bound column tasks with memory-bound row tasks, but
// a round's first tasks wait for the previous round's last). A task waits
// only for tasks with lower tickets. A workgroup claims its next ticket
// when it starts a task (the atomic's latency hides behind the task), so
// the lowest unfinished ticket is always held by a workgroup that is
// running it, whose own waits are for lower, finished tickets: every wait
// ends. Waits are bounded anyway: past spin_limit sleeps a workgroup raises
// its XCD's abort flag and stops, the other workgroups of the XCD stop at
// their next wait, the launch drains, and the check kernel records the
// failure for the host.
//
// Registers. Run in a loop, the three task bodies would let the optimiser
// hoist what they derive from the arguments and the thread index out of
// the loop and hold it live across all of them (256 VGPRs and spills where
// each kernel alone takes 74-80). So the bodies take an opaque thread index,
// the arguments are re-read per task through an opaque pointer to the
// kernel-argument segment, this file is compiled without machine-code
// loop-invariant motion (-mllvm -disable-machine-licm), and the kernel is
// capped at the column pass's occupancy.
//
// Counters (PersistArgs::ctr, persist_ctr_words(units) words): word
// x * kPersistLine the ticket of XCD x, the next word its abort flag;
// kPersistStatus the sticky status (bit 0: a wait gave up, bit 1: a queue
// was not drained; cleared only by the host); kPersistUnits + u *
// kPersistLine the forward-row tasks done of unit u, the next word its
// column tasks done.
#include "../../include/wos_hip.h"
#include "wos_engine.hpp"

namespace wos {

__device__ __forceinline__ int xcc_id() { return __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & 15; }

__device__ __forceinline__ int persist_units_of(const PersistArgs& a, int x) {
  return a.units > x ? (a.units - x + 7) / 8 : 0;
}

// Task `tk` of XCD x's queue: phase (0 forward rows, 1 columns, 2 inverse
// rows), the unit, and the workgroup index within the unit's pass; false
// past the end of the queue.
__device__ __forceinline__ bool persist_task(const PersistArgs& a, int x, unsigned tk, int* phase, int* unit,
                                             int* blk) {
  const int nu = persist_units_of(a, x);
  const bool inter = a.interleave != 0;
  unsigned base = 0;
  for (int r = 0; r < nu + 2; ++r) {
    const bool vi = r >= 2 && r - 2 < nu, vc = r >= 1 && r - 1 < nu, vf = r < nu;
    const int ni = vi ? a.rb : 0, nc = vc ? a.cb : 0, nf = vf ? a.rb : 0;
    const unsigned size = (unsigned)(ni + nc + nf);
    if (tk < base + size) {
      unsigned o = tk - base;
      int ph = 0;
      if (inter) {  // task by task: inverse, column, forward
        const int nv = (vi ? 1 : 0) + (vc ? 1 : 0) + (vf ? 1 : 0);
        const int q = (int)(o % (unsigned)nv);
        const int order[3] = {vi ? 2 : -1, vc ? 1 : -1, vf ? 0 : -1};
        int seen = -1;
        for (int k = 0; k < 3; ++k)
          if (order[k] >= 0 && ++seen == q) ph = order[k];
        *blk = (int)(o / (unsigned)nv);
      } else {  // whole passes in turn
        if (o < (unsigned)ni) {
          ph = 2;
        } else if (o < (unsigned)(ni + nc)) {
          ph = 1;
          o -= ni;
        } else {
          ph = 0;
          o -= ni + nc;
        }
        *blk = (int)o;
      }
      *phase = ph;
      *unit = x + 8 * (ph == 2 ? r - 2 : ph == 1 ? r - 1 : r);
      return true;
    }
    base += size;
  }
  return false;
}

// Counter value by an atomic add of an opaque zero (an idempotent add the
// optimiser would otherwise turn into a plain load).
__device__ __forceinline__ unsigned ctr_read(unsigned* p) {
  unsigned zero = 0;
  asm volatile("" : "+v"(zero));
  return __hip_atomic_fetch_add(p, zero, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename R, int N, int ROWS, int OCC>
__global__ __launch_bounds__((ROWS * Plan<N, 16>::T)) __attribute__((amdgpu_waves_per_eu(OCC))) void k_transfer_persist(
    const PersistArgs a) {
  constexpr int LR = rows_lds_elems<R, N, ROWS, true>(), LC = cols_lds_elems<R, N, ROWS, true>();
  __shared__ cpx<R> lds[LR > LC ? LR : LC];
  __shared__ double red[(ROWS * Plan<N, 16>::T + 63) / 64];
  __shared__ int task[4];
  const int x = xcc_id();
  if (x >= 8) return;  // no queue (the host checks the XCD map before it enables this kernel)
  unsigned* const ticket = a.ctr + x * kPersistLine;
  unsigned* const abort_x = ticket + 1;
  unsigned next = 0;  // thread 0: the ticket claimed for the next task
  if (threadIdx.x == 0) next = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  for (;;) {
    if (threadIdx.x == 0) {
      int ph = -1, u = 0, b = 0, ok = 1;
      if (!persist_task(a, x, next, &ph, &u, &b)) {
        ok = 0;
      } else if (ph != 0) {  // wait for the unit's previous pass
        unsigned* const cnt = a.ctr + kPersistUnits + u * kPersistLine + (ph == 1 ? 0 : 1);
        const unsigned want = (unsigned)(ph == 1 ? a.rb : a.cb);
        int it = 0;
        while (ctr_read(cnt) < want) {
          if (++it > a.spin_limit || ctr_read(abort_x) != 0) {
            __hip_atomic_fetch_or(abort_x, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            ok = 0;
            break;
          }
          __builtin_amdgcn_s_sleep(2);
        }
      }
      if (ok) next = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      task[0] = ok;
      task[1] = ph;
      task[2] = u;
      task[3] = b;
    }
    __syncthreads();
    // task[] is rewritten only after the end-of-task barrier below
    const int ok = __builtin_amdgcn_readfirstlane(task[0]), ph = __builtin_amdgcn_readfirstlane(task[1]);
    const int u = __builtin_amdgcn_readfirstlane(task[2]), b = __builtin_amdgcn_readfirstlane(task[3]);
    if (!ok) break;
    // acquire: drop L1 lines that producers on other CUs of this XCD rewrote
    asm volatile("buffer_inv sc0\n\ts_waitcnt vmcnt(0)" ::: "memory");
    auto ka = __builtin_amdgcn_kernarg_segment_ptr();
    asm volatile("" : "+s"(ka));
    const PersistArgs& t = *(const PersistArgs*)ka;
    if (ph == 0) {
      rows_body<R, N, -1, ROWS, 0, true, 16, true>(t.fwd, u * t.rb + b, lds, red);
    } else if (ph == 1) {
      cols_body<R, N, ROWS, 1, 1, 16, true, true>(t.col, u * t.cb + b, t.units * t.cb, lds);
    } else {
      rows_body<R, N, 1, ROWS, 0, true, 16, true>(t.inv, u * t.rb + b, lds, red);
    }
    // release: this workgroup's stores reach the L2 before the count
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    if (threadIdx.x == 0 && ph != 2)
      __hip_atomic_fetch_add(a.ctr + kPersistUnits + u * kPersistLine + ph, 1u, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Zeroes the tickets, the abort flags and the done counts (a kernel, not a
// memset, so that the step can be captured into a graph).
__global__ void k_persist_reset(unsigned* ctr, int units) {
  for (int i = threadIdx.x; i < 8; i += blockDim.x) ctr[i * kPersistLine] = ctr[i * kPersistLine + 1] = 0;
  for (int i = threadIdx.x; i < units; i += blockDim.x)
    ctr[kPersistUnits + i * kPersistLine] = ctr[kPersistUnits + i * kPersistLine + 1] = 0;
}

// Records in the sticky status word whether the launch gave up a wait or
// left a queue with unclaimed tasks.
__global__ void k_persist_check(const PersistArgs a) {
  if (threadIdx.x != 0) return;
  unsigned st = 0;
  for (int x = 0; x < 8; ++x) {
    const unsigned tasks = (unsigned)persist_units_of(a, x) * (unsigned)(2 * a.rb + a.cb);
    if (a.ctr[x * kPersistLine + 1]) st |= 1u;
    if (a.ctr[x * kPersistLine] < tasks) st |= 2u;
  }
  if (st) a.ctr[kPersistStatus] |= st;
}

// XCD of each workgroup of a small launch.
__global__ void k_xcc_map(int* out) {
  if (threadIdx.x == 0) out[blockIdx.x] = xcc_id();
}

template <typename R, int N>
struct PersistShape;
template <int N>
struct PersistShape<float, N> {
  static constexpr int ROWS = 4, OCC = 6;
};
template <int N>
struct PersistShape<double, N> {
  static constexpr int ROWS = 2, OCC = 3;
};

template <typename R, int N>
static auto persist_kernel() {
  return k_transfer_persist<R, N, PersistShape<R, N>::ROWS, PersistShape<R, N>::OCC>;
}

// Sides built: FP32 512, 1024, 2048; FP64 1024, 2048 (two-column FP64
// tiles below 1024 points are not built for the column pass).
bool persist_built(int prec, int N) {
  if (prec == WOS_FP64) return N == 1024 || N == 2048;
  return N == 512 || N == 1024 || N == 2048;
}

int persist_rows(int prec) { return prec == WOS_FP64 ? PersistShape<double, 2048>::ROWS : PersistShape<float, 2048>::ROWS; }

template <typename R, int N>
static int persist_grid_T(int* grid) {
  constexpr int NT = PersistShape<R, N>::ROWS * Plan<N, 16>::T;
  int per_cu = 0, dev = 0, cus = 0;
  WOS_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, persist_kernel<R, N>(), NT, 0));
  WOS_HIP_TRY(hipGetDevice(&dev));
  WOS_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const int g = per_cu * cus / 8 * 8;
  if (g < 8) {
    set_error("persistent transfer: no resident workgroup per XCD");
    return WOS_ERR_UNSUPPORTED;
  }
  *grid = g;
  return 0;
}

// Workgroups of one persistent launch: as many as the GPU holds at once
// (a multiple of 8, so each XCD gets the same number).
int persist_grid(int prec, int N, int* grid) {
  if (prec == WOS_FP64) {
    if (N == 1024) return persist_grid_T<double, 1024>(grid);
    if (N == 2048) return persist_grid_T<double, 2048>(grid);
  } else {
    if (N == 512) return persist_grid_T<float, 512>(grid);
    if (N == 1024) return persist_grid_T<float, 1024>(grid);
    if (N == 2048) return persist_grid_T<float, 2048>(grid);
  }
  set_error("persistent transfer not built for this side and precision");
  return WOS_ERR_UNSUPPORTED;
}

template <typename R, int N>
static int persist_launch_T(int grid, const PersistArgs& a, hipStream_t s) {
  constexpr int NT = PersistShape<R, N>::ROWS * Plan<N, 16>::T;
  hipLaunchKernelGGL(k_persist_reset, dim3(1), dim3(256), 0, s, a.ctr, a.units);
  hipLaunchKernelGGL((persist_kernel<R, N>()), dim3(grid), dim3(NT), 0, s, a);
  hipLaunchKernelGGL(k_persist_check, dim3(1), dim3(64), 0, s, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error(std::string("persistent transfer launch: ") + hipGetErrorString(e));
    return WOS_ERR_HIP;
  }
  return 0;
}

// The reset, the persistent kernel on `grid` workgroups, and the check, on
// stream s. a.ctr holds persist_ctr_words(a.units) words.
int launch_transfer_persist(int prec, int N, int grid, const PersistArgs& a, hipStream_t s) {
  if (prec == WOS_FP64) {
    if (N == 1024) return persist_launch_T<double, 1024>(grid, a, s);
    if (N == 2048) return persist_launch_T<double, 2048>(grid, a, s);
  } else {
    if (N == 512) return persist_launch_T<float, 512>(grid, a, s);
    if (N == 1024) return persist_launch_T<float, 1024>(grid, a, s);
    if (N == 2048) return persist_launch_T<float, 2048>(grid, a, s);
  }
  set_error("persistent transfer not built for this side and precision");
  return WOS_ERR_UNSUPPORTED;
}

// Whether the XCC_ID register reads 0..7 and the workgroups of a launch go
// to the XCDs in turn (workgroup i to XCD (x0 + i) % 8: the dispatcher's
// turn carries over from the previous launch, so x0 varies), which gives
// every XCD grid / 8 workgroups of a persistent launch: one launch of 64
// workgroups.
int persist_xcc_map_ok(hipStream_t s, bool* ok) {
  int* d = nullptr;
  WOS_HIP_TRY(hipMalloc(&d, 64 * sizeof(int)));
  hipLaunchKernelGGL(k_xcc_map, dim3(64), dim3(64), 0, s, d);
  int h[64];
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(d);
  if (e != hipSuccess) {
    set_error(std::string("XCD map: ") + hipGetErrorString(e));
    return WOS_ERR_HIP;
  }
  *ok = true;
  for (int i = 0; i < 64; ++i) *ok = *ok && h[i] >= 0 && h[i] < 8 && h[i] == (h[0] + i) % 8;
  if (!*ok) {
    std::string m = "XCD of workgroups 0..15:";
    for (int i = 0; i < 16; ++i) m += " " + std::to_string(h[i]);
    set_error(m);
  }
  return 0;
}

}  // namespace wos
