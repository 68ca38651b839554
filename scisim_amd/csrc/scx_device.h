// scx_device.h -- device helpers (wave64 reductions/scans, hashing, packed cell keys).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace scx {

__device__ __forceinline__ uint32_t wave_inclusive_scan_u32(uint32_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t t = __shfl_up(v, off, 64);
    if (lane >= off) v += t;
  }
  return v;
}

__device__ __forceinline__ uint32_t wave_reduce_sum_u32(uint32_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block-wide sum over blockDim.x == 256 threads; result valid in every thread.
__device__ __forceinline__ uint32_t block_reduce_sum_u32(uint32_t v) {
  __shared__ uint32_t s[4];
  v = wave_reduce_sum_u32(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
  __syncthreads();
  return s[0] + s[1] + s[2] + s[3];
}

// Block-wide exclusive scan over 256 threads; *total = block sum.
__device__ __forceinline__ uint32_t block_exclusive_scan_u32(uint32_t v, uint32_t* total) {
  __shared__ uint32_t s[4];
  const uint32_t incl = wave_inclusive_scan_u32(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 63) s[threadIdx.x >> 6] = incl;
  __syncthreads();
  const int w = threadIdx.x >> 6;
  uint32_t before = 0;
  for (int k = 0; k < w; ++k) before += s[k];
  *total = s[0] + s[1] + s[2] + s[3];
  return before + incl - v;
}

__device__ __forceinline__ double wave_reduce_max_f64(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}
__device__ __forceinline__ double wave_reduce_min_f64(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
  return v;
}
__device__ __forceinline__ float wave_reduce_sum_f32(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ double wave_reduce_sum_f64(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Logical workgroup index for launches whose neighbouring workgroups share data. The
// dispatcher hands workgroups to the XCDs round-robin (workgroup b to XCD b % 8) and every
// XCD has its own L2; renumbering gives each XCD one contiguous range of logical indices
// (a bijection for any grid size), so lines shared by neighbouring tiles are fetched
// into one L2 instead of up to eight.
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_block() {
  const uint32_t b = blockIdx.x, nb = gridDim.x;
  const uint32_t q = nb / kXcds, rem = nb % kXcds;
  const uint32_t x = b % kXcds, i = b / kXcds;
  return x * q + min(x, rem) + i;
}

// Waits until this wave's outstanding memory operations (stores included) have completed.
__device__ __forceinline__ void wait_stores() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// Grid-wide barrier for persistent kernels (launched cooperatively, so that every
// workgroup is resident). One wave per workgroup arrives on a counter with an
// agent-scope release-acquire increment; the last arriver resets the counter and bumps
// the generation word, which the others poll with plain atomic loads and follow with one
// agent-scope acquire fence. The counter returns to zero, so the state is reusable across
// launches of any width without a reset. Every wave first waits for its own stores to
// complete (vmcnt 0; a workgroup barrier alone does not wait for them on CDNA). The
// agent-scope release (the L2 write-back that publishes the workgroup's stores to the
// other XCDs) and acquire (the invalidation that drops stale lines for the whole CU) are
// each issued once per workgroup: on MI355X their cost, not the atomics, is what grows
// with the workgroup count. Measured with 512-thread workgroups: 1.5 us at 8 workgroups,
// 2.9 us at 64, 4.8 us at 128 (a two-level tree of counters with a fence on each side:
// 3.4, 7.1 and 14.0 us).
// Layout (u32 words): [0] generation, [32] arrival counter (separate 128-byte lines); the
// coherent barrier below: [64] 64-bit arrival count of the groups, [96] barriers
// completed, [128] exit count, [160 + 32 g] 64-bit arrival count of group g (8 groups);
// [144..146] free for a kernel's own counts.
constexpr uint32_t kBarrierWords = 160 + 32 * 8;
constexpr uint32_t kBarrierGroups = 8;
constexpr uint32_t kBarrierTwoLevel = 128;

__device__ __forceinline__ void grid_barrier(uint32_t* bar) {
  // one workgroup: its waves share the CU's L1, a workgroup barrier orders their
  // global-memory accesses (no device-scope fence or atomics needed)
  if (gridDim.x == 1) {
    __syncthreads();
    return;
  }
  wait_stores();
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t* gen = bar;
    uint32_t* cnt = bar + 32;
    // the generation before this workgroup arrives (it cannot advance before then); the
    // release increment below orders this load before the arrival
    const uint32_t g = __hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (__hip_atomic_fetch_add(cnt, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == gridDim.x - 1) {
      __hip_atomic_store(cnt, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(gen, g + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      while (__hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == g) __builtin_amdgcn_s_sleep(1);
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
  }
  __syncthreads();
}

// The same barrier without agent-scope cache maintenance (no L2 writeback on arrival, no
// L2 invalidate on departure): for intervals in which every access to data that another
// workgroup wrote or will read since the previous barrier is a coherent (SC1) load or
// store, which goes to the memory side past the per-XCD L2s. Each thread's stores complete
// (s_waitcnt) before its workgroup arrives, so they are in memory when any workgroup
// leaves; loads issued after the departure read memory. Data read only by the thread that
// wrote it, or not written during the solve, may use ordinary accesses. Measured with
// 512-thread workgroups doing two coherent gathers and two coherent stores per interval:
// 6.4 us at 128 workgroups against 9.8 us for grid_barrier with ordinary accesses.
// Split in two: the arrival (after which this workgroup may issue loads of data that no
// workgroup writes before the barrier, overlapping their latency with the wait) and the
// wait. Thread 0 keeps the ticket.
// Two levels of arrival counts that only grow during a launch: workgroup w arrives in
// group w mod 8 (the workgroups of one XCD, as the dispatcher deals them round-robin), and
// the last arrival of each group arrives at the level above; so no count takes more than 32
// of the 256 arrivals of a full chip. At its i-th barrier (from 0) a workgroup finds its
// group's count in [i n, (i + 1) n) for the group's n workgroups (every workgroup has passed
// barrier i - 1, not every one has arrived at barrier i), so the arrival that finds
// (i + 1) n - 1 is its group's last; likewise for the groups above, whose last arrival sets
// the completed-barrier word to i + 1, which the others poll (on a line of its own). Nothing
// is reset between barriers. Every workgroup of a launch that uses this barrier calls
// grid_exit_coherent before it ends: the last one resets the words for the next launch.
struct GridTicket {
  uint32_t target;
};

__device__ __forceinline__ unsigned long long* grid_coherent_count(uint32_t* bar) {
  return reinterpret_cast<unsigned long long*>(bar + 64);
}
__device__ __forceinline__ unsigned long long* grid_group_count(uint32_t* bar, uint32_t g) {
  return reinterpret_cast<unsigned long long*>(bar + 160 + 32 * g);
}

__device__ __forceinline__ GridTicket grid_arrive_coherent(uint32_t* bar) {
  wait_stores();
  __syncthreads();
  GridTicket t{0u};
  if (gridDim.x > 1 && threadIdx.x == 0) {
    const uint32_t G = gridDim.x;
    // below kBarrierTwoLevel workgroups one group (the second level would only add a hop)
    const uint32_t ng = G < kBarrierTwoLevel ? 1u : kBarrierGroups;
    const uint32_t g = blockIdx.x % ng;
    const uint32_t n = (G - g + ng - 1) / ng;  // workgroups g, g + ng, ... below G
    const unsigned long long v =
        __hip_atomic_fetch_add(grid_group_count(bar, g), 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    t.target = (uint32_t)(v / n) + 1u;
    if (v % n == n - 1) {
      if (ng == 1) {
        __hip_atomic_store(bar + 96, t.target, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      } else {
        const unsigned long long u =
            __hip_atomic_fetch_add(grid_coherent_count(bar), 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (u % ng == ng - 1) __hip_atomic_store(bar + 96, t.target, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
  }
  return t;
}

__device__ __forceinline__ void grid_wait_coherent(uint32_t* bar, const GridTicket& t) {
  if (gridDim.x > 1 && threadIdx.x == 0)
    while (__hip_atomic_load(bar + 96, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < t.target)
      __builtin_amdgcn_s_sleep(1);
  // keeps the compiler from moving the loads that follow above the wait
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  __syncthreads();
}

__device__ __forceinline__ void grid_sync_coherent(uint32_t* bar) { grid_wait_coherent(bar, grid_arrive_coherent(bar)); }

// The end of a launch that used grid_sync_coherent, in every workgroup: the last workgroup
// to get here (all others have passed their last barrier) resets the arrival count and the
// completed-barrier word.
__device__ __forceinline__ void grid_exit_coherent(uint32_t* bar) {
  if (gridDim.x > 1 && threadIdx.x == 0) {
    if (__hip_atomic_fetch_add(bar + 128, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gridDim.x - 1) {
      for (uint32_t g = 0; g < kBarrierGroups; ++g)
        __hip_atomic_store(grid_group_count(bar, g), 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(grid_coherent_count(bar), 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(bar + 96, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(bar + 128, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// Time (us) of one barrier interval of a persistent launch over `width` items with `b`
// workgroups of `threads`, from timings on MI355X. One workgroup: ~0.2 us barrier, ~1.9 us
// per pass (caches stay warm). Several: ~1.2 us + 0.035 us per workgroup of barrier, ~4.5 us
// for the first pass after the cache invalidation and ~3.5 us for each further one.
// Coherent intervals (grid_sync_coherent, coherent loads and stores): ~1.5 us + 0.004 us per
// workgroup of barrier, ~6 us for the first pass (coherent gathers and the completion of
// coherent stores) and ~5 us for each further one (the coloured sweeps of a 262,144-sphere
// pile: 273, 195 and 176 us per sweep at 64, 128 and 256 workgroups).
inline double interval_cost(uint32_t width, int threads, int b, bool coherent = false) {
  double passes = (double)((width + (uint64_t)threads * b - 1) / ((uint64_t)threads * b));
  if (passes < 1.0) passes = 1.0;
  if (b == 1) return 0.2 + 1.9 * passes;
  if (coherent) return (1.5 + 0.004 * b) + 6.0 + 5.0 * (passes - 1.0);
  return (1.2 + 0.035 * b) + 4.5 + 3.5 * (passes - 1.0);
}

// Workgroups (a power of two up to max_blocks) minimising the summed time of barrier
// intervals over the given widths (one launch runs them all with the same width).
inline int persistent_blocks(const uint32_t* widths, int nwidths, int threads, int max_blocks,
                             bool coherent = false) {
  double best = 1e30;
  int pick = 1;
  for (int b = 1; b <= max_blocks; b *= 2) {
    double t = 0.0;
    for (int k = 0; k < nwidths; ++k) t += interval_cost(widths[k], threads, b, coherent);
    if (t < best) {
      best = t;
      pick = b;
    }
  }
  return pick;
}

inline int persistent_blocks(uint32_t width, int threads, int max_blocks) {
  return persistent_blocks(&width, 1, threads, max_blocks);
}

// Contact priority for the colouring rounds: (fmix32(c * golden + salt) << 32) | c.
__device__ __forceinline__ uint32_t fmix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}
__device__ __forceinline__ unsigned long long prio_key(uint32_t c) {
  return ((unsigned long long)fmix32(c * 0x9E3779B9u + 0x632BE5ABu) << 32) | c;
}

// Cell index of a coordinate: floor((x - min) / h) as the reference's unsigned cast
// (ball2d/SpatialGridDetector.cpp:29-34), clamped to [0, dim - 1] (monotone in x, which
// is all the emission rule of the sweep needs).
__device__ __forceinline__ uint32_t cell_of(double x, double mn, double h, uint32_t dim) {
  const double t = (x - mn) / h;
  if (!(t > 0.0)) return 0u;
  if (!(t < (double)(dim - 1))) return dim - 1;
  return (uint32_t)t;
}

// FP32 cell index, clamped to [0, dim - 1]. Non-decreasing in x (every step rounds
// monotonically), which is all the binning needs; NaN maps to cell 0.
__device__ __forceinline__ uint32_t cell_of_f(float x, float mn, float inv_h, uint32_t dim) {
  const float t = (x - mn) * inv_h;
  if (!(t > 0.0f)) return 0u;
  if (!(t < (float)(dim - 1))) return dim - 1;
  return (uint32_t)t;
}

}  // namespace scx
