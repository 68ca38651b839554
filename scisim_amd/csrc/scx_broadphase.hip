// scx_broadphase.hip -- predictor + AABBs, uniform-grid binning and the per-owner
// neighbour sweep with the exact FP64 narrow phase, for gfx950.
//
// Output contract (SpatialGridDetector::getPotentialOverlaps,
// ball2d/SpatialGridDetector.cpp:106-133): every pair i<j of boxes that overlap in
// FP64 (AABB::overlaps, :15-26), once, in lexicographic order. The reference's
// grid finds exactly that set (its own test asserts equality with the all-pairs
// loop, ball2dtests/collision_detection_tests.cpp:34-40), so any binning that
// cannot miss a pair and emits it once is a drop-in.
//
// Sweep order. Every per-body kernel of the broad phase runs over a permutation of
// the bodies, order[t] = body id, taken from the previous binning's home-cell sort
// (the cell of a box's lower corner). Bodies move little in one step, so consecutive
// threads handle neighbouring bodies: the cell-count atomics hit nearby counters and
// the neighbour sweep gathers from nearby memory, without moving the resident state.
// The predictor writes compact records in sweep order: a conservative FP32 box
// (lower corner rounded down, upper corner rounded up) and the exact FP64 data of the
// narrow phase. Any permutation gives the same pairs; the order only sets the speed.
//
// Binning: the reference's cell size and origin (initializeSpatialGrid, :76-104),
// coarsened only when the dense grid would exceed its cell budget. Cells are
// assigned from the FP32 boxes with one monotone FP32 map, so two boxes that overlap
// in FP64 overlap in FP32 and share every cell between their lower-corner cells'
// component-wise max and their upper-corner cells' min. Each covered cell holds a
// 4-byte entry: the sweep position t plus "first cell along x/y/z" flags. A pair is
// emitted only from the covered cell equal to the component-wise max of the two
// boxes' first cells, so each pair is found exactly once without a global sort; the
// FP32 boxes reject most candidates, the FP64 test and the narrow phase decide; each
// owner (the smaller body id) sorts its short list.
//
// No FMA contraction in this file: the predictor, box bounds and narrow-phase tests
// must round exactly like the FP64 CPU restatement so that pair lists are bit-exact.
#pragma clang fp contract(off)

#include "scx_device.h"
#include "scx_flow.h"
#include "scx_internal.h"

namespace scx {

constexpr uint32_t kEntIndexMask = (1u << 29) - 1;

// Block partials of the grid bounds (min of mins, max of maxs, sum of extents) from
// per-thread box values; partials[blk * 9 + 3 a + {0, 1, 2}].
template <int D>
__device__ __forceinline__ void bounds_partials(bool valid, const double* lo_v, const double* hi_v,
                                                double* __restrict__ partials) {
  __shared__ double s[9][kBlock / 64];
  const int wave = threadIdx.x >> 6;
#pragma unroll
  for (int a = 0; a < D; ++a) {
    double lo = valid ? lo_v[a] : INFINITY;
    double hi = valid ? hi_v[a] : -INFINITY;
    double ext = valid ? hi_v[a] - lo_v[a] : 0.0;
    lo = wave_reduce_min_f64(lo);
    hi = wave_reduce_max_f64(hi);
    ext = wave_reduce_sum_f64(ext);
    if ((threadIdx.x & 63) == 0) {
      s[3 * a + 0][wave] = lo;
      s[3 * a + 1][wave] = hi;
      s[3 * a + 2][wave] = ext;
    }
  }
  __syncthreads();
  if (threadIdx.x < 3 * D) {
    const int kind = threadIdx.x % 3;
    double v = s[threadIdx.x][0];
    for (int w = 1; w < kBlock / 64; ++w) {
      const double o = s[threadIdx.x][w];
      v = (kind == 0) ? fmin(v, o) : (kind == 1) ? fmax(v, o) : v + o;
    }
    partials[(size_t)blockIdx.x * 9 + threadIdx.x] = v;
  }
}

// Largest float <= x and smallest float >= x.
__device__ __forceinline__ float f32_down(double x) {
  const float f = (float)x;
  return ((double)f > x) ? nextafterf(f, -INFINITY) : f;
}
__device__ __forceinline__ float f32_up(double x) {
  const float f = (float)x;
  return ((double)f < x) ? nextafterf(f, INFINITY) : f;
}

// Conservative FP32 box in sweep order: 2D one float4 (lo.x, lo.y, hi.x, hi.y),
// 3D two float4 (lo.xyz, hi.xyz).
template <int D>
__device__ __forceinline__ void store_box32(float4* __restrict__ tb, uint32_t t, const double* lo, const double* hi) {
  if (D == 2) {
    tb[t] = make_float4(f32_down(lo[0]), f32_down(lo[1]), f32_up(hi[0]), f32_up(hi[1]));
  } else {
    tb[2 * (size_t)t] = make_float4(f32_down(lo[0]), f32_down(lo[1]), f32_down(lo[D - 1]), 0.f);
    tb[2 * (size_t)t + 1] = make_float4(f32_up(hi[0]), f32_up(hi[1]), f32_up(hi[D - 1]), 0.f);
  }
}

template <int D>
__device__ __forceinline__ void load_box32(const float4* __restrict__ tb, uint32_t t, float* lo, float* hi) {
  if (D == 2) {
    const float4 b = tb[t];
    lo[0] = b.x;
    lo[1] = b.y;
    hi[0] = b.z;
    hi[1] = b.w;
  } else {
    const float4 a = tb[2 * (size_t)t], b = tb[2 * (size_t)t + 1];
    lo[0] = a.x;
    lo[1] = a.y;
    lo[D - 1] = a.z;
    hi[0] = b.x;
    hi[1] = b.y;
    hi[D - 1] = b.z;
  }
}

// Predictor (UnconstrainedMap::flow) fused with the box computation, the sweep-order
// records and the bounds partials. Thread t handles body order[t].
// Sweep-order narrow data: 2D tnar[t] = (q0.x, q0.y, q1.x, q1.y), tr[t] = r;
// 3D tnar[t] = (q1.x, q1.y, q1.z, r).
template <int D>
__global__ void __launch_bounds__(kBlock) k_predict_aabb(uint32_t n, const uint32_t* __restrict__ order,
                                                         const double* __restrict__ q0, const double* __restrict__ v0,
                                                         const double* __restrict__ m, const double* __restrict__ r,
                                                         double dt, double gx, double gy, double gz, int integrator,
                                                         double* __restrict__ q1p, double* __restrict__ v1p,
                                                         float4* __restrict__ tbox,
                                                         double4* __restrict__ tnar, double* __restrict__ tr,
                                                         const uint32_t* __restrict__ bscene,
                                                         const double* __restrict__ shift,
                                                         double* __restrict__ partials) {
  const uint32_t t = xcd_block() * kBlock + threadIdx.x;
  const bool valid = t < n;
  const double g[3] = {gx, gy, gz};
  double lo[D], hi[D];
  if (valid) {
    const uint32_t i = order[t];
    const double mi = m[i], ri = r[i];
    double qa[D], va[D], q1[D], v1[D];
    if (D == 2) {
      const double2 qq = reinterpret_cast<const double2*>(q0)[i];
      const double2 vv = reinterpret_cast<const double2*>(v0)[i];
      qa[0] = qq.x;
      qa[1] = qq.y;
      va[0] = vv.x;
      va[1] = vv.y;
    } else {
#pragma unroll
      for (int a = 0; a < D; ++a) {
        qa[a] = q0[(size_t)D * i + a];
        va[a] = v0[(size_t)D * i + a];
      }
    }
#pragma unroll
    for (int a = 0; a < D; ++a) {
      flow_component<D>(integrator, mi, g[a], dt, qa[a], va[a], &q1[a], &v1[a]);
      if (D == 2) {
        // swept box of the linear motion q0 -> q1 (ball2d/Ball2DSim.cpp:565-573)
        lo[a] = fmin(q1[a], qa[a]) - ri;
        hi[a] = fmax(q1[a], qa[a]) + ri;
      } else {
        // end-of-step box (rigidbody3d/RigidBody3DSim.cpp:1057-1069, RigidBodySphere.cpp:45-49)
        lo[a] = q1[a] - ri;
        hi[a] = q1[a] + ri;
      }
    }
    if (v1p) {
#pragma unroll
      for (int a = 0; a < D; ++a) v1p[(size_t)D * i + a] = v1[a];
    }
    if (D == 2) {
      reinterpret_cast<double2*>(q1p)[i] = make_double2(q1[0], q1[1]);
      tnar[t] = make_double4(qa[0], qa[1], q1[0], q1[1]);
      tr[t] = ri;
    } else {
#pragma unroll
      for (int a = 0; a < D; ++a) q1p[(size_t)D * i + a] = q1[a];
      tnar[t] = make_double4(q1[0], q1[1], q1[D - 1], ri);
    }
    if (bscene) {
      // ensembles: each scene binned at its own place; the FP64 narrow data stays
      // unshifted, and rounding is monotone, so the shifted FP32 box still contains the
      // shifted FP64 box and rejects no pair of the scene
      const double* sh = shift + 3 * (size_t)bscene[i];
#pragma unroll
      for (int a = 0; a < D; ++a) {
        lo[a] = lo[a] + sh[a];
        hi[a] = hi[a] + sh[a];
      }
    }
    store_box32<D>(tbox, t, lo, hi);
  }
  bounds_partials<D>(valid, lo, hi, partials);
}

// Unconstrained flow of given host arrays (scx_unconstrained_flow).
template <int D>
__global__ void __launch_bounds__(kBlock) k_flow(uint32_t n, const double* __restrict__ q0, const double* __restrict__ v0,
                                                 const double* __restrict__ m, double dt, double gx, double gy, double gz,
                                                 int integrator, double* __restrict__ q1, double* __restrict__ v1) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const double g[3] = {gx, gy, gz};
#pragma unroll
  for (int a = 0; a < D; ++a) {
    const size_t k = (size_t)D * i + a;
    flow_component<D>(integrator, m[i], g[a], dt, q0[k], v0[k], &q1[k], &v1[k]);
  }
}

// Bounds partials of given boxes (scx_potential_overlaps).
template <int D>
__global__ void __launch_bounds__(kBlock) k_bounds_partials(uint32_t n, const double* __restrict__ amin,
                                                            const double* __restrict__ amax,
                                                            double* __restrict__ partials) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  const bool valid = i < n;
  double lo[D], hi[D];
  if (valid)
#pragma unroll
    for (int a = 0; a < D; ++a) {
      lo[a] = amin[(size_t)D * i + a];
      hi[a] = amax[(size_t)D * i + a];
    }
  bounds_partials<D>(valid, lo, hi, partials);
}

// Given boxes into sweep order (order == nullptr: identity). Narrow data = the FP64
// box: 2D tnar[t] = (lo.x, lo.y, hi.x, hi.y); 3D tnar[2t] = lo, tnar[2t+1] = hi.
template <int D>
__global__ void __launch_bounds__(kBlock) k_boxes_to_sweep(uint32_t n, const uint32_t* __restrict__ order,
                                                           const double* __restrict__ amin,
                                                           const double* __restrict__ amax, float4* __restrict__ tbox,
                                                           double4* __restrict__ tnar) {
  const uint32_t t = xcd_block() * kBlock + threadIdx.x;
  if (t >= n) return;
  const uint32_t i = order ? order[t] : t;
  double lo[D], hi[D];
#pragma unroll
  for (int a = 0; a < D; ++a) {
    lo[a] = amin[(size_t)D * i + a];
    hi[a] = amax[(size_t)D * i + a];
  }
  store_box32<D>(tbox, t, lo, hi);
  if (D == 2) {
    tnar[t] = make_double4(lo[0], lo[1], hi[0], hi[1]);
  } else {
    tnar[2 * (size_t)t] = make_double4(lo[0], lo[1], lo[D - 1], 0.0);
    tnar[2 * (size_t)t + 1] = make_double4(hi[0], hi[1], hi[D - 1], 0.0);
  }
}

__global__ void __launch_bounds__(kBlock) k_iota(uint32_t n, uint32_t* __restrict__ order) {
  const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
  if (t < n) order[t] = t;
}

// min_coord = min - 2e-6, max_coord = max + 2e-6, h = max over axes of summed extents / n,
// dims = ceil((max - min) / h) (ball2d/SpatialGridDetector.cpp:76-104); h grows until the
// dense grid fits cell_cap cells.
template <int D>
__global__ void __launch_bounds__(kBlock) k_grid_params(const double* __restrict__ partials, uint32_t nblk, uint32_t n,
                                                        uint32_t cell_cap, GridParams* __restrict__ grid,
                                                        StepCounters* __restrict__ ctr) {
  __shared__ double s_lo[D][kBlock / 64], s_hi[D][kBlock / 64], s_ext[D][kBlock / 64];
  const int wave = threadIdx.x >> 6;
#pragma unroll
  for (int a = 0; a < D; ++a) {
    double lo = INFINITY, hi = -INFINITY, ext = 0.0;
    for (uint32_t b = threadIdx.x; b < nblk; b += kBlock) {
      lo = fmin(lo, partials[(size_t)b * 9 + 3 * a + 0]);
      hi = fmax(hi, partials[(size_t)b * 9 + 3 * a + 1]);
      ext += partials[(size_t)b * 9 + 3 * a + 2];
    }
    lo = wave_reduce_min_f64(lo);
    hi = wave_reduce_max_f64(hi);
    ext = wave_reduce_sum_f64(ext);
    if ((threadIdx.x & 63) == 0) {
      s_lo[a][wave] = lo;
      s_hi[a][wave] = hi;
      s_ext[a][wave] = ext;
    }
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double dmax = -INFINITY, mn[3] = {0, 0, 0}, mx[3] = {0, 0, 0};
  for (int a = 0; a < D; ++a) {
    double lo = s_lo[a][0], hi = s_hi[a][0], ext = s_ext[a][0];
    for (int w = 1; w < kBlock / 64; ++w) {
      lo = fmin(lo, s_lo[a][w]);
      hi = fmax(hi, s_hi[a][w]);
      ext += s_ext[a][w];
    }
    mn[a] = lo - 2.0e-6;
    mx[a] = hi + 2.0e-6;
    dmax = fmax(dmax, ext);
  }
  double h = dmax / (double)n;
  if (!(h > 0.0) || !(h < INFINITY)) h = 1.0;
  uint32_t dims[3] = {1, 1, 1};
  for (int guard = 0; guard < 200; ++guard) {
    double cells = 1.0;
    for (int a = 0; a < D; ++a) {
      const double d = ceil((mx[a] - mn[a]) / h);
      cells *= fmax(1.0, d);
    }
    if (cells <= (double)cell_cap) break;
    h *= fmax(1.01, pow(cells / (double)cell_cap, 1.0 / D));
  }
  for (int a = 0; a < D; ++a) dims[a] = (uint32_t)fmax(1.0, ceil((mx[a] - mn[a]) / h));
  for (int a = 0; a < 3; ++a) grid->min_coord[a] = (a < D) ? mn[a] : 0.0;
  grid->h = h;
  for (int a = 0; a < 3; ++a) grid->dims[a] = dims[a];
  grid->ncells = dims[0] * dims[1] * dims[2];
  ctr->ncells = grid->ncells;
  for (int a = 0; a < 3; ++a) grid->min_f[a] = (a < D && mn[a] > -INFINITY && mn[a] < INFINITY) ? (float)mn[a] : 0.f;
  grid->inv_h_f = (float)(1.0 / h);
}

// Covered cells [lo, hi] of an FP32 box. One monotone map for every box: cells of
// boxes that overlap are never disjoint.
template <int D>
__device__ __forceinline__ void cell_range_f(const float* mn, const float* mx, const GridParams& g, uint32_t* lo,
                                             uint32_t* hi) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (a < D) {
      lo[a] = cell_of_f(mn[a], g.min_f[a], g.inv_h_f, g.dims[a]);
      hi[a] = cell_of_f(mx[a], g.min_f[a], g.inv_h_f, g.dims[a]);
    } else {
      lo[a] = 0;
      hi[a] = 0;
    }
  }
}

__device__ __forceinline__ uint32_t cell_key(const GridParams& g, uint32_t x, uint32_t y, uint32_t z) {
  return x + g.dims[0] * (y + g.dims[1] * z);
}

// Counts over sweep positions: home cells (one per body) and covered cells (one per entry).
template <int D>
__global__ void __launch_bounds__(kBlock) k_cells_count(uint32_t n, const float4* __restrict__ tbox,
                                                        const GridParams* __restrict__ gp, uint32_t* __restrict__ ccnt,
                                                        uint32_t* __restrict__ hcnt, StepCounters* __restrict__ ctr) {
  const uint32_t t = xcd_block() * kBlock + threadIdx.x;
  const GridParams g = *gp;
  uint32_t cnt = 0;
  if (t < n) {
    float bl[3], bh[3];
    load_box32<D>(tbox, t, bl, bh);
    uint32_t lo[3], hi[3];
    cell_range_f<D>(bl, bh, g, lo, hi);
    if (hcnt) atomicAdd(&hcnt[cell_key(g, lo[0], lo[1], lo[2])], 1u);
    if (ccnt)
      for (uint32_t z = lo[2]; z <= hi[2]; ++z)
        for (uint32_t y = lo[1]; y <= hi[1]; ++y)
          for (uint32_t x = lo[0]; x <= hi[0]; ++x) {
            atomicAdd(&ccnt[cell_key(g, x, y, z)], 1u);
            ++cnt;
          }
  }
  if (!ccnt) return;
  cnt = block_reduce_sum_u32(cnt);
  if (threadIdx.x == 0 && cnt) atomicAdd(&ctr->nentries, (unsigned long long)cnt);
}

// Entries in cell order (sweep position | first-cell flags, bits 29..31) and the
// home-cell order of the bodies for the next binning (order_out[k] = body id).
// ccnt == nullptr: home order only.
template <int D>
__global__ void __launch_bounds__(kBlock) k_entries_scatter(uint32_t n, const float4* __restrict__ tbox,
                                                            const GridParams* __restrict__ gp,
                                                            uint32_t* __restrict__ ccnt,
                                                            const uint32_t* __restrict__ cstart,
                                                            uint32_t* __restrict__ hcnt,
                                                            const uint32_t* __restrict__ hstart,
                                                            const uint32_t* __restrict__ order_in,
                                                            uint32_t* __restrict__ order_out, uint32_t* __restrict__ ent,
                                                            uint32_t cap, StepCounters* __restrict__ ctr) {
  const uint32_t t = xcd_block() * kBlock + threadIdx.x;
  if (t >= n) return;
  const GridParams g = *gp;
  float bl[3], bh[3];
  load_box32<D>(tbox, t, bl, bh);
  uint32_t lo[3], hi[3];
  cell_range_f<D>(bl, bh, g, lo, hi);
  const uint32_t home = cell_key(g, lo[0], lo[1], lo[2]);
  const uint32_t k = hstart[home] + atomicSub(&hcnt[home], 1u) - 1u;
  order_out[k] = order_in ? order_in[t] : t;
  if (!ccnt) return;
  for (uint32_t z = lo[2]; z <= hi[2]; ++z)
    for (uint32_t y = lo[1]; y <= hi[1]; ++y)
      for (uint32_t x = lo[0]; x <= hi[0]; ++x) {
        const uint32_t key = cell_key(g, x, y, z);
        const uint32_t pos = cstart[key] + atomicSub(&ccnt[key], 1u) - 1u;
        if (pos >= cap) {
          atomicOr(&ctr->overflow, 1u);
          continue;
        }
        ent[pos] = t | ((uint32_t)(x == lo[0]) << 29) | ((uint32_t)(y == lo[1]) << 30) |
                   ((uint32_t)(z == lo[2]) << 31);
      }
}

// Narrow-phase filters fused into the sweep.
constexpr int kFilterNone = 0;    // boxes only (getPotentialOverlaps)
constexpr int kFilterCCD2 = 1;    // ball2d: CCD on (q0, q1) (Ball2DSim.cpp:577-607)
constexpr int kFilterSphere3 = 2; // rigidbody3d: overlap at q1 (RigidBody3DSim.cpp:793-800)

// ballBallCCDCollisionHappens (scisim/CollisionDetection/CollisionDetectionUtilities.cpp:3-126).
__device__ __forceinline__ bool ccd2(const double* q0a, const double* q1a, double ra, const double* q0b,
                                     const double* q1b, double rb) {
  const double d0x = q0a[0] - q0b[0], d0y = q0a[1] - q0b[1];
  const double dqx = (q1a[0] - q1b[0]) - d0x, dqy = (q1a[1] - q1b[1]) - d0y;
  const double rs = ra + rb;
  const double c0 = (d0x * d0x + d0y * d0y) - rs * rs;
  const double c1 = 2.0 * (d0x * dqx + d0y * dqy);
  const double c2 = dqx * dqx + dqy * dqy;
  if (c2 != 0.0) {
    const double c1c1 = c1 * c1;
    const double fc2c0 = 4 * c2 * c0;
    if (c1c1 < fc2c0) return false;
    const double ds = sqrt(c1c1 - fc2c0);
    const double root1 = (c1 > 0.0) ? (2.0 * c0) / (-c1 - ds) : (-c1 + ds) / (2.0 * c2);
    if (root1 < 0.0) return false;
    const double root0 = (c1 >= 0.0) ? (-c1 - ds) / (2.0 * c2) : (2.0 * c0) / (-c1 + ds);
    return !(root0 > 1.0);
  }
  return c0 <= 0.0;
}

// Sweep-order records read by the neighbour sweep.
struct SweepBodies {
  const uint32_t* order;
  const float4* box;
  const double4* nar;
  const double* r;
  const uint32_t* scene;  // ensembles: scene of each body id (pairs only within a scene), else nullptr
};

// Exact FP64 data of one body: its box as the FP64 test sees it, plus the narrow data.
// The 2D swept box is recomputed from (q0, q1, r) with the predictor's expression, so it
// is the same value the box test of the reference would use.
template <int D, int FILTER>
struct Exact {
  double lo[D], hi[D];
  double q0[2], q1[D], r;
};

template <int D, int FILTER>
__device__ __forceinline__ void load_exact(const SweepBodies& sb, uint32_t t, Exact<D, FILTER>& e) {
  if (FILTER == kFilterCCD2) {
    const double4 a = sb.nar[t];
    e.r = sb.r[t];
    e.q0[0] = a.x;
    e.q0[1] = a.y;
    e.q1[0] = a.z;
    e.q1[1] = a.w;
#pragma unroll
    for (int k = 0; k < D; ++k) {
      e.lo[k] = fmin(e.q1[k], e.q0[k]) - e.r;
      e.hi[k] = fmax(e.q1[k], e.q0[k]) + e.r;
    }
  } else if (FILTER == kFilterSphere3) {
    const double4 a = sb.nar[t];
    e.q1[0] = a.x;
    e.q1[1] = a.y;
    e.q1[D - 1] = a.z;
    e.r = a.w;
#pragma unroll
    for (int k = 0; k < D; ++k) {
      e.lo[k] = e.q1[k] - e.r;
      e.hi[k] = e.q1[k] + e.r;
    }
  } else if (D == 2) {
    const double4 a = sb.nar[t];
    e.lo[0] = a.x;
    e.lo[1] = a.y;
    e.hi[0] = a.z;
    e.hi[1] = a.w;
  } else {
    const double4 a = sb.nar[2 * (size_t)t], b = sb.nar[2 * (size_t)t + 1];
    e.lo[0] = a.x;
    e.lo[1] = a.y;
    e.lo[D - 1] = a.z;
    e.hi[0] = b.x;
    e.hi[1] = b.y;
    e.hi[D - 1] = b.z;
  }
}

// The FP64 box test (AABB::overlaps) and the filter of the active-set builders.
template <int D, int FILTER>
__device__ __forceinline__ bool exact_test(const Exact<D, FILTER>& a, const Exact<D, FILTER>& b) {
  bool sep = false;
#pragma unroll
  for (int k = 0; k < D; ++k) sep = sep || (a.hi[k] < b.lo[k]) || (b.hi[k] < a.lo[k]);
  if (sep) return false;
  if (FILTER == kFilterCCD2) return ccd2(a.q0, a.q1, a.r, b.q0, b.q1, b.r);
  if (FILTER == kFilterSphere3) {
    const double dx = a.q1[0] - b.q1[0], dy = a.q1[1] - b.q1[1], dz = a.q1[D - 1] - b.q1[D - 1];
    const double rs = a.r + b.r;
    return (dx * dx + dy * dy + dz * dz) <= rs * rs;
  }
  return true;
}

// Per-owner sweep over sweep positions. MODE 0: count emitted partners of owner
// i = order[t] into cnt[i]. MODE 1: write them at off[i] (pa = i, pb = j) and sort the
// owner's segment ascending.
template <int D, int MODE, int FILTER>
__global__ void __launch_bounds__(kBlock) k_pairs(uint32_t n, SweepBodies sb, const GridParams* __restrict__ gp,
                                                  const uint32_t* __restrict__ cstart,
                                                  const uint32_t* __restrict__ ent, uint32_t ecap,
                                                  uint32_t* __restrict__ cnt_off, uint32_t* __restrict__ pa,
                                                  uint32_t* __restrict__ pb, uint32_t cap,
                                                  StepCounters* __restrict__ ctr) {
  const uint32_t t = xcd_block() * kBlock + threadIdx.x;
  if (t >= n) return;
  const GridParams g = *gp;
  const uint32_t i = sb.order[t];
  float bl[3], bh[3];
  load_box32<D>(sb.box, t, bl, bh);
  uint32_t lo[3], hi[3];
  cell_range_f<D>(bl, bh, g, lo, hi);
  Exact<D, FILTER> me;
  load_exact<D, FILTER>(sb, t, me);
  const uint32_t my_scene = sb.scene ? sb.scene[i] : 0u;
  uint32_t base = 0, count = 0;
  if (MODE == 1) base = cnt_off[i];
  for (uint32_t z = lo[2]; z <= hi[2]; ++z)
    for (uint32_t y = lo[1]; y <= hi[1]; ++y) {
      // the cells of one row are consecutive keys: their starts are consecutive words
      const uint32_t key0 = cell_key(g, lo[0], y, z);
      // entries beyond the buffer (overflow, the step is redone) are never read
      uint32_t e0 = min(cstart[key0], ecap);
      for (uint32_t x = lo[0]; x <= hi[0]; ++x) {
        const uint32_t e1 = min(cstart[key0 + (x - lo[0]) + 1], ecap);
        for (uint32_t e = e0; e < e1; ++e) {
          const uint32_t v = ent[e];
          // emit only from the cell max(first cell of i, first cell of j), per axis
          const bool fx = (v >> 29) & 1u || x == lo[0];
          const bool fy = (v >> 30) & 1u || y == lo[1];
          const bool fz = (D < 3) || ((v >> 31) & 1u) || z == lo[2];
          if (!(fx && fy && fz)) continue;
          const uint32_t tt = v & kEntIndexMask;
          // conservative FP32 boxes: a rejection here is a rejection in FP64
          float ol[3], oh[3];
          load_box32<D>(sb.box, tt, ol, oh);
          bool sep = false;
#pragma unroll
          for (int a = 0; a < D; ++a) sep = sep || (bh[a] < ol[a]) || (oh[a] < bl[a]);
          if (sep) continue;
          const uint32_t j = sb.order[tt];
          if (j <= i) continue;
          if (sb.scene && sb.scene[j] != my_scene) continue;
          Exact<D, FILTER> other;
          load_exact<D, FILTER>(sb, tt, other);
          if (!exact_test<D, FILTER>(me, other)) continue;
          if (MODE == 1) {
            const uint32_t pos = base + count;
            if (pos < cap) {
              pa[pos] = i;
              pb[pos] = j;
            }
          }
          ++count;
        }
        e0 = e1;
      }
    }
  if (MODE == 0) {
    cnt_off[i] = count;
    return;
  }
  if (base + count > cap) {
    atomicOr(&ctr->overflow, 2u);
    return;
  }
  // insertion sort of the owner's partners (short lists)
  for (uint32_t u = 1; u < count; ++u) {
    const uint32_t vv = pb[base + u];
    uint32_t w = u;
    while (w > 0 && pb[base + w - 1] > vv) {
      pb[base + w] = pb[base + w - 1];
      --w;
    }
    pb[base + w] = vv;
  }
}

// Active-plane mask per body at the predictor position:
// n . (q1 - x) <= r (ball2d/Constraints/BallStaticPlaneConstraint.cpp:10-15,
// rigidbody3d/Constraints/StaticPlaneSphereConstraint.cpp:13-17).
template <int D>
__global__ void __launch_bounds__(kBlock) k_plane_masks(uint32_t n, const double* __restrict__ q1,
                                                        const double* __restrict__ r, uint32_t nplanes,
                                                        const double* __restrict__ px, const double* __restrict__ pn,
                                                        const uint32_t* __restrict__ bscene,
                                                        const uint32_t* __restrict__ scene_poff,
                                                        unsigned long long* __restrict__ mask) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  double x[D];
#pragma unroll
  for (int a = 0; a < D; ++a) x[a] = q1[(size_t)D * i + a];
  const double ri = r[i];
  unsigned long long m = 0;
  // ensembles: the planes of the body's scene, bits scene-local
  uint32_t p0 = 0, p1 = nplanes;
  if (bscene) {
    const uint32_t sc = bscene[i];
    p0 = scene_poff[sc];
    p1 = scene_poff[sc + 1];
  }
  for (uint32_t p = p0; p < p1; ++p) {
    double d = pn[D * p] * (x[0] - px[D * p]);
#pragma unroll
    for (int a = 1; a < D; ++a) d = d + pn[D * p + a] * (x[a] - px[D * p + a]);
    if (d <= ri) m |= 1ull << (p - p0);
  }
  mask[i] = m;
}

// ---------------------------------------------------------------------------
// Host launchers.
// ---------------------------------------------------------------------------
hipError_t launch_predict_aabb(scx_ctx* c, const double* q0, const double* v0, const double* m, double dt,
                               const double* g, int integrator, double* q1p, double* v1p) {
  const uint32_t n = c->n;
  const uint32_t nblk = (n + kBlock - 1) / kBlock;
  if (nblk == 0) return hipSuccess;
  DeviceBuffers& d = c->d;
  const uint32_t* bscene = c->nscenes ? d.bscene : nullptr;
  if (c->dim == 2)
    k_predict_aabb<2><<<nblk, kBlock, 0, c->stream>>>(n, d.order, q0, v0, m, d.r, dt, g[0], g[1], g[2], integrator,
                                                       q1p, v1p, d.tbox, d.tnar, d.tr, bscene, d.scene_shift, d.red);
  else
    k_predict_aabb<3><<<nblk, kBlock, 0, c->stream>>>(n, d.order, q0, v0, m, d.r, dt, g[0], g[1], g[2], integrator,
                                                       q1p, v1p, d.tbox, d.tnar, d.tr, bscene, d.scene_shift, d.red);
  return hipGetLastError();
}

hipError_t launch_flow(scx_ctx* c, uint32_t n, int integrator, const double* m, const double* g, double dt,
                       const double* q0, const double* v0, double* q1, double* v1) {
  const uint32_t nblk = (n + kBlock - 1) / kBlock;
  if (nblk == 0) return hipSuccess;
  if (c->dim == 2)
    k_flow<2><<<nblk, kBlock, 0, c->stream>>>(n, q0, v0, m, dt, g[0], g[1], g[2], integrator, q1, v1);
  else
    k_flow<3><<<nblk, kBlock, 0, c->stream>>>(n, q0, v0, m, dt, g[0], g[1], g[2], integrator, q1, v1);
  return hipGetLastError();
}

hipError_t launch_order_reset(scx_ctx* c, uint32_t n) {
  const uint32_t nblk = (n + kBlock - 1) / kBlock;
  if (nblk == 0) return hipSuccess;
  k_iota<<<nblk, kBlock, 0, c->stream>>>(n, c->d.order);
  return hipGetLastError();
}

__global__ void __launch_bounds__(kBlock) k_slots(uint32_t n, const uint32_t* __restrict__ order,
                                                  uint32_t* __restrict__ slot_of) {
  const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
  if (s < n) slot_of[order[s]] = s;
}

hipError_t launch_slots(scx_ctx* c, uint32_t n) {
  const uint32_t nblk = (n + kBlock - 1) / kBlock;
  c->d.sid = c->d.order;
  c->d.slot_map = c->d.slot_of;
  if (nblk == 0) return hipSuccess;
  k_slots<<<nblk, kBlock, 0, c->stream>>>(n, c->d.order, c->d.slot_of);
  return hipGetLastError();
}

// Home-cell sort (home_only) or full binning of the sweep-order boxes in d.tbox.
// order_in == nullptr: the sweep order is the identity.
template <int D>
static hipError_t bin_sweep_boxes(scx_ctx* c, uint32_t n, bool home_only, const uint32_t* order_in,
                                  uint32_t* order_out) {
  hipStream_t s = c->stream;
  DeviceBuffers& d = c->d;
  const uint32_t nblk = (n + kBlock - 1) / kBlock;
  const uint32_t cap = c->cap_cells;
  uint32_t* ccnt = home_only ? nullptr : d.cells;
  uint32_t* cstart = d.cells + (cap + 1);
  uint32_t* hcnt = d.cells + 2 * (size_t)(cap + 1);
  uint32_t* hstart = d.cells + 3 * (size_t)(cap + 1);
  hipError_t e = hipMemsetAsync(hcnt, 0, sizeof(uint32_t) * (cap + 1), s);
  if (e == hipSuccess && ccnt) e = hipMemsetAsync(ccnt, 0, sizeof(uint32_t) * (cap + 1), s);
  if (e != hipSuccess) return e;
  k_cells_count<D><<<nblk, kBlock, 0, s>>>(n, d.tbox, d.grid, ccnt, hcnt, d.ctr);
  if (ccnt) e = launch_exclusive_scan(ccnt, cstart, cap, d.scan_tmp, s);
  if (e == hipSuccess) e = launch_exclusive_scan(hcnt, hstart, cap, d.scan_tmp, s);
  if (e != hipSuccess) return e;
  k_entries_scatter<D><<<nblk, kBlock, 0, s>>>(n, d.tbox, d.grid, ccnt, cstart, hcnt, hstart, order_in, order_out,
                                                d.ent, c->cap_entries, d.ctr);
  return hipGetLastError();
}

template <int D>
static hipError_t grid_build_dim(scx_ctx* c, uint32_t n, bool given_boxes) {
  hipStream_t s = c->stream;
  DeviceBuffers& d = c->d;
  const uint32_t nblk = (n + kBlock - 1) / kBlock;
  hipError_t e = hipSuccess;
  if (given_boxes) {
    // uploaded boxes in caller order: sort them by home cell first, so that the
    // binning and the sweep below run in a spatially coherent order
    k_bounds_partials<D><<<nblk, kBlock, 0, s>>>(n, d.aabb_min, d.aabb_max, d.red);
    k_grid_params<D><<<1, kBlock, 0, s>>>(d.red, nblk, n, c->cap_cells, d.grid, d.ctr);
    k_boxes_to_sweep<D><<<nblk, kBlock, 0, s>>>(n, nullptr, d.aabb_min, d.aabb_max, d.tbox, d.tnar);
    e = bin_sweep_boxes<D>(c, n, true, nullptr, d.order);
    if (e != hipSuccess) return e;
    k_boxes_to_sweep<D><<<nblk, kBlock, 0, s>>>(n, d.order, d.aabb_min, d.aabb_max, d.tbox, d.tnar);
  } else {
    k_grid_params<D><<<1, kBlock, 0, s>>>(d.red, nblk, n, c->cap_cells, d.grid, d.ctr);
  }
  e = bin_sweep_boxes<D>(c, n, false, d.order, d.order_next);
  if (e != hipSuccess) return e;
  return hipGetLastError();
}

// Bounds -> grid parameters -> cell entries of the sweep-order boxes, and the
// home-cell order of this binning in d.order_next.
hipError_t launch_grid_build(scx_ctx* c, uint32_t n, bool given_boxes) {
  if (n == 0) return hipSuccess;
  return c->dim == 2 ? grid_build_dim<2>(c, n, given_boxes) : grid_build_dim<3>(c, n, given_boxes);
}

template <int D, int FILTER>
static void pairs_pass(scx_ctx* c, uint32_t n, int mode, uint32_t nblk, uint32_t* out_a, uint32_t* out_b,
                       uint32_t cap) {
  DeviceBuffers& d = c->d;
  const SweepBodies sb{d.order, d.tbox, d.tnar, d.tr, (FILTER != kFilterNone && c->nscenes) ? d.bscene : nullptr};
  const uint32_t* cstart = d.cells + (c->cap_cells + 1);
  if (mode == 0)
    k_pairs<D, 0, FILTER><<<nblk, kBlock, 0, c->stream>>>(n, sb, d.grid, cstart, d.ent, c->cap_entries, d.pcount,
                                                          out_a, out_b, cap, d.ctr);
  else
    k_pairs<D, 1, FILTER><<<nblk, kBlock, 0, c->stream>>>(n, sb, d.grid, cstart, d.ent, c->cap_entries, d.pcount,
                                                          out_a, out_b, cap, d.ctr);
}

// mode 0: counts into pcount; mode 1: pairs at the scanned offsets in pcount.
// filter: 0 boxes only, 1 ball2d CCD, 2 sphere overlap at q1.
hipError_t launch_pairs(scx_ctx* c, uint32_t n, int mode, int filter, uint32_t* out_a, uint32_t* out_b,
                        uint32_t cap) {
  const uint32_t nblk = (n + kBlock - 1) / kBlock;
  if (nblk == 0) return hipSuccess;
  if (c->dim == 2) {
    if (filter == kFilterNone)
      pairs_pass<2, kFilterNone>(c, n, mode, nblk, out_a, out_b, cap);
    else
      pairs_pass<2, kFilterCCD2>(c, n, mode, nblk, out_a, out_b, cap);
  } else {
    if (filter == kFilterNone)
      pairs_pass<3, kFilterNone>(c, n, mode, nblk, out_a, out_b, cap);
    else
      pairs_pass<3, kFilterSphere3>(c, n, mode, nblk, out_a, out_b, cap);
  }
  return hipGetLastError();
}

hipError_t launch_plane_masks(scx_ctx* c, uint32_t n) {
  const uint32_t nblk = (n + kBlock - 1) / kBlock;
  if (nblk == 0) return hipSuccess;
  const uint32_t* bscene = c->nscenes ? c->d.bscene : nullptr;
  if (c->dim == 2)
    k_plane_masks<2><<<nblk, kBlock, 0, c->stream>>>(n, c->d.q1p, c->d.r, c->nplanes, c->d.px, c->d.pn, bscene,
                                                     c->d.scene_poff, c->d.pmask);
  else
    k_plane_masks<3><<<nblk, kBlock, 0, c->stream>>>(n, c->d.q1p, c->d.r, c->nplanes, c->d.px, c->d.pn, bscene,
                                                     c->d.scene_poff, c->d.pmask);
  return hipGetLastError();
}

}  // namespace scx
