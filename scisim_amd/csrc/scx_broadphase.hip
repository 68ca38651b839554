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
// Binning: the reference's cell size and origin (initializeSpatialGrid, :76-104),
// coarsened only when the dense grid would exceed its cell budget, and its
// rasterisation of the 1e-6-inflated box into every covered cell (:42-74). Bodies are
// counting-sorted by home cell (the cell of their lower corner) and their boxes and
// narrow-phase data copied into that order, so the sweep, which runs over bodies in
// home-cell order, reads its neighbours from nearby memory. Each covered cell holds
// 4-byte entries: the sorted index of the body plus three "first cell along x/y/z"
// flags. A pair is emitted only from the covered cell equal to the component-wise max
// of the two boxes' first cells, which both boxes cover when they overlap, so each
// pair is found exactly once without a global sort; each owner sorts its short list.
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

// Predictor (UnconstrainedMap::flow) fused with the box computation and the bounds partials.
template <int D>
__global__ void __launch_bounds__(kBlock) k_predict_aabb(uint32_t n, const double* __restrict__ q0,
                                                         const double* __restrict__ v0, const double* __restrict__ m,
                                                         const double* __restrict__ r, double dt, double gx, double gy,
                                                         double gz, int integrator, double* __restrict__ q1p,
                                                         double* __restrict__ amin, double* __restrict__ amax,
                                                         double* __restrict__ partials) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  const bool valid = i < n;
  const double g[3] = {gx, gy, gz};
  double lo[D], hi[D];
  if (valid) {
    const double mi = m[i], ri = r[i];
#pragma unroll
    for (int a = 0; a < D; ++a) {
      const size_t k = (size_t)D * i + a;
      double q1, v1;
      flow_component<D>(integrator, mi, g[a], dt, q0[k], v0[k], &q1, &v1);
      q1p[k] = q1;
      if (D == 2) {
        // swept box of the linear motion q0 -> q1 (ball2d/Ball2DSim.cpp:565-573)
        lo[a] = fmin(q1, q0[k]) - ri;
        hi[a] = fmax(q1, q0[k]) + ri;
      } else {
        // end-of-step box (rigidbody3d/RigidBody3DSim.cpp:1057-1069, RigidBodySphere.cpp:45-49)
        lo[a] = q1 - ri;
        hi[a] = q1 + ri;
      }
      amin[k] = lo[a];
      amax[k] = hi[a];
    }
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

// min_coord = min - 2e-6, max_coord = max + 2e-6, h = max over axes of summed extents / n,
// dims = ceil((max - min) / h) (ball2d/SpatialGridDetector.cpp:76-104); h grows until the
// dense grid fits cell_cap cells.
template <int D>
__global__ void __launch_bounds__(kBlock) k_grid_params(const double* __restrict__ partials, uint32_t nblk, uint32_t n,
                                                        uint32_t cell_cap, GridParams* __restrict__ grid) {
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
}

// Covered cells [lo, hi] of a box inflated by 1e-6 (rasterizeAABBs, :42-74), clamped to the grid.
template <int D>
__device__ __forceinline__ void cell_range(const double* mn, const double* mx, const GridParams& g, uint32_t* lo,
                                           uint32_t* hi) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (a < D) {
      lo[a] = cell_of(mn[a] - 1.0e-6, g.min_coord[a], g.h, g.dims[a]);
      hi[a] = cell_of(mx[a] + 1.0e-6, g.min_coord[a], g.h, g.dims[a]);
    } else {
      lo[a] = 0;
      hi[a] = 0;
    }
  }
}

__device__ __forceinline__ uint32_t cell_key(const GridParams& g, uint32_t x, uint32_t y, uint32_t z) {
  return x + g.dims[0] * (y + g.dims[1] * z);
}

// Counts: home cells (one per body) and covered cells (one per entry).
template <int D>
__global__ void __launch_bounds__(kBlock) k_cells_count(uint32_t n, const double* __restrict__ amin,
                                                        const double* __restrict__ amax,
                                                        const GridParams* __restrict__ gp, uint32_t* __restrict__ ccnt,
                                                        uint32_t* __restrict__ hcnt, StepCounters* __restrict__ ctr) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  const GridParams g = *gp;
  uint32_t cnt = 0;
  if (i < n) {
    uint32_t lo[3], hi[3];
    cell_range<D>(amin + (size_t)D * i, amax + (size_t)D * i, g, lo, hi);
    atomicAdd(&hcnt[cell_key(g, lo[0], lo[1], lo[2])], 1u);
    for (uint32_t z = lo[2]; z <= hi[2]; ++z)
      for (uint32_t y = lo[1]; y <= hi[1]; ++y)
        for (uint32_t x = lo[0]; x <= hi[0]; ++x) {
          atomicAdd(&ccnt[cell_key(g, x, y, z)], 1u);
          ++cnt;
        }
  }
  cnt = block_reduce_sum_u32(cnt);
  if (threadIdx.x == 0 && cnt) atomicAdd(&ctr->nentries, (unsigned long long)cnt);
}

// Bodies into home-cell order: sorted id and copies of the boxes and narrow-phase data.
template <int D>
__global__ void __launch_bounds__(kBlock) k_home_scatter(uint32_t n, const double* __restrict__ amin,
                                                         const double* __restrict__ amax,
                                                         const GridParams* __restrict__ gp,
                                                         uint32_t* __restrict__ hcnt, const uint32_t* __restrict__ hstart,
                                                         uint32_t* __restrict__ sid, double* __restrict__ smin,
                                                         double* __restrict__ smax, const double* __restrict__ q0,
                                                         const double* __restrict__ q1, const double* __restrict__ r,
                                                         double* __restrict__ sq0, double* __restrict__ sq1,
                                                         double* __restrict__ sr) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const GridParams g = *gp;
  double mn[D], mx[D];
#pragma unroll
  for (int a = 0; a < D; ++a) {
    mn[a] = amin[(size_t)D * i + a];
    mx[a] = amax[(size_t)D * i + a];
  }
  uint32_t lo[3], hi[3];
  cell_range<D>(mn, mx, g, lo, hi);
  const uint32_t home = cell_key(g, lo[0], lo[1], lo[2]);
  const uint32_t k = hstart[home] + atomicSub(&hcnt[home], 1u) - 1u;
  sid[k] = i;
#pragma unroll
  for (int a = 0; a < D; ++a) {
    smin[(size_t)D * k + a] = mn[a];
    smax[(size_t)D * k + a] = mx[a];
  }
  if (q1) {
#pragma unroll
    for (int a = 0; a < D; ++a) {
      if (D == 2) sq0[(size_t)D * k + a] = q0[(size_t)D * i + a];
      sq1[(size_t)D * k + a] = q1[(size_t)D * i + a];
    }
    sr[k] = r[i];
  }
}

// Entries in cell order: sorted body index | first-cell flags (bits 29..31).
template <int D>
__global__ void __launch_bounds__(kBlock) k_entries_scatter(uint32_t n, const double* __restrict__ smin,
                                                            const double* __restrict__ smax,
                                                            const GridParams* __restrict__ gp,
                                                            uint32_t* __restrict__ ccnt,
                                                            const uint32_t* __restrict__ cstart,
                                                            uint32_t* __restrict__ ent, uint32_t cap,
                                                            StepCounters* __restrict__ ctr) {
  const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= n) return;
  const GridParams g = *gp;
  uint32_t lo[3], hi[3];
  cell_range<D>(smin + (size_t)D * k, smax + (size_t)D * k, g, lo, hi);
  for (uint32_t z = lo[2]; z <= hi[2]; ++z)
    for (uint32_t y = lo[1]; y <= hi[1]; ++y)
      for (uint32_t x = lo[0]; x <= hi[0]; ++x) {
        const uint32_t key = cell_key(g, x, y, z);
        const uint32_t pos = cstart[key] + atomicSub(&ccnt[key], 1u) - 1u;
        if (pos >= cap) {
          atomicOr(&ctr->overflow, 1u);
          continue;
        }
        ent[pos] = k | ((uint32_t)(x == lo[0]) << 29) | ((uint32_t)(y == lo[1]) << 30) |
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

struct SortedBodies {
  const uint32_t* sid;
  const double* smin;
  const double* smax;
  const double* sq0;
  const double* sq1;
  const double* sr;
};

template <int D, int FILTER>
__device__ __forceinline__ bool narrow(const SortedBodies& sb, uint32_t k, uint32_t kk) {
  if (FILTER == kFilterCCD2) {
    return ccd2(sb.sq0 + 2 * (size_t)k, sb.sq1 + 2 * (size_t)k, sb.sr[k], sb.sq0 + 2 * (size_t)kk,
                sb.sq1 + 2 * (size_t)kk, sb.sr[kk]);
  } else if (FILTER == kFilterSphere3) {
    const double* xi = sb.sq1 + 3 * (size_t)k;
    const double* xj = sb.sq1 + 3 * (size_t)kk;
    const double dx = xi[0] - xj[0], dy = xi[1] - xj[1], dz = xi[2] - xj[2];
    const double rs = sb.sr[k] + sb.sr[kk];
    return (dx * dx + dy * dy + dz * dz) <= rs * rs;
  }
  return true;
}

// Per-owner sweep over bodies in home-cell order. MODE 0: count emitted partners of
// owner i into cnt[i]. MODE 1: write them at off[i] (pa = i, pb = j) and sort the
// owner's segment ascending.
template <int D, int MODE, int FILTER>
__global__ void __launch_bounds__(kBlock) k_pairs(uint32_t n, SortedBodies sb, const GridParams* __restrict__ gp,
                                                  const uint32_t* __restrict__ cstart,
                                                  const uint32_t* __restrict__ ent, uint32_t ecap,
                                                  uint32_t* __restrict__ cnt_off, uint32_t* __restrict__ pa,
                                                  uint32_t* __restrict__ pb, uint32_t cap,
                                                  StepCounters* __restrict__ ctr) {
  const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= n) return;
  const GridParams g = *gp;
  const uint32_t i = sb.sid[k];
  double mi[D], xi[D];
#pragma unroll
  for (int a = 0; a < D; ++a) {
    mi[a] = sb.smin[(size_t)D * k + a];
    xi[a] = sb.smax[(size_t)D * k + a];
  }
  uint32_t lo[3], hi[3];
  cell_range<D>(mi, xi, g, lo, hi);
  uint32_t base = 0, count = 0;
  if (MODE == 1) base = cnt_off[i];
  for (uint32_t z = lo[2]; z <= hi[2]; ++z)
    for (uint32_t y = lo[1]; y <= hi[1]; ++y)
      for (uint32_t x = lo[0]; x <= hi[0]; ++x) {
        const uint32_t key = cell_key(g, x, y, z);
        // entries beyond the buffer (overflow, the step is redone) are never read
        const uint32_t e0 = min(cstart[key], ecap), e1 = min(cstart[key + 1], ecap);
        for (uint32_t e = e0; e < e1; ++e) {
          const uint32_t v = ent[e];
          // emit only from the cell max(first cell of i, first cell of j), per axis
          const bool fx = (v >> 29) & 1u || x == lo[0];
          const bool fy = (v >> 30) & 1u || y == lo[1];
          const bool fz = (D < 3) || ((v >> 31) & 1u) || z == lo[2];
          if (!(fx && fy && fz)) continue;
          const uint32_t kk = v & kEntIndexMask;
          const uint32_t j = sb.sid[kk];
          if (j <= i) continue;
          bool sep = false;
#pragma unroll
          for (int a = 0; a < D; ++a) {
            const double mj = sb.smin[(size_t)D * kk + a], xj = sb.smax[(size_t)D * kk + a];
            sep = sep || (xi[a] < mj) || (xj < mi[a]);
          }
          if (sep) continue;
          if (!narrow<D, FILTER>(sb, k, kk)) continue;
          if (MODE == 1) {
            const uint32_t pos = base + count;
            if (pos < cap) {
              pa[pos] = i;
              pb[pos] = j;
            }
          }
          ++count;
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
                                                        unsigned long long* __restrict__ mask) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  double x[D];
#pragma unroll
  for (int a = 0; a < D; ++a) x[a] = q1[(size_t)D * i + a];
  const double ri = r[i];
  unsigned long long m = 0;
  for (uint32_t p = 0; p < nplanes; ++p) {
    double d = pn[D * p] * (x[0] - px[D * p]);
#pragma unroll
    for (int a = 1; a < D; ++a) d = d + pn[D * p + a] * (x[a] - px[D * p + a]);
    if (d <= ri) m |= 1ull << p;
  }
  mask[i] = m;
}

// ---------------------------------------------------------------------------
// Host launchers.
// ---------------------------------------------------------------------------
hipError_t launch_predict_aabb(scx_ctx* c, const double* q0, const double* v0, const double* m, double dt,
                               const double* g, int integrator, double* q1p) {
  const uint32_t n = c->n;
  const uint32_t nblk = (n + kBlock - 1) / kBlock;
  if (nblk == 0) return hipSuccess;
  if (c->dim == 2)
    k_predict_aabb<2><<<nblk, kBlock, 0, c->stream>>>(n, q0, v0, m, c->d.r, dt, g[0], g[1], g[2], integrator, q1p,
                                                       c->d.aabb_min, c->d.aabb_max, c->d.red);
  else
    k_predict_aabb<3><<<nblk, kBlock, 0, c->stream>>>(n, q0, v0, m, c->d.r, dt, g[0], g[1], g[2], integrator, q1p,
                                                       c->d.aabb_min, c->d.aabb_max, c->d.red);
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

template <int D>
static hipError_t grid_build_dim(scx_ctx* c, uint32_t n, bool given_boxes, bool narrow_data) {
  hipStream_t s = c->stream;
  DeviceBuffers& d = c->d;
  const uint32_t nblk = (n + kBlock - 1) / kBlock;
  const uint32_t cap = c->cap_cells;
  uint32_t* ccnt = d.cells;
  uint32_t* cstart = d.cells + (cap + 1);
  uint32_t* hcnt = d.cells + 2 * (size_t)(cap + 1);
  uint32_t* hstart = d.cells + 3 * (size_t)(cap + 1);
  hipError_t e = hipMemsetAsync(ccnt, 0, sizeof(uint32_t) * (cap + 1), s);
  if (e == hipSuccess) e = hipMemsetAsync(hcnt, 0, sizeof(uint32_t) * (cap + 1), s);
  if (e != hipSuccess) return e;
  if (given_boxes) k_bounds_partials<D><<<nblk, kBlock, 0, s>>>(n, d.aabb_min, d.aabb_max, d.red);
  k_grid_params<D><<<1, kBlock, 0, s>>>(d.red, nblk, n, cap, d.grid);
  k_cells_count<D><<<nblk, kBlock, 0, s>>>(n, d.aabb_min, d.aabb_max, d.grid, ccnt, hcnt, d.ctr);
  e = launch_exclusive_scan(ccnt, cstart, cap, d.scan_tmp, s);
  if (e == hipSuccess) e = launch_exclusive_scan(hcnt, hstart, cap, d.scan_tmp, s);
  if (e != hipSuccess) return e;
  k_home_scatter<D><<<nblk, kBlock, 0, s>>>(n, d.aabb_min, d.aabb_max, d.grid, hcnt, hstart, d.sid, d.smin, d.smax,
                                             d.q0, narrow_data ? d.q1p : nullptr, d.r, d.sq0, d.sq1, d.sr);
  k_entries_scatter<D><<<nblk, kBlock, 0, s>>>(n, d.smin, d.smax, d.grid, ccnt, cstart, d.ent, c->cap_entries, d.ctr);
  return hipGetLastError();
}

// Bounds -> grid parameters -> home-cell sort of the bodies -> cell entries.
// given_boxes: the boxes were uploaded (bounds partials not yet computed);
// narrow_data: copy q0/q1/r into home order for the fused narrow phase.
hipError_t launch_grid_build(scx_ctx* c, uint32_t n, bool given_boxes, bool narrow_data) {
  if (n == 0) return hipSuccess;
  return c->dim == 2 ? grid_build_dim<2>(c, n, given_boxes, narrow_data)
                     : grid_build_dim<3>(c, n, given_boxes, narrow_data);
}

template <int D, int FILTER>
static void pairs_pass(scx_ctx* c, uint32_t n, int mode, uint32_t nblk, uint32_t* out_a, uint32_t* out_b,
                       uint32_t cap) {
  DeviceBuffers& d = c->d;
  const SortedBodies sb{d.sid, d.smin, d.smax, d.sq0, d.sq1, d.sr};
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
  if (c->dim == 2)
    k_plane_masks<2><<<nblk, kBlock, 0, c->stream>>>(n, c->d.q1p, c->d.r, c->nplanes, c->d.px, c->d.pn, c->d.pmask);
  else
    k_plane_masks<3><<<nblk, kBlock, 0, c->stream>>>(n, c->d.q1p, c->d.r, c->nplanes, c->d.px, c->d.pn, c->d.pmask);
  return hipGetLastError();
}

}  // namespace scx
