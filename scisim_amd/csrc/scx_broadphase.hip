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
// Binning: the reference's cell size and origin (initializeSpatialGrid, :76-104) and
// its rasterisation of the 1e-6-inflated box into every covered cell (:42-74), but
// cells are hashed into a table sized to the entry count (no dense row-major key,
// which overflows for large scenes) and binned by a counting sort. A pair is emitted
// only from the covered cell equal to the component-wise max of the two boxes' first
// cells, which both boxes cover when they overlap, so each pair is found exactly once
// without a global sort; each owner sorts its own short j-list.
//
// No FMA contraction in this file: the predictor, box bounds and narrow-phase tests
// must round exactly like the FP64 CPU restatement so that pair lists are bit-exact.
#pragma clang fp contract(off)

#include "scx_device.h"
#include "scx_flow.h"
#include "scx_internal.h"

namespace scx {

// Predictor (UnconstrainedMap::flow) fused with the box computation.
template <int D>
__global__ void __launch_bounds__(kBlock) k_predict_aabb(uint32_t n, const double* __restrict__ q0,
                                                         const double* __restrict__ v0, const double* __restrict__ m,
                                                         const double* __restrict__ r, double dt, double gx, double gy,
                                                         double gz, int integrator, double* __restrict__ q1p,
                                                         double* __restrict__ amin, double* __restrict__ amax) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const double g[3] = {gx, gy, gz};
  const double mi = m[i], ri = r[i];
#pragma unroll
  for (int a = 0; a < D; ++a) {
    const size_t k = (size_t)D * i + a;
    double q1, v1;
    flow_component<D>(integrator, mi, g[a], dt, q0[k], v0[k], &q1, &v1);
    q1p[k] = q1;
    if (D == 2) {
      // swept box of the linear motion q0 -> q1 (ball2d/Ball2DSim.cpp:565-573)
      amin[k] = fmin(q1, q0[k]) - ri;
      amax[k] = fmax(q1, q0[k]) + ri;
    } else {
      // end-of-step box (rigidbody3d/RigidBody3DSim.cpp:1057-1069, RigidBodySphere.cpp:45-49)
      amin[k] = q1 - ri;
      amax[k] = q1 + ri;
    }
  }
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

// Per-block partials of the grid bounds: min of mins, max of maxs, sum of extents.
template <int D>
__global__ void __launch_bounds__(kBlock) k_bounds_partials(uint32_t n, const double* __restrict__ amin,
                                                            const double* __restrict__ amax,
                                                            double* __restrict__ partials) {
  __shared__ double s[3 * 3][kBlock / 64];
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  const int wave = threadIdx.x >> 6;
#pragma unroll
  for (int a = 0; a < D; ++a) {
    double lo = INFINITY, hi = -INFINITY, ext = 0.0;
    if (i < n) {
      lo = amin[(size_t)D * i + a];
      hi = amax[(size_t)D * i + a];
      ext = hi - lo;
    }
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
    const int a = threadIdx.x / 3, kind = threadIdx.x % 3;
    double v = s[threadIdx.x][0];
    for (int w = 1; w < kBlock / 64; ++w) {
      const double o = s[threadIdx.x][w];
      v = (kind == 0) ? fmin(v, o) : (kind == 1) ? fmax(v, o) : v + o;
    }
    partials[(size_t)blockIdx.x * 9 + 3 * a + kind] = v;
  }
}

// min_coord = min - 2e-6, h = max over axes of summed extents / n
// (ball2d/SpatialGridDetector.cpp:76-104).
template <int D>
__global__ void k_grid_params(const double* __restrict__ partials, uint32_t nblk, uint32_t n, uint32_t table_mask,
                              GridParams* __restrict__ grid) {
  if (threadIdx.x != 0) return;
  double dmax = -INFINITY;
  for (int a = 0; a < D; ++a) {
    double lo = INFINITY, ext = 0.0;
    for (uint32_t b = 0; b < nblk; ++b) {
      lo = fmin(lo, partials[(size_t)b * 9 + 3 * a + 0]);
      ext += partials[(size_t)b * 9 + 3 * a + 2];
    }
    grid->min_coord[a] = lo - 2.0e-6;
    dmax = fmax(dmax, ext);
  }
  for (int a = D; a < 3; ++a) grid->min_coord[a] = 0.0;
  grid->h = dmax / (double)n;
  grid->table_mask = table_mask;
}

template <int D>
__device__ __forceinline__ void cell_range(const double* mn, const double* mx, const GridParams& g, uint32_t* lo,
                                           uint32_t* hi) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (a < D) {
      lo[a] = cell_of(mn[a] - 1.0e-6, g.min_coord[a], g.h);
      hi[a] = cell_of(mx[a] + 1.0e-6, g.min_coord[a], g.h);
    } else {
      lo[a] = 0;
      hi[a] = 0;
    }
  }
}

template <int D>
__global__ void __launch_bounds__(kBlock) k_grid_count(uint32_t n, const double* __restrict__ amin,
                                                       const double* __restrict__ amax,
                                                       const GridParams* __restrict__ gp, uint32_t* __restrict__ bucket,
                                                       StepCounters* __restrict__ ctr) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  const GridParams g = *gp;
  uint32_t cnt = 0;
  if (i < n) {
    uint32_t lo[3], hi[3];
    cell_range<D>(amin + (size_t)D * i, amax + (size_t)D * i, g, lo, hi);
    for (uint32_t z = lo[2]; z <= hi[2]; ++z)
      for (uint32_t y = lo[1]; y <= hi[1]; ++y)
        for (uint32_t x = lo[0]; x <= hi[0]; ++x) {
          atomicAdd(&bucket[hash_cell(pack_cell(x, y, z)) & g.table_mask], 1u);
          ++cnt;
        }
  }
  cnt = block_reduce_sum_u32(cnt);
  if (threadIdx.x == 0 && cnt) atomicAdd(&ctr->nentries, (unsigned long long)cnt);
}

template <int D>
__global__ void __launch_bounds__(kBlock) k_grid_scatter(uint32_t n, const double* __restrict__ amin,
                                                         const double* __restrict__ amax,
                                                         const GridParams* __restrict__ gp,
                                                         uint32_t* __restrict__ fill, const uint32_t* __restrict__ start,
                                                         unsigned long long* __restrict__ ekey,
                                                         uint32_t* __restrict__ eid, uint32_t cap,
                                                         StepCounters* __restrict__ ctr) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const GridParams g = *gp;
  uint32_t lo[3], hi[3];
  cell_range<D>(amin + (size_t)D * i, amax + (size_t)D * i, g, lo, hi);
  for (uint32_t z = lo[2]; z <= hi[2]; ++z)
    for (uint32_t y = lo[1]; y <= hi[1]; ++y)
      for (uint32_t x = lo[0]; x <= hi[0]; ++x) {
        const unsigned long long key = pack_cell(x, y, z);
        const uint32_t b = hash_cell(key) & g.table_mask;
        const uint32_t pos = start[b] + atomicSub(&fill[b], 1u) - 1u;
        if (pos >= cap) {
          atomicOr(&ctr->overflow, 1u);
          continue;
        }
        const unsigned long long flags = (unsigned long long)(x == lo[0]) << 60 |
                                         (unsigned long long)(y == lo[1]) << 61 |
                                         (unsigned long long)(z == lo[2]) << 62;
        ekey[pos] = key | flags;
        eid[pos] = i;
      }
}

// Narrow-phase filters fused into the sweep.
constexpr int kFilterNone = 0;    // boxes only (getPotentialOverlaps)
constexpr int kFilterCCD2 = 1;    // ball2d: CCD on (q0, q1) (Ball2DSim.cpp:577-607)
constexpr int kFilterSphere3 = 2; // rigidbody3d: overlap at q1 (RigidBody3DSim.cpp:793-800)

struct NarrowArgs {
  const double* q0;
  const double* q1;
  const double* r;
};

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

template <int D, int FILTER>
__device__ __forceinline__ bool narrow(const NarrowArgs& na, uint32_t i, uint32_t j) {
  if (FILTER == kFilterCCD2) {
    return ccd2(na.q0 + 2 * (size_t)i, na.q1 + 2 * (size_t)i, na.r[i], na.q0 + 2 * (size_t)j, na.q1 + 2 * (size_t)j,
                na.r[j]);
  } else if (FILTER == kFilterSphere3) {
    const double* xi = na.q1 + 3 * (size_t)i;
    const double* xj = na.q1 + 3 * (size_t)j;
    const double dx = xi[0] - xj[0], dy = xi[1] - xj[1], dz = xi[2] - xj[2];
    const double rs = na.r[i] + na.r[j];
    return (dx * dx + dy * dy + dz * dz) <= rs * rs;
  }
  return true;
}

// Per-owner sweep. MODE 0: count emitted partners of i into cnt[i]. MODE 1: write them
// at off[i] (pair arrays pa = i, pb = j) and sort the owner's segment ascending.
template <int D, int MODE, int FILTER>
__global__ void __launch_bounds__(kBlock) k_pairs(uint32_t n, const double* __restrict__ amin,
                                                  const double* __restrict__ amax, const GridParams* __restrict__ gp,
                                                  const uint32_t* __restrict__ start,
                                                  const unsigned long long* __restrict__ ekey,
                                                  const uint32_t* __restrict__ eid, NarrowArgs na,
                                                  uint32_t* __restrict__ cnt_off, uint32_t* __restrict__ pa,
                                                  uint32_t* __restrict__ pb, uint32_t cap, uint32_t ecap,
                                                  StepCounters* __restrict__ ctr) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const GridParams g = *gp;
  double mi[D], xi[D];
#pragma unroll
  for (int a = 0; a < D; ++a) {
    mi[a] = amin[(size_t)D * i + a];
    xi[a] = amax[(size_t)D * i + a];
  }
  uint32_t lo[3], hi[3];
  cell_range<D>(mi, xi, g, lo, hi);
  uint32_t base = 0, count = 0;
  if (MODE == 1) base = cnt_off[i];
  for (uint32_t z = lo[2]; z <= hi[2]; ++z)
    for (uint32_t y = lo[1]; y <= hi[1]; ++y)
      for (uint32_t x = lo[0]; x <= hi[0]; ++x) {
        const unsigned long long key = pack_cell(x, y, z);
        const uint32_t b = hash_cell(key) & g.table_mask;
        // entries beyond the buffer (overflow, the step is redone) are never read
        const uint32_t e0 = start[b], e1 = min(start[b + 1], ecap);
        for (uint32_t e = e0; e < e1; ++e) {
          const unsigned long long ek = ekey[e];
          if ((ek & kKeyMask) != key) continue;
          const uint32_t j = eid[e];
          if (j <= i) continue;
          // emit only from the cell max(first cell of i, first cell of j), per axis
          const bool fx = ((ek >> 60) & 1ull) || x == lo[0];
          const bool fy = (D < 2) || ((ek >> 61) & 1ull) || y == lo[1];
          const bool fz = (D < 3) || ((ek >> 62) & 1ull) || z == lo[2];
          if (!(fx && fy && fz)) continue;
          bool sep = false;
#pragma unroll
          for (int a = 0; a < D; ++a) {
            const double mj = amin[(size_t)D * j + a], xj = amax[(size_t)D * j + a];
            sep = sep || (xi[a] < mj) || (xj < mi[a]);
          }
          if (sep) continue;
          if (!narrow<D, FILTER>(na, i, j)) continue;
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
    const uint32_t v = pb[base + u];
    uint32_t w = u;
    while (w > 0 && pb[base + w - 1] > v) {
      pb[base + w] = pb[base + w - 1];
      --w;
    }
    pb[base + w] = v;
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
                                                       c->d.aabb_min, c->d.aabb_max);
  else
    k_predict_aabb<3><<<nblk, kBlock, 0, c->stream>>>(n, q0, v0, m, c->d.r, dt, g[0], g[1], g[2], integrator, q1p,
                                                       c->d.aabb_min, c->d.aabb_max);
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

// Bounds -> grid parameters -> counting sort of the (cell, body) entries by hash bucket.
hipError_t launch_grid_build(scx_ctx* c, uint32_t n) {
  hipStream_t s = c->stream;
  const uint32_t nblk = (n + kBlock - 1) / kBlock;
  const uint32_t T = c->table_size;
  uint32_t* fill = c->d.bucket;             // T counts
  uint32_t* start = c->d.bucket + T;        // T + 1 starts
  uint32_t* tmp = c->d.scan_tmp;
  hipError_t e = hipMemsetAsync(fill, 0, sizeof(uint32_t) * T, s);
  if (e != hipSuccess) return e;
  if (c->dim == 2) {
    k_bounds_partials<2><<<nblk, kBlock, 0, s>>>(n, c->d.aabb_min, c->d.aabb_max, c->d.red);
    k_grid_params<2><<<1, 64, 0, s>>>(c->d.red, nblk, n, T - 1, c->d.grid);
    k_grid_count<2><<<nblk, kBlock, 0, s>>>(n, c->d.aabb_min, c->d.aabb_max, c->d.grid, fill, c->d.ctr);
  } else {
    k_bounds_partials<3><<<nblk, kBlock, 0, s>>>(n, c->d.aabb_min, c->d.aabb_max, c->d.red);
    k_grid_params<3><<<1, 64, 0, s>>>(c->d.red, nblk, n, T - 1, c->d.grid);
    k_grid_count<3><<<nblk, kBlock, 0, s>>>(n, c->d.aabb_min, c->d.aabb_max, c->d.grid, fill, c->d.ctr);
  }
  e = launch_exclusive_scan(fill, start, T, tmp, s);
  if (e != hipSuccess) return e;
  if (c->dim == 2)
    k_grid_scatter<2><<<nblk, kBlock, 0, s>>>(n, c->d.aabb_min, c->d.aabb_max, c->d.grid, fill, start, c->d.ekey,
                                               c->d.eid, c->cap_entries, c->d.ctr);
  else
    k_grid_scatter<3><<<nblk, kBlock, 0, s>>>(n, c->d.aabb_min, c->d.aabb_max, c->d.grid, fill, start, c->d.ekey,
                                               c->d.eid, c->cap_entries, c->d.ctr);
  return hipGetLastError();
}

template <int D, int FILTER>
static void pairs_pass(scx_ctx* c, uint32_t n, int mode, uint32_t nblk, const NarrowArgs& na, uint32_t* out_a,
                       uint32_t* out_b, uint32_t cap) {
  const uint32_t T = c->table_size;
  const uint32_t* start = c->d.bucket + T;
  if (mode == 0)
    k_pairs<D, 0, FILTER><<<nblk, kBlock, 0, c->stream>>>(n, c->d.aabb_min, c->d.aabb_max, c->d.grid, start,
                                                          c->d.ekey, c->d.eid, na, c->d.pcount, out_a, out_b, cap,
                                                          c->cap_entries, c->d.ctr);
  else
    k_pairs<D, 1, FILTER><<<nblk, kBlock, 0, c->stream>>>(n, c->d.aabb_min, c->d.aabb_max, c->d.grid, start,
                                                          c->d.ekey, c->d.eid, na, c->d.pcount, out_a, out_b, cap,
                                                          c->cap_entries, c->d.ctr);
}

// mode 0: counts into pcount; mode 1: pairs at the scanned offsets in pcount.
// filter: 0 boxes only, 1 ball2d CCD, 2 sphere overlap at q1.
hipError_t launch_pairs(scx_ctx* c, uint32_t n, int mode, int filter, uint32_t* out_a, uint32_t* out_b,
                        uint32_t cap) {
  const uint32_t nblk = (n + kBlock - 1) / kBlock;
  if (nblk == 0) return hipSuccess;
  NarrowArgs na{c->d.q0, c->d.q1p, c->d.r};
  if (c->dim == 2) {
    if (filter == kFilterNone)
      pairs_pass<2, kFilterNone>(c, n, mode, nblk, na, out_a, out_b, cap);
    else
      pairs_pass<2, kFilterCCD2>(c, n, mode, nblk, na, out_a, out_b, cap);
  } else {
    if (filter == kFilterNone)
      pairs_pass<3, kFilterNone>(c, n, mode, nblk, na, out_a, out_b, cap);
    else
      pairs_pass<3, kFilterSphere3>(c, n, mode, nblk, na, out_a, out_b, cap);
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
