// scx_contacts.hip -- active-set assembly after the broad phase: contact frames,
// Delassus diagonal blocks and forcing terms (reference order), the deterministic
// graph colouring of the contact graph and the colour-major record layout used by
// the Gauss-Seidel sweeps.
//
// Contact frames and arms are computed in FP64 without contraction so that the
// 3D tangent choice (a branch on |n x u|^2 < 1e-9) rounds like the CPU restatement.
#pragma clang fp contract(off)

#include "scx_device.h"
#include "scx_internal.h"

namespace scx {

__device__ __forceinline__ void cross3(const double* a, const double* b, double* o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

template <int D>
__device__ __forceinline__ void normalize(double* v) {
  double z = v[0] * v[0];
#pragma unroll
  for (int a = 1; a < D; ++a) z = z + v[a] * v[a];
  if (z > 0.0) {
    const double s = sqrt(z);
#pragma unroll
    for (int a = 0; a < D; ++a) v[a] = v[a] / s;
  }
}

// FrictionUtilities::orthogonalVector (rigidbody3d/Constraints/FrictionUtilities.cpp:19-29).
__device__ __forceinline__ void orthogonal_vector(const double* n, double* o) {
  double e[3] = {0.0, 0.0, 0.0};
  if (fabs(n[0]) <= fabs(n[1]) && fabs(n[0]) <= fabs(n[2]))
    e[0] = 1.0;
  else if (fabs(n[1]) <= fabs(n[2]))
    e[1] = 1.0;
  else
    e[2] = 1.0;
  const double d = n[0] * e[0] + n[1] * e[1] + n[2] * e[2];
  for (int a = 0; a < 3; ++a) o[a] = e[a] - d * n[a];
  normalize<3>(o);
}

// Tangent s opposing the tangential relative velocity u at the contact, or the
// least-aligned axis if |n x u|^2 < 1e-9; t = normalize(n x s)
// (SphereSphereConstraint.cpp:273-300, StaticPlaneSphereConstraint.cpp:255-283).
__device__ __forceinline__ void frame3(const double* nn, double* s, double* t) {
  double nxs[3];
  cross3(nn, s, nxs);
  if (nxs[0] * nxs[0] + nxs[1] * nxs[1] + nxs[2] * nxs[2] < 1.0e-9) {
    orthogonal_vector(nn, s);
  } else {
    const double d = s[0] * nn[0] + s[1] * nn[1] + s[2] * nn[2];
    for (int a = 0; a < 3; ++a) s[a] = s[a] - d * nn[a];
    normalize<3>(s);
  }
  for (int a = 0; a < 3; ++a) s[a] *= -1.0;
  cross3(nn, s, t);
  normalize<3>(t);
}

struct SetupArgs {
  uint32_t nbb, npl;
  const uint32_t* plane_body;    // plane contact k - nbb -> body (plane-major order)
  const uint32_t* plane_off;     // nplanes + 1 offsets into plane_body
  uint32_t nplanes;              // plane buckets (ensembles: planes per scene, scene-local)
  const uint32_t* bscene;        // ensembles: scene of each body (else nullptr)
  const uint32_t* scene_poff;    // ensembles: first plane of each scene
  const double* q0;
  const double* v0;  // velocity of the frames and of the CoR term (the map's choice)
  const double* w0;
  const double* r;
  const double* m;
  const double* inertia;
  const double* px;
  const double* pn;
  double CoR;
  double stab_dt;        // > 0: stabilization term min(0, gap(q0)) / stab_dt in the normal forcing
  double pen_threshold;  // gaps of smaller magnitude count as 0
  uint32_t* ca;
  uint32_t* cb;
  int32_t* cplane;
  double* basis;
  double* la;
  double* lb;
  double* wn;
  double* wt;
  double* forcing;
};

// Contact k of the active set, in the reference's order: body pairs first
// (lexicographic), then plane contacts plane-major
// (Ball2DSim::computeActiveSet ball2d/Ball2DSim.cpp:151-173;
//  RigidBody3DSim::computeActiveSet rigidbody3d/RigidBody3DSim.cpp:250-262).
template <int D>
__global__ void __launch_bounds__(kBlock) k_contact_setup(SetupArgs A) {
  const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t C = A.nbb + A.npl;
  if (k >= C) return;
  uint32_t a, b;
  int32_t plane = -1;
  if (k < A.nbb) {
    a = A.ca[k];
    b = A.cb[k];
  } else {
    const uint32_t kp = k - A.nbb;
    a = A.plane_body[kp];
    b = kNoBody;
    plane = 0;
    while ((uint32_t)(plane + 1) < A.nplanes && A.plane_off[plane + 1] <= kp) ++plane;
    if (A.bscene) plane += (int32_t)A.scene_poff[A.bscene[a]];
    A.ca[k] = a;
    A.cb[k] = b;
  }
  A.cplane[k] = plane;
  const double minv_a = 1.0 / A.m[a];
  const double minv_b = (b != kNoBody) ? 1.0 / A.m[b] : 0.0;
  double nn[3] = {0.0, 0.0, 0.0};
  double gap = 0.0;
  if (b != kNoBody) {
    for (int d = 0; d < D; ++d) nn[d] = A.q0[(size_t)D * a + d] - A.q0[(size_t)D * b + d];
    if (A.stab_dt > 0.0) {
      // evaluateGapFunction at q0 (ball2d/Constraints/BallBallConstraint.cpp:76-86: minus the
      // radius sum; rigidbody3d/Constraints/SphereSphereConstraint.cpp:28-31: each radius)
      double d2 = nn[0] * nn[0];
      for (int d = 1; d < D; ++d) d2 = d2 + nn[d] * nn[d];
      gap = (D == 2) ? sqrt(d2) - (A.r[a] + A.r[b]) : (sqrt(d2) - A.r[a]) - A.r[b];
    }
    normalize<D>(nn);
  } else {
    for (int d = 0; d < D; ++d) nn[d] = A.pn[(size_t)D * plane + d];
    if (A.stab_dt > 0.0) {
      // n . (x - x_plane) - r (BallStaticPlaneConstraint.cpp:58-63, StaticPlaneSphereConstraint.cpp:31-35)
      double s = 0.0;
      for (int d = 0; d < D; ++d) {
        const double dd = A.q0[(size_t)D * a + d] - A.px[(size_t)D * plane + d];
        s = (d == 0) ? nn[0] * dd : s + nn[d] * dd;
      }
      gap = s - A.r[a];
    }
  }
  double ndotv = 0.0;
  for (int d = 0; d < D; ++d) {
    const double dv = A.v0[(size_t)D * a + d] - ((b != kNoBody) ? A.v0[(size_t)D * b + d] : 0.0);
    ndotv = (d == 0) ? nn[0] * dv : ndotv + nn[d] * dv;
  }
  double forcing = A.CoR * ndotv;
  if (A.stab_dt > 0.0) {
    // SymplecticEulerImpactFrictionMap.cpp:309-327: g0 = min(0, g(q0)), 0 below the threshold
    double g0 = fmin(0.0, gap);
    if (fabs(g0) < A.pen_threshold) g0 = 0.0;
    forcing = forcing + g0 / A.stab_dt;
  }
  A.forcing[k] = forcing;
  if (D == 2) {
    // basis [n, t], t = (-n_y, n_x) (BallBallConstraint.cpp:244-258, StaticPlane.cpp:10-20)
    double* B = A.basis + 4 * (size_t)k;
    B[0] = nn[0];
    B[1] = nn[1];
    B[2] = -nn[1];
    B[3] = nn[0];
    // W_ii = (1/ma + 1/mb) [n;t][n;t]^T, diagonal entries
    const double wn = (nn[0] * minv_a * nn[0] + nn[1] * minv_a * nn[1]) +
                      ((b != kNoBody) ? (nn[0] * minv_b * nn[0] + nn[1] * minv_b * nn[1]) : 0.0);
    const double wt = (nn[1] * minv_a * nn[1] + nn[0] * minv_a * nn[0]) +
                      ((b != kNoBody) ? (nn[1] * minv_b * nn[1] + nn[0] * minv_b * nn[0]) : 0.0);
    A.wn[k] = wn;
    A.wt[k] = wt;
    A.la[k] = 0.0;
    A.lb[k] = 0.0;
  } else {
    double pa[3], pb[3], s[3], t[3];
    const double* xa = A.q0 + 3 * (size_t)a;
    const double* va = A.v0 + 3 * (size_t)a;
    const double* wa = A.w0 + 3 * (size_t)a;
    double ua[3];
    if (b != kNoBody) {
      // contact point at q0 (RigidBody3DSim.cpp:803-809)
      const double* xb = A.q0 + 3 * (size_t)b;
      const double f = A.r[a] / (A.r[a] + A.r[b]);
      for (int d = 0; d < 3; ++d) {
        const double p = xa[d] + f * (xb[d] - xa[d]);
        pa[d] = p - xa[d];
        pb[d] = p - xb[d];
      }
      double cb_[3];
      cross3(wa, pa, ua);
      cross3(A.w0 + 3 * (size_t)b, pb, cb_);
      for (int d = 0; d < 3; ++d) s[d] = ((va[d] + ua[d]) - A.v0[3 * (size_t)b + d]) - cb_[d];
    } else {
      for (int d = 0; d < 3; ++d) {
        pa[d] = (-A.r[a]) * nn[d];
        pb[d] = 0.0;
      }
      cross3(wa, pa, ua);
      for (int d = 0; d < 3; ++d) s[d] = (va[d] + ua[d]) - 0.0;
    }
    frame3(nn, s, t);
    double* B = A.basis + 9 * (size_t)k;
    for (int d = 0; d < 3; ++d) {
      B[d] = nn[d];
      B[3 + d] = s[d];
      B[6 + d] = t[d];
    }
    // arms are parallel to n: r_a = -la n, r_b = +lb n
    const double la = sqrt(pa[0] * pa[0] + pa[1] * pa[1] + pa[2] * pa[2]);
    const double lb = sqrt(pb[0] * pb[0] + pb[1] * pb[1] + pb[2] * pb[2]);
    const double Iinv_a = 1.0 / A.inertia[a];
    const double Iinv_b = (b != kNoBody) ? 1.0 / A.inertia[b] : 0.0;
    A.la[k] = la;
    A.lb[k] = lb;
    A.wn[k] = minv_a + minv_b;
    A.wt[k] = (minv_a + Iinv_a * la * la) + ((b != kNoBody) ? (minv_b + Iinv_b * lb * lb) : 0.0);
  }
}

hipError_t launch_contact_setup(scx_ctx* c, uint32_t nbb, uint32_t npl, double CoR, double stab_dt,
                                double pen_threshold) {
  const uint32_t C = nbb + npl;
  const uint32_t nblk = (C + kBlock - 1) / kBlock;
  if (nblk == 0) return hipSuccess;
  SetupArgs A;
  A.nbb = nbb;
  A.npl = npl;
  A.plane_body = c->d.perm;  // plane contacts' bodies, written by the plane compaction
  A.plane_off = c->d.coloff; // plane offsets, likewise
  A.nplanes = c->nscenes ? c->scene_planes : c->nplanes;
  A.bscene = c->nscenes ? c->d.bscene : nullptr;
  A.scene_poff = c->d.scene_poff;
  A.q0 = c->d.q0;
  A.v0 = c->vbasis;
  A.w0 = c->d.w0;
  A.r = c->d.r;
  A.m = c->d.m;
  A.inertia = c->d.inertia;
  A.px = c->d.px;
  A.pn = c->d.pn;
  A.CoR = CoR;
  A.stab_dt = stab_dt;
  A.pen_threshold = pen_threshold;
  A.ca = c->d.ca;
  A.cb = c->d.cb;
  A.cplane = c->d.cplane;
  A.basis = c->d.basis;
  A.la = c->d.cla;
  A.lb = c->d.clb;
  A.wn = c->d.cwn;
  A.wt = c->d.cwt;
  A.forcing = c->d.cforcing;
  if (c->dim == 2)
    k_contact_setup<2><<<nblk, kBlock, 0, c->stream>>>(A);
  else
    k_contact_setup<3><<<nblk, kBlock, 0, c->stream>>>(A);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Body -> contact adjacency (CSR), then Jones-Plassmann colouring rounds.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_adj_count(uint32_t C, const uint32_t* __restrict__ ca,
                                                      const uint32_t* __restrict__ cb, uint32_t* __restrict__ deg) {
  const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= C) return;
  atomicAdd(&deg[ca[k]], 1u);
  const uint32_t b = cb[k];
  if (b != kNoBody) atomicAdd(&deg[b], 1u);
}

__global__ void __launch_bounds__(kBlock) k_adj_fill(uint32_t C, const uint32_t* __restrict__ ca,
                                                     const uint32_t* __restrict__ cb, uint32_t* __restrict__ deg,
                                                     const uint32_t* __restrict__ off, uint32_t* __restrict__ adj) {
  const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= C) return;
  const uint32_t a = ca[k];
  adj[off[a] + atomicSub(&deg[a], 1u) - 1u] = k;
  const uint32_t b = cb[k];
  if (b != kNoBody) adj[off[b] + atomicSub(&deg[b], 1u) - 1u] = k;
}

// Colouring specification (shared with the CPU restatement): in round k every
// uncoloured contact whose priority key exceeds that of every uncoloured neighbour
// takes the smallest colour unused by its coloured neighbours. Neighbours coloured
// during round k still count as uncoloured in round k, so decisions never depend
// on scheduling. state = (round << 8) | colour, or -1.
//
// Each workgroup owns one contiguous range of contacts and keeps the ones still
// uncoloured in an LDS list (when the range fits kColourLds), so a round visits only
// those. Every access to state and to the round counts is an atomic that goes to memory,
// so the rounds are separated by grid_sync_coherent alone, one per round: the contacts
// coloured in round r are added to count r % 3, which every workgroup reads after the
// barrier (the same value everywhere, hence the same exit), and count (r + 2) % 3, read a
// round earlier, is cleared for round r + 2.
constexpr uint32_t kColourLds = 4096;

__global__ void __launch_bounds__(kBlock) k_colour(uint32_t C, const uint32_t* __restrict__ ca,
                                                   const uint32_t* __restrict__ cb, const uint32_t* __restrict__ off,
                                                   const uint32_t* __restrict__ adj, int32_t* state,
                                                   StepCounters* ctr, uint32_t* bar) {
  __shared__ uint16_t s_list[2][kColourLds];
  __shared__ uint32_t s_n[2];
  uint32_t* cnt = bar + 144;  // three round counts
  const uint32_t per = (C + gridDim.x - 1) / gridDim.x;
  const uint32_t lo = min(C, blockIdx.x * per), hi = min(C, lo + per);
  const bool listed = per <= kColourLds;
  for (uint32_t k = lo + threadIdx.x; k < hi; k += kBlock) {
    __hip_atomic_store(&state[k], -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (listed) s_list[0][k - lo] = (uint16_t)(k - lo);
  }
  if (threadIdx.x == 0) {
    s_n[0] = hi - lo;
    s_n[1] = 0;
  }
  if (blockIdx.x == 0 && threadIdx.x < 3) __hip_atomic_store(cnt + threadIdx.x, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  grid_sync_coherent(bar);
  uint32_t done = 0, cur = 0;
  for (uint32_t round = 0; round <= C && done < C; ++round) {
    uint32_t mine = 0;
    uint32_t maxcol = 0;
    const uint32_t nk = listed ? s_n[cur] : hi - lo;
    for (uint32_t i = threadIdx.x; i < nk; i += kBlock) {
      const uint32_t c = lo + (listed ? (uint32_t)s_list[cur][i] : i);
      if (!listed && __hip_atomic_load(&state[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= 0) continue;
      const unsigned long long kc = prio_key(c);
      const uint32_t bodies[2] = {ca[c], cb[c]};
      bool is_max = true;
      unsigned long long used = 0;
      for (int q = 0; q < 2 && is_max; ++q) {
        if (bodies[q] == kNoBody) continue;
        for (uint32_t e = off[bodies[q]]; e < off[bodies[q] + 1]; ++e) {
          const uint32_t d = adj[e];
          if (d == c) continue;
          const int32_t sd = __hip_atomic_load(&state[d], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          const bool unc = (sd < 0) || ((uint32_t)sd >> 8) == round;
          if (unc) {
            if (prio_key(d) > kc) {
              is_max = false;
              break;
            }
          } else {
            used |= 1ull << (sd & 0xff);
          }
        }
      }
      if (!is_max) {
        if (listed) s_list[cur ^ 1][atomicAdd(&s_n[cur ^ 1], 1u)] = (uint16_t)(c - lo);
        continue;
      }
      const int col = (used == ~0ull) ? 64 : __ffsll((long long)~used) - 1;
      if (col >= kMaxColours) {
        atomicOr(&ctr->overflow, 8u);
        __hip_atomic_store(&state[c], (int32_t)((round << 8) | 63), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      } else {
        __hip_atomic_store(&state[c], (int32_t)((round << 8) | col), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        maxcol = max(maxcol, (uint32_t)col + 1);
      }
      ++mine;
    }
    mine = block_reduce_sum_u32(mine);
    if (threadIdx.x == 0 && mine) atomicAdd(cnt + round % 3, mine);
    if (maxcol) atomicMax(&ctr->ncolours, maxcol);
    grid_sync_coherent(bar);
    if (blockIdx.x == 0 && threadIdx.x == 0)
      __hip_atomic_store(cnt + (round + 2) % 3, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    done += __hip_atomic_load(cnt + round % 3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (blockIdx.x == 0 && threadIdx.x == 0 && done >= C) ctr->nrounds = round + 1;
    // the next round's list is complete (every thread passed the barrier); the old one is
    // emptied for the round after
    if (threadIdx.x == 0) s_n[cur] = 0;
    cur ^= 1;
    __syncthreads();
  }
  grid_exit_coherent(bar);
}

// Colour masks for the stable colour-major ordering.
__global__ void __launch_bounds__(kBlock) k_colour_masks(uint32_t C, const int32_t* __restrict__ state,
                                                         unsigned long long* __restrict__ mask) {
  const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= C) return;
  mask[k] = 1ull << (state[k] & 0xff);
}

hipError_t launch_colouring(scx_ctx* c, uint32_t C) {
  hipStream_t s = c->stream;
  const uint32_t n = c->n;
  const uint32_t nblk = (C + kBlock - 1) / kBlock;
  uint32_t* deg = c->d.deg;            // n counts
  uint32_t* off = c->d.deg + n + 1;    // n + 1 offsets
  hipError_t e = hipMemsetAsync(deg, 0, sizeof(uint32_t) * (n + 1), s);
  if (e != hipSuccess) return e;
  if (nblk) k_adj_count<<<nblk, kBlock, 0, s>>>(C, c->d.ca, c->d.cb, deg);
  e = launch_exclusive_scan(deg, off, n, c->d.scan_tmp, s);
  if (e != hipSuccess) return e;
  if (nblk) k_adj_fill<<<nblk, kBlock, 0, s>>>(C, c->d.ca, c->d.cb, deg, off, c->d.adj);
  if (C == 0) return hipGetLastError();
  // every round passes over the uncoloured contacts between two barriers
  const int want = c->force_blocks > 0 ? c->force_blocks : persistent_blocks(C, kBlock, c->colour_blocks);
  const int blocks = std::max(1, std::min(c->colour_blocks, want));
  uint32_t Cv = C;
  const uint32_t* pa = c->d.ca;
  const uint32_t* pb = c->d.cb;
  const uint32_t* poff = off;
  const uint32_t* padj = c->d.adj;
  int32_t* pstate = c->d.cstate;
  StepCounters* pctr = c->d.ctr;
  uint32_t* pbar = c->d.barrier;
  void* args[] = {&Cv, &pa, &pb, &poff, &padj, &pstate, &pctr, &pbar};
  bool launched = false;
  const hipError_t e2 = launch_cooperative(c, (const void*)k_colour, blocks, kBlock, args, 0, s, &launched);
  if (e2 != hipSuccess || launched) return e2;
  // one workgroup: its grid barrier is a workgroup barrier, no cooperative launch needed
  return hipLaunchKernel((const void*)k_colour, dim3(1), dim3(kBlock), args, 0, s);
}

int colour_max_blocks(scx_ctx* c) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_colour, kBlock, 0) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  return per_cu * c->num_cus;
}

// perm: colour order -> reference order (stable), coloff: colour offsets.
// Contacts by the sweep slot of their first body (counting sort; contacts that share a
// body, and so a slot, lie in different colours).
__global__ void __launch_bounds__(kBlock) k_slot_count(uint32_t C, const uint32_t* __restrict__ ca,
                                                       const uint32_t* __restrict__ slot_map,
                                                       uint32_t* __restrict__ cnt) {
  const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k < C) atomicAdd(&cnt[slot_map[ca[k]]], 1u);
}

__global__ void __launch_bounds__(kBlock) k_slot_scatter(uint32_t C, const uint32_t* __restrict__ ca,
                                                         const uint32_t* __restrict__ slot_map,
                                                         const uint32_t* __restrict__ start,
                                                         uint32_t* __restrict__ cnt, uint32_t* __restrict__ list) {
  const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= C) return;
  const uint32_t sl = slot_map[ca[k]];
  list[start[sl] + atomicSub(&cnt[sl], 1u) - 1u] = k;
}

__global__ void __launch_bounds__(kBlock) k_list_colour_masks(uint32_t C, const uint32_t* __restrict__ list,
                                                              const int32_t* __restrict__ state,
                                                              unsigned long long* __restrict__ mask) {
  const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
  if (j < C) mask[j] = 1ull << (state[list[j]] & 0xff);
}

__global__ void __launch_bounds__(kBlock) k_compose(uint32_t C, const uint32_t* __restrict__ list,
                                                    uint32_t* __restrict__ perm) {
  const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k < C) perm[k] = list[perm[k]];
}

// perm: colour order -> reference order, coloff: colour offsets. Within a colour the
// contacts follow the sweep slot of their first body when slots are in use (the order
// inside a colour does not change the sweep's result: its contacts share no body), else
// the reference order.
hipError_t launch_colour_sort(scx_ctx* c, uint32_t C) {
  const uint32_t nblk = (C + kBlock - 1) / kBlock;
  if (nblk == 0) return hipSuccess;
  hipStream_t s = c->stream;
  DeviceBuffers& d = c->d;
  if (!d.slot_map) {
    k_colour_masks<<<nblk, kBlock, 0, s>>>(C, d.cstate, d.pmask);
    return launch_multibucket(c, d.pmask, C, kMaxColours, d.perm, c->cap_perm, d.coloff);
  }
  const uint32_t n = c->n;
  uint32_t* cnt = d.deg;              // n + 1 (free after the colouring)
  uint32_t* start = d.deg + n + 1;    // n + 1
  uint32_t* list = d.adj;             // C of 2C (free after the colouring)
  hipError_t e = hipMemsetAsync(cnt, 0, sizeof(uint32_t) * (n + 1), s);
  if (e != hipSuccess) return e;
  k_slot_count<<<nblk, kBlock, 0, s>>>(C, d.ca, d.slot_map, cnt);
  e = launch_exclusive_scan(cnt, start, n, d.scan_tmp, s);
  if (e != hipSuccess) return e;
  k_slot_scatter<<<nblk, kBlock, 0, s>>>(C, d.ca, d.slot_map, start, cnt, list);
  k_list_colour_masks<<<nblk, kBlock, 0, s>>>(C, list, d.cstate, d.pmask);
  e = launch_multibucket(c, d.pmask, C, kMaxColours, d.perm, c->cap_perm, d.coloff);
  if (e != hipSuccess) return e;
  k_compose<<<nblk, kBlock, 0, s>>>(C, list, d.perm);
  return hipGetLastError();
}

// Records in colour order for the sweeps (FP64, and FP32 when requested).
__global__ void __launch_bounds__(kBlock) k_gather(uint32_t C, int dim, const uint32_t* __restrict__ perm,
                                                   const uint32_t* __restrict__ ca, const uint32_t* __restrict__ cb,
                                                   const double* __restrict__ basis, const double* __restrict__ la,
                                                   const double* __restrict__ lb, const double* __restrict__ wn,
                                                   const double* __restrict__ wt, const double* __restrict__ forcing,
                                                   const double* __restrict__ rinit, double mu, const double* __restrict__ mu_arr,
                                                   const uint32_t* __restrict__ slot_map, ContactSoA<double> o64,
                                                   ContactSoA<float> o32, int fp32) {
  const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= C) return;
  const uint32_t c = perm[k];
  // the sweeps address bodies by slot
  const uint32_t a = slot_map ? slot_map[ca[c]] : ca[c];
  const uint32_t b = (cb[c] == kNoBody) ? kNoBody : (slot_map ? slot_map[cb[c]] : cb[c]);
  o64.a[k] = a;
  o64.b[k] = b;
  double n[3] = {0, 0, 0}, s[3] = {0, 0, 0}, r[3] = {0, 0, 0};
  const int DD = dim * dim;
  for (int d = 0; d < dim; ++d) {
    n[d] = basis[(size_t)DD * c + d];
    if (dim == 3) s[d] = basis[(size_t)DD * c + 3 + d];
    r[d] = rinit ? rinit[(size_t)dim * c + d] : 0.0;
  }
  for (int d = 0; d < 3; ++d) {
    o64.n[3 * (size_t)k + d] = n[d];
    o64.s[3 * (size_t)k + d] = s[d];
    o64.r[3 * (size_t)k + d] = r[d];
  }
  o64.la[k] = la[c];
  o64.lb[k] = lb[c];
  o64.wn[k] = wn[c];
  o64.wt[k] = wt[c];
  const double muc = mu_arr ? mu_arr[c] : mu;
  o64.mu[k] = muc;
  o64.forcing[k] = forcing[c];
  if (fp32) {
    o32.a[k] = a;
    o32.b[k] = b;
    for (int d = 0; d < 3; ++d) {
      o32.n[3 * (size_t)k + d] = (float)n[d];
      o32.s[3 * (size_t)k + d] = (float)s[d];
      o32.r[3 * (size_t)k + d] = (float)r[d];
    }
    o32.la[k] = (float)la[c];
    o32.lb[k] = (float)lb[c];
    o32.wn[k] = (float)wn[c];
    o32.wt[k] = (float)wt[c];
    o32.mu[k] = (float)muc;
    o32.forcing[k] = (float)forcing[c];
  }
}

// Cache keys in active-set order, ascending: body pairs (i << 32 | j), then plane
// contacts (1 << 63 | plane << 32 | i) -- the (i, j) and (plane, i) keys of
// ball2d/ConstraintCache.cpp:20-123. In an ensemble the plane contacts are ordered by
// the scene-local plane index (all scenes' plane-0 contacts, then plane 1, ...), so the
// key takes that local index: with global body ids the keys stay unique and ascend in
// that order, which the binary search of k_warm_lookup needs.
__global__ void __launch_bounds__(kBlock) k_contact_keys(uint32_t C, const uint32_t* __restrict__ ca,
                                                         const uint32_t* __restrict__ cb,
                                                         const int32_t* __restrict__ cplane,
                                                         const uint32_t* __restrict__ bscene,
                                                         const uint32_t* __restrict__ scene_poff,
                                                         unsigned long long* __restrict__ key) {
  const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= C) return;
  if (cb[k] != kNoBody) {
    key[k] = ((unsigned long long)ca[k] << 32) | cb[k];
  } else {
    uint32_t plane = (uint32_t)cplane[k];
    if (bscene) plane -= scene_poff[bscene[ca[k]]];
    key[k] = (1ull << 63) | ((unsigned long long)plane << 32) | ca[k];
  }
}

// Cached world impulse r_old rotated by the minimal rotation taking the old normal n_old
// to the current normal n (the quaternion FromTwoVectors / vector product that
// SymplecticEulerImpactFrictionMap.cpp:96-115 applies); all-zero entries give zero. 2D
// vectors rotate in the xy-plane.
__device__ void transport_impulse(int dim, const double* n_old, const double* r_old, const double* n, double* f) {
  double a[3] = {0, 0, 0}, b[3] = {0, 0, 0}, r[3] = {0, 0, 0};
  bool zero = true;
  for (int k = 0; k < dim; ++k) {
    a[k] = n_old[k];
    b[k] = n[k];
    r[k] = r_old[k];
    zero = zero && n_old[k] == 0.0 && r_old[k] == 0.0;
  }
  f[0] = f[1] = f[2] = 0.0;
  if (zero) return;
  const double na = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
  const double nb = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
  for (int k = 0; k < 3; ++k) {
    a[k] = a[k] / na;
    b[k] = b[k] / nb;
  }
  const double c = b[0] * a[0] + b[1] * a[1] + b[2] * a[2];
  double v[3], w;
  if (c < -1.0 + 1e-12) {
    // opposite normals: half turn about an axis orthogonal to the old normal
    double ax[3];
    orthogonal_vector(a, ax);
    const double w2 = (1.0 + fmax(c, -1.0)) * 0.5;
    w = sqrt(w2);
    const double sv = sqrt(1.0 - w2);
    for (int k = 0; k < 3; ++k) v[k] = ax[k] * sv;
  } else {
    const double ax[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    const double s = sqrt((1.0 + c) * 2.0);
    const double invs = 1.0 / s;
    for (int k = 0; k < 3; ++k) v[k] = ax[k] * invs;
    w = s * 0.5;
  }
  double uv[3] = {v[1] * r[2] - v[2] * r[1], v[2] * r[0] - v[0] * r[2], v[0] * r[1] - v[1] * r[0]};
  for (int k = 0; k < 3; ++k) uv[k] = uv[k] + uv[k];
  const double vu[3] = {v[1] * uv[2] - v[2] * uv[1], v[2] * uv[0] - v[0] * uv[2], v[0] * uv[1] - v[1] * uv[0]};
  for (int k = 0; k < 3; ++k) f[k] = (r[k] + w * uv[k]) + vu[k];
}

// initializeImpulses (GeometricImpactFrictionMap.cpp:174-220): the cached impulse of the
// same key (binary search in the previous step's ascending keys), zero on a miss
// (ball2d/ConstraintCache.cpp:122). World impulse alpha n + sum_k beta_k t_k
// (Sobogus.cpp:301, :518; beta = 0 for NORMAL), or the transported world impulse, then
// local coordinates E^T r. Cache entries are kCacheWords doubles: (alpha, beta...) or
// (old normal, world impulse).
constexpr int kCacheWords = 6;

__global__ void __launch_bounds__(kBlock) k_warm_lookup(uint32_t C, int dim, int mode,
                                                        const unsigned long long* __restrict__ key,
                                                        const unsigned long long* __restrict__ pkey, uint32_t P,
                                                        const double* __restrict__ pr,
                                                        const double* __restrict__ basis,
                                                        double* __restrict__ rinit) {
  const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= C) return;
  const unsigned long long kk = key[k];
  uint32_t lo = 0, hi = P;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (pkey[mid] < kk)
      lo = mid + 1;
    else
      hi = mid;
  }
  double r[3] = {0.0, 0.0, 0.0};
  if (lo < P && pkey[lo] == kk) {
    const double* B = basis + (size_t)dim * dim * k;
    const double* c = pr + (size_t)kCacheWords * lo;
    double rw[3] = {0.0, 0.0, 0.0};
    if (mode == SCX_WARM_TRANSPORT) {
      transport_impulse(dim, c, c + 3, B, rw);
    } else {
      for (int a = 0; a < dim; ++a) rw[a] = c[0] * B[a];
      if (mode == SCX_WARM_CACHE)
        for (int j = 1; j < dim; ++j)
          for (int a = 0; a < dim; ++a) rw[a] = rw[a] + c[j] * B[j * dim + a];
    }
    for (int j = 0; j < dim; ++j) {
      double acc = B[j * dim] * rw[0];
      for (int a = 1; a < dim; ++a) acc = acc + B[j * dim + a] * rw[a];
      r[j] = acc;
    }
  }
  for (int j = 0; j < dim; ++j) rinit[(size_t)dim * k + j] = r[j];
}

// cacheImpulses (GeometricImpactFrictionMap.cpp:222-264): alpha, and beta with friction;
// for the transported cache the normal and the world impulse alpha n + sum_k beta_k t_k.
__global__ void __launch_bounds__(kBlock) k_cache_store(uint32_t C, int dim, int mode,
                                                        const unsigned long long* __restrict__ key,
                                                        const double* __restrict__ basis,
                                                        const double* __restrict__ alpha,
                                                        const double* __restrict__ beta,
                                                        unsigned long long* __restrict__ pkey,
                                                        double* __restrict__ pr) {
  const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= C) return;
  pkey[k] = key[k];
  double* o = pr + (size_t)kCacheWords * k;
  for (int j = 0; j < kCacheWords; ++j) o[j] = 0.0;
  if (mode == SCX_WARM_TRANSPORT) {
    const double* B = basis + (size_t)dim * dim * k;
    for (int a = 0; a < dim; ++a) {
      o[a] = B[a];
      double w = alpha[k] * B[a];
      for (int j = 1; j < dim; ++j) w = w + beta[(size_t)(dim - 1) * k + j - 1] * B[j * dim + a];
      o[3 + a] = w;
    }
  } else {
    o[0] = alpha[k];
    if (mode == SCX_WARM_CACHE)
      for (int j = 1; j < dim; ++j) o[j] = beta[(size_t)(dim - 1) * k + j - 1];
  }
}

// Per-contact inputs of the learned warm start (16 FP32, reference order), with the
// impulse scale s = vref / w_n (the impulse that stops a relative velocity vref = |g| dt):
//   0-2 normal n (world; 2D z = 0)
//   3-5 pre-solve relative velocity in the contact frame, u = H v0 + w, over vref
//   6   w_t / w_n (tangential over normal Delassus diagonal)
//   7   mu
//   8-10 the transported previous impulse (local frame) over s, 11: 1 if it is non-zero
//   12  overlap at q0 over the radius sum (plane: over the radius)
//   13  1 for a plane contact, 14: 0, 15: 1
template <int D>
__global__ void __launch_bounds__(kBlock) k_contact_features(uint32_t C, const uint32_t* __restrict__ ca,
                                                             const uint32_t* __restrict__ cb,
                                                             const int32_t* __restrict__ cplane,
                                                             const double* __restrict__ basis,
                                                             const double* __restrict__ la,
                                                             const double* __restrict__ lb,
                                                             const double* __restrict__ wn,
                                                             const double* __restrict__ wt, double mu,
                                                             const double* __restrict__ forcing,
                                                             const double* __restrict__ q0,
                                                             const double* __restrict__ v0,
                                                             const double* __restrict__ w0,
                                                             const double* __restrict__ r,
                                                             const double* __restrict__ px,
                                                             const double* __restrict__ rprev, double vref,
                                                             const double* __restrict__ mu_arr,
                                                             float* __restrict__ X, float* __restrict__ S) {
  const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= C) return;
  const uint32_t a = ca[k], b = cb[k];
  const bool hasb = b != kNoBody;
  const double* B = basis + (size_t)D * D * k;
  const double s = vref / wn[k];
  float f[16];
  for (int i = 0; i < 16; ++i) f[i] = 0.f;
  for (int i = 0; i < D; ++i) f[i] = (float)B[i];
  // u = H v0 + w: dv projected on the frame, minus the angular terms (arms -la n, +lb n)
  double dv[3] = {0, 0, 0};
  for (int i = 0; i < D; ++i) dv[i] = v0[(size_t)D * a + i] - (hasb ? v0[(size_t)D * b + i] : 0.0);
  if (D == 3) {
    const double* n = B;
    const double* wa = w0 + 3 * (size_t)a;
    double ca_[3] = {wa[1] * n[2] - wa[2] * n[1], wa[2] * n[0] - wa[0] * n[2], wa[0] * n[1] - wa[1] * n[0]};
    double cb_[3] = {0, 0, 0};
    if (hasb) {
      const double* wb = w0 + 3 * (size_t)b;
      cb_[0] = wb[1] * n[2] - wb[2] * n[1];
      cb_[1] = wb[2] * n[0] - wb[0] * n[2];
      cb_[2] = wb[0] * n[1] - wb[1] * n[0];
    }
    for (int i = 0; i < 3; ++i) dv[i] = dv[i] - la[k] * ca_[i] - lb[k] * cb_[i];
  }
  for (int j = 0; j < D; ++j) {
    double u = 0.0;
    for (int i = 0; i < D; ++i) u += B[j * D + i] * dv[i];
    if (j == 0) u += forcing[k];
    f[3 + j] = (float)(u / vref);
  }
  f[6] = (float)(wt[k] / wn[k]);
  f[7] = (float)(mu_arr ? mu_arr[k] : mu);
  bool prev = false;
  for (int j = 0; j < D; ++j) {
    const double rp = rprev[(size_t)D * k + j];
    f[8 + j] = (float)(rp / s);
    prev = prev || rp != 0.0;
  }
  f[11] = prev ? 1.f : 0.f;
  if (hasb) {
    double d2 = 0.0;
    for (int i = 0; i < D; ++i) {
      const double dx = q0[(size_t)D * a + i] - q0[(size_t)D * b + i];
      d2 += dx * dx;
    }
    const double rs = r[a] + r[b];
    f[12] = (float)((rs - sqrt(d2)) / rs);
  } else {
    const int p = cplane[k];
    double h = 0.0;
    for (int i = 0; i < D; ++i) h += B[i] * (q0[(size_t)D * a + i] - px[(size_t)D * p + i]);
    f[12] = (float)((r[a] - h) / r[a]);
    f[13] = 1.f;
  }
  f[15] = 1.f;
  float4* xo = reinterpret_cast<float4*>(X + (size_t)16 * k);
  for (int i = 0; i < 4; ++i) xo[i] = make_float4(f[4 * i], f[4 * i + 1], f[4 * i + 2], f[4 * i + 3]);
  S[k] = (float)s;
}

// Predicted impulses (over s) into the warm start: r = s y, normal clamped at zero and
// the tangential part projected onto the friction cone |r_T| <= mu r_n.
__global__ void __launch_bounds__(kBlock) k_apply_prediction(uint32_t C, int dim, double mu,
                                                             const double* __restrict__ mu_arr,
                                                             const float* __restrict__ Y,
                                                             const float* __restrict__ S,
                                                             double* __restrict__ rinit) {
  const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= C) return;
  const double s = S[k];
  const double rn = fmax(0.0, s * (double)Y[(size_t)3 * k]);
  double rt[2] = {0.0, 0.0}, nt = 0.0;
  for (int j = 1; j < dim; ++j) {
    rt[j - 1] = s * (double)Y[(size_t)3 * k + j];
    nt += rt[j - 1] * rt[j - 1];
  }
  nt = sqrt(nt);
  const double lim = (mu_arr ? mu_arr[k] : mu) * rn;
  const double sc = (nt > lim && nt > 0.0) ? lim / nt : 1.0;
  rinit[(size_t)dim * k] = rn;
  for (int j = 1; j < dim; ++j) rinit[(size_t)dim * k + j] = rt[j - 1] * sc;
}

// Training targets of the last solve: the converged local impulses over s.
__global__ void __launch_bounds__(kBlock) k_sample_targets(uint32_t C, int dim, const double* __restrict__ alpha,
                                                           const double* __restrict__ beta,
                                                           const float* __restrict__ S, float* __restrict__ T) {
  const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= C) return;
  const double s = S[k];
  T[(size_t)3 * k] = (float)(alpha[k] / s);
  T[(size_t)3 * k + 1] = (dim > 1) ? (float)(beta[(size_t)(dim - 1) * k] / s) : 0.f;
  T[(size_t)3 * k + 2] = (dim > 2) ? (float)(beta[(size_t)(dim - 1) * k + 1] / s) : 0.f;
}

// Friction coefficient of each contact of an ensemble: that of the scene of body a.
__global__ void __launch_bounds__(kBlock) k_contact_mu(uint32_t C, const uint32_t* __restrict__ ca,
                                                       const uint32_t* __restrict__ bscene,
                                                       const double* __restrict__ scene_mu, double* __restrict__ cmu) {
  const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= C) return;
  cmu[k] = scene_mu[bscene[ca[k]]];
}

hipError_t launch_contact_mu(scx_ctx* c, uint32_t C) {
  const uint32_t nblk = (C + kBlock - 1) / kBlock;
  if (nblk == 0 || c->nscenes == 0) return hipSuccess;
  k_contact_mu<<<nblk, kBlock, 0, c->stream>>>(C, c->d.ca, c->d.bscene, c->d.scene_mu, c->d.cmu);
  return hipGetLastError();
}

hipError_t launch_contact_features(scx_ctx* c, uint32_t C, double mu, double vref) {
  const uint32_t nblk = (C + kBlock - 1) / kBlock;
  if (nblk == 0) return hipSuccess;
  DeviceBuffers& d = c->d;
  if (c->dim == 2)
    k_contact_features<2><<<nblk, kBlock, 0, c->stream>>>(C, d.ca, d.cb, d.cplane, d.basis, d.cla, d.clb, d.cwn,
                                                           d.cwt, mu, d.cforcing, d.q0, c->vsolve, d.w0, d.r, d.px,
                                                           d.crinit, vref, c->nscenes ? d.cmu : nullptr, d.mlp_x,
                                                           d.mlp_s);
  else
    k_contact_features<3><<<nblk, kBlock, 0, c->stream>>>(C, d.ca, d.cb, d.cplane, d.basis, d.cla, d.clb, d.cwn,
                                                           d.cwt, mu, d.cforcing, d.q0, c->vsolve, d.w0, d.r, d.px,
                                                           d.crinit, vref, c->nscenes ? d.cmu : nullptr, d.mlp_x,
                                                           d.mlp_s);
  return hipGetLastError();
}

hipError_t launch_apply_prediction(scx_ctx* c, uint32_t C, double mu) {
  const uint32_t nblk = (C + kBlock - 1) / kBlock;
  if (nblk == 0) return hipSuccess;
  k_apply_prediction<<<nblk, kBlock, 0, c->stream>>>(C, c->dim, mu, c->nscenes ? c->d.cmu : nullptr, c->d.mlp_y,
                                                     c->d.mlp_s, c->d.crinit);
  return hipGetLastError();
}

hipError_t launch_sample_targets(scx_ctx* c, uint32_t C) {
  const uint32_t nblk = (C + kBlock - 1) / kBlock;
  if (nblk == 0) return hipSuccess;
  k_sample_targets<<<nblk, kBlock, 0, c->stream>>>(C, c->dim, c->d.calpha, c->d.cbeta, c->d.mlp_s, c->d.mlp_y);
  return hipGetLastError();
}

hipError_t launch_warm_start(scx_ctx* c, uint32_t C, int mode) {
  const uint32_t nblk = (C + kBlock - 1) / kBlock;
  if (nblk == 0) return hipSuccess;
  k_contact_keys<<<nblk, kBlock, 0, c->stream>>>(C, c->d.ca, c->d.cb, c->d.cplane,
                                                 c->nscenes ? c->d.bscene : nullptr, c->d.scene_poff,
                                                 c->d.cur_key);
  // entries written in the other layout are not read (a context that switches modes
  // starts that step cold)
  const int layout = (mode >= SCX_WARM_TRANSPORT) ? 1 : 0;
  const uint32_t P = (layout == c->d.prev_layout) ? c->d.prev_count : 0u;
  k_warm_lookup<<<nblk, kBlock, 0, c->stream>>>(C, c->dim, mode, c->d.cur_key, c->d.prev_key, P, c->d.prev_r,
                                                c->d.basis, c->d.crinit);
  return hipGetLastError();
}

// The cache must hold C entries of kCacheWords (the caller grows it).
hipError_t launch_cache_store(scx_ctx* c, uint32_t C, int mode) {
  const uint32_t nblk = (C + kBlock - 1) / kBlock;
  if (nblk == 0) return hipSuccess;
  k_cache_store<<<nblk, kBlock, 0, c->stream>>>(C, c->dim, mode, c->d.cur_key, c->d.basis, c->d.calpha,
                                                c->d.cbeta, c->d.prev_key, c->d.prev_r);
  return hipGetLastError();
}

hipError_t launch_gather_records(scx_ctx* c, uint32_t C, bool fp32, double mu, const double* mu_arr,
                                 const double* rinit) {
  const uint32_t nblk = (C + kBlock - 1) / kBlock;
  if (nblk == 0) return hipSuccess;
  k_gather<<<nblk, kBlock, 0, c->stream>>>(C, c->dim, c->d.perm, c->d.ca, c->d.cb, c->d.basis, c->d.cla, c->d.clb,
                                           c->d.cwn, c->d.cwt, c->d.cforcing, rinit, mu, mu_arr, c->d.slot_map,
                                           c->d.c64, c->d.c32, fp32 ? 1 : 0);
  return hipGetLastError();
}

}  // namespace scx
