// scx_solver.hip -- graph-coloured projected Gauss-Seidel Coulomb-friction solve
// (the So-bogus GaussSeidel + SOCLaw<D,double,true> path of
// scisim/ConstrainedMaps/bogus/Ball2DSobogusInterface.cpp:102-148 and
// RigidBody3DSobogusInterface.cpp:99-145), Fischer-Burmeister residual, and the
// end-of-step update (GeometricImpactFrictionMap.cpp:441).
//
// One persistent cooperative launch runs the whole solve: per sweep, one pass per
// colour (contacts of a colour share no body, so their updates are independent and
// the result does not depend on scheduling), a grid barrier between colours, and the
// residual (max over contacts of ||FB||^2, in FP64) every eval_every sweeps; all
// workgroups read the same reduced residual after a barrier and leave the loop
// together. The sweep works in primal-velocity form: the body velocities
// v = v0 + M^-1 H^T r are kept current, so a contact's pre-impulse velocity is
// H v + w and no Delassus off-diagonal block is ever formed.
// FMA contraction per source expression only: the backend forms no further fusions, so
// the same expression rounds the same way in every inlined copy (a contact updated as
// the first of its thread's colour pass or as a later one, by the main loop or the tail),
// which keeps the iterates independent of the launch width.
#pragma clang fp contract(on)

#include "scx_device.h"
#include "scx_flow.h"
#include "scx_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <type_traits>

namespace scx {


template <typename Real>
struct Vec4 {
  using type = double4;
};
template <>
struct Vec4<float> {
  using type = float4;
};

template <int D, typename Real>
struct Rec {
  uint32_t a, b;
  Real n[3], s[3], t[3];
  Real la, lb, wn, wt, mu, f;
};

// The rest of the contact frame from n (and s in 3D): t = n x s (3D), t = (-n_y, n_x) (2D).
template <int D, typename Real>
__device__ __forceinline__ void frame_from_ns(Rec<D, Real>& R) {
  if (D == 3) {
    R.t[0] = R.n[1] * R.s[2] - R.n[2] * R.s[1];
    R.t[1] = R.n[2] * R.s[0] - R.n[0] * R.s[2];
    R.t[2] = R.n[0] * R.s[1] - R.n[1] * R.s[0];
  } else {
    R.t[0] = -R.n[1];
    R.t[1] = R.n[0];
    R.t[2] = 0;
    R.s[0] = R.s[1] = R.s[2] = 0;
    R.la = R.lb = 0;
  }
}

template <int D, typename Real>
__device__ __forceinline__ void load_rec(const ContactSoA<Real>& c, uint32_t k, Rec<D, Real>& R) {
  R.a = c.a[k];
  R.b = c.b[k];
#pragma unroll
  for (int d = 0; d < 3; ++d) R.n[d] = c.n[3 * (size_t)k + d];
  if (D == 3) {
#pragma unroll
    for (int d = 0; d < 3; ++d) R.s[d] = c.s[3 * (size_t)k + d];
    R.la = c.la[k];
    R.lb = c.lb[k];
  }
  frame_from_ns<D, Real>(R);
  R.wn = c.wn[k];
  R.wt = c.wt[k];
  R.mu = c.mu[k];
  R.f = c.forcing[k];
}

// Pre-impulse relative velocity in the contact frame, u = H_a v_a - H_b v_b + w,
// computed in precision T from bodies stored in precision Real.
template <int D, typename T, typename Real>
__device__ __forceinline__ void rel_vel(const Rec<D, Real>& R, const typename Vec4<Real>::type& va,
                                        const typename Vec4<Real>::type& wa, const typename Vec4<Real>::type& vb,
                                        const typename Vec4<Real>::type& wb, T u[3]) {
  const T dv[3] = {(T)va.x - (T)vb.x, (T)va.y - (T)vb.y, (T)va.z - (T)vb.z};
  const T n[3] = {(T)R.n[0], (T)R.n[1], (T)R.n[2]};
  const T t[3] = {(T)R.t[0], (T)R.t[1], (T)R.t[2]};
  u[0] = n[0] * dv[0] + n[1] * dv[1] + n[2] * dv[2] + (T)R.f;
  if (D == 3) {
    const T s[3] = {(T)R.s[0], (T)R.s[1], (T)R.s[2]};
    const T la = (T)R.la, lb = (T)R.lb;
    const T ca[3] = {(T)wa.y * n[2] - (T)wa.z * n[1], (T)wa.z * n[0] - (T)wa.x * n[2],
                     (T)wa.x * n[1] - (T)wa.y * n[0]};
    const T cb[3] = {(T)wb.y * n[2] - (T)wb.z * n[1], (T)wb.z * n[0] - (T)wb.x * n[2],
                     (T)wb.x * n[1] - (T)wb.y * n[0]};
    T uv[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) uv[d] = dv[d] - la * ca[d] - lb * cb[d];
    u[1] = s[0] * uv[0] + s[1] * uv[1] + s[2] * uv[2];
    u[2] = t[0] * uv[0] + t[1] * uv[1] + t[2] * uv[2];
  } else {
    u[1] = t[0] * dv[0] + t[1] * dv[1];
    u[2] = 0;
  }
}

// Exact Coulomb solution for the block diag(w_n, w_t, w_t): separate, stick, or slide
// on the cone boundary against the tangential pre-velocity.
template <int D, typename Real>
__device__ __forceinline__ void local_solve(const Rec<D, Real>& R, const Real u[3], const Real r[3], Real rn[3]) {
  const Real bn = u[0] - R.wn * r[0];
  const Real n0 = -bn / R.wn;
  if (!(n0 > (Real)0)) {
    rn[0] = rn[1] = rn[2] = 0;
    return;
  }
  rn[0] = n0;
  Real nrm2 = 0;
#pragma unroll
  for (int a = 1; a < D; ++a) {
    rn[a] = -(u[a] - R.wt * r[a]) / R.wt;
    nrm2 += rn[a] * rn[a];
  }
  if (D == 2) rn[2] = 0;
  const Real lim = R.mu * n0;
  const Real nrm = sqrt(nrm2);
  if (nrm > lim) {
    const Real sc = (nrm > (Real)0) ? lim / nrm : (Real)0;
#pragma unroll
    for (int a = 1; a < D; ++a) rn[a] *= sc;
  }
}

// FB residual of one contact (De Saxce change of variable), ||phi||^2, in precision T
// (FP64 for every decision on the tolerance; FP32 only for the estimate of the FP32 phase).
template <int D, typename T = double>
__device__ __forceinline__ T fb_err(T mu, const T* r, const T* u) {
  T uT2 = (T)0;
#pragma unroll
  for (int a = 1; a < D; ++a) uT2 += u[a] * u[a];
  const T un_hat = u[0] + mu * sqrt(uT2);
  T xh[3], yh[3];
  xh[0] = mu * r[0];
  yh[0] = un_hat;
#pragma unroll
  for (int a = 1; a < D; ++a) {
    xh[a] = r[a];
    yh[a] = mu * u[a];
  }
  T z0 = (T)0;
#pragma unroll
  for (int a = 0; a < D; ++a) z0 += xh[a] * xh[a] + yh[a] * yh[a];
  T z1[2] = {(T)0, (T)0}, nz1 = (T)0;
#pragma unroll
  for (int a = 1; a < D; ++a) {
    z1[a - 1] = (T)2 * (xh[0] * xh[a] + yh[0] * yh[a]);
    nz1 += z1[a - 1] * z1[a - 1];
  }
  nz1 = sqrt(nz1);
  const T s1 = sqrt(z0 + nz1);
  const T s2 = sqrt(fmax((T)0, z0 - nz1));
  T e = xh[0] + yh[0] - (T)0.5 * (s1 + s2);
  e = e * e;
#pragma unroll
  for (int a = 1; a < D; ++a) {
    const T dir = nz1 > (T)0 ? z1[a - 1] / nz1 : (T)0;
    const T p = xh[a] + yh[a] - (T)0.5 * (s1 - s2) * dir;
    e += p * p;
  }
  // a NaN residual (e.g. from non-finite velocities) counts as infinitely far from the
  // solution: the max reductions drop NaN operands, so it must never read as converged
  return __builtin_isnan(e) ? (T)INFINITY : e;
}

// Body records per slot: 2D one vector [v (3) | minv], 3D two adjacent vectors
// [v (3) | minv][omega (3) | Iinv].
template <int D>
constexpr int kBodyVecs = (D == 3) ? 2 : 1;

// Coherent accesses to the body records for the colour passes between grid_sync_coherent
// barriers: buffer loads and stores with the agent-scope coherence bit (SC1), which go to
// the memory side past the per-XCD L2s, so that a record another workgroup updated in the
// previous pass is seen without an L2 invalidate. The records are addressed by their
// 16-byte (float) or 32-byte (double) vectors; the buffer covers nbodies records (the host
// keeps it below 2 GiB).
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
constexpr int kCoherentAux = 16;  // cache policy bits: SC1

template <typename Real>
struct BodyPort;

template <>
struct BodyPort<float> {
  __amdgpu_buffer_rsrc_t rs;
  __device__ __forceinline__ float4 ld(size_t v) const {
    const u32x4 x = __builtin_amdgcn_raw_buffer_load_b128(rs, (uint32_t)v * 16u, 0, kCoherentAux);
    return make_float4(__uint_as_float(x[0]), __uint_as_float(x[1]), __uint_as_float(x[2]), __uint_as_float(x[3]));
  }
  __device__ __forceinline__ void st(size_t v, const float4& f) const {
    const u32x4 x = {__float_as_uint(f.x), __float_as_uint(f.y), __float_as_uint(f.z), __float_as_uint(f.w)};
    __builtin_amdgcn_raw_buffer_store_b128(x, rs, (uint32_t)v * 16u, 0, kCoherentAux);
  }
};

template <>
struct BodyPort<double> {
  __amdgpu_buffer_rsrc_t rs;
  static __device__ __forceinline__ double join(uint32_t lo, uint32_t hi) {
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
  }
  __device__ __forceinline__ double4 ld(size_t v) const {
    const u32x4 a = __builtin_amdgcn_raw_buffer_load_b128(rs, (uint32_t)v * 32u, 0, kCoherentAux);
    const u32x4 b = __builtin_amdgcn_raw_buffer_load_b128(rs, (uint32_t)v * 32u + 16u, 0, kCoherentAux);
    return make_double4(join(a[0], a[1]), join(a[2], a[3]), join(b[0], b[1]), join(b[2], b[3]));
  }
  __device__ __forceinline__ void st(size_t v, const double4& d) const {
    const unsigned long long x = (unsigned long long)__double_as_longlong(d.x);
    const unsigned long long y = (unsigned long long)__double_as_longlong(d.y);
    const unsigned long long z = (unsigned long long)__double_as_longlong(d.z);
    const unsigned long long w = (unsigned long long)__double_as_longlong(d.w);
    const u32x4 a = {(uint32_t)x, (uint32_t)(x >> 32), (uint32_t)y, (uint32_t)(y >> 32)};
    const u32x4 b = {(uint32_t)z, (uint32_t)(z >> 32), (uint32_t)w, (uint32_t)(w >> 32)};
    __builtin_amdgcn_raw_buffer_store_b128(a, rs, (uint32_t)v * 32u, 0, kCoherentAux);
    __builtin_amdgcn_raw_buffer_store_b128(b, rs, (uint32_t)v * 32u + 16u, 0, kCoherentAux);
  }
};

template <int D, typename Real>
__device__ __forceinline__ BodyPort<Real> body_port(Real* body, uint32_t nbodies) {
  BodyPort<Real> P;
  const uint32_t bytes = nbodies * (uint32_t)(kBodyVecs<D> * 4 * sizeof(Real));
  P.rs = __builtin_amdgcn_make_buffer_rsrc(body, (short)0, (int)bytes, 0x00020000);
  return P;
}

// Coherent (SC1) load and store of one impulse component: with coherent colour passes the
// impulses are read by the residual evaluation in another thread's mapping, so they take
// the same path to memory as the body records.
template <bool Coh, typename Real>
__device__ __forceinline__ Real ld_imp(const Real* p) {
  return Coh ? __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : *p;
}
template <bool Coh, typename Real>
__device__ __forceinline__ void st_imp(Real* p, Real x) {
  if (Coh) __hip_atomic_store(p, x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  else *p = x;
}


template <typename Real>
struct GSArgs {
  ContactSoA<Real> c;
  Real* body;               // body records by slot
  const uint32_t* sid;      // slot -> body id (nullptr: identity)
  const double* v0;         // dim per body
  const double* w0;         // 3 per body (3D)
  uint32_t nbodies;
  uint32_t C;
  const uint32_t* coloff;   // colour offsets
  Real* rbest;              // 3 per contact
  double* partial;          // per-workgroup residual partials
  StepCounters* ctr;
  uint32_t* bar;             // grid barrier state
  uint32_t max_iters;
  uint32_t eval_every;
  double tol;
  int phase;                // 0: only phase, 1: FP32 phase, 2: FP64 polish after FP32
  int apply_init;           // reset velocities to v0 and apply the records' impulses
  uint32_t* lmap;           // per body slot: owning endpoint in the LDS solve (all ones when idle)
  uint32_t tail_col;        // colours [tail_col, ncolours) are swept by workgroup 0 alone (small)
  // ensembles (else cscene == nullptr): per contact (colour order) its scene; per scene the
  // residual of the last two evaluations (max ||FB||^2 bits, double-buffered) and the sweep
  // at which it met the tolerance (kNoBody until then). A scene that meets the tolerance is
  // frozen: its contacts are skipped until the solve ends.
  const uint32_t* cscene;
  unsigned long long* scene_acc;  // 2 x nscenes
  uint32_t* scene_sweeps;         // nscenes
  uint32_t nscenes;
  // ensembles: the contacts of scenes not yet frozen, compacted colour by colour once
  // enough scenes have frozen (two lists of C entries, alternately rebuilt), and per-colour
  // per-workgroup counts for the compaction (kMaxColours x gridDim); nullptr: no compaction
  uint32_t* act_idx;
  uint32_t* act_part;
  // FP32 phase of one scene: contacts whose residual at the last full evaluation was at
  // least the capture threshold (up to kViolCap, in any order; counts of the two last
  // evaluations, double-buffered), and the word through which workgroup 0 tells the others
  // that one of them is still above the stopping value (nullptr: every sweep is evaluated)
  uint32_t* viol;  // 2 counts, then kViolCap contacts
  uint32_t* vflag;
  // One scene with coherent passes on several workgroups: colours [comp_col, ncolours) are
  // swept as connected components (contacts linked through shared bodies), every component
  // by one workgroup through its L2. comp_parent: per body slot, its parent in a union-find
  // forest over the bodies of those contacts; comp_cnt: an overflow flag, then per
  // workgroup the number of contacts it was given; comp_list: per workgroup kCompCap
  // contacts. comp_parent == nullptr: no component phase.
  uint32_t* comp_parent;
  uint32_t* comp_cnt;
  uint32_t* comp_list;
  uint32_t comp_col;
  int comp_lds;  // hold a workgroup's components in LDS when they fit (else through its L2)
};


// True if contact k belongs to a scene whose last evaluated residual is below stop.
template <typename Real>
__device__ __forceinline__ bool frozen(const GSArgs<Real>& A, uint32_t k, uint32_t eb, double stop) {
  if (!A.cscene) return false;
  const unsigned long long bits =
      __hip_atomic_load(A.scene_acc + (size_t)eb * A.nscenes + A.cscene[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return __longlong_as_double((long long)bits) < stop;
}

// Frozen scenes as a bit mask in LDS, rebuilt by every workgroup after each evaluation,
// so that the sweeps test a contact with one LDS read (ensembles of up to kSceneMaskBits
// scenes; larger ones read the residuals from global memory).
constexpr uint32_t kSceneMaskBits = 16384;

template <typename Real>
__device__ __forceinline__ void build_frozen_mask(const GSArgs<Real>& A, uint32_t eb, double stop, uint32_t* mask) {
  // one scene per thread, one 64-bit ballot per wave (blockDim is a multiple of 64)
  const uint32_t words = (A.nscenes + 31) / 32;
  for (uint32_t b0 = 0; b0 < A.nscenes; b0 += blockDim.x) {
    const uint32_t sc = b0 + threadIdx.x;
    bool f = false;
    if (sc < A.nscenes) {
      const unsigned long long v = __hip_atomic_load(A.scene_acc + (size_t)eb * A.nscenes + sc, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT);
      f = __longlong_as_double((long long)v) < stop;
    }
    const unsigned long long bal = __ballot(f);
    if ((threadIdx.x & 63) == 0) {
      const uint32_t w = (b0 + threadIdx.x) >> 5;
      if (w < words) mask[w] = (uint32_t)bal;
      if (w + 1 < words) mask[w + 1] = (uint32_t)(bal >> 32);
    }
  }
  __syncthreads();
}

template <typename Real>
__device__ __forceinline__ bool frozen_fast(const GSArgs<Real>& A, uint32_t k, uint32_t eb, double stop,
                                            const uint32_t* mask) {
  if (!A.cscene) return false;
  if (A.nscenes > kSceneMaskBits) return frozen(A, k, eb, stop);
  const uint32_t sc = A.cscene[k];
  return (mask[sc >> 5] >> (sc & 31)) & 1u;
}

// A frozen contact keeps its impulse; the lazy best-iterate copy still stores it.
template <bool Coh = false, typename Real>
__device__ __forceinline__ void keep(const GSArgs<Real>& A, uint32_t k, bool copy_best) {
  if (!copy_best) return;
#pragma unroll
  for (int d = 0; d < 3; ++d) A.rbest[3 * (size_t)k + d] = ld_imp<Coh>(A.c.r + 3 * (size_t)k + d);
}

// Frozen scenes in the mask (every workgroup computes the same count).
__device__ __forceinline__ uint32_t count_frozen(uint32_t nscenes, const uint32_t* mask, uint32_t* s_tmp) {
  uint32_t n = 0;
  for (uint32_t w = threadIdx.x; w < (nscenes + 31) / 32; w += blockDim.x) n += __popc(mask[w]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) n += __shfl_xor(n, off, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) s_tmp[threadIdx.x >> 6] = n;
  __syncthreads();
  uint32_t t = 0;
  for (uint32_t w = 0; w < blockDim.x / 64; ++w) t += s_tmp[w];
  __syncthreads();
  return t;
}

// Ensembles: the contacts of the scenes not frozen in `mask`, compacted colour by colour
// from the list src (nullptr: every contact) with colour offsets soff into dst, with its
// colour offsets in doff (LDS, ncol + 1 entries). Each workgroup takes one contiguous
// chunk of each colour and keeps the order of its contacts (spatial within a colour), so
// the chunks land in order too. Two grid barriers: after the per-chunk counts and after
// the writes. All threads of the grid take part.
template <typename Real>
__device__ void compact_active(const GSArgs<Real>& A, const uint32_t* src, const uint32_t* soff, uint32_t ncol,
                               const uint32_t* mask, uint32_t* dst, uint32_t* doff) {
  __shared__ uint32_t s_cnt[kMaxColours];  // this workgroup's chunk count, then the colour totals
  __shared__ uint32_t s_pre[kMaxColours];  // active contacts of the chunks before this one
  __shared__ uint32_t s_wave[kGSBlock / 64];
  const uint32_t G = gridDim.x, b = blockIdx.x;
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  auto active = [&](uint32_t j) {
    const uint32_t sc = A.cscene[src ? src[j] : j];
    return ((mask[sc >> 5] >> (sc & 31)) & 1u) == 0u;
  };
  for (uint32_t c = threadIdx.x; c < kMaxColours; c += kGSBlock) {
    s_cnt[c] = 0u;
    s_pre[c] = 0u;
  }
  __syncthreads();
  for (uint32_t col = 0; col < ncol; ++col) {
    const uint32_t n = soff[col + 1] - soff[col], per = (n + G - 1) / G;
    const uint32_t lo = soff[col] + min(n, b * per), hi = soff[col] + min(n, (b + 1) * per);
    for (uint32_t j0 = lo; j0 < hi; j0 += kGSBlock) {
      const uint32_t j = j0 + threadIdx.x;
      const unsigned long long bal = __ballot(j < hi && active(j));
      if (lane == 0 && bal) atomicAdd(s_cnt + col, (uint32_t)__popcll(bal));
    }
  }
  __syncthreads();
  for (uint32_t col = threadIdx.x; col < ncol; col += kGSBlock) A.act_part[(size_t)col * G + b] = s_cnt[col];
  grid_barrier(A.bar);
  for (uint32_t c = threadIdx.x; c < kMaxColours; c += kGSBlock) s_cnt[c] = 0u;
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < ncol * G; i += kGSBlock) {
    const uint32_t v = A.act_part[i], col = i / G;
    if (v) {
      atomicAdd(s_cnt + col, v);
      if (i % G < b) atomicAdd(s_pre + col, v);
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    doff[0] = 0u;
    for (uint32_t col = 0; col < ncol; ++col) doff[col + 1] = doff[col] + s_cnt[col];
  }
  __syncthreads();
  for (uint32_t col = 0; col < ncol; ++col) {
    const uint32_t n = soff[col + 1] - soff[col], per = (n + G - 1) / G;
    const uint32_t lo = soff[col] + min(n, b * per), hi = soff[col] + min(n, (b + 1) * per);
    uint32_t pos = doff[col] + s_pre[col];
    for (uint32_t j0 = lo; j0 < hi; j0 += kGSBlock) {
      const uint32_t j = j0 + threadIdx.x;
      const bool a = j < hi && active(j);
      const unsigned long long bal = __ballot(a);
      if (lane == 0) s_wave[wave] = (uint32_t)__popcll(bal);
      __syncthreads();
      uint32_t before = 0, total = 0;
      for (uint32_t w = 0; w < kGSBlock / 64; ++w) {
        before += (w < wave) ? s_wave[w] : 0u;
        total += s_wave[w];
      }
      if (a) dst[pos + before + (uint32_t)__popcll(bal & ((1ull << lane) - 1ull))] = src ? src[j] : j;
      pos += total;
      __syncthreads();
    }
  }
  grid_barrier(A.bar);
}

// A contact's record and impulse, loaded ahead of its colour pass. Only the thread that
// updates a contact ever writes its impulse, and the record is read-only during the
// solve, so the load may be issued before the grid barrier that opens the pass.
template <int D, typename Real>
struct Loaded {
  Rec<D, Real> R;
  Real r[3];
};

template <int D, typename Real, bool Coh = false>
__device__ __forceinline__ void load_contact(const GSArgs<Real>& A, uint32_t k, Loaded<D, Real>& L) {
  load_rec<D, Real>(A.c, k, L.R);
  const Real* rp = A.c.r + 3 * (size_t)k;
  L.r[0] = ld_imp<Coh>(rp);
  L.r[1] = ld_imp<Coh>(rp + 1);
  L.r[2] = ld_imp<Coh>(rp + 2);
}

// One contact of a colour pass from its loaded record, on body vectors ia and ib (3D: ia + 1
// and ib + 1 are the angular halves of the same records) read and written through ld/st,
// the new impulse stored through st_r(d, x). copy_best: first store the contact's impulse
// as it was before this sweep into rbest (the previous sweep's iterate was the best one so
// far). Every update of a contact goes through this code, wherever its bodies are held.
template <int D, typename Real, typename Ld, typename St, typename StR>
__device__ __forceinline__ void update_core(const GSArgs<Real>& A, uint32_t k, const Loaded<D, Real>& L, bool init,
                                            bool copy_best, size_t ia, size_t ib, Ld&& ld, St&& st, StR&& st_r) {
  using V = typename Vec4<Real>::type;
  const Rec<D, Real>& R = L.R;
  const bool hasb = R.b != kNoBody;
  V va = ld(ia), wa = {0, 0, 0, 0}, vb = {0, 0, 0, 0}, wb = {0, 0, 0, 0};
  if (D == 3) wa = ld(ia + 1);
  if (hasb) {
    vb = ld(ib);
    if (D == 3) wb = ld(ib + 1);
  }
  const Real r[3] = {L.r[0], L.r[1], L.r[2]};
  if (copy_best) {
    Real* rb = A.rbest + 3 * (size_t)k;
    rb[0] = r[0];
    rb[1] = r[1];
    rb[2] = r[2];
  }
  Real dr[3];
  if (init) {
    dr[0] = r[0];
    dr[1] = r[1];
    dr[2] = r[2];
  } else {
    Real u[3], rn[3];
    rel_vel<D, Real, Real>(R, va, wa, vb, wb, u);
    local_solve<D, Real>(R, u, r, rn);
    dr[0] = rn[0] - r[0];
    dr[1] = rn[1] - r[1];
    dr[2] = rn[2] - r[2];
    st_r(0, rn[0]);
    st_r(1, rn[1]);
    if (D == 3) st_r(2, rn[2]);
  }
  Real Pi[3];  // the impulse increment in world coordinates
#pragma unroll
  for (int d = 0; d < 3; ++d) Pi[d] = R.n[d] * dr[0] + R.t[d] * dr[D - 1] + ((D == 3) ? R.s[d] * dr[1] : (Real)0);
  va.x += va.w * Pi[0];
  va.y += va.w * Pi[1];
  va.z += va.w * Pi[2];
  st(ia, va);
  Real tau[3];
  if (D == 3) {
#pragma unroll
    for (int d = 0; d < 3; ++d) tau[d] = dr[1] * R.t[d] - dr[2] * R.s[d];
    const Real ka = -R.la * wa.w;
    wa.x += ka * tau[0];
    wa.y += ka * tau[1];
    wa.z += ka * tau[2];
    st(ia + 1, wa);
  }
  if (hasb) {
    vb.x -= vb.w * Pi[0];
    vb.y -= vb.w * Pi[1];
    vb.z -= vb.w * Pi[2];
    st(ib, vb);
    if (D == 3) {
      const Real kb = R.lb * wb.w;
      wb.x -= kb * tau[0];
      wb.y -= kb * tau[1];
      wb.z -= kb * tau[2];
      st(ib + 1, wb);
    }
  }
}

// One contact of a colour pass on the body records in global memory (Coh: coherently).
template <int D, typename Real, bool Coh = false>
__device__ __forceinline__ void update(const GSArgs<Real>& A, uint32_t k, const Loaded<D, Real>& L, bool init,
                                       bool copy_best, const BodyPort<Real>& P = BodyPort<Real>{}) {
  using V = typename Vec4<Real>::type;
  V* bv = reinterpret_cast<V*>(A.body);
  Real* rp = A.c.r + 3 * (size_t)k;
  update_core<D, Real>(
      A, k, L, init, copy_best, (size_t)kBodyVecs<D> * L.R.a, (size_t)kBodyVecs<D> * L.R.b,
      [&](size_t i) { return Coh ? P.ld(i) : bv[i]; },
      [&](size_t i, const V& x) {
        if (Coh) P.st(i, x);
        else bv[i] = x;
      },
      [&](int d, Real x) { st_imp<Coh>(rp + d, x); });
}

// One contact of a colour pass (record loaded here).
template <int D, typename Real, bool Coh = false>
__device__ __forceinline__ void process(const GSArgs<Real>& A, uint32_t k, bool init, bool copy_best = false,
                                        const BodyPort<Real>& P = BodyPort<Real>{}) {
  Loaded<D, Real> L;
  load_contact<D, Real, Coh>(A, k, L);
  update<D, Real, Coh>(A, k, L, init, copy_best, P);
}

template <int D, typename Real, typename T = double, bool Coh = false>
__device__ __forceinline__ T contact_err(const GSArgs<Real>& A, uint32_t k, const BodyPort<Real>& P = BodyPort<Real>{}) {
  using V = typename Vec4<Real>::type;
  Rec<D, Real> R;
  load_rec<D, Real>(A.c, k, R);
  const V* bv = reinterpret_cast<const V*>(A.body);
  const size_t ia = (size_t)kBodyVecs<D> * R.a, ib = (size_t)kBodyVecs<D> * R.b;
  auto ld = [&](size_t i) { return Coh ? P.ld(i) : bv[i]; };
  const V z = {0, 0, 0, 0};
  const V va = ld(ia);
  const V wa = (D == 3) ? ld(ia + 1) : z;
  const V vb = (R.b != kNoBody) ? ld(ib) : z;
  const V wb = (D == 3 && R.b != kNoBody) ? ld(ib + 1) : z;
  T u[3];
  rel_vel<D, T, Real>(R, va, wa, vb, wb, u);
  const Real* rp = A.c.r + 3 * (size_t)k;
  const T r[3] = {(T)ld_imp<Coh>(rp), (T)ld_imp<Coh>(rp + 1), (T)ld_imp<Coh>(rp + 2)};
  return fb_err<D, T>((T)R.mu, r, u);
}

// Scenes whose per-scene residual max is combined in LDS during an evaluation.
constexpr uint32_t kSceneLdsMax = 2048;

// Per-scene max of one contact's residual (ensembles): lanes of a wave that share a scene
// (the usual case: contacts are in spatial order within a colour) combine before one atomic.
__device__ __forceinline__ void scene_max(unsigned long long* acc, uint32_t scene, double e, bool valid) {
  const unsigned long long bits = valid ? (unsigned long long)__double_as_longlong(e) : 0ull;
  const uint32_t first = __shfl(valid ? scene : 0xFFFFFFFFu, __ffsll((long long)__ballot(valid)) - 1, 64);
  const bool same = __all(!valid || scene == first);
  if (same) {
    unsigned long long m = bits;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const unsigned long long o = __shfl_xor(m, off, 64);
      m = o > m ? o : m;
    }
    if ((threadIdx.x & 63) == 0 && __ballot(valid) != 0) atomicMax(acc + first, m);
  } else if (valid) {
    atomicMax(acc + scene, bits);
  }
}

// Grid-wide max of the residual; identical in every thread after the barrier. Ensembles:
// also the per-scene max into buffer ev & 1 (the other buffer is cleared for the next
// evaluation), and the sweep at which each scene meets `stop` (it sweeps done in this
// launch, base those of earlier phases).
// Coh: after coherent colour passes; the body records and impulses are read coherently and
// the reduction's barrier is grid_sync_coherent.
template <int D, typename Real, bool Coh = false>
__device__ double eval_residual(const GSArgs<Real>& A, uint32_t gtid, uint32_t gsize, uint32_t ev = 0,
                                double stop = 0.0, uint32_t base = 0, uint32_t it = 0,
                                const uint32_t* lst = nullptr, uint32_t nlst = 0,
                                const BodyPort<Real>& P = BodyPort<Real>{}, double capture = -1.0) {
  // ensembles after a compaction: only the contacts of scenes not frozen are evaluated (a
  // frozen scene met the tolerance and is no longer updated; its buffer entry stays 0)
  const uint32_t nk = lst ? nlst : A.C;
  __shared__ double s_red[kGSBlock / 64];
  double m = 0.0;
  if (A.cscene) {
    unsigned long long* acc = A.scene_acc + (size_t)(ev & 1) * A.nscenes;
    unsigned long long* other = A.scene_acc + (size_t)((ev + 1) & 1) * A.nscenes;
    for (uint32_t sc = gtid; sc < A.nscenes; sc += gsize)
      __hip_atomic_store(other + sc, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (A.nscenes <= kSceneLdsMax) {
      // each workgroup takes one contiguous range of contacts (in a colour, contacts are in
      // spatial order, so a range touches few scenes), combines per scene in LDS, then
      // publishes the scenes it touched
      __shared__ unsigned long long s_acc[kSceneLdsMax];
      for (uint32_t sc = threadIdx.x; sc < A.nscenes; sc += kGSBlock) s_acc[sc] = 0ull;
      __syncthreads();
      const uint32_t per = (nk + gridDim.x - 1) / gridDim.x;
      const uint32_t lo = min(nk, blockIdx.x * per), hi = min(nk, lo + per);
#pragma unroll 4
      for (uint32_t j = lo + threadIdx.x; j < hi; j += kGSBlock) {
        const uint32_t k = lst ? lst[j] : j;
        const double e = (A.phase == 1) ? (double)contact_err<D, Real, float, Coh>(A, k, P) : contact_err<D, Real, double, Coh>(A, k, P);
        m = fmax(m, e);
        if (e > 0.0) atomicMax(s_acc + A.cscene[k], (unsigned long long)__double_as_longlong(e));
      }
      __syncthreads();
      for (uint32_t sc = threadIdx.x; sc < A.nscenes; sc += kGSBlock)
        if (s_acc[sc] != 0ull) atomicMax(acc + sc, s_acc[sc]);
    } else {
      const uint32_t nkw = (nk + gsize - 1) / gsize * gsize;  // whole waves take part in scene_max
      for (uint32_t j = gtid; j < nkw; j += gsize) {
        const bool valid = j < nk;
        const uint32_t k = valid ? (lst ? lst[j] : j) : 0u;
        double e = 0.0;
        if (valid) e = (A.phase == 1) ? (double)contact_err<D, Real, float, Coh>(A, k, P) : contact_err<D, Real, double, Coh>(A, k, P);
        m = fmax(m, e);
        scene_max(acc, valid ? A.cscene[k] : 0u, e, valid);
      }
    }
  } else if (A.phase == 1) {  // FP32 phase: an FP32 estimate (the FP64 phase decides on the tolerance)
    if (capture >= 0.0) {
      // also list the contacts at or above `capture` (the violators that the sweeps after
      // this evaluation check first); the other buffer's count starts the next list
      if (gtid == 0) __hip_atomic_store(A.viol + ((ev + 1) & 1), 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll 4
      for (uint32_t k = gtid; k < A.C; k += gsize) {
        const double e = (double)contact_err<D, Real, float, Coh>(A, k, P);
        m = fmax(m, e);
        if (e >= capture) {
          const uint32_t i = atomicAdd(A.viol + (ev & 1), 1u);
          if (i < kViolCap) __hip_atomic_store(A.viol + 2 + i, k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
      }
    } else {
      // read-only loop: unrolled so that the loads of several contacts are in flight together
#pragma unroll 4
      for (uint32_t k = gtid; k < A.C; k += gsize) m = fmax(m, (double)contact_err<D, Real, float, Coh>(A, k, P));
    }
  } else if (capture >= 0.0) {
    // FP64: the same list
    if (gtid == 0) __hip_atomic_store(A.viol + ((ev + 1) & 1), 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll 4
    for (uint32_t k = gtid; k < A.C; k += gsize) {
      const double e = contact_err<D, Real, double, Coh>(A, k, P);
      m = fmax(m, e);
      if (e >= capture) {
        const uint32_t i = atomicAdd(A.viol + (ev & 1), 1u);
        if (i < kViolCap) __hip_atomic_store(A.viol + 2 + i, k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
  } else {
#pragma unroll 4
    for (uint32_t k = gtid; k < A.C; k += gsize) m = fmax(m, contact_err<D, Real, double, Coh>(A, k, P));
  }
  m = wave_reduce_max_f64(m);
  if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = s_red[0];
    for (int w = 1; w < kGSBlock / 64; ++w) b = fmax(b, s_red[w]);
    __hip_atomic_store(A.partial + blockIdx.x, b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  if (Coh) grid_sync_coherent(A.bar);
  else grid_barrier(A.bar);
  if (A.cscene) {
    unsigned long long* acc = A.scene_acc + (size_t)(ev & 1) * A.nscenes;
    for (uint32_t sc = gtid; sc < A.nscenes; sc += gsize) {
      const double es = __longlong_as_double(
          (long long)__hip_atomic_load(acc + sc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
      if (es < stop) {
        if (A.scene_sweeps[sc] == kNoBody) A.scene_sweeps[sc] = base + it;
      } else if (A.phase == 2) {
        A.scene_sweeps[sc] = kNoBody;  // an FP32 estimate below the tolerance was not confirmed
      }
    }
  }
  double e = 0.0;
  for (uint32_t b = threadIdx.x; b < gridDim.x; b += kGSBlock)
    e = fmax(e, __hip_atomic_load(A.partial + b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
  e = wave_reduce_max_f64(e);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = e;
  __syncthreads();
  double out = s_red[0];
  for (int w = 1; w < kGSBlock / 64; ++w) out = fmax(out, s_red[w]);
  __syncthreads();
  return out;
}

// v = v0 + M^-1 H^T r for the current impulses, applied colour by colour. Only bodies
// touched by a contact can differ from v0 (every body starts at v0, k_body_init), so
// a reset touches those bodies only; two contacts sharing a body write the same v0.
template <int D, typename Real>
__device__ void apply_impulses(const GSArgs<Real>& A, const uint32_t* off, uint32_t ncol,
                               uint32_t gtid, uint32_t gsize, bool reset) {
  if (reset) {
    for (uint32_t k = gtid; k < A.C; k += gsize) {
      const uint32_t bodies[2] = {A.c.a[k], A.c.b[k]};
      for (int q = 0; q < 2; ++q) {
        const uint32_t sl = bodies[q];
        if (sl == kNoBody) continue;
        const uint32_t i = A.sid ? A.sid[sl] : sl;
        Real* rec = A.body + 4 * kBodyVecs<D> * (size_t)sl;
#pragma unroll
        for (int d = 0; d < D; ++d) rec[d] = (Real)A.v0[(size_t)D * i + d];
        if (D == 3)
#pragma unroll
          for (int d = 0; d < 3; ++d) rec[4 + d] = (Real)A.w0[3 * (size_t)i + d];
      }
    }
    grid_barrier(A.bar);
  }
  for (uint32_t col = 0; col < ncol; ++col) {
    for (uint32_t k = off[col] + gtid; k < off[col + 1]; k += gsize) process<D, Real>(A, k, true);
    grid_barrier(A.bar);
  }
}

// Union-find over body slots for the component phase. Every access to the parent array is
// an agent-scope atomic (the workgroups of all XCDs share it). A parent is always a smaller
// slot: a root is linked only under a smaller root (compare-and-swap, which fails when
// another thread linked it first), and path halving replaces a parent by the parent's parent,
// another ancestor; so the forest stays acyclic under any interleaving, and once every link
// is made (after a grid barrier) a find returns the component's smallest slot.
__device__ __forceinline__ uint32_t uf_ld(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t uf_find(uint32_t* parent, uint32_t x) {
  for (;;) {
    const uint32_t p = uf_ld(parent + x);
    if (p == x) return x;
    const uint32_t gp = uf_ld(parent + p);
    if (gp == p) return p;
    __hip_atomic_store(parent + x, gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    x = gp;
  }
}
__device__ __forceinline__ void uf_union(uint32_t* parent, uint32_t a, uint32_t b) {
  for (;;) {
    a = uf_find(parent, a);
    b = uf_find(parent, b);
    if (a == b) return;
    if (a > b) {
      const uint32_t t = a;
      a = b;
      b = t;
    }
    uint32_t expect = b;
    if (__hip_atomic_compare_exchange_strong(parent + b, &expect, a, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT))
      return;
  }
}

// Component phase set-up, once per launch (one scene, coherent passes, several workgroups):
// the contacts of colours [first, ncol) are grouped into connected components through their
// bodies, every component is given to the workgroup hash(root) % gridDim.x, and each
// workgroup lists its contacts in LDS sorted by colour (s_list, colour offsets s_coff).
// The order inside a colour is free: contacts of one colour share no body. Returns `first`,
// or ncol (no component phase) when some workgroup was given more than kCompCap contacts.
template <int D, typename Real>
__device__ uint32_t component_setup(const GSArgs<Real>& A, const uint32_t* off, uint32_t ncol, uint32_t first,
                                    uint32_t gtid, uint32_t gsize, uint32_t* s_list, uint32_t* s_coff) {
  uint32_t* par = A.comp_parent;
  const uint32_t k0 = off[first], k1 = off[ncol];
  for (uint32_t k = k0 + gtid; k < k1; k += gsize) {
    const uint32_t a = A.c.a[k], b = A.c.b[k];
    __hip_atomic_store(par + a, a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (b != kNoBody) __hip_atomic_store(par + b, b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  for (uint32_t w = gtid; w <= gridDim.x; w += gsize)
    __hip_atomic_store(A.comp_cnt + w, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  grid_sync_coherent(A.bar);
  for (uint32_t k = k0 + gtid; k < k1; k += gsize)
    if (A.c.b[k] != kNoBody) uf_union(par, A.c.a[k], A.c.b[k]);
  grid_sync_coherent(A.bar);
  for (uint32_t k = k0 + gtid; k < k1; k += gsize) {
    const uint32_t w = fmix32(uf_find(par, A.c.a[k])) % gridDim.x;
    const uint32_t pos = __hip_atomic_fetch_add(A.comp_cnt + 1 + w, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (pos < kCompCap)
      __hip_atomic_store(A.comp_list + (size_t)w * kCompCap + pos, k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else
      __hip_atomic_store(A.comp_cnt, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  grid_sync_coherent(A.bar);
  if (__hip_atomic_load(A.comp_cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) {
    if (gtid == 0) ++A.ctr->comp_overflows;
    return ncol;
  }
  // counting sort of this workgroup's contacts by colour
  const uint32_t n = __hip_atomic_load(A.comp_cnt + 1 + blockIdx.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const uint32_t* src = A.comp_list + (size_t)blockIdx.x * kCompCap;
  __shared__ uint32_t s_pos[kMaxColours + 1];
  for (uint32_t c = threadIdx.x; c <= kMaxColours; c += kGSBlock) s_pos[c] = 0u;
  __syncthreads();
  // colour of contact k: the last c in [first, ncol) with off[c] <= k
  auto colour_of = [&](uint32_t k) {
    uint32_t lo = first, hi = ncol - 1;
    while (lo < hi) {
      const uint32_t mid = (lo + hi + 1) >> 1;
      if (off[mid] <= k) lo = mid;
      else hi = mid - 1;
    }
    return lo;
  };
  for (uint32_t j = threadIdx.x; j < n; j += kGSBlock)
    atomicAdd(s_pos + colour_of(__hip_atomic_load(src + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)), 1u);
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t sum = 0;
    for (uint32_t c = first; c < ncol; ++c) {
      const uint32_t x = s_pos[c];
      s_coff[c] = s_pos[c] = sum;
      sum += x;
    }
    s_coff[ncol] = sum;
  }
  __syncthreads();
  for (uint32_t j = threadIdx.x; j < n; j += kGSBlock) {
    const uint32_t k = __hip_atomic_load(src + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s_list[atomicAdd(s_pos + colour_of(k), 1u)] = k;
  }
  if (threadIdx.x == 0) {
    // the widest colour of this workgroup's list (for the choice of the phase's colours)
    uint32_t w = 0;
    for (uint32_t c = first; c < ncol; ++c) w = max(w, s_coff[c + 1] - s_coff[c]);
    atomicMax(&A.ctr->comp_maxw, w);
  }
  __syncthreads();
  return first;
}

// A workgroup's components held in LDS for the whole launch when its list has at most
// kCompLds contacts touching at most kCompLds bodies: the contact records and impulses of
// its contacts (only this workgroup updates those impulses; every new value is also stored
// coherently for the evaluations) and the body records, loaded from memory at the start of
// the phase of every sweep and stored back at its end, both coherently. A colour of the
// phase is then LDS traffic and a workgroup barrier. FP32 sweeps hold 1024 contacts, the
// FP64 polish 512 (the same LDS bytes).
template <typename Real>
constexpr uint32_t kCompLds = (sizeof(Real) == 4) ? 1024u : 512u;
template <int D>
constexpr int kRecWords = (D == 3) ? 12 : 7;  // n (3), s (3), la, lb (3D), wn, wt, mu, f
constexpr uint16_t kNoLocal = 0xFFFF;

template <int D, typename Real>
struct CompLds {
  using V = typename Vec4<Real>::type;
  V* body;          // kCompLds x kBodyVecs: the records of the bodies held
  Real* r;          // 3 per listed contact: its impulse
  Real* rec;        // kRecWords per listed contact: its record
  uint16_t* eidx;   // 2 per listed contact: the LDS body of each side (kNoLocal: a plane)
  uint32_t* slot;   // per LDS body: its slot
  uint32_t* nb;     // bodies held; 0: the lists are swept through the L2 instead
};

template <int D, typename Real>
__device__ __forceinline__ CompLds<D, Real> comp_lds_arrays() {
  using V = typename Vec4<Real>::type;
  constexpr uint32_t cap = kCompLds<Real>;
  __shared__ V s_body[cap * kBodyVecs<D>];
  __shared__ Real s_r[3 * cap];
  __shared__ Real s_rec[kRecWords<D> * cap];
  __shared__ uint16_t s_eidx[2 * cap];
  __shared__ uint32_t s_slot[cap];
  __shared__ uint32_t s_nb;
  return CompLds<D, Real>{s_body, s_r, s_rec, s_eidx, s_slot, &s_nb};
}

// After component_setup (its last grid barrier ends every read of the union-find parents):
// numbers the bodies of this workgroup's n listed contacts in LDS. A body of a component
// belongs to that component's workgroup alone, so the workgroup reuses the body's parent
// word as a claim: every endpoint clears it, the first endpoint to swap in its own number
// owns the body and takes the next LDS slot, the others read the owner's slot. Loads the
// listed contacts' records and impulses. Leaves *L.nb = 0 when the list or its bodies exceed
// kCompLds.
template <int D, typename Real>
__device__ void comp_lds_setup(const GSArgs<Real>& A, const uint32_t* s_list, uint32_t n, const CompLds<D, Real>& L) {
  uint32_t* par = A.comp_parent;
  __shared__ uint32_t s_fail;
  if (threadIdx.x == 0) {
    *L.nb = 0;
    s_fail = (n > kCompLds<Real> || !A.comp_lds) ? 1u : 0u;
  }
  __syncthreads();
  if (s_fail) {
    __syncthreads();
    return;
  }
  auto body_of = [&](uint32_t e) {
    const uint32_t k = s_list[e >> 1];
    return (e & 1) ? A.c.b[k] : A.c.a[k];
  };
  for (uint32_t e = threadIdx.x; e < 2 * n; e += kGSBlock) {
    const uint32_t x = body_of(e);
    if (x != kNoBody) __hip_atomic_store(par + x, kNoBody, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  wait_stores();
  __syncthreads();
  for (uint32_t e = threadIdx.x; e < 2 * n; e += kGSBlock) {
    const uint32_t x = body_of(e);
    uint32_t expect = kNoBody;
    if (x != kNoBody)
      __hip_atomic_compare_exchange_strong(par + x, &expect, e, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
  }
  wait_stores();
  __syncthreads();
  for (uint32_t e = threadIdx.x; e < 2 * n; e += kGSBlock) {
    const uint32_t x = body_of(e);
    L.eidx[e] = kNoLocal;
    if (x != kNoBody && __hip_atomic_load(par + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == e) {
      const uint32_t i = atomicAdd(L.nb, 1u);
      if (i < kCompLds<Real>) {
        L.eidx[e] = (uint16_t)i;
        L.slot[i] = x;
      } else {
        s_fail = 1u;
      }
    }
  }
  __syncthreads();
  for (uint32_t e = threadIdx.x; e < 2 * n; e += kGSBlock) {
    const uint32_t x = body_of(e);
    if (x == kNoBody) continue;
    const uint32_t owner = __hip_atomic_load(par + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (owner != e) L.eidx[e] = L.eidx[owner];
  }
  for (uint32_t j = threadIdx.x; j < n; j += kGSBlock) {
    const uint32_t k = s_list[j];
    const Real* rp = A.c.r + 3 * (size_t)k;
#pragma unroll
    for (int d = 0; d < 3; ++d) L.r[3 * j + d] = ld_imp<true>(rp + d);
    Rec<D, Real> R;
    load_rec<D, Real>(A.c, k, R);
    Real* q = L.rec + (size_t)kRecWords<D> * j;
#pragma unroll
    for (int d = 0; d < 3; ++d) q[d] = R.n[d];
    if (D == 3) {
#pragma unroll
      for (int d = 0; d < 3; ++d) q[3 + d] = R.s[d];
      q[6] = R.la;
      q[7] = R.lb;
    }
    Real* w = q + kRecWords<D> - 4;
    w[0] = R.wn;
    w[1] = R.wt;
    w[2] = R.mu;
    w[3] = R.f;
  }
  __syncthreads();
  if (threadIdx.x == 0 && s_fail) *L.nb = 0;
  __syncthreads();
}

// One sweep of this workgroup's components held in LDS (comp_lds_setup): the body records in,
// the colours [ccol, ncol) of its list with workgroup barriers, the records out.
template <int D, typename Real>
__device__ void comp_lds_sweep(const GSArgs<Real>& A, const uint32_t* s_list, const uint32_t* s_coff, uint32_t ccol,
                               uint32_t ncol, const CompLds<D, Real>& L, const BodyPort<Real>& P, bool copy) {
  using V = typename Vec4<Real>::type;
  const uint32_t nv = *L.nb * kBodyVecs<D>;
  for (uint32_t i = threadIdx.x; i < nv; i += kGSBlock)
    L.body[i] = P.ld((size_t)kBodyVecs<D> * L.slot[i / kBodyVecs<D>] + i % kBodyVecs<D>);
  __syncthreads();
  for (uint32_t col = ccol; col < ncol; ++col) {
    const uint32_t lo = s_coff[col], hi = s_coff[col + 1];
    if (lo == hi) continue;  // the same in every thread of the workgroup
    for (uint32_t j = lo + threadIdx.x; j < hi; j += kGSBlock) {
      const uint32_t k = s_list[j];
      const uint16_t ea = L.eidx[2 * j], eb = L.eidx[2 * j + 1];
      Loaded<D, Real> ld;
      // the record as load_rec builds it (a and b: only whether b is a body matters here)
      Rec<D, Real>& R = ld.R;
      const Real* q = L.rec + (size_t)kRecWords<D> * j;
      R.a = 0;
      R.b = (eb == kNoLocal) ? kNoBody : 0u;
#pragma unroll
      for (int d = 0; d < 3; ++d) R.n[d] = q[d];
      if (D == 3) {
#pragma unroll
        for (int d = 0; d < 3; ++d) R.s[d] = q[3 + d];
        R.la = q[6];
        R.lb = q[7];
      }
      frame_from_ns<D, Real>(R);
      const Real* w = q + kRecWords<D> - 4;
      R.wn = w[0];
      R.wt = w[1];
      R.mu = w[2];
      R.f = w[3];
      Real* rl = L.r + 3 * j;
      ld.r[0] = rl[0];
      ld.r[1] = rl[1];
      ld.r[2] = rl[2];
      Real* rp = A.c.r + 3 * (size_t)k;
      update_core<D, Real>(
          A, k, ld, false, copy, (size_t)kBodyVecs<D> * ea, (size_t)kBodyVecs<D> * (eb == kNoLocal ? 0u : eb),
          [&](size_t i) { return L.body[i]; }, [&](size_t i, const V& x) { L.body[i] = x; },
          [&](int d, Real x) {
            rl[d] = x;
            st_imp<true>(rp + d, x);
          });
    }
    __syncthreads();
  }
  for (uint32_t i = threadIdx.x; i < nv; i += kGSBlock)
    P.st((size_t)kBodyVecs<D> * L.slot[i / kBodyVecs<D>] + i % kBodyVecs<D>, L.body[i]);
}

// Workgroup 0: is one of the contacts listed by the last full evaluation (one per thread)
// still at or above `th`? Each residual is the one the full evaluation computes (FP32
// estimate in the FP32 phase, FP64 otherwise). The answer goes to A.vflag for every
// workgroup after the next grid barrier.
template <int D, typename Real, bool Coh>
__device__ void violator_check(const GSArgs<Real>& A, uint32_t ev, double th, const BodyPort<Real>& P,
                               double* s_fm) {
  const uint32_t cnt =
      min(__hip_atomic_load(A.viol + ((ev - 1) & 1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), kViolCap);
  double e = 0.0;
  if (threadIdx.x < cnt) {
    const uint32_t k = __hip_atomic_load(A.viol + 2 + threadIdx.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    e = (A.phase == 1) ? (double)contact_err<D, Real, float, Coh>(A, k, P) : contact_err<D, Real, double, Coh>(A, k, P);
  }
  e = wave_reduce_max_f64(e);
  if ((threadIdx.x & 63) == 0) s_fm[threadIdx.x >> 6] = e;
  __syncthreads();
  if (threadIdx.x == 0) {
    double mx = s_fm[0];
    for (uint32_t w = 1; w < kGSBlock / 64; ++w) mx = fmax(mx, s_fm[w]);
    __hip_atomic_store(A.vflag, (cnt > 0 && mx >= th) ? 1u : 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
}

// Coh: the colour passes and the residual evaluations access the body records and the
// impulses coherently (BodyPort, ld_imp/st_imp) and are separated by grid_sync_coherent;
// the small colours that workgroup 0 sweeps alone go through its L2 between an agent-scope
// acquire and release; the initial application of the impulses, the compaction and the
// best-iterate return keep full barriers and ordinary accesses.
template <int D, typename Real, bool Coh>
__global__ void __launch_bounds__(kGSBlock) k_gs(GSArgs<Real> A) {
  const uint32_t gsize = gridDim.x * kGSBlock;
  const uint32_t gtid = blockIdx.x * kGSBlock + threadIdx.x;
  __shared__ uint32_t off[kMaxColours + 1];
  __shared__ uint32_t s_frozen[kSceneMaskBits / 32];
  __shared__ uint32_t s_aoff[2][kMaxColours + 1];  // colour offsets of the compacted lists
  __shared__ uint32_t s_tmp[kGSBlock / 64];
  __shared__ double s_fm[kGSBlock / 64];
  const uint32_t ncol = A.ctr->ncolours;
  for (uint32_t c = threadIdx.x; c <= ncol; c += kGSBlock) off[c] = A.coloff[c];
  __syncthreads();
  uint32_t max_iters = A.max_iters;
  // FP32 phase: at most three quarters of the sweep budget, the FP64 polish gets the rest
  if (A.phase == 1) max_iters = max(1u, max_iters - max_iters / 4);
  if (A.phase == 2) max_iters = (A.ctr->sweeps_fp32 < max_iters) ? max_iters - A.ctr->sweeps_fp32 : 0u;
  const bool track_best = A.phase != 1;

  if (A.apply_init) apply_impulses<D, Real>(A, off, ncol, gtid, gsize, false);
  // the FP32 phase stops on an FP32 estimate with a factor-2 margin; the FP64 phase
  // that follows decides on the tolerance
  const double stop = (A.phase == 1) ? 0.5 * A.tol : A.tol;
  const uint32_t base = (A.phase == 2) ? A.ctr->sweeps_fp32 : 0u;
  if (A.cscene) {
    for (uint32_t sc = gtid; sc < 2 * A.nscenes; sc += gsize)
      __hip_atomic_store(A.scene_acc + sc, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    grid_barrier(A.bar);
  }
  uint32_t ev = 0;  // evaluations so far; the last one's scene residuals are in buffer (ev - 1) & 1
  const BodyPort<Real> P = body_port<D, Real>(A.body, A.nbodies);
  // Component phase: after the wide colours [0, ccol), each workgroup sweeps the components
  // it was given colour by colour, through its L2 between an agent-scope acquire and
  // release, and one grid barrier closes the sweep (instead of one per colour and a
  // one-workgroup tail). Components share no body, so this is the sweep order of the
  // colour passes for every contact.
  __shared__ uint32_t s_clist[kCompCap];
  __shared__ uint32_t s_coff[kMaxColours + 1];
  uint32_t ccol = ncol;
  CompLds<D, Real> CL{};
  if constexpr (Coh) {
    if (A.comp_parent && !A.cscene && gridDim.x > 1 && A.comp_col < ncol) {
      ccol = component_setup<D, Real>(A, off, ncol, A.comp_col, gtid, gsize, s_clist, s_coff);
      if (ccol < ncol) {
        CL = comp_lds_arrays<D, Real>();
        comp_lds_setup<D, Real>(A, s_clist, s_coff[ncol], CL);
      }
    }
  }
  if (ccol < ncol && gtid == 0) ++A.ctr->comp_launches;
  // One scene with a one-workgroup tail or a component phase: after a full evaluation that
  // listed the contacts near its maximum, a sweep's full evaluation is skipped while
  // workgroup 0 finds one of them still at or above a threshold th, with a full evaluation
  // at least every kFullEvery sweeps. th = the best residual so far: then the maximum is at
  // least the best, so the full evaluation could only have said "go on, no new best" (the
  // solve is not converged, as the best is at or above the stopping value while it runs,
  // and neither the best iterate nor the best residual changes). The FP32 phase's
  // stagnation test (every 1024 sweeps) compares best residuals, which skipped sweeps thus
  // leave as they are; its sweeps are evaluated in full. In the first kSkipUntil sweeps of
  // the FP32 phase, which keeps no best iterate, th = the stopping value (a warm-started
  // step converges well within them, and the best residual by the first stagnation test at
  // sweep 1024 is set by the sweeps after them).
  constexpr uint32_t kFullEvery = 256, kSkipUntil = 1024 - kFullEvery;
  const bool subset = A.viol && !A.cscene && (min(A.tail_col, ncol) < ncol || ccol < ncol);
  double err = eval_residual<D, Real, Coh>(A, gtid, gsize, ev++, stop, base, 0, nullptr, 0, P,
                                           subset ? stop : -1.0);
  uint32_t last_full = 0;
  const bool masked = A.cscene && A.nscenes <= kSceneMaskBits;
  if (masked) build_frozen_mask(A, (ev - 1) & 1, stop, s_frozen);
  // Ensembles: once enough scenes have frozen, the sweeps and evaluations walk a compacted
  // list of the other scenes' contacts (lst with colour offsets o) instead of skipping the
  // frozen ones one by one. A compacted-away scene is no longer evaluated (its residual
  // reads 0 and it stays frozen); frozen_max keeps the largest residual such scenes had,
  // so that the reported residual is the maximum over all contacts still.
  const bool compact = masked && A.act_idx;
  const uint32_t* lst = nullptr;
  const uint32_t* o = off;
  uint32_t nlst = A.C, sel = 0, nfrozen = 0;
  double frozen_max = 0.0;
  double best = err;
  // Best iterate (So-bogus returns it when the sweep limit is reached): after an
  // evaluation that lowers the best residual, the next sweep stores every contact's
  // impulse into rbest as it reads it, before updating it; while that copy is pending the
  // current iterate is the best one.
  bool copy_pending = track_best;
  // FP32 phase stagnation: leave when 1024 sweeps lowered the best residual by less than
  // 0.01% (FP32 has reached its rounding floor; FP64 continues from there). Slow but
  // steady progress stays in FP32, whose sweeps cost about half as much.
  double ref = best;
  uint32_t ref_it = 0;
  uint32_t it = 0;
  if (!(err < stop)) {
    for (it = 1; it <= max_iters; ++it) {
      const bool copy = copy_pending;
      copy_pending = false;
      const uint32_t eb = (ev - 1) & 1;
      const uint32_t tail = min(A.tail_col, ncol);
      const uint32_t nwide = (ccol < ncol) ? ccol : tail;  // colours swept in grid-wide passes
      const bool eval_now = A.eval_every > 0 && it % A.eval_every == 0;
      const bool stagnation_due = A.phase == 1 && it - ref_it >= 1024;
      const bool may_skip = subset && eval_now && it - last_full < kFullEvery && !stagnation_due;
      const double th = (A.phase == 1 && it < kSkipUntil) ? stop : fmax(stop, best);
      // the first contact of each colour pass is loaded before the barrier that opens it
      // completes (between the arrival and the wait of a coherent barrier), so that only the
      // body gathers wait for the barrier
      Loaded<D, Real> next;
      if (nwide > 0 && o[0] + gtid < o[1]) load_contact<D, Real, Coh>(A, lst ? lst[o[0] + gtid] : o[0] + gtid, next);
      for (uint32_t col = 0; col < nwide; ++col) {
        uint32_t j = o[col] + gtid;
        if (j < o[col + 1]) {
          uint32_t k = lst ? lst[j] : j;
          if (frozen_fast(A, k, eb, stop, s_frozen)) keep<Coh>(A, k, copy);
          else update<D, Real, Coh>(A, k, next, false, copy, P);
          for (j += gsize; j < o[col + 1]; j += gsize) {
            k = lst ? lst[j] : j;
            if (frozen_fast(A, k, eb, stop, s_frozen)) keep<Coh>(A, k, copy);
            else process<D, Real, Coh>(A, k, false, copy, P);
          }
        }
        const bool pre = col + 1 < nwide && o[col + 1] + gtid < o[col + 2];
        if (Coh) {
          const GridTicket t = grid_arrive_coherent(A.bar);
          if (pre) load_contact<D, Real, Coh>(A, lst ? lst[o[col + 1] + gtid] : o[col + 1] + gtid, next);
          grid_wait_coherent(A.bar, t);
        } else {
          if (pre) load_contact<D, Real>(A, lst ? lst[o[col + 1] + gtid] : o[col + 1] + gtid, next);
          grid_barrier(A.bar);
        }
      }
      if (ccol < ncol) {
        if (Coh && *CL.nb > 0) {
          if constexpr (Coh) comp_lds_sweep<D, Real>(A, s_clist, s_coff, ccol, ncol, CL, P, copy);
        } else {
          // through the L2 between an agent-scope acquire and release
          if (threadIdx.x == 0) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
          __syncthreads();
          for (uint32_t col = ccol; col < ncol; ++col) {
            const uint32_t lo = s_coff[col], hi = s_coff[col + 1];
            if (lo == hi) continue;  // the same in every thread of the workgroup
            for (uint32_t j = lo + threadIdx.x; j < hi; j += kGSBlock) process<D, Real>(A, s_clist[j], false, copy);
            __syncthreads();
          }
          wait_stores();
          __syncthreads();
          if (threadIdx.x == 0) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
          __syncthreads();
        }
        grid_sync_coherent(A.bar);
        if (may_skip) {
          // the listed contacts may lie in any workgroup's components: checked after the
          // barrier, and one more barrier publishes the answer
          if (blockIdx.x == 0) violator_check<D, Real, Coh>(A, ev, th, P, s_fm);
          grid_sync_coherent(A.bar);
        }
      } else if (tail < ncol) {
        // the last, small colours (typically the contacts of a few bodies of high degree):
        // one workgroup sweeps them in order with workgroup barriers, then one grid barrier,
        // instead of a grid barrier per colour; the sweep order is unchanged
        // With coherent passes the workgroup works through its L2 with ordinary accesses:
        // an agent-scope acquire first (drops stale lines of records other workgroups
        // updated), an agent-scope release at the end (writes its updates back to memory,
        // where the coherent loads of the next pass and of the evaluation read them).
        if (blockIdx.x == 0) {
          if (Coh) {
            if (threadIdx.x == 0) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            __syncthreads();
          }
          for (uint32_t col = tail; col < ncol; ++col) {
            for (uint32_t j = o[col] + threadIdx.x; j < o[col + 1]; j += kGSBlock) {
              const uint32_t k = lst ? lst[j] : j;
              if (frozen_fast(A, k, eb, stop, s_frozen)) keep(A, k, copy);
              else process<D, Real>(A, k, false, copy);
            }
            __syncthreads();
          }
          if (Coh) {
            wait_stores();
            __syncthreads();
            if (threadIdx.x == 0) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            __syncthreads();
          }
          if (may_skip) violator_check<D, Real, Coh>(A, ev, th, P, s_fm);
        }
        if (Coh) grid_sync_coherent(A.bar);
        else grid_barrier(A.bar);
      }
      if (may_skip && __hip_atomic_load(A.vflag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 1u) {
        if (gtid == 0) ++A.ctr->evals_skipped;
        continue;  // a listed contact is still at or above th: the full evaluation would change nothing
      }
      if (eval_now) {
        last_full = it;
        err = fmax(eval_residual<D, Real, Coh>(A, gtid, gsize, ev++, stop, base, it, lst, nlst, P,
                                               subset ? fmax(stop, 0.1 * err) : -1.0),
                   frozen_max);
        if (masked) build_frozen_mask(A, (ev - 1) & 1, stop, s_frozen);
        if (err < best) {
          best = err;
          copy_pending = track_best;
        }
        if (err < stop) break;
        if (A.phase == 1 && it - ref_it >= 1024) {
          if (best > 0.9999 * ref) break;
          ref = best;
          ref_it = it;
        }
        if (compact) {
          // compact when the frozen scenes grew by an eighth of those that were left
          const uint32_t nf = count_frozen(A.nscenes, s_frozen, s_tmp);
          if (nf > nfrozen && nf - nfrozen >= max(1u, (A.nscenes - nfrozen) / 8)) {
            // the residuals of the scenes compacted away now (0 for those compacted earlier)
            double fm = 0.0;
            for (uint32_t sc = threadIdx.x; sc < A.nscenes; sc += kGSBlock)
              if ((s_frozen[sc >> 5] >> (sc & 31)) & 1u)
                fm = fmax(fm, __longlong_as_double((long long)__hip_atomic_load(
                                  A.scene_acc + (size_t)((ev - 1) & 1) * A.nscenes + sc, __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT)));
            fm = wave_reduce_max_f64(fm);
            if ((threadIdx.x & 63) == 0) s_fm[threadIdx.x >> 6] = fm;
            __syncthreads();
            for (uint32_t w = 0; w < kGSBlock / 64; ++w) frozen_max = fmax(frozen_max, s_fm[w]);
            uint32_t* dst = A.act_idx + (size_t)sel * A.C;
            compact_active(A, lst, o, ncol, s_frozen, dst, s_aoff[sel]);
            lst = dst;
            o = s_aoff[sel];
            nlst = o[ncol];
            sel ^= 1u;
            nfrozen = nf;
          }
        }
      }
    }
    if (it > max_iters && track_best && !copy_pending) {
      // not converged and the current iterate is not the best: return the best iterate
      // and its velocities (So-bogus keeps x_best). Ensembles: a frozen scene met the
      // tolerance and keeps its impulses (each scene is a solve of its own). A full barrier
      // first: the copies and impulses are read here in another thread's mapping.
      grid_barrier(A.bar);
      for (uint32_t col = 0; col < ncol; ++col)
        for (uint32_t j = o[col] + gtid; j < o[col + 1]; j += gsize) {
          const uint32_t k = lst ? lst[j] : j;
          if (frozen_fast(A, k, (ev - 1) & 1, stop, s_frozen)) continue;
#pragma unroll
          for (int d = 0; d < 3; ++d) A.c.r[3 * (size_t)k + d] = A.rbest[3 * (size_t)k + d];
        }
      grid_barrier(A.bar);
      apply_impulses<D, Real>(A, off, ncol, gtid, gsize, true);
    }
    if (it > max_iters) it = max_iters;
  }
  if (gtid == 0) {
    if (A.phase == 1) {
      A.ctr->sweeps_fp32 = it;
      A.ctr->error_fp32 = best;
      A.ctr->evals_fp32 = ev;
    }
    A.ctr->sweeps += it;
    A.ctr->error = best;
    A.ctr->evals += ev;
  }
  if (Coh) grid_exit_coherent(A.bar);
}

// ---------------------------------------------------------------------------
// The same solve for a contact set that one workgroup holds in LDS: the records, the
// impulses and the velocity records of every touched body live in LDS for the whole
// solve, so a colour pass is LDS traffic and a workgroup barrier instead of dependent
// global-memory round trips. Bodies are addressed by endpoint: each touched body is
// owned by the smallest endpoint 2k + q (contact k, side q) that references it (an
// atomicMin over the per-slot map lmap, restored to all ones afterwards).
// ---------------------------------------------------------------------------
template <int D, typename Real>
struct LdsLayout {
  using V = typename Vec4<Real>::type;
  // bytes per contact: 2 endpoints x body record, ids, frame, scalars, impulse
  static constexpr size_t kBodyBytes = 2 * kBodyVecs<D> * sizeof(V);
  static constexpr size_t kRecBytes = 2 * sizeof(uint32_t) + (D == 3 ? 8 : 2) * sizeof(Real) + 4 * sizeof(Real) +
                                      3 * sizeof(Real);
  static size_t bytes(uint32_t C) { return (size_t)C * (kBodyBytes + kRecBytes) + 64; }
};

template <int D, typename Real>
__global__ void __launch_bounds__(kGSBlock) k_gs_lds(GSArgs<Real> A) {
  using V = typename Vec4<Real>::type;
  extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
  const uint32_t C = A.C;
  V* sb = reinterpret_cast<V*>(lds);                                       // 2C x kBodyVecs
  uint32_t* ea = reinterpret_cast<uint32_t*>(sb + (size_t)2 * C * kBodyVecs<D>);
  uint32_t* eb = ea + C;
  Real* sn = reinterpret_cast<Real*>(eb + C);                              // D == 3: n(3) s(3) la lb; 2D: n(2)
  Real* ss = sn + (size_t)(D == 3 ? 3 : 2) * C;
  Real* sl = (D == 3) ? ss + (size_t)3 * C : ss;                          // la, lb (3D)
  Real* sw = (D == 3) ? sl + (size_t)2 * C : ss;                          // wn, wt, mu, f
  Real* sr = sw + (size_t)4 * C;                                           // impulses, 3 per contact
  __shared__ uint32_t off[kMaxColours + 1];
  __shared__ double s_red[kGSBlock / 64];
  const uint32_t tid = threadIdx.x;
  const uint32_t ncol = A.ctr->ncolours;
  for (uint32_t c = tid; c <= ncol; c += kGSBlock) off[c] = A.coloff[c];
  // endpoint owners
  for (uint32_t k = tid; k < C; k += kGSBlock) {
    atomicMin(&A.lmap[A.c.a[k]], 2 * k);
    if (A.c.b[k] != kNoBody) atomicMin(&A.lmap[A.c.b[k]], 2 * k + 1);
  }
  __syncthreads();
  const V* gb = reinterpret_cast<const V*>(A.body);
  for (uint32_t k = tid; k < C; k += kGSBlock) {
    const uint32_t a = A.c.a[k], b = A.c.b[k];
    const uint32_t oa = __hip_atomic_load(&A.lmap[a], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t ob = (b != kNoBody) ? __hip_atomic_load(&A.lmap[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                       : kNoBody;
    ea[k] = oa;
    eb[k] = ob;
    if (oa == 2 * k)
      for (int h = 0; h < kBodyVecs<D>; ++h) sb[(size_t)kBodyVecs<D> * oa + h] = gb[(size_t)kBodyVecs<D> * a + h];
    if (ob == 2 * k + 1)
      for (int h = 0; h < kBodyVecs<D>; ++h) sb[(size_t)kBodyVecs<D> * ob + h] = gb[(size_t)kBodyVecs<D> * b + h];
    for (int d = 0; d < (D == 3 ? 3 : 2); ++d) sn[(size_t)(D == 3 ? 3 : 2) * k + d] = A.c.n[3 * (size_t)k + d];
    if (D == 3) {
      for (int d = 0; d < 3; ++d) ss[3 * (size_t)k + d] = A.c.s[3 * (size_t)k + d];
      sl[2 * (size_t)k] = A.c.la[k];
      sl[2 * (size_t)k + 1] = A.c.lb[k];
    }
    sw[4 * (size_t)k] = A.c.wn[k];
    sw[4 * (size_t)k + 1] = A.c.wt[k];
    sw[4 * (size_t)k + 2] = A.c.mu[k];
    sw[4 * (size_t)k + 3] = A.c.forcing[k];
    for (int d = 0; d < 3; ++d) sr[3 * (size_t)k + d] = A.c.r[3 * (size_t)k + d];
  }
  __syncthreads();

  auto rec = [&](uint32_t k, Rec<D, Real>& R) {
    R.a = ea[k];
    R.b = eb[k];
    if (D == 3) {
      for (int d = 0; d < 3; ++d) {
        R.n[d] = sn[3 * (size_t)k + d];
        R.s[d] = ss[3 * (size_t)k + d];
      }
      R.t[0] = R.n[1] * R.s[2] - R.n[2] * R.s[1];
      R.t[1] = R.n[2] * R.s[0] - R.n[0] * R.s[2];
      R.t[2] = R.n[0] * R.s[1] - R.n[1] * R.s[0];
      R.la = sl[2 * (size_t)k];
      R.lb = sl[2 * (size_t)k + 1];
    } else {
      R.n[0] = sn[2 * (size_t)k];
      R.n[1] = sn[2 * (size_t)k + 1];
      R.n[2] = 0;
      R.t[0] = -R.n[1];
      R.t[1] = R.n[0];
      R.t[2] = 0;
      R.s[0] = R.s[1] = R.s[2] = 0;
      R.la = R.lb = 0;
    }
    R.wn = sw[4 * (size_t)k];
    R.wt = sw[4 * (size_t)k + 1];
    R.mu = sw[4 * (size_t)k + 2];
    R.f = sw[4 * (size_t)k + 3];
  };
  // one contact: the k_gs update on LDS data (init: apply the stored impulse)
  auto update = [&](uint32_t k, bool init) {
    Rec<D, Real> R;
    rec(k, R);
    const bool hasb = R.b != kNoBody;
    V* bv = sb;
    const size_t ia = (size_t)kBodyVecs<D> * R.a, ib = (size_t)kBodyVecs<D> * R.b;
    V va = bv[ia], wa = {0, 0, 0, 0}, vb = {0, 0, 0, 0}, wb = {0, 0, 0, 0};
    if (D == 3) wa = bv[ia + 1];
    if (hasb) {
      vb = bv[ib];
      if (D == 3) wb = bv[ib + 1];
    }
    Real* rp = sr + 3 * (size_t)k;
    const Real r[3] = {rp[0], rp[1], rp[2]};
    Real dr[3];
    if (init) {
      dr[0] = r[0];
      dr[1] = r[1];
      dr[2] = r[2];
    } else {
      Real u[3], rn[3];
      rel_vel<D, Real, Real>(R, va, wa, vb, wb, u);
      local_solve<D, Real>(R, u, r, rn);
      dr[0] = rn[0] - r[0];
      dr[1] = rn[1] - r[1];
      dr[2] = rn[2] - r[2];
      rp[0] = rn[0];
      rp[1] = rn[1];
      if (D == 3) rp[2] = rn[2];
    }
    Real P[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) P[d] = R.n[d] * dr[0] + R.t[d] * dr[D - 1] + ((D == 3) ? R.s[d] * dr[1] : (Real)0);
    va.x += va.w * P[0];
    va.y += va.w * P[1];
    va.z += va.w * P[2];
    bv[ia] = va;
    Real tau[3];
    if (D == 3) {
#pragma unroll
      for (int d = 0; d < 3; ++d) tau[d] = dr[1] * R.t[d] - dr[2] * R.s[d];
      const Real ka = -R.la * wa.w;
      wa.x += ka * tau[0];
      wa.y += ka * tau[1];
      wa.z += ka * tau[2];
      bv[ia + 1] = wa;
    }
    if (hasb) {
      vb.x -= vb.w * P[0];
      vb.y -= vb.w * P[1];
      vb.z -= vb.w * P[2];
      bv[ib] = vb;
      if (D == 3) {
        const Real kb = R.lb * wb.w;
        wb.x -= kb * tau[0];
        wb.y -= kb * tau[1];
        wb.z -= kb * tau[2];
        bv[ib + 1] = wb;
      }
    }
  };
  auto residual = [&]() -> double {
    double m = 0.0;
    for (uint32_t k = tid; k < C; k += kGSBlock) {
      Rec<D, Real> R;
      rec(k, R);
      const V z = {0, 0, 0, 0};
      const size_t ia = (size_t)kBodyVecs<D> * R.a, ib = (size_t)kBodyVecs<D> * R.b;
      const V va = sb[ia];
      const V wa = (D == 3) ? sb[ia + 1] : z;
      const V vb = (R.b != kNoBody) ? sb[ib] : z;
      const V wb = (D == 3 && R.b != kNoBody) ? sb[ib + 1] : z;
      if (A.phase == 1) {  // FP32 phase: an FP32 estimate
        float u[3];
        rel_vel<D, float, Real>(R, va, wa, vb, wb, u);
        const float r[3] = {(float)sr[3 * (size_t)k], (float)sr[3 * (size_t)k + 1], (float)sr[3 * (size_t)k + 2]};
        m = fmax(m, (double)fb_err<D, float>((float)R.mu, r, u));
      } else {
        double u[3];
        rel_vel<D, double, Real>(R, va, wa, vb, wb, u);
        const double r[3] = {(double)sr[3 * (size_t)k], (double)sr[3 * (size_t)k + 1],
                             (double)sr[3 * (size_t)k + 2]};
        m = fmax(m, fb_err<D>((double)R.mu, r, u));
      }
    }
    m = wave_reduce_max_f64(m);
    if ((tid & 63) == 0) s_red[tid >> 6] = m;
    __syncthreads();
    double out = s_red[0];
    for (int w = 1; w < kGSBlock / 64; ++w) out = fmax(out, s_red[w]);
    __syncthreads();
    return out;
  };
  // v = v0 + M^-1 H^T r: reset the owned bodies to v0, then apply every impulse colour by colour
  auto apply_all = [&](bool reset) {
    if (reset) {
      for (uint32_t e = tid; e < 2 * C; e += kGSBlock) {
        const uint32_t k = e >> 1;
        const uint32_t owner = (e & 1) ? eb[k] : ea[k];
        if (owner != e) continue;
        const uint32_t slot = (e & 1) ? A.c.b[k] : A.c.a[k];
        const uint32_t i = A.sid ? A.sid[slot] : slot;
        V* rb = sb + (size_t)kBodyVecs<D> * e;
        rb[0].x = (Real)A.v0[(size_t)D * i];
        rb[0].y = (Real)A.v0[(size_t)D * i + 1];
        rb[0].z = (D == 3) ? (Real)A.v0[(size_t)D * i + D - 1] : (Real)0;
        if (D == 3) {
          rb[1].x = (Real)A.w0[3 * (size_t)i];
          rb[1].y = (Real)A.w0[3 * (size_t)i + 1];
          rb[1].z = (Real)A.w0[3 * (size_t)i + 2];
        }
      }
      __syncthreads();
    }
    for (uint32_t col = 0; col < ncol; ++col) {
      for (uint32_t k = off[col] + tid; k < off[col + 1]; k += kGSBlock) update(k, true);
      __syncthreads();
    }
  };

  uint32_t max_iters = A.max_iters;
  if (A.phase == 1) max_iters = max(1u, max_iters - max_iters / 4);
  if (A.phase == 2) max_iters = (A.ctr->sweeps_fp32 < max_iters) ? max_iters - A.ctr->sweeps_fp32 : 0u;
  const bool track_best = A.phase != 1;
  if (A.apply_init) apply_all(false);
  double err = residual();
  uint32_t nev = 1;  // full residual evaluations
  double best = err;
  if (track_best)
    for (uint32_t k = tid; k < C; k += kGSBlock)
      for (int d = 0; d < 3; ++d) A.rbest[3 * (size_t)k + d] = sr[3 * (size_t)k + d];
  double ref = best;
  uint32_t ref_it = 0, it = 0;
  // the FP32 phase stops on an FP32 estimate with a factor-2 margin; the FP64 phase
  // that follows decides on the tolerance
  const double stop = (A.phase == 1) ? 0.5 * A.tol : A.tol;
  if (!(err < stop)) {
    for (it = 1; it <= max_iters; ++it) {
      for (uint32_t col = 0; col < ncol; ++col) {
        for (uint32_t k = off[col] + tid; k < off[col + 1]; k += kGSBlock) update(k, false);
        __syncthreads();
      }
      if (A.eval_every > 0 && it % A.eval_every == 0) {
        err = residual();
        ++nev;
        if (err < best) {
          best = err;
          if (track_best)
            for (uint32_t k = tid; k < C; k += kGSBlock)
              for (int d = 0; d < 3; ++d) A.rbest[3 * (size_t)k + d] = sr[3 * (size_t)k + d];
        }
        if (err < stop) break;
        if (A.phase == 1 && it - ref_it >= 1024) {
          if (best > 0.9999 * ref) break;
          ref = best;
          ref_it = it;
        }
      }
    }
    if (it > max_iters && track_best) {
      // not converged: the best iterate and its velocities
      __syncthreads();
      for (uint32_t k = tid; k < C; k += kGSBlock)
        for (int d = 0; d < 3; ++d) sr[3 * (size_t)k + d] = A.rbest[3 * (size_t)k + d];
      __syncthreads();
      apply_all(true);
    }
    if (it > max_iters) it = max_iters;
  }
  __syncthreads();
  // results back: impulses, owned bodies; then the owner map back to all ones
  V* gbw = reinterpret_cast<V*>(A.body);
  for (uint32_t k = tid; k < C; k += kGSBlock) {
    for (int d = 0; d < 3; ++d) A.c.r[3 * (size_t)k + d] = sr[3 * (size_t)k + d];
    if (ea[k] == 2 * k)
      for (int h = 0; h < kBodyVecs<D>; ++h) gbw[(size_t)kBodyVecs<D> * A.c.a[k] + h] = sb[(size_t)kBodyVecs<D> * ea[k] + h];
    if (eb[k] == 2 * k + 1)
      for (int h = 0; h < kBodyVecs<D>; ++h) gbw[(size_t)kBodyVecs<D> * A.c.b[k] + h] = sb[(size_t)kBodyVecs<D> * eb[k] + h];
  }
  __syncthreads();
  for (uint32_t k = tid; k < C; k += kGSBlock) {
    if (ea[k] == 2 * k) A.lmap[A.c.a[k]] = kNoBody;
    if (eb[k] == 2 * k + 1) A.lmap[A.c.b[k]] = kNoBody;
  }
  if (tid == 0) {
    if (A.phase == 1) {
      A.ctr->sweeps_fp32 = it;
      A.ctr->error_fp32 = best;
      A.ctr->evals_fp32 = nev;
    }
    A.ctr->sweeps += it;
    A.ctr->error = best;
    A.ctr->evals += nev;
  }
}

// Body records [v | minv] ([omega | Iinv]) by slot, in both precisions.
template <int D>
__global__ void __launch_bounds__(kBlock) k_body_init(uint32_t n, const uint32_t* __restrict__ sid,
                                                      const double* __restrict__ v0, const double* __restrict__ w0,
                                                      const double* __restrict__ m,
                                                      const double* __restrict__ inertia, double* __restrict__ body64,
                                                      float* __restrict__ body32, int fp32) {
  const uint32_t sl = blockIdx.x * kBlock + threadIdx.x;
  if (sl >= n) return;
  const uint32_t i = sid ? sid[sl] : sl;
  double4* bv64 = reinterpret_cast<double4*>(body64) + (size_t)kBodyVecs<D> * sl;
  float4* bv32 = reinterpret_cast<float4*>(body32) + (size_t)kBodyVecs<D> * sl;
  double v[3] = {0, 0, 0}, w[3] = {0, 0, 0};
#pragma unroll
  for (int d = 0; d < D; ++d) v[d] = v0[(size_t)D * i + d];
  if (D == 3)
#pragma unroll
    for (int d = 0; d < 3; ++d) w[d] = w0[3 * (size_t)i + d];
  const double minv = 1.0 / m[i];
  const double iinv = (D == 3) ? 1.0 / inertia[i] : 0.0;
  bv64[0] = make_double4(v[0], v[1], v[2], minv);
  if (D == 3) bv64[1] = make_double4(w[0], w[1], w[2], iinv);
  if (fp32) {
    bv32[0] = make_float4((float)v[0], (float)v[1], (float)v[2], (float)minv);
    if (D == 3) bv32[1] = make_float4((float)w[0], (float)w[1], (float)w[2], (float)iinv);
  }
}

__global__ void __launch_bounds__(kBlock) k_promote_r(uint32_t C, const float* __restrict__ r32,
                                                      double* __restrict__ r64) {
  const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= 3 * C) return;
  r64[k] = (double)r32[k];
}

// End of step: v2 from the sweep velocities, impulses back to the reference order,
// and the final flow from q0 with v2 (GeometricImpactFrictionMap.cpp:441).
template <int D>
__global__ void __launch_bounds__(kBlock) k_finish_bodies(uint32_t n, const uint32_t* __restrict__ slot_of,
                                                          const double* __restrict__ body64,
                                                          const double* __restrict__ q0, const double* __restrict__ m,
                                                          double dt, double gx, double gy, double gz, int integrator,
                                                          double* __restrict__ q, double* __restrict__ v,
                                                          double* __restrict__ w, double* __restrict__ v2,
                                                          double* __restrict__ w2) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const double g[3] = {gx, gy, gz};
  const double4* rec = reinterpret_cast<const double4*>(body64) + (size_t)kBodyVecs<D> * (slot_of ? slot_of[i] : i);
  const double4 bvi = rec[0];
  const double vv[3] = {bvi.x, bvi.y, bvi.z};
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const size_t k = (size_t)D * i + d;
    v2[k] = vv[d];
    flow_component<D>(integrator, m[i], g[d], dt, q0[k], vv[d], &q[k], &v[k]);
  }
  if (D == 3) {
    const double4 bwi = rec[1];
    w2[3 * (size_t)i] = w[3 * (size_t)i] = bwi.x;
    w2[3 * (size_t)i + 1] = w[3 * (size_t)i + 1] = bwi.y;
    w2[3 * (size_t)i + 2] = w[3 * (size_t)i + 2] = bwi.z;
  }
}

__global__ void __launch_bounds__(kBlock) k_impulses_out(uint32_t C, int dim, const uint32_t* __restrict__ perm,
                                                         const double* __restrict__ r64, double* __restrict__ alpha,
                                                         double* __restrict__ beta, double mu) {
  const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= C) return;
  const uint32_t c = perm[k];
  alpha[c] = r64[3 * (size_t)k];
  for (int d = 1; d < dim; ++d) beta[(size_t)(dim - 1) * c + d - 1] = (mu == 0.0) ? 0.0 : r64[3 * (size_t)k + d];
}

// Persistent launch sized by the barrier/pass trade-off of persistent_blocks over one
// sweep's intervals (the barrier's cost grows with the workgroups): one workgroup (a
// workgroup barrier, no device fence) for small contact sets, more as colours widen.
// Largest dynamic LDS of the single-workgroup solve: sweeps measured on MI355X ran
// about three times slower per sweep once the allocation passed 128 KB, so the LDS
// solve stays below that.
constexpr size_t kGSLdsMax = 124 * 1024;

// ---------------------------------------------------------------------------
// Online autotuning of the sweep launch width (SweepTuner in scx_internal.h).
// ---------------------------------------------------------------------------
constexpr int kTuneWords = 100;  // host_u32[100..101]: the step's sweeps and FP32 sweeps

// Folds the timed launches of the previous solve into the table (their events and the
// copy of the sweep counters completed at the stream synchronisations since).
static void tune_resolve(scx_ctx* c) {
  SweepTuner& T = c->tuner;
  if (!T.counters_live) return;
  T.counters_live = false;
  const bool ok = hipEventSynchronize(T.copied) == hipSuccess;
  const uint32_t total = c->d.host_u32[kTuneWords], fp32 = c->d.host_u32[kTuneWords + 1];
  for (int ph = 0; ph < 2; ++ph) {
    SweepTuner::Pending& P = T.pend[ph];
    if (!P.live) continue;
    P.live = false;
    float ms = 0.f;
    if (!ok || hipEventElapsedTime(&ms, P.ev[0], P.ev[1]) != hipSuccess) continue;
    T.table.record(P.key, P.exp, ms, (ph == 0) ? fp32 : total - fp32);
  }
}

// Class of a sweep launch: bits 0-7 contacts in half-octaves, 8 3D, 9 FP64 phase,
// 10 ensemble, 11-15 barrier intervals per sweep in pairs.
static uint32_t tune_key(int D, int phase, bool ens, uint32_t C, int nintervals) {
  const uint32_t cb = (uint32_t)std::min(255.0, std::floor(2.0 * std::log2((double)std::max(1u, C))));
  const uint32_t wb = (uint32_t)std::min(31, (nintervals + 1) / 2);
  return cb | ((uint32_t)(D == 3) << 8) | ((uint32_t)phase << 9) | ((uint32_t)ens << 10) | (wb << 11);
}

// Workgroups of the next launch of a class (the exponent is remembered for its timing).
static int tune_pick(scx_ctx* c, int phase, uint32_t key, int model, int max_blocks) {
  SweepTuner& T = c->tuner;
  const int e = T.table.pick(key, model, max_blocks);
  SweepTuner::Pending& P = T.pend[phase];
  P.live = true;
  P.key = key;
  P.exp = e;
  return TuneTable::blocks(model, e, max_blocks);
}

// After a solve with timed launches: the sweep counters for their measurement.
static hipError_t tune_after_solve(scx_ctx* c) {
  SweepTuner& T = c->tuner;
  if (!T.enabled || !(T.pend[0].live || T.pend[1].live)) return hipSuccess;
  hipError_t e = hipMemcpyAsync(c->d.host_u32 + kTuneWords, &c->d.ctr->sweeps, 2 * sizeof(uint32_t),
                                hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipEventRecord(T.copied, c->stream);
  T.counters_live = e == hipSuccess;
  return e;
}

template <int D, typename Real>
static hipError_t coop_gs(scx_ctx* c, GSArgs<Real> A, int max_blocks, const uint32_t* col_off) {
  // Tail: the suffix of colours no wider than one workgroup, swept by one workgroup (no
  // grid barrier between them) when the launch has several workgroups and the suffix has
  // at least two colours.
  int ncol = 0;
  for (int b = 0; b < kMaxColours; ++b)
    if (col_off[b + 1] > col_off[b]) ncol = b + 1;
  int tail = ncol;
  while (tail > 0 && col_off[tail] - col_off[tail - 1] <= (uint32_t)kGSBlock) --tail;
  if (ncol - tail < 2) tail = ncol;
  // coherent colour passes unless switched off or the body records exceed the 32-bit
  // byte offsets of the buffer accesses
  const bool coh = c->coherent_sweeps && (uint64_t)A.nbodies * kBodyVecs<D> * 4 * sizeof(Real) < (1ull << 31);
  // Component phase (one scene, coherent passes): the colours after which at most
  // comp_share percent of the contacts remain. In the contact graphs of packings and piles
  // the late colours' contacts form many small components, of a few dozen contacts (on a
  // settled 262k-sphere pile at 20%; at 25% they join into one too large for a workgroup);
  // in sparser graphs (a 2D packing as it collapses) far more of the colours do. The share
  // adapts from step to step (comp_adapt_share); the phase starts before the tail and spans
  // two colours at least.
  int comp = ncol;
  uint32_t comp_contacts = 0;
  if (coh && A.comp_parent && !A.cscene) {
    int k = ncol;
    while (k > 0 && (double)(col_off[ncol] - col_off[k - 1]) * 100.0 <= (double)A.C * c->comp_share) --k;
    if (k < tail && ncol - k >= 2) {
      comp = k;
      comp_contacts = col_off[ncol] - col_off[k];
    }
  }
  // per sweep: one barrier interval per colour before the tail (or the component phase),
  // one for the tail (or the component phase) and one for the residual over all contacts
  uint32_t widths[kMaxColours + 2];
  int nw = 0;
  for (int b = 0; b < std::min(tail, comp); ++b)
    if (col_off[b + 1] > col_off[b]) widths[nw++] = col_off[b + 1] - col_off[b];
  if (std::min(tail, comp) < ncol) widths[nw++] = 1;
  widths[nw++] = A.C;
  const int model =
      c->force_blocks > 0 ? c->force_blocks : persistent_blocks(widths, nw, kGSBlock, max_blocks, coh);
  const size_t lds = LdsLayout<D, Real>::bytes(A.C);
  const bool lds_path = std::min(max_blocks, model) <= 1 && c->force_blocks <= 0 && lds <= kGSLdsMax && A.lmap &&
                        !A.cscene;
  // the autotuner picks among widths around the model's (not for the LDS solve)
  const int phase = std::is_same<Real, double>::value ? 1 : 0;
  const bool tune = c->tuner.enabled && c->force_blocks <= 0 && !lds_path && !c->coop_unavailable;
  const int want = tune ? tune_pick(c, phase, tune_key(D, phase, A.cscene != nullptr, A.C, nw), model, max_blocks)
                        : model;
  const int blocks = c->coop_unavailable ? 1 : std::max(1, std::min(max_blocks, want));
  A.tail_col = (blocks > 1) ? (uint32_t)tail : (uint32_t)kMaxColours;
  // the component phase needs several workgroups, and lists that hold the phase's contacts
  // with room for an uneven split (else the tail as before; a launch whose split still
  // overflows a list falls back to it by itself)
  if (blocks > 1 && comp < ncol && comp_contacts <= (uint64_t)blocks * kCompCap / 2) A.comp_col = (uint32_t)comp;
  else A.comp_parent = nullptr;
  if (lds_path) {
    // the whole solve in one workgroup's LDS
    hipError_t e = hipFuncSetAttribute((const void*)k_gs_lds<D, Real>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds);
    if (e != hipSuccess) return e;
    k_gs_lds<D, Real><<<1, kGSBlock, lds, c->stream>>>(A);
    return hipGetLastError();
  }
  void* args[] = {&A};
  hipError_t e = tune ? hipEventRecord(c->tuner.pend[phase].ev[0], c->stream) : hipSuccess;
  if (e != hipSuccess) return e;
  const void* fn = coh ? (const void*)k_gs<D, Real, true> : (const void*)k_gs<D, Real, false>;
  bool launched = false;
  e = launch_cooperative(c, fn, blocks, kGSBlock, args, 0, c->stream, &launched);
  if (e == hipSuccess && !launched) {
    // one workgroup (chosen, or after a refused cooperative launch: see
    // scx_ctx::coop_unavailable) sweeps every colour itself, so there is no tail split
    A.tail_col = (uint32_t)kMaxColours;
    e = hipLaunchKernel(fn, dim3(1), dim3(kGSBlock), args, 0, c->stream);
  }
  if (e == hipSuccess && tune) e = hipEventRecord(c->tuner.pend[phase].ev[1], c->stream);
  return e;
}

template <int D, typename Real>
int gs_max_blocks(scx_ctx* c) {
  int per_cu = 0;
  int per_cu_coh = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_gs<D, Real, false>, kGSBlock, 0) !=
          hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_coh, (const void*)k_gs<D, Real, true>, kGSBlock, 0) !=
          hipSuccess)
    per_cu = per_cu_coh = 1;
  return std::max(1, std::min(per_cu, per_cu_coh)) * c->num_cus;
}

// Ensembles: the scene of each contact in colour order (the scene of its body a).
__global__ void __launch_bounds__(kBlock) k_contact_scene(uint32_t C, const uint32_t* __restrict__ a_slot,
                                                          const uint32_t* __restrict__ sid,
                                                          const uint32_t* __restrict__ bscene,
                                                          uint32_t* __restrict__ cscene) {
  const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= C) return;
  const uint32_t sl = a_slot[k];
  cscene[k] = bscene[sid ? sid[sl] : sl];
}

template <int D>
static hipError_t solve_dim(scx_ctx* c, uint32_t C, const uint32_t* col_off, const scx_step_params* p) {
  hipStream_t s = c->stream;
  if (c->tuner.enabled) tune_resolve(c);
  const bool fp32 = p->precision == SCX_PRECISION_FP32_POLISH;
  const bool warm = p->warm_start != SCX_WARM_NONE;
  const uint32_t nb = (c->n + kBlock - 1) / kBlock;
  if (nb)
    k_body_init<D><<<nb, kBlock, 0, s>>>(c->n, c->d.sid, c->vsolve, c->d.w0, c->d.m, c->d.inertia, c->d.bv64,
                                          c->d.bv32, fp32 ? 1 : 0);
  hipError_t e = hipGetLastError();
  timing_mark(c, "gs_init");
  if (e != hipSuccess || C == 0) return e;
  const bool ens = c->nscenes > 0;
  if (ens) {
    k_contact_scene<<<(C + kBlock - 1) / kBlock, kBlock, 0, s>>>(C, c->d.c64.a, c->d.sid, c->d.bscene, c->d.cscene);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemsetAsync(c->d.scene_sweeps, 0xFF, sizeof(uint32_t) * c->nscenes, s);
    if (e != hipSuccess) return e;
  }
  if (fp32) {
    GSArgs<float> A{c->d.c32, c->d.bv32, c->d.sid, c->vsolve, c->d.w0, c->n, C, c->d.coloff,
                    reinterpret_cast<float*>(c->d.rbest), c->d.partial, c->d.ctr, c->d.barrier, p->max_iters,
                    p->eval_every, p->tol, 1, warm ? 1 : 0};
    A.lmap = c->d.lmap;
    if (!ens && c->d.viol && c->violator_skip) {
      // the violator list of the FP32 phase starts empty
      e = hipMemsetAsync(c->d.viol, 0, sizeof(uint32_t) * (kViolCap + 3), s);
      if (e != hipSuccess) return e;
      A.viol = c->d.viol;
      A.vflag = c->d.viol + 2 + kViolCap;
    }
    if (!ens && c->comp_sweeps) {
      A.comp_parent = c->d.comp_parent;
      A.comp_cnt = c->d.comp_cnt;
      A.comp_list = c->d.comp_list;
      A.comp_lds = c->comp_lds ? 1 : 0;
    }
    if (ens) {
      A.cscene = c->d.cscene;
      A.scene_acc = c->d.scene_acc;
      A.scene_sweeps = c->d.scene_sweeps;
      A.nscenes = c->nscenes;
      if (c->ens_compact) {
        A.act_idx = c->d.act_idx;
        A.act_part = c->d.act_part;
      }
    }
    e = coop_gs<D, float>(c, A, c->gs_blocks, col_off);
    if (e != hipSuccess) return e;
    timing_mark(c, "gs_fp32");
    const uint32_t nr = (3 * C + kBlock - 1) / kBlock;
    k_promote_r<<<nr, kBlock, 0, s>>>(C, c->d.c32.r, c->d.c64.r);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    timing_mark(c, "gs_promote");
  }
  GSArgs<double> B{c->d.c64, c->d.bv64, c->d.sid, c->vsolve, c->d.w0, c->n, C, c->d.coloff, c->d.rbest,
                   c->d.partial, c->d.ctr, c->d.barrier, p->max_iters, p->eval_every, p->tol, fp32 ? 2 : 0,
                   (fp32 || warm) ? 1 : 0};
  B.lmap = c->d.lmap;
  if (!ens && c->d.viol && c->violator_skip) {
    e = hipMemsetAsync(c->d.viol, 0, sizeof(uint32_t) * (kViolCap + 3), s);
    if (e != hipSuccess) return e;
    B.viol = c->d.viol;
    B.vflag = c->d.viol + 2 + kViolCap;
  }
  if (!ens && c->comp_sweeps) {
    B.comp_parent = c->d.comp_parent;
    B.comp_cnt = c->d.comp_cnt;
    B.comp_list = c->d.comp_list;
    B.comp_lds = c->comp_lds ? 1 : 0;
  }
  if (ens) {
    B.cscene = c->d.cscene;
    B.scene_acc = c->d.scene_acc;
    B.scene_sweeps = c->d.scene_sweeps;
    B.nscenes = c->nscenes;
    if (c->ens_compact) {
      B.act_idx = c->d.act_idx;
      B.act_part = c->d.act_part;
    }
  }
  e = coop_gs<D, double>(c, B, c->gs_blocks, col_off);
  if (e != hipSuccess) return e;
  return tune_after_solve(c);
}

hipError_t launch_solve(scx_ctx* c, uint32_t C, const uint32_t* col_off, const scx_step_params* p) {
  return c->dim == 2 ? solve_dim<2>(c, C, col_off, p) : solve_dim<3>(c, C, col_off, p);
}

hipError_t launch_finish(scx_ctx* c, uint32_t C, const scx_step_params* p) {
  hipStream_t s = c->stream;
  // geometric map: the unconstrained map from q0 with v2 (GeometricImpactFrictionMap.cpp:441);
  // symplectic Euler and stabilized: q1 = q0 + dt v1 with v1 the solve's result
  // (SymplecticEulerImpactFrictionMap.cpp:405, StabilizedImpactFrictionMap.cpp:181)
  const int integ = (p->staggering == SCX_STAGGER_GEOMETRIC) ? p->integrator : kFlowLinear;
  const uint32_t nb = (c->n + kBlock - 1) / kBlock;
  if (nb) {
    if (c->dim == 2)
      k_finish_bodies<2><<<nb, kBlock, 0, s>>>(c->n, c->d.slot_map, c->d.bv64, c->d.q0, c->d.m, p->dt, p->g[0], p->g[1],
                                               p->g[2], integ, c->d.q, c->d.v, c->d.w, c->d.v2, c->d.w2);
    else
      k_finish_bodies<3><<<nb, kBlock, 0, s>>>(c->n, c->d.slot_map, c->d.bv64, c->d.q0, c->d.m, p->dt, p->g[0], p->g[1],
                                               p->g[2], integ, c->d.q, c->d.v, c->d.w, c->d.v2, c->d.w2);
  }
  const uint32_t nc = (C + kBlock - 1) / kBlock;
  if (nc) k_impulses_out<<<nc, kBlock, 0, s>>>(C, c->dim, c->d.perm, c->d.c64.r, c->d.calpha, c->d.cbeta, p->mu);
  return hipGetLastError();
}

hipError_t launch_impulses(scx_ctx* c, uint32_t C) {
  const uint32_t nc = (C + kBlock - 1) / kBlock;
  if (nc) k_impulses_out<<<nc, kBlock, 0, c->stream>>>(C, c->dim, c->d.perm, c->d.c64.r, c->d.calpha, c->d.cbeta, 1.0);
  return hipGetLastError();
}

// Per-scene contacts and residual (max ||FB||^2 over the scene's contacts) of the last
// solve of an ensemble, from the FP64 sweep records and body records it left.
template <int D>
__global__ void __launch_bounds__(kBlock) k_scene_stats(GSArgs<double> A, const uint32_t* __restrict__ bscene,
                                                        double* __restrict__ err, uint32_t* __restrict__ cnt) {
  const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= A.C) return;
  const uint32_t sl = A.c.a[k];
  const uint32_t sc = bscene[A.sid ? A.sid[sl] : sl];
  const double e = contact_err<D, double>(A, k);
  // residuals are >= 0, so their bit patterns order like their values
  atomicMax(reinterpret_cast<unsigned long long*>(err) + sc, (unsigned long long)__double_as_longlong(e));
  atomicAdd(cnt + sc, 1u);
}

hipError_t launch_scene_stats(scx_ctx* c, uint32_t C) {
  DeviceBuffers& d = c->d;
  hipError_t e = hipMemsetAsync(d.scene_err, 0, sizeof(double) * c->nscenes, c->stream);
  if (e == hipSuccess) e = hipMemsetAsync(d.scene_cnt, 0, sizeof(uint32_t) * c->nscenes, c->stream);
  const uint32_t nc = (C + kBlock - 1) / kBlock;
  if (e != hipSuccess || nc == 0) return e;
  GSArgs<double> A{};
  A.c = d.c64;
  A.body = d.bv64;
  A.sid = d.sid;
  A.nbodies = c->n;
  A.C = C;
  if (c->dim == 2)
    k_scene_stats<2><<<nc, kBlock, 0, c->stream>>>(A, d.bscene, d.scene_err, d.scene_cnt);
  else
    k_scene_stats<3><<<nc, kBlock, 0, c->stream>>>(A, d.bscene, d.scene_err, d.scene_cnt);
  return hipGetLastError();
}

int solver_max_blocks(scx_ctx* c) {
  return c->dim == 2 ? std::min(gs_max_blocks<2, double>(c), gs_max_blocks<2, float>(c))
                     : std::min(gs_max_blocks<3, double>(c), gs_max_blocks<3, float>(c));
}

}  // namespace scx
