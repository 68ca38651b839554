// scx_api.hip -- host side of libscx: context, device buffers, the per-step
// orchestration of ImpactFrictionMap::flow (GeometricImpactFrictionMap.cpp:267-442)
// and the C ABI declared in include/scx.h.
#pragma clang fp contract(off)

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "scx_device.h"
#include "scx_flow.h"
#include "scx_internal.h"

using namespace scx;

// Phase boundary for scx_get_timing: an event on the context stream.
void scx::timing_mark(scx_ctx* c, const char* name) {
  if (!c->timing) return;
  hipEvent_t ev;
  if (hipEventCreate(&ev) != hipSuccess) return;
  hipEventRecord(ev, c->stream);
  c->events.push_back(ev);
  c->event_names.push_back(name);
}

// Scratch of scx_potential_overlaps_noalloc: file-scope device memory, one copy on every
// device the library runs on, held until the process ends. In host code `scratch` is only a
// stand-in for the symbol; its device address on each device comes from
// hipGetSymbolAddress (scratch_address below) and is the only way host code refers to it.
constexpr size_t kNoAllocScratchBytes = (size_t)40 << 20;  // 40 MiB
constexpr uint32_t kNoAllocMaxBoxes = 1u << 16;            // 65,536 boxes
constexpr uint32_t kNoAllocMaxPairs = 1u << 20;            // 1,048,576 pairs
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

const char* kVersion = "scx 0.1 (gfx950)";

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_addr[kMaxDevices] = {};

// The scratch array's address on `device`, which must be current; looked up once per device.
hipError_t scratch_address(int device, unsigned char** out) {
  *out = nullptr;
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_addr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_addr[device] = static_cast<unsigned char*>(p);
  }
  *out = g_scratch_addr[device];
  return hipSuccess;
}

int fail(scx_ctx* c, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}

#define HIPC(call)                                                                                      \
  do {                                                                                                  \
    hipError_t e_ = (call);                                                                             \
    if (e_ != hipSuccess) return fail(c, SCX_E_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                                      __FILE__, __LINE__);                                              \
  } while (0)

template <typename T>
hipError_t dalloc(T** p, size_t count) {
  if (*p) {
    hipFree(*p);
    *p = nullptr;
  }
  if (count == 0) count = 1;
  return hipMalloc((void**)p, sizeof(T) * count);
}

template <typename T>
void dfree(T** p) {
  if (*p) hipFree(*p);
  *p = nullptr;
}

uint32_t next_pow2(uint64_t v) {
  uint64_t p = 1;
  while (p < v) p <<= 1;
  return (uint32_t)std::min<uint64_t>(p, 1ull << 31);
}

// After a step with stats: the share of the contacts in the component phase for the next
// step (one scene; comp_share_update).
void comp_adapt_share(scx_ctx* c, const StepCounters& h) {
  if (!c->comp_adapt || !c->comp_sweeps || c->nscenes || h.sweeps == 0) return;
  comp_share_update(c->comp_share, c->comp_ceiling, h.comp_overflows, h.comp_launches, h.comp_maxw,
                    (uint32_t)kGSBlock);
}

// Body-sized buffers.
int alloc_bodies(scx_ctx* c, uint32_t N) {
  const int D = c->dim;
  DeviceBuffers& d = c->d;
  // grid entries carry a 29-bit body position
  if (N >= (1u << 29)) return fail(c, SCX_E_CAPACITY, "at most %u bodies", (1u << 29) - 1);
  HIPC(dalloc(&d.q, (size_t)D * N));
  HIPC(dalloc(&d.v, (size_t)D * N));
  HIPC(dalloc(&d.w, 3 * (size_t)N));
  HIPC(dalloc(&d.r, N));
  HIPC(dalloc(&d.m, N));
  HIPC(dalloc(&d.inertia, N));
  HIPC(dalloc(&d.q0, (size_t)D * N));
  HIPC(dalloc(&d.v0, (size_t)D * N));
  HIPC(dalloc(&d.w0, 3 * (size_t)N));
  HIPC(dalloc(&d.q1p, (size_t)D * N));
  HIPC(dalloc(&d.v1p, (size_t)D * N));
  HIPC(dalloc(&d.v2, (size_t)D * N));
  HIPC(dalloc(&d.w2, 3 * (size_t)N));
  HIPC(dalloc(&d.bv64, 8 * (size_t)N));
  HIPC(dalloc(&d.bv32, 8 * (size_t)N));
  HIPC(dalloc(&d.slot_of, N));
  HIPC(dalloc(&d.lmap, N));
  HIPC(hipMemset(d.lmap, 0xFF, sizeof(uint32_t) * N));
  HIPC(dalloc(&d.comp_parent, N));
  d.sid = nullptr;
  d.slot_map = nullptr;
  HIPC(dalloc(&d.aabb_min, (size_t)D * N));
  HIPC(dalloc(&d.aabb_max, (size_t)D * N));
  HIPC(dalloc(&d.red, 9 * (size_t)((N + kBlock - 1) / kBlock + 1)));
  HIPC(dalloc(&d.pcount, (size_t)N + 1));
  HIPC(dalloc(&d.pmask, std::max<size_t>(N, c->cap_contacts)));
  HIPC(dalloc(&d.deg, 2 * ((size_t)N + 1)));
  HIPC(dalloc(&d.order, N));
  HIPC(dalloc(&d.order_next, N));
  HIPC(dalloc(&d.tbox, 2 * (size_t)N));
  HIPC(dalloc(&d.tnar, 2 * (size_t)N));
  HIPC(dalloc(&d.tr, N));
  c->order_valid = false;
  // dense grid budget: cells of the reference's size (mean box extent) number about N
  // for a packed scene; the grid is coarsened on the device beyond this budget
  c->cap_cells = 2 * N + 1024;
  HIPC(dalloc(&d.cells, 4 * ((size_t)c->cap_cells + 1)));
  c->cap_bodies = N;
  return SCX_OK;
}

int alloc_entries(scx_ctx* c, uint32_t cap_entries) {
  DeviceBuffers& d = c->d;
  c->cap_entries = cap_entries;
  HIPC(dalloc(&d.ent, cap_entries));
  return SCX_OK;
}

template <typename Real>
int alloc_soa(scx_ctx* c, ContactSoA<Real>& s, size_t C) {
  HIPC(dalloc(&s.a, C));
  HIPC(dalloc(&s.b, C));
  HIPC(dalloc(&s.n, 3 * C));
  HIPC(dalloc(&s.s, 3 * C));
  HIPC(dalloc(&s.la, C));
  HIPC(dalloc(&s.lb, C));
  HIPC(dalloc(&s.wn, C));
  HIPC(dalloc(&s.wt, C));
  HIPC(dalloc(&s.mu, C));
  HIPC(dalloc(&s.forcing, C));
  HIPC(dalloc(&s.r, 3 * C));
  return SCX_OK;
}

template <typename Real>
void free_soa(ContactSoA<Real>& s) {
  dfree(&s.a);
  dfree(&s.b);
  dfree(&s.n);
  dfree(&s.s);
  dfree(&s.la);
  dfree(&s.lb);
  dfree(&s.wn);
  dfree(&s.wt);
  dfree(&s.mu);
  dfree(&s.forcing);
  dfree(&s.r);
}

int alloc_contacts(scx_ctx* c, uint32_t C) {
  const int D = c->dim;
  DeviceBuffers& d = c->d;
  c->cap_contacts = C;
  HIPC(dalloc(&d.ca, C));
  HIPC(dalloc(&d.cb, C));
  HIPC(dalloc(&d.cplane, C));
  HIPC(dalloc(&d.basis, (size_t)D * D * C));
  HIPC(dalloc(&d.cla, C));
  HIPC(dalloc(&d.clb, C));
  HIPC(dalloc(&d.cwn, C));
  HIPC(dalloc(&d.cwt, C));
  HIPC(dalloc(&d.cforcing, C));
  HIPC(dalloc(&d.cmu, C));
  HIPC(dalloc(&d.cscene, C));
  HIPC(dalloc(&d.calpha, C));
  HIPC(dalloc(&d.cbeta, (size_t)(D - 1) * C));
  HIPC(dalloc(&d.crinit, (size_t)D * C));
  HIPC(dalloc(&d.cur_key, C));
  HIPC(dalloc(&d.mlp_x, (size_t)16 * C));
  HIPC(dalloc(&d.mlp_y, (size_t)3 * C));
  HIPC(dalloc(&d.mlp_s, C));
  d.features_valid = false;
  HIPC(dalloc(&d.adj, 2 * (size_t)C));
  HIPC(dalloc(&d.cstate, C));
  c->cap_perm = std::max<uint32_t>(C, c->cap_bodies);
  HIPC(dalloc(&d.perm, c->cap_perm));
  HIPC(dalloc(&d.rbest, 3 * (size_t)C));
  HIPC(dalloc(&d.act_idx, 2 * (size_t)C));
  if (alloc_soa(c, d.c64, C) != SCX_OK) return SCX_E_HIP;
  if (alloc_soa(c, d.c32, C) != SCX_OK) return SCX_E_HIP;
  HIPC(dalloc(&d.pmask, std::max<size_t>(c->cap_bodies, C)));
  const size_t nblk = (std::max<size_t>(C, c->cap_bodies) + kBlock - 1) / kBlock;
  HIPC(dalloc(&d.mbcount, (size_t)kMaxColours * nblk + 2));
  return SCX_OK;
}

int alloc_scan_tmp(scx_ctx* c) {
  const size_t big = std::max<size_t>({(size_t)c->cap_cells + 1, (size_t)c->cap_bodies + 1, (size_t)c->cap_pairs,
                                       (size_t)kMaxColours * ((std::max<size_t>(c->cap_contacts, c->cap_bodies) +
                                                               kBlock - 1) / kBlock) + 2});
  HIPC(dalloc(&c->d.scan_tmp, scan_tmp_words((uint32_t)std::min<size_t>(big, 0xFFFFFFFFu)) + 4));
  return SCX_OK;
}

void free_all(scx_ctx* c) {
  DeviceBuffers& d = c->d;
  double* dbl[] = {d.q, d.v, d.w, d.r, d.m, d.inertia, d.q0, d.v0, d.w0, d.q1p, d.v1p, d.v2, d.w2, d.bv64,
                   d.px, d.pn, d.aabb_min, d.aabb_max, d.red, d.basis, d.cla, d.clb, d.cwn, d.cwt, d.cforcing,
                   d.calpha, d.cbeta, d.crinit, d.rbest, d.partial, d.prev_r, d.scene_mu, d.scene_shift,
                   d.scene_err, d.cmu};
  for (double* p : dbl)
    if (p) hipFree(p);
  if (d.bv32) hipFree(d.bv32);
  dfree(&d.slot_of);
  dfree(&d.lmap);
  dfree(&d.bscene);
  dfree(&d.scene_poff);
  dfree(&d.scene_cnt);
  dfree(&d.cscene);
  dfree(&d.scene_acc);
  dfree(&d.scene_sweeps);
  dfree(&d.act_idx);
  dfree(&d.act_part);
  dfree(&d.viol);
  dfree(&d.comp_parent);
  dfree(&d.comp_cnt);
  dfree(&d.comp_list);
  uint32_t* u32[] = {d.cells, d.ent, d.pcount, d.scan_tmp, d.pa, d.pb, d.ca, d.cb, d.mbcount, d.deg, d.adj,
                     d.perm, d.coloff};
  for (uint32_t* p : u32)
    if (p) hipFree(p);
  if (d.cplane) hipFree(d.cplane);
  if (d.cstate) hipFree(d.cstate);
  dfree(&d.order);
  dfree(&d.order_next);
  dfree(&d.tbox);
  dfree(&d.tnar);
  dfree(&d.tr);
  if (d.pmask) hipFree(d.pmask);
  float* mlpf[] = {d.mlp_params, d.mlp_grad, d.mlp_m, d.mlp_v, d.mlp_b, d.mlp_sse, d.mlp_x, d.mlp_y, d.mlp_s,
                   d.mlp_pool_x, d.mlp_pool_t};
  for (float* p : mlpf)
    if (p) hipFree(p);
  if (d.mlp_w) hipFree(d.mlp_w);
  if (d.mlp_act) hipFree(d.mlp_act);
  if (d.mlp_idx) hipFree(d.mlp_idx);
  if (d.prev_key) hipFree(d.prev_key);
  if (d.cur_key) hipFree(d.cur_key);
  if (d.grid) hipFree(d.grid);
  if (d.ctr) hipFree(d.ctr);
  if (d.barrier) hipFree(d.barrier);
  if (d.ctr_host) hipHostFree(d.ctr_host);
  if (d.host_u32) hipHostFree(d.host_u32);
  free_soa(d.c64);
  free_soa(d.c32);
  d = DeviceBuffers{};
}

void timing_collect(scx_ctx* c) {
  if (!c->timing || c->events.size() < 2) {
    for (hipEvent_t e : c->events) hipEventDestroy(e);
    c->events.clear();
    c->event_names.clear();
    return;
  }
  hipEventSynchronize(c->events.back());
  c->phase_ms.clear();
  c->phase_names.clear();
  for (size_t k = 1; k < c->events.size(); ++k) {
    float ms = 0.f;
    hipEventElapsedTime(&ms, c->events[k - 1], c->events[k]);
    c->phase_ms.push_back(ms);
    c->phase_names.push_back(c->event_names[k]);
  }
  for (hipEvent_t e : c->events) hipEventDestroy(e);
  c->events.clear();
  c->event_names.clear();
}

// One geometric-map step. Returns SCX_OK, or SCX_E_CAPACITY after growing a buffer
// (the caller retries the step from the untouched start-of-step state).
int step_once(scx_ctx* c, const scx_step_params* p, scx_step_stats* st, bool want_stats) {
  hipStream_t s = c->stream;
  DeviceBuffers& d = c->d;
  const uint32_t n = c->n;
  const int D = c->dim;
  if (c->events.empty()) timing_mark(c, "start");
  HIPC(hipMemsetAsync(d.ctr, 0, sizeof(StepCounters), s));
  d.features_valid = false;
  if (!c->order_valid) {
    HIPC(launch_order_reset(c, n));
    c->order_valid = true;
  }
  // start-of-step state q0, v0, omega0 (the resident state becomes the step input)
  std::swap(d.q, d.q0);
  std::swap(d.v, d.v0);
  std::swap(d.w, d.w0);
  // predictor + boxes: the unconstrained map (GeometricImpactFrictionMap.cpp:279,
  // StabilizedImpactFrictionMap.cpp:60) or q0 + dt (v0 + dt M^-1 F)
  // (SymplecticEulerImpactFrictionMap.cpp:222-232). The symplectic Euler map's solve on v0
  // with forcing N^T vdelta, D^T vdelta (:296-308) is the same problem as a solve that starts
  // from v0 + vdelta with no extra forcing, whose result is already v1 = v2 + vdelta (:368):
  // the force increment is folded into the start velocity, so the sweep records carry no
  // tangential forcing. The stabilized map solves on the predictor velocity (:117).
  const int stag = p->staggering;
  HIPC(launch_predict_aabb(c, d.q0, d.v0, d.m, p->dt, p->g,
                           stag == SCX_STAGGER_SYMPLECTIC_EULER ? kFlowForceIncrement : p->integrator, d.q1p,
                           stag == SCX_STAGGER_GEOMETRIC ? nullptr : d.v1p));
  c->vbasis = (stag == SCX_STAGGER_STABILIZED) ? d.v1p : d.v0;
  c->vsolve = (stag == SCX_STAGGER_GEOMETRIC) ? d.v0 : d.v1p;
  timing_mark(c, "predict");
  // broad phase binning
  HIPC(launch_grid_build(c, n, false));
  timing_mark(c, "grid");
  // body-body active set (Ball2DSim.cpp:553-608 / RigidBody3DSim.cpp:1072-1260)
  const int filter = (D == 2) ? 1 : 2;
  HIPC(launch_pairs(c, n, 0, filter, d.ca, d.cb, c->cap_contacts));
  timing_mark(c, "pairs_count");
  HIPC(launch_exclusive_scan(d.pcount, d.pcount, n, d.scan_tmp, s));
  timing_mark(c, "pairs_scan");
  HIPC(launch_pairs(c, n, 1, filter, d.ca, d.cb, c->cap_contacts));
  timing_mark(c, "pairs_write");
  // this binning's home-cell order is the next step's sweep order, and this step's body
  // slots for the Gauss-Seidel sweeps
  std::swap(d.order, d.order_next);
  HIPC(launch_slots(c, n));
  // plane contacts, plane-major (Ball2DSim.cpp:747-764 / RigidBody3DSim.cpp:1414-1502)
  // (ensembles: buckets are scene-local plane indices, so every scene's plane contacts
  // keep their own plane-major order)
  uint32_t npl = 0;
  const uint32_t nbuckets = c->nscenes ? c->scene_planes : c->nplanes;
  if (nbuckets > 0) {
    HIPC(launch_plane_masks(c, n));
    HIPC(launch_multibucket(c, d.pmask, n, nbuckets, d.perm, c->cap_perm, d.coloff));
  }
  timing_mark(c, "planes");
  // sync point: sizes of the active set
  HIPC(hipMemcpyAsync(d.ctr_host, d.ctr, sizeof(StepCounters), hipMemcpyDeviceToHost, s));
  HIPC(hipMemcpyAsync(d.host_u32, d.pcount + n, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  if (nbuckets > 0)
    HIPC(hipMemcpyAsync(d.host_u32 + 1, d.coloff + nbuckets, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  HIPC(hipStreamSynchronize(s));
  const uint32_t nbb = d.host_u32[0];
  if (nbuckets > 0) npl = d.host_u32[1];
  const uint64_t nentries = d.ctr_host->nentries;
  if ((d.ctr_host->overflow & 1u) || nentries > c->cap_entries) {
    std::swap(d.q, d.q0);
    std::swap(d.v, d.v0);
    std::swap(d.w, d.w0);
    const uint32_t want = (uint32_t)std::min<uint64_t>(0xFFFFFFFFull, nentries + nentries / 4 + 1024);
    if (alloc_entries(c, want) != SCX_OK || alloc_scan_tmp(c) != SCX_OK) return SCX_E_HIP;
    return SCX_E_CAPACITY;
  }
  const uint64_t C64 = (uint64_t)nbb + npl;
  if ((d.ctr_host->overflow & 6u) || C64 > c->cap_contacts || npl > c->cap_perm) {
    std::swap(d.q, d.q0);
    std::swap(d.v, d.v0);
    std::swap(d.w, d.w0);
    if (C64 > 0x7FFFFFFFull) return fail(c, SCX_E_CAPACITY, "active set too large (%llu)", (unsigned long long)C64);
    if (alloc_contacts(c, (uint32_t)(C64 + C64 / 4 + 1024)) != SCX_OK || alloc_scan_tmp(c) != SCX_OK)
      return SCX_E_HIP;
    return SCX_E_CAPACITY;
  }
  const uint32_t C = (uint32_t)C64;
  if (C > 0) {
    const bool stab = stag == SCX_STAGGER_SYMPLECTIC_EULER && p->stabilize;
    HIPC(launch_contact_setup(c, nbb, npl, p->CoR, stab ? p->dt : 0.0, p->penetration_threshold));
    HIPC(launch_contact_mu(c, C));
    // warm start from the impulse cache (GeometricImpactFrictionMap.cpp:306); the learned
    // warm start reads the transported cache as an input and predicts the impulses
    if (p->warm_start != SCX_WARM_NONE) {
      const bool learned = p->warm_start == SCX_WARM_LEARNED;
      HIPC(launch_warm_start(c, C, learned ? SCX_WARM_TRANSPORT : p->warm_start));
      if (p->warm_start >= SCX_WARM_TRANSPORT) {
        double g2 = 0.0;
        for (int a = 0; a < D; ++a) g2 += p->g[a] * p->g[a];
        const double vref = (g2 > 0.0 ? std::sqrt(g2) : 1.0) * p->dt;
        HIPC(launch_contact_features(c, C, p->mu, vref));
        d.features_valid = true;
        if (learned) {
          timing_mark(c, "features");
          HIPC(launch_mlp_infer(c, C, d.mlp_x, d.mlp_y));
          timing_mark(c, "mlp");
          HIPC(launch_apply_prediction(c, C, p->mu));
        }
      }
    }
    timing_mark(c, "setup");
    HIPC(launch_colouring(c, C));
    HIPC(launch_colour_sort(c, C));
    // second sync point: colour sizes size the persistent sweep launch; a contact graph
    // that needs more colours than the sweeps support stops the step here, before any
    // impulse or body state changes
    HIPC(hipMemcpyAsync(d.host_u32 + 2, d.coloff, sizeof(uint32_t) * (kMaxColours + 1), hipMemcpyDeviceToHost, s));
    HIPC(hipMemcpyAsync(d.host_u32 + kOverflowWord, &d.ctr->overflow, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    if (d.host_u32[kOverflowWord] & 8u) {
      std::swap(d.q, d.q0);
      std::swap(d.v, d.v0);
      std::swap(d.w, d.w0);
      timing_collect(c);
      return fail(c, SCX_E_COLOURS, "contact graph needs more than %d colours (a body with too many contacts)",
                  kMaxColours);
    }
    timing_mark(c, "colour");
    HIPC(launch_gather_records(c, C, p->precision == SCX_PRECISION_FP32_POLISH, p->mu, c->nscenes ? d.cmu : nullptr,
                               p->warm_start != SCX_WARM_NONE ? d.crinit : nullptr));
    timing_mark(c, "gather");
  }
  HIPC(launch_solve(c, C, d.host_u32 + 2, p));
  timing_mark(c, "gs_fp64");
  HIPC(launch_finish(c, C, p));
  // the stabilized map with an empty active set keeps the unconstrained predictor
  // (StabilizedImpactFrictionMap.cpp:66-78); its velocity is already v1p
  if (stag == SCX_STAGGER_STABILIZED && C == 0)
    HIPC(hipMemcpyAsync(d.q, d.q1p, sizeof(double) * D * n, hipMemcpyDeviceToDevice, s));
  // cache this step's impulses (GeometricImpactFrictionMap.cpp:419); an empty active set
  // empties the cache (:286-288); otherwise a step without caching (NONE) leaves it
  // untouched, as the reference's map never touches its cache in that mode; the
  // stabilized map has no cache
  if (stag != SCX_STAGGER_STABILIZED && (p->warm_start != SCX_WARM_NONE || C == 0)) d.prev_count = 0;
  if (p->warm_start != SCX_WARM_NONE && C > 0) {
    if (C > d.cap_cache) {
      const uint32_t want = C + C / 4 + 1024;
      HIPC(dalloc(&d.prev_key, want));
      HIPC(dalloc(&d.prev_r, (size_t)6 * want));  // kCacheWords per entry
      d.cap_cache = want;
    }
    HIPC(launch_cache_store(c, C, p->warm_start == SCX_WARM_LEARNED ? SCX_WARM_TRANSPORT : p->warm_start));
    d.prev_count = C;
    d.prev_layout = (p->warm_start >= SCX_WARM_TRANSPORT) ? 1 : 0;
  }
  timing_mark(c, "finish");
  c->last_ncontacts = C;
  c->last_npairs = nbb;
  c->last_nplane = npl;
  if (want_stats || c->timing) {
    HIPC(hipMemcpyAsync(d.ctr_host, d.ctr, sizeof(StepCounters), hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    c->last_ncolours = d.ctr_host->ncolours;
    comp_adapt_share(c, *d.ctr_host);
    if (c->violator_skip == 2 && p->precision == SCX_PRECISION_FP32_POLISH && c->nscenes == 0 &&
        d.ctr_host->sweeps_fp32 >= 64 && d.ctr_host->evals_skipped == 0)
      return fail(c, SCX_E_STATE, "no FP32 evaluation skipped in %u sweeps", d.ctr_host->sweeps_fp32);
    if (st) {
      st->nentries = nentries;
      st->npairs = nbb;
      st->nplane_contacts = npl;
      st->ncontacts = C;
      st->ncolours = d.ctr_host->ncolours;
      st->sweeps = d.ctr_host->sweeps;
      st->sweeps_fp32 = d.ctr_host->sweeps_fp32;
      st->error = (C == 0) ? 0.0 : d.ctr_host->error;
      st->succeeded = (C == 0) ? 1 : (d.ctr_host->error < p->tol ? 1 : 0);
      st->ncells = d.ctr_host->ncells;
    }
  }
  timing_collect(c);
  return SCX_OK;
}

int check_params(scx_ctx* c, const scx_step_params* p) {
  if (!p) return fail(c, SCX_E_ARG, "null step parameters");
  if (!c->has_bodies) return fail(c, SCX_E_STATE, "scx_set_bodies must be called before stepping");
  if (!(p->dt > 0.0)) return fail(c, SCX_E_ARG, "dt must be positive");
  if (!(p->mu >= 0.0)) return fail(c, SCX_E_ARG, "mu must be non-negative");
  if (!(p->CoR >= 0.0 && p->CoR <= 1.0)) return fail(c, SCX_E_ARG, "CoR must lie in [0, 1]");
  if (!(p->tol >= 0.0)) return fail(c, SCX_E_ARG, "tol must be non-negative");
  if (p->integrator != SCX_INTEGRATOR_SYMPLECTIC_EULER && p->integrator != SCX_INTEGRATOR_VERLET)
    return fail(c, SCX_E_ARG, "unknown integrator %d", p->integrator);
  if (p->precision != SCX_PRECISION_FP64 && p->precision != SCX_PRECISION_FP32_POLISH)
    return fail(c, SCX_E_ARG, "unknown precision mode %d", p->precision);
  if (p->warm_start < SCX_WARM_NONE || p->warm_start > SCX_WARM_LEARNED)
    return fail(c, SCX_E_ARG, "unknown warm start mode %d", p->warm_start);
  if (p->staggering < SCX_STAGGER_GEOMETRIC || p->staggering > SCX_STAGGER_STABILIZED)
    return fail(c, SCX_E_ARG, "unknown staggering %d", p->staggering);
  // RigidBody3DSceneParser.cpp:1881-1888
  if (p->staggering == SCX_STAGGER_SYMPLECTIC_EULER && p->CoR != 0.0)
    return fail(c, SCX_E_ARG, "the symplectic Euler staggering needs CoR = 0");
  if (p->staggering == SCX_STAGGER_STABILIZED && p->warm_start != SCX_WARM_NONE)
    return fail(c, SCX_E_ARG, "the stabilized staggering has no impulse cache (warm_start must be SCX_WARM_NONE)");
  if (!(p->penetration_threshold >= 0.0)) return fail(c, SCX_E_ARG, "penetration_threshold must be non-negative");
  if (p->warm_start == SCX_WARM_LEARNED && !c->d.mlp_ready)
    return fail(c, SCX_E_STATE, "learned warm start needs a model (scx_set_warm_model or scx_train_warm_model)");
  // GeometricImpactFrictionMap.cpp:258-262: "Caching in 3 space not yet supported."
  if (p->warm_start == SCX_WARM_CACHE && c->dim == 3)
    return fail(c, SCX_E_ARG, "normal and friction impulse caching is 2D only (use SCX_WARM_NORMAL)");
  if (c->nscenes == 0 && c->nplanes > kMaxScenePlanes)
    return fail(c, SCX_E_ARG, "at most %u static planes per scene (use scx_set_scenes for ensembles)",
                kMaxScenePlanes);
  return SCX_OK;
}

}  // namespace

extern "C" {

const char* scx_version(void) { return kVersion; }

const char* scx_last_error(const scx_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int scx_create(const scx_config* cfg, scx_ctx** out) {
  if (!out) return SCX_E_ARG;
  *out = nullptr;
  if (!cfg || (cfg->dim != 2 && cfg->dim != 3) || cfg->max_bodies == 0 || cfg->max_planes > kMaxPlanesTotal)
    return SCX_E_ARG;
  scx_ctx* c = new scx_ctx;
  c->dim = cfg->dim;
  c->device = cfg->device;
  int rc = SCX_OK;
  do {
    hipError_t e = hipSetDevice(cfg->device);
    if (e != hipSuccess) {
      rc = fail(c, SCX_E_HIP, "hipSetDevice(%d): %s", cfg->device, hipGetErrorString(e));
      break;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, cfg->device) != hipSuccess) {
      rc = fail(c, SCX_E_HIP, "hipGetDeviceProperties failed");
      break;
    }
    c->num_cus = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
      rc = fail(c, SCX_E_HIP, "hipStreamCreate failed");
      break;
    }
    const uint32_t N = cfg->max_bodies;
    c->cap_contacts = cfg->max_contacts ? cfg->max_contacts : 8u * N;
    c->cap_pairs = cfg->max_pairs ? cfg->max_pairs : 16u * N;
    c->cap_planes = std::max<uint32_t>(1, cfg->max_planes);
    if ((rc = alloc_bodies(c, N)) != SCX_OK) break;
    if ((rc = alloc_entries(c, (c->dim == 2 ? 8u : 16u) * N + 1024)) != SCX_OK) break;
    if ((rc = alloc_contacts(c, c->cap_contacts)) != SCX_OK) break;
    if ((rc = alloc_scan_tmp(c)) != SCX_OK) break;
    DeviceBuffers& d = c->d;
    if ((e = dalloc(&d.px, 3 * (size_t)c->cap_planes)) != hipSuccess ||
        (e = dalloc(&d.pn, 3 * (size_t)c->cap_planes)) != hipSuccess ||
        (e = dalloc(&d.coloff, (size_t)kMaxColours + 2)) != hipSuccess ||
        (e = dalloc(&d.grid, 1)) != hipSuccess || (e = dalloc(&d.ctr, 1)) != hipSuccess ||
        (e = dalloc(&d.barrier, kBarrierWords)) != hipSuccess ||
        (e = hipMemset(d.barrier, 0, sizeof(uint32_t) * kBarrierWords)) != hipSuccess) {
      rc = fail(c, SCX_E_HIP, "hipMalloc failed: %s", hipGetErrorString(e));
      break;
    }
    if (hipHostMalloc((void**)&d.ctr_host, sizeof(StepCounters)) != hipSuccess ||
        hipHostMalloc((void**)&d.host_u32, kHostWords * sizeof(uint32_t)) != hipSuccess) {
      rc = fail(c, SCX_E_HIP, "hipHostMalloc failed");
      break;
    }
    // Persistent (cooperative) kernels: the grid barrier's cost grows with the number of
    // workgroups, so they run one workgroup per CU unless SCX_PERSISTENT_PER_CU says more.
    int per_cu = 1;
    if (const char* env = getenv("SCX_PERSISTENT_PER_CU")) per_cu = std::min(4, std::max(1, atoi(env)));
    c->gs_blocks = std::min(solver_max_blocks(c), per_cu * c->num_cus);
    c->colour_blocks = std::min(colour_max_blocks(c), per_cu * c->num_cus);
    // a fixed workgroup count for the persistent kernels, so that tests can run the
    // device-wide barrier paths on small problems
    if (const char* env = getenv("SCX_FORCE_PERSISTENT_BLOCKS")) c->force_blocks = std::max(0, atoi(env));
    // SCX_ENSEMBLE_COMPACT=0: ensembles skip frozen contacts one by one instead of sweeping
    // compacted lists (the same iterates bit for bit; for A/B measurements and tests)
    if (const char* env = getenv("SCX_ENSEMBLE_COMPACT")) c->ens_compact = atoi(env) != 0;
    // SCX_COHERENT_SWEEPS=0: the colour passes use ordinary body accesses and full grid
    // barriers (the same iterates bit for bit; for A/B measurements and tests)
    if (const char* env = getenv("SCX_COHERENT_SWEEPS")) c->coherent_sweeps = atoi(env) != 0;
    // SCX_VIOLATOR_SKIP=0: every FP32 sweep gets a full residual evaluation (the same
    // iterates and sweep counts; for A/B measurements and tests); 2: a step whose FP32 phase
    // ran 64 sweeps or more without skipping one fails (tests)
    if (const char* env = getenv("SCX_VIOLATOR_SKIP")) c->violator_skip = atoi(env);
    // SCX_COMPONENT_SWEEPS=0: the late colours of one scene are swept in grid-wide passes and
    // a one-workgroup tail instead of as components (the same iterates bit for bit; for A/B
    // measurements and tests)
    if (const char* env = getenv("SCX_COMPONENT_SWEEPS")) c->comp_sweeps = atoi(env) != 0;
    // SCX_COMPONENT_LDS=0: every workgroup sweeps its components through its L2, none in LDS
    // (the same iterates bit for bit; for A/B measurements and tests)
    if (const char* env = getenv("SCX_COMPONENT_LDS")) c->comp_lds = atoi(env) != 0;
    // SCX_COMPONENT_PERCENT: a fixed share of the contacts (in percent, 1 to 100) in the
    // colours of the component phase at most (default: 20, then adapted step by step)
    if (const char* env = getenv("SCX_COMPONENT_PERCENT")) {
      c->comp_share = std::min(100.0, std::max(1.0, atof(env)));
      c->comp_adapt = false;
    }
    if ((e = dalloc(&d.partial, (size_t)std::max(c->gs_blocks, 1) + 1)) != hipSuccess ||
        (e = dalloc(&d.viol, (size_t)3 + kViolCap)) != hipSuccess ||
        (e = dalloc(&d.comp_cnt, (size_t)std::max(c->gs_blocks, 1) + 1)) != hipSuccess ||
        (e = dalloc(&d.comp_list, (size_t)kCompCap * std::max(c->gs_blocks, 1))) != hipSuccess ||
        (e = dalloc(&d.act_part, (size_t)kMaxColours * std::max(c->gs_blocks, 1))) != hipSuccess) {
      rc = fail(c, SCX_E_HIP, "hipMalloc failed: %s", hipGetErrorString(e));
      break;
    }
    if ((e = scratch_address(cfg->device, &c->scratch)) != hipSuccess) {
      rc = fail(c, SCX_E_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
      break;
    }
  } while (0);
  if (rc != SCX_OK) {
    // keep the context so that scx_last_error can report; caller must destroy it
    *out = c;
    return rc;
  }
  *out = c;
  return SCX_OK;
}

void scx_destroy(scx_ctx* c) {
  if (!c) return;
  hipSetDevice(c->device);
  if (c->stream) hipStreamSynchronize(c->stream);
  free_all(c);
  for (hipEvent_t e : c->events) hipEventDestroy(e);
  for (auto& P : c->tuner.pend)
    for (hipEvent_t e : P.ev)
      if (e) hipEventDestroy(e);
  if (c->tuner.copied) hipEventDestroy(c->tuner.copied);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
}

int scx_set_bodies(scx_ctx* c, uint32_t n, const double* q, const double* v, const double* omega, const double* r,
                   const double* m, const double* inertia) {
  if (!c) return SCX_E_ARG;
  if (n == 0 || !q || !v || !r || !m) return fail(c, SCX_E_ARG, "scx_set_bodies: null array or n == 0");
  HIPC(hipSetDevice(c->device));
  const int D = c->dim;
  for (uint32_t i = 0; i < n; ++i)
    if (!(r[i] > 0.0) || !(m[i] > 0.0)) return fail(c, SCX_E_ARG, "body %u: radius and mass must be positive", i);
  if (n > c->cap_bodies) {
    if (alloc_bodies(c, n) != SCX_OK) return SCX_E_HIP;
    if (alloc_entries(c, (D == 2 ? 8u : 16u) * n + 1024) != SCX_OK) return SCX_E_HIP;
    if (alloc_contacts(c, std::max(c->cap_contacts, 8u * n)) != SCX_OK) return SCX_E_HIP;
    if (alloc_scan_tmp(c) != SCX_OK) return SCX_E_HIP;
  }
  hipStream_t s = c->stream;
  DeviceBuffers& d = c->d;
  HIPC(hipMemcpyAsync(d.q, q, sizeof(double) * D * n, hipMemcpyHostToDevice, s));
  HIPC(hipMemcpyAsync(d.v, v, sizeof(double) * D * n, hipMemcpyHostToDevice, s));
  HIPC(hipMemcpyAsync(d.r, r, sizeof(double) * n, hipMemcpyHostToDevice, s));
  HIPC(hipMemcpyAsync(d.m, m, sizeof(double) * n, hipMemcpyHostToDevice, s));
  std::vector<double> tmp;
  if (D == 3) {
    if (omega)
      HIPC(hipMemcpyAsync(d.w, omega, sizeof(double) * 3 * n, hipMemcpyHostToDevice, s));
    else
      HIPC(hipMemsetAsync(d.w, 0, sizeof(double) * 3 * n, s));
    tmp.resize(n);
    for (uint32_t i = 0; i < n; ++i) {
      if (inertia) {
        tmp[i] = inertia[i];
      } else {
        // RigidBodySphere::computeMassAndInertia (rigidbody3d/Geometry/RigidBodySphere.cpp:10-18)
        tmp[i] = 2.0 * m[i] * r[i] * r[i] / 5.0;
      }
      if (!(tmp[i] > 0.0)) return fail(c, SCX_E_ARG, "body %u: inertia must be positive", i);
    }
    HIPC(hipMemcpyAsync(d.inertia, tmp.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
  } else {
    HIPC(hipMemsetAsync(d.w, 0, sizeof(double) * 3 * n, s));
    HIPC(hipMemsetAsync(d.inertia, 0, sizeof(double) * n, s));
  }
  HIPC(hipStreamSynchronize(s));
  // The impulse cache belongs to the scene, as the reference's constraint cache belongs to
  // the simulation: new positions and velocities of the same bodies (a caller feeding
  // each step's q0, v0) keep it, a different body count empties it.
  if (n != c->n) {
    d.prev_count = 0;
    c->nscenes = 0;  // an ensemble's layout refers to the old bodies
  }
  c->n = n;
  c->has_bodies = true;
  c->order_valid = false;
  c->last_ncontacts = 0;
  return SCX_OK;
}

int scx_set_planes(scx_ctx* c, uint32_t nplanes, const double* x, const double* nrm) {
  if (!c) return SCX_E_ARG;
  if (nplanes > kMaxPlanesTotal) return fail(c, SCX_E_ARG, "at most %u static planes", kMaxPlanesTotal);
  if (nplanes > 0 && (!x || !nrm)) return fail(c, SCX_E_ARG, "null plane arrays");
  HIPC(hipSetDevice(c->device));
  const int D = c->dim;
  if (nplanes > c->cap_planes) {
    HIPC(dalloc(&c->d.px, 3 * (size_t)nplanes));
    HIPC(dalloc(&c->d.pn, 3 * (size_t)nplanes));
    c->cap_planes = nplanes;
  }
  std::vector<double> nn((size_t)D * nplanes);
  for (uint32_t p = 0; p < nplanes; ++p) {
    // StaticPlane normalises its normal (ball2d/StaticGeometry/StaticPlane.cpp:10-20)
    double z = nrm[(size_t)D * p] * nrm[(size_t)D * p];
    for (int a = 1; a < D; ++a) z = z + nrm[(size_t)D * p + a] * nrm[(size_t)D * p + a];
    if (!(z > 0.0)) return fail(c, SCX_E_ARG, "plane %u has a zero normal", p);
    const double sq = std::sqrt(z);
    for (int a = 0; a < D; ++a) nn[(size_t)D * p + a] = nrm[(size_t)D * p + a] / sq;
  }
  if (nplanes > 0) {
    HIPC(hipMemcpyAsync(c->d.px, x, sizeof(double) * D * nplanes, hipMemcpyHostToDevice, c->stream));
    HIPC(hipMemcpyAsync(c->d.pn, nn.data(), sizeof(double) * D * nplanes, hipMemcpyHostToDevice, c->stream));
  }
  HIPC(hipStreamSynchronize(c->stream));
  if (nplanes != c->nplanes) c->nscenes = 0;
  c->nplanes = nplanes;
  return SCX_OK;
}

int scx_set_scenes(scx_ctx* c, uint32_t nscenes, const uint32_t* body_off, const uint32_t* plane_off,
                   const double* mu) {
  if (!c) return SCX_E_ARG;
  if (nscenes == 0) {
    c->nscenes = 0;
    return SCX_OK;
  }
  if (!c->has_bodies) return fail(c, SCX_E_STATE, "scx_set_bodies must be called before scx_set_scenes");
  if (!body_off || !plane_off || !mu) return fail(c, SCX_E_ARG, "scx_set_scenes: null array");
  if (body_off[0] != 0 || body_off[nscenes] != c->n)
    return fail(c, SCX_E_ARG, "scene body offsets must run from 0 to the body count %u", c->n);
  if (plane_off[0] != 0 || plane_off[nscenes] != c->nplanes)
    return fail(c, SCX_E_ARG, "scene plane offsets must run from 0 to the plane count %u", c->nplanes);
  uint32_t most = 0;
  for (uint32_t k = 0; k < nscenes; ++k) {
    if (body_off[k + 1] < body_off[k] || plane_off[k + 1] < plane_off[k])
      return fail(c, SCX_E_ARG, "scene %u: offsets must not decrease", k);
    if (plane_off[k + 1] - plane_off[k] > kMaxScenePlanes)
      return fail(c, SCX_E_ARG, "scene %u: at most %u static planes per scene", k, kMaxScenePlanes);
    if (!(mu[k] >= 0.0)) return fail(c, SCX_E_ARG, "scene %u: mu must be non-negative", k);
    most = std::max(most, plane_off[k + 1] - plane_off[k]);
  }
  HIPC(hipSetDevice(c->device));
  DeviceBuffers& d = c->d;
  hipStream_t s = c->stream;
  const int D = c->dim;
  const uint32_t n = c->n;
  // binning shifts: each scene's region (bounds of its bodies' spheres, widened) placed
  // on shelves along x, the shelves stacked along the last axis, with a gap between
  // regions; the grid then keeps scenes apart and stays about as dense as the scenes
  std::vector<double> q((size_t)D * n), r(n);
  HIPC(hipMemcpyAsync(q.data(), d.q, sizeof(double) * D * n, hipMemcpyDeviceToHost, s));
  HIPC(hipMemcpyAsync(r.data(), d.r, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  HIPC(hipStreamSynchronize(s));
  std::vector<double> lo((size_t)3 * nscenes, 0.0), hi((size_t)3 * nscenes, 0.0);
  double rmax = 0.0;
  for (uint32_t k = 0; k < nscenes; ++k) {
    double l[3] = {INFINITY, INFINITY, INFINITY}, h[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = body_off[k]; i < body_off[k + 1]; ++i) {
      for (int a = 0; a < D; ++a) {
        l[a] = std::min(l[a], q[(size_t)D * i + a] - r[i]);
        h[a] = std::max(h[a], q[(size_t)D * i + a] + r[i]);
      }
      rmax = std::max(rmax, r[i]);
    }
    for (int a = 0; a < D; ++a) {
      if (!(l[a] <= h[a])) l[a] = h[a] = 0.0;  // a scene without bodies
      const double w = 0.1 * (h[a] - l[a]);
      lo[3 * (size_t)k + a] = l[a] - w;
      hi[3 * (size_t)k + a] = h[a] + w;
    }
  }
  const double gap = 4.0 * rmax + 1e-3;
  const int ax = 0, az = D - 1;  // shelf axis, stacking axis
  std::vector<uint32_t> idx(nscenes);
  double area = 0.0, widest = 0.0;
  for (uint32_t k = 0; k < nscenes; ++k) {
    idx[k] = k;
    const double wx = hi[3 * (size_t)k + ax] - lo[3 * (size_t)k + ax] + gap;
    const double wz = hi[3 * (size_t)k + az] - lo[3 * (size_t)k + az] + gap;
    area += wx * wz;
    widest = std::max(widest, wx);
  }
  std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) {
    return hi[3 * (size_t)a + az] - lo[3 * (size_t)a + az] > hi[3 * (size_t)b + az] - lo[3 * (size_t)b + az];
  });
  const double row = std::max(widest, 1.2 * std::sqrt(area));
  std::vector<double> shift((size_t)3 * nscenes, 0.0);
  double x = 0.0, z = 0.0, depth = 0.0;
  for (uint32_t k : idx) {
    const double wx = hi[3 * (size_t)k + ax] - lo[3 * (size_t)k + ax] + gap;
    const double wz = hi[3 * (size_t)k + az] - lo[3 * (size_t)k + az] + gap;
    if (x > 0.0 && x + wx > row) {
      z += depth;
      x = 0.0;
      depth = 0.0;
    }
    for (int a = 0; a < D; ++a) shift[3 * (size_t)k + a] = -lo[3 * (size_t)k + a];
    shift[3 * (size_t)k + ax] += x;
    if (az != ax) shift[3 * (size_t)k + az] += z;
    x += wx;
    depth = std::max(depth, wz);
  }
  std::vector<uint32_t> bs(n);
  for (uint32_t k = 0; k < nscenes; ++k)
    for (uint32_t i = body_off[k]; i < body_off[k + 1]; ++i) bs[i] = k;
  if (nscenes > c->cap_scenes) {
    HIPC(dalloc(&d.scene_poff, (size_t)nscenes + 1));
    HIPC(dalloc(&d.scene_mu, nscenes));
    HIPC(dalloc(&d.scene_shift, (size_t)3 * nscenes));
    HIPC(dalloc(&d.scene_err, nscenes));
    HIPC(dalloc(&d.scene_cnt, nscenes));
    HIPC(dalloc(&d.scene_acc, (size_t)2 * nscenes));
    HIPC(dalloc(&d.scene_sweeps, nscenes));
    c->cap_scenes = nscenes;
  }
  HIPC(dalloc(&d.bscene, c->cap_bodies));
  HIPC(hipMemcpyAsync(d.bscene, bs.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, s));
  HIPC(hipMemcpyAsync(d.scene_poff, plane_off, sizeof(uint32_t) * (nscenes + 1), hipMemcpyHostToDevice, s));
  HIPC(hipMemcpyAsync(d.scene_mu, mu, sizeof(double) * nscenes, hipMemcpyHostToDevice, s));
  HIPC(hipMemcpyAsync(d.scene_shift, shift.data(), sizeof(double) * 3 * nscenes, hipMemcpyHostToDevice, s));
  HIPC(hipStreamSynchronize(s));
  c->nscenes = nscenes;
  c->scene_planes = most;
  c->last_ncontacts = 0;
  return SCX_OK;
}

int scx_get_scene_stats(scx_ctx* c, uint32_t nscenes, uint32_t* ncontacts, double* error, uint32_t* sweeps) {
  if (!c) return SCX_E_ARG;
  if (c->nscenes == 0 || nscenes != c->nscenes)
    return fail(c, SCX_E_ARG, "scx_get_scene_stats: the context holds %u scenes", c->nscenes);
  HIPC(hipSetDevice(c->device));
  DeviceBuffers& d = c->d;
  HIPC(launch_scene_stats(c, c->last_ncontacts));
  if (ncontacts)
    HIPC(hipMemcpyAsync(ncontacts, d.scene_cnt, sizeof(uint32_t) * nscenes, hipMemcpyDeviceToHost, c->stream));
  if (error) HIPC(hipMemcpyAsync(error, d.scene_err, sizeof(double) * nscenes, hipMemcpyDeviceToHost, c->stream));
  if (sweeps) {
    if (c->last_ncontacts > 0)
      HIPC(hipMemcpyAsync(sweeps, d.scene_sweeps, sizeof(uint32_t) * nscenes, hipMemcpyDeviceToHost, c->stream));
    else
      for (uint32_t k = 0; k < nscenes; ++k) sweeps[k] = 0;
  }
  StepCounters ctr{};
  HIPC(hipMemcpyAsync(&ctr, d.ctr, sizeof(StepCounters), hipMemcpyDeviceToHost, c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  if (sweeps && c->last_ncontacts > 0)  // a scene that never met the tolerance ran every sweep
    for (uint32_t k = 0; k < nscenes; ++k)
      if (sweeps[k] == kNoBody) sweeps[k] = ctr.sweeps;
  return SCX_OK;
}

int scx_get_state(scx_ctx* c, double* q, double* v, double* omega) {
  if (!c) return SCX_E_ARG;
  if (!c->has_bodies) return fail(c, SCX_E_STATE, "no bodies");
  HIPC(hipSetDevice(c->device));
  const int D = c->dim;
  if (q) HIPC(hipMemcpyAsync(q, c->d.q, sizeof(double) * D * c->n, hipMemcpyDeviceToHost, c->stream));
  if (v) HIPC(hipMemcpyAsync(v, c->d.v, sizeof(double) * D * c->n, hipMemcpyDeviceToHost, c->stream));
  if (omega && D == 3)
    HIPC(hipMemcpyAsync(omega, c->d.w, sizeof(double) * 3 * c->n, hipMemcpyDeviceToHost, c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  return SCX_OK;
}

int scx_step(scx_ctx* c, const scx_step_params* p, scx_step_stats* stats) {
  if (!c) return SCX_E_ARG;
  int rc = check_params(c, p);
  if (rc != SCX_OK) return rc;
  HIPC(hipSetDevice(c->device));
  for (int attempt = 0; attempt < 8; ++attempt) {
    rc = step_once(c, p, stats, true);
    if (rc != SCX_E_CAPACITY) return rc;
  }
  return fail(c, SCX_E_CAPACITY, "buffers kept overflowing");
}

int scx_run(scx_ctx* c, const scx_step_params* p, uint32_t nsteps, scx_step_stats* stats) {
  if (!c) return SCX_E_ARG;
  int rc = check_params(c, p);
  if (rc != SCX_OK) return rc;
  HIPC(hipSetDevice(c->device));
  for (uint32_t k = 0; k < nsteps; ++k) {
    scx_step_params q = *p;
    q.iteration = p->iteration + k;
    const bool last = (k + 1 == nsteps);
    int attempt = 0;
    for (; attempt < 8; ++attempt) {
      rc = step_once(c, &q, last ? stats : nullptr, last);
      if (rc != SCX_E_CAPACITY) break;
    }
    if (rc != SCX_OK) return attempt >= 8 ? fail(c, SCX_E_CAPACITY, "buffers kept overflowing") : rc;
  }
  return SCX_OK;
}

int scx_get_contacts(scx_ctx* c, uint32_t cap, int32_t* body0, int32_t* body1, int32_t* plane, double* bases,
                     double* alpha, double* beta, int32_t* colour, uint32_t* ncontacts) {
  if (!c) return SCX_E_ARG;
  HIPC(hipSetDevice(c->device));
  const uint32_t C = c->last_ncontacts;
  if (ncontacts) *ncontacts = C;
  if (C > cap) return fail(c, SCX_E_CAPACITY, "need room for %u contacts", C);
  const int D = c->dim;
  hipStream_t s = c->stream;
  if (C == 0) return SCX_OK;
  HIPC(hipStreamSynchronize(s));  // the synchronous copies below do not wait for the context stream
  std::vector<uint32_t> tmp(C);
  if (body0) {
    HIPC(hipMemcpy(tmp.data(), c->d.ca, sizeof(uint32_t) * C, hipMemcpyDeviceToHost));
    for (uint32_t k = 0; k < C; ++k) body0[k] = (int32_t)tmp[k];
  }
  if (body1) {
    HIPC(hipMemcpy(tmp.data(), c->d.cb, sizeof(uint32_t) * C, hipMemcpyDeviceToHost));
    for (uint32_t k = 0; k < C; ++k) body1[k] = (tmp[k] == kNoBody) ? -1 : (int32_t)tmp[k];
  }
  if (plane) HIPC(hipMemcpyAsync(plane, c->d.cplane, sizeof(int32_t) * C, hipMemcpyDeviceToHost, s));
  if (bases) HIPC(hipMemcpyAsync(bases, c->d.basis, sizeof(double) * D * D * C, hipMemcpyDeviceToHost, s));
  if (alpha) HIPC(hipMemcpyAsync(alpha, c->d.calpha, sizeof(double) * C, hipMemcpyDeviceToHost, s));
  if (beta) HIPC(hipMemcpyAsync(beta, c->d.cbeta, sizeof(double) * (D - 1) * C, hipMemcpyDeviceToHost, s));
  if (colour) {
    HIPC(hipMemcpyAsync(colour, c->d.cstate, sizeof(int32_t) * C, hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    for (uint32_t k = 0; k < C; ++k) colour[k] &= 0xff;
  }
  HIPC(hipStreamSynchronize(s));
  return SCX_OK;
}

}  // extern "C"

namespace {

// The reference's force output of one contact (ImpactFrictionMap::exportConstraintForcesToBinaryFile,
// scisim/ConstrainedMaps/ImpactFrictionMap.cpp:58-163): body indices with a static plane as
// -plane - 2 (:58-66); the world-space contact point at q0 (2D x_a - r_a n:
// ball2d/Constraints/BallBallConstraint.cpp:219-222, BallStaticPlaneConstraint.cpp:171-174;
// 3D sphere pairs the point x_a + r_a / (r_a + r_b) (x_b - x_a) of the narrow phase,
// rigidbody3d/RigidBody3DSim.cpp:803-809, plane x_a - r_a n:
// StaticPlaneSphereConstraint.cpp:314-317); the normal; the impulse alpha n + sum beta_k t_k.
template <int D>
__global__ void __launch_bounds__(kBlock) k_contact_forces(uint32_t C, const uint32_t* __restrict__ ca,
                                                           const uint32_t* __restrict__ cb,
                                                           const int32_t* __restrict__ cplane,
                                                           const double* __restrict__ basis,
                                                           const double* __restrict__ alpha,
                                                           const double* __restrict__ beta,
                                                           const double* __restrict__ q0, const double* __restrict__ r,
                                                           int32_t* __restrict__ idx, double* __restrict__ pts,
                                                           double* __restrict__ nrm, double* __restrict__ frc) {
#pragma clang fp contract(off)
  const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= C) return;
  const uint32_t a = ca[k], b = cb[k];
  const double* B = basis + (size_t)D * D * k;
  idx[2 * (size_t)k] = (int32_t)a;
  idx[2 * (size_t)k + 1] = (b == kNoBody) ? -cplane[k] - 2 : (int32_t)b;
  const double al = alpha[k];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const double xa = q0[(size_t)D * a + d];
    double p;
    if (D == 3 && b != kNoBody) {
      const double f = r[a] / (r[a] + r[b]);
      p = xa + f * (q0[(size_t)D * b + d] - xa);
    } else {
      p = xa - r[a] * B[d];
    }
    pts[(size_t)D * k + d] = p;
    nrm[(size_t)D * k + d] = B[d];
    double fo = al * B[d];
    for (int t = 1; t < D; ++t) fo = fo + beta[(size_t)(D - 1) * k + t - 1] * B[(size_t)D * t + d];
    frc[(size_t)D * k + d] = fo;
  }
}

}  // namespace

extern "C" {

int scx_get_contact_forces(scx_ctx* c, uint32_t cap, int32_t* indices, double* points, double* normals,
                           double* forces, uint32_t* ncontacts) {
  if (!c) return SCX_E_ARG;
  HIPC(hipSetDevice(c->device));
  const uint32_t C = c->last_ncontacts;
  if (ncontacts) *ncontacts = C;
  if (C > cap) return fail(c, SCX_E_CAPACITY, "need room for %u contacts", C);
  if (C == 0) return SCX_OK;
  if (!indices || !points || !normals || !forces) return fail(c, SCX_E_ARG, "scx_get_contact_forces: null array");
  const int D = c->dim;
  hipStream_t s = c->stream;
  DeviceBuffers& d = c->d;
  // staging in the contact-capacity buffers that the next step overwrites anyway: the
  // records' FP64 arrays (3 per contact) hold points, normals and forces, pmask the indices
  double* pts = d.c64.n;
  double* nrm = d.c64.s;
  double* frc = d.c64.r;
  int32_t* idx = reinterpret_cast<int32_t*>(d.pmask);
  const uint32_t nb = (C + kBlock - 1) / kBlock;
  if (D == 2)
    k_contact_forces<2><<<nb, kBlock, 0, s>>>(C, d.ca, d.cb, d.cplane, d.basis, d.calpha, d.cbeta, d.q0, d.r, idx,
                                               pts, nrm, frc);
  else
    k_contact_forces<3><<<nb, kBlock, 0, s>>>(C, d.ca, d.cb, d.cplane, d.basis, d.calpha, d.cbeta, d.q0, d.r, idx,
                                               pts, nrm, frc);
  HIPC(hipGetLastError());
  HIPC(hipMemcpyAsync(indices, idx, sizeof(int32_t) * 2 * C, hipMemcpyDeviceToHost, s));
  HIPC(hipMemcpyAsync(points, pts, sizeof(double) * D * C, hipMemcpyDeviceToHost, s));
  HIPC(hipMemcpyAsync(normals, nrm, sizeof(double) * D * C, hipMemcpyDeviceToHost, s));
  HIPC(hipMemcpyAsync(forces, frc, sizeof(double) * D * C, hipMemcpyDeviceToHost, s));
  HIPC(hipStreamSynchronize(s));
  return SCX_OK;
}

int scx_get_v2(scx_ctx* c, double* v2, double* omega2) {
  if (!c) return SCX_E_ARG;
  HIPC(hipSetDevice(c->device));
  if (v2) HIPC(hipMemcpy(v2, c->d.v2, sizeof(double) * c->dim * c->n, hipMemcpyDeviceToHost));
  if (omega2 && c->dim == 3) HIPC(hipMemcpy(omega2, c->d.w2, sizeof(double) * 3 * c->n, hipMemcpyDeviceToHost));
  return SCX_OK;
}

int scx_friction_solve(scx_ctx* c, uint32_t nbodies, const double* m, const double* inertia, const double* v0,
                       const double* omega0, const scx_contact_set* cs, double tol, uint32_t max_iters,
                       uint32_t eval_every, int32_t precision, double* alpha, double* beta, double* v2,
                       double* omega2, double* f, int32_t* succeeded, double* error, uint32_t* sweeps) {
  if (!c) return SCX_E_ARG;
  if (!cs || !m || !v0 || !alpha || !beta) return fail(c, SCX_E_ARG, "scx_friction_solve: null argument");
  if (precision != SCX_PRECISION_FP64 && precision != SCX_PRECISION_FP32_POLISH)
    return fail(c, SCX_E_ARG, "unknown precision mode %d", precision);
  HIPC(hipSetDevice(c->device));
  const int D = c->dim;
  const uint32_t C = cs->ncontacts;
  if (nbodies == 0) return fail(c, SCX_E_ARG, "no bodies");
  if (C > 0 && (!cs->body0 || !cs->body1 || !cs->bases || !cs->mu))
    return fail(c, SCX_E_ARG, "scx_friction_solve: incomplete contact set");
  if (D == 3 && C > 0 && (!cs->arm0 || !cs->arm1)) return fail(c, SCX_E_ARG, "3D contacts need arms");
  if (cs->drel_extra) return fail(c, SCX_E_ARG, "tangential forcing terms (drel_extra) are not supported");
  if (D == 3 && !inertia) return fail(c, SCX_E_ARG, "3D bodies need their inertia");
  for (uint32_t i = 0; i < nbodies; ++i) {
    if (!(m[i] > 0.0)) return fail(c, SCX_E_ARG, "body %u: mass must be positive", i);
    if (D == 3 && !(inertia[i] > 0.0)) return fail(c, SCX_E_ARG, "body %u: inertia must be positive", i);
  }
  for (uint32_t k = 0; k < C; ++k) {
    if (cs->body0[k] < 0 || (uint32_t)cs->body0[k] >= nbodies || cs->body1[k] < -1 ||
        cs->body1[k] >= (int32_t)nbodies || cs->body0[k] == cs->body1[k])
      return fail(c, SCX_E_ARG, "contact %u: invalid bodies (%d, %d)", k, cs->body0[k], cs->body1[k]);
    if (!(cs->mu[k] >= 0.0)) return fail(c, SCX_E_ARG, "contact %u: mu must be non-negative", k);
  }
  if (nbodies > c->cap_bodies) {
    if (alloc_bodies(c, nbodies) != SCX_OK || alloc_scan_tmp(c) != SCX_OK) return SCX_E_HIP;
  }
  if (C > c->cap_contacts) {
    if (alloc_contacts(c, C) != SCX_OK || alloc_scan_tmp(c) != SCX_OK) return SCX_E_HIP;
  }
  // the explicit problem replaces the resident scene; its bodies keep their own indices
  c->has_bodies = false;
  c->n = nbodies;
  c->nscenes = 0;
  hipStream_t s = c->stream;
  DeviceBuffers& d = c->d;
  d.sid = nullptr;
  d.slot_map = nullptr;
  d.prev_count = 0;
  std::vector<double> Ivec(nbodies, 0.0), w0(3 * (size_t)nbodies, 0.0);
  if (D == 3) {
    for (uint32_t i = 0; i < nbodies; ++i) Ivec[i] = inertia ? inertia[i] : 0.0;
    if (omega0) std::copy(omega0, omega0 + 3 * (size_t)nbodies, w0.begin());
  }
  // per-contact frames, arms, Delassus diagonal, forcing term and warm start (host; this
  // entry point serves the reference's FrictionSolver boundary, not the step loop)
  std::vector<uint32_t> ca(C), cb(C);
  std::vector<double> la(C, 0.0), lb(C, 0.0), wn(C), wt(C), forcing(C), rinit((size_t)D * C), mu(cs->mu, cs->mu + C);
  for (uint32_t k = 0; k < C; ++k) {
    const uint32_t a = (uint32_t)cs->body0[k];
    const int32_t b = cs->body1[k];
    ca[k] = a;
    cb[k] = b >= 0 ? (uint32_t)b : kNoBody;
    const double* B = cs->bases + (size_t)D * D * k;
    const double minv_a = 1.0 / m[a], minv_b = b >= 0 ? 1.0 / m[b] : 0.0;
    double ndotv = 0.0;
    for (int q = 0; q < D; ++q) ndotv += B[q] * (v0[(size_t)D * a + q] - (b >= 0 ? v0[(size_t)D * b + q] : 0.0));
    const double cor = cs->CoR ? cs->CoR[k] : 0.0;
    forcing[k] = cor * ndotv + (cs->nrel_extra ? cs->nrel_extra[k] : 0.0);
    if (D == 2) {
      wn[k] = minv_a + minv_b;
      wt[k] = minv_a + minv_b;
    } else {
      const double* pa = cs->arm0 + 3 * (size_t)k;
      const double* pb = cs->arm1 + 3 * (size_t)k;
      la[k] = std::sqrt(pa[0] * pa[0] + pa[1] * pa[1] + pa[2] * pa[2]);
      lb[k] = b >= 0 ? std::sqrt(pb[0] * pb[0] + pb[1] * pb[1] + pb[2] * pb[2]) : 0.0;
      const double ia = 1.0 / Ivec[a], ib = b >= 0 ? 1.0 / Ivec[b] : 0.0;
      wn[k] = minv_a + minv_b;
      wt[k] = (minv_a + ia * la[k] * la[k]) + (b >= 0 ? (minv_b + ib * lb[k] * lb[k]) : 0.0);
    }
    rinit[(size_t)D * k] = alpha[k];
    for (int q = 1; q < D; ++q) rinit[(size_t)D * k + q] = beta[(size_t)(D - 1) * k + q - 1];
  }
  bool warm = false;
  for (double x : rinit) warm = warm || x != 0.0;
  HIPC(hipMemcpyAsync(d.m, m, sizeof(double) * nbodies, hipMemcpyHostToDevice, s));
  HIPC(hipMemcpyAsync(d.inertia, Ivec.data(), sizeof(double) * nbodies, hipMemcpyHostToDevice, s));
  HIPC(hipMemcpyAsync(d.v0, v0, sizeof(double) * D * nbodies, hipMemcpyHostToDevice, s));
  HIPC(hipMemcpyAsync(d.w0, w0.data(), sizeof(double) * 3 * nbodies, hipMemcpyHostToDevice, s));
  HIPC(hipMemsetAsync(d.ctr, 0, sizeof(StepCounters), s));
  c->vbasis = c->vsolve = d.v0;
  if (C > 0) {
    HIPC(hipMemcpyAsync(d.ca, ca.data(), sizeof(uint32_t) * C, hipMemcpyHostToDevice, s));
    HIPC(hipMemcpyAsync(d.cb, cb.data(), sizeof(uint32_t) * C, hipMemcpyHostToDevice, s));
    HIPC(hipMemcpyAsync(d.basis, cs->bases, sizeof(double) * D * D * C, hipMemcpyHostToDevice, s));
    HIPC(hipMemcpyAsync(d.cla, la.data(), sizeof(double) * C, hipMemcpyHostToDevice, s));
    HIPC(hipMemcpyAsync(d.clb, lb.data(), sizeof(double) * C, hipMemcpyHostToDevice, s));
    HIPC(hipMemcpyAsync(d.cwn, wn.data(), sizeof(double) * C, hipMemcpyHostToDevice, s));
    HIPC(hipMemcpyAsync(d.cwt, wt.data(), sizeof(double) * C, hipMemcpyHostToDevice, s));
    HIPC(hipMemcpyAsync(d.cforcing, forcing.data(), sizeof(double) * C, hipMemcpyHostToDevice, s));
    HIPC(hipMemcpyAsync(d.crinit, rinit.data(), sizeof(double) * D * C, hipMemcpyHostToDevice, s));
    HIPC(hipMemcpyAsync(d.calpha, mu.data(), sizeof(double) * C, hipMemcpyHostToDevice, s));  // mu staging
    HIPC(launch_colouring(c, C));
    HIPC(launch_colour_sort(c, C));
    HIPC(hipMemcpyAsync(d.host_u32 + 2, d.coloff, sizeof(uint32_t) * (kMaxColours + 1), hipMemcpyDeviceToHost, s));
    HIPC(hipMemcpyAsync(d.host_u32 + kOverflowWord, &d.ctr->overflow, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    if (d.host_u32[kOverflowWord] & 8u)
      return fail(c, SCX_E_COLOURS, "contact graph needs more than %d colours (a body with too many contacts)",
                  kMaxColours);
    HIPC(launch_gather_records(c, C, precision == SCX_PRECISION_FP32_POLISH, 0.0, d.calpha, d.crinit));
  }
  scx_step_params p{};
  p.dt = 1.0;
  p.tol = tol;
  p.max_iters = max_iters;
  p.eval_every = eval_every;
  p.precision = precision;
  p.warm_start = warm ? SCX_WARM_CACHE : SCX_WARM_NONE;
  p.mu = 1.0;  // beta is taken as solved (per-contact mu is in the records)
  HIPC(launch_solve(c, C, d.host_u32 + 2, &p));
  if (C > 0) {
    const uint32_t nc = C;
    (void)nc;
    HIPC(launch_impulses(c, C));
  }
  // body records by slot (identity slots here): [v | minv] (2D), [v | minv][omega | Iinv] (3D)
  const size_t rec = (D == 3) ? 8 : 4;
  std::vector<double> bv(rec * (size_t)nbodies);
  HIPC(hipMemcpyAsync(bv.data(), d.bv64, sizeof(double) * bv.size(), hipMemcpyDeviceToHost, s));
  if (C > 0) {
    HIPC(hipMemcpyAsync(alpha, d.calpha, sizeof(double) * C, hipMemcpyDeviceToHost, s));
    HIPC(hipMemcpyAsync(beta, d.cbeta, sizeof(double) * (D - 1) * C, hipMemcpyDeviceToHost, s));
  }
  HIPC(hipMemcpyAsync(d.ctr_host, d.ctr, sizeof(StepCounters), hipMemcpyDeviceToHost, s));
  HIPC(hipStreamSynchronize(s));
  for (uint32_t i = 0; i < nbodies; ++i) {
    if (v2)
      for (int q = 0; q < D; ++q) v2[(size_t)D * i + q] = bv[rec * i + q];
    if (omega2 && D == 3)
      for (int q = 0; q < 3; ++q) omega2[3 * (size_t)i + q] = bv[rec * i + 4 + q];
  }
  // friction-only generalized impulse (Sobogus.cpp:335-346, :543-566): layout [lin | ang]
  if (f) {
    const size_t nv = (size_t)(D == 2 ? 2 : 6) * nbodies;
    std::fill(f, f + nv, 0.0);
    for (uint32_t k = 0; k < C; ++k) {
      const double* B = cs->bases + (size_t)D * D * k;
      double ft[3] = {0.0, 0.0, 0.0};
      for (int col = 1; col < D; ++col)
        for (int q = 0; q < D; ++q) ft[q] += beta[(size_t)(D - 1) * k + col - 1] * B[col * D + q];
      const int32_t bodies[2] = {cs->body0[k], cs->body1[k]};
      for (int side = 0; side < 2; ++side) {
        const int32_t i = bodies[side];
        if (i < 0) continue;
        const double sg = side == 0 ? 1.0 : -1.0;
        for (int q = 0; q < D; ++q) f[(size_t)D * i + q] += sg * ft[q];
        if (D == 3) {
          const double* r = (side == 0 ? cs->arm0 : cs->arm1) + 3 * (size_t)k;
          const double tq[3] = {r[1] * ft[2] - r[2] * ft[1], r[2] * ft[0] - r[0] * ft[2], r[0] * ft[1] - r[1] * ft[0]};
          for (int q = 0; q < 3; ++q) f[3 * (size_t)nbodies + 3 * (size_t)i + q] += sg * tq[q];
        }
      }
    }
  }
  const double err = (C == 0) ? 0.0 : d.ctr_host->error;
  if (error) *error = err;
  if (succeeded) *succeeded = (C == 0 || err < tol) ? 1 : 0;
  if (sweeps) *sweeps = d.ctr_host->sweeps;
  return SCX_OK;
}

int scx_potential_overlaps(scx_ctx* c, uint32_t n, const double* aabb_min, const double* aabb_max, uint64_t cap,
                           uint32_t* pairs_out, uint64_t* npairs) {
  if (!c) return SCX_E_ARG;
  if (npairs) *npairs = 0;
  if (n == 0) return SCX_OK;
  if (!aabb_min || !aabb_max) return fail(c, SCX_E_ARG, "null box arrays");
  HIPC(hipSetDevice(c->device));
  const int D = c->dim;
  if (n > c->cap_bodies) {
    if (alloc_bodies(c, n) != SCX_OK || alloc_contacts(c, std::max(c->cap_contacts, 8u * n)) != SCX_OK ||
        alloc_scan_tmp(c) != SCX_OK)
      return SCX_E_HIP;
    c->has_bodies = false;
  }
  hipStream_t s = c->stream;
  DeviceBuffers& d = c->d;
  HIPC(hipMemcpyAsync(d.aabb_min, aabb_min, sizeof(double) * D * n, hipMemcpyHostToDevice, s));
  HIPC(hipMemcpyAsync(d.aabb_max, aabb_max, sizeof(double) * D * n, hipMemcpyHostToDevice, s));
  for (int attempt = 0; attempt < 8; ++attempt) {
    HIPC(hipMemsetAsync(d.ctr, 0, sizeof(StepCounters), s));
    c->order_valid = false;  // the sweep order now belongs to these boxes
    HIPC(launch_grid_build(c, n, true));
    HIPC(launch_pairs(c, n, 0, 0, nullptr, nullptr, 0));
    HIPC(launch_exclusive_scan(d.pcount, d.pcount, n, d.scan_tmp, s));
    HIPC(hipMemcpyAsync(d.ctr_host, d.ctr, sizeof(StepCounters), hipMemcpyDeviceToHost, s));
    HIPC(hipMemcpyAsync(d.host_u32, d.pcount + n, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    const uint64_t ne = d.ctr_host->nentries;
    if ((d.ctr_host->overflow & 1u) || ne > c->cap_entries) {
      if (alloc_entries(c, (uint32_t)std::min<uint64_t>(0xFFFFFFFFull, ne + ne / 4 + 1024)) != SCX_OK ||
          alloc_scan_tmp(c) != SCX_OK)
        return SCX_E_HIP;
      continue;
    }
    const uint32_t P = d.host_u32[0];
    if (npairs) *npairs = P;
    if (P > cap) return fail(c, SCX_E_CAPACITY, "need room for %u pairs", P);
    if (P == 0) return SCX_OK;
    if (!d.pa || P > c->cap_pairs) {
      c->cap_pairs = std::max<uint32_t>(P, c->cap_pairs);
      HIPC(dalloc(&d.pa, c->cap_pairs));
      HIPC(dalloc(&d.pb, c->cap_pairs));
    }
    HIPC(launch_pairs(c, n, 1, 0, d.pa, d.pb, c->cap_pairs));
    std::vector<uint32_t> a(P), b(P);
    HIPC(hipMemcpyAsync(a.data(), d.pa, sizeof(uint32_t) * P, hipMemcpyDeviceToHost, s));
    HIPC(hipMemcpyAsync(b.data(), d.pb, sizeof(uint32_t) * P, hipMemcpyDeviceToHost, s));
    HIPC(hipMemcpyAsync(d.ctr_host, d.ctr, sizeof(StepCounters), hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    if (d.ctr_host->overflow & 2u) return fail(c, SCX_E_CAPACITY, "pair buffer overflow");
    for (uint32_t k = 0; k < P; ++k) {
      pairs_out[2 * (size_t)k] = a[k];
      pairs_out[2 * (size_t)k + 1] = b[k];
    }
    return SCX_OK;
  }
  return fail(c, SCX_E_CAPACITY, "grid entry buffer kept overflowing");
}

}  // extern "C"

namespace {

// Interleaves the pair lists into (a, b) pairs for one copy to the caller.
__global__ void __launch_bounds__(kBlock) k_interleave_pairs(uint32_t P, const uint32_t* __restrict__ a,
                                                             const uint32_t* __restrict__ b, uint32_t* __restrict__ out) {
  const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= P) return;
  out[2 * (size_t)k] = a[k];
  out[2 * (size_t)k + 1] = b[k];
}

// Bump allocator over the scratch array (256-byte aligned pieces).
struct Carve {
  unsigned char* base;
  size_t used = 0;
  template <typename T>
  T* take(size_t count) {
    T* p = reinterpret_cast<T*>(base + used);
    used += (sizeof(T) * std::max<size_t>(count, 1) + 255) / 256 * 256;
    return p;
  }
};

// The context's broad-phase buffers and capacities, restored when the no-allocation
// entry point returns (it points them at the scratch array for the call).
struct BroadphaseSwap {
  scx_ctx* c;
  DeviceBuffers saved;
  uint32_t cap_cells, cap_entries;
  bool order_valid;
  explicit BroadphaseSwap(scx_ctx* ctx)
      : c(ctx), saved(ctx->d), cap_cells(ctx->cap_cells), cap_entries(ctx->cap_entries), order_valid(ctx->order_valid) {}
  ~BroadphaseSwap() {
    c->d = saved;
    c->cap_cells = cap_cells;
    c->cap_entries = cap_entries;
    c->order_valid = order_valid;
  }
};

}  // namespace

extern "C" {

int scx_potential_overlaps_noalloc(scx_ctx* c, uint32_t n, const double* aabb_min, const double* aabb_max,
                                   uint64_t cap, uint32_t* pairs_out, uint64_t* npairs) {
  if (!c) return SCX_E_ARG;
  if (npairs) *npairs = 0;
  if (n == 0) return SCX_OK;
  if (!aabb_min || !aabb_max) return fail(c, SCX_E_ARG, "null box arrays");
  if (n > kNoAllocMaxBoxes)
    return fail(c, SCX_E_CAPACITY, "scx_potential_overlaps_noalloc: at most %u boxes (got %u)", kNoAllocMaxBoxes, n);
  if (!c->scratch) return fail(c, SCX_E_STATE, "no scratch address for device %d", c->device);
  HIPC(hipSetDevice(c->device));
  const int D = c->dim;
  hipStream_t s = c->stream;
  BroadphaseSwap guard(c);
  DeviceBuffers& d = c->d;
  // the broad phase's working set, carved from the scratch array (same sizes as a context
  // of n bodies: 2n + 1024 cells, 8n / 16n + 1024 cell entries)
  const uint32_t cap_cells = 2 * n + 1024;
  const uint32_t cap_entries = (D == 2 ? 8u : 16u) * n + 1024;
  Carve sc{c->scratch};
  d.aabb_min = sc.take<double>((size_t)D * n);
  d.aabb_max = sc.take<double>((size_t)D * n);
  d.red = sc.take<double>(9 * (size_t)((n + kBlock - 1) / kBlock + 1));
  d.grid = sc.take<GridParams>(1);
  d.ctr = sc.take<StepCounters>(1);
  d.cells = sc.take<uint32_t>(4 * ((size_t)cap_cells + 1));
  d.ent = sc.take<uint32_t>(cap_entries);
  d.order = sc.take<uint32_t>(n);
  d.order_next = sc.take<uint32_t>(n);
  d.tbox = sc.take<float4>(2 * (size_t)n);
  d.tnar = sc.take<double4>(2 * (size_t)n);
  d.tr = sc.take<double>(n);
  d.pcount = sc.take<uint32_t>((size_t)n + 1);
  d.scan_tmp = sc.take<uint32_t>(scan_tmp_words(std::max<uint32_t>(cap_cells + 1, kNoAllocMaxPairs)) + 4);
  d.pa = sc.take<uint32_t>(kNoAllocMaxPairs);
  d.pb = sc.take<uint32_t>(kNoAllocMaxPairs);
  uint32_t* pout = sc.take<uint32_t>(2 * (size_t)kNoAllocMaxPairs);
  if (sc.used > kNoAllocScratchBytes) return fail(c, SCX_E_CAPACITY, "scratch layout exceeds the scratch array");
  c->cap_cells = cap_cells;
  c->cap_entries = cap_entries;
  c->order_valid = false;
  HIPC(hipMemcpyAsync(d.aabb_min, aabb_min, sizeof(double) * D * n, hipMemcpyHostToDevice, s));
  HIPC(hipMemcpyAsync(d.aabb_max, aabb_max, sizeof(double) * D * n, hipMemcpyHostToDevice, s));
  HIPC(hipMemsetAsync(d.ctr, 0, sizeof(StepCounters), s));
  HIPC(launch_grid_build(c, n, true));
  HIPC(launch_pairs(c, n, 0, 0, nullptr, nullptr, 0));
  HIPC(launch_exclusive_scan(d.pcount, d.pcount, n, d.scan_tmp, s));
  HIPC(hipMemcpyAsync(d.ctr_host, d.ctr, sizeof(StepCounters), hipMemcpyDeviceToHost, s));
  HIPC(hipMemcpyAsync(d.host_u32, d.pcount + n, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  HIPC(hipStreamSynchronize(s));
  if ((d.ctr_host->overflow & 1u) || d.ctr_host->nentries > cap_entries)
    return fail(c, SCX_E_CAPACITY, "scx_potential_overlaps_noalloc: %llu cell entries exceed its %u",
                (unsigned long long)d.ctr_host->nentries, cap_entries);
  const uint32_t P = d.host_u32[0];
  if (npairs) *npairs = P;
  if (P > kNoAllocMaxPairs)
    return fail(c, SCX_E_CAPACITY, "scx_potential_overlaps_noalloc: %u pairs exceed its %u", P, kNoAllocMaxPairs);
  if (P > cap) return fail(c, SCX_E_CAPACITY, "need room for %u pairs", P);
  if (P == 0) return SCX_OK;
  if (!pairs_out) return fail(c, SCX_E_ARG, "null pair array");
  HIPC(launch_pairs(c, n, 1, 0, d.pa, d.pb, kNoAllocMaxPairs));
  k_interleave_pairs<<<(P + kBlock - 1) / kBlock, kBlock, 0, s>>>(P, d.pa, d.pb, pout);
  HIPC(hipGetLastError());
  HIPC(hipMemcpyAsync(pairs_out, pout, sizeof(uint32_t) * 2 * (size_t)P, hipMemcpyDeviceToHost, s));
  HIPC(hipMemcpyAsync(d.ctr_host, d.ctr, sizeof(StepCounters), hipMemcpyDeviceToHost, s));
  HIPC(hipStreamSynchronize(s));
  if (d.ctr_host->overflow & 2u) return fail(c, SCX_E_CAPACITY, "pair buffer overflow");
  return SCX_OK;
}

// ---------------------------------------------------------------------------
// Learned warm start: model I/O, training samples, training and gradients.
// ---------------------------------------------------------------------------
namespace {

int mlp_alloc(scx_ctx* c) {
  DeviceBuffers& d = c->d;
  if (d.mlp_params) return SCX_OK;
  HIPC(dalloc(&d.mlp_params, kMlpParams));
  HIPC(dalloc(&d.mlp_grad, kMlpParams));
  HIPC(dalloc(&d.mlp_m, kMlpParams));
  HIPC(dalloc(&d.mlp_v, kMlpParams));
  HIPC(dalloc(&d.mlp_w, (size_t)kMlpHid * kMlpIn + 2 * kMlpHid * kMlpHid + 2 * 16 * kMlpHid));
  HIPC(dalloc(&d.mlp_b, (size_t)2 * kMlpHid + 16));
  HIPC(dalloc(&d.mlp_sse, kMlpSseBlocks));
  HIPC(hipMemsetAsync(d.mlp_m, 0, sizeof(float) * kMlpParams, c->stream));
  HIPC(hipMemsetAsync(d.mlp_v, 0, sizeof(float) * kMlpParams, c->stream));
  d.mlp_step = 0;
  return SCX_OK;
}

// Device pool of n samples and activation space for batches of `batch`.
int mlp_pool(scx_ctx* c, uint32_t n, uint32_t batch, const float* X, const float* T) {
  DeviceBuffers& d = c->d;
  if (n > d.mlp_pool_cap) {
    HIPC(dalloc(&d.mlp_pool_x, (size_t)16 * n));
    HIPC(dalloc(&d.mlp_pool_t, (size_t)3 * n));
    HIPC(dalloc(&d.mlp_idx, n));
    d.mlp_pool_cap = n;
  }
  const uint32_t tiles = (batch + 15) / 16;
  if (tiles > d.mlp_cap_tiles) {
    // H1T, H2T, G1T, G2T: 128 x 16 each per tile; GY, XT: 16 x 16 each
    HIPC(dalloc(&d.mlp_act, (size_t)tiles * (4 * kMlpHid * 16 + 2 * 256)));
    d.mlp_cap_tiles = tiles;
  }
  HIPC(hipMemcpyAsync(d.mlp_pool_x, X, sizeof(float) * 16 * (size_t)n, hipMemcpyHostToDevice, c->stream));
  HIPC(hipMemcpyAsync(d.mlp_pool_t, T, sizeof(float) * 3 * (size_t)n, hipMemcpyHostToDevice, c->stream));
  return SCX_OK;
}

// Sum of the squared-error partials of the last gradient pass.
int mlp_sse(scx_ctx* c, double* out) {
  const uint32_t nb = std::max<uint32_t>(1u, c->d.mlp_sse_blocks);
  std::vector<float> part(nb);
  HIPC(hipMemcpyAsync(part.data(), c->d.mlp_sse, sizeof(float) * nb, hipMemcpyDeviceToHost, c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  double s = 0.0;
  for (float x : part) s += x;
  *out = s;
  return SCX_OK;
}

}  // namespace

int scx_set_warm_model(scx_ctx* c, const float* params, uint32_t n) {
  if (!c) return SCX_E_ARG;
  if (!params || n != (uint32_t)kMlpParams) return fail(c, SCX_E_ARG, "expected %d parameters", kMlpParams);
  HIPC(hipSetDevice(c->device));
  if (mlp_alloc(c) != SCX_OK) return SCX_E_HIP;
  HIPC(hipMemcpyAsync(c->d.mlp_params, params, sizeof(float) * kMlpParams, hipMemcpyHostToDevice, c->stream));
  HIPC(hipMemsetAsync(c->d.mlp_m, 0, sizeof(float) * kMlpParams, c->stream));
  HIPC(hipMemsetAsync(c->d.mlp_v, 0, sizeof(float) * kMlpParams, c->stream));
  c->d.mlp_step = 0;
  HIPC(launch_mlp_pack(c));
  HIPC(hipStreamSynchronize(c->stream));
  c->d.mlp_ready = true;
  return SCX_OK;
}

int scx_get_warm_model(scx_ctx* c, float* params, uint32_t n) {
  if (!c) return SCX_E_ARG;
  if (!params || n != (uint32_t)kMlpParams) return fail(c, SCX_E_ARG, "expected %d parameters", kMlpParams);
  if (!c->d.mlp_ready) return fail(c, SCX_E_STATE, "no warm-start model");
  HIPC(hipSetDevice(c->device));
  HIPC(hipMemcpy(params, c->d.mlp_params, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost));
  return SCX_OK;
}

int scx_get_samples(scx_ctx* c, uint32_t cap, float* features, float* targets, uint32_t* count) {
  if (!c) return SCX_E_ARG;
  if (count) *count = 0;
  if (!c->d.features_valid) return fail(c, SCX_E_STATE, "the last step computed no features (warm start 3 or 4)");
  HIPC(hipSetDevice(c->device));
  const uint32_t C = c->last_ncontacts;
  if (count) *count = C;
  if (C > cap) return fail(c, SCX_E_CAPACITY, "need room for %u samples", C);
  if (C == 0) return SCX_OK;
  HIPC(launch_sample_targets(c, C));
  if (features)
    HIPC(hipMemcpyAsync(features, c->d.mlp_x, sizeof(float) * 16 * (size_t)C, hipMemcpyDeviceToHost, c->stream));
  if (targets)
    HIPC(hipMemcpyAsync(targets, c->d.mlp_y, sizeof(float) * 3 * (size_t)C, hipMemcpyDeviceToHost, c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  return SCX_OK;
}

int scx_warm_model_gradients(scx_ctx* c, uint32_t n, const float* features, const float* targets, float* grads,
                             double* loss) {
  if (!c) return SCX_E_ARG;
  if (n == 0 || !features || !targets) return fail(c, SCX_E_ARG, "empty batch");
  if (!c->d.mlp_ready) return fail(c, SCX_E_STATE, "no warm-start model");
  HIPC(hipSetDevice(c->device));
  if (mlp_pool(c, n, n, features, targets) != SCX_OK) return SCX_E_HIP;
  HIPC(launch_mlp_gradients(c, n, nullptr, 0, c->d.mlp_pool_x, c->d.mlp_pool_t));
  if (grads) HIPC(hipMemcpyAsync(grads, c->d.mlp_grad, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, c->stream));
  double sse = 0.0;
  if (mlp_sse(c, &sse) != SCX_OK) return SCX_E_HIP;
  if (loss) *loss = sse / ((double)n * kMlpOut);
  return SCX_OK;
}

int scx_train_warm_model(scx_ctx* c, uint32_t n, const float* features, const float* targets, uint32_t epochs,
                         uint32_t batch, float lr, uint64_t seed, double* loss) {
  if (!c) return SCX_E_ARG;
  if (n == 0 || !features || !targets || batch == 0) return fail(c, SCX_E_ARG, "empty training set or batch");
  if (!c->d.mlp_ready) return fail(c, SCX_E_STATE, "no warm-start model to train (scx_set_warm_model first)");
  HIPC(hipSetDevice(c->device));
  batch = std::min(batch, n);
  if (mlp_pool(c, n, batch, features, targets) != SCX_OK) return SCX_E_HIP;
  std::vector<uint32_t> idx(n);
  uint64_t state = seed;
  for (uint32_t e = 0; e < epochs; ++e) {
    for (uint32_t i = 0; i < n; ++i) idx[i] = i;
    for (uint32_t i = n; i > 1; --i) {  // Fisher-Yates, splitmix64
      uint64_t z = (state += 0x9E3779B97F4A7C15ull);
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      z ^= z >> 31;
      std::swap(idx[i - 1], idx[z % i]);
    }
    HIPC(hipMemcpyAsync(c->d.mlp_idx, idx.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, c->stream));
    for (uint32_t b0 = 0; b0 + batch <= n; b0 += batch) {
      HIPC(launch_mlp_gradients(c, batch, c->d.mlp_idx, b0, c->d.mlp_pool_x, c->d.mlp_pool_t));
      HIPC(launch_mlp_adam(c, lr, ++c->d.mlp_step));
      HIPC(launch_mlp_pack(c));
    }
    // idx is reused by the next epoch's upload only after these launches (same stream)
    HIPC(hipStreamSynchronize(c->stream));
  }
  if (loss) {
    // mean squared error over (up to) the first batch-sized slice of the pool
    HIPC(launch_mlp_gradients(c, batch, nullptr, 0, c->d.mlp_pool_x, c->d.mlp_pool_t));
    double sse = 0.0;
    if (mlp_sse(c, &sse) != SCX_OK) return SCX_E_HIP;
    *loss = sse / ((double)batch * kMlpOut);
  }
  return SCX_OK;
}

int scx_unconstrained_flow(scx_ctx* c, int32_t integrator, uint32_t n, const double* m, const double* g, double dt,
                           const double* q0, const double* v0, double* q1, double* v1) {
  if (!c) return SCX_E_ARG;
  if (n == 0) return SCX_OK;
  if (!m || !g || !q0 || !v0 || !q1 || !v1) return fail(c, SCX_E_ARG, "null array");
  if (integrator != SCX_INTEGRATOR_SYMPLECTIC_EULER && integrator != SCX_INTEGRATOR_VERLET)
    return fail(c, SCX_E_ARG, "unknown integrator %d", integrator);
  HIPC(hipSetDevice(c->device));
  if (n > c->cap_bodies) {
    if (alloc_bodies(c, n) != SCX_OK || alloc_contacts(c, std::max(c->cap_contacts, 8u * n)) != SCX_OK ||
        alloc_scan_tmp(c) != SCX_OK)
      return SCX_E_HIP;
    c->has_bodies = false;
  }
  const int D = c->dim;
  hipStream_t s = c->stream;
  DeviceBuffers& d = c->d;
  // stage through the step scratch buffers (q0/v0 hold the start state, q1p/v2 the result)
  HIPC(hipMemcpyAsync(d.q1p, q0, sizeof(double) * D * n, hipMemcpyHostToDevice, s));
  HIPC(hipMemcpyAsync(d.v2, v0, sizeof(double) * D * n, hipMemcpyHostToDevice, s));
  HIPC(hipMemcpyAsync(d.aabb_min, m, sizeof(double) * n, hipMemcpyHostToDevice, s));
  double gg[3] = {g[0], g[1], D == 3 ? g[2] : 0.0};
  HIPC(launch_flow(c, n, integrator, d.aabb_min, gg, dt, d.q1p, d.v2, d.q0, d.aabb_max));
  HIPC(hipMemcpyAsync(q1, d.q0, sizeof(double) * D * n, hipMemcpyDeviceToHost, s));
  HIPC(hipMemcpyAsync(v1, d.aabb_max, sizeof(double) * D * n, hipMemcpyDeviceToHost, s));
  HIPC(hipStreamSynchronize(s));
  return SCX_OK;
}

int scx_set_timing(scx_ctx* c, int32_t enable) {
  if (!c) return SCX_E_ARG;
  c->timing = enable != 0;
  return SCX_OK;
}

int scx_set_autotune(scx_ctx* c, int32_t enable) {
  if (!c) return SCX_E_ARG;
  HIPC(hipSetDevice(c->device));
  scx::SweepTuner& T = c->tuner;
  if (enable && !T.copied) {
    for (auto& P : T.pend)
      for (hipEvent_t& e : P.ev) HIPC(hipEventCreate(&e));
    HIPC(hipEventCreate(&T.copied));
  }
  T.enabled = enable != 0;
  if (!T.enabled) {
    T.pend[0].live = T.pend[1].live = false;
    T.counters_live = false;
  }
  return SCX_OK;
}

int scx_get_autotune(scx_ctx* c, uint32_t cap, uint32_t* keys, int32_t* exponent, int32_t* blocks,
                     float* us_per_sweep, int32_t* settled, uint32_t* count) {
  if (!c) return SCX_E_ARG;
  const scx::SweepTuner& T = c->tuner;
  uint32_t k = 0;
  for (const auto& kv : T.table.entries()) {
    if (k < cap) {
      const scx::TuneEntry& E = kv.second;
      if (keys) keys[k] = kv.first;
      if (exponent) exponent[k] = E.cur;
      if (blocks) blocks[k] = scx::TuneTable::blocks(E.model, E.cur, E.max_blocks);
      if (us_per_sweep) us_per_sweep[k] = E.at(E.cur);
      if (settled) settled[k] = E.dir == 0 ? 1 : 0;
    }
    ++k;
  }
  if (count) *count = k;
  return SCX_OK;
}

int scx_get_timing(scx_ctx* c, uint32_t cap, const char** names, double* ms, uint32_t* count) {
  if (!c) return SCX_E_ARG;
  const uint32_t k = (uint32_t)c->phase_ms.size();
  if (count) *count = k;
  for (uint32_t i = 0; i < k && i < cap; ++i) {
    if (names) names[i] = c->phase_names[i];
    if (ms) ms[i] = c->phase_ms[i];
  }
  return SCX_OK;
}

int scx_get_sweep_counters(const scx_ctx* c, scx_sweep_counters* out) {
  if (!c || !out) return SCX_E_ARG;
  // the step's counters, copied to the host at the end of a step that returned stats
  const StepCounters* h = c->d.ctr_host;
  *out = scx_sweep_counters{};
  if (h) {
    out->full_evals = h->evals;
    out->full_evals_fp32 = h->evals_fp32;
    out->skipped_evals = h->evals_skipped;
    out->component_launches = h->comp_launches;
    out->component_overflows = h->comp_overflows;
    out->component_width = h->comp_maxw;
  }
  out->component_share = c->comp_share;
  return SCX_OK;
}


// ---------------------------------------------------------------------------
// Checkpoint / resume: the state that the next step's result depends on, as one byte
// image (the serializeSystem / deserializeSystem pair of ball2dcli/ball2d_cli.cpp:199-316
// for this path: bodies, static geometry, the impulse cache of ConstraintCache::serialize
// (ball2d/ConstraintCache.cpp:125-175), and this library's own ensemble layout and
// warm-start model). Layout: a fixed header, then the arrays in the order below, each as
// raw little-endian words; an FNV-1a hash of the payload guards against corruption.
// The sweep order of the broad phase and the tuner's table are not saved: no result
// depends on them.
// ---------------------------------------------------------------------------
namespace {

constexpr char kCkptMagic[8] = {'S', 'C', 'X', 'C', 'K', 'P', 'T', '1'};
constexpr uint32_t kCkptVersion = 1;
constexpr int kCacheEntryWords = 6;  // doubles per impulse-cache entry

struct CkptHeader {
  char magic[8];
  uint32_t version;
  int32_t dim;
  uint32_t n, nplanes, nscenes, scene_planes;
  uint32_t cache_count;
  int32_t cache_layout;
  uint32_t has_model, model_steps;
  uint64_t payload;  // bytes after the header
  uint64_t hash;     // FNV-1a of the payload
};
static_assert(sizeof(CkptHeader) == 64, "checkpoint header is 64 bytes");

uint64_t fnv1a(const unsigned char* p, uint64_t len) {
  uint64_t h = 1469598103934665603ull;
  for (uint64_t k = 0; k < len; ++k) {
    h ^= p[k];
    h *= 1099511628211ull;
  }
  return h;
}

// Array sizes (bytes) of the payload in order, from the header.
struct CkptLayout {
  uint64_t sizes[16];
  int count = 0;
  uint64_t total = 0;
  void add(uint64_t b) {
    sizes[count++] = b;
    total += b;
  }
};

void ckpt_layout(const CkptHeader& h, CkptLayout& L) {
  L = CkptLayout{};
  const uint64_t D = (uint64_t)h.dim, n = h.n, P = h.nplanes, S = h.nscenes, K = h.cache_count;
  L.add(8 * D * n);  // q
  L.add(8 * D * n);  // v
  L.add(8 * 3 * n);  // omega (3D; zeros in 2D)
  L.add(8 * n);      // r
  L.add(8 * n);      // m
  L.add(8 * n);      // inertia (3D; zeros in 2D)
  L.add(8 * D * P);  // plane points
  L.add(8 * D * P);  // plane normals (normalised as stored)
  if (S) {
    L.add(4 * n);            // scene of each body
    L.add(4 * (S + 1));      // scene plane offsets
    L.add(8 * S);            // scene mu
    L.add(8 * 3 * S);        // scene binning shifts
  }
  if (K) {
    L.add(8 * K);                     // cache keys
    L.add(8 * kCacheEntryWords * K);  // cache entries
  }
  if (h.has_model) {
    L.add(4 * (uint64_t)kMlpParams);  // FP32 parameters
    L.add(4 * (uint64_t)kMlpParams);  // Adam first moments
    L.add(4 * (uint64_t)kMlpParams);  // Adam second moments
  }
}

int ckpt_check(scx_ctx* c, const void* buf, uint64_t len, CkptHeader* out) {
  if (!buf || len < sizeof(CkptHeader)) return fail(c, SCX_E_ARG, "checkpoint shorter than its header");
  CkptHeader h;
  std::memcpy(&h, buf, sizeof(h));
  if (std::memcmp(h.magic, kCkptMagic, 8) != 0) return fail(c, SCX_E_ARG, "not an scx checkpoint (bad magic)");
  if (h.version != kCkptVersion) return fail(c, SCX_E_ARG, "checkpoint version %u, expected %u", h.version, kCkptVersion);
  if (h.dim != 2 && h.dim != 3) return fail(c, SCX_E_ARG, "checkpoint dimension %d", h.dim);
  if (h.n == 0 || h.n >= (1u << 29)) return fail(c, SCX_E_ARG, "checkpoint body count %u", h.n);
  if (h.nplanes > kMaxPlanesTotal || h.scene_planes > kMaxScenePlanes || h.cache_layout < 0 || h.cache_layout > 1 ||
      h.has_model > 1)
    return fail(c, SCX_E_ARG, "checkpoint header out of range");
  CkptLayout L;
  ckpt_layout(h, L);
  if (h.payload != L.total || len != sizeof(CkptHeader) + L.total)
    return fail(c, SCX_E_ARG, "checkpoint length %llu does not match its header (%llu)", (unsigned long long)len,
                (unsigned long long)(sizeof(CkptHeader) + L.total));
  if (fnv1a(static_cast<const unsigned char*>(buf) + sizeof(CkptHeader), L.total) != h.hash)
    return fail(c, SCX_E_ARG, "checkpoint payload is corrupt (hash mismatch)");
  *out = h;
  return SCX_OK;
}

}  // namespace

int scx_checkpoint_info(const void* buf, uint64_t len, int32_t* dim, uint32_t* nbodies, uint32_t* nplanes,
                        uint32_t* nscenes, uint32_t* cache_entries, int32_t* has_model) {
  CkptHeader h;
  const int rc = ckpt_check(nullptr, buf, len, &h);
  if (rc != SCX_OK) return rc;
  if (dim) *dim = h.dim;
  if (nbodies) *nbodies = h.n;
  if (nplanes) *nplanes = h.nplanes;
  if (nscenes) *nscenes = h.nscenes;
  if (cache_entries) *cache_entries = h.cache_count;
  if (has_model) *has_model = (int32_t)h.has_model;
  return SCX_OK;
}

int scx_save_checkpoint(scx_ctx* c, void* buf, uint64_t cap, uint64_t* len) {
  if (!c) return SCX_E_ARG;
  if (!len) return fail(c, SCX_E_ARG, "scx_save_checkpoint: null length");
  if (!c->has_bodies) return fail(c, SCX_E_STATE, "no bodies to checkpoint");
  HIPC(hipSetDevice(c->device));
  DeviceBuffers& d = c->d;
  CkptHeader h;
  std::memset(&h, 0, sizeof(h));
  std::memcpy(h.magic, kCkptMagic, 8);
  h.version = kCkptVersion;
  h.dim = c->dim;
  h.n = c->n;
  h.nplanes = c->nplanes;
  h.nscenes = c->nscenes;
  h.scene_planes = c->nscenes ? c->scene_planes : 0;
  h.cache_count = d.prev_count;
  h.cache_layout = d.prev_layout;
  h.has_model = d.mlp_ready ? 1 : 0;
  h.model_steps = d.mlp_ready ? d.mlp_step : 0;
  CkptLayout L;
  ckpt_layout(h, L);
  h.payload = L.total;
  *len = sizeof(CkptHeader) + L.total;
  if (!buf) return SCX_OK;
  if (cap < *len) return fail(c, SCX_E_CAPACITY, "checkpoint needs %llu bytes", (unsigned long long)*len);
  unsigned char* out = static_cast<unsigned char*>(buf) + sizeof(CkptHeader);
  const void* src[16];
  int k = 0;
  src[k++] = d.q;
  src[k++] = d.v;
  src[k++] = d.w;
  src[k++] = d.r;
  src[k++] = d.m;
  src[k++] = d.inertia;
  src[k++] = d.px;
  src[k++] = d.pn;
  if (h.nscenes) {
    src[k++] = d.bscene;
    src[k++] = d.scene_poff;
    src[k++] = d.scene_mu;
    src[k++] = d.scene_shift;
  }
  if (h.cache_count) {
    src[k++] = d.prev_key;
    src[k++] = d.prev_r;
  }
  if (h.has_model) {
    src[k++] = d.mlp_params;
    src[k++] = d.mlp_m;
    src[k++] = d.mlp_v;
  }
  for (int a = 0; a < L.count; ++a) {
    if (L.sizes[a]) HIPC(hipMemcpyAsync(out, src[a], L.sizes[a], hipMemcpyDeviceToHost, c->stream));
    out += L.sizes[a];
  }
  HIPC(hipStreamSynchronize(c->stream));
  h.hash = fnv1a(static_cast<unsigned char*>(buf) + sizeof(CkptHeader), L.total);
  std::memcpy(buf, &h, sizeof(h));
  return SCX_OK;
}

int scx_load_checkpoint(scx_ctx* c, const void* buf, uint64_t len) {
  if (!c) return SCX_E_ARG;
  CkptHeader h;
  if (ckpt_check(c, buf, len, &h) != SCX_OK) return SCX_E_ARG;
  if (h.dim != c->dim) return fail(c, SCX_E_ARG, "checkpoint of a %dD scene, context is %dD", h.dim, c->dim);
  HIPC(hipSetDevice(c->device));
  DeviceBuffers& d = c->d;
  const int D = c->dim;
  const uint32_t n = h.n;
  if (n > c->cap_bodies) {
    if (alloc_bodies(c, n) != SCX_OK) return SCX_E_HIP;
    if (alloc_entries(c, (D == 2 ? 8u : 16u) * n + 1024) != SCX_OK) return SCX_E_HIP;
    if (alloc_contacts(c, std::max(c->cap_contacts, 8u * n)) != SCX_OK) return SCX_E_HIP;
    if (alloc_scan_tmp(c) != SCX_OK) return SCX_E_HIP;
  }
  if (h.nplanes > c->cap_planes) {
    HIPC(dalloc(&d.px, 3 * (size_t)h.nplanes));
    HIPC(dalloc(&d.pn, 3 * (size_t)h.nplanes));
    c->cap_planes = h.nplanes;
  }
  if (h.nscenes) {
    if (h.nscenes > c->cap_scenes) {
      HIPC(dalloc(&d.scene_poff, (size_t)h.nscenes + 1));
      HIPC(dalloc(&d.scene_mu, h.nscenes));
      HIPC(dalloc(&d.scene_shift, (size_t)3 * h.nscenes));
      HIPC(dalloc(&d.scene_err, h.nscenes));
      HIPC(dalloc(&d.scene_cnt, h.nscenes));
      HIPC(dalloc(&d.scene_acc, (size_t)2 * h.nscenes));
      HIPC(dalloc(&d.scene_sweeps, h.nscenes));
      c->cap_scenes = h.nscenes;
    }
    HIPC(dalloc(&d.bscene, c->cap_bodies));
  }
  if (h.cache_count > d.cap_cache) {
    const uint32_t want = h.cache_count + h.cache_count / 4 + 1024;
    HIPC(dalloc(&d.prev_key, want));
    HIPC(dalloc(&d.prev_r, (size_t)kCacheEntryWords * want));
    d.cap_cache = want;
  }
  if (h.has_model && mlp_alloc(c) != SCX_OK) return SCX_E_HIP;
  CkptLayout L;
  ckpt_layout(h, L);
  void* dst[16];
  int k = 0;
  dst[k++] = d.q;
  dst[k++] = d.v;
  dst[k++] = d.w;
  dst[k++] = d.r;
  dst[k++] = d.m;
  dst[k++] = d.inertia;
  dst[k++] = d.px;
  dst[k++] = d.pn;
  if (h.nscenes) {
    dst[k++] = d.bscene;
    dst[k++] = d.scene_poff;
    dst[k++] = d.scene_mu;
    dst[k++] = d.scene_shift;
  }
  if (h.cache_count) {
    dst[k++] = d.prev_key;
    dst[k++] = d.prev_r;
  }
  if (h.has_model) {
    dst[k++] = d.mlp_params;
    dst[k++] = d.mlp_m;
    dst[k++] = d.mlp_v;
  }
  const unsigned char* in = static_cast<const unsigned char*>(buf) + sizeof(CkptHeader);
  for (int a = 0; a < L.count; ++a) {
    if (L.sizes[a]) HIPC(hipMemcpyAsync(dst[a], in, L.sizes[a], hipMemcpyHostToDevice, c->stream));
    in += L.sizes[a];
  }
  if (h.has_model) HIPC(launch_mlp_pack(c));
  HIPC(hipStreamSynchronize(c->stream));
  c->n = n;
  c->nplanes = h.nplanes;
  c->nscenes = h.nscenes;
  c->scene_planes = h.scene_planes;
  c->has_bodies = true;
  c->order_valid = false;
  c->last_ncontacts = 0;
  d.prev_count = h.cache_count;
  d.prev_layout = h.cache_layout;
  d.features_valid = false;
  // the model is part of the saved state: a checkpoint without one leaves none
  d.mlp_ready = h.has_model != 0;
  d.mlp_step = h.has_model ? h.model_steps : 0;
  return SCX_OK;
}

}  // extern "C"
