// scx_internal.h -- device data layout, context and launch helpers shared by the
// HIP translation units of libscx.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/scx.h"
#include "scx_tune.h"

namespace scx {

constexpr int kBlock = 256;           // threads per workgroup (4 wave64s)
constexpr int kScanItems = 8;         // items per thread in scans
constexpr int kScanTile = kBlock * kScanItems;
constexpr uint32_t kNoBody = 0xFFFFFFFFu;
constexpr int kMaxColours = 64;
// Contacts listed per full residual evaluation of an FP32 phase (one per thread of the
// sweep kernel's workgroup 0).
constexpr uint32_t kViolCap = 512;
constexpr int kGSBlock = 512;  // persistent sweep workgroup: 8 wave64s
// Contacts of the component phase one workgroup of the persistent sweep kernel holds in LDS.
constexpr uint32_t kCompCap = 2048;
constexpr uint32_t kMaxScenePlanes = 64;        // static planes of one scene (64-bit plane masks)
constexpr uint32_t kMaxPlanesTotal = 1u << 24;  // static planes of a context (ensembles)

// Learned warm start: 16-128-128-3 perceptron; FP32 parameter vector layout
// W1 [16][128], b1 [128], W2 [128][128], b2 [128], W3 [128][3], b3 [3].
constexpr int kMlpIn = 16, kMlpHid = 128, kMlpOut = 3;
constexpr int kMlpOffW1 = 0;
constexpr int kMlpOffB1 = kMlpOffW1 + kMlpIn * kMlpHid;
constexpr int kMlpOffW2 = kMlpOffB1 + kMlpHid;
constexpr int kMlpOffB2 = kMlpOffW2 + kMlpHid * kMlpHid;
constexpr int kMlpOffW3 = kMlpOffB2 + kMlpHid;
constexpr int kMlpOffB3 = kMlpOffW3 + kMlpHid * kMlpOut;
constexpr int kMlpParams = kMlpOffB3 + kMlpOut;
constexpr int kMlpSseBlocks = 1024;   // most workgroups of a training pass (loss partials)
constexpr int kHostWords = 128;       // pinned host scratch words (DeviceBuffers::host_u32)
constexpr int kOverflowWord = 80;     // host_u32 word that receives StepCounters::overflow

// Grid parameters, computed on the device each step (reference formula of
// ball2d/SpatialGridDetector.cpp:76-104: min corner - 2e-6, h = max axis of the summed
// extents / n). Device-resident so that no host round trip is needed.
struct GridParams {
  double min_coord[3];
  double h;
  uint32_t dims[3];      // cells per axis (dense row-major grid)
  uint32_t ncells;
  float min_f[3];        // FP32 copies used to bin the FP32 boxes
  float inv_h_f;
};

// Device-side counters of one step, read back by the host at the sync points.
struct StepCounters {
  unsigned long long nentries;   // grid entries
  uint32_t overflow;             // bit 0 entries, bit 1 pairs, bit 2 contacts, bit 3 colours
  uint32_t ncolours;
  uint32_t nrounds;
  uint32_t remaining;            // uncoloured contacts
  uint32_t sweeps;               // total sweeps
  uint32_t sweeps_fp32;
  uint32_t ncells;               // cells of this step's grid
  uint32_t evals_skipped;        // FP32 residual evaluations replaced by the violator check
  uint32_t comp_launches;        // sweep launches that ran the component phase
  uint32_t evals;                // full residual evaluations (passes over the contacts), all phases
  uint32_t evals_fp32;           // of which in the FP32 phase
  uint32_t comp_overflows;       // sweep launches whose components overflowed a workgroup's list
  uint32_t comp_maxw;            // largest number of one colour's contacts a workgroup swept in the phase
  uint32_t pad;
  double error;                  // best residual of the last phase
  double error_fp32;
};

// Per-contact record in colour order used by the sweeps. Real = float or double.
// 2D: t = (-n_y, n_x); 3D: t = n x s; arms are r_a = -la n, r_b = +lb n (sphere arms are
// parallel to the normal: rigidbody3d/Constraints/SphereSphereConstraint.cpp:132,253).
template <typename Real>
struct ContactSoA {
  uint32_t* a;
  uint32_t* b;
  Real* n;      // 3 per contact (2D: 2 used, stride 3 kept for one layout)
  Real* s;      // 3 per contact (3D)
  Real* la;
  Real* lb;
  Real* wn;     // W_nn (Delassus diagonal, normal)
  Real* wt;     // W_tt = W_ss (tangential)
  Real* mu;
  Real* forcing;
  Real* r;      // 3 per contact
};

// Online autotuner of the persistent sweep launch width (scx_set_autotune): the search
// (TuneTable, scx_tune.h) plus the HIP events that time the launches. Every width gives
// the same iterates (the contacts of one colour are independent and the residual is a
// maximum), so tuning changes only time.
struct SweepTuner {
  struct Pending {
    bool live = false;
    uint32_t key = 0;
    int exp = 0;
    hipEvent_t ev[2] = {nullptr, nullptr};
  };
  bool enabled = false;
  TuneTable table;
  Pending pend[2];              // last launch of the FP32 (0) and FP64 (1) sweep phases
  hipEvent_t copied = nullptr;  // after the copy of the step's sweep counters
  bool counters_live = false;
};

struct DeviceBuffers {
  // bodies
  double* q = nullptr;       // dim*N
  double* v = nullptr;       // dim*N
  double* w = nullptr;       // 3*N (3D)
  double* r = nullptr;       // N
  double* m = nullptr;       // N
  double* inertia = nullptr; // N (3D)
  double* q0 = nullptr;      // dim*N start-of-step copy
  double* v0 = nullptr;
  double* w0 = nullptr;
  double* q1p = nullptr;     // predictor positions
  double* v1p = nullptr;     // predictor velocities (symplectic Euler and stabilized staggerings)
  double* v2 = nullptr;      // post-impact velocity (dim*N)
  double* w2 = nullptr;      // (3D)
  // working velocities for the sweeps, per body slot: 2D [v (3) | minv], 3D
  // [v (3) | minv][omega (3) | Iinv] adjacent, so that a body is one or two vector loads
  // from one cache line
  double* bv64 = nullptr;    // 8*N
  float* bv32 = nullptr;     // 8*N
  // Body slots of the sweeps: slot_of[id] = slot, sid[slot] = id. A step uses the
  // broad phase's home-cell order (spatially sorted), so that the bodies of neighbouring
  // contacts sit in neighbouring memory; scx_friction_solve uses the identity.
  uint32_t* slot_of = nullptr;  // N
  uint32_t* lmap = nullptr;     // N, all ones between solves (endpoint owners of the LDS solve)
  const uint32_t* sid = nullptr;        // slot -> id of the current solve, nullptr = identity
  const uint32_t* slot_map = nullptr;   // id -> slot of the current solve, nullptr = identity
  // planes
  double* px = nullptr;      // dim*P
  double* pn = nullptr;      // dim*P (normalised)
  // ensemble of independent scenes (scx_set_scenes); unused when nscenes == 0
  uint32_t* bscene = nullptr;      // N: scene of each body
  uint32_t* scene_poff = nullptr;  // S + 1: plane ranges
  double* scene_mu = nullptr;      // S: friction coefficients
  double* scene_shift = nullptr;   // 3 S: binning shift of each scene (keeps scenes apart in the grid)
  double* scene_err = nullptr;     // S: per-scene residual (scx_get_scene_stats)
  uint32_t* scene_cnt = nullptr;   // S: per-scene contacts
  double* cmu = nullptr;           // per contact friction coefficient (cap_contacts)
  uint32_t* cscene = nullptr;      // per contact in colour order: its scene (cap_contacts)
  unsigned long long* scene_acc = nullptr;  // 2 S: per-scene residual bits of the sweeps
  uint32_t* scene_sweeps = nullptr;         // S: sweep at which each scene met the tolerance
  uint32_t* act_idx = nullptr;   // 2 x cap_contacts: compacted lists of the unfrozen scenes' contacts
  uint32_t* act_part = nullptr;  // kMaxColours x gs_blocks: per-colour per-workgroup counts
  // broad phase
  double* aabb_min = nullptr;  // dim*N
  double* aabb_max = nullptr;
  double* red = nullptr;       // reduction partials
  GridParams* grid = nullptr;
  uint32_t* cells = nullptr;   // 4 x (cap_cells + 1): entry counts, entry starts, home counts, home starts
  uint32_t* ent = nullptr;     // cell entries: sweep position t | first-cell flags
  // Sweep order: order[t] = body id, the home-cell order of the previous binning (any
  // permutation is correct; a spatially sorted one keeps the atomics and the neighbour
  // gathers local). order_next receives this step's home-cell order.
  uint32_t* order = nullptr;
  uint32_t* order_next = nullptr;
  float4* tbox = nullptr;      // conservative FP32 boxes in sweep order (2D: 1, 3D: 2 per body)
  double4* tnar = nullptr;     // exact FP64 narrow-phase data in sweep order (2 per body, see k_pairs)
  double* tr = nullptr;        // radii in sweep order (2D)
  uint32_t* pcount = nullptr;  // N + 1 per-owner counts / offsets
  uint32_t* scan_tmp = nullptr; // scratch of the device-wide scans
  uint32_t* pa = nullptr;      // pair lists of scx_potential_overlaps (capacity cap_pairs)
  uint32_t* pb = nullptr;
  // active set in reference order
  uint32_t* ca = nullptr;      // body a
  uint32_t* cb = nullptr;      // body b or kNoBody
  int32_t* cplane = nullptr;   // plane index or -1
  unsigned long long* pmask = nullptr;  // per-body active plane mask / per-contact colour mask
  uint32_t* mbcount = nullptr; // multi-bucket counts
  // contact setup in reference order (FP64)
  double* basis = nullptr;     // dim*dim per contact
  double* cla = nullptr;
  double* clb = nullptr;
  double* cwn = nullptr;
  double* cwt = nullptr;
  double* cforcing = nullptr;
  double* calpha = nullptr;    // outputs in reference order
  double* cbeta = nullptr;
  double* crinit = nullptr;    // warm start impulses (local) in reference order
  // colouring
  uint32_t* deg = nullptr;     // N + 1 adjacency offsets
  uint32_t* adj = nullptr;     // 2C
  int32_t* cstate = nullptr;   // (round << 8) | colour, -1 uncoloured
  uint32_t* perm = nullptr;    // colour order -> reference order
  uint32_t* coloff = nullptr;  // colour offsets (kMaxColours + 1)
  // sweep records
  ContactSoA<double> c64{};
  ContactSoA<float> c32{};
  double* rbest = nullptr;     // 3 per contact
  double* partial = nullptr;   // per-block residual partials
  uint32_t* viol = nullptr;    // FP32 phase: 2 counts, 512 listed contacts, a flag (k_gs)
  uint32_t* comp_parent = nullptr;  // N: union-find parents of the component phase (k_gs)
  uint32_t* comp_cnt = nullptr;     // 1 + gs_blocks: overflow flag, per-workgroup counts
  uint32_t* comp_list = nullptr;    // gs_blocks x kCompCap: per-workgroup contacts
  // warm-start cache (previous step's active set, sorted keys + impulses)
  // keys ascend in active-set order: (i << 32 | j) for body pairs, then
  // (1 << 63 | plane << 32 | i) for plane contacts
  unsigned long long* prev_key = nullptr;  // cap_cache
  double* prev_r = nullptr;    // cached (alpha, beta...) per key, dim each (cap_cache)
  uint32_t prev_count = 0;
  int prev_layout = 0;         // layout of the cached entries: 0 (alpha, beta...) for CACHE /
                               // NORMAL, 1 (old normal, world impulse) for TRANSPORT / LEARNED;
                               // a lookup in the other layout finds no entry
  uint32_t cap_cache = 0;
  unsigned long long* cur_key = nullptr;   // this step's keys (cap_contacts)
  // learned warm start
  float* mlp_params = nullptr;   // kMlpParams FP32 master
  float* mlp_grad = nullptr;     // kMlpParams
  float* mlp_m = nullptr;        // Adam moments
  float* mlp_v = nullptr;
  uint16_t* mlp_w = nullptr;     // packed BF16 weights (k_mlp_pack layouts)
  float* mlp_b = nullptr;        // packed biases
  uint16_t* mlp_act = nullptr;   // training activations/gradients per sample tile
  uint32_t mlp_cap_tiles = 0;    // sample tiles mlp_act holds
  float* mlp_sse = nullptr;      // kMlpSseBlocks loss partials
  uint32_t mlp_sse_blocks = 0;   // partials written by the last gradient pass
  bool mlp_ready = false;        // weights set
  uint32_t mlp_step = 0;         // Adam steps taken
  float* mlp_x = nullptr;        // per contact: 16 features (cap_contacts)
  float* mlp_y = nullptr;        // per contact: 3 predictions / targets
  float* mlp_s = nullptr;        // per contact: impulse scale
  float* mlp_pool_x = nullptr;   // training pool (cap mlp_pool_cap)
  float* mlp_pool_t = nullptr;
  uint32_t* mlp_idx = nullptr;
  uint32_t mlp_pool_cap = 0;
  bool features_valid = false;   // the last step computed mlp_x / mlp_s
  StepCounters* ctr = nullptr;
  uint32_t* barrier = nullptr;       // grid barrier state of the persistent kernels (zeroed once)
  StepCounters* ctr_host = nullptr;  // pinned
  uint32_t* host_u32 = nullptr;      // pinned scratch (kHostWords): [0..1] active-set sizes,
                                     // [2..66] colour offsets, [kOverflowWord] overflow bits,
                                     // [100..101] sweep counters of the autotuner
};

}  // namespace scx

struct scx_ctx {
  int dim = 0;
  int device = 0;
  int num_cus = 0;
  hipStream_t stream = nullptr;
  std::string err;
  uint32_t cap_bodies = 0, cap_planes = 0, cap_contacts = 0, cap_entries = 0, cap_pairs = 0, cap_cells = 0;
  uint32_t cap_perm = 0;  // perm buffer length (plane compaction and colour order)
  uint32_t n = 0, nplanes = 0;
  uint32_t nscenes = 0, cap_scenes = 0;
  uint32_t scene_planes = 0;  // most planes of one scene
  bool has_bodies = false;
  bool order_valid = false;  // d.order holds a permutation of the current n bodies
  uint32_t last_ncontacts = 0, last_npairs = 0, last_nplane = 0, last_ncolours = 0;
  bool timing = false;
  std::vector<hipEvent_t> events;
  std::vector<const char*> event_names;
  std::vector<double> phase_ms;
  std::vector<const char*> phase_names;
  int gs_blocks = 0;
  int colour_blocks = 0;
  int force_blocks = 0;  // > 0: persistent kernels use this many workgroups (capped), for tests
  bool ens_compact = true;  // ensembles: sweep compacted lists of the unfrozen scenes' contacts
  bool coherent_sweeps = true;  // colour passes with coherent body accesses and light barriers
  int violator_skip = 1;        // FP32 phase: skip full evaluations while a listed violator remains (2: required)
  bool comp_sweeps = true;      // one scene: sweep the late colours as components, one workgroup each
  // The component phase starts where at most comp_share percent of the contacts remain.
  // Adapted after every step with stats (comp_adapt): lowered when the phase overflowed a
  // workgroup's list or one colour of a workgroup took more than one pass of its threads,
  // raised while those stay well below one pass. SCX_COMPONENT_PERCENT fixes it.
  double comp_share = 20.0;
  bool comp_lds = true;         // component phase: hold a workgroup's components in LDS when they fit
  double comp_ceiling = 100.0;  // below the last share that overflowed (drifts back up)
  bool comp_adapt = true;
  // The device's cooperative queues are limited in number (HSA_QUEUE_TYPE_COOPERATIVE,
  // hsa/hsa.h: "Queues of this type may be limited in number"); while another process holds
  // them a cooperative launch fails with hipErrorIllegalState. The persistent kernels need
  // only co-resident workgroups, and one workgroup needs no cooperative launch at all (its
  // grid barrier is a workgroup barrier), so after such a failure the context runs them on
  // one workgroup with ordinary launches: slower, same results (the sweeps do not depend on
  // the launch width).
  bool coop_unavailable = false;
  // Device address of the library's file-scope scratch array on this context's device
  // (scx_potential_overlaps_noalloc), looked up once per device at scx_create.
  unsigned char* scratch = nullptr;
  // Velocities of the current step (device, dim per body): the one the contact frames and
  // forcing terms are evaluated at, and the one the solve starts from. Geometric map: both
  // v0; symplectic Euler: frames at v0, solve from v0 + dt M^-1 F (the force increment
  // folded into the start velocity); stabilized: both the predictor velocity.
  const double* vbasis = nullptr;
  const double* vsolve = nullptr;
  scx::SweepTuner tuner;
  scx::DeviceBuffers d;
};

// Launch wrappers implemented in the .hip translation units.
namespace scx {
hipError_t launch_exclusive_scan(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* tmp, hipStream_t s);
size_t scan_tmp_words(uint32_t n);
// Records a phase boundary (HIP event on the context stream) when timing is enabled.
void timing_mark(scx_ctx* c, const char* name);

// integrator: SCX_INTEGRATOR_* or kFlowForceIncrement; v1p (may be null) receives the
// predictor velocities.
hipError_t launch_predict_aabb(scx_ctx* c, const double* q0, const double* v0, const double* m, double dt,
                               const double* g, int integrator, double* q1p, double* v1p);
hipError_t launch_flow(scx_ctx* c, uint32_t n, int integrator, const double* m, const double* g, double dt,
                       const double* q0, const double* v0, double* q1, double* v1);
// given_boxes: bin the uploaded boxes d.aabb_min/max (scx_potential_overlaps);
// otherwise the predictor has written the sweep-order records of this step.
hipError_t launch_grid_build(scx_ctx* c, uint32_t n, bool given_boxes);
hipError_t launch_order_reset(scx_ctx* c, uint32_t n);
// Body slots of the sweeps from the sweep order in d.order (d.sid, d.slot_map).
hipError_t launch_slots(scx_ctx* c, uint32_t n);
hipError_t launch_pairs(scx_ctx* c, uint32_t n, int mode, int filter, uint32_t* out_a, uint32_t* out_b, uint32_t cap);
hipError_t launch_plane_masks(scx_ctx* c, uint32_t n);
hipError_t launch_multibucket(scx_ctx* c, const unsigned long long* masks, uint32_t n, uint32_t nbuckets,
                              uint32_t* out_idx, uint32_t out_cap, uint32_t* bucket_offsets);
// stab_dt > 0: add min(0, gap(q0)) / stab_dt to the normal forcing of every contact whose
// gap magnitude is at least pen_threshold.
hipError_t launch_contact_setup(scx_ctx* c, uint32_t nbb, uint32_t npl, double CoR, double stab_dt,
                                double pen_threshold);
// Ensembles: d.cmu[k] = mu of the scene of contact k (no-op for one scene).
hipError_t launch_contact_mu(scx_ctx* c, uint32_t C);
// Ensembles: per-scene contacts and residual of the last solve.
hipError_t launch_scene_stats(scx_ctx* c, uint32_t C);
hipError_t launch_colouring(scx_ctx* c, uint32_t C);
hipError_t launch_colour_sort(scx_ctx* c, uint32_t C);
hipError_t launch_gather_records(scx_ctx* c, uint32_t C, bool fp32, double mu, const double* mu_arr,
                                 const double* rinit);
// col_off: the host copy of the kMaxColours + 1 colour offsets.
hipError_t launch_solve(scx_ctx* c, uint32_t C, const uint32_t* col_off, const scx_step_params* p);
hipError_t launch_finish(scx_ctx* c, uint32_t C, const scx_step_params* p);
hipError_t launch_impulses(scx_ctx* c, uint32_t C);
// Impulse cache: keys of this step's active set and the warm start from the previous
// step's cache into d.crinit; then this step's alpha/beta into the cache.
hipError_t launch_warm_start(scx_ctx* c, uint32_t C, int mode);
hipError_t launch_cache_store(scx_ctx* c, uint32_t C, int mode);
// Learned warm start: per-contact features and impulse scales (d.mlp_x, d.mlp_s) from
// the transported cache in d.crinit; predictions d.mlp_y into d.crinit; the converged
// impulses over the scales (training targets) into d.mlp_y.
hipError_t launch_contact_features(scx_ctx* c, uint32_t C, double mu, double vref);
hipError_t launch_apply_prediction(scx_ctx* c, uint32_t C, double mu);
hipError_t launch_sample_targets(scx_ctx* c, uint32_t C);
// scx_mlp.hip
hipError_t launch_mlp_pack(scx_ctx* c);
hipError_t launch_mlp_infer(scx_ctx* c, uint32_t n, const float* X, float* Y);
hipError_t launch_mlp_gradients(scx_ctx* c, uint32_t nb, const uint32_t* idx, uint32_t b0, const float* X,
                                const float* Tgt);
hipError_t launch_mlp_adam(scx_ctx* c, float lr, uint32_t step);
int solver_max_blocks(scx_ctx* c);
int colour_max_blocks(scx_ctx* c);
}  // namespace scx

namespace scx {
// Launches a persistent kernel (workgroups that meet at grid_barrier) on `blocks` > 1
// workgroups cooperatively, so that all of them are resident. If the launch is refused
// because the device's cooperative queues are held elsewhere (hipErrorIllegalState),
// nothing runs: *launched is false, the context is marked (scx_ctx::coop_unavailable) and
// the caller launches the kernel on one workgroup instead (with the arguments of that
// width).
inline hipError_t launch_cooperative(scx_ctx* c, const void* fn, int blocks, int threads, void** args, size_t lds,
                                     hipStream_t s, bool* launched) {
  *launched = false;
  if (blocks <= 1 || c->coop_unavailable) return hipSuccess;
  const hipError_t e = hipLaunchCooperativeKernel(fn, dim3(blocks), dim3(threads), args, (unsigned)lds, s);
  if (e != hipErrorIllegalState) {
    *launched = e == hipSuccess;
    return e;
  }
  (void)hipGetLastError();
  c->coop_unavailable = true;
  return hipSuccess;
}
}  // namespace scx
