// Host side of the C ABI declared in include/lbm_amd.h.
//
// One lbm_solver owns its device state (allocated once in lbm_create):
//   Fb[2], Gb[2]  distribution sets, ping-pong (post-collision after a step)
//   u             velocity (D components), updated in place by pass B
//   phi[2]        start-of-step phi, by step parity
//   meta, wmask, normal, eps, dist, surf   geometry (set_geometry)
//   ctrl          M0, per-parity mass and mass factor, reduction counters
// lbm_step enqueues two kernels per step on the caller's stream and never
// synchronises, so step sequences are graph-capturable.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/lbm_amd.h"
#include "lbm_kernels.hpp"
#include "lbm_lists.hpp"
#include "lbm_geometry.hpp"
#include "lbm_measure.hpp"

using namespace lbm;
static_assert(kListXCD == kXCD, "host chunk lists and device block mapping disagree on XCDs");

static thread_local std::string g_err;

static int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(expr)                                                                   \
  do {                                                                                 \
    hipError_t _e = (expr);                                                            \
    if (_e != hipSuccess)                                                              \
      return fail(LBM_EHIP, std::string(#expr " failed: ") + hipGetErrorString(_e));   \
  } while (0)

constexpr int BS = 256;
constexpr int kTuneRounds = 27;  // autotuner rounds (steps: rounds * (n + 1))

struct lbm_solver {
  lbm_params p;
  int dim = 3, Q = 19;
  Grid g{};
  long long N = 0;
  size_t esz = 4;            // bytes per computed / field element
  size_t dsz = 4;            // bytes per stored distribution element
  int wet = WET_NONE;
  double hk = 0.0, mu_wall = 0.0, pck = 0.0, target = 0.0;
  bool has_target = false;
  bool corr_on = false;
  void* Fb[2] = {nullptr, nullptr};
  void* Gb[2] = {nullptr, nullptr};
  void* u = nullptr;
  void* phi[2] = {nullptr, nullptr};
  void* eps = nullptr;
  void* ampf = nullptr;
  uint8_t* meta = nullptr;
  uint8_t* bnd = nullptr;
  uint8_t* dist[2] = {nullptr, nullptr};
  uint8_t* solid = nullptr;
  uint32_t* wmask = nullptr;
  float* normal = nullptr;
  int32_t* nsw = nullptr;     // direct wetting: next-interior node of each boundary node
  void* mu = nullptr;         // Cahn-Hilliard / mu drive: mu of every node
  int gmu = GMU_NONE;
  int16_t* surf = nullptr;
  double* partials = nullptr;
  void* mparts = nullptr;
  double* dout = nullptr;     // device scratch for measure outputs
  Ctrl* ctrl = nullptr;
  unsigned int nblocks = 0;
  // host-side state machine
  bool geom = false, inited = false;
  long long k = 0;            // completed steps
  int cur = 0;                // buffer index holding the latest distributions
  bool local = true;          // buffers at cur hold post-stream values
  bool prev_local = true;     // form of buffers at cur^1 (source of P)
  bool phiu_valid = false;
  bool has_frac = false;
  int bsA = 256, bsB = 256;   // block sizes of pass A / pass B (autotuned)
  unsigned int nbA = 0, nbB = 0;
  // pass-B chunk lists (64-node chunks): plain-fluid chunks first, then the
  // rest; built at set_geometry
  uint32_t* chunks = nullptr;
  uint8_t* chunk_flag = nullptr;
  unsigned int n_plain = 0, n_wall = 0;
  // the same chunks as one list per pass-B block size (128, 256, 512) for the
  // single-launch form: per XCD range, plain chunks then wall chunks
  uint32_t* mixed[3] = {nullptr, nullptr, nullptr};
  unsigned int n_mixed[3] = {0, 0, 0};
  // pass B: 0 one launch over all nodes, 1 the two chunk-list launches,
  // 2 one launch over the per-XCD mixed list
  int split = 1;
  int capA = 0, capB = 0;     // occupancy caps, waves per SIMD (0: none)
  unsigned int ldsA = 0, ldsB = 0;   // dynamic LDS per block enforcing the caps
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph = nullptr;
  int gsteps = 0;
  // droplet injection: pressure and chemical potential of the last step
  // (internal order) kept for the observables until the next step, then the
  // drop indicator; allocated at the first injection
  void* keep = nullptr;
  bool kept = false;
  bool batched = false;       // owned by a batch engine
};

// ---------------------------------------------------------------------------
static void fast_div(uint32_t d, uint32_t* m, uint32_t* sh) {
  if (d <= 1) { *m = 0; *sh = 0; return; }
  int cl = 0;
  while ((1u << cl) < d) ++cl;
  const int p = 31 + cl;
  *m = (uint32_t)(((1ull << p) + d - 1) / d);
  *sh = (uint32_t)(p - 32);
}

// Internal node order: wall-normal (last) axis slowest -- 3D (z, x, y) with y
// fastest, 2D (y, x) with x fastest.
static Grid make_grid(int nx, int ny, int nz) {
  Grid g{};
  g.nx = nx; g.ny = ny; g.nz = nz;
  if (nz > 1) {
    g.sy = 1; g.sx = ny; g.sz = nx * ny;
    g.n1 = (uint32_t)ny; g.n2 = (uint32_t)nx;
  } else {
    g.sx = 1; g.sy = nx; g.sz = 0;
    g.n1 = (uint32_t)nx; g.n2 = 1;
  }
  g.wx = nx * g.sx; g.wy = ny * g.sy; g.wz = nz * g.sz;
  g.N = (long long)nx * ny * nz;
  fast_div(g.n1, &g.m1, &g.sh1);
  fast_div(g.n2, &g.m2, &g.sh2);
  return g;
}

template <typename T> static T TT(double v) { return (T)v; }

template <int DIM> static float w32(int i) { return W32<DIM>(i); }

// Physics constants in the run precision (see lbm_phys.hpp).
template <typename T, int DIM> static Phys<T> make_phys(const lbm_solver* s) {
  const lbm_params& p = s->p;
  constexpr int Q = Lat<DIM>::Q;
  Phys<T> ph;
  std::memset(&ph, 0, sizeof(ph));
  const double cs2 = 1.0 / 3.0;
  ph.cs2 = TT<T>(cs2);
  ph.two_cs2 = TT<T>(2.0 * cs2);
  ph.rcp_cs2 = T(1) / ph.cs2;
  ph.rcp_two_cs2 = T(1) / ph.two_cs2;
  for (int i = 0; i < Q; ++i) {
    const int c = wclass<DIM>(i);
    ph.wc[c] = (T)w32<DIM>(i);
    ph.woc[c] = ph.wc[c] / ph.cs2;                   // self._w / self.cs2
    ph.lc[c] = T(2) * ph.woc[c];                     // 2.0 * wi
    ph.lic[c] = T(2) * ph.wc[c];                     // 2.0 * self._w[i]
  }
  ph.four_beta = TT<T>(4.0 * p.beta);
  ph.kappa = TT<T>(p.kappa);
  ph.rho_l = TT<T>(p.rho_l);
  ph.rho_g = TT<T>(p.rho_g);
  ph.rho_floor = TT<T>(0.5 * p.rho_l);
  ph.drho = TT<T>(p.rho_l - p.rho_g);
  ph.visc = TT<T>((0.5 - p.tau_g) / p.tau_g);
  ph.xi = TT<T>(p.xi);
  ph.rcp_xi = T(1) / ph.xi;
  ph.ac_scale = TT<T>(p.ac_scale);
  ph.wet = s->wet;
  ph.mu_wall_add = TT<T>(s->mu_wall * p.mu_wall_scale);
  ph.phi_c = TT<T>(p.phi_c);
  ph.hk32 = (float)s->hk;
  ph.pck32 = (float)s->pck;
  const double th = p.theta_eq * M_PI / 180.0;
  ph.cos_t = TT<T>(std::cos(th));
  ph.sin_t = TT<T>(std::sin(th));
  ph.amp = TT<T>(p.geo_amplification);
  ph.amp_on = p.geo_amplification > 1.0 ? 1 : 0;
  ph.corr_on = s->corr_on ? 1 : 0;
  // sharpening factor per distance class: 1 - reduce * T(wall_weight)
  for (int c = 0; c < 16; ++c) {
    if (s->wet == WET_NONE) { ph.sfac[c] = T(1); continue; }
    const float arg = -(float)(c <= 11 ? c : 100) / 2.5f;
    const float ww = (float)std::exp((double)arg);   // correctly rounded fp32 exp
    ph.sfac[c] = T(1) - TT<T>(p.sharpen_reduce) * (T)ww;
  }
  const double tau_f = p.tau_h > 0 ? p.tau_h : p.tau_l;
  ph.acoef = TT<T>(1.0 - 0.5 / tau_f);
  ph.om_f = TT<T>(1.0 / tau_f);
  const T tl = TT<T>(p.tau_l);
  ph.ns_c = T(1) - T(0.5) / tl;
  ph.omg_c = T(1) / tl;
  ph.var_tau = p.tau_g != p.tau_l ? 1 : 0;
  ph.tl_m = TT<T>(p.tau_l - 0.5);
  ph.tg_m = TT<T>(p.tau_g - 0.5);
  bool body = false;
  for (int d = 0; d < DIM; ++d) {
    ph.gforce[d] = TT<T>(p.g_force[d]);
    if (p.g_force[d] != 0.0) body = true;
  }
  ph.body_on = body ? 1 : 0;
  ph.stab = p.stab_mode;
  ph.direct_on = (p.direct_wetting && s->wet != WET_NONE) ? 1 : 0;
  ph.s12_on = (s->wet != WET_NONE && p.boundary_relax > 0 &&
               (ph.direct_on || (s->has_target && std::fabs(s->target - 0.5) > 0.05))) ? 1 : 0;
  if (ph.direct_on) {
    // a = 6 sigma cos(theta) / kappa with math.radians' rounding
    // (lbm/fe_wetting.py:361-364)
    const double a = 6.0 * p.sigma * std::cos(p.theta_eq * (M_PI / 180.0)) / p.kappa;
    ph.direct = std::fabs(a) >= 1e-10 ? 1 : 0;
    ph.cub_4a = TT<T>(4.0 * a);
    ph.cub_1ma2 = TT<T>((1.0 - a) * (1.0 - a));
    ph.cub_m1a = TT<T>(-(1.0 - a));
    ph.cub_2a = TT<T>(2.0 * a);
    ph.lo01 = TT<T>(0.01);
    ph.hi99 = TT<T>(0.99);
    ph.nsw = s->nsw;
  }
  ph.gmu_mode = s->gmu;
  if (s->gmu != GMU_NONE) {
    const double mob = (1.0 / 3.0) * (tau_f - 0.5);
    ph.mob = TT<T>(mob);
    ph.guo = TT<T>(tau_f / (tau_f - 0.5));
    ph.mu_drive_a = TT<T>(p.mu_drive * mob);
  }
  ph.s12_r = TT<T>(p.boundary_relax);
  ph.s12_t = TT<T>(s->target);
  ph.thr_lo = TT<T>(0.15);
  ph.thr_hi = TT<T>(0.85);
  return ph;
}

static unsigned int nblk(long long n) { return (unsigned int)((n + BS - 1) / BS); }

// Dynamic LDS per workgroup that caps residency at `cap` waves per SIMD for
// blocks of `bs` threads (4 SIMDs and 160 KiB of LDS per CU; `stat` bytes of
// static LDS per block).  The step kernels are memory bound, and more resident
// waves than the memory system can keep in flight only add contention, so the
// autotuner can trade occupancy for fewer concurrent streams.  0: no cap.
static unsigned int lds_for_cap(int bs, int cap, unsigned int stat) {
  if (cap <= 0) return 0;
  const unsigned int cu = 160u * 1024u, gran = 512u;
  const int wpb = bs / 64;
  int bpc = cap * 4 / wpb;                 // blocks per CU allowed
  if (bpc < 1) bpc = 1;
  // smallest per-block total (static + dynamic, in LDS granules) that fits bpc
  // blocks but not bpc + 1
  unsigned int tot = cu / (unsigned int)(bpc + 1) + 1;
  tot = (tot + gran - 1) / gran * gran;
  if ((unsigned long long)tot * (unsigned int)bpc > cu) return 0;   // cannot express
  return tot > stat ? tot - stat : 0;
}

// ---------------------------------------------------------------------------
// Exported entry points: C linkage comes from the declarations in lbm_amd.h.

const char* lbm_last_error(void) { return g_err.c_str(); }

void lbm_params_default(lbm_params* p, int nx, int ny, int nz) {
  std::memset(p, 0, sizeof(*p));
  p->nx = nx; p->ny = ny; p->nz = nz;
  p->precision = LBM_F64;
  p->rho_l = 828.0; p->rho_g = 1.0; p->sigma = 0.01; p->xi = 5.0;
  p->tau_l = 0.0; p->tau_g = 0.6; p->tau_h = 1.0;
  p->theta_eq = 162.0;
  p->stab_mode = LBM_STAB_FAKHARI;
  p->boundary_relax = 0.3; p->sharpen_reduce = 0.8; p->mu_wall_scale = 1.0;
  p->ac_scale = 1.0; p->ghost_scale = 0.0;
  p->geometric_wetting = 0; p->geo_amplification = 1.0;
}

static int create_impl(const lbm_params* pin, lbm_solver** out, bool async, hipStream_t st);

int lbm_create(const lbm_params* pin, lbm_solver** out) {
  return create_impl(pin, out, false, nullptr);
}

// async: zero the buffers on `st` and return without synchronising (used by
// the batch engine while other cases run).
static int create_impl(const lbm_params* pin, lbm_solver** out, bool async, hipStream_t st) {
  if (!pin || !out) return fail(LBM_EINVAL, "null argument");
  const lbm_params& p = *pin;
  if (p.nx < 2 || p.ny < 2 || p.nz < 1)
    return fail(LBM_EINVAL, "grid must be at least 2 x 2 (x 2 in 3D)");
  if (p.nz == 2) return fail(LBM_EINVAL, "nz must be 1 (2D) or >= 3");
  if (p.precision < LBM_F32 || p.precision > LBM_F16)
    return fail(LBM_EINVAL, "unknown precision");
  if (!(p.tau_l > 0.5) || !(p.tau_g > 0.5))
    return fail(LBM_EINVAL, "tau_l and tau_g must exceed 0.5");
  if (p.stab_mode < 0 || p.stab_mode > 2) return fail(LBM_EINVAL, "bad stab_mode");
  if (p.interface_mode < 0 || p.interface_mode > 1)
    return fail(LBM_EINVAL, "bad interface_mode");
  const long long N = (long long)p.nx * p.ny * p.nz;
  {
    // every device array is addressed with 32-bit byte offsets: the largest,
    // a distribution set, must stay below 4 GiB
    const long long q = p.nz == 1 ? 9 : 19;
    const long long ds = p.precision == LBM_F64 ? 8 : (p.precision == LBM_F32 ? 4 : 2);
    if (q * padded_nodes(N) * ds >= (1ll << 32) || 3 * N * 8 >= (1ll << 32))
      return fail(LBM_EINVAL, "grid too large: a distribution set must stay below 4 GiB");
  }
  lbm_solver* s = new lbm_solver();
  s->p = p;
  s->dim = p.nz == 1 ? 2 : 3;
  s->Q = s->dim == 3 ? 19 : 9;
  s->g = make_grid(p.nx, p.ny, p.nz);
  s->N = N;
  s->esz = p.precision == LBM_F64 ? 8 : 4;
  s->dsz = p.precision == LBM_F64 ? 8 : (p.precision == LBM_F32 ? 4 : 2);
  s->gmu = p.interface_mode == LBM_INTERFACE_CH ? GMU_CH
           : (p.mu_drive > 0.0 ? GMU_DRIVE : GMU_NONE);
  // wetting mode (lbm/fe_ac_solver.py:223-236, lbm/fe_wetting.py:85-115)
  if (p.theta_eq != 90.0 && p.kappa > 0) {
    s->wet = (p.theta_eq > 130.0 || p.theta_eq < 50.0) ? WET_LINEAR : WET_QUADRATIC;
    if (s->wet == WET_LINEAR) {
      const double h = p.sigma * std::cos(p.theta_eq * M_PI / 180.0);
      double gs = p.ghost_scale;
      if (!(gs > 0)) gs = std::fabs(p.theta_eq - 90.0) > 60.0 ? 3.0 : 2.0;
      s->hk = h / p.kappa * gs;
      s->mu_wall = -h * 2.0;
    } else {
      s->pck = p.phi_c / p.kappa;
    }
  }
  // boundary-relax target (lbm/fe_ac_solver.py:165-177)
  if (p.theta_eq != 90.0) {
    const double c = std::cos(p.theta_eq * M_PI / 180.0);
    const double disc = 1.0 - 4.0 * (-c / 3.0);
    s->target = disc >= 0 ? 0.5 * (1.0 - std::sqrt(disc))
                          : 0.5 * (1.0 - std::cos(M_PI - p.theta_eq * M_PI / 180.0));
    s->has_target = true;
  }
  const size_t es = s->esz, Q = s->Q, D = s->dim;
  s->nblocks = nblk(N);
  s->nbA = s->nbB = s->nblocks;
  // default pass-B form: the single mixed launch for fp32 (its kernel keeps
  // the plain code's occupancy there), the two chunk-list launches otherwise
  // (16-bit storage loses occupancy to the general code's registers)
  s->split = p.precision == LBM_F32 ? 2 : 1;
  hipError_t e = hipSuccess;
#define ALLOC(ptr, bytes)                                                          \
  if (e == hipSuccess) {                                                           \
    e = hipMalloc((void**)&(ptr), (bytes));                                        \
    if (e == hipSuccess)                                                           \
      e = async ? hipMemsetAsync((ptr), 0, (bytes), st) : hipMemset((ptr), 0, (bytes)); \
  }
  for (int b = 0; b < 2; ++b) {
    ALLOC(s->Fb[b], Q * (size_t)padded_nodes(N) * s->dsz);
    ALLOC(s->Gb[b], Q * (size_t)padded_nodes(N) * s->dsz);
    ALLOC(s->phi[b], N * es);
    ALLOC(s->dist[b], N);
  }
  ALLOC(s->u, D * N * es);
  ALLOC(s->eps, N * es);
  ALLOC(s->meta, N);
  ALLOC(s->bnd, N);
  ALLOC(s->solid, N);
  ALLOC(s->wmask, N * 4);
  ALLOC(s->normal, D * N * 4);
  ALLOC(s->surf, (size_t)p.nx * p.ny * 2);
  ALLOC(s->partials, (size_t)((N + 63) / 64 + 1) * 8);   // room for 64-thread pass A blocks
  ALLOC(s->chunks, (size_t)((N + 63) / 64) * 4);
  ALLOC(s->chunk_flag, (size_t)((N + 63) / 64));
  ALLOC(s->mparts, (size_t)s->nblocks * sizeof(MeasRec));
  ALLOC(s->dout, 16 * 8);
  ALLOC(s->ctrl, sizeof(Ctrl));
  if (p.direct_wetting && s->wet != WET_NONE) ALLOC(s->nsw, N * 4);
  if (s->gmu != GMU_NONE) ALLOC(s->mu, N * es);
#undef ALLOC
  if (e == hipSuccess && !async) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    lbm_destroy(s);
    return fail(e == hipErrorOutOfMemory ? LBM_ENOMEM : LBM_EHIP,
                std::string("allocation failed: ") + hipGetErrorString(e));
  }
  *out = s;
  return LBM_OK;
}

int lbm_destroy(lbm_solver* s) {
  if (!s) return LBM_OK;
  if (s->gexec) (void)hipGraphExecDestroy(s->gexec);
  if (s->graph) (void)hipGraphDestroy(s->graph);
  void* ptrs[] = {s->Fb[0], s->Fb[1], s->Gb[0], s->Gb[1], s->u, s->phi[0], s->phi[1],
                  s->eps, s->ampf, s->meta, s->bnd, s->dist[0], s->dist[1], s->solid,
                  s->wmask, s->normal, s->surf, s->partials, s->mparts, s->dout, s->ctrl,
                  s->chunks, s->chunk_flag, s->keep, s->nsw, s->mu,
                  s->mixed[0], s->mixed[1], s->mixed[2]};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  delete s;
  return LBM_OK;
}

template <typename T, int DIM>
static int set_geometry_impl(lbm_solver* s, const void* frac, int fdt, hipStream_t st) {
  const Grid g = s->g;
  const unsigned nb = s->nblocks;
  k_boundary<DIM, BS><<<nb, BS, 0, st>>>(g, s->solid, s->bnd, s->wmask);
  if (fdt == LBM_FRAC_F64)
    k_normals_frac<double, DIM, BS><<<nb, BS, 0, st>>>(g, (const double*)frac, s->bnd, s->normal);
  else if (fdt == LBM_FRAC_F32)
    k_normals_frac<float, DIM, BS><<<nb, BS, 0, st>>>(g, (const float*)frac, s->bnd, s->normal);
  else
    k_normals_discrete<DIM, BS><<<nb, BS, 0, st>>>(g, s->solid, s->bnd, s->normal);
  k_dist_seed<BS><<<nb, BS, 0, st>>>(s->N, s->bnd, s->dist[0]);
  int a = 0;
  for (int pass = 0; pass < 11; ++pass) {   // int(2 * xi_wall) + 1, xi_wall = 5
    k_dist_pass<DIM, BS><<<nb, BS, 0, st>>>(g, s->dist[a], s->dist[a ^ 1]);
    a ^= 1;
  }
  if (a != 0)
    HIPCHK(hipMemcpyAsync(s->dist[0], s->dist[1], s->N, hipMemcpyDeviceToDevice, st));
  const int wet = s->wet != WET_NONE;
  if (fdt == LBM_FRAC_F64)
    k_meta<T, double, DIM, BS><<<nb, BS, 0, st>>>(g, s->solid, s->bnd, s->dist[0],
                                                  (const double*)frac, s->normal, s->meta,
                                                  (T*)s->eps, wet);
  else
    k_meta<T, float, DIM, BS><<<nb, BS, 0, st>>>(g, s->solid, s->bnd, s->dist[0],
                                                 (const float*)frac, s->normal, s->meta,
                                                 (T*)s->eps, wet);
  if (s->nsw) k_nsw<DIM, BS><<<nb, BS, 0, st>>>(g, s->bnd, s->solid, s->normal, s->nsw);
  const long long ncol = DIM == 3 ? (long long)g.nx * g.ny : g.nx;
  k_surface<DIM, BS><<<nblk(ncol), BS, 0, st>>>(g, s->solid, s->surf);
  // pass-B chunk lists: plain-fluid chunks, then the others (set-up only: one
  // read-back of a byte per 64 nodes)
  const size_t nch = (size_t)((s->N + 63) / 64);
  k_chunk_plain<BS><<<nblk((long long)nch * 64), BS, 0, st>>>(s->N, s->meta, s->chunk_flag);
  HIPCHK(hipGetLastError());
  std::vector<uint8_t> flag(nch);
  HIPCHK(hipMemcpyAsync(flag.data(), s->chunk_flag, nch, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  std::vector<uint32_t> list;
  list.reserve(nch);
  for (size_t c = 0; c < nch; ++c)
    if (flag[c]) list.push_back((uint32_t)c);
  s->n_plain = (unsigned int)list.size();
  for (size_t c = 0; c < nch; ++c)
    if (!flag[c]) list.push_back((uint32_t)c);
  s->n_wall = (unsigned int)(nch - s->n_plain);
  HIPCHK(hipMemcpyAsync(s->chunks, list.data(), nch * 4, hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));
  std::vector<uint32_t> mx;
  for (int k = 0; k < 3; ++k) {
    if (!mixed_list(list, s->n_plain, s->n_wall, 2u << k, mx))
      return fail(LBM_EHIP, "pass-B chunk list construction failed");
    if (s->mixed[k]) HIPCHK(hipFree(s->mixed[k]));
    s->mixed[k] = nullptr;
    HIPCHK(hipMalloc(&s->mixed[k], mx.size() * 4));
    HIPCHK(hipMemcpy(s->mixed[k], mx.data(), mx.size() * 4, hipMemcpyHostToDevice));
    s->n_mixed[k] = (unsigned int)mx.size();
  }
  return LBM_OK;
}

// (compute type, storage type, dimension) of a solver
#define DISPATCH3(s, CALL)                                                           \
  ((s)->p.precision == LBM_F64                                                       \
       ? (((s)->dim == 3) ? CALL(double, double, 3) : CALL(double, double, 2))       \
   : (s)->p.precision == LBM_F32                                                     \
       ? (((s)->dim == 3) ? CALL(float, float, 3) : CALL(float, float, 2))           \
   : (s)->p.precision == LBM_BF16                                                    \
       ? (((s)->dim == 3) ? CALL(float, __hip_bfloat16, 3)                           \
                          : CALL(float, __hip_bfloat16, 2))                          \
       : (((s)->dim == 3) ? CALL(float, __half, 3) : CALL(float, __half, 2)))
#define DISPATCH(s, CALL)                                                  \
  (((s)->esz == 8) ? (((s)->dim == 3) ? CALL(double, 3) : CALL(double, 2)) \
                   : (((s)->dim == 3) ? CALL(float, 3) : CALL(float, 2)))

// Geometry precompute from s->solid and a fraction array, both already in the
// internal node order.
static int set_geometry_internal(lbm_solver* s, const void* frac_dev, int frac_dtype,
                                 hipStream_t st) {
  if (frac_dtype == LBM_FRAC_NONE) frac_dev = nullptr;
  s->has_frac = frac_dev != nullptr;
#define CALL(T, D) set_geometry_impl<T, D>(s, frac_dev, frac_dev ? frac_dtype : 0, st)
  int rc = DISPATCH(s, CALL);
#undef CALL
  if (rc) return rc;
  // The correction acts only at boundary nodes (the only nodes with a nonzero
  // wall normal), so "no near-wall node" needs no separate flag.
  const double th = s->p.theta_eq * M_PI / 180.0;
  s->corr_on = s->wet != WET_NONE && s->p.geometric_wetting &&
               std::fabs(std::sin(th)) >= 1e-10;
  s->geom = true;
  return LBM_OK;
}

// Public-order array -> internal order (element size es bytes, comps arrays).
static int to_internal(const lbm_solver* s, const void* pub, void* in, size_t es, int comps,
                       hipStream_t st) {
  const Grid g = s->g;
#define TOI(E)                                                                              \
  (s->dim == 3 ? (k_to_internal<E, 3, BS><<<s->nblocks, BS, 0, st>>>(g, (const E*)pub, (E*)in, comps), 0) \
               : (k_to_internal<E, 2, BS><<<s->nblocks, BS, 0, st>>>(g, (const E*)pub, (E*)in, comps), 0))
  if (es == 1) TOI(uint8_t);
  else if (es == 2) TOI(uint16_t);
  else if (es == 4) TOI(uint32_t);
  else TOI(uint64_t);
#undef TOI
  HIPCHK(hipGetLastError());
  return LBM_OK;
}

static int to_public(const lbm_solver* s, const void* in, void* pub, size_t es, int comps,
                     hipStream_t st) {
  const Grid g = s->g;
#define TOP(E)                                                                              \
  (s->dim == 3 ? (k_to_public<E, 3, BS><<<s->nblocks, BS, 0, st>>>(g, (const E*)in, (E*)pub, comps), 0) \
               : (k_to_public<E, 2, BS><<<s->nblocks, BS, 0, st>>>(g, (const E*)in, (E*)pub, comps), 0))
  if (es == 1) TOP(uint8_t);
  else if (es == 2) TOP(uint16_t);
  else if (es == 4) TOP(uint32_t);
  else TOP(uint64_t);
#undef TOP
  HIPCHK(hipGetLastError());
  return LBM_OK;
}

int lbm_set_geometry(lbm_solver* s, const uint8_t* solid_dev, const void* frac_dev,
                     int frac_dtype, void* stream) {
  if (!s || !solid_dev) return fail(LBM_EINVAL, "null argument");
  if (frac_dtype != LBM_FRAC_NONE && !frac_dev) return fail(LBM_EINVAL, "null fraction");
  if (frac_dtype < 0 || frac_dtype > 2) return fail(LBM_EINVAL, "bad fraction dtype");
  if (frac_dtype == LBM_FRAC_NONE) frac_dev = nullptr;
  hipStream_t st = (hipStream_t)stream;
  // inputs arrive in the public order: reorder solid into place and the
  // fraction into a temporary (set-up only)
  if (solid_dev == s->solid) return fail(LBM_EINVAL, "solid must not alias solver memory");
  int rc = to_internal(s, solid_dev, s->solid, 1, 1, st);
  if (rc) return rc;
  void* tmp = nullptr;
  if (frac_dev) {
    const size_t fes = frac_dtype == LBM_FRAC_F64 ? 8 : 4;
    HIPCHK(hipMalloc(&tmp, (size_t)s->N * fes));
    rc = to_internal(s, frac_dev, tmp, fes, 1, st);
  }
  if (!rc) rc = set_geometry_internal(s, tmp, frac_dtype, st);
  if (tmp) {
    const hipError_t e = hipStreamSynchronize(st);
    (void)hipFree(tmp);
    if (!rc && e != hipSuccess) rc = fail(LBM_EHIP, "stream synchronize failed");
  }
  return rc;
}

int lbm_set_amplification_field(lbm_solver* s, const void* amp_dev, void* stream) {
  if (!s) return fail(LBM_EINVAL, "null solver");
  hipStream_t st = (hipStream_t)stream;
  if (!amp_dev) {
    if (s->ampf) {
      HIPCHK(hipStreamSynchronize(st));
      HIPCHK(hipFree(s->ampf));
    }
    s->ampf = nullptr;
    return LBM_OK;
  }
  if (!s->ampf) HIPCHK(hipMalloc(&s->ampf, s->N * s->esz));
  return to_internal(s, amp_dev, s->ampf, s->esz, 1, st);
}

template <typename T, typename S, int DIM>
static int init_impl(lbm_solver* s, const void* phi, const void* u, hipStream_t st) {
  const Phys<T> ph = make_phys<T, DIM>(s);
  k_init<T, S, DIM, BS><<<s->nblocks, BS, 0, st>>>(s->g, ph, (const T*)phi, (const T*)u,
                                                   (T*)s->phi[0], (T*)s->u, (S*)s->Fb[0],
                                                   (S*)s->Gb[0], s->partials);
  k_sum_partials<BS><<<1, BS, 0, st>>>(s->partials, (unsigned int)((s->N + 63) / 64),
                                       &s->ctrl->M0);
  HIPCHK(hipGetLastError());
  return LBM_OK;
}

int lbm_init_fields(lbm_solver* s, const void* phi_dev, const void* u_dev, void* stream) {
  if (!s || !phi_dev) return fail(LBM_EINVAL, "null argument");
  if (!s->geom) {
    // the reference allows init before set_solid: an all-fluid domain
    HIPCHK(hipMemsetAsync(s->solid, 0, s->N, (hipStream_t)stream));
    int rc = set_geometry_internal(s, nullptr, LBM_FRAC_NONE, (hipStream_t)stream);
    if (rc) return rc;
  }
#define CALL(T, S, D) init_impl<T, S, D>(s, phi_dev, u_dev, (hipStream_t)stream)
  int rc = DISPATCH3(s, CALL);
#undef CALL
  if (rc) return rc;
  s->inited = true;
  s->k = 0;
  s->cur = 0;
  s->local = true;
  s->prev_local = true;
  s->phiu_valid = false;
  s->kept = false;
  return LBM_OK;
}

// Block sizes the step kernels are instantiated for (the autotuner's space).
#define LBM_BS_SWITCH(bs, LAUNCH) \
  switch (bs) {                   \
    case 128: LAUNCH(128); break; \
    case 512: LAUNCH(512); break; \
    default: LAUNCH(256); break;  \
  }

template <typename T, typename S, int DIM>
static int step_impl(lbm_solver* s, int nsteps, hipStream_t st, hipEvent_t* ev = nullptr) {
  const Phys<T> ph = make_phys<T, DIM>(s);
  for (int it = 0; it < nsteps; ++it) {
    if (ev) HIPCHK(hipEventRecord(ev[3 * it + 0], st));
    const int src = s->cur, dst = s->cur ^ 1;
    const int pslot = (int)(s->k & 1), nslot = (int)((s->k + 1) & 1);
#define LAUNCH_A(B)                                                                        \
  if (s->local)                                                                             \
    k_phi<T, S, DIM, B, true><<<s->nbA, B, s->ldsA, st>>>(                                  \
        s->g, ph, (const S*)s->Fb[src], (const T*)s->phi[pslot], (T*)s->phi[nslot], s->meta, \
        (const T*)s->eps, s->partials, s->ctrl, pslot);                                      \
  else                                                                                      \
    k_phi<T, S, DIM, B, false><<<s->nbA, B, s->ldsA, st>>>(                                 \
        s->g, ph, (const S*)s->Fb[src], (const T*)s->phi[pslot], (T*)s->phi[nslot], s->meta, \
        (const T*)s->eps, s->partials, s->ctrl, pslot)
    LBM_BS_SWITCH(s->bsA, LAUNCH_A)
#undef LAUNCH_A
    if (ev) HIPCHK(hipEventRecord(ev[3 * it + 1], st));
    CollideArgs<T, S, DIM> a;
    a.g = s->g;
    a.Fin = (const S*)s->Fb[src];
    a.Gin = (const S*)s->Gb[src];
    a.Fout = (S*)s->Fb[dst];
    a.Gout = (S*)s->Gb[dst];
    a.u = (T*)s->u;
    a.phi_cur = (const T*)s->phi[nslot];
    a.phi_prev = (const T*)s->phi[pslot];
    a.meta = s->meta;
    a.wmask = s->wmask;
    a.normal = s->normal;
    a.eps = (const T*)s->eps;
    a.ampf = (const T*)s->ampf;
    a.ctrl = s->ctrl;
    a.ctrl_w = s->ctrl;
    a.partials = s->partials;
    a.nparts = (unsigned int)((s->N + 63) / 64);
    a.slot_new = nslot;
    a.slot_prev = pslot;
    a.local = s->local ? 1 : 0;
    a.phiu_valid = s->phiu_valid ? 1 : 0;
    a.mu = (const T*)s->mu;
    if (s->gmu != GMU_NONE) {
      // grad mu needs mu at every neighbour: one more pass, then pass B in
      // its grad-mu form (one launch, 256 threads)
      k_mu<T, DIM, 256><<<nblk(s->N), 256, 0, st>>>(s->g, ph, (const T*)s->phi[nslot], s->meta,
                                                    s->wmask, s->normal, (T*)s->mu);
      if (s->local)
        k_collide<T, S, DIM, 256, true, true><<<nblk(s->N), 256, 0, st>>>(a, ph);
      else
        k_collide<T, S, DIM, 256, false, true><<<nblk(s->N), 256, 0, st>>>(a, ph);
    } else if (!s->split) {
#define LAUNCH_B(B)                                               \
  if (s->local) k_collide<T, S, DIM, B, true><<<s->nbB, B, s->ldsB, st>>>(a, ph); \
  else k_collide<T, S, DIM, B, false><<<s->nbB, B, s->ldsB, st>>>(a, ph)
      LBM_BS_SWITCH(s->bsB, LAUNCH_B)
#undef LAUNCH_B
    } else {
      // plain-fluid chunks, then the rest, as two launches (the first forms
      // the mass factor) or as one launch over the mixed list (the first step
      // after init_fields, in the local form, always takes the two launches)
      const uint32_t* cp = s->chunks;
      const uint32_t* cw = s->chunks + s->n_plain;
      const unsigned int np = s->n_plain, nw = s->n_wall;
#define LAUNCH_C(B)                                                                         \
  {                                                                                         \
    const unsigned int gp = (np + B / 64 - 1) / (B / 64), gw = (nw + B / 64 - 1) / (B / 64); \
    if (s->local) {                                                                         \
      if (np) k_collide_chunks<T, S, DIM, B, true, true><<<gp, B, s->ldsB, st>>>(a, ph, cp, np, 1); \
      if (nw) k_collide_chunks<T, S, DIM, B, true, false><<<gw, B, s->ldsB, st>>>(a, ph, cw, nw, np == 0); \
    } else if (s->split == 2) {                                                             \
      const int k = B == 128 ? 0 : (B == 256 ? 1 : 2);                                      \
      k_collide_mixed<T, S, DIM, B, false><<<s->n_mixed[k] / (B / 64), B, s->ldsB, st>>>(   \
          a, ph, s->mixed[k], s->n_mixed[k]);                                                \
    } else {                                                                                \
      if (np) k_collide_chunks<T, S, DIM, B, false, true><<<gp, B, s->ldsB, st>>>(a, ph, cp, np, 1); \
      if (nw) k_collide_chunks<T, S, DIM, B, false, false><<<gw, B, s->ldsB, st>>>(a, ph, cw, nw, np == 0); \
    }                                                                                       \
  }
      LBM_BS_SWITCH(s->bsB, LAUNCH_C)
#undef LAUNCH_C
    }
    if (ev) HIPCHK(hipEventRecord(ev[3 * it + 2], st));
    s->prev_local = s->local;
    s->local = false;
    s->cur = dst;
    s->k += 1;
    s->phiu_valid = true;
    s->kept = false;
  }
  HIPCHK(hipGetLastError());
  return LBM_OK;
}

int lbm_step(lbm_solver* s, int nsteps, void* stream) {
  if (!s) return fail(LBM_EINVAL, "null solver");
  if (!s->inited) return fail(LBM_ESTATE, "step() before init_fields()");
  if (nsteps < 0) return fail(LBM_EINVAL, "negative step count");
#define CALL(T, S, D) step_impl<T, S, D>(s, nsteps, (hipStream_t)stream)
  return DISPATCH3(s, CALL);
#undef CALL
}

int lbm_graph_capture(lbm_solver* s, int nsteps) {
  if (!s) return fail(LBM_EINVAL, "null solver");
  if (!s->inited || s->k < 1 || s->local)
    return fail(LBM_ESTATE, "graph capture needs at least one eager step after init");
  if (nsteps < 2 || (nsteps & 1)) return fail(LBM_EINVAL, "nsteps must be even and >= 2");
  if (s->gexec) { (void)hipGraphExecDestroy(s->gexec); s->gexec = nullptr; }
  if (s->graph) { (void)hipGraphDestroy(s->graph); s->graph = nullptr; }
  hipStream_t cs;
  HIPCHK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  // snapshot the host state machine: capture records launches only
  const long long k0 = s->k;
  const int cur0 = s->cur;
  const bool l0 = s->local, pl0 = s->prev_local, pv0 = s->phiu_valid;
  hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  int rc = LBM_OK;
  if (e == hipSuccess) {
    rc = lbm_step(s, nsteps, cs);
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(cs, &g);
    if (rc == LBM_OK && e2 == hipSuccess) {
      s->graph = g;
      e = hipGraphInstantiate(&s->gexec, g, nullptr, nullptr, 0);
    } else {
      if (g) (void)hipGraphDestroy(g);
      if (e2 != hipSuccess) e = e2;
    }
  }
  s->k = k0; s->cur = cur0; s->local = l0; s->prev_local = pl0; s->phiu_valid = pv0;
  (void)hipStreamDestroy(cs);
  if (rc != LBM_OK) return rc;
  if (e != hipSuccess) return fail(LBM_EHIP, std::string("graph capture: ") + hipGetErrorString(e));
  s->gsteps = nsteps;
  return LBM_OK;
}

int lbm_graph_launch(lbm_solver* s, void* stream) {
  if (!s) return fail(LBM_EINVAL, "null solver");
  if (!s->gexec) return fail(LBM_ESTATE, "no captured graph");
  if (s->local || !s->phiu_valid) return fail(LBM_ESTATE, "state changed since capture");
  HIPCHK(hipGraphLaunch(s->gexec, (hipStream_t)stream));
  s->k += s->gsteps;        // even: buffer parity unchanged
  s->prev_local = false;
  return LBM_OK;
}

int lbm_profile_steps(lbm_solver* s, int nsteps, void* stream, double* out) {
  if (!s || !out) return fail(LBM_EINVAL, "null argument");
  if (!s->inited) return fail(LBM_ESTATE, "step() before init_fields()");
  if (nsteps < 1) return fail(LBM_EINVAL, "nsteps must be >= 1");
  hipStream_t st = (hipStream_t)stream;
  std::vector<hipEvent_t> ev((size_t)nsteps * 3, nullptr);
  int rc = LBM_OK;
  for (auto& e : ev)
    if (hipEventCreate(&e) != hipSuccess) rc = fail(LBM_EHIP, "hipEventCreate failed");
  if (rc == LBM_OK) {
#define CALL(T, S, D) step_impl<T, S, D>(s, nsteps, st, ev.data())
    rc = DISPATCH3(s, CALL);
#undef CALL
  }
  if (rc == LBM_OK && hipStreamSynchronize(st) != hipSuccess)
    rc = fail(LBM_EHIP, "stream synchronize failed");
  double ta = 0, tb = 0, tstep = 0;
  if (rc == LBM_OK) {
    for (int i = 0; i < nsteps; ++i) {
      float a = 0, b = 0;
      (void)hipEventElapsedTime(&a, ev[3 * i], ev[3 * i + 1]);
      (void)hipEventElapsedTime(&b, ev[3 * i + 1], ev[3 * i + 2]);
      ta += a; tb += b;
      if (i > 0) {
        float c = 0;
        (void)hipEventElapsedTime(&c, ev[3 * (i - 1)], ev[3 * i]);
        tstep += c;
      }
    }
    out[0] = ta / nsteps;
    out[1] = tb / nsteps;
    out[2] = nsteps > 1 ? tstep / (nsteps - 1) : ta + tb;
  }
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
  return rc;
}

// Exhaustive check of divk against IEEE division for one fp32 divisor: every
// one of the 2^32 fp32 bit patterns x (NaN inputs excepted) must give the
// bit-identical quotient.
__global__ void k_check_div(float d, float r, unsigned long long* bad) {
  const unsigned long long t = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned int cnt = 0;
  for (unsigned int j = 0; j < 64; ++j) {
    const unsigned int bits = (unsigned int)(t * 64ull + j);
    const float x = __builtin_bit_cast(float, bits);
    if (x != x) continue;
    const float a = divk(x, d, r);
    const float b = x / d;
    if (__builtin_bit_cast(unsigned int, a) != __builtin_bit_cast(unsigned int, b)) ++cnt;
  }
  if (cnt) atomicAdd(bad, (unsigned long long)cnt);
}

int lbm_check_division(float d, unsigned long long* mismatches, void* stream) {
  if (!mismatches) return fail(LBM_EINVAL, "null argument");
  hipStream_t st = (hipStream_t)stream;
  unsigned long long* dev = nullptr;
  HIPCHK(hipMalloc(&dev, sizeof(unsigned long long)));
  hipError_t e = hipMemsetAsync(dev, 0, sizeof(unsigned long long), st);
  const float r = 1.0f / d;
  if (e == hipSuccess) {
    // 2^32 values, 64 per thread, 256 threads per block -> 2^18 blocks
    k_check_div<<<1u << 18, 256, 0, st>>>(d, r, dev);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(mismatches, dev, sizeof(*mismatches), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipFree(dev);
  HIPCHK(e);
  return LBM_OK;
}

int lbm_set_block_sizes(lbm_solver* s, int bs_phi, int bs_collide) {
  if (!s) return fail(LBM_EINVAL, "null solver");
  auto ok = [](int b) { return b == 128 || b == 256 || b == 512; };
  if (!ok(bs_phi) || !ok(bs_collide)) return fail(LBM_EINVAL, "block size must be 128, 256 or 512");
  s->bsA = bs_phi;
  s->bsB = bs_collide;
  s->nbA = (unsigned int)((s->N + bs_phi - 1) / bs_phi);
  s->nbB = (unsigned int)((s->N + bs_collide - 1) / bs_collide);
  s->ldsA = lds_for_cap(s->bsA, s->capA, 0);
  s->ldsB = lds_for_cap(s->bsB, s->capB, 1024);   // finalize_mass's 128 doubles
  if (s->gexec) { (void)hipGraphExecDestroy(s->gexec); s->gexec = nullptr; }
  if (s->graph) { (void)hipGraphDestroy(s->graph); s->graph = nullptr; }
  return LBM_OK;
}

int lbm_set_occupancy(lbm_solver* s, int cap_phi, int cap_collide) {
  if (!s) return fail(LBM_EINVAL, "null solver");
  if (cap_phi < 0 || cap_phi > 8 || cap_collide < 0 || cap_collide > 8)
    return fail(LBM_EINVAL, "occupancy cap must be 0 (none) or 1..8 waves per SIMD");
  s->capA = cap_phi;
  s->capB = cap_collide;
  return lbm_set_block_sizes(s, s->bsA, s->bsB);
}

int lbm_set_split(lbm_solver* s, int split) {
  if (!s) return fail(LBM_EINVAL, "null solver");
  if (split < 0 || split > 2) return fail(LBM_EINVAL, "split must be 0, 1 or 2");
  s->split = split;
  return lbm_set_block_sizes(s, s->bsA, s->bsB);
}

int lbm_get_launch(const lbm_solver* s, int* out5) {
  if (!s || !out5) return fail(LBM_EINVAL, "null argument");
  out5[0] = s->bsA; out5[1] = s->bsB; out5[2] = s->capA; out5[3] = s->capB;
  out5[4] = s->split;
  return LBM_OK;
}

int lbm_autotune(lbm_solver* s, int nsteps_each, void* stream, int* best_phi,
                 int* best_collide) {
  if (!s) return fail(LBM_EINVAL, "null solver");
  if (!s->inited) return fail(LBM_ESTATE, "autotune before init_fields()");
  if (nsteps_each < 2) nsteps_each = 2;
  // (block size, occupancy cap, pass-B form) candidates; pass A and pass B
  // are timed independently (events around each pass), so round k times
  // candidate k of pass B together with candidate k % 6 of pass A
  int candB[kTuneRounds][3];
  {
    const int bsz[3] = {128, 256, 512}, caps[3] = {0, 5, 4};
    for (int k = 0; k < kTuneRounds; ++k) {
      candB[k][0] = bsz[k % 3];
      candB[k][1] = caps[(k / 3) % 3];
      candB[k][2] = k / 9;
    }
  }
  const int candA[6][2] = {{128, 0}, {256, 0}, {512, 0}, {128, 6}, {256, 6}, {512, 6}};
  double bestA = 1e30, bestB = 1e30;
  int bA[2] = {256, 0}, bB[3] = {256, 0, 1};
  for (int c = 0; c < kTuneRounds; ++c) {
    const int* a = candA[c % 6];
    const int* b = candB[c];
    s->capA = a[1];
    s->capB = b[1];
    s->split = b[2];
    int rc = lbm_set_block_sizes(s, a[0], b[0]);
    if (rc) return rc;
    double t[3];
    rc = lbm_profile_steps(s, 1, stream, t);   // first launch of this shape: untimed
    if (rc) return rc;
    rc = lbm_profile_steps(s, nsteps_each, stream, t);
    if (rc) return rc;
    if (c < 6 && t[0] < bestA) { bestA = t[0]; bA[0] = a[0]; bA[1] = a[1]; }
    if (t[1] < bestB) { bestB = t[1]; bB[0] = b[0]; bB[1] = b[1]; bB[2] = b[2]; }
  }
  s->capA = bA[1];
  s->capB = bB[1];
  s->split = bB[2];
  int rc = lbm_set_block_sizes(s, bA[0], bB[0]);
  if (rc) return rc;
  if (best_phi) *best_phi = bA[0];
  if (best_collide) *best_collide = bB[0];
  return LBM_OK;
}

size_t lbm_field_bytes(const lbm_solver* s, int which) {
  if (!s) return 0;
  const size_t N = (size_t)s->N, es = s->esz, Q = s->Q, D = s->dim;
  switch (which) {
    case LBM_FIELD_PHI: case LBM_FIELD_RHO: case LBM_FIELD_P: case LBM_FIELD_MU:
    case LBM_FIELD_PHI_RAW:
      return N * es;
    case LBM_FIELD_U: return D * N * es;
    case LBM_FIELD_F: case LBM_FIELD_G: return Q * N * es;
    case LBM_FIELD_IS_BOUNDARY: case LBM_FIELD_NEAR_WALL: case LBM_FIELD_PARTIAL:
    case LBM_FIELD_SOLID:
      return N;
    case LBM_FIELD_WALL_WEIGHT: return N * 4;
    case LBM_FIELD_WALL_NORMAL: return D * N * 4;
  }
  return 0;
}

// geometry observables from the packed byte
template <int BS_>
__global__ void k_geom_obs(Grid g, int dim, const uint8_t* __restrict__ meta,
                           const uint8_t* __restrict__ bnd, int which, void* out_) {
  const long long x = (long long)blockIdx.x * BS_ + threadIdx.x;
  if (x >= g.N) return;
  const uint32_t ex = dim == 3 ? ext_index<3>(g, decompose<3>(g, (uint32_t)x))
                               : ext_index<2>(g, decompose<2>(g, (uint32_t)x));
  // outputs in the public order
  uint8_t* out8 = (uint8_t*)out_ + ex;
  float* outf = (float*)out_ + ex;
  const uint8_t m = meta[x];
  const bool solid = m & M_SOLID;
  const int cls = dist_class(m);
  switch (which) {
    case LBM_FIELD_IS_BOUNDARY: *out8 = bnd[x]; break;
    case LBM_FIELD_NEAR_WALL: *out8 = (!solid && cls < 10) ? 1 : 0; break;
    case LBM_FIELD_PARTIAL: *out8 = (m & M_PARTIAL) ? 1 : 0; break;
    case LBM_FIELD_SOLID: *out8 = solid ? 1 : 0; break;
    case LBM_FIELD_WALL_WEIGHT: {
      float w = 0.0f;
      if (!solid) {
        const float arg = -(float)(cls <= 11 ? cls : 100) / 2.5f;
        w = (float)exp((double)arg);
      }
      *outf = w;
      break;
    }
  }
}

template <typename T, typename S, int DIM>
static int get_field_impl(lbm_solver* s, int which, void* dst, hipStream_t st,
                          bool internal = false) {
  const Phys<T> ph = make_phys<T, DIM>(s);
  ObsArgs<T, S, DIM> a;
  std::memset(&a, 0, sizeof(a));
  a.g = s->g;
  a.phi_cur = (const T*)s->phi[s->k & 1];
  a.pref = a.phi_cur;
  a.phi_prev = (const T*)s->phi[(s->k + 1) & 1];
  a.meta = s->meta;
  a.wmask = s->wmask;
  a.normal = s->normal;
  a.eps = (const T*)s->eps;
  a.ctrl = s->ctrl;
  a.out = (T*)dst;
  a.slot = (int)(s->k & 1);
  // buffers in the local form (after init_fields or an injection) hold the
  // observable state as it is; otherwise the last step's mass factor and
  // boundary relax still apply
  a.scaled = s->local ? 0 : 1;
  a.local = s->local ? 1 : 0;
  a.internal = internal ? 1 : 0;
  const size_t N = (size_t)s->N;
  if (s->kept && (which == LBM_FIELD_P || which == LBM_FIELD_MU)) {
    // an injection leaves P and mu of the last step as they were
    const char* kp = (const char*)s->keep + (which == LBM_FIELD_MU ? N * s->esz : 0);
    if (internal) {
      HIPCHK(hipMemcpyAsync(dst, kp, N * s->esz, hipMemcpyDeviceToDevice, st));
      return LBM_OK;
    }
    return to_public(s, kp, dst, s->esz, 1, st);
  }
  switch (which) {
    case LBM_FIELD_PHI: a.kind = OBS_PHI; break;
    case LBM_FIELD_PHI_RAW:
      return to_public(s, a.phi_cur, dst, s->esz, 1, st);
    case LBM_FIELD_RHO: a.kind = OBS_RHO; break;
    case LBM_FIELD_U:
      return to_public(s, s->u, dst, s->esz, s->dim, st);
    case LBM_FIELD_MU:
      a.kind = OBS_MU;
      a.init_lap = s->k == 0 ? 1 : 0;
      break;
    case LBM_FIELD_F: a.kind = OBS_F; a.Fsrc = (const S*)s->Fb[s->cur]; break;
    case LBM_FIELD_G: a.kind = OBS_G; a.Fsrc = (const S*)s->Gb[s->cur]; break;
    case LBM_FIELD_P:
      if (s->k == 0) {
        HIPCHK(hipMemsetAsync(dst, 0, N * s->esz, st));
        return LBM_OK;
      }
      a.kind = OBS_P;
      a.Fsrc = (const S*)s->Gb[s->cur ^ 1];
      a.local = s->prev_local ? 1 : 0;
      break;
    case LBM_FIELD_WALL_NORMAL:
      return to_public(s, s->normal, dst, 4, s->dim, st);
    case LBM_FIELD_IS_BOUNDARY: case LBM_FIELD_NEAR_WALL: case LBM_FIELD_WALL_WEIGHT:
    case LBM_FIELD_PARTIAL: case LBM_FIELD_SOLID:
      k_geom_obs<BS><<<s->nblocks, BS, 0, st>>>(s->g, s->dim, s->meta, s->bnd, which, dst);
      HIPCHK(hipGetLastError());
      return LBM_OK;
    default:
      return fail(LBM_EINVAL, "unknown field");
  }
  k_observe<T, S, DIM, BS><<<s->nblocks, BS, 0, st>>>(a, ph);
  HIPCHK(hipGetLastError());
  return LBM_OK;
}

int lbm_get_field(lbm_solver* s, int which, void* dst_dev, void* stream) {
  if (!s || !dst_dev) return fail(LBM_EINVAL, "null argument");
  if (!s->geom) return fail(LBM_ESTATE, "geometry not set");
  const bool geo = which >= LBM_FIELD_IS_BOUNDARY;
  if (!geo && !s->inited) return fail(LBM_ESTATE, "fields not initialised");
#define CALL(T, S, D) get_field_impl<T, S, D>(s, which, dst_dev, (hipStream_t)stream)
  return DISPATCH3(s, CALL);
#undef CALL
}

template <typename T, int DIM>
static int gradlap_impl(lbm_solver* s, const void* phi, void* grad, void* lap, hipStream_t st) {
  const Phys<T> ph = make_phys<T, DIM>(s);
  // phi arrives in the public order: reorder into a temporary first
  void* tmp = nullptr;
  HIPCHK(hipMalloc(&tmp, (size_t)s->N * sizeof(T)));
  int rc = to_internal(s, phi, tmp, sizeof(T), 1, st);
  if (!rc) {
    k_gradlap<T, DIM, BS><<<s->nblocks, BS, 0, st>>>(s->g, ph, (const T*)tmp,
                                                     s->geom ? s->meta : nullptr, s->wmask,
                                                     s->normal, (T*)grad, (T*)lap);
    if (hipGetLastError() != hipSuccess) rc = fail(LBM_EHIP, "k_gradlap launch failed");
  }
  const hipError_t e = hipStreamSynchronize(st);
  (void)hipFree(tmp);
  if (!rc && e != hipSuccess) rc = fail(LBM_EHIP, "stream synchronize failed");
  return rc;
}

int lbm_grad_lap(lbm_solver* s, const void* phi_dev, void* grad_dev, void* lap_dev,
                 void* stream) {
  if (!s || !phi_dev || !grad_dev || !lap_dev) return fail(LBM_EINVAL, "null argument");
#define CALL(T, D) gradlap_impl<T, D>(s, phi_dev, grad_dev, lap_dev, (hipStream_t)stream)
  return DISPATCH(s, CALL);
#undef CALL
}

int lbm_get_mass(lbm_solver* s, double* m0, double* m_last, void* stream) {
  if (!s) return fail(LBM_EINVAL, "null solver");
  Ctrl c;
  hipStream_t st = (hipStream_t)stream;
  HIPCHK(hipMemcpyAsync(&c, s->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (m0) *m0 = c.M0;
  if (m_last) *m_last = s->local ? c.M0 : c.mass[s->k & 1];
  return LBM_OK;
}

int lbm_set_mass_target(lbm_solver* s, double m0) {
  if (!s) return fail(LBM_EINVAL, "null solver");
  HIPCHK(hipMemcpy(&s->ctrl->M0, &m0, sizeof(double), hipMemcpyHostToDevice));
  return LBM_OK;
}

long long lbm_step_count(const lbm_solver* s) { return s ? s->k : -1; }

// ---------------------------------------------------------------------------
// Mid-run droplet injection (lbm/fe_ac_solver.py:645-732).
// ---------------------------------------------------------------------------
template <typename T, typename S, int DIM>
static int inject_impl(lbm_solver* s, const void* c_pub, const double* uimp, hipStream_t st) {
  const Phys<T> ph = make_phys<T, DIM>(s);
  const size_t fb = (size_t)s->N * s->esz;
  if (!s->keep) HIPCHK(hipMalloc(&s->keep, 3 * fb));
  char* kp = (char*)s->keep;
  if (!s->kept) {
    // the injection changes phi; P and mu stay those of the last step
    int rc = get_field_impl<T, S, DIM>(s, LBM_FIELD_P, kp, st, true);
    if (!rc) rc = get_field_impl<T, S, DIM>(s, LBM_FIELD_MU, kp + fb, st, true);
    if (rc) return rc;
  }
  int rc = to_internal(s, c_pub, kp + 2 * fb, s->esz, 1, st);
  if (rc) return rc;
  const int pslot = (int)(s->k & 1), oslot = pslot ^ 1;
  InjectArgs<T, S, DIM> a;
  std::memset(&a, 0, sizeof(a));
  a.g = s->g;
  a.Fin = (const S*)s->Fb[s->cur];
  a.Gin = (const S*)s->Gb[s->cur];
  a.Fout = (S*)s->Fb[s->cur ^ 1];
  a.Gout = (S*)s->Gb[s->cur ^ 1];
  a.pref = (const T*)s->phi[pslot];
  a.C = (const T*)(kp + 2 * fb);
  a.P = (const T*)kp;
  a.u = (T*)s->u;
  // neighbours read phi[pslot] as their 16-bit reference: the new phi goes
  // to the other slot first and is copied over after the pass
  a.phi_out = (T*)s->phi[oslot];
  a.meta = s->meta;
  a.eps = (const T*)s->eps;
  a.ctrl = s->ctrl;
  a.partials = s->partials;
  for (int d = 0; d < DIM; ++d) a.uimp[d] = uimp ? (T)uimp[d] : T(0);
  a.slot = pslot;
  if (s->local)
    k_inject<T, S, DIM, BS, true><<<s->nblocks, BS, 0, st>>>(a, ph);
  else
    k_inject<T, S, DIM, BS, false><<<s->nblocks, BS, 0, st>>>(a, ph);
  k_sum_partials<BS><<<1, BS, 0, st>>>(s->partials, (unsigned int)((s->N + 63) / 64),
                                       &s->ctrl->M0);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(s->phi[pslot], s->phi[oslot], fb, hipMemcpyDeviceToDevice, st));
  return LBM_OK;
}

int lbm_inject_droplet(lbm_solver* s, const void* c_dev, const double* u_impact, void* stream) {
  if (!s || !c_dev) return fail(LBM_EINVAL, "null argument");
  if (!s->inited) return fail(LBM_ESTATE, "inject_droplet() before init_fields()");
  if (s->batched) return fail(LBM_EINVAL, "cases of a batch engine take no injections");
#define CALL(T, S, D) inject_impl<T, S, D>(s, c_dev, u_impact, (hipStream_t)stream)
  int rc = DISPATCH3(s, CALL);
#undef CALL
  if (rc) return rc;
  s->cur ^= 1;
  s->local = true;
  s->phiu_valid = false;
  s->kept = true;
  // a captured graph assumes the post-collision form of the buffers
  if (s->gexec) { (void)hipGraphExecDestroy(s->gexec); s->gexec = nullptr; }
  if (s->graph) { (void)hipGraphDestroy(s->graph); s->graph = nullptr; }
  s->gsteps = 0;
  return LBM_OK;
}

// ---------------------------------------------------------------------------
// Solver state snapshot (preemption / checkpoint).  Layout of the byte image:
//   [0, 256)   StateHeader
//   then       F[cur], G[cur], G[cur^1] (AoSoA with the header's block size,
//              the step's distributions and
//              the previous g set the P observable is pulled from), u,
//              phi[0], phi[1], Ctrl, kept P, kept mu
// each array starting at a 256-byte boundary.  The kept P and mu (observables
// right after a droplet injection) are copied only when the header says so.  Geometry is not included: the
// restoring solver gets the same set_geometry first.
// ---------------------------------------------------------------------------
namespace {
struct StateHeader {
  uint32_t magic, version;
  int32_t nx, ny, nz, precision;
  long long k;
  int32_t cur, local, prev_local, phiu_valid, kept, block;   // block: AoSoA nodes per block
  uint64_t bytes;
};
constexpr uint32_t kStateMagic = 0x4c424d53u;   // "LBMS"
constexpr uint32_t kStateVersion = 3;

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

struct StateLayout {
  size_t dist, u, phi, off[9], total;
};

StateLayout state_layout(const lbm_solver* s) {
  StateLayout L;
  L.dist = (size_t)s->Q * (size_t)padded_nodes(s->N) * s->dsz;
  L.u = (size_t)s->dim * (size_t)s->N * s->esz;
  L.phi = (size_t)s->N * s->esz;
  const size_t sz[9] = {L.dist, L.dist, L.dist, L.u, L.phi, L.phi, sizeof(Ctrl), L.phi, L.phi};
  size_t o = 256;
  for (int i = 0; i < 9; ++i) {
    L.off[i] = o;
    o += align256(sz[i]);
  }
  L.total = o;
  return L;
}
}  // namespace

size_t lbm_state_bytes(const lbm_solver* s) { return s ? state_layout(s).total : 0; }

int lbm_save_state(lbm_solver* s, void* dst, void* stream) {
  if (!s || !dst) return fail(LBM_EINVAL, "null argument");
  if (!s->inited) return fail(LBM_ESTATE, "nothing to save before init_fields()");
  hipStream_t st = (hipStream_t)stream;
  const StateLayout L = state_layout(s);
  StateHeader h{};
  h.magic = kStateMagic;
  h.version = kStateVersion;
  h.nx = s->p.nx; h.ny = s->p.ny; h.nz = s->p.nz; h.precision = s->p.precision;
  h.k = s->k;
  h.cur = s->cur; h.local = s->local; h.prev_local = s->prev_local; h.phiu_valid = s->phiu_valid;
  h.kept = s->kept;
  h.block = kAB;
  h.bytes = L.total;
  char* d = (char*)dst;
  const void* src[9] = {s->Fb[s->cur], s->Gb[s->cur], s->Gb[s->cur ^ 1], s->u, s->phi[0],
                        s->phi[1], s->ctrl, s->keep, (const char*)s->keep + L.phi};
  const size_t sz[9] = {L.dist, L.dist, L.dist, L.u, L.phi, L.phi, sizeof(Ctrl), L.phi, L.phi};
  for (int i = 0; i < (s->kept ? 9 : 7); ++i)
    HIPCHK(hipMemcpyAsync(d + L.off[i], src[i], sz[i], hipMemcpyDefault, st));
  // the header comes from host memory that goes out of scope: synchronous copy
  HIPCHK(hipStreamSynchronize(st));
  HIPCHK(hipMemcpy(d, &h, sizeof(h), hipMemcpyDefault));
  return LBM_OK;
}

int lbm_load_state(lbm_solver* s, const void* src, void* stream) {
  if (!s || !src) return fail(LBM_EINVAL, "null argument");
  if (!s->geom) return fail(LBM_ESTATE, "set_geometry before load_state");
  hipStream_t st = (hipStream_t)stream;
  StateHeader h{};
  HIPCHK(hipStreamSynchronize(st));
  HIPCHK(hipMemcpy(&h, src, sizeof(h), hipMemcpyDefault));
  const StateLayout L = state_layout(s);
  if (h.magic != kStateMagic) return fail(LBM_EINVAL, "not a solver state image");
  if (h.version != kStateVersion || h.block != kAB)
    return fail(LBM_EINVAL, "state image written with another distribution layout");
  if (h.nx != s->p.nx || h.ny != s->p.ny || h.nz != s->p.nz || h.precision != s->p.precision ||
      h.bytes != L.total)
    return fail(LBM_EINVAL, "state image is for another grid or precision");
  if (h.cur < 0 || h.cur > 1) return fail(LBM_EINVAL, "corrupt state image");
  if (h.kept && !s->keep) HIPCHK(hipMalloc(&s->keep, 3 * L.phi));
  const char* d = (const char*)src;
  void* dstp[9] = {s->Fb[h.cur], s->Gb[h.cur], s->Gb[h.cur ^ 1], s->u, s->phi[0], s->phi[1],
                   s->ctrl, s->keep, (char*)s->keep + L.phi};
  const size_t sz[9] = {L.dist, L.dist, L.dist, L.u, L.phi, L.phi, sizeof(Ctrl), L.phi, L.phi};
  for (int i = 0; i < (h.kept ? 9 : 7); ++i)
    HIPCHK(hipMemcpyAsync(dstp[i], d + L.off[i], sz[i], hipMemcpyDefault, st));
  s->k = h.k;
  s->cur = h.cur;
  s->local = h.local != 0;
  s->prev_local = h.prev_local != 0;
  s->phiu_valid = h.phiu_valid != 0;
  s->kept = h.kept != 0;
  s->inited = true;
  if (s->gexec) { (void)hipGraphExecDestroy(s->gexec); s->gexec = nullptr; }
  if (s->graph) { (void)hipGraphDestroy(s->graph); s->graph = nullptr; }
  return LBM_OK;
}


// dev_out: 10 doubles of device memory receive the result (no copy back, no
// synchronisation); otherwise the solver's own buffer is copied to `out`.
template <typename T, int DIM>
static int measure_impl(lbm_solver* s, double* out, hipStream_t st, bool dev_out = false) {
  const Phys<T> ph = make_phys<T, DIM>(s);
  const unsigned int nb = (unsigned int)((s->N + (long long)BS * kMeasPer - 1) / ((long long)BS * kMeasPer));
  k_measure<T, DIM, BS><<<nb, BS, 0, st>>>(
      s->g, ph, (const T*)s->phi[s->k & 1], s->meta, s->ctrl, (int)(s->k & 1),
      s->local ? 0 : 1, (MeasRec*)s->mparts);
  k_measure_final<DIM, BS><<<1, BS, 0, st>>>((const MeasRec*)s->mparts, nb,
                                             dev_out ? out : s->dout);
  HIPCHK(hipGetLastError());
  if (dev_out) return LBM_OK;
  HIPCHK(hipMemcpyAsync(out, s->dout, 10 * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return LBM_OK;
}

int lbm_measure(lbm_solver* s, double* out, void* stream) {
  if (!s || !out) return fail(LBM_EINVAL, "null argument");
  if (!s->inited) return fail(LBM_ESTATE, "fields not initialised");
#define CALL(T, D) measure_impl<T, D>(s, out, (hipStream_t)stream)
  return DISPATCH(s, CALL);
#undef CALL
}

int lbm_measure_async(lbm_solver* s, double* out_dev, void* stream) {
  if (!s || !out_dev) return fail(LBM_EINVAL, "null argument");
  if (!s->inited) return fail(LBM_ESTATE, "fields not initialised");
#define CALL(T, D) measure_impl<T, D>(s, out_dev, (hipStream_t)stream, true)
  return DISPATCH(s, CALL);
#undef CALL
}

template <typename T, int DIM>
static int contact_impl(lbm_solver* s, double cx, double cy, double* out, hipStream_t st) {
  const Phys<T> ph = make_phys<T, DIM>(s);
  const long long ncol = (long long)s->g.nx * s->g.ny;
  k_contact<T, DIM, BS><<<nblk(ncol), BS, 0, st>>>(
      s->g, ph, (const T*)s->phi[s->k & 1], s->meta, s->surf, s->ctrl, (int)(s->k & 1),
      s->local ? 0 : 1, cx, cy, (double*)s->mparts);
  k_contact_final<BS><<<1, BS, 0, st>>>((const double*)s->mparts, nblk(ncol), s->dout + 12);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, s->dout + 12, sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return LBM_OK;
}

int lbm_contact_radius(lbm_solver* s, double cx, double cy, double* out, void* stream) {
  if (!s || !out) return fail(LBM_EINVAL, "null argument");
  if (!s->inited) return fail(LBM_ESTATE, "fields not initialised");
  if (s->dim != 3) return fail(LBM_EINVAL, "contact radius is defined for 3D grids");
#define CALL(T, D) contact_impl<T, D>(s, cx, cy, out, (hipStream_t)stream)
  return DISPATCH(s, CALL);
#undef CALL
}

// ===========================================================================
// Batch engine: many cases of uneven size advanced together (lbm_batch_*).
// ===========================================================================
#include "lbm_batch.hpp"

#include <vector>

struct lbm_batch {
  int precision = LBM_F32;
  size_t budget = 0, used = 0;
  std::vector<lbm_solver*> cs;        // slot -> case (nullptr: free)
  std::vector<long long> k_end;
  std::vector<size_t> bytes;
  int bsA = 256, bsB = 256;
  bool dirty = true;
  // device tables
  void* dcases = nullptr;             // BatchCase<T,S>[cap]
  size_t dcases_cap = 0;
  uint32_t* mapA = nullptr;           // [2 * nblocksA]: case, base
  uint32_t* mapB = nullptr;
  uint32_t* mapP = nullptr;           // pass B over plain-fluid chunks
  uint32_t* mapW = nullptr;           // pass B over the other chunks
  size_t mapA_cap = 0, mapB_cap = 0, mapP_cap = 0, mapW_cap = 0;
  unsigned int nA = 0, nB = 0, nP = 0, nW = 0;   // grid sizes
  bool split = true;                  // pass B as the two chunk-list launches
  int ncase = 0;                      // descriptor count (== cs.size())
  // scratch for device-side case generation
  void* gen = nullptr;
  size_t gen_cap = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int gsteps = 0;
};

static size_t case_bytes(const lbm_params& p) {
  const size_t N = (size_t)p.nx * p.ny * p.nz;
  const int dim = p.nz == 1 ? 2 : 3, Q = dim == 3 ? 19 : 9;
  const size_t es = p.precision == LBM_F64 ? 8 : 4;
  const size_t ds = p.precision == LBM_F64 ? 8 : (p.precision == LBM_F32 ? 4 : 2);
  const size_t nb = (N + 255) / 256;
  return 4 * Q * (size_t)padded_nodes((long long)N) * ds + 2 * N * es + 2 * N + dim * N * es + N * es + 3 * N + 4 * N +
         dim * N * 4 + (size_t)p.nx * p.ny * 2 + ((N + 63) / 64 + 1) * 8 +
         nb * sizeof(MeasRec) + 128 + sizeof(Ctrl);
}

size_t lbm_case_footprint(const lbm_params* p) { return p ? case_bytes(*p) : 0; }

int lbm_batch_create(int precision, size_t budget, lbm_batch** out) {
  if (!out) return fail(LBM_EINVAL, "null argument");
  if (precision < LBM_F32 || precision > LBM_F16) return fail(LBM_EINVAL, "unknown precision");
  lbm_batch* b = new lbm_batch();
  b->precision = precision;
  b->budget = budget;
  *out = b;
  return LBM_OK;
}

static void batch_drop_graph(lbm_batch* b) {
  if (b->gexec) { (void)hipGraphExecDestroy(b->gexec); b->gexec = nullptr; }
  if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
}

int lbm_batch_destroy(lbm_batch* b) {
  if (!b) return LBM_OK;
  (void)hipDeviceSynchronize();
  batch_drop_graph(b);
  for (lbm_solver* s : b->cs)
    if (s) lbm_destroy(s);
  void* ptrs[] = {b->dcases, b->mapA, b->mapB, b->mapP, b->mapW, b->gen};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  delete b;
  return LBM_OK;
}

// Device-side case generation into the batch scratch, then the normal
// set_geometry / init_fields path of the case's solver.
template <int DIM>
static int gen_case(lbm_batch* b, lbm_solver* s, const lbm_case_spec& sp, hipStream_t st) {
  const size_t N = (size_t)s->N;
  const size_t need = N * (1 + 8 + 4 + 4 + 3 * 4) + 256;
  if (need > b->gen_cap) {
    if (b->gen) {
      HIPCHK(hipStreamSynchronize(st));
      HIPCHK(hipFree(b->gen));
      b->gen = nullptr;
    }
    HIPCHK(hipMalloc(&b->gen, need));
    b->gen_cap = need;
  }
  char* base = (char*)b->gen;
  double* frac64 = (double*)base;
  float* phi = (float*)(base + 8 * N);
  float* u = (float*)(base + 12 * N);
  float* frac32 = (float*)(base + 24 * N);
  uint8_t* solid = (uint8_t*)(base + 28 * N);
  const double Rg = std::fabs(sp.R_star) * sp.R_d;
  const int kind = sp.substrate;
  k_gen_substrate<DIM, BS><<<s->nblocks, BS, 0, st>>>(
      s->g, kind, Rg, kind == SUB_FLAT ? nullptr : frac64, kind == SUB_FLAT ? frac32 : nullptr,
      solid);
  double c[3] = {sp.center[0], sp.center[1], DIM == 3 ? sp.center[2] : 0.0};
  const int ax = DIM - 1;   // wall-normal axis
  if (sp.center[ax] < 0) {
    // placement rule of the reference's sweep driver
    // (scripts/run_thesis_sweep.py:146-157): gap 2 above the substrate top at
    // the centre column (concave: the highest rim), clamped below the lid
    HIPCHK(hipGetLastError());
    const long long ncol = DIM == 3 ? (long long)s->g.nx * s->g.ny : s->g.nx;
    k_surface<DIM, BS><<<nblk(ncol), BS, 0, st>>>(s->g, solid, s->surf);
    std::vector<int16_t> h((size_t)ncol);
    HIPCHK(hipMemcpyAsync(h.data(), s->surf, ncol * 2, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    int top;
    if (kind == SUB_CONCAVE) {
      top = -1;
      for (int16_t v : h) top = v > top ? v : top;
    } else {
      const long long col = DIM == 3 ? (long long)(s->g.nx / 2) * s->g.ny + s->g.ny / 2
                                     : (long long)(s->g.nx / 2);
      top = h[(size_t)col];
    }
    const int nax = DIM == 3 ? s->g.nz : s->g.ny;
    double cz = top + 2.0 + sp.radius;
    const double mx = nax - sp.radius - 2;
    if (cz > mx) cz = mx;
    c[ax] = cz;
  }
  k_gen_droplet<DIM, BS><<<s->nblocks, BS, 0, st>>>(
      s->g, c[0], c[1], c[2], sp.radius, s->p.xi, (float)sp.u_impact[0],
      (float)sp.u_impact[1], (float)sp.u_impact[2], phi, u);
  HIPCHK(hipGetLastError());
  // the generated substrate is in the internal order already
  HIPCHK(hipMemcpyAsync(s->solid, solid, N, hipMemcpyDeviceToDevice, st));
  int rc = set_geometry_internal(s, kind == SUB_FLAT ? (const void*)frac32 : (const void*)frac64,
                                 kind == SUB_FLAT ? LBM_FRAC_F32 : LBM_FRAC_F64, st);
  if (rc) return rc;
  if (s->esz == 8) {
    // fp64 runs take fp32-quantised inputs promoted to fp64, as the
    // reference's generators hand float32 arrays to init_fields
    return fail(LBM_EINVAL, "batch engine runs fp32 compute (f32, bf16, fp16 storage)");
  }
  return lbm_init_fields(s, phi, u, st);
}

int lbm_batch_add(lbm_batch* b, const lbm_case_spec* sp, void* stream, int* case_id) {
  if (!b || !sp || !case_id) return fail(LBM_EINVAL, "null argument");
  if (sp->p.precision != b->precision)
    return fail(LBM_EINVAL, "case precision differs from the batch precision");
  if (sp->p.precision == LBM_F64) return fail(LBM_EINVAL, "batch engine runs fp32 compute");
  if (sp->p.interface_mode != LBM_INTERFACE_AC || sp->p.mu_drive > 0.0)
    return fail(LBM_EINVAL, "batch engine runs the Allen-Cahn source without the mu drive");
  if (sp->steps < 0) return fail(LBM_EINVAL, "negative step target");
  if (sp->substrate < 0 || sp->substrate > 3) return fail(LBM_EINVAL, "bad substrate");
  const size_t need = case_bytes(sp->p);
  if (b->budget && b->used + need > b->budget)
    return fail(LBM_ENOMEM, "case does not fit the batch HBM budget");
  hipStream_t st = (hipStream_t)stream;
  lbm_solver* s = nullptr;
  int rc = create_impl(&sp->p, &s, true, st);
  if (!rc) s->batched = true;
  if (rc) return rc;
  rc = s->dim == 3 ? gen_case<3>(b, s, *sp, st) : gen_case<2>(b, s, *sp, st);
  if (rc) { (void)hipStreamSynchronize(st); lbm_destroy(s); return rc; }
  int slot = -1;
  for (size_t i = 0; i < b->cs.size(); ++i)
    if (!b->cs[i]) { slot = (int)i; break; }
  if (slot < 0) {
    slot = (int)b->cs.size();
    b->cs.push_back(nullptr);
    b->k_end.push_back(0);
    b->bytes.push_back(0);
  }
  b->cs[slot] = s;
  b->k_end[slot] = sp->steps;
  b->bytes[slot] = need;
  b->used += need;
  b->dirty = true;
  *case_id = slot;
  return LBM_OK;
}

int lbm_batch_remove(lbm_batch* b, int id) {
  if (!b || id < 0 || id >= (int)b->cs.size() || !b->cs[id]) return fail(LBM_EINVAL, "bad case id");
  (void)hipDeviceSynchronize();
  lbm_destroy(b->cs[id]);
  b->cs[id] = nullptr;
  b->used -= b->bytes[id];
  b->bytes[id] = 0;
  b->dirty = true;
  return LBM_OK;
}

int lbm_batch_solver(lbm_batch* b, int id, lbm_solver** s) {
  if (!b || !s || id < 0 || id >= (int)b->cs.size() || !b->cs[id]) return fail(LBM_EINVAL, "bad case id");
  *s = b->cs[id];
  return LBM_OK;
}

int lbm_batch_load_state(lbm_batch* b, int id, const void* src, void* stream) {
  if (!b || id < 0 || id >= (int)b->cs.size() || !b->cs[id]) return fail(LBM_EINVAL, "bad case id");
  lbm_solver* s = b->cs[id];
  int rc = lbm_load_state(s, src, stream);
  if (rc) return rc;
  if (s->k > b->k_end[id]) return fail(LBM_EINVAL, "state is past the case's step target");
  // the per-case descriptors (step counter, buffer parity, form) are rebuilt
  b->dirty = true;
  return LBM_OK;
}

int lbm_batch_status(lbm_batch* b, int id, long long* done, long long* target) {
  if (!b || id < 0 || id >= (int)b->cs.size() || !b->cs[id]) return fail(LBM_EINVAL, "bad case id");
  if (done) *done = b->cs[id]->k;
  if (target) *target = b->k_end[id];
  return LBM_OK;
}

size_t lbm_batch_bytes_used(const lbm_batch* b) { return b ? b->used : 0; }

int lbm_batch_set_split(lbm_batch* b, int split) {
  if (!b) return fail(LBM_EINVAL, "null batch");
  b->split = split != 0;
  b->dirty = true;
  return LBM_OK;
}

int lbm_batch_set_block_sizes(lbm_batch* b, int bsA, int bsB) {
  if (!b) return fail(LBM_EINVAL, "null batch");
  auto ok = [](int x) { return x == 128 || x == 256 || x == 512; };
  if (!ok(bsA) || !ok(bsB)) return fail(LBM_EINVAL, "block size must be 128, 256 or 512");
  b->bsA = bsA;
  b->bsB = bsB;
  b->dirty = true;
  return LBM_OK;
}

template <typename T, typename S>
static int batch_upload(lbm_batch* b, hipStream_t st) {
  const int n = (int)b->cs.size();
  std::vector<BatchCase<T, S>> h((size_t)std::max(n, 1));
  std::vector<uint32_t> mA, mB, mP, mW;
  for (int c = 0; c < n; ++c) {
    BatchCase<T, S>& d = h[c];
    std::memset(&d, 0, sizeof(d));
    lbm_solver* s = b->cs[c];
    if (!s) { d.k = 0; d.k_end = 0; continue; }   // free slot: never runs
    d.g = s->g;
    d.ph = s->dim == 3 ? make_phys<T, 3>(s) : make_phys<T, 2>(s);
    for (int i = 0; i < 2; ++i) {
      d.F[i] = (S*)s->Fb[i];
      d.G[i] = (S*)s->Gb[i];
      d.phi[i] = (T*)s->phi[i];
    }
    d.u = (T*)s->u;
    d.eps = (T*)s->eps;
    d.ampf = (T*)s->ampf;
    d.meta = s->meta;
    d.wmask = s->wmask;
    d.normal = s->normal;
    d.partials = s->partials;
    d.ctrl = s->ctrl;
    d.nbA = (unsigned int)((s->N + b->bsA - 1) / b->bsA);
    d.nbB = (unsigned int)((s->N + b->bsB - 1) / b->bsB);
    d.nbA8 = (d.nbA + kXCD - 1) / kXCD * kXCD;
    d.nbB8 = (d.nbB + kXCD - 1) / kXCD * kXCD;
    d.chunks = s->chunks;
    d.n_plain = s->n_plain;
    d.n_wall = s->n_wall;
    const unsigned int wpb = (unsigned int)(b->bsB / 64);
    d.nbP8 = ((s->n_plain + wpb - 1) / wpb + kXCD - 1) / kXCD * kXCD;
    d.nbW8 = ((s->n_wall + wpb - 1) / wpb + kXCD - 1) / kXCD * kXCD;
    d.k = s->k;
    d.k_end = b->k_end[c];
    d.local = s->local ? 1 : 0;
    d.phiu_valid = s->phiu_valid ? 1 : 0;
    if (s->k >= b->k_end[c]) continue;   // finished: no blocks
    const uint32_t baseA = (uint32_t)(mA.size() / 2), baseB = (uint32_t)(mB.size() / 2);
    for (unsigned int i = 0; i < d.nbA8; ++i) { mA.push_back((uint32_t)c); mA.push_back(baseA); }
    for (unsigned int i = 0; i < d.nbB8; ++i) { mB.push_back((uint32_t)c); mB.push_back(baseB); }
    const uint32_t baseP = (uint32_t)(mP.size() / 2), baseW = (uint32_t)(mW.size() / 2);
    if (s->n_plain)
      for (unsigned int i = 0; i < d.nbP8; ++i) { mP.push_back((uint32_t)c); mP.push_back(baseP); }
    if (s->n_wall)
      for (unsigned int i = 0; i < d.nbW8; ++i) { mW.push_back((uint32_t)c); mW.push_back(baseW); }
  }
  // split the interleaved (case, base) pairs into two arrays per pass
  auto upload_map = [&](std::vector<uint32_t>& m, uint32_t** dptr, size_t* cap,
                        unsigned int* nblocks) -> hipError_t {
    const size_t nb = m.size() / 2;
    std::vector<uint32_t> flat(2 * std::max<size_t>(nb, 1));
    for (size_t i = 0; i < nb; ++i) {
      flat[i] = m[2 * i];
      flat[nb + i] = m[2 * i + 1];
    }
    *nblocks = (unsigned int)nb;
    if (flat.size() > *cap) {
      if (*dptr) { hipError_t e = hipStreamSynchronize(st); if (e) return e; (void)hipFree(*dptr); }
      hipError_t e = hipMalloc((void**)dptr, flat.size() * 4);
      if (e) return e;
      *cap = flat.size();
    }
    return hipMemcpyAsync(*dptr, flat.data(), flat.size() * 4, hipMemcpyHostToDevice, st);
  };
  HIPCHK(upload_map(mA, &b->mapA, &b->mapA_cap, &b->nA));
  HIPCHK(upload_map(mB, &b->mapB, &b->mapB_cap, &b->nB));
  HIPCHK(upload_map(mP, &b->mapP, &b->mapP_cap, &b->nP));
  HIPCHK(upload_map(mW, &b->mapW, &b->mapW_cap, &b->nW));
  const size_t bytes = h.size() * sizeof(BatchCase<T, S>);
  if (bytes > b->dcases_cap) {
    if (b->dcases) { HIPCHK(hipStreamSynchronize(st)); (void)hipFree(b->dcases); }
    HIPCHK(hipMalloc(&b->dcases, bytes));
    b->dcases_cap = bytes;
  }
  HIPCHK(hipMemcpyAsync(b->dcases, h.data(), bytes, hipMemcpyHostToDevice, st));
  // pageable source: the copy has been staged when hipMemcpyAsync returns
  b->ncase = n;
  b->dirty = false;
  return LBM_OK;
}

template <typename T, typename S, int DIM>
static void batch_launch(lbm_batch* b, hipStream_t st) {
  const BatchCase<T, S>* dc = (const BatchCase<T, S>*)b->dcases;
  const unsigned int nA = b->nA, nB = b->nB;
  BlockMap ma{b->mapA, b->mapA + nA}, mb{b->mapB, b->mapB + nB};
  if (nA) {
#define LA(B) k_phi_batch<T, S, DIM, B><<<nA, B, 0, st>>>(dc, ma)
    LBM_BS_SWITCH(b->bsA, LA)
#undef LA
  }
  if (!b->split) {
    if (nB) {
#define LB(B) k_collide_batch<T, S, DIM, B><<<nB, B, 0, st>>>(dc, mb)
      LBM_BS_SWITCH(b->bsB, LB)
#undef LB
    }
  } else {
    // plain-fluid chunks of every case, then the other chunks
    const unsigned int nP = b->nP, nW = b->nW;
    BlockMap mp{b->mapP, b->mapP + nP}, mw{b->mapW, b->mapW + nW};
    if (nP) {
#define LP(B) k_collide_batch_chunks<T, S, DIM, B, true><<<nP, B, 0, st>>>(dc, mp)
      LBM_BS_SWITCH(b->bsB, LP)
#undef LP
    }
    if (nW) {
#define LW(B) k_collide_batch_chunks<T, S, DIM, B, false><<<nW, B, 0, st>>>(dc, mw)
      LBM_BS_SWITCH(b->bsB, LW)
#undef LW
    }
  }
  k_advance_batch<T, S><<<(b->ncase + 63) / 64, 64, 0, st>>>((BatchCase<T, S>*)b->dcases,
                                                           b->ncase);
}

// host mirror of k_advance_batch
static void batch_host_advance(lbm_batch* b) {
  for (size_t c = 0; c < b->cs.size(); ++c) {
    lbm_solver* s = b->cs[c];
    if (!s || s->k >= b->k_end[c]) continue;
    s->prev_local = s->local;
    s->local = false;
    s->cur ^= 1;
    s->k += 1;
    s->phiu_valid = true;
    s->kept = false;
  }
}

static int batch_dim(const lbm_batch* b) {
  for (lbm_solver* s : b->cs)
    if (s) return s->dim;
  return 3;
}

#define BATCH_DISPATCH(b, CALL)                                                        \
  ((b)->precision == LBM_F32                                                           \
       ? (batch_dim(b) == 3 ? CALL(float, float, 3) : CALL(float, float, 2))           \
   : (b)->precision == LBM_BF16                                                        \
       ? (batch_dim(b) == 3 ? CALL(float, __hip_bfloat16, 3)                           \
                            : CALL(float, __hip_bfloat16, 2))                          \
       : (batch_dim(b) == 3 ? CALL(float, __half, 3) : CALL(float, __half, 2)))

template <typename T, typename S, int DIM>
static int batch_step_impl(lbm_batch* b, int nsteps, hipStream_t st) {
  for (int it = 0; it < nsteps; ++it) {
    if (b->dirty) {
      int rc = batch_upload<T, S>(b, st);
      if (rc) return rc;
    }
    // a case that reaches its target drops out of the block map next step
    bool finishing = false;
    for (size_t c = 0; c < b->cs.size(); ++c)
      if (b->cs[c] && b->cs[c]->k + 1 == b->k_end[c]) finishing = true;
    batch_launch<T, S, DIM>(b, st);
    batch_host_advance(b);
    if (finishing) b->dirty = true;
  }
  HIPCHK(hipGetLastError());
  return LBM_OK;
}

int lbm_batch_step(lbm_batch* b, int nsteps, void* stream) {
  if (!b) return fail(LBM_EINVAL, "null batch");
  if (nsteps < 0) return fail(LBM_EINVAL, "negative step count");
  for (lbm_solver* s : b->cs)
    if (s && s->dim != batch_dim(b)) return fail(LBM_EINVAL, "a batch holds 2D or 3D cases, not both");
#define CALL(T, S, D) batch_step_impl<T, S, D>(b, nsteps, (hipStream_t)stream)
  return BATCH_DISPATCH(b, CALL);
#undef CALL
}

int lbm_batch_graph_capture(lbm_batch* b, int nsteps, void* stream) {
  if (!b) return fail(LBM_EINVAL, "null batch");
  if (nsteps < 2 || (nsteps & 1)) return fail(LBM_EINVAL, "nsteps must be even and >= 2");
  hipStream_t st0 = (hipStream_t)stream;
  // every active case must have nsteps left and have left its init form, so
  // each replay is identical
  for (size_t c = 0; c < b->cs.size(); ++c) {
    lbm_solver* s = b->cs[c];
    if (!s || s->k >= b->k_end[c]) continue;
    if (s->local || b->k_end[c] - s->k < nsteps)
      return fail(LBM_ESTATE, "graph capture needs one eager step and nsteps left per case");
  }
  if (b->dirty) {
#define CALL(T, S, D) batch_upload<T, S>(b, st0)
    int rc = BATCH_DISPATCH(b, CALL);
#undef CALL
    if (rc) return rc;
  }
  HIPCHK(hipStreamSynchronize(st0));
  batch_drop_graph(b);
  hipStream_t cs;
  HIPCHK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    for (int it = 0; it < nsteps; ++it) {
#define CALL(T, S, D) (batch_launch<T, S, D>(b, cs), 0)
      (void)BATCH_DISPATCH(b, CALL);
#undef CALL
    }
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(cs, &g);
    if (e == hipSuccess) {
      b->graph = g;
      e = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
    }
  }
  (void)hipStreamDestroy(cs);
  HIPCHK(e);
  b->gsteps = nsteps;
  return LBM_OK;
}

int lbm_batch_graph_launch(lbm_batch* b, void* stream) {
  if (!b || !b->gexec) return fail(LBM_ESTATE, "no captured batch graph");
  if (b->dirty) return fail(LBM_ESTATE, "case set changed since capture");
  for (size_t c = 0; c < b->cs.size(); ++c) {
    lbm_solver* s = b->cs[c];
    if (s && s->k < b->k_end[c] && b->k_end[c] - s->k < b->gsteps)
      return fail(LBM_ESTATE, "a case has fewer steps left than the captured graph");
  }
  HIPCHK(hipGraphLaunch(b->gexec, (hipStream_t)stream));
  for (int it = 0; it < b->gsteps; ++it) batch_host_advance(b);
  return LBM_OK;
}

int lbm_batch_autotune(lbm_batch* b, int nsteps_each, void* stream, int* bsA, int* bsB) {
  if (!b) return fail(LBM_EINVAL, "null batch");
  if (nsteps_each < 1) nsteps_each = 1;
  hipStream_t st = (hipStream_t)stream;
  const int cand[3] = {128, 256, 512};
  float best = 1e30f;
  int bA = 256, bB = 256;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  int rc = LBM_OK;
  for (int i = 0; i < 3 && rc == LBM_OK; ++i)
    for (int j = 0; j < 3 && rc == LBM_OK; ++j) {
      rc = lbm_batch_set_block_sizes(b, cand[i], cand[j]);
      if (rc) break;
      rc = lbm_batch_step(b, 1, stream);   // upload + first launch of this shape
      if (rc) break;
      (void)hipEventRecord(e0, st);
      rc = lbm_batch_step(b, nsteps_each, stream);
      (void)hipEventRecord(e1, st);
      (void)hipEventSynchronize(e1);
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      if (ms < best) { best = ms; bA = cand[i]; bB = cand[j]; }
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  rc = lbm_batch_set_block_sizes(b, bA, bB);
  if (bsA) *bsA = bA;
  if (bsB) *bsB = bB;
  return rc;
}
