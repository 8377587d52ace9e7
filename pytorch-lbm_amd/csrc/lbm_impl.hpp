// Internal header of the library: the solver and batch objects, the launch
// sequences of the step and its neighbours (templates over compute type,
// storage type and dimension) and the helpers they share.  lbm_api.hip holds
// the C entry points; lbm_inst.hip instantiates the launch sequences of one
// storage type per object file, so the four precisions compile in parallel.
#pragma once
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

// message of the last failed call (one per thread; defined in lbm_api.hip)
extern thread_local std::string g_err;

inline int fail(int code, const std::string& msg) {
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
  void* u[2] = {nullptr, nullptr};   // velocity by step parity: after k steps in u[k & 1]
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
  uint8_t* chunk_flag = nullptr;     // kind of every chunk (k_chunk_kind)
  unsigned int n_plain = 0, n_wall = 0;
  // the same chunks as one list per pass-B block size (128, 256, 512) for the
  // single-launch form: per XCD range, plain chunks then wall chunks
  uint32_t* mixed[3] = {nullptr, nullptr, nullptr};
  unsigned int n_mixed[3] = {0, 0, 0};
  // pass B: 0 one launch over all nodes, 1 the two chunk-list launches,
  // 2 one launch over the per-XCD mixed list; 3 the fused step (one launch
  // for both passes; 3D with distributions stored in the compute type and no
  // grad-mu form, else form 2; the first step after init_fields always takes
  // the two chunk-list launches)
  int split = 1;
  // fused step: requested tile (0: automatic) and the plan in use
  int freq_ty = 0, freq_tx = 0, freq_zlen = 0;
  int f_ty = 0, f_tx = 0, f_nty = 0, f_ntx = 0, f_zlen = 0, f_nseg = 0;
  unsigned int f_blocks = 0, ldsF = 0;
  double* fpart = nullptr;    // one fp64 sum of phi per fused block
  unsigned int fpart_cap = 0;
  int capA = 0, capB = 0;     // occupancy caps, waves per SIMD (0: none)
  unsigned int ldsA = 0, ldsB = 0;   // dynamic LDS per block enforcing the caps
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph = nullptr;
  int gsteps = 0;
  int gcur = -1;              // buffer index and step parity the graph was captured at
  int gparity = -1;
  // droplet injection: pressure and chemical potential of the last step
  // (internal order) kept for the observables until the next step, then the
  // drop indicator; allocated at the first injection
  void* keep = nullptr;
  bool kept = false;
  bool batched = false;       // owned by a batch engine
};

// ---------------------------------------------------------------------------
inline void fast_div(uint32_t d, uint32_t* m, uint32_t* sh) {
  if (d <= 1) { *m = 0; *sh = 0; return; }
  int cl = 0;
  while ((1u << cl) < d) ++cl;
  const int p = 31 + cl;
  *m = (uint32_t)(((1ull << p) + d - 1) / d);
  *sh = (uint32_t)(p - 32);
}

// Internal node order: wall-normal (last) axis slowest -- 3D (z, x, y) with y
// fastest, 2D (y, x) with x fastest.
inline Grid make_grid(int nx, int ny, int nz) {
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

template <typename T> inline T TT(double v) { return (T)v; }

template <int DIM> inline float w32(int i) { return W32<DIM>(i); }

// Physics constants in the run precision (see lbm_phys.hpp).
template <typename T, int DIM> inline Phys<T> make_phys(const lbm_solver* s) {
  const lbm_params& p = s->p;
  constexpr int Q = Lat<DIM>::Q;
  Phys<T> ph;
  std::memset(&ph, 0, sizeof(ph));
  const double cs2 = 1.0 / 3.0;
  ph.cs2 = TT<T>(cs2);
  ph.two_cs2 = TT<T>(2.0 * cs2);
  // fp64 reciprocals of the divisors as the run holds them (the divk forms)
  ph.rcp_cs2 = 1.0 / (double)ph.cs2;
  ph.rcp_two_cs2 = 1.0 / (double)ph.two_cs2;
  for (int i = 0; i < Q; ++i) {
    const int c = wclass<DIM>(i);
    ph.wc[c] = (T)w32<DIM>(i);
    ph.woc[c] = ph.wc[c] / ph.cs2;                   // self._w / self.cs2
    ph.lc[c] = T(2) * ph.woc[c];                     // 2.0 * wi
    ph.lic[c] = T(2) * ph.wc[c];                     // 2.0 * self._w[i]
    ph.w3c[c] = TT<T>(3.0 * (double)w32<DIM>(i));
  }
  ph.four_beta = TT<T>(4.0 * p.beta);
  ph.kappa = TT<T>(p.kappa);
  ph.rho_l = TT<T>(p.rho_l);
  ph.rho_g = TT<T>(p.rho_g);
  ph.rho_floor = TT<T>(0.5 * p.rho_l);
  ph.drho = TT<T>(p.rho_l - p.rho_g);
  ph.visc = TT<T>((0.5 - p.tau_g) / p.tau_g);
  ph.xi = TT<T>(p.xi);
  ph.rcp_xi = 1.0 / (double)ph.xi;
  // rho_eff cs2 stays below max(rho_l, rho_g) cs2 (phi is clamped to [0, 1])
  ph.rc_small = std::max(p.rho_l, p.rho_g) * cs2 < 1.9;
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

inline unsigned int nblk(long long n) { return (unsigned int)((n + BS - 1) / BS); }

// Dynamic LDS per workgroup that caps residency at `cap` waves per SIMD for
// blocks of `bs` threads (4 SIMDs and 160 KiB of LDS per CU; `stat` bytes of
// static LDS per block).  The step kernels are memory bound, and more resident
// waves than the memory system can keep in flight only add contention, so the
// autotuner can trade occupancy for fewer concurrent streams.  0: no cap.
inline unsigned int lds_for_cap(int bs, int cap, unsigned int stat) {
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

// Can this solver take the fused step?
inline bool fused_ok(const lbm_solver* s) {
  return s->dim == 3 && s->esz == s->dsz && s->gmu == GMU_NONE;
}

// Tile and z segments of the fused step for blocks of bs threads.  ty, tx:
// tile extent along y and x (0: choose), zlen: planes per segment (0:
// choose).  The automatic tile minimises the collision lanes launched
// (ragged edge tiles leave lanes idle) plus the ring work (a ring node costs
// about 0.14 of a collision: pass A's share of the step's VALU work), with a
// 30% penalty on rows that are neither whole grid rows nor 64 nodes long
// (a wave's accesses then split into more, partly used cache lines: on the
// 150x150x107 grid 50x5 tiles ran 338 us, whole-row 150x3 tiles 255 us); the
// automatic segments give about four blocks per CU (256 CUs).  Returns false
// for a tile that does not fit the block (ty * tx > bs or a ring plane of
// more than 2 bs values).
inline bool fused_plan(const Grid& g, int bs, int ty, int tx, int zlen, int* out6) {
  auto ceil_div = [](int a, int b) { return (a + b - 1) / b; };
  if (ty <= 0 || tx <= 0) {
    double best = 1e300;
    int bty = 0, btx = 0;
    for (int cy = 1; cy <= (g.ny < bs ? g.ny : bs); ++cy) {
      int cx = bs / cy;
      if (cx > g.nx) cx = g.nx;
      if (cx < 1 || (cy + 2) * (cx + 2) > 2 * bs) continue;
      const double tiles = (double)ceil_div(g.ny, cy) * ceil_div(g.nx, cx);
      double cost = tiles * (bs + 0.14 * (cy + 2) * (cx + 2));
      if (cy < 64 && cy < g.ny) cost *= 1.3;
      if (cost < best) { best = cost; bty = cy; btx = cx; }
    }
    if (!bty) return false;
    ty = bty;
    tx = btx;
  }
  if (ty > g.ny) ty = g.ny;
  if (tx > g.nx) tx = g.nx;
  if (ty * tx > bs || (ty + 2) * (tx + 2) > 2 * bs) return false;
  const int nty = ceil_div(g.ny, ty), ntx = ceil_div(g.nx, tx);
  if (zlen <= 0) {
    const int target = 16 * 256 / (bs / 64);
    int nseg = (int)((double)target / (nty * ntx) + 0.5);
    if (nseg < 1) nseg = 1;
    if (nseg > g.nz) nseg = g.nz;
    zlen = ceil_div(g.nz, nseg);
  }
  if (zlen > g.nz) zlen = g.nz;
  out6[0] = ty; out6[1] = tx; out6[2] = nty; out6[3] = ntx; out6[4] = zlen;
  out6[5] = ceil_div(g.nz, zlen);
  return true;
}

// A captured graph holds the buffer pointers, chunk lists and amplification
// field of the moment of capture: every call that changes any of them drops it.
inline void drop_graph(lbm_solver* s) {
  // a replay may still be in flight on some stream: let it finish first
  if (s->gexec) {
    (void)hipDeviceSynchronize();
    (void)hipGraphExecDestroy(s->gexec);
    s->gexec = nullptr;
  }
  if (s->graph) { (void)hipGraphDestroy(s->graph); s->graph = nullptr; }
  s->gsteps = 0;
  s->gcur = -1;
  s->gparity = -1;
}

// Calls that would change a batched case behind its batch's back (the batch
// keeps a device descriptor of every case: step counter, buffers, fields).
#define REFUSE_BATCHED(s)                                                            \
  do {                                                                               \
    if ((s)->batched)                                                                \
      return fail(LBM_EINVAL, "this call is not available for a case of a batch "    \
                              "engine (use the lbm_batch_* entry points)");          \
  } while (0)

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

// Public-order array -> internal order (element size es bytes, comps arrays).
inline int to_internal(const lbm_solver* s, const void* pub, void* in, size_t es, int comps,
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

inline int to_public(const lbm_solver* s, const void* in, void* pub, size_t es, int comps,
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

template <typename T, typename S, int DIM>
int init_impl(lbm_solver* s, const void* phi, const void* u, hipStream_t st) {
  const Phys<T> ph = make_phys<T, DIM>(s);
  k_init<T, S, DIM, BS><<<s->nblocks, BS, 0, st>>>(s->g, ph, (const T*)phi, (const T*)u,
                                                   (T*)s->phi[0], (T*)s->u[0], (S*)s->Fb[0],
                                                   (S*)s->Gb[0], s->partials);
  k_sum_partials<BS><<<1, BS, 0, st>>>(s->partials, (unsigned int)((s->N + 63) / 64),
                                       &s->ctrl->M0);
  HIPCHK(hipGetLastError());
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
int step_impl(lbm_solver* s, int nsteps, hipStream_t st, hipEvent_t* ev = nullptr) {
  const Phys<T> ph = make_phys<T, DIM>(s);
  for (int it = 0; it < nsteps; ++it) {
    if (ev) HIPCHK(hipEventRecord(ev[3 * it + 0], st));
    const int src = s->cur, dst = s->cur ^ 1;
    const int pslot = (int)(s->k & 1), nslot = (int)((s->k + 1) & 1);
    const bool fused = s->split == 3 && !s->local && s->f_blocks > 0 && fused_ok(s);
    if (!fused) {
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
    }
    CollideArgs<T, S, DIM> a;
    a.g = s->g;
    a.Fin = (const S*)s->Fb[src];
    a.Gin = (const S*)s->Gb[src];
    a.Fout = (S*)s->Fb[dst];
    a.Gout = (S*)s->Gb[dst];
    a.u_in = (const T*)s->u[pslot];
    a.u_out = (T*)s->u[nslot];
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
    if (fused) {
      if constexpr (DIM == 3 && std::is_same<T, S>::value) {
        FusedTile ft;
        ft.ty = s->f_ty;
        ft.tx = s->f_tx;
        ft.nty = s->f_nty;
        ft.ntx = s->f_ntx;
        ft.zlen = s->f_zlen;
        ft.sy = s->f_ty + 2;
        ft.box = (s->f_ty + 2) * (s->f_tx + 2);
        ft.bpart = s->fpart;
#define LAUNCH_F(B) \
  k_fused<T, B><<<s->f_blocks, B, s->ldsF, st>>>(a, ph, ft, (T*)s->phi[nslot])
        LBM_BS_SWITCH(s->bsB, LAUNCH_F)
#undef LAUNCH_F
        if (ev) HIPCHK(hipEventRecord(ev[3 * it + 1], st));
        k_fused_mass<256><<<1, 256, 0, st>>>(s->fpart, s->f_blocks, s->ctrl, nslot);
      }
    } else if (s->gmu != GMU_NONE) {
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
    } else if (s->split >= 2) {                                                             \
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
int get_field_impl(lbm_solver* s, int which, void* dst, hipStream_t st,
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
      return to_public(s, s->u[s->k & 1], dst, s->esz, s->dim, st);
    case LBM_FIELD_MU:
      a.kind = OBS_MU;
      a.init_lap = s->k == 0 ? 1 : 0;
      break;
    case LBM_FIELD_F: a.kind = OBS_F; a.Fsrc = (const S*)s->Fb[s->cur]; break;
    case LBM_FIELD_G:
      a.kind = OBS_G;
      a.Fsrc = (const S*)s->Gb[s->cur];
      a.u_src = (const T*)s->u[s->k & 1];
      break;
    case LBM_FIELD_P:
      if (s->k == 0) {
        HIPCHK(hipMemsetAsync(dst, 0, N * s->esz, st));
        return LBM_OK;
      }
      a.kind = OBS_P;
      a.Fsrc = (const S*)s->Gb[s->cur ^ 1];
      a.u_src = (const T*)s->u[(s->k + 1) & 1];   // the velocity of step k - 1
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

// ---------------------------------------------------------------------------
// Mid-run droplet injection (lbm/fe_ac_solver.py:645-732).
// ---------------------------------------------------------------------------
template <typename T, typename S, int DIM>
int inject_impl(lbm_solver* s, const void* c_pub, const double* uimp, hipStream_t st) {
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
  // the new velocity goes to the spare buffer (neighbours still read the old
  // one for the bf16 g references) and is copied over after the pass
  a.u_in = (const T*)s->u[pslot];
  a.u_out = (T*)s->u[oslot];
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
  HIPCHK(hipMemcpyAsync(s->u[pslot], s->u[oslot], (size_t)s->dim * fb, hipMemcpyDeviceToDevice,
                        st));
  return LBM_OK;
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
  uint32_t* mapM = nullptr;           // pass B over the single-launch lists
  size_t mapA_cap = 0, mapB_cap = 0, mapP_cap = 0, mapW_cap = 0, mapM_cap = 0;
  unsigned int nA = 0, nB = 0, nP = 0, nW = 0, nM = 0;   // grid sizes
  // form of pass B: 0 one launch over all nodes, 1 the two chunk-list
  // launches, 2 one launch over the single-launch lists
  int split = 1;
  int capA = 0, capB = 0;             // occupancy caps (waves per SIMD, 0: none)
  unsigned int ldsA = 0, ldsB = 0;    // dynamic LDS per block enforcing them
  int ncase = 0;                      // descriptor count (== cs.size())
  // scratch for device-side case generation
  void* gen = nullptr;
  size_t gen_cap = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int gsteps = 0;
};

inline void batch_drop_graph(lbm_batch* b) {
  // a replay may still be in flight on some stream: let it finish first
  if (b->gexec) {
    (void)hipDeviceSynchronize();
    (void)hipGraphExecDestroy(b->gexec);
    b->gexec = nullptr;
  }
  if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
  b->gsteps = 0;
}

template <typename T, typename S>
int batch_upload(lbm_batch* b, hipStream_t st) {
  // a captured graph holds the old grid sizes, block maps and descriptor
  // table (which may be reallocated below): it cannot be replayed any more
  batch_drop_graph(b);
  const int n = (int)b->cs.size();
  std::vector<BatchCase<T, S>> h((size_t)std::max(n, 1));
  std::vector<uint32_t> mA, mB, mP, mW, mM;
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
    d.u[0] = (T*)s->u[0];
    d.u[1] = (T*)s->u[1];
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
    const int mk = b->bsB == 128 ? 0 : (b->bsB == 256 ? 1 : 2);
    d.mixed = s->mixed[mk];
    d.n_mixed = s->n_mixed[mk];
    d.nbM8 = (d.n_mixed / wpb + kXCD - 1) / kXCD * kXCD;
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
    const uint32_t baseM = (uint32_t)(mM.size() / 2);
    for (unsigned int i = 0; i < d.nbM8; ++i) { mM.push_back((uint32_t)c); mM.push_back(baseM); }
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
  HIPCHK(upload_map(mM, &b->mapM, &b->mapM_cap, &b->nM));
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
void batch_launch(lbm_batch* b, hipStream_t st) {
  const BatchCase<T, S>* dc = (const BatchCase<T, S>*)b->dcases;
  const unsigned int nA = b->nA, nB = b->nB;
  BlockMap ma{b->mapA, b->mapA + nA}, mb{b->mapB, b->mapB + nB};
  const unsigned int lA = b->ldsA, lB = b->ldsB;
  if (nA) {
#define LA(B) k_phi_batch<T, S, DIM, B><<<nA, B, lA, st>>>(dc, ma)
    LBM_BS_SWITCH(b->bsA, LA)
#undef LA
  }
  if (b->split == 0) {
    if (nB) {
#define LB(B) k_collide_batch<T, S, DIM, B><<<nB, B, lB, st>>>(dc, mb)
      LBM_BS_SWITCH(b->bsB, LB)
#undef LB
    }
  } else if (b->split == 2) {
    const unsigned int nM = b->nM;
    BlockMap mm{b->mapM, b->mapM + nM};
    if (nM) {
#define LM(B) k_collide_batch_mixed<T, S, DIM, B><<<nM, B, lB, st>>>(dc, mm)
      LBM_BS_SWITCH(b->bsB, LM)
#undef LM
    }
  } else {
    // plain-fluid chunks of every case, then the other chunks
    const unsigned int nP = b->nP, nW = b->nW;
    BlockMap mp{b->mapP, b->mapP + nP}, mw{b->mapW, b->mapW + nW};
    if (nP) {
#define LP(B) k_collide_batch_chunks<T, S, DIM, B, true><<<nP, B, lB, st>>>(dc, mp)
      LBM_BS_SWITCH(b->bsB, LP)
#undef LP
    }
    if (nW) {
#define LW(B) k_collide_batch_chunks<T, S, DIM, B, false><<<nW, B, lB, st>>>(dc, mw)
      LBM_BS_SWITCH(b->bsB, LW)
#undef LW
    }
  }
  k_advance_batch<T, S><<<(b->ncase + 63) / 64, 64, 0, st>>>((BatchCase<T, S>*)b->dcases,
                                                           b->ncase);
}

// host mirror of k_advance_batch
inline void batch_host_advance(lbm_batch* b) {
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

inline int batch_dim(const lbm_batch* b) {
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
int batch_step_impl(lbm_batch* b, int nsteps, hipStream_t st) {
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


// Launch sequences instantiated per storage type in lbm_inst.hip (compiled
// once per precision); every other translation unit only declares them.
#define LBM_SOLVER_TEMPLATES(X, T, S, D)                                                  \
  X template int init_impl<T, S, D>(lbm_solver*, const void*, const void*, hipStream_t);   \
  X template int step_impl<T, S, D>(lbm_solver*, int, hipStream_t, hipEvent_t*);           \
  X template int get_field_impl<T, S, D>(lbm_solver*, int, void*, hipStream_t, bool);      \
  X template int inject_impl<T, S, D>(lbm_solver*, const void*, const double*, hipStream_t);
#define LBM_BATCH_TEMPLATES(X, T, S)                                                      \
  X template int batch_upload<T, S>(lbm_batch*, hipStream_t);                              \
  X template void batch_launch<T, S, 2>(lbm_batch*, hipStream_t);                          \
  X template void batch_launch<T, S, 3>(lbm_batch*, hipStream_t);                          \
  X template int batch_step_impl<T, S, 2>(lbm_batch*, int, hipStream_t);                   \
  X template int batch_step_impl<T, S, 3>(lbm_batch*, int, hipStream_t);

#ifndef LBM_INST_TU
LBM_SOLVER_TEMPLATES(extern, double, double, 2)
LBM_SOLVER_TEMPLATES(extern, double, double, 3)
LBM_SOLVER_TEMPLATES(extern, float, float, 2)
LBM_SOLVER_TEMPLATES(extern, float, float, 3)
LBM_SOLVER_TEMPLATES(extern, float, __hip_bfloat16, 2)
LBM_SOLVER_TEMPLATES(extern, float, __hip_bfloat16, 3)
LBM_SOLVER_TEMPLATES(extern, float, __half, 2)
LBM_SOLVER_TEMPLATES(extern, float, __half, 3)
LBM_BATCH_TEMPLATES(extern, float, float)
LBM_BATCH_TEMPLATES(extern, float, __hip_bfloat16)
LBM_BATCH_TEMPLATES(extern, float, __half)
#endif
