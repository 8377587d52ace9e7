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
#include "lbm_impl.hpp"

#include <mutex>

thread_local std::string g_err;

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
static int destroy_impl(lbm_solver* s);

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
  ALLOC(s->u[0], D * N * es);
  ALLOC(s->u[1], D * N * es);
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
    destroy_impl(s);
    return fail(e == hipErrorOutOfMemory ? LBM_ENOMEM : LBM_EHIP,
                std::string("allocation failed: ") + hipGetErrorString(e));
  }
  *out = s;
  return LBM_OK;
}

static int destroy_impl(lbm_solver* s) {
  if (!s) return LBM_OK;
  drop_graph(s);
  void* ptrs[] = {s->Fb[0], s->Fb[1], s->Gb[0], s->Gb[1], s->u[0], s->u[1], s->phi[0], s->phi[1],
                  s->eps, s->ampf, s->meta, s->bnd, s->dist[0], s->dist[1], s->solid,
                  s->wmask, s->normal, s->surf, s->partials, s->mparts, s->dout, s->ctrl,
                  s->chunks, s->chunk_flag, s->keep, s->nsw, s->mu,
                  s->mixed[0], s->mixed[1], s->mixed[2], s->fpart};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  delete s;
  return LBM_OK;
}

int lbm_destroy(lbm_solver* s) {
  if (!s) return LBM_OK;
  REFUSE_BATCHED(s);
  return destroy_impl(s);
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
  // pass-B chunk lists: plain-fluid chunks, then the others, all-solid ones
  // tagged (kChunkSolid) (set-up only: one read-back of a byte per 64 nodes)
  const size_t nch = (size_t)((s->N + 63) / 64);
  k_chunk_kind<BS><<<nblk((long long)nch * 64), BS, 0, st>>>(s->N, s->meta, s->chunk_flag);
  HIPCHK(hipGetLastError());
  std::vector<uint8_t> flag(nch);
  HIPCHK(hipMemcpyAsync(flag.data(), s->chunk_flag, nch, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  std::vector<uint32_t> list;
  list.reserve(nch);
  for (size_t c = 0; c < nch; ++c)
    if (flag[c] == 1) list.push_back((uint32_t)c);
  s->n_plain = (unsigned int)list.size();
  for (size_t c = 0; c < nch; ++c)
    if (flag[c] != 1) list.push_back((uint32_t)c | (flag[c] == 2 ? kChunkSolid : 0u));
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

// Geometry precompute from s->solid and a fraction array, both already in the
// internal node order.
static int set_geometry_internal(lbm_solver* s, const void* frac_dev, int frac_dtype,
                                 hipStream_t st) {
  // the chunk lists a captured graph reads are rebuilt (and reallocated)
  drop_graph(s);
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

int lbm_set_geometry(lbm_solver* s, const uint8_t* solid_dev, const void* frac_dev,
                     int frac_dtype, void* stream) {
  if (!s || !solid_dev) return fail(LBM_EINVAL, "null argument");
  REFUSE_BATCHED(s);
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
  REFUSE_BATCHED(s);
  hipStream_t st = (hipStream_t)stream;
  // a captured graph holds the field's pointer (or its absence)
  drop_graph(s);
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

static int init_fields_internal(lbm_solver* s, const void* phi_dev, const void* u_dev,
                                void* stream) {
  if (!s || !phi_dev) return fail(LBM_EINVAL, "null argument");
  // the buffers return to the post-stream form at index 0
  drop_graph(s);
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

int lbm_init_fields(lbm_solver* s, const void* phi_dev, const void* u_dev, void* stream) {
  if (!s) return fail(LBM_EINVAL, "null argument");
  REFUSE_BATCHED(s);
  return init_fields_internal(s, phi_dev, u_dev, stream);
}

int lbm_step(lbm_solver* s, int nsteps, void* stream) {
  if (!s) return fail(LBM_EINVAL, "null solver");
  REFUSE_BATCHED(s);
  if (!s->inited) return fail(LBM_ESTATE, "step() before init_fields()");
  if (nsteps < 0) return fail(LBM_EINVAL, "negative step count");
#define CALL(T, S, D) step_impl<T, S, D>(s, nsteps, (hipStream_t)stream)
  return DISPATCH3(s, CALL);
#undef CALL
}

int lbm_graph_capture(lbm_solver* s, int nsteps) {
  if (!s) return fail(LBM_EINVAL, "null solver");
  REFUSE_BATCHED(s);
  if (!s->inited || s->k < 1 || s->local)
    return fail(LBM_ESTATE, "graph capture needs at least one eager step after init");
  if (nsteps < 2 || (nsteps & 1)) return fail(LBM_EINVAL, "nsteps must be even and >= 2");
  drop_graph(s);
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
  if (e != hipSuccess) {
    drop_graph(s);
    return fail(LBM_EHIP, std::string("graph capture: ") + hipGetErrorString(e));
  }
  s->gsteps = nsteps;
  s->gcur = cur0;
  s->gparity = (int)(k0 & 1);
  return LBM_OK;
}

int lbm_graph_launch(lbm_solver* s, void* stream) {
  if (!s) return fail(LBM_EINVAL, "null solver");
  REFUSE_BATCHED(s);
  if (!s->gexec) return fail(LBM_ESTATE, "no captured graph");
  // the graph reads the buffer set and mass-factor slot of its capture: an odd
  // number of eager steps (or an injection) since then moves both
  if (s->local || !s->phiu_valid || s->cur != s->gcur || (int)(s->k & 1) != s->gparity)
    return fail(LBM_ESTATE, "state changed since capture (capture the graph again)");
  HIPCHK(hipGraphLaunch(s->gexec, (hipStream_t)stream));
  s->k += s->gsteps;        // even: buffer parity unchanged
  s->prev_local = false;
  return LBM_OK;
}

int lbm_profile_steps(lbm_solver* s, int nsteps, void* stream, double* out) {
  if (!s || !out) return fail(LBM_EINVAL, "null argument");
  REFUSE_BATCHED(s);
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

// Exhaustive check of the kernels' division forms against IEEE division for
// one fp32 divisor: for every one of the 2^32 fp32 bit patterns x (NaN inputs
// excepted) divk_any -- and divk, unless d is an even integer -- must give the
// bit-identical quotient.
__global__ void k_check_div(float d, double y, int plain, unsigned long long* bad) {
  const unsigned long long t = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned int cnt = 0;
  for (unsigned int j = 0; j < 64; ++j) {
    const unsigned int bits = (unsigned int)(t * 64ull + j);
    const float x = __builtin_bit_cast(float, bits);
    if (x != x) continue;
    const unsigned int b = __builtin_bit_cast(unsigned int, x / d);
    if (__builtin_bit_cast(unsigned int, divk_any(x, d, y)) != b) ++cnt;
    if (plain && __builtin_bit_cast(unsigned int, divk(x, d, y)) != b) ++cnt;
  }
  if (cnt) atomicAdd(bad, (unsigned long long)cnt);
}

int lbm_check_division(float d, unsigned long long* mismatches, void* stream) {
  if (!mismatches) return fail(LBM_EINVAL, "null argument");
  if (!(d > 0.0f) || !std::isfinite(d)) return fail(LBM_EINVAL, "divisor must be positive and finite");
  hipStream_t st = (hipStream_t)stream;
  unsigned long long* dev = nullptr;
  HIPCHK(hipMalloc(&dev, sizeof(unsigned long long)));
  hipError_t e = hipMemsetAsync(dev, 0, sizeof(unsigned long long), st);
  const double y = 1.0 / (double)d;
  // divk is claimed exact for every divisor that is not an even integer
  const int plain = !(d >= 2.0f && std::floor(d * 0.5f) == d * 0.5f);
  if (e == hipSuccess) {
    // 2^32 values, 64 per thread, 256 threads per block -> 2^18 blocks
    k_check_div<<<1u << 18, 256, 0, st>>>(d, y, plain, dev);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(mismatches, dev, sizeof(*mismatches), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipFree(dev);
  HIPCHK(e);
  return LBM_OK;
}

// Streaming kernels for lbm_hbm_stream: 16 bytes per lane, grid-stride,
// default or non-temporal accesses.
typedef float f4v __attribute__((ext_vector_type(4)));
template <int NT>
__global__ __launch_bounds__(256) void k_stream_read(const f4v* __restrict__ a, size_t n,
                                                     float* __restrict__ sink) {
  float acc = 0.f;
  for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += (size_t)gridDim.x * 256ull) {
    const f4v v = NT ? __builtin_nontemporal_load(a + i) : a[i];
    acc += (v.x + v.y) + (v.z + v.w);
  }
  if (acc == -1.0f) sink[0] = acc;   // never true for the zero-filled source; keeps the loads
}
template <int NT>
__global__ __launch_bounds__(256) void k_stream_write(f4v* __restrict__ a, size_t n) {
  for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += (size_t)gridDim.x * 256ull) {
    const f4v v = {(float)i, 1.f, 2.f, 3.f};
    if (NT) __builtin_nontemporal_store(v, a + i);
    else a[i] = v;
  }
}
template <int NT>
__global__ __launch_bounds__(256) void k_stream_copy(const f4v* __restrict__ a,
                                                     f4v* __restrict__ b, size_t n) {
  for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += (size_t)gridDim.x * 256ull) {
    const f4v v = NT ? __builtin_nontemporal_load(a + i) : a[i];
    if (NT) __builtin_nontemporal_store(v, b + i);
    else b[i] = v;
  }
}

int lbm_hbm_stream(size_t bytes, int reps, void* stream, double* out6) {
  if (!out6) return fail(LBM_EINVAL, "null argument");
  if (bytes < (1u << 20) || reps < 1) return fail(LBM_EINVAL, "bytes >= 1 MiB and reps >= 1");
  bytes &= ~(size_t)15;
  const size_t n = bytes / 16;
  hipStream_t st = (hipStream_t)stream;
  void *a = nullptr, *b = nullptr, *sink = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = LBM_OK;
  auto chk = [&](hipError_t e, const char* what) {
    if (rc == LBM_OK && e != hipSuccess) rc = fail(LBM_EHIP, std::string(what) + ": " + hipGetErrorString(e));
  };
  chk(hipMalloc(&a, bytes), "hipMalloc");
  chk(hipMalloc(&b, bytes), "hipMalloc");
  chk(hipMalloc(&sink, 64), "hipMalloc");
  chk(hipEventCreate(&e0), "hipEventCreate");
  chk(hipEventCreate(&e1), "hipEventCreate");
  if (rc == LBM_OK) chk(hipMemsetAsync(a, 0, bytes, st), "hipMemsetAsync");
  if (rc == LBM_OK) chk(hipMemsetAsync(b, 0, bytes, st), "hipMemsetAsync");
  const unsigned int grids[3] = {1024u, 4096u, 16384u};
  for (int kind = 0; kind < 6 && rc == LBM_OK; ++kind) {
    double best = 0.0;
    for (unsigned int G : grids) {
      auto launch = [&]() {
        switch (kind) {
          case 0: k_stream_read<0><<<G, 256, 0, st>>>((const f4v*)a, n, (float*)sink); break;
          case 1: k_stream_read<1><<<G, 256, 0, st>>>((const f4v*)a, n, (float*)sink); break;
          case 2: k_stream_write<0><<<G, 256, 0, st>>>((f4v*)b, n); break;
          case 3: k_stream_write<1><<<G, 256, 0, st>>>((f4v*)b, n); break;
          case 4: k_stream_copy<0><<<G, 256, 0, st>>>((const f4v*)a, (f4v*)b, n); break;
          default: k_stream_copy<1><<<G, 256, 0, st>>>((const f4v*)a, (f4v*)b, n); break;
        }
      };
      launch();
      chk(hipEventRecord(e0, st), "hipEventRecord");
      for (int r = 0; r < reps; ++r) launch();
      chk(hipEventRecord(e1, st), "hipEventRecord");
      chk(hipEventSynchronize(e1), "hipEventSynchronize");
      float ms = 0.f;
      chk(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
      if (rc != LBM_OK) break;
      const double moved = (kind >= 4 ? 2.0 : 1.0) * (double)(n * 16) * reps;
      const double tbs = moved / (ms * 1e-3) / 1e12;
      if (tbs > best) best = tbs;
    }
    out6[kind] = best;
  }
  if (rc == LBM_OK) chk(hipGetLastError(), "stream kernel launch");
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  for (void* q : {a, b, sink})
    if (q) (void)hipFree(q);
  return rc;
}

int lbm_set_block_sizes(lbm_solver* s, int bs_phi, int bs_collide) {
  if (!s) return fail(LBM_EINVAL, "null solver");
  REFUSE_BATCHED(s);
  auto ok = [](int b) { return b == 128 || b == 256 || b == 512; };
  if (!ok(bs_phi) || !ok(bs_collide)) return fail(LBM_EINVAL, "block size must be 128, 256 or 512");
  // the fused form's plan (and its partials buffer) first: a tile that does
  // not fit leaves every setting as it was
  const bool fused = s->split == 3 && fused_ok(s);
  int pl[6] = {0, 0, 0, 0, 0, 0};
  unsigned int nbF = 0;
  if (fused) {
    if (!fused_plan(s->g, bs_collide, s->freq_ty, s->freq_tx, s->freq_zlen, pl))
      return fail(LBM_EINVAL, "fused tile does not fit the pass-B block size "
                              "(ty * tx <= block and (ty + 2) * (tx + 2) <= 2 * block)");
    nbF = (unsigned int)(pl[2] * pl[3] * pl[5]);
    if (nbF > s->fpart_cap) {
      double* np = nullptr;
      HIPCHK(hipMalloc(&np, (size_t)nbF * sizeof(double)));
      if (s->fpart) (void)hipFree(s->fpart);
      s->fpart = np;
      s->fpart_cap = nbF;
    }
  }
  drop_graph(s);
  s->bsA = bs_phi;
  s->bsB = bs_collide;
  s->nbA = (unsigned int)((s->N + bs_phi - 1) / bs_phi);
  s->nbB = (unsigned int)((s->N + bs_collide - 1) / bs_collide);
  s->ldsA = lds_for_cap(s->bsA, s->capA, 0);
  s->ldsB = lds_for_cap(s->bsB, s->capB, 1024);   // finalize_mass's 128 doubles
  s->f_ty = pl[0]; s->f_tx = pl[1]; s->f_nty = pl[2]; s->f_ntx = pl[3];
  s->f_zlen = pl[4]; s->f_nseg = pl[5];
  s->f_blocks = nbF;
  s->ldsF = 0;
  if (fused) {
    // four ring planes; the occupancy cap pads the same dynamic allocation
    const unsigned int ring = 4u * (unsigned int)((pl[0] + 2) * (pl[1] + 2)) * (unsigned int)s->esz;
    const unsigned int tot = lds_for_cap(s->bsB, s->capB, 0);
    s->ldsF = ring > tot ? ring : tot;
  }
  return LBM_OK;
}

int lbm_set_fused_tile(lbm_solver* s, int ty, int tx, int zlen) {
  if (!s) return fail(LBM_EINVAL, "null solver");
  REFUSE_BATCHED(s);
  if (ty < 0 || tx < 0 || zlen < 0) return fail(LBM_EINVAL, "tile sizes must be >= 0 (0: automatic)");
  const int o[3] = {s->freq_ty, s->freq_tx, s->freq_zlen};
  s->freq_ty = ty;
  s->freq_tx = tx;
  s->freq_zlen = zlen;
  const int rc = lbm_set_block_sizes(s, s->bsA, s->bsB);   // changes nothing on failure
  if (rc) {
    s->freq_ty = o[0]; s->freq_tx = o[1]; s->freq_zlen = o[2];
  }
  return rc;
}

int lbm_get_fused_tile(const lbm_solver* s, int* out6) {
  if (!s || !out6) return fail(LBM_EINVAL, "null argument");
  out6[0] = s->f_ty; out6[1] = s->f_tx; out6[2] = s->f_nty; out6[3] = s->f_ntx;
  out6[4] = s->f_zlen; out6[5] = s->f_nseg;
  return LBM_OK;
}

int lbm_set_occupancy(lbm_solver* s, int cap_phi, int cap_collide) {
  if (!s) return fail(LBM_EINVAL, "null solver");
  REFUSE_BATCHED(s);
  if (cap_phi < 0 || cap_phi > 8 || cap_collide < 0 || cap_collide > 8)
    return fail(LBM_EINVAL, "occupancy cap must be 0 (none) or 1..8 waves per SIMD");
  const int oa = s->capA, ob = s->capB;
  s->capA = cap_phi;
  s->capB = cap_collide;
  const int rc = lbm_set_block_sizes(s, s->bsA, s->bsB);
  if (rc) {
    s->capA = oa;
    s->capB = ob;
  }
  return rc;
}

int lbm_set_split(lbm_solver* s, int split) {
  if (!s) return fail(LBM_EINVAL, "null solver");
  REFUSE_BATCHED(s);
  if (split < 0 || split > 3) return fail(LBM_EINVAL, "split must be 0, 1, 2 or 3");
  const int old = s->split;
  s->split = split;
  const int rc = lbm_set_block_sizes(s, s->bsA, s->bsB);   // changes nothing on failure
  if (rc) s->split = old;
  return rc;
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
  REFUSE_BATCHED(s);
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

// ---------------------------------------------------------------------------
// The stencil operator without allocations (lbm_grad_lap_noalloc).  Its one
// scratch buffer -- phi reordered into the internal node order -- is a
// file-scope device array: every GPU holds its own copy for the life of the
// process, so the entry point calls no allocator.  Host code never uses the
// array's name as an address: its device address is looked up once per
// device with hipGetSymbolAddress and kept in g_scratch_dev.  Calls that
// overlap on one GPU share the array and race.
// ---------------------------------------------------------------------------
constexpr size_t kScratchBytes = size_t(32) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
static void* g_scratch_dev[kMaxDevices];
static std::mutex g_scratch_mu;

// Device address of this device's scratch array (looked up on first use).
static int scratch_on_current_device(void** out) {
  int dev = -1;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(LBM_EINVAL, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(LBM_EHIP, std::string("hipGetSymbolAddress(scratch) failed: ") +
                                hipGetErrorString(e));
    g_scratch_dev[dev] = p;
  }
  *out = g_scratch_dev[dev];
  return LBM_OK;
}

template <typename T, int DIM>
static int gradlap_noalloc_impl(lbm_solver* s, const void* phi, void* grad, void* lap,
                                hipStream_t st) {
  if ((size_t)s->N * sizeof(T) > kScratchBytes)
    return fail(LBM_EINVAL, "grid too large for lbm_grad_lap_noalloc: phi must fit its " +
                                std::to_string(kScratchBytes) + "-byte scratch");
  void* tmp = nullptr;
  int rc = scratch_on_current_device(&tmp);
  if (rc) return rc;
  const Phys<T> ph = make_phys<T, DIM>(s);
  rc = to_internal(s, phi, tmp, sizeof(T), 1, st);
  if (rc) return rc;
  k_gradlap<T, DIM, BS><<<s->nblocks, BS, 0, st>>>(s->g, ph, (const T*)tmp,
                                                   s->geom ? s->meta : nullptr, s->wmask,
                                                   s->normal, (T*)grad, (T*)lap);
  HIPCHK(hipGetLastError());
  return LBM_OK;
}

int lbm_grad_lap_noalloc(lbm_solver* s, const void* phi_dev, void* grad_dev, void* lap_dev,
                         void* stream) {
  if (!s || !phi_dev || !grad_dev || !lap_dev) return fail(LBM_EINVAL, "null argument");
#define CALL(T, D) gradlap_noalloc_impl<T, D>(s, phi_dev, grad_dev, lap_dev, (hipStream_t)stream)
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
  drop_graph(s);
  return LBM_OK;
}

// ---------------------------------------------------------------------------
// Solver state snapshot (preemption / checkpoint).  Layout of the byte image:
//   [0, 256)   StateHeader
//   then       F[cur], G[cur], G[cur^1] (AoSoA with the header's block size,
//              the step's distributions and
//              the previous g set the P observable is pulled from),
//              u[0], u[1] (by step parity: the last and the previous step's
//              velocity, which bf16 g slots are read against), phi[0],
//              phi[1], Ctrl, kept P, kept mu
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
constexpr uint32_t kStateVersion = 4;

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

constexpr int kStateArrays = 10;   // the last two only after an injection

struct StateLayout {
  size_t dist, u, phi, off[kStateArrays], total;
};

StateLayout state_layout(const lbm_solver* s) {
  StateLayout L;
  L.dist = (size_t)s->Q * (size_t)padded_nodes(s->N) * s->dsz;
  L.u = (size_t)s->dim * (size_t)s->N * s->esz;
  L.phi = (size_t)s->N * s->esz;
  const size_t sz[kStateArrays] = {L.dist, L.dist, L.dist, L.u, L.u, L.phi, L.phi,
                                   sizeof(Ctrl), L.phi, L.phi};
  size_t o = 256;
  for (int i = 0; i < kStateArrays; ++i) {
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
  const void* src[kStateArrays] = {s->Fb[s->cur], s->Gb[s->cur], s->Gb[s->cur ^ 1], s->u[0],
                                   s->u[1], s->phi[0], s->phi[1], s->ctrl, s->keep,
                                   (const char*)s->keep + L.phi};
  const size_t sz[kStateArrays] = {L.dist, L.dist, L.dist, L.u, L.u, L.phi, L.phi,
                                   sizeof(Ctrl), L.phi, L.phi};
  for (int i = 0; i < (s->kept ? kStateArrays : kStateArrays - 2); ++i)
    HIPCHK(hipMemcpyAsync(d + L.off[i], src[i], sz[i], hipMemcpyDefault, st));
  // the header comes from host memory that goes out of scope: synchronous copy
  HIPCHK(hipStreamSynchronize(st));
  HIPCHK(hipMemcpy(d, &h, sizeof(h), hipMemcpyDefault));
  return LBM_OK;
}

// Header of a state image in device or host memory (synchronises `st` first,
// so an image written by an earlier call on the stream is complete).
static int read_state_header(const void* src, hipStream_t st, StateHeader* h) {
  HIPCHK(hipStreamSynchronize(st));
  HIPCHK(hipMemcpy(h, src, sizeof(*h), hipMemcpyDefault));
  return LBM_OK;
}

static int load_state_internal(lbm_solver* s, const void* src, hipStream_t st) {
  if (!s || !src) return fail(LBM_EINVAL, "null argument");
  if (!s->geom) return fail(LBM_ESTATE, "set_geometry before load_state");
  StateHeader h{};
  int rc0 = read_state_header(src, st, &h);
  if (rc0) return rc0;
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
  void* dstp[kStateArrays] = {s->Fb[h.cur], s->Gb[h.cur], s->Gb[h.cur ^ 1], s->u[0], s->u[1],
                              s->phi[0], s->phi[1], s->ctrl, s->keep, (char*)s->keep + L.phi};
  const size_t sz[kStateArrays] = {L.dist, L.dist, L.dist, L.u, L.u, L.phi, L.phi,
                                   sizeof(Ctrl), L.phi, L.phi};
  for (int i = 0; i < (h.kept ? kStateArrays : kStateArrays - 2); ++i)
    HIPCHK(hipMemcpyAsync(dstp[i], d + L.off[i], sz[i], hipMemcpyDefault, st));
  s->k = h.k;
  s->cur = h.cur;
  s->local = h.local != 0;
  s->prev_local = h.prev_local != 0;
  s->phiu_valid = h.phiu_valid != 0;
  s->kept = h.kept != 0;
  s->inited = true;
  drop_graph(s);
  // the source may be host memory the caller reuses as soon as this returns
  HIPCHK(hipStreamSynchronize(st));
  return LBM_OK;
}

int lbm_load_state(lbm_solver* s, const void* src, void* stream) {
  if (!s) return fail(LBM_EINVAL, "null argument");
  REFUSE_BATCHED(s);
  return load_state_internal(s, src, (hipStream_t)stream);
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

template <typename T, int DIM>
static int base_impl(lbm_solver* s, double cx, double cy, double* out, hipStream_t st) {
  const Phys<T> ph = make_phys<T, DIM>(s);
  const unsigned int nb = (unsigned int)((s->N + (long long)BS * kMeasPer - 1) / ((long long)BS * kMeasPer));
  const int slot = (int)(s->k & 1), scaled = s->local ? 0 : 1;
  int* zmin = (int*)(s->dout + 14);
  k_measure<T, DIM, BS><<<nb, BS, 0, st>>>(s->g, ph, (const T*)s->phi[slot], s->meta, s->ctrl,
                                           slot, scaled, (MeasRec*)s->mparts);
  k_measure_final<DIM, BS><<<1, BS, 0, st>>>((const MeasRec*)s->mparts, nb, s->dout, zmin);
  const long long ncol = (long long)s->g.nx * s->g.ny;
  // the records are consumed: their buffer takes the per-block maxima
  k_base<T, DIM, BS><<<nblk(ncol), BS, 0, st>>>(s->g, ph, (const T*)s->phi[slot], s->meta,
                                                s->ctrl, slot, scaled, cx, cy, zmin,
                                                (double*)s->mparts);
  k_contact_final<BS><<<1, BS, 0, st>>>((const double*)s->mparts, nblk(ncol), s->dout + 12);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, s->dout + 12, sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return LBM_OK;
}

int lbm_base_radius(lbm_solver* s, double cx, double cy, double* out, void* stream) {
  if (!s || !out) return fail(LBM_EINVAL, "null argument");
  if (!s->inited) return fail(LBM_ESTATE, "fields not initialised");
  if (s->dim != 3) return fail(LBM_EINVAL, "base radius is defined for 3D grids");
#define CALL(T, D) base_impl<T, D>(s, cx, cy, out, (hipStream_t)stream)
  return DISPATCH(s, CALL);
#undef CALL
}

int lbm_contact_radius(lbm_solver* s, double cx, double cy, double* out, void* stream) {
  if (!s || !out) return fail(LBM_EINVAL, "null argument");
  if (!s->inited) return fail(LBM_ESTATE, "fields not initialised");
  if (s->dim != 3) return fail(LBM_EINVAL, "contact radius is defined for 3D grids");
#define CALL(T, D) contact_impl<T, D>(s, cx, cy, out, (hipStream_t)stream)
  return DISPATCH(s, CALL);
#undef CALL
}

static size_t case_bytes(const lbm_params& p) {
  const size_t N = (size_t)p.nx * p.ny * p.nz;
  const int dim = p.nz == 1 ? 2 : 3, Q = dim == 3 ? 19 : 9;
  const size_t es = p.precision == LBM_F64 ? 8 : 4;
  const size_t ds = p.precision == LBM_F64 ? 8 : (p.precision == LBM_F32 ? 4 : 2);
  const size_t nb = (N + 255) / 256;
  return 4 * Q * (size_t)padded_nodes((long long)N) * ds + 2 * N * es + 2 * N + 2 * dim * N * es + N * es + 3 * N + 4 * N +
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

int lbm_batch_destroy(lbm_batch* b) {
  if (!b) return LBM_OK;
  (void)hipDeviceSynchronize();
  batch_drop_graph(b);
  for (lbm_solver* s : b->cs)
    if (s) destroy_impl(s);
  void* ptrs[] = {b->dcases, b->mapA, b->mapB, b->mapP, b->mapW, b->mapM, b->gen};
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
  return init_fields_internal(s, phi, u, st);
}

// Inputs of a case handed over by the caller (lbm_batch_add_fields), in the
// public order, as AllenCahnSolver.set_solid / init_fields take them.
struct CaseInputs {
  const uint8_t* solid;
  const void* frac;
  int frac_dtype;
  const void* phi;
  const void* u;
};

static int case_from_inputs(lbm_solver* s, const CaseInputs& in, hipStream_t st) {
  if (in.solid == s->solid) return fail(LBM_EINVAL, "solid must not alias solver memory");
  int rc = to_internal(s, in.solid, s->solid, 1, 1, st);
  if (rc) return rc;
  void* tmp = nullptr;
  const void* frac = nullptr;
  if (in.frac && in.frac_dtype != LBM_FRAC_NONE) {
    const size_t fes = in.frac_dtype == LBM_FRAC_F64 ? 8 : 4;
    HIPCHK(hipMalloc(&tmp, (size_t)s->N * fes));
    rc = to_internal(s, in.frac, tmp, fes, 1, st);
    frac = tmp;
  }
  if (!rc) rc = set_geometry_internal(s, frac, frac ? in.frac_dtype : LBM_FRAC_NONE, st);
  if (!rc) rc = init_fields_internal(s, in.phi, in.u, st);
  const hipError_t e = hipStreamSynchronize(st);
  if (tmp) (void)hipFree(tmp);
  if (!rc && e != hipSuccess) rc = fail(LBM_EHIP, "stream synchronize failed");
  return rc;
}

static int batch_add_impl(lbm_batch* b, const lbm_case_spec* sp, const CaseInputs* in,
                          void* stream, int* case_id) {
  if (!b || !sp || !case_id) return fail(LBM_EINVAL, "null argument");
  if (sp->p.precision != b->precision)
    return fail(LBM_EINVAL, "case precision differs from the batch precision");
  if (sp->p.precision == LBM_F64) return fail(LBM_EINVAL, "batch engine runs fp32 compute");
  if (sp->p.interface_mode != LBM_INTERFACE_AC || sp->p.mu_drive > 0.0)
    return fail(LBM_EINVAL, "batch engine runs the Allen-Cahn source without the mu drive");
  if (sp->steps < 0) return fail(LBM_EINVAL, "negative step target");
  if (!in && (sp->substrate < 0 || sp->substrate > 3)) return fail(LBM_EINVAL, "bad substrate");
  if (in && (!in->solid || !in->phi)) return fail(LBM_EINVAL, "null solid or phi");
  if (in && (in->frac_dtype < 0 || in->frac_dtype > 2 ||
             (in->frac_dtype != LBM_FRAC_NONE && !in->frac)))
    return fail(LBM_EINVAL, "bad fraction");
  const size_t need = case_bytes(sp->p);
  if (b->budget && b->used + need > b->budget)
    return fail(LBM_ENOMEM, "case does not fit the batch HBM budget");
  hipStream_t st = (hipStream_t)stream;
  lbm_solver* s = nullptr;
  int rc = create_impl(&sp->p, &s, true, st);
  if (rc) return rc;
  s->batched = true;
  if (in) rc = case_from_inputs(s, *in, st);
  else rc = s->dim == 3 ? gen_case<3>(b, s, *sp, st) : gen_case<2>(b, s, *sp, st);
  if (rc) { (void)hipStreamSynchronize(st); destroy_impl(s); return rc; }
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
  batch_drop_graph(b);
  *case_id = slot;
  return LBM_OK;
}

int lbm_batch_add(lbm_batch* b, const lbm_case_spec* sp, void* stream, int* case_id) {
  return batch_add_impl(b, sp, nullptr, stream, case_id);
}

int lbm_batch_add_fields(lbm_batch* b, const lbm_case_spec* sp, const uint8_t* solid_dev,
                         const void* frac_dev, int frac_dtype, const void* phi_dev,
                         const void* u_dev, void* stream, int* case_id) {
  const CaseInputs in{solid_dev, frac_dev, frac_dtype, phi_dev, u_dev};
  return batch_add_impl(b, sp, &in, stream, case_id);
}

int lbm_batch_remove(lbm_batch* b, int id) {
  if (!b || id < 0 || id >= (int)b->cs.size() || !b->cs[id]) return fail(LBM_EINVAL, "bad case id");
  (void)hipDeviceSynchronize();
  destroy_impl(b->cs[id]);
  b->cs[id] = nullptr;
  b->used -= b->bytes[id];
  b->bytes[id] = 0;
  b->dirty = true;
  batch_drop_graph(b);
  return LBM_OK;
}

int lbm_batch_solver(lbm_batch* b, int id, lbm_solver** s) {
  if (!b || !s || id < 0 || id >= (int)b->cs.size() || !b->cs[id]) return fail(LBM_EINVAL, "bad case id");
  *s = b->cs[id];
  return LBM_OK;
}

int lbm_batch_load_state(lbm_batch* b, int id, const void* src, void* stream) {
  if (!b || id < 0 || id >= (int)b->cs.size() || !b->cs[id]) return fail(LBM_EINVAL, "bad case id");
  if (!src) return fail(LBM_EINVAL, "null argument");
  lbm_solver* s = b->cs[id];
  hipStream_t st = (hipStream_t)stream;
  // batch kernels take the buffer set from the step parity (set = k & 1); an
  // image whose set differs (a stand-alone solver's after an injection) or
  // that lies past the case's target cannot resume here
  StateHeader h{};
  int rc = read_state_header(src, st, &h);
  if (rc) return rc;
  if (h.magic != kStateMagic) return fail(LBM_EINVAL, "not a solver state image");
  if (h.cur != (int)(h.k & 1))
    return fail(LBM_EINVAL, "state image's buffer set does not follow its step parity");
  if (h.k > b->k_end[id]) return fail(LBM_EINVAL, "state is past the case's step target");
  rc = load_state_internal(s, src, st);
  if (rc) return rc;
  // the per-case descriptors (step counter, buffer parity, form) are rebuilt
  b->dirty = true;
  batch_drop_graph(b);
  return LBM_OK;
}

int lbm_batch_status(lbm_batch* b, int id, long long* done, long long* target) {
  if (!b || id < 0 || id >= (int)b->cs.size() || !b->cs[id]) return fail(LBM_EINVAL, "bad case id");
  if (done) *done = b->cs[id]->k;
  if (target) *target = b->k_end[id];
  return LBM_OK;
}

size_t lbm_batch_bytes_used(const lbm_batch* b) { return b ? b->used : 0; }

int lbm_batch_set_occupancy(lbm_batch* b, int cap_phi, int cap_collide) {
  if (!b) return fail(LBM_EINVAL, "null batch");
  if (cap_phi < 0 || cap_phi > 8 || cap_collide < 0 || cap_collide > 8)
    return fail(LBM_EINVAL, "occupancy cap must be 0 (none) or 1..8 waves per SIMD");
  b->capA = cap_phi;
  b->capB = cap_collide;
  return lbm_batch_set_block_sizes(b, b->bsA, b->bsB);
}

int lbm_batch_get_launch(const lbm_batch* b, int* out5) {
  if (!b || !out5) return fail(LBM_EINVAL, "null argument");
  out5[0] = b->bsA; out5[1] = b->bsB; out5[2] = b->capA; out5[3] = b->capB;
  out5[4] = b->split;
  return LBM_OK;
}

int lbm_batch_set_split(lbm_batch* b, int split) {
  if (!b) return fail(LBM_EINVAL, "null batch");
  if (split < 0 || split > 2) return fail(LBM_EINVAL, "split must be 0, 1 or 2");
  b->split = split;
  b->dirty = true;
  batch_drop_graph(b);
  return LBM_OK;
}

int lbm_batch_set_block_sizes(lbm_batch* b, int bsA, int bsB) {
  if (!b) return fail(LBM_EINVAL, "null batch");
  auto ok = [](int x) { return x == 128 || x == 256 || x == 512; };
  if (!ok(bsA) || !ok(bsB)) return fail(LBM_EINVAL, "block size must be 128, 256 or 512");
  b->bsA = bsA;
  b->bsB = bsB;
  b->ldsA = lds_for_cap(bsA, b->capA, 0);
  b->ldsB = lds_for_cap(bsB, b->capB, 1024);   // finalize_mass's 128 doubles
  b->dirty = true;
  batch_drop_graph(b);
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
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  // shape: {block A, block B, cap A, cap B, form of pass B}
  int best[5] = {b->bsA, b->bsB, b->capA, b->capB, b->split};
  float best_ms = 1e30f;
  auto trial = [&](const int* c) -> int {
    b->split = c[4];
    b->capA = c[2];
    b->capB = c[3];
    int rc = lbm_batch_set_block_sizes(b, c[0], c[1]);
    if (!rc) rc = lbm_batch_step(b, 1, stream);   // upload + first launch of this shape
    if (rc) return rc;
    (void)hipEventRecord(e0, st);
    rc = lbm_batch_step(b, nsteps_each, stream);
    (void)hipEventRecord(e1, st);
    (void)hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (!rc && ms < best_ms) {
      best_ms = ms;
      for (int k = 0; k < 5; ++k) best[k] = c[k];
    }
    return rc;
  };
  int rc = LBM_OK;
  // stage 1: form of pass B and its occupancy cap at 256-thread blocks;
  // stage 2: the block sizes of both passes with the stage-1 winner
  const int caps[3] = {0, 5, 4}, bsz[3] = {128, 256, 512};
  for (int f = 0; f < 3 && !rc; ++f)
    for (int k = 0; k < 3 && !rc; ++k) {
      const int c[5] = {256, 256, 0, caps[k], f};
      rc = trial(c);
    }
  const int form = best[4], capb = best[3];
  for (int i = 0; i < 3 && !rc; ++i)
    for (int k = 0; k < 3 && !rc; ++k) {
      if (bsz[i] == 256 && bsz[k] == 256) continue;
      const int c[5] = {bsz[i], bsz[k], 0, capb, form};
      rc = trial(c);
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  b->split = best[4];
  b->capA = best[2];
  b->capB = best[3];
  rc = lbm_batch_set_block_sizes(b, best[0], best[1]);
  if (bsA) *bsA = best[0];
  if (bsB) *bsB = best[1];
  return rc;
}
