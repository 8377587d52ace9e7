/*
 * lbm_amd.h -- C ABI of the MI355X Allen-Cahn phase-field LBM library.
 *
 * The reference (JunjunQing/pytorch-lbm) exposes this path as a Python class,
 * not as an FFI.  Each entry point below replaces one method (or one
 * observable attribute) of that class; the citation names the reference code
 * it stands in for.  All pointers named *_dev are device pointers; "stream"
 * is a hipStream_t passed as void* (NULL = the legacy default stream).  No
 * entry point synchronises the stream unless its comment says so, so a
 * sequence of lbm_step calls can be captured into a hipGraph.
 *
 * Precision: LBM_F32 and LBM_F64 compute and store in that type (the
 * reference's dtype argument).  LBM_BF16 / LBM_F16 store both distribution
 * sets in 16 bits and collide in fp32: f as the deviation from the
 * lattice-weight equilibrium w_i phi, and under LBM_BF16 g as the deviation
 * from its velocity-carrying part w_i (e_i . u) / cs2 (fp16 stores g as is).
 * Mass reductions are always fp64.
 *
 * Return value: 0 on success, a negative LBM_E* code otherwise;
 * lbm_last_error() gives a message for the calling thread.
 */
#ifndef LBM_AMD_H
#define LBM_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum lbm_precision { LBM_F32 = 0, LBM_F64 = 1, LBM_BF16 = 2, LBM_F16 = 3 };
enum lbm_stab { LBM_STAB_FAKHARI = 0, LBM_STAB_LIANG = 1, LBM_STAB_NONE = 2 };
enum lbm_frac_dtype { LBM_FRAC_NONE = 0, LBM_FRAC_F32 = 1, LBM_FRAC_F64 = 2 };
enum lbm_interface { LBM_INTERFACE_AC = 0, LBM_INTERFACE_CH = 1 };

enum lbm_status {
  LBM_OK = 0,
  LBM_EINVAL = -1,     /* bad argument / unsupported option */
  LBM_ESTATE = -2,     /* call out of order (e.g. step before init) */
  LBM_EHIP = -3,       /* HIP runtime error */
  LBM_ENOMEM = -4      /* device allocation failed */
};

/* Observable fields for lbm_get_field.  Layouts are C-contiguous with the
 * last grid axis fastest: scalar fields (nx,ny[,nz]); vectors (D,nx,ny[,nz]);
 * distributions (Q,nx,ny[,nz]).  Element type is the run dtype (fp32 for the
 * 16-bit storage forms), except the masks (uint8 0/1). */
enum lbm_field {
  LBM_FIELD_PHI = 0,        /* solver.phi   (mass-corrected, S12-relaxed)    */
  LBM_FIELD_RHO = 1,        /* solver.rho   (from the start-of-step phi)     */
  LBM_FIELD_U = 2,          /* solver.u                                      */
  LBM_FIELD_P = 3,          /* solver.P     (sum of g at start of step)      */
  LBM_FIELD_MU = 4,         /* solver.mu_phi                                 */
  LBM_FIELD_F = 5,          /* solver.f     (post-stream, mass-corrected)    */
  LBM_FIELD_G = 6,          /* solver.g                                      */
  LBM_FIELD_PHI_RAW = 7,    /* start-of-step phi (clamped, 0 at solid)       */
  LBM_FIELD_IS_BOUNDARY = 16, /* wetting.is_boundary (uint8)                 */
  LBM_FIELD_NEAR_WALL = 17,   /* wetting.near_wall_mask (uint8)              */
  LBM_FIELD_WALL_WEIGHT = 18, /* wetting.wall_weight (fp32 always)           */
  LBM_FIELD_WALL_NORMAL = 19, /* wetting.wall_normal (fp32 always, D comps)  */
  LBM_FIELD_PARTIAL = 20,     /* partial-solid mask (uint8)                  */
  LBM_FIELD_SOLID = 21        /* solid mask (uint8)                          */
};

/* Constructor arguments: FEConfig (lbm/fe_config.py:27-150, after its
 * __post_init__ derivations) plus the AllenCahnSolver keyword arguments
 * (lbm/fe_ac_solver.py:54-58). */
typedef struct lbm_params {
  int nx, ny, nz;            /* nz == 1 selects D2Q9, otherwise D3Q19 */
  int precision;             /* enum lbm_precision */
  double rho_l, rho_g, sigma, xi, beta, kappa;
  double tau_l, tau_g, tau_h;
  double theta_eq, phi_c;
  double g_force[3];
  int stab_mode;             /* enum lbm_stab */
  double boundary_relax;     /* default 0.3 */
  double sharpen_reduce;     /* default 0.8 */
  double mu_wall_scale;      /* default 1.0 */
  double ac_scale;           /* default 1.0 */
  double ghost_scale;        /* <= 0: automatic (3 if |theta-90|>60 else 2) */
  int geometric_wetting;     /* 0/1 */
  double geo_amplification;  /* scalar alpha_geo (default 1.0) */
  int direct_wetting;        /* 0/1: S12 relaxes towards the cubic wall-energy
                                phi (lbm/fe_ac_solver.py:609-613,
                                lbm/fe_wetting.py:313-416) */
  int interface_mode;        /* enum lbm_interface: Allen-Cahn sharpening
                                source, or the Cahn-Hilliard source
                                M (grad mu - grad phi) tau_f/(tau_f-0.5)
                                (lbm/fe_ac_solver.py:457-461) */
  double mu_drive;           /* > 0 (Allen-Cahn only): adds
                                mu_drive * M * grad mu to the source
                                (lbm/fe_ac_solver.py:484-486).  Either form
                                adds one pass per step (mu of every node). */
} lbm_params;

typedef struct lbm_solver lbm_solver;

/* Fill *p with the reference's defaults for an nx*ny*nz grid. */
void lbm_params_default(lbm_params* p, int nx, int ny, int nz);

/* AllenCahnSolver.__init__ (lbm/fe_ac_solver.py:54-191): allocates the
 * device state on the current HIP device. */
int lbm_create(const lbm_params* p, lbm_solver** out);

/* Releases everything lbm_create allocated. */
int lbm_destroy(lbm_solver* s);

/* AllenCahnSolver.set_solid (lbm/fe_ac_solver.py:193-237) including the
 * BounceBack masks (lbm/boundary.py:15-34) and the ConningtonLeeWetting
 * precompute (lbm/fe_wetting.py:122-311): boundary classification (periodic),
 * wall normals (np.gradient of the fraction, or neighbour counting when
 * frac_dev is NULL), near-wall distance/weight, per-direction wetting masks,
 * partial-solid set.  solid_dev: uint8 0/1 per node.  Bit-exact with the
 * reference for fp32 and fp64 fractions. */
int lbm_set_geometry(lbm_solver* s, const uint8_t* solid_dev, const void* frac_dev,
                     int frac_dtype, void* stream);

/* Optional per-node alpha_geo field (tensor geo_amplification,
 * lbm/fe_wetting.py:503-512); amp_dev in the run dtype (fp32 for 16-bit
 * storage), NULL restores the scalar. */
int lbm_set_amplification_field(lbm_solver* s, const void* amp_dev, void* stream);

/* AllenCahnSolver.init_fields (lbm/fe_ac_solver.py:239-261).  phi_dev and
 * u_dev (nullable: zero velocity) are in the run dtype (fp32 for 16-bit
 * storage).  Computes M0 = sum(phi) in fp64 on the device. */
int lbm_init_fields(lbm_solver* s, const void* phi_dev, const void* u_dev, void* stream);

/* AllenCahnSolver.step() x nsteps (lbm/fe_ac_solver.py:380-639).  Two kernel
 * launches per step, no host synchronisation. */
int lbm_step(lbm_solver* s, int nsteps, void* stream);

/* Capture `nsteps` steps (even, >= 2) into a HIP graph owned by the solver,
 * replacing any earlier capture.  The solver must have stepped at least once
 * since init_fields (so every replay starts from the same buffer parity and
 * state form).  Nothing executes during capture.  The reference's own graph
 * experiment is scripts/bench_cuda_graphs.py:58-91 (torch.cuda.graph around
 * step()). */
int lbm_graph_capture(lbm_solver* s, int nsteps);

/* Launch the captured graph on `stream`: advances the solver by the captured
 * number of steps.  LBM_ESTATE when no graph is held or the state moved since
 * the capture (an odd number of eager steps, an injection); lbm_set_geometry,
 * lbm_set_amplification_field, lbm_init_fields, lbm_load_state,
 * lbm_inject_droplet and the launch-shape setters drop a captured graph
 * (waiting for the device first, in case a replay is in flight). */
int lbm_graph_launch(lbm_solver* s, void* stream);

/* Measurement harness: runs `nsteps` steps on `stream` with HIP events
 * around every kernel launch and writes the mean duration in milliseconds of
 * pass A (out[0], phi + mass) and pass B (out[1], collide) and the mean step
 * time (out[2]).  Synchronises `stream`.  No counterpart in the reference,
 * which only times whole steps with time.time()
 * (scripts/run_thesis_sweep.py:173,248). */
int lbm_profile_steps(lbm_solver* s, int nsteps, void* stream, double* out);

/* Block sizes (threads per workgroup) of pass A and pass B: 128, 256 or 512
 * (default 256).  Invalidates a captured graph. */
int lbm_set_block_sizes(lbm_solver* s, int bs_phi, int bs_collide);

/* Occupancy caps of pass A and pass B in waves per SIMD (1..8; 0 = no cap),
 * enforced with dynamic LDS per workgroup.  Invalidates a captured graph.
 * No counterpart in the reference. */
int lbm_set_occupancy(lbm_solver* s, int cap_phi, int cap_collide);

/* Form of pass B.  1: two launches over 64-node chunk lists, plain-fluid
 * chunks with a wall-free kernel (fewer registers, more resident waves), the
 * other chunks with the general kernel (the default for 16-bit and fp64
 * storage).  2: one launch over both lists, arranged so that each XCD runs
 * its plain chunks first and its wall chunks last (the wall chunks fill the
 * launch's tail; the kernel carries the general code's registers; the
 * default for fp32).  0: one launch over all nodes.  3: the fused step --
 * pass A and pass B in one launch, each workgroup marching a tile of columns
 * through a segment of planes with the phi stencil in an LDS ring (3D with
 * fp32 or fp64 storage and no grad-mu form; elsewhere form 2 runs; the first
 * step after init_fields takes form 1).  Results are identical in every form
 * (the fp64 mass sum of a fused step adds the same values in another order);
 * the autotuner picks one.  LBM_EINVAL for other values.  No counterpart in
 * the reference. */
int lbm_set_split(lbm_solver* s, int split);

/* Tile of the fused step: ty x tx columns (y, x) per workgroup, zlen planes
 * per segment; 0 chooses automatically (the default; sizes above the grid's
 * are clipped to it).  Checked whenever the fused form is (re)planned -- here
 * when form 3 is selected, and by lbm_set_split / lbm_set_block_sizes /
 * lbm_set_occupancy: LBM_EINVAL, with the previous setting kept, when ty * tx
 * exceeds the pass-B block size or (ty + 2) * (tx + 2) twice it.  Invalidates
 * a captured graph.  No counterpart in the reference. */
int lbm_set_fused_tile(lbm_solver* s, int ty, int tx, int zlen);

/* Fused-step plan in use: out6 = {ty, tx, tiles along y, tiles along x,
 * zlen, segments}; zeros when the fused form is not selected. */
int lbm_get_fused_tile(const lbm_solver* s, int* out6);

/* Current launch shape: out5 = {block size A, block size B, cap A, cap B,
 * form of pass B (lbm_set_split)}. */
int lbm_get_launch(const lbm_solver* s, int* out5);

/* Autotuner: times candidate (block size, occupancy cap[, form]) shapes for
 * pass A and pass B over `nsteps_each` real steps each (the solver advances
 * by 27*(nsteps_each+1) steps) and keeps the fastest of each; reports the
 * block sizes in *best_phi / *best_collide (lbm_get_launch gives the rest).
 * Synchronises `stream`.  No counterpart in the reference. */
int lbm_autotune(lbm_solver* s, int nsteps_each, void* stream, int* best_phi,
                 int* best_collide);

/* Observable attributes of the solver object (lbm/fe_ac_solver.py:97-105,
 * lbm/fe_wetting.py:140,147,305-306): writes field `which` into dst_dev. */
int lbm_get_field(lbm_solver* s, int which, void* dst_dev, void* stream);

/* Size in bytes of lbm_get_field's output for `which` (0 if unknown). */
size_t lbm_field_bytes(const lbm_solver* s, int which);

/* AllenCahnSolver._lattice_gradient_and_laplacian (lbm/fe_ac_solver.py:
 * 296-320) with the wetting ghost substitution (lbm/fe_wetting.py:418-446):
 * phi_dev (N) -> grad_dev (D*N), lap_dev (N), run dtype. */
int lbm_grad_lap(lbm_solver* s, const void* phi_dev, void* grad_dev, void* lap_dev,
                 void* stream);

/* lbm_grad_lap without allocations, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  Same inputs and
 * results as lbm_grad_lap; asynchronous on `stream` (no synchronisation).
 * Its scratch -- phi reordered into the library's node order -- is a static
 * device array of 32 MiB held by every GPU the library runs on until the
 * process ends: LBM_EINVAL when N * sizeof(run dtype) exceeds it (8,388,608
 * nodes in fp32 and the 16-bit forms, 4,194,304 in fp64).  Calls that overlap
 * on one GPU share that array and race; calls on one stream are ordered. */
int lbm_grad_lap_noalloc(lbm_solver* s, const void* phi_dev, void* grad_dev, void* lap_dev,
                         void* stream);

/* Mass bookkeeping: M0 (phi_mass_init) and the fp64 sum of the last
 * start-of-step phi.  Synchronises `stream`. */
int lbm_get_mass(lbm_solver* s, double* m0, double* m_last, void* stream);

/* Replace the mass target (inject_droplet resets it, lbm/fe_ac_solver.py:723). */
int lbm_set_mass_target(lbm_solver* s, double m0);

/* Number of completed steps (solver.step_count). */
long long lbm_step_count(const lbm_solver* s);

/* Mid-run droplet injection, AllenCahnSolver.inject_droplet
 * (lbm/fe_ac_solver.py:645-732), between two steps.  `c_dev` is the new
 * drop's indicator C in [0, 1] (N elements of the compute dtype, public
 * order; the reference builds it as 0.5 + 0.5 tanh(2 (R - |x - c|) / xi),
 * clipped, in float32); `u_impact` holds `dim` doubles (NULL: at rest).
 * Where C > 0.1: phi = max(phi, C) and f, g = f_eq(phi, u), g_eq(P, u);
 * everywhere u = (1 - C) u + C u_impact; solid nodes keep phi = u = 0; the
 * mass target becomes the fp64 sum of the new phi and the next step has no
 * d(phi u) term.  Until the next step, LBM_FIELD_P and LBM_FIELD_MU report
 * the values of the last step, as the reference keeps them.  Drops a captured
 * graph.  Not for cases of a batch engine.  Asynchronous on `stream`. */
int lbm_inject_droplet(lbm_solver* s, const void* c_dev, const double* u_impact, void* stream);

/* Solver state snapshot for preemption and checkpoints (no counterpart in
 * the reference, which has no solver checkpoint: its sweep resume only skips
 * finished cases, scripts/run_pending_sweep.py:207-240).  The image holds
 * everything step() reads and the observables need -- distributions, both
 * velocity and both phi buffers, mass factors, step count -- but not the
 * geometry: restore
 * into a solver of the same grid and precision after the same
 * lbm_set_geometry.  `dst` / `src` may be device or host memory of
 * lbm_state_bytes() bytes.  Both synchronise `stream`.  lbm_load_state
 * returns LBM_EINVAL, leaving the solver unchanged, for an image of another
 * grid, precision, format version or distribution layout. */
size_t lbm_state_bytes(const lbm_solver* s);
int lbm_save_state(lbm_solver* s, void* dst, void* stream);
int lbm_load_state(lbm_solver* s, const void* src, void* stream);

/* Diagnostics of lbm/fe_measure.py:11-82 and scripts/run_thesis_sweep.py:
 * 185-206 on the current phi, reduced on the device:
 * out[0..9] = Dx, Dy, Dz, k, volume, C_max, z_com, n_liquid, phi_sum, nan_flag.
 * Synchronises `stream`. */
int lbm_measure(lbm_solver* s, double* out, void* stream);

/* lbm_measure without synchronising: the same 10 doubles are written to
 * device memory out_dev in stream order, so a driver can watch the NaN flag
 * (out_dev[9]) and the extents between launches without blocking the stream
 * (the reference checks nothing inside step() and copies phi to the host
 * every 50 steps, scripts/run_thesis_sweep.py:185-206).  Graph-capturable. */
int lbm_measure_async(lbm_solver* s, double* out_dev, void* stream);

/* Contact-line radius: max horizontal distance from (cx, cy) of nodes with
 * phi > 0.5 in the first non-solid layer above the substrate at each column
 * (3D only).  Synchronises `stream`. */
int lbm_contact_radius(lbm_solver* s, double cx, double cy, double* out, void* stream);

/* Base radius of the drop: max horizontal distance from (cx, cy) of liquid
 * nodes (phi > 0.5, not solid) in the lowest layer that holds liquid (3D
 * only; 0 without liquid).  The reference's 3D contact-angle estimate takes
 * r_contact on the first fluid layer above the wall (lbm/fe_measure.py:
 * 188-241); a drop that spreads on a gas cushion leaves that layer empty, and
 * this radius follows the drop's own base instead.  Synchronises `stream`. */
int lbm_base_radius(lbm_solver* s, double cx, double cy, double* out, void* stream);

/* Self-check of the kernels' division by a divisor d > 0: counts the fp32
 * inputs (all 2^32 bit patterns, NaNs excepted) whose quotient by d differs
 * from IEEE division, over both division forms the kernels use (the general
 * one always, the one-product form unless d is an even integer).
 * Synchronises `stream`.  No counterpart in the reference (whose tensor
 * divisions are IEEE). */
int lbm_check_division(float d, unsigned long long* mismatches, void* stream);

/* Streaming bandwidth of the device (the practical ceiling the step's traffic
 * is compared with): best over three grid sizes of `reps` passes over
 * `bytes` (>= 1 MiB; two buffers of that size are allocated and freed), in
 * TB/s: out6 = {read, read non-temporal, write, write non-temporal, copy,
 * copy non-temporal} (a copy counts its reads and its writes).  Synchronises
 * `stream`.  No counterpart in the reference. */
int lbm_hbm_stream(size_t bytes, int reps, void* stream, double* out6);

/* ------------------------------------------------------------------------
 * Batch engine: many cases of uneven grid size advanced together, one launch
 * per pass per step for all of them (no reference counterpart: the reference
 * runs sweep cases one after another, scripts/run_thesis_sweep.py:57-284,
 * scripts/night_queue.py:28-50).  Cases run fp32 compute; precision selects
 * fp32, bf16 or fp16 distribution storage.
 * ------------------------------------------------------------------------ */
typedef struct lbm_batch lbm_batch;

typedef struct lbm_case_spec {
  lbm_params p;            /* grid, physics and solver options; p.precision must
                              equal the batch precision */
  int substrate;           /* 0 flat, 1 ridge, 2 convex, 3 concave
                              (geometry/substrate.py:162-267) */
  double R_star, R_d;      /* substrate radius |R_star| * R_d (lattice units) */
  double center[3];        /* droplet centre (x, y[, z]); a negative last
                              component places the drop by the reference's
                              sweep rule (scripts/run_thesis_sweep.py:146-157) */
  double radius;           /* droplet radius (lbm/fe_droplet.py:22-112) */
  double u_impact[3];      /* impact velocity */
  long long steps;         /* step target of the case */
} lbm_case_spec;

/* New empty batch; hbm_budget_bytes = 0 means unlimited. */
int lbm_batch_create(int precision, size_t hbm_budget_bytes, lbm_batch** out);
int lbm_batch_destroy(lbm_batch* b);

/* Device bytes one case of these parameters occupies (admission control). */
size_t lbm_case_footprint(const lbm_params* p);

/* Build a case on the device (substrate, droplet, geometry precompute,
 * init_fields) and add it to the batch.  LBM_ENOMEM when it does not fit
 * the budget.  Does not synchronise other cases' work. */
int lbm_batch_add(lbm_batch* b, const lbm_case_spec* spec, void* stream, int* case_id);

/* lbm_batch_add with the case's inputs handed over instead of generated:
 * solid (uint8 0/1), the solid fraction (frac_dtype, nullable) and the initial
 * phi and velocity (fp32, u nullable: at rest), device arrays in the public
 * order -- exactly what AllenCahnSolver.set_solid / init_fields take
 * (lbm/fe_ac_solver.py:193-261).  A case built from the reference's
 * generators (geometry/substrate.py:162-267, lbm/fe_droplet.py:22-112) then
 * starts from the same bits as a stand-alone solver.  The spec's grid,
 * physics and step target are used; its substrate, R_star, R_d, center,
 * radius and u_impact are not read.  Synchronises `stream`. */
int lbm_batch_add_fields(lbm_batch* b, const lbm_case_spec* spec, const uint8_t* solid_dev,
                         const void* frac_dev, int frac_dtype, const void* phi_dev,
                         const void* u_dev, void* stream, int* case_id);

/* Release a case (synchronises the device). */
int lbm_batch_remove(lbm_batch* b, int case_id);

/* The solver handle of a case, for the calls that only read it
 * (lbm_get_field, lbm_field_bytes, lbm_measure[_async], lbm_contact_radius,
 * lbm_get_mass, lbm_step_count, lbm_state_bytes, lbm_save_state,
 * lbm_grad_lap).  The calls that would change the case behind the batch's
 * back (lbm_step, lbm_graph_*, lbm_set_geometry, lbm_set_amplification_field,
 * lbm_init_fields, lbm_load_state, lbm_inject_droplet, the launch-shape
 * setters, lbm_autotune, lbm_profile_steps, lbm_destroy) return LBM_EINVAL for
 * such a handle. */
int lbm_batch_solver(lbm_batch* b, int case_id, lbm_solver** s);

/* Restore a case from an lbm_save_state image of its solver (taken with
 * lbm_batch_solver + lbm_save_state): the case resumes bit for bit at the
 * image's step.  For crash-safe sweep queues (the reference's sweep resume
 * only skips finished cases, scripts/run_pending_sweep.py:207-240).
 * LBM_EINVAL for an image past the case's step target or one whose buffer set
 * does not follow its step parity (a stand-alone solver's image taken after
 * an injection).  Synchronises `stream`; drops a captured batch graph. */
int lbm_batch_load_state(lbm_batch* b, int case_id, const void* src, void* stream);

/* Steps done and step target of a case. */
int lbm_batch_status(lbm_batch* b, int case_id, long long* steps_done, long long* steps_target);

/* Device bytes held by the batch's cases. */
size_t lbm_batch_bytes_used(const lbm_batch* b);

/* Advance every unfinished case by up to nsteps steps (cases stop at their
 * target): 3 or 4 launches per step for the whole batch, no host
 * synchronisation. */
int lbm_batch_step(lbm_batch* b, int nsteps, void* stream);

/* Block sizes of the batch kernels (128/256/512). */
int lbm_batch_set_block_sizes(lbm_batch* b, int bs_phi, int bs_collide);

/* Form of pass B for the whole batch, as lbm_set_split for one solver: 1
 * (default) plain-fluid / other chunk-list launches, 2 one launch over every
 * case's single-launch list (per XCD range plain chunks first, wall chunks
 * last), 0 one launch over all nodes; results are identical.  LBM_EINVAL for
 * other values.  Drops a captured batch graph. */
int lbm_batch_set_split(lbm_batch* b, int split);

/* Occupancy caps of the batch's pass A and pass B (waves per SIMD, 1..8; 0:
 * none), as lbm_set_occupancy.  Drops a captured batch graph. */
int lbm_batch_set_occupancy(lbm_batch* b, int cap_phi, int cap_collide);

/* Current batch launch shape: out5 = {block size A, block size B, cap A,
 * cap B, form of pass B}. */
int lbm_batch_get_launch(const lbm_batch* b, int* out5);

/* Autotuner of the batch launch shape: times the 9 (form of pass B, pass-B
 * occupancy cap) pairs at 256-thread blocks, then the other 8 block-size
 * pairs with the winner, over nsteps_each steps each (the batch advances by
 * 17*(nsteps_each+1) steps), and keeps the fastest.  Synchronises. */
int lbm_batch_autotune(lbm_batch* b, int nsteps_each, void* stream, int* bs_phi,
                       int* bs_collide);

/* Capture nsteps (even) batch steps into a graph for the current case set;
 * every active case must have stepped once and have nsteps left.  Any change
 * of the case set or launch shape (add, remove, a case reaching its target,
 * set_split, set_block_sizes, autotune, load_state) drops the graph, and
 * lbm_batch_graph_launch then returns LBM_ESTATE until it is captured again. */
int lbm_batch_graph_capture(lbm_batch* b, int nsteps, void* stream);
int lbm_batch_graph_launch(lbm_batch* b, void* stream);

/* Thread-local message for the last failed call. */
const char* lbm_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* LBM_AMD_H */
