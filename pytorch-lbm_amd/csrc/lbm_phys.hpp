// Physics constants of one solver, prepared once on the host in the run
// precision T so that every scalar enters the device arithmetic exactly as
// the reference's tensor ops see it (a Python float meeting a float32 tensor
// is rounded to float32 first; values the reference derives with tensor
// arithmetic, such as w/cs2 or 1/tau_eff, are derived here with T arithmetic).
#pragma once

#include "lbm_lattice.hpp"

namespace lbm {

enum WetMode : int { WET_NONE = 0, WET_LINEAR = 1, WET_QUADRATIC = 2 };
// interface sources that use grad mu
enum GmuMode : int { GMU_NONE = 0, GMU_CH = 1, GMU_DRIVE = 2 };

// Direction-dependent constants are stored once per weight class (rest, axis,
// diagonal): with 19 per-direction copies the kernels ran out of scalar
// registers.  The stencil's (w_i / cs2) * e_id is +-woc[class] (e_id = +-1),
// applied as an exact sign on the product.
template <typename T> struct Phys {
  T wc[3];            // T(w32) per class
  T woc[3];           // w / cs2                 (lbm/fe_ac_solver.py:306,318)
  T lc[3];            // 2 * (w / cs2)           (lbm/fe_ac_solver.py:319)
  T lic[3];           // 2 * w                   (lbm/fe_ac_solver.py:293)
  T w3c[3];           // T(3 w), the bf16 g reference's weight
  T cs2, two_cs2;
  double rcp_cs2, rcp_two_cs2, rcp_xi;   // RN64(1/d) of the fp32 cs2, 2 cs2, xi
  int rc_small;       // rho_eff cs2 < 2 for every density the run can form
  T four_beta, kappa, rho_l, rho_g, rho_floor, drho, visc, xi, ac_scale;
  T mu_wall_add;      // linear mode: mu_wall * mu_wall_scale
  T phi_c;            // quadratic mode
  float hk32, pck32;  // h/kappa*ghost_scale, phi_c/kappa (fp32 n.e products)
  T cos_t, sin_t, amp;
  int amp_on;         // scalar alpha > 1
  int corr_on;        // geometric gradient correction active
  T sfac[16];         // 1 - sharpen_reduce * wall_weight per distance class
  T acoef, om_f;      // 1 - 0.5/tau_f, 1/tau_f
  T ns_c, omg_c;      // 1 - 0.5/tau_eff, 1/tau_eff when tau_eff is constant
  int var_tau;
  T tl_m, tg_m;       // tau_l - 0.5, tau_g - 0.5
  T gforce[3];
  int body_on;
  int stab;           // 0 fakhari, 1 liang, 2 none
  int wet;            // WetMode
  int s12_on;
  T s12_r, s12_t;
  T thr_lo, thr_hi;   // 0.15, 0.85
  // direct wetting: S12 target = cubic wall-energy phi of each top boundary
  // node, from the phi of its next-interior node nsw[x] (-1: none / solid)
  int direct;         // direct_wetting and a != 0
  int direct_on;      // direct_wetting (S12 runs even without a cubic target)
  T cub_4a, cub_1ma2, cub_m1a, cub_2a;   // 4a, (1-a)^2, -(1-a), 2a in T
  T lo01, hi99;       // 0.01, 0.99
  const int32_t* nsw;
  int gmu_mode;       // GmuMode
  T mob, guo;         // cs2 (tau_f - 0.5), tau_f / (tau_f - 0.5)
  T mu_drive_a;       // mu_drive * mobility

  template <int DIM> __host__ __device__ T w(int i) const { return wc[wclass<DIM>(i)]; }
  template <int DIM> __host__ __device__ T w3(int i) const { return w3c[wclass<DIM>(i)]; }
};

}  // namespace lbm
