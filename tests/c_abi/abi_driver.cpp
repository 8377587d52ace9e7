// A caller of the C ABI with no Python and no torch: what a cgo, JNI or
// N-API binding does.  Host code only (built with g++); device memory comes
// from the HIP runtime API.  Runs a 24 x 24 x 20 flat-wall impact in fp32
// through the whole life cycle and checks what can be checked without a
// reference solution:
//   * mass: lbm_get_mass's M0 equals the fp64 sum of the clamped phi input,
//     and sum(phi) after 20 steps stays within 1e-4 of it (the mass factor,
//     with the boundary relax off as in the reference's validation gate);
//   * determinism: two solvers stepped alike give bit-identical phi;
//   * state image: save, step 5, load, step 5 reproduces phi bit for bit;
//   * injection: M0 grows by the injected liquid; u stays finite;
//   * errors: null and out-of-order calls return negative codes and set
//     lbm_last_error.
// Prints "abi_driver ok" and exits 0 on success.
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/lbm_amd.h"

static int failures = 0;
#define CHECK(cond, ...)                                   \
  do {                                                     \
    if (!(cond)) {                                         \
      std::fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
      std::fprintf(stderr, __VA_ARGS__);                   \
      std::fprintf(stderr, "\n");                          \
      ++failures;                                          \
    }                                                      \
  } while (0)
#define OK(call)                                                              \
  do {                                                                        \
    int rc_ = (call);                                                         \
    if (rc_ != LBM_OK) {                                                      \
      std::fprintf(stderr, "FAIL %s -> %d (%s)\n", #call, rc_, lbm_last_error()); \
      std::exit(2);                                                           \
    }                                                                         \
  } while (0)
#define HIP(call)                                                    \
  do {                                                               \
    hipError_t e_ = (call);                                          \
    if (e_ != hipSuccess) {                                          \
      std::fprintf(stderr, "HIP %s: %s\n", #call, hipGetErrorString(e_)); \
      std::exit(3);                                                  \
    }                                                                \
  } while (0)

static const int NX = 24, NY = 24, NZ = 20;
static const size_t N = (size_t)NX * NY * NZ;

static size_t at(int x, int y, int z) { return ((size_t)x * NY + y) * NZ + z; }

static lbm_params params() {
  lbm_params p;
  lbm_params_default(&p, NX, NY, NZ);
  p.precision = LBM_F32;
  p.rho_l = 1.0;
  p.rho_g = 1.0 / 828.0;
  p.sigma = 0.01424;
  p.xi = 4.0;
  p.beta = 12.0 * p.sigma / p.xi;                      // FEConfig derivations
  p.kappa = p.beta * p.xi * p.xi / 8.0;
  p.tau_l = p.tau_g = 0.53;
  p.tau_h = 1.0;
  p.theta_eq = 162.0;
  p.phi_c = -std::cos(p.theta_eq * M_PI / 180.0) * std::sqrt(2.0 * p.kappa * p.beta);
  p.geometric_wetting = 1;
  p.geo_amplification = 1.5;
  p.boundary_relax = 0.0;   // no S12: the mass factor alone holds sum(phi)
  return p;
}

struct Case {
  lbm_solver* s = nullptr;
  void *solid = nullptr, *phi = nullptr, *u = nullptr;
};

static double host_sum(const std::vector<float>& v) {
  double t = 0.0;
  for (float x : v) t += (double)x;
  return t;
}

static std::vector<float> read_phi(lbm_solver* s) {
  std::vector<float> out(N);
  void* d = nullptr;
  HIP(hipMalloc(&d, N * 4));
  OK(lbm_get_field(s, LBM_FIELD_PHI, d, nullptr));
  HIP(hipMemcpy(out.data(), d, N * 4, hipMemcpyDeviceToHost));
  HIP(hipFree(d));
  return out;
}

static Case make_case(std::vector<float>& phi_h) {
  Case c;
  lbm_params p = params();
  OK(lbm_create(&p, &c.s));
  std::vector<unsigned char> solid(N, 0);
  for (int x = 0; x < NX; ++x)
    for (int y = 0; y < NY; ++y) solid[at(x, y, 0)] = 1;   // flat wall, no fraction
  phi_h.assign(N, 0.0f);
  std::vector<float> u(3 * N, 0.0f);
  for (int x = 0; x < NX; ++x)
    for (int y = 0; y < NY; ++y)
      for (int z = 0; z < NZ; ++z) {
        const double r = std::sqrt((x - 12.0) * (x - 12.0) + (y - 12.0) * (y - 12.0) +
                                   (z - 9.0) * (z - 9.0));
        const float v = (float)(0.5 + 0.5 * std::tanh(2.0 * (5.0 - r) / 4.0));
        phi_h[at(x, y, z)] = v;
        u[2 * N + at(x, y, z)] = -0.05f * v;
      }
  HIP(hipMalloc(&c.solid, N));
  HIP(hipMalloc(&c.phi, N * 4));
  HIP(hipMalloc(&c.u, 3 * N * 4));
  HIP(hipMemcpy(c.solid, solid.data(), N, hipMemcpyHostToDevice));
  HIP(hipMemcpy(c.phi, phi_h.data(), N * 4, hipMemcpyHostToDevice));
  HIP(hipMemcpy(c.u, u.data(), 3 * N * 4, hipMemcpyHostToDevice));
  OK(lbm_set_geometry(c.s, (const uint8_t*)c.solid, nullptr, LBM_FRAC_NONE, nullptr));
  OK(lbm_init_fields(c.s, c.phi, c.u, nullptr));
  return c;
}

static void free_case(Case& c) {
  OK(lbm_destroy(c.s));
  HIP(hipFree(c.solid));
  HIP(hipFree(c.phi));
  HIP(hipFree(c.u));
}

int main() {
  // error paths before any GPU work
  CHECK(lbm_create(nullptr, nullptr) == LBM_EINVAL, "null create accepted");
  CHECK(std::strlen(lbm_last_error()) > 0, "no error message");
  CHECK(lbm_step(nullptr, 1, nullptr) == LBM_EINVAL, "null step accepted");
  {
    lbm_params p = params();
    lbm_solver* s = nullptr;
    OK(lbm_create(&p, &s));
    CHECK(lbm_step(s, 1, nullptr) == LBM_ESTATE, "step before init accepted");
    int launch[5] = {0, 0, 0, 0, 0};
    OK(lbm_get_launch(s, launch));
    CHECK(launch[4] == 2, "fp32 default pass-B form %d", launch[4]);
    CHECK(lbm_set_split(s, 4) == LBM_EINVAL && lbm_set_split(s, -1) == LBM_EINVAL,
          "bad pass-B form accepted");
    // the fused form plans a tile; an oversized tile is refused and the plan stays
    int plan[6] = {0, 0, 0, 0, 0, 0};
    OK(lbm_set_split(s, 3));
    OK(lbm_get_fused_tile(s, plan));
    CHECK(plan[0] > 0 && plan[1] > 0 && plan[0] * plan[1] <= 256 && plan[5] >= 1,
          "fused plan %d %d %d %d %d %d", plan[0], plan[1], plan[2], plan[3], plan[4], plan[5]);
    CHECK(lbm_set_fused_tile(s, 64, 64, 0) == LBM_EINVAL, "oversized fused tile accepted");
    CHECK(lbm_set_fused_tile(s, -1, 0, 0) == LBM_EINVAL, "negative fused tile accepted");
    int plan2[6] = {0, 0, 0, 0, 0, 0};
    OK(lbm_get_fused_tile(s, plan2));
    CHECK(plan2[0] == plan[0] && plan2[1] == plan[1], "fused plan changed by a refused tile");
    OK(lbm_set_split(s, 1));
    OK(lbm_get_fused_tile(s, plan2));
    CHECK(plan2[0] == 0, "fused plan kept after leaving form 3");
    OK(lbm_get_launch(s, launch));
    CHECK(launch[4] == 1, "pass-B form not set: %d", launch[4]);
    OK(lbm_destroy(s));
  }

  std::vector<float> phi0;
  Case a = make_case(phi0), b = make_case(phi0);

  // M0: fp64 sum of the (already clamped) input
  double m0 = 0.0, mlast = 0.0;
  OK(lbm_get_mass(a.s, &m0, &mlast, nullptr));
  CHECK(std::fabs(m0 - host_sum(phi0)) <= 1e-9 * m0, "M0 %.17g vs %.17g", m0, host_sum(phi0));

  OK(lbm_step(a.s, 20, nullptr));
  OK(lbm_step(b.s, 20, nullptr));
  HIP(hipDeviceSynchronize());
  std::vector<float> pa = read_phi(a.s), pb = read_phi(b.s);
  CHECK(std::memcmp(pa.data(), pb.data(), N * 4) == 0, "two identical runs differ");
  CHECK(std::fabs(host_sum(pa) - m0) <= 1e-4 * m0, "mass drift %.3e",
        std::fabs(host_sum(pa) - m0) / m0);
  CHECK(lbm_step_count(a.s) == 20, "step count %lld", lbm_step_count(a.s));

  // state image in host memory: save, advance, restore, advance again
  const size_t nb = lbm_state_bytes(a.s);
  std::vector<unsigned char> img(nb);
  OK(lbm_save_state(a.s, img.data(), nullptr));
  OK(lbm_step(a.s, 5, nullptr));
  std::vector<float> p25 = read_phi(a.s);
  OK(lbm_load_state(a.s, img.data(), nullptr));
  CHECK(lbm_step_count(a.s) == 20, "restored step count %lld", lbm_step_count(a.s));
  OK(lbm_step(a.s, 5, nullptr));
  std::vector<float> p25b = read_phi(a.s);
  CHECK(std::memcmp(p25.data(), p25b.data(), N * 4) == 0, "resumed run differs");
  {
    // a damaged magic word or an older format version is refused
    std::vector<unsigned char> bad(img);
    bad[0] ^= 0xff;
    CHECK(lbm_load_state(a.s, bad.data(), nullptr) == LBM_EINVAL, "bad magic accepted");
    bad = img;
    uint32_t v = 0;
    std::memcpy(&v, bad.data() + 4, 4);
    v -= 1;
    std::memcpy(bad.data() + 4, &v, 4);
    CHECK(lbm_load_state(a.s, bad.data(), nullptr) == LBM_EINVAL, "old version accepted");
    CHECK(std::strstr(lbm_last_error(), "layout") != nullptr, "message: %s", lbm_last_error());
    CHECK(lbm_step_count(a.s) == 25, "refused image changed the solver");
  }

  // a second drop between steps: the mass target grows
  double m_before = 0.0;
  OK(lbm_get_mass(a.s, &m_before, nullptr, nullptr));
  std::vector<float> c(N, 0.0f);
  for (int x = 0; x < NX; ++x)
    for (int y = 0; y < NY; ++y)
      for (int z = 0; z < NZ; ++z) {
        const double r = std::sqrt((x - 5.0) * (x - 5.0) + (y - 18.0) * (y - 18.0) +
                                   (z - 15.0) * (z - 15.0));
        c[at(x, y, z)] = (float)std::fmin(1.0, std::fmax(0.0, 0.5 + 0.5 * std::tanh(2.0 * (3.0 - r) / 4.0)));
      }
  void* cd = nullptr;
  HIP(hipMalloc(&cd, N * 4));
  HIP(hipMemcpy(cd, c.data(), N * 4, hipMemcpyHostToDevice));
  const double uimp[3] = {0.0, 0.0, -0.04};
  OK(lbm_inject_droplet(a.s, cd, uimp, nullptr));
  double m_after = 0.0;
  OK(lbm_get_mass(a.s, &m_after, nullptr, nullptr));
  CHECK(m_after > m_before + 50.0, "injection added %.3f", m_after - m_before);
  OK(lbm_step(a.s, 10, nullptr));
  std::vector<float> u(3 * N);
  void* ud = nullptr;
  HIP(hipMalloc(&ud, 3 * N * 4));
  OK(lbm_get_field(a.s, LBM_FIELD_U, ud, nullptr));
  HIP(hipMemcpy(u.data(), ud, 3 * N * 4, hipMemcpyDeviceToHost));
  bool finite = true;
  for (float v : u) finite = finite && std::isfinite(v);
  CHECK(finite, "non-finite velocity after injection");
  double meas[10];
  OK(lbm_measure(a.s, meas, nullptr));
  CHECK(meas[9] == 0.0 && meas[0] > 0.0, "measure: Dx %.1f nan %.0f", meas[0], meas[9]);

  HIP(hipFree(cd));
  HIP(hipFree(ud));
  free_case(a);
  free_case(b);
  if (failures) {
    std::fprintf(stderr, "abi_driver: %d failure(s)\n", failures);
    return 1;
  }
  std::printf("abi_driver ok\n");
  return 0;
}
