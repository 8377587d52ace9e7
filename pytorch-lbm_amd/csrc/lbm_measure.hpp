// Diagnostics reduced on the device (lbm/fe_measure.py:11-82 and the driver
// metrics of scripts/run_thesis_sweep.py:185-206): extents of the liquid
// (phi > 0.5 and not solid), volume (sum of phi with solid set to 0), C_max,
// z_com (mean last-axis coordinate of the liquid nodes), NaN flag; the
// contact-line radius on the first fluid layer above each column's surface;
// and the base radius of the drop on its lowest liquid layer.
// Each block writes one partial record; a one-block kernel combines the
// records in a fixed order, so results do not depend on scheduling.
#pragma once

#include "lbm_kernels.hpp"

namespace lbm {

struct MeasRec {
  int mn[3], mx[3];
  unsigned long long cnt, zsum;
  double vol, cmax;
  int nan;
  int pad;
};

template <typename T, int DIM>
__device__ __forceinline__ T phi_observed(const Phys<T>& ph, const T* __restrict__ phi,
                                          uint32_t x, uint8_t m, T cf, int scaled) {
  const T p = phi[x];
  if (!scaled) return p;
  return s12_phi<T>(ph, m, clamp01(p * cf), nsw_phi<T>(ph, m, phi, x), cf);
}

template <int BS>
__device__ void combine_rec(MeasRec* s, int n) {
  // s[0..n) in LDS -> s[0], fixed tree order
  for (int w = 1; w < n; w <<= 1) {
    if ((int)threadIdx.x % (2 * w) == 0 && (int)threadIdx.x + w < n) {
      MeasRec& a = s[threadIdx.x];
      const MeasRec& b = s[threadIdx.x + w];
      for (int d = 0; d < 3; ++d) {
        a.mn[d] = b.mn[d] < a.mn[d] ? b.mn[d] : a.mn[d];
        a.mx[d] = b.mx[d] > a.mx[d] ? b.mx[d] : a.mx[d];
      }
      a.cnt += b.cnt;
      a.zsum += b.zsum;
      a.vol += b.vol;
      a.cmax = b.cmax > a.cmax ? b.cmax : a.cmax;
      a.nan |= b.nan;
    }
    __syncthreads();
  }
}

__device__ __forceinline__ void rec_empty(MeasRec& r) {
  for (int d = 0; d < 3; ++d) { r.mn[d] = 0x7fffffff; r.mx[d] = -1; }
  r.cnt = 0; r.zsum = 0; r.vol = 0.0; r.cmax = 0.0; r.nan = 0; r.pad = 0;
}

// Per-block records (block order is the reduction order of k_measure_final).
// Thread t of block b folds nodes b*BS + t + k*(gridDim.x*BS), k = 0..kMeasPer-1,
// so loads stay coalesced and the grid (a fixed function of N) is 16x smaller
// than one node per thread.
constexpr int kMeasPer = 16;
template <typename T, int DIM, int BS>
__global__ __launch_bounds__(BS) void k_measure(Grid g, Phys<T> ph, const T* __restrict__ phi,
                                                const uint8_t* __restrict__ meta,
                                                const Ctrl* ctrl, int slot, int scaled,
                                                MeasRec* __restrict__ parts) {
  __shared__ MeasRec s[BS];
  MeasRec r;
  rec_empty(r);
  const T cf = scaled ? (T)ctrl->c[slot] : T(1);
  const uint32_t stride = gridDim.x * BS;
  for (int k = 0; k < kMeasPer; ++k) {
    const uint32_t x = blockIdx.x * BS + threadIdx.x + (uint32_t)k * stride;
    if (x >= g.N) break;
    const uint8_t m = meta[x];
    T p = phi_observed<T, DIM>(ph, phi, x, m, cf, scaled);
    if (p != p) r.nan = 1;
    if (m & M_SOLID) p = T(0);
    r.vol += (double)p;
    r.cmax = (double)p > r.cmax ? (double)p : r.cmax;
    if (p > T(0.5)) {
      const Coord<DIM> c = decompose<DIM>(g, x);
      const int cc[3] = {c.x, c.y, c.z};
      for (int d = 0; d < DIM; ++d) {
        r.mn[d] = cc[d] < r.mn[d] ? cc[d] : r.mn[d];
        r.mx[d] = cc[d] > r.mx[d] ? cc[d] : r.mx[d];
      }
      r.cnt += 1;
      r.zsum += (unsigned long long)cc[DIM - 1];
    }
  }
  s[threadIdx.x] = r;
  __syncthreads();
  combine_rec<BS>(s, BS);
  if (threadIdx.x == 0) parts[blockIdx.x] = s[0];
}

// One block folds the n block records and writes
// out[0..9] = Dx, Dy, Dz, k, volume, C_max, z_com, n_liquid, phi_sum, nan.
// Thread t folds records t, t + BS, t + 2 BS, ... in that order, then the
// block combines the BS partial records in a fixed tree: the result does not
// depend on scheduling.
__device__ __forceinline__ void fold_rec(MeasRec& a, const MeasRec& b) {
  for (int d = 0; d < 3; ++d) {
    a.mn[d] = b.mn[d] < a.mn[d] ? b.mn[d] : a.mn[d];
    a.mx[d] = b.mx[d] > a.mx[d] ? b.mx[d] : a.mx[d];
  }
  a.cnt += b.cnt;
  a.zsum += b.zsum;
  a.vol += b.vol;
  a.cmax = b.cmax > a.cmax ? b.cmax : a.cmax;
  a.nan |= b.nan;
}

template <int DIM, int BS>
__global__ __launch_bounds__(BS) void k_measure_final(const MeasRec* __restrict__ parts,
                                                      unsigned int n, double* out,
                                                      int* zmin_out = nullptr) {
  __shared__ MeasRec s[BS];
  MeasRec acc;
  rec_empty(acc);
  for (unsigned int i = threadIdx.x; i < n; i += BS) fold_rec(acc, parts[i]);
  s[threadIdx.x] = acc;
  __syncthreads();
  combine_rec<BS>(s, BS);
  if (threadIdx.x == 0) {
    acc = s[0];
    double D[3] = {0, 0, 0};
    if (acc.cnt > 0)
      for (int d = 0; d < DIM; ++d) D[d] = (double)(acc.mx[d] - acc.mn[d] + 1);
    out[0] = D[0];
    out[1] = D[1];
    out[2] = DIM == 3 ? D[2] : 0.0;
    out[3] = (acc.cnt > 0 && D[1] > 0) ? D[0] / D[1] : 0.0;
    out[4] = acc.vol;
    out[5] = acc.cmax;
    out[6] = acc.cnt > 0 ? (double)acc.zsum / (double)acc.cnt : 0.0;
    out[7] = (double)acc.cnt;
    out[8] = acc.vol;
    out[9] = (double)acc.nan;
    // lowest liquid layer (last axis), -1 without liquid
    if (zmin_out) *zmin_out = acc.cnt > 0 ? acc.mn[DIM - 1] : -1;
  }
}

// Base radius of the drop: one thread per column of the lowest liquid layer
// *zmin (from k_measure_final); per-block max of r^2 over its liquid nodes.
template <typename T, int DIM, int BS>
__global__ __launch_bounds__(BS) void k_base(Grid g, Phys<T> ph, const T* __restrict__ phi,
                                             const uint8_t* __restrict__ meta, const Ctrl* ctrl,
                                             int slot, int scaled, double cx, double cy,
                                             const int* __restrict__ zmin,
                                             double* __restrict__ parts) {
  __shared__ double s[BS];
  const long long col = (long long)blockIdx.x * BS + threadIdx.x;
  const long long ncol = (long long)g.nx * g.ny;
  double r2 = -1.0;
  const int z = *zmin;
  if (DIM == 3 && col < ncol && z >= 0 && z < g.nz) {
    const long long x = int_index<DIM>(g, (int)(col / g.ny), (int)(col % g.ny), z);
    const uint8_t m = meta[x];
    const T cf = scaled ? (T)ctrl->c[slot] : T(1);
    const T p = phi_observed<T, DIM>(ph, phi, x, m, cf, scaled);
    if (!(m & M_SOLID) && p > T(0.5)) {
      const double dx = (double)(col / g.ny) - cx, dy = (double)(col % g.ny) - cy;
      r2 = dx * dx + dy * dy;
    }
  }
  s[threadIdx.x] = r2;
  __syncthreads();
  for (int w = BS / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w && s[threadIdx.x + w] > s[threadIdx.x]) s[threadIdx.x] = s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) parts[blockIdx.x] = s[0];
}

// Contact-line radius: one thread per column; per-block max of r^2.
template <typename T, int DIM, int BS>
__global__ __launch_bounds__(BS) void k_contact(Grid g, Phys<T> ph, const T* __restrict__ phi,
                                                const uint8_t* __restrict__ meta,
                                                const int16_t* __restrict__ surf,
                                                const Ctrl* ctrl, int slot, int scaled,
                                                double cx, double cy,
                                                double* __restrict__ parts) {
  __shared__ double s[BS];
  const long long col = (long long)blockIdx.x * BS + threadIdx.x;
  const long long ncol = (long long)g.nx * g.ny;
  double r2 = -1.0;
  if (DIM == 3 && col < ncol) {
    const int h = surf[col];
    const int z = h + 1;
    if (z < g.nz) {
      const long long x = int_index<DIM>(g, (int)(col / g.ny), (int)(col % g.ny), z);
      const uint8_t m = meta[x];
      const T cf = scaled ? (T)ctrl->c[slot] : T(1);
      const T p = phi_observed<T, DIM>(ph, phi, x, m, cf, scaled);
      if (!(m & M_SOLID) && p > T(0.5)) {
        const double dx = (double)(col / g.ny) - cx, dy = (double)(col % g.ny) - cy;
        r2 = dx * dx + dy * dy;
      }
    }
  }
  s[threadIdx.x] = r2;
  __syncthreads();
  for (int w = BS / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w && s[threadIdx.x + w] > s[threadIdx.x]) s[threadIdx.x] = s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) parts[blockIdx.x] = s[0];
}

template <int BS>
__global__ __launch_bounds__(BS) void k_contact_final(const double* __restrict__ parts,
                                                      unsigned int n, double* out) {
  if (threadIdx.x != 0) return;
  double best = -1.0;
  for (unsigned int i = 0; i < n; ++i) best = parts[i] > best ? parts[i] : best;
  *out = best < 0 ? 0.0 : sqrt(best);
}

}  // namespace lbm
