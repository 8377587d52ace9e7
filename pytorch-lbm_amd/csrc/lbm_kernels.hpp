// Device kernels of the Allen-Cahn step.
//
// State between steps (pull scheme): Fin/Gin hold the post-collision sets
// f*, g* of the previous step.  Every kernel that needs the post-stream
// distributions of the current step pulls them node by node:
//   fluid x   : f_i(x) = f*_i(x - e_i)                 (periodic, all axes)
//   solid x   : f_i(x) = f*_opp(i)(x)                  (bounce-back)
//   partial x : f_i(x) = (1-eps) f*_i(x-e_i) + eps f*_opp(i)(x)
// and then applies the lagged mass factor c and the boundary-relax scale s,
// so all writes of a step are node-local.  After init_fields the buffers hold
// post-stream values instead ("local" form) and are read in place.
//
// Pass A (k_phi): phi = clamp(sum_i f_i, 0, 1), 0 at solid; fp64 block
// partials of sum(phi); the last block to finish adds the partials in a fixed
// order and writes the step's mass factor.
// Pass B (k_collide): 1-ring stencil of phi with wetting ghosts, chemical
// potential, geometric gradient correction, stress, forces, velocity, AC
// source, both BGK collisions with Guo forcing; stores f*, g*, u.
//
// Arithmetic mirrors lbm/fe_ac_solver.py:380-639 operation by operation
// (same association order, same scalar roundings); sums over directions use
// the order of torch's CPU reduction (runs of 16, then run totals).
#pragma once

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <type_traits>

#include "lbm_lattice.hpp"
#include "lbm_phys.hpp"

namespace lbm {

// Cache-policy operand of buffer loads and stores: the NT (non-temporal) bit.
constexpr int kNT = 2;

// Pass-B chunk-list entries: a 64-node chunk index plus two flags -- bit 31
// marks a wall chunk in the single-launch list, bit 30 a chunk whose 64 nodes
// are all solid (no fluid pull of it is ever used).
constexpr uint32_t kChunkWall = 0x80000000u, kChunkSolid = 0x40000000u, kChunkIdx = 0x3FFFFFFFu;

struct Ctrl {
  double M0;
  double mass[2];     // fp64 sum of the start-of-step phi, by step parity
  double c[2];        // mass factor c_k, by step parity
  unsigned int counter;
  unsigned int counter2;
  double stats[16];
};

// Distribution storage: fp32/fp64 as is; bf16/fp16 with round-to-nearest-even
// conversion, collided in fp32.  Accesses use a per-direction base pointer
// (uniform, so it lives in SGPRs) plus a 32-bit element index, which maps to
// the SGPR-base + 32-bit VGPR-offset global addressing mode.
template <typename S> struct Stor {
  template <typename T> __device__ static __forceinline__ T to(S v) { return (T)v; }
  template <typename T> __device__ static __forceinline__ S from(T v) { return (S)v; }
};
template <> struct Stor<__hip_bfloat16> {
  template <typename T> __device__ static __forceinline__ T to(__hip_bfloat16 v) {
    return (T)__bfloat162float(v);
  }
  template <typename T> __device__ static __forceinline__ __hip_bfloat16 from(T v) {
    return __float2bfloat16((float)v);
  }
};
template <> struct Stor<__half> {
  template <typename T> __device__ static __forceinline__ T to(__half v) {
    return (T)__half2float(v);
  }
  template <typename T> __device__ static __forceinline__ __half from(T v) {
    return __float2half((float)v);
  }
};

template <typename E> __device__ __forceinline__ const E& at(const E* base, uint32_t idx) {
  return *(const E*)((const char*)base + (size_t)(idx * (uint32_t)sizeof(E)));
}
template <typename E> __device__ __forceinline__ E& at(E* base, uint32_t idx) {
  return *(E*)((char*)base + (size_t)(idx * (uint32_t)sizeof(E)));
}

// Raw buffer resource over one device array: loads/stores take a 32-bit
// per-lane byte offset and a uniform (SGPR) byte offset, so the per-direction
// base i*N never costs vector address arithmetic.  Out-of-range offsets read
// as 0 and drop writes (num_records bound).
struct Buf {
  __amdgpu_buffer_rsrc_t r;
  __device__ __forceinline__ Buf(const void* p, long long bytes)
      : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0,
                                            bytes >= 0x7fffffffLL ? -1 : (int)bytes,
                                            0x00020000)) {}
};

template <typename E> struct BufIO;
template <> struct BufIO<float> {
  template <int CP = 0> __device__ static __forceinline__ float ld(const Buf& b, uint32_t vo, uint32_t so) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(b.r, (int)vo, (int)so, CP));
  }
  template <int CP = 0> __device__ static __forceinline__ void st(const Buf& b, uint32_t vo, uint32_t so, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), b.r, (int)vo,
                                          (int)so, CP);
  }
};
template <> struct BufIO<double> {
  template <int CP = 0> __device__ static __forceinline__ double ld(const Buf& b, uint32_t vo, uint32_t so) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(b.r, (int)vo, (int)so, CP));
  }
  template <int CP = 0> __device__ static __forceinline__ void st(const Buf& b, uint32_t vo, uint32_t so, double v) {
    typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), b.r, (int)vo, (int)so, CP);
  }
};
template <> struct BufIO<__hip_bfloat16> {
  template <int CP = 0> __device__ static __forceinline__ __hip_bfloat16 ld(const Buf& b, uint32_t vo, uint32_t so) {
    return __builtin_bit_cast(__hip_bfloat16,
                              __builtin_amdgcn_raw_buffer_load_b16(b.r, (int)vo, (int)so, CP));
  }
  template <int CP = 0> __device__ static __forceinline__ void st(const Buf& b, uint32_t vo, uint32_t so,
                                            __hip_bfloat16 v) {
    __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, v), b.r, (int)vo,
                                          (int)so, CP);
  }
};
template <> struct BufIO<__half> {
  template <int CP = 0> __device__ static __forceinline__ __half ld(const Buf& b, uint32_t vo, uint32_t so) {
    return __builtin_bit_cast(__half, __builtin_amdgcn_raw_buffer_load_b16(b.r, (int)vo, (int)so, CP));
  }
  template <int CP = 0> __device__ static __forceinline__ void st(const Buf& b, uint32_t vo, uint32_t so, __half v) {
    __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, v), b.r, (int)vo,
                                          (int)so, CP);
  }
};

// Element idx of sub-array i (each of N elements) of a buffer holding S.
template <typename T, typename S, int CP = 0>
__device__ __forceinline__ T bget(const Buf& b, uint32_t N, int i, uint32_t idx) {
  return Stor<S>::template to<T>(BufIO<S>::template ld<CP>(
      b, idx * (uint32_t)sizeof(S), (uint32_t)i * N * (uint32_t)sizeof(S)));
}
template <typename T, typename S, int CP = 0>
__device__ __forceinline__ void bput(const Buf& b, uint32_t N, int i, uint32_t idx, T v) {
  BufIO<S>::template st<CP>(b, idx * (uint32_t)sizeof(S), (uint32_t)i * N * (uint32_t)sizeof(S),
                            Stor<S>::template from<T>(v));
}

// Distribution layout (AoSoA): nodes are grouped in blocks of kAB consecutive
// nodes; a block stores its Q directions one after another, kAB values each,
// so element (i, x) sits at ((x / kAB) * Q + i) * kAB + x % kAB.  A wave's
// node-local stores of all directions fill one contiguous region instead of
// Q regions N elements apart, and pulls read 64-value runs of one direction.
// Buffers hold padded_nodes(N) nodes.
#ifndef LBM_AOSOA_BLOCK
#define LBM_AOSOA_BLOCK 256
#endif
constexpr int kAB = LBM_AOSOA_BLOCK;
static_assert(kAB >= 64 && (kAB & (kAB - 1)) == 0, "AoSoA block: a power of two >= 64");
constexpr int kABsh = __builtin_ctz(kAB);
__host__ __device__ constexpr long long padded_nodes(long long N) {
  return (N + (kAB - 1)) & ~(long long)(kAB - 1);
}

template <int Q, typename S> __device__ __forceinline__ uint32_t dvoff(uint32_t idx) {
  // = ((idx / kAB) * Q * kAB + idx % kAB) * sizeof(S), with a full-rate 24-bit
  // multiply (idx / kAB < 2^24; the product stays below 2^32 because a
  // distribution set is smaller than 4 GiB)
  return __umul24(idx >> kABsh, (uint32_t)((Q - 1) * kAB * sizeof(S))) + idx * (uint32_t)sizeof(S);
}
// load of direction i at the node whose dvoff is vo
template <int Q, typename T, typename S, int CP = 0>
__device__ __forceinline__ T dload(const Buf& b, int i, uint32_t vo) {
  return Stor<S>::template to<T>(
      BufIO<S>::template ld<CP>(b, vo, (uint32_t)i * (uint32_t)kAB * (uint32_t)sizeof(S)));
}
template <int Q, typename T, typename S>
__device__ __forceinline__ T dget(const Buf& b, int i, uint32_t idx) {
  return Stor<S>::template to<T>(
      BufIO<S>::ld(b, dvoff<Q, S>(idx), (uint32_t)i * (uint32_t)kAB * (uint32_t)sizeof(S)));
}
// store of direction i at the node whose dvoff is vo
template <int Q, typename T, typename S, int CP = 0>
__device__ __forceinline__ void dput(const Buf& b, int i, uint32_t vo, T v) {
  BufIO<S>::template st<CP>(b, vo, (uint32_t)i * (uint32_t)kAB * (uint32_t)sizeof(S),
                            Stor<S>::template from<T>(v));
}

// Correctly rounded fp32 division x / d by a divisor whose fp64 reciprocal
// y = RN64(1/d) is known, without a branch or a range test (fp32 and fp64
// FMAs issue at the same rate on gfx950 when packed fp32 math is off).
//
// divk: RN32(RN64(x y)).  x y is within 2^-52 (relative) of x / d.  Write
// x = a 2^ex, d = b 2^ed (a, b integers below 2^24).  Unless x / d lies
// exactly on a rounding boundary m of fp32, x - m d is a nonzero multiple of
// 2^min(ex, ed - 150), so x / d is at least 2^-49 (relative) from every
// boundary of a normal result and at least min(2^(ex-ed-24), 2^-174) from
// every boundary of a subnormal one -- both beyond the 2^-52 -- and rounding
// x y to fp32 gives RN32(x / d).  x / d = m exactly needs m d to be an fp32
// value with m a midpoint (odd multiple of 2^-150 or 25 significant bits),
// which happens only for d an even integer.  So divk is exact for every fp32
// x (zeros, subnormals and infinities included) whenever d is not an even
// integer: cs2, 2 cs2 and rho_eff cs2 below 2.
// divk_any: any d > 0.  Markstein's correction in fp64 (q = RN64(x y)
// faithful, residual q d - x exact by FMA, no under/overflow for fp32
// operands) gives RN64(x / d); rounding that to fp32 equals RN32(x / d)
// (double rounding through 53 >= 2*24 + 2 bits is innocuous for division).
// The residual is formed as q d - x and subtracted, so that x = -0 keeps its
// sign (-0 + -0); infinite x takes q itself (the residual is inf - inf).
// Both are checked against IEEE division over all 2^32 inputs by
// lbm_check_division.  fp64 runs keep the IEEE division.
__device__ __forceinline__ float divk(float x, float, double y) {
  return (float)((double)x * y);
}
__device__ __forceinline__ float divk_any(float x, float d, double y) {
  const double xd = (double)x, dd = (double)d;
  const double q = xd * y;
  const double e = __builtin_fma(q, dd, -xd);
  const double r = __builtin_fma(-e, y, q);
  return (float)(__builtin_isinf(xd) ? q : r);
}
__device__ __forceinline__ double divk(double x, double d, double) { return x / d; }
__device__ __forceinline__ double divk_any(double x, double d, double) { return x / d; }
// RN64(1/d) for the divk forms (unused by fp64 runs)
__device__ __forceinline__ double rcp64(float d) { return 1.0 / (double)d; }
__device__ __forceinline__ double rcp64(double) { return 0.0; }

template <typename T> __device__ __forceinline__ T clamp01(T v) {
  return v < T(0) ? T(0) : (v > T(1) ? T(1) : v);
}
template <typename T> __device__ __forceinline__ T clamp_lo(T v, T lo) {
  return v < lo ? lo : v;
}

template <int Q, typename T> __device__ __forceinline__ T qsum(const T (&a)[Q]) {
  T tot = a[0];
#pragma unroll
  for (int i = 1; i < (Q < 16 ? Q : 16); ++i) tot = tot + a[i];
  if constexpr (Q > 16) {
    T r = a[16];
#pragma unroll
    for (int i = 17; i < Q; ++i) r = r + a[i];
    tot = tot + r;
  }
  return tot;
}

// e_i . v, components summed in index order; zero components skipped (the
// skipped terms are exact signed zeros in the reference).
template <int DIM, typename T> __device__ __forceinline__ T edot(int i, const T* v) {
  T acc = T(0);
  bool first = true;
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const int e = E<DIM>(i, d);
    if (e != 0) {
      T t = e > 0 ? v[d] : -v[d];
      acc = first ? t : acc + t;
      first = false;
    }
  }
  return acc;
}

template <int DIM, typename T> __device__ __forceinline__ T vsq(const T* v) {
  T s = v[0] * v[0];
#pragma unroll
  for (int d = 1; d < DIM; ++d) s = s + v[d] * v[d];
  return s;
}

// ---------------------------------------------------------------------------
// XCD-aware block order.  Workgroups are dispatched round-robin over the 8 XCDs
// (block b runs on XCD b % 8) and every XCD has its own L2, so consecutive
// blocks -- which share the cache lines of the shifted pulls and the phi
// stencil -- would land in different L2s.  Remap so that each XCD works on a
// contiguous range of nodes: XCD j gets logical blocks [start_j, start_j+n_j).
// ---------------------------------------------------------------------------
constexpr unsigned int kXCD = 8;
__device__ __forceinline__ unsigned int xcd_block(unsigned int b, unsigned int nb) {
  const unsigned int j = b % kXCD, i = b / kXCD;
  const unsigned int chunk = nb / kXCD, rem = nb % kXCD;
  return j * chunk + (j < rem ? j : rem) + i;
}
// The same ranges, each walked from its end.
__device__ __forceinline__ unsigned int xcd_block_rev(unsigned int b, unsigned int nb) {
  const unsigned int j = b % kXCD, i = b / kXCD;
  const unsigned int chunk = nb / kXCD, rem = nb % kXCD;
  const unsigned int nj = chunk + (j < rem ? 1u : 0u);
  return j * chunk + (j < rem ? j : rem) + (nj - 1u - i);
}

// ---------------------------------------------------------------------------
// Neighbour indices of one node: i[k] = index of x + e_k (periodic), i[0] = x.
// The pull source of direction k is x - e_k = i[OPP(k)]; the phi stencil reads
// i[k].  Built once per node from the six wrapped axis offsets.
// ---------------------------------------------------------------------------
template <int DIM> struct Nbrs {
  uint32_t i[Lat<DIM>::Q];
};

template <int DIM>
__device__ __forceinline__ Nbrs<DIM> neighbours(const Grid& g, uint32_t x, const Coord<DIM>& c) {
  constexpr int Q = Lat<DIM>::Q;
  const AxisOff<DIM> o = axis_offsets<DIM>(g, c);
  Nbrs<DIM> n;
  n.i[0] = x;
#pragma unroll
  for (int k = 1; k < Q; ++k) n.i[k] = (uint32_t)((int)x + dir_offset<DIM>(o, k, 1));
  return n;
}

// ---------------------------------------------------------------------------
// Pull of the distribution sets at node x: f from F and, with WITH_G, g from G
// (same layout, so both share the per-direction source offsets).
//
// DEV (16-bit f storage): slot (i, y) holds f*_i(y) - w_i phi_ref(y + e_i),
// phi_ref being the start-of-step phi of the step that wrote it, so a fluid
// pull adds back w_i phi_ref(x) of the pulling node itself; bounce-back and
// penalization slots (i, x) read at solid/partial x add w_i phi_ref(x - e_i).
// g is stored as is.
// ---------------------------------------------------------------------------
// LOCAL: the buffers hold post-stream values (first step after init_fields)
// and are read in place.  It is a template parameter so that the common form
// has no branch between the geometry-byte load and the pulls.
// BB = false: the caller discards the values of solid nodes (pass A sets
// their phi to 0), so their bounce-back loads are skipped.
// spec = false (wave-uniform): every node of the wave is solid, so the
// speculative fluid pulls are not issued.
template <typename T, typename S, int DIM, bool DEV, bool WITH_G, bool LOCAL, int CP = 0,
          bool BB = true>
__device__ __forceinline__ void pull(const Buf& F, const Buf& G, uint32_t x, const Nbrs<DIM>& nb,
                                     uint8_t m, const T* __restrict__ eps_arr,
                                     const Buf& pref, const Phys<T>& ph, T (&f)[Lat<DIM>::Q],
                                     T (&gg)[Lat<DIM>::Q], bool spec = true) {
  constexpr int Q = Lat<DIM>::Q;
  const uint32_t vx = dvoff<Q, S>(x);
  if constexpr (LOCAL) {
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      f[i] = dload<Q, T, S>(F, i, vx);
      if constexpr (WITH_G) gg[i] = dload<Q, T, S>(G, i, vx);
    }
    if constexpr (DEV) {
      const T r = bget<T, T>(pref, 0, 0, x);
#pragma unroll
      for (int i = 0; i < Q; ++i) f[i] = f[i] + ph.template w<DIM>(i) * r;
    }
    return;
  }
  // Fluid pulls are issued for every lane before the node's geometry byte is
  // known (one memory latency instead of two in series); solid lanes then
  // take bounce-back values instead.  The choice is a select after the
  // branch, so the compiler cannot sink the fluid loads behind the branch.
  if (spec) {
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const uint32_t vs = dvoff<Q, S>(nb.i[OPP<DIM>(i)]);
      f[i] = dload<Q, T, S, CP>(F, i, vs);
      if constexpr (WITH_G) gg[i] = dload<Q, T, S, CP>(G, i, vs);
    }
  } else {
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      f[i] = T(0);
      if constexpr (WITH_G) gg[i] = T(0);
    }
  }
  // nothing that reads m may be scheduled above the loads
  __builtin_amdgcn_sched_barrier(0);
  const bool solid = (m & M_SOLID) != 0;
  if (BB && solid) {
    T bf[Lat<DIM>::Q], bg[Lat<DIM>::Q];
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      bf[i] = dload<Q, T, S>(F, OPP<DIM>(i), vx);
      if constexpr (WITH_G) bg[i] = dload<Q, T, S>(G, OPP<DIM>(i), vx);
      if constexpr (DEV) bf[i] = bf[i] + ph.template w<DIM>(i) * bget<T, T>(pref, 0, 0, nb.i[OPP<DIM>(i)]);
    }
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      f[i] = bf[i];
      if constexpr (WITH_G) gg[i] = bg[i];
    }
  }
  if (solid) return;
  if constexpr (DEV) {
    const T r = bget<T, T>(pref, 0, 0, x);
#pragma unroll
    for (int i = 0; i < Q; ++i) f[i] = f[i] + ph.template w<DIM>(i) * r;
  }
  if (m & M_PARTIAL) {
    const T eps = eps_arr[x];
    const T a = T(1) - eps;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      T o = dload<Q, T, S>(F, OPP<DIM>(i), vx);
      if constexpr (DEV) o = o + ph.template w<DIM>(i) * bget<T, T>(pref, 0, 0, nb.i[OPP<DIM>(i)]);
      f[i] = a * f[i] + eps * o;
      if constexpr (WITH_G) {
        const T og = dload<Q, T, S>(G, OPP<DIM>(i), vx);
        gg[i] = a * gg[i] + eps * og;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// bf16 g storage.  u is carried by the odd part of g (u = sum e_i g_i plus the
// force term), which bf16's 8-bit mantissa rounds away, so a bf16 g slot
// written at node y for direction i holds g*_i(y) - r_i(y), with the odd
// velocity term of g_eq, r_i(y) = (3 w_i) (e_i . u(y)), formed from y's new
// velocity (e_i . u summed over the nonzero components in index order).  A
// reader adds back the same value, formed from the velocity of the node that
// wrote the slot: the pull source for fluid nodes, the node itself for
// bounce-back and penalization slots and in the post-stream ("local") form.
// fp16 (10-bit mantissa) stores g as is.
// ---------------------------------------------------------------------------
template <typename S> constexpr bool kGRef = std::is_same<S, __hip_bfloat16>::value;
#ifndef LBM_GREF_EARLY
#define LBM_GREF_EARLY 0
#endif

// r_i(y) from the velocity field bU (DIM arrays of N32 values).
template <typename T, int DIM>
__device__ __forceinline__ T gref(const Phys<T>& ph, int i, const Buf& bU, uint32_t N32,
                                  uint32_t y) {
  T v[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) v[d] = E<DIM>(i, d) != 0 ? bget<T, T>(bU, N32, d, y) : T(0);
  return ph.template w3<DIM>(i) * edot<DIM>(i, v);
}

// Pull of the g set at node x (the pull() rules), with the bf16 references
// added back when GREF.  bU: the velocity the slots were written with.
template <typename T, typename S, int DIM, bool GREF, bool LOCAL, int CP = 0>
__device__ __forceinline__ void pull_g(const Buf& G, const Buf& bU, uint32_t N32, uint32_t x,
                                       const Nbrs<DIM>& nb, uint8_t m,
                                       const T* __restrict__ eps_arr, const Phys<T>& ph,
                                       T (&g)[Lat<DIM>::Q], bool spec = true) {
  constexpr int Q = Lat<DIM>::Q;
  if constexpr (!GREF) {
    T unused[Q];
    pull<T, S, DIM, false, false, LOCAL, CP>(G, G, x, nb, m, eps_arr, G, ph, g, unused, spec);
  } else {
    const uint32_t vx = dvoff<Q, S>(x);
    if constexpr (LOCAL) {
#pragma unroll
      for (int i = 0; i < Q; ++i) g[i] = dload<Q, T, S>(G, i, vx) + gref<T, DIM>(ph, i, bU, N32, x);
      return;
    }
    if (spec) {
#pragma unroll
      for (int i = 0; i < Q; ++i) g[i] = dload<Q, T, S, CP>(G, i, dvoff<Q, S>(nb.i[OPP<DIM>(i)]));
    } else {
#pragma unroll
      for (int i = 0; i < Q; ++i) g[i] = T(0);
    }
#if LBM_GREF_EARLY
    // the sources' velocities are loaded together with the g slots (one
    // memory latency instead of two in series)
    T r[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i) r[i] = gref<T, DIM>(ph, i, bU, N32, nb.i[OPP<DIM>(i)]);
#endif
    __builtin_amdgcn_sched_barrier(0);
    if (m & M_SOLID) {
#pragma unroll
      for (int i = 0; i < Q; ++i)
        g[i] = dload<Q, T, S>(G, OPP<DIM>(i), vx) + gref<T, DIM>(ph, OPP<DIM>(i), bU, N32, x);
      return;
    }
#if LBM_GREF_EARLY
#pragma unroll
    for (int i = 0; i < Q; ++i) g[i] = g[i] + r[i];
#else
#pragma unroll
    for (int i = 0; i < Q; ++i) g[i] = g[i] + gref<T, DIM>(ph, i, bU, N32, nb.i[OPP<DIM>(i)]);
#endif
    if (m & M_PARTIAL) {
      const T eps = eps_arr[x];
      const T a = T(1) - eps;
#pragma unroll
      for (int i = 0; i < Q; ++i) {
        const T og = dload<Q, T, S>(G, OPP<DIM>(i), vx) + gref<T, DIM>(ph, OPP<DIM>(i), bU, N32, x);
        g[i] = a * g[i] + eps * og;
      }
    }
  }
}

template <int DIM, typename S> __device__ __forceinline__ Buf dist_buf(const S* F, long long N) {
  return Buf(F, (long long)Lat<DIM>::Q * padded_nodes(N) * (long long)sizeof(S));
}
template <typename T> __device__ __forceinline__ Buf field_buf(const T* p, long long N, int comps = 1) {
  return Buf(p, (long long)comps * N * (long long)sizeof(T));
}

// phi at the next-interior node of a top boundary node x (start-of-step phi of
// the previous step), or -1 where the cubic target does not apply; read only
// under direct wetting.
template <typename T>
__device__ __forceinline__ T nsw_phi(const Phys<T>& ph, uint8_t m, const T* __restrict__ phi_prev,
                                     uint32_t x) {
  if (!ph.direct || !(m & M_TOP)) return T(-1);
  const int32_t q = ph.nsw[x];
  return q >= 0 ? phi_prev[q] : T(-1);
}

// S12 target of a top boundary node whose mass-corrected phi is pc: the wall
// phi of the contact angle, or under direct wetting the cubic wall-energy phi
// a phi_b^2 + (1 - a) phi_b = phi_nsw (lbm/fe_wetting.py:369-416) where the
// next-interior node is fluid and 0.01 < pc < 0.99, else pc itself.
// pns: raw phi of the next-interior node (-1: none), cfac: the mass factor.
template <typename T>
__device__ __forceinline__ T s12_target(const Phys<T>& ph, T pc, T pns, T cfac) {
  if (!ph.direct_on) return ph.s12_t;
  if (ph.direct && pns >= T(0) && pc > ph.lo01 && pc < ph.hi99) {
    const T pn = clamp01(pns * cfac);
    T disc = ph.cub_1ma2 + ph.cub_4a * pn;
    disc = disc < T(0) ? T(0) : disc;
    return clamp01((ph.cub_m1a + (T)sqrt(disc)) / ph.cub_2a);
  }
  return pc;
}

// Relaxed phi of a top boundary node (the S12 update), pc mass-corrected.
template <typename T>
__device__ __forceinline__ T s12_phi(const Phys<T>& ph, uint8_t m, T pc, T pns, T cfac) {
  if (ph.s12_on && (m & M_TOP) && pc > ph.thr_lo && pc < ph.thr_hi)
    return clamp01(pc + ph.s12_r * (s12_target<T>(ph, pc, pns, cfac) - pc));
  return pc;
}

// Mass factor and boundary relax of the previous step applied to a pulled f
// (lbm/fe_ac_solver.py:582-591, 594-636).  phi_prev is that step's
// start-of-step phi at the node, pns the one at its next-interior node (see
// nsw_phi); the relaxed phi is written to *phi_relaxed when given.
template <typename T, int DIM>
__device__ __forceinline__ void scale_f(const Phys<T>& ph, T (&f)[Lat<DIM>::Q], T cfac,
                                        uint8_t m, T phi_prev, T* phi_relaxed, T pns = T(-1)) {
  constexpr int Q = Lat<DIM>::Q;
#pragma unroll
  for (int i = 0; i < Q; ++i) f[i] = f[i] * cfac;
  if (ph.s12_on && (m & M_TOP)) {
    const T pc = clamp01(phi_prev * cfac);
    if (pc > ph.thr_lo && pc < ph.thr_hi) {
      const T pn = clamp01(pc + ph.s12_r * (s12_target<T>(ph, pc, pns, cfac) - pc));
      const T tot = clamp_lo(qsum<Q>(f), T(1e-10));
      T s = pn / tot;
      s = s < T(0.5) ? T(0.5) : (s > T(2) ? T(2) : s);
#pragma unroll
      for (int i = 0; i < Q; ++i) f[i] = f[i] * s;
      if (phi_relaxed) *phi_relaxed = pn;
    }
  }
}

// ---------------------------------------------------------------------------
// Deterministic block / grid reductions (fp64).
// ---------------------------------------------------------------------------
template <int BS> __device__ __forceinline__ double block_sum(double v, double* lds) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) lds[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < BS / 64; ++i) r += lds[i];
  }
  __syncthreads();
  return r;  // valid in thread 0
}

// Sum of the 64 nodes of a wave, written to partials[first node / 64]: the
// reduction tree depends only on node indices, not on block size or order.
// Waves that lie wholly past the last node write nothing.
__device__ __forceinline__ void wave_partial(double v, uint32_t x, long long N,
                                             double* __restrict__ partials) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  if ((threadIdx.x & 63) == 0 && (long long)x < N) partials[x >> 6] = v;
}

// Fixed-order sum of n wave partials by the first 128 threads of a block
// (valid in thread 0); every thread of the block must call it.
__device__ __forceinline__ double sum_partials(const double* __restrict__ partials,
                                               unsigned int n, double* lds128) {
  if (threadIdx.x < 128) {
    double a0 = 0.0, a1 = 0.0;
    unsigned int i = threadIdx.x;
    for (; i + 128 < n; i += 256) {
      a0 += partials[i];
      a1 += partials[i + 128];
    }
    if (i < n) a0 += partials[i];
    lds128[threadIdx.x] = a0 + a1;
  }
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < 128; ++i) r += lds128[i];
  __syncthreads();
  return r;
}

// ---------------------------------------------------------------------------
// init_fields: phi = clamp(phi_init), u, f = f_eq(phi, u), g = g_eq(0, u),
// M0 = sum(phi) in fp64 (lbm/fe_ac_solver.py:239-261, 346-367).
// ---------------------------------------------------------------------------
// (e_i . v) / cs2 for every direction.  Opposite directions are exact
// negatives of each other (same terms, negated, summed in the same order, and
// round-to-nearest is symmetric), so only one of each pair is divided.
template <typename T, int DIM>
__device__ __forceinline__ void ediv(const Phys<T>& ph, const T* v, T d, double r,
                                     T (&out)[Lat<DIM>::Q]) {
  constexpr int Q = Lat<DIM>::Q;
  out[0] = T(0);
#pragma unroll
  for (int i = 1; i < Q; ++i) {
    if (OPP<DIM>(i) > i) out[i] = divk(edot<DIM>(i, v), d, r);
  }
#pragma unroll
  for (int i = 1; i < Q; ++i) {
    if (OPP<DIM>(i) < i) out[i] = -out[OPP<DIM>(i)];
  }
}

// f_eq and g_eq (lbm/fe_ac_solver.py:346-367) from ec = (e.u)/cs2 and
// usq_t = |u|^2 / (2 cs2).
template <typename T, int DIM>
__device__ __forceinline__ T feq_of(const Phys<T>& ph, int i, T phi, T ec, T usq_t) {
  const T sq = (ec * ec) / T(2);
  return phi * (ph.template w<DIM>(i) * (((T(1) + ec) + sq) - usq_t));
}
template <typename T, int DIM>
__device__ __forceinline__ T geq_of(const Phys<T>& ph, int i, T P, T ec, T usq_t) {
  const T sq = (ec * ec) / T(2);
  return P * ph.template w<DIM>(i) + ph.template w<DIM>(i) * ((ec + sq) - usq_t);
}

template <typename T, typename S, int DIM, int BS>
__global__ __launch_bounds__(BS) void k_init(Grid g, Phys<T> ph, const T* __restrict__ phi_in,
                                             const T* __restrict__ u_in, T* __restrict__ phi,
                                             T* __restrict__ u, S* __restrict__ F,
                                             S* __restrict__ G, double* __restrict__ partials) {
  constexpr bool DEV = !std::is_same<T, S>::value;
  constexpr int Q = Lat<DIM>::Q;
  const long long N = g.N;
  const uint32_t x = blockIdx.x * BS + threadIdx.x;
  double v = 0.0;
  if (x < N) {
    // phi_in, u_in in the public (last axis fastest) order
    const uint32_t e = ext_index<DIM>(g, decompose<DIM>(g, x));
    const T p = clamp01(phi_in[e]);
    T uu[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) uu[d] = u_in ? u_in[(long long)d * N + e] : T(0);
    phi[x] = p;
#pragma unroll
    for (int d = 0; d < DIM; ++d) u[(long long)d * N + x] = uu[d];
    const T usq_t = divk(vsq<DIM>(uu), ph.two_cs2, ph.rcp_two_cs2);
    T ec[Q];
    ediv<T, DIM>(ph, uu, ph.cs2, ph.rcp_cs2, ec);
    const Buf bF = dist_buf<DIM>(F, N), bG = dist_buf<DIM>(G, N);
    const uint32_t vx = dvoff<Q, S>(x);
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      T fe = feq_of<T, DIM>(ph, i, p, ec[i], usq_t);
      T ge = geq_of<T, DIM>(ph, i, T(0), ec[i], usq_t);
      if constexpr (DEV) fe = fe - ph.template w<DIM>(i) * p;
      if constexpr (kGRef<S>) ge = ge - ph.template w3<DIM>(i) * edot<DIM>(i, uu);
      dput<Q, T, S>(bF, i, vx, fe);
      dput<Q, T, S>(bG, i, vx, ge);
    }
    v = (double)p;
  }
  wave_partial(v, x, N, partials);
}

// One block: fixed-order sum of n wave partials -> *out (M0 after k_init).
template <int BS>
__global__ __launch_bounds__(BS) void k_sum_partials(const double* __restrict__ partials,
                                                     unsigned int n, double* out) {
  __shared__ double lds[128];
  const double t = sum_partials(partials, n, lds);
  if (threadIdx.x == 0) *out = t;
}

// ---------------------------------------------------------------------------
// Pass A: start-of-step phi and its fp64 sum -> mass factor of this step.
// ---------------------------------------------------------------------------
template <typename T, typename S, int DIM, bool LOCAL>
__device__ __forceinline__ T phi_node(const Grid& g, const Phys<T>& ph, const S* __restrict__ Fin,
                                      const T* __restrict__ phi_prev,
                                      const uint8_t* __restrict__ meta,
                                      const T* __restrict__ eps, T cprev, uint32_t x) {
  constexpr int Q = Lat<DIM>::Q;
  const long long N = g.N;
  const uint8_t m = meta[x];
  const Coord<DIM> c = decompose<DIM>(g, x);
  const Nbrs<DIM> nb = neighbours<DIM>(g, x, c);
  T f[Q], unused[Q];
  const Buf bF = dist_buf<DIM>(Fin, N), bP = field_buf(phi_prev, N);
  // a solid node's phi is 0 whatever its distributions: no bounce-back loads
  pull<T, S, DIM, !std::is_same<T, S>::value, false, LOCAL, 0, false>(bF, bF, x, nb, m, eps, bP,
                                                                      ph, f, unused);
  if constexpr (!LOCAL)
    scale_f<T, DIM>(ph, f, cprev, m, (m & M_TOP) ? phi_prev[x] : T(0), nullptr,
                    nsw_phi<T>(ph, m, phi_prev, x));
  T p = clamp01(qsum<Q>(f));
  if (m & M_SOLID) p = T(0);
  return p;
}

// LOCAL (first step after init_fields) is a kernel template parameter: the
// host launches the matching instantiation, so no branch sits between the
// geometry-byte load and the pulls.
template <typename T, typename S, int DIM, int BS, bool LOCAL>
__global__ __launch_bounds__(BS) void k_phi(Grid g, Phys<T> ph, const S* __restrict__ Fin,
                                            const T* __restrict__ phi_prev,
                                            T* __restrict__ phi_out,
                                            const uint8_t* __restrict__ meta,
                                            const T* __restrict__ eps,
                                            double* __restrict__ partials, const Ctrl* ctrl,
                                            int slot_prev) {
  const long long N = g.N;
  // each XCD walks its range from the end: the distributions pass B wrote
  // last are still in the memory-side cache and are read first (measured on
  // the 150x150x107 case: 50.5 -> 44.5 us)
  const unsigned int lb = xcd_block_rev(blockIdx.x, gridDim.x);
  const uint32_t x = lb * BS + threadIdx.x;
  double v = 0.0;
  if (x < N) {
    const T cf = LOCAL ? T(1) : (T)ctrl->c[slot_prev];
    const T p = phi_node<T, S, DIM, LOCAL>(g, ph, Fin, phi_prev, meta, eps, cf, x);
    at(phi_out, x) = p;
    v = (double)p;
  }
  // wave partials of sum(phi); pass B's first block adds them up (after this
  // kernel has finished, so no inter-block fence is needed here)
  wave_partial(v, x, N, partials);
}

// Mass factor of a step from pass A's partials (fixed order): called by every
// thread of one block of pass B.
template <int BS>
__device__ __forceinline__ void finalize_mass(const double* __restrict__ partials,
                                              unsigned int n, Ctrl* ctrl, int slot) {
  __shared__ double lds[128];
  const double m = sum_partials(partials, n, lds);
  if (threadIdx.x == 0) {
    ctrl->mass[slot] = m;
    const double M0 = ctrl->M0;
    ctrl->c[slot] = m > 1e-6 ? M0 / m : 1.0;
  }
}

// ---------------------------------------------------------------------------
// The 1-ring stencil with wetting ghosts (lbm/fe_ac_solver.py:296-320,
// lbm/fe_wetting.py:418-446).  init_lap selects the stand-alone Laplacian of
// init_fields (lbm/fe_ac_solver.py:282-294).
// ---------------------------------------------------------------------------
// Sources of the start-of-step phi for the stencil: the phi array of every
// node, or (fused step) an LDS ring of phi planes around the block's tile.
template <typename T> struct PhiGlobal {
  const T* __restrict__ p;
  Buf b;
  __device__ __forceinline__ PhiGlobal(const T* ptr, long long N) : p(ptr), b(field_buf(ptr, N)) {}
  __device__ __forceinline__ T self(uint32_t x) const { return at(p, x); }
  template <int DIM> __device__ __forceinline__ T nbr(const Nbrs<DIM>& nb, int i) const {
    return bget<T, T>(b, 0, 0, nb.i[i]);
  }
};
// lm, l0, lp: the node's position in the ring planes z-1, z, z+1; sy: the
// ring's row stride (one step along x; one step along y is 1).  3D only.
template <typename T> struct PhiRing {
  const T* lm;
  const T* l0;
  const T* lp;
  int sy;
  __device__ __forceinline__ T self(uint32_t) const { return l0[0]; }
  template <int DIM> __device__ __forceinline__ T nbr(const Nbrs<DIM>&, int i) const {
    static_assert(DIM == 3, "the LDS phi ring is 3D only");
    const int ez = E<3>(i, 2);
    const T* pl = ez > 0 ? lp : (ez < 0 ? lm : l0);
    return pl[E<3>(i, 0) * sy + E<3>(i, 1)];
  }
};

template <typename T, int DIM, typename PS>
__device__ __forceinline__ void stencil_src(const Phys<T>& ph, const PS& ps, const Nbrs<DIM>& nb,
                                            T p, uint32_t wm, const float* n32, T* grad, T* lap,
                                            bool init_lap) {
  constexpr int Q = Lat<DIM>::Q;
#pragma unroll
  for (int d = 0; d < DIM; ++d) grad[d] = T(0);
  T l = T(0);
#pragma unroll
  for (int i = 1; i < Q; ++i) {
    T psi = ps.template nbr<DIM>(nb, i);
    if (wm & (1u << i)) {
      // n . e_i accumulated in fp32 from zero (lbm/fe_wetting.py:237-239)
      float nde = 0.0f;
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        const int e = E<DIM>(i, d);
        if (e != 0) nde = nde + (e > 0 ? n32[d] : -n32[d]);
        else nde = nde + 0.0f * n32[d];
      }
      if (ph.wet == WET_LINEAR) {
        psi = p - (T)(nde * ph.hk32);
      } else {
        psi = p + ((T)(nde * ph.pck32) * p) * (T(1) - p);
      }
    }
    const T diff = psi - p;
    if (!init_lap) {
#pragma unroll
      for (int d = 0; d < DIM; ++d)
        if (E<DIM>(i, d) != 0) {
          const T t = ph.woc[wclass<DIM>(i)] * psi;
          grad[d] = E<DIM>(i, d) > 0 ? grad[d] + t : grad[d] - t;
        }
      l = l + ph.lc[wclass<DIM>(i)] * diff;
    } else {
      l = l + (ph.lic[wclass<DIM>(i)] * diff) / ph.cs2;
    }
  }
  *lap = l;
}

template <typename T, int DIM>
__device__ __forceinline__ void stencil(const Phys<T>& ph, const Grid& g,
                                        const T* __restrict__ phi, const Nbrs<DIM>& nb, T p,
                                        uint32_t wm, const float* n32, T* grad, T* lap,
                                        bool init_lap) {
  stencil_src<T, DIM>(ph, PhiGlobal<T>(phi, g.N), nb, p, wm, n32, grad, lap, init_lap);
}

template <typename T>
__device__ __forceinline__ T chem_pot(const Phys<T>& ph, T p, T lap, bool bnd) {
  T mu = (((ph.four_beta * p) * (p - T(1))) * (p - T(0.5))) - ph.kappa * lap;
  if (bnd) {
    if (ph.wet == WET_LINEAR) mu = mu + ph.mu_wall_add;
    else if (ph.wet == WET_QUADRATIC) mu = mu + (ph.phi_c * p) * (T(1) - p);
  }
  return mu;
}

// Stand-alone gradient (lbm/fe_ac_solver.py:267-280) of the chemical
// potential, as the step takes it: ((w_i e_id) psi) / cs2 accumulated in
// direction order, psi = mu(x + e_i) or the wetting ghost formula of phi
// evaluated on mu at wall directions.
template <typename T, int DIM>
__device__ __forceinline__ void grad_mu(const Phys<T>& ph, const T* __restrict__ mu_arr,
                                        const Nbrs<DIM>& nb, T mc, uint32_t wm, const float* n32,
                                        T* gmu) {
  constexpr int Q = Lat<DIM>::Q;
#pragma unroll
  for (int d = 0; d < DIM; ++d) gmu[d] = T(0);
#pragma unroll
  for (int i = 1; i < Q; ++i) {
    T psi = mu_arr[nb.i[i]];
    if (wm & (1u << i)) {
      float nde = 0.0f;
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        const int e = E<DIM>(i, d);
        if (e != 0) nde = nde + (e > 0 ? n32[d] : -n32[d]);
        else nde = nde + 0.0f * n32[d];
      }
      if (ph.wet == WET_LINEAR) psi = mc - (T)(nde * ph.hk32);
      else psi = mc + ((T)(nde * ph.pck32) * mc) * (T(1) - mc);
    }
    const T t = divk(ph.template w<DIM>(i) * psi, ph.cs2, ph.rcp_cs2);
#pragma unroll
    for (int d = 0; d < DIM; ++d)
      if (E<DIM>(i, d) != 0) gmu[d] = E<DIM>(i, d) > 0 ? gmu[d] + t : gmu[d] - t;
  }
}

// Chemical potential of every node from the start-of-step phi (the input of
// grad mu in the Cahn-Hilliard and mu-drive forms).
template <typename T, int DIM, int BS>
__global__ __launch_bounds__(BS) void k_mu(Grid g, Phys<T> ph, const T* __restrict__ phi,
                                           const uint8_t* __restrict__ meta,
                                           const uint32_t* __restrict__ wmask,
                                           const float* __restrict__ normal,
                                           T* __restrict__ mu) {
  const uint32_t x = blockIdx.x * BS + threadIdx.x;
  if (x >= g.N) return;
  const uint8_t m = meta[x];
  const bool bnd = (m & M_BOUNDARY) && ph.wet != WET_NONE;
  float n32[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) n32[d] = bnd ? normal[(long long)d * g.N + x] : 0.0f;
  T gr[DIM], lap;
  const T p = phi[x];
  stencil<T, DIM>(ph, g, phi, neighbours<DIM>(g, x, decompose<DIM>(g, x)), p,
                  bnd ? wmask[x] : 0u, n32, gr, &lap, false);
  mu[x] = chem_pot<T>(ph, p, lap, bnd);
}

// Geometric contact-angle correction at a wall-normal-carrying node
// (lbm/fe_wetting.py:448-539); blend weight is 1 there (distance 0).
template <typename T, int DIM>
__device__ __forceinline__ void correct_grad(const Phys<T>& ph, T* gr, const float* n32,
                                             T amp_node, bool amp_field) {
  T n[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) n[d] = (T)n32[d];
  T gn = n[0] * gr[0];
#pragma unroll
  for (int d = 1; d < DIM; ++d) gn = gn + n[d] * gr[d];
  T gt[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) gt[d] = gr[d] - gn * n[d];
  const T gtm = clamp_lo((T)sqrt(vsq<DIM>(gt)), T(1e-10));
  T gd = ((-gtm) * ph.cos_t) / ph.sin_t;
  if (amp_field) {
    if (amp_node > T(1)) gd = gn + (gd - gn) * amp_node;
  } else if (ph.amp_on) {
    gd = gn + (gd - gn) * ph.amp;
  }
  const T gm = (T)sqrt(vsq<DIM>(gr));
  if (gm > T(1e-6)) {
    const T blend = T(1);
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      const T gc = gt[d] + gd * n[d];
      gr[d] = gr[d] + blend * (gc - gr[d]);
    }
  }
}

// ---------------------------------------------------------------------------
// Pass B: everything from S2 to S8 at one node, then node-local stores.
// ---------------------------------------------------------------------------
#ifndef LBM_EARLY_F
#define LBM_EARLY_F 0
#endif

template <typename T, typename S, int DIM> struct CollideArgs {
  Grid g;
  const S* __restrict__ Fin;
  const S* __restrict__ Gin;
  S* __restrict__ Fout;
  S* __restrict__ Gout;
  const T* __restrict__ u_in;   // velocity of the last step (by step parity)
  T* __restrict__ u_out;        // velocity of this step (the other buffer)
  const T* __restrict__ phi_cur;
  const T* __restrict__ phi_prev;
  const uint8_t* __restrict__ meta;
  const uint32_t* __restrict__ wmask;
  const float* __restrict__ normal;
  const T* __restrict__ eps;
  const T* __restrict__ ampf;
  const Ctrl* ctrl;
  Ctrl* ctrl_w;              // same block; written only by block 0 (slot_new)
  const double* partials;    // pass A block partials of sum(phi)
  unsigned int nparts;
  int slot_prev;
  int slot_new;
  int local;
  int phiu_valid;
  const T* __restrict__ mu;  // chemical potential of every node (GMU kernels)
};

// FLUID: the caller guarantees a plain fluid node (no solid, partial, wall or
// top flag), so every wall path compiles away.  GMU: the interface source
// uses grad mu (Cahn-Hilliard or the mu drive), from the mu of every node.
// PS: source of the start-of-step phi (PhiGlobal or PhiRing).
// spec = false: every node of the wave is solid (a flagged chunk), so the
// speculative fluid pulls are skipped; the bounce-back collision is the same.
template <typename T, typename S, int DIM, bool LOCAL, bool FLUID, bool GMU, typename PS>
__device__ __forceinline__ void collide_node_src(const CollideArgs<T, S, DIM>& a,
                                                 const Phys<T>& ph, uint32_t x, const PS& ps,
                                                 bool spec = true) {
  constexpr int Q = Lat<DIM>::Q;
  constexpr bool DEV = !std::is_same<T, S>::value;
  const Grid& g = a.g;
  const long long N = g.N;
  const uint8_t m = FLUID ? (uint8_t)(a.meta[x] & 0xF0u) : a.meta[x];
  const bool solid = m & M_SOLID;
  const bool bnd = m & M_BOUNDARY;
  const Coord<DIM> c = decompose<DIM>(g, x);
  const Nbrs<DIM> nb = neighbours<DIM>(g, x, c);

  // g is pulled first and f only for the final f collision, after g* and u
  // are stored: f's 19 values are then never live together with g's, which
  // keeps the register footprint (and so the occupancy) of this pass down.
  // Memory-side cache (256 MB on MI355X): pass A reads back the f set this
  // pass writes, so f stores keep the default policy; the g set written here
  // is not read before the next pass B, and the source sets die with their
  // last read here, so those accesses are non-temporal and leave the cache to
  // the new f.  (For 16-bit storage, whose whole state nearly fits, the NT
  // f reads measured slower and stay default.)
  // For 4-byte storage the old velocity is also read non-temporally (its
  // last read; the new one is stored with the default policy): the cache
  // keeps more of f for pass A (fp32 step 173.9 -> 172.7 us; with the new
  // velocity stored non-temporally too, pass B lost what pass A gained).
  // Measured alternatives, fp32 step: g pulls with the default policy 186.9
  // us, f pulls with it 186.6, both 192.3.
  constexpr int CP_G = kNT, CP_FLD = sizeof(S) >= 4 ? kNT : 0, CP_UL = sizeof(S) >= 4 ? kNT : 0;
  T gg[Q], unused[Q];
  const uint32_t N32 = (uint32_t)N;
  const Buf bFi = dist_buf<DIM>(a.Fin, N), bGi = dist_buf<DIM>(a.Gin, N);
  const Buf bPp = field_buf(a.phi_prev, N);
  const Buf bU = field_buf(a.u_in, N, DIM), bUo = field_buf(a.u_out, N, DIM);
  // 16-bit storage (VALU-bound): a wave issuing its pulls runs at raised
  // priority, so its loads leave early while other waves compute (measured
  // 142.1 -> 140.6 us/step for bf16; no gain for 4-byte storage)
  if constexpr (DEV) __builtin_amdgcn_s_setprio(1);
  pull_g<T, S, DIM, kGRef<S>, LOCAL, CP_G>(bGi, bU, N32, x, nb, m, a.eps, ph, gg, spec);
  T f[Q];
#if LBM_EARLY_F
  // the f pulls leave together with the g pulls: their latency hides behind
  // the g collision (f and g then live together: more registers)
  pull<T, S, DIM, DEV, false, LOCAL, CP_FLD>(bFi, bFi, x, nb, m, a.eps, bPp, ph, f, unused);
#endif
  if constexpr (DEV) __builtin_amdgcn_s_setprio(0);
  T pprev = T(0);
  if (((m & M_TOP) && !LOCAL) || a.phiu_valid) pprev = a.phi_prev[x];

  // S1/S2: phi, rho, stencil, mu, correction
  const T p = ps.self(x);
  const T rho = p * ph.rho_l + (T(1) - p) * ph.rho_g;
  const T rho_eff = rho < ph.rho_floor ? ph.rho_floor : rho;
  float n32[DIM];
  uint32_t wm = 0;
  if (bnd && ph.wet != WET_NONE) {
    wm = a.wmask[x];
#pragma unroll
    for (int d = 0; d < DIM; ++d) n32[d] = a.normal[(long long)d * N + x];
  } else {
#pragma unroll
    for (int d = 0; d < DIM; ++d) n32[d] = 0.0f;
  }
  T gr[DIM], lap;
  stencil_src<T, DIM>(ph, ps, nb, p, wm, n32, gr, &lap, false);
  const T mu = chem_pot<T>(ph, p, lap, bnd && ph.wet != WET_NONE);
  T gmu[DIM];
  if constexpr (GMU) grad_mu<T, DIM>(ph, a.mu, nb, mu, wm, n32, gmu);
  if (ph.corr_on && bnd) {
    const T an = a.ampf ? a.ampf[x] : T(0);
    correct_grad<T, DIM>(ph, gr, n32, an, a.ampf != nullptr);
  }

  // S3/S4: P, stress with the old velocity
  const T P = qsum<Q>(gg);
  T uo[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) uo[d] = bget<T, T, CP_UL>(bU, N32, d, x);
  constexpr int NS = DIM == 3 ? 6 : 3;
  T sig[NS];
#pragma unroll
  for (int k = 0; k < NS; ++k) sig[k] = T(0);
  {
    const T usq_t = divk(vsq<DIM>(uo), ph.two_cs2, ph.rcp_two_cs2);
    T ec[Q];
    ediv<T, DIM>(ph, uo, ph.cs2, ph.rcp_cs2, ec);
    bool first[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) first[k] = true;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const T gn = gg[i] - geq_of<T, DIM>(ph, i, P, ec[i], usq_t);
      int k = 0;
#pragma unroll
      for (int aa = 0; aa < DIM; ++aa) {
#pragma unroll
        for (int bb = aa; bb < DIM; ++bb, ++k) {
          const int ee = E<DIM>(i, aa) * E<DIM>(i, bb);
          if (ee != 0) {
            const T t = ee > 0 ? gn : -gn;
            sig[k] = first[k] ? t : sig[k] + t;
            first[k] = false;
          }
        }
      }
    }
  }
  auto SIG = [&](int aa, int bb) -> T {
    if (aa > bb) { int t = aa; aa = bb; bb = t; }
    // index of (aa,bb) in the packed upper triangle
    int k = 0;
    for (int r = 0; r < aa; ++r) k += DIM - r;
    return sig[k + (bb - aa)];
  };

  // S5: forces
  T F[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const T Fs = mu * gr[d];
    const T Fp = (((-P) * ph.drho) * ph.cs2) * gr[d];
    T sg = SIG(d, 0) * gr[0];
#pragma unroll
    for (int b = 1; b < DIM; ++b) sg = sg + SIG(d, b) * gr[b];
    const T Fm = (ph.visc * sg) * ph.drho;
    const T Fb = ph.body_on ? ph.gforce[d] * rho : T(0);
    T t;
    if (ph.stab == 0) t = ((Fs + Fp) + Fm) + Fb;
    else if (ph.stab == 1) t = (Fs + Fp) + Fb;
    else t = Fs + Fb;
    F[d] = t;
  }

  // S6: velocity
  T un[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    T acc = T(0);
    bool first = true;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const int e = E<DIM>(i, d);
      if (e != 0) {
        const T t = e > 0 ? gg[i] : -gg[i];
        acc = first ? t : acc + t;
        first = false;
      }
    }
    un[d] = acc + (T(0.5) * F[d]) / rho_eff;
    if (solid) un[d] = T(0);
  }

  // S7: interface source
  T src[DIM];
  {
    T gm = (T)sqrt(vsq<DIM>(gr));
    gm = clamp_lo(gm, T(1e-10));
    const T lam = divk_any((T(4) * p) * (T(1) - p), ph.xi, ph.rcp_xi);
    const T sf = ph.sfac[dist_class(m)];
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      const T now = p * un[d];
      const T prev = a.phiu_valid ? pprev * uo[d] : now;
      const T dphiu = now - prev;
      const T nv = gr[d] / gm;
      T sh = ((ph.cs2 * lam) * nv) * ph.ac_scale;
      if (ph.wet != WET_NONE) sh = sh * sf;
      src[d] = dphiu + sh;
      if constexpr (GMU) {
        // lbm/fe_ac_solver.py:457-461 (Cahn-Hilliard) and :484-486 (mu drive)
        if (ph.gmu_mode == GMU_CH) src[d] = dphiu + (ph.mob * (gmu[d] - gr[d])) * ph.guo;
        else src[d] = src[d] + ph.mu_drive_a * gmu[d];
      }
    }
  }

  // S8: collisions
  T ns, omg;
  if (ph.var_tau) {
    const T eta = ((p * ph.rho_l) * ph.tl_m) + (((T(1) - p) * ph.rho_g) * ph.tg_m);
    const T rd = p * ph.rho_l + (T(1) - p) * ph.rho_g;
    const T teff = T(0.5) + eta / rd;
    ns = T(1) - T(0.5) / teff;
    omg = T(1) / teff;
  } else {
    ns = ph.ns_c;
    omg = ph.omg_c;
  }
  const Buf bFo = dist_buf<DIM>(a.Fout, N), bGo = dist_buf<DIM>(a.Gout, N);
  const T usq_n = divk(vsq<DIM>(un), ph.two_cs2, ph.rcp_two_cs2);
  const uint32_t vx = dvoff<Q, S>(x);
  T ec[Q];
  ediv<T, DIM>(ph, un, ph.cs2, ph.rcp_cs2, ec);
  // g collision with the Guo term (odd in e: opposite directions are exact
  // negatives), then g* and u are stored
  {
    const T rc = rho_eff * ph.cs2;
    const double rrc = rcp64(rc);
    T G[Q];
    G[0] = T(0);
    // rc < 2 (densities below 5.7, the usual case) cannot be an even integer
    if (ph.rc_small) {
#pragma unroll
      for (int i = 1; i < Q; ++i)
        if (OPP<DIM>(i) > i) G[i] = divk((ns * ph.template w<DIM>(i)) * edot<DIM>(i, F), rc, rrc);
    } else {
#pragma unroll
      for (int i = 1; i < Q; ++i)
        if (OPP<DIM>(i) > i)
          G[i] = divk_any((ns * ph.template w<DIM>(i)) * edot<DIM>(i, F), rc, rrc);
    }
#pragma unroll
    for (int i = 1; i < Q; ++i)
      if (OPP<DIM>(i) < i) G[i] = -G[OPP<DIM>(i)];
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const T ge = geq_of<T, DIM>(ph, i, P, ec[i], usq_n);
      T gs = (gg[i] - omg * (gg[i] - ge)) + G[i];
      if constexpr (kGRef<S>) gs = gs - ph.template w3<DIM>(i) * edot<DIM>(i, un);
      dput<Q, T, S, CP_G>(bGo, i, vx, gs);
    }
  }
#pragma unroll
  for (int d = 0; d < DIM; ++d) bput<T, T>(bUo, N32, d, x, un[d]);

  // f: pull, mass factor and boundary relax of the previous step, collision
  // with the interface source (odd in e as well)
  // the neighbour indices are rebuilt here rather than kept live from the
  // top (18 registers across the whole node); the empty asm hides x's
  // identity so the compiler cannot merge the two computations
#if LBM_EARLY_F
  const Nbrs<DIM>& nb2 = nb;
#else
  uint32_t x2 = x;
  asm volatile("" : "+v"(x2));
  const Nbrs<DIM> nb2 = neighbours<DIM>(g, x2, decompose<DIM>(g, x2));
  if constexpr (DEV) __builtin_amdgcn_s_setprio(1);
  pull<T, S, DIM, DEV, false, LOCAL, CP_FLD>(bFi, bFi, x, nb2, m, a.eps, bPp, ph, f, unused,
                                              spec);
  if constexpr (DEV) __builtin_amdgcn_s_setprio(0);
#endif
  if constexpr (!LOCAL)
    scale_f<T, DIM>(ph, f, (T)a.ctrl->c[a.slot_prev], m, pprev, nullptr,
                    nsw_phi<T>(ph, m, a.phi_prev, x));
  T fsrc[Q];
  fsrc[0] = T(0);
#pragma unroll
  for (int i = 1; i < Q; ++i)
    if (OPP<DIM>(i) > i)
      fsrc[i] = divk((ph.acoef * ph.template w<DIM>(i)) * edot<DIM>(i, src), ph.cs2, ph.rcp_cs2);
#pragma unroll
  for (int i = 1; i < Q; ++i)
    if (OPP<DIM>(i) < i) fsrc[i] = -fsrc[OPP<DIM>(i)];
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const T fe = feq_of<T, DIM>(ph, i, p, ec[i], usq_n);
    const T fs = (f[i] - ph.om_f * (f[i] - fe)) + fsrc[i];
    T fst = fs;
    if constexpr (DEV) fst = fs - ph.template w<DIM>(i) * ps.template nbr<DIM>(nb2, i);
    dput<Q, T, S>(bFo, i, vx, fst);
  }
}

template <typename T, typename S, int DIM, bool LOCAL, bool FLUID = false, bool GMU = false>
__device__ __forceinline__ void collide_node(const CollideArgs<T, S, DIM>& a,
                                             const Phys<T>& ph, uint32_t x, bool spec = true) {
  collide_node_src<T, S, DIM, LOCAL, FLUID, GMU>(a, ph, x, PhiGlobal<T>(a.phi_cur, a.g.N), spec);
}

// Occupancy target of pass B (waves per SIMD); a build-time tuning knob.
#ifdef LBM_COLLIDE_WPE
#define LBM_COLLIDE_ATTR __attribute__((amdgpu_waves_per_eu(LBM_COLLIDE_WPE)))
#else
#define LBM_COLLIDE_ATTR
#endif

template <typename T, typename S, int DIM, int BS, bool LOCAL, bool GMU = false>
__global__ __launch_bounds__(BS) LBM_COLLIDE_ATTR void k_collide(CollideArgs<T, S, DIM> a,
                                                                 Phys<T> ph) {
  if (blockIdx.x == 0) finalize_mass<BS>(a.partials, a.nparts, a.ctrl_w, a.slot_new);
  const uint32_t x = xcd_block(blockIdx.x, gridDim.x) * BS + threadIdx.x;
  if (x >= a.g.N) return;
  collide_node<T, S, DIM, LOCAL, false, GMU>(a, ph, x);
}

// Pass B over a list of 64-node chunks, one wave per chunk.  The chunks of a
// grid are split once at set-up into plain ones (every node plain fluid),
// run with FLUID = true -- no wall code, fewer registers, more resident
// waves -- and the rest, run with the general code.  `finalize`: this launch
// is the step's first pass-B launch and its block 0 forms the mass factor.
template <typename T, typename S, int DIM, int BS, bool LOCAL, bool FLUID>
__global__ __launch_bounds__(BS) LBM_COLLIDE_ATTR void k_collide_chunks(
    CollideArgs<T, S, DIM> a, Phys<T> ph, const uint32_t* __restrict__ chunks,
    unsigned int nchunks, int finalize) {
  if (finalize && blockIdx.x == 0) finalize_mass<BS>(a.partials, a.nparts, a.ctrl_w, a.slot_new);
  const unsigned int w = xcd_block(blockIdx.x, gridDim.x) * (BS / 64) + (threadIdx.x >> 6);
  if (w >= nchunks) return;
  // one entry per wave: a scalar
  const uint32_t c = __builtin_amdgcn_readfirstlane(chunks[w]);
  const uint32_t x = ((c & kChunkIdx) << 6) + (threadIdx.x & 63u);
  if (x >= a.g.N) return;
  // plain chunks are never all solid: the flag is a constant there
  collide_node<T, S, DIM, LOCAL, FLUID>(a, ph, x, FLUID || !(c & kChunkSolid));
}

// Pass B in one launch over a chunk list arranged on the host so that each
// XCD's range of logical blocks (xcd_block) holds its share of the plain
// chunks first and its share of the wall chunks (bit 31 set) last: the wall
// chunks, whose waves live as long as plain ones but run the general code,
// fill the tail of the launch instead of forming a launch of their own.
// 0xFFFFFFFF pads the list to whole blocks.  Block 0 forms the mass factor.
template <typename T, typename S, int DIM, int BS, bool LOCAL>
__global__ __launch_bounds__(BS) LBM_COLLIDE_ATTR void k_collide_mixed(
    CollideArgs<T, S, DIM> a, Phys<T> ph, const uint32_t* __restrict__ list,
    unsigned int nlist) {
  if (blockIdx.x == 0) finalize_mass<BS>(a.partials, a.nparts, a.ctrl_w, a.slot_new);
  const unsigned int w = xcd_block(blockIdx.x, gridDim.x) * (BS / 64) + (threadIdx.x >> 6);
  if (w >= nlist) return;
  // one entry per wave: a scalar, so the branch below is uniform
  const uint32_t c = __builtin_amdgcn_readfirstlane(list[w]);
  if (c == 0xFFFFFFFFu) return;
  const uint32_t x = ((c & kChunkIdx) << 6) + (threadIdx.x & 63u);
  if (x >= a.g.N) return;
  if (c & kChunkWall) collide_node<T, S, DIM, LOCAL, false>(a, ph, x, !(c & kChunkSolid));
  else collide_node<T, S, DIM, LOCAL, true>(a, ph, x);
}

// ---------------------------------------------------------------------------
// Fused step (3D, distributions stored in the compute type): pass A and pass
// B in one launch, without pass A's second read of the f set.  A block owns a
// tile of ty x tx columns (y fastest) and a segment of zlen planes and marches
// through them.  For every plane it first computes the start-of-step phi
// (phi_node, as pass A) of the tile and of a one-node ring around it into a
// four-plane LDS ring -- the ring nodes are the neighbouring tiles' own nodes,
// recomputed here from the same pulls, so no block ever waits for another --
// and then collides the tile's plane one behind, with the phi stencil read from
// the ring.  Segments alternate their marching direction, so two neighbouring
// segments reach the planes they share at the same time (both at their start
// or both at their end) and the second reads of those planes hit the cache.
// Each block adds up the phi of its own nodes in fp64 in a fixed order into
// bpart[block]; k_fused_mass forms the step's mass factor from those sums.
// ---------------------------------------------------------------------------
struct FusedTile {
  int ty, tx;        // tile extent along y (fastest) and x
  int nty, ntx;      // tiles along y and x
  int zlen;          // planes per z segment
  int sy;            // ring row stride: ty + 2
  int box;           // values per ring plane: (ty + 2) * (tx + 2) <= 2 * block size
  double* bpart;     // one fp64 sum of phi per block
};

template <typename T, int BS>
__global__ __launch_bounds__(BS) LBM_COLLIDE_ATTR void k_fused(CollideArgs<T, T, 3> a,
                                                               Phys<T> ph, FusedTile ft,
                                                               T* __restrict__ phi_out) {
  extern __shared__ __align__(16) unsigned char fused_lds[];
  T* ring = reinterpret_cast<T*>(fused_lds);
  const Grid& g = a.g;
  const unsigned int L = xcd_block(blockIdx.x, gridDim.x);
  const unsigned int ntile = (unsigned int)(ft.nty * ft.ntx);
  const int seg = (int)(L / ntile);
  const int r = (int)(L - (unsigned int)seg * ntile);
  const int txi = r / ft.nty, tyi = r - txi * ft.nty;
  const int y0 = tyi * ft.ty, x0 = txi * ft.tx;
  const int wy = min(ft.ty, g.ny - y0), wx = min(ft.tx, g.nx - x0);
  const int z0 = seg * ft.zlen, z1 = min(g.nz, z0 + ft.zlen);
  const int nown = z1 - z0;
  const int dir = (seg & 1) ? -1 : 1;
  const int zfirst = dir > 0 ? z0 : z1 - 1;
  const int sy = ft.sy, box = ft.box;
  const int tid = (int)threadIdx.x;

  // ring positions of this thread (at most two per plane): in-plane base
  // index, and flags 1 = inside the tile's box, 2 = a node of the tile itself
  uint32_t qb[2];
  int qf[2];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int q = tid + j * BS;
    const int qx = q / sy, qy = q - qx * sy;
    int yy = y0 - 1 + qy, xx = x0 - 1 + qx;
    yy = yy < 0 ? yy + g.ny : (yy >= g.ny ? yy - g.ny : yy);
    xx = xx < 0 ? xx + g.nx : (xx >= g.nx ? xx - g.nx : xx);
    const bool in = q < box && qy <= wy + 1 && qx <= wx + 1;
    const bool own = qy >= 1 && qy <= wy && qx >= 1 && qx <= wx;
    qb[j] = (uint32_t)(xx * g.sx + yy * g.sy);
    qf[j] = in ? (own ? 3 : 1) : 0;
  }
  // the node this thread collides: (cy, cx) of the tile
  const int cx = tid / wy, cy = tid - cx * wy;
  const bool cvalid = cx < wx;
  const uint32_t cb = (uint32_t)((x0 + cx) * g.sx + (y0 + cy) * g.sy);
  const int li = (cx + 1) * sy + (cy + 1);

  const T cprev = (T)a.ctrl->c[a.slot_prev];
  double acc = 0.0;
  // march step s holds plane zfirst + (s - 1) dir (steps 0 and nown + 1 are
  // the neighbouring segments' planes), in ring slot s & 3
  auto plane = [&](int s) {
    const int p = zfirst + (s - 1) * dir;
    return p < 0 ? p + g.nz : (p >= g.nz ? p - g.nz : p);
  };
  auto phi_plane = [&](int s) {
    const uint32_t zb = (uint32_t)plane(s) * (uint32_t)g.sz;
    const bool own = s >= 1 && s <= nown;
    T* slot = ring + (s & 3) * box;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      if (qf[j]) {
        const uint32_t x = qb[j] + zb;
        const T v = phi_node<T, T, 3, false>(g, ph, a.Fin, a.phi_prev, a.meta, a.eps, cprev, x);
        slot[tid + j * BS] = v;
        if (own && qf[j] == 3) {
          at(phi_out, x) = v;
          acc += (double)v;
        }
      }
    }
  };
  phi_plane(0);
  phi_plane(1);
  for (int s = 1; s <= nown; ++s) {
    phi_plane(s + 1);
    // slot (s + 1) & 3 is written before this barrier and slots s - 1 .. s + 1
    // are read after it; the next write, to slot (s + 2) & 3, is none of them
    __syncthreads();
    if (cvalid) {
      const uint32_t x = cb + (uint32_t)plane(s) * (uint32_t)g.sz;
      const PhiRing<T> pr{ring + ((s - dir) & 3) * box + li, ring + (s & 3) * box + li,
                          ring + ((s + dir) & 3) * box + li, sy};
      // one wave-uniform choice: the plain-fluid code when no node of the
      // wave carries a geometry flag
      const bool wall = __ballot((a.meta[x] & 0x0Fu) != 0) != 0ull;
      if (wall) collide_node_src<T, T, 3, false, false, false>(a, ph, x, pr);
      else collide_node_src<T, T, 3, false, true, false>(a, ph, x, pr);
    }
  }
  __syncthreads();   // the ring is no longer read: reuse it for the block sum
  const double tot = block_sum<BS>(acc, reinterpret_cast<double*>(fused_lds));
  if (tid == 0) ft.bpart[L] = tot;
}

// Mass factor of a fused step from the block sums (fixed order), one block.
template <int BS>
__global__ __launch_bounds__(BS) void k_fused_mass(const double* __restrict__ bpart,
                                                   unsigned int n, Ctrl* ctrl, int slot) {
  finalize_mass<BS>(bpart, n, ctrl, slot);
}

// Chunk classification: 1 when all 64 nodes of the chunk exist and are plain
// fluid.
// Kind of every 64-node chunk: 1 plain (every node fluid with no solid,
// partial, wall or top flag), 2 all solid (every node solid, none beyond N),
// 0 the rest.
template <int BS>
__global__ __launch_bounds__(BS) void k_chunk_kind(long long N, const uint8_t* __restrict__ meta,
                                                  uint8_t* __restrict__ kind) {
  const long long x = (long long)blockIdx.x * BS + threadIdx.x;
  const bool in = x < N;
  const uint8_t m = in ? meta[x] : (uint8_t)0;
  const unsigned long long plain = __ballot(in && (m & 0x0Fu) == 0);
  const unsigned long long solid = __ballot(in && (m & M_SOLID) != 0);
  if ((threadIdx.x & 63) == 0 && (x >> 6) < (N + 63) / 64)
    kind[x >> 6] = plain == ~0ull ? 1 : (solid == ~0ull ? 2 : 0);
}

// ---------------------------------------------------------------------------
// Observables (lazily materialised attributes of the solver object).
// ---------------------------------------------------------------------------
enum ObsKind : int {
  OBS_PHI = 0, OBS_RHO = 1, OBS_P = 3, OBS_MU = 4, OBS_F = 5, OBS_G = 6,
};

template <typename T, typename S, int DIM> struct ObsArgs {
  Grid g;
  const S* __restrict__ Fsrc;   // F/G source buffer
  const T* __restrict__ pref;   // phi reference of 16-bit f slots
  const T* __restrict__ phi_cur;
  const T* __restrict__ phi_prev;
  const uint8_t* __restrict__ meta;
  const uint32_t* __restrict__ wmask;
  const float* __restrict__ normal;
  const T* __restrict__ eps;
  const T* __restrict__ u_src;  // velocity the Fsrc slots were written with (bf16 g)
  const Ctrl* ctrl;
  T* __restrict__ out;
  int kind;
  int local;       // source buffer in post-stream form
  int slot;        // mass factor slot of the step that produced phi_cur
  int scaled;      // apply mass factor / relax (phi, f); 0 right after init
  int init_lap;    // mu with the init_fields Laplacian
  int internal;    // write in the internal node order instead of the public one
};

template <typename T, typename S, int DIM, int BS>
__global__ __launch_bounds__(BS) void k_observe(ObsArgs<T, S, DIM> a, Phys<T> ph) {
  constexpr int Q = Lat<DIM>::Q;
  const Grid& g = a.g;
  const long long N = g.N;
  const uint32_t x = blockIdx.x * BS + threadIdx.x;
  if (x >= N) return;
  const uint8_t m = a.meta[x];
  const Coord<DIM> c = decompose<DIM>(g, x);
  // outputs in the public order unless asked otherwise
  const uint32_t ex = a.internal ? x : ext_index<DIM>(g, c);
  switch (a.kind) {
    case OBS_PHI: {
      T p = a.phi_cur[x];
      if (a.scaled) {
        const T cf = (T)a.ctrl->c[a.slot];
        p = s12_phi<T>(ph, m, clamp01(p * cf), nsw_phi<T>(ph, m, a.phi_cur, x), cf);
      }
      a.out[ex] = p;
      break;
    }
    case OBS_RHO: {
      const T p = a.phi_cur[x];
      a.out[ex] = p * ph.rho_l + (T(1) - p) * ph.rho_g;
      break;
    }
    case OBS_MU: {
      const T p = a.phi_cur[x];
      const bool bnd = (m & M_BOUNDARY) && ph.wet != WET_NONE;
      float n32[DIM];
      uint32_t wm = 0;
#pragma unroll
      for (int d = 0; d < DIM; ++d) n32[d] = bnd ? a.normal[(long long)d * N + x] : 0.0f;
      if (bnd) wm = a.wmask[x];
      T gr[DIM], lap;
      stencil<T, DIM>(ph, g, a.phi_cur, neighbours<DIM>(g, x, c), p, wm, n32, gr, &lap,
                      a.init_lap != 0);
      a.out[ex] = chem_pot<T>(ph, p, lap, bnd);
      break;
    }
    case OBS_P:
    case OBS_F:
    case OBS_G: {
      T f[Q], unused[Q];
      const Buf bF = dist_buf<DIM>(a.Fsrc, N), bR = field_buf(a.pref, N);
      const Nbrs<DIM> nb = neighbours<DIM>(g, x, c);
      constexpr bool DV = !std::is_same<T, S>::value;
      if (a.kind == OBS_F) {
        if (a.local) pull<T, S, DIM, DV, false, true>(bF, bF, x, nb, m, a.eps, bR, ph, f, unused);
        else pull<T, S, DIM, DV, false, false>(bF, bF, x, nb, m, a.eps, bR, ph, f, unused);
      } else {
        const Buf bU = field_buf(a.u_src, N, DIM);
        const uint32_t N32 = (uint32_t)N;
        if (a.local) pull_g<T, S, DIM, kGRef<S>, true>(bF, bU, N32, x, nb, m, a.eps, ph, f);
        else pull_g<T, S, DIM, kGRef<S>, false>(bF, bU, N32, x, nb, m, a.eps, ph, f);
      }
      if (a.kind == OBS_P) {
        a.out[ex] = qsum<Q>(f);
        break;
      }
      if (a.kind == OBS_F && a.scaled && !a.local)
        scale_f<T, DIM>(ph, f, (T)a.ctrl->c[a.slot], m, (m & M_TOP) ? a.phi_cur[x] : T(0),
                        nullptr, nsw_phi<T>(ph, m, a.phi_cur, x));
#pragma unroll
      for (int i = 0; i < Q; ++i) a.out[(long long)i * N + ex] = f[i];
      break;
    }
  }
}

// Gradient and Laplacian of an arbitrary phi (the private operator the
// reference's MMS script calls); phi in internal order, outputs in the public
// order.
template <typename T, int DIM, int BS>
__global__ __launch_bounds__(BS) void k_gradlap(Grid g, Phys<T> ph, const T* __restrict__ phi,
                                                const uint8_t* __restrict__ meta,
                                                const uint32_t* __restrict__ wmask,
                                                const float* __restrict__ normal,
                                                T* __restrict__ grad, T* __restrict__ lap) {
  const long long N = g.N;
  const uint32_t x = blockIdx.x * BS + threadIdx.x;
  if (x >= N) return;
  const Coord<DIM> c = decompose<DIM>(g, x);
  const uint8_t m = meta ? meta[x] : 0;
  const bool bnd = (m & M_BOUNDARY) && ph.wet != WET_NONE;
  float n32[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) n32[d] = bnd ? normal[(long long)d * N + x] : 0.0f;
  T gr[DIM], l;
  stencil<T, DIM>(ph, g, phi, neighbours<DIM>(g, x, c), phi[x], bnd ? wmask[x] : 0u, n32, gr,
                  &l, false);
  const uint32_t ex = ext_index<DIM>(g, c);
#pragma unroll
  for (int d = 0; d < DIM; ++d) grad[(long long)d * N + ex] = gr[d];
  lap[ex] = l;
}

// ---------------------------------------------------------------------------
// Mid-run droplet injection (lbm/fe_ac_solver.py:645-732), one pass over the
// nodes.  The current state is observed exactly as the phi / u / f / g
// observables report it (pull, mass factor and boundary relax of the last
// step), edited, and stored back in the post-stream ("local") form at the
// other buffer set, so the next step reads it in place as it does after
// init_fields:
//   core = C > 0.1;  phi' = core ? max(phi, C) : phi, clamped, 0 in solid;
//   u' = (1 - C) u + C u_impact, 0 in solid;
//   core nodes: f = f_eq(phi', u'), g = g_eq(P, u') with the P of the last
//   step; other nodes keep their f, g.
// phi' goes to phi_out; per-wave fp64 partial sums of phi' give the new mass
// reference.  16-bit f slots are stored relative to w_i phi'(x), the
// reference the next step's pull adds back.
// ---------------------------------------------------------------------------
template <typename T, typename S, int DIM> struct InjectArgs {
  Grid g;
  const S* __restrict__ Fin;    // buffer set holding the current state
  const S* __restrict__ Gin;
  S* __restrict__ Fout;         // the other set, written in the local form
  S* __restrict__ Gout;
  const T* __restrict__ pref;   // start-of-step phi of the last step (16-bit f reference)
  const T* __restrict__ C;      // new drop indicator, internal order
  const T* __restrict__ P;      // pressure of the last step, internal order
  const T* __restrict__ u_in;   // velocity of the last step
  T* __restrict__ u_out;        // velocity after the injection (spare buffer)
  T* __restrict__ phi_out;
  const uint8_t* __restrict__ meta;
  const T* __restrict__ eps;
  const Ctrl* ctrl;
  double* __restrict__ partials;
  T uimp[3];
  int slot;                     // mass factor slot of the last step
};

template <typename T, typename S, int DIM, int BS, bool LOCAL>
__global__ __launch_bounds__(BS) void k_inject(InjectArgs<T, S, DIM> a, Phys<T> ph) {
  constexpr int Q = Lat<DIM>::Q;
  constexpr bool DEV = !std::is_same<T, S>::value;
  const Grid& g = a.g;
  const long long N = g.N;
  const uint32_t x = blockIdx.x * BS + threadIdx.x;
  double v = 0.0;
  if (x < N) {
    const uint8_t m = a.meta[x];
    const Coord<DIM> c = decompose<DIM>(g, x);
    const Nbrs<DIM> nb = neighbours<DIM>(g, x, c);
    const Buf bF = dist_buf<DIM>(a.Fin, N), bG = dist_buf<DIM>(a.Gin, N);
    const Buf bR = field_buf(a.pref, N), bU = field_buf(a.u_in, N, DIM);
    T f[Q], gg[Q], unused[Q];
    pull<T, S, DIM, DEV, false, LOCAL>(bF, bF, x, nb, m, a.eps, bR, ph, f, unused);
    pull_g<T, S, DIM, kGRef<S>, LOCAL>(bG, bU, (uint32_t)N, x, nb, m, a.eps, ph, gg);
    T p = a.pref[x];
    if constexpr (!LOCAL) {
      const T cf = (T)a.ctrl->c[a.slot];
      const T praw = p;
      p = clamp01(praw * cf);
      scale_f<T, DIM>(ph, f, cf, m, (m & M_TOP) ? praw : T(0), &p,
                      nsw_phi<T>(ph, m, a.pref, x));
    }
    const T C = a.C[x];
    const bool core = C > T(0.1);
    const bool solid = (m & M_SOLID) != 0;
    T pn = (core && p < C) ? C : p;
    pn = clamp01(pn);
    if (solid) pn = T(0);
    T uu[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      T un = (T(1) - C) * a.u_in[(long long)d * N + x] + C * a.uimp[d];
      if (solid) un = T(0);
      uu[d] = un;
      a.u_out[(long long)d * N + x] = un;
    }
    if (core) {
      const T usq_t = divk(vsq<DIM>(uu), ph.two_cs2, ph.rcp_two_cs2);
      T ec[Q];
      ediv<T, DIM>(ph, uu, ph.cs2, ph.rcp_cs2, ec);
      const T P = a.P[x];
#pragma unroll
      for (int i = 0; i < Q; ++i) {
        f[i] = feq_of<T, DIM>(ph, i, pn, ec[i], usq_t);
        gg[i] = geq_of<T, DIM>(ph, i, P, ec[i], usq_t);
      }
    }
    const Buf oF = dist_buf<DIM>(a.Fout, N), oG = dist_buf<DIM>(a.Gout, N);
    const uint32_t vx = dvoff<Q, S>(x);
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      T fv = f[i], gv = gg[i];
      if constexpr (DEV) fv = fv - ph.template w<DIM>(i) * pn;
      // local form: the slot is read back at this node, against its new u
      if constexpr (kGRef<S>) gv = gv - ph.template w3<DIM>(i) * edot<DIM>(i, uu);
      dput<Q, T, S>(oF, i, vx, fv);
      dput<Q, T, S>(oG, i, vx, gv);
    }
    a.phi_out[x] = pn;
    v = (double)pn;
  }
  wave_partial(v, x, N, a.partials);
}

// Reordering between the public layout (last axis fastest) and the internal
// one, for `comps` stacked arrays of N elements each.
template <typename E, int DIM, int BS>
__global__ __launch_bounds__(BS) void k_to_internal(Grid g, const E* __restrict__ pub,
                                                    E* __restrict__ in, int comps) {
  const uint32_t x = blockIdx.x * BS + threadIdx.x;
  if (x >= g.N) return;
  const uint32_t ex = ext_index<DIM>(g, decompose<DIM>(g, x));
  for (int k = 0; k < comps; ++k) in[(long long)k * g.N + x] = pub[(long long)k * g.N + ex];
}
template <typename E, int DIM, int BS>
__global__ __launch_bounds__(BS) void k_to_public(Grid g, const E* __restrict__ in,
                                                  E* __restrict__ pub, int comps) {
  const uint32_t x = blockIdx.x * BS + threadIdx.x;
  if (x >= g.N) return;
  const uint32_t ex = ext_index<DIM>(g, decompose<DIM>(g, x));
  for (int k = 0; k < comps; ++k) pub[(long long)k * g.N + ex] = in[(long long)k * g.N + x];
}

}  // namespace lbm
