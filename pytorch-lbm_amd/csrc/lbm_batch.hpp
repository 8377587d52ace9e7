// Multi-case step kernels: many solver cases of different grid sizes advanced
// by one launch per pass.  Each case keeps its own device state (the buffers
// of its lbm_solver) and its own physics constants; a per-block table maps a
// workgroup to (case, block within the case).  Every case's block range starts
// at a multiple of the XCD count, so the XCD-aware block order of the
// single-case kernels carries over unchanged.
//
// The per-case step counter lives on the device: pass A and pass B derive the
// buffer parity from it and skip cases that reached their step target, and a
// one-block kernel advances it after pass B, so a sequence of batch steps can
// be captured into a graph and replayed.
#pragma once

#include "lbm_kernels.hpp"

namespace lbm {

template <typename T, typename S> struct BatchCase {
  Grid g;
  Phys<T> ph;
  S* F[2];
  S* G[2];
  T* u[2];                     // velocity by step parity
  T* phi[2];
  T* eps;
  T* ampf;
  uint8_t* meta;
  uint32_t* wmask;
  float* normal;
  double* partials;
  Ctrl* ctrl;
  unsigned int nbA, nbB;       // blocks of this case in pass A / pass B
  unsigned int nbA8, nbB8;     // rounded up to a multiple of the XCD count
  const uint32_t* chunks;      // pass-B chunk lists: plain chunks, then the rest
  unsigned int n_plain, n_wall;
  unsigned int nbP8, nbW8;     // blocks of the two chunk lists (XCD multiples)
  const uint32_t* mixed;       // single-launch list for the batch's pass-B block size
  unsigned int n_mixed, nbM8;  // its entries; its blocks rounded up to an XCD multiple
  long long k, k_end;          // steps done, step target
  int local, phiu_valid;
  int pad[2];
};

// blockIdx -> (case, block within the case) for one pass.
struct BlockMap {
  const uint32_t* cs;    // case index per block
  const uint32_t* base;  // first block of that case
};

template <typename T, typename S, int DIM, int BS>
__global__ __launch_bounds__(BS) void k_phi_batch(const BatchCase<T, S>* __restrict__ cases,
                                                  BlockMap map) {
  const uint32_t ci = map.cs[blockIdx.x];
  const BatchCase<T, S>& cs = cases[ci];
  if (cs.k >= cs.k_end) return;
  const unsigned int i = blockIdx.x - map.base[blockIdx.x];
  const unsigned int lb = xcd_block_rev(i, cs.nbA8);   // pass B's last writes first
  if (lb >= cs.nbA) return;
  const long long N = cs.g.N;
  const uint32_t x = lb * BS + threadIdx.x;
  const int pslot = (int)(cs.k & 1), nslot = (int)((cs.k + 1) & 1);
  const int src = pslot;   // buffers alternate with the step count from init
  double v = 0.0;
  if (x < N) {
    const T p = cs.local
                    ? phi_node<T, S, DIM, true>(cs.g, cs.ph, cs.F[src], cs.phi[pslot], cs.meta,
                                                cs.eps, T(1), x)
                    : phi_node<T, S, DIM, false>(cs.g, cs.ph, cs.F[src], cs.phi[pslot], cs.meta,
                                                 cs.eps, (T)cs.ctrl->c[pslot], x);
    at(cs.phi[nslot], x) = p;
    v = (double)p;
  }
  wave_partial(v, x, N, cs.partials);
}

template <typename T, typename S, int DIM, int BS>
__global__ __launch_bounds__(BS) void k_collide_batch(const BatchCase<T, S>* __restrict__ cases,
                                                      BlockMap map) {
  const uint32_t ci = map.cs[blockIdx.x];
  const BatchCase<T, S>& cs = cases[ci];
  if (cs.k >= cs.k_end) return;
  const unsigned int i = blockIdx.x - map.base[blockIdx.x];
  const int pslot = (int)(cs.k & 1), nslot = (int)((cs.k + 1) & 1);
  const long long N = cs.g.N;
  if (i == 0)
    finalize_mass<BS>(cs.partials, (unsigned int)((N + 63) / 64), cs.ctrl, nslot);
  const unsigned int lb = xcd_block(i, cs.nbB8);
  if (lb >= cs.nbB) return;
  const uint32_t x = lb * BS + threadIdx.x;
  if (x >= N) return;
  CollideArgs<T, S, DIM> a;
  a.g = cs.g;
  a.Fin = cs.F[pslot];
  a.Gin = cs.G[pslot];
  a.Fout = cs.F[nslot];
  a.Gout = cs.G[nslot];
  a.u_in = cs.u[pslot];
  a.u_out = cs.u[nslot];
  a.phi_cur = cs.phi[nslot];
  a.phi_prev = cs.phi[pslot];
  a.meta = cs.meta;
  a.wmask = cs.wmask;
  a.normal = cs.normal;
  a.eps = cs.eps;
  a.ampf = cs.ampf;
  a.ctrl = cs.ctrl;
  a.ctrl_w = cs.ctrl;
  a.partials = cs.partials;
  a.nparts = 0;
  a.slot_prev = pslot;
  a.slot_new = nslot;
  a.local = cs.local;
  a.phiu_valid = cs.phiu_valid;
  if (cs.local) collide_node<T, S, DIM, true>(a, cs.ph, x);
  else collide_node<T, S, DIM, false>(a, cs.ph, x);
}

// Pass B over the chunk lists of every case (one wave per 64-node chunk):
// FLUID launches cover the plain-fluid chunks with the wall-free code, the
// other launch the remaining chunks.  A case's mass factor is formed by its
// block 0 in the first of the two launches that has blocks for it.
template <typename T, typename S, int DIM, int BS, bool FLUID>
__global__ __launch_bounds__(BS) void k_collide_batch_chunks(
    const BatchCase<T, S>* __restrict__ cases, BlockMap map) {
  const uint32_t ci = map.cs[blockIdx.x];
  const BatchCase<T, S>& cs = cases[ci];
  if (cs.k >= cs.k_end) return;
  const unsigned int i = blockIdx.x - map.base[blockIdx.x];
  const int pslot = (int)(cs.k & 1), nslot = (int)((cs.k + 1) & 1);
  const long long N = cs.g.N;
  if (i == 0 && (FLUID || cs.n_plain == 0))
    finalize_mass<BS>(cs.partials, (unsigned int)((N + 63) / 64), cs.ctrl, nslot);
  const unsigned int nch = FLUID ? cs.n_plain : cs.n_wall;
  const unsigned int nb = (nch + BS / 64 - 1) / (BS / 64);
  const unsigned int lb = xcd_block(i, FLUID ? cs.nbP8 : cs.nbW8);
  if (lb >= nb) return;
  const unsigned int w = lb * (BS / 64) + (threadIdx.x >> 6);
  if (w >= nch) return;
  const uint32_t ce = __builtin_amdgcn_readfirstlane(cs.chunks[(FLUID ? 0u : cs.n_plain) + w]);
  const uint32_t x = ((ce & kChunkIdx) << 6) + (threadIdx.x & 63u);
  if (x >= N) return;
  CollideArgs<T, S, DIM> a;
  a.g = cs.g;
  a.Fin = cs.F[pslot];
  a.Gin = cs.G[pslot];
  a.Fout = cs.F[nslot];
  a.Gout = cs.G[nslot];
  a.u_in = cs.u[pslot];
  a.u_out = cs.u[nslot];
  a.phi_cur = cs.phi[nslot];
  a.phi_prev = cs.phi[pslot];
  a.meta = cs.meta;
  a.wmask = cs.wmask;
  a.normal = cs.normal;
  a.eps = cs.eps;
  a.ampf = cs.ampf;
  a.ctrl = cs.ctrl;
  a.ctrl_w = cs.ctrl;
  a.partials = cs.partials;
  a.nparts = 0;
  a.slot_prev = pslot;
  a.slot_new = nslot;
  a.local = cs.local;
  a.phiu_valid = cs.phiu_valid;
  if (cs.local) collide_node<T, S, DIM, true, FLUID>(a, cs.ph, x);
  else collide_node<T, S, DIM, false, FLUID>(a, cs.ph, x, FLUID || !(ce & kChunkSolid));
}

// Pass B of every case in one launch over its single-launch chunk list (each
// XCD's range of the case's blocks holds its plain chunks first and its wall
// chunks last, as in k_collide_mixed; a case's blocks start at a multiple of
// the XCD count, so block i of the case runs on XCD i % 8 as there).  A case
// in the post-stream form (its first step) runs the general code throughout.
template <typename T, typename S, int DIM, int BS>
__global__ __launch_bounds__(BS) void k_collide_batch_mixed(
    const BatchCase<T, S>* __restrict__ cases, BlockMap map) {
  const uint32_t ci = map.cs[blockIdx.x];
  const BatchCase<T, S>& cs = cases[ci];
  if (cs.k >= cs.k_end) return;
  const unsigned int i = blockIdx.x - map.base[blockIdx.x];
  const int pslot = (int)(cs.k & 1), nslot = (int)((cs.k + 1) & 1);
  const long long N = cs.g.N;
  if (i == 0)
    finalize_mass<BS>(cs.partials, (unsigned int)((N + 63) / 64), cs.ctrl, nslot);
  const unsigned int nb = cs.n_mixed / (BS / 64);
  if (i >= nb) return;
  const unsigned int w = xcd_block(i, nb) * (BS / 64) + (threadIdx.x >> 6);
  const uint32_t c = __builtin_amdgcn_readfirstlane(cs.mixed[w]);
  if (c == 0xFFFFFFFFu) return;
  const uint32_t x = ((c & kChunkIdx) << 6) + (threadIdx.x & 63u);
  if (x >= N) return;
  CollideArgs<T, S, DIM> a;
  a.g = cs.g;
  a.Fin = cs.F[pslot];
  a.Gin = cs.G[pslot];
  a.Fout = cs.F[nslot];
  a.Gout = cs.G[nslot];
  a.u_in = cs.u[pslot];
  a.u_out = cs.u[nslot];
  a.phi_cur = cs.phi[nslot];
  a.phi_prev = cs.phi[pslot];
  a.meta = cs.meta;
  a.wmask = cs.wmask;
  a.normal = cs.normal;
  a.eps = cs.eps;
  a.ampf = cs.ampf;
  a.ctrl = cs.ctrl;
  a.ctrl_w = cs.ctrl;
  a.partials = cs.partials;
  a.nparts = 0;
  a.slot_prev = pslot;
  a.slot_new = nslot;
  a.local = cs.local;
  a.phiu_valid = cs.phiu_valid;
  if (cs.local) collide_node<T, S, DIM, true, false>(a, cs.ph, x);
  else if (c & kChunkWall) collide_node<T, S, DIM, false, false>(a, cs.ph, x, !(c & kChunkSolid));
  else collide_node<T, S, DIM, false, true>(a, cs.ph, x);
}

// After pass B: every case that stepped advances its counter.
template <typename T, typename S>
__global__ void k_advance_batch(BatchCase<T, S>* __restrict__ cases, int n) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  BatchCase<T, S>& cs = cases[c];
  if (cs.k >= cs.k_end) return;
  cs.k += 1;
  cs.local = 0;
  cs.phiu_valid = 1;
}

}  // namespace lbm
