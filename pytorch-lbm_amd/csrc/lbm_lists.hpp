// Host-side construction of the pass-B chunk list of the single-launch form
// (k_collide_mixed).  Plain C++ (no HIP), so that the host tests can build
// it with g++ and check its properties.
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <vector>

namespace lbm {

// XCDs of an MI355X; the device-side block mapping (xcd_block) uses the same
// count.
constexpr unsigned int kListXCD = 8;

// The single-launch pass-B list for blocks of `wpb` waves: the launch has
// G = ceil(n / wpb) blocks, and xcd_block gives XCD j the logical blocks
// [j*(G/8) + min(j, G%8), ...) of G/8 (+1 for j < G%8).  Each XCD's range
// gets a contiguous slice of the plain chunks, then a slice of the wall
// chunks (bit 31 set) in proportion to its capacity; padding entries are
// 0xFFFFFFFF.  `list` holds the plain chunks, then the wall chunks; other
// flag bits of its entries (bit 30: all nodes solid) are carried over.
inline bool mixed_list(const std::vector<uint32_t>& list, unsigned int np, unsigned int nw,
                       unsigned int wpb, std::vector<uint32_t>& mx) {
  const unsigned int n = np + nw, G = (n + wpb - 1) / wpb;
  const unsigned int per = G / kListXCD, rem = G % kListXCD;
  mx.assign((size_t)G * wpb, 0xFFFFFFFFu);
  size_t pp = 0, ww = 0, left = mx.size();
  for (unsigned int j = 0; j < kListXCD; ++j) {
    const size_t start = ((size_t)j * per + (j < rem ? j : rem)) * wpb;
    const size_t cap = ((size_t)per + (j < rem ? 1 : 0)) * wpb;
    size_t wj = left ? (size_t)((double)(nw - ww) * (double)cap / (double)left + 0.5) : 0;
    wj = std::min(std::min(wj, (size_t)nw - ww), cap);
    size_t pj = std::min(cap - wj, (size_t)np - pp);
    wj = std::min(cap - pj, (size_t)nw - ww);   // fill what the plain share leaves
    for (size_t t = 0; t < pj; ++t) mx[start + t] = list[pp + t];
    for (size_t t = 0; t < wj; ++t) mx[start + pj + t] = list[np + ww + t] | 0x80000000u;
    pp += pj;
    ww += wj;
    left -= cap;
  }
  // capacity G * wpb >= n and every range takes min(its capacity, what is
  // left): everything is placed
  return pp == np && ww == nw;
}

}  // namespace lbm
