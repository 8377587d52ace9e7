// Property check of the single-launch pass-B chunk list (lbm_lists.hpp),
// host only: for many plain / wall chunk counts and block sizes, every chunk
// appears exactly once, wall chunks carry bit 31, and the logical block range
// each XCD receives from the device mapping (restated below) holds plain
// chunks, then wall chunks, then padding.  Prints "mixed_list ok".
#include <cstdio>
#include <vector>

#include "../../pytorch-lbm_amd/csrc/lbm_lists.hpp"

// the device mapping of a launch of nb blocks: physical block b -> logical
static unsigned int xcd_block(unsigned int b, unsigned int nb) {
  const unsigned int j = b % 8, i = b / 8, per = nb / 8, rem = nb % 8;
  return j * per + (j < rem ? j : rem) + i;
}

static int fails = 0;
#define CHECK(c, ...)                                  \
  do {                                                 \
    if (!(c)) {                                        \
      if (fails < 20) {                                \
        std::printf("FAIL line %d: ", __LINE__);       \
        std::printf(__VA_ARGS__);                      \
        std::printf("\n");                             \
      }                                                \
      ++fails;                                         \
    }                                                  \
  } while (0)

static void check(unsigned int np, unsigned int nw, unsigned int wpb) {
  std::vector<uint32_t> list;
  for (unsigned int c = 0; c < np + nw; ++c) list.push_back((c * 7919u) % (np + nw));   // any order
  std::vector<uint32_t> mx;
  CHECK(lbm::mixed_list(list, np, nw, wpb, mx), "np %u nw %u wpb %u: not all placed", np, nw, wpb);
  const unsigned int n = np + nw, G = (n + wpb - 1) / wpb;
  CHECK(mx.size() == (size_t)G * wpb, "size");
  // every chunk once, with the right kind
  std::vector<int> seen(n, 0);
  for (uint32_t e : mx) {
    if (e == 0xFFFFFFFFu) continue;
    const uint32_t c = e & 0x7FFFFFFFu;
    CHECK(c < n, "bad chunk %u", c);
    if (c >= n) continue;
    ++seen[c];
  }
  for (unsigned int k = 0; k < n; ++k) {
    CHECK(seen[list[k]] == 1, "chunk %u seen %d times", list[k], seen[list[k]]);
  }
  for (uint32_t e : mx) {
    if (e == 0xFFFFFFFFu) continue;
    // position of the chunk in `list` says its kind
    unsigned int pos = 0;
    while (pos < n && list[pos] != (e & 0x7FFFFFFFu)) ++pos;
    CHECK((pos >= np) == ((e >> 31) != 0), "kind of chunk %u", e & 0x7FFFFFFFu);
  }
  // the mapping is a bijection and each XCD's range reads plain, wall, padding
  std::vector<int> hit(G, 0);
  for (unsigned int b = 0; b < G; ++b) ++hit[xcd_block(b, G)];
  for (unsigned int lb = 0; lb < G; ++lb) CHECK(hit[lb] == 1, "mapping not a bijection");
  for (unsigned int j = 0; j < 8; ++j) {
    int stage = 0;   // 0 plain, 1 wall, 2 padding
    for (unsigned int b = j; b < G; b += 8) {
      const unsigned int lb = xcd_block(b, G);
      for (unsigned int w = 0; w < wpb; ++w) {
        const uint32_t e = mx[(size_t)lb * wpb + w];
        const int kind = e == 0xFFFFFFFFu ? 2 : ((e >> 31) ? 1 : 0);
        CHECK(kind >= stage, "np %u nw %u wpb %u: XCD %u out of order", np, nw, wpb, j);
        stage = kind;
      }
    }
  }
}

int main() {
  const unsigned int counts[] = {0, 1, 2, 3, 7, 8, 9, 15, 16, 17, 31, 63, 64, 65, 100, 257, 1000};
  for (unsigned int wpb : {2u, 4u, 8u})
    for (unsigned int np : counts)
      for (unsigned int nw : counts)
        if (np + nw > 0) check(np, nw, wpb);
  check(34620, 3000, 4);   // the 150 x 150 x 107 ridge case
  check(34620, 3000, 2);
  check(34620, 3000, 8);
  if (fails) {
    std::printf("%d failure(s)\n", fails);
    return 1;
  }
  std::printf("mixed_list ok\n");
  return 0;
}
