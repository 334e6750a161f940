// CPU check of the device scratch layouts (built and run by
// tests/test_layouts.py with the host compiler; no GPU).
//
// Eppstein's per-automaton scratch (epp_layout): the 64-bit pair keys take
// 64-bit atomics, which must be naturally aligned, so every region starts on
// an 8-byte boundary, regions do not overlap, and the total is a multiple of
// 256 (consecutive automata's scratch stays aligned).  Checked for every
// n <= 512 and k <= 8.
#include <cstdio>

#include "../synchrowords_amd/csrc/sw_kernels.hpp"

int main() {
  long bad = 0, checked = 0;
  for (uint32_t n = 1; n <= 512; ++n)
    for (uint32_t k = 1; k <= 8; ++k) {
      const sw::EppLayout l = sw::epp_layout(n, k);
      const size_t kn = size_t(k) * n, nn = size_t(n) * n, all = size_t(n) * (n - 1) / 2;
      const size_t starts[] = {l.edges, l.info, l.keys, l.queue, l.cnt};
      bool ok = l.total % 256 == 0;
      for (size_t s : starts) ok = ok && s % 8 == 0;
      // each region holds what it must before the next one starts
      ok = ok && l.edges >= (kn + 1) * 4 && l.info >= l.edges + kn * 2 && l.keys >= l.info + nn * 4 &&
           l.queue >= l.keys + nn * 8 && l.cnt >= l.queue + (all + kn) * 4 && l.total >= l.cnt + (all + kn) * 4;
      if (!ok) {
        if (bad < 10) std::printf("bad layout n=%u k=%u\n", n, k);
        ++bad;
      }
      ++checked;
    }
  std::printf("checked %ld layouts, %ld bad\n", checked, bad);
  return bad ? 1 : 0;
}
