// Prints the compile-time shape-function tables of the FP64 element kernels
// (redynela_amd/csrc/dnlg_tables.hpp) as hexadecimal doubles, one per line:
// first the 196 entries of the hex8 table, then the 48 of the quad4 table.
// tests/test_capi.py compares them with the oracle's tables.
#include <cstdio>

#include "dnlg_tables.hpp"

int main()
{
  constexpr dnl::HexTab h = dnl::make_hex_tab();
  constexpr dnl::QuadTab q = dnl::make_quad_tab();
  for (double v : h.v) std::printf("%a\n", v);
  for (double v : q.v) std::printf("%a\n", v);
  return 0;
}
