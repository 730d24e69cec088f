// Compile-time shape-function tables of the FP64 element kernels, in the
// layouts of their LDS copies (HexLds::tab, QuadLds::tab).  Plain C++17
// (constexpr, no device code), so that a host program can build the same
// tables and compare them with another source of the values.
#pragma once

namespace dnl {

// The hex8 table of dN0(ip, node, d) in the layout of HexLds::tab, built at
// compile time: the magnitudes are the closed forms (2 - sqrt 3)/12, 1/12 and
// (2 + sqrt 3)/12 of FE/ElHex8n3D.C:81-87, correctly rounded -- the doubles
// the host computes for ElemArgs::c0..c2 -- and the signs and the padding
// after point 3 are those of hex_dn0 and hex_tab.
constexpr double kHexC0 = 0x1.6dd707e2911cbp-6;
constexpr double kHexC1 = 0x1.5555555555555p-4;
constexpr double kHexC2 = 0x1.3e77e4d72c439p-2;
struct HexTab { double v[8 * 8 * 3 + 4]; };
constexpr int hex_csg(int k, int d)
{
  return d == 0 ? (((k & 3) == 1 || (k & 3) == 2) ? 1 : -1) : (d == 1 ? (((k & 3) >= 2) ? 1 : -1) : (k >= 4 ? 1 : -1));
}
constexpr HexTab make_hex_tab()
{
  HexTab t{};
  for (int ip = 0; ip < 8; ip++)
    for (int k = 0; k < 8; k++)
      for (int d = 0; d < 3; d++)
      {
        const int shared = (hex_csg(ip, (d + 1) % 3) == hex_csg(k, (d + 1) % 3)) +
                           (hex_csg(ip, (d + 2) % 3) == hex_csg(k, (d + 2) % 3));
        const double m = shared == 2 ? kHexC2 : (shared == 1 ? kHexC1 : kHexC0);
        t.v[(ip * 8 + k) * 3 + (ip >= 4 ? 4 : 0) + d] = hex_csg(k, d) > 0 ? m : -m;
      }
  return t;
}
// The quad4 table of QuadLds::tab (dN0(ip, k, 0..1), N(ip, k)) built the same
// way: (3 -+ sqrt 3)/12 and (2 -+ sqrt 3)/6, 1/6 of FE/ElQua4nAx.C:51-73,
// correctly rounded -- the doubles of ElemArgs::c0, c1 and n0..n2 -- with
// the signs of quad_dn0 and the choice of quad_n.
constexpr double kQuadC0 = 0x1.b0cb174df99c8p-4;
constexpr double kQuadC1 = 0x1.93cd3a2c8198dp-2;
constexpr double kQuadN0 = 0x1.6dd707e2911cbp-5;
constexpr double kQuadN1 = 0x1.5555555555555p-3;
constexpr double kQuadN2 = 0x1.3e77e4d72c439p-1;
struct QuadTab { double v[4 * 4 * 3]; };
constexpr QuadTab make_quad_tab()
{
  QuadTab t{};
  for (int ip = 0; ip < 4; ip++)
    for (int k = 0; k < 4; k++)
    {
      const bool s0 = hex_csg(ip, 0) == hex_csg(k, 0), s1 = hex_csg(ip, 1) == hex_csg(k, 1);
      const double m0 = s1 ? kQuadC1 : kQuadC0, m1 = s0 ? kQuadC1 : kQuadC0;
      t.v[(ip * 4 + k) * 3 + 0] = hex_csg(k, 0) > 0 ? m0 : -m0;
      t.v[(ip * 4 + k) * 3 + 1] = hex_csg(k, 1) > 0 ? m1 : -m1;
      const int ag = s0 + s1;
      t.v[(ip * 4 + k) * 3 + 2] = ag == 2 ? kQuadN2 : (ag == 1 ? kQuadN1 : kQuadN0);
    }
  return t;
}
}  // namespace dnl
