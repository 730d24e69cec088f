// Mesh generators of libdnlgpu (host code, no GPU call).
//
// dnlg_mesh_axi_bar follows the node/element recipe of the reference's sample
// deck Samples/Taylor/Taylor_Axi_Parametric.dyn:39-117 (node (i,j) at
// (i*dx, j*dy), quads (i,j),(i+1,j),(i+1,j+1),(i,j+1), base restrained along
// y, axis restrained along x and z, initial speed (0,-v0,0)), with 0-based
// indices.  dnlg_mesh_ogrid_bar is build-defined (the reference has no 3D
// mesher): an O-grid cylinder of hexahedra in the ElHex8n3D local node order
// (FE/ElHex8n3D.C:41-72).
#include "../../include/dnlg.h"

#include <cmath>

extern "C" int dnlg_mesh_axi_bar(int nr, int nz, double radius, double length, double v0, double *coords, int *conn,
                                 unsigned char *restrain, double *vel)
{
  if (nr < 1 || nz < 1 || !coords || !conn) return DNLG_E_ARG;
  const double dx = radius / (nr);
  const double dy = length / (nz);
  long long id = 0;
  for (int j = 0; j <= nz; j++)
    for (int i = 0; i <= nr; i++, id++)
    {
      coords[3 * id + 0] = (double)i * dx;
      coords[3 * id + 1] = (double)j * dy;
      coords[3 * id + 2] = 0.0;
      if (restrain)
      {
        unsigned char b = 0;
        if (j == 0) b |= 2;           // base: BoundaryRestrain (0,1,0)
        if (i == 0) b |= 1 | 4;       // axis: BoundaryRestrain (1,0,1)
        restrain[id] = b;
      }
      if (vel) { vel[3 * id] = 0.0; vel[3 * id + 1] = -v0; vel[3 * id + 2] = 0.0; }
    }
  long long e = 0;
  for (int j = 0; j < nz; j++)
    for (int i = 0; i < nr; i++, e++)
    {
      conn[4 * e + 0] = i + j * (nr + 1);
      conn[4 * e + 1] = i + 1 + j * (nr + 1);
      conn[4 * e + 2] = i + 1 + (j + 1) * (nr + 1);
      conn[4 * e + 3] = i + (j + 1) * (nr + 1);
    }
  return DNLG_OK;
}

extern "C" int dnlg_mesh_ogrid_sizes(int n, int m, int nz, long long *n_nodes, long long *n_elems)
{
  if (n < 1 || m < 1 || nz < 1) return DNLG_E_ARG;
  const long long npl = (long long)(n + 1) * (n + 1) + 4LL * n * m;
  const long long epl = (long long)n * n + 4LL * n * m;
  if (n_nodes) *n_nodes = npl * (nz + 1);
  if (n_elems) *n_elems = epl * nz;
  return DNLG_OK;
}

namespace {
// index of perimeter point p (0..4n-1, counter-clockwise from (-s,-s)) of the
// central block, and its position on the square of half-width s
long long perim_node(int n, int p)
{
  const int side = p / n, i = p % n;
  switch (side)
  {
  case 0: return i;                                         // y = -s, x increasing
  case 1: return (long long)i * (n + 1) + n;                // x = +s, y increasing
  case 2: return (long long)n * (n + 1) + (n - i);          // y = +s, x decreasing
  default: return (long long)(n - i) * (n + 1);             // x = -s, y decreasing
  }
}
void perim_pos(int n, int p, double s, double *x, double *y)
{
  const int side = p / n, i = p % n;
  const double t = 2.0 * s * (double)i / (double)n;
  switch (side)
  {
  case 0: *x = -s + t; *y = -s; break;
  case 1: *x = s; *y = -s + t; break;
  case 2: *x = s - t; *y = s; break;
  default: *x = -s; *y = s - t; break;
  }
}
}  // namespace

extern "C" int dnlg_mesh_ogrid_bar(int n, int m, int nz, double radius, double length, double s_frac, double v0,
                                   double *coords, int *conn, unsigned char *restrain, double *vel)
{
  long long NN, NE;
  if (dnlg_mesh_ogrid_sizes(n, m, nz, &NN, &NE) != DNLG_OK || !coords || !conn) return DNLG_E_ARG;
  if (!(s_frac > 0.0 && s_frac < 0.70)) return DNLG_E_ARG;
  if (NN > 2147483647LL || NE * 8 > 2147483647LL) return DNLG_E_ARG;
  const long long npl = (long long)(n + 1) * (n + 1) + 4LL * n * m;
  const long long epl = (long long)n * n + 4LL * n * m;
  const long long ring0 = (long long)(n + 1) * (n + 1);
  const double s = s_frac * radius;
  // in-plane node positions of one layer
  double *px = new double[npl], *py = new double[npl];
  for (int j = 0; j <= n; j++)
    for (int i = 0; i <= n; i++)
    {
      px[j * (n + 1) + i] = -s + 2.0 * s * (double)i / (double)n;
      py[j * (n + 1) + i] = -s + 2.0 * s * (double)j / (double)n;
    }
  for (int r = 1; r <= m; r++)
    for (int p = 0; p < 4 * n; p++)
    {
      double xs, ys;
      perim_pos(n, p, s, &xs, &ys);
      const double rr = std::sqrt(xs * xs + ys * ys);
      const double xc = radius * xs / rr, yc = radius * ys / rr;
      const double f = (double)r / (double)m;
      const long long id = ring0 + (long long)(r - 1) * 4 * n + p;
      px[id] = xs + f * (xc - xs);
      py[id] = ys + f * (yc - ys);
    }
  // in-plane quads (counter-clockwise seen from +z)
  int *q = new int[epl * 4];
  long long qe = 0;
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++, qe++)
    {
      q[4 * qe + 0] = j * (n + 1) + i;
      q[4 * qe + 1] = j * (n + 1) + i + 1;
      q[4 * qe + 2] = (j + 1) * (n + 1) + i + 1;
      q[4 * qe + 3] = (j + 1) * (n + 1) + i;
    }
  auto ring_node = [&](int r, int p) -> long long {
    p %= 4 * n;
    return r == 0 ? perim_node(n, p) : ring0 + (long long)(r - 1) * 4 * n + p;
  };
  for (int r = 0; r < m; r++)
    for (int p = 0; p < 4 * n; p++, qe++)
    {
      q[4 * qe + 0] = (int)ring_node(r, p);
      q[4 * qe + 1] = (int)ring_node(r + 1, p);
      q[4 * qe + 2] = (int)ring_node(r + 1, p + 1);
      q[4 * qe + 3] = (int)ring_node(r, p + 1);
    }
  for (int k = 0; k <= nz; k++)
  {
    const double z = (double)k * (length / (double)nz);
    for (long long i = 0; i < npl; i++)
    {
      const long long id = (long long)k * npl + i;
      coords[3 * id + 0] = px[i];
      coords[3 * id + 1] = py[i];
      coords[3 * id + 2] = z;
      if (restrain) restrain[id] = (k == 0) ? 4 : 0;
      if (vel) { vel[3 * id] = 0.0; vel[3 * id + 1] = 0.0; vel[3 * id + 2] = -v0; }
    }
  }
  for (int k = 0; k < nz; k++)
    for (long long t = 0; t < epl; t++)
    {
      const long long e = (long long)k * epl + t;
      for (int c = 0; c < 4; c++)
      {
        conn[8 * e + c] = (int)(q[4 * t + c] + (long long)k * npl);
        conn[8 * e + 4 + c] = (int)(q[4 * t + c] + (long long)(k + 1) * npl);
      }
    }
  delete[] px;
  delete[] py;
  delete[] q;
  return DNLG_OK;
}
