// Linear-element (Tri3 / Tet4) geometry and integrals, shared by host and
// device code.  Operation order follows the reference element classes so
// that, compiled without floating-point contraction, every value equals the
// reference's bit for bit:
//   Tet4: tet4.cc:113-171 (coefficients), :348-419 (IntFF/IntFFF/IntPFF/
//         IntFDD/IntPDD), :819-832 (BndIntPFF), :1011-1014 (size),
//         :1039-1068 + :1109-1122 (boundary mass), tet4.cc:19-29 (side areas)
//   Tri3: tri3.cc:79-90, :184-189, :285-353, :561-582, tri3.h:196-218
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#include "mesh.hpp"

namespace tmi {

#define TMI_HD __host__ __device__ inline

// side -> local node tables (tet4.cc:177, tri3.cc:121-127)
TMI_HD int tet_side_node(int s, int k)
{
    const int t[4][3] = {{0, 1, 2}, {0, 3, 1}, {0, 2, 3}, {1, 3, 2}};
    return t[s][k];
}
TMI_HD int tri_side_node(int s, int k)
{
    const int t[3][2] = {{0, 1}, {1, 2}, {2, 0}};
    return t[s][k];
}
TMI_HD bool tet_in_side(int s, int i)
{
    return i == tet_side_node(s, 0) || i == tet_side_node(s, 1) || i == tet_side_node(s, 2);
}

// numerator of the integral of l_i l_j l_k over a simplex (x 1/120 tet, x 2A/120 triangle)
TMI_HD int mult3(int i, int j, int k)
{
    if (i == j && j == k) return 6;
    if (i == j || j == k || i == k) return 2;
    return 1;
}

TMI_HD double pdist3(const double *p, const double *q)
{
    double d2 = 0.0;
    for (int i = 0; i < 3; i++) d2 += (p[i] - q[i]) * (p[i] - q[i]);
    return sqrt(d2);
}

TMI_HD double tri_area3(const double *p1, const double *p2, const double *p3)
{
    double a = pdist3(p1, p2), b = pdist3(p2, p3), c = pdist3(p3, p1);
    double s = 0.5 * (a + b + c);
    return sqrt(s * (s - a) * (s - b) * (s - c));
}

// Fill the geometry record g of a tetrahedron with nodes P0..P3 (3 doubles each)
TMI_HD void tet_geometry(const double *P0, const double *P1, const double *P2, const double *P3, double *g)
{
    double x0 = P0[0], y0 = P0[1], z0 = P0[2];
    double x1 = P1[0], y1 = P1[1], z1 = P1[2];
    double x2 = P2[0], y2 = P2[1], z2 = P2[2];
    double x3 = P3[0], y3 = P3[1], z3 = P3[2];
    double *a = g, *b = g + 4, *c = g + 8, *d = g + 12;
    a[0] = x1 * (y2 * z3 - y3 * z2) - x2 * (y1 * z3 - y3 * z1) + x3 * (y1 * z2 - y2 * z1);
    a[1] = x0 * (y3 * z2 - y2 * z3) - x2 * (y3 * z0 - y0 * z3) + x3 * (y2 * z0 - y0 * z2);
    a[2] = x0 * (y1 * z3 - y3 * z1) - x1 * (y0 * z3 - y3 * z0) + x3 * (y0 * z1 - y1 * z0);
    a[3] = x0 * (y2 * z1 - y1 * z2) - x1 * (y2 * z0 - y0 * z2) + x2 * (y1 * z0 - y0 * z1);
    b[0] = y1 * (z3 - z2) + y2 * (z1 - z3) + y3 * (z2 - z1);
    b[1] = y0 * (z2 - z3) + y2 * (z3 - z0) + y3 * (z0 - z2);
    b[2] = y0 * (z3 - z1) + y1 * (z0 - z3) + y3 * (z1 - z0);
    b[3] = y0 * (z1 - z2) + y1 * (z2 - z0) + y2 * (z0 - z1);
    c[0] = x1 * (z2 - z3) + x2 * (z3 - z1) + x3 * (z1 - z2);
    c[1] = x0 * (z3 - z2) + x2 * (z0 - z3) + x3 * (z2 - z0);
    c[2] = x0 * (z1 - z3) + x1 * (z3 - z0) + x3 * (z0 - z1);
    c[3] = x0 * (z2 - z1) + x1 * (z0 - z2) + x2 * (z1 - z0);
    d[0] = x1 * (y3 - y2) + x2 * (y1 - y3) + x3 * (y2 - y1);
    d[1] = x0 * (y2 - y3) + x2 * (y3 - y0) + x3 * (y0 - y2);
    d[2] = x0 * (y3 - y1) + x1 * (y0 - y3) + x3 * (y1 - y0);
    d[3] = x0 * (y1 - y2) + x1 * (y2 - y0) + x2 * (y0 - y1);
    const double *P[4] = {P0, P1, P2, P3};
    for (int s = 0; s < 4; s++)
        g[17 + s] = tri_area3(P[tet_side_node(s, 0)], P[tet_side_node(s, 1)], P[tet_side_node(s, 2)]);
    g[16] = (1.0 / 6.0) * (a[0] + a[1] + a[2] + a[3]);
}

TMI_HD void tri_geometry(const double *P0, const double *P1, const double *P2, double *g)
{
    double x0 = P0[0], y0 = P0[1], x1 = P1[0], y1 = P1[1], x2 = P2[0], y2 = P2[1];
    double *a = g, *b = g + 3, *c = g + 6;
    a[0] = x1 * y2 - x2 * y1;  b[0] = y1 - y2;  c[0] = x2 - x1;
    a[1] = x2 * y0 - x0 * y2;  b[1] = y2 - y0;  c[1] = x0 - x2;
    a[2] = x0 * y1 - x1 * y0;  b[2] = y0 - y1;  c[2] = x1 - x0;
    g[9] = 0.5 * (a[0] + a[1] + a[2]);
    for (int k = 10; k < 16; k++) g[k] = 0.0;
}

// index of (i,j) in the packed lower triangle of a 3x3 symmetric matrix
TMI_HD int tri_lt(int i, int j)
{
    if (i < j) { int t = i; i = j; j = t; }
    return i * (i + 1) / 2 + j;
}

// Tri3 boundary mass matrix (ComputeBndIntFF) from the node coordinates;
// writes g[10..15].  side_mask: bit s set for boundary side s.
TMI_HD void tri_bndff(const double *P0, const double *P1, const double *P2, int side_mask, double *g)
{
    const double *P[3] = {P0, P1, P2};
    for (int s = 0; s < 3; s++) {
        if (!(side_mask & (1 << s))) continue;
        int n0 = tri_side_node(s, 0), n1 = tri_side_node(s, 1);
        double dx = P[n0][0] - P[n1][0];
        double dy = P[n0][1] - P[n1][1];
        double dl = hypot(dx, dy);
        for (int i = 0; i < 3; i++) {
            if (i != n0 && i != n1) continue;
            g[10 + tri_lt(i, i)] += dl / 3.0;
            for (int j = 0; j < i; j++)
                if (j == n0 || j == n1) g[10 + tri_lt(i, j)] += dl / 6.0;
        }
    }
}

// ---- Tet4 integrals (g: geometry record) ----------------------------------
TMI_HD double tet_intff(const double *g, int i, int j) { return ((i == j ? 2.0 : 1.0) * (1.0 / 20.0)) * g[16]; }

TMI_HD double tet_intfff(const double *g, int i, int j, int k) { return ((double)mult3(i, j, k) / 120.0) * g[16]; }

TMI_HD double tet_intfdd(const double *g, int j, int k)   // IntFDD(.,j,k): independent of the first index
{
    if (j > k) { int t = j; j = k; k = t; }
    const double *b = g + 4, *c = g + 8, *d = g + 12;
    double scale = 1.0 / (144.0 * g[16]);
    return (b[j] * b[k] + c[j] * c[k] + d[j] * d[k]) * scale;
}

// IntDD (the stored intdd, tet4.cc:1016-1031): (b_i b_j + c_i c_j + d_i d_j) / (36 V)
TMI_HD double tet_intdd(const double *g, int i, int j)
{
    if (i < j) { int t = i; i = j; j = t; }
    const double *b = g + 4, *c = g + 8, *d = g + 12;
    double scale = 1.0 / (36.0 * g[16]);
    return scale * (b[j] * b[i] + c[j] * c[i] + d[j] * d[i]);
}

TMI_HD double tet_intpff(const double *g, int i, int j, double P0, double P1, double P2, double P3)
{
    return ((double)mult3(0, i, j) / 120.0 * P0 + (double)mult3(1, i, j) / 120.0 * P1 +
            (double)mult3(2, i, j) / 120.0 * P2 + (double)mult3(3, i, j) / 120.0 * P3) * g[16];
}

TMI_HD double tet_intpdd(const double *g, int i, int j, double P0, double P1, double P2, double P3)
{
    double res = 0.0;
    res += P0; res += P1; res += P2; res += P3;
    return res * tet_intfdd(g, i, j);
}

TMI_HD double tet_bndintpff(const double *g, int flags, int i, int j, const double *P4)
{
    double res = 0.0;
    for (int s = 0; s < 4; s++) {
        if (!(flags & (1 << s))) continue;
        double sres = 0.0;
        for (int k = 0; k < 4; k++) {
            double t = (tet_in_side(s, i) && tet_in_side(s, j) && tet_in_side(s, k))
                           ? (double)mult3(k, i, j) * (2.0 / 120.0) : 0.0;
            sres += P4[k] * t;
        }
        res += sres * g[17 + s];
    }
    return res;
}

// boundary mass matrix entry (ComputeBndIntFF, tet4.cc:1109-1122)
TMI_HD double tet_bndff(const double *g, int flags, int i, int j)
{
    if (i < j) { int t = i; i = j; j = t; }
    double bf = 0.0;
    for (int s = 0; s < 4; s++) {
        if (!(flags & (1 << s))) continue;
        double v = (tet_in_side(s, i) && tet_in_side(s, j)) ? (i == j ? 2.0 : 1.0) * (2.0 / 24.0) : 0.0;
        bf += v * g[17 + s];
    }
    return bf;
}

// ---- Tri3 integrals ---------------------------------------------------------
TMI_HD double tri_intff(const double *g, int i, int j) { return ((i == j ? 2.0 : 1.0) * (1.0 / 12.0)) * g[9]; }

TMI_HD double tri_intfff(const double *g, int i, int j, int k)
{
    int idx = (i == j && j == k) ? 2 : ((i == j || j == k || i == k) ? 1 : 0);
    double sc = idx == 2 ? 1.0 / 10.0 : (idx == 1 ? 1.0 / 30.0 : 1.0 / 60.0);
    return g[9] * sc;
}

TMI_HD double tri_intdd(const double *g, int i, int j)
{
    const double *b = g + 3, *c = g + 6;
    double scale = 0.25 / g[9];
    if (i < j) { int t = i; i = j; j = t; }   // stored lower triangle: dd(i,j), i >= j
    return scale * (b[i] * b[j] + c[i] * c[j]);
}

TMI_HD double tri_intfdd(const double *g, int j, int k) { return tri_intdd(g, j, k) * 0.3333333333; }

TMI_HD double tri_intpff(const double *g, int i, int j, double p0, double p1, double p2)
{
    if (i < j) { int t = i; i = j; j = t; }
    const double i30 = 1.0 / 30.0;
    double s;
    if (i == 0) s = 3.0 * p0 + p1 + p2;
    else if (i == 1) s = (j == 0) ? p0 + p1 + 0.5 * p2 : p0 + 3.0 * p1 + p2;
    else s = (j == 0) ? p0 + 0.5 * p1 + p2 : (j == 1) ? 0.5 * p0 + p1 + p2 : p0 + p1 + 3.0 * p2;
    return g[9] * i30 * s;
}

TMI_HD double tri_intpdd(const double *g, int i, int j, double p0, double p1, double p2)
{
    return tri_intdd(g, i, j) * ((p0 + p1 + p2) / 3.0);
}

TMI_HD double tri_bndff(const double *g, int i, int j) { return g[10 + tri_lt(i, j)]; }

TMI_HD double tri_bndintpff(const double *g, int i, int j, double p0, double p1, double p2)
{
    return tri_bndff(g, i, j) * ((p0 + p1 + p2) / 3.0);
}

}  // namespace tmi
