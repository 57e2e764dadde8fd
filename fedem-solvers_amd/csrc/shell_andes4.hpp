// 4-noded ANDES shell element (ADD_ANDES4, vpmReducer/inadd.f90:1407-1552):
// projected stiffness P^T (T^T K_loc T) P and the FQS35 lumped mass.
//
// The element is computed by one 64-lane wave.  Every step below is a
// function of (lane, work area): the kernel runs each step on all lanes and
// synchronises between steps; the same steps run lane by lane on the host
// in the element test harness, so the device arithmetic can be checked
// without a GPU.
//
// Formulation (Femlib/andes4ShellModule.f90:1054-1103, the element's
// parameters alpha_NR = 1.5, beta = 0.9 and the quadratic bending lumping):
// in the element's local frame, with the 2x2 Gauss rule,
//   K_loc = sum_g |J_g| B_g^T C B_g,   C = diag(t E_ps, t^3/12 E_ps),
//   B_g = [ L_m / A + a_H ((Q(g) - <Q>) H)  ]    membrane (u, v, theta_z)
//         [ L_b / A + a_H  (K(g) - <K>)     ]    bending  (w, theta_x, theta_y)
// where a_H = sqrt(beta), A the element area, L_m / L_b the constant-strain
// lumping tables, Q_n / K_n nodal higher-order strain tables interpolated
// bilinearly (Q(g) = sum_n N_n(g) Q_n) and <.> their area mean, H the map of
// the membrane DOFs to the seven higher-order membrane coordinates
// (four deviatoric drilling rotations, the torsional rotation and two
// hourglass translations).
//
// Nodal bending tables: at corner n the second derivative of the edge-wise
// cubic Hermite deflection along the three directions to the other corners,
//   w''_s = 6/l^2 (w_j - w_n) - 4/l (nu . theta_n) - 2/l (nu . theta_j),
// nu = (s_y, -s_x), turned into Cartesian curvatures (k_xx, k_yy, 2 k_xy)
// by solving [s_x^2, s_y^2, s_x s_y] k = w''_s for the three directions.
// Nodal membrane tables: natural strains along the two median directions
// Xi, Eta and one diagonal, linear in the higher-order coordinates with the
// element's strain-distribution factors, likewise turned Cartesian.
#pragma once
#include <cmath>

namespace fgpu {
namespace shell {

#ifndef SHELL_HD
#define SHELL_HD __host__ __device__
#endif

constexpr double ALPHA_NR = 1.5;  // normal-rotation scaling of the membrane lumping
constexpr double RHO_HO = 0.1;    // higher-order strain factor of the deviatoric rotations
constexpr double ALPHA_HO = 0.1;  // torsional-rotation factor
constexpr double BETA1 = 0.6;     // hourglass factor (0.5 + RHO_HO)
constexpr double DIAG_HO = 0.5;   // diagonal-strain factor of the deviatoric rotations (BETA1 - RHO_HO)

struct Work {
  // inputs
  double G[4][3];        // global corner coordinates
  double E, nu, rho, thk;  // material, mean thickness
  // frame and local geometry
  double ax[3][3];       // ax[k] = local axis k in global components
  double p[4][3];        // local coordinates of the corners (corner 0 at the origin)
  // membrane geometry
  double xi[2], eta[2], lxi, leta;  // unit median directions and half lengths
  double u13[2], u24[2];            // unit diagonals (corner 0 -> 2, corner 1 -> 3)
  double fxi[4], feta[4], mxi, meta, f13, f24;
  double area, hm_tor;              // element area; 1 / |Xi x Eta| (3D medians)
  double hq[4];                     // hourglass coordinate of the corner opposite to each corner
  int bad;
  // Gauss rule
  double wJ[4], N[4][4];
  // edges e = corner e -> e+1: d = p_e - p_{e+1}, squared direction cosines
  double ed[4][2], ecc[4], ess[4], ecs[4];
  // bending directions of each corner: to the previous, next and opposite corner
  double bdir[4][3][3];  // (s_x, s_y, l)
  double binv[4][9], minv[4][9];     // inverses of the direction matrices (row-major)
  double bnat[4][3][12];             // directional second derivatives (cols w0..3, tx0..3, ty0..3)
  double mnat[4][3][7];              // natural membrane strains over the higher-order coordinates
  double Kn[4][3][12], Qn[4][3][7];  // Cartesian nodal tables
  double Kbar[3][12], Qbar[3][7];
  double H[7][12];                   // membrane DOFs (u0..3, v0..3, rz0..3) -> higher-order coordinates
  double Lm[3][12], Lb[3][12];       // lumping tables (membrane: u, v, rz; bending: w, tx, ty)
  double B[4][6][24], CB[4][6][24];  // strain-displacement at the Gauss points; C |J| B (later K P)
  double K[24][25];                  // global-frame stiffness before projection
  double R[24][6], Rs[24][6];        // rigid-body basis (unit columns) and R (R^T R)^-1
  double cr[4][3];                   // corners relative to the centroid
  double Si[9];
  double P[24][25];
  double em[24];                     // lumped mass diagonal
};

SHELL_HD inline double cross2(const double* a, const double* b) { return a[0] * b[1] - a[1] * b[0]; }

// inverse of the 3x3 matrix a (row-major) into b; false when singular
// relative to the size of its rows
SHELL_HD inline bool inverse3(const double* a, double* b) {
  const double c0 = a[4] * a[8] - a[5] * a[7], c1 = a[5] * a[6] - a[3] * a[8], c2 = a[3] * a[7] - a[4] * a[6];
  const double det = a[0] * c0 + a[1] * c1 + a[2] * c2;
  double s = 0.0;
  for (int k = 0; k < 9; k++) s = fmax(s, fabs(a[k]));
  if (!(fabs(det) > 1e-14 * s * s * s)) return false;
  const double r = 1.0 / det;
  b[0] = c0 * r; b[1] = (a[2] * a[7] - a[1] * a[8]) * r; b[2] = (a[1] * a[5] - a[2] * a[4]) * r;
  b[3] = c1 * r; b[4] = (a[0] * a[8] - a[2] * a[6]) * r; b[5] = (a[2] * a[3] - a[0] * a[5]) * r;
  b[6] = c2 * r; b[7] = (a[1] * a[6] - a[0] * a[7]) * r; b[8] = (a[0] * a[4] - a[1] * a[3]) * r;
  return true;
}

// bilinear shape functions and |g_xi x g_eta| of the corners q at (s, t)
SHELL_HD inline double bilinear(double s, double t, const double (*q)[3], double* n) {
  const double ms = 1.0 - s, ps = 1.0 + s, mt = 1.0 - t, pt = 1.0 + t;
  if (n) { n[0] = 0.25 * ms * mt; n[1] = 0.25 * ps * mt; n[2] = 0.25 * ps * pt; n[3] = 0.25 * ms * pt; }
  double gs[3], gt[3];
  for (int c = 0; c < 3; c++) {
    gs[c] = 0.25 * (mt * (q[1][c] - q[0][c]) + pt * (q[2][c] - q[3][c]));
    gt[c] = 0.25 * (ms * (q[3][c] - q[0][c]) + ps * (q[2][c] - q[1][c]));
  }
  const double a = gs[1] * gt[2] - gs[2] * gt[1], b = gs[2] * gt[0] - gs[0] * gt[2], c = gs[0] * gt[1] - gs[1] * gt[0];
  return sqrt(a * a + b * b + c * c);
}

// natural corner coordinates of the bilinear map
SHELL_HD inline double corner_xi(int n) { return (n == 1 || n == 2) ? 1.0 : -1.0; }
SHELL_HD inline double corner_eta(int n) { return n >= 2 ? 1.0 : -1.0; }

// ---- step 1 (lane 0): element frame (trans3P): x along corner 0 -> 1, z
// normal to it and to the line from corner 0 to the midpoint of corners 2, 3
SHELL_HD inline void step_frame(int lane, Work& w) {
  if (lane != 0) return;
  double a[3], m[3], z[3];
  for (int c = 0; c < 3; c++) {
    a[c] = w.G[1][c] - w.G[0][c];
    m[c] = 0.5 * (w.G[2][c] + w.G[3][c]) - w.G[0][c];
  }
  z[0] = a[1] * m[2] - a[2] * m[1];
  z[1] = a[2] * m[0] - a[0] * m[2];
  z[2] = a[0] * m[1] - a[1] * m[0];
  double la = a[0] * a[0] + a[1] * a[1] + a[2] * a[2], lz = z[0] * z[0] + z[1] * z[1] + z[2] * z[2];
  if (!(la > 1e-16) || !(lz > 1e-16)) {
    w.bad = 1;
    la = lz = 1.0;
  }
  la = sqrt(la);
  lz = sqrt(lz);
  for (int c = 0; c < 3; c++) {
    w.ax[0][c] = a[c] / la;
    w.ax[2][c] = z[c] / lz;
  }
  w.ax[1][0] = w.ax[2][1] * w.ax[0][2] - w.ax[2][2] * w.ax[0][1];
  w.ax[1][1] = w.ax[2][2] * w.ax[0][0] - w.ax[2][0] * w.ax[0][2];
  w.ax[1][2] = w.ax[2][0] * w.ax[0][1] - w.ax[2][1] * w.ax[0][0];
}

// ---- step 2 (lanes 0-11): local corner coordinates; rigid-body offsets
SHELL_HD inline void step_local(int lane, Work& w) {
  if (lane < 12) {
    const int n = lane / 3, c = lane % 3;
    const double* e = w.ax[c];
    w.p[n][c] = e[0] * (w.G[n][0] - w.G[0][0]) + e[1] * (w.G[n][1] - w.G[0][1]) + e[2] * (w.G[n][2] - w.G[0][2]);
  } else if (lane >= 16 && lane < 28) {
    // corner minus corner 0, minus the mean of those (pmatStiff.f90:60-72)
    const int n = (lane - 16) / 3, c = (lane - 16) % 3;
    const double mean = 0.25 * ((w.G[0][c] - w.G[0][c]) + (w.G[1][c] - w.G[0][c]) + (w.G[2][c] - w.G[0][c]) +
                                (w.G[3][c] - w.G[0][c]));
    w.cr[n][c] = (w.G[n][c] - w.G[0][c]) - mean;
  }
}

// ---- step 3: element geometry, split over lanes by role
SHELL_HD inline void step_geometry(int lane, Work& w) {
  if (lane == 0) {
    // median directions, diagonals and the membrane strain-distribution
    // factors (andes4ShellModule.f90:479-600)
    double r[4][2];
    const double cx = 0.25 * (w.p[1][0] + w.p[2][0] + w.p[3][0]), cy = 0.25 * (w.p[1][1] + w.p[2][1] + w.p[3][1]);
    for (int n = 0; n < 4; n++) { r[n][0] = w.p[n][0] - cx; r[n][1] = w.p[n][1] - cy; }
    double X[2], Y[2];
    for (int c = 0; c < 2; c++) {
      X[c] = (r[1][c] + r[2][c]) - (r[0][c] + r[3][c]);
      Y[c] = (r[2][c] + r[3][c]) - (r[0][c] + r[1][c]);
    }
    const double nx = sqrt(X[0] * X[0] + X[1] * X[1]), ny = sqrt(Y[0] * Y[0] + Y[1] * Y[1]);
    if (!(nx > 0.0) || !(ny > 0.0)) { w.bad = 1; return; }
    w.xi[0] = X[0] / nx; w.xi[1] = X[1] / nx; w.lxi = 0.5 * nx;
    w.eta[0] = Y[0] / ny; w.eta[1] = Y[1] / ny; w.leta = 0.5 * ny;
    const double d13[2] = {w.p[2][0] - w.p[0][0], w.p[2][1] - w.p[0][1]};
    const double d24[2] = {w.p[3][0] - w.p[1][0], w.p[3][1] - w.p[1][1]};
    const double l13 = sqrt(d13[0] * d13[0] + d13[1] * d13[1]), l24 = sqrt(d24[0] * d24[0] + d24[1] * d24[1]);
    w.u13[0] = d13[0] / l13; w.u13[1] = d13[1] / l13;
    w.u24[0] = d24[0] / l24; w.u24[1] = d24[1] / l24;
    w.f13 = 0.5 * cross2(w.u13, d24) / l13;
    w.f24 = -0.5 * cross2(w.u24, d13) / l24;
    // distance factors of the corners across each median, centred, then
    // signed so that they are positive on a parallelogram
    double fx[4], fe[4], sx = 0.0, se = 0.0;
    for (int n = 0; n < 4; n++) {
      fx[n] = cross2(w.xi, r[n]) / w.lxi;
      fe[n] = cross2(w.eta, r[n]) / w.leta;
      sx += fx[n];
      se += fe[n];
    }
    double mx = 0.0, me = 0.0;
    for (int n = 0; n < 4; n++) {
      w.fxi[n] = corner_eta(n) * (fx[n] - 0.25 * sx);
      w.feta[n] = -corner_xi(n) * (fe[n] - 0.25 * se);
      mx += w.fxi[n];
      me += w.feta[n];
    }
    w.mxi = 0.25 * mx;
    w.meta = 0.25 * me;
  } else if (lane == 1) {
    // element area (two triangles 0-1-2 and 2-3-0)
    const double* q0 = w.p[0];
    const double* q1 = w.p[1];
    const double* q2 = w.p[2];
    const double* q3 = w.p[3];
    w.area = 0.5 * ((q1[0] - q0[0]) * (q2[1] - q0[1]) - (q2[0] - q0[0]) * (q1[1] - q0[1]) +
                    (q3[0] - q2[0]) * (q0[1] - q2[1]) - (q0[0] - q2[0]) * (q3[1] - q2[1]));
    if (!(w.area > 0.0)) w.bad = 1;
  } else if (lane == 2) {
    // higher-order membrane map scalars (andes4ShellModule.f90:753-863):
    // the medians in 3D for the torsional row, the natural coordinates of
    // the centred corners under the centre Jacobian for the hourglass rows
    double r[4][3], c[3];
    for (int k = 0; k < 3; k++) c[k] = 0.25 * (w.p[1][k] + w.p[2][k] + w.p[3][k]);
    for (int n = 0; n < 4; n++)
      for (int k = 0; k < 3; k++) r[n][k] = w.p[n][k] - c[k];
    const double jc = bilinear(0.0, 0.0, r, nullptr);
    if (!(jc > 0.0)) { w.bad = 1; return; }
    w.hm_tor = 0.0625 / jc;
    double gs[2], gt[2];
    for (int k = 0; k < 2; k++) {
      gs[k] = 0.25 * ((r[1][k] + r[2][k]) - (r[0][k] + r[3][k]));
      gt[k] = 0.25 * ((r[2][k] + r[3][k]) - (r[0][k] + r[1][k]));
    }
    // (s, t) of r solves [gs gt] (s, t) = r
    const double det = gs[0] * gt[1] - gs[1] * gt[0];
    if (!(fabs(det) > 0.0)) { w.bad = 1; return; }
    double st[4], mean = 0.0;
    for (int n = 0; n < 4; n++) {
      const double s = (gt[1] * r[n][0] - gt[0] * r[n][1]) / det;
      const double t = (gs[0] * r[n][1] - gs[1] * r[n][0]) / det;
      st[n] = s * t;
      mean += st[n];
    }
    mean *= 0.25;
    for (int n = 0; n < 4; n++) w.hq[n] = st[(n + 2) % 4] - mean;
  } else if (lane >= 4 && lane < 8) {
    // Gauss point g = 2 i + j at (xi_i, eta_j), unit weights
    const int g = lane - 4;
    const double q = sqrt(1.0 / 3.0);
    w.wJ[g] = bilinear(g < 2 ? -q : q, (g & 1) ? q : -q, w.p, w.N[g]);
  } else if (lane >= 8 && lane < 12) {
    const int e = lane - 8, f = (e + 1) % 4;
    const double dx = w.p[e][0] - w.p[f][0], dy = w.p[e][1] - w.p[f][1], l2 = dx * dx + dy * dy;
    w.ed[e][0] = dx;
    w.ed[e][1] = dy;
    w.ecc[e] = dx * dx / l2;
    w.ess[e] = dy * dy / l2;
    w.ecs[e] = dx * dy / l2;
  } else if (lane >= 16 && lane < 28) {
    // bending direction k of corner n: previous, next, opposite corner
    const int n = (lane - 16) / 3, k = (lane - 16) % 3;
    const int j = (n + (k == 0 ? 3 : k == 1 ? 1 : 2)) % 4;
    const double dx = w.p[j][0] - w.p[n][0], dy = w.p[j][1] - w.p[n][1];
    const double l = sqrt(dx * dx + dy * dy);
    w.bdir[n][k][0] = dx / l;
    w.bdir[n][k][1] = dy / l;
    w.bdir[n][k][2] = l;
  }
}

// ---- step 4: direction-matrix inverses, natural tables, H, lumping
SHELL_HD inline void step_tables(int lane, Work& w) {
  if (lane < 4) {
    const int n = lane;
    double a[9];
    for (int k = 0; k < 3; k++) {
      const double sx = w.bdir[n][k][0], sy = w.bdir[n][k][1];
      a[3 * k] = sx * sx; a[3 * k + 1] = sy * sy; a[3 * k + 2] = sx * sy;
    }
    if (!inverse3(a, w.binv[n])) w.bad = 1;
  } else if (lane < 8) {
    // membrane directions of corner n: Xi, Eta and the diagonal not
    // through n
    const int n = lane - 4;
    const double* d = (n % 2 == 0) ? w.u24 : w.u13;
    const double a[9] = {w.xi[0] * w.xi[0], w.xi[1] * w.xi[1], w.xi[0] * w.xi[1],
                         w.eta[0] * w.eta[0], w.eta[1] * w.eta[1], w.eta[0] * w.eta[1],
                         d[0] * d[0], d[1] * d[1], d[0] * d[1]};
    if (!inverse3(a, w.minv[n])) w.bad = 1;
  } else if (lane < 20) {
    // natural membrane strain r of corner n over (rd_0..3, r_tor, t_xi, t_eta)
    const int n = (lane - 8) / 3, r = (lane - 8) % 3;
    double* m = w.mnat[n][r];
    const double cx = corner_xi(n), ce = corner_eta(n), alt = cx * ce;
    for (int k = 0; k < 7; k++) m[k] = 0.0;
    if (r == 0) {  // along Xi
      for (int k = 0; k < 4; k++) m[k] = RHO_HO * w.fxi[n] * ce * corner_xi(k);
      m[4] = alt * ALPHA_HO * w.leta / w.lxi;
      m[5] = BETA1 * ce * w.fxi[n] / (w.mxi * w.lxi);
    } else if (r == 1) {  // along Eta
      for (int k = 0; k < 4; k++) m[k] = -RHO_HO * w.feta[n] * cx * corner_eta(k);
      m[4] = -alt * ALPHA_HO * w.lxi / w.leta;
      m[6] = BETA1 * cx * w.feta[n] / (w.meta * w.leta);
    } else {  // along the diagonal not through n: the two neighbours
      const double f = (n % 2 == 0) ? w.f24 : w.f13;
      m[(n + 1) % 4] = DIAG_HO * f;
      m[(n + 3) % 4] = -DIAG_HO * f;
    }
  } else if (lane < 32) {
    // second derivative along bending direction k of corner n (cubic
    // Hermite deflection of the edge to corner j)
    const int n = (lane - 20) / 3, k = (lane - 20) % 3;
    const int j = (n + (k == 0 ? 3 : k == 1 ? 1 : 2)) % 4;
    const double sx = w.bdir[n][k][0], sy = w.bdir[n][k][1], l = w.bdir[n][k][2];
    double* b = w.bnat[n][k];
    for (int c = 0; c < 12; c++) b[c] = 0.0;
    b[n] = -6.0 / (l * l);
    b[j] = 6.0 / (l * l);
    b[4 + n] = -4.0 * sy / l;  // nu = (s_y, -s_x) against (theta_x, theta_y)
    b[8 + n] = 4.0 * sx / l;
    b[4 + j] = -2.0 * sy / l;
    b[8 + j] = 2.0 * sx / l;
  } else {
    // H (84 entries) and the lumping tables (2 x 36), 32 lanes
    for (int idx = lane - 32; idx < 84 + 72; idx += 32) {
      if (idx < 84) {
        const int k = idx / 12, c = idx % 12, n = c % 4;
        double v = 0.0;
        if (k < 4) {  // deviatoric drilling rotation of corner k
          if (c >= 8) v = (n == k ? 1.0 : 0.0) - 0.25;
        } else if (k == 4) {  // torsion: mean drilling rotation minus the centre rotation
          const int a = (n + 3) % 4, b = (n + 1) % 4, comp = c / 4;
          v = c >= 8 ? 0.25 : (w.p[a][comp] - w.p[b][comp]) * w.hm_tor;
        } else if (c < 8) {  // hourglass translation along Xi (k = 5) or Eta (k = 6)
          const double* dir = k == 5 ? w.xi : w.eta;
          v = w.hq[n] * dir[c / 4];
        }
        w.H[k][c] = v;
      } else {
        const int q = idx - 84, tab = q / 36, r = (q % 36) / 12, c = q % 12, n = c % 4, kind = c / 4;
        // edges at corner n: en from n to n+1 (vector p_n - p_{n+1}), ep
        // from n-1 to n (vector p_{n-1} - p_n)
        const int en = n, ep = (n + 3) % 4;
        const double* dn = w.ed[en];
        const double* dp = w.ed[ep];
        double v;
        if (tab == 0) {
          // membrane lumping (lump4mem): u, v by the opposite-neighbour
          // difference, drilling rotation by the two edges at n
          const double jx = w.p[(n + 1) % 4][0] - w.p[(n + 3) % 4][0];
          const double jy = w.p[(n + 1) % 4][1] - w.p[(n + 3) % 4][1];
          if (kind == 0) v = r == 0 ? 0.5 * jy : r == 1 ? 0.0 : -0.5 * jx;
          else if (kind == 1) v = r == 0 ? 0.0 : r == 1 ? -0.5 * jx : 0.5 * jy;
          else if (r == 0) v = (dp[1] * dp[1] - dn[1] * dn[1]) * ALPHA_NR / 12.0;
          else if (r == 1) v = (dp[0] * dp[0] - dn[0] * dn[0]) * ALPHA_NR / 12.0;
          else v = (dn[0] * dn[1] - dp[0] * dp[1]) * ALPHA_NR / 6.0;
          w.Lm[r][c] = v;
        } else {
          // bending lumping (lump4ben, quadratic rotations along the edges)
          if (kind == 0) {  // w
            v = r == 0 ? -w.ecs[en] + w.ecs[ep] : r == 1 ? w.ecs[en] - w.ecs[ep]
                                                         : (w.ess[ep] - w.ecc[ep]) - (w.ess[en] - w.ecc[en]);
          } else if (kind == 1) {  // theta_x
            v = r == 0 ? 0.5 * (w.ess[en] * dn[0] + w.ess[ep] * dp[0])
                : r == 1 ? 0.5 * (w.ecc[en] * dn[0] + w.ecc[ep] * dp[0])
                         : -w.ecc[en] * dn[1] - w.ecc[ep] * dp[1];
          } else {  // theta_y
            v = r == 0 ? 0.5 * (w.ess[en] * dn[1] + w.ess[ep] * dp[1])
                : r == 1 ? 0.5 * (w.ecc[en] * dn[1] + w.ecc[ep] * dp[1])
                         : -w.ess[en] * dn[0] - w.ess[ep] * dp[0];
          }
          w.Lb[r][c] = v;
        }
      }
    }
  }
}

// ---- step 5: Cartesian nodal tables (84 membrane + 144 bending entries)
SHELL_HD inline void step_cartesian(int lane, Work& w) {
  for (int idx = lane; idx < 84 + 144; idx += 64) {
    if (idx < 84) {
      const int n = idx / 21, r = (idx % 21) / 7, k = idx % 7;
      const double* a = w.minv[n] + 3 * r;
      w.Qn[n][r][k] = a[0] * w.mnat[n][0][k] + a[1] * w.mnat[n][1][k] + a[2] * w.mnat[n][2][k];
    } else {
      const int q = idx - 84, n = q / 36, r = (q % 36) / 12, c = q % 12;
      const double* a = w.binv[n] + 3 * r;
      w.Kn[n][r][c] = a[0] * w.bnat[n][0][c] + a[1] * w.bnat[n][1][c] + a[2] * w.bnat[n][2][c];
    }
  }
}

// ---- step 6: area means of the interpolated tables (21 + 36 entries)
SHELL_HD inline void step_means(int lane, Work& w) {
  if (lane >= 57) return;
  const bool mem = lane < 21;
  const int r = mem ? lane / 7 : (lane - 21) / 12, c = mem ? lane % 7 : (lane - 21) % 12;
  double s = 0.0, a = 0.0;
  for (int g = 0; g < 4; g++) {
    double v = 0.0;
    for (int n = 0; n < 4; n++) v += w.N[g][n] * (mem ? w.Qn[n][r][c] : w.Kn[n][r][c]);
    s += w.wJ[g] * v;
    a += w.wJ[g];
  }
  if (mem) w.Qbar[r][c] = s / a;
  else w.Kbar[r][c] = s / a;
}

// ---- step 7: B at the four Gauss points (576 entries)
SHELL_HD inline void step_strain(int lane, Work& w) {
  const double aH = sqrt(0.9);
  for (int idx = lane; idx < 4 * 6 * 24; idx += 64) {
    const int g = idx / 144, row = (idx % 144) / 24, d = idx % 24, n = d / 6, comp = d % 6;
    double v = 0.0;
    if (row < 3) {
      const int c = comp == 0 ? n : comp == 1 ? 4 + n : comp == 5 ? 8 + n : -1;
      if (c >= 0) {
        double h = 0.0;
        for (int k = 0; k < 7; k++) {
          double q = -w.Qbar[row][k];
          for (int m = 0; m < 4; m++) q += w.N[g][m] * w.Qn[m][row][k];
          h += q * w.H[k][c];
        }
        v = w.Lm[row][c] / w.area + aH * h;
      }
    } else {
      const int r = row - 3, c = comp == 2 ? n : comp == 3 ? 4 + n : comp == 4 ? 8 + n : -1;
      if (c >= 0) {
        double q = -w.Kbar[r][c];
        for (int m = 0; m < 4; m++) q += w.N[g][m] * w.Kn[m][r][c];
        v = w.Lb[r][c] / w.area + aH * q;
      }
    }
    w.B[g][row][d] = v;
  }
}

// ---- step 8: |J_g| C B_g (plane-stress C scaled by t and t^3/12)
SHELL_HD inline void step_cb(int lane, Work& w) {
  const double c11 = w.E / (1.0 - w.nu * w.nu), c12 = w.nu * c11, c33 = 0.5 * w.E / (1.0 + w.nu);
  const double t = w.thk, tb = t * t * t / 12.0;
  for (int idx = lane; idx < 4 * 6 * 24; idx += 64) {
    const int g = idx / 144, row = (idx % 144) / 24, d = idx % 24, r0 = row < 3 ? 0 : 3, rr = row - r0;
    const double f = (row < 3 ? t : tb) * w.wJ[g];
    const double e1 = w.B[g][r0][d], e2 = w.B[g][r0 + 1][d], e3 = w.B[g][r0 + 2][d];
    w.CB[g][row][d] = rr == 0 ? f * (c11 * e1 + c12 * e2) : rr == 1 ? f * (c12 * e1 + c11 * e2) : f * c33 * e3;
  }
}

// ---- step 9: 3x3 block (a, b) of K_loc = sum_g B_g^T (C|J| B_g), rotated
// to the global axes (lane = 8 a + b)
SHELL_HD inline void step_stiffness(int lane, Work& w) {
  const int a = lane / 8, b = lane % 8;
  double k[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double s = 0.0;
      for (int g = 0; g < 4; g++)
        for (int r = 0; r < 6; r++) s += w.B[g][r][3 * a + i] * w.CB[g][r][3 * b + j];
      k[i][j] = s;
    }
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double s = 0.0;
      for (int p = 0; p < 3; p++)
        s += w.ax[p][i] * (k[p][0] * w.ax[0][j] + k[p][1] * w.ax[1][j] + k[p][2] * w.ax[2][j]);
      w.K[3 * a + i][3 * b + j] = s;
    }
}

// ---- step 10: rigid-body basis of the projector (pmatStiff.f90:23-127):
// three translations and three rotations about the centroid, each column
// scaled to unit length (lanes 0-5: one column each)
SHELL_HD inline void step_rigid(int lane, Work& w) {
  if (lane >= 6) return;
  const int k = lane;
  double col[24];
  for (int i = 0; i < 24; i++) col[i] = 0.0;
  for (int n = 0; n < 4; n++) {
    const double* x = w.cr[n];
    double* u = col + 6 * n;
    if (k < 3) {
      u[k] = 1.0;
    } else {
      const int a = k - 3;  // rotation about axis a: displacement e_a x x
      const int b = (a + 1) % 3, c = (a + 2) % 3;
      u[b] = -x[c];
      u[c] = x[b];
      u[3 + a] = 1.0;
    }
  }
  double s = 0.0;
  for (int i = 0; i < 24; i++) s += col[i] * col[i];
  s = sqrt(s);
  for (int i = 0; i < 24; i++) w.R[i][k] = col[i] / s;
}

// ---- step 11 (lane 0): inverse Gram matrix of the rotation columns
SHELL_HD inline void step_gram(int lane, Work& w) {
  if (lane != 0) return;
  double g[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double s = 0.0;
      for (int q = 0; q < 24; q++) s += w.R[q][3 + i] * w.R[q][3 + j];
      g[3 * i + j] = s;
    }
  if (!inverse3(g, w.Si)) w.bad = 1;
}

// ---- step 12: P = I - [R_t, R_r S] R^T (576 entries; rows 24 lanes first)
SHELL_HD inline void step_projector(int lane, Work& w) {
  if (lane < 24) {
    const int i = lane;
    for (int j = 0; j < 3; j++) {
      w.Rs[i][j] = w.R[i][j];
      w.Rs[i][3 + j] = w.R[i][3] * w.Si[j] + w.R[i][4] * w.Si[3 + j] + w.R[i][5] * w.Si[6 + j];
    }
  }
}
SHELL_HD inline void step_projector2(int lane, Work& w) {
  for (int idx = lane; idx < 576; idx += 64) {
    const int i = idx / 24, j = idx % 24;
    double s = 0.0;
    for (int k = 0; k < 6; k++) s += w.Rs[i][k] * w.R[j][k];
    w.P[i][j] = (i == j ? 1.0 : 0.0) - s;
  }
}

// ---- step 13: K P into CB (as a 24 x 24 array)
SHELL_HD inline void step_kp(int lane, Work& w) {
  double* KP = &w.CB[0][0][0];
  for (int idx = lane; idx < 576; idx += 64) {
    const int i = idx / 24, j = idx % 24;
    double s = 0.0;
    for (int k = 0; k < 24; k++) s += w.K[i][k] * w.P[k][j];
    KP[i * 24 + j] = s;
  }
}

// entry idx (< 360) of the packed lower node-pair blocks of P^T (K P):
// block (a, b), a >= b, at a (a + 1) / 2 + b, row-major 6 x 6
SHELL_HD inline double projected_entry(int idx, const Work& w, int& r, int& c) {
  const int blk = idx / 36, q = idx % 36;
  int a = 0;
  while ((a + 1) * (a + 2) / 2 <= blk) a++;
  const int b = blk - a * (a + 1) / 2;
  r = 6 * a + q / 6;
  c = 6 * b + q % 6;
  const double* KP = &w.CB[0][0][0];
  double s = 0.0;
  for (int k = 0; k < 24; k++) s += w.P[k][r] * KP[k * 24 + c];
  return s;
}

// ---- lumped mass (FQS35 = QMRF35 + QBEK35, fqs.f:550-583): each triangle
// (0, 1, 2) and (0, 2, 3) gives a third of its mass and a twentieth of its
// centroidal moments of inertia to each of its corners; the element then
// gives every corner the mean over its corners.  The triangle's moments
// about its local axes (x along its first edge, z normal) are those of the
// two right triangles the apex height splits it into, each shifted to the
// triangle's centroid; the thickness adds m t^2 / 12 about the in-plane
// axes.  The fractions 2/3, 1/3 and 1/20 are the single-precision values
// the reference multiplies by (qmrf.f:1566-1597, qbek.f:1845).
SHELL_HD inline bool triangle_inertia(const Work& w, int tri, double& m, double* I) {
  const float f23 = 2.0f / 3.0f, f13 = 1.0f / 3.0f;
  const double c23 = f23, c13 = f13;
  const double* x0 = w.G[0];
  const double* x1 = w.G[tri + 1];
  const double* x2 = w.G[tri + 2];
  double e[3], f[3];
  for (int c = 0; c < 3; c++) { e[c] = x1[c] - x0[c]; f[c] = x2[c] - x0[c]; }
  const double le = sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]), lf = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
  const double cg = (f[0] * e[0] + f[1] * e[1] + f[2] * e[2]) / (lf * le);
  const double sg2 = 1.0 - cg * cg, sg = sg2 > 0.0 ? sqrt(sg2) : 0.0;
  const double base = le, b1 = lf * cg, h = lf * sg, b2 = base - b1;
  const double area = base * h / 2.0;
  if (!(area > 0.0)) return false;
  const double rt = w.rho * w.thk, t2 = w.thk * w.thk;
  const double xc = (b1 + b2 / 2.0) * c23;  // centroid along the base
  double iyy = 0.0, izz = 0.0;
  for (int piece = 0; piece < 2; piece++) {
    const double wd = piece == 0 ? b1 : b2;
    const double xp = piece == 0 ? c23 * b1 : b1 + c13 * b2;  // centroid of the right triangle
    const double mp = wd * h / 2.0 * rt, dx = xp - xc;
    const double s2 = (c23 * wd) * (c23 * wd) + (c13 * h) * (c13 * h);  // acute corner to centroid
    iyy += rt * wd * h * (wd * wd / 36.0 + t2 / 24.0) + dx * dx * mp;
    izz += rt * h * wd * (wd * wd / 4.0 + h * h / 12.0) - s2 * mp + dx * dx * mp;
  }
  const double ixx = rt * base * h * (h * h / 36.0 + t2 / 24.0);
  // local axes of the triangle in global components
  double a1[3], a3[3], a2[3];
  for (int c = 0; c < 3; c++) a1[c] = e[c] / le;
  a3[0] = a1[1] * f[2] - f[1] * a1[2];
  a3[1] = f[0] * a1[2] - a1[0] * f[2];
  a3[2] = a1[0] * f[1] - f[0] * a1[1];
  const double l3 = sqrt(a3[0] * a3[0] + a3[1] * a3[1] + a3[2] * a3[2]);
  for (int c = 0; c < 3; c++) a3[c] /= l3;
  a2[0] = a3[1] * a1[2] - a1[1] * a3[2];
  a2[1] = a1[0] * a3[2] - a3[0] * a1[2];
  a2[2] = a3[0] * a1[1] - a1[0] * a3[1];
  for (int c = 0; c < 3; c++) I[c] = fabs(a1[c] * ixx + a2[c] * iyy + a3[c] * izz);
  m = rt * area / 3.0;
  return true;
}

// ---- step 14 (lanes 0-1 then lane 0): lumped mass diagonal
SHELL_HD inline void step_mass(int lane, Work& w) {
  if (lane != 0) return;
  double m[2], I[2][3];
  if (!triangle_inertia(w, 0, m[0], I[0]) || !triangle_inertia(w, 1, m[1], I[1])) {
    w.bad = 1;
    for (int k = 0; k < 24; k++) w.em[k] = 0.0;
    return;
  }
  const float f120 = 1.0f / 20.0f;
  const double g = f120;
  // corners 0 and 2 belong to both triangles, 1 to the first, 3 to the second
  auto mean4 = [](double a, double b) { return ((a + b) + a + (a + b) + b) / 4.0; };
  const double mt = mean4(m[0], m[1]);
  const double ir[3] = {mean4(g * I[0][0], g * I[1][0]), mean4(g * I[0][1], g * I[1][1]),
                        mean4(g * I[0][2], g * I[1][2])};
  for (int n = 0; n < 4; n++)
    for (int c = 0; c < 3; c++) {
      w.em[6 * n + c] = mt;
      w.em[6 * n + 3 + c] = ir[c];
    }
}

// The element as a sequence of phases: within a phase no lane reads what
// another lane of the same phase writes, so a barrier between phases is the
// only synchronisation (the host harness runs the lanes of a phase one
// after another).  Inputs G, E, nu, rho, thk and bad = 0 are set before
// phase 0; afterwards CB holds K P (see projected_entry) and em the mass.
constexpr int NPHASE = 11;
SHELL_HD inline void phase(int k, int lane, Work& w) {
  switch (k) {
    case 0: step_frame(lane, w); break;
    case 1: step_local(lane, w); break;
    case 2:
      step_geometry(lane, w);
      step_rigid(lane >= 58 ? lane - 58 : 64, w);  // lanes 58-63 (read cr only)
      break;
    case 3:
      step_tables(lane, w);
      step_gram(lane == 63 ? 0 : 1, w);  // lane 63 (reads R only)
      break;
    case 4:
      step_cartesian(lane, w);
      step_projector(lane, w);
      step_mass(lane == 62 ? 0 : 1, w);  // lane 62 (reads G and the material only)
      break;
    case 5:
      step_means(lane, w);
      step_projector2(lane, w);
      break;
    case 6: step_strain(lane, w); break;
    case 7: step_cb(lane, w); break;
    case 8: step_stiffness(lane, w); break;
    case 9: step_kp(lane, w); break;
    default: break;
  }
}

}  // namespace shell
}  // namespace fgpu
