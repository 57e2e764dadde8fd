// gfx950 kernel of the 4-noded ANDES shell element (ADD_ANDES4,
// vpmReducer/inadd.f90:1407-1552): stiffness P^T (T^T K_loc T) P and the
// FQS35 lumped mass, written as packed 6x6 node-pair blocks.
//
// One 64-lane workgroup per element.  The element-level geometry of the
// ANDES formulation (Femlib/andes4ShellModule.f90: nodal membrane and
// bending strain tables, their Gauss means, the hierarchical membrane map)
// is small serial work that four lanes split; the B matrices at the 2x2
// Gauss points, C*B, the 64 3x3 blocks of the local stiffness, the rigid-body
// projector and P^T K P are spread over all 64 lanes through LDS.
#include <algorithm>
#include "kernels.hpp"

namespace fgpu {
namespace {

#define D4(a, i, j) (a)[((i) - 1) + 4 * ((j) - 1)]

__device__ int inv33(const double* a, double* b) {
  // column-major 3x3 inverse by cofactors (manipMatrixModule.f90:373-415)
#define A(i, j) a[((i) - 1) + 3 * ((j) - 1)]
#define B(i, j) b[((i) - 1) + 3 * ((j) - 1)]
  double det = A(1, 1) * (A(2, 2) * A(3, 3) - A(3, 2) * A(2, 3)) - A(1, 2) * (A(2, 1) * A(3, 3) - A(3, 1) * A(2, 3)) +
               A(1, 3) * (A(2, 1) * A(3, 2) - A(3, 1) * A(2, 2));
  if (fabs(det) < 1.0e-16) return -1;
  B(1, 1) = (A(2, 2) * A(3, 3) - A(3, 2) * A(2, 3)) / det;
  B(1, 2) = -(A(1, 2) * A(3, 3) - A(3, 2) * A(1, 3)) / det;
  B(1, 3) = (A(1, 2) * A(2, 3) - A(2, 2) * A(1, 3)) / det;
  B(2, 1) = -(A(2, 1) * A(3, 3) - A(3, 1) * A(2, 3)) / det;
  B(2, 2) = (A(1, 1) * A(3, 3) - A(3, 1) * A(1, 3)) / det;
  B(2, 3) = -(A(1, 1) * A(2, 3) - A(2, 1) * A(1, 3)) / det;
  B(3, 1) = (A(2, 1) * A(3, 2) - A(3, 1) * A(2, 2)) / det;
  B(3, 2) = -(A(1, 1) * A(3, 2) - A(3, 1) * A(1, 2)) / det;
  B(3, 3) = (A(1, 1) * A(2, 2) - A(2, 1) * A(1, 2)) / det;
#undef A
#undef B
  return 0;
}

__device__ void shape4(double xsi, double eta, double* nv) {
  nv[0] = (1.0 - xsi) * (1.0 - eta) * 0.25;
  nv[1] = (1.0 + xsi) * (1.0 - eta) * 0.25;
  nv[2] = (1.0 + xsi) * (1.0 + eta) * 0.25;
  nv[3] = (1.0 - xsi) * (1.0 + eta) * 0.25;
}

// |g_xsi x g_eta| of the bilinear map of the four points (andes4ShellModule.f90:392-438)
__device__ double jacobi_quad(double xsi, double eta, const double (*xyz)[3], double (*jac)[3]) {
  const double nx[4] = {-(1.0 - eta) * 0.25, (1.0 - eta) * 0.25, (1.0 + eta) * 0.25, -(1.0 + eta) * 0.25};
  const double ne[4] = {-(1.0 - xsi) * 0.25, -(1.0 + xsi) * 0.25, (1.0 + xsi) * 0.25, (1.0 - xsi) * 0.25};
  double gx[3], ge[3];
  for (int c = 0; c < 3; c++) {
    gx[c] = nx[0] * xyz[0][c] + nx[1] * xyz[1][c] + nx[2] * xyz[2][c] + nx[3] * xyz[3][c];
    ge[c] = ne[0] * xyz[0][c] + ne[1] * xyz[1][c] + ne[2] * xyz[2][c] + ne[3] * xyz[3][c];
  }
  if (jac)
    for (int c = 0; c < 3; c++) { jac[0][c] = gx[c]; jac[1][c] = ge[c]; }
  const double n0 = gx[1] * ge[2] - gx[2] * ge[1];
  const double n1 = gx[2] * ge[0] - gx[0] * ge[2];
  const double n2 = gx[0] * ge[1] - gx[1] * ge[0];
  return sqrt(n0 * n0 + n1 * n1 + n2 * n2);
}

__device__ double cross2d(const double* a, const double* b) { return a[0] * b[1] - b[0] * a[1]; }

// nodal bending strains bb[n*36 + r*12 + c] (andes4ShellModule.f90:189-290)
__device__ int bben_nod(const double* xij, const double* yij, double* bb) {
  const int stnod[4][3] = {{4, 2, 3}, {1, 3, 4}, {2, 4, 1}, {3, 1, 2}};
  double lij[4][4], sij[4][4][2], nij[4][4][2];
  for (int k = 0; k < 144; k++) bb[k] = 0.0;
  for (int i = 1; i <= 4; i++)
    for (int k = 1; k <= 3; k++) {
      const int j = 1 + (i + k - 1) % 4;
      const double l = sqrt(D4(xij, i, j) * D4(xij, i, j) + D4(yij, i, j) * D4(yij, i, j));
      lij[i - 1][j - 1] = l;
      lij[j - 1][i - 1] = l;
      sij[i - 1][j - 1][0] = D4(xij, j, i) / l;
      sij[i - 1][j - 1][1] = D4(yij, j, i) / l;
      nij[i - 1][j - 1][0] = sij[i - 1][j - 1][1];
      nij[i - 1][j - 1][1] = -sij[i - 1][j - 1][0];
      sij[j - 1][i - 1][0] = -sij[i - 1][j - 1][0];
      sij[j - 1][i - 1][1] = -sij[i - 1][j - 1][1];
      nij[j - 1][i - 1][0] = -nij[i - 1][j - 1][0];
      nij[j - 1][i - 1][1] = -nij[i - 1][j - 1][1];
    }
  for (int in = 1; in <= 4; in++)
    for (int r = 1; r <= 3; r++) {
      const int jn = stnod[in - 1][r - 1];
      const double l = lij[in - 1][jn - 1];
      double* b = bb + (in - 1) * 36 + (r - 1) * 12;
      b[in - 1] = -6.0 / (l * l);
      b[in + 3] = -4.0 * nij[in - 1][jn - 1][0] / l;
      b[in + 7] = -4.0 * nij[in - 1][jn - 1][1] / l;
      b[jn - 1] = -b[in - 1];
      b[jn + 3] = b[in + 3] * 0.5;
      b[jn + 7] = b[in + 7] * 0.5;
    }
  for (int in = 1; in <= 4; in++) {
    double tinv[9], tran[9];
    for (int r = 1; r <= 3; r++) {
      const double* s = sij[in - 1][stnod[in - 1][r - 1] - 1];
      tinv[(r - 1) + 0] = s[0] * s[0];
      tinv[(r - 1) + 3] = s[1] * s[1];
      tinv[(r - 1) + 6] = s[0] * s[1];
    }
    if (inv33(tinv, tran)) return -1;
    double* b = bb + (in - 1) * 36;
    for (int c = 0; c < 12; c++) {
      double v[3];
      for (int r = 0; r < 3; r++) v[r] = tran[r] * b[c] + tran[r + 3] * b[12 + c] + tran[r + 6] * b[24 + c];
      for (int r = 0; r < 3; r++) b[r * 12 + c] = v[r];
    }
  }
  return 0;
}

// nodal higher-order membrane strains bm[n*21 + r*7 + c] (andes4ShellModule.f90:479-750)
__device__ int bmqh_nod(const double* xij, const double* yij, double* bm) {
  const double rho1 = 0.1, rho2 = -rho1, rho3 = rho2, rho4 = rho1, alpha = 0.1;
  const double beta1 = 0.5 + rho1, beta2 = 0.0, rho5 = 0.0, rho6 = beta1 - rho1, rho7 = 0.0, rho8 = -rho6;
  const double rc1 = (D4(xij, 2, 1) + D4(xij, 3, 1) + D4(xij, 4, 1)) * 0.25;
  const double rc2 = (D4(yij, 2, 1) + D4(yij, 3, 1) + D4(yij, 4, 1)) * 0.25;
  double rv[4][2];
  for (int n = 1; n <= 4; n++) {
    rv[n - 1][0] = D4(xij, n, 1) - rc1;
    rv[n - 1][1] = D4(yij, n, 1) - rc2;
  }
  double Xi[2], Eta[2];
  Xi[0] = (rv[1][0] + rv[2][0]) - (rv[0][0] + rv[3][0]);
  Xi[1] = (rv[1][1] + rv[2][1]) - (rv[0][1] + rv[3][1]);
  double len = sqrt(Xi[0] * Xi[0] + Xi[1] * Xi[1]);
  Xi[0] /= len; Xi[1] /= len;
  const double lXi = len * 0.5;
  Eta[0] = (rv[2][0] + rv[3][0]) - (rv[0][0] + rv[1][0]);
  Eta[1] = (rv[2][1] + rv[3][1]) - (rv[0][1] + rv[1][1]);
  len = sqrt(Eta[0] * Eta[0] + Eta[1] * Eta[1]);
  Eta[0] /= len; Eta[1] /= len;
  const double lEta = len * 0.5;
  const double dv13[2] = {D4(xij, 3, 1), D4(yij, 3, 1)};
  const double l13 = sqrt(dv13[0] * dv13[0] + dv13[1] * dv13[1]);
  const double u13[2] = {dv13[0] / l13, dv13[1] / l13};
  const double dv24[2] = {D4(xij, 4, 2), D4(yij, 4, 2)};
  const double l24 = sqrt(dv24[0] * dv24[0] + dv24[1] * dv24[1]);
  const double u24[2] = {dv24[0] / l24, dv24[1] / l24};
  double fX[4], fE[4];
  for (int i = 0; i < 4; i++) {
    fX[i] = cross2d(Xi, rv[i]) / lXi;
    fE[i] = cross2d(Eta, rv[i]) / lEta;
  }
  double mX = (fX[0] + fX[1] + fX[2] + fX[3]) * 0.25, mE = (fE[0] + fE[1] + fE[2] + fE[3]) * 0.25;
  for (int i = 0; i < 4; i++) { fX[i] -= mX; fE[i] -= mE; }
  fX[0] = -fX[0]; fX[1] = -fX[1];
  fE[1] = -fE[1]; fE[2] = -fE[2];
  mX = (fX[0] + fX[1] + fX[2] + fX[3]) * 0.25;
  mE = (fE[0] + fE[1] + fE[2] + fE[3]) * 0.25;
  const double f13 = cross2d(u13, dv24) * 0.5 / l13;
  const double f24 = -cross2d(u24, dv13) * 0.5 / l24;
  const double d13x = u13[0] * Xi[0] + u13[1] * Xi[1], d13e = u13[0] * Eta[0] + u13[1] * Eta[1];
  const double d24x = u24[0] * Xi[0] + u24[1] * Xi[1], d24e = u24[0] * Eta[0] + u24[1] * Eta[1];
  const double aE = alpha * lEta / lXi, aX = alpha * lXi / lEta;
  const double t[84] = {
      rho1 * fX[0], rho2 * fX[0], rho3 * fX[0], rho4 * fX[0], aE, -beta1 * fX[0] / (mX * lXi), 0.0,
      -rho1 * fE[0], -rho4 * fE[0], -rho3 * fE[0], -rho2 * fE[0], -aX, 0.0, -beta1 * fE[0] / (mE * lEta),
      rho5 * f24, rho6 * f24, rho7 * f24, rho8 * f24, 0.0, beta2 * d24x / l24, -beta2 * d24e / l24,
      -rho2 * fX[1], -rho1 * fX[1], -rho4 * fX[1], -rho3 * fX[1], -aE, -beta1 * fX[1] / (mX * lXi), 0.0,
      rho4 * fE[1], rho1 * fE[1], rho2 * fE[1], rho3 * fE[1], aX, 0.0, beta1 * fE[1] / (mE * lEta),
      rho8 * f13, rho5 * f13, rho6 * f13, rho7 * f13, 0.0, -beta2 * d13x / l13, beta2 * d13e / l13,
      rho3 * fX[2], rho4 * fX[2], rho1 * fX[2], rho2 * fX[2], aE, beta1 * fX[2] / (mX * lXi), 0.0,
      -rho3 * fE[2], -rho2 * fE[2], -rho1 * fE[2], -rho4 * fE[2], -aX, 0.0, beta1 * fE[2] / (mE * lEta),
      rho7 * f24, rho8 * f24, rho5 * f24, rho6 * f24, 0.0, -beta2 * d24x / l24, beta2 * d24e / l24,
      -rho4 * fX[3], -rho3 * fX[3], -rho2 * fX[3], -rho1 * fX[3], -aE, beta1 * fX[3] / (mX * lXi), 0.0,
      rho2 * fE[3], rho3 * fE[3], rho4 * fE[3], rho1 * fE[3], aX, 0.0, -beta1 * fE[3] / (mE * lEta),
      rho6 * f13, rho7 * f13, rho8 * f13, rho5 * f13, 0.0, beta2 * d13x / l13, -beta2 * d13e / l13};
  double ti13[9], ti24[9], tr13[9], tr24[9];
  ti24[0] = Xi[0] * Xi[0]; ti24[3] = Xi[1] * Xi[1]; ti24[6] = Xi[0] * Xi[1];
  ti24[1] = Eta[0] * Eta[0]; ti24[4] = Eta[1] * Eta[1]; ti24[7] = Eta[0] * Eta[1];
  for (int k = 0; k < 9; k++) ti13[k] = ti24[k];
  ti13[2] = u24[0] * u24[0]; ti13[5] = u24[1] * u24[1]; ti13[8] = u24[0] * u24[1];
  ti24[2] = u13[0] * u13[0]; ti24[5] = u13[1] * u13[1]; ti24[8] = u13[0] * u13[1];
  if (inv33(ti13, tr13) || inv33(ti24, tr24)) return -1;
  for (int n = 0; n < 4; n++) {
    const double* tr = (n % 2 == 0) ? tr13 : tr24;  // nodes 1,3 and 2,4
    for (int c = 0; c < 7; c++) {
      const double b0 = t[n * 21 + c], b1 = t[n * 21 + 7 + c], b2 = t[n * 21 + 14 + c];
      for (int r = 0; r < 3; r++) bm[n * 21 + r * 7 + c] = tr[r] * b0 + tr[r + 3] * b1 + tr[r + 6] * b2;
    }
  }
  return 0;
}

// hierarchical membrane map hm[k*12 + c], 7 x 12 (andes4ShellModule.f90:753-863)
__device__ int hmqv(const double* xij, const double* yij, const double* zij, double* hm) {
  const double rc1 = (D4(xij, 2, 1) + D4(xij, 3, 1) + D4(xij, 4, 1)) * 0.25;
  const double rc2 = (D4(yij, 2, 1) + D4(yij, 3, 1) + D4(yij, 4, 1)) * 0.25;
  const double rc3 = (D4(zij, 2, 1) + D4(zij, 3, 1) + D4(zij, 4, 1)) * 0.25;
  double rv[4][3];
  for (int n = 1; n <= 4; n++) {
    rv[n - 1][0] = D4(xij, n, 1) - rc1;
    rv[n - 1][1] = D4(yij, n, 1) - rc2;
    rv[n - 1][2] = D4(zij, n, 1) - rc3;
  }
  double Xi1 = (rv[1][0] + rv[2][0]) - (rv[0][0] + rv[3][0]);
  double Xi2 = (rv[1][1] + rv[2][1]) - (rv[0][1] + rv[3][1]);
  double fac = sqrt(Xi1 * Xi1 + Xi2 * Xi2);
  Xi1 /= fac; Xi2 /= fac;
  double Eta1 = (rv[2][0] + rv[3][0]) - (rv[0][0] + rv[1][0]);
  double Eta2 = (rv[2][1] + rv[3][1]) - (rv[0][1] + rv[1][1]);
  fac = sqrt(Eta1 * Eta1 + Eta2 * Eta2);
  Eta1 /= fac; Eta2 /= fac;
  double jac[2][3];
  fac = 0.0625 / jacobi_quad(0.0, 0.0, rv, jac);
  for (int k = 0; k < 84; k++) hm[k] = 0.0;
  for (int i = 0; i < 4; i++) {
    for (int c = 8; c < 12; c++) hm[i * 12 + c] = -0.25;
    hm[i * 12 + i + 8] = 0.75;
  }
  hm[48 + 0] = -D4(xij, 2, 4) * fac; hm[48 + 1] = -D4(xij, 3, 1) * fac;
  hm[48 + 2] = -D4(xij, 4, 2) * fac; hm[48 + 3] = -D4(xij, 1, 3) * fac;
  hm[48 + 4] = -D4(yij, 2, 4) * fac; hm[48 + 5] = -D4(yij, 3, 1) * fac;
  hm[48 + 6] = -D4(yij, 4, 2) * fac; hm[48 + 7] = -D4(yij, 1, 3) * fac;
  for (int c = 8; c < 12; c++) hm[48 + c] = 0.25;
  const double a11 = jac[0][0], a12 = jac[1][0], a21 = jac[0][1], a22 = jac[1][1];
  const double det = a11 * a22 - a21 * a12;
  if (fabs(det) < 1.0e-16) return -1;
  const double t11 = a22 / det, t22 = a11 / det, t12 = -a12 / det, t21 = -a21 / det;
  double rn[4][2];
  for (int n = 0; n < 4; n++) {
    rn[n][0] = t11 * rv[n][0] + t12 * rv[n][1];
    rn[n][1] = t21 * rv[n][0] + t22 * rv[n][1];
  }
  fac = (rn[0][0] * rn[0][1] + rn[1][0] * rn[1][1] + rn[2][0] * rn[2][1] + rn[3][0] * rn[3][1]) * 0.25;
  const int nodeof[4] = {2, 3, 0, 1};
  for (int c = 0; c < 4; c++) {
    const double q = rn[nodeof[c]][0] * rn[nodeof[c]][1] - fac;
    hm[60 + c] = q * Xi1; hm[60 + c + 4] = q * Xi2;
    hm[72 + c] = q * Eta1; hm[72 + c + 4] = q * Eta2;
  }
  return 0;
}

// 2x2-Gauss mean of nodal values bn[n*3*ncol + r*ncol + c]
// (andes4ShellModule.f90:342-389), with the four point weights wJ[g] and
// shape functions nv[g][n] precomputed
__device__ void bmean(int ncol, const double* bn, const double (*nv)[4], const double* wJ, double* mean) {
  double area = 0.0;
  for (int k = 0; k < 3 * ncol; k++) mean[k] = 0.0;
  for (int g = 0; g < 4; g++) {
    double n[4];
    for (int k = 0; k < 4; k++) n[k] = nv[g][k] * wJ[g];
    for (int k = 0; k < 3 * ncol; k++)
      mean[k] += n[0] * bn[k] + n[1] * bn[3 * ncol + k] + n[2] * bn[6 * ncol + k] + n[3] * bn[9 * ncol + k];
    area += wJ[g];
  }
  for (int k = 0; k < 3 * ncol; k++) mean[k] /= area;
}

// single-precision literals of qmrf.f / qbek.f (REAL*4 promoted to double)
#define F23 ((double)(2.0f / 3.0f))
#define F13 ((double)(1.0f / 3.0f))
#define F120 ((double)(1.0f / 20.0f))

// triangle (1, I+1, I+2) of QMRF35/QBEK35 (qmrf.f:1527-1640, qbek.f:1729-1845)
__device__ int tri_inertia(const double* X, const double* Y, const double* Z, int I, double thk, double rho,
                           double* m, double* ix, double* iy, double* iz) {
  const double X2 = X[I] - X[0], Y2 = Y[I] - Y[0], Z2 = Z[I] - Z[0];
  const double X3 = X[I + 1] - X[0], Y3 = Y[I + 1] - Y[0], Z3 = Z[I + 1] - Z[0];
  const double SL21 = sqrt(X2 * X2 + Y2 * Y2 + Z2 * Z2), SL31 = sqrt(X3 * X3 + Y3 * Y3 + Z3 * Z3);
  const double COSG = (X3 * X2 + Y3 * Y2 + Z3 * Z2) / (SL31 * SL21);
  const double AUX1 = 1. - COSG * COSG, SING = AUX1 > 0 ? sqrt(AUX1) : 0.0;
  const double XL2 = SL21, XL3 = SL31 * COSG, YL3 = SL31 * SING;
  const double AREA = (XL2 * YL3) / 2.;
  if (!(AREA > 0)) return 0;
  const double XTP = ((XL3 + ((XL2 - XL3) / 2.))) * F23;
  const double H = YL3, B = XL2, B1 = XL3, B2 = XL2 - XL3;
  const double S1 = sqrt((F23 * B1) * (F23 * B1) + (F13 * H) * (F13 * H));
  const double S2 = sqrt((F23 * B2) * (F23 * B2) + (F13 * H) * (F13 * H));
  const double IXX = (rho * thk * B * H) * (((H * H) / 36.) + ((thk * thk) / 24.));
  double IYY1 = (rho * thk * B1 * H) * (((B1 * B1) / 36.) + ((thk * thk) / 24.));
  const double d1 = XTP - F23 * B1;
  IYY1 = IYY1 + ((d1 * d1) * (((B1 * H) / 2.) * rho * thk));
  double IYY2 = (rho * thk * B2 * H) * (((B2 * B2) / 36.) + ((thk * thk) / 24.));
  const double d2 = B1 - XTP + F13 * B2;
  IYY2 = IYY2 + ((d2 * d2) * (((B2 * H) / 2.) * rho * thk));
  const double IYY = IYY1 + IYY2;
  double IZZ1 = (rho * thk * H * B1) * (((B1 * B1) / 4.) + ((H * H) / 12.));
  IZZ1 = IZZ1 - ((S1 * S1) * rho * thk * (B1 * H) / 2.);
  IZZ1 = IZZ1 + ((d1 * d1) * ((B1 * H) / 2.) * rho * thk);
  double IZZ2 = (rho * thk * H * B2) * (((B2 * B2) / 4.) + ((H * H) / 12.));
  IZZ2 = IZZ2 - ((S2 * S2) * rho * thk * (B2 * H) / 2.);
  const double d3 = B1 + F13 * B2 - XTP;
  IZZ2 = IZZ2 + ((d3 * d3) * ((B2 * H) / 2.) * rho * thk);
  const double IZZ = IZZ1 + IZZ2;
  double T[3][3];
  T[0][0] = X2; T[1][0] = Y2; T[2][0] = Z2;
  double SL = sqrt(T[0][0] * T[0][0] + T[1][0] * T[1][0] + T[2][0] * T[2][0]);
  T[0][0] /= SL; T[1][0] /= SL; T[2][0] /= SL;
  T[0][1] = X3; T[1][1] = Y3; T[2][1] = Z3;
  T[0][2] = T[1][0] * T[2][1] - T[1][1] * T[2][0];
  T[1][2] = T[0][1] * T[2][0] - T[0][0] * T[2][1];
  T[2][2] = T[0][0] * T[1][1] - T[0][1] * T[1][0];
  SL = sqrt(T[0][2] * T[0][2] + T[1][2] * T[1][2] + T[2][2] * T[2][2]);
  T[0][2] /= SL; T[1][2] /= SL; T[2][2] /= SL;
  T[0][1] = T[1][2] * T[2][0] - T[1][0] * T[2][2];
  T[1][1] = T[0][0] * T[2][2] - T[0][2] * T[2][0];
  T[2][1] = T[0][2] * T[1][0] - T[0][0] * T[1][2];
  *ix = fabs(T[0][0] * IXX + T[0][1] * IYY + T[0][2] * IZZ);
  *iy = fabs(T[1][0] * IXX + T[1][1] * IYY + T[1][2] * IZZ);
  *iz = fabs(T[2][0] * IXX + T[2][1] * IYY + T[2][2] * IZZ);
  *m = (rho * AREA * thk) / 3.;
  return 1;
}

}  // namespace

__global__ __launch_bounds__(64) void k_andes4(int nel, const int* __restrict__ elist, const int* __restrict__ eptr,
                                               const int* __restrict__ enodes, const double* __restrict__ xyz,
                                               const double* __restrict__ eprop, const int64_t* __restrict__ eoff,
                                               double* __restrict__ Kp, double* __restrict__ Mp,
                                               int* __restrict__ ebad) {
  __shared__ double sG[4][3];           // global corner coordinates
  __shared__ double sT[9];              // TelT: columns = local axes (trans3P)
  __shared__ double sD[3][16];          // local coordinate differences d(i,j)
  __shared__ double sNv[4][4], sWJ[4];  // shape functions and weights at the 2x2 Gauss points
  __shared__ double sBm[84], sBb[144], sHm[84], sMm[21], sMb[36];
  __shared__ double sLm[36], sLb[36];   // lumped strain tables, already divided by the area factor
  __shared__ double sBq[4][21];         // membrane nodal strains interpolated at each point minus mean
  __shared__ double sB[4][6][24];       // strain-displacement matrices
  __shared__ double sCB[4][6][24];      // C * wJ * B, later reused for K P
  __shared__ double sK[24][25];         // global-frame stiffness before projection
  __shared__ double sR[24][6], sRs[24][6], sSi[9];
  __shared__ double sP[24][25];
  __shared__ double sEM[24];
  __shared__ int sBad;
  const int lane = threadIdx.x;
  const int e = elist[blockIdx.x];
  const double* pr = eprop + (int64_t)NPROP * e;
  if (lane == 0) sBad = 0;
  if (lane < 12) {
    const int node = enodes[eptr[e] + lane / 3];
    sG[lane / 3][lane % 3] = xyz[3 * node + lane % 3];
  }
  __syncthreads();
  // ---- element frame and local coordinates (ADD_ANDES4, inadd.f90:1470-1492)
  if (lane == 0) {
    double t1[3], t2[3], t3[3];
    for (int k = 0; k < 3; k++) {
      t1[k] = sG[1][k] - sG[0][k];
      t2[k] = 0.5 * (sG[2][k] + sG[3][k]) - sG[0][k];
    }
    t3[0] = t1[1] * t2[2] - t1[2] * t2[1];
    t3[1] = t1[2] * t2[0] - t1[0] * t2[2];
    t3[2] = t1[0] * t2[1] - t1[1] * t2[0];
    double v = t1[0] * t1[0] + t1[1] * t1[1] + t1[2] * t1[2];
    double w = t3[0] * t3[0] + t3[1] * t3[1] + t3[2] * t3[2];
    if (!(v > 1.0e-16) || !(w > 1.0e-16)) {
      sBad = 1;
      v = w = 1.0;
    }
    v = sqrt(v);
    w = sqrt(w);
    for (int k = 0; k < 3; k++) { t1[k] /= v; t3[k] /= w; }
    t2[0] = t3[1] * t1[2] - t3[2] * t1[1];
    t2[1] = t3[2] * t1[0] - t3[0] * t1[2];
    t2[2] = t3[0] * t1[1] - t3[1] * t1[0];
    for (int k = 0; k < 3; k++) { sT[k] = t1[k]; sT[3 + k] = t2[k]; sT[6 + k] = t3[k]; }
  }
  __syncthreads();
  if (lane < 12) {
    // local coordinate c of node n: (column c of TelT) . (x_n - x_1)
    const int n = lane / 3, c = lane % 3;
    const double* ax = sT + 3 * c;
    const double v0 = sG[n][0] - sG[0][0], v1 = sG[n][1] - sG[0][1], v2 = sG[n][2] - sG[0][2];
    sD[c][n] = ax[0] * v0 + ax[1] * v1 + ax[2] * v2;  // temporarily the coordinates
  }
  __syncthreads();
  double loc[3][4];
  for (int c = 0; c < 3; c++)
    for (int n = 0; n < 4; n++) loc[c][n] = sD[c][n];
  __syncthreads();
  if (lane < 48) {
    const int c = lane / 16, i = (lane % 16) % 4, j = (lane % 16) / 4;
    sD[c][i + 4 * j] = loc[c][i] - loc[c][j];
  }
  __syncthreads();
  const double g = sqrt(1.0 / 3.0), gp[2] = {-g, g};
  // ---- element-level ANDES tables: four lanes in parallel
  if (lane < 4) {
    // Gauss point lane: xsi = gp[lane/2], eta = gp[lane%2] (i outer, j inner)
    double xl[4][3];
    for (int n = 0; n < 4; n++)
      for (int c = 0; c < 3; c++) xl[n][c] = loc[c][n];
    const double xs = gp[lane / 2], et = gp[lane % 2];
    sWJ[lane] = jacobi_quad(xs, et, xl, nullptr) * 1.0 * 1.0;
    shape4(xs, et, sNv[lane]);
  } else if (lane == 4) {
    if (bmqh_nod(sD[0], sD[1], sBm)) sBad = 1;
  } else if (lane == 5) {
    if (bben_nod(sD[0], sD[1], sBb)) sBad = 1;
  } else if (lane == 6) {
    if (hmqv(sD[0], sD[1], sD[2], sHm)) sBad = 1;
  } else if (lane == 7) {
    // lumped tables (lump4mem alphaNR = 1.5, lump4ben ltype 2) over the area factor a
    const double* x = sD[0];
    const double* y = sD[1];
    const double a = 0.5 * (D4(x, 2, 1) * D4(y, 3, 1) - D4(x, 3, 1) * D4(y, 2, 1) + D4(x, 4, 3) * D4(y, 1, 3) -
                            D4(x, 1, 3) * D4(y, 4, 3));
    const double alpha = 1.5;
    for (int i = 1; i <= 4; i++) {
      const int j = 1 + i % 4, l = 1 + (i + 2) % 4;
      const int c1 = i - 1, c2 = i + 3, c3 = i + 7;
      sLm[0 * 12 + c1] = D4(y, j, l) * 0.5 / a;
      sLm[1 * 12 + c1] = 0.0 / a;
      sLm[2 * 12 + c1] = -D4(x, j, l) * 0.5 / a;
      sLm[0 * 12 + c2] = 0.0 / a;
      sLm[1 * 12 + c2] = -D4(x, j, l) * 0.5 / a;
      sLm[2 * 12 + c2] = D4(y, j, l) * 0.5 / a;
      sLm[0 * 12 + c3] = (D4(y, l, i) * D4(y, l, i) - D4(y, j, i) * D4(y, j, i)) * alpha / 12.0 / a;
      sLm[1 * 12 + c3] = (D4(x, l, i) * D4(x, l, i) - D4(x, j, i) * D4(x, j, i)) * alpha / 12.0 / a;
      sLm[2 * 12 + c3] = (D4(x, j, i) * D4(y, j, i) - D4(x, l, i) * D4(y, l, i)) * alpha / 6.0 / a;
      double len2 = D4(x, i, j) * D4(x, i, j) + D4(y, i, j) * D4(y, i, j);
      const double ccij = D4(x, i, j) * D4(x, i, j) / len2, ssij = D4(y, i, j) * D4(y, i, j) / len2;
      const double csij = D4(x, i, j) * D4(y, i, j) / len2;
      len2 = D4(x, l, i) * D4(x, l, i) + D4(y, l, i) * D4(y, l, i);
      const double ccli = D4(x, l, i) * D4(x, l, i) / len2, ssli = D4(y, l, i) * D4(y, l, i) / len2;
      const double csli = D4(x, l, i) * D4(y, l, i) / len2;
      sLb[0 * 12 + c1] = (ssij * D4(x, i, j) + ssli * D4(x, l, i)) * 0.5 / a;
      sLb[1 * 12 + c1] = (ccij * D4(x, i, j) + ccli * D4(x, l, i)) * 0.5 / a;
      sLb[2 * 12 + c1] = (-ccij * D4(y, i, j) - ccli * D4(y, l, i)) / a;
      sLb[0 * 12 + c2] = (ssij * D4(y, i, j) + ssli * D4(y, l, i)) * 0.5 / a;
      sLb[1 * 12 + c2] = (ccij * D4(y, i, j) + ccli * D4(y, l, i)) * 0.5 / a;
      sLb[2 * 12 + c2] = (-ssij * D4(x, i, j) - ssli * D4(x, l, i)) / a;
      sLb[0 * 12 + c3] = (-csij + csli) / a;
      sLb[1 * 12 + c3] = (csij - csli) / a;
      sLb[2 * 12 + c3] = ((ssli - ccli) - (ssij - ccij)) / a;
    }
  } else if (lane == 8) {
    // rigid-body basis of the projector (pmatStiff.f90:23-127), coordinates
    // relative to node 1 then to the centroid
    double cr[4][3];
    for (int c = 0; c < 3; c++) {
      double v[4], s = 0.0;
      for (int n = 0; n < 4; n++) { v[n] = sG[n][c] - sG[0][c]; s += v[n]; }
      s /= 4.0;
      for (int n = 0; n < 4; n++) cr[n][c] = v[n] - s;
    }
    for (int i = 0; i < 24; i++)
      for (int k = 0; k < 6; k++) sR[i][k] = 0.0;
    for (int n = 0; n < 4; n++) {
      const int j = 6 * n;
      sR[j + 0][0] = 1.0; sR[j + 1][1] = 1.0; sR[j + 2][2] = 1.0;
      sR[j + 1][3] = -cr[n][2]; sR[j + 2][3] = cr[n][1]; sR[j + 3][3] = 1.0;
      sR[j + 0][4] = cr[n][2]; sR[j + 2][4] = -cr[n][0]; sR[j + 4][4] = 1.0;
      sR[j + 0][5] = -cr[n][1]; sR[j + 1][5] = cr[n][0]; sR[j + 5][5] = 1.0;
    }
    for (int k = 0; k < 6; k++) {
      double s = 0.0;
      for (int i = 0; i < 24; i++) s += sR[i][k] * sR[i][k];
      s = sqrt(s);
      for (int i = 0; i < 24; i++) sR[i][k] /= s;
    }
    double sub[9];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        double s = 0.0;
        for (int k = 0; k < 24; k++) s += sR[k][i + 3] * sR[k][j + 3];
        sub[i + 3 * j] = s;
      }
    if (inv33(sub, sSi)) sBad = 1;
  } else if (lane == 9) {
    // FQS35 lumped mass (fqs.f: QMRF35 membrane + QBEK35 bending)
    const double X[4] = {sG[0][0], sG[1][0], sG[2][0], sG[3][0]};
    const double Y[4] = {sG[0][1], sG[1][1], sG[2][1], sG[3][1]};
    const double Z[4] = {sG[0][2], sG[1][2], sG[2][2], sG[3][2]};
    const double thk = (pr[P_T1] + pr[P_T1 + 1] + pr[P_T1 + 2] + pr[P_T1 + 3]) * 0.25;
    const double rho = pr[P_RHO];
    double m[2], ix[2], iy[2], iz[2];
    const int ok1 = tri_inertia(X, Y, Z, 1, thk, rho, &m[0], &ix[0], &iy[0], &iz[0]);
    const int ok2 = tri_inertia(X, Y, Z, 2, thk, rho, &m[1], &ix[1], &iy[1], &iz[1]);
    for (int k = 0; k < 24; k++) sEM[k] = 0.0;
    if (ok1 && ok2) {
      const double q = F120;
      const double e1 = ((m[0] + m[1]) + m[0] + (m[0] + m[1]) + m[1]) / 4.0;
      const double e6 = ((q * iz[0] + q * iz[1]) + q * iz[0] + (q * iz[0] + q * iz[1]) + q * iz[1]) / 4.0;
      const double e4 = ((q * ix[0] + q * ix[1]) + q * ix[0] + (q * ix[0] + q * ix[1]) + q * ix[1]) / 4.0;
      const double e5 = ((q * iy[0] + q * iy[1]) + q * iy[0] + (q * iy[0] + q * iy[1]) + q * iy[1]) / 4.0;
      for (int n = 0; n < 4; n++) {
        sEM[6 * n + 0] = e1; sEM[6 * n + 1] = e1; sEM[6 * n + 2] = e1;
        sEM[6 * n + 3] = e4; sEM[6 * n + 4] = e5; sEM[6 * n + 5] = e6;
      }
    } else {
      sBad = 1;
    }
  }
  __syncthreads();
  // ---- Gauss means of the nodal strain tables (bmean_quad)
  if (lane == 0) bmean(7, sBm, sNv, sWJ, sMm);
  if (lane == 1) bmean(12, sBb, sNv, sWJ, sMb);
  // projector helper rows rs(i,:) (pmatStiff.f90:108-120)
  if (lane >= 32 && lane < 56) {
    const int i = lane - 32;
    sRs[i][0] = sR[i][0]; sRs[i][1] = sR[i][1]; sRs[i][2] = sR[i][2];
    for (int j = 0; j < 3; j++)
      sRs[i][3 + j] = sR[i][3] * sSi[0 + 3 * j] + sR[i][4] * sSi[1 + 3 * j] + sR[i][5] * sSi[2 + 3 * j];
  }
  __syncthreads();
  const double alpH = sqrt(0.9);
  // membrane point values minus mean: sBq[g][r*7 + k]
  for (int idx = lane; idx < 84; idx += 64) {
    const int gq = idx / 21, rk = idx % 21;
    const double* nv = sNv[gq];
    sBq[gq][rk] = (nv[0] * sBm[rk] + nv[1] * sBm[21 + rk] + nv[2] * sBm[42 + rk] + nv[3] * sBm[63 + rk]) - sMm[rk];
  }
  // projector P = I - R_s R^T (pmatStiff.f90:121-126)
  for (int idx = lane; idx < 576; idx += 64) {
    const int i = idx / 24, j = idx % 24;
    double s = 0.0;
    for (int k = 0; k < 6; k++) s += sRs[i][k] * sR[j][k];
    sP[i][j] = (i == j ? 1.0 : 0.0) - s;
  }
  __syncthreads();
  // ---- B at the four points (Bmatrix, andes4ShellModule.f90:917-967)
  for (int idx = lane; idx < 4 * 6 * 24; idx += 64) (&sB[0][0][0])[idx] = 0.0;
  __syncthreads();
  const int dofm[12] = {0, 6, 12, 18, 1, 7, 13, 19, 5, 11, 17, 23};
  const int dofb[12] = {3, 9, 15, 21, 4, 10, 16, 22, 2, 8, 14, 20};
  for (int idx = lane; idx < 4 * 3 * 12 * 2; idx += 64) {
    const int gq = idx / 72, rem = idx % 72, ben = rem / 36, r = (rem % 36) / 12, c = rem % 12;
    if (!ben) {
      double s = 0.0;
      for (int k = 0; k < 7; k++) s += sBq[gq][r * 7 + k] * sHm[k * 12 + c];
      sB[gq][r][dofm[c]] = sLm[r * 12 + c] + s * alpH;
    } else {
      // bh4ben: column j of the nodal table lands in column 1+mod(7+j,12)
      const int j = c + 1, k = 1 + (7 + j) % 12;
      const double* nv = sNv[gq];
      const double bh = (nv[0] * sBb[r * 12 + c] + nv[1] * sBb[36 + r * 12 + c] + nv[2] * sBb[72 + r * 12 + c] +
                         nv[3] * sBb[108 + r * 12 + c]) -
                        sMb[r * 12 + c];
      sB[gq][3 + r][dofb[k - 1]] = sLb[r * 12 + k - 1] + bh * alpH;
    }
  }
  __syncthreads();
  // ---- C * wJ * B (block-diagonal C: membrane t*E2, bending t^3/12*E2)
  {
    const double E = pr[P_E], nu = pr[P_NU];
    const double t = (pr[P_T1] + pr[P_T1 + 1] + pr[P_T1 + 2] + pr[P_T1 + 3]) / 4.0;
    const double c11 = E / (1.0 - nu * nu), c12 = nu * c11, c33 = 0.5 * E / (1.0 + nu);
    const double fb = t * t * t / 12.0;
    for (int idx = lane; idx < 4 * 6 * 24; idx += 64) {
      const int gq = idx / 144, r = (idx % 144) / 24, c = idx % 24;
      const int blk = r / 3, rr = r % 3;
      const double f = blk == 0 ? t : fb, w = sWJ[gq];
      const double* b = &sB[gq][3 * blk][c];
      double v;
      if (rr == 0) v = (c11 * f) * w * b[0] + (c12 * f) * w * b[24];
      else if (rr == 1) v = (c12 * f) * w * b[0] + (c11 * f) * w * b[24];
      else v = (c33 * f) * w * b[48];
      sCB[gq][r][c] = v;
    }
  }
  __syncthreads();
  // ---- local stiffness 3x3 block (a, b) of this lane, rotated to global axes
  {
    const int a = lane / 8, b = lane % 8;
    double k[3][3];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        double s = 0.0;
        for (int gq = 0; gq < 4; gq++) {
          double t = 0.0;
          for (int r = 0; r < 6; r++) t += sB[gq][r][3 * a + i] * sCB[gq][r][3 * b + j];
          s += t;
        }
        k[i][j] = s;
      }
    // TelT * k * Tel; TelT(i,m) = sT[i + 3m]
    double tmp[3][3];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) tmp[i][j] = k[i][0] * sT[j] + k[i][1] * sT[3 + j] + k[i][2] * sT[6 + j];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++)
        sK[3 * a + i][3 * b + j] = sT[i] * tmp[0][j] + sT[i + 3] * tmp[1][j] + sT[i + 6] * tmp[2][j];
  }
  __syncthreads();
  // ---- K P into sCB (reused as a 24x24 buffer)
  double* KP = &sCB[0][0][0];
  for (int idx = lane; idx < 576; idx += 64) {
    const int i = idx / 24, j = idx % 24;
    double s = 0.0;
    for (int k = 0; k < 24; k++) s += sK[i][k] * sP[k][j];
    KP[i * 24 + j] = s;
  }
  __syncthreads();
  // ---- P^T (K P): the ten lower node-pair 6x6 blocks, packed row-major
  const int64_t off = eoff[e];
  for (int idx = lane; idx < 360; idx += 64) {
    const int blk = idx / 36, q = idx % 36, i = q / 6, j = q % 6;
    int na = 0;
    while ((na + 1) * (na + 2) / 2 <= blk) na++;
    const int nb = blk - na * (na + 1) / 2;
    const int r = 6 * na + i, c = 6 * nb + j;
    double s = 0.0;
    for (int k = 0; k < 24; k++) s += sP[k][r] * KP[k * 24 + c];
    Kp[off + idx] = s;
    Mp[off + idx] = (na == nb && i == j) ? sEM[r] : 0.0;
  }
  if (lane == 0 && sBad) ebad[e] = -3;
}

void launch_andes4(hipStream_t st, int nel, const int* elist, const int* eptr, const int* enodes, const double* xyz,
                   const double* eprop, const int64_t* eoff, double* Kp, double* Mp, int* ebad) {
  if (nel > 0) k_andes4<<<nel, 64, 0, st>>>(nel, elist, eptr, enodes, xyz, eprop, eoff, Kp, Mp, ebad);
}

}  // namespace fgpu
