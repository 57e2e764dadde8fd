#include "dense_small.hpp"
#include <algorithm>
#include <cmath>

namespace fgpu {

bool chol_upper(int n, const std::vector<double>& A, std::vector<double>& R) {
  R.assign((size_t)n * n, 0.0);
  for (int j = 0; j < n; j++) {
    double s = A[(size_t)j * n + j];
    for (int k = 0; k < j; k++) s -= R[(size_t)k * n + j] * R[(size_t)k * n + j];
    if (!(s > 0.0)) return false;
    double d = std::sqrt(s);
    R[(size_t)j * n + j] = d;
    for (int i = j + 1; i < n; i++) {
      double t = A[(size_t)j * n + i];
      for (int k = 0; k < j; k++) t -= R[(size_t)k * n + j] * R[(size_t)k * n + i];
      R[(size_t)j * n + i] = t / d;
    }
  }
  return true;
}

void inv_upper(int n, const std::vector<double>& R, std::vector<double>& Ri) {
  Ri.assign((size_t)n * n, 0.0);
  for (int j = 0; j < n; j++) {
    Ri[(size_t)j * n + j] = 1.0 / R[(size_t)j * n + j];
    for (int i = j - 1; i >= 0; i--) {
      double s = 0.0;
      for (int k = i + 1; k <= j; k++) s += R[(size_t)i * n + k] * Ri[(size_t)k * n + j];
      Ri[(size_t)i * n + j] = -s / R[(size_t)i * n + i];
    }
  }
}

void sym_eig(int n, const std::vector<double>& Ain, std::vector<double>& w, std::vector<double>& V) {
  std::vector<double> A(Ain);
  // Zt = V^T is accumulated row-major, so both the Householder products and
  // the QL rotations below update contiguous rows.
  std::vector<double> Zt((size_t)n * n, 0.0);
  for (int i = 0; i < n; i++) Zt[(size_t)i * n + i] = 1.0;
  std::vector<double> d(n), e(n, 0.0), v(n), p(n), q(n), zv(n);
  for (int k = 0; k + 2 < n; k++) {
    double nrm = 0.0;
    for (int i = k + 1; i < n; i++) nrm += A[(size_t)i * n + k] * A[(size_t)i * n + k];
    nrm = std::sqrt(nrm);
    double x0 = A[(size_t)(k + 1) * n + k];
    if (nrm == 0.0) continue;
    double alpha = x0 > 0 ? -nrm : nrm;
    for (int i = 0; i < n; i++) v[i] = 0.0;
    for (int i = k + 1; i < n; i++) v[i] = A[(size_t)i * n + k];
    v[k + 1] -= alpha;
    double vn = 0.0;
    for (int i = k + 1; i < n; i++) vn += v[i] * v[i];
    if (vn == 0.0) continue;
    vn = std::sqrt(vn);
    for (int i = k + 1; i < n; i++) v[i] /= vn;
    // A <- H A H with H = I - 2 v v^T (rows/cols k..n-1; A symmetric)
    for (int i = k; i < n; i++) {
      const double* Ai = &A[(size_t)i * n];
      double s = 0.0;
      for (int j = k + 1; j < n; j++) s += Ai[j] * v[j];
      p[i] = 2.0 * s;
    }
    double K = 0.0;
    for (int i = k + 1; i < n; i++) K += v[i] * p[i];
    for (int i = k; i < n; i++) q[i] = p[i] - K * v[i];
    for (int i = k; i < n; i++) {
      double* Ai = &A[(size_t)i * n];
      const double vi = v[i], qi = q[i];
      for (int j = k; j < n; j++) Ai[j] -= vi * q[j] + qi * v[j];
    }
    // V <- V H  <=>  Zt <- H Zt: rows k+1.. of Zt minus 2 v (v^T Zt)
    for (int j = 0; j < n; j++) zv[j] = 0.0;
    for (int i = k + 1; i < n; i++) {
      const double* Zi = &Zt[(size_t)i * n];
      const double vi = v[i];
      for (int j = 0; j < n; j++) zv[j] += vi * Zi[j];
    }
    for (int i = k + 1; i < n; i++) {
      double* Zi = &Zt[(size_t)i * n];
      const double vi2 = 2.0 * v[i];
      for (int j = 0; j < n; j++) Zi[j] -= vi2 * zv[j];
    }
  }
  for (int i = 0; i < n; i++) d[i] = A[(size_t)i * n + i];
  for (int i = 0; i + 1 < n; i++) e[i] = A[(size_t)(i + 1) * n + i];
  e[n - 1] = 0.0;
  // implicit QL on the tridiagonal (d, e), e[i] couples i and i+1
  for (int l = 0; l < n; l++) {
    int iter = 0;
    while (true) {
      int m = l;
      for (; m < n - 1; m++) {
        double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (std::fabs(e[m]) <= 1e-300 + 2.2e-16 * dd) break;
      }
      if (m == l) break;
      if (++iter > 60) break;
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = std::sqrt(g * g + 1.0);
      g = d[m] - d[l] + e[l] / (g + (g >= 0 ? std::fabs(r) : -std::fabs(r)));
      double s = 1.0, c = 1.0, pp = 0.0;
      int i = m - 1;
      bool early = false;
      for (; i >= l; i--) {
        double f = s * e[i], b = c * e[i];
        r = std::sqrt(f * f + g * g);
        e[i + 1] = r;
        if (r == 0.0) {
          d[i + 1] -= pp;
          e[m] = 0.0;
          early = true;
          break;
        }
        s = f / r;
        c = g / r;
        g = d[i + 1] - pp;
        r = (d[i] - g) * s + 2.0 * c * b;
        pp = s * r;
        d[i + 1] = g + pp;
        g = c * r - b;
        double* z0 = &Zt[(size_t)i * n];
        double* z1 = &Zt[(size_t)(i + 1) * n];
        for (int k = 0; k < n; k++) {
          const double fz = z1[k];
          z1[k] = s * z0[k] + c * fz;
          z0[k] = c * z0[k] - s * fz;
        }
      }
      if (early) continue;
      d[l] -= pp;
      e[l] = g;
      e[m] = 0.0;
    }
  }
  std::vector<int> idx(n);
  for (int i = 0; i < n; i++) idx[i] = i;
  std::sort(idx.begin(), idx.end(), [&](int a, int b) { return d[a] < d[b]; });
  w.resize(n);
  V.assign((size_t)n * n, 0.0);
  for (int j = 0; j < n; j++) {
    w[j] = d[idx[j]];
    const double* Zr = &Zt[(size_t)idx[j] * n];
    for (int i = 0; i < n; i++) V[(size_t)i * n + j] = Zr[i];
  }
}

namespace {

// Householder tridiagonalisation of A (row-major, destroyed); reflector k is
// stored in R row k (entries k+1..n-1, unit 2-norm), beta[k] = 0 when skipped.
void tridiag(int n, std::vector<double>& A, std::vector<double>& R, std::vector<char>& used, std::vector<double>& d,
             std::vector<double>& e) {
  // Householder reduction working on the lower triangle only (row-major
  // storage, A[i*n + j] with j <= i): p = A v as a symmetric product, then
  // the rank-2 update A -= v q^T + q v^T on the lower triangle.
  R.assign((size_t)n * n, 0.0);
  used.assign(n, 0);
  std::vector<double> p(n), q(n);
  for (int k = 0; k + 2 < n; k++) {
    double nrm = 0.0;
    for (int i = k + 1; i < n; i++) nrm += A[(size_t)i * n + k] * A[(size_t)i * n + k];
    nrm = std::sqrt(nrm);
    if (nrm == 0.0) continue;
    double x0 = A[(size_t)(k + 1) * n + k];
    double alpha = x0 > 0 ? -nrm : nrm;
    double* v = &R[(size_t)k * n];
    for (int i = k + 1; i < n; i++) v[i] = A[(size_t)i * n + k];
    v[k + 1] -= alpha;
    double vn = 0.0;
    for (int i = k + 1; i < n; i++) vn += v[i] * v[i];
    if (vn == 0.0) continue;
    vn = std::sqrt(vn);
    for (int i = k + 1; i < n; i++) v[i] /= vn;
    used[k] = 1;
    // column k itself becomes (alpha, 0, ...); the trailing block k+1.. is
    // transformed by I - 2 v v^T
    A[(size_t)(k + 1) * n + k] = alpha;
    for (int i = k + 2; i < n; i++) A[(size_t)i * n + k] = 0.0;
    for (int i = k + 1; i < n; i++) p[i] = 0.0;
    for (int i = k + 1; i < n; i++) {
      const double* Ai = &A[(size_t)i * n];
      const double vi = v[i];
      double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;  // four partial sums (vectorisable)
      int j = k + 1;
      for (; j + 3 < i; j += 4) {
        s0 += Ai[j] * v[j];
        s1 += Ai[j + 1] * v[j + 1];
        s2 += Ai[j + 2] * v[j + 2];
        s3 += Ai[j + 3] * v[j + 3];
        p[j] += Ai[j] * vi;
        p[j + 1] += Ai[j + 1] * vi;
        p[j + 2] += Ai[j + 2] * vi;
        p[j + 3] += Ai[j + 3] * vi;
      }
      for (; j < i; j++) {
        s0 += Ai[j] * v[j];
        p[j] += Ai[j] * vi;
      }
      p[i] += ((s0 + s1) + (s2 + s3)) + Ai[i] * vi;
    }
    double K = 0.0;
    for (int i = k + 1; i < n; i++) {
      p[i] *= 2.0;
      K += v[i] * p[i];
    }
    for (int i = k + 1; i < n; i++) q[i] = p[i] - K * v[i];
    for (int i = k + 1; i < n; i++) {
      double* Ai = &A[(size_t)i * n];
      const double vi = v[i], qi = q[i];
      for (int j = k + 1; j <= i; j++) Ai[j] -= vi * q[j] + qi * v[j];
    }
  }
  d.resize(n);
  e.assign(n, 0.0);
  for (int i = 0; i < n; i++) d[i] = A[(size_t)i * n + i];
  for (int i = 0; i + 1 < n; i++) e[i] = A[(size_t)(i + 1) * n + i];
}

// eigenvalues of the symmetric tridiagonal (d, e) by implicit QL (no vectors)
void tql_values(int n, std::vector<double> d, std::vector<double> e, std::vector<double>& w) {
  for (int l = 0; l < n; l++) {
    int iter = 0;
    while (true) {
      int m = l;
      for (; m < n - 1; m++) {
        double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (std::fabs(e[m]) <= 1e-300 + 2.2e-16 * dd) break;
      }
      if (m == l || ++iter > 60) break;
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = std::sqrt(g * g + 1.0);
      g = d[m] - d[l] + e[l] / (g + (g >= 0 ? std::fabs(r) : -std::fabs(r)));
      double s = 1.0, c = 1.0, pp = 0.0;
      bool early = false;
      for (int i = m - 1; i >= l; i--) {
        double f = s * e[i], b = c * e[i];
        r = std::sqrt(f * f + g * g);
        e[i + 1] = r;
        if (r == 0.0) { d[i + 1] -= pp; e[m] = 0.0; early = true; break; }
        s = f / r;
        c = g / r;
        g = d[i + 1] - pp;
        r = (d[i] - g) * s + 2.0 * c * b;
        pp = s * r;
        d[i + 1] = g + pp;
        g = c * r - b;
      }
      if (early) continue;
      d[l] -= pp;
      e[l] = g;
      e[m] = 0.0;
    }
  }
  w = d;
  std::sort(w.begin(), w.end());
}

// solve (T - sigma I) x = b for tridiagonal T with partial pivoting
void trisolve(int n, const std::vector<double>& d, const std::vector<double>& e, double sigma, std::vector<double>& x) {
  // rows: l (sub), dm (diag), u1 (super), u2 (second super after pivoting)
  std::vector<double> dm(n), u1(n, 0.0), u2(n, 0.0), l(n, 0.0);
  std::vector<char> swp(n, 0);
  for (int i = 0; i < n; i++) dm[i] = d[i] - sigma;
  for (int i = 0; i + 1 < n; i++) u1[i] = e[i];
  std::vector<double> sub(n, 0.0);
  for (int i = 0; i + 1 < n; i++) sub[i] = e[i];
  const double tiny = 1e-300;
  for (int i = 0; i + 1 < n; i++) {
    // eliminate sub[i] (row i+1, col i)
    if (std::fabs(sub[i]) > std::fabs(dm[i])) {
      swp[i] = 1;
      // swap rows i and i+1 (columns i, i+1, i+2)
      double a0 = dm[i], a1 = u1[i], a2 = u2[i];
      double b0 = sub[i], b1 = dm[i + 1], b2 = (i + 1 < n - 1) ? u1[i + 1] : 0.0;
      dm[i] = b0; u1[i] = b1; u2[i] = b2;
      double m = a0 / b0;
      l[i] = m;
      dm[i + 1] = a1 - m * b1;
      if (i + 1 < n - 1) u1[i + 1] = a2 - m * b2;
    } else {
      double piv = std::fabs(dm[i]) > tiny ? dm[i] : tiny;
      dm[i] = piv;
      double m = sub[i] / piv;
      l[i] = m;
      dm[i + 1] -= m * u1[i];
    }
  }
  if (std::fabs(dm[n - 1]) < tiny) dm[n - 1] = tiny;
  for (int i = 0; i + 1 < n; i++) {
    if (swp[i]) std::swap(x[i], x[i + 1]);
    x[i + 1] -= l[i] * x[i];
  }
  for (int i = n - 1; i >= 0; i--) {
    double s = x[i];
    if (i + 1 < n) s -= u1[i] * x[i + 1];
    if (i + 2 < n) s -= u2[i] * x[i + 2];
    x[i] = s / dm[i];
  }
}

}  // namespace

void sym_eig_top(int n, const std::vector<double>& Ain, int k, std::vector<double>& w, std::vector<double>& V) {
  std::vector<double> A(Ain), R, d, e;
  std::vector<char> used;
  tridiag(n, A, R, used, d, e);
  tql_values(n, d, e, w);
  k = std::min(k, n);
  V.assign((size_t)n * k, 0.0);
  double scale = 0.0;
  for (int i = 0; i < n; i++) scale = std::max(scale, std::fabs(w[i]));
  std::vector<std::vector<double>> Z;  // tridiagonal eigenvectors
  for (int c = 0; c < k; c++) {
    const double lam = w[n - 1 - c];
    std::vector<double> x(n);
    for (int i = 0; i < n; i++) x[i] = 1.0 + 0.01 * ((i * 7919) % 101);
    const double sigma = lam + 1e-14 * std::max(scale, std::fabs(lam));
    for (int it = 0; it < 4; it++) {
      trisolve(n, d, e, sigma, x);
      // keep members of a cluster orthogonal to the ones already computed
      for (int q = 0; q < c; q++) {
        if (std::fabs(w[n - 1 - q] - lam) > 1e-6 * std::max(scale, 1e-300)) continue;
        double s = 0.0;
        for (int i = 0; i < n; i++) s += Z[q][i] * x[i];
        for (int i = 0; i < n; i++) x[i] -= s * Z[q][i];
      }
      double nr = 0.0;
      for (int i = 0; i < n; i++) nr += x[i] * x[i];
      nr = std::sqrt(nr);
      for (int i = 0; i < n; i++) x[i] /= nr;
    }
    Z.push_back(x);
    // back-transform: y = H_0 H_1 ... H_{n-3} x
    for (int r = n - 3; r >= 0; r--) {
      if (!used[r]) continue;
      const double* v = &R[(size_t)r * n];
      double s = 0.0;
      for (int i = r + 1; i < n; i++) s += v[i] * x[i];
      s *= 2.0;
      for (int i = r + 1; i < n; i++) x[i] -= s * v[i];
    }
    for (int i = 0; i < n; i++) V[(size_t)i * k + c] = x[i];
  }
}

int sym_gen_eig_lowest(int n, const double* A, const double* B, int nval, double* lambda) {
  // R upper with B = R^T R, from the upper triangle (column-major B(i,j), i <= j)
  auto up = [n](const double* M, int i, int j) { return i <= j ? M[i + (size_t)j * n] : M[j + (size_t)i * n]; };
  std::vector<double> R((size_t)n * n, 0.0);  // row-major R[i*n+j]
  for (int j = 0; j < n; j++) {
    double s = up(B, j, j);
    for (int k = 0; k < j; k++) s -= R[(size_t)k * n + j] * R[(size_t)k * n + j];
    if (!(s > 0.0)) return n + j + 1;
    const double d = std::sqrt(s);
    R[(size_t)j * n + j] = d;
    for (int i = j + 1; i < n; i++) {
      double t = up(B, j, i);
      for (int k = 0; k < j; k++) t -= R[(size_t)k * n + j] * R[(size_t)k * n + i];
      R[(size_t)j * n + i] = t / d;
    }
  }
  std::vector<double> Ri;
  inv_upper(n, R, Ri);
  // C = Ri^T A Ri (row-major), T = A Ri first
  std::vector<double> T((size_t)n * n, 0.0), C((size_t)n * n, 0.0);
  for (int i = 0; i < n; i++)
    for (int k = 0; k < n; k++) {
      const double a = up(A, i, k);
      if (a == 0.0) continue;
      for (int j = k; j < n; j++) T[(size_t)i * n + j] += a * Ri[(size_t)k * n + j];
    }
  for (int k = 0; k < n; k++)
    for (int i = 0; i <= k; i++) {
      const double r = Ri[(size_t)i * n + k];
      if (r == 0.0) continue;
      for (int j = 0; j < n; j++) C[(size_t)k * n + j] += r * T[(size_t)i * n + j];
    }
  for (int i = 0; i < n; i++)  // symmetrise the rounding
    for (int j = 0; j < i; j++) C[(size_t)i * n + j] = C[(size_t)j * n + i] = 0.5 * (C[(size_t)i * n + j] + C[(size_t)j * n + i]);
  std::vector<double> w, V;
  sym_eig(n, C, w, V);
  for (int i = 0; i < nval && i < n; i++) lambda[i] = w[i];
  return 0;
}

}  // namespace fgpu
