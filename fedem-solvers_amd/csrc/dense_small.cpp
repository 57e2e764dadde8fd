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
  V.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; i++) V[(size_t)i * n + i] = 1.0;
  std::vector<double> d(n), e(n, 0.0), v(n), p(n), q(n);
  // Householder reduction: for column k, annihilate A[k+2:, k]
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
    // A <- H A H with H = I - 2 v v^T (acting on rows/cols k..n-1)
    for (int i = k; i < n; i++) {
      double s = 0.0;
      for (int j = k + 1; j < n; j++) s += A[(size_t)i * n + j] * v[j];
      p[i] = 2.0 * s;
    }
    double K = 0.0;
    for (int i = k + 1; i < n; i++) K += v[i] * p[i];
    for (int i = k; i < n; i++) q[i] = p[i] - K * v[i];
    for (int i = k; i < n; i++)
      for (int j = k; j < n; j++) A[(size_t)i * n + j] -= v[i] * q[j] + q[i] * v[j];
    // accumulate V <- V H
    for (int i = 0; i < n; i++) {
      double s = 0.0;
      for (int j = k + 1; j < n; j++) s += V[(size_t)i * n + j] * v[j];
      s *= 2.0;
      for (int j = k + 1; j < n; j++) V[(size_t)i * n + j] -= s * v[j];
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
      double r = std::hypot(g, 1.0);
      g = d[m] - d[l] + e[l] / (g + (g >= 0 ? std::fabs(r) : -std::fabs(r)));
      double s = 1.0, c = 1.0, pp = 0.0;
      int i = m - 1;
      bool early = false;
      for (; i >= l; i--) {
        double f = s * e[i], b = c * e[i];
        r = std::hypot(f, g);
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
        for (int k = 0; k < n; k++) {
          double fz = V[(size_t)k * n + i + 1];
          V[(size_t)k * n + i + 1] = s * V[(size_t)k * n + i] + c * fz;
          V[(size_t)k * n + i] = c * V[(size_t)k * n + i] - s * fz;
        }
      }
      if (early) continue;
      d[l] -= pp;
      e[l] = g;
      e[m] = 0.0;
    }
  }
  // sort ascending
  std::vector<int> idx(n);
  for (int i = 0; i < n; i++) idx[i] = i;
  std::sort(idx.begin(), idx.end(), [&](int a, int b) { return d[a] < d[b]; });
  w.resize(n);
  std::vector<double> Vs((size_t)n * n);
  for (int j = 0; j < n; j++) {
    w[j] = d[idx[j]];
    for (int i = 0; i < n; i++) Vs[(size_t)i * n + j] = V[(size_t)i * n + idx[j]];
  }
  V.swap(Vs);
}

}  // namespace fgpu
