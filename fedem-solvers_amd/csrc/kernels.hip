// gfx950 kernels of the part reducer: element matrices, pull assembly,
// multifrontal LDL^T (panel factorisation on VALU, trailing updates on FP64
// MFMA), front-based multi-RHS solves and the tall-skinny reductions.
//
// Front storage: the front of supernode s is an nrow x nrow column-major
// FP64 block at pool + foff[s]; only its lower triangle is meaningful.  After
// factorisation its first ncol columns hold L (unit diagonal implied) with D
// on the diagonal, and the trailing (nrow-ncol)^2 block holds the update
// (contribution) matrix the parent adds in.
#include <algorithm>
#include "kernels.hpp"

namespace fgpu {

using f64x4 = __attribute__((__vector_size__(4 * sizeof(double)))) double;

// ---------------------------------------------------------------- HEX8
// One 64-lane workgroup per element.  Lanes 0..26 evaluate the 27 Gauss
// points (Femlib/hexa.f HEXA31 with NIP=3 and HEXA35/PRIMAS share the same
// 3x3x3 rule, whose abscissa and weights are single-precision literals in
// the Fortran source), lanes 0..35 then accumulate one 3x3 node-pair block
// of K and of the consistent M each.
__global__ __launch_bounds__(64) void k_hex8(int nhex, const int* __restrict__ elist,
                                             const int* __restrict__ eptr,
                                             const int* __restrict__ enodes,
                                             const double* __restrict__ xyz,
                                             const double* __restrict__ eprop,
                                             const int64_t* __restrict__ eoff,
                                             double* __restrict__ Kp, double* __restrict__ Mp,
                                             int* __restrict__ ebad) {
  __shared__ double sX[3][8];
  __shared__ double sD[27][24];  // dN/dx: [n*8 + node]
  __shared__ double sN[27][8];
  __shared__ double sW[27];
  __shared__ int sBad;
  const int lane = threadIdx.x;
  const int e = elist[blockIdx.x];
  if (lane == 0) sBad = 0;
  if (lane < 24) {
    int node = enodes[eptr[e] + (lane & 7)];
    sX[lane >> 3][lane & 7] = xyz[3 * node + (lane >> 3)];
  }
  __syncthreads();
  const double a3 = (double)0.77459666924148f;
  const double wt1 = (double)0.55555555555556f, wt2 = (double)0.88888888888889f;
  const double* X = sX[0];
  const double* Y = sX[1];
  const double* Z = sX[2];
  auto jac = [&](double xi, double eta, double zeta, double* BJ, bool inv) -> double {
    double A1 = 1. + xi, A2 = 1. - xi, B1 = 1. + eta, B2 = 1. - eta, C1 = 1. + zeta, C2 = 1. - zeta;
    double J[3][3];
    const double* V[3] = {X, Y, Z};
    for (int c = 0; c < 3; c++) {
      const double* v = V[c];
      J[0][c] = 0.125 * (B1 * C1 * (v[6] - v[7]) - B2 * C1 * (v[4] - v[5]) + B1 * C2 * (v[2] - v[3]) - B2 * C2 * (v[0] - v[1]));
      J[1][c] = 0.125 * (A1 * C1 * (v[6] - v[5]) + A2 * C1 * (v[7] - v[4]) + A1 * C2 * (v[2] - v[1]) + A2 * C2 * (v[3] - v[0]));
      J[2][c] = 0.125 * (A1 * B1 * (v[6] - v[2]) + A2 * B1 * (v[7] - v[3]) + A2 * B2 * (v[4] - v[0]) + A1 * B2 * (v[5] - v[1]));
    }
    double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                 J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    if (inv)
      for (int q = 0; q < 9; q++) BJ[q] = 0.0;  // hexa.f:814-816 zeroes BJ first
    if (inv && det > 0.0) {
      double r = 1. / det;
      BJ[0] = r * (J[1][1] * J[2][2] - J[2][1] * J[1][2]);
      BJ[1] = -r * (J[0][1] * J[2][2] - J[0][2] * J[2][1]);
      BJ[2] = r * (J[0][1] * J[1][2] - J[0][2] * J[1][1]);
      BJ[3] = -r * (J[1][0] * J[2][2] - J[1][2] * J[2][0]);
      BJ[4] = r * (J[0][0] * J[2][2] - J[0][2] * J[2][0]);
      BJ[5] = -r * (J[0][0] * J[1][2] - J[0][2] * J[1][0]);
      BJ[6] = r * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
      BJ[7] = -r * (J[0][0] * J[2][1] - J[0][1] * J[2][0]);
      BJ[8] = r * (J[0][0] * J[1][1] - J[0][1] * J[1][0]);
    }
    return det;
  };
  if (lane < 27) {
    int ix = lane % 3, iy = (lane / 3) % 3, iz = lane / 9;
    double gp[3] = {-a3, 0.0, a3}, gw[3] = {wt1, wt2, wt1};
    double xi = gp[ix], eta = gp[iy], zeta = gp[iz];
    double BJ[9];
    double det = jac(xi, eta, zeta, BJ, true);
    sW[lane] = gw[ix] * gw[iy] * gw[iz] * det;
    double xip = 1. + xi, xim = 1. - xi, etap = 1. + eta, etam = 1. - eta, zetp = 1. + zeta, zetm = 1. - zeta;
    for (int n = 0; n < 3; n++) {
      // BJ row n = (BJ(n,1), BJ(n,2), BJ(n,3)) of the Fortran inverse Jacobian
      double b1 = BJ[3 * n + 0], b2 = BJ[3 * n + 1], b3 = BJ[3 * n + 2];
      double* D = &sD[lane][n * 8];
      D[0] = .125 * (-b1 * etam * zetm - b2 * xim * zetm - b3 * xim * etam);
      D[1] = .125 * (b1 * etam * zetm - b2 * xip * zetm - b3 * xip * etam);
      D[2] = .125 * (b1 * etap * zetm + b2 * xip * zetm - b3 * xip * etap);
      D[3] = .125 * (-b1 * etap * zetm + b2 * xim * zetm - b3 * xim * etap);
      D[4] = .125 * (-b1 * etam * zetp - b2 * xim * zetp + b3 * xim * etam);
      D[5] = .125 * (b1 * etam * zetp - b2 * xip * zetp + b3 * xip * etam);
      D[6] = .125 * (b1 * etap * zetp + b2 * xip * zetp + b3 * xip * etap);
      D[7] = .125 * (-b1 * etap * zetp + b2 * xim * zetp + b3 * xim * etap);
    }
    sN[lane][0] = .125 * xim * etam * zetm;
    sN[lane][1] = .125 * xip * etam * zetm;
    sN[lane][2] = .125 * xip * etap * zetm;
    sN[lane][3] = .125 * xim * etap * zetm;
    sN[lane][4] = .125 * xim * etam * zetp;
    sN[lane][5] = .125 * xip * etam * zetp;
    sN[lane][6] = .125 * xip * etap * zetp;
    sN[lane][7] = .125 * xim * etap * zetp;
  } else if (lane < 35) {
    // Jacobian sign check at the nodes (hexa.f:75-80)
    const double cx[8] = {-1, 1, 1, -1, -1, 1, 1, -1}, cy[8] = {-1, -1, 1, 1, -1, -1, 1, 1},
                 cz[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
    int i = lane - 27;
    if (jac(cx[i], cy[i], cz[i], nullptr, false) < 0.0) sBad = 1;
  }
  __syncthreads();
  if (lane == 0 && sBad) ebad[e] = -2;
  if (lane >= 36) return;
  const double E = eprop[NPROP * e + P_E], nu = eprop[NPROP * e + P_NU], rho = eprop[NPROP * e + P_RHO];
  const double fac = E / ((1.0 + nu) * (1.0 - nu - nu));
  const double C11 = (1.0 - nu) * fac, C12 = nu * fac, G = (0.5 - nu) * fac;
  int a = (int)((sqrt(8.0 * lane + 1.0) - 1.0) * 0.5);
  while ((a + 1) * (a + 2) / 2 <= lane) a++;
  while (a * (a + 1) / 2 > lane) a--;
  const int b = lane - a * (a + 1) / 2;
  double acc[9], macc = 0.0;
#pragma unroll
  for (int q = 0; q < 9; q++) acc[q] = 0.0;
  for (int g = 0; g < 27; g++) {
    double da[3] = {sD[g][a], sD[g][8 + a], sD[g][16 + a]};
    double db[3] = {sD[g][b], sD[g][8 + b], sD[g][16 + b]};
    double wd = sW[g];
    double dot = da[0] * db[0] + da[1] * db[1] + da[2] * db[2];
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
      for (int c = 0; c < 3; c++) {
        double v = (r == c) ? C11 * da[r] * db[r] + G * (dot - da[r] * db[r])
                            : C12 * da[r] * db[c] + G * da[c] * db[r];
        acc[3 * r + c] += wd * v;
      }
    macc += wd * sN[g][a] * sN[g][b];
  }
  const int64_t off = eoff[e] + 9 * lane;
#pragma unroll
  for (int q = 0; q < 9; q++) {
    Kp[off + q] = acc[q];
    Mp[off + q] = (q % 4 == 0) ? rho * macc : 0.0;
  }
}

// ---------------------------------------------------------------- assembly
// Owner-computes (pull) assembly: every CSR entry sums its element
// contributions in a fixed order, so the result is deterministic and needs no
// atomics or colouring passes.  Lower-triangle stiffness entries are also
// written straight into their front position.
__global__ __launch_bounds__(256) void k_assemble(int64_t nnz, const int* __restrict__ cptr,
                                                  const int64_t* __restrict__ cidx,
                                                  const double* __restrict__ Kp,
                                                  const double* __restrict__ Mp,
                                                  double* __restrict__ Kv, double* __restrict__ Mv,
                                                  const int64_t* __restrict__ fmap,
                                                  double* __restrict__ pool) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnz) return;
  double kv = 0.0, mv = 0.0;
  for (int q = cptr[k]; q < cptr[k + 1]; q++) {
    int64_t c = cidx[q];
    kv += Kp[c];
    mv += Mp[c];
  }
  Kv[k] = kv;
  Mv[k] = mv;
  int64_t f = fmap[k];
  if (f >= 0) pool[f] = kv;
}

__global__ void k_fill(double* __restrict__ p, int64_t n, double v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

// ---------------------------------------------------------------- extend-add
// Task: sn = parent front, a = first parent column of a 64-column block.  The
// workgroup adds, child after child, every child update-matrix column whose
// parent column falls in its block.  Blocks are disjoint, so no two
// workgroups touch the same parent entry; children are added in list order.
__global__ __launch_bounds__(256) void k_extend_add(const FTask* __restrict__ tasks, DevSym S,
                                                    double* __restrict__ pool) {
  const FTask t = tasks[blockIdx.x];
  const int sn = t.sn, j0 = t.a, j1 = t.a + 64;
  const int ldp = S.nrow[sn];
  double* P = pool + S.foff[sn];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int q = S.chptr[sn]; q < S.chptr[sn + 1]; q++) {
    const int c = S.chlist[q];
    const int nc = S.ncol[c], ldc = S.nrow[c], nu = ldc - nc;
    const int* rel = S.rel + S.relptr[c];
    int lo = 0, hi = nu;
    while (lo < hi) { int m = (lo + hi) >> 1; if (rel[m] < j0) lo = m + 1; else hi = m; }
    int a = lo;
    hi = nu;
    while (lo < hi) { int m = (lo + hi) >> 1; if (rel[m] < j1) lo = m + 1; else hi = m; }
    const int b = lo;
    const double* C = pool + S.foff[c] + (int64_t)nc * ldc + nc;
    for (int jj = a + w; jj < b; jj += 4) {
      const int pc = rel[jj];
      double* dst = P + (int64_t)pc * ldp;
      const double* src = C + (int64_t)jj * ldc;
      for (int ii = jj + lane; ii < nu; ii += 64) dst[rel[ii]] += src[ii];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------- panel
// Task: sn, a = row chunk, k0 = first panel column.  Every workgroup factors
// the (kb x kb, kb <= 64) diagonal block in LDS (redundantly, it is cheap),
// chunk 0 writes it back, and each workgroup then solves its 256 rows of the
// panel: L_r = F_r L_kk^{-T} D^{-1}.
__global__ __launch_bounds__(256) void k_panel(const FTask* __restrict__ tasks, DevSym S,
                                               double* __restrict__ pool, int* __restrict__ piv) {
  __shared__ double sA[64][65];
  const FTask t = tasks[blockIdx.x];
  const int sn = t.sn, k0 = t.k0, ld = S.nrow[sn];
  const int kb = min(64, S.ncol[sn] - k0);
  double* F = pool + S.foff[sn];
  const int tid = threadIdx.x;
  for (int idx = tid; idx < kb * kb; idx += 256) {
    int i = idx % kb, j = idx / kb;
    sA[i][j] = (i >= j) ? F[(int64_t)(k0 + i) + (int64_t)(k0 + j) * ld] : 0.0;
  }
  __syncthreads();
  for (int j = 0; j < kb; j++) {
    const double d = sA[j][j];
    for (int i = j + 1 + tid; i < kb; i += 256) sA[i][j] = sA[i][j] / d;
    __syncthreads();
    const int m = kb - j - 1;
    for (int idx = tid; idx < m * m; idx += 256) {
      int i = j + 1 + idx / m, q = j + 1 + idx % m;
      if (q <= i) sA[i][q] -= sA[i][j] * d * sA[q][j];
    }
    __syncthreads();
  }
  if (t.a == 0) {
    for (int idx = tid; idx < kb * kb; idx += 256) {
      int i = idx % kb, j = idx / kb;
      if (i >= j) F[(int64_t)(k0 + i) + (int64_t)(k0 + j) * ld] = sA[i][j];
    }
    if (tid < kb) {
      double d = sA[tid][tid];
      if (!(d > 0.0)) {  // K_ii must be positive definite
        atomicAdd(&piv[0], 1);
        atomicMin(&piv[1], S.col0[sn] + k0 + tid);
      }
    }
  }
  const int r = k0 + kb + t.a * 256 + tid;
  if (r >= ld) return;
  double w[64];
#pragma unroll
  for (int j = 0; j < 64; j++) {
    if (j < kb) {
      double f = F[(int64_t)r + (int64_t)(k0 + j) * ld];
#pragma unroll
      for (int m = 0; m < 64; m++)
        if (m < j) f -= w[m] * sA[j][m];
      w[j] = f;
    }
  }
#pragma unroll
  for (int j = 0; j < 64; j++)
    if (j < kb) F[(int64_t)r + (int64_t)(k0 + j) * ld] = w[j] / sA[j][j];
}

// ---------------------------------------------------------------- update
// Task: sn, a = ti, b = tj (64x64 tiles of the trailing lower triangle),
// k0 = panel.  F[I,J] -= (L_I D) L_J^T with the panel of width kb <= 64,
// on v_mfma_f64_16x16x4_f64.  Four waves, each a 32x32 quarter of the tile.
// The MFMA's A operand takes L_J rows and its B operand the D-scaled L_I rows,
// so the 16x16 result lands with consecutive lanes on consecutive rows i of
// one column j of the column-major front (coalesced stores).
__global__ __launch_bounds__(256) void k_update(const FTask* __restrict__ tasks, DevSym S,
                                                double* __restrict__ pool) {
  __shared__ double sI[64][65];
  __shared__ double sJ[64][65];
  const FTask t = tasks[blockIdx.x];
  const int sn = t.sn, k0 = t.k0, ld = S.nrow[sn];
  const int kb = min(64, S.ncol[sn] - k0);
  const int t0 = k0 + kb;
  const int I0 = t0 + 64 * t.a, J0 = t0 + 64 * t.b;
  double* F = pool + S.foff[sn];
  const int tid = threadIdx.x;
  const int kbr = (kb + 3) & ~3;
  for (int idx = tid; idx < 64 * kbr; idx += 256) {
    int i = idx & 63, m = idx >> 6;
    double vi = 0.0, vj = 0.0;
    if (m < kb) {
      double d = F[(int64_t)(k0 + m) * (ld + 1)];
      if (I0 + i < ld) vi = F[(int64_t)(I0 + i) + (int64_t)(k0 + m) * ld] * d;
      if (J0 + i < ld) vj = F[(int64_t)(J0 + i) + (int64_t)(k0 + m) * ld];
    }
    sI[i][m] = vi;
    sJ[i][m] = vj;
  }
  __syncthreads();
  const int lane = tid & 63, wv = tid >> 6;
  const int wr = (wv >> 1) * 32, wc = (wv & 1) * 32;
  f64x4 acc[2][2];
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++) acc[x][y] = f64x4{0.0, 0.0, 0.0, 0.0};
  const int lr = lane & 15, lk = lane >> 4;
  for (int m0 = 0; m0 < kbr; m0 += 4) {
    double aj[2], bi[2];
#pragma unroll
    for (int x = 0; x < 2; x++) {
      aj[x] = sJ[wc + 16 * x + lr][m0 + lk];
      bi[x] = sI[wr + 16 * x + lr][m0 + lk];
    }
#pragma unroll
    for (int x = 0; x < 2; x++)      // i blocks
#pragma unroll
      for (int y = 0; y < 2; y++)    // j blocks
        acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(aj[y], bi[x], acc[x][y], 0, 0, 0);
  }
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        // v_mfma_f64_16x16x4f64 result D[p][q] sits in lane q + 16*(p % 4),
        // register p / 4; here p runs over j and q over i
        const int i = I0 + wr + 16 * x + lr;
        const int j = J0 + wc + 16 * y + lk + 4 * r;
        if (i < ld && j < ld && i >= j) F[(int64_t)i + (int64_t)j * ld] -= acc[x][y][r];
      }
}

// ---------------------------------------------------------------- solves
// Work buffers: front s owns rows woff[s] .. woff[s]+nrow[s]-1 of a
// row-major (rows x nrhs) array W.  Forward mode uses L, backward mode L^T.
//
// k_solve_gemm task: W[r0+i][j0+jj] -= sum_k op(L)[r0+i][k0+k] W[k0+k][j0+jj]
//   forward : op(L)[i][k] = F[i + k*ld]   (rows below the source block)
//   backward: op(L)[i][k] = F[k + i*ld]   (rows above the source block)
__global__ __launch_bounds__(256) void k_solve_gemm(const SolveTask* __restrict__ tasks, DevSym S,
                                                    const double* __restrict__ pool,
                                                    double* __restrict__ W, int nrhs, int backward) {
  __shared__ double sL[64][65];
  __shared__ double sW[64][64];
  const SolveTask t = tasks[blockIdx.x];
  const int sn = t.sn, ld = S.nrow[sn];
  const double* F = pool + S.foff[sn];
  double* Wf = W + S.woff[sn] * (int64_t)nrhs;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int nj = min(64, nrhs - t.j0);
  double acc[16];
#pragma unroll
  for (int q = 0; q < 16; q++) acc[q] = 0.0;
  for (int kc = 0; kc < t.kn; kc += 64) {
    const int kn = min(64, t.kn - kc);
    __syncthreads();
    for (int idx = tid; idx < 64 * 64; idx += 256) {
      int i, k;
      double v = 0.0;
      if (!backward) {
        i = idx & 63; k = idx >> 6;
        if (i < t.rn && k < kn) v = F[(int64_t)(t.r0 + i) + (int64_t)(t.k0 + kc + k) * ld];
      } else {
        k = idx & 63; i = idx >> 6;
        if (i < t.rn && k < kn) v = F[(int64_t)(t.k0 + kc + k) + (int64_t)(t.r0 + i) * ld];
      }
      sL[i][k] = v;
      int kk = idx >> 6, jj = idx & 63;
      sW[kk][jj] = (kk < kn && jj < nj) ? Wf[(int64_t)(t.k0 + kc + kk) * nrhs + t.j0 + jj] : 0.0;
    }
    __syncthreads();
    for (int k = 0; k < kn; k++) {
      const double wk = sW[k][lane];
#pragma unroll
      for (int q = 0; q < 16; q++) acc[q] += sL[wv + 4 * q][k] * wk;
    }
  }
  if (lane < nj) {
#pragma unroll
    for (int q = 0; q < 16; q++) {
      int i = wv + 4 * q;
      if (i < t.rn) Wf[(int64_t)(t.r0 + i) * nrhs + t.j0 + lane] -= acc[q];
    }
  }
}

// k_solve_diag task: unit-triangular solve of the rn x rn diagonal block at
// rows r0 for right-hand sides j0..j0+63 (thread per right-hand side).
__global__ __launch_bounds__(64) void k_solve_diag(const SolveTask* __restrict__ tasks, DevSym S,
                                                   const double* __restrict__ pool,
                                                   double* __restrict__ W, int nrhs, int backward) {
  __shared__ double sL[64][65];
  const SolveTask t = tasks[blockIdx.x];
  const int sn = t.sn, ld = S.nrow[sn], n = t.rn;
  const double* F = pool + S.foff[sn];
  double* Wf = W + S.woff[sn] * (int64_t)nrhs;
  const int lane = threadIdx.x;
  for (int idx = lane; idx < n * n; idx += 64) {
    int i = idx % n, j = idx / n;
    sL[i][j] = F[(int64_t)(t.r0 + i) + (int64_t)(t.r0 + j) * ld];
  }
  __syncthreads();
  const int j = t.j0 + lane;
  if (j >= nrhs) return;
  double w[64];
#pragma unroll
  for (int i = 0; i < 64; i++) w[i] = (i < n) ? Wf[(int64_t)(t.r0 + i) * nrhs + j] : 0.0;
  if (!backward) {
#pragma unroll
    for (int i = 1; i < 64; i++) {
      double s = w[i];
#pragma unroll
      for (int m = 0; m < 64; m++)
        if (m < i) s -= sL[i][m] * w[m];
      w[i] = s;
    }
  } else {
#pragma unroll
    for (int i = 62; i >= 0; i--) {
      double s = w[i];
#pragma unroll
      for (int m = 1; m < 64; m++)
        if (m > i && m < n) s -= sL[m][i] * w[m];
      w[i] = s;
    }
  }
#pragma unroll
  for (int i = 0; i < 64; i++)
    if (i < n) Wf[(int64_t)(t.r0 + i) * nrhs + j] = w[i];
}

// Forward-solve front initialisation (one workgroup per front): pivot rows
// from the global right-hand side R (row-major, perm rows), update rows
// zero, then the children's update parts added in list order.
__global__ __launch_bounds__(256) void k_fw_init(const int* __restrict__ sns, DevSym S,
                                                 const double* __restrict__ R,
                                                 double* __restrict__ W, int nrhs, int ndof1) {
  const int sn = sns[blockIdx.x];
  const int nc = S.ncol[sn], nr = S.nrow[sn], c0 = S.col0[sn];
  double* Wf = W + S.woff[sn] * (int64_t)nrhs;
  const int tid = threadIdx.x;
  for (int64_t idx = tid; idx < (int64_t)nr * nrhs; idx += 256) {
    int i = (int)(idx / nrhs), j = (int)(idx % nrhs);
    double v = 0.0;
    if (i < nc) v = R[(int64_t)(c0 + i) * nrhs + j];
    else if (nc == 0) v = R[(int64_t)(ndof1 + i) * nrhs + j];  // Schur front: boundary rows
    Wf[idx] = v;
  }
  for (int q = S.chptr[sn]; q < S.chptr[sn + 1]; q++) {
    __syncthreads();
    const int c = S.chlist[q];
    const int ncc = S.ncol[c], nu = S.nrow[c] - ncc;
    const int* rel = S.rel + S.relptr[c];
    const double* Wc = W + (S.woff[c] + ncc) * (int64_t)nrhs;
    for (int64_t idx = tid; idx < (int64_t)nu * nrhs; idx += 256) {
      int i = (int)(idx / nrhs), j = (int)(idx % nrhs);
      Wf[(int64_t)rel[i] * nrhs + j] += Wc[idx];
    }
  }
}

// Copy the solved pivot rows of each front out to Y (optionally scaled by
// 1/D, the diagonal of the factor), and for the Schur front (ncol == 0) its
// rows to the boundary rows of Y.
__global__ __launch_bounds__(256) void k_fw_out(const int* __restrict__ sns, DevSym S,
                                                const double* __restrict__ pool,
                                                const double* __restrict__ W,
                                                double* __restrict__ Y, int nrhs, int scale, int ndof1) {
  const int sn = sns[blockIdx.x];
  const int nc = S.ncol[sn], nr = S.nrow[sn], c0 = S.col0[sn];
  const double* F = pool + S.foff[sn];
  const double* Wf = W + S.woff[sn] * (int64_t)nrhs;
  const int n = nc > 0 ? nc : nr;
  const int base = nc > 0 ? c0 : ndof1;
  for (int64_t idx = threadIdx.x; idx < (int64_t)n * nrhs; idx += 256) {
    int i = (int)(idx / nrhs);
    double v = Wf[idx];
    if (scale && nc > 0) v /= F[(int64_t)i * (nr + 1)];
    Y[(int64_t)(base + i) * nrhs + idx % nrhs] = v;
  }
}

// Backward-solve front initialisation: pivot rows from Y, update rows
// gathered from the already final rows of X (ancestors and boundary).
__global__ __launch_bounds__(256) void k_bw_init(const int* __restrict__ sns, DevSym S,
                                                 const double* __restrict__ Y,
                                                 const double* __restrict__ X,
                                                 double* __restrict__ W, int nrhs) {
  const int sn = sns[blockIdx.x];
  const int nc = S.ncol[sn], nr = S.nrow[sn], c0 = S.col0[sn];
  const int* rows = S.rows + S.rptr[sn];
  double* Wf = W + S.woff[sn] * (int64_t)nrhs;
  for (int64_t idx = threadIdx.x; idx < (int64_t)nr * nrhs; idx += 256) {
    int i = (int)(idx / nrhs), j = (int)(idx % nrhs);
    Wf[idx] = (i < nc) ? Y[(int64_t)(c0 + i) * nrhs + j] : X[(int64_t)rows[i] * nrhs + j];
  }
}

__global__ __launch_bounds__(256) void k_bw_out(const int* __restrict__ sns, DevSym S,
                                                const double* __restrict__ W,
                                                double* __restrict__ X, int nrhs) {
  const int sn = sns[blockIdx.x];
  const int nc = S.ncol[sn], c0 = S.col0[sn];
  const double* Wf = W + S.woff[sn] * (int64_t)nrhs;
  for (int64_t idx = threadIdx.x; idx < (int64_t)nc * nrhs; idx += 256)
    X[(int64_t)c0 * nrhs + idx] = Wf[idx];
}

// ---------------------------------------------------------------- products
// Z = A X for the full symmetric CSR matrix A (perm space), X, Z row-major.
__global__ __launch_bounds__(256) void k_spmm(int n, int nrhs, const int* __restrict__ rowptr,
                                              const int* __restrict__ colind,
                                              const double* __restrict__ val,
                                              const double* __restrict__ X, int ldx,
                                              double* __restrict__ Z, int ldz) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int r = (int)(g / nrhs), j = (int)(g % nrhs);
  if (r >= n) return;
  double s = 0.0;
  for (int k = rowptr[r]; k < rowptr[r + 1]; k++) s += val[k] * X[(int64_t)colind[k] * ldx + j];
  Z[(int64_t)r * ldz + j] = s;
}

// Partial C = X^T Y: blockIdx.x = chunk of XTY_ROWS rows, blockIdx.y/z = 64x64
// tile of C.  part[chunk][a][b] (na x nb row-major per chunk).
#define XTY_ROWS 1024
__global__ __launch_bounds__(256) void k_xty_part(int n, int na, int nb, const double* __restrict__ X,
                                                  int ldx, const double* __restrict__ Y, int ldy,
                                                  double* __restrict__ part) {
  __shared__ double sX[32][64];
  __shared__ double sY[32][64];
  const int r0 = blockIdx.x * XTY_ROWS, a0 = blockIdx.y * 64, b0 = blockIdx.z * 64;
  const int tid = threadIdx.x, lane = tid & 63, bg = tid >> 6;
  double acc[16];
#pragma unroll
  for (int q = 0; q < 16; q++) acc[q] = 0.0;
  const int rend = min(n, r0 + XTY_ROWS);
  for (int rc = r0; rc < rend; rc += 32) {
    __syncthreads();
    for (int idx = tid; idx < 32 * 64; idx += 256) {
      int rr = idx >> 6, c = idx & 63, r = rc + rr;
      bool ok = r < rend;
      sX[rr][c] = (ok && a0 + c < na) ? X[(int64_t)r * ldx + a0 + c] : 0.0;
      sY[rr][c] = (ok && b0 + c < nb) ? Y[(int64_t)r * ldy + b0 + c] : 0.0;
    }
    __syncthreads();
#pragma unroll 4
    for (int rr = 0; rr < 32; rr++) {
      const double xv = sX[rr][lane];
#pragma unroll
      for (int q = 0; q < 16; q++) acc[q] += xv * sY[rr][bg + 4 * q];
    }
  }
  double* P = part + (int64_t)blockIdx.x * na * nb;
  const int a = a0 + lane;
  if (a < na) {
#pragma unroll
    for (int q = 0; q < 16; q++) {
      int b = b0 + bg + 4 * q;
      if (b < nb) P[(int64_t)a * nb + b] = acc[q];
    }
  }
}

__global__ void k_xty_sum(int nchunk, int nab, const double* __restrict__ part, double* __restrict__ C) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nab) return;
  double s = 0.0;
  for (int c = 0; c < nchunk; c++) s += part[(int64_t)c * nab + e];
  C[e] = s;
}

// Y[r][j] = alpha*Y[r][j] + beta * sum_a X[r][a] S[a][j]; S is (na x nb) row-major.
__global__ __launch_bounds__(256) void k_gemm_small(int n, int na, int nb, const double* __restrict__ X,
                                                    int ldx, const double* __restrict__ Sm,
                                                    double* __restrict__ Y, int ldy, double alpha,
                                                    double beta) {
  extern __shared__ double sS[];
  for (int i = threadIdx.x; i < na * nb; i += blockDim.x) sS[i] = Sm[i];
  __syncthreads();
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int r = (int)(g / nb), j = (int)(g % nb);
  if (r >= n) return;
  double s = 0.0;
  for (int a = 0; a < na; a++) s += X[(int64_t)r * ldx + a] * sS[a * nb + j];
  double y = (alpha == 0.0) ? 0.0 : alpha * Y[(int64_t)r * ldy + j];
  Y[(int64_t)r * ldy + j] = y + beta * s;
}

// X[(ndof1 + b) * nrhs + b] = 1, all other entries 0 (the [B; I] identity rows)
__global__ void k_identity_rows(double* __restrict__ X, int64_t nrows, int nrhs, int ndof1) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nrows * nrhs) return;
  int64_t r = g / nrhs;
  int j = (int)(g % nrhs);
  X[g] = (r >= ndof1 && r - ndof1 == j) ? 1.0 : 0.0;
}

// ---------------------------------------------------------------- launchers
static inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

void launch_hex8(hipStream_t st, int nhex, const int* elist, const int* eptr, const int* enodes,
                 const double* xyz, const double* eprop, const int64_t* eoff, double* Kp, double* Mp,
                 int* ebad) {
  if (nhex > 0) k_hex8<<<nhex, 64, 0, st>>>(nhex, elist, eptr, enodes, xyz, eprop, eoff, Kp, Mp, ebad);
}
void launch_assemble(hipStream_t st, int64_t nnz, const int* cptr, const int64_t* cidx, const double* Kp,
                     const double* Mp, double* Kv, double* Mv, const int64_t* fmap, double* pool) {
  if (nnz > 0) k_assemble<<<nblk(nnz, 256), 256, 0, st>>>(nnz, cptr, cidx, Kp, Mp, Kv, Mv, fmap, pool);
}
void launch_fill(hipStream_t st, double* p, int64_t n, double v) {
  if (n > 0) k_fill<<<(unsigned)std::min<int64_t>(nblk(n, 256), 16384), 256, 0, st>>>(p, n, v);
}
void launch_extend_add(hipStream_t st, const FTask* t, int n, const DevSym& S, double* pool) {
  if (n > 0) k_extend_add<<<n, 256, 0, st>>>(t, S, pool);
}
void launch_panel(hipStream_t st, const FTask* t, int n, const DevSym& S, double* pool, int* piv) {
  if (n > 0) k_panel<<<n, 256, 0, st>>>(t, S, pool, piv);
}
void launch_update(hipStream_t st, const FTask* t, int n, const DevSym& S, double* pool) {
  if (n > 0) k_update<<<n, 256, 0, st>>>(t, S, pool);
}
void launch_solve_gemm(hipStream_t st, const SolveTask* t, int n, const DevSym& S, const double* pool,
                       double* W, int nrhs, int backward) {
  if (n > 0) k_solve_gemm<<<n, 256, 0, st>>>(t, S, pool, W, nrhs, backward);
}
void launch_solve_diag(hipStream_t st, const SolveTask* t, int n, const DevSym& S, const double* pool,
                       double* W, int nrhs, int backward) {
  if (n > 0) k_solve_diag<<<n, 64, 0, st>>>(t, S, pool, W, nrhs, backward);
}
void launch_fw_init(hipStream_t st, const int* sns, int n, const DevSym& S, const double* R, double* W,
                    int nrhs, int ndof1) {
  if (n > 0) k_fw_init<<<n, 256, 0, st>>>(sns, S, R, W, nrhs, ndof1);
}
void launch_fw_out(hipStream_t st, const int* sns, int n, const DevSym& S, const double* pool,
                   const double* W, double* Y, int nrhs, int scale, int ndof1) {
  if (n > 0) k_fw_out<<<n, 256, 0, st>>>(sns, S, pool, W, Y, nrhs, scale, ndof1);
}
void launch_bw_init(hipStream_t st, const int* sns, int n, const DevSym& S, const double* Y,
                    const double* X, double* W, int nrhs) {
  if (n > 0) k_bw_init<<<n, 256, 0, st>>>(sns, S, Y, X, W, nrhs);
}
void launch_bw_out(hipStream_t st, const int* sns, int n, const DevSym& S, const double* W, double* X,
                   int nrhs) {
  if (n > 0) k_bw_out<<<n, 256, 0, st>>>(sns, S, W, X, nrhs);
}
void launch_spmm(hipStream_t st, int n, int nrhs, const int* rowptr, const int* colind, const double* val,
                 const double* X, int ldx, double* Z, int ldz) {
  if (n > 0 && nrhs > 0)
    k_spmm<<<nblk((int64_t)n * nrhs, 256), 256, 0, st>>>(n, nrhs, rowptr, colind, val, X, ldx, Z, ldz);
}
int xty_chunks(int n) { return (n + XTY_ROWS - 1) / XTY_ROWS; }
void launch_xty(hipStream_t st, int n, int na, int nb, const double* X, int ldx, const double* Y, int ldy,
                double* part, double* C) {
  if (na <= 0 || nb <= 0) return;
  int nch = xty_chunks(n);
  if (nch == 0) { launch_fill(st, C, (int64_t)na * nb, 0.0); return; }
  dim3 g(nch, (na + 63) / 64, (nb + 63) / 64);
  k_xty_part<<<g, 256, 0, st>>>(n, na, nb, X, ldx, Y, ldy, part);
  k_xty_sum<<<nblk((int64_t)na * nb, 256), 256, 0, st>>>(nch, na * nb, part, C);
}
void launch_gemm_small(hipStream_t st, int n, int na, int nb, const double* X, int ldx, const double* Sm,
                       double* Y, int ldy, double alpha, double beta) {
  if (n > 0 && nb > 0)
    k_gemm_small<<<nblk((int64_t)n * nb, 256), 256, sizeof(double) * na * nb, st>>>(n, na, nb, X, ldx, Sm, Y,
                                                                                   ldy, alpha, beta);
}
void launch_identity_rows(hipStream_t st, double* X, int64_t nrows, int nrhs, int ndof1) {
  if (nrows * nrhs > 0) k_identity_rows<<<nblk(nrows * nrhs, 256), 256, 0, st>>>(X, nrows, nrhs, ndof1);
}

}  // namespace fgpu
