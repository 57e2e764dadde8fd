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
#include <climits>
#include <type_traits>
#include "kernels.hpp"

namespace fgpu {

using f64x4 = __attribute__((__vector_size__(4 * sizeof(double)))) double;

// Workgroups of a launch are dispatched round-robin over the eight XCDs of
// the MI355X (workgroup b on XCD b mod 8), each with its own L2.  xcd_task
// maps workgroup b to a task index so that every XCD takes one contiguous
// range of the task list: neighbouring tasks, which share panel columns or
// source rows, then share an L2.  A bijection of [0, n).
__device__ __forceinline__ int xcd_task(int b, int n) {
  const int q = n >> 3, r = n & 7, c = b & 7, m = b >> 3;
  return c < r ? c * (q + 1) + m : r * (q + 1) + (c - r) * q + m;
}

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
template <bool COEF>
__global__ __launch_bounds__(256) void k_assemble(int64_t nnz, const int* __restrict__ cptr,
                                                  const int* __restrict__ cidx,
                                                  const double* __restrict__ ccoef,
                                                  const double* __restrict__ Kp,
                                                  const double* __restrict__ Mp,
                                                  double* __restrict__ Kv, double* __restrict__ Mv,
                                                  const int64_t* __restrict__ fmap,
                                                  double* __restrict__ pool) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnz) return;
  double kv = 0.0, mv = 0.0;
  for (int q = cptr[k]; q < cptr[k + 1]; q++) {
    const int c = cidx[q];
    if (COEF) {  // rigid-element constraint coefficient of this contribution
      const double w = ccoef[q];
      kv += w * Kp[c];
      mv += w * Mp[c];
    } else {
      kv += Kp[c];
      mv += Mp[c];
    }
  }
  Kv[k] = kv;
  Mv[k] = mv;
  if (pool) {  // FP64 factor: original entries straight into the fronts
    int64_t f = fmap[k];
    if (f >= 0) pool[f] = kv;
  }
}

// original diagonal a_jj of the first n permuted equations (times rs_j^2 when
// rs is given: the equilibrated matrix) and the largest |a_jj| in *amax
// (non-negative doubles order as their bit patterns; *amax starts at 0)
__global__ void k_diag_extract(int n, const int* __restrict__ rowptr, const int* __restrict__ colind,
                               const double* __restrict__ val, const double* __restrict__ rs,
                               double* __restrict__ adiag, unsigned long long* __restrict__ amax) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double a = 0.0;
  for (int k = rowptr[r]; k < rowptr[r + 1]; k++)
    if (colind[k] == r) a = val[k];
  if (rs) a *= rs[r] * rs[r];
  adiag[r] = a;
  atomicMax(amax, (unsigned long long)__double_as_longlong(fabs(a)));
}

// pivot statistics of the factor: out[0] negative pivots, out[1] true-zero
// and out[2] reduced-to-zero pivots replaced, out[3..5] the first such
// pivot columns (INT_MAX: none); out starts at {0, 0, 0, INT_MAX x 3}
__global__ void k_pivot_stats(int n, const double* __restrict__ dpiv, const int* __restrict__ sing,
                              int* __restrict__ out) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
    if (sing[c]) {
      atomicAdd(&out[sing[c]], 1);
      atomicMin(&out[3 + sing[c]], c);
    } else if (dpiv[c] < 0.0) {
      atomicAdd(&out[0], 1);
      atomicMin(&out[3], c);
    }
  }
}

__global__ void k_fill(double* __restrict__ p, int64_t n, double v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

// ---------------------------------------------------------------- extend-add
#define XADD_REL 4096
// Task: sn = parent front, a = first parent column of a 16-column block.  The
// workgroup adds, child after child, every child update-matrix column whose
// parent column falls in its block; the task's b/kb index the (child,
// first, end) column ranges in xrange, found when the schedule was built.
// Blocks are disjoint, so no two workgroups touch the same parent entry;
// children are added in list order.
template <class T>
__global__ __launch_bounds__(256) void k_extend_add(const FTask* __restrict__ tasks, DevSym S,
                                                    const int* __restrict__ xrange, T* __restrict__ pool) {
  // the child's parent positions rel[a..nu) staged in LDS (when they fit),
  // so that the parent loads of a batch wait on one memory round, not two
  __shared__ int srel[XADD_REL];
  const FTask t = tasks[xcd_task(blockIdx.x, gridDim.x)];
  const int sn = t.sn;
  const int ldp = S.nrow[sn];
  T* P = pool + S.foff[sn];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int q = 0; q < t.kb; q++) {
    const int* xr = xrange + 3 * (t.b + q);
    const int c = xr[0], a = xr[1], b = xr[2];
    const int nc = S.ncol[c], ldc = S.nrow[c], nu = ldc - nc;
    const int* rel = S.rel + S.relptr[c];
    const T* C = pool + S.foff[c] + (int64_t)nc * ldc + nc;
    const bool staged = nu - a <= XADD_REL;
    if (staged) {
      for (int i = a + (int)threadIdx.x; i < nu; i += 256) srel[i - a] = rel[i];
      __syncthreads();
    }
    // rl(i) = rel[i] for i >= a, from LDS or from memory
    auto columns = [&](auto rl) {
      for (int jj = a + w; jj < b; jj += 4) {
        const int pc = rl(jj);
        T* dst = P + (int64_t)pc * ldp;
        const T* src = C + (int64_t)jj * ldc;
        // the rows of one child column land in distinct parent rows, so
        // eight iterations' loads go out before their stores
        for (int i0 = jj + lane; i0 < nu; i0 += 64 * 8) {
          int pr[8];
          T v[8], o[8];
#pragma unroll
          for (int u = 0; u < 8; u++) {
            const int ii = i0 + 64 * u;
            pr[u] = ii < nu ? rl(ii) : -1;
            v[u] = ii < nu ? src[ii] : (T)0;
          }
#pragma unroll
          for (int u = 0; u < 8; u++) o[u] = pr[u] >= 0 ? dst[pr[u]] : (T)0;
#pragma unroll
          for (int u = 0; u < 8; u++)
            if (pr[u] >= 0) dst[pr[u]] = o[u] + v[u];
        }
      }
    };
    if (staged) columns([&](int i) { return srel[i - a]; });
    else columns([&](int i) { return rel[i]; });
    __syncthreads();
  }
}

// ---------------------------------------------------------------- MFMA tile
// acc[x][y] += sum_m sA[wr+16x+(lane&15)][m] * sB[wc+16y+...][m] on the
// 16x16x4 MFMA of the storage type (v_mfma_f64_16x16x4f64 for the FP64
// factor, v_mfma_f32_16x16x4f32 for the mixed-precision one) for a 64x64
// tile split in four 32x32 wave quarters.  The A operand of the instruction
// takes the sB rows and its B operand the sA rows, so for fixed register
// consecutive lanes hold consecutive rows of sA (coalesced stores into
// column-major fronts).  The two instructions place the 16x16 result
// differently: D[p][q] sits in lane q + 16*(p % 4), register p / 4 for FP64
// and in lane q + 16*(p / 4), register p % 4 for FP32.  Result element of
// acc[x][y][r]: row wr + 16x + (lane & 15), column wc + 16y + MF<T>::col(lane >> 4, r).
using f32x4 = __attribute__((__vector_size__(4 * sizeof(float)))) float;

template <class T> struct MF;
template <> struct MF<double> {
  using acc = f64x4;
  static __device__ __forceinline__ acc op(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int col(int lk, int r) { return lk + 4 * r; }
};
template <> struct MF<float> {
  using acc = f32x4;
  static __device__ __forceinline__ acc op(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int col(int lk, int r) { return 4 * lk + r; }
};

template <class T, int LD>
__device__ __forceinline__ void mfma_tile(const T (*sA)[LD], const T (*sB)[LD], int kbr,
                                          typename MF<T>::acc (&acc)[2][2], int lane, int wr, int wc) {
  const int lr = lane & 15, lk = lane >> 4;
  for (int m0 = 0; m0 < kbr; m0 += 4) {
    T ab[2], aa[2];
#pragma unroll
    for (int x = 0; x < 2; x++) {
      ab[x] = sB[wc + 16 * x + lr][m0 + lk];
      aa[x] = sA[wr + 16 * x + lr][m0 + lk];
    }
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
      for (int y = 0; y < 2; y++) acc[x][y] = MF<T>::op(ab[y], aa[x], acc[x][y]);
  }
}

template <class A>
__device__ __forceinline__ void zero_acc(A (&acc)[2][2]) {
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++) acc[x][y] = A{};
}

// ---------------------------------------------------------------- panel
// Value of v in lane src (uniform src), through the scalar unit: no LDS
// round trip and no memory wait between the broadcasts of a sweep.
__device__ __forceinline__ double lane_bcast(double v, int src) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), src);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), src);
  return __hiloint2double(hi, lo);
}

// k_diag task: sn, k0 = first panel column.  Blocked LDL^T of the kb x kb
// (kb <= 64) diagonal block in LDS by four waves, in 16-column steps:
// (1) wave 0 factors the 16 x 16 diagonal sub-block with its rows in
// registers and readlane broadcasts (L11 unit lower, D1, X11 = L11^-1);
// (2) the rows below form L21 = A21 X11^T D1^-1 on FP64 MFMA, one 16-row
// tile per wave; (3) the lower trailing block takes A22 -= (L21 D1) L21^T,
// 16 x 16 MFMA tiles shared by the waves.  The off-diagonal tiles of
// X = L^-1 then follow by distance from the diagonal:
// X_ij = -X_ii sum_{k=j}^{i-1} L_ik X_kj.  L and D go back into the front,
// X into the diagonal tile of the pivot-block inverse pool.  The arithmetic
// is FP64 for both storage types (the FP32 factor of the mixed-precision
// path keeps its diagonal blocks' factorisation in FP64 and rounds only the
// stored result).  Rows beyond kb are carried as identity rows.
//
// diag_block: that factorisation in LDS; sA holds L (unit lower) with D on
// its diagonal, sD = D, sX = L^-1, all in FP64.
template <class T>
__device__ __forceinline__ void diag_block(const T* __restrict__ F, int ld, int k0, int kb, double (&sA)[64][65],
                                           double (&sX)[64][65], double (&sD)[64], double (&sT)[4][16][17],
                                           const DevSym& S, int colbase) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  // the singular-pivot test's |a_jj| of the block's columns, lane j of wave
  // 0 holding column j: read once here, broadcast in the sweeps below (a
  // load per pivot would put a memory latency on every step of the chain)
  const double adv = (wv == 0 && lane < kb) ? S.adiag[colbase + lane] : 0.0;
  for (int idx = tid; idx < 64 * 64; idx += 256) {
    const int i = idx & 63, m = idx >> 6;
    sA[i][m] = (i < kb && m <= i) ? (double)F[(int64_t)(k0 + i) + (int64_t)(k0 + m) * ld] : (i == m ? 1.0 : 0.0);
    sX[i][m] = (i == m) ? 1.0 : 0.0;
  }
  __syncthreads();
  const int nsb = (kb + 15) >> 4;
  for (int p = 0; p < nsb; p++) {
    const int c0 = 16 * p;
    // (1) the 16 x 16 diagonal sub-block: lane i < 16 holds row c0 + i
    if (wv == 0) {
      double a[16], x[16];
#pragma unroll
      for (int m = 0; m < 16; m++) {
        a[m] = (lane < 16 && m <= lane) ? sA[c0 + lane][c0 + m] : 0.0;
        x[m] = (m == lane) ? 1.0 : 0.0;
      }
#pragma unroll
      for (int j = 0; j < 16; j++) {
        double dj = lane_bcast(a[j], j);
        if (c0 + j < kb) {  // singular pivot: replaced, column flagged (same decision on every lane)
          const int col = colbase + c0 + j;
          const double aj = lane_bcast(adv, c0 + j);
          if (!(fabs(dj) > S.tolf * fabs(aj))) {
            const bool truez = aj == 0.0 || fabs(dj) <= 1.0e3 * 2.220446049250313e-16 * fabs(aj);
            dj = aj != 0.0 ? fabs(aj) : *S.amax;
            if (lane == j) a[j] = dj;
            if (lane == 0) S.sing[col] = truez ? 1 : 2;
          }
        }
        const bool below = lane > j && lane < 16;
        const double l = below ? a[j] / dj : 0.0;
        if (below) a[j] = l;
        const double ldj = l * dj;
#pragma unroll
        for (int m = j + 1; m < 16; m++) {
          const double lm = lane_bcast(l, m);  // L[m][j]
          if (m <= lane) a[m] -= ldj * lm;
        }
#pragma unroll
        for (int m = 0; m <= j; m++) x[m] -= l * lane_bcast(x[m], j);  // X[j][m]
      }
      if (lane < 16) {
        double dd = 0.0;
#pragma unroll
        for (int m = 0; m < 16; m++) {
          if (m <= lane) {
            sA[c0 + lane][c0 + m] = a[m];
            sX[c0 + lane][c0 + m] = x[m];
          }
          dd = (m == lane) ? a[m] : dd;
        }
        sD[c0 + lane] = dd;
      }
    }
    __syncthreads();
    // (2) L21 = A21 X11^T D1^-1, wave w the row tile at c0 + 16 (w + 1)
    const int rb0 = c0 + 16, nrt = nsb - p - 1;
    if (wv < nrt) {
      const int r0 = rb0 + 16 * wv;
      f64x4 acc = f64x4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int m0 = 0; m0 < 16; m0 += 4)
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(sX[c0 + lr][c0 + m0 + lk], sA[r0 + lr][c0 + m0 + lk], acc, 0, 0, 0);
      // acc[r]: row r0 + lr, column c0 + lk + 4 r
#pragma unroll
      for (int r = 0; r < 4; r++) sA[r0 + lr][c0 + lk + 4 * r] = acc[r] / sD[c0 + lk + 4 * r];
    }
    __syncthreads();
    // (3) A22 -= (L21 D1) L21^T over the lower 16 x 16 tiles (ti >= tj)
    for (int q = wv; q < nrt * (nrt + 1) / 2; q += 4) {
      const int ti = q == 0 ? 0 : (q < 3 ? 1 : (q < 6 ? 2 : 3));
      const int tj = q - ti * (ti + 1) / 2;
      const int ri = rb0 + 16 * ti, rj = rb0 + 16 * tj;
      f64x4 acc = f64x4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int m0 = 0; m0 < 16; m0 += 4) {
        const int m = c0 + m0 + lk;
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(sA[rj + lr][m], sA[ri + lr][m] * sD[m], acc, 0, 0, 0);
      }
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int i = ri + lr, j = rj + lk + 4 * r;
        if (i >= j) sA[i][j] -= acc[r];
      }
    }
    __syncthreads();
  }
  // off-diagonal 16 x 16 tiles of X = L^-1, by distance d from the diagonal
  for (int d = 1; d < nsb; d++) {
    const int j = wv, i = wv + d;
    if (i < nsb) {
      f64x4 acc = f64x4{0.0, 0.0, 0.0, 0.0};
      for (int kt = j; kt < i; kt++)
#pragma unroll
        for (int m0 = 0; m0 < 16; m0 += 4)
          acc = __builtin_amdgcn_mfma_f64_16x16x4f64(sX[16 * kt + m0 + lk][16 * j + lr],
                                                     sA[16 * i + lr][16 * kt + m0 + lk], acc, 0, 0, 0);
      // T (rows of tile i, columns of tile j) through this wave's LDS tile
#pragma unroll
      for (int r = 0; r < 4; r++) sT[wv][lr][lk + 4 * r] = acc[r];
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      f64x4 acc2 = f64x4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int m0 = 0; m0 < 16; m0 += 4)
        acc2 = __builtin_amdgcn_mfma_f64_16x16x4f64(sT[wv][m0 + lk][lr], sX[16 * i + lr][16 * i + m0 + lk], acc2, 0,
                                                    0, 0);
#pragma unroll
      for (int r = 0; r < 4; r++) sX[16 * i + lr][16 * j + lk + 4 * r] = -acc2[r];
    }
    __syncthreads();
  }
}

// X = L^-1 into the diagonal tile of the pivot-block inverse pool and D into
// the pivot array, plus the positive-definiteness check of the pivots.  The
// front's diagonal block itself is left alone (later kernels read D from the
// pivot array, the solves L^-1 from the inverse pool), so that other
// workgroups of the same launch can still read it.
template <class T>
__device__ __forceinline__ void diag_store(T* __restrict__ Xg, double* __restrict__ dpiv, int nc, int k0, int kb,
                                           int col0, const double (&sA)[64][65], const double (&sX)[64][65],
                                           int* __restrict__ piv) {
  const int tid = threadIdx.x;
  for (int idx = tid; idx < 64 * 64; idx += 256) {
    const int i = idx & 63, m = idx >> 6;
    if (i < kb && m <= i) Xg[(int64_t)(k0 + i) + (int64_t)(k0 + m) * nc] = (T)sX[i][m];
  }
  if (tid < kb) dpiv[col0 + k0 + tid] = sA[tid][tid];
  (void)piv;
}

template <class T>
__global__ __launch_bounds__(256) void k_diag(const FTask* __restrict__ tasks, DevSym S,
                                              T* __restrict__ pool, const int64_t* __restrict__ ioff,
                                              T* __restrict__ ipool, int* __restrict__ piv) {
  __shared__ double sA[64][65], sX[64][65], sD[64], sT[4][16][17];
  const FTask t = tasks[blockIdx.x];
  const int sn = t.sn, k0 = t.k0, ld = S.nrow[sn], nc = S.ncol[sn];
  const int kb = min(64, nc - k0);
  T* F = pool + S.foff[sn];
  T* Xg = ipool + ioff[sn];
  diag_block<T>(F, ld, k0, kb, sA, sX, sD, sT, S, S.col0[sn] + k0);
  diag_store<T>(Xg, S.dpiv, nc, k0, kb, S.col0[sn], sA, sX, piv);
}

// Diagonal block and triangular solve in one launch, for panels with few
// row tiles below the diagonal block (the upper levels of the tree, where a
// separate diagonal-block launch would run on a handful of CUs).  Task: sn,
// a = row tile (-1: none below), k0 = panel.  Every workgroup factors the
// diagonal block in LDS (the same FP64 arithmetic as k_diag, so the same
// bits); the one with a <= 0 stores it.  The row tile then takes
// L_r = F_r L_kk^{-T} D^{-1} with the factor values rounded to the storage
// type, exactly as k_trsm reads them back from memory.  The row tile's
// loads are issued before the factorisation.
template <class T>
__global__ __launch_bounds__(256) void k_diag_trsm(const FTask* __restrict__ tasks, DevSym S,
                                                   T* __restrict__ pool, const int64_t* __restrict__ ioff,
                                                   T* __restrict__ ipool, int* __restrict__ piv) {
  __shared__ double sA[64][65], sX[64][65], sD[64], sT[4][16][17];
  const FTask t = tasks[xcd_task(blockIdx.x, gridDim.x)];
  const int sn = t.sn, k0 = t.k0, ld = S.nrow[sn], nc = S.ncol[sn];
  const int kb = min(64, nc - k0);
  T* F = pool + S.foff[sn];
  T* Xg = ipool + ioff[sn];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int R0 = k0 + kb + 64 * max(t.a, 0);
  T ra[16];
  if (t.a >= 0) {
#pragma unroll
    for (int q = 0; q < 16; q++) {
      const int idx = tid + 256 * q, i = idx & 63, m = idx >> 6;
      ra[q] = (m < kb && R0 + i < ld) ? F[(int64_t)(R0 + i) + (int64_t)(k0 + m) * ld] : (T)0;
    }
  }
  diag_block<T>(F, ld, k0, kb, sA, sX, sD, sT, S, S.col0[sn] + k0);
  if (t.a <= 0) diag_store<T>(Xg, S.dpiv, nc, k0, kb, S.col0[sn], sA, sX, piv);
  if (t.a < 0) return;
  __syncthreads();  // every read of sA by the store is done
  // operand tiles in the storage type: the row tile in sA's space, L^-1 in
  // sX's space (in place for double: each thread rewrites only its entries)
  T(*sF)[65] = reinterpret_cast<T(*)[65]>(&sA[0][0]);
  T(*sL)[65] = reinterpret_cast<T(*)[65]>(&sX[0][0]);
  T xv[16];
#pragma unroll
  for (int q = 0; q < 16; q++) {
    const int idx = tid + 256 * q, i = idx & 63, m = idx >> 6;
    xv[q] = (m < kb && i < kb && i >= m) ? (T)sX[i][m] : (T)0;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 16; q++) {
    const int idx = tid + 256 * q, i = idx & 63, m = idx >> 6;
    sF[i][m] = ra[q];
    sL[i][m] = xv[q];
  }
  __syncthreads();
  const int kbr = (kb + 3) & ~3;
  const int wr = (wv >> 1) * 32, wc = (wv & 1) * 32, lr = lane & 15, lk = lane >> 4;
  typename MF<T>::acc acc[2][2];
  zero_acc(acc);
  mfma_tile<T>(sF, sL, kbr, acc, lane, wr, wc);
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int i = R0 + wr + 16 * x + lr;
        const int j = wc + 16 * y + MF<T>::col(lk, r);
        if (i < ld && j < kb) F[(int64_t)i + (int64_t)(k0 + j) * ld] = acc[x][y][r] / (T)sD[j];
      }
}

// k_trsm task: sn, a = 64-row tile below the diagonal block, k0 = panel.
// L_r = F_r L_kk^{-T} D^{-1}, as an FP64 MFMA product with the inverse of the
// diagonal block built by k_diag.
template <class T>
__global__ __launch_bounds__(256) void k_trsm(const FTask* __restrict__ tasks, DevSym S,
                                              T* __restrict__ pool, const int64_t* __restrict__ ioff,
                                              const T* __restrict__ ipool) {
  __shared__ T sA[64][65];
  __shared__ T sB[64][65];
  const FTask t = tasks[xcd_task(blockIdx.x, gridDim.x)];
  const int sn = t.sn, k0 = t.k0, ld = S.nrow[sn], nc = S.ncol[sn];
  const int kb = min(64, nc - k0);
  const int R0 = k0 + kb + 64 * t.a;
  T* F = pool + S.foff[sn];
  const T* Xg = ipool + ioff[sn];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int kbr = (kb + 3) & ~3;
  {
    T ra[16], rb[16];
#pragma unroll
    for (int q = 0; q < 16; q++) {
      const int idx = tid + 256 * q, i = idx & 63, m = idx >> 6;
      ra[q] = (m < kb && R0 + i < ld) ? F[(int64_t)(R0 + i) + (int64_t)(k0 + m) * ld] : (T)0;
      // sB[j][m] = Linv_kk[j][m] (lower)
      rb[q] = (m < kb && i < kb && i >= m) ? Xg[(int64_t)(k0 + i) + (int64_t)(k0 + m) * nc] : (T)0;
    }
#pragma unroll
    for (int q = 0; q < 16; q++) {
      const int idx = tid + 256 * q, i = idx & 63, m = idx >> 6;
      sA[i][m] = ra[q];
      sB[i][m] = rb[q];
    }
  }
  __syncthreads();
  const int wr = (wv >> 1) * 32, wc = (wv & 1) * 32, lr = lane & 15, lk = lane >> 4;
  typename MF<T>::acc acc[2][2];
  zero_acc(acc);
  mfma_tile<T>(sA, sB, kbr, acc, lane, wr, wc);
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int i = R0 + wr + 16 * x + lr;
        const int j = wc + 16 * y + MF<T>::col(lk, r);
        if (i < ld && j < kb) F[(int64_t)i + (int64_t)(k0 + j) * ld] = acc[x][y][r] / (T)S.dpiv[S.col0[sn] + k0 + j];
      }
}

// ---------------------------------------------------------------- update
// Task: sn, a = ti, b = tj (64x64 tiles of the trailing lower triangle),
// k0 = panel, kb = panel width (<= 128).  F[I,J] -= (L_I D) L_J^T over the
// panel, staged through LDS in UCH-column chunks whose products accumulate
// in the MFMA registers, so a 128-wide panel reads and writes the trailing
// tile once for both halves.  FP64: the accumulators start from the tile's
// current values (loaded with the first chunk, so the two latencies
// overlap) and the staged L_I D is negated, so the MFMA chain forms
// F - sum (L_I D) L_J^T directly and the tile is only stored at the end.
// FP32 (mixed precision) sums the products from zero and subtracts them
// from the tile at the end: accumulating onto the tile through the FP32
// MFMA left the factor too inexact for the FP64 refinement of a thin
// 200x200 plate to converge.
template <class T>
__global__ __launch_bounds__(256) void k_update(const FTask* __restrict__ tasks, DevSym S,
                                                T* __restrict__ pool) {
  constexpr int UCH = 16, UQ = 64 * UCH / 256;  // chunk width, loads per thread per operand
  // two LDS buffers: chunk c is multiplied from buffer c & 1 while chunk
  // c + 1 is staged into the other, one barrier per chunk
  __shared__ T sI[2][64][UCH + 1];
  __shared__ T sJ[2][64][UCH + 1];
  const FTask t = tasks[xcd_task(blockIdx.x, gridDim.x)];
  const int sn = t.sn, k0 = t.k0, ld = S.nrow[sn];
  const int kb = t.kb;
  const int t0 = k0 + kb;
  const int I0 = t0 + 64 * t.a, J0 = t0 + 64 * t.b;
  T* F = pool + S.foff[sn];
  const int c0s = S.col0[sn];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wv = tid >> 6;
  const int wr = (wv >> 1) * 32, wc = (wv & 1) * 32;
  const int lr = lane & 15, lk = lane >> 4;
  T ri[UQ], rj[UQ], rd[UQ];
  // global loads of the chunk at c0 (zero beyond the panel)
  auto load = [&](int c0) {
    const int kc = min(UCH, kb - c0), p0 = k0 + c0;
#pragma unroll
    for (int q = 0; q < UQ; q++) {
      const int idx = tid + 256 * q, i = idx & 63, m = idx >> 6;
      const bool okm = m < kc;
      rd[q] = okm ? (T)S.dpiv[c0s + p0 + m] : (T)0;
      ri[q] = (okm && I0 + i < ld) ? F[(int64_t)(I0 + i) + (int64_t)(p0 + m) * ld] : (T)0;
      rj[q] = (okm && J0 + i < ld) ? F[(int64_t)(J0 + i) + (int64_t)(p0 + m) * ld] : (T)0;
    }
  };
  load(0);
  // FP64: the stored elements of the tile (lower triangle inside the
  // front) seed the accumulators; the others start at zero, never stored
  constexpr bool seed = std::is_same<T, double>::value;
  typename MF<T>::acc acc[2][2];
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int i = I0 + wr + 16 * x + lr;
        const int j = J0 + wc + 16 * y + MF<T>::col(lk, r);
        acc[x][y][r] = (seed && i < ld && j < ld && i >= j) ? F[(int64_t)i + (int64_t)j * ld] : (T)0;
      }
  auto stage = [&](int buf) {
#pragma unroll
    for (int q = 0; q < UQ; q++) {
      const int idx = tid + 256 * q, i = idx & 63, m = idx >> 6;
      sI[buf][i][m] = seed ? -(ri[q] * rd[q]) : ri[q] * rd[q];
      sJ[buf][i][m] = rj[q];
    }
  };
  stage(0);
  __syncthreads();
  for (int c0 = 0, buf = 0; c0 < kb; c0 += UCH, buf ^= 1) {
    const int kc = min(UCH, kb - c0);
    const bool more = c0 + UCH < kb;
    if (more) load(c0 + UCH);  // the next chunk's loads are in flight during the products
    mfma_tile<T>(sI[buf], sJ[buf], (kc + 3) & ~3, acc, lane, wr, wc);
    // the other buffer was last read by the products of the chunk before
    // this one, which every wave finished before the previous barrier
    if (more) {
      stage(buf ^ 1);
      __syncthreads();
    }
  }
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int i = I0 + wr + 16 * x + lr;
        const int j = J0 + wc + 16 * y + MF<T>::col(lk, r);
        if (i < ld && j < ld && i >= j) {
          if (seed) F[(int64_t)i + (int64_t)j * ld] = acc[x][y][r];
          else F[(int64_t)i + (int64_t)j * ld] -= acc[x][y][r];
        }
      }
}

// ---------------------------------------------------------------- inverse
// Explicit inverse X = L_PP^{-1} of the unit lower-triangular pivot block of
// every front, kept in its own pool (ioff[s], ncol x ncol column-major, lower
// triangle).  With it each triangular solve of a front is a GEMM, which turns
// the forward/backward passes of the block Lanczos and of the B computation
// into a few wide launches per tree level instead of a sequence of dependent
// 64-column substitutions.
//
// k_inv_off task: sn, a = tile column j, b = tile row i > j:
//   X_ij = - X_ii * sum_{k=j}^{i-1} L_ik X_kj
// (all tiles at smaller distance i-j are complete).
template <class T>
__global__ __launch_bounds__(256) void k_inv_off(const FTask* __restrict__ tasks, DevSym S,
                                                 const T* __restrict__ pool, const int64_t* __restrict__ ioff,
                                                 T* __restrict__ ipool) {
  __shared__ T sA[64][65];
  __shared__ T sB[64][65];
  const FTask t = tasks[blockIdx.x];
  const int sn = t.sn, ld = S.nrow[sn], nc = S.ncol[sn];
  const int rj = 64 * t.a, nj = min(64, nc - rj);
  const int ri = 64 * t.b, ni = min(64, nc - ri);
  const T* F = pool + S.foff[sn];
  T* X = ipool + ioff[sn];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int wr = (wv >> 1) * 32, wc = (wv & 1) * 32, lr = lane & 15, lk = lane >> 4;
  typename MF<T>::acc acc[2][2];
  zero_acc(acc);
  T ra[16], rb[16];
  // global loads of tile k (sA rows of tile i, sB the transposed column tile of X)
  auto load = [&](int kt) {
    const int rk = 64 * kt, nk = min(64, nc - rk);
#pragma unroll
    for (int q = 0; q < 16; q++) {
      const int idx = tid + 256 * q, r = idx & 63, m = idx >> 6;
      // sA[r][m] = L[ri+r][rk+m]  (rows of tile i, columns of tile k)
      ra[q] = (r < ni && m < nk) ? F[(int64_t)(ri + r) + (int64_t)(rk + m) * ld] : (T)0;
      // sB[c][m] = X[rk+m][rj+c]  (transposed, lower part only on the diagonal tile)
      const int c = idx >> 6, mm = idx & 63;
      const bool ok = c < nj && mm < nk && (kt > t.a || mm >= c);
      rb[q] = ok ? X[(int64_t)(rk + mm) + (int64_t)(rj + c) * nc] : (T)0;
    }
  };
  load(t.a);
  for (int kt = t.a; kt < t.b; kt++) {
    __syncthreads();  // the previous tile's products are done with the LDS tiles
#pragma unroll
    for (int q = 0; q < 16; q++) {
      const int idx = tid + 256 * q;
      sA[idx & 63][idx >> 6] = ra[q];
      sB[idx >> 6][idx & 63] = rb[q];
    }
    __syncthreads();
    if (kt + 1 < t.b) load(kt + 1);  // the next tile's loads are in flight during the products
    mfma_tile<T>(sA, sB, 64, acc, lane, wr, wc);
  }
  // T = acc (rows i, cols j) -> LDS as sB[c][m] = T[m][c] for the second product
  __syncthreads();
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++)
#pragma unroll
      for (int r = 0; r < 4; r++) sB[wc + 16 * y + MF<T>::col(lk, r)][wr + 16 * x + lr] = acc[x][y][r];
  for (int idx = tid; idx < 64 * 64; idx += 256) {
    int r = idx & 63, m = idx >> 6;
    sA[r][m] = (r < ni && m < ni && r >= m) ? X[(int64_t)(ri + r) + (int64_t)(ri + m) * nc] : (T)0;
  }
  __syncthreads();
  zero_acc(acc);
  mfma_tile<T>(sA, sB, 64, acc, lane, wr, wc);
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int i = wr + 16 * x + lr, c = wc + 16 * y + MF<T>::col(lk, r);
        if (i < ni && c < nj) X[(int64_t)(ri + i) + (int64_t)(rj + c) * nc] = -acc[x][y][r];
      }
}

// Split inverse tiles.  k_inv_part task: sn, a = tile column j, b = tile
// row i, k0 = first k tile of the chunk, kb = partial slot:
//   part[slot] = sum_{k=k0}^{min(k0+INV_CHUNK,i)-1} L_ik X_kj   (64 x 64, row-major)
// k_inv_fin task: sn, a = j, b = i, k0 = first slot, kb = number of slots:
//   X_ij = - X_ii * (sum of the partials in slot order).
template <class T>
__global__ __launch_bounds__(256) void k_inv_part(const FTask* __restrict__ tasks, DevSym S,
                                                  const T* __restrict__ pool, const int64_t* __restrict__ ioff,
                                                  const T* __restrict__ ipool, T* __restrict__ part) {
  __shared__ T sA[64][65];
  __shared__ T sB[64][65];
  const FTask t = tasks[blockIdx.x];
  const int sn = t.sn, ld = S.nrow[sn], nc = S.ncol[sn];
  const int rj = 64 * t.a, nj = min(64, nc - rj);
  const int ri = 64 * t.b, ni = min(64, nc - ri);
  const T* F = pool + S.foff[sn];
  const T* X = ipool + ioff[sn];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int wr = (wv >> 1) * 32, wc = (wv & 1) * 32, lr = lane & 15, lk = lane >> 4;
  const int kend = min(t.k0 + INV_CHUNK, t.b);
  typename MF<T>::acc acc[2][2];
  zero_acc(acc);
  T ra[16], rb[16];
  auto load = [&](int kt) {
    const int rk = 64 * kt, nk = min(64, nc - rk);
#pragma unroll
    for (int q = 0; q < 16; q++) {
      const int idx = tid + 256 * q, r = idx & 63, m = idx >> 6;
      ra[q] = (r < ni && m < nk) ? F[(int64_t)(ri + r) + (int64_t)(rk + m) * ld] : (T)0;
      const int c = idx >> 6, mm = idx & 63;
      const bool ok = c < nj && mm < nk && (kt > t.a || mm >= c);
      rb[q] = ok ? X[(int64_t)(rk + mm) + (int64_t)(rj + c) * nc] : (T)0;
    }
  };
  load(t.k0);
  for (int kt = t.k0; kt < kend; kt++) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 16; q++) {
      const int idx = tid + 256 * q;
      sA[idx & 63][idx >> 6] = ra[q];
      sB[idx >> 6][idx & 63] = rb[q];
    }
    __syncthreads();
    if (kt + 1 < kend) load(kt + 1);
    mfma_tile<T>(sA, sB, 64, acc, lane, wr, wc);
  }
  T* P = part + (int64_t)t.kb * 4096;
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++)
#pragma unroll
      for (int r = 0; r < 4; r++) P[(wr + 16 * x + lr) * 64 + wc + 16 * y + MF<T>::col(lk, r)] = acc[x][y][r];
}

template <class T>
__global__ __launch_bounds__(1024) void k_inv_fin(const FTask* __restrict__ tasks, DevSym S,
                                                  const int64_t* __restrict__ ioff, T* __restrict__ ipool,
                                                  const T* __restrict__ part) {
  __shared__ T sA[64][65];
  __shared__ T sB[64][65];
  const FTask t = tasks[blockIdx.x];
  const int sn = t.sn, nc = S.ncol[sn];
  const int rj = 64 * t.a, nj = min(64, nc - rj);
  const int ri = 64 * t.b, ni = min(64, nc - ri);
  T* X = ipool + ioff[sn];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  // sum of the partials (slot order) for four entries per thread, every
  // slot's loads issued before the additions; X_ii lower -> sA
  {
    T v[4];
#pragma unroll
    for (int e = 0; e < 4; e++) v[e] = (T)0;
    const T* p = part + (int64_t)t.k0 * 4096 + tid;
    int q = 0;
    for (; q + 1 < t.kb; q += 2) {
      T a0[4], a1[4];
#pragma unroll
      for (int e = 0; e < 4; e++) {
        a0[e] = p[(int64_t)q * 4096 + 1024 * e];
        a1[e] = p[(int64_t)(q + 1) * 4096 + 1024 * e];
      }
#pragma unroll
      for (int e = 0; e < 4; e++) {
        v[e] += a0[e];
        v[e] += a1[e];
      }
    }
    if (q < t.kb) {
#pragma unroll
      for (int e = 0; e < 4; e++) v[e] += p[(int64_t)q * 4096 + 1024 * e];
    }
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const int idx = tid + 1024 * e, m = idx >> 6, c = idx & 63;  // T[m][c] -> sB[c][m]
      sB[c][m] = v[e];
      const int r = idx & 63, mm = idx >> 6;  // X_ii lower -> sA[r][mm]
      sA[r][mm] = (r < ni && mm < ni && r >= mm) ? X[(int64_t)(ri + r) + (int64_t)(ri + mm) * nc] : (T)0;
    }
  }
  __syncthreads();
  if (wv >= 4) return;
  const int wr = (wv >> 1) * 32, wc = (wv & 1) * 32, lr = lane & 15, lk = lane >> 4;
  typename MF<T>::acc acc[2][2];
  zero_acc(acc);
  mfma_tile<T>(sA, sB, 64, acc, lane, wr, wc);
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int i = wr + 16 * x + lr, c = wc + 16 * y + MF<T>::col(lk, r);
        if (i < ni && c < nj) X[(int64_t)(ri + i) + (int64_t)(rj + c) * nc] = -acc[x][y][r];
      }
}

// ---------------------------------------------------------------- solves
// Multi-vectors are row-major (rows x nrhs).  Work buffer W holds, for front
// s, rows woff[s] .. woff[s]+nrow[s]-1; V holds the solved pivot rows before
// D-scaling.  A k_mv task multiplies target rows r0..r0+rn-1 (rn <= 64) and
// right-hand sides j0..j0+63 of front sn over the source slice k0..k0+kn-1:
//   mode 0 FW_INV : V[r]      =  sum_{k<=r}   Linv[r][k]   W[k]
//   mode 1 FW_UPD : W[nc+r]  -=  sum_{k<nc}   L[nc+r][k]   V[k]
//   mode 2 BW_UPD : W[r]     -=  sum_{k<nu}   L[nc+k][r]   X[rows[nc+k]]
//   mode 3 BW_INV : X[c0+r]   =  sum_{k>=r}   Linv[k][r]   W[k]
// Long K ranges are split over several tasks (slot >= 0) whose 64x64
// partials k_mv_red then sums in slot order (for the forward update rows
// with at most 64 right-hand sides, the parent's k_fw_init does), so the
// result does not depend on scheduling.
__device__ __forceinline__ void mv_apply(int mode, const DevSym& S, int sn, int nc, int r, int j, int nrhs,
                                         double v, double* Wf, double* Vf, double* X) {
  switch (mode) {
    case 0: Vf[(int64_t)r * nrhs + j] = v; break;
    case 1: Wf[(int64_t)(nc + r) * nrhs + j] -= v; break;
    case 2: Wf[(int64_t)r * nrhs + j] -= v; break;
    default: X[(int64_t)(S.col0[sn] + r) * nrhs + j] = v; break;
  }
}

template <int MODE, int NJ, class T>
__global__ __launch_bounds__(256) void k_mv(const SolveTask* __restrict__ tasks, DevSym S,
                                            const T* __restrict__ pool, const int64_t* __restrict__ ioff,
                                            const T* __restrict__ ipool, double* __restrict__ W,
                                            double* __restrict__ V, double* __restrict__ X,
                                            double* __restrict__ part, int nrhs) {
  constexpr int CH = 32;            // K chunk staged in LDS (small tiles: more workgroups per CU)
  __shared__ double sA[64][CH + 1];
  __shared__ double sB[NJ][CH + 1];  // B tile transposed: sB[j][k]
  const SolveTask t = tasks[xcd_task(blockIdx.x, gridDim.x)];
  const int sn = t.sn, ld = S.nrow[sn], nc = S.ncol[sn];
  const T* F = pool + S.foff[sn];
  const T* Xi = ipool + ioff[sn];
  double* Wf = W + S.woff[sn] * (int64_t)nrhs;
  double* Vf = V + S.woff[sn] * (int64_t)nrhs;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int nj = min(NJ, nrhs - t.j0);
  const int r0 = t.r0, rn = t.rn;
  // B rows: mode 1 the solved pivot rows V, mode 2 the final solution rows
  // of the front's update rows (X[rows[nc + k]]), modes 0 and 3 W
  const double* Bsrc = MODE == 1 ? Vf : Wf;
  const int* urows = S.rows + S.rptr[sn] + nc;
  constexpr bool rowmajorA = (MODE == 0 || MODE == 1);
  constexpr int NT = NJ / 16;   // 16x16 MFMA result tiles per wave: rows 16 wv .. 16 wv + 15, all NJ columns
  constexpr int BPT = NJ * CH / 256;  // B-tile values loaded per thread (CH x NJ tile)
  f64x4 acc[NT];
#pragma unroll
  for (int y = 0; y < NT; y++) acc[y] = f64x4{0.0, 0.0, 0.0, 0.0};
  const int kend = t.k0 + t.kn;
  double ra[64 * CH / 256], rb[BPT];
  // mode 2: the X row of each B value this thread loads, fetched one chunk
  // ahead of the B loads that use them
  int ur[BPT];
  auto load_rows = [&](int kc) {
#pragma unroll
    for (int q = 0; q < BPT; q++) {
      const int kb = (tid + 256 * q) / NJ;
      ur[q] = kc + kb < kend ? urows[kc + kb] : 0;
    }
  };
  if (MODE == 2) load_rows(t.k0);
  // global loads of the CH-wide chunk starting at kc (zero outside the task)
  auto load = [&](int kc) {
    const int kn = min(CH, kend - kc);
#pragma unroll
    for (int q = 0; q < 64 * CH / 256; q++) {
      const int idx = tid + 256 * q;
      const int i = rowmajorA ? (idx & 63) : (idx / CH);
      const int k = rowmajorA ? (idx >> 6) : (idx % CH);
      const int r = r0 + i, kk = kc + k;
      double v = 0.0;
      if (i < rn && k < kn) {
        if (MODE == 0) v = kk <= r ? (double)Xi[(int64_t)r + (int64_t)kk * nc] : 0.0;
        else if (MODE == 1) v = F[(int64_t)(nc + r) + (int64_t)kk * ld];
        else if (MODE == 2) v = F[(int64_t)(nc + kk) + (int64_t)r * ld];
        else v = kk >= r ? (double)Xi[(int64_t)kk + (int64_t)r * nc] : 0.0;
      }
      ra[q] = v;
    }
#pragma unroll
    for (int q = 0; q < BPT; q++) {
      const int idx = tid + 256 * q, kb = idx / NJ, jb = idx % NJ;
      if (MODE == 2) rb[q] = (kb < kn && jb < nj) ? X[(int64_t)ur[q] * nrhs + t.j0 + jb] : 0.0;
      else rb[q] = (kb < kn && jb < nj) ? Bsrc[(int64_t)(kc + kb) * nrhs + t.j0 + jb] : 0.0;
    }
    if (MODE == 2) load_rows(kc + CH);
  };
  if (t.k0 < kend) load(t.k0);
  for (int kc = t.k0; kc < kend; kc += CH) {
    const int kn = min(CH, kend - kc);
    __syncthreads();  // the previous chunk's products are done with the LDS tiles
#pragma unroll
    for (int q = 0; q < 64 * CH / 256; q++) {
      const int idx = tid + 256 * q;
      const int i = rowmajorA ? (idx & 63) : (idx / CH);
      const int k = rowmajorA ? (idx >> 6) : (idx % CH);
      sA[i][k] = ra[q];
    }
#pragma unroll
    for (int q = 0; q < BPT; q++) {
      const int idx = tid + 256 * q;
      sB[idx % NJ][idx / NJ] = rb[q];
    }
    __syncthreads();
    // the next chunk's loads are in flight while this one is multiplied
    if (kc + CH < kend) load(kc + CH);
    // v_mfma_f64_16x16x4f64 over k in steps of 4 (sA and sB are zero beyond
    // kn and nj): the A operand takes the sB rows (columns j), the B operand
    // the sA rows (target rows), as in mfma_tile
    const int lr = lane & 15, lk = lane >> 4;
    for (int m0 = 0; m0 < kn; m0 += 4) {
      const double aa = sA[16 * wv + lr][m0 + lk];
#pragma unroll
      for (int y = 0; y < NT; y++)
        acc[y] = __builtin_amdgcn_mfma_f64_16x16x4f64(sB[16 * y + lr][m0 + lk], aa, acc[y], 0, 0, 0);
    }
  }
  // lane holds target row 16 wv + (lane & 15), columns 16 y + (lane >> 4) + 4 r
  const int i = 16 * wv + (lane & 15), lk = lane >> 4;
  if (t.slot >= 0) {
    double* P = part + (int64_t)t.slot * 4096;
#pragma unroll
    for (int y = 0; y < NT; y++)
#pragma unroll
      for (int r = 0; r < 4; r++) P[i * 64 + 16 * y + lk + 4 * r] = acc[y][r];
    return;
  }
  if (i >= rn) return;
#pragma unroll
  for (int y = 0; y < NT; y++)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int jj = 16 * y + lk + 4 * r;
      if (jj < nj) mv_apply(MODE, S, sn, nc, r0 + i, t.j0 + jj, nrhs, acc[y][r], Wf, Vf, X);
    }
}

// Reduce task: slots k0 .. k0+kn-1 hold the partials of rows r0.. (rn) and
// right-hand sides j0.. of front sn; sum them in slot order and apply.
__global__ __launch_bounds__(1024) void k_mv_red(const SolveTask* __restrict__ tasks, DevSym S,
                                                double* __restrict__ W, double* __restrict__ V,
                                                double* __restrict__ X, const double* __restrict__ part,
                                                int nrhs, int mode) {
  const SolveTask t = tasks[blockIdx.x];
  const int sn = t.sn, nc = S.ncol[sn];
  double* Wf = W + S.woff[sn] * (int64_t)nrhs;
  double* Vf = V + S.woff[sn] * (int64_t)nrhs;
  const int nj = min(64, nrhs - t.j0);
  for (int idx = threadIdx.x; idx < t.rn * nj; idx += blockDim.x) {
    const int i = idx / nj, jj = idx % nj;
    const double* p = part + (int64_t)t.k0 * 4096 + i * 64 + jj;
    // slots summed in order; the loads of four slots are issued together
    double v = 0.0;
    int q = 0;
    for (; q + 3 < t.kn; q += 4) {
      const double a0 = p[(int64_t)q * 4096], a1 = p[(int64_t)(q + 1) * 4096];
      const double a2 = p[(int64_t)(q + 2) * 4096], a3 = p[(int64_t)(q + 3) * 4096];
      v += a0;
      v += a1;
      v += a2;
      v += a3;
    }
    for (; q < t.kn; q++) v += p[(int64_t)q * 4096];
    mv_apply(mode, S, sn, nc, t.r0 + i, t.j0 + jj, nrhs, v, Wf, Vf, X);
  }
}

// Forward-solve front initialisation, task = (front, 64-row chunk r0): pivot
// rows from the global right-hand side R (row-major, perm rows; the Schur
// front takes the boundary rows), update rows zero, plus the children's
// update rows that land in each row.  Work row w adds the rows
// gsrc[gptr[w]] .. gsrc[gptr[w+1]-1] of W in child order.  A gathered row
// whose update product was split (ugrp[row] = g >= 0) is completed here:
// W row minus the sum of its group's partials in slot order (group g:
// first slot, slot count, first work row of the tile at ugroups[3 g ..]).
__global__ __launch_bounds__(1024) void k_fw_init(const SolveTask* __restrict__ tasks, DevSym S,
                                                 const int* __restrict__ gptr, const int* __restrict__ gsrc,
                                                 const int* __restrict__ ugrp, const int* __restrict__ ugroups,
                                                 const double* __restrict__ part, const double* __restrict__ R,
                                                 double* __restrict__ W, int nrhs, int ndof1) {
  const SolveTask t = tasks[blockIdx.x];
  const int sn = t.sn, nc = S.ncol[sn], c0 = S.col0[sn];
  const int64_t wo = S.woff[sn];
  double* Wf = W + wo * nrhs;
  const int base = nc > 0 ? c0 : ndof1;
  for (int idx = threadIdx.x; idx < t.rn * nrhs; idx += blockDim.x) {
    const int i = t.r0 + idx / nrhs, j = idx % nrhs;
    double v = 0.0;
    if (i < nc || nc == 0) v = R[(int64_t)(base + i) * nrhs + j];
    const int e1 = gptr[wo + i + 1];
    for (int e = gptr[wo + i]; e < e1; e++) {
      const int w = gsrc[e];
      double t = W[(int64_t)w * nrhs + j];
      const int g = ugrp ? ugrp[w] : -1;
      if (g >= 0) {
        const int first = ugroups[3 * g], ns = ugroups[3 * g + 1], w0 = ugroups[3 * g + 2];
        const double* p = part + (int64_t)first * 4096 + (w - w0) * 64 + j;
        double sum = 0.0;
        int q = 0;
        for (; q + 3 < ns; q += 4) {
          const double a0 = p[(int64_t)q * 4096], a1 = p[(int64_t)(q + 1) * 4096];
          const double a2 = p[(int64_t)(q + 2) * 4096], a3 = p[(int64_t)(q + 3) * 4096];
          sum += a0;
          sum += a1;
          sum += a2;
          sum += a3;
        }
        for (; q < ns; q++) sum += p[(int64_t)q * 4096];
        t -= sum;
      }
      v += t;
    }
    Wf[(int64_t)i * nrhs + j] = v;
  }
}

// Copy solved rows out to Y, task = (front, 64-row chunk): internal fronts
// give their pivot rows (V, optionally scaled by 1/D), the Schur front its
// rows (W) to the boundary rows of Y.
template <class T>
__global__ __launch_bounds__(256) void k_fw_out(const SolveTask* __restrict__ tasks, DevSym S,
                                                const T* __restrict__ pool,
                                                const double* __restrict__ W, const double* __restrict__ V,
                                                double* __restrict__ Y, int nrhs, int scale, int ndof1) {
  const SolveTask t = tasks[blockIdx.x];
  const int sn = t.sn, nc = S.ncol[sn], c0 = S.col0[sn];
  (void)pool;  // the storage type of the factor sets the rounding of the pivots
  const double* Wf = (nc > 0 ? V : W) + S.woff[sn] * (int64_t)nrhs;
  const int base = nc > 0 ? c0 : ndof1;
  for (int idx = threadIdx.x; idx < t.rn * nrhs; idx += 256) {
    const int i = t.r0 + idx / nrhs, j = idx % nrhs;
    double v = Wf[(int64_t)i * nrhs + j];
    if (scale && nc > 0) v /= (double)(T)S.dpiv[c0 + i];
    Y[(int64_t)(base + i) * nrhs + j] = v;
  }
}

// Backward-solve front initialisation, task = (front, 64-row chunk of the
// pivot rows): pivot rows from Y (zero when Y is null).
__global__ __launch_bounds__(256) void k_bw_init(const SolveTask* __restrict__ tasks, DevSym S,
                                                 const double* __restrict__ Y, double* __restrict__ W, int nrhs) {
  const SolveTask t = tasks[blockIdx.x];
  const int sn = t.sn, c0 = S.col0[sn];
  double* Wf = W + S.woff[sn] * (int64_t)nrhs;
  for (int idx = threadIdx.x; idx < t.rn * nrhs; idx += 256) {
    const int i = t.r0 + idx / nrhs, j = idx % nrhs;
    Wf[(int64_t)i * nrhs + j] = Y ? Y[(int64_t)(c0 + i) * nrhs + j] : 0.0;
  }
}

// ---------------------------------------------------------------- products
// Z = A X for the full symmetric CSR matrix A (perm space), X, Z row-major.
template <bool SKIP>
__global__ __launch_bounds__(256) void k_spmm(int n, int nrhs, const int* __restrict__ rowptr,
                                              const int* __restrict__ colind,
                                              const double* __restrict__ val,
                                              const double* __restrict__ X, int ldx,
                                              double* __restrict__ Z, int ldz) {
  // threads sweep all rows for one block of columns (all of them up to 128,
  // else 64) before the next block, so the rows of X that a sweep gathers
  // stay in cache
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int wcol = nrhs <= 128 ? nrhs : 64;
  const int64_t per = (int64_t)n * wcol;
  const int jb = (int)(g / per);
  const int64_t rem = g - (int64_t)jb * per;
  const int r = (int)(rem / wcol), j = jb * wcol + (int)(rem % wcol);
  if (r >= n || j >= nrhs) return;
  const int kb = rowptr[r], ke = rowptr[r + 1];
  // SKIP: entries that are zero in value (the off-diagonal pattern of a
  // lumped mass) skip their gather of X
  // entry k goes to sum k mod 4 (relative to kb), each sum in entry order;
  // sixteen entries' index, value and gather loads are issued together
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  int k = kb;
  for (; k + 16 <= ke; k += 16) {
    int c[16];
    double v[16], x[16];
#pragma unroll
    for (int u = 0; u < 16; u++) {
      c[u] = colind[k + u];
      v[u] = val[k + u];
    }
#pragma unroll
    for (int u = 0; u < 16; u++) x[u] = (!SKIP || v[u] != 0.0) ? X[(int64_t)c[u] * ldx + j] : 0.0;
#pragma unroll
    for (int u = 0; u < 16; u += 4) {
      if (!SKIP || v[u] != 0.0) s0 += v[u] * x[u];
      if (!SKIP || v[u + 1] != 0.0) s1 += v[u + 1] * x[u + 1];
      if (!SKIP || v[u + 2] != 0.0) s2 += v[u + 2] * x[u + 2];
      if (!SKIP || v[u + 3] != 0.0) s3 += v[u + 3] * x[u + 3];
    }
  }
  for (; k + 4 <= ke; k += 4) {
    const int c0 = colind[k], c1 = colind[k + 1], c2 = colind[k + 2], c3 = colind[k + 3];
    const double v0 = val[k], v1 = val[k + 1], v2 = val[k + 2], v3 = val[k + 3];
    if (!SKIP || v0 != 0.0) s0 += v0 * X[(int64_t)c0 * ldx + j];
    if (!SKIP || v1 != 0.0) s1 += v1 * X[(int64_t)c1 * ldx + j];
    if (!SKIP || v2 != 0.0) s2 += v2 * X[(int64_t)c2 * ldx + j];
    if (!SKIP || v3 != 0.0) s3 += v3 * X[(int64_t)c3 * ldx + j];
  }
  for (; k < ke; k++) {
    const double v = val[k];
    if (!SKIP || v != 0.0) s0 += v * X[(int64_t)colind[k] * ldx + j];
  }
  Z[(int64_t)r * ldz + j] = (s0 + s1) + (s2 + s3);
}

// k_spmm with four columns per thread (nrhs, ldx and ldz multiples of 4):
// each entry's gather is two 16-byte loads and its index and value are
// loaded once for the four columns; every column keeps k_spmm's four sums
// (entry k to sum k mod 4 relative to kb, in entry order), so the results
// are the same bits.
template <bool SKIP>
__global__ __launch_bounds__(256) void k_spmm4(int n, int nrhs, const int* __restrict__ rowptr,
                                               const int* __restrict__ colind, const double* __restrict__ val,
                                               const double* __restrict__ X, int ldx, double* __restrict__ Z, int ldz) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int wcol = nrhs <= 128 ? nrhs : 64, q4 = wcol / 4;
  const int64_t per = (int64_t)n * q4;
  const int jb = (int)(g / per);
  const int64_t rem = g - (int64_t)jb * per;
  const int r = (int)(rem / q4), j0 = jb * wcol + 4 * (int)(rem % q4);
  if (r >= n || j0 >= nrhs) return;
  const int kb = rowptr[r], ke = rowptr[r + 1];
  double s[4][4];
#pragma unroll
  for (int c = 0; c < 4; c++)
#pragma unroll
    for (int m = 0; m < 4; m++) s[c][m] = 0.0;
  auto gather = [&](int c, double2& a, double2& b) {
    const double2* p = reinterpret_cast<const double2*>(X + (int64_t)c * ldx + j0);
    a = p[0];
    b = p[1];
  };
  int k = kb;
  for (; k + 8 <= ke; k += 8) {
    int cc[8];
    double v[8];
    double2 xa[8], xb[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      cc[u] = colind[k + u];
      v[u] = val[k + u];
    }
#pragma unroll
    for (int u = 0; u < 8; u++) {
      if (!SKIP || v[u] != 0.0) gather(cc[u], xa[u], xb[u]);
      else xa[u] = xb[u] = double2{0.0, 0.0};
    }
#pragma unroll
    for (int u = 0; u < 8; u++) {
      if (SKIP && v[u] == 0.0) continue;
      s[0][u & 3] += v[u] * xa[u].x;
      s[1][u & 3] += v[u] * xa[u].y;
      s[2][u & 3] += v[u] * xb[u].x;
      s[3][u & 3] += v[u] * xb[u].y;
    }
  }
  for (; k + 4 <= ke; k += 4) {
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const double vv = val[k + u];
      if (SKIP && vv == 0.0) continue;
      double2 a, b;
      gather(colind[k + u], a, b);
      s[0][u] += vv * a.x;
      s[1][u] += vv * a.y;
      s[2][u] += vv * b.x;
      s[3][u] += vv * b.y;
    }
  }
  for (; k < ke; k++) {
    const double vv = val[k];
    if (SKIP && vv == 0.0) continue;
    double2 a, b;
    gather(colind[k], a, b);
    s[0][0] += vv * a.x;
    s[1][0] += vv * a.y;
    s[2][0] += vv * b.x;
    s[3][0] += vv * b.y;
  }
  double* z = Z + (int64_t)r * ldz + j0;
#pragma unroll
  for (int c = 0; c < 4; c++) z[c] = (s[c][0] + s[c][1]) + (s[c][2] + s[c][3]);
}

// Partial C = X^T Y: blockIdx.x = chunk of crows rows, blockIdx.y = 64
// columns of X (a), blockIdx.z = 64 columns of Y (b); lane -> a, wave -> b
// (b = 16*wave + u).  part[chunk][a][b] (na x nb row-major per chunk).
#define XTY_ROWS 128
__global__ __launch_bounds__(256) void k_xty_part(int n, int crows, int na, int nb, const double* __restrict__ X,
                                                  int ldx, const double* __restrict__ Y, int ldy,
                                                  double* __restrict__ part) {
  __shared__ double sX[32][64];
  __shared__ double sY[32][64];
  const int r0 = blockIdx.x * crows, a0 = blockIdx.y * 64, b0 = blockIdx.z * 64;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int nbt = min(64, nb - b0), nat = min(64, na - a0);
  const bool narrow = nb <= 16;
  double acc[16];
#pragma unroll
  for (int q = 0; q < 16; q++) acc[q] = 0.0;
  const int rend = min(n, r0 + crows);
  for (int rc = r0; rc < rend; rc += 32) {
    double rx[8], ry[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int idx = tid + 256 * q, rr = idx >> 6, c = idx & 63, r = rc + rr;
      const bool ok = r < rend;
      rx[q] = (ok && c < nat) ? X[(int64_t)r * ldx + a0 + c] : 0.0;
      ry[q] = (ok && c < nbt) ? Y[(int64_t)r * ldy + b0 + c] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int idx = tid + 256 * q;
      sX[idx >> 6][idx & 63] = rx[q];
      sY[idx >> 6][idx & 63] = ry[q];
    }
    __syncthreads();
    // sY is zero beyond nbt: same unguarded loop in every thread.  Up to 16
    // columns of Y (the Lanczos block) the four waves split the rows of the
    // chunk instead of the columns, so none of them multiplies zeros.
    if (narrow) {
#pragma unroll 2
      for (int rr = wv; rr < 32; rr += 4) {
        const double xv = sX[rr][lane];
#pragma unroll
        for (int q = 0; q < 16; q++) acc[q] += xv * sY[rr][q];
      }
    } else {
#pragma unroll 4
      for (int rr = 0; rr < 32; rr++) {
        const double xv = sX[rr][lane];
#pragma unroll
        for (int q = 0; q < 16; q++) acc[q] += xv * sY[rr][wv * 16 + q];
      }
    }
  }
  double* P = part + (int64_t)blockIdx.x * na * nb;
  if (narrow) {
    // the waves' partial sums, added in wave order by wave 0: waves 1 and 2
    // leave theirs in rows 0-15 and 16-31 of sX, wave 3 in rows 0-15 of sY
    __syncthreads();
    if (wv == 1 || wv == 2) {
#pragma unroll
      for (int q = 0; q < 16; q++) sX[(wv - 1) * 16 + q][lane] = acc[q];
    } else if (wv == 3) {
#pragma unroll
      for (int q = 0; q < 16; q++) sY[q][lane] = acc[q];
    }
    __syncthreads();
    if (wv == 0 && lane < nat) {
#pragma unroll
      for (int q = 0; q < 16; q++) {
        const double v = ((acc[q] + sX[q][lane]) + sX[16 + q][lane]) + sY[q][lane];
        if (q < nbt) P[(int64_t)(a0 + lane) * nb + b0 + q] = v;
      }
    }
    return;
  }
  if (lane < nat) {
#pragma unroll
    for (int q = 0; q < 16; q++) {
      int b = wv * 16 + q;
      if (b < nbt) P[(int64_t)(a0 + lane) * nb + b0 + b] = acc[q];
    }
  }
}

// Partial C = X^T Y for wide products (both sides >= 64 columns), on FP64
// MFMA: the workgroup's 64 x 64 output tile accumulates over its chunk of
// crows rows in steps of 32; X and Y rows are read along their columns
// (coalesced) and stored transposed in LDS (sA[a][r], sB[b][r]) for
// mfma_tile.  Same grid and partial layout as k_xty_part.
__global__ __launch_bounds__(256) void k_xty_mfma(int n, int crows, int na, int nb, const double* __restrict__ X,
                                                  int ldx, const double* __restrict__ Y, int ldy,
                                                  double* __restrict__ part) {
  __shared__ double sA[64][33];
  __shared__ double sB[64][33];
  const int r0 = blockIdx.x * crows, a0 = blockIdx.y * 64, b0 = blockIdx.z * 64;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int nbt = min(64, nb - b0), nat = min(64, na - a0);
  const int wr = (wv >> 1) * 32, wc = (wv & 1) * 32;
  f64x4 acc[2][2];
  zero_acc(acc);
  const int rend = min(n, r0 + crows);
  double rx[8], ry[8];
  auto load = [&](int rc) {
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int idx = tid + 256 * q, rr = idx >> 6, c = idx & 63, r = rc + rr;
      const bool ok = r < rend;
      rx[q] = (ok && c < nat) ? X[(int64_t)r * ldx + a0 + c] : 0.0;
      ry[q] = (ok && c < nbt) ? Y[(int64_t)r * ldy + b0 + c] : 0.0;
    }
  };
  if (r0 < rend) load(r0);
  for (int rc = r0; rc < rend; rc += 32) {
    if (rc > r0) __syncthreads();  // the previous rows' products are done with the tiles
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int idx = tid + 256 * q;
      sA[idx & 63][idx >> 6] = rx[q];
      sB[idx & 63][idx >> 6] = ry[q];
    }
    __syncthreads();
    if (rc + 32 < rend) load(rc + 32);
    mfma_tile<double, 33>(sA, sB, 32, acc, lane, wr, wc);
  }
  double* P = part + (int64_t)blockIdx.x * na * nb;
  const int lr = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int a = wr + 16 * x + lr, b = wc + 16 * y + lk + 4 * r;
        if (a < nat && b < nbt) P[(int64_t)(a0 + a) * nb + b0 + b] = acc[x][y][r];
      }
}

// One wave per output element: lane l sums chunks l, l+64, ... then a fixed
// butterfly combines the lanes (same order on every run).
// C = the sum of the chunk partials; acc (optional) = C, or acc + C with
// acc_add.
__global__ __launch_bounds__(256) void k_xty_sum(int nchunk, int nab, const double* __restrict__ part,
                                                 double* __restrict__ C, double* __restrict__ acc, int acc_add) {
  const int e = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (e >= nab) return;
  double s = 0.0;
  for (int c = lane; c < nchunk; c += 64) s += part[(int64_t)c * nab + e];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (lane == 0) {
    C[e] = s;
    if (acc) acc[e] = acc_add ? acc[e] + s : s;
  }
}

// Sign convention of the component modes: for column i of Phi (rows in
// permuted order, ld columns), the component of largest magnitude over the
// internal equations taken in equation order (q -> row p1[q]; the first
// maximum wins) is made positive.  Pass 1: workgroup g scans equations
// g*span .. with lane -> column (64 columns at a time) and wave -> every
// fourth equation, and leaves per column its (|v|, v, q) best; pass 2: one
// wave per column combines the workgroups' bests (larger |v|, then smaller
// q), so the choice does not depend on scheduling.
__global__ __launch_bounds__(256) void k_mode_signs_part(int n1, int span, const int* __restrict__ p1,
                                                         const double* __restrict__ Phi, int ld, int nev,
                                                         double* __restrict__ best) {
  __shared__ double sa[4][64], sv[4][64];
  __shared__ int sq[4][64];
  const int g = blockIdx.x, lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int q0 = g * span, q1 = min(n1, q0 + span);
  for (int j0 = 0; j0 < nev; j0 += 64) {
    const int i = j0 + lane;
    double ba = -1.0, bv = 0.0;
    int bq = INT_MAX;
    if (i < nev)
      for (int q = q0 + wv; q < q1; q += 4) {  // q ascends within a lane
        const double v = Phi[(int64_t)p1[q] * ld + i], a = fabs(v);
        if (a > ba) { ba = a; bv = v; bq = q; }
      }
    sa[wv][lane] = ba;
    sv[wv][lane] = bv;
    sq[wv][lane] = bq;
    __syncthreads();
    if (wv == 0 && i < nev) {
      for (int w = 1; w < 4; w++) {
        const double a2 = sa[w][lane];
        const int q2 = sq[w][lane];
        if (a2 > ba || (a2 == ba && q2 < bq)) { ba = a2; bv = sv[w][lane]; bq = q2; }
      }
      double* o = best + ((int64_t)g * nev + i) * 3;
      o[0] = ba;
      o[1] = bv;
      o[2] = (double)bq;
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(64) void k_mode_signs_fin(int ngroups, int nev, const double* __restrict__ best,
                                                       double* __restrict__ sign) {
  const int i = blockIdx.x, lane = threadIdx.x;
  double ba = -1.0, bv = 0.0, bq = 1e300;
  for (int g = lane; g < ngroups; g += 64) {
    const double* o = best + ((int64_t)g * nev + i) * 3;
    if (o[0] > ba || (o[0] == ba && o[2] < bq)) { ba = o[0]; bv = o[1]; bq = o[2]; }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double a2 = __shfl_xor(ba, off), v2 = __shfl_xor(bv, off), q2 = __shfl_xor(bq, off);
    if (a2 > ba || (a2 == ba && q2 < bq)) { ba = a2; bv = v2; bq = q2; }
  }
  if (lane == 0) sign[i] = bv < 0.0 ? -1.0 : 1.0;
}

__global__ void k_scale_cols(int64_t n, int nc, const double* __restrict__ sign, double* __restrict__ X) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g < n * nc) X[g] *= sign[g % nc];
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

// Y = alpha Y + beta X S for S with at most 16 columns, on FP64 MFMA: a
// workgroup takes 64 rows of X, staged through LDS in 32-column chunks read
// in row order (coalesced), the next chunk's loads in flight during the
// current chunk's products; wave w accumulates rows 16 w .. 16 w + 15.
// blockIdx.y = 1 takes the second pair (X2, Y2) with the same S.  S rows
// are lds apart (the nb columns may be a slice of a wider S).  Y may be X
// (in place: a workgroup has read its rows of X before it writes them).
__global__ __launch_bounds__(256) void k_gemm_rows16(int n, int na, int nb, const double* X1, int ldx1,
                                                     const double* __restrict__ Sm, int lds, double* Y1, int ldy1,
                                                     double alpha, double beta, const double* X2, int ldx2,
                                                     double* Y2, int ldy2) {
  const double* X = blockIdx.y ? X2 : X1;
  double* Y = blockIdx.y ? Y2 : Y1;
  const int ldx = blockIdx.y ? ldx2 : ldx1, ldy = blockIdx.y ? ldy2 : ldy1;
  __shared__ double sX[64][33];
  __shared__ double sS[16][33];  // S chunk transposed: sS[j][k]
  const int r0 = blockIdx.x * 64;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, lr = lane & 15, lk = lane >> 4;
  f64x4 acc = f64x4{0.0, 0.0, 0.0, 0.0};
  double rx[8], rs[2];
  auto load = [&](int kc) {
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int idx = tid + 256 * q, k = idx & 31, r = r0 + (idx >> 5);
      rx[q] = (r < n && kc + k < na) ? X[(int64_t)r * ldx + kc + k] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < 2; q++) {
      const int idx = tid + 256 * q, k = idx >> 4, j = idx & 15;
      rs[q] = (kc + k < na && j < nb) ? Sm[(int64_t)(kc + k) * lds + j] : 0.0;
    }
  };
  if (na > 0) load(0);
  for (int kc = 0; kc < na; kc += 32) {
    __syncthreads();  // the previous chunk's products are done with the LDS tiles
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int idx = tid + 256 * q;
      sX[idx >> 5][idx & 31] = rx[q];
    }
#pragma unroll
    for (int q = 0; q < 2; q++) {
      const int idx = tid + 256 * q;
      sS[idx & 15][idx >> 4] = rs[q];
    }
    __syncthreads();
    if (kc + 32 < na) load(kc + 32);
#pragma unroll
    for (int m0 = 0; m0 < 32; m0 += 4)
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(sS[lr][m0 + lk], sX[16 * wv + lr][m0 + lk], acc, 0, 0, 0);
  }
  // acc[q]: row r0 + 16 wv + lr, column lk + 4 q
  const int r = r0 + 16 * wv + lr;
  if (r >= n) return;
  double* y = Y + (int64_t)r * ldy;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int j = lk + 4 * q;
    if (j < nb) y[j] = ((alpha == 0.0) ? 0.0 : alpha * y[j]) + beta * acc[q];
  }
}

// X[(ndof1 + b) * nrhs + b] = 1, all other entries 0 (the [B; I] identity rows)
__global__ void k_identity_rows(double* __restrict__ X, int64_t nrows, int nrhs, int ndof1) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nrows * nrhs) return;
  int64_t r = g / nrhs;
  int j = (int)(g % nrhs);
  X[g] = (r >= ndof1 && r - ndof1 == j) ? 1.0 : 0.0;
}

// Element mass EMASS of lhsINADD (inadd.f90:258-265): the x-translation mass
// u_x^T M_e u_x of each element's packed node-pair blocks (diagonal blocks
// once, off-diagonal blocks twice; entry (0,0) of a block couples the two
// nodes' x translations).  One lane per element of one type list.
__global__ void k_emass(int n, const int* __restrict__ elist, int nen, int d, const int64_t* __restrict__ eoff,
                        const double* __restrict__ Mp, double* __restrict__ emass) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int e = elist[t];
  const double* m = Mp + eoff[e];
  double s = 0.0;
  for (int a = 0; a < nen; a++)
    for (int b = 0; b <= a; b++) s += (a == b ? 1.0 : 2.0) * m[(int64_t)(a * (a + 1) / 2 + b) * d * d];
  emass[e] = s;
}

// out[0] = sum of x[0..n-1] by one workgroup in a fixed order (strided lane
// sums, then a tree), so the total is the same bits on every run
__global__ __launch_bounds__(256) void k_sum_fixed(int n, const double* __restrict__ x, double* __restrict__ out) {
  __shared__ double red[256];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) s += x[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

// ---------------------------------------------------------------- Lanczos block orthonormalisation
// Cholesky factor of the symmetrised n x n Gram matrix G = R^T R (R upper,
// row-major) and R^{-1}, by one workgroup (n <= 64): at step j every thread
// forms the pivot (same value in each), thread i > j the entry R[j][i]; the
// inverse is solved column by column, thread t owning column t.  A
// non-positive pivot sets *bad (the block has broken down).
__global__ __launch_bounds__(64) void k_chol_inv(int n, const double* __restrict__ G, double* __restrict__ R,
                                                 double* __restrict__ Ri, int* __restrict__ bad) {
  __shared__ double A[64][65], U[64][65], V[64][65];
  const int t = threadIdx.x;
  if (t < n)
    for (int i = 0; i < n; i++) {
      A[i][t] = i == t ? G[i * n + t] : 0.5 * (G[i * n + t] + G[t * n + i]);
      U[i][t] = 0.0;
      V[i][t] = 0.0;
    }
  __syncthreads();
  for (int j = 0; j < n; j++) {
    double s = A[j][j];
    for (int k = 0; k < j; k++) s -= U[k][j] * U[k][j];
    if (!(s > 0.0)) {
      if (t == 0) *bad = 1;
      s = 1.0;
    }
    const double d = sqrt(s);
    if (t == j) U[j][j] = d;
    if (t > j && t < n) {
      double v = A[j][t];
      for (int k = 0; k < j; k++) v -= U[k][j] * U[k][t];
      U[j][t] = v / d;
    }
    __syncthreads();
  }
  if (t < n) {
    V[t][t] = 1.0 / U[t][t];
    for (int i = t - 1; i >= 0; i--) {
      double v = 0.0;
      for (int k = i + 1; k <= t; k++) v += U[i][k] * V[k][t];
      V[i][t] = -v / U[i][i];
    }
    for (int i = 0; i < n; i++) {
      R[i * n + t] = U[i][t];
      Ri[i * n + t] = V[i][t];
    }
  }
}

// k_chol_inv for n <= 16 in one wave with the rows in registers (lane i <
// 16 holds row i): the LDL^T readlane sweep of k_diag gives the unit lower
// L1, D and X1 = L1^-1, then G = R^T R with R = D^1/2 L1^T and
// R^-1 = X1^T D^-1/2.  Rows and columns beyond n are identity.
__global__ __launch_bounds__(64) void k_chol16(int n, const double* __restrict__ G, double* __restrict__ R,
                                               double* __restrict__ Ri, int* __restrict__ bad) {
  const int lane = threadIdx.x;
  double a[16], x[16];
#pragma unroll
  for (int m = 0; m < 16; m++) {
    double v = lane == m ? 1.0 : 0.0;
    if (lane < n && m < n) v = lane == m ? G[lane * n + m] : 0.5 * (G[lane * n + m] + G[m * n + lane]);
    a[m] = m <= lane ? v : 0.0;
    x[m] = lane == m ? 1.0 : 0.0;
  }
  bool nonpos = false;
#pragma unroll
  for (int j = 0; j < 16; j++) {
    double dj = lane_bcast(a[j], j);
    if (!(dj > 0.0)) {
      nonpos = true;
      dj = 1.0;
    }
    const bool below = lane > j && lane < 16;
    const double l = below ? a[j] / dj : 0.0;
    if (below) a[j] = l;
    if (lane == j) a[j] = dj;
    const double ldj = l * dj;
#pragma unroll
    for (int m = j + 1; m < 16; m++) {
      const double lm = lane_bcast(l, m);  // L1[m][j]
      if (m <= lane) a[m] -= ldj * lm;
    }
#pragma unroll
    for (int m = 0; m <= j; m++) x[m] -= l * lane_bcast(x[m], j);  // X1[j][m]
  }
  if (nonpos && lane == 0) *bad = 1;
  // lane i: D_i = a[i]; L1[i][m] = a[m] (m < i), X1[i][m] = x[m] (m <= i)
  double sq[16];
#pragma unroll
  for (int m = 0; m < 16; m++) sq[m] = sqrt(lane_bcast(a[m], m));
  if (lane >= n) return;
  double si = 1.0;
#pragma unroll
  for (int m = 0; m < 16; m++) si = m == lane ? sq[m] : si;
  const double ri = 1.0 / si;
#pragma unroll
  for (int m = 0; m < 16; m++) {
    if (m >= n) break;
    const double sm = sq[m];
    // R[m][i] = sqrt(D_m) L1[i][m] (m < i), sqrt(D_i) on the diagonal, 0 above
    R[m * n + lane] = m < lane ? sm * a[m] : (m == lane ? si : 0.0);
    // R^-1[m][i] = X1[i][m] / sqrt(D_i) (m <= i)
    Ri[m * n + lane] = m <= lane ? x[m] * ri : 0.0;
  }
}

// ---------------------------------------------------------------- mixed precision
// Symmetric diagonal equilibration of the stiffness matrix for the FP32
// factor: s[r] = 1/sqrt(K[r][r]) for every perm row r of the CSR matrix
// (1 for a non-positive diagonal, which the factorisation then reports).
__global__ void k_diag_scale(int n, const int* __restrict__ rowptr, const int* __restrict__ colind,
                             const double* __restrict__ val, double* __restrict__ s) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double d = 0.0;
  for (int k = rowptr[r]; k < rowptr[r + 1]; k++)
    if (colind[k] == r) d = val[k];
  s[r] = d != 0.0 ? 1.0 / sqrt(fabs(d)) : 1.0;
}

// The equilibrated stiffness s_r K_rc s_c of every lower CSR entry into its
// FP32 front position (the FP64 factor takes the entries from k_assemble).
__global__ void k_front_scatter32(int n, const int* __restrict__ rowptr, const int* __restrict__ colind,
                                  const double* __restrict__ val, const double* __restrict__ s,
                                  const int64_t* __restrict__ fmap, float* __restrict__ pool) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  for (int k = rowptr[r]; k < rowptr[r + 1]; k++) {
    const int64_t f = fmap[k];
    if (f >= 0) pool[f] = (float)(s[r] * val[k] * s[colind[k]]);
  }
}

// X[r][j] *= rs[r] (rows 0..nrows-1) and, with cs, *= cs[j] (row-major, nrhs columns)
__global__ void k_scale_rc(int64_t nrows, int nrhs, const double* __restrict__ rs, const double* __restrict__ cs,
                           double* __restrict__ X) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nrows * nrhs) return;
  const int64_t r = g / nrhs;
  const int j = (int)(g % nrhs);
  double v = X[g] * rs[r];
  if (cs) v *= cs[j];
  X[g] = v;
}

// y[i] = 1 / x[i]
__global__ void k_recip(int n, const double* __restrict__ x, double* __restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = 1.0 / x[i];
}

// Y[g] = a * X[g] + b * Y[g] over the first n entries
// Y = a X + b Y (Y not read when b = 0; X and Y may be the same array)
__global__ void k_axpby(int64_t n, double a, const double* X, double b, double* Y) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g < n) Y[g] = b == 0.0 ? a * X[g] : a * X[g] + b * Y[g];
}

// part[blk] = (sum x^2, sum y^2) over the block's slice of the first n
// entries; k_sumsq_fin adds the blocks in order (deterministic norms)
__global__ __launch_bounds__(256) void k_sumsq2(int64_t n, const double* __restrict__ x,
                                                const double* __restrict__ y, double* __restrict__ part) {
  __shared__ double r0[256], r1[256];
  double a = 0.0, b = 0.0;
  const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
  const int64_t lo = (int64_t)blockIdx.x * chunk, hi = min(n, lo + chunk);
  for (int64_t i = lo + threadIdx.x; i < hi; i += 256) {
    a += x[i] * x[i];
    b += y[i] * y[i];
  }
  r0[threadIdx.x] = a;
  r1[threadIdx.x] = b;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      r0[threadIdx.x] += r0[threadIdx.x + w];
      r1[threadIdx.x] += r1[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = r0[0];
    part[2 * blockIdx.x + 1] = r1[0];
  }
}

// Column dot products of up to three pairs (A_k, B_k) of row-major n x nr
// arrays: block b sums rows [b COLDOT_ROWS, (b+1) COLDOT_ROWS) of each
// column in row order into part[(b * npair + k) * nr + c]; k_coldot_sum
// then adds the blocks in block order (a fixed order: same bits every run).
__global__ __launch_bounds__(256) void k_coldot_part(int64_t n, int nr, int npair, const double* __restrict__ A0,
                                                     const double* __restrict__ B0, const double* __restrict__ A1,
                                                     const double* __restrict__ B1, const double* __restrict__ A2,
                                                     const double* __restrict__ B2, double* __restrict__ part) {
  const int64_t r0 = (int64_t)blockIdx.x * COLDOT_ROWS, r1 = min(n, r0 + COLDOT_ROWS);
  for (int c = threadIdx.x; c < nr; c += 256) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int64_t r = r0; r < r1; r++) {
      const int64_t o = r * nr + c;
      s0 += A0[o] * B0[o];
      if (npair > 1) s1 += A1[o] * B1[o];
      if (npair > 2) s2 += A2[o] * B2[o];
    }
    double* P = part + (int64_t)blockIdx.x * npair * nr + c;
    P[0] = s0;
    if (npair > 1) P[nr] = s1;
    if (npair > 2) P[2 * nr] = s2;
  }
}

__global__ __launch_bounds__(256) void k_coldot_sum(int nb, int nr, int npair, const double* __restrict__ part,
                                                    double* __restrict__ out) {
  const int idx = blockIdx.x * 256 + threadIdx.x;
  if (idx >= npair * nr) return;
  double s = 0.0;
  for (int b = 0; b < nb; b++) s += part[(int64_t)b * npair * nr + idx];
  out[idx] = s;
}

// Y[r][c] = a[c] X[r][c] + b[c] Y[r][c] over n rows of row-major n x nr
// arrays (a or b null: coefficient 1).
__global__ __launch_bounds__(256) void k_col_axpby(int64_t n, int nr, const double* __restrict__ a,
                                                   const double* __restrict__ X, const double* __restrict__ b,
                                                   double* __restrict__ Y) {
  const int64_t total = n * nr;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < total; i += (int64_t)gridDim.x * 256) {
    const int c = (int)(i % nr);
    const double ya = b ? b[c] * Y[i] : Y[i];
    Y[i] = (a ? a[c] * X[i] : X[i]) + ya;
  }
}

// ---------------------------------------------------------------- launchers
static inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

void launch_hex8(hipStream_t st, int nhex, const int* elist, const int* eptr, const int* enodes,
                 const double* xyz, const double* eprop, const int64_t* eoff, double* Kp, double* Mp,
                 int* ebad) {
  if (nhex > 0) k_hex8<<<nhex, 64, 0, st>>>(nhex, elist, eptr, enodes, xyz, eprop, eoff, Kp, Mp, ebad);
}
void launch_assemble(hipStream_t st, int64_t nnz, const int* cptr, const int* cidx, const double* ccoef,
                     const double* Kp, const double* Mp, double* Kv, double* Mv, const int64_t* fmap, double* pool) {
  if (nnz <= 0) return;
  if (ccoef) k_assemble<true><<<nblk(nnz, 256), 256, 0, st>>>(nnz, cptr, cidx, ccoef, Kp, Mp, Kv, Mv, fmap, pool);
  else k_assemble<false><<<nblk(nnz, 256), 256, 0, st>>>(nnz, cptr, cidx, ccoef, Kp, Mp, Kv, Mv, fmap, pool);
}
void launch_diag_extract(hipStream_t st, int n, const int* rowptr, const int* colind, const double* val,
                         const double* rs, double* adiag, double* amax) {
  if (n > 0)
    k_diag_extract<<<nblk(n, 256), 256, 0, st>>>(n, rowptr, colind, val, rs, adiag,
                                                 reinterpret_cast<unsigned long long*>(amax));
}
void launch_pivot_stats(hipStream_t st, int n, const double* dpiv, const int* sing, int* out) {
  if (n > 0) k_pivot_stats<<<std::min(1024u, nblk(n, 256)), 256, 0, st>>>(n, dpiv, sing, out);
}
void launch_fill(hipStream_t st, double* p, int64_t n, double v) {
  if (n > 0) k_fill<<<(unsigned)std::min<int64_t>(nblk(n, 256), 16384), 256, 0, st>>>(p, n, v);
}
#define FGPU_FACTOR_LAUNCHERS(T)                                                                           \
  void launch_extend_add(hipStream_t st, const FTask* t, int n, const DevSym& S, const int* xrange,        \
                         T* pool) {                                                                        \
    if (n > 0) k_extend_add<T><<<n, 256, 0, st>>>(t, S, xrange, pool);                                       \
  }                                                                                                        \
  void launch_diag(hipStream_t st, const FTask* t, int n, const DevSym& S, T* pool, const int64_t* ioff,   \
                   T* ipool, int* piv) {                                                                   \
    if (n > 0) k_diag<T><<<n, 256, 0, st>>>(t, S, pool, ioff, ipool, piv);                                   \
  }                                                                                                        \
  void launch_diag_trsm(hipStream_t st, const FTask* t, int n, const DevSym& S, T* pool,                   \
                        const int64_t* ioff, T* ipool, int* piv) {                                          \
    if (n > 0) k_diag_trsm<T><<<n, 256, 0, st>>>(t, S, pool, ioff, ipool, piv);                             \
  }                                                                                                        \
  void launch_trsm(hipStream_t st, const FTask* t, int n, const DevSym& S, T* pool, const int64_t* ioff,   \
                   const T* ipool) {                                                                       \
    if (n > 0) k_trsm<T><<<n, 256, 0, st>>>(t, S, pool, ioff, ipool);                                        \
  }                                                                                                        \
  void launch_update(hipStream_t st, const FTask* t, int n, const DevSym& S, T* pool) {                    \
    if (n > 0) k_update<T><<<n, 256, 0, st>>>(t, S, pool);                                                   \
  }                                                                                                        \
  void launch_inv_off(hipStream_t st, const FTask* t, int n, const DevSym& S, const T* pool,               \
                      const int64_t* ioff, T* ipool) {                                                     \
    if (n > 0) k_inv_off<T><<<n, 256, 0, st>>>(t, S, pool, ioff, ipool);                                     \
  }                                                                                                        \
  void launch_inv_part(hipStream_t st, const FTask* t, int n, const DevSym& S, const T* pool,              \
                       const int64_t* ioff, const T* ipool, T* ipart) {                                    \
    if (n > 0) k_inv_part<T><<<n, 256, 0, st>>>(t, S, pool, ioff, ipool, ipart);                             \
  }                                                                                                        \
  void launch_inv_fin(hipStream_t st, const FTask* t, int n, const DevSym& S, const int64_t* ioff,         \
                      T* ipool, const T* ipart) {                                                          \
    if (n > 0) k_inv_fin<T><<<n, 1024, 0, st>>>(t, S, ioff, ipool, ipart);                                    \
  }                                                                                                        \
  void launch_mv(hipStream_t st, const SolveTask* t, int n, const DevSym& S, const T* pool,                \
                 const int64_t* ioff, const T* ipool, double* W, double* V, double* X, double* part,       \
                 int nrhs, int mode) {                                                                     \
    if (n <= 0) return;                                                                                    \
    switch (mode * 2 + (nrhs <= 16 ? 0 : 1)) {                                                             \
      case 0: k_mv<0, 16, T><<<n, 256, 0, st>>>(t, S, pool, ioff, ipool, W, V, X, part, nrhs); break;      \
      case 1: k_mv<0, 64, T><<<n, 256, 0, st>>>(t, S, pool, ioff, ipool, W, V, X, part, nrhs); break;      \
      case 2: k_mv<1, 16, T><<<n, 256, 0, st>>>(t, S, pool, ioff, ipool, W, V, X, part, nrhs); break;      \
      case 3: k_mv<1, 64, T><<<n, 256, 0, st>>>(t, S, pool, ioff, ipool, W, V, X, part, nrhs); break;      \
      case 4: k_mv<2, 16, T><<<n, 256, 0, st>>>(t, S, pool, ioff, ipool, W, V, X, part, nrhs); break;      \
      case 5: k_mv<2, 64, T><<<n, 256, 0, st>>>(t, S, pool, ioff, ipool, W, V, X, part, nrhs); break;      \
      case 6: k_mv<3, 16, T><<<n, 256, 0, st>>>(t, S, pool, ioff, ipool, W, V, X, part, nrhs); break;      \
      default: k_mv<3, 64, T><<<n, 256, 0, st>>>(t, S, pool, ioff, ipool, W, V, X, part, nrhs); break;     \
    }                                                                                                      \
  }                                                                                                        \
  void launch_fw_out(hipStream_t st, const SolveTask* t, int n, const DevSym& S, const T* pool,            \
                     const double* W, const double* V, double* Y, int nrhs, int scale, int ndof1) {        \
    if (n > 0) k_fw_out<T><<<n, 256, 0, st>>>(t, S, pool, W, V, Y, nrhs, scale, ndof1);                      \
  }
FGPU_FACTOR_LAUNCHERS(double)
FGPU_FACTOR_LAUNCHERS(float)
#undef FGPU_FACTOR_LAUNCHERS
void launch_mv_red(hipStream_t st, const SolveTask* t, int n, const DevSym& S, double* W, double* V, double* X,
                   const double* part, int nrhs, int mode) {
  if (n > 0) k_mv_red<<<n, nrhs <= 16 ? 64 * nrhs : 1024, 0, st>>>(t, S, W, V, X, part, nrhs, mode);
}
void launch_fw_init(hipStream_t st, const SolveTask* t, int n, const DevSym& S, const int* gptr, const int* gsrc,
                    const int* ugrp, const int* ugroups, const double* part, const double* R, double* W, int nrhs,
                    int ndof1) {
  if (n > 0)
    k_fw_init<<<n, nrhs <= 16 ? 64 * nrhs : 1024, 0, st>>>(t, S, gptr, gsrc, ugrp, ugroups, part, R, W, nrhs, ndof1);
}
void launch_bw_init(hipStream_t st, const SolveTask* t, int n, const DevSym& S, const double* Y, double* W,
                    int nrhs) {
  if (n > 0) k_bw_init<<<n, 256, 0, st>>>(t, S, Y, W, nrhs);
}
void launch_spmm(hipStream_t st, int n, int nrhs, const int* rowptr, const int* colind, const double* val,
                 const double* X, int ldx, double* Z, int ldz, bool skip_zeros) {
  if (n > 0 && nrhs > 0) {
    const int wcol = nrhs <= 128 ? nrhs : 64;
    const int64_t threads = (int64_t)n * wcol * ((nrhs + wcol - 1) / wcol);
    if (nrhs % 4 == 0 && ldx % 4 == 0 && ldz % 4 == 0 && (reinterpret_cast<uintptr_t>(X) & 31) == 0) {
      const int64_t t4 = threads / 4;
      if (skip_zeros) k_spmm4<true><<<nblk(t4, 256), 256, 0, st>>>(n, nrhs, rowptr, colind, val, X, ldx, Z, ldz);
      else k_spmm4<false><<<nblk(t4, 256), 256, 0, st>>>(n, nrhs, rowptr, colind, val, X, ldx, Z, ldz);
    } else if (skip_zeros) {
      k_spmm<true><<<nblk(threads, 256), 256, 0, st>>>(n, nrhs, rowptr, colind, val, X, ldx, Z, ldz);
    } else {
      k_spmm<false><<<nblk(threads, 256), 256, 0, st>>>(n, nrhs, rowptr, colind, val, X, ldx, Z, ldz);
    }
  }
}
int xty_chunks(int n) { return (n + XTY_ROWS - 1) / XTY_ROWS; }
void launch_xty(hipStream_t st, int n, int na, int nb, const double* X, int ldx, const double* Y, int ldy,
                double* part, double* C, double* acc, bool acc_add) {
  if (na <= 0 || nb <= 0) return;
  // wide products (many 64x64 output tiles) take longer row chunks: fewer
  // partial matrices to write and add, still thousands of workgroups
  const int tiles = ((na + 63) / 64) * ((nb + 63) / 64);
  const int crows = tiles >= 16 ? 8 * XTY_ROWS : XTY_ROWS;
  const int nch = (n + crows - 1) / crows;
  if (nch == 0) {
    launch_fill(st, C, (int64_t)na * nb, 0.0);
    if (acc && !acc_add) launch_fill(st, acc, (int64_t)na * nb, 0.0);
    return;
  }
  dim3 g(nch, (na + 63) / 64, (nb + 63) / 64);
  if (tiles >= 16) k_xty_mfma<<<g, 256, 0, st>>>(n, crows, na, nb, X, ldx, Y, ldy, part);
  else k_xty_part<<<g, 256, 0, st>>>(n, crows, na, nb, X, ldx, Y, ldy, part);
  k_xty_sum<<<nblk((int64_t)na * nb, 4), 256, 0, st>>>(nch, na * nb, part, C, acc, acc_add ? 1 : 0);
}
void launch_gemm_small2(hipStream_t st, int n, int na, int nb, const double* X1, int ldx1, double* Y1, int ldy1,
                        const double* X2, int ldx2, double* Y2, int ldy2, const double* Sm, double alpha,
                        double beta) {
  if (n <= 0 || nb <= 0) return;
  if (nb > 16) {
    launch_gemm_small(st, n, na, nb, X1, ldx1, Sm, Y1, ldy1, alpha, beta);
    launch_gemm_small(st, n, na, nb, X2, ldx2, Sm, Y2, ldy2, alpha, beta);
    return;
  }
  k_gemm_rows16<<<dim3((unsigned)((n + 63) / 64), 2), 256, 0, st>>>(n, na, nb, X1, ldx1, Sm, nb, Y1, ldy1, alpha,
                                                                     beta, X2, ldx2, Y2, ldy2);
}
void launch_gemm_small(hipStream_t st, int n, int na, int nb, const double* X, int ldx, const double* Sm,
                       double* Y, int ldy, double alpha, double beta) {
  if (n <= 0 || nb <= 0) return;
  if (nb <= 64) {  // MFMA over 64-row tiles, 16 columns of S per launch
    for (int c0 = 0; c0 < nb; c0 += 16)
      k_gemm_rows16<<<(unsigned)((n + 63) / 64), 256, 0, st>>>(n, na, std::min(16, nb - c0), X, ldx, Sm + c0, nb,
                                                               Y + c0, ldy, alpha, beta, nullptr, 0, nullptr, 0);
    return;
  }
  // S is staged in LDS in row chunks of at most 64 KiB; chunk k adds
  // beta X_k S_k to the result of the chunks before it (fixed order)
  const int rows = std::max(1, (int)(65536 / (sizeof(double) * nb)));
  for (int a0 = 0; a0 < na || a0 == 0; a0 += rows) {
    const int ka = std::min(rows, na - a0);
    k_gemm_small<<<nblk((int64_t)n * nb, 256), 256, sizeof(double) * std::max(ka, 1) * nb, st>>>(
        n, ka, nb, X + a0, ldx, Sm + (size_t)a0 * nb, Y, ldy, a0 == 0 ? alpha : 1.0, beta);
    if (na <= 0) break;
  }
}
void launch_mode_signs(hipStream_t st, int n1, const int* p1, double* Phi, int64_t nrows, int nev, double* sign) {
  if (nev <= 0) return;
  // sign[0 .. nev) is the result; the groups' bests follow it
  double* best = sign + nev;
  const int span = std::max(1, (n1 + MODE_SIGN_GROUPS - 1) / MODE_SIGN_GROUPS);
  const int ng = std::max(1, (n1 + span - 1) / span);
  k_mode_signs_part<<<ng, 256, 0, st>>>(n1, span, p1, Phi, nev, nev, best);
  k_mode_signs_fin<<<nev, 64, 0, st>>>(ng, nev, best, sign);
  if (nrows > 0) k_scale_cols<<<nblk(nrows * nev, 256), 256, 0, st>>>(nrows, nev, sign, Phi);
}
void launch_emass(hipStream_t st, int n, const int* elist, int nen, int d, const int64_t* eoff, const double* Mp,
                  double* emass) {
  if (n > 0) k_emass<<<nblk(n, 256), 256, 0, st>>>(n, elist, nen, d, eoff, Mp, emass);
}
void launch_sum_fixed(hipStream_t st, int n, const double* x, double* out) {
  k_sum_fixed<<<1, 256, 0, st>>>(n, x, out);
}
void launch_chol_inv(hipStream_t st, int n, const double* G, double* R, double* Ri, int* bad) {
  if (n <= 16) k_chol16<<<1, 64, 0, st>>>(n, G, R, Ri, bad);
  else k_chol_inv<<<1, 64, 0, st>>>(n, G, R, Ri, bad);
}
void launch_diag_scale(hipStream_t st, int n, const int* rowptr, const int* colind, const double* val, double* s) {
  if (n > 0) k_diag_scale<<<nblk(n, 256), 256, 0, st>>>(n, rowptr, colind, val, s);
}
void launch_front_scatter32(hipStream_t st, int n, const int* rowptr, const int* colind, const double* val,
                            const double* s, const int64_t* fmap, float* pool) {
  if (n > 0) k_front_scatter32<<<nblk(n, 256), 256, 0, st>>>(n, rowptr, colind, val, s, fmap, pool);
}
void launch_scale_rc(hipStream_t st, int64_t nrows, int nrhs, const double* rs, const double* cs, double* X) {
  if (nrows * nrhs > 0) k_scale_rc<<<nblk(nrows * nrhs, 256), 256, 0, st>>>(nrows, nrhs, rs, cs, X);
}
void launch_recip(hipStream_t st, int n, const double* x, double* y) {
  if (n > 0) k_recip<<<nblk(n, 256), 256, 0, st>>>(n, x, y);
}
void launch_axpby(hipStream_t st, int64_t n, double a, const double* X, double b, double* Y) {
  if (n > 0) k_axpby<<<nblk(n, 256), 256, 0, st>>>(n, a, X, b, Y);
}
int coldot_blocks(int64_t n) { return (int)((n + COLDOT_ROWS - 1) / COLDOT_ROWS); }
void launch_coldots(hipStream_t st, int64_t n, int nr, int npair, const double* A0, const double* B0,
                    const double* A1, const double* B1, const double* A2, const double* B2, double* part,
                    double* out) {
  const int nb = coldot_blocks(n);
  if (nb > 0) k_coldot_part<<<nb, 256, 0, st>>>(n, nr, npair, A0, B0, A1, B1, A2, B2, part);
  k_coldot_sum<<<nblk((int64_t)npair * nr, 256), 256, 0, st>>>(nb, nr, npair, part, out);
}
void launch_col_axpby(hipStream_t st, int64_t n, int nr, const double* a, const double* X, const double* b,
                      double* Y) {
  const int64_t total = n * nr;
  if (total > 0) k_col_axpby<<<(unsigned)std::min<int64_t>(nblk(total, 256), 8192), 256, 0, st>>>(n, nr, a, X, b, Y);
}
void launch_sumsq2(hipStream_t st, int64_t n, const double* x, const double* y, double* part) {
  k_sumsq2<<<SUMSQ_BLOCKS, 256, 0, st>>>(n, x, y, part);
}
void launch_identity_rows(hipStream_t st, double* X, int64_t nrows, int nrhs, int ndof1) {
  if (nrows * nrhs > 0) k_identity_rows<<<nblk(nrows * nrhs, 256), 256, 0, st>>>(X, nrows, nrhs, ndof1);
}

}  // namespace fgpu
