// Device construction of a part's assembled-matrix structure (the per-part
// symbolic assembly of csAllocPointerArrays / csPreassemble,
// asmExtensionModule.f90:227-875, after the host has chosen the ordering and
// the fronts):
//
//   k_nbr_sort      each node's neighbour list in permuted-equation order and
//                   the DOF offset of every neighbour within the node's rows
//   k_csr_cols      CSR column indices (full symmetric pattern)
//   k_contrib       every element-matrix entry's CSR position: counted
//                   (pass 0), then written (pass 1) into the pull-assembly
//                   contribution lists; a dependent DOF of a constraint
//                   element (rigid or weighted-average) contributes through
//                   the terms of its constraint equation with their
//                   coefficients
//   k_seg_sort      each CSR entry's contributions in a fixed order (element
//                   matrix position, then coefficient), so that the assembly
//                   sums in the same order whatever the atomics did
//   k_fmap          front position of every lower CSR entry
//   scan kernels    exclusive prefix sum of the per-entry counts
#include "kernels.hpp"

namespace fgpu {

// ---------------------------------------------------------------- neighbours
__global__ void k_nbr_sort(int nnod, const int* __restrict__ nbptr, int* __restrict__ nbl,
                           const int* __restrict__ nfirst, const int* __restrict__ neqd, int* __restrict__ nbo) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= nnod) return;
  const int b = nbptr[u], n = nbptr[u + 1] - b;
  int* l = nbl + b;
  for (int i = 1; i < n; i++) {  // insertion sort by first equation (lists are short)
    const int v = l[i], key = nfirst[v];
    int j = i - 1;
    while (j >= 0 && nfirst[l[j]] > key) {
      l[j + 1] = l[j];
      j--;
    }
    l[j + 1] = v;
  }
  int off = 0;
  for (int i = 0; i < n; i++) {
    nbo[b + i] = off;
    off += neqd[l[i]];
  }
}

// position of node v in node u's sorted neighbour list (-1: absent)
__device__ inline int find_nbr(int u, int v, const int* __restrict__ nbptr, const int* __restrict__ nbl,
                               const int* __restrict__ nfirst) {
  int lo = nbptr[u], hi = nbptr[u + 1] - 1;
  const int key = nfirst[v];
  while (lo <= hi) {
    const int mid = (lo + hi) >> 1, k = nfirst[nbl[mid]];
    if (k == key) return mid;
    if (k < key) lo = mid + 1;
    else hi = mid - 1;
  }
  return -1;
}

// ---------------------------------------------------------------- CSR columns
// one 64-lane workgroup per row: lane q writes the columns of neighbour q
__global__ __launch_bounds__(64) void k_csr_cols(int neq, const int* __restrict__ rnode, const int* __restrict__ rowptr,
                                                 const int* __restrict__ nbptr, const int* __restrict__ nbl,
                                                 const int* __restrict__ nbo, const int* __restrict__ nfirst,
                                                 const int* __restrict__ neqd, int* __restrict__ colind) {
  const int r = blockIdx.x;
  const int u = rnode[r], base = rowptr[r];
  for (int q = nbptr[u] + threadIdx.x; q < nbptr[u + 1]; q += 64) {
    const int v = nbl[q], f = nfirst[v], o = base + nbo[q];
    for (int j = 0; j < neqd[v]; j++) colind[o + j] = f + j;
  }
}

// ---------------------------------------------------------------- contributions
// The equation terms of a DOF: a free DOF is its own equation (one term,
// coefficient 1), a dependent DOF of a constraint element has the terms of
// its constraint equation (ceq_* tables), a fixed DOF none.
struct ETerm {
  int node, rank;
  double c;
};

__device__ inline ETerm dof_term(int self_node, int d, int base, int k, const int* __restrict__ drank,
                                 const int* __restrict__ ceq_node, const int* __restrict__ ceq_dof,
                                 const double* __restrict__ ceq_coef) {
  if (base < 0) return {self_node, drank[d], 1.0};
  return {ceq_node[base + k], drank[ceq_dof[base + k]], ceq_coef[base + k]};
}

// One 64-lane workgroup per element.  Pass 0 counts the contributions of
// every CSR entry; pass 1 appends them at cptr[k] + (running count).
__global__ __launch_bounds__(64) void k_contrib(
    int pass, const int* __restrict__ etype, const int* __restrict__ eptr, const int* __restrict__ enodes,
    const int64_t* __restrict__ eoff, const int* __restrict__ madof, const int* __restrict__ msc,
    const int* __restrict__ nndof, const int* __restrict__ ceq_ptr, const int* __restrict__ ceq_node,
    const int* __restrict__ ceq_dof, const double* __restrict__ ceq_coef, const int* __restrict__ drank,
    const int* __restrict__ nfirst, const int* __restrict__ rowptr, const int* __restrict__ nbptr,
    const int* __restrict__ nbl, const int* __restrict__ nbo, int* __restrict__ cnt, const int* __restrict__ cptr,
    int* __restrict__ cidx, double* __restrict__ ccoef, int* __restrict__ bad) {
  // per local DOF: its node, global DOF, term-table base (-1: the DOF itself)
  // and term count
  __shared__ int lnode[48], ldof[48], lbase[48], nt[48];
  const int e = blockIdx.x, lane = threadIdx.x;
  const int t = etype[e], d = t == 44 ? 3 : 6, nn = t == 44 ? 8 : t == 24 ? 4 : 2, nd = nn * d;
  const int* en = enodes + eptr[e];
  if (lane < nd) {
    const int a = lane / d, i = lane % d, u = en[a];
    int n = 0, base = -1, dof = 0;
    if (i < nndof[u]) {
      dof = madof[u] - 1 + i;
      const int m = msc[dof];
      if (m > 0) n = 1;
      else if (m == -1) { base = ceq_ptr[dof]; n = ceq_ptr[dof + 1] - base; }
    }
    lnode[lane] = u;
    ldof[lane] = dof;
    lbase[lane] = base;
    nt[lane] = n;
  }
  __syncthreads();
  const int64_t e0 = eoff[e];
  for (int idx = lane; idx < nd * nd; idx += 64) {
    const int li = idx / nd, lj = idx % nd, a = li / d, i = li % d, b = lj / d, j = lj % d;
    const int off = a >= b ? ((a * (a + 1) / 2 + b) * d + i) * d + j : ((b * (b + 1) / 2 + a) * d + j) * d + i;
    for (int ti = 0; ti < nt[li]; ti++) {
      const ETerm P = dof_term(lnode[li], ldof[li], lbase[li], ti, drank, ceq_node, ceq_dof, ceq_coef);
      const int prow = rowptr[nfirst[P.node] + P.rank];
      for (int tj = 0; tj < nt[lj]; tj++) {
        const ETerm Q = dof_term(lnode[lj], ldof[lj], lbase[lj], tj, drank, ceq_node, ceq_dof, ceq_coef);
        const int q = find_nbr(P.node, Q.node, nbptr, nbl, nfirst);
        if (q < 0) {
          atomicOr(bad, 1);
          continue;
        }
        const int k = prow + nbo[q] + Q.rank;
        if (pass == 0) {
          atomicAdd(&cnt[k], 1);
        } else {
          const int slot = cptr[k] + atomicAdd(&cnt[k], 1);
          cidx[slot] = (int)(e0 + off);
          if (ccoef) ccoef[slot] = P.c * Q.c;
        }
      }
    }
  }
}

// contributions of each CSR entry sorted by (element-matrix position, coefficient)
__global__ void k_seg_sort(int nnz, const int* __restrict__ cptr, int* __restrict__ cidx, double* __restrict__ ccoef) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnz) return;
  const int b = cptr[k], n = cptr[k + 1] - b;
  int* c = cidx + b;
  double* w = ccoef ? ccoef + b : nullptr;
  for (int i = 1; i < n; i++) {
    const int key = c[i];
    const double kw = w ? w[i] : 0.0;
    int j = i - 1;
    while (j >= 0 && (c[j] > key || (c[j] == key && w && w[j] > kw))) {
      c[j + 1] = c[j];
      if (w) w[j + 1] = w[j];
      j--;
    }
    c[j + 1] = key;
    if (w) w[j + 1] = kw;
  }
}

// ---------------------------------------------------------------- front map
// front position of every lower entry (c <= r) of row r, -1 for the others
__global__ void k_fmap(int neq, const int* __restrict__ rowptr, const int* __restrict__ colind,
                       const int* __restrict__ colsn, DevSym S, int64_t* __restrict__ fmap, int* __restrict__ bad) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= neq) return;
  int last = -1;
  int64_t rpos = 0;
  for (int k = rowptr[r]; k < rowptr[r + 1]; k++) {
    const int c = colind[k];
    if (c > r) {
      fmap[k] = -1;
      continue;
    }
    const int sn = colsn[c];
    if (sn != last) {
      const int* rw = S.rows + S.rptr[sn];
      int lo = 0, hi = S.nrow[sn] - 1, pos = -1;
      while (lo <= hi) {
        const int mid = (lo + hi) >> 1;
        if (rw[mid] == r) { pos = mid; break; }
        if (rw[mid] < r) lo = mid + 1;
        else hi = mid - 1;
      }
      if (pos < 0) {
        atomicOr(bad, 2);
        fmap[k] = -1;
        continue;
      }
      last = sn;
      rpos = pos;
    }
    fmap[k] = S.foff[sn] + rpos + (int64_t)(c - S.col0[sn]) * S.nrow[sn];
  }
}

// ---------------------------------------------------------------- scan
constexpr int SCAN_BLOCK = 1024, SCAN_ITEMS = 4, SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;

// exclusive scan of one value per thread across the workgroup; returns the
// workgroup total in *total
__device__ inline int block_exclusive(int v, int* lds, int* total) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int x = v;
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) lds[wv] = x;
  __syncthreads();
  if (threadIdx.x < 64) {
    int s = threadIdx.x < SCAN_BLOCK / 64 ? lds[threadIdx.x] : 0;
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(s, o, 64);
      if (threadIdx.x >= o) s += y;
    }
    if (threadIdx.x < SCAN_BLOCK / 64) lds[64 + threadIdx.x] = s;  // inclusive over waves
  }
  __syncthreads();
  const int before = wv ? lds[64 + wv - 1] : 0;
  *total = lds[64 + SCAN_BLOCK / 64 - 1];
  return before + x - v;
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_tiles(int n, const int* __restrict__ in, int* __restrict__ out,
                                                           int* __restrict__ tsum) {
  __shared__ int lds[128];
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
  int v[SCAN_ITEMS], s = 0;
  for (int q = 0; q < SCAN_ITEMS; q++) {
    v[q] = base + q < n ? in[base + q] : 0;
    s += v[q];
  }
  int total;
  int run = block_exclusive(s, lds, &total);
  for (int q = 0; q < SCAN_ITEMS; q++) {
    if (base + q < n) out[base + q] = run;
    run += v[q];
  }
  if (threadIdx.x == 0) tsum[blockIdx.x] = total;
}

// one workgroup: exclusive scan of the tile sums (in place), total at [ntiles]
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_sums(int ntiles, int* __restrict__ tsum) {
  __shared__ int lds[128];
  int carry = 0;
  for (int b0 = 0; b0 < ntiles; b0 += SCAN_BLOCK) {
    const int i = b0 + threadIdx.x;
    const int v = i < ntiles ? tsum[i] : 0;
    int total;
    const int ex = block_exclusive(v, lds, &total);
    __syncthreads();
    if (i < ntiles) tsum[i] = carry + ex;
    carry += total;
    __syncthreads();
  }
  if (threadIdx.x == 0) tsum[ntiles] = carry;
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_add(int n, int ntiles, const int* __restrict__ tsum,
                                                         int* __restrict__ out) {
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
  const int add = blockIdx.x < ntiles ? tsum[blockIdx.x] : 0;
  for (int q = threadIdx.x; q < SCAN_TILE; q += SCAN_BLOCK)
    if (base + q < n) out[base + q] += add;
  if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = tsum[ntiles];
}

// ---------------------------------------------------------------- launchers
void launch_nbr_sort(hipStream_t st, int nnod, const int* nbptr, int* nbl, const int* nfirst, const int* neqd,
                     int* nbo) {
  if (nnod > 0) k_nbr_sort<<<(nnod + 255) / 256, 256, 0, st>>>(nnod, nbptr, nbl, nfirst, neqd, nbo);
}

void launch_csr_cols(hipStream_t st, int neq, const int* rnode, const int* rowptr, const int* nbptr, const int* nbl,
                     const int* nbo, const int* nfirst, const int* neqd, int* colind) {
  if (neq > 0) k_csr_cols<<<neq, 64, 0, st>>>(neq, rnode, rowptr, nbptr, nbl, nbo, nfirst, neqd, colind);
}

void launch_contrib(hipStream_t st, int pass, int nel, const int* etype, const int* eptr, const int* enodes,
                    const int64_t* eoff, const int* madof, const int* msc, const int* nndof, const int* ceq_ptr,
                    const int* ceq_node, const int* ceq_dof, const double* ceq_coef, const int* drank,
                    const int* nfirst, const int* rowptr, const int* nbptr, const int* nbl, const int* nbo, int* cnt,
                    const int* cptr, int* cidx, double* ccoef, int* bad) {
  if (nel > 0)
    k_contrib<<<nel, 64, 0, st>>>(pass, etype, eptr, enodes, eoff, madof, msc, nndof, ceq_ptr, ceq_node, ceq_dof,
                                  ceq_coef, drank, nfirst, rowptr, nbptr, nbl, nbo, cnt, cptr, cidx, ccoef, bad);
}

void launch_seg_sort(hipStream_t st, int nnz, const int* cptr, int* cidx, double* ccoef) {
  if (nnz > 0) k_seg_sort<<<(nnz + 255) / 256, 256, 0, st>>>(nnz, cptr, cidx, ccoef);
}

void launch_fmap(hipStream_t st, int neq, const int* rowptr, const int* colind, const int* colsn, const DevSym& S,
                 int64_t* fmap, int* bad) {
  if (neq > 0) k_fmap<<<(neq + 255) / 256, 256, 0, st>>>(neq, rowptr, colind, colsn, S, fmap, bad);
}

int scan_tiles(int n) { return (n + SCAN_TILE - 1) / SCAN_TILE; }

void launch_exclusive_scan(hipStream_t st, int n, const int* in, int* out, int* tsum) {
  const int nt = scan_tiles(n);
  if (nt > 0) k_scan_tiles<<<nt, SCAN_BLOCK, 0, st>>>(n, in, out, tsum);
  k_scan_sums<<<1, SCAN_BLOCK, 0, st>>>(nt, tsum);
  k_scan_add<<<nt > 0 ? nt : 1, SCAN_BLOCK, 0, st>>>(n, nt, tsum, out);
}

}  // namespace fgpu
