// Host symbolic analysis: fill-reducing ordering with the external DOFs last,
// supernode (front) structure, CSR pattern, pull-assembly contribution lists
// and the maps the device kernels use.
//
// Reference roles: csAllocPointerArrays / SPRSAS (symbolic assembly),
// csRenumber / SPRRNM (fill-reducing renumbering) and csPreassemble /
// SPRTRS+SPRSMB+SPRPMP (elimination tree and supernode partition) in
// asmExtensionModule.f90:227-875, with the status-2 equations ordered last
// (mspar(57)=1, solExtensionModule.f90:711) so that the trailing block of the
// factorisation is the Schur complement.  SAM's own ordering code is not in
// the reference tree; this is a geometric nested dissection on the node
// graph (separators become supernodes), which suits the grid-like parts the
// reducer sees.  The ordering does not change the reduced matrices, only the
// rounding.
#include <algorithm>
#include <numeric>
#include "fgpu.hpp"

namespace fgpu {
namespace {

struct NDBuilder {
  const Part& p;
  const std::vector<int>& adjptr;
  const std::vector<int>& adj;
  int leaf;
  std::vector<char> mark;            // scratch: 1 = in current set, 2 = in L
  std::vector<std::vector<int>> sn;  // node lists of the supernodes (postorder)
  std::vector<int> parent;

  NDBuilder(const Part& part, const std::vector<int>& ap, const std::vector<int>& a, int lf)
      : p(part), adjptr(ap), adj(a), leaf(lf), mark(part.nnod, 0) {}

  int make(std::vector<int>&& nodes) {
    sn.push_back(std::move(nodes));
    parent.push_back(-1);
    return (int)sn.size() - 1;
  }

  // returns the roots of the supernode forest spanning S
  std::vector<int> nd(std::vector<int> S) {
    std::vector<int> roots;
    if (S.empty()) return roots;
    if ((int)S.size() <= leaf) {
      roots.push_back(make(std::move(S)));
      return roots;
    }
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int v : S)
      for (int d = 0; d < 3; d++) {
        lo[d] = std::min(lo[d], p.xyz[3 * v + d]);
        hi[d] = std::max(hi[d], p.xyz[3 * v + d]);
      }
    int ax = 0;
    for (int d = 1; d < 3; d++)
      if (hi[d] - lo[d] > hi[ax] - lo[ax]) ax = d;
    std::stable_sort(S.begin(), S.end(), [&](int a, int b) {
      return p.xyz[3 * a + ax] < p.xyz[3 * b + ax];
    });
    int n = (int)S.size();
    // split at the widest coordinate gap near the median
    int k0 = std::max(1, (int)(0.4 * n)), k1 = std::min(n - 1, (int)(0.6 * n));
    if (k1 < k0) k1 = k0;
    int ks = n / 2;
    double best = -1.0;
    for (int k = k0; k <= k1; k++) {
      double g = p.xyz[3 * S[k] + ax] - p.xyz[3 * S[k - 1] + ax];
      if (g > best * (1.0 + 1e-9)) { best = g; ks = k; }
    }
    for (int i = 0; i < n; i++) mark[S[i]] = (i < ks) ? 2 : 1;
    std::vector<int> L(S.begin(), S.begin() + ks), R, sep;
    for (int i = ks; i < n; i++) {
      int v = S[i];
      bool touch = false;
      for (int k = adjptr[v]; k < adjptr[v + 1] && !touch; k++) touch = mark[adj[k]] == 2;
      (touch ? sep : R).push_back(v);
    }
    for (int v : S) mark[v] = 0;
    std::vector<int> rl = nd(std::move(L));
    std::vector<int> rr = nd(std::move(R));
    rl.insert(rl.end(), rr.begin(), rr.end());
    if (sep.empty()) return rl;
    int s = make(std::move(sep));
    for (int c : rl) parent[c] = s;
    roots.push_back(s);
    return roots;
  }
};

}  // namespace

int build_symbolic(const Part& p, const Sam& s, const SymbolicOptions& o,
                   Symbolic& y, std::string& err) {
  y = Symbolic();
  y.neq = s.neq;
  y.ndof1 = s.ndof1;
  y.ndof2 = s.ndof2;
  const int nnod = p.nnod;
  // node classes: 1 internal (has status-1 DOFs), 2 boundary, 0 no equations
  std::vector<int> ncls(nnod, 0), neqd(nnod, 0);
  for (int i = 0; i < nnod; i++) {
    for (int j = 0; j < s.nndof[i]; j++) {
      int m = s.msc[s.madof[i] - 1 + j];
      if (m > 0) { neqd[i]++; ncls[i] = m; }
    }
  }
  // Equation terms of every DOF: a free DOF is its own equation; a dependent
  // DOF of a rigid element is u_g + theta_g x (x_u - x_g) (translations) or
  // theta_g (rotations) of its independent node g; a fixed DOF has none.
  struct Term { int node, j; double c; };
  auto dof_terms = [&](int u, int j, Term* t) -> int {
    int d = s.madof[u] - 1 + j;
    if (s.msc[d] > 0) { t[0] = {u, j, 1.0}; return 1; }
    if (s.msc[d] != -1) return 0;
    int g = s.dep_indep[d], n = 0;
    double r[3];
    for (int c = 0; c < 3; c++) r[c] = p.xyz[3 * u + c] - p.xyz[3 * g + c];
    Term cand[3];
    int nc = 1;
    cand[0] = {g, j, 1.0};
    if (j == 0) { cand[1] = {g, 4, r[2]}; cand[2] = {g, 5, -r[1]}; nc = 3; }
    if (j == 1) { cand[1] = {g, 5, r[0]}; cand[2] = {g, 3, -r[2]}; nc = 3; }
    if (j == 2) { cand[1] = {g, 3, r[1]}; cand[2] = {g, 4, -r[0]}; nc = 3; }
    for (int k = 0; k < nc; k++) {
      if (cand[k].j >= s.nndof[g] || cand[k].c == 0.0) continue;
      if (s.msc[s.madof[g] - 1 + cand[k].j] <= 0) continue;
      t[n++] = cand[k];
    }
    return n;
  };
  // equation nodes each element couples (its own nodes with equations and
  // the independent nodes of its dependent DOFs)
  std::vector<int> ecptr(p.nel + 1, 0), ecn;
  {
    std::vector<int> stamp(nnod, -1);
    for (int e = 0; e < p.nel; e++) {
      const int d = elem_ndn(p.etype[e]);
      for (int q = p.eptr[e]; q < p.eptr[e + 1]; q++) {
        int u = p.enodes[q];
        for (int j = 0; j < d && j < s.nndof[u]; j++) {
          Term t[3];
          int nt = dof_terms(u, j, t);
          for (int k = 0; k < nt; k++)
            if (stamp[t[k].node] != e) { stamp[t[k].node] = e; ecn.push_back(t[k].node); }
        }
      }
      ecptr[e + 1] = (int)ecn.size();
    }
  }
  // node adjacency through elements (equation nodes only, no self loops)
  std::vector<int> nel_of(nnod + 1, 0);
  for (int e = 0; e < p.nel; e++)
    for (int k = ecptr[e]; k < ecptr[e + 1]; k++) nel_of[ecn[k] + 1]++;
  for (int i = 0; i < nnod; i++) nel_of[i + 1] += nel_of[i];
  std::vector<int> elist(nel_of[nnod]);
  {
    std::vector<int> fill(nel_of.begin(), nel_of.end() - 1);
    for (int e = 0; e < p.nel; e++)
      for (int k = ecptr[e]; k < ecptr[e + 1]; k++) elist[fill[ecn[k]]++] = e;
  }
  std::vector<int> adjptr(nnod + 1, 0), adj;
  {
    std::vector<int> stamp(nnod, -1), tmp;
    adj.reserve((size_t)nnod * 16);
    for (int u = 0; u < nnod; u++) {
      tmp.clear();
      if (ncls[u]) {
        for (int k = nel_of[u]; k < nel_of[u + 1]; k++) {
          int e = elist[k];
          for (int q = ecptr[e]; q < ecptr[e + 1]; q++) {
            int v = ecn[q];
            if (v != u && ncls[v] && stamp[v] != u) { stamp[v] = u; tmp.push_back(v); }
          }
        }
      }
      std::sort(tmp.begin(), tmp.end());
      adj.insert(adj.end(), tmp.begin(), tmp.end());
      adjptr[u + 1] = (int)adj.size();
    }
  }
  // internal-node graph (boundary nodes removed) for the dissection
  std::vector<int> iadjptr(nnod + 1, 0), iadj;
  iadj.reserve(adj.size());
  for (int u = 0; u < nnod; u++) {
    if (ncls[u] == 1)
      for (int k = adjptr[u]; k < adjptr[u + 1]; k++)
        if (ncls[adj[k]] == 1) iadj.push_back(adj[k]);
    iadjptr[u + 1] = (int)iadj.size();
  }
  NDBuilder nb(p, iadjptr, iadj, std::max(1, o.leaf_nodes));
  std::vector<int> inodes;
  for (int i = 0; i < nnod; i++)
    if (ncls[i] == 1) inodes.push_back(i);
  std::vector<int> roots = nb.nd(inodes);
  const int nsi = (int)nb.sn.size();  // internal supernodes
  y.nsup = nsi + 1;                   // + Schur front
  // node positions
  std::vector<int> npos(nnod, -1);
  int pos = 0;
  for (int sidx = 0; sidx < nsi; sidx++)
    for (int v : nb.sn[sidx]) npos[v] = pos++;
  for (int i = 0; i < nnod; i++)
    if (ncls[i] == 2) npos[i] = pos++;
  std::vector<int> node_at(pos);
  for (int i = 0; i < nnod; i++)
    if (npos[i] >= 0) node_at[npos[i]] = i;
  // DOF permutation: perm index of the first equation DOF of each node
  std::vector<int> nfirst(nnod, -1);
  y.eq_to_perm.assign(s.neq, -1);
  y.perm_to_eq.assign(s.neq, -1);
  int pp = 0;
  for (int q = 0; q < pos; q++) {
    int i = node_at[q];
    nfirst[i] = pp;
    for (int j = 0; j < s.nndof[i]; j++) {
      int d = s.madof[i] - 1 + j;
      if (s.msc[d] > 0) {
        y.eq_to_perm[s.meqn[d] - 1] = pp;
        y.perm_to_eq[pp] = s.meqn[d] - 1;
        pp++;
      }
    }
  }
  if (pp != s.neq) { err = "symbolic: equation count mismatch"; return -1; }
  // supernode parents
  y.parent.assign(y.nsup, -1);
  for (int sidx = 0; sidx < nsi; sidx++) y.parent[sidx] = nb.parent[sidx] >= 0 ? nb.parent[sidx] : nsi;
  // node-level update structures (children are complete before parents)
  std::vector<std::vector<int>> ust(nsi);
  std::vector<std::vector<int>> kids(y.nsup);
  for (int sidx = 0; sidx < nsi; sidx++) kids[y.parent[sidx]].push_back(sidx);
  {
    std::vector<int> stamp(pos, -1);
    for (int sidx = 0; sidx < nsi; sidx++) {
      int maxp = -1;
      for (int v : nb.sn[sidx]) maxp = std::max(maxp, npos[v]);
      std::vector<int>& u = ust[sidx];
      u.clear();
      for (int v : nb.sn[sidx])
        for (int k = adjptr[v]; k < adjptr[v + 1]; k++) {
          int w = npos[adj[k]];
          if (w > maxp && stamp[w] != sidx) { stamp[w] = sidx; u.push_back(w); }
        }
      for (int c : kids[sidx])
        for (int w : ust[c])
          if (w > maxp && stamp[w] != sidx) { stamp[w] = sidx; u.push_back(w); }
      std::sort(u.begin(), u.end());
    }
  }
  // DOF-level rows
  y.col0.assign(y.nsup, 0);
  y.ncol.assign(y.nsup, 0);
  y.nrow.assign(y.nsup, 0);
  y.rptr.assign(y.nsup + 1, 0);
  for (int sidx = 0; sidx < nsi; sidx++) {
    int c0 = nfirst[nb.sn[sidx][0]], nc = 0;
    for (int v : nb.sn[sidx]) nc += neqd[v];
    y.col0[sidx] = c0;
    y.ncol[sidx] = nc;
    int nr = nc;
    for (int w : ust[sidx]) nr += neqd[node_at[w]];
    y.nrow[sidx] = nr;
    y.rptr[sidx + 1] = y.rptr[sidx] + nr;
  }
  y.col0[nsi] = s.ndof1;
  y.ncol[nsi] = 0;
  y.nrow[nsi] = s.ndof2;
  y.rptr[nsi + 1] = y.rptr[nsi] + s.ndof2;
  y.rows.resize(y.rptr[y.nsup]);
  for (int sidx = 0; sidx < nsi; sidx++) {
    int* r = &y.rows[y.rptr[sidx]];
    int k = 0;
    for (int j = 0; j < y.ncol[sidx]; j++) r[k++] = y.col0[sidx] + j;
    for (int w : ust[sidx]) {
      int i = node_at[w];
      for (int j = 0; j < neqd[i]; j++) r[k++] = nfirst[i] + j;
    }
  }
  for (int j = 0; j < s.ndof2; j++) y.rows[y.rptr[nsi] + j] = s.ndof1 + j;
  std::vector<std::vector<int>>().swap(ust);
  // children, relative maps, levels
  y.chptr.assign(y.nsup + 1, 0);
  for (int sidx = 0; sidx < y.nsup; sidx++) y.chptr[sidx + 1] = y.chptr[sidx] + (int)kids[sidx].size();
  y.chlist.resize(y.chptr[y.nsup]);
  for (int sidx = 0; sidx < y.nsup; sidx++)
    std::copy(kids[sidx].begin(), kids[sidx].end(), y.chlist.begin() + y.chptr[sidx]);
  y.relptr.assign(y.nsup + 1, 0);
  for (int sidx = 0; sidx < y.nsup; sidx++)
    y.relptr[sidx + 1] = y.relptr[sidx] + (y.nrow[sidx] - y.ncol[sidx]) * (sidx < nsi ? 1 : 0);
  y.rel.resize(y.relptr[y.nsup]);
  for (int sidx = 0; sidx < nsi; sidx++) {
    int par = y.parent[sidx];
    const int* pr = &y.rows[y.rptr[par]];
    const int* pe = pr + y.nrow[par];
    int* rr = &y.rel[y.relptr[sidx]];
    const int* cr = &y.rows[y.rptr[sidx]];
    const int* it = pr;
    for (int k = y.ncol[sidx]; k < y.nrow[sidx]; k++) {
      it = std::lower_bound(it, pe, cr[k]);
      if (it == pe || *it != cr[k]) { err = "symbolic: update row missing in parent front"; return -1; }
      rr[k - y.ncol[sidx]] = (int)(it - pr);
    }
  }
  y.level.assign(y.nsup, 0);
  for (int sidx = 0; sidx < y.nsup; sidx++)
    for (int c : kids[sidx]) y.level[sidx] = std::max(y.level[sidx], y.level[c] + 1);
  y.nlevels = y.level[nsi] + 1;
  y.lvptr.assign(y.nlevels + 1, 0);
  for (int sidx = 0; sidx < y.nsup; sidx++) y.lvptr[y.level[sidx] + 1]++;
  for (int l = 0; l < y.nlevels; l++) y.lvptr[l + 1] += y.lvptr[l];
  y.lvsup.resize(y.nsup);
  {
    std::vector<int> fill(y.lvptr.begin(), y.lvptr.end() - 1);
    for (int sidx = 0; sidx < y.nsup; sidx++) y.lvsup[fill[y.level[sidx]]++] = sidx;
  }
  // fronts, work
  y.foff.assign(y.nsup + 1, 0);
  for (int sidx = 0; sidx < y.nsup; sidx++) {
    y.foff[sidx + 1] = y.foff[sidx] + (int64_t)y.nrow[sidx] * y.nrow[sidx];
    y.maxfront = std::max(y.maxfront, y.nrow[sidx]);
    for (int j = 0; j < y.ncol[sidx]; j++) {
      int64_t c = y.nrow[sidx] - 1 - j;
      y.factor_flops += c * c + 2 * c;
      y.nnzL += c + 1;
    }
  }
  y.fpool = y.foff[y.nsup];
  // column -> supernode
  std::vector<int> colsn(s.neq, nsi);
  for (int sidx = 0; sidx < nsi; sidx++)
    for (int j = 0; j < y.ncol[sidx]; j++) colsn[y.col0[sidx] + j] = sidx;
  // element DOF maps and packed element matrix offsets
  y.edof_ptr.assign(p.nel + 1, 0);
  y.elem_off.assign(p.nel + 1, 0);
  for (int e = 0; e < p.nel; e++) {
    int t = p.etype[e];
    y.edof_ptr[e + 1] = y.edof_ptr[e] + elem_nodes(t) * elem_ndn(t);
    y.elem_off[e + 1] = y.elem_off[e] + elem_psize(t);
  }
  y.elem_pool = y.elem_off[p.nel];
  y.edof_perm.assign(y.edof_ptr[p.nel], -1);
  for (int e = 0; e < p.nel; e++) {
    int t = p.etype[e], d = elem_ndn(t), k = 0;
    for (int q = p.eptr[e]; q < p.eptr[e + 1]; q++) {
      int i = p.enodes[q];
      for (int j = 0; j < d; j++, k++) {
        if (j >= s.nndof[i]) continue;
        int dof = s.madof[i] - 1 + j;
        if (s.msc[dof] > 0) y.edof_perm[y.edof_ptr[e] + k] = y.eq_to_perm[s.meqn[dof] - 1];
      }
    }
  }
  // CSR pattern: node neighbour lists (incl. self) sorted by position
  std::vector<int> nbptr(nnod + 1, 0), nbl;
  nbl.reserve(adj.size() + nnod);
  std::vector<int> coloff;  // DOF offset of each neighbour within the row
  coloff.reserve(adj.size() + nnod);
  std::vector<int> rowlen(nnod, 0);
  for (int u = 0; u < nnod; u++) {
    if (ncls[u]) {
      size_t b = nbl.size();
      nbl.push_back(u);
      for (int k = adjptr[u]; k < adjptr[u + 1]; k++) nbl.push_back(adj[k]);
      std::sort(nbl.begin() + b, nbl.end(), [&](int a, int c) { return npos[a] < npos[c]; });
      int off = 0;
      for (size_t k = b; k < nbl.size(); k++) { coloff.push_back(off); off += neqd[nbl[k]]; }
      rowlen[u] = off;
    }
    nbptr[u + 1] = (int)nbl.size();
  }
  y.rowptr.assign(s.neq + 1, 0);
  for (int q = 0; q < pos; q++) {
    int i = node_at[q];
    for (int j = 0; j < neqd[i]; j++) y.rowptr[nfirst[i] + j + 1] = rowlen[i];
  }
  for (int r = 0; r < s.neq; r++) y.rowptr[r + 1] += y.rowptr[r];
  const int64_t nnz = y.rowptr[s.neq];
  y.colind.resize(nnz);
  for (int u = 0; u < nnod; u++) {
    if (!ncls[u]) continue;
    for (int j = 0; j < neqd[u]; j++) {
      int r = nfirst[u] + j, k = y.rowptr[r];
      for (int q = nbptr[u]; q < nbptr[u + 1]; q++) {
        int v = nbl[q];
        for (int jj = 0; jj < neqd[v]; jj++) y.colind[k++] = nfirst[v] + jj;
      }
    }
  }
  // local DOF index -> equation-DOF rank within the node (-1 if no equation)
  std::vector<int> drank(s.ndof, -1);
  for (int i = 0; i < nnod; i++) {
    int k = 0;
    for (int j = 0; j < s.nndof[i]; j++)
      if (s.msc[s.madof[i] - 1 + j] > 0) drank[s.madof[i] - 1 + j] = k++;
  }
  // contributions: two passes over (element, node pair, dof pair, term pair);
  // a coefficient is kept only when the part has rigid elements
  const bool coef = !p.rgd_indep.empty();
  std::vector<int> ccount(nnz + 1, 0);
  auto find_nb = [&](int u, int v) -> int {
    const int* b = &nbl[nbptr[u]];
    const int* e = &nbl[nbptr[u + 1]];
    const int* it = std::lower_bound(b, e, v, [&](int a, int c) { return npos[a] < npos[c]; });
    return (int)(it - &nbl[0]);
  };
  for (int pass = 0; pass < 2; pass++) {
    std::vector<int> fill;
    if (pass == 1) {
      for (int64_t k = 0; k < nnz; k++) ccount[k + 1] += ccount[k];
      y.cptr.assign(ccount.begin(), ccount.end());
      y.cidx.resize(ccount[nnz]);
      if (coef) y.ccoef.resize(ccount[nnz]);
      fill.assign(ccount.begin(), ccount.end() - 1);
    }
    // column offset of node q in node pn's CSR row, for the last pair asked
    // (the DOF and term pairs of one node pair repeat it)
    int last_p = -1, last_q = -1, last_co = 0;
    auto col_of = [&](int pn, int qn) -> int {
      if (pn != last_p || qn != last_q) {
        last_p = pn;
        last_q = qn;
        last_co = coloff[find_nb(pn, qn)];
      }
      return last_co;
    };
    for (int e = 0; e < p.nel; e++) {
      int t = p.etype[e], d = elem_ndn(t), nn = elem_nodes(t);
      const int* en = &p.enodes[p.eptr[e]];
      // equation terms of the element's local DOFs, each with the CSR start
      // of its row and its rank within its node
      Term tl[8 * 6][3];
      int nt[8 * 6], trk[8 * 6][3];
      int64_t trow[8 * 6][3];
      for (int a = 0; a < nn; a++)
        for (int i = 0; i < d; i++) {
          const int li = a * d + i;
          nt[li] = i < s.nndof[en[a]] ? dof_terms(en[a], i, tl[li]) : 0;
          for (int ti = 0; ti < nt[li]; ti++) {
            const Term& T = tl[li][ti];
            trk[li][ti] = drank[s.madof[T.node] - 1 + T.j];
            trow[li][ti] = y.rowptr[nfirst[T.node] + trk[li][ti]];
          }
        }
      for (int a = 0; a < nn; a++)
        for (int b = 0; b < nn; b++) {
          int blk = a >= b ? a * (a + 1) / 2 + b : b * (b + 1) / 2 + a;
          for (int i = 0; i < d; i++)
            for (int j = 0; j < d; j++) {
              int off = a >= b ? (blk * d * d + i * d + j) : (blk * d * d + j * d + i);
              const int li = a * d + i, lj = b * d + j;
              for (int ti = 0; ti < nt[li]; ti++) {
                const Term& P = tl[li][ti];
                for (int tj = 0; tj < nt[lj]; tj++) {
                  const Term& Q = tl[lj][tj];
                  const int co = col_of(P.node, Q.node);
                  int64_t k = trow[li][ti] + co + trk[lj][tj];
                  if (pass == 0) { ccount[k + 1]++; continue; }
                  int64_t slot = fill[k]++;
                  y.cidx[slot] = y.elem_off[e] + off;
                  if (coef) y.ccoef[slot] = P.c * Q.c;
                }
              }
            }
        }
    }
  }
  // lower CSR entries -> front positions
  y.fmap.assign(nnz, -1);
  for (int r = 0; r < s.neq; r++) {
    int last_sn = -1;  // consecutive columns of a row mostly share a front
    int64_t rpos = 0;
    for (int64_t k = y.rowptr[r]; k < y.rowptr[r + 1]; k++) {
      int c = y.colind[k];
      if (c > r) continue;
      int sn = colsn[c];
      if (sn != last_sn) {
        const int* rb = &y.rows[y.rptr[sn]];
        const int* re = rb + y.nrow[sn];
        const int* it = std::lower_bound(rb, re, r);
        if (it == re || *it != r) { err = "symbolic: matrix entry outside its front"; return -1; }
        last_sn = sn;
        rpos = (int64_t)(it - rb);
      }
      int64_t ld = y.nrow[sn];
      y.fmap[k] = y.foff[sn] + rpos + (int64_t)(c - y.col0[sn]) * ld;
    }
  }
  (void)roots;
  return 0;
}

}  // namespace fgpu
