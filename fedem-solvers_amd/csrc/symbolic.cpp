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
#include <cstdint>
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
  // Equation terms of every DOF: a free DOF is its own equation, a dependent
  // DOF the terms of its constraint equation, a fixed DOF has none.
  auto nterms = [&](int u, int j) -> int {
    const int d = s.madof[u] - 1 + j, m = s.msc[d];
    return m > 0 ? 1 : m == -1 ? s.ceq_ptr[d + 1] - s.ceq_ptr[d] : 0;
  };
  auto term_node = [&](int u, int j, int k) -> int {
    const int d = s.madof[u] - 1 + j;
    return s.msc[d] > 0 ? u : s.ceq_node[s.ceq_ptr[d] + k];
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
          const int nt = nterms(u, j);
          for (int k = 0; k < nt; k++) {
            const int v = term_node(u, j, k);
            if (stamp[v] != e) { stamp[v] = e; ecn.push_back(v); }
          }
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
  // packed element matrix offsets
  y.elem_off.assign(p.nel + 1, 0);
  for (int e = 0; e < p.nel; e++) y.elem_off[e + 1] = y.elem_off[e] + elem_psize(p.etype[e]);
  y.elem_pool = y.elem_off[p.nel];
  if (y.elem_pool > INT32_MAX) {
    err = "symbolic: the part's element matrices exceed 32-bit pool indices";
    return -1;
  }
  // Inputs of the device construction of the CSR matrix, the pull-assembly
  // contribution lists and the front map (structure.hip): per node the
  // first permuted equation and the equation count, its neighbour list
  // (itself included, any order), per row its node, per column its front,
  // per DOF its rank among the node's equation DOFs.  The CSR rows follow
  // from the neighbour lists: a row of node u holds the equations of its
  // neighbours in permuted order.
  y.nfirst = nfirst;
  y.neqd = neqd;
  y.nbptr.assign(nnod + 1, 0);
  y.nbl.clear();
  y.nbl.reserve(adj.size() + nnod);
  y.rowptr.assign(s.neq + 1, 0);
  std::vector<int> rowlen(nnod, 0);
  for (int u = 0; u < nnod; u++) {
    if (ncls[u]) {
      y.nbl.push_back(u);
      int len = neqd[u];
      for (int k = adjptr[u]; k < adjptr[u + 1]; k++) {
        y.nbl.push_back(adj[k]);
        len += neqd[adj[k]];
      }
      rowlen[u] = len;
    }
    y.nbptr[u + 1] = (int)y.nbl.size();
  }
  y.rnode.assign(s.neq, -1);
  for (int q = 0; q < pos; q++) {
    const int i = node_at[q];
    for (int j = 0; j < neqd[i]; j++) {
      y.rnode[nfirst[i] + j] = i;
      y.rowptr[nfirst[i] + j + 1] = rowlen[i];
    }
  }
  for (int r = 0; r < s.neq; r++) y.rowptr[r + 1] += y.rowptr[r];
  y.colsn.swap(colsn);
  y.drank.assign(s.ndof, -1);
  for (int i = 0; i < nnod; i++) {
    int k = 0;
    for (int j = 0; j < s.nndof[i]; j++)
      if (s.msc[s.madof[i] - 1 + j] > 0) y.drank[s.madof[i] - 1 + j] = k++;
  }
  // contributions: every pair of equation terms of an element's DOFs
  y.has_rigid = !s.ceq_dof.empty();
  y.ncontrib = 0;
  for (int e = 0; e < p.nel; e++) {
    const int t = p.etype[e], d = elem_ndn(t);
    int64_t nt = 0;
    for (int q = p.eptr[e]; q < p.eptr[e + 1]; q++)
      for (int j = 0; j < d && j < s.nndof[p.enodes[q]]; j++) nt += nterms(p.enodes[q], j);
    y.ncontrib += nt * nt;
  }
  if (y.rowptr[s.neq] < 0 || y.ncontrib > INT32_MAX) {
    err = "symbolic: the part's matrix is too large for 32-bit contribution indices";
    return -1;
  }
  (void)roots;
  return 0;
}

}  // namespace fgpu
