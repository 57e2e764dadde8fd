// C ABI of the plug-in seam (include/fedem_reducer_gpu.h, part 2).
#include <algorithm>
#include <cstring>
#include <vector>
#include <exception>
#include <new>
#include "../../include/fedem_reducer_gpu.h"
#include "reducer_gpu.hpp"

struct fgpu_reducer {
  fgpu::DeviceReducer red;
  fgpu::Part part;
  bool have_part = false, analyzed = false, reduced = false;
  int ngen = 0;
  fgpu::ReduceOptions opts;
  std::string err;
};

#define GUARD(h)                         \
  if (!(h)) return -1;                   \
  try {
#define END_GUARD(h)                     \
  }                                      \
  catch (const std::exception& ex) {     \
    (h)->err = ex.what();                \
    return -99;                          \
  }

namespace fgpu {
// The numeric options of fgpu_set_option (and of the server's requests).
int set_reduce_option(ReduceOptions& o, const std::string& n, double value) {
  if (n == "lanczosBlock") o.lanczos_block = (int)value;
  else if (n == "lanczosMaxVectors") o.lanczos_maxvec = (int)value;
  else if (n == "lanczosTol") o.tol_eig = value;
  else if (n == "eigenshift") o.eig_shift = value;
  else if (n == "seed") o.seed = (uint64_t)value;
  else if (n == "leafNodes") o.sym.leaf_nodes = (int)value;
  else if (n == "timeUpdates") o.time_updates = value != 0.0;
  else if (n == "mixedPrecision") o.mixed = value != 0.0;
  else if (n == "solveSplitK") o.solve_split = (int)value;
  else if (n == "factorLookahead") o.factor_lookahead = value != 0.0;
  else if (n == "factorPanel") o.factor_panel = (int)value;
  else if (n == "refineTol") o.ir_tol = value;
  else if (n == "refineMaxSteps") o.ir_maxit = (int)value;
  else if (n == "tolFactorize") o.tol_factorize = value;
  else if (n == "singularityHandler") o.sing_handler = (int)value;
  else return -1;
  return 0;
}
}  // namespace fgpu

extern "C" {

int fgpu_create(fgpu_reducer** h) {
  if (!h) return -1;
  *h = new (std::nothrow) fgpu_reducer();
  return *h ? 0 : -1;
}

void fgpu_destroy(fgpu_reducer* h) { delete h; }

const char* fgpu_last_error(const fgpu_reducer* h) { return h ? h->err.c_str() : "null handle"; }

int fgpu_set_part(fgpu_reducer* h, int nnod, const double* xyz, const int* status, int nel, const int* melcon,
                  const int* mpmnpc, const int* mmnpc, const double* eprop) {
  GUARD(h)
  if (nnod <= 0 || nel < 0 || !xyz || !status || (nel > 0 && (!melcon || !mpmnpc || !mmnpc || !eprop))) {
    h->err = "fgpu_set_part: invalid arguments";
    return -1;
  }
  fgpu::Part& p = h->part;
  p = fgpu::Part();
  p.nnod = nnod;
  p.nel = nel;
  p.xyz.assign(xyz, xyz + 3 * (size_t)nnod);
  p.status.assign(status, status + nnod);
  p.etype.assign(melcon, melcon + nel);
  p.eptr.resize(nel + 1);
  for (int e = 0; e <= nel; e++) p.eptr[e] = mpmnpc[e] - 1;
  for (int e = 0; e < nel; e++) {
    int t = melcon[e];
    if (t != fgpu::HEX8 && t != fgpu::ANDES4 && t != fgpu::BEAM2) {
      h->err = "fgpu_set_part: unsupported element type " + std::to_string(t);
      return -2;
    }
    if (p.eptr[e + 1] - p.eptr[e] != fgpu::elem_nodes(t)) {
      h->err = "fgpu_set_part: element " + std::to_string(e + 1) + " has a wrong node count";
      return -2;
    }
  }
  p.enodes.resize(p.eptr[nel]);
  for (int k = 0; k < p.eptr[nel]; k++) {
    int n = mmnpc[k] - 1;
    if (n < 0 || n >= nnod) {
      h->err = "fgpu_set_part: node number out of range";
      return -2;
    }
    p.enodes[k] = n;
  }
  p.eprop.assign(eprop, eprop + (size_t)fgpu::NPROP * nel);
  h->have_part = true;
  h->analyzed = h->reduced = false;
  return 0;
  END_GUARD(h)
}

int fgpu_set_rigid(fgpu_reducer* h, int nrgd, const int* indep, const int* dofmask, const int* dptr,
                   const int* dep) {
  GUARD(h)
  if (!h->have_part || nrgd < 0 || (nrgd > 0 && (!indep || !dofmask || !dptr || !dep))) {
    h->err = "fgpu_set_rigid: invalid arguments (set the part first)";
    return -1;
  }
  fgpu::Part& p = h->part;
  p.rgd_indep.resize(nrgd);
  p.rgd_mask.assign(dofmask, dofmask + nrgd);
  p.rgd_ptr.resize(nrgd + 1);
  for (int r = 0; r <= nrgd; r++) p.rgd_ptr[r] = dptr[r] - 1;
  for (int r = 0; r < nrgd; r++) {
    p.rgd_indep[r] = indep[r] - 1;
    if (p.rgd_ptr[r + 1] < p.rgd_ptr[r]) { h->err = "fgpu_set_rigid: bad pointers"; return -2; }
  }
  p.rgd_dep.resize(nrgd ? p.rgd_ptr[nrgd] : 0);
  for (size_t k = 0; k < p.rgd_dep.size(); k++) p.rgd_dep[k] = dep[k] - 1;
  h->analyzed = h->reduced = false;
  return 0;
  END_GUARD(h)
}

int fgpu_set_wavgm(fgpu_reducer* h, int nwv, const int* ref, const int* refc, const int* indc, const int* nptr,
                   const int* nodes, const int* wptr, const double* weights) {
  GUARD(h)
  if (!h->have_part || nwv < 0 || (nwv > 0 && (!ref || !refc || !indc || !nptr || !nodes || !wptr))) {
    h->err = "fgpu_set_wavgm: invalid arguments (set the part first)";
    return -1;
  }
  fgpu::Part& p = h->part;
  p.wv_ref.resize(nwv);
  p.wv_refc.assign(refc, refc + nwv);
  p.wv_indc.assign(indc, indc + 6 * (size_t)nwv);
  p.wv_nptr.resize(nwv + 1);
  p.wv_wptr.resize(nwv + 1);
  for (int w = 0; w <= nwv; w++) {
    p.wv_nptr[w] = nwv ? nptr[w] - 1 : 0;
    p.wv_wptr[w] = nwv ? wptr[w] - 1 : 0;
  }
  for (int w = 0; w < nwv; w++) {
    p.wv_ref[w] = ref[w] - 1;
    if (p.wv_nptr[w + 1] < p.wv_nptr[w] || p.wv_wptr[w + 1] < p.wv_wptr[w] || p.wv_nptr[0] != 0 ||
        p.wv_wptr[0] != 0) {
      h->err = "fgpu_set_wavgm: bad pointers";
      return -2;
    }
  }
  p.wv_nodes.resize(nwv ? p.wv_nptr[nwv] : 0);
  for (size_t k = 0; k < p.wv_nodes.size(); k++) p.wv_nodes[k] = nodes[k] - 1;
  const size_t nw = nwv ? p.wv_wptr[nwv] : 0;
  if (nw && !weights) {
    h->err = "fgpu_set_wavgm: weights missing";
    return -1;
  }
  p.wv_w.assign(weights, weights + nw);
  h->analyzed = h->reduced = false;
  return 0;
  END_GUARD(h)
}

int fgpu_analyze(fgpu_reducer* h, int ngen, int neval, int flags) {
  GUARD(h)
  if (!h->have_part) { h->err = "fgpu_analyze: no part"; return -1; }
  fgpu::ReduceOptions o = h->opts;
  o.ngen = ngen < 0 ? 0 : ngen;
  h->ngen = o.ngen;
  o.neval = neval < 0 ? 0 : neval;
  o.want_vgi = flags & 1;
  o.no_mass = (flags & 2) != 0;
  int rc = h->red.analyze(h->part, o, h->err, true);
  h->analyzed = rc == 0;
  h->reduced = false;
  return rc;
  END_GUARD(h)
}

int fgpu_set_option(fgpu_reducer* h, const char* name, double value) {
  if (!h || !name) return -1;
  if (fgpu::set_reduce_option(h->opts, name, value)) {
    h->err = std::string("fgpu_set_option: unknown option ") + name;
    return -1;
  }
  return 0;
}

int fgpu_get_sizes(const fgpu_reducer* h, int* s) {
  if (!h || !h->analyzed || !s) return -1;
  const fgpu::Sam& m = h->red.sam();
  s[0] = m.nnod; s[1] = m.ndof; s[2] = m.neq; s[3] = m.ndof1; s[4] = m.ndof2;
  s[5] = m.ndim_ext + h->ngen;
  s[6] = m.nenod; s[7] = h->part.nel;
  return 0;
}

int fgpu_get_sam(const fgpu_reducer* h, int* madof, int* msc, int* meqn, int* meqn1, int* meqn2, int* pos2) {
  if (!h || !h->analyzed) return -1;
  const fgpu::Sam& m = h->red.sam();
  auto cp = [](int* dst, const std::vector<int>& v) { if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(int)); };
  cp(madof, m.madof); cp(msc, m.msc); cp(meqn, m.meqn); cp(meqn1, m.meqn1); cp(meqn2, m.meqn2); cp(pos2, m.dofPosIn2);
  return 0;
}

int fgpu_get_constraints(const fgpu_reducer* h, int* sizes, int* mpmceq, int* mmceq, double* ttcc) {
  if (!h || !h->analyzed || !sizes) return -1;
  const fgpu::Sam& m = h->red.sam();
  int nceq = 0, nmm = 0;
  for (int d = 0; d < m.ndof; d++)
    if (m.msc[d] == -1) {
      if (mpmceq) mpmceq[nceq] = nmm + 1;
      if (mmceq) mmceq[nmm] = d + 1;
      if (ttcc) ttcc[nmm] = 0.0;
      nmm++;
      for (int k = m.ceq_ptr[d]; k < m.ceq_ptr[d + 1]; k++) {
        if (mmceq) mmceq[nmm] = m.ceq_dof[k] + 1;
        if (ttcc) ttcc[nmm] = m.ceq_coef[k];
        nmm++;
      }
      nceq++;
    }
  if (mpmceq) mpmceq[nceq] = nmm + 1;
  sizes[0] = nceq;
  sizes[1] = nmm;
  return 0;
}

int fgpu_reduce(fgpu_reducer* h) {
  GUARD(h)
  if (!h->analyzed) { h->err = "fgpu_reduce: not analyzed"; return -1; }
  int rc = h->red.reduce(h->err);
  h->reduced = rc == 0;
  return rc;
  END_GUARD(h)
}

int fgpu_get_reduced(const fgpu_reducer* h, double* SK, double* SM, double* G) {
  if (!h || !h->reduced) return -1;
  const fgpu::Results& r = h->red.results();
  if (SK) std::memcpy(SK, r.SK.data(), r.SK.size() * sizeof(double));
  if (SM && !r.SM.empty()) std::memcpy(SM, r.SM.data(), r.SM.size() * sizeof(double));
  if (G && !r.G.empty()) std::memcpy(G, r.G.data(), r.G.size() * sizeof(double));
  return 0;
}

int fgpu_get_eigen(const fgpu_reducer* h, int* nev, double* lambda, double* evl) {
  if (!h || !h->reduced) return -1;
  const fgpu::Results& r = h->red.results();
  if (nev) *nev = r.nev;
  if (lambda) std::memcpy(lambda, r.lambda.data(), r.lambda.size() * sizeof(double));
  if (evl) std::memcpy(evl, r.eval.data(), r.eval.size() * sizeof(double));
  return 0;
}

int fgpu_get_B(fgpu_reducer* h, double* B) {
  GUARD(h)
  if (!h->reduced || !B) return -1;
  std::vector<double> b;
  if (h->red.get_B(b)) { h->err = "fgpu_get_B: copy failed"; return -3; }
  std::memcpy(B, b.data(), b.size() * sizeof(double));
  return 0;
  END_GUARD(h)
}

int fgpu_get_phi(fgpu_reducer* h, double* phi) {
  GUARD(h)
  if (!h->reduced || !phi) return -1;
  std::vector<double> p;
  if (h->red.get_phi(p)) { h->err = "fgpu_get_phi: copy failed"; return -3; }
  std::memcpy(phi, p.data(), p.size() * sizeof(double));
  return 0;
  END_GUARD(h)
}

int fgpu_get_vgi(const fgpu_reducer* h, double* vgi) {
  if (!h || !h->reduced || !vgi) return -1;
  const fgpu::Results& r = h->red.results();
  if (r.vgi.empty()) return -1;
  std::memcpy(vgi, r.vgi.data(), r.vgi.size() * sizeof(double));
  return 0;
}

int fgpu_autotune(fgpu_reducer* h, int* panel, int* split) {
  GUARD(h)
  if (!h->analyzed) { h->err = "fgpu_autotune: not analyzed"; return -1; }
  int P = 0, K = 0;
  int rc = h->red.autotune(P, K, h->err);
  h->reduced = false;
  if (rc) return rc;
  h->opts.factor_panel = P;
  h->opts.solve_split = K;
  if (panel) *panel = P;
  if (split) *split = K;
  return 0;
  END_GUARD(h)
}

int fgpu_get_fe_mass(const fgpu_reducer* h, double* smass) {
  if (!h || !h->reduced || !smass) return -1;
  *smass = h->red.results().mass_fe;
  return 0;
}

int fgpu_get_ext_nodes(const fgpu_reducer* h, int* nenod, double* xyz, int* medof) {
  if (!h || !h->analyzed || !nenod) return -1;
  const fgpu::Sam& s = h->red.sam();
  const fgpu::Part& p = h->red.part();
  *nenod = s.nenod;
  int k = 0;
  if (medof) medof[0] = 1;
  for (int n = 0; n < s.nnod; n++) {
    if (p.status[n] != 1) continue;
    if (xyz)
      for (int c = 0; c < 3; c++) xyz[3 * k + c] = p.xyz[3 * (size_t)n + c];
    if (medof) medof[k + 1] = medof[k] + s.nndof[n];
    k++;
  }
  return 0;
}

int fgpu_solve_internal(fgpu_reducer* h, int nrhs, const double* R, double* X) {
  GUARD(h)
  if (!h->reduced || nrhs <= 0 || !R || !X) return -1;
  return h->red.solve_kii(nrhs, R, X, h->err);
  END_GUARD(h)
}

int fgpu_get_gravity_internal(fgpu_reducer* h, double* gi) {
  GUARD(h)
  if (!h->reduced || !gi) return -1;
  std::vector<double> g;
  int rc = h->red.gravity_internal(g, h->err);
  if (rc == 0) std::copy(g.begin(), g.end(), gi);
  return rc;
  END_GUARD(h)
}

int fgpu_reduce_loads(fgpu_reducer* h, int nrhs, const double* F, double* Fred) {
  GUARD(h)
  if (!h->reduced || nrhs <= 0 || !F || !Fred) return -1;
  std::vector<double> r;
  int rc = h->red.reduce_loads(nrhs, F, r, h->err);
  if (rc == 0) std::copy(r.begin(), r.end(), Fred);
  return rc;
  END_GUARD(h)
}

int fgpu_reduce_loads_noalloc(fgpu_reducer* h, int nrhs, const double* F, double* Fred) {
  GUARD(h)
  if (!h->reduced || nrhs <= 0 || !F || !Fred) return -1;
  return h->red.reduce_loads_noalloc(nrhs, F, Fred, h->err);
  END_GUARD(h)
}

int fgpu_recover(fgpu_reducer* h, int nrhs, const double* Ub, const double* Fi, double* Ui) {
  GUARD(h)
  if (!h->reduced || nrhs <= 0 || !Ui) return -1;
  return h->red.recover(nrhs, Ub, Fi, Ui, h->err);
  END_GUARD(h)
}

int fgpu_get_times(const fgpu_reducer* h, double* t) {
  if (!h || !h->reduced || !t) return -1;
  const fgpu::PhaseTimes& p = h->red.times();
  double v[13] = {p.elements, p.assembly, p.factor, p.bsolve, p.modes, p.mass, p.gravity, p.total,
                  p.update_kernel_ms, (double)p.update_launches, (double)p.refine_steps, p.analyze_host, p.upload};
  std::memcpy(t, v, sizeof v);
  return 0;
}

int fgpu_get_info(const fgpu_reducer* h, int64_t* info) {
  if (!h || !h->analyzed || !info) return -1;
  const fgpu::Symbolic& y = h->red.sym();
  info[0] = y.nsup; info[1] = y.nlevels; info[2] = y.fpool; info[3] = y.nnzL;
  info[4] = y.factor_flops; info[5] = y.maxfront; info[6] = y.rowptr.empty() ? 0 : y.rowptr.back();
  info[7] = h->red.results().lanczos_vectors; info[8] = y.elem_pool;
  info[9] = (int64_t)h->red.update_kernel_flops();
  return 0;
}

int fgpu_get_csr(fgpu_reducer* h, int* rowptr, int* colind, double* K, double* M, int* perm_to_eq) {
  GUARD(h)
  if (!h->reduced) return -1;
  std::vector<int> rp, ci;
  std::vector<double> k, m;
  if (h->red.get_csr(rp, ci, k, m)) { h->err = "fgpu_get_csr: copy failed"; return -3; }
  if (rowptr) std::memcpy(rowptr, rp.data(), rp.size() * sizeof(int));
  if (colind) std::memcpy(colind, ci.data(), ci.size() * sizeof(int));
  if (K) std::memcpy(K, k.data(), k.size() * sizeof(double));
  if (M) std::memcpy(M, m.data(), m.size() * sizeof(double));
  const std::vector<int>& pe = h->red.sym().perm_to_eq;
  if (perm_to_eq) std::memcpy(perm_to_eq, pe.data(), pe.size() * sizeof(int));
  return 0;
  END_GUARD(h)
}

int fgpu_get_element_matrices(fgpu_reducer* h, double* Kp, double* Mp, int64_t* offsets) {
  GUARD(h)
  if (!h->reduced) return -1;
  std::vector<double> k, m;
  if (h->red.get_element_matrices(k, m)) { h->err = "copy failed"; return -3; }
  if (Kp) std::memcpy(Kp, k.data(), k.size() * sizeof(double));
  if (Mp) std::memcpy(Mp, m.data(), m.size() * sizeof(double));
  const std::vector<int64_t>& o = h->red.sym().elem_off;
  if (offsets) std::memcpy(offsets, o.data(), o.size() * sizeof(int64_t));
  return 0;
  END_GUARD(h)
}

int fgpu_get_supernodes(const fgpu_reducer* h, int* col0, int* ncol, int* nrow, int* parent, int* rows) {
  if (!h || !h->analyzed) return -1;
  const fgpu::Symbolic& y = h->red.sym();
  auto cp = [](int* dst, const std::vector<int>& v) { if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(int)); };
  cp(col0, y.col0); cp(ncol, y.ncol); cp(nrow, y.nrow); cp(parent, y.parent); cp(rows, y.rows);
  return 0;
}

int fgpu_get_perm(const fgpu_reducer* h, int* perm_to_eq) {
  if (!h || !h->analyzed || !perm_to_eq) return -1;
  const std::vector<int>& pe = h->red.sym().perm_to_eq;
  std::memcpy(perm_to_eq, pe.data(), pe.size() * sizeof(int));
  return 0;
}

int fgpu_get_pivots(const fgpu_reducer* h, int* counts, int* node, int* ldof, int* eqn, int* kind, int max) {
  if (!h || !h->analyzed) return -1;
  const fgpu::Results& r = h->red.results();
  if (counts) {
    counts[0] = r.n_negative;
    counts[1] = r.n_true_zero;
    counts[2] = r.n_reduced_zero;
  }
  const fgpu::Sam& s = h->red.sam();
  std::vector<int> dof_of(s.neq, -1);
  for (int d = 0; d < s.ndof; d++)
    if (s.meqn[d] > 0) dof_of[s.meqn[d] - 1] = d;
  const int n = (int)r.singular.size();
  for (int k = 0; k < n && k < max; k++) {
    const int e = r.singular[k].first, d = dof_of[e];
    int i = 0;
    while (i + 1 < s.nnod && s.madof[i + 1] - 1 <= d) i++;
    if (node) node[k] = i + 1;
    if (ldof) ldof[k] = d - (s.madof[i] - 1) + 1;
    if (eqn) eqn[k] = e + 1;
    if (kind) kind[k] = r.singular[k].second;
  }
  return n;
}

int fgpu_get_lanczos(const fgpu_reducer* h, int* nvec, double* resid) {
  if (!h || !h->reduced) return -1;
  if (nvec) *nvec = h->red.results().lanczos_vectors;
  if (resid) *resid = h->red.results().lanczos_resid;
  return 0;
}

}  // extern "C"
