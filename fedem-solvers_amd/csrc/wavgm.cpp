// Weighted-average-motion constraint elements (WAVGM, Nastran RBE3): the
// constraint equations that tie dependent DOFs of a reference node to the
// DOFs of its independent nodes.
//
// Reference: wavgmConstrEqn (wavgmModule.f90:58-367) with the reducer's
// default switches (reducerInterface.C:99,121-124: tolWAVGM 1e-4, useRotCpl
// and useEccCpl on, useUnWeightedCG and useOldWAVGM off), the element loop
// of MultiPointConstraints (rigidModule.f90:89-338) and defineConstraintEqn
// (rigidModule.f90:1575-1674).
//
// Every dependent DOF follows
//   translation i:  the weighted mean of the independent translations i,
//                   plus (theta x e)_i with e the reference node's offset
//                   from the weight set's centre and theta the rotation
//                   fitted to the independent translations;
//   rotation i:     the weighted mean of the independent rotations i (when
//                   the element gives rotational weights) plus the fitted
//                   rotation about axis i.
// The fit about axis r from the translations p, q (r, p, q cyclic) is
//   theta_r = sum w (dX_p u_q - dX_q u_p) / sum w (dX_p^2 + dX_q^2),
// the decoupled (diagonal-inertia) least-squares fit; with uniform weights
// the reference takes w = 1 in the numerator and nM * sum (dX_p^2 + dX_q^2)
// as the denominator (wavgmModule.f90:132,150,203), kept here.  Each term
// uses the centre and tolerances of the weight set of the DOF it couples to.
#include <cmath>
#include <limits>
#include <map>
#include "fgpu.hpp"

namespace fgpu {

namespace {

constexpr double kEps = std::numeric_limits<double>::epsilon();  // epsDiv0_p
constexpr double kTolX = 1.0e-4;                                 // tolWAVGM

class WavgmElement {
 public:
  WavgmElement(const Part& p, int w) : p_(p) {
    ref_ = p.wv_ref[w];
    nodes_.assign(p.wv_nodes.begin() + p.wv_nptr[w], p.wv_nodes.begin() + p.wv_nptr[w + 1]);
    nM_ = (int)nodes_.size();
    const int nw = p.wv_wptr[w + 1] - p.wv_wptr[w];
    if (nw == 0) {  // no properties: the reference's defaults (rigidModule.f90:195-200)
      refc_ = 123456;
      for (int c = 0; c < 6; c++) indc_[c] = c < 3 ? -1 : 0;
    } else {
      refc_ = p.wv_refc[w];
      for (int c = 0; c < 6; c++) indc_[c] = p.wv_indc[6 * w + c];
      w_.assign(p.wv_w.begin() + p.wv_wptr[w], p.wv_w.begin() + p.wv_wptr[w + 1]);
    }
  }

  int refc() const { return refc_; }
  int ref() const { return ref_; }
  const std::vector<int>& nodes() const { return nodes_; }

  // weights of set ic (1-based start), checked against the weight count
  bool valid() const {
    for (int c = 0; c < 6; c++)
      if (indc_[c] > 0 && indc_[c] - 1 + nM_ > (int)w_.size()) return false;
    return true;
  }

  // coefficients om[6*n + c] of dependent DOF ld (0..5) on DOF c of node n
  void omega(int ld, std::vector<double>& om) {
    om.assign(6 * (size_t)nM_, 0.0);
    const int i = ld % 3, j = (i + 1) % 3, k = (i + 2) % 3;
    // direct coupling to the same component
    if (indc_[ld] > 0) {
      const double* w = wset(ld);
      double s = 0.0;
      for (int n = 0; n < nM_; n++) s += w[n];
      if (s > kEps)
        for (int n = 0; n < nM_; n++) om[6 * n + ld] += w[n] / s;
    } else if (indc_[ld] < 0) {
      for (int n = 0; n < nM_; n++) om[6 * n + ld] += 1.0 / nM_;
    }
    if (ld >= 3) {
      // fitted rotation about axis i from the translations j (p) and k (q)
      double tolwm = 0.0;
      fit_term(i, j, 1.0, nullptr, tolwm, true, om);
      fit_term(i, k, 1.0, nullptr, tolwm, true, om);
      return;
    }
    // u_i += (theta x e)_i = theta_j e_k - theta_k e_j: first the rotation
    // about k through e_j, then the rotation about j through e_k
    double tolwm = 0.0;
    const int rot[2] = {k, j}, ecc[2] = {j, k};
    const double sgn[2] = {-1.0, 1.0};
    for (int b = 0; b < 2; b++) {
      const int r = rot[b], a = ecc[b];
      const int p = (r + 1) % 3, q = (r + 2) % 3;  // the translations of the fit
      // the term on u_i (i is p in the first block, q in the second) sets
      // the denominator tolerance; the term on the other translation reuses
      // it in the first block (wavgmModule.f90:234) and sets it in the second
      const int first = i, other = (i == p) ? q : p;
      fit_term(r, first, sgn[b], &a, tolwm, true, om);
      fit_term(r, other, sgn[b], &a, tolwm, b == 1, om);
      // mean rotation r of the independent nodes times the eccentricity
      if (indc_[3 + r] > 0) {
        const Centre& C = centre(indc_[3 + r]);
        const double e = C.dX[a][0];
        if (std::fabs(e) > C.tol[a]) {
          const double* w = wset(3 + r);
          double s = 0.0;
          for (int n = 0; n < nM_; n++) s += w[n];
          if (s > kEps)
            for (int n = 0; n < nM_; n++) om[6 * n + 3 + r] += sgn[b] * e / s * w[n];
        }
      }
    }
  }

 private:
  struct Centre {
    std::vector<double> dX[3];  // offsets from the centre: [0] reference node, [1..nM] independent nodes
    double tol[3];
  };

  const double* wset(int c) const { return w_.data() + indc_[c] - 1; }

  // centre and tolerances of weight set ic (computePointCoords,
  // wavgmModule.f90:381-418); ic < 0 is the uniform set
  const Centre& centre(int ic) {
    auto it = cache_.find(ic);
    if (it != cache_.end()) return it->second;
    Centre C;
    double x0[3] = {0.0, 0.0, 0.0}, sw = 0.0;
    for (int n = 0; n < nM_; n++) {
      const double wn = ic > 0 ? w_[ic - 1 + n] : 1.0;
      for (int a = 0; a < 3; a++) x0[a] += p_.xyz[3 * nodes_[n] + a] * wn;
      sw += wn;
    }
    for (int a = 0; a < 3; a++) {
      x0[a] /= sw;
      C.dX[a].resize(nM_ + 1);
      C.dX[a][0] = p_.xyz[3 * ref_ + a] - x0[a];
      for (int n = 0; n < nM_; n++) C.dX[a][n + 1] = p_.xyz[3 * nodes_[n] + a] - x0[a];
      // tolerance over the reference node and the first nM-1 independent
      // nodes, as the reference takes it (wavgmModule.f90:408-411)
      double lo = -std::numeric_limits<double>::infinity(), hi = lo;
      for (int n = 0; n < nM_; n++) {
        lo = std::max(lo, -C.dX[a][n] * kTolX);
        hi = std::max(hi, C.dX[a][n] * kTolX);
      }
      C.tol[a] = std::max(std::max(lo, hi), kEps);
    }
    return cache_.emplace(ic, std::move(C)).first->second;
  }

  // Adds the fitted-rotation term on translation t of the rotation about
  // axis r, scaled by sgn * e_a (e_a the reference node's offset along a) or
  // by 1 for a rotational dependent DOF (a == nullptr).  The weight set is
  // the one of translation t.  set_tol: recompute the denominator tolerance
  // from this set before the check, else use the one carried in tolwm.
  void fit_term(int r, int t, double sgn, const int* a, double& tolwm, bool set_tol, std::vector<double>& om) {
    if (indc_[t] == 0) return;
    const int p = (r + 1) % 3, q = (r + 2) % 3;
    const Centre& C = centre(indc_[t]);
    double scale = sgn;
    if (a) {
      const double e = C.dX[*a][0];
      if (!(std::fabs(e) > C.tol[*a])) return;
      scale *= e;
    }
    // theta_r = sum w (dX_p u_q - dX_q u_p) / S: on u_q the numerator is
    // w dX_p, on u_p it is -w dX_q
    const int num = (t == q) ? p : q;
    const double tsign = (t == q) ? 1.0 : -1.0;
    const bool uniform = indc_[t] < 0;
    const double* w = uniform ? nullptr : wset(t);
    double S = 0.0;
    for (int n = 1; n <= nM_; n++) {
      const double v = C.dX[p][n] * C.dX[p][n] + C.dX[q][n] * C.dX[q][n];
      S += uniform ? v : w[n - 1] * v;
    }
    if (uniform) S *= nM_;
    if (set_tol) tolwm = C.tol[p] * C.tol[p] + C.tol[q] * C.tol[q];
    if (!(S > tolwm)) return;
    for (int n = 0; n < nM_; n++) {
      const double wn = uniform ? 1.0 : w[n];
      om[6 * n + t] += scale * tsign * wn * C.dX[num][n + 1] / S;
    }
  }

  const Part& p_;
  int ref_ = 0, nM_ = 0, refc_ = 0, indc_[6] = {0, 0, 0, 0, 0, 0};
  std::vector<int> nodes_;
  std::vector<double> w_;
  std::map<int, Centre> cache_;
};

}  // namespace

int wavgm_equations(const Part& p, const std::vector<int>& madof, const std::vector<int>& nndof,
                    std::vector<int>& msc, std::vector<WavgmEquation>& eqs, std::string& err) {
  eqs.clear();
  const int nw = (int)p.wv_ref.size();
  std::vector<double> om;
  for (int w = 0; w < nw; w++) {
    WavgmElement el(p, w);
    const std::string tag = "WAVGM element " + std::to_string(w + 1);
    if (el.ref() < 0 || el.ref() >= p.nnod) { err = tag + ": bad reference node"; return -1; }
    for (int n : el.nodes())
      if (n < 0 || n >= p.nnod || n == el.ref()) { err = tag + ": bad independent node"; return -1; }
    if (el.nodes().empty()) { err = tag + ": no independent nodes"; return -1; }
    if (!el.valid()) { err = tag + ": weight index beyond the element's weights"; return -1; }
    // dependent components (procDofComp, rigidModule.f90:1694-1730): the
    // digits 1..6 of refC, all six for refC = 0
    bool comp[6] = {false, false, false, false, false, false};
    if (el.refc() < 0) { err = tag + ": explicit constraint equations (negative refC) are not supported"; return -1; }
    if (el.refc() == 0) {
      for (bool& c : comp) c = true;
    } else {
      int x = el.refc();
      for (int div = 100000; div > 0; div /= 10) {
        const int dg = x / div;
        if (dg >= 1 && dg <= 6) comp[dg - 1] = true;
        else if (dg != 0) { err = tag + ": invalid dependent components " + std::to_string(el.refc()); return -1; }
        x -= dg * div;
      }
    }
    for (int ld = 0; ld < 6; ld++) {
      if (!comp[ld] || ld >= nndof[el.ref()]) continue;  // no such DOF at the reference node
      const int dep = madof[el.ref()] - 1 + ld;
      if (msc[dep] != 1) continue;  // external, fixed or dependent already
      el.omega(ld, om);
      // defineConstraintEqn: free DOFs enter the equation; a fixed DOF first
      // in line fixes the dependent DOF until a free one follows; DOFs that
      // are dependent already are left out
      WavgmEquation q;
      q.dep = dep;
      bool first = true;
      for (size_t a = 0; a < el.nodes().size(); a++) {
        const int n = el.nodes()[a];
        for (int c = 0; c < nndof[n]; c++) {
          const double v = om[6 * a + c];
          if (!(std::fabs(v) > kEps)) continue;
          const int d = madof[n] - 1 + c;
          if (msc[d] > 0) {
            q.node.push_back(n);
            q.dof.push_back(d);
            q.coef.push_back(v);
            first = false;
            msc[dep] = -3;
          } else if (msc[d] == 0 && first) {
            msc[dep] = 0;
          }
        }
      }
      if (!q.dof.empty()) eqs.push_back(std::move(q));
    }
  }
  return 0;
}

}  // namespace fgpu
