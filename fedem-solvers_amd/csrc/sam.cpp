// SAM DOF numbering for a part: MADOF, MSC, MEQN, MEQN1, MEQN2, dofPosIn2,
// and the constraint equations of the dependent DOFs.
//
// Reference: samReducerModule.f90:227-569 (initiateSAM, nodal arrays from
// FFl, constraint elements), asmExtensionModule.f90:365-377 (SPRPRP turns unconnected internal
// DOFs into fixed ones), asmExtensionModule.f90:910-980 (csCreateEqPartition)
// and samModule.f90:892-991 (invertEqPartition).
//
// MEQN itself is produced by SAM's SPRPRP, whose source is not part of the
// reference tree.  The numbering used here follows the one property the
// reference code relies on (csCreateEqPartition expects the status-2
// equations after all status-1 ones): status-1 DOFs are numbered 1..ndof1 in
// DOF order, status-2 DOFs ndof1+1..neq in DOF order, fixed DOFs get 0.
#include <algorithm>
#include "fgpu.hpp"

namespace fgpu {

int build_sam(const Part& p, Sam& s, std::string& err) {
  s = Sam();
  s.nnod = p.nnod;
  // DOFs per node: 6 for external nodes (superelement triads carry 6 DOFs) and
  // for nodes of beams/shells, 3 for nodes of solids only.
  std::vector<int> maxdn(p.nnod, 0);
  for (int e = 0; e < p.nel; e++) {
    int d = elem_ndn(p.etype[e]);
    for (int k = p.eptr[e]; k < p.eptr[e + 1]; k++) {
      int n = p.enodes[k];
      if (n < 0 || n >= p.nnod) { err = "element node index out of range"; return -1; }
      maxdn[n] = std::max(maxdn[n], d);
    }
  }
  s.nndof.resize(p.nnod);
  s.madof.resize(p.nnod + 1);
  s.madof[0] = 1;
  for (int i = 0; i < p.nnod; i++) {
    int nd = (p.status[i] == 1 || maxdn[i] == 0) ? 6 : maxdn[i];
    s.nndof[i] = nd;
    s.madof[i + 1] = s.madof[i] + nd;
  }
  s.ndof = s.madof[p.nnod] - 1;
  // nodal status codes: 0 fixed, 1 internal, 2 external
  s.msc.assign(s.ndof, 0);
  for (int i = 0; i < p.nnod; i++) {
    const int st = p.status[i];
    if (st == 1) { s.nenod++; s.ndim_ext += s.nndof[i]; }
    for (int j = 0; j < s.nndof[i]; j++)
      s.msc[s.madof[i] - 1 + j] = (st < 0 && ((-st) >> j) & 1) ? 0 : st == 1 ? 2 : 1;
  }
  // constraint elements, WAVGM (RBE3) first on the nodal status codes, then
  // the rigid elements (samReducerModule.f90:385-459)
  std::vector<WavgmEquation> weq;
  if (wavgm_equations(p, s.madof, s.nndof, s.msc, weq, err)) return -1;
  std::vector<int> rgd_of(s.ndof, -1);  // independent node of a rigid element's dependent DOF
  std::vector<char> indep(p.nnod, 0);   // independent nodes of constraint elements
  const int nrgd = (int)p.rgd_indep.size();
  for (int r = 0; r < nrgd; r++) {
    int g = p.rgd_indep[r];
    if (g < 0 || g >= p.nnod) { err = "rigid element " + std::to_string(r + 1) + ": bad independent node"; return -1; }
    indep[g] = 1;
  }
  for (int r = 0; r < nrgd; r++) {
    int g = p.rgd_indep[r];
    for (int k = p.rgd_ptr[r]; k < p.rgd_ptr[r + 1]; k++) {
      int u = p.rgd_dep[k];
      if (u < 0 || u >= p.nnod || u == g) { err = "rigid element " + std::to_string(r + 1) + ": bad dependent node"; return -1; }
      if (indep[u]) { err = "rigid element chains (a dependent node that is independent elsewhere) are not supported"; return -1; }
      for (int j = 0; j < s.nndof[u]; j++)
        if ((p.rgd_mask[r] >> j) & 1) {
          int d = s.madof[u] - 1 + j;
          if (rgd_of[d] >= 0 || s.msc[d] < 0) { err = "DOF constrained by two constraint elements"; return -1; }
          rgd_of[d] = g;
        }
    }
  }
  for (int w = 0; w < (int)p.wv_ref.size(); w++)
    for (int k = p.wv_nptr[w]; k < p.wv_nptr[w + 1]; k++) indep[p.wv_nodes[k]] = 1;
  for (int i = 0; i < p.nnod; i++)
    for (int j = 0; j < s.nndof[i]; j++) {
      const int idof = s.madof[i] - 1 + j;
      int m = s.msc[idof];
      // unconnected internal DOF -> fixed (SPRPRP behaviour); independent
      // nodes of constraint elements are connected through the constraint
      if (m == 1 && j >= maxdn[i] && !indep[i]) m = 0;
      if (m == -3 || rgd_of[idof] >= 0) m = -1;
      s.msc[idof] = m;
    }
  // equation terms of the dependent DOFs: a rigid element's u_g + theta_g x
  // (x_u - x_g) (translations) or theta_g (rotations) over the equation DOFs
  // of its independent node g, a WAVGM element's weighted terms
  std::vector<int> weq_of(s.ndof, -1);
  for (int q = 0; q < (int)weq.size(); q++) {
    for (int d : weq[q].dof)
      if (s.msc[d] <= 0) {
        err = "constraint chains (an independent DOF that is dependent elsewhere) are not supported";
        return -1;
      }
    weq_of[weq[q].dep] = q;
  }
  s.ceq_ptr.assign(s.ndof + 1, 0);
  for (int u = 0; u < p.nnod; u++)
    for (int j = 0; j < s.nndof[u]; j++) {
      const int d = s.madof[u] - 1 + j;
      s.ceq_ptr[d + 1] = s.ceq_ptr[d];
      if (s.msc[d] != -1) continue;
      auto add = [&](int n, int dof, double c) {
        s.ceq_node.push_back(n);
        s.ceq_dof.push_back(dof);
        s.ceq_coef.push_back(c);
        s.ceq_ptr[d + 1]++;
      };
      if (weq_of[d] >= 0) {
        const WavgmEquation& q = weq[weq_of[d]];
        for (size_t k = 0; k < q.dof.size(); k++) add(q.node[k], q.dof[k], q.coef[k]);
        continue;
      }
      const int g = rgd_of[d];
      for (int k = 0; k < s.nndof[g]; k++)
        if (s.msc[s.madof[g] - 1 + k] < 0) {
          err = "constraint chains (an independent DOF that is dependent elsewhere) are not supported";
          return -1;
        }
      double r[3];
      for (int c = 0; c < 3; c++) r[c] = p.xyz[3 * u + c] - p.xyz[3 * g + c];
      int cj[3] = {j, -1, -1};
      double cc[3] = {1.0, 0.0, 0.0};
      if (j == 0) { cj[1] = 4; cc[1] = r[2]; cj[2] = 5; cc[2] = -r[1]; }
      if (j == 1) { cj[1] = 5; cc[1] = r[0]; cj[2] = 3; cc[2] = -r[2]; }
      if (j == 2) { cj[1] = 3; cc[1] = r[1]; cj[2] = 4; cc[2] = -r[0]; }
      for (int k = 0; k < 3; k++) {
        if (cj[k] < 0 || cj[k] >= s.nndof[g] || cc[k] == 0.0) continue;
        const int dg = s.madof[g] - 1 + cj[k];
        if (s.msc[dg] > 0) add(g, dg, cc[k]);
      }
    }
  s.meqn.assign(s.ndof, 0);
  int ieq = 0;
  for (int d = 0; d < s.ndof; d++)
    if (s.msc[d] == 1) s.meqn[d] = ++ieq;
  s.ndof1 = ieq;
  for (int d = 0; d < s.ndof; d++)
    if (s.msc[d] == 2) s.meqn[d] = ++ieq;
  s.neq = ieq;
  s.ndof2 = s.neq - s.ndof1;
  // csCreateEqPartition: MEQN1/MEQN2 list the status-1/2 equations ascending
  std::vector<int> eqst(s.neq + 1, 0);
  for (int d = 0; d < s.ndof; d++)
    if (s.msc[d] > 0) eqst[s.meqn[d]] = s.msc[d];
  for (int q = 1; q <= s.neq; q++) {
    if (eqst[q] == 1) s.meqn1.push_back(q);
    else s.meqn2.push_back(q);
  }
  // invertEqPartition: dofPosIn2(i2) = findloc(MEQN2, MEQN(idof)) over the
  // status-2 DOFs in DOF order
  std::vector<int> pos2(s.neq + 1, 0);
  for (int k = 0; k < s.ndof2; k++) pos2[s.meqn2[k]] = k + 1;
  for (int d = 0; d < s.ndof; d++)
    if (s.msc[d] == 2) s.dofPosIn2.push_back(pos2[s.meqn[d]]);
  if ((int)s.meqn1.size() != s.ndof1 || (int)s.meqn2.size() != s.ndof2) {
    err = "equation partition mismatch";
    return -1;
  }
  return 0;
}

}  // namespace fgpu
