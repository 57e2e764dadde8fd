// Internal declarations of the MI355X part reducer (host side).
//
// Index conventions used throughout the library:
//   * node / element / DOF indices are 0-based internally; the C ABI takes and
//     returns the reference's 1-based SAM arrays (MADOF, MEQN, ...).
//   * "perm space": the order the device works in.  Positions
//     0 .. ndof1-1 are the internal (status-1) equations in elimination order
//     (nested dissection, children before parents); positions
//     ndof1 .. neq-1 are the external (status-2) equations in MEQN2 order, so
//     the trailing block of the factorisation is the Schur complement.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace fgpu {

// Element type codes are the reference's MELCON values (inadd.f90:108-181).
enum ElemType : int { BEAM2 = 11, ANDES4 = 24, HEX8 = 44 };

// Per-element property record (doubles), shared by all element types.
enum PropSlot : int {
  P_E = 0, P_NU = 1, P_RHO = 2, P_G = 3,
  P_T1 = 4, P_T2 = 5, P_T3 = 6, P_T4 = 7,        // shell thickness per node
  P_A = 4, P_IY = 5, P_IZ = 6, P_IT = 7,         // beam section
  P_DX = 8, P_DY = 9, P_DZ = 10,                 // beam orientation vector v (local x-y plane, CBAR)
  P_KY = 11, P_KZ = 12,                          // beam shear reduction factors
  P_NSM = 13,                                    // non-structural mass
  NPROP = 16
};

inline int elem_nodes(int t) { return t == HEX8 ? 8 : t == ANDES4 ? 4 : 2; }
inline int elem_ndn(int t) { return t == HEX8 ? 3 : 6; }
// Packed element matrix: node-pair blocks (a >= b) of ndn x ndn, row-major
// within a block, block index a*(a+1)/2 + b.
inline int elem_psize(int t) {
  int n = elem_nodes(t), d = elem_ndn(t);
  return n * (n + 1) / 2 * d * d;
}

struct Part {
  int nnod = 0, nel = 0;
  std::vector<double> xyz;     // 3*nnod
  std::vector<int> status;     // per node: 1 external, 0 internal, <0 fixed bitmask
  std::vector<int> etype;      // nel
  std::vector<int> eptr;       // nel+1
  std::vector<int> enodes;     // 0-based node indices
  std::vector<double> eprop;   // NPROP*nel
  // rigid elements (RBE2): the DOFs in rgd_mask of every dependent node
  // follow the rigid-body motion of the independent node
  std::vector<int> rgd_indep;  // per rigid element, 0-based node
  std::vector<int> rgd_mask;   // bitmask over the six local DOFs
  std::vector<int> rgd_ptr;    // nrgd+1 pointers into rgd_dep (0-based)
  std::vector<int> rgd_dep;    // 0-based dependent nodes
  // weighted-average-motion elements (WAVGM / RBE3): element w ties the
  // components refc (digits 1..6, 0 = all) of reference node wv_ref[w] to
  // its independent nodes wv_nodes[wv_nptr[w] .. wv_nptr[w+1]); wv_indc[6w+c]
  // is the 1-based start of component c's weights within the element's
  // weights wv_w[wv_wptr[w] .. wv_wptr[w+1]) (< 0 uniform, 0 no coupling);
  // an element without weights takes refc 123456, uniform translations
  std::vector<int> wv_ref, wv_refc, wv_nptr, wv_nodes, wv_indc, wv_wptr;
  std::vector<double> wv_w;
};

// One constraint equation of a WAVGM element: u(dep) = sum coef[k] u(dof[k])
struct WavgmEquation {
  int dep = -1;                  // 0-based dependent DOF
  std::vector<int> node, dof;    // independent node and 0-based DOF of each term
  std::vector<double> coef;
};

// SAM arrays (samModule.f90:27-64), 0-based storage of 1-based values where
// the reference stores equation numbers (meqn* hold 1-based equation numbers).
struct Sam {
  int nnod = 0, ndof = 0, neq = 0, ndof1 = 0, ndof2 = 0;
  std::vector<int> madof;      // nnod+1, 1-based accumulated DOF pointers
  std::vector<int> msc;        // ndof, 0 fixed / 1 internal / 2 external / -1 dependent (constraint element)
  std::vector<int> meqn;       // ndof, 0 for fixed, else 1-based equation
  std::vector<int> meqn1, meqn2, dofPosIn2;
  std::vector<int> nndof;      // per node
  // constraint equations of the dependent DOFs (msc -1): DOF d follows
  // sum ceq_coef[k] u(ceq_dof[k]) over k in [ceq_ptr[d], ceq_ptr[d+1]),
  // ceq_node[k] being the node of DOF ceq_dof[k]; every ceq_dof[k] has an
  // equation
  std::vector<int> ceq_ptr, ceq_node, ceq_dof;
  std::vector<double> ceq_coef;
  int nenod = 0;               // number of external nodes
  int ndim_ext = 0;            // sum of DOFs of external nodes
};

int build_sam(const Part& p, Sam& s, std::string& err);

// Equations of the part's WAVGM elements on the nodal status codes msc
// (1 internal, 2 external, 0 fixed), in element order; msc of a dependent
// DOF becomes -3 (or 0 when its coupled DOFs are fixed).
int wavgm_equations(const Part& p, const std::vector<int>& madof, const std::vector<int>& nndof,
                    std::vector<int>& msc, std::vector<WavgmEquation>& eqs, std::string& err);

// ---------------------------------------------------------------- symbolic
struct Symbolic {
  int neq = 0, ndof1 = 0, ndof2 = 0;
  std::vector<int> eq_to_perm;        // 0-based equation -> perm index
  std::vector<int> perm_to_eq;
  std::vector<int64_t> elem_off;      // offset of each element's packed matrix
  int64_t elem_pool = 0;              // total doubles of packed element matrices

  // CSR row pointers (full symmetric pattern in perm space); the column
  // indices, the pull-assembly contribution lists and the front map are
  // built on the device from the arrays below (structure.hip)
  std::vector<int> rowptr;
  std::vector<int> nfirst, neqd;      // per node: first perm equation (-1: none), equation count
  std::vector<int> nbptr, nbl;        // per node: neighbour nodes through elements, itself included
  std::vector<int> rnode;             // per perm row: its node
  std::vector<int> colsn;             // per perm column: the front that eliminates it (Schur front last)
  std::vector<int> drank;             // per DOF: rank among its node's equation DOFs (-1: none)
  int64_t ncontrib = 0;               // element-matrix contributions to the CSR entries
  bool has_rigid = false;             // contributions carry constraint coefficients

  // supernodes in postorder; the last one (index nsup-1) is the Schur front
  int nsup = 0;
  std::vector<int> col0, ncol, nrow, parent, level;
  std::vector<int> rptr;              // nsup+1 -> rows (perm indices)
  std::vector<int> rows;
  std::vector<int64_t> foff;          // front pool offsets (nrow^2 each)
  int64_t fpool = 0;
  std::vector<int> chptr, chlist;     // children lists
  std::vector<int> relptr, rel;       // per supernode: update rows -> parent front row
  int nlevels = 0;                    // levels of factored supernodes
  std::vector<int> lvptr, lvsup;      // supernodes of each level
  int64_t factor_flops = 0, nnzL = 0; // algorithmic factor work and nnz(L)
  int maxfront = 0;
};

struct SymbolicOptions {
  int leaf_nodes = 24;  // stop nested dissection below this many nodes
};

int build_symbolic(const Part& p, const Sam& s, const SymbolicOptions& o,
                   Symbolic& sym, std::string& err);

}  // namespace fgpu
