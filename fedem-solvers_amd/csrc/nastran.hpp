#pragma once
#include <string>
#include <vector>
#include "fgpu.hpp"

namespace fgpu {

// Rigid element (RBE2): the listed DOFs of the dependent nodes follow the
// rigid-body motion of the independent node.
struct Rbe2 {
  int indep = -1;
  int dofs = 0;  // bitmask over the six local DOFs
  std::vector<int> dep;
};

// Weighted-average-motion element (RBE3) in the WAVGM form the reducer
// takes: dependent components refc of the reference node, independent nodes,
// per component the 1-based start of its weights in w (0: no coupling).
struct Rbe3 {
  int ref = -1;
  int refc = 0;
  std::vector<int> nodes;
  int indc[6] = {0, 0, 0, 0, 0, 0};
  std::vector<double> w;
};

// Nodal load of a load case (FORCE: DOFs 0-2, MOMENT: DOFs 3-5)
struct NodalLoad {
  int lc = 0;    // load set id
  int node = 0;  // 0-based node index
  int dof = 0;   // 0..5
  double value = 0.0;
};

struct NastranPart {
  Part part;
  std::vector<int> node_ids;
  std::vector<Rbe2> rbe2;
  std::vector<Rbe3> rbe3;
  std::vector<NodalLoad> loads;
};

int read_nastran(const std::string& path, NastranPart& out, std::string& err);

}  // namespace fgpu
