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
  std::vector<NodalLoad> loads;
};

int read_nastran(const std::string& path, NastranPart& out, std::string& err);

}  // namespace fgpu
