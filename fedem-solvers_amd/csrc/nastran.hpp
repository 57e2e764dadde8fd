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

struct NastranPart {
  Part part;
  std::vector<int> node_ids;
  std::vector<Rbe2> rbe2;
};

int read_nastran(const std::string& path, NastranPart& out, std::string& err);

}  // namespace fgpu
