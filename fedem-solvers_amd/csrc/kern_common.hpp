// Shared device-side declarations for the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace fgpu {

// Dense multi-vectors (right-hand sides, B columns, Lanczos blocks) are stored
// row-major: V[row * nrhs + j], so that one matrix row is contiguous.
//
// A factorisation task list entry.  The host builds one list per level and
// panel step; see reducer_gpu.cpp.
struct FTask {
  int sn;      // supernode
  int a, b;    // meaning depends on the kernel (row chunk, tile indices)
  int k0;      // first pivot column of the panel
  int kb;      // trailing update: panel width (<= 128); the trailing block starts at k0 + kb
};

struct SolveTask {
  int sn;
  int r0, rn;  // target rows of the front work buffer
  int k0, kn;  // source rows (K dimension); for reduce tasks: first slot, count
  int j0;      // first right-hand side of this task
  int slot;    // -1: apply to the target directly; >= 0: write a 64x64 partial
  int pad;
};

struct DevSym {
  const int* col0;
  const int* ncol;
  const int* nrow;
  const int64_t* foff;
  const int* rptr;
  const int* rows;
  const int* chptr;
  const int* chlist;
  const int* relptr;
  const int* rel;
  const int64_t* woff;  // per-front offsets of solve work buffers (rows)
  double* dpiv;         // pivots D of the LDL^T factor by pivot column (col0[front] + column)
  // singular pivots: a pivot d_j with |d_j| <= tolf |a_jj| (a_jj the
  // original diagonal, adiag) is replaced by |a_jj| (by amax when a_jj = 0)
  // and its column flagged in sing: 1 a true zero (|d_j| at rounding level
  // of a_jj, or a_jj = 0), 2 a pivot reduced to near zero
  const double* adiag;
  const double* amax;
  int* sing;
  double tolf;
};

}  // namespace fgpu
