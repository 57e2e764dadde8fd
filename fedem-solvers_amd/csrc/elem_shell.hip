// gfx950 kernel of the 4-noded ANDES shell element: one 64-lane workgroup
// per element runs the phases of shell_andes4.hpp (each phase on all lanes,
// a barrier between phases) and writes the packed lower node-pair 6x6
// blocks of the projected stiffness and of the lumped mass.
#include "kernels.hpp"
#include "shell_andes4.hpp"

namespace fgpu {

__global__ __launch_bounds__(64) void k_andes4(int nel, const int* __restrict__ elist, const int* __restrict__ eptr,
                                               const int* __restrict__ enodes, const double* __restrict__ xyz,
                                               const double* __restrict__ eprop, const int64_t* __restrict__ eoff,
                                               double* __restrict__ Kp, double* __restrict__ Mp,
                                               int* __restrict__ ebad) {
  __shared__ shell::Work w;
  const int lane = threadIdx.x;
  const int e = elist[blockIdx.x];
  const double* pr = eprop + (int64_t)NPROP * e;
  if (lane < 12) {
    const int node = enodes[eptr[e] + lane / 3];
    w.G[lane / 3][lane % 3] = xyz[3 * node + lane % 3];
  } else if (lane == 12) {
    w.E = pr[P_E];
    w.nu = pr[P_NU];
    w.rho = pr[P_RHO];
    w.thk = (pr[P_T1] + pr[P_T1 + 1] + pr[P_T1 + 2] + pr[P_T1 + 3]) * 0.25;
    w.bad = 0;
  }
  __syncthreads();
  for (int k = 0; k < shell::NPHASE; k++) {
    shell::phase(k, lane, w);
    __syncthreads();
  }
  const int64_t off = eoff[e];
  for (int idx = lane; idx < 360; idx += 64) {
    int r, c;
    Kp[off + idx] = shell::projected_entry(idx, w, r, c);
    Mp[off + idx] = r == c ? w.em[r] : 0.0;
  }
  if (lane == 0 && w.bad) ebad[e] = -3;
}

void launch_andes4(hipStream_t st, int nel, const int* elist, const int* eptr, const int* enodes, const double* xyz,
                   const double* eprop, const int64_t* eoff, double* Kp, double* Mp, int* ebad) {
  if (nel > 0) k_andes4<<<nel, 64, 0, st>>>(nel, elist, eptr, enodes, xyz, eprop, eoff, Kp, Mp, ebad);
}

}  // namespace fgpu
