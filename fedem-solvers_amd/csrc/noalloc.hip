// The reduced load vectors without any device allocation: the work buffers
// of fgpu_reduce_loads_noalloc (the permuted load vectors, the partial sums
// of the B^T F and Phi^T F products and their results) live in one
// file-scope device array, `scratch`, that every GPU holds from library
// load until the process ends.  Host code never uses the array's name as an
// address: the per-device address comes from hipGetSymbolAddress at the
// first call on that device.  Calls on one GPU share the array, so calls
// that overlap on one GPU race; the entry point is synchronous.
#include <mutex>
#include <string>
#include "kernels.hpp"

namespace fgpu {

constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(32) << 20;  // 32 MiB per device

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

// Device address of this GPU's scratch array (looked up once per device,
// with that device current).
int noalloc_scratch(double** ptr, size_t* bytes, std::string& err) {
  static std::mutex mu;
  static void* addr[64] = {};
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess || dev < 0 || dev >= 64) {
    err = std::string("no current device for the scratch array: ") + hipGetErrorString(e);
    return -3;
  }
  std::lock_guard<std::mutex> g(mu);
  if (!addr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      err = std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e);
      return -3;
    }
    addr[dev] = p;
  }
  *ptr = static_cast<double*>(addr[dev]);
  *bytes = NOALLOC_SCRATCH_BYTES;
  return 0;
}

// Fp[r * nrhs + c] = F[p2e[r] + c * neq]: column-major load vectors in
// equation order to row-major permuted rows
__global__ void k_perm_rows(int neq, int nrhs, const int* __restrict__ p2e, const double* __restrict__ F,
                            double* __restrict__ Fp) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (int64_t)neq * nrhs) return;
  const int r = (int)(g / nrhs), c = (int)(g % nrhs);
  Fp[g] = F[p2e[r] + (int64_t)c * neq];
}

// Fred (nd x nrhs, column-major) = [r2 rows in dofPosIn2 order; r1; 0]
__global__ void k_scatter_reduced(int nd, int n2, int ngen, int nrhs, const int* __restrict__ pos2,
                                  const double* __restrict__ r2, const double* __restrict__ r1,
                                  double* __restrict__ Fred) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (int64_t)nd * nrhs) return;
  const int i = (int)(g % nd), c = (int)(g / nd);
  double v = 0.0;
  if (i < n2) v = r2[(int64_t)(pos2[i] - 1) * nrhs + c];
  else if (i < n2 + ngen) v = r1[(int64_t)(i - n2) * nrhs + c];
  Fred[g] = v;
}

void launch_perm_rows(hipStream_t st, int neq, int nrhs, const int* p2e, const double* F, double* Fp) {
  const int64_t n = (int64_t)neq * nrhs;
  if (n > 0) k_perm_rows<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(neq, nrhs, p2e, F, Fp);
}

void launch_scatter_reduced(hipStream_t st, int nd, int n2, int ngen, int nrhs, const int* pos2, const double* r2,
                            const double* r1, double* Fred) {
  const int64_t n = (int64_t)nd * nrhs;
  if (n > 0) k_scatter_reduced<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(nd, n2, ngen, nrhs, pos2, r2, r1, Fred);
}

}  // namespace fgpu
