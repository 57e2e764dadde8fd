// Device-side part reduction (Guyan + fixed-interface component modes).
#pragma once
#include <hip/hip_runtime.h>
#include <atomic>
#include <map>
#include <memory>
#include "fgpu.hpp"
#include "kern_common.hpp"

namespace fgpu {

// Device buffers that reducers allocate on the calling thread from now on
// are counted in *a (bytes held; nullptr: not counted).
void set_alloc_account(std::atomic<int64_t>* a);

struct ReduceOptions {
  int ngen = 0;            // component modes kept in the reduced matrices
  int neval = 0;           // eigenpairs computed (max(neval, ngen) when ngen > 0)
  double eig_shift = 0.0;  // eigenshift
  double tol_eig = 1e-11;  // relative residual ||Op phi - theta phi||_M / theta
  int lanczos_block = 16;
  int lanczos_maxvec = 0;  // 0: min(ndof1, 10*nev + 40)
  bool want_vgi = false;   // K_ii^{-1} g_i (dispfile)
  bool no_mass = false;
  bool time_updates = false;  // bracket every trailing-update launch with HIP events
  bool use_graphs = true;     // replay the factorisation and Lanczos solve steps as HIP graphs
  // Mixed precision: FP32 fronts of the equilibrated K (FP32 MFMA trailing
  // and Schur updates, FP64 diagonal-block factorisation), every solve with
  // K_ii refined against the FP64 matrix until the correction falls below
  // ir_tol relative (at most ir_maxit steps; no convergence is an error).
  int factor_panel = 128;   // pivot columns per trailing update (64..256, multiple of 64)
  // top levels: trailing tiles past the next panel on a side stream (off by
  // default: on the 48x24x24 block the side updates take the CUs that the
  // critical-path diagonal launches need, 12.8 -> 13.1 ms)
  bool factor_lookahead = false;
  int solve_split = 128;    // longest K range of one solve GEMM task (longer ones are split and reduced)
  bool mixed = false;
  // singular pivots (solExtensionModule.f90:716-760; reducerInterface.C:100,108):
  // a pivot below tol_factorize |a_jj| is replaced in any case; with
  // sing_handler 0 any such pivot is an error, with 1 only those reduced to
  // near zero by the elimination (true zeros are suppressed), with > 1 none
  double tol_factorize = 1e-12;
  int sing_handler = 1;
  double ir_tol = 1e-14;
  int ir_maxit = 40;
  uint64_t seed = 0x5eed;  // Lanczos start block
  SymbolicOptions sym;
};

// Sets a numeric option by its fgpu_set_option name; -1 for an unknown name.
int set_reduce_option(ReduceOptions& o, const std::string& name, double value);

struct PhaseTimes {  // milliseconds, measured with HIP events on the work stream
  double elements = 0, assembly = 0, factor = 0, bsolve = 0, modes = 0, mass = 0, gravity = 0, total = 0;
  double update_kernel_ms = 0;   // sum of trailing-update (MFMA) launches, when timed
  int refine_steps = 0;          // mixed precision: refinement steps of the B solve
  int64_t update_launches = 0;
  double analyze_host = 0;       // host analysis of the last analyze(): SAM numbering, ordering, fronts
  double upload = 0;             // last upload: part and structure transfer, device structure build, schedules
};

struct Results {
  int ndim = 0, ndof1 = 0, ndof2 = 0, ngen = 0, nev = 0;
  std::vector<double> SK, SM, G;   // ndim x ndim, ndim x ndim, ndim x 3 (column-major)
  std::vector<double> eval;        // nev angular frequencies sqrt(lambda) (EIGVAL)
  std::vector<double> lambda;      // nev eigenvalues
  std::vector<double> vgi;         // ndof1 x 3 (meqn1 order)
  int lanczos_vectors = 0;
  double lanczos_resid = 0;        // worst relative Ritz residual at the last check
  double mass_fe = 0;              // sum of element masses
  // pivots of the K_ii factor: negative ones, replaced true zeros and
  // replaced reduced-to-zero ones; the replaced ones as (0-based equation,
  // kind 1 true / 2 reduced), ascending
  int n_negative = 0, n_true_zero = 0, n_reduced_zero = 0;
  std::vector<std::pair<int, int>> singular;
};

class DeviceReducer {
 public:
  DeviceReducer();
  ~DeviceReducer();
  // Host analysis (SAM, symbolic, schedules) and upload of the part and all
  // structure arrays.  Called once per part topology.
  // upload_now: also upload the part when a device is present (otherwise the
  // first reduce() does).
  int analyze(const Part& p, const ReduceOptions& o, std::string& err, bool upload_now = false);
  // Device upload of the part and structure (first reduce() does it too).
  int upload(std::string& err);
  // The solve plans and work buffers of the reduction (end of upload).
  int prepare_solves(std::string& err);
  // One full numeric reduction on the device from the resident part data.
  int reduce(std::string& err);
  // Copies of device results in the reference layout (on demand).
  int get_B(std::vector<double>& B);     // ndof1 x ndof2, MEQN1 rows, MEQN2 columns
  int get_phi(std::vector<double>& P);   // ndof1 x nev, MEQN1 rows
  int get_schur(std::vector<double>& S); // ndof2 x ndof2 (MEQN2 order)
  int get_csr(std::vector<int>& rowptr, std::vector<int>& colind, std::vector<double>& K,
              std::vector<double>& M);     // perm-space CSR
  int get_element_matrices(std::vector<double>& Kp, std::vector<double>& Mp);
  // Extra solves on the resident factor (load cases / recovery)
  int solve_kii(int nrhs, const double* R_eq, double* X_eq, std::string& err);  // MEQN1 rows
  // g_i = (M a) on the internal equations for unit accelerations along x, y, z
  int gravity_internal(std::vector<double>& gi, std::string& err);  // ndof1 x 3, MEQN1 rows
  // reduced load vectors [f_e + B^T f_i (dofPosIn2 order); Phi^T f_i], ndim x nrhs,
  // from full load vectors F_eq (neq x nrhs, equation order)
  int reduce_loads(int nrhs, const double* F_eq, std::vector<double>& out, std::string& err);
  // the same on device arrays (F neq x nrhs, Fred ndim x nrhs, column-major)
  // with the work buffers in the library's static scratch array
  int reduce_loads_noalloc(int nrhs, const double* F_dev, double* Fred_dev, std::string& err);
  // internal displacements from boundary displacements (and internal loads):
  // U_i = B U_b (+ K_ii^{-1} F_i); U_b ndof2 x nrhs (MEQN2 rows), F_i and U_i
  // ndof1 x nrhs (MEQN1 rows), F_i may be null
  int recover(int nrhs, const double* Ub, const double* Fi, double* Ui, std::string& err);
  // device memory one reduction of the analysed part needs (bytes, estimate)
  size_t footprint_bytes() const;
  // Times this part's reduction for every candidate factor panel width and
  // solve split and keeps the fastest pair (results then differ from the
  // defaults' at rounding level only); call after analyze.
  int autotune(int& panel, int& split, std::string& err);

  const Sam& sam() const { return sam_; }
  const Symbolic& sym() const { return sym_; }
  const Part& part() const { return part_; }
  const Results& results() const { return res_; }
  const PhaseTimes& times() const { return times_; }
  hipStream_t stream() const { return st_; }
  double update_kernel_flops() const { return upd_flops_; }
  int64_t update_launches() const { return upd_launches_; }

  struct Impl;

 private:
  int build_structure(std::string& err);
  Part part_;
  Sam sam_;
  Symbolic sym_;
  ReduceOptions opt_;
  Results res_;
  PhaseTimes times_;
  hipStream_t st_ = nullptr;
  // device-to-host copy ordered after the handle's stream work
  hipError_t d2h(void* dst, const void* src, size_t bytes) const;
  bool uploaded_ = false;
  double upd_flops_ = 0;
  int64_t upd_launches_ = 0;
  std::unique_ptr<Impl> d_;
};

}  // namespace fgpu
