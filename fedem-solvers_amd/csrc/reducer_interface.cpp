// The drop-in entry points of libfedem_reducer_core (reducerInterface.C):
// initSolverArgs, reducePart, solvePart, solvePartDis and freeSingeltons,
// plus the small C services the Fortran host driver (reducer_driver.f90)
// calls through ISO_C_BINDING: option values, the FE part ("link") read from
// -linkfile, and the .fmx writer.
//
// As in the reference, the library is not re-entrant: the option registry and
// the link are process-wide singletons, and one reduction runs at a time.
#include <unistd.h>
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>
#include "../../include/fedem_reducer_gpu.h"
#include "dense_small.hpp"
#include "nastran.hpp"
#include "options.hpp"

namespace fgpu {
int write_fmx(const std::string& file, const std::string& dtype, unsigned checksum, const double* data, size_t n);
int write_tag(FILE* f, const std::string& tag, unsigned cs);
}

extern "C" int fgpu_reducer_run(int* ierr);                             // reducer_driver.f90
extern "C" int fgpu_solver_run(int n, const int* nodes, double* displ, int* ierr);  // reducer_driver.f90

namespace {

using fgpu::Options;

// Options of reducerInterface.C:63-129 (initSolverArgs) that the device
// reducer reads or that callers pass; development-only switches of the
// reference that select other element formulations are not registered.
void define_solver_options() {
  Options& o = Options::instance();
  o.define("printArray", "0", "Additional debug print switch for some arrays");
  o.define("cachesize", "0", "Cache size (KB), unused by the device solver");
  o.define("linkId", "1", "Link base-ID number");
  o.define("linkfile", "", "Name of link input file (must be specified)");
  o.define("samfile", "", "Name of SAM data file");
  o.define("extNodes", "", "List of external nodes (in addition to those specified in the link file)");
  o.define("resfile", "", "Name of result output file");
  o.define("frsfile", "", "Name of results database file");
  o.define("neval", "0", "Number of eigenvalues/eigenvectors to compute");
  o.define("ngen", "0", "Number of generalized modes");
  o.define("lumpedmass", "false", "Use lumped element mass matrices");
  o.define("tolFactorize", "1.0e-12", "Equation solver singularity criterion");
  o.define("tolEigval", "1.0e-8", "Max acceptable relative error in eigenvalues");
  o.define("eigenshift", "0.0", "Shift value for eigenvalue analysis");
  o.define("factorMass", "false", "Factorize mass matrix in the eigenvalue solver");
  o.define("singularityHandler", "1", "Option on how to treat singular matrices");
  o.define("drillingStiff", "1.0e-6", "Fictitious drilling DOF stiffness");
  o.define("cwd", "", "Working directory of the run");
  // device-solver controls (no counterpart in the reference)
  o.define("lanczosBlock", "16", "Block size of the device Lanczos eigensolver");
  o.define("mixedPrecision", "false",
           "FP32 factorisation (FP32 MFMA updates, FP64 diagonal blocks) with FP64 iterative refinement");
  o.define("lanczosMaxVectors", "0", "Basis size cap of the device Lanczos (0: automatic)");
}

// reducerInterface.C:165-204 (initReducerArgs)
void define_reducer_options() {
  Options& o = Options::instance();
  o.define("datacheck", "false", "Do data check only (exit after data input)");
  o.define("Bmatprecision", "2", "Storage precision of the B-matrix on disk");
  o.define("stiffile", "", "Name of stiffness matrix file");
  o.define("massfile", "", "Name of mass matrix file");
  o.define("gravfile", "", "Name of gravity force vector file");
  o.define("dispfile", "", "Name of displacement vector file");
  o.define("loadfile", "", "Name of load vector file");
  o.define("Bmatfile", "", "Name of B-matrix file");
  o.define("eigfile", "", "Name of eigenvector file");
  o.define("nevred", "12", "Number of eigenvalues to compute for reduced system");
  o.define("nomass", "false", "Skip mass matrix reduction");
  o.define("nograv", "false", "Skip gravity force and mass matrix calculation");
}

// the FE part of the current run (the reference's FFl link singleton)
struct Link {
  bool loaded = false;
  fgpu::NastranPart np;
  std::map<int, int> index;  // node id -> 0-based node index
  std::string err;
};
Link& link() {
  static Link l;
  return l;
}

// working-directory switch of -cwd, undone by reset_cwd (resetCWD)
std::string g_saved_cwd;
void enter_cwd() {
  std::string d = Options::instance().get("cwd");
  if (d.empty() || !g_saved_cwd.empty()) return;
  char buf[4096];
  if (getcwd(buf, sizeof buf) && chdir(d.c_str()) == 0) g_saved_cwd = buf;
}
void reset_cwd() {
  if (!g_saved_cwd.empty() && chdir(g_saved_cwd.c_str()) == 0) g_saved_cwd.clear();
}

// Option files of the program (readOptionFiles): "<program>.fco" and
// "<program>.fop" in the working directory are read when present; explicit
// -fco/-fop files are parsed with the command line.
void read_option_files(const char* program) {
  for (const char* ext : {".fco", ".fop"}) {
    std::string f = std::string(program) + ext;
    if (access(f.c_str(), R_OK) != 0) continue;
    std::string tok = "-fco=" + f, err;
    char* av[1] = {&tok[0]};
    Options::instance().parse(1, av, err);
  }
}

std::string cstr(const char* s) { return s ? std::string(s) : std::string(); }

// Fortran CHARACTER output: blank padded, returns the trimmed length
int to_fortran(const std::string& v, char* buf, int len) {
  if (!buf || len <= 0) return -1;
  int n = std::min<int>((int)v.size(), len);
  std::memcpy(buf, v.data(), n);
  std::memset(buf + n, ' ', len - n);
  return (int)v.size() <= len ? n : -(int)v.size();
}

// FNV-1a over the part data; the reference's FFl checksum (ffl_calcs) lives
// in fedem-foundation, which is not available, so files written here carry
// this checksum instead.
struct Fnv {
  uint32_t h = 2166136261u;
  void add(const void* p, size_t n) {
    const unsigned char* c = (const unsigned char*)p;
    for (size_t i = 0; i < n; i++) h = (h ^ c[i]) * 16777619u;
  }
};

}  // namespace

// ------------------------------------------------------------- entry points
extern "C" {

int initSolverArgs(int argc, char** argv, bool first, bool /*muted: ignored, callers may omit it*/) {
  Options& o = Options::instance();
  std::string err;
  if (first || !o.defined()) {
    define_solver_options();
  } else {
    // later parts reuse the definitions and take new values only
    o.clear_values();
  }
  return o.parse(argc, argv, err) == 0 ? 0 : -1;
}

int reducePart(bool first, bool last) {
  if (first || !Options::instance().defined()) {
    if (!Options::instance().defined()) define_solver_options();
    define_reducer_options();
  }
  enter_cwd();
  read_option_files("fedem_reducer");
  int ierr = 0;
  fgpu_reducer_run(&ierr);
  reset_cwd();
  // releaseSingeltons(last): the link is released after every part, the
  // option registry after the last one
  link() = Link();
  if (last) Options::instance().clear_values();
  return ierr;
}

int solvePartDis(int n, const int* nodes, double* displ) {
  if (n < 0 || (n > 0 && (!nodes || !displ))) return -1;
  if (!Options::instance().defined()) define_solver_options();
  Options::instance().define("gvec", "", "Gravity vector in global coordinates");
  enter_cwd();
  read_option_files("fedem_partsol");
  int ierr = 0;
  fgpu_solver_run(n, nodes, displ, &ierr);
  reset_cwd();
  return ierr;
}

int solvePart(void) {
  Options::instance().define("gvec", "", "Gravity vector in global coordinates");
  int ierr = solvePartDis(0, nullptr, nullptr);
  link() = Link();
  return ierr;
}

void freeSingeltons(void) {
  link() = Link();
  Options::instance().clear_values();
}

// ----------------------------------------------- services for the Fortran host
int fgpu_opt_int(const char* name) { return Options::instance().get_int(cstr(name)); }
double fgpu_opt_double(const char* name) { return Options::instance().get_double(cstr(name)); }
int fgpu_opt_bool(const char* name) { return Options::instance().get_bool(cstr(name)) ? 1 : 0; }
int fgpu_opt_isset(const char* name) { return Options::instance().is_set(cstr(name)) ? 1 : 0; }
int fgpu_opt_string(const char* name, char* buf, int len) {
  return to_fortran(Options::instance().get(cstr(name)), buf, len);
}
int fgpu_opt_doubles(const char* name, double* v, int n) {
  std::vector<double> d = Options::instance().get_doubles(cstr(name));
  int m = std::min<int>(n, (int)d.size());
  for (int i = 0; i < m; i++) v[i] = d[i];
  return (int)d.size();
}

// Reads the FE part; nodes of -extNodes become external in addition to the
// link file's ASET nodes.  0 on success, <0 with fgpu_link_error text.
int fgpu_link_read(const char* path) {
  Link& L = link();
  L = Link();
  int rc = fgpu::read_nastran(cstr(path), L.np, L.err);
  if (rc) return rc;
  for (size_t i = 0; i < L.np.node_ids.size(); i++) L.index[L.np.node_ids[i]] = (int)i;
  L.loaded = true;
  return 0;
}

int fgpu_link_error(char* buf, int len) { return to_fortran(link().err, buf, len); }

int fgpu_link_set_external(int node_id) {
  Link& L = link();
  auto it = L.index.find(node_id);
  if (!L.loaded || it == L.index.end()) {
    L.err = "external node " + std::to_string(node_id) + " is not in the part";
    return -1;
  }
  L.np.part.status[it->second] = 1;
  return 0;
}

// External nodes become fixed (all six DOFs) for the direct solve of the
// part, which has no boundary (solver.f90 assembles with linearSolve).
int fgpu_link_clamp_external(void) {
  Link& L = link();
  if (!L.loaded) return -1;
  for (int& s : L.np.part.status)
    if (s == 1) s = -63;
  return 0;
}

// sizes[0..3] = nnod, nel, number of rigid (RBE2) elements, load cases
int fgpu_link_sizes(int* sizes) {
  Link& L = link();
  if (!L.loaded || !sizes) return -1;
  sizes[0] = L.np.part.nnod;
  sizes[1] = L.np.part.nel;
  sizes[2] = (int)L.np.rbe2.size();
  std::vector<int> lc;
  for (const auto& l : L.np.loads)
    if (std::find(lc.begin(), lc.end(), l.lc) == lc.end()) lc.push_back(l.lc);
  sizes[3] = (int)lc.size();
  return 0;
}

// Nodal loads of the part: case(k) = 1-based load case in ascending set-id
// order, node(k) = 1-based node, dof(k) = 1..6, val(k).  Returns the count;
// NULL arrays only count.
int fgpu_link_loads(int* lcase, int* node, int* dof, double* val) {
  Link& L = link();
  if (!L.loaded) return -1;
  std::vector<int> ids;
  for (const auto& l : L.np.loads) ids.push_back(l.lc);
  std::sort(ids.begin(), ids.end());
  ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
  int k = 0;
  for (const auto& l : L.np.loads) {
    if (lcase) lcase[k] = (int)(std::lower_bound(ids.begin(), ids.end(), l.lc) - ids.begin()) + 1;
    if (node) node[k] = l.node + 1;
    if (dof) dof[k] = l.dof + 1;
    if (val) val[k] = l.value;
    k++;
  }
  return k;
}

int fgpu_link_node_ids(int* ids) {
  Link& L = link();
  if (!L.loaded || !ids) return -1;
  std::copy(L.np.node_ids.begin(), L.np.node_ids.end(), ids);
  return 0;
}

int fgpu_link_attach(fgpu_reducer* h) {
  Link& L = link();
  if (!L.loaded || !h) return -1;
  const fgpu::Part& p = L.np.part;
  std::vector<int> mpmnpc(p.nel + 1), mmnpc(p.enodes.size());
  for (int e = 0; e <= p.nel; e++) mpmnpc[e] = p.eptr[e] + 1;
  for (size_t k = 0; k < p.enodes.size(); k++) mmnpc[k] = p.enodes[k] + 1;
  int rc = fgpu_set_part(h, p.nnod, p.xyz.data(), p.status.data(), p.nel, p.etype.data(), mpmnpc.data(),
                         mmnpc.data(), p.eprop.data());
  if (rc == 0 && !L.np.rbe2.empty()) {
    std::vector<int> indep, mask, dptr(1, 1), dep;
    for (const auto& r : L.np.rbe2) {
      indep.push_back(r.indep + 1);
      mask.push_back(r.dofs);
      for (int u : r.dep) dep.push_back(u + 1);
      dptr.push_back((int)dep.size() + 1);
    }
    rc = fgpu_set_rigid(h, (int)indep.size(), indep.data(), mask.data(), dptr.data(), dep.data());
  }
  if (rc == 0 && !L.np.rbe3.empty()) {
    std::vector<int> ref, refc, indc, nptr(1, 1), nodes, wptr(1, 1);
    std::vector<double> w;
    for (const auto& r : L.np.rbe3) {
      ref.push_back(r.ref + 1);
      refc.push_back(r.refc);
      indc.insert(indc.end(), r.indc, r.indc + 6);
      for (int u : r.nodes) nodes.push_back(u + 1);
      nptr.push_back((int)nodes.size() + 1);
      w.insert(w.end(), r.w.begin(), r.w.end());
      wptr.push_back((int)w.size() + 1);
    }
    rc = fgpu_set_wavgm(h, (int)ref.size(), ref.data(), refc.data(), indc.data(), nptr.data(), nodes.data(),
                        wptr.data(), w.data());
  }
  if (rc) L.err = fgpu_last_error(h);
  return rc;
}

// checksum of the part and of the reduction options that change the result
// (the reference adds ngen, neval and the tolerances, reducer.f90:160-172)
int fgpu_link_checksum(int ngen, int neval, double tol_eig, double tol_fac) {
  Link& L = link();
  Fnv f;
  const fgpu::Part& p = L.np.part;
  f.add(p.xyz.data(), p.xyz.size() * sizeof(double));
  f.add(p.status.data(), p.status.size() * sizeof(int));
  f.add(p.etype.data(), p.etype.size() * sizeof(int));
  f.add(p.enodes.data(), p.enodes.size() * sizeof(int));
  f.add(p.eprop.data(), p.eprop.size() * sizeof(double));
  for (const auto& r : L.np.rbe2) {
    f.add(&r.indep, sizeof r.indep);
    f.add(&r.dofs, sizeof r.dofs);
    f.add(r.dep.data(), r.dep.size() * sizeof(int));
  }
  for (const auto& r : L.np.rbe3) {
    f.add(&r.ref, sizeof r.ref);
    f.add(&r.refc, sizeof r.refc);
    f.add(r.indc, sizeof r.indc);
    f.add(r.nodes.data(), r.nodes.size() * sizeof(int));
    f.add(r.w.data(), r.w.size() * sizeof(double));
  }
  f.add(&ngen, sizeof ngen);
  f.add(&neval, sizeof neval);
  f.add(&tol_eig, sizeof tol_eig);
  f.add(&tol_fac, sizeof tol_fac);
  return (int)(f.h & 0x7fffffff);
}

// writeDoubleDBfile (binaryDBInterface.f90:452-500) for the Fortran host
int fgpu_write_fmx(const char* file, const char* dtype, int checksum, const double* data, int64_t n) {
  if (!file || !dtype || n < 0 || (n > 0 && !data)) return -1;
  return fgpu::write_fmx(cstr(file), cstr(dtype), (unsigned)checksum, data, (size_t)n);
}

// _SAM.fsm (saveSAM, samReducerModule.f90:586-660): the "#SAM data" tag
// (the 40-byte file tag of write_tag) and the ";1.0;\n" version field,
// then int32 records: size(mpar), mpar(50), madof, minex, mnnn, msc, mpmnpc,
// mmnpc, melcon, meqn, meqn1, meqn2, dofPosIn2.  MPAR entries follow
// samModule.f90:267-280 and samReducerModule.f90:516 / inputReducerModule
// (18 link id, 19 neval, 22 ngen, 24 ndim, 29 external nodes, 50 version 3).
int fgpu_write_sam(fgpu_reducer* h, const char* file, int checksum, int link_id, int neval, int ngen) {
  Link& L = link();
  int sz[8];
  if (!h || !file || !L.loaded || fgpu_get_sizes(h, sz)) return -1;
  const int nnod = sz[0], ndof = sz[1], neq = sz[2], ndof1 = sz[3], ndof2 = sz[4], ndim = sz[5], nenod = sz[6],
            nel = sz[7];
  std::vector<int> madof(nnod + 1), msc(ndof), meqn(ndof), meqn1(ndof1), meqn2(ndof2), pos2(ndof2);
  if (fgpu_get_sam(h, madof.data(), msc.data(), meqn.data(), meqn1.data(), meqn2.data(), pos2.data())) return -1;
  const fgpu::Part& p = L.np.part;
  std::vector<int> mpar(50, 0), mnnn(nnod), mpmnpc(nel + 1), mmnpc(p.enodes.size());
  int nfix = 0;
  for (int v : msc) nfix += v == 0;
  mpar[0] = nnod;
  mpar[1] = nel;
  mpar[2] = ndof;
  mpar[3] = ndof1;
  mpar[4] = ndof2;
  mpar[5] = nfix;  // specified DOFs, all of them suppressed
  mpar[7] = nfix;
  mpar[10] = neq;
  mpar[14] = (int)mmnpc.size();
  mpar[17] = link_id;
  mpar[18] = neval;
  mpar[21] = ngen;
  mpar[23] = ndim;
  mpar[28] = nenod;
  mpar[49] = 3;
  for (int i = 0; i < nnod; i++) mnnn[i] = p.status[i] == 1 ? 2 : 1;
  for (int e = 0; e <= nel; e++) mpmnpc[e] = p.eptr[e] + 1;
  for (size_t k = 0; k < mmnpc.size(); k++) mmnpc[k] = p.enodes[k] + 1;
  FILE* f = std::fopen(file, "wb");
  if (!f) return -2;
  static const char ver[] = ";1.0;\n";
  bool ok = fgpu::write_tag(f, "#SAM data", (unsigned)checksum) == 0;
  ok = ok && std::fwrite(ver, 1, sizeof ver - 1, f) == sizeof ver - 1;
  auto put = [&](const std::vector<int>& v) {
    if (ok && !v.empty()) ok = std::fwrite(v.data(), sizeof(int), v.size(), f) == v.size();
  };
  put(std::vector<int>{50});
  put(mpar);
  put(madof);
  put(L.np.node_ids);
  put(mnnn);
  put(msc);
  put(mpmnpc);
  put(mmnpc);
  put(p.etype);
  put(meqn);
  put(meqn1);
  put(meqn2);
  put(pos2);
  ok = (std::fclose(f) == 0) && ok;
  return ok ? 0 : -2;
}

int fgpu_copy_error(const fgpu_reducer* h, char* buf, int len) { return to_fortran(fgpu_last_error(h), buf, len); }

int fgpu_reduced_eigen(int n, const double* SK, const double* SM, int nval, double* lambda) {
  if (n <= 0 || nval <= 0 || nval > n || !SK || !SM || !lambda) return -1;
  try {
    return fgpu::sym_gen_eig_lowest(n, SK, SM, nval, lambda);
  } catch (...) {
    return -1;
  }
}

}  // extern "C"
