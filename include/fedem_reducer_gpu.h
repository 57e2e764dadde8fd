/*
 * C ABI of the MI355X FE part reducer (libfedem_reducer_core.so).
 *
 * Three layers are exported:
 *
 * 1. The drop-in entry points of the reference's reducer library, with the
 *    same names, arguments, meaning and return conventions
 *    (src/vpmReducer/reducerInterface.C:52-222), so that an unmodified
 *    fedempy.reducer.FedemReducer (PythonAPI/src/fedempy/reducer.py:39-119)
 *    can load this library.
 *
 * 2. The storage-type plug-in seam the Fortran host drives (the reference
 *    dispatches its equation-solver work on SysMatrixType%storageType,
 *    sysMatrixTypeModule.f90:122-127; this library is the GPU storage type).
 *    Plain pointers and sizes only; 0 on success, negative on error (the
 *    reference's ierr convention), message via fgpu_last_error().
 *
 * 3. The bind(C) services of the Fortran host driver behind layer 1
 *    (options, the FE part, file writers).
 *
 * Index conventions follow SAM: node numbers and equation numbers are
 * 1-based, matrices are column-major (Fortran order).
 */
#ifndef FEDEM_REDUCER_GPU_H
#define FEDEM_REDUCER_GPU_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- 1. drop-in reducer entry points ---------------------------------- */

/* reducerInterface.C:52 initSolverArgs: registers (first=true) or updates the
 * command-line options.  Returns 0. */
int initSolverArgs(int argc, char** argv, bool first, bool muted);
/* reducerInterface.C:207 reducePart: reduces the part named by -linkfile and
 * writes the _S/_M/_G/_B/_E/_V/_L .fmx files.  Returns 0 or ierr < 0. */
int reducePart(bool first, bool last);
/* reducerInterface.C:147 solvePart: direct static solve of the part. */
int solvePart(void);
/* reducerInterface.C:136 solvePartDis: as solvePart, returning the
 * displacements of n nodes in displ (3*n doubles). */
int solvePartDis(int n, const int* nodes, double* displ);
/* reducerInterface.C:159 freeSingeltons */
void freeSingeltons(void);

/* ---- 2. plug-in seam -------------------------------------------------- */

typedef struct fgpu_reducer fgpu_reducer;

/* element type codes are the reference's MELCON values */
enum { FGPU_BEAM2 = 11, FGPU_ANDES4 = 24, FGPU_HEX8 = 44 };
/* per-element property record, FGPU_NPROP doubles per element */
enum {
  FGPU_P_E = 0, FGPU_P_NU = 1, FGPU_P_RHO = 2, FGPU_P_G = 3,
  FGPU_P_T1 = 4, /* shell thickness of nodes 1..4 at 4..7 */
  FGPU_P_A = 4, FGPU_P_IY = 5, FGPU_P_IZ = 6, FGPU_P_IT = 7, /* beam section */
  FGPU_P_DX = 8, FGPU_P_DY = 9, FGPU_P_DZ = 10,              /* beam orientation v (local x-y plane) */
  FGPU_P_KY = 11, FGPU_P_KZ = 12, FGPU_P_NSM = 13,
  FGPU_NPROP = 16
};

/* Creates / destroys a reducer handle (one GPU stream, device buffers). */
int fgpu_create(fgpu_reducer** h);
void fgpu_destroy(fgpu_reducer* h);
const char* fgpu_last_error(const fgpu_reducer* h);

/* The FE part: the data samReducerModule.f90:227-569 (initiateSAM) gets from
 * FFl.  status per node: 1 external, 0 internal, negative = fixed-DOF bitmask
 * -(sum of 2^i over fixed local DOFs i) (testModels.C:512-525).
 * mpmnpc: nel+1 1-based pointers into mmnpc (1-based node numbers). */
int fgpu_set_part(fgpu_reducer* h, int nnod, const double* xyz, const int* status, int nel,
                  const int* melcon, const int* mpmnpc, const int* mmnpc, const double* eprop);

/* Rigid elements (RBE2 / RGD, samReducerModule.f90:385-459) of the part set
 * last: element r has independent node indep[r], constrains the local DOFs
 * in the bitmask dofmask[r] (bit i = DOF i+1) of its dependent nodes
 * dep[dptr[r]-1 .. dptr[r+1]-2]; all node numbers and pointers 1-based.
 * Dependent DOFs get MSC -1 and follow u_g + theta_g x (x - x_g). */
int fgpu_set_rigid(fgpu_reducer* h, int nrgd, const int* indep, const int* dofmask, const int* dptr,
                   const int* dep);

/* Weighted-average-motion elements (WAVGM / Nastran RBE3; rigidModule.f90:
 * 89-338 MultiPointConstraints with wavgmModule.f90:58-367 wavgmConstrEqn)
 * of the part, set last, like the rigid elements: element w has reference
 * node ref[w] and independent nodes nodes[nptr[w]-1 .. nptr[w+1]-2];
 * refc[w] names its dependent components (digits 1..6, 0 = all);
 * indc[6w .. 6w+5] give, per component, the 1-based start of that
 * component's weights within the element's own weights
 * weights[wptr[w]-1 .. wptr[w+1]-2] (< 0 uniform weights, 0 no coupling),
 * the refC / indC / weight triple of ffl_getwavgm (rigidModule.f90:194).  An
 * element without weights takes refC 123456 with uniform translational
 * weights (rigidModule.f90:195-200).  Node numbers and pointers are 1-based.
 * Dependent DOFs get MSC -1; WAVGM elements are processed before the rigid
 * elements, and chains of constraints are rejected at fgpu_analyze. */
int fgpu_set_wavgm(fgpu_reducer* h, int nwavgm, const int* ref, const int* refc, const int* indc, const int* nptr,
                   const int* nodes, const int* wptr, const double* weights);

/* Options of the reduction: ngen (component modes kept), neval (eigenpairs
 * computed), flags bit0 = dispfile (vgi), bit1 = nomass.  Performs SAM
 * numbering (csAllocPointerArrays/csPreassemble, asmExtensionModule.f90:227,
 * 678) and the symbolic analysis, and uploads the part. */
int fgpu_analyze(fgpu_reducer* h, int ngen, int neval, int flags);

/* Numeric options applied at the next fgpu_analyze: "lanczosBlock",
 * "lanczosMaxVectors", "lanczosTol" (relative Ritz residual), "eigenshift",
 * "seed" (Lanczos start block), "leafNodes" (nested-dissection leaf size),
 * "timeUpdates" (1: time every trailing-update launch with HIP events),
 * "factorPanel" (pivot columns per trailing update: 64, 128, 192 or 256,
 * default 128), "solveSplitK" (longest contraction range of one solve task, a multiple
 * of 64, default 128; longer ranges are split and reduced in order),
 * "factorLookahead" (1: on the levels of few large fronts, the trailing
 * tiles past the next panel are updated on a second stream while the next
 * panel is factored; default 0),
 * "mixedPrecision" (1: FP32 factor of the equilibrated K with FP32 MFMA
 * trailing and Schur updates and FP64 diagonal blocks; B and every K_ii
 * solve refined against the FP64 matrix by conjugate gradients
 * preconditioned by the FP32 factor, stationary refinement where K_ii is
 * indefinite), "refineTol" (relative correction that ends the refinement,
 * default 1e-14), "refineMaxSteps" (default 40, and a correction still above
 * 0.1 after 12 steps ends it; reduce returns -6 when the refinement does not
 * converge). */
int fgpu_set_option(fgpu_reducer* h, const char* name, double value);

/* Constraint equations of the dependent DOFs in SAM's layout (samModule.f90:
 * 35-37, MPMCEQ/MMCEQ/TTCC): equation c (1-based) occupies positions
 * mpmceq[c-1] .. mpmceq[c]-1 (1-based) of mmceq/ttcc, the first holding the
 * dependent DOF and the constant term (0 here), the others an independent
 * DOF and its coefficient; DOF numbers are 1-based.  sizes[0] = nceq,
 * sizes[1] = nmmceq; call with NULL arrays to get the sizes.  Equations come
 * in ascending dependent-DOF order (the reference stores them in the order
 * they are made: WAVGM elements first, then rigid elements). */
int fgpu_get_constraints(const fgpu_reducer* h, int* sizes, int* mpmceq, int* mmceq, double* ttcc);

/* sizes[0..7] = nnod, ndof, neq, ndof1, ndof2, ndim, nenod, nel */
int fgpu_get_sizes(const fgpu_reducer* h, int* sizes);
/* SAM arrays (samModule.f90:27-64): madof(nnod+1), msc(ndof), meqn(ndof),
 * meqn1(ndof1), meqn2(ndof2), dofPosIn2(ndof2); any pointer may be NULL. */
int fgpu_get_sam(const fgpu_reducer* h, int* madof, int* msc, int* meqn, int* meqn1, int* meqn2,
                 int* dofPosIn2);

/* One numeric reduction on the GPU: element matrices, assembly, LDL^T with
 * the Schur complement (csSolveSupEl, solExtensionModule.f90:558-1033), B,
 * component modes (EIGVAL, cmstrs.f90:53), CMSTRS mass/stiffness reduction
 * (cmstrs.f90:511) and GRAV (cmstrs.f90:369). */
int fgpu_reduce(fgpu_reducer* h);

/* Reduced matrices in the reference layout (ndim x ndim column-major;
 * G is ndim x 3).  Any pointer may be NULL. */
int fgpu_get_reduced(const fgpu_reducer* h, double* SK, double* SM, double* G);
/* nev eigenvalues lambda and angular frequencies sqrt(lambda) (EIGVAL). */
int fgpu_get_eigen(const fgpu_reducer* h, int* nev, double* lambda, double* evl);
/* B (ndof1 x ndof2, MEQN1 rows, MEQN2 columns), the _B.fmx content. */
int fgpu_get_B(fgpu_reducer* h, double* B);
/* component modes (ndof1 x nev, MEQN1 rows), the _E.fmx content. */
int fgpu_get_phi(fgpu_reducer* h, double* phi);
/* vgi = K_ii^-1 g_i (ndof1 x 3), the _V.fmx content (flags bit0). */
int fgpu_get_vgi(const fgpu_reducer* h, double* vgi);
/* Tile autotuner (no counterpart in the reference: device tuning).  After
 * fgpu_analyze, times this part's reduction for the factor panel widths
 * 64/128/192/256 and the solve splits 64/128/256/512 and keeps the fastest
 * ("factorPanel", "solveSplitK"); the choice is returned so that a run can
 * be repeated with the same settings.  Results of tuned settings differ from
 * the defaults' at rounding level only. */
int fgpu_autotune(fgpu_reducer* h, int* panel, int* split);

/* Total FE mass: the sum over elements of the element mass EMASS (SMASS of
 * lhsINADD, inadd.f90:258-265 and 319), computed on the device from the
 * element mass matrices as the x-translation mass of each element. */
int fgpu_get_fe_mass(const fgpu_reducer* h, double* smass);
/* External nodes in node order, as JCMS needs them (tenc/medof of
 * readReducerData, inputReducerModule.f90:216-233): *nenod, then (when not
 * NULL) xyz(3, nenod) and medof(nenod+1), 1-based DOF pointers into the
 * reduced matrices.  Call with NULL arrays first to get nenod. */
int fgpu_get_ext_nodes(const fgpu_reducer* h, int* nenod, double* xyz, int* medof);
/* X = K_ii^-1 R on the resident factor, R and X ndof1 x nrhs (MEQN1 rows). */
int fgpu_solve_internal(fgpu_reducer* h, int nrhs, const double* R, double* X);

/* g_i = rows of M a on the internal equations for unit accelerations along
 * x, y, z (GRAVfull, cmstrs.f90:286-315): ndof1 x 3, MEQN1 rows. */
int fgpu_get_gravity_internal(fgpu_reducer* h, double* gi);
/* Reduced load vectors (CMSTRS rhs, cmstrs.f90:598-640,763-782,951-955):
 * Fred (ndim x nrhs) = [f_e + B^T f_i in dofPosIn2 order; Phi^T f_i; 0]
 * from full load vectors F (neq x nrhs, equation order), the _L.fmx
 * content. */
int fgpu_reduce_loads(fgpu_reducer* h, int nrhs, const double* F, double* Fred);

/* fgpu_reduce_loads on device memory, allocating nothing: no counterpart in
 * the reference; for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops).  F (neq x nrhs) and Fred (ndim x
 * nrhs) are column-major device arrays; the work buffers live in a static
 * 32 MiB device array that each GPU holds until the process ends.  Returns
 * -1 when nrhs needs more than that; synchronous on the handle's stream.
 * Calls that overlap on one GPU share that array and race. */
int fgpu_reduce_loads_noalloc(fgpu_reducer* h, int nrhs, const double* F, double* Fred);

/* Internal displacements on the resident factor and B (displacement
 * recovery, displacementModule.f90:555-578): Ui = B Ub + K_ii^-1 Fi, Ub
 * ndof2 x nrhs (MEQN2 rows, may be NULL), Fi and Ui ndof1 x nrhs (MEQN1
 * rows, Fi may be NULL). */
int fgpu_recover(fgpu_reducer* h, int nrhs, const double* Ub, const double* Fi, double* Ui);

/* ---- continuous-batching server ----------------------------------------
 * No counterpart in the reference (its reducer handles one part per call):
 * the batched/served form of the same reduction.  `slots` worker threads,
 * each with its own reducer and HIP stream, take requests in arrival order;
 * a request starts when its estimated device footprint, with the buffers
 * every slot holds (slots keep them between requests), fits the budget;
 * otherwise idle slots release their buffers, the largest holders first,
 * and a request that still does not fit waits for running ones (a request
 * above the whole budget runs alone).  A slot releases its own buffers
 * after a request only when all slots together hold more than the budget.
 * After the reduction each load case gives the
 * reduced load vector (as fgpu_reduce_loads) and the recovered internal
 * displacements (as fgpu_recover with Fi = the case's internal loads).
 * Results of a request are bitwise the same alone or among others. */
typedef struct fgpu_server fgpu_server;
int fgpu_server_create(fgpu_server** s, int slots, double budget_gb);
void fgpu_server_destroy(fgpu_server* s);
/* Part as fgpu_set_part, ngen component modes, nlc load cases: case c has
 * nodal loads k in [lptr[c], lptr[c+1]) (0-based) on node lnode[k] (1-based),
 * DOF ldof[k] (1..6), value lval[k]; ub (nub = ndof2 values per case, reduced
 * dofPosIn2 order) the boundary displacements or NULL.  Returns the request
 * id (>= 0) or a negative error. */
int fgpu_server_submit(fgpu_server* s, int nnod, const double* xyz, const int* status, int nel, const int* melcon,
                       const int* mpmnpc, const int* mmnpc, const double* eprop, int ngen, int nlc, const int* lptr,
                       const int* lnode, const int* ldof, const double* lval, int nub, const double* ub);
/* Numeric option (the names of fgpu_set_option, e.g. "mixedPrecision") for
 * the requests submitted after the call; not to be called while another
 * thread submits.  0, or -1 for an unknown name. */
int fgpu_server_set_option(fgpu_server* s, const char* name, double value);
/* Blocks until request id is done; its status (0 or negative). */
int fgpu_server_wait(fgpu_server* s, int id);
/* Arithmetic a finished request ran in: 1 mixed precision, 0 FP64 (a
 * mixed-precision request whose refinement did not converge is redone in
 * FP64 on its slot), -1 no such finished request. */
int fgpu_server_precision(fgpu_server* s, int id);
/* sizes[0..4] = ndim, ndof1, ndof2, nlc, nev of a finished request. */
int fgpu_server_sizes(fgpu_server* s, int id, int* sizes);
/* SK, SM (ndim^2), G (ndim x 3), Fred (ndim x nlc), Ui (ndof1 x nlc);
 * times[0..3] = queueing ms, service ms, latency ms, footprint bytes. */
int fgpu_server_result(fgpu_server* s, int id, double* SK, double* SM, double* G, double* Fred, double* Ui,
                       double* times);
int fgpu_server_error(fgpu_server* s, int id, char* buf, int len);
/* Frees a finished request's results. */
int fgpu_server_release(fgpu_server* s, int id);
/* Device bytes the slots' buffers hold now (running and idle slots). */
int64_t fgpu_server_device_bytes(fgpu_server* s);

/* Diagnostics: times[0..12] (ms) = elements, assembly, factor, B, modes,
 * mass, gravity, total, summed duration of the trailing-update (MFMA)
 * launches and their count (option "timeUpdates"), refinement steps of the
 * mixed-precision B solve, host analysis of the last fgpu_analyze (SAM
 * numbering, ordering, fronts), the last upload (part and structure
 * transfer, device construction of the CSR pattern, contribution lists
 * and front map, schedules; done by the first fgpu_reduce after
 * fgpu_analyze); info[0..9] = nsup, nlevels, front pool doubles,
 * nnz(L), factor flops, max front, nnz(K), Lanczos vectors, packed element
 * pool doubles, trailing-update (MFMA) flops. */
int fgpu_get_times(const fgpu_reducer* h, double* times);
int fgpu_get_info(const fgpu_reducer* h, int64_t* info);
/* Assembled K and M in the device's permuted CSR (rows/cols 0-based perm
 * indices), perm_to_eq maps a perm index to a 0-based equation number. */
int fgpu_get_csr(fgpu_reducer* h, int* rowptr, int* colind, double* K, double* M, int* perm_to_eq);
/* Packed element matrices as computed on the device (node-pair blocks). */
int fgpu_get_element_matrices(fgpu_reducer* h, double* Kp, double* Mp, int64_t* offsets);
/* Symbolic structure for host-side verification: per supernode col0, ncol,
 * nrow, parent; rows (nrow each, perm indices). */
int fgpu_get_supernodes(const fgpu_reducer* h, int* col0, int* ncol, int* nrow, int* parent, int* rows);
/* Pivots of the last factorisation of K_ii (also after fgpu_reduce failed
 * on them): counts[0..2] = negative pivots, true zero pivots and pivots
 * reduced to near zero, the last two replaced by |a_jj| (the reference's
 * SPRFC3 modification, solExtensionModule.f90:725-747; the replacement value
 * itself is this library's choice); for the replaced ones, in equation
 * order, the 1-based node, local DOF, equation and kind (1 true, 2
 * reduced) of at most max of them (reportSingEqns, cmstrs.f90:1324-1352).
 * Returns how many were replaced.  Options "tolFactorize" (default 1e-12)
 * and "singularityHandler" (0: any replaced pivot makes fgpu_reduce fail
 * with -5; 1, the default: only reduced-to-zero ones do; > 1: none). */
int fgpu_get_pivots(const fgpu_reducer* h, int* counts, int* node, int* ldof, int* eqn, int* kind, int max);
/* Lanczos statistics of the last reduction: basis vectors used and the worst
 * relative Ritz residual ||K^-1 M phi - theta phi||_M / theta at the end. */
int fgpu_get_lanczos(const fgpu_reducer* h, int* nvec, double* resid);
/* perm_to_eq[p] = 0-based equation number at device position p (neq). */
int fgpu_get_perm(const fgpu_reducer* h, int* perm_to_eq);

/* ---- 3. services of the Fortran host driver --------------------------- */
/* The drop-in entry points run the Fortran phase driver
 * (reducer_driver.f90: fgpu_reducer_run mirrors reducer.f90:98-716,
 * fgpu_solver_run mirrors solver.f90:30-383).  It reaches options, the FE
 * part and the file writers through these bind(C) services. */

/* Option values by name (FFaCmdLineArgInterface: ffa_cmdlinearg_getint/
 * getdouble/getbool/isSet/getstring/getdoubles).  fgpu_opt_string fills a
 * blank-padded Fortran buffer and returns the trimmed length (negative: the
 * value needs that many characters); fgpu_opt_doubles returns the count. */
int fgpu_opt_int(const char* name);
double fgpu_opt_double(const char* name);
int fgpu_opt_bool(const char* name);
int fgpu_opt_isset(const char* name);
int fgpu_opt_string(const char* name, char* buf, int len);
int fgpu_opt_doubles(const char* name, double* v, int n);

/* The FE part of the run (the FFl link singleton, ffl_init/ffl_done):
 * fgpu_link_read parses a Nastran bulk-data file; set_external adds a node id
 * to the external nodes (-extNodes); clamp_external fixes all DOFs of the
 * external nodes (fixExternalDOFs for the direct solve,
 * samReducerModule.f90:339); sizes[0..3] = nnod, nel, rigid elements, load
 * cases; node_ids = external ids in node order (MINEX); attach hands the
 * part to a reducer handle; loads returns nodal loads (1-based load case,
 * node, DOF 1..6, value) and their count (NULL arrays only count). */
int fgpu_link_read(const char* path);
int fgpu_link_error(char* buf, int len);
int fgpu_link_set_external(int node_id);
int fgpu_link_clamp_external(void);
int fgpu_link_sizes(int* sizes);
int fgpu_link_node_ids(int* ids);
int fgpu_link_attach(fgpu_reducer* h);
int fgpu_link_loads(int* lcase, int* node, int* dof, double* val);
/* Checksum carried by the output files (stands in for ffl_calcs +
 * ffl_addcs_*, reducer.f90:160-172): FNV-1a over the part and options. */
int fgpu_link_checksum(int ngen, int neval, double tol_eig, double tol_fac);

/* writeDoubleDBfile (binaryDBInterface.f90:452-500): the 40-byte file tag
 * "#FEDEM <dtype>" with the checksum (FFa_writeTag layout: 30 blank-padded
 * characters, int32 endian marker 0x1234 and two zero bytes, uint32
 * checksum), the ";1.0;\n" version field, then n column-major doubles. */
int fgpu_write_fmx(const char* file, const char* dtype, int checksum, const double* data, int64_t n);
/* writetagdb (binaryDB.c:643-674) as a file service: a new file holding the
 * file tag (tag of at most 30 characters) and the given version field. */
int fgpu_write_tag(const char* file, const char* tag, unsigned checksum, const char* version);
/* readtagdb (binaryDB.c:676-733): tag text (trailing blanks removed),
 * checksum (byte-swapped when the endian marker is the other order) and
 * version line of a tagged binary file; returns the offset of the data
 * after the version line, or < 0 (-3: not a tagged file, -4: short). */
int64_t fgpu_read_tag(const char* file, char* tag, int tag_len, unsigned* checksum, char* version, int ver_len);
/* saveSAM (samReducerModule.f90:586-660): the _SAM.fsm file. */
int fgpu_write_sam(fgpu_reducer* h, const char* file, int checksum, int link_id, int neval, int ngen);
/* Last error text of a handle into a blank-padded Fortran buffer. */
int fgpu_copy_error(const fgpu_reducer* h, char* buf, int len);
/* Lowest nval eigenvalues (ascending) of SK z = lambda SM z for the n x n
 * reduced matrices (column-major, upper triangles read): the
 * solveEigenvalues / DSYGVX call of EIGCMS (cmstrs.f90:1215-1310;
 * denseMatrixModule.f90:314-340).  Returns 0, n + j when the leading minor
 * of order j of SM is not positive definite (DSYGVX's INFO), -1 on bad
 * arguments.  Host arithmetic on matrices of at most a few hundred rows. */
int fgpu_reduced_eigen(int n, const double* SK, const double* SM, int nval, double* lambda);
/* The Fortran phase drivers themselves (ierr-style status in *ierr). */
int fgpu_reducer_run(int* ierr);
int fgpu_solver_run(int n, const int* nodes, double* displ, int* ierr);

/* fmxWriter.f90 (loaded by fedempy.write_fmx): itype 1 _K.fmx, 2 _M.fmx,
 * 3 _G.fmx next to the prefix; the trailing argument is the Fortran length
 * of prefix.  0 on success. */
int writefmx_(const char* prefix, const int* itype, double* data, const int* nval, size_t len);
int readfmx_(const char* prefix, const int* itype, double* data, const int* nval, size_t len);

#ifdef __cplusplus
}
#endif

#endif
