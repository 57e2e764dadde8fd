// Device orchestration of one part reduction on a single MI355X.
//
// Phase map to the reference driver (reducer.f90:22-748):
//   elements + assembly   INADD / csAddEM                    (inadd.f90:26-464)
//   factor + Schur        csSolveSupEl SPR branch, SPRFC1 +
//                         SPREXK                              (solExtensionModule.f90:708-775)
//   B                     SPRBS1/2 per external DOF           (solExtensionModule.f90:809-843)
//   component modes       EIGVAL -> csLanczosEigenSolve       (cmstrs.f90:53-263)
//   mass / stiffness      CMSTRS K11, M11, M12, M22           (cmstrs.f90:728-944)
//   gravity               GRAVfull + GRAV                     (cmstrs.f90:286-474)
//   dispfile              vgi = K_ii^{-1} g_i                 (cmstrs.f90:674-712)
#include "reducer_gpu.hpp"
#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <future>
#include "dense_small.hpp"
#include "kernels.hpp"

namespace fgpu {

#define HIPCHK(x)                                                                          \
  do {                                                                                     \
    hipError_t e_ = (x);                                                                   \
    if (e_ != hipSuccess) {                                                                \
      err = std::string("HIP error '") + hipGetErrorString(e_) + "' in " + #x;             \
      return -3;                                                                           \
    }                                                                                      \
  } while (0)

// Device bytes held by the buffers that the calling thread allocates (the
// server gives every slot's worker thread its own counter).
static thread_local std::atomic<int64_t>* tl_account = nullptr;
void set_alloc_account(std::atomic<int64_t>* a) { tl_account = a; }
// Stream the uploads of this thread go through (the handle's own stream while
// one of its methods runs), so that they never wait for other handles' work
// the way copies on the null stream would.
static thread_local hipStream_t tl_stream = nullptr;
struct StreamScope {
  hipStream_t prev;
  explicit StreamScope(hipStream_t s) : prev(tl_stream) { tl_stream = s; }
  ~StreamScope() { tl_stream = prev; }
};

template <class T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  std::atomic<int64_t>* acct = nullptr;  // counter charged with this buffer
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) {
      (void)hipFree(p);
      if (acct) *acct -= (int64_t)(std::max<size_t>(1, n) * sizeof(T));
    }
    p = nullptr;
    n = 0;
    acct = nullptr;
  }
  hipError_t alloc(size_t count) {
    if (count <= n && p) return hipSuccess;
    release();
    hipError_t e = hipMalloc((void**)&p, std::max<size_t>(1, count) * sizeof(T));
    if (e != hipSuccess) {
      p = nullptr;
      return e;
    }
    n = count;
    acct = tl_account;
    if (acct) *acct += (int64_t)(std::max<size_t>(1, n) * sizeof(T));
    return hipSuccess;
  }
  hipError_t upload(const std::vector<T>& v) {
    hipError_t e = alloc(v.size());
    if (e != hipSuccess || v.empty()) return e;
    if (!tl_stream) return hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
    e = hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, tl_stream);
    return e == hipSuccess ? hipStreamSynchronize(tl_stream) : e;  // v may go out of scope after the call
  }
};

// Pinned host memory (hipHostMalloc) for asynchronous device-to-host copies.
template <class T>
struct HBuf {
  T* p = nullptr;
  size_t n = 0;
  HBuf() = default;
  HBuf(const HBuf&) = delete;
  HBuf& operator=(const HBuf&) = delete;
  ~HBuf() { release(); }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    n = 0;
  }
  hipError_t alloc(size_t count) {
    if (count <= n && p) return hipSuccess;
    release();
    n = count;
    return hipHostMalloc((void**)&p, std::max<size_t>(1, count) * sizeof(T), hipHostMallocDefault);
  }
};

enum OpKind {
  EXTADD, DIAG, TRSM, DIAG_TRSM, UPDATE, INV_OFF, INV_PART, INV_FIN,
  UPDATE_SIDE, FORK, JOIN,  // lookahead: trailing tiles on the side stream, fork/join with the main one
  FW_INIT, FW_INV, FW_UPD, FW_OUT, BW_INIT, BW_UPD, BW_INV,
  FW_INV_R, FW_UPD_R, BW_UPD_R, BW_INV_R  // ordered reductions of split-K partials
};
struct Op {
  int kind, off, n;
};

struct SolvePlan {
  bool valid = false;  // built for the current structure
  int nrhs = 0;
  int nslots = 0;  // 64x64 partial buffers of the split-K solve tasks
  std::vector<Op> fw_int, fw_all, bw;
  DBuf<SolveTask> tasks;
  DBuf<int> gptr, gsrc;  // forward init: the child update rows added to each work row, in child order
  // Split-K update-row products of the forward pass (nrhs <= 64) are reduced
  // by the parent's forward init as it gathers the rows: ugrp_int/ugrp_all
  // (per work row, -1: none) index the groups (first slot, slots, first
  // work row of the tile) of the internal-only and the full forward pass.
  DBuf<int> ugrp_int, ugrp_all, ugroups;
  DBuf<double> part;
};

// Host lists of a solve plan before their upload.
struct PlanHost {
  std::vector<SolveTask> tasks;
  std::vector<int> gptr, gsrc, ugrp[2], ugroups;
};

// A captured launch sequence (HIP graph), replayed while the buffers it
// was captured with are unchanged (key = their addresses).
struct GraphCache {
  hipGraphExec_t exec = nullptr;
  std::vector<const void*> key;
  void reset() {
    if (exec) (void)hipGraphExecDestroy(exec);
    exec = nullptr;
    key.clear();
  }
  ~GraphCache() { reset(); }
};

// Replays gc when its key matches, else captures body() on stream st (kernel
// launches only: no memset or host-synchronous call inside) and launches it.
template <class F>
static hipError_t run_graph(hipStream_t st, GraphCache& gc, const std::vector<const void*>& key, F&& body) {
  if (gc.exec && gc.key == key) return hipGraphLaunch(gc.exec, st);
  if (gc.exec) {
    (void)hipGraphExecDestroy(gc.exec);
    gc.exec = nullptr;
  }
  hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return e;
  body();
  hipGraph_t g = nullptr;
  e = hipStreamEndCapture(st, &g);
  if (e != hipSuccess) return e;
  e = hipGraphInstantiate(&gc.exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    gc.exec = nullptr;
    return e;
  }
  gc.key = key;
  return hipGraphLaunch(gc.exec, st);
}

struct DeviceReducer::Impl {
  GraphCache g_factor, g_lanczos;
  int factor_runs = 0;  // factorisations since the last upload
  // part
  DBuf<double> xyz, eprop;
  DBuf<int> eptr, enodes, hexlist, shlist, bmlist, p1;
  DBuf<int64_t> eoff;
  // structure
  DBuf<int> col0, ncol, nrow, rptr, rows, chptr, chlist, relptr, rel, rowptr, colind, cptr, cidx;
  DBuf<int64_t> foff, woff, fmap, ioff;
  // inputs of the device structure build: SAM arrays, element types, per
  // node first equation / equation count / neighbour lists (+ DOF offsets),
  // per row its node, per column its front
  DBuf<int> madof, msc, nndof, drank, etype, nfirst, neqd, nbptr, nbl, nbo, rnode, colsn;
  DBuf<double> signs;                    // mode signs and the per-group bests of their search
  // grow-only work arrays of the load-case calls (freeing device memory
  // waits for the whole device, so they are kept between calls)
  DBuf<double> sk_r, sk_t, sk_x, rl_f, rc_s, rc_y, gi_g;
  DBuf<int> ceq_ptr, ceq_node, ceq_dof;  // constraint equations of the dependent DOFs
  DBuf<double> ceq_coef;
  DBuf<int> p2e, pos2;  // perm row -> 0-based equation; dofPosIn2 (for the device-side reduced loads)
  DevSym S{};
  DBuf<FTask> ftasks;
  DBuf<int> xrange;  // extend-add tasks: (child, first, end) column ranges
  DBuf<double> dpiv;   // pivots D of the factor by pivot column (FP64 for both factor types)
  DBuf<double> adiag, amax;  // original (mixed: equilibrated) diagonal of K_ii and its largest magnitude
  DBuf<int> sing;            // per pivot column: 0, 1 true zero replaced, 2 reduced-to-zero replaced
  DBuf<double> ipart;  // partial sums of split pivot-block inverse tiles (64 x 64 each; FP32 ones use half)
  std::vector<Op> fops;
  std::map<int, std::unique_ptr<SolvePlan>> plans;
  // values
  DBuf<double> Kp, Mp, Kv, Mv, pool, W, V, ipool;
  DBuf<int> ebad, piv;
  DBuf<double> X, Z, Y, part, small, acc, grav, ccoef;
  DBuf<double> emass;  // nel element masses, then their total
  // mixed precision: FP32 fronts and pivot-block inverses, equilibration
  // scale, refinement residual and norm partials
  bool mixed = false;
  bool m_lumped = false;  // only lumped-mass (ANDES4) elements: M is diagonal in a K-shaped pattern
  int solve_split = 128;
  DBuf<float> pool32, ipool32;
  DBuf<double> dscale, dcs, Rr, irpart;
  DBuf<double> cg_p, cg_q, cg_z, cg_t, cg_part, cg_dot, cg_co;  // conjugate-gradient refinement
  std::vector<double> s22;  // mixed: Schur complement from the refined B (n2 x n2, column-major)
  // component modes
  DBuf<double> Q, P, Wl, Zl, Yl, Cc, Phi, Tmp;
  DBuf<double> Cacc, Rdev;  // reorthogonalisation coefficients of a step; the two CholQR factors
  DBuf<int> lzbad;
  HBuf<double> hC, hR;  // two slots each (step parity)
  HBuf<int> hbad;
  hipEvent_t lz_ev[2]{};
  hipStream_t side = nullptr;  // lookahead stream of the factorisation
  hipEvent_t fj[2]{};          // its fork and join events
  std::vector<int> hexl, shl, bml;  // elements by type: HEX8, ANDES4, BEAM2
  int64_t wrows = 0;
  hipEvent_t ev[12]{};
  std::vector<hipEvent_t> upd_ev;  // start/stop pairs of the trailing-update launches
};

DeviceReducer::DeviceReducer() : d_(new Impl) {}
DeviceReducer::~DeviceReducer() {
  if (st_) {
    (void)hipStreamSynchronize(st_);
    for (auto& e : d_->ev)
      if (e) (void)hipEventDestroy(e);
    for (auto& e : d_->upd_ev) (void)hipEventDestroy(e);
    for (auto& e : d_->lz_ev)
      if (e) (void)hipEventDestroy(e);
    for (auto& e : d_->fj)
      if (e) (void)hipEventDestroy(e);
    if (d_->side) {
      (void)hipStreamSynchronize(d_->side);
      (void)hipStreamDestroy(d_->side);
    }
    (void)hipStreamDestroy(st_);
  }
}

static void build_factor_schedule(const Symbolic& y, std::vector<FTask>& tasks, std::vector<Op>& ops,
                                  std::vector<int>& xrange, double& flops, int panel, int& inv_slots,
                                  bool lookahead) {
  xrange.clear();
  inv_slots = 0;
  flops = 0;
  for (int l = 0; l < y.nlevels; l++) {
    const int* sups = &y.lvsup[y.lvptr[l]];
    const int ns = y.lvptr[l + 1] - y.lvptr[l];
    int off = (int)tasks.size();
    for (int q = 0; q < ns; q++) {
      int s = sups[q];
      if (y.chptr[s + 1] == y.chptr[s]) continue;
      // extend-add task (front s, 16-column block j0): the (child, first,
      // end) ranges of child update columns landing in the block, in child
      // order, at xrange[3 b ..], kb of them
      for (int j0 = 0; j0 < y.nrow[s]; j0 += 16) {
        FTask t{s, j0, (int)xrange.size() / 3, 0, 0};
        for (int q = y.chptr[s]; q < y.chptr[s + 1]; q++) {
          const int c = y.chlist[q], nu = y.nrow[c] - y.ncol[c];
          const int* rel = y.rel.data() + y.relptr[c];
          const int ka = (int)(std::lower_bound(rel, rel + nu, j0) - rel);
          const int kb = (int)(std::lower_bound(rel, rel + nu, j0 + 16) - rel);
          if (kb > ka) {
            xrange.push_back(c);
            xrange.push_back(ka);
            xrange.push_back(kb);
            t.kb++;
          }
        }
        tasks.push_back(t);
      }
    }
    if ((int)tasks.size() > off) ops.push_back({EXTADD, off, (int)tasks.size() - off});
    int maxnc = 0;
    for (int q = 0; q < ns; q++) maxnc = std::max(maxnc, y.ncol[sups[q]]);
    // Pivot columns go in panels of P = 64 h columns (h sub-panels of 64):
    // for each sub-panel the diagonal block and the triangular solve, then
    // its update of the panel's remaining columns only; after the last
    // sub-panel one trailing update over the whole panel, so the trailing
    // block is read and written once per panel.  A front with fewer than P
    // pivot columns left takes plain 64-column panels.
    // Panels with at most FUSE_TRSM row tiles below their diagonal blocks
    // (summed over the level) take one fused launch instead.
    constexpr int FUSE_TRSM = 256;
    auto add_diag_trsm = [&](int k0) {
      int ntr = 0;
      for (int q = 0; q < ns; q++) {
        int s = sups[q];
        if (y.ncol[s] > k0) ntr += (y.nrow[s] - k0 - std::min(64, y.ncol[s] - k0) + 63) / 64;
      }
      if (ntr <= FUSE_TRSM) {
        int off = (int)tasks.size();
        for (int q = 0; q < ns; q++) {
          int s = sups[q];
          if (y.ncol[s] <= k0) continue;
          int kb = std::min(64, y.ncol[s] - k0), rem = y.nrow[s] - k0 - kb;
          if (rem <= 0) tasks.push_back({s, -1, 0, k0, 0});
          for (int a = 0; 64 * a < rem; a++) tasks.push_back({s, a, 0, k0, 0});
        }
        if ((int)tasks.size() > off) ops.push_back({DIAG_TRSM, off, (int)tasks.size() - off});
        return;
      }
      int off = (int)tasks.size();
      for (int q = 0; q < ns; q++) {
        int s = sups[q];
        if (y.ncol[s] > k0) tasks.push_back({s, 0, 0, k0, 0});
      }
      if ((int)tasks.size() > off) ops.push_back({DIAG, off, (int)tasks.size() - off});
      off = (int)tasks.size();
      for (int q = 0; q < ns; q++) {
        int s = sups[q];
        if (y.ncol[s] <= k0) continue;
        int kb = std::min(64, y.ncol[s] - k0), rem = y.nrow[s] - k0 - kb;
        for (int a = 0; 64 * a < rem; a++) tasks.push_back({s, a, 0, k0, 0});
      }
      if ((int)tasks.size() > off) ops.push_back({TRSM, off, (int)tasks.size() - off});
    };
    // 64x64 lower tiles of the trailing block from k0 + kb, tile columns in
    // [tlo, thi) (thi < 0: all), appended to `out`
    auto add_update_to = [&](std::vector<FTask>& out, int s, int k0, int kb, int tlo, int thi) {
      const int rem = y.nrow[s] - k0 - kb;
      if (rem <= 0) return;
      const int T = (rem + 63) / 64;
      for (int ti = 0; ti < T; ti++)
        for (int tj = tlo; tj <= ti && (thi < 0 || tj < thi); tj++) out.push_back({s, ti, tj, k0, kb});
      // lower-triangle entries (incl. diagonal) of the updated columns c in
      // [clo, chi), rem - c each, 2 flops per FMA
      const double clo = std::min(rem, 64 * tlo), chi = thi < 0 ? rem : std::min(rem, 64 * thi);
      if (chi > clo) flops += 2.0 * kb * ((chi - clo) * rem - (chi * (chi - 1) - clo * (clo - 1)) / 2);
    };
    auto add_update = [&](int s, int k0, int kb, int ntc) { add_update_to(tasks, s, k0, kb, 0, ntc); };
    const int P = panel, NA = P / 64;
    // Lookahead on the levels of few large fronts (those that take the fused
    // diagonal/TRSM launch): the trailing update after a panel is split into
    // the tiles of the next panel's columns (main stream) and the rest (side
    // stream), so the next panel's diagonal blocks, TRSM and narrow update
    // run while the rest of the trailing block is updated.  Any update that
    // reaches past the next panel joins the side stream first, and so does
    // the end of the level.
    int ntr0 = 0;
    for (int q = 0; q < ns; q++) {
      int s = sups[q];
      if (y.ncol[s] > 0) ntr0 += (y.nrow[s] - std::min(64, y.ncol[s]) + 63) / 64;
    }
    const bool look = lookahead && ntr0 <= FUSE_TRSM;
    bool side_pending = false;
    for (int k0 = 0; k0 < maxnc; k0 += P) {
      for (int h = 0; h < P / 64 && k0 + 64 * h < maxnc; h++) {
        const int kh = k0 + 64 * h;
        add_diag_trsm(kh);
        const int off = (int)tasks.size();
        std::vector<FTask> side;
        bool past = false;  // some task updates columns past the next panel
        for (int q = 0; q < ns; q++) {
          int s = sups[q];
          const int rc = y.ncol[s] - k0;
          if (rc <= 64 * h) continue;
          if (rc >= P) {
            if (h + 1 < P / 64) {
              add_update(s, kh, 64, P / 64 - h - 1);  // the panel's remaining columns
            } else if (look) {                        // the trailing block, whole panel: split
              add_update_to(tasks, s, k0, P, 0, NA);
              add_update_to(side, s, k0, P, NA, -1);
              past = true;
            } else {
              add_update(s, k0, P, -1);
              past = true;
            }
          } else {
            add_update(s, kh, std::min(64, rc - 64 * h), -1);
            past = true;
          }
        }
        if ((int)tasks.size() > off) {
          if (past && side_pending) {
            ops.push_back({JOIN, 0, 0});
            side_pending = false;
          }
          ops.push_back({UPDATE, off, (int)tasks.size() - off});
        }
        if (!side.empty()) {
          const int soff = (int)tasks.size();
          tasks.insert(tasks.end(), side.begin(), side.end());
          ops.push_back({FORK, 0, 0});
          ops.push_back({UPDATE_SIDE, soff, (int)side.size()});
          side_pending = true;
        }
      }
    }
    if (side_pending) ops.push_back({JOIN, 0, 0});
  }
  // Explicit inverse of every pivot block: the diagonal tiles come from
  // k_diag, the tiles at growing distance from the diagonal follow.  Only the
  // solves read the inverses, so they are formed after the whole
  // factorisation, distance by distance over all fronts at once.
  int maxT = 0;
  for (int s = 0; s < y.nsup; s++) maxT = std::max(maxT, (y.ncol[s] + 63) / 64);
  // A tile whose sum has more than INV_CHUNK terms is split: partial sums
  // over INV_CHUNK-tile chunks in parallel (k_inv_part), then their sum in
  // chunk order and the product with X_ii (k_inv_fin).
  for (int dist = 1; dist < maxT; dist++) {
    int off = (int)tasks.size();
    std::vector<FTask> part, fin;
    int slot = 0;
    for (int s = 0; s < y.nsup; s++) {
      const int T = (y.ncol[s] + 63) / 64;
      for (int j = 0; j + dist < T; j++) {
        if (dist <= INV_CHUNK) {
          tasks.push_back({s, j, j + dist, 0, 0});
          continue;
        }
        const int first = slot;
        for (int k = j; k < j + dist; k += INV_CHUNK) part.push_back({s, j, j + dist, k, slot++});
        fin.push_back({s, j, j + dist, first, slot - first});
      }
    }
    if ((int)tasks.size() > off) ops.push_back({INV_OFF, off, (int)tasks.size() - off});
    if (!part.empty()) {
      off = (int)tasks.size();
      tasks.insert(tasks.end(), part.begin(), part.end());
      ops.push_back({INV_PART, off, (int)part.size()});
      off = (int)tasks.size();
      tasks.insert(tasks.end(), fin.begin(), fin.end());
      ops.push_back({INV_FIN, off, (int)fin.size()});
    }
    inv_slots = std::max(inv_slots, slot);
  }
}

static void build_solve_plan(const Symbolic& y, SolvePlan& sp, std::vector<SolveTask>& tasks,
                             std::vector<int>& gptr, std::vector<int>& gsrc, std::vector<int> (&ugrp)[2],
                             std::vector<int>& ugroups, int KSplit) {
  const int nrhs = sp.nrhs, schur = y.nsup - 1;
  // gather lists of the forward init: work row woff[s] + i of front s adds
  // the update rows of s's children that land in its row i, child after
  // child (a child adds at most one row to a parent row)
  std::vector<int64_t> woff(y.nsup + 1, 0);
  for (int s = 0; s < y.nsup; s++) woff[s + 1] = woff[s] + y.nrow[s];
  gptr.assign(woff[y.nsup] + 1, 0);
  for (int s = 0; s < y.nsup; s++)
    for (int q = y.chptr[s]; q < y.chptr[s + 1]; q++) {
      const int c = y.chlist[q], nu = y.nrow[c] - y.ncol[c];
      const int* rel = y.rel.data() + y.relptr[c];
      for (int k = 0; k < nu; k++) gptr[woff[s] + rel[k] + 1]++;
    }
  for (size_t i = 1; i < gptr.size(); i++) gptr[i] += gptr[i - 1];
  gsrc.assign(gptr.back(), 0);
  {
    std::vector<int> fill(gptr.begin(), gptr.end() - 1);
    for (int s = 0; s < y.nsup; s++)
      for (int q = y.chptr[s]; q < y.chptr[s + 1]; q++) {
        const int c = y.chlist[q], nu = y.nrow[c] - y.ncol[c];
        const int* rel = y.rel.data() + y.relptr[c];
        for (int k = 0; k < nu; k++) gsrc[fill[woff[s] + rel[k]]++] = (int)(woff[c] + y.ncol[c] + k);
      }
  }
  // (front, 64-row chunk) tasks over all rows (forward init), the solved
  // rows (forward output, the Schur front: all its rows) or the pivot rows
  // (backward init)
  auto add_sns = [&](std::vector<Op>& ops, int kind, const std::vector<int>& list) {
    int off = (int)tasks.size();
    for (int s : list) {
      const bool piv = (kind == FW_OUT && s != schur) || kind == BW_INIT;
      const int nr = piv ? y.ncol[s] : y.nrow[s];
      for (int r0 = 0; r0 < nr; r0 += 64) tasks.push_back({s, r0, std::min(64, nr - r0), 0, 0, 0, -1, 0});
    }
    if ((int)tasks.size() > off) ops.push_back({kind, off, (int)tasks.size() - off});
  };
  // GEMM tasks over the row tiles of a front and all right-hand-side tiles;
  // K ranges longer than KS are split and reduced in slot order
  const bool fuse_upd = nrhs <= 64;
  ugroups.clear();
  for (auto& u : ugrp) u.assign(woff[y.nsup], -1);
  int pass = 0;  // forward pass being built (0: internal fronts only, 1: all)
  auto add_tiles = [&](std::vector<Op>& ops, int kind, const std::vector<int>& list, bool upd_rows) {
    const int mode = kind == FW_INV ? 0 : kind == FW_UPD ? 1 : kind == BW_UPD ? 2 : 3;
    // an op with enough unsplit tasks to fill the device is not split
    int64_t ntask = 0;
    for (int s : list) {
      const int nc = y.ncol[s], nu = y.nrow[s] - nc;
      if (kind == BW_UPD && nu == 0) continue;
      ntask += (int64_t)((upd_rows ? nu : nc) + 63) / 64 * ((nrhs + 63) / 64);
    }
    const int KS = ntask >= 1024 ? INT_MAX : KSplit;
    const int rkind = kind == FW_INV ? FW_INV_R : kind == FW_UPD ? FW_UPD_R : kind == BW_UPD ? BW_UPD_R : BW_INV_R;
    std::vector<SolveTask> red;
    int off = (int)tasks.size();
    for (int s : list) {
      const int nc = y.ncol[s], nu = y.nrow[s] - nc;
      int nr = upd_rows ? nu : nc;
      if (kind == BW_UPD && nu == 0) continue;
      for (int r0 = 0; r0 < nr; r0 += 64) {
        const int rn = std::min(64, nr - r0);
        int kb = 0, ke = 0;
        switch (mode) {
          case 0: kb = 0; ke = r0 + rn; break;
          case 1: kb = 0; ke = nc; break;
          case 2: kb = 0; ke = nu; break;
          default: kb = r0; ke = nc; break;
        }
        for (int j0 = 0; j0 < nrhs; j0 += 64) {
          if (ke - kb <= KS) {
            tasks.push_back({s, r0, rn, kb, ke - kb, j0, -1, 0});
            continue;
          }
          int first = sp.nslots;
          for (int k = kb; k < ke; k += KS) tasks.push_back({s, r0, rn, k, std::min(KS, ke - k), j0, sp.nslots++, 0});
          if (mode == 1 && fuse_upd && pass < 2) {
            const int w0 = (int)(woff[s] + nc + r0), g = (int)ugroups.size() / 3;
            ugroups.insert(ugroups.end(), {first, sp.nslots - first, w0});
            for (int i = 0; i < rn; i++) ugrp[pass][w0 + i] = g;
            continue;
          }
          red.push_back({s, r0, rn, first, sp.nslots - first, j0, -2, 0});
        }
      }
    }
    if ((int)tasks.size() > off) ops.push_back({kind, off, (int)tasks.size() - off});
    if (!red.empty()) {
      off = (int)tasks.size();
      tasks.insert(tasks.end(), red.begin(), red.end());
      ops.push_back({rkind, off, (int)red.size()});
    }
  };
  // forward: levels bottom-up; fw_all also carries the Schur front.  The
  // solved pivot rows of a front stay in V, so one output launch after the
  // last level copies them all.
  for (pass = 0; pass < 2; pass++) {
    std::vector<Op>& ops = pass == 0 ? sp.fw_int : sp.fw_all;
    std::vector<int> all;
    for (int l = 0; l < y.nlevels; l++) {
      std::vector<int> list, ilist;
      for (int q = y.lvptr[l]; q < y.lvptr[l + 1]; q++) {
        int s = y.lvsup[q];
        if (s == schur && pass == 0) continue;
        list.push_back(s);
        if (s != schur) ilist.push_back(s);
      }
      add_sns(ops, FW_INIT, list);
      add_tiles(ops, FW_INV, ilist, false);
      add_tiles(ops, FW_UPD, ilist, true);
      all.insert(all.end(), list.begin(), list.end());
    }
    add_sns(ops, FW_OUT, all);
  }
  // backward: the pivot rows of every internal front from Y in one launch,
  // then levels top-down (the update-row products read the final rows of X
  // directly)
  {
    std::vector<int> all;
    for (int s = 0; s < y.nsup; s++)
      if (s != schur) all.push_back(s);
    add_sns(sp.bw, BW_INIT, all);
  }
  for (int l = y.nlevels - 1; l >= 0; l--) {
    std::vector<int> list;
    for (int q = y.lvptr[l]; q < y.lvptr[l + 1]; q++)
      if (y.lvsup[q] != schur) list.push_back(y.lvsup[q]);
    if (list.empty()) continue;
    add_tiles(sp.bw, BW_UPD, list, false);
    add_tiles(sp.bw, BW_INV, list, false);
  }
}

int DeviceReducer::analyze(const Part& p, const ReduceOptions& o, std::string& err, bool upload_now) {
  const auto t0 = std::chrono::steady_clock::now();
  part_ = p;
  opt_ = o;
  if (build_sam(part_, sam_, err)) return -1;
  // inputReducerModule.f90:403-409: NGEN may not exceed NDOF1; the number of
  // eigenpairs is at most NDOF1
  if (opt_.ngen > sam_.ndof1) {
    err = "The number of generalized DOFs (NGEN) is larger than the number of internal nodal DOFs (NDOF1): "
          "NGEN = " + std::to_string(opt_.ngen) + ", NDOF1 = " + std::to_string(sam_.ndof1) +
          ". Reduce NGEN to less than NDOF1.";
    return -1;
  }
  opt_.neval = std::min(opt_.neval, sam_.ndof1);
  if (build_symbolic(part_, sam_, opt_.sym, sym_, err)) return -1;
  d_->hexl.clear();
  d_->shl.clear();
  d_->bml.clear();
  for (int e = 0; e < part_.nel; e++) {
    if (part_.etype[e] == HEX8) d_->hexl.push_back(e);
    else if (part_.etype[e] == ANDES4) d_->shl.push_back(e);
    else if (part_.etype[e] == BEAM2) d_->bml.push_back(e);
    else {
      err = "element type " + std::to_string(part_.etype[e]) + " not supported on the device path yet";
      return -2;
    }
  }
  uploaded_ = false;
  times_.analyze_host = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  // upload_now with a device: the upload (and the device structure build and
  // the solve plans) happens here, on the caller's thread, so that a caller
  // that analyses the next part while another handle reduces overlaps all of
  // it; otherwise it waits for the first reduce()
  if (!upload_now) return 0;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) == hipSuccess && ndev > 0) return upload(err);
  (void)hipGetLastError();
  return 0;
}

// The lookahead side stream and its events, created only for a schedule that
// forks (an idle extra stream per reducer would share the hardware queues
// with the streams of other reducers).
static hipError_t ensure_side(DeviceReducer::Impl& d) {
  bool fork = false;
  for (const Op& op : d.fops) fork = fork || op.kind == FORK;
  if (!fork || d.side) return hipSuccess;
  hipError_t e = hipStreamCreateWithFlags(&d.side, hipStreamNonBlocking);
  for (auto& ev : d.fj)
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
  return e;
}

// The assembled-matrix structure built on the device from the host's
// ordering and fronts: CSR columns, the pull-assembly contribution lists
// (in a fixed order) and the front map (structure.hip).  Checks the
// contribution count against the host's and reports entries outside the
// pattern.
int DeviceReducer::build_structure(std::string& err) {
  Impl& d = *d_;
  const Symbolic& y = sym_;
  const int nnz = y.rowptr[y.neq];
  HIPCHK(d.madof.upload(sam_.madof));
  HIPCHK(d.msc.upload(sam_.msc));
  HIPCHK(d.nndof.upload(sam_.nndof));
  HIPCHK(d.ceq_ptr.upload(sam_.ceq_ptr));
  HIPCHK(d.ceq_node.upload(sam_.ceq_node));
  HIPCHK(d.ceq_dof.upload(sam_.ceq_dof));
  HIPCHK(d.ceq_coef.upload(sam_.ceq_coef));
  HIPCHK(d.drank.upload(y.drank));
  HIPCHK(d.etype.upload(part_.etype));
  HIPCHK(d.nfirst.upload(y.nfirst));
  HIPCHK(d.neqd.upload(y.neqd));
  HIPCHK(d.nbptr.upload(y.nbptr));
  HIPCHK(d.nbl.upload(y.nbl));
  HIPCHK(d.rnode.upload(y.rnode));
  HIPCHK(d.colsn.upload(y.colsn));
  HIPCHK(d.rowptr.upload(y.rowptr));
  HIPCHK(d.nbo.alloc(std::max<size_t>(1, y.nbl.size())));
  HIPCHK(d.colind.alloc(nnz));
  HIPCHK(d.cptr.alloc((size_t)nnz + 1));
  HIPCHK(d.cidx.alloc(y.ncontrib));
  if (y.has_rigid) HIPCHK(d.ccoef.alloc(y.ncontrib));
  else d.ccoef.release();
  HIPCHK(d.fmap.alloc(nnz));
  DBuf<int> cnt, tsum, bad;
  HIPCHK(cnt.alloc((size_t)nnz));
  HIPCHK(tsum.alloc((size_t)scan_tiles(nnz) + 1));
  HIPCHK(bad.alloc(1));
  hipStream_t st = st_;
  HIPCHK(hipMemsetAsync(bad.p, 0, sizeof(int), st));
  launch_nbr_sort(st, part_.nnod, d.nbptr.p, d.nbl.p, d.nfirst.p, d.neqd.p, d.nbo.p);
  launch_csr_cols(st, y.neq, d.rnode.p, d.rowptr.p, d.nbptr.p, d.nbl.p, d.nbo.p, d.nfirst.p, d.neqd.p, d.colind.p);
  double* ccoef = y.has_rigid ? d.ccoef.p : nullptr;
  for (int pass = 0; pass < 2; pass++) {
    HIPCHK(hipMemsetAsync(cnt.p, 0, sizeof(int) * (size_t)nnz, st));
    launch_contrib(st, pass, part_.nel, d.etype.p, d.eptr.p, d.enodes.p, d.eoff.p, d.madof.p, d.msc.p, d.nndof.p,
                   d.ceq_ptr.p, d.ceq_node.p, d.ceq_dof.p, d.ceq_coef.p, d.drank.p, d.nfirst.p, d.rowptr.p, d.nbptr.p,
                   d.nbl.p, d.nbo.p, cnt.p, d.cptr.p, d.cidx.p, ccoef, bad.p);
    if (pass == 0) launch_exclusive_scan(st, nnz, cnt.p, d.cptr.p, tsum.p);
  }
  launch_seg_sort(st, nnz, d.cptr.p, d.cidx.p, ccoef);
  launch_fmap(st, y.neq, d.rowptr.p, d.colind.p, d.colsn.p, d.S, d.fmap.p, bad.p);
  HIPCHK(hipGetLastError());
  int hb = 0, total = 0;
  HIPCHK(hipMemcpyAsync(&hb, bad.p, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(&total, d.cptr.p + nnz, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (hb || total != y.ncontrib) {
    err = "device structure build: " + std::string(hb & 1 ? "element contribution outside the CSR pattern" :
                                                   hb & 2 ? "matrix entry outside its front" :
                                                            "contribution count mismatch");
    return -3;
  }
  return 0;
}

// Device upload of the part and of every structure array the kernels read.
static void lanczos_sizes(int n1, int nev, const ReduceOptions& o, int& bs, int& mcap);
static SolvePlan* plan_slot(DeviceReducer::Impl& d, int nrhs);
static void plan_host(const Symbolic& y, SolvePlan& sp, PlanHost& h, int split);
static int plan_upload(SolvePlan& sp, const PlanHost& h, std::string& err);

int DeviceReducer::upload(std::string& err) {
  const auto t0 = std::chrono::steady_clock::now();
  // FGPU_TRACE_REDUCE=1: host-side timeline of the upload on stderr
  const bool trace = std::getenv("FGPU_TRACE_REDUCE") != nullptr;
  auto mark = [&](const char* what) {
    if (trace)
      std::fprintf(stderr, "upload +%.2f ms: %s\n",
                   std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count(), what);
  };
  Impl& d = *d_;
  if (!st_) {
    // non-blocking: no implicit ordering with the null stream (and so with
    // other threads' null-stream copies); default priority
    HIPCHK(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
    for (auto& e : d.ev) HIPCHK(hipEventCreate(&e));
  }
  StreamScope scope(st_);
  HIPCHK(d.xyz.upload(part_.xyz));
  HIPCHK(d.eprop.upload(part_.eprop));
  HIPCHK(d.eptr.upload(part_.eptr));
  HIPCHK(d.enodes.upload(part_.enodes));
  HIPCHK(d.hexlist.upload(d.hexl));
  {
    std::vector<int> p1(sym_.ndof1);  // permuted row of every internal equation, in MEQN1 order
    for (int q = 0; q < sym_.ndof1; q++) p1[q] = sym_.eq_to_perm[sam_.meqn1[q] - 1];
    HIPCHK(d.p1.upload(p1));
  }
  HIPCHK(d.shlist.upload(d.shl));
  HIPCHK(d.bmlist.upload(d.bml));
  HIPCHK(d.eoff.upload(sym_.elem_off));
  mark("part arrays");
  const Symbolic& y = sym_;
  HIPCHK(d.col0.upload(y.col0));
  HIPCHK(d.ncol.upload(y.ncol));
  HIPCHK(d.nrow.upload(y.nrow));
  HIPCHK(d.rptr.upload(y.rptr));
  HIPCHK(d.rows.upload(y.rows));
  HIPCHK(d.chptr.upload(y.chptr));
  HIPCHK(d.chlist.upload(y.chlist));
  HIPCHK(d.relptr.upload(y.relptr));
  HIPCHK(d.rel.upload(y.rel));
  HIPCHK(d.foff.upload(y.foff));
  std::vector<int64_t> woff(y.nsup + 1, 0);
  for (int s = 0; s < y.nsup; s++) woff[s + 1] = woff[s] + y.nrow[s];
  d.wrows = woff[y.nsup];
  HIPCHK(d.woff.upload(woff));
  std::vector<int64_t> ioff(y.nsup + 1, 0);
  for (int s = 0; s < y.nsup; s++) ioff[s + 1] = ioff[s] + (int64_t)y.ncol[s] * y.ncol[s];
  HIPCHK(d.ioff.upload(ioff));
  d.mixed = opt_.mixed;
  d.m_lumped = d.hexl.empty() && d.bml.empty();
  d.solve_split = std::max(64, (opt_.solve_split / 64) * 64);
  if (d.mixed) {
    d.ipool.release();
    HIPCHK(d.ipool32.alloc(ioff[y.nsup]));
  } else {
    d.ipool32.release();
    HIPCHK(d.ipool.alloc(ioff[y.nsup]));
  }
  HIPCHK(d.dpiv.alloc((size_t)std::max(1, y.ndof1)));
  HIPCHK(d.adiag.alloc((size_t)std::max(1, y.ndof1)));
  HIPCHK(d.amax.alloc(1));
  HIPCHK(d.sing.alloc((size_t)std::max(1, y.ndof1)));
  d.S = DevSym{d.col0.p, d.ncol.p, d.nrow.p, d.foff.p, d.rptr.p, d.rows.p, d.chptr.p,
               d.chlist.p, d.relptr.p, d.rel.p, d.woff.p, d.dpiv.p,
               d.adiag.p, d.amax.p, d.sing.p, std::max(opt_.tol_factorize, 1e-20)};
  // The host-side schedules (the factorisation's and the solve plans the
  // reduction uses) are built on host threads while the device builds the
  // structure.  The solve plans keep their device buffers: freeing them
  // would wait for the whole device.
  for (auto& kv : d.plans) kv.second->valid = false;
  std::vector<int> want;
  {
    const int n1 = y.ndof1, n2 = y.ndof2;
    const int nev = opt_.ngen > 0 ? std::max(opt_.neval, opt_.ngen) : 0;
    if (n1 > 0 && n2 > 0) want.push_back(n2);
    if (n1 > 0 && nev > 0) {
      int bs = 0, mcap = 0;
      lanczos_sizes(n1, nev, opt_, bs, mcap);
      if (std::find(want.begin(), want.end(), bs) == want.end()) want.push_back(bs);
    }
    if (n1 > 0 && opt_.want_vgi && std::find(want.begin(), want.end(), 3) == want.end()) want.push_back(3);
  }
  std::vector<SolvePlan*> sps;
  for (int nr : want) sps.push_back(plan_slot(d, nr));
  std::vector<PlanHost> phs(sps.size());
  std::vector<FTask> ft;
  d.fops.clear();
  std::vector<int> xr;
  int inv_slots = 0;
  const int panel = std::min(256, std::max(64, (opt_.factor_panel / 64) * 64));
  std::vector<std::future<void>> host_jobs;
  host_jobs.push_back(std::async(std::launch::async, [&]() {
    build_factor_schedule(y, ft, d.fops, xr, upd_flops_, panel, inv_slots, opt_.factor_lookahead);
  }));
  for (size_t i = 0; i < sps.size(); i++)
    host_jobs.push_back(std::async(std::launch::async, [&, i]() { plan_host(y, *sps[i], phs[i], d.solve_split); }));
  const int src = build_structure(err);
  for (auto& f : host_jobs) f.get();
  if (src) return src;
  mark("structure built");
  for (size_t i = 0; i < sps.size(); i++)
    if (int rc = plan_upload(*sps[i], phs[i], err)) return rc;
  HIPCHK(ensure_side(d));
  HIPCHK(d.ftasks.upload(ft));
  mark("factor schedule built");
  HIPCHK(d.xrange.upload(xr));
  HIPCHK(d.ipart.alloc((size_t)std::max(1, inv_slots) * 4096));
  upd_launches_ = 0;
  for (auto& op : d.fops)
    if (op.kind == UPDATE || op.kind == UPDATE_SIDE) upd_launches_++;
  // a new structure: every captured launch sequence is stale
  d.g_factor.reset();
  d.factor_runs = 0;
  d.g_lanczos.reset();
  // value buffers
  const int64_t nnz = y.rowptr[y.neq];
  HIPCHK(d.Kp.alloc(y.elem_pool));
  HIPCHK(d.Mp.alloc(y.elem_pool));
  HIPCHK(d.Kv.alloc(nnz));
  HIPCHK(d.Mv.alloc(nnz));
  if (d.mixed) {
    d.pool.release();
    HIPCHK(d.pool32.alloc(y.fpool));
    HIPCHK(d.dscale.alloc(y.neq));
    HIPCHK(d.dcs.alloc(std::max(1, y.ndof2)));
    HIPCHK(d.irpart.alloc(2 * SUMSQ_BLOCKS));
  } else {
    d.pool32.release();
    HIPCHK(d.pool.alloc(y.fpool));
  }
  HIPCHK(d.ebad.alloc(part_.nel));
  HIPCHK(d.emass.alloc(part_.nel + 1));
  HIPCHK(d.piv.alloc(6));
  mark("value buffers");
  // gravity directions: unit accelerations on translational equation DOFs
  std::vector<double> a((size_t)y.neq * 3, 0.0);
  for (int i = 0; i < sam_.nnod; i++)
    for (int j = 0; j < 3 && j < sam_.nndof[i]; j++) {
      int dof = sam_.madof[i] - 1 + j;
      if (sam_.meqn[dof] > 0) a[(size_t)y.eq_to_perm[sam_.meqn[dof] - 1] * 3 + j] = 1.0;
    }
  HIPCHK(d.grav.upload(a));
  HIPCHK(d.p2e.upload(y.perm_to_eq));
  HIPCHK(d.pos2.upload(sam_.dofPosIn2.empty() ? std::vector<int>(1, 1) : sam_.dofPosIn2));
  mark("gravity and maps");
  // the solve plans and work buffers the reduction will use, made here (on
  // the caller's analysis thread) instead of on the reduction's path
  if (int rc = prepare_solves(err)) return rc;
  mark("solve plans");
  uploaded_ = true;
  times_.upload = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return 0;
}

// The slot of the plan for nrhs right-hand sides: the existing plan (its
// device buffers reused, its host lists cleared) or a new one in the map.
static SolvePlan* plan_slot(DeviceReducer::Impl& d, int nrhs) {
  auto it = d.plans.find(nrhs);
  SolvePlan* sp = nullptr;
  if (it != d.plans.end()) {
    sp = it->second.get();
    sp->nslots = 0;
    sp->fw_int.clear();
    sp->fw_all.clear();
    sp->bw.clear();
  } else {
    auto fresh = std::make_unique<SolvePlan>();
    sp = fresh.get();
    d.plans[nrhs] = std::move(fresh);
  }
  sp->valid = false;
  sp->nrhs = nrhs;
  return sp;
}

// Host part of a plan (pure host work, safe on another thread).
static void plan_host(const Symbolic& y, SolvePlan& sp, PlanHost& h, int split) {
  build_solve_plan(y, sp, h.tasks, h.gptr, h.gsrc, h.ugrp, h.ugroups, split);
}

static int plan_upload(SolvePlan& sp, const PlanHost& h, std::string& err) {
  HIPCHK(sp.tasks.upload(h.tasks));
  HIPCHK(sp.gptr.upload(h.gptr));
  HIPCHK(sp.gsrc.upload(h.gsrc.empty() ? std::vector<int>(1, 0) : h.gsrc));
  HIPCHK(sp.ugrp_int.upload(h.ugrp[0]));
  HIPCHK(sp.ugrp_all.upload(h.ugrp[1]));
  HIPCHK(sp.ugroups.upload(h.ugroups.empty() ? std::vector<int>(3, 0) : h.ugroups));
  HIPCHK(sp.part.alloc((size_t)std::max(1, sp.nslots) * 4096));
  sp.valid = true;
  return 0;
}

static int get_plan(const Symbolic& y, DeviceReducer::Impl& d, int nrhs, SolvePlan*& out, std::string& err) {
  auto it = d.plans.find(nrhs);
  if (it != d.plans.end() && it->second->valid) { out = it->second.get(); return 0; }
  SolvePlan* sp = plan_slot(d, nrhs);
  PlanHost h;
  plan_host(y, *sp, h, d.solve_split);
  if (int rc = plan_upload(*sp, h, err)) return rc;
  out = sp;
  return 0;
}

static int ensure_work(DeviceReducer::Impl& d, int nrhs, std::string& err) {
  HIPCHK(d.W.alloc((size_t)d.wrows * nrhs));
  HIPCHK(d.V.alloc((size_t)d.wrows * nrhs));
  return 0;
}

// Lanczos block size and basis capacity for nev modes of n1 internal
// equations: blocks of lanczos_block vectors (<= 64, the Gram matrices' limit),
// lanczosMaxVectors or max(10 nev + 40, 16 blocks) vectors, at least nev, in
// whole blocks; a basis that would span the whole internal space takes a
// block size dividing n1, so that it does so exactly.
static void lanczos_sizes(int n1, int nev, const ReduceOptions& o, int& bs, int& mcap) {
  bs = std::max(1, std::min(std::min(o.lanczos_block, 64), n1));
  mcap = o.lanczos_maxvec > 0 ? o.lanczos_maxvec : std::max(10 * nev + 40, 16 * bs);
  mcap = (std::max(mcap, nev) + bs - 1) / bs * bs;
  if (mcap >= n1) {
    while (n1 % bs) bs--;
    mcap = n1;
  }
}

int DeviceReducer::prepare_solves(std::string& err) {
  Impl& d = *d_;
  const Symbolic& y = sym_;
  const int n1 = y.ndof1, n2 = y.ndof2, neq = y.neq;
  const int nev = opt_.ngen > 0 ? std::max(opt_.neval, opt_.ngen) : 0;
  if (n1 <= 0) return 0;
  SolvePlan* p = nullptr;
  if (n2 > 0) {
    if (get_plan(y, d, n2, p, err)) return -3;
    if (ensure_work(d, n2, err)) return -3;
    HIPCHK(d.X.alloc((size_t)neq * n2));
    HIPCHK(d.Z.alloc((size_t)neq * n2));
    HIPCHK(d.Y.alloc((size_t)neq * n2));
  }
  if (nev > 0) {
    int bs = 0, mcap = 0;
    lanczos_sizes(n1, nev, opt_, bs, mcap);
    if (get_plan(y, d, bs, p, err)) return -3;
    if (ensure_work(d, std::max(bs, n2), err)) return -3;
    HIPCHK(d.Q.alloc((size_t)neq * mcap));
    HIPCHK(d.P.alloc((size_t)neq * mcap));
    HIPCHK(d.Wl.alloc((size_t)neq * bs));
    HIPCHK(d.Zl.alloc((size_t)neq * bs));
    HIPCHK(d.Yl.alloc((size_t)neq * bs));
    HIPCHK(d.Tmp.alloc((size_t)neq * bs));
    HIPCHK(d.Cc.alloc((size_t)mcap * std::max(bs, nev) + 64));
    HIPCHK(d.part.alloc((size_t)xty_chunks(neq) * mcap * std::max(bs, nev)));
    HIPCHK(d.small.alloc((size_t)mcap * mcap + 64));
    HIPCHK(d.Cacc.alloc((size_t)mcap * bs));
    HIPCHK(d.Rdev.alloc((size_t)2 * bs * bs));
    HIPCHK(d.signs.alloc((size_t)nev * (1 + 3 * MODE_SIGN_GROUPS)));
  }
  if (opt_.want_vgi && get_plan(y, d, 3, p, err)) return -3;
  return 0;
}

static void run_mv_op(hipStream_t st, DeviceReducer::Impl& d, const SolvePlan& sp, const Op& op, double* X) {
  static const int mode_of[] = {0, 1, 2, 3};
  const SolveTask* t = sp.tasks.p + op.off;
  switch (op.kind) {
    case FW_INV: case FW_UPD: case BW_UPD: case BW_INV: {
      int mode = op.kind == FW_INV ? 0 : op.kind == FW_UPD ? 1 : op.kind == BW_UPD ? 2 : 3;
      if (d.mixed)
        launch_mv(st, t, op.n, d.S, d.pool32.p, d.ioff.p, d.ipool32.p, d.W.p, d.V.p, X, sp.part.p, sp.nrhs,
                  mode_of[mode]);
      else
        launch_mv(st, t, op.n, d.S, d.pool.p, d.ioff.p, d.ipool.p, d.W.p, d.V.p, X, sp.part.p, sp.nrhs,
                  mode_of[mode]);
      break;
    }
    case FW_INV_R: case FW_UPD_R: case BW_UPD_R: case BW_INV_R: {
      int mode = op.kind == FW_INV_R ? 0 : op.kind == FW_UPD_R ? 1 : op.kind == BW_UPD_R ? 2 : 3;
      launch_mv_red(st, t, op.n, d.S, d.W.p, d.V.p, X, sp.part.p, sp.nrhs, mode);
      break;
    }
  }
}

// Y = L^{-1} R (scale: D^{-1} L^{-1} R); all: also the Schur front (reduced RHS rows)
static void run_forward(hipStream_t st, DeviceReducer::Impl& d, const SolvePlan& sp, const double* R, double* Yout,
                        int ndof1, bool scale, bool all) {
  const std::vector<Op>& ops = all ? sp.fw_all : sp.fw_int;
  for (const Op& op : ops) {
    switch (op.kind) {
      case FW_INIT:
        launch_fw_init(st, sp.tasks.p + op.off, op.n, d.S, sp.gptr.p, sp.gsrc.p, all ? sp.ugrp_all.p : sp.ugrp_int.p,
                       sp.ugroups.p, sp.part.p, R, d.W.p, sp.nrhs, ndof1);
        break;
      case FW_OUT:
        if (d.mixed)
          launch_fw_out(st, sp.tasks.p + op.off, op.n, d.S, d.pool32.p, d.W.p, d.V.p, Yout, sp.nrhs, scale, ndof1);
        else
          launch_fw_out(st, sp.tasks.p + op.off, op.n, d.S, d.pool.p, d.W.p, d.V.p, Yout, sp.nrhs, scale, ndof1);
        break;
      default: run_mv_op(st, d, sp, op, nullptr); break;
    }
  }
}

// X = L^{-T} Y on the internal rows; X's boundary rows must be set by the caller
static void run_backward(hipStream_t st, DeviceReducer::Impl& d, const SolvePlan& sp, const double* Yin, double* X) {
  for (const Op& op : sp.bw) {
    if (op.kind == BW_INIT) launch_bw_init(st, sp.tasks.p + op.off, op.n, d.S, Yin, d.W.p, sp.nrhs);
    else run_mv_op(st, d, sp, op, X);
  }
}

// Relative size ||D|| / ||X|| of a correction D against the solution X
// over their first n entries (device partial sums, added in block order).
static int rel_correction(hipStream_t st, DeviceReducer::Impl& d, int64_t n, const double* D, const double* X,
                          double& rel, std::string& err) {
  launch_sumsq2(st, n, D, X, d.irpart.p);
  double h[2 * SUMSQ_BLOCKS];
  HIPCHK(hipMemcpyAsync(h, d.irpart.p, sizeof h, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  double a = 0.0, b = 0.0;
  for (int q = 0; q < SUMSQ_BLOCKS; q++) {
    a += h[2 * q];
    b += h[2 * q + 1];
  }
  rel = b > 0.0 ? std::sqrt(a / b) : (a > 0.0 ? INFINITY : 0.0);
  return 0;
}

// Stopping rule of the mixed-precision refinement, from the relative size of
// the latest correction: done when it is below tol, or when it no longer
// halves (the residual has reached FP64 rounding level, where the
// correction is noise of the size cond(K) * eps64); failure when it stays
// large past the step limit, or above 0.1 after 12 steps.
static int refine_state(double rel, double prev, int steps, const ReduceOptions& opt, std::string& err,
                        const char* what) {
  if (rel <= opt.ir_tol) return 1;
  if (steps >= 2 && rel > 0.5 * prev && rel < 1e-6) return 1;
  // a correction still above 0.1 after 12 steps has contracted by less than
  // 0.83 a step: it cannot reach ir_tol within the step limit, so stop now
  if (steps >= opt.ir_maxit || !std::isfinite(rel) || (steps >= 12 && rel > 0.1)) {
    char buf[160];
    std::snprintf(buf, sizeof buf, "mixed-precision iterative refinement of %s did not converge "
                  "(correction %.3e after %d steps)", what, rel, steps);
    err = buf;
    return -1;
  }
  return 0;
}

// One approximate solve D = K_ii^{-1} Rin on the mixed-precision factor:
// D = S (K~_ii)^{-1} S Rin with K~ = S K S the equilibrated matrix whose FP32
// LDL^T the fronts hold.  Rin (neq x nrhs, internal rows read) is scaled in
// place; D's boundary rows are zero.
static void approx_solve(hipStream_t st, DeviceReducer::Impl& d, const SolvePlan& sp, int neq, int ndof1,
                         double* Rin, double* Yw, double* D) {
  launch_scale_rc(st, ndof1, sp.nrhs, d.dscale.p, nullptr, Rin);
  run_forward(st, d, sp, Rin, Yw, ndof1, true, false);
  launch_fill(st, D + (size_t)ndof1 * sp.nrhs, (int64_t)(neq - ndof1) * sp.nrhs, 0.0);  // boundary rows
  run_backward(st, d, sp, Yw, D);
  launch_scale_rc(st, ndof1, sp.nrhs, d.dscale.p, nullptr, D);
}

// Mixed precision: conjugate gradients on K_ii X = b, one independent CG
// per column, preconditioned by the FP32 factor (z = K~_ii^-1 r through
// approx_solve), residuals and updates in FP64.  The stationary refinement
// X += K~^-1 r contracts each step by the spread rho of the preconditioned
// spectrum about 1; CG by about rho / (1 + sqrt(1 - rho^2)) and faster
// where the spectrum clusters.  On entry X holds the starting solution and
// R (neq x nr, internal rows) its residual b - K_ii X; X's boundary rows
// are left alone.  A column whose curvature p^T K_ii p or r^T z is not
// positive (an indefinite K_ii or factor) turns the remaining steps into
// stationary refinement for all columns.  Stopping: refine_state on the
// relative size of the step's correction.
static int pcg_int(hipStream_t st, DeviceReducer::Impl& d, const Symbolic& y, const ReduceOptions& opt,
                   const SolvePlan& sp, double* X, double* R, double* Yw, int* steps, const char* what,
                   std::string& err) {
  const int neq = y.neq, n1 = y.ndof1, nr = sp.nrhs;
  const int64_t nall = (int64_t)neq * nr, nint = (int64_t)n1 * nr;
  HIPCHK(d.cg_p.alloc(nall));
  HIPCHK(d.cg_q.alloc(nall));
  HIPCHK(d.cg_z.alloc(nall));
  HIPCHK(d.cg_t.alloc(nall));
  HIPCHK(d.cg_part.alloc((size_t)std::max(1, coldot_blocks(n1)) * 3 * nr));
  HIPCHK(d.cg_dot.alloc((size_t)3 * nr));
  HIPCHK(d.cg_co.alloc((size_t)3 * nr));
  std::vector<double> h((size_t)3 * nr), rz(nr), xx(nr), co((size_t)3 * nr);
  // column dot products of npair pairs over the internal rows, to h
  auto dots = [&](int npair, const double* A0, const double* B0, const double* A1, const double* B1,
                  const double* A2, const double* B2) -> int {
    launch_coldots(st, n1, nr, npair, A0, B0, A1, B1, A2, B2, d.cg_part.p, d.cg_dot.p);
    HIPCHK(hipMemcpyAsync(h.data(), d.cg_dot.p, sizeof(double) * npair * nr, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return 0;
  };
  // z = K~^-1 r (approx_solve scales its input in place: it gets a copy)
  auto precond = [&]() -> int {
    HIPCHK(hipMemcpyAsync(d.cg_t.p, R, sizeof(double) * nint, hipMemcpyDeviceToDevice, st));
    approx_solve(st, d, sp, neq, n1, d.cg_t.p, Yw, d.cg_z.p);
    return 0;
  };
  bool stationary = false;
  double prev = INFINITY;
  if (int rc = precond()) return rc;
  if (int rc = dots(2, R, d.cg_z.p, X, X, nullptr, nullptr)) return rc;
  for (int c = 0; c < nr; c++) {
    rz[c] = h[c];
    xx[c] = h[nr + c];  // |X_c|^2, carried through the updates for the stopping rule
    if (!(rz[c] >= 0.0)) stationary = true;
  }
  // p = z (boundary rows zero, so K p has K_ii p in its internal rows)
  HIPCHK(hipMemcpyAsync(d.cg_p.p, d.cg_z.p, sizeof(double) * nall, hipMemcpyDeviceToDevice, st));
  for (int it = 0;; it++) {
    double rel = 0.0;
    if (stationary) {
      // X += z; r = b - K_ii X = r - K_ii z
      launch_axpby(st, nint, 1.0, d.cg_z.p, 1.0, X);
      launch_spmm(st, neq, nr, d.rowptr.p, d.colind.p, d.Kv.p, d.cg_z.p, nr, d.cg_q.p, nr);
      launch_axpby(st, nint, -1.0, d.cg_q.p, 1.0, R);
      if (rel_correction(st, d, nint, d.cg_z.p, X, rel, err)) return -3;
    } else {
      launch_spmm(st, neq, nr, d.rowptr.p, d.colind.p, d.Kv.p, d.cg_p.p, nr, d.cg_q.p, nr);
      if (int rc = dots(3, d.cg_p.p, d.cg_q.p, d.cg_p.p, d.cg_p.p, d.cg_p.p, X)) return rc;
      double num = 0.0, den = 0.0;
      for (int c = 0; c < nr; c++) {
        const double pq = h[c], pp = h[nr + c], px = h[2 * nr + c];
        double a = 0.0;
        if (rz[c] > 0.0) {
          if (pq > 0.0) a = rz[c] / pq;
          else stationary = true;
        }
        co[c] = a;
        co[nr + c] = -a;
        num += a * a * pp;
        xx[c] += a * (2.0 * px + a * pp);  // |X_c + a p_c|^2
        den += xx[c];
      }
      if (stationary) {  // fall back from the current X and r: redo this step as a stationary one
        it--;
        continue;
      }
      HIPCHK(hipMemcpyAsync(d.cg_co.p, co.data(), sizeof(double) * 2 * nr, hipMemcpyHostToDevice, st));
      launch_col_axpby(st, n1, nr, d.cg_co.p, d.cg_p.p, nullptr, X);           // X += a p
      launch_col_axpby(st, n1, nr, d.cg_co.p + nr, d.cg_q.p, nullptr, R);      // r -= a K p
      rel = den > 0.0 ? std::sqrt(num / den) : (num > 0.0 ? INFINITY : 0.0);
    }
    if (steps) *steps = it + 1;
    if (std::getenv("FGPU_TRACE_REFINE"))  // FGPU_TRACE_REFINE=1: each step's correction on stderr
      std::fprintf(stderr, "refine %s step %d: |dX|/|X| %.3e%s\n", what, it + 1, rel,
                   stationary ? " (stationary)" : "");
    const int state = refine_state(rel, prev, it + 1, opt, err, what);
    if (state > 0) return 0;
    if (state < 0) return -6;
    prev = rel;
    if (int rc = precond()) return rc;
    if (stationary) continue;
    if (int rc = dots(1, R, d.cg_z.p, nullptr, nullptr, nullptr, nullptr)) return rc;
    for (int c = 0; c < nr; c++) {
      const double rzn = h[c];
      if (!(rzn >= 0.0)) stationary = true;
      co[2 * nr + c] = rz[c] > 0.0 ? rzn / rz[c] : 0.0;
      rz[c] = rzn;
    }
    if (stationary) continue;
    HIPCHK(hipMemcpyAsync(d.cg_co.p + 2 * nr, co.data() + 2 * nr, sizeof(double) * nr, hipMemcpyHostToDevice, st));
    launch_col_axpby(st, n1, nr, nullptr, d.cg_z.p, d.cg_co.p + 2 * nr, d.cg_p.p);  // p = z + beta p
  }
}

// X = K_ii^{-1} R on the internal rows (X's boundary rows zero), R and X
// neq x nrhs row-major in perm order; Yw is neq x nrhs work.  FP64 factor:
// one forward and one backward pass.  Mixed precision: conjugate gradients
// preconditioned by the FP32 factor with residuals from the FP64 CSR
// matrix (pcg_int), until the correction is below ir_tol relative to X.
static int solve_int(hipStream_t st, DeviceReducer::Impl& d, const Symbolic& y, const ReduceOptions& opt,
                     const SolvePlan& sp, const double* R, double* Yw, double* X, int* steps, std::string& err) {
  const int neq = y.neq, n1 = y.ndof1, nr = sp.nrhs;
  const int64_t nall = (int64_t)neq * nr, nint = (int64_t)n1 * nr;
  if (!d.mixed) {
    run_forward(st, d, sp, R, Yw, n1, true, false);
    launch_fill(st, X + nint, nall - nint, 0.0);  // boundary rows; the backward pass writes the others
    run_backward(st, d, sp, Yw, X);
    if (steps) *steps = 0;
    return 0;
  }
  // conjugate gradients from X = 0 (residual R)
  HIPCHK(d.Rr.alloc(nall));
  HIPCHK(hipMemcpyAsync(d.Rr.p, R, sizeof(double) * nall, hipMemcpyDeviceToDevice, st));
  launch_fill(st, X, nall, 0.0);
  return pcg_int(st, d, y, opt, sp, X, d.Rr.p, Yw, steps, "a K_ii solve", err);
}

static double ev_ms(hipEvent_t a, hipEvent_t b) {
  float ms = 0;
  (void)hipEventElapsedTime(&ms, a, b);
  return ms;
}

int DeviceReducer::reduce(std::string& err) {
  if (!uploaded_) {
    int rc = upload(err);
    if (rc) return rc;
  }
  StreamScope scope(st_);
  // FGPU_TRACE_REDUCE=1: host-side timeline of the phases on stderr
  const bool trace = std::getenv("FGPU_TRACE_REDUCE") != nullptr;
  const auto tr0 = std::chrono::steady_clock::now();
  auto mark = [&](const char* what) {
    if (trace)
      std::fprintf(stderr, "reduce +%.2f ms: %s\n",
                   std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tr0).count(), what);
  };
  Impl& d = *d_;
  const Symbolic& y = sym_;
  const int neq = y.neq, n1 = y.ndof1, n2 = y.ndof2, schur = y.nsup - 1;
  const int64_t nnz = y.rowptr[neq];
  const int nev = opt_.ngen > 0 ? std::max(opt_.neval, opt_.ngen) : 0;
  const int ngen = opt_.ngen;
  hipStream_t st = st_;
  res_ = Results();
  times_.refine_steps = 0;
  res_.ndof1 = n1;
  res_.ndof2 = n2;
  res_.ngen = ngen;
  res_.nev = nev;
  res_.ndim = sam_.ndim_ext + ngen;
  HIPCHK(hipEventRecord(d.ev[0], st));
  mark("start issued");
  // ---- element matrices
  HIPCHK(hipMemsetAsync(d.ebad.p, 0, sizeof(int) * part_.nel, st));
  launch_hex8(st, (int)d.hexl.size(), d.hexlist.p, d.eptr.p, d.enodes.p, d.xyz.p, d.eprop.p, d.eoff.p, d.Kp.p,
              d.Mp.p, d.ebad.p);
  launch_andes4(st, (int)d.shl.size(), d.shlist.p, d.eptr.p, d.enodes.p, d.xyz.p, d.eprop.p, d.eoff.p, d.Kp.p,
                d.Mp.p, d.ebad.p);
  launch_beam(st, (int)d.bml.size(), d.bmlist.p, d.eptr.p, d.enodes.p, d.xyz.p, d.eprop.p, d.eoff.p, d.Kp.p,
              d.Mp.p, d.ebad.p);
  if (!opt_.no_mass) {  // total FE mass (SMASS of lhsINADD) for the reduced-mass check
    launch_emass(st, (int)d.hexl.size(), d.hexlist.p, 8, 3, d.eoff.p, d.Mp.p, d.emass.p);
    launch_emass(st, (int)d.shl.size(), d.shlist.p, 4, 6, d.eoff.p, d.Mp.p, d.emass.p);
    launch_emass(st, (int)d.bml.size(), d.bmlist.p, 2, 6, d.eoff.p, d.Mp.p, d.emass.p);
    launch_sum_fixed(st, part_.nel, d.emass.p, d.emass.p + part_.nel);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(d.ev[1], st));
  // ---- assembly (CSR K, M and the K entries of the fronts)
  // fronts start at zero (eager, outside any graph capture)
  if (d.mixed) HIPCHK(hipMemsetAsync(d.pool32.p, 0, sizeof(float) * y.fpool, st));
  else HIPCHK(hipMemsetAsync(d.pool.p, 0, sizeof(double) * y.fpool, st));
  launch_assemble(st, nnz, d.cptr.p, d.cidx.p, y.has_rigid ? d.ccoef.p : nullptr, d.Kp.p, d.Mp.p, d.Kv.p, d.Mv.p,
                  d.fmap.p, d.mixed ? nullptr : d.pool.p);
  if (d.mixed) {  // equilibrated stiffness into the FP32 fronts
    launch_diag_scale(st, neq, d.rowptr.p, d.colind.p, d.Kv.p, d.dscale.p);
    launch_front_scatter32(st, neq, d.rowptr.p, d.colind.p, d.Kv.p, d.dscale.p, d.fmap.p, d.pool32.p);
  }
  // original diagonal of K_ii for the singular-pivot test (equilibrated for
  // the mixed-precision factor), no pivot flagged yet
  HIPCHK(hipMemsetAsync(d.amax.p, 0, sizeof(double), st));
  HIPCHK(hipMemsetAsync(d.sing.p, 0, sizeof(int) * (size_t)std::max(1, n1), st));
  launch_diag_extract(st, n1, d.rowptr.p, d.colind.p, d.Kv.p, d.mixed ? d.dscale.p : nullptr, d.adiag.p, d.amax.p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(d.ev[2], st));
  mark("elements and assembly issued");
  // ---- multifrontal LDL^T; the Schur front ends up holding K_bb - K_bi K_ii^-1 K_ib
  {
    int init[6] = {0, 0, 0, INT_MAX, INT_MAX, INT_MAX};
    HIPCHK(hipMemcpyAsync(d.piv.p, init, sizeof init, hipMemcpyHostToDevice, st));
  }
  if (opt_.time_updates && d.upd_ev.size() < 2 * (size_t)upd_launches_) {  // another schedule may launch more
    const size_t have = d.upd_ev.size();
    d.upd_ev.resize(2 * (size_t)upd_launches_);
    for (size_t q = have; q < d.upd_ev.size(); q++) HIPCHK(hipEventCreate(&d.upd_ev[q]));
  }
  int64_t ui = 0;
  hipError_t fork_err = hipSuccess;  // first failure of a fork or join of the side stream
  auto note = [&](hipError_t e) {
    if (e != hipSuccess && fork_err == hipSuccess) fork_err = e;
  };
  auto factor_with = [&](auto* pool, auto* ipool) {
    for (const Op& op : d.fops) {
      switch (op.kind) {
        case EXTADD: launch_extend_add(st, d.ftasks.p + op.off, op.n, d.S, d.xrange.p, pool); break;
        case DIAG: launch_diag(st, d.ftasks.p + op.off, op.n, d.S, pool, d.ioff.p, ipool, d.piv.p); break;
        case TRSM: launch_trsm(st, d.ftasks.p + op.off, op.n, d.S, pool, d.ioff.p, ipool); break;
        case DIAG_TRSM: launch_diag_trsm(st, d.ftasks.p + op.off, op.n, d.S, pool, d.ioff.p, ipool, d.piv.p); break;
        case UPDATE:
        case UPDATE_SIDE: {
          // with per-launch timing everything stays on the main stream
          hipStream_t us = (op.kind == UPDATE_SIDE && !opt_.time_updates) ? d.side : st;
          if (opt_.time_updates) (void)hipEventRecord(d.upd_ev[2 * ui], st);
          launch_update(us, d.ftasks.p + op.off, op.n, d.S, pool);
          if (opt_.time_updates) (void)hipEventRecord(d.upd_ev[2 * ui + 1], st);
          ui++;
          break;
        }
        case FORK:
          if (!opt_.time_updates) {
            note(hipEventRecord(d.fj[0], st));
            note(hipStreamWaitEvent(d.side, d.fj[0], 0));
          }
          break;
        case JOIN:
          if (!opt_.time_updates) {
            note(hipEventRecord(d.fj[1], d.side));
            note(hipStreamWaitEvent(st, d.fj[1], 0));
          }
          break;
        case INV_OFF: launch_inv_off(st, d.ftasks.p + op.off, op.n, d.S, pool, d.ioff.p, ipool); break;
        case INV_PART:
          launch_inv_part(st, d.ftasks.p + op.off, op.n, d.S, pool, d.ioff.p, ipool,
                          reinterpret_cast<decltype(pool)>(d.ipart.p));
          break;
        case INV_FIN:
          launch_inv_fin(st, d.ftasks.p + op.off, op.n, d.S, d.ioff.p, ipool,
                         reinterpret_cast<decltype(pool)>(d.ipart.p));
          break;
      }
    }
  };
  auto factor_ops = [&]() {
    if (d.mixed) factor_with(d.pool32.p, d.ipool32.p);
    else factor_with(d.pool.p, d.ipool.p);
  };
  // The first factorisation after an upload runs eagerly (a part reduced
  // once gains nothing from a capture); later ones replay one captured graph.
  if (opt_.time_updates || !opt_.use_graphs || d.factor_runs++ == 0) {
    factor_ops();  // per-launch events need the plain stream
  } else {
    // the factor schedule is fixed per symbolic structure: one graph
    HIPCHK(run_graph(st, d.g_factor,
                     {d.ftasks.p, d.xrange.p, d.ipart.p, d.pool.p, d.ipool.p, d.pool32.p, d.ipool32.p, d.ioff.p,
                      d.piv.p, d.S.foff, d.S.rows, d.S.dpiv},
                     factor_ops));
  }
  HIPCHK(fork_err);
  launch_pivot_stats(st, n1, d.dpiv.p, d.sing.p, d.piv.p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(d.ev[3], st));
  mark("factor issued");
  // ---- B = -K_ii^{-1} K_ib: back-substitution of [0; I].  With the FP64
  // factor and component modes it is queued after the Lanczos steps instead,
  // where it runs on the device while the host solves the final projected
  // eigenproblem (events 8 and 9 bracket it there).
  SolvePlan* pb = nullptr;
  const bool defer_b = n2 > 0 && !d.mixed && nev > 0;
  if (n2 > 0) {
    if (get_plan(y, d, n2, pb, err)) return -3;
    if (ensure_work(d, n2, err)) return -3;
    HIPCHK(d.X.alloc((size_t)neq * n2));
    HIPCHK(d.Z.alloc((size_t)neq * n2));
    HIPCHK(d.Y.alloc((size_t)neq * n2));
  }
  // the deferred pass also forms the mass phase's Z = M [B; I] (events 9-11)
  const bool defer_z = defer_b && !opt_.no_mass;
  auto issue_bsolve = [&]() -> int {
    HIPCHK(hipEventRecord(d.ev[8], st));
    launch_identity_rows(st, d.X.p + (size_t)n1 * n2, neq - n1, n2, 0);  // the backward pass writes rows < n1
    run_backward(st, d, *pb, nullptr, d.X.p);
    HIPCHK(hipEventRecord(d.ev[9], st));
    if (defer_z) {
      launch_spmm(st, neq, n2, d.rowptr.p, d.colind.p, d.Mv.p, d.X.p, n2, d.Z.p, n2, d.m_lumped);
      HIPCHK(hipEventRecord(d.ev[11], st));
    }
    return 0;
  };
  if (n2 > 0 && !defer_b) {
    launch_identity_rows(st, d.X.p + (size_t)n1 * n2, neq - n1, n2, 0);  // the backward pass writes rows < n1
    run_backward(st, d, *pb, nullptr, d.X.p);
    if (d.mixed) {
      // the backward pass of the equilibrated factor gives S_i^-1 B S_b
      launch_recip(st, n2, d.dscale.p + n1, d.dcs.p);
      launch_scale_rc(st, n1, n2, d.dscale.p, d.dcs.p, d.X.p);
      // refinement against the FP64 matrix: Z = K [B; I] has the residual
      // -(K_ii B + K_ib) in its internal rows and K_bb + K_bi B in its
      // boundary rows (the Schur complement once B has converged)
      // conjugate gradients on K_ii B = -K_ib from this B: the internal
      // rows of K [B; I] are -(the residual), its boundary rows K_bb + K_bi B
      // (the Schur complement once B has converged)
      launch_spmm(st, neq, n2, d.rowptr.p, d.colind.p, d.Kv.p, d.X.p, n2, d.Z.p, n2);
      launch_axpby(st, (int64_t)n1 * n2, -1.0, d.Z.p, 0.0, d.Z.p);
      int steps = 0;
      if (int rc = pcg_int(st, d, y, opt_, *pb, d.X.p, d.Z.p, d.Y.p, &steps, "B", err)) return rc;
      launch_spmm(st, neq, n2, d.rowptr.p, d.colind.p, d.Kv.p, d.X.p, n2, d.Z.p, n2);
      times_.refine_steps = steps;
      std::vector<double> zb((size_t)n2 * n2);
      HIPCHK(hipMemcpyAsync(zb.data(), d.Z.p + (size_t)n1 * n2, sizeof(double) * zb.size(), hipMemcpyDeviceToHost,
                            st));
      HIPCHK(hipStreamSynchronize(st));
      d.s22.assign((size_t)n2 * n2, 0.0);
      for (int b = 0; b < n2; b++)
        for (int a = 0; a < n2; a++) d.s22[a + (size_t)b * n2] = 0.5 * (zb[(size_t)a * n2 + b] + zb[(size_t)b * n2 + a]);
    }
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(d.ev[4], st));
  mark("B issued");
  // error checks of the numeric phase
  int hbad[6] = {0, 0, 0, 0, 0, 0};
  HIPCHK(hipMemcpyAsync(hbad, d.piv.p, sizeof hbad, hipMemcpyDeviceToHost, st));
  std::vector<int> eb(part_.nel);
  HIPCHK(hipMemcpyAsync(eb.data(), d.ebad.p, sizeof(int) * part_.nel, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  for (int e = 0; e < part_.nel; e++)
    if (eb[e]) {
      err = "element " + std::to_string(e + 1) +
            (eb[e] == -3 ? " has a degenerate shell geometry (zero area or length)"
                         : " has a non-positive Jacobian determinant");
      return -4;
    }
  res_.n_negative = hbad[0];
  res_.n_true_zero = hbad[1];
  res_.n_reduced_zero = hbad[2];
  if (hbad[1] + hbad[2] > 0) {
    std::vector<int> sg(n1);
    HIPCHK(hipMemcpyAsync(sg.data(), d.sing.p, sizeof(int) * n1, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (int c = 0; c < n1; c++)
      if (sg[c]) res_.singular.push_back({y.perm_to_eq[c], sg[c]});
    std::sort(res_.singular.begin(), res_.singular.end());
    const int h = opt_.sing_handler;
    const bool abort = h == 0 || (h == 1 && hbad[2] > 0);
    if (abort) {
      const int c = h == 0 ? std::min(hbad[4], hbad[5]) : hbad[5];
      err = "singular K_ii: " + std::to_string(hbad[1]) + " zero and " + std::to_string(hbad[2]) +
            " reduced-to-zero pivot(s) (tolFactorize " + std::to_string(opt_.tol_factorize) + ", singularityHandler " +
            std::to_string(h) + "); first at equation " + std::to_string(y.perm_to_eq[c] + 1);
      return -5;
    }
  }
  // ---- component modes (block Lanczos, shift-invert on the resident factor)
  std::vector<double> Sres;
  if (nev > 0) {
    int rc = 0;
    int bs = 0, mcap = 0;
    lanczos_sizes(n1, nev, opt_, bs, mcap);
    SolvePlan* pl = nullptr;
    if (get_plan(y, d, bs, pl, err)) return -3;
    if (ensure_work(d, std::max(bs, n2), err)) return -3;
    HIPCHK(d.Q.alloc((size_t)neq * mcap));
    HIPCHK(d.P.alloc((size_t)neq * mcap));
    HIPCHK(d.Wl.alloc((size_t)neq * bs));
    HIPCHK(d.Zl.alloc((size_t)neq * bs));
    HIPCHK(d.Yl.alloc((size_t)neq * bs));
    HIPCHK(d.Tmp.alloc((size_t)neq * bs));
    HIPCHK(d.Cc.alloc((size_t)mcap * std::max(bs, nev) + 64));
    HIPCHK(d.part.alloc((size_t)xty_chunks(neq) * mcap * std::max(bs, nev)));
    HIPCHK(d.small.alloc((size_t)mcap * mcap + 64));
    HIPCHK(d.Cacc.alloc((size_t)mcap * bs));
    HIPCHK(d.Rdev.alloc((size_t)2 * bs * bs));
    HIPCHK(d.lzbad.alloc(1));
    HIPCHK(d.hC.alloc((size_t)2 * mcap * bs));
    HIPCHK(d.hR.alloc((size_t)4 * bs * bs));
    HIPCHK(d.hbad.alloc(2));
    for (auto& e : d.lz_ev)
      if (!e) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    // start block: deterministic pseudo-random values on the internal rows
    {
      std::vector<double> q0((size_t)neq * bs, 0.0);
      uint64_t s = opt_.seed;
      for (int r = 0; r < n1; r++)
        for (int j = 0; j < bs; j++) {
          s += 0x9E3779B97F4A7C15ull;
          uint64_t z = s;
          z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
          z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
          z ^= z >> 31;
          q0[(size_t)r * bs + j] = (double)(z >> 11) * (1.0 / 9007199254740992.0) - 0.5;
        }
      HIPCHK(hipMemcpyAsync(d.Wl.p, q0.data(), q0.size() * sizeof(double), hipMemcpyHostToDevice, st));
    }
    std::vector<double> Cfull;
    std::vector<double> H((size_t)mcap * mcap, 0.0);
    // M-orthonormalise W into block j of Q (and M W into P); returns R with W = Q_j R
    // Block M-orthonormalisation of W (CholQR twice): G = W^T M W = R^T R on
    // the device, W <- W R^-1, M W <- (M W) R^-1, then Q_j = W and P_j = M W.
    // orth_issue queues the work and the copies of both R factors and the
    // breakdown flag to host slot `par`; orth_finish (after the stream has
    // passed them) forms R_tot = R_2 R_1.
    auto orth_issue = [&](int j, int par) -> int {
      HIPCHK(hipMemsetAsync(d.lzbad.p, 0, sizeof(int), st));
      // M W is formed once: the second pass's M (W R_1^-1) is (M W) R_1^-1
      launch_spmm(st, neq, bs, d.rowptr.p, d.colind.p, d.Mv.p, d.Wl.p, bs, d.Zl.p, bs, d.m_lumped);
      // the first pass rescales W and M W (in place for blocks of up to 16
      // columns), the second writes block j of Q and P
      const bool inplace = bs <= 16;
      for (int pass = 0; pass < 2; pass++) {
        launch_xty(st, neq, bs, bs, d.Wl.p, bs, d.Zl.p, bs, d.part.p, d.Cc.p);
        launch_chol_inv(st, bs, d.Cc.p, d.Rdev.p + (size_t)pass * bs * bs, d.small.p, d.lzbad.p);
        double* qo = pass ? d.Q.p + (size_t)j * bs : (inplace ? d.Wl.p : d.Tmp.p);
        double* po = pass ? d.P.p + (size_t)j * bs : (inplace ? d.Zl.p : d.Yl.p);
        const int ldo = pass ? mcap : bs;
        launch_gemm_small2(st, neq, bs, bs, d.Wl.p, bs, qo, ldo, d.Zl.p, bs, po, ldo, d.small.p, 0.0, 1.0);
        if (pass == 0 && !inplace) {
          HIPCHK(hipMemcpyAsync(d.Wl.p, d.Tmp.p, sizeof(double) * neq * bs, hipMemcpyDeviceToDevice, st));
          HIPCHK(hipMemcpyAsync(d.Zl.p, d.Yl.p, sizeof(double) * neq * bs, hipMemcpyDeviceToDevice, st));
        }
      }
      HIPCHK(hipMemcpyAsync(d.hR.p + (size_t)par * 2 * bs * bs, d.Rdev.p, sizeof(double) * 2 * bs * bs,
                            hipMemcpyDeviceToHost, st));
      HIPCHK(hipMemcpyAsync(d.hbad.p + par, d.lzbad.p, sizeof(int), hipMemcpyDeviceToHost, st));
      return 0;
    };
    auto orth_finish = [&](int par, std::vector<double>& Rtot) -> int {
      if (d.hbad.p[par]) { err = "Lanczos: block breakdown (Gram matrix not positive definite)"; return -6; }
      const double* R1 = d.hR.p + (size_t)par * 2 * bs * bs;
      const double* R2 = R1 + (size_t)bs * bs;
      Rtot.assign((size_t)bs * bs, 0.0);
      for (int a = 0; a < bs; a++)
        for (int b = 0; b < bs; b++) {
          double s = 0;
          for (int k = 0; k < bs; k++) s += R2[(size_t)a * bs + k] * R1[(size_t)k * bs + b];
          Rtot[(size_t)a * bs + b] = s;
        }
      return 0;
    };
    std::vector<double> R0;
    if ((rc = orth_issue(0, 0))) return rc;
    HIPCHK(hipStreamSynchronize(st));
    if ((rc = orth_finish(0, R0))) return rc;
    int nb = 0;  // number of completed blocks in H
    std::vector<double> theta, Svec, Rlast;
    bool conv = false;
    // symmetrised block-tridiagonal projection of the first mj Lanczos vectors
    auto projected = [&](int mj) {
      std::vector<double> Hs((size_t)mj * mj);
      for (int a = 0; a < mj; a++)
        for (int b = 0; b < mj; b++) {
          double h1 = H[(size_t)a * mcap + b], h2 = H[(size_t)b * mcap + a];
          Hs[(size_t)a * mj + b] = (a <= b + bs && b <= a + bs) ? 0.5 * (h1 + h2) : 0.0;
        }
      return Hs;
    };
    // Ritz pairs of the projection and their relative residuals
    // ||R_{j+1} s_last|| / |theta| (theta = 1 / (lambda - shift))
    struct Check {
      int mj = 0;
      bool ok = false;
      double worst = 0.0;
      std::vector<double> w, V;
    };
    const double tol = opt_.tol_eig;
    // G = the M-Gram matrix of the next (unnormalised) Lanczos block: the
    // residual of Ritz vector s is sqrt(s_last^T G s_last)
    auto run_check = [bs, nev, tol](int mj, std::vector<double> Hs, std::vector<double> G) {
      Check c;
      c.mj = mj;
      sym_eig_top(mj, Hs, nev, c.w, c.V);  // V column i: i-th largest theta
      c.ok = true;
      for (int i = 0; i < nev; i++) {
        double th = c.w[mj - 1 - i], r2 = 0.0;
        const double* v = &c.V[(size_t)(mj - bs) * nev + i];
        for (int a = 0; a < bs; a++) {
          double s = 0;
          for (int b = 0; b < bs; b++) s += G[(size_t)a * bs + b] * v[(size_t)b * nev];
          r2 += v[(size_t)a * nev] * s;
        }
        double rel = std::sqrt(std::max(r2, 0.0)) / std::fabs(th);
        c.worst = std::max(c.worst, rel);
        if (!(rel <= tol)) c.ok = false;
      }
      return c;
    };
    auto gram_of = [&](const std::vector<double>& Rm) {  // R^T R
      std::vector<double> Gm((size_t)bs * bs, 0.0);
      for (int a = 0; a < bs; a++)
        for (int b = 0; b < bs; b++) {
          double s = 0.0;
          for (int k = 0; k < bs; k++) s += Rm[(size_t)k * bs + a] * Rm[(size_t)k * bs + b];
          Gm[(size_t)a * bs + b] = s;
        }
      return Gm;
    };
    std::future<Check> pending;
    int pending_step = 0;
    Check done;
    // Step j (device): W = K_ii^{-1} (M Q_j), two passes of full
    // reorthogonalisation against Q_0..Q_j (coefficients added up on the
    // device and copied to host slot j % 2), then the next block Q_{j+1}.
    // Steps are queued one ahead of the host's bookkeeping, so the device
    // never waits for the host between blocks.
    auto issue_step = [&](int j) -> int {
      const int mj = (j + 1) * bs, par = j & 1;
      HIPCHK(hipMemcpy2DAsync(d.Tmp.p, sizeof(double) * bs, d.P.p + (size_t)j * bs, sizeof(double) * mcap,
                              sizeof(double) * bs, neq, hipMemcpyDeviceToDevice, st));
      auto solve_step = [&]() {
        run_forward(st, d, *pl, d.Tmp.p, d.Yl.p, n1, true, false);
        launch_fill(st, d.Wl.p + (size_t)n1 * bs, (int64_t)(neq - n1) * bs, 0.0);  // boundary rows stay zero
        run_backward(st, d, *pl, d.Yl.p, d.Wl.p);
      };
      if (d.mixed) {
        if (int rc2 = solve_int(st, d, y, opt_, *pl, d.Tmp.p, d.Yl.p, d.Wl.p, nullptr, err)) return rc2;
      } else if (!opt_.use_graphs) solve_step();
      else HIPCHK(run_graph(st, d.g_lanczos,
                       {pl, pl->tasks.p, pl->part.p, d.Tmp.p, d.Yl.p, d.Wl.p, d.W.p, d.V.p, d.pool.p, d.ipool.p,
                        d.S.foff, d.S.woff, d.S.dpiv}, solve_step));
      for (int pass = 0; pass < 2; pass++) {
        // coefficients of the pass into Cc, their running sum into Cacc
        launch_xty(st, neq, mj, bs, d.P.p, mcap, d.Wl.p, bs, d.part.p, d.Cc.p, d.Cacc.p, pass > 0);
        launch_gemm_small(st, neq, mj, bs, d.Q.p, mcap, d.Cc.p, d.Wl.p, bs, 1.0, -1.0);
      }
      HIPCHK(hipMemcpyAsync(d.hC.p + (size_t)par * mcap * bs, d.Cacc.p, sizeof(double) * mj * bs,
                            hipMemcpyDeviceToHost, st));
      if (mj + bs <= mcap)
        if (int rc2 = orth_issue(j + 1, par)) return rc2;
      HIPCHK(hipEventRecord(d.lz_ev[par], st));
      return 0;
    };
    if ((rc = issue_step(0))) return rc;
    for (int j = 0; (j + 1) * bs <= mcap; j++) {
      const int mj = (j + 1) * bs, par = j & 1;
      const bool last = mj + bs > mcap;
      if (!last && (rc = issue_step(j + 1))) return rc;
      HIPCHK(hipEventSynchronize(d.lz_ev[par]));
      const double* hc = d.hC.p + (size_t)par * mcap * bs;
      for (int a = 0; a < mj; a++)
        for (int b = 0; b < bs; b++) H[(size_t)a * mcap + j * bs + b] = hc[(size_t)a * bs + b];
      nb = j + 1;
      if (last) break;
      std::vector<double> Rn;
      if ((rc = orth_finish(par, Rn))) return rc;
      for (int a = 0; a < bs; a++)
        for (int b = 0; b < bs; b++) H[(size_t)(mj + a) * mcap + j * bs + b] = Rn[(size_t)a * bs + b];
      // Convergence tests on the projected problem run on a host thread while
      // the device computes the next blocks.  A test started after block j is
      // collected after block j + 2 whatever the timing, so where the
      // iteration stops (and so every result bit) depends on the data only.
      if (pending.valid() && j >= pending_step + 2) {
        const auto tw = std::chrono::steady_clock::now();
        Check c = pending.get();
        if (std::getenv("FGPU_TRACE_LANCZOS"))
          std::fprintf(stderr, "lanczos step %d: check of %d vectors joined after %.3f ms wait, worst %.2e\n", j,
                       c.mj, std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tw).count(),
                       c.worst);
        if (c.ok) {
          done = std::move(c);
          conv = true;
          break;
        }
      }
      if (!pending.valid() && mj >= std::max(2 * nev, 3 * bs)) {
        pending = std::async(std::launch::async, run_check, mj, projected(mj), gram_of(Rn));
        pending_step = j;
      }
    }
    if (!conv && pending.valid()) {
      Check c = pending.get();
      if (c.ok) {
        done = std::move(c);
        conv = true;
      }
    }
    // without a converged test the Ritz pairs come from the full basis; the
    // M-Gram matrix of the last (reorthogonalised) block gives their residuals
    if (!conv) {
      launch_spmm(st, neq, bs, d.rowptr.p, d.colind.p, d.Mv.p, d.Wl.p, bs, d.Zl.p, bs, d.m_lumped);
      launch_xty(st, neq, bs, bs, d.Wl.p, bs, d.Zl.p, bs, d.part.p, d.Cc.p);
      std::vector<double> Gl((size_t)bs * bs);
      HIPCHK(hipMemcpyAsync(Gl.data(), d.Cc.p, sizeof(double) * bs * bs, hipMemcpyDeviceToHost, st));
      HIPCHK(hipEventRecord(d.ev[10], st));
      if (defer_b && (rc = issue_bsolve())) return rc;
      HIPCHK(hipEventSynchronize(d.ev[10]));
      const auto tw = std::chrono::steady_clock::now();
      done = run_check(nb * bs, projected(nb * bs), Gl);
      if (std::getenv("FGPU_TRACE_LANCZOS"))
        std::fprintf(stderr, "lanczos: final check of %d vectors on the host took %.3f ms, worst %.2e\n", done.mj,
                     std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tw).count(),
                     done.worst);
    } else if (defer_b && (rc = issue_bsolve())) {
      return rc;
    }
    res_.lanczos_resid = done.worst;
    if (!done.ok) {
      char buf[200];
      std::snprintf(buf, sizeof buf, "component modes did not converge: worst relative Ritz residual %.3e after %d "
                    "Lanczos vectors (tolerance %.1e; raise lanczosMaxVectors)", done.worst, done.mj, tol);
      err = buf;
      return -7;
    }
    const int m = done.mj;
    std::vector<double> w, V;
    w.swap(done.w);
    V.swap(done.V);
    (void)conv;
    res_.lanczos_vectors = m;
    // Ritz vectors for the nev largest theta (lowest lambda), lambda ascending
    std::vector<double> S((size_t)m * nev);
    res_.lambda.resize(nev);
    res_.eval.resize(nev);
    for (int i = 0; i < nev; i++) {
      double lam = 1.0 / w[m - 1 - i] + opt_.eig_shift;
      res_.lambda[i] = lam;
      res_.eval[i] = lam < 0 ? -std::sqrt(-lam) : std::sqrt(lam);
      for (int a = 0; a < m; a++) S[(size_t)a * nev + i] = V[(size_t)a * nev + i];
    }
    HIPCHK(hipMemcpyAsync(d.small.p, S.data(), sizeof(double) * S.size(), hipMemcpyHostToDevice, st));
    HIPCHK(d.Phi.alloc((size_t)neq * nev));
    launch_gemm_small(st, neq, m, nev, d.Q.p, mcap, d.small.p, d.Phi.p, nev, 0.0, 1.0);
    // sign convention: the largest-magnitude component of every mode is positive
    HIPCHK(d.signs.alloc((size_t)nev * (1 + 3 * MODE_SIGN_GROUPS)));
    launch_mode_signs(st, n1, d.p1.p, d.Phi.p, neq, nev, d.signs.p);
  }
  HIPCHK(hipEventRecord(d.ev[5], st));
  mark("modes done on host");
  // ---- reduced mass: Z = M [B; I], M22 = [B; I]^T Z ; modes: K11, M11, M12
  std::vector<double> M22, K11, M11, M12, Ge, Gg, S22((size_t)n2 * n2);
  const int maxab = std::max(std::max(n2, 3), std::max(nev, 1));
  HIPCHK(d.part.alloc((size_t)xty_chunks(neq) * maxab * maxab));
  HIPCHK(d.acc.alloc((size_t)maxab * maxab));
  auto xty_host = [&](int na, int nb, const double* X, int ldx, const double* Yv, int ldy,
                      std::vector<double>& out) -> int {
    out.assign((size_t)na * nb, 0.0);
    if (na == 0 || nb == 0) return 0;
    launch_xty(st, neq, na, nb, X, ldx, Yv, ldy, d.part.p, d.acc.p);
    HIPCHK(hipMemcpyAsync(out.data(), d.acc.p, sizeof(double) * na * nb, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return 0;
  };
  if (!opt_.no_mass && n2 > 0) {
    if (!defer_z) launch_spmm(st, neq, n2, d.rowptr.p, d.colind.p, d.Mv.p, d.X.p, n2, d.Z.p, n2, d.m_lumped);
    if (xty_host(n2, n2, d.X.p, n2, d.Z.p, n2, M22)) return -3;
  }
  if (ngen > 0) {
    HIPCHK(d.Yl.alloc((size_t)neq * std::max(ngen, opt_.lanczos_block)));
    launch_spmm(st, neq, ngen, d.rowptr.p, d.colind.p, d.Kv.p, d.Phi.p, nev, d.Yl.p, ngen);
    if (xty_host(ngen, ngen, d.Phi.p, nev, d.Yl.p, ngen, K11)) return -3;
    if (!opt_.no_mass) {
      launch_spmm(st, neq, ngen, d.rowptr.p, d.colind.p, d.Mv.p, d.Phi.p, nev, d.Yl.p, ngen, d.m_lumped);
      if (xty_host(ngen, ngen, d.Phi.p, nev, d.Yl.p, ngen, M11)) return -3;
      if (n2 > 0 && xty_host(ngen, n2, d.Phi.p, nev, d.Z.p, n2, M12)) return -3;
    }
  }
  HIPCHK(hipEventRecord(d.ev[6], st));
  // ---- gravity: g = M a ; G_e = [B; I]^T g ; G_g = Phi^T g
  if (!opt_.no_mass) {
    HIPCHK(d.Tmp.alloc((size_t)neq * std::max(3, std::max(opt_.lanczos_block, nev))));
    launch_spmm(st, neq, 3, d.rowptr.p, d.colind.p, d.Mv.p, d.grav.p, 3, d.Tmp.p, 3, d.m_lumped);
    if (n2 > 0 && xty_host(n2, 3, d.X.p, n2, d.Tmp.p, 3, Ge)) return -3;
    if (ngen > 0 && xty_host(ngen, 3, d.Phi.p, nev, d.Tmp.p, 3, Gg)) return -3;
    if (opt_.want_vgi && n1 > 0) {
      SolvePlan* p3 = nullptr;
      if (get_plan(y, d, 3, p3, err)) return -3;
      if (ensure_work(d, std::max(std::max(n2, 3), opt_.lanczos_block), err)) return -3;
      HIPCHK(d.Zl.alloc((size_t)neq * std::max(3, opt_.lanczos_block)));
      HIPCHK(d.Wl.alloc((size_t)neq * std::max(3, opt_.lanczos_block)));
      if (int rc2 = solve_int(st, d, y, opt_, *p3, d.Tmp.p, d.Zl.p, d.Wl.p, nullptr, err)) return rc2;
      std::vector<double> v((size_t)neq * 3);
      HIPCHK(hipMemcpyAsync(v.data(), d.Wl.p, sizeof(double) * v.size(), hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      res_.vgi.assign((size_t)n1 * 3, 0.0);
      for (int q = 0; q < n1; q++) {  // MEQN1 order: equation q+1
        int r = y.eq_to_perm[sam_.meqn1[q] - 1];
        for (int c = 0; c < 3; c++) res_.vgi[(size_t)c * n1 + q] = v[(size_t)r * 3 + c];
      }
    }
  }
  HIPCHK(hipEventRecord(d.ev[7], st));
  mark("mass and gravity issued");
  // ---- Schur complement out of the root front
  if (n2 > 0 && d.mixed) S22 = d.s22;
  else if (n2 > 0) {
    HIPCHK(hipMemcpyAsync(S22.data(), d.pool.p + y.foff[schur], sizeof(double) * n2 * n2, hipMemcpyDeviceToHost, st));
  }
  if (!opt_.no_mass)
    HIPCHK(hipMemcpyAsync(&res_.mass_fe, d.emass.p + part_.nel, sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  mark("device done");
  times_.elements = ev_ms(d.ev[0], d.ev[1]);
  times_.assembly = ev_ms(d.ev[1], d.ev[2]);
  times_.factor = ev_ms(d.ev[2], d.ev[3]);
  times_.bsolve = defer_b ? ev_ms(d.ev[8], d.ev[9]) : ev_ms(d.ev[3], d.ev[4]);
  const double zms = defer_z ? ev_ms(d.ev[9], d.ev[11]) : 0.0;
  times_.modes = ev_ms(d.ev[4], d.ev[5]) - (defer_b ? times_.bsolve : 0.0) - zms;
  times_.mass = ev_ms(d.ev[5], d.ev[6]) + zms;
  times_.gravity = ev_ms(d.ev[6], d.ev[7]);
  times_.total = ev_ms(d.ev[0], d.ev[7]);
  times_.update_kernel_ms = 0;
  times_.update_launches = 0;
  if (opt_.time_updates) {
    for (int64_t q = 0; q < ui; q++) times_.update_kernel_ms += ev_ms(d.upd_ev[2 * q], d.upd_ev[2 * q + 1]);
    times_.update_launches = ui;
  }
  // ---- reduced matrices in the reference layout (CMSTRS, GRAV): first ndof2
  // rows in DOF order through dofPosIn2, then the ngen component modes, then
  // zero padding up to ndim
  const int nd = res_.ndim;
  const std::vector<int>& pos2 = sam_.dofPosIn2;
  res_.SK.assign((size_t)nd * nd, 0.0);
  res_.SM.assign(opt_.no_mass ? 0 : (size_t)nd * nd, 0.0);
  res_.G.assign(opt_.no_mass ? 0 : (size_t)nd * 3, 0.0);
  for (int j = 0; j < n2; j++)
    for (int i = 0; i < n2; i++) {
      int a = pos2[i] - 1, b = pos2[j] - 1;
      double v = a >= b ? S22[a + (size_t)b * n2] : S22[b + (size_t)a * n2];
      res_.SK[i + (size_t)j * nd] = v;
      if (!opt_.no_mass) res_.SM[i + (size_t)j * nd] = 0.5 * (M22[(size_t)a * n2 + b] + M22[(size_t)b * n2 + a]);
    }
  for (int a = 0; a < ngen; a++)
    for (int b = 0; b < ngen; b++) {
      res_.SK[(n2 + a) + (size_t)(n2 + b) * nd] = 0.5 * (K11[(size_t)a * ngen + b] + K11[(size_t)b * ngen + a]);
      if (!opt_.no_mass) res_.SM[(n2 + a) + (size_t)(n2 + b) * nd] = 0.5 * (M11[(size_t)a * ngen + b] + M11[(size_t)b * ngen + a]);
    }
  if (!opt_.no_mass) {
    for (int a = 0; a < ngen; a++)
      for (int i = 0; i < n2; i++) {
        double v = M12[(size_t)a * n2 + (pos2[i] - 1)];
        res_.SM[(n2 + a) + (size_t)i * nd] = v;
        res_.SM[i + (size_t)(n2 + a) * nd] = v;
      }
    for (int c = 0; c < 3; c++) {
      for (int i = 0; i < n2; i++) res_.G[i + (size_t)c * nd] = Ge[(size_t)(pos2[i] - 1) * 3 + c];
      for (int a = 0; a < ngen; a++) res_.G[(n2 + a) + (size_t)c * nd] = Gg[(size_t)a * 3 + c];
    }
  }
  return 0;
}

hipError_t DeviceReducer::d2h(void* dst, const void* src, size_t bytes) const {
  if (!st_) return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost);
  hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st_);
  return e == hipSuccess ? hipStreamSynchronize(st_) : e;
}

int DeviceReducer::get_B(std::vector<double>& B) {
  const Symbolic& y = sym_;
  const int n1 = y.ndof1, n2 = y.ndof2;
  std::vector<double> x((size_t)y.neq * n2);
  if (n2 == 0) { B.clear(); return 0; }
  if (d2h(x.data(), d_->X.p, sizeof(double) * x.size()) != hipSuccess) return -3;
  B.assign((size_t)n1 * n2, 0.0);
  for (int q = 0; q < n1; q++) {
    int r = y.eq_to_perm[sam_.meqn1[q] - 1];
    for (int c = 0; c < n2; c++) B[q + (size_t)c * n1] = x[(size_t)r * n2 + c];
  }
  return 0;
}

int DeviceReducer::get_phi(std::vector<double>& P) {
  const Symbolic& y = sym_;
  const int n1 = y.ndof1, nev = res_.nev;
  P.assign((size_t)n1 * nev, 0.0);
  if (nev == 0) return 0;
  std::vector<double> x((size_t)y.neq * nev);
  if (d2h(x.data(), d_->Phi.p, sizeof(double) * x.size()) != hipSuccess) return -3;
  for (int q = 0; q < n1; q++) {
    int r = y.eq_to_perm[sam_.meqn1[q] - 1];
    for (int c = 0; c < nev; c++) P[q + (size_t)c * n1] = x[(size_t)r * nev + c];
  }
  return 0;
}

int DeviceReducer::get_schur(std::vector<double>& S) {
  const Symbolic& y = sym_;
  const int n2 = y.ndof2;
  std::vector<double> f((size_t)n2 * n2);
  S.assign((size_t)n2 * n2, 0.0);
  if (n2 == 0) return 0;
  if (d_->mixed) {  // from the refined B (symmetric)
    S = d_->s22;
    return 0;
  }
  if (d2h(f.data(), d_->pool.p + y.foff[y.nsup - 1], sizeof(double) * f.size()) != hipSuccess)
    return -3;
  for (int j = 0; j < n2; j++)
    for (int i = 0; i < n2; i++) S[i + (size_t)j * n2] = i >= j ? f[i + (size_t)j * n2] : f[j + (size_t)i * n2];
  return 0;
}

int DeviceReducer::get_csr(std::vector<int>& rowptr, std::vector<int>& colind, std::vector<double>& K,
                           std::vector<double>& M) {
  const Symbolic& y = sym_;
  rowptr = y.rowptr;
  size_t nnz = y.rowptr[y.neq];
  colind.resize(nnz);
  K.resize(nnz);
  M.resize(nnz);
  if (d2h(colind.data(), d_->colind.p, sizeof(int) * nnz) != hipSuccess) return -3;
  if (d2h(K.data(), d_->Kv.p, sizeof(double) * nnz) != hipSuccess) return -3;
  if (d2h(M.data(), d_->Mv.p, sizeof(double) * nnz) != hipSuccess) return -3;
  return 0;
}

int DeviceReducer::get_element_matrices(std::vector<double>& Kp, std::vector<double>& Mp) {
  Kp.resize(sym_.elem_pool);
  Mp.resize(sym_.elem_pool);
  if (d2h(Kp.data(), d_->Kp.p, sizeof(double) * Kp.size()) != hipSuccess) return -3;
  if (d2h(Mp.data(), d_->Mp.p, sizeof(double) * Mp.size()) != hipSuccess) return -3;
  return 0;
}

int DeviceReducer::solve_kii(int nrhs, const double* R_eq, double* X_eq, std::string& err) {
  StreamScope scope(st_);
  Impl& d = *d_;
  const Symbolic& y = sym_;
  const int neq = y.neq, n1 = y.ndof1;
  SolvePlan* p = nullptr;
  if (get_plan(y, d, nrhs, p, err)) return -3;
  if (ensure_work(d, std::max(nrhs, std::max(y.ndof2, opt_.lanczos_block)), err)) return -3;
  DBuf<double>& r = d.sk_r;
  DBuf<double>& t = d.sk_t;
  DBuf<double>& x = d.sk_x;
  std::vector<double> h((size_t)neq * nrhs, 0.0);
  for (int q = 0; q < n1; q++) {
    int rr = y.eq_to_perm[sam_.meqn1[q] - 1];
    for (int c = 0; c < nrhs; c++) h[(size_t)rr * nrhs + c] = R_eq[q + (size_t)c * n1];
  }
  HIPCHK(r.upload(h));
  HIPCHK(t.alloc(h.size()));
  HIPCHK(x.alloc(h.size()));
  if (int rc = solve_int(st_, d, y, opt_, *p, r.p, t.p, x.p, nullptr, err)) return rc;
  HIPCHK(hipMemcpyAsync(h.data(), x.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, st_));
  HIPCHK(hipStreamSynchronize(st_));
  for (int q = 0; q < n1; q++) {
    int rr = y.eq_to_perm[sam_.meqn1[q] - 1];
    for (int c = 0; c < nrhs; c++) X_eq[q + (size_t)c * n1] = h[(size_t)rr * nrhs + c];
  }
  return 0;
}

int DeviceReducer::gravity_internal(std::vector<double>& gi, std::string& err) {
  StreamScope scope(st_);
  Impl& d = *d_;
  const Symbolic& y = sym_;
  const int neq = y.neq, n1 = y.ndof1;
  if (opt_.no_mass) { err = "gravity forces need the mass matrix"; return -1; }
  DBuf<double>& g = d.gi_g;
  HIPCHK(g.alloc((size_t)neq * 3));
  launch_spmm(st_, neq, 3, d.rowptr.p, d.colind.p, d.Mv.p, d.grav.p, 3, g.p, 3, d.m_lumped);
  std::vector<double> h((size_t)neq * 3);
  HIPCHK(hipMemcpyAsync(h.data(), g.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, st_));
  HIPCHK(hipStreamSynchronize(st_));
  gi.assign((size_t)n1 * 3, 0.0);
  for (int q = 0; q < n1; q++) {
    int rr = y.eq_to_perm[sam_.meqn1[q] - 1];
    for (int c = 0; c < 3; c++) gi[q + (size_t)c * n1] = h[(size_t)rr * 3 + c];
  }
  return 0;
}

int DeviceReducer::reduce_loads(int nrhs, const double* F_eq, std::vector<double>& out, std::string& err) {
  StreamScope scope(st_);
  Impl& d = *d_;
  const Symbolic& y = sym_;
  const int neq = y.neq, n2 = y.ndof2, ngen = res_.ngen, nev = res_.nev, nd = res_.ndim;
  out.assign((size_t)nd * std::max(nrhs, 0), 0.0);
  if (nrhs <= 0) return 0;
  std::vector<double> h((size_t)neq * nrhs);
  for (int e = 0; e < neq; e++) {
    int rr = y.eq_to_perm[e];
    for (int c = 0; c < nrhs; c++) h[(size_t)rr * nrhs + c] = F_eq[e + (size_t)c * neq];
  }
  DBuf<double>& f = d.rl_f;
  HIPCHK(f.upload(h));
  const int maxab = std::max(std::max(n2, ngen), nrhs);
  HIPCHK(d.part.alloc((size_t)xty_chunks(neq) * maxab * maxab));
  HIPCHK(d.acc.alloc((size_t)maxab * maxab));
  std::vector<double> r2((size_t)n2 * nrhs), r1((size_t)ngen * nrhs);
  if (n2 > 0) {  // rhs2 = f_e + B^T f_i = [B; I]^T f
    launch_xty(st_, neq, n2, nrhs, d.X.p, n2, f.p, nrhs, d.part.p, d.acc.p);
    HIPCHK(hipMemcpyAsync(r2.data(), d.acc.p, sizeof(double) * r2.size(), hipMemcpyDeviceToHost, st_));
    HIPCHK(hipStreamSynchronize(st_));
  }
  if (ngen > 0) {  // rhs1 = Phi^T f_i
    launch_xty(st_, neq, ngen, nrhs, d.Phi.p, nev, f.p, nrhs, d.part.p, d.acc.p);
    HIPCHK(hipMemcpyAsync(r1.data(), d.acc.p, sizeof(double) * r1.size(), hipMemcpyDeviceToHost, st_));
    HIPCHK(hipStreamSynchronize(st_));
  }
  const std::vector<int>& pos2 = sam_.dofPosIn2;
  for (int c = 0; c < nrhs; c++) {
    for (int i = 0; i < n2; i++) out[i + (size_t)c * nd] = r2[(size_t)(pos2[i] - 1) * nrhs + c];
    for (int a = 0; a < ngen; a++) out[(n2 + a) + (size_t)c * nd] = r1[(size_t)a * nrhs + c];
  }
  (void)err;
  return 0;
}

// reduce_loads on device buffers, with every work buffer in the library's
// static scratch array: no device allocation.  F (neq x nrhs) and Fred
// (ndim x nrhs) are column-major device arrays.
int DeviceReducer::reduce_loads_noalloc(int nrhs, const double* F, double* Fred, std::string& err) {
  StreamScope scope(st_);
  Impl& d = *d_;
  const Symbolic& y = sym_;
  const int neq = y.neq, n2 = y.ndof2, ngen = res_.ngen, nev = res_.nev, nd = res_.ndim;
  double* scr = nullptr;
  size_t cap = 0;
  if (int rc = noalloc_scratch(&scr, &cap, err)) return rc;
  const size_t nF = (size_t)neq * nrhs, maxa = (size_t)std::max(n2, ngen);
  const size_t npart = (size_t)xty_chunks(neq) * maxa * nrhs;
  const size_t need = 8 * (nF + npart + (size_t)(n2 + ngen) * nrhs);
  if (need > cap) {
    err = "reduced loads without allocation: " + std::to_string(nrhs) + " load cases need " + std::to_string(need) +
          " bytes of scratch, more than the " + std::to_string(cap) + " available";
    return -1;
  }
  double* Fp = scr;
  double* part = Fp + nF;
  double* r2 = part + npart;
  double* r1 = r2 + (size_t)n2 * nrhs;
  launch_perm_rows(st_, neq, nrhs, d.p2e.p, F, Fp);
  if (n2 > 0) launch_xty(st_, neq, n2, nrhs, d.X.p, n2, Fp, nrhs, part, r2);
  if (ngen > 0) launch_xty(st_, neq, ngen, nrhs, d.Phi.p, nev, Fp, nrhs, part, r1);
  launch_scatter_reduced(st_, nd, n2, ngen, nrhs, d.pos2.p, r2, r1, Fred);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(st_));
  return 0;
}

int DeviceReducer::recover(int nrhs, const double* Ub, const double* Fi, double* Ui, std::string& err) {
  StreamScope scope(st_);
  Impl& d = *d_;
  const Symbolic& y = sym_;
  const int neq = y.neq, n1 = y.ndof1, n2 = y.ndof2;
  if (nrhs <= 0) return 0;
  std::vector<double> acc((size_t)n1 * nrhs, 0.0);
  if (n2 > 0 && Ub) {
    // Y = [B; I] U_b in permuted rows, the boundary block of U_b in LDS-sized
    // column chunks
    const int cj = std::max(1, std::min(nrhs, 8192 / std::max(n2, 1)));
    DBuf<double>& S = d.rc_s;
    DBuf<double>& Yd = d.rc_y;
    HIPCHK(Yd.alloc((size_t)neq * cj));
    std::vector<double> hs((size_t)n2 * cj), hy((size_t)neq * cj);
    for (int j0 = 0; j0 < nrhs; j0 += cj) {
      const int nj = std::min(cj, nrhs - j0);
      for (int a = 0; a < n2; a++)
        for (int j = 0; j < nj; j++) hs[(size_t)a * nj + j] = Ub[a + (size_t)(j0 + j) * n2];
      std::vector<double> hsn(hs.begin(), hs.begin() + (size_t)n2 * nj);
      HIPCHK(S.upload(hsn));
      launch_gemm_small(st_, neq, n2, nj, d.X.p, n2, S.p, Yd.p, nj, 0.0, 1.0);
      HIPCHK(hipMemcpyAsync(hy.data(), Yd.p, sizeof(double) * (size_t)neq * nj, hipMemcpyDeviceToHost, st_));
      HIPCHK(hipStreamSynchronize(st_));
      for (int q = 0; q < n1; q++) {
        const int r = y.eq_to_perm[sam_.meqn1[q] - 1];
        for (int j = 0; j < nj; j++) acc[q + (size_t)(j0 + j) * n1] = hy[(size_t)r * nj + j];
      }
    }
  }
  if (Fi) {
    std::vector<double> x((size_t)n1 * nrhs);
    int rc = solve_kii(nrhs, Fi, x.data(), err);
    if (rc) return rc;
    for (size_t k = 0; k < x.size(); k++) acc[k] += x[k];
  }
  std::copy(acc.begin(), acc.end(), Ui);
  return 0;
}

int DeviceReducer::autotune(int& panel, int& split, std::string& err) {
  Impl& d = *d_;
  if (!uploaded_ && upload(err)) return -3;
  auto set_panel = [&](int P) -> int {
    std::vector<FTask> ft;
    std::vector<int> xr;
    d.fops.clear();
    int inv_slots = 0;
    build_factor_schedule(sym_, ft, d.fops, xr, upd_flops_, P, inv_slots, opt_.factor_lookahead);
    HIPCHK(ensure_side(d));
    HIPCHK(d.ipart.alloc((size_t)std::max(1, inv_slots) * 4096));
    HIPCHK(d.ftasks.upload(ft));
    HIPCHK(d.xrange.upload(xr));
    upd_launches_ = 0;
    for (auto& op : d.fops)
      if (op.kind == UPDATE || op.kind == UPDATE_SIDE) upd_launches_++;
    d.g_factor.reset();
    opt_.factor_panel = P;
    return 0;
  };
  auto set_split = [&](int KS) {
    d.solve_split = KS;
    opt_.solve_split = KS;
    d.plans.clear();
    d.g_lanczos.reset();
  };
  // warm-up (allocations, first captures), then one timed reduction per candidate
  if (int rc = reduce(err)) return rc;
  double best = INFINITY;
  int bestP = opt_.factor_panel;
  for (int P : {64, 128, 192, 256}) {
    if (set_panel(P)) return -3;
    if (int rc = reduce(err)) return rc;
    if (int rc = reduce(err)) return rc;  // the second run replays the captured factor graph
    if (times_.factor < best) { best = times_.factor; bestP = P; }
  }
  if (set_panel(bestP)) return -3;
  best = INFINITY;
  int bestK = d.solve_split;
  for (int KS : {64, 128, 256, 512}) {
    set_split(KS);
    if (int rc = reduce(err)) return rc;
    if (int rc = reduce(err)) return rc;
    const double t = times_.bsolve + times_.modes;
    if (t < best) { best = t; bestK = KS; }
  }
  set_split(bestK);
  panel = bestP;
  split = bestK;
  return 0;
}

size_t DeviceReducer::footprint_bytes() const {
  const Symbolic& y = sym_;
  const size_t nnz = y.rowptr.empty() ? 0 : (size_t)y.rowptr.back();
  size_t ipool = 0, wrows = 0;
  for (int s = 0; s < y.nsup; s++) {
    ipool += (size_t)y.ncol[s] * y.ncol[s];
    wrows += y.nrow[s];
  }
  const int nev = opt_.ngen > 0 ? std::max(opt_.neval, opt_.ngen) : 0;
  const int bs = opt_.lanczos_block;
  const size_t mcap = nev > 0 ? (size_t)std::max(10 * nev + 40, 16 * bs) : 0;
  const size_t rhs = (size_t)std::max(std::max(y.ndof2, bs), 3);
  const size_t fb = opt_.mixed ? 4 : 8;  // bytes per factor entry
  size_t b = fb * ((size_t)y.fpool + ipool) + 8 * (2 * nnz + 2 * (size_t)y.elem_pool);  // fronts, inverses, K, M, elements
  b += (y.has_rigid ? 12 : 4) * (size_t)y.ncontrib + 4 * (nnz + 1) + 8 * nnz + 4 * nnz;  // contributions, fmap, colind
  b += 8 * nnz;                                                                     // structure-build counts
  b += 20 * sam_.ceq_dof.size() + 4 * sam_.ceq_ptr.size();                          // constraint equations
  b += 8 * (size_t)y.neq * (3 * (size_t)y.ndof2 + 2 * mcap + 8 * (size_t)bs);     // B solves and Lanczos basis
  b += 8 * 2 * wrows * rhs;                                                        // solve work rows
  if (opt_.mixed) b += 8 * 5 * (size_t)y.neq * rhs;                                // refinement work
  return b;
}

}  // namespace fgpu
