// Continuous-batching reduction server: parts arrive as requests, a fixed
// number of slots (host worker threads, each owning a DeviceReducer with its
// own HIP stream) take them in arrival order, and a request starts on the
// device only when its estimated device footprint fits next to the requests
// already running.  Work of different slots overlaps on the GPU.  Each
// request is a full reduction followed by its load cases on the resident
// factor: the reduced load vectors and the internal displacements
// U_i = B U_b + K_ii^-1 F_i (displacement recovery from boundary
// displacements).
//
// Every kernel of a reduction is deterministic and a reduction touches only
// its own slot's buffers and stream, so a request's results are the same bit
// for bit whether it runs alone or next to others.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <thread>
#include "../../include/fedem_reducer_gpu.h"
#include "reducer_gpu.hpp"

namespace fgpu {
namespace {

using Clock = std::chrono::steady_clock;

struct Request {
  Part part;
  ReduceOptions opt;
  int nlc = 0;
  std::vector<int> lptr, lnode, ldof;
  std::vector<double> lval, ub;
  // results (written by the worker before `done` is set)
  int status = 0;
  std::atomic<bool> done{false};
  std::string err;
  Results res;
  std::vector<double> fred, ui;
  size_t footprint = 0;
  Clock::time_point t_submit, t_start, t_end;
};

double ms(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

}  // namespace

// Device memory accounting of the admission control.  Every slot's reducer
// keeps its buffers between requests (they only grow), so a slot holds
// `bytes` of device memory whether it is running or idle; a running slot
// holds at least its request's estimated footprint.  A request is admitted
// when the memory of all slots, with its own slot at max(held, footprint),
// fits the budget; otherwise idle slots release their buffers, the largest
// holders first, until it fits; if it still does not, it waits for running
// requests, and when nothing else runs it is admitted anyway (a request
// larger than the whole budget runs alone).  Slots keep their buffers
// between requests (releasing device memory waits for the whole device),
// except that a slot whose request leaves all slots together above the
// budget releases its own.
class Server {
 public:
  Server(int slots, size_t budget) : budget_(budget), state_(slots, IDLE), fp_(slots, 0), bytes_(slots) {
    for (auto& b : bytes_) b = 0;
    for (int k = 0; k < slots; k++) reducers_.emplace_back(new DeviceReducer());
    for (int k = 0; k < slots; k++) threads_.emplace_back(&Server::worker, this, k);
  }
  ~Server() {
    {
      std::lock_guard<std::mutex> g(mu_);
      stop_ = true;
    }
    cv_queue_.notify_all();
    cv_mem_.notify_all();
    for (auto& t : threads_) t.join();
  }
  int submit(std::unique_ptr<Request> r) {
    std::lock_guard<std::mutex> g(mu_);
    const int id = next_id_++;
    r->t_submit = Clock::now();
    reqs_[id] = std::move(r);
    queue_.push_back(id);
    cv_queue_.notify_one();
    return id;
  }
  Request* wait(int id) {
    std::unique_lock<std::mutex> g(mu_);
    auto it = reqs_.find(id);
    if (it == reqs_.end()) return nullptr;
    Request* r = it->second.get();
    cv_done_.wait(g, [&] { return r->done.load(); });
    return r;
  }
  // a finished request, or nullptr (unknown id or still running)
  Request* find_done(int id) {
    std::lock_guard<std::mutex> g(mu_);
    auto it = reqs_.find(id);
    return it == reqs_.end() || !it->second->done.load() ? nullptr : it->second.get();
  }
  void release(int id) {
    std::lock_guard<std::mutex> g(mu_);
    auto it = reqs_.find(id);
    if (it != reqs_.end() && it->second->done) reqs_.erase(it);
  }
  // device bytes held by all slots' buffers (for tests of the budget)
  int64_t held_bytes() {
    std::lock_guard<std::mutex> g(mu_);
    int64_t b = 0;
    for (auto& x : bytes_) b += x.load();
    return b;
  }

 private:
  enum SlotState { IDLE, WAITING, RUNNING };

  // memory the slots other than `me` hold (under mu_)
  size_t others_held(int me) const {
    size_t b = 0;
    for (size_t k = 0; k < bytes_.size(); k++)
      if ((int)k != me) b += std::max<size_t>((size_t)std::max<int64_t>(0, bytes_[k].load()), state_[k] == RUNNING ? fp_[k] : 0);
    return b;
  }
  bool others_running(int me) const {
    for (size_t k = 0; k < state_.size(); k++)
      if ((int)k != me && state_[k] == RUNNING) return true;
    return false;
  }
  // drops the buffers of slot k (its reducer is not in use: idle, or the
  // calling slot itself)
  void trim(int k) {
    set_alloc_account(&bytes_[k]);
    reducers_[k].reset(new DeviceReducer());
    set_alloc_account(&bytes_[me_of_thread()]);
  }
  static int& me_of_thread() {
    static thread_local int me = -1;
    return me;
  }

  void worker(int slot) {
    me_of_thread() = slot;
    set_alloc_account(&bytes_[slot]);
    while (true) {
      Request* r = nullptr;
      {
        std::unique_lock<std::mutex> g(mu_);
        cv_queue_.wait(g, [&] { return stop_ || !queue_.empty(); });
        if (stop_ && queue_.empty()) return;
        r = reqs_[queue_.front()].get();
        queue_.pop_front();
        state_[slot] = WAITING;
      }
      std::string err;
      int rc;
      try {
        rc = run(slot, *r, err);
        if (rc == -6 && r->opt.mixed) {
          // the mixed-precision refinement did not converge (an FP32 factor
          // too inexact for this part): the request is redone in FP64
          r->opt.mixed = false;
          err.clear();
          rc = run(slot, *r, err);
        }
      } catch (const std::exception& e) {
        err = std::string("request failed: ") + e.what();
        rc = -99;
      } catch (...) {
        err = "request failed with an unknown exception";
        rc = -99;
      }
      bool shrink;
      {
        std::lock_guard<std::mutex> g(mu_);
        int64_t all = 0;
        for (auto& b : bytes_) all += b.load();
        shrink = all > (int64_t)budget_;
      }
      if (shrink) trim(slot);  // outside the lock: the slot is still WAITING/RUNNING for the others
      std::lock_guard<std::mutex> g(mu_);
      state_[slot] = IDLE;
      fp_[slot] = 0;
      r->status = rc;
      r->err = err;
      r->t_end = Clock::now();
      r->done = true;
      cv_done_.notify_all();
      cv_mem_.notify_all();
    }
  }

  int run(int slot, Request& r, std::string& err) {
    int rc = reducers_[slot]->analyze(r.part, r.opt, err);  // host only
    const size_t fp = rc == 0 ? reducers_[slot]->footprint_bytes() : 0;
    if (rc == 0) {
      std::unique_lock<std::mutex> g(mu_);
      cv_mem_.wait(g, [&] {
        if (stop_) return true;
        const size_t mine = std::max<size_t>((size_t)std::max<int64_t>(0, bytes_[slot].load()), fp);
        if (others_held(slot) + mine <= budget_) return true;
        // idle slots give their buffers back, the largest holders first
        std::vector<int> idle;
        for (size_t k = 0; k < state_.size(); k++)
          if ((int)k != slot && state_[k] == IDLE && bytes_[k].load() > 0) idle.push_back((int)k);
        std::sort(idle.begin(), idle.end(), [&](int a, int b) { return bytes_[a].load() > bytes_[b].load(); });
        for (int k : idle) {
          trim(k);
          if (others_held(slot) + mine <= budget_) return true;
        }
        // still too large: wait for the running requests, or run alone
        return !others_running(slot);
      });
      state_[slot] = RUNNING;
      fp_[slot] = fp;
    }
    DeviceReducer& R = *reducers_[slot];
    r.t_start = Clock::now();
    r.footprint = fp;
    if (rc == 0) rc = R.reduce(err);
    if (rc == 0) {
      r.res = R.results();
      const Sam& s = R.sam();
      const int neq = s.neq, n1 = s.ndof1, n2 = s.ndof2;
      if (r.nlc > 0) {
        // full load vectors (equation order) from the nodal loads
        std::vector<double> F((size_t)neq * r.nlc, 0.0);
        for (int c = 0; c < r.nlc && rc == 0; c++)
          for (int k = r.lptr[c]; k < r.lptr[c + 1]; k++) {
            const int n = r.lnode[k] - 1, j = r.ldof[k] - 1;
            if (n < 0 || n >= s.nnod || j < 0 || j >= s.nndof[n]) {
              err = "load on a node or DOF outside the part";
              rc = -1;
              break;
            }
            const int eq = s.meqn[s.madof[n] - 1 + j];
            if (eq > 0) F[(eq - 1) + (size_t)c * neq] += r.lval[k];
          }
        if (rc == 0) rc = R.reduce_loads(r.nlc, F.data(), r.fred, err);
        if (rc == 0) {
          std::vector<double> Fi((size_t)n1 * r.nlc), Ub((size_t)n2 * r.nlc, 0.0);
          for (int c = 0; c < r.nlc; c++)
            for (int q = 0; q < n1; q++) Fi[q + (size_t)c * n1] = F[(s.meqn1[q] - 1) + (size_t)c * neq];
          if (!r.ub.empty()) {
            if (r.ub.size() != (size_t)n2 * r.nlc) {
              err = "boundary displacements: expected ndof2 values per load case";
              rc = -1;
            } else {
              // reduced (dofPosIn2) order -> MEQN2 order
              for (int c = 0; c < r.nlc; c++)
                for (int i = 0; i < n2; i++) Ub[(s.dofPosIn2[i] - 1) + (size_t)c * n2] = r.ub[i + (size_t)c * n2];
            }
          }
          r.ui.assign((size_t)n1 * r.nlc, 0.0);
          if (rc == 0) rc = R.recover(r.nlc, Ub.data(), Fi.data(), r.ui.data(), err);
        }
      }
    }
    return rc;
  }

  size_t budget_;
  bool stop_ = false;
  int next_id_ = 0;
  std::mutex mu_;
  std::condition_variable cv_queue_, cv_done_, cv_mem_;
  std::deque<int> queue_;
  std::map<int, std::unique_ptr<Request>> reqs_;
  std::vector<SlotState> state_;
  std::vector<size_t> fp_;
  std::vector<std::atomic<int64_t>> bytes_;
  std::vector<std::unique_ptr<DeviceReducer>> reducers_;
  std::vector<std::thread> threads_;
};

}  // namespace fgpu

struct fgpu_server {
  std::unique_ptr<fgpu::Server> s;
  fgpu::ReduceOptions base;  // options of the requests submitted from now on
};

extern "C" {

int fgpu_server_create(fgpu_server** out, int slots, double budget_gb) {
  if (!out || slots < 1 || !(budget_gb > 0)) return -1;
  try {
    *out = new fgpu_server{std::unique_ptr<fgpu::Server>(new fgpu::Server(slots, (size_t)(budget_gb * 1073741824.0)))};
  } catch (...) {
    return -1;
  }
  return 0;
}

void fgpu_server_destroy(fgpu_server* s) { delete s; }

int fgpu_server_submit(fgpu_server* s, int nnod, const double* xyz, const int* status, int nel, const int* melcon,
                       const int* mpmnpc, const int* mmnpc, const double* eprop, int ngen, int nlc, const int* lptr,
                       const int* lnode, const int* ldof, const double* lval, int nub, const double* ub) {
  if (!s || nnod <= 0 || nel < 0 || !xyz || !status || nlc < 0 || (nlc > 0 && !lptr)) return -1;
  std::unique_ptr<fgpu::Request> r(new fgpu::Request());
  fgpu::Part& p = r->part;
  p.nnod = nnod;
  p.nel = nel;
  p.xyz.assign(xyz, xyz + 3 * (size_t)nnod);
  p.status.assign(status, status + nnod);
  p.etype.assign(melcon, melcon + nel);
  p.eptr.resize(nel + 1);
  for (int e = 0; e <= nel; e++) p.eptr[e] = mpmnpc[e] - 1;
  p.enodes.resize(p.eptr[nel]);
  for (int k = 0; k < p.eptr[nel]; k++) {
    p.enodes[k] = mmnpc[k] - 1;
    if (p.enodes[k] < 0 || p.enodes[k] >= nnod) return -2;
  }
  for (int e = 0; e < nel; e++)
    if (p.eptr[e + 1] - p.eptr[e] != fgpu::elem_nodes(p.etype[e])) return -2;
  p.eprop.assign(eprop, eprop + (size_t)fgpu::NPROP * nel);
  r->opt = s->base;
  r->opt.ngen = ngen > 0 ? ngen : 0;
  // no stream capture in the server: another slot's allocations could
  // synchronise the device in the middle of a capture
  r->opt.use_graphs = false;
  r->nlc = nlc;
  if (nlc > 0) {
    r->lptr.assign(lptr, lptr + nlc + 1);
    const int nl = lptr[nlc];
    if (nl > 0 && (!lnode || !ldof || !lval)) return -1;
    r->lnode.assign(lnode, lnode + nl);
    r->ldof.assign(ldof, ldof + nl);
    r->lval.assign(lval, lval + nl);
    if (ub && nub > 0) r->ub.assign(ub, ub + (size_t)nub * nlc);
  }
  return s->s->submit(std::move(r));
}

int fgpu_server_set_option(fgpu_server* s, const char* name, double value) {
  if (!s || !name) return -1;
  return fgpu::set_reduce_option(s->base, name, value);
}

int fgpu_server_precision(fgpu_server* s, int id) {
  fgpu::Request* r = s ? s->s->find_done(id) : nullptr;
  if (!r) return -1;
  return r->opt.mixed ? 1 : 0;
}

int fgpu_server_wait(fgpu_server* s, int id) {
  if (!s) return -1;
  fgpu::Request* r = s->s->wait(id);
  return r ? r->status : -1;
}

int fgpu_server_sizes(fgpu_server* s, int id, int* sizes) {
  fgpu::Request* r = s ? s->s->find_done(id) : nullptr;
  if (!r || !sizes) return -1;
  sizes[0] = r->res.ndim;
  sizes[1] = r->res.ndof1;
  sizes[2] = r->res.ndof2;
  sizes[3] = r->nlc;
  sizes[4] = r->res.nev;
  return 0;
}

int fgpu_server_result(fgpu_server* s, int id, double* SK, double* SM, double* G, double* Fred, double* Ui,
                       double* times) {
  fgpu::Request* r = s ? s->s->find_done(id) : nullptr;
  if (!r) return -1;
  if (r->status) return r->status;
  auto put = [](double* dst, const std::vector<double>& v) {
    if (dst && !v.empty()) std::memcpy(dst, v.data(), sizeof(double) * v.size());
  };
  put(SK, r->res.SK);
  put(SM, r->res.SM);
  put(G, r->res.G);
  put(Fred, r->fred);
  put(Ui, r->ui);
  if (times) {
    times[0] = fgpu::ms(r->t_submit, r->t_start);  // queueing and admission
    times[1] = fgpu::ms(r->t_start, r->t_end);     // service on the device
    times[2] = fgpu::ms(r->t_submit, r->t_end);    // latency
    times[3] = (double)r->footprint;
  }
  return 0;
}

int fgpu_server_error(fgpu_server* s, int id, char* buf, int len) {
  fgpu::Request* r = s ? s->s->find_done(id) : nullptr;
  if (!r || !buf || len <= 0) return -1;
  std::snprintf(buf, (size_t)len, "%s", r->err.c_str());
  return 0;
}

int fgpu_server_release(fgpu_server* s, int id) {
  if (!s) return -1;
  s->s->release(id);
  return 0;
}

int64_t fgpu_server_device_bytes(fgpu_server* s) { return s ? s->s->held_bytes() : -1; }

}  // extern "C"
