// serve.hip -- the serving runtime: a persistent request loop that runs many
// load-stepped quasi-static solves at once on one GPU (declared in
// include/lpm_gpu.h, lpmg_server_*).  The reference has no counterpart: it runs
// one Assembly per process.
//
// Threads:
//   workers  one per slot.  A worker admits the head of the FIFO when the
//            request's device-memory estimate fits the budget next to the
//            admitted requests, then runs SolverStatic::solveProblem
//            (solver_static.h:64-95) for it on its own handle and stream:
//            bond/conn setup, forward passes, assembly and state updates of
//            different requests overlap on the device.
//   batcher  runs the CG iterations of every request that is in a linear
//            solve.  A worker's CG start (cg_begin) is enqueued on its own
//            stream; the worker then hands the system to the batcher and
//            sleeps.  Per chunk of kCgChunk iterations the batcher launches,
//            for each batch class (dimension, precision, chunks per pass) with
//            members, one fused SpMV, one fused reduction (large systems only)
//            and one fused vector launch per iteration, then reads every
//            member's done flag.  Systems join at chunk boundaries (after
//            their start event) and leave when their flag is seen; chunks are
//            pipelined two deep, so the flags of chunk k are read while chunk
//            k+1 runs.  A leaving system's stream waits on an event recorded
//            after the last chunk that still lists it.
//            A chunk is a captured hipGraph: the fused kernels read the live
//            system count and workgroup totals from a device header uploaded
//            before each chunk, and their grids are rounded up to a bucket
//            (<= 25% above the live count), so one graph per bucket combination
//            serves every membership that fits it; graphs are cached.
#include <algorithm>
#include <array>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/lpm_gpu.h"
#include "lpmg_internal.h"

namespace lpmg {

namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

struct CgJob {
  Context* c = nullptr;
  CgSys q{};
  bool finished = false, retired = false;
  int err = 0;
  std::string msg;
};

class Batcher {
 public:
  int start(int device, int cap) {
    device_ = device;
    cap_ = cap;
    LPMG_HIP(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
    LPMG_HIP(hipStreamCreateWithFlags(&capst_, hipStreamNonBlocking));
    for (int s = 0; s < 2; ++s) {
      for (int k = 0; k < kCgClasses; ++k) {
        LPMG_HIP(hipMalloc((void**)&dbuf_[s][k], kCgHdrBytes + sizeof(CgSys) * cap));
        LPMG_HIP(hipHostMalloc((void**)&hbuf_[s][k], kCgHdrBytes + sizeof(CgSys) * cap, hipHostMallocDefault));
      }
      LPMG_HIP(hipMalloc((void**)&dflags_[s], sizeof(int) * 2 * kCgClasses * cap));
      LPMG_HIP(hipHostMalloc((void**)&hflags_[s], sizeof(int) * 2 * kCgClasses * cap, hipHostMallocDefault));
      LPMG_HIP(hipEventCreateWithFlags(&cev_[s], hipEventDisableTiming));
    }
    th_ = std::thread([this] { loop(); });
    started_ = true;
    return 0;
  }

  void stop() {
    if (started_) {
      {
        std::lock_guard<std::mutex> lk(m_);
        stopping_ = true;
      }
      cv_.notify_all();
      th_.join();
      started_ = false;
    }
    if (st_) (void)hipStreamSynchronize(st_);
    for (auto& g : graphs_) (void)hipGraphExecDestroy(g.second);
    graphs_.clear();
    for (int s = 0; s < 2; ++s) {
      for (int k = 0; k < kCgClasses; ++k) {
        if (dbuf_[s][k]) (void)hipFree(dbuf_[s][k]);
        if (hbuf_[s][k]) (void)hipHostFree(hbuf_[s][k]);
      }
      if (dflags_[s]) (void)hipFree(dflags_[s]);
      if (hflags_[s]) (void)hipHostFree(hflags_[s]);
      if (cev_[s]) (void)hipEventDestroy(cev_[s]);
    }
    if (st_) (void)hipStreamDestroy(st_);
    if (capst_) (void)hipStreamDestroy(capst_);
    st_ = capst_ = nullptr;
  }

  // Worker side: the system's start is enqueued on c->stream.  Returns once
  // the system has converged (or hit its iteration limit), with c->stream
  // ordered after the last fused launch that lists it.
  int exec(Context* c, const CgSys& q) {
    if (!c->cg_ready) {
      LPMG_HIP(hipEventCreateWithFlags(&c->cg_ready, hipEventDisableTiming));
      LPMG_HIP(hipEventCreateWithFlags(&c->cg_released, hipEventDisableTiming));
    }
    LPMG_HIP(hipEventRecord(c->cg_ready, c->stream));
    auto j = std::make_shared<CgJob>();
    j->c = c;
    j->q = q;
    {
      std::unique_lock<std::mutex> lk(m_);
      if (fatal_) {
        set_error("CG batcher stopped after an error: " + fatal_msg_);
        return -2;
      }
      incoming_.push_back(j);
      cv_.notify_all();
      done_cv_.wait(lk, [&] { return j->finished; });
    }
    if (j->err) {
      set_error(j->msg);
      return j->err;
    }
    LPMG_HIP(hipStreamWaitEvent(c->stream, c->cg_released, 0));
    return 0;
  }

  void info(double* chunks, double* sys_chunks, double* max_batch, double* graphs) {
    std::lock_guard<std::mutex> lk(m_);
    *chunks = (double)chunks_;
    *sys_chunks = (double)sys_chunks_;
    *max_batch = (double)max_batch_;
    *graphs = (double)graphs_made_;
  }

 private:
  struct Flight {
    bool valid = false;
    std::vector<std::shared_ptr<CgJob>> members;
  };

  // launch grid for a live workgroup count: rounded up to a multiple of a
  // quarter of its leading power of two (at least 16)
  static int bucket(int x) {
    if (x <= 0) return 0;
    int p = 1;
    while (p * 2 <= x) p *= 2;
    const int g = std::max(16, p / 4);
    return (x + g - 1) / g * g;
  }

  // the captured chunk for this slot and these grids (captured on first use)
  int chunk_graph(int slot, const CgBatchClass* cls, hipGraphExec_t* out) {
    std::array<int, 1 + 3 * kCgClasses> key{};
    key[0] = slot;
    for (int k = 0; k < kCgClasses; ++k) {
      key[1 + 3 * k] = cls[k].sp_grid;
      key[2 + 3 * k] = cls[k].red_grid;
      key[3 + 3 * k] = cls[k].vec_grid;
    }
    auto it = graphs_.find(key);
    if (it != graphs_.end()) {
      *out = it->second;
      return 0;
    }
    if (graphs_.size() >= 1024) {  // bound the cache; every graph is idle once the stream drains
      LPMG_HIP(hipStreamSynchronize(st_));
      for (auto& g : graphs_) (void)hipGraphExecDestroy(g.second);
      graphs_.clear();
    }
    // captured on a stream that never runs work: the chunk stream itself is
    // never in capture mode, so events recorded on it stay plain events for
    // the workers' streams to wait on
    hipGraph_t graph = nullptr;
    LPMG_HIP(hipStreamBeginCapture(capst_, hipStreamCaptureModeThreadLocal));
    const int rc = cg_batch_enqueue_chunk(capst_, cls, kCgClasses);
    const hipError_t e = hipStreamEndCapture(capst_, &graph);
    if (rc) return rc;
    LPMG_HIP(e);
    hipGraphExec_t exec = nullptr;
    const hipError_t e2 = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    LPMG_HIP(e2);
    graphs_[key] = exec;
    {
      std::lock_guard<std::mutex> lk(m_);
      ++graphs_made_;
    }
    *out = exec;
    return 0;
  }

  int n_active() const {
    int n = 0;
    for (const auto& a : active_) n += (int)a.size();
    return n;
  }

  // finish a job: its stream will wait for everything enqueued so far
  void retire(const std::shared_ptr<CgJob>& j, int err, const std::string& msg) {
    auto& a = active_[j->q.cls];
    a.erase(std::remove(a.begin(), a.end(), j), a.end());
    j->retired = true;
    const bool rec_failed = !err && hipEventRecord(j->c->cg_released, st_) != hipSuccess;
    std::lock_guard<std::mutex> lk(m_);
    if (rec_failed) {
      err = -2;
      fatal_ = true;
      fatal_msg_ = "hipEventRecord failed";
    }
    j->err = err;
    j->msg = msg;
    j->finished = true;
  }

  // a HIP error in the batcher ends every job in it; later exec calls fail
  void fail_all(const std::string& what) {
    {
      std::lock_guard<std::mutex> lk(m_);
      fatal_ = true;
      fatal_msg_ = what;
    }
    std::vector<std::shared_ptr<CgJob>> all;
    for (auto& a : active_)
      for (auto& j : a) all.push_back(j);
    for (auto& f : flight_) f = Flight{};
    for (auto& j : all) retire(j, -2, "CG batcher: " + what);
    std::vector<std::shared_ptr<CgJob>> inc;
    {
      std::lock_guard<std::mutex> lk(m_);
      inc.swap(incoming_);
      for (auto& j : inc) {
        j->err = -2;
        j->msg = "CG batcher: " + what;
        j->finished = true;
      }
    }
    done_cv_.notify_all();
  }

  // one pass: admit joiners, launch a chunk, read the previous chunk's flags
  bool step(int* chunk) {
    std::vector<std::shared_ptr<CgJob>> joining;
    {
      std::unique_lock<std::mutex> lk(m_);
      auto idle = [&] { return n_active() == 0 && !flight_[0].valid && !flight_[1].valid; };
      cv_.wait(lk, [&] { return stopping_ || !incoming_.empty() || !idle() || fatal_; });
      if (fatal_ || (stopping_ && incoming_.empty() && idle())) return false;
      joining.swap(incoming_);
    }
    for (auto& j : joining) {
      if (hipStreamWaitEvent(st_, j->c->cg_ready, 0) != hipSuccess) {
        active_[j->q.cls].push_back(j);
        fail_all("hipStreamWaitEvent failed");
        return false;
      }
      active_[j->q.cls].push_back(j);
    }
    const int slot = *chunk & 1;  // its buffers' previous chunk was read already
    bool launched = false;
    if (n_active() > 0) {
      CgBatchClass cls[kCgClasses];
      Flight& f = flight_[slot];
      f.members.clear();
      int nflag = 0;
      for (int k = 0; k < kCgClasses; ++k) {
        const auto& a = active_[k];
        cls[k].cls = k;
        if (a.empty()) continue;
        CgBatchHdr* hh = reinterpret_cast<CgBatchHdr*>(hbuf_[slot][k]);
        CgSys* ht = reinterpret_cast<CgSys*>(hbuf_[slot][k] + kCgHdrBytes);
        for (size_t t = 0; t < a.size(); ++t) ht[t] = a[t]->q;
        cg_batch_layout(ht, (int)a.size(), hh);
        cls[k].hdr = reinterpret_cast<const CgBatchHdr*>(dbuf_[slot][k]);
        cls[k].live = *hh;
        cls[k].sp_grid = bucket(hh->sp_blocks);
        cls[k].red_grid = bucket(hh->red_blocks);
        cls[k].vec_grid = bucket(hh->vec_blocks);
        if (hipMemcpyAsync(dbuf_[slot][k], hbuf_[slot][k], kCgHdrBytes + sizeof(CgSys) * a.size(),
                           hipMemcpyHostToDevice, st_) != hipSuccess) {
          fail_all("table upload failed");
          return false;
        }
      }
      hipGraphExec_t g = nullptr;
      if (chunk_graph(slot, cls, &g) || hipGraphLaunch(g, st_) != hipSuccess) {
        fail_all("fused chunk launch failed");
        return false;
      }
      for (int k = 0; k < kCgClasses; ++k) {
        const auto& a = active_[k];
        if (a.empty()) continue;
        const CgSys* dt = reinterpret_cast<const CgSys*>(dbuf_[slot][k] + kCgHdrBytes);
        if (cg_batch_flags(st_, dt, (int)a.size(), dflags_[slot] + 2 * nflag)) {
          fail_all("flag gather failed");
          return false;
        }
        for (auto& j : a) f.members.push_back(j);
        nflag += (int)a.size();
      }
      if (hipMemcpyAsync(hflags_[slot], dflags_[slot], sizeof(int) * 2 * nflag, hipMemcpyDeviceToHost, st_) !=
              hipSuccess ||
          hipEventRecord(cev_[slot], st_) != hipSuccess) {
        fail_all("flag read-back failed");
        return false;
      }
      f.valid = true;
      launched = true;
      std::lock_guard<std::mutex> lk(m_);
      chunks_ += 1;
      sys_chunks_ += nflag;
      max_batch_ = std::max(max_batch_, nflag);
    }
    Flight& pf = flight_[slot ^ 1];
    if (pf.valid) {
      if (hipEventSynchronize(cev_[slot ^ 1]) != hipSuccess) {
        fail_all("chunk synchronization failed");
        return false;
      }
      bool any = false;
      for (size_t t = 0; t < pf.members.size(); ++t) {
        auto& j = pf.members[t];
        if (j->retired || !hflags_[slot ^ 1][2 * t]) continue;
        retire(j, 0, std::string());
        any = true;
      }
      pf.valid = false;
      pf.members.clear();
      if (any) done_cv_.notify_all();
    }
    if (launched) ++*chunk;
    return true;
  }

  void loop() {
    if (hipSetDevice(device_) != hipSuccess) {
      fail_all("hipSetDevice failed");
      return;
    }
    int chunk = 0;
    while (step(&chunk)) {
    }
    if (st_) (void)hipStreamSynchronize(st_);
  }

  int device_ = 0, cap_ = 0;
  hipStream_t st_ = nullptr, capst_ = nullptr;  // chunk stream, capture-only stream
  char* dbuf_[2][kCgClasses] = {};  // header + table per chunk slot and class
  char* hbuf_[2][kCgClasses] = {};
  std::map<std::array<int, 1 + 3 * kCgClasses>, hipGraphExec_t> graphs_;
  int64_t graphs_made_ = 0;
  int* dflags_[2] = {};
  int* hflags_[2] = {};
  hipEvent_t cev_[2] = {};
  Flight flight_[2];
  std::vector<std::shared_ptr<CgJob>> active_[kCgClasses];
  std::mutex m_;
  std::condition_variable cv_, done_cv_;
  std::vector<std::shared_ptr<CgJob>> incoming_;
  bool stopping_ = false, fatal_ = false, started_ = false;
  std::string fatal_msg_;
  std::thread th_;
  int64_t chunks_ = 0, sys_chunks_ = 0;
  int max_batch_ = 0;
};

int batcher_exec(void* arg, Context* c, const CgSys& q) { return static_cast<Batcher*>(arg)->exec(c, q); }

// A request with its own copies of the inputs
struct Req {
  int64_t ticket = 0;
  lpmg_request r{};
  std::vector<double> xyz, params;
  std::vector<int32_t> p1, p2, layer, type, gsize, gidx;
  std::vector<lpmg_bc> bc;
  std::vector<lpmg_out_field> out;
  int64_t est = 0;
  Clock::time_point t_submit, t_admit, t_done;
  bool finished = false;
  int status = 0;
  std::string msg;
  double stats[LPMG_RS_N] = {0};
};

int check_request(const lpmg_request* r) {
  if (!r) { set_error("null request"); return -1; }
  Cell cell;
  if (!make_cell(r->lattice, r->radius, &cell) || !(r->radius > 0)) { set_error("bad lattice or radius"); return -1; }
  if (r->n <= 0 || !r->xyz) { set_error("bad particle count or coordinates"); return -1; }
  if (r->ptype < 0 || r->ptype > 2) { set_error("bad particle type"); return -1; }
  if (r->nbond < 0 || (r->nbond > 0 && (!r->p1 || !r->p2 || !r->layer))) { set_error("bad bond list"); return -1; }
  static const int need[5] = {5, 5, 8, 8, 9};  // lpmg_set_material's parameter counts per model
  if (r->model < 0 || r->model > 4 || r->n_params < need[r->model] || !r->params) {
    set_error("bad material model or parameters");
    return -1;
  }
  if (r->n_groups < 0 || (r->n_groups > 0 && !r->group_size)) { set_error("bad groups"); return -1; }
  int64_t tot = 0;
  for (int g = 0; g < r->n_groups; ++g) {
    if (r->group_size[g] < 0) { set_error("negative group size"); return -1; }
    tot += r->group_size[g];
  }
  if (tot > 0 && !r->group_idx) { set_error("missing group indices"); return -1; }
  if (r->n_bc < 0 || (r->n_bc > 0 && !r->bc) || r->n_steps < 0) { set_error("bad load steps"); return -1; }
  for (int b = 0; b < r->n_bc; ++b) {
    const lpmg_bc& x = r->bc[b];
    if (x.kind < 0 || x.kind > 1 || x.group < 0 || x.group >= r->n_groups || x.mode < 0 || x.mode > 1 ||
        (x.kind == 0 && (x.flag < 'x' || x.flag > 'z'))) {
      set_error("bad boundary condition " + std::to_string(b));
      return -1;
    }
  }
  if (r->n_out < 0 || (r->n_out > 0 && !r->out)) { set_error("bad output fields"); return -1; }
  for (int k = 0; k < r->n_out; ++k)
    if (!r->out[k].data || r->out[k].capacity < 0) { set_error("bad output field " + std::to_string(k)); return -1; }
  if (r->step_stats_cap < 0 || (r->step_stats_cap > 0 && !r->step_stats)) { set_error("bad step statistics"); return -1; }
  if (r->max_iter <= 0 || r->max_halvings < 0) { set_error("bad Newton settings"); return -1; }
  return 0;
}

int64_t estimate_bytes(const lpmg_request* r) {
  Cell cell;
  make_cell(r->lattice, r->radius, &cell);
  const double N = r->n, dim = cell.dim, DD = dim * dim;
  const double NB = r->nbond > 0 ? (double)r->nbond : N * cell.nneighbors;
  // block-CSR blocks: the lattices' interior rows hold 1 + nb + (2-hop) <= 1 + 4 nb
  const double nnzb = N + 4.0 * NB;
  const bool mixed = r->opts[4] != 0;
  const double n_partials = std::max(std::ceil(N / 256), 4096.0) + std::ceil(N / 8) + 4096 + 64;
  double b = N * (680.0 + 92.0 * dim) + 16.0 * n_partials;     // particle and solver arrays, partials
  b += NB * 120.0;                                              // bond arrays
  b += nnzb * (12.0 + DD * 16.0 + (mixed ? DD * 4.0 : 0.0));   // conns, BSR values (tangent + symmetrised)
  b += N * 48.0;                                                // setup scratch
  for (int g = 0; g < r->n_groups; ++g) b += 4.0 * std::max(1, (int)r->group_size[g]) + 256;
  if (r->opts[0] == 4 || (r->opts[0] < 0 && (dim == 2 || r->lattice == LPMG_BCC3D))) {
    // smoothed-aggregation multigrid: per fine row about Pr prolongator
    // blocks (values, FP64 and FP32, in row and column order, and five index
    // arrays), 3 Pr blocks of A P (FP64 values), a 2-byte slot per term of A P
    // (A's blocks of the row times the P blocks of their columns) and of
    // P^T (A P) (P's blocks of the row times the row's A P blocks), and a
    // 4-byte offset per source of either; the coarser levels add up to 20%.
    // Counted on the CT plates: Pr 8.0, A P 21.3, 244 and 172 terms per row;
    // the 3D figure holds the 1.0M-particle BCC polycrystal's hierarchy (5.5 GB
    // measured) with a 1.8x margin.
    const double pr = dim == 2 ? 9.0 : 12.0, yr = 3.0 * pr, k = nnzb / N;
    b += 1.2 * N * (2.0 * (k * pr + pr * yr) + 4.0 * (k + pr) + yr * DD * 8.0 + pr * (DD * 24.0 + 24.0));
  }
  if (r->nonlocal_ratio > 0 && r->n_params > 0) {
    // nonlocal neighbours within ratio * L (params[0] is nonlocal_L)
    const double R = std::max(r->nonlocal_ratio * r->params[0], 1.01 * cell.cut[1]);
    const double per = dim == 3 ? 4.19 * std::pow(R / (2 * r->radius), 3) * 1.5
                                : 3.15 * std::pow(R / (2 * r->radius), 2) * 1.2;
    b += N * (4.0 + 4.0 * (per + 8));
  }
  return (int64_t)(b + 1e6);
}

}  // namespace
}  // namespace lpmg

using namespace lpmg;

struct lpmg_server_s {
  int device = 0, max_active = 1, batching = 1;
  int64_t fuse_max_rows = -1;  // systems with more block rows run their CG on their own stream
  int64_t budget = 0;
  std::mutex m;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<Req>> queue;
  std::map<int64_t, std::shared_ptr<Req>> reqs;
  int64_t next_ticket = 1, reserved = 0, peak_reserved = 0, finished = 0;
  int running = 0, peak_running = 0;
  bool stopping = false;
  std::vector<std::thread> workers;
  Batcher batcher;
};

namespace {

// SolverStatic::solveProblem (solver_static.h:64-95) for one request: each
// load step applies its force BCs then its displacement BCs
// (solver_static.h:101-102) and runs solveProblemStep; a step that does not
// converge resets the state (positions included), is cut in half and runs twice.
int run_request(lpmg_server_s* S, Req* q, int64_t* bytes) {
  const lpmg_request& r = q->r;
  lpmg_handle h = nullptr;
  int rc = r.nbond > 0 ? lpmg_create_with_bonds(r.lattice, r.radius, r.ptype, r.n, r.xyz, r.nbond, r.p1, r.p2,
                                                r.layer, &h)
                       : lpmg_create(r.lattice, r.radius, r.ptype, r.n, r.xyz, &h);
  if (rc) return rc;
  struct Guard {
    lpmg_handle h;
    ~Guard() { lpmg_destroy(h); }
  } guard{h};
  Context* c = handle_context(h);
  if (r.type && (rc = lpmg_set_types(h, r.type))) return rc;
  if ((rc = lpmg_set_material(h, 0, nullptr, r.model, r.params)) < 0) return rc;
  std::vector<int> gid(r.n_groups);
  int64_t off = 0;
  for (int g = 0; g < r.n_groups; ++g) {
    gid[g] = lpmg_define_group(h, r.group_size[g], r.group_idx ? r.group_idx + off : nullptr);
    if (gid[g] < 0) return gid[g];
    off += r.group_size[g];
  }
  if (r.nonlocal_ratio > 0 && (rc = lpmg_search_nonlocal(h, r.nonlocal_ratio))) return rc;
  // the fused launches run the Jacobi CG; multigrid solves use the handle's
  // stream.  opts[0] < 0 lets the server choose: fused Jacobi for the systems
  // it fuses; for the larger ones smoothed-aggregation multigrid in 2D
  // (Jacobi needs thousands of iterations per solve there; the tentative
  // aggregation measured 3.5-4x faster than Jacobi at 1e5-2.5e5 particles, the
  // smoothed one about 2x faster again on the CT stream's large requests) and
  // on the BCC lattice (mixed precision: 1.0M-particle polycrystal 2.5 s
  // against 5.5 s, 108k-particle J2 bar 0.92 against 1.83 s), Jacobi on the
  // other 3D lattices (SC 40^3 and 60^3 need so few Jacobi iterations that
  // the V-cycle does not pay: 0.34 against 0.15 s, 0.70 against 0.40 s).
  const bool fusable = S->batching && (S->fuse_max_rows < 0 || (int64_t)c->N <= S->fuse_max_rows);
  double opts[8];
  std::memcpy(opts, r.opts, sizeof(opts));
  if (opts[0] < 0) opts[0] = (!fusable && (c->dim == 2 || r.lattice == LPMG_BCC3D)) ? 4 : 1;
  if (fusable && opts[0] <= 1) {
    c->cg_exec = batcher_exec;
    c->cg_exec_arg = &S->batcher;
  }
  if ((rc = lpmg_forward(h))) return rc;
  if (r.prepare_state && (rc = lpmg_update_state_var(h))) return rc;
  std::vector<double> scales(r.n_steps, 1.0);
  double last_scale = 1.0;  // size of the last converged step, for the predictor
  int halvings = 0, done_steps = 0;
  double newton = 0, cg = 0, passes = 0;
  for (size_t i = 0; i < scales.size();) {
    const double sc = scales[i];
    for (int b = 0; b < r.n_bc; ++b) {
      const lpmg_bc& x = r.bc[b];
      if (x.kind == 1 &&
          (rc = lpmg_force_bc(h, gid[x.group], x.mode, x.value[0] * sc, x.value[1] * sc, x.value[2] * sc)))
        return rc;
    }
    for (int b = 0; b < r.n_bc; ++b) {
      const lpmg_bc& x = r.bc[b];
      if (x.kind == 0 && (rc = lpmg_disp_bc(h, gid[x.group], x.mode, x.flag, x.value[0] * sc))) return rc;
    }
    double st[3] = {0, 0, 0};
    if (r.opts[6] > 0) opts[6] = r.opts[6] * sc / last_scale;
    rc = lpmg_solve_step(h, r.undamaged, r.max_iter, r.tol, opts, st);
    if (rc < 0) return rc;
    if (rc == 0) {  // not converged: restore the last converged state, halve the step
      if (halvings >= r.max_halvings) {
        set_error("load step " + std::to_string(i) + " did not converge after " + std::to_string(halvings) +
                  " halvings");
        return -4;
      }
      ++halvings;
      if ((rc = lpmg_reset_state(h, 1))) return rc;
      scales[i] = sc / 2;
      scales.insert(scales.begin() + i, sc / 2);
      continue;
    }
    if (done_steps < r.step_stats_cap) {
      r.step_stats[3 * done_steps] = st[2];
      r.step_stats[3 * done_steps + 1] = st[0];
      r.step_stats[3 * done_steps + 2] = st[1];
    }
    newton += st[2];
    cg += st[0];
    passes += st[1];
    last_scale = sc;
    ++done_steps;
    ++i;
  }
  for (int k = 0; k < r.n_out; ++k) {
    const lpmg_out_field& o = r.out[k];
    const int64_t len = lpmg_field_len(h, o.field);
    if (len < 0 || len > o.capacity) {
      set_error("output field " + std::to_string(o.field) + " needs " + std::to_string(len) + " doubles");
      return -1;
    }
    if ((rc = lpmg_get_field(h, o.field, o.data))) return rc;
  }
  q->stats[LPMG_RS_STEPS] = done_steps;
  q->stats[LPMG_RS_HALVINGS] = halvings;
  q->stats[LPMG_RS_NEWTON] = newton;
  q->stats[LPMG_RS_CG] = cg;
  q->stats[LPMG_RS_PASSES] = passes;
  *bytes = (int64_t)c->bytes;
  return 0;
}

void worker_main(lpmg_server_s* S) {
  const bool dev_ok = hipSetDevice(S->device) == hipSuccess;
  for (;;) {
    std::shared_ptr<Req> q;
    {
      std::unique_lock<std::mutex> lk(S->m);
      S->cv_work.wait(lk, [&] {
        return (S->stopping && S->queue.empty()) ||
               (!S->queue.empty() && S->reserved + S->queue.front()->est <= S->budget);
      });
      if (S->queue.empty()) return;
      q = S->queue.front();
      S->queue.pop_front();
      S->reserved += q->est;
      S->running += 1;
      S->peak_reserved = std::max(S->peak_reserved, S->reserved);
      S->peak_running = std::max(S->peak_running, S->running);
      q->t_admit = Clock::now();
    }
    S->cv_work.notify_all();  // the next request may fit too
    int64_t bytes = 0;
    int rc = dev_ok ? run_request(S, q.get(), &bytes) : -2;
    std::string msg = rc ? std::string(lpmg_last_error()) : std::string();
    if (!dev_ok) msg = "hipSetDevice failed in a worker";
    {
      std::lock_guard<std::mutex> lk(S->m);
      S->reserved -= q->est;
      S->running -= 1;
      S->finished += 1;
      q->t_done = Clock::now();
      q->status = rc;
      q->msg = msg;
      q->stats[LPMG_RS_STATUS] = rc;
      q->stats[LPMG_RS_QUEUE_MS] = ms_between(q->t_submit, q->t_admit);
      q->stats[LPMG_RS_RUN_MS] = ms_between(q->t_admit, q->t_done);
      q->stats[LPMG_RS_TOTAL_MS] = ms_between(q->t_submit, q->t_done);
      q->stats[LPMG_RS_BYTES_EST] = (double)q->est;
      q->stats[LPMG_RS_BYTES] = (double)bytes;
      q->finished = true;
    }
    S->cv_done.notify_all();
    S->cv_work.notify_all();
  }
}

}  // namespace

extern "C" {

int64_t lpmg_request_bytes(const lpmg_request* r) {
  if (check_request(r)) return -1;
  return estimate_bytes(r);
}

int lpmg_server_create(int max_active, int64_t mem_budget, int64_t fuse_max_rows, lpmg_server* out) {
  if (!out) { set_error("null output server"); return -1; }
  *out = nullptr;
  if (max_active <= 0 || max_active > 256) { set_error("max_active must be in 1..256"); return -1; }
  auto* S = new lpmg_server_s;
  S->max_active = max_active;
  S->batching = fuse_max_rows != 0 ? 1 : 0;
  S->fuse_max_rows = fuse_max_rows;
  hipError_t e = hipGetDevice(&S->device);
  size_t fr = 0, tot = 0;
  if (e == hipSuccess) e = hipMemGetInfo(&fr, &tot);
  if (e != hipSuccess) {
    delete S;
    return hip_fail(e, "hipGetDevice/hipMemGetInfo");
  }
  S->budget = mem_budget > 0 ? mem_budget : (int64_t)(0.9 * (double)fr);
  if (S->batching)
    if (int rc = S->batcher.start(S->device, max_active)) {
      S->batcher.stop();
      delete S;
      return rc;
    }
  for (int k = 0; k < max_active; ++k) S->workers.emplace_back(worker_main, S);
  *out = S;
  return 0;
}

int lpmg_server_submit(lpmg_server S, const lpmg_request* r, int64_t* ticket) {
  if (!S || !ticket) { set_error("null server or ticket"); return -1; }
  if (int rc = check_request(r)) return rc;
  auto q = std::make_shared<Req>();
  q->r = *r;
  q->xyz.assign(r->xyz, r->xyz + 3 * (size_t)r->n);
  q->r.xyz = q->xyz.data();
  if (r->nbond > 0) {
    q->p1.assign(r->p1, r->p1 + r->nbond);
    q->p2.assign(r->p2, r->p2 + r->nbond);
    q->layer.assign(r->layer, r->layer + r->nbond);
    q->r.p1 = q->p1.data();
    q->r.p2 = q->p2.data();
    q->r.layer = q->layer.data();
  }
  if (r->type) {
    q->type.assign(r->type, r->type + r->n);
    q->r.type = q->type.data();
  }
  q->params.assign(r->params, r->params + r->n_params);
  q->r.params = q->params.data();
  q->gsize.assign(r->group_size, r->group_size + r->n_groups);
  q->r.group_size = q->gsize.data();
  int64_t tot = 0;
  for (int g = 0; g < r->n_groups; ++g) tot += r->group_size[g];
  for (int64_t t = 0; t < tot; ++t)
    if (r->group_idx[t] < 0 || r->group_idx[t] >= r->n) { set_error("group index out of range"); return -1; }
  q->gidx.assign(r->group_idx, r->group_idx + tot);
  q->r.group_idx = q->gidx.data();
  q->bc.assign(r->bc, r->bc + r->n_bc);
  q->r.bc = q->bc.data();
  q->out.assign(r->out, r->out + r->n_out);
  q->r.out = q->out.data();
  q->est = estimate_bytes(r);
  {
    std::lock_guard<std::mutex> lk(S->m);
    if (q->est > S->budget) {
      set_error("request needs about " + std::to_string(q->est) + " bytes, more than the server's budget of " +
                std::to_string(S->budget));
      return -3;
    }
    if (S->stopping) { set_error("server is stopping"); return -1; }
    q->ticket = S->next_ticket++;
    q->t_submit = Clock::now();
    S->queue.push_back(q);
    S->reqs[q->ticket] = q;
    *ticket = q->ticket;
  }
  S->cv_work.notify_all();
  return 0;
}

int lpmg_server_wait(lpmg_server S, int64_t ticket, double timeout_s, double* stats) {
  if (!S) { set_error("null server"); return -1; }
  std::shared_ptr<Req> q;
  {
    std::unique_lock<std::mutex> lk(S->m);
    auto it = S->reqs.find(ticket);
    if (it == S->reqs.end()) { set_error("unknown ticket"); return -1; }
    q = it->second;
    auto fin = [&] { return q->finished; };
    if (timeout_s < 0) S->cv_done.wait(lk, fin);
    else if (!S->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), fin)) return 1;
    S->reqs.erase(ticket);
  }
  if (stats) std::memcpy(stats, q->stats, sizeof(q->stats));
  if (q->status) set_error(q->msg);
  return q->status;
}

int lpmg_server_info(lpmg_server S, double* out) {
  if (!S || !out) { set_error("null server or output"); return -1; }
  {
    std::lock_guard<std::mutex> lk(S->m);
    out[0] = (double)S->queue.size();
    out[1] = S->running;
    out[2] = (double)S->reserved;
    out[3] = (double)S->budget;
    out[4] = S->peak_running;
    out[5] = (double)S->peak_reserved;
    out[6] = (double)S->finished;
  }
  out[7] = out[8] = out[9] = out[10] = 0;
  if (S->batching) S->batcher.info(&out[7], &out[8], &out[9], &out[10]);
  return 0;
}

void lpmg_server_destroy(lpmg_server S) {
  if (!S) return;
  {
    std::lock_guard<std::mutex> lk(S->m);
    S->stopping = true;
  }
  S->cv_work.notify_all();
  for (auto& t : S->workers) t.join();
  if (S->batching) S->batcher.stop();
  delete S;
}

}  // extern "C"
