// Batched objective + gradient over many independent designs (include/tmx.h, tmx_batch_*).
//
// A batch step runs in one of two ways per design:
//
//  * merged (the default for every design that qualifies): each design's objective step is
//    recorded once as launch sequences, one per phase (tmx_multi.h); the sequences of all the
//    designs are merged into one launch per kernel instance and position, and the step runs as
//    those merged launches on the batch's stream.  The PCG iterations of all designs run in
//    lockstep, one merged launch per PCG phase and V-cycle stage, so the launches of a batch
//    step follow the deepest hierarchy and the largest iteration count, not the number of
//    designs.  A design that has converged goes through the remaining iterations of its solve
//    as the single-design solver goes through the iteration it queued behind convergence: its
//    stop flag has set the step sizes to zero, so its iterate and residual do not change, and
//    its results equal those of a single-design call bit for bit.
//  * alone: the design's step is a single-design call on its own stream, driven by a worker
//    thread.  Designs that do not qualify run this way, concurrently with the merged step, and
//    so do designs whose merged solve stopped without converging (breakdown, non-finite data,
//    iteration limit): they are run again alone, so that they take the single-design solver's
//    fallbacks and report exactly as a single call does.
//
// A design qualifies for the merged step once an objective step in the same precision has
// completed on its handle (the smoother bound and the filter preconditioner are then built),
// with the P1 element, FP64 or mixed precision and no option that the merged step does not
// record (graph replay, kernel timing).  The recorded launches are checked: each phase is
// recorded under stream capture on a private stream, and a phase that would launch anything
// outside the recorder sends the design to the single-design path.
#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

#include "tmx_handle.h"
#include "tmx_multi.h"

namespace {

// phases of one objective step, in order; the PCG iterations of each solve come in two
// parities and the solve's end (the last deferred x update) in two variants
enum Phase : int {
  kP0 = 0,                    // M rho, filter PCG start
  kF1It0, kF1It1,             // filter iterations
  kF1End0, kF1End1,           // filter end
  kP1,                        // xi out, elasticity operator rebuild, elasticity PCG start
  kElIt0, kElIt1,             // elasticity iterations
  kElEnd0, kElEnd1,           // elasticity end
  kP2,                        // u out, true residual, J, sensitivity, adjoint filter start
  kF2It0, kF2It1,             // adjoint filter iterations
  kF2End0, kF2End1,           // adjoint filter end
  kP3,                        // w out
  kPhases
};

// one merged launch: the kernel, its grid, LDS and the byte offsets of its argument records,
// block -> design table and first-block table in the batch's table blob
struct Merged {
  tmx::MultiFn fn;
  unsigned grid;
  size_t lds;
  size_t rec, first, owner;
  size_t skip;  // offset of the designs' skip-flag pointers, or 0: no skipping
};

size_t align_up(size_t v, size_t a) { return (v + a - 1) & ~(a - 1); }

// Merges the launch sequences of several designs (one recorder each, same phase) into one
// launch per kernel instance at each step: at every step the kernel at the head of the longest
// remaining sequence is taken (ties: the one at the head of the most sequences), and the heads
// of all sequences showing it are merged.  Each sequence keeps its order, and the designs are
// independent, so any such interleaving is a valid schedule.  Taking the critical path first
// lets the shorter hierarchies' V-cycles meet the deeper ones at the common levels: on the
// 64-design workload it needs 21 launches per elasticity iteration (18 for the deepest
// design alone) where taking the most common head needed 45.
// skip: null, or each design's skip flag (iteration phases: a design whose solve has stopped
// and applied its last update sits the remaining iterations out).
void merge_phase(const std::vector<const tmx::Recorder*>& recs,
                 const std::vector<const double*>* skip, std::vector<Merged>& out,
                 std::vector<unsigned char>& blob) {
  const size_t D = recs.size();
  std::vector<size_t> pos(D, 0);
  std::vector<tmx::MultiFn> fns;
  std::vector<size_t> count, left;
  std::vector<size_t> take;
  for (;;) {
    fns.clear();
    count.clear();
    left.clear();
    for (size_t d = 0; d < D; ++d) {
      if (pos[d] >= recs[d]->recs.size()) continue;
      const tmx::MultiFn f = recs[d]->recs[pos[d]].fn;
      const size_t rem = recs[d]->recs.size() - pos[d];
      size_t k = 0;
      while (k < fns.size() && fns[k] != f) ++k;
      if (k == fns.size()) {
        fns.push_back(f);
        count.push_back(0);
        left.push_back(0);
      }
      ++count[k];
      left[k] = std::max(left[k], rem);
    }
    if (fns.empty()) break;
    size_t best = 0;
    for (size_t k = 1; k < fns.size(); ++k)
      if (left[k] > left[best] || (left[k] == left[best] && count[k] > count[best])) best = k;
    const tmx::MultiFn f = fns[best];
    take.clear();
    for (size_t d = 0; d < D; ++d)
      if (pos[d] < recs[d]->recs.size() && recs[d]->recs[pos[d]].fn == f) take.push_back(d);
    Merged m{f, 0, 0, 0, 0, 0, 0};
    const size_t rsize = recs[take[0]]->recs[pos[take[0]]].size;
    m.rec = align_up(blob.size(), 16);
    blob.resize(m.rec + rsize * take.size());
    for (size_t i = 0; i < take.size(); ++i) {
      const tmx::LaunchRec& r = recs[take[i]]->recs[pos[take[i]]];
      std::memcpy(blob.data() + m.rec + i * rsize, recs[take[i]]->arena.data() + r.off, rsize);
      m.grid += r.nblk;
      m.lds = std::max(m.lds, r.lds);
    }
    m.first = align_up(blob.size(), 4);
    blob.resize(m.first + sizeof(int) * (take.size() + 1));
    // (the owner table starts right after: tables_valid reads the design count from the gap)
    int* first = reinterpret_cast<int*>(blob.data() + m.first);
    first[0] = 0;
    for (size_t i = 0; i < take.size(); ++i)
      first[i + 1] = first[i] + (int)recs[take[i]]->recs[pos[take[i]]].nblk;
    m.owner = blob.size();  // (4-byte aligned already)
    blob.resize(m.owner + sizeof(unsigned short) * m.grid);
    unsigned short* owner = reinterpret_cast<unsigned short*>(blob.data() + m.owner);
    first = reinterpret_cast<int*>(blob.data() + m.first);  // (blob may have moved)
    for (size_t i = 0; i < take.size(); ++i)
      for (int b = first[i]; b < first[i + 1]; ++b) owner[b] = (unsigned short)i;
    if (skip) {
      m.skip = align_up(blob.size(), 8);
      blob.resize(m.skip + sizeof(const double*) * take.size());
      const double** sp = reinterpret_cast<const double**>(blob.data() + m.skip);
      for (size_t i = 0; i < take.size(); ++i) sp[i] = (*skip)[take[i]];
    }
    if (m.grid > 0) out.push_back(m);
    for (size_t d : take) ++pos[d];
  }
}

// Checks a phase's tables before they reach the device: each launch's first-block table
// rises from 0 to its grid, and every block's owner is the design whose range holds it.
bool tables_valid(const std::vector<Merged>& ms, const std::vector<unsigned char>& blob,
                  size_t designs) {
  for (const Merged& m : ms) {
    if (m.owner + sizeof(unsigned short) * m.grid > blob.size()) return false;
    const int* first = reinterpret_cast<const int*>(blob.data() + m.first);
    const unsigned short* owner = reinterpret_cast<const unsigned short*>(blob.data() + m.owner);
    const size_t nd = (m.owner - m.first) / sizeof(int) - 1;  // (first holds nd + 1 entries)
    if (m.skip && m.skip + sizeof(const double*) * nd > blob.size()) return false;
    if (nd < 1 || nd > designs || first[0] != 0 || first[nd] != (int)m.grid) return false;
    for (size_t d = 0; d < nd; ++d) {
      if (first[d + 1] < first[d]) return false;
      for (int b = first[d]; b < first[d + 1]; ++b)
        if (owner[b] != d) return false;
    }
  }
  return true;
}

// PCG progress of one design in a merged solve (the single-design host loop's state)
struct LoopState {
  double bb = -1.0, tol2 = 0.0, flag = 0.0, rr_last = 0.0, ratio = 1.0;
  double rr_end = 0.0;  // r.r of the last report
  bool near = false;
  int iters = 0;
  // the last report this design makes: once stopped (flag seen in report s), iteration s + 1
  // still runs (and reports) if the flag was set in iteration s itself, to apply the deferred
  // x update; the iterations after it are skipped on the device and make no report
  int last = 1 << 30;
};

}  // namespace

struct tmx_batch {
  std::vector<tmx_problem*> probs;
  std::vector<int> order;  // largest first
  // host threads for designs run alone: up to 16 (one per core of a one-GPU share); on the
  // 64-design workload 16 measured 5% faster than 8 and 30% faster than 4
  int workers = (int)std::min(16u, std::max(4u, std::thread::hardware_concurrency()));
  int mode = TMX_BATCH_MERGED;
  int device = 0;
  hipStream_t stream = nullptr;   // the merged step's stream
  hipStream_t capture = nullptr;  // recording: launches that escape the recorder land here
  hipEvent_t ev_in = nullptr, ev_done = nullptr;
  // the recorded and merged plan, reused while the designs, their handles' state, the
  // design buffers and the options stay the same
  bool plan_valid = false;
  std::vector<tmx_problem*> plan_designs;
  std::vector<unsigned long long> plan_versions;
  tmx_solve_opts plan_opts{};
  bool plan_has_opts = false;
  std::vector<int> plan_rejected;  // designs whose recording failed (run alone)
  std::vector<Merged> phase[kPhases];
  std::vector<unsigned char> blob;
  DevBuf<unsigned char> dev_blob;
  unsigned char* pinned = nullptr;
  size_t pinned_size = 0;
  // the merged designs read rho from, and write w to, the batch's own vectors (the recorded
  // launches keep these addresses); each call copies the callers' vectors in and out with one
  // launch each, from a task table in pinned memory
  std::vector<DevBuf<double>> rho_b, w_b;
  DevBuf<CopyTask> dev_tasks;
  CopyTask* tasks = nullptr;
  size_t tasks_cap = 0;
  tmx_batch_report report{};
  ~tmx_batch() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (pinned) (void)hipHostFree(pinned);
    if (tasks) (void)hipHostFree(tasks);
    if (ev_in) (void)hipEventDestroy(ev_in);
    if (ev_done) (void)hipEventDestroy(ev_done);
    if (capture) (void)hipStreamDestroy(capture);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

// the schemes (kSolveFP64 or kSolveMixed, else -1) of a design's elasticity and filter solves
void schemes(const tmx_problem* h, const tmx_solve_opts* o, int* el, int* fl) {
  const SolveParams ep = merge(h->el_prm, o);
  const SolveParams fp = filter_params(h, o);
  *el = (ep.precision == kSolveFP64 || ep.precision == kSolveMixed) ? ep.precision : -1;
  *fl = (fp.precision == kSolveFP64 || fp.precision == kSolveMixed) ? fp.precision : -1;
}

bool qualifies(const tmx_problem* h, const tmx_solve_opts* o, int device) {
  if (h->device != device || h->degree != 1 || !h->has_penalty) return false;
  int el = -1, fl = -1;
  schemes(h, o, &el, &fl);
  if (el < 0 || fl < 0) return false;
  if (!h->merge_ready[el][fl]) return false;
  if (h->el.use_graphs_ || h->fl.use_graphs_ || h->el.au_timing_) return false;
  if (h->el.cheb_auto_ && !h->el.cheb_done_) return false;
  return true;
}

// The launches of one design's objective step, phase by phase, as tmx_objective_and_gradient
// makes them (with xi, u and M w left to the handle's scratch and J and the true residual
// published to pinned host memory by a kernel instead of a copy).
template <typename PF, typename PE>
int record_phase(tmx_problem* h, int ph, const double* rho, double* w, const SolveParams& ep,
                 const SolveParams& fp) {
  hipStream_t s = h->stream;
  Workspace& ws = h->ws;
  const size_t nv = nverts(h);
  switch (ph) {
    case kP0:
      TMX_TRY(mass_apply(h, rho, h->b_s.get()));
      return h->fl.pcg_enqueue_begin<double, PF>(ws, h->b_s.get(), fp.rtol, true, s);
    case kF1It0:
    case kF1It1:
      return h->fl.pcg_enqueue_iteration<double, PF>(ws, ph - kF1It0, s);
    case kF1End0:
    case kF1End1:
      return h->fl.pcg_enqueue_end<double>(ws, 1 + (ph - kF1End0), s);
    case kP1:
      TMX_HIP(copy_vec<double>(nv, h->fl.x64_.get(), h->xi_s.get(), s));
      TMX_TRY(h->el.update_elasticity(h->xi_s.get(), h->p, h->m, s));
      return h->el.pcg_enqueue_begin<double, PE>(ws, h->Fm.get(), ep.rtol, true, s);
    case kElIt0:
    case kElIt1:
      return h->el.pcg_enqueue_iteration<double, PE>(ws, ph - kElIt0, s);
    case kElEnd0:
    case kElEnd1:
      return h->el.pcg_enqueue_end<double>(ws, 1 + (ph - kElEnd0), s);
    case kP2: {
      const size_t nd = ndofs(h);
      RedSlot gen = ws.region(3);
      TMX_HIP(copy_vec<double>(nd, h->el.x64_.get(), h->u_s.get(), s));
      if (ep.verify) {
        TMX_TRY(h->el.residual<double>(h->u_s.get(), h->Fm.get(), h->el.rr64_.get(), s));
        TMX_HIP(dot(nd, h->el.rr64_.get(), h->el.rr64_.get(), gen, ws.sc() + kScVerify, s));
        TMX_HIP(publish_scalar(ws.sc() + kScVerify, ws.host_dev + kScVerify, s));
      }
      TMX_HIP(dot(nd, load_F(h), h->u_s.get(), gen, ws.sc() + kScJ, s));
      TMX_HIP(publish_scalar(ws.sc() + kScJ, ws.host_dev + kScJ, s));
      TMX_TRY(sensitivity(h, h->xi_s.get(), h->u_s.get(), h->b_s.get()));
      return h->fl.pcg_enqueue_begin<double, PF>(ws, h->b_s.get(), fp.rtol, true, s);
    }
    case kF2It0:
    case kF2It1:
      return h->fl.pcg_enqueue_iteration<double, PF>(ws, ph - kF2It0, s);
    case kF2End0:
    case kF2End1:
      return h->fl.pcg_enqueue_end<double>(ws, 1 + (ph - kF2End0), s);
    case kP3:
      TMX_HIP(copy_vec<double>(nv, h->fl.x64_.get(), w, s));
      return 0;
  }
  return -1;
}

// Records every phase of design h into rec[]; false when a phase fails or would launch
// anything outside the recorder (the design then runs alone).
bool record_design(tmx_batch* b, tmx_problem* h, const double* rho, double* w,
                   const tmx_solve_opts* o, std::array<tmx::Recorder, kPhases>& rec) {
  const SolveParams ep = merge(h->el_prm, o);
  const SolveParams fp = filter_params(h, o);
  int el = -1, fl = -1;
  schemes(h, o, &el, &fl);
  hipStream_t own = h->stream;
  h->stream = b->capture;  // anything not recorded is captured on this stream, never run
  bool ok = true;
  for (int ph = 0; ph < kPhases && ok; ++ph) {
    rec[ph].clear();
    if (hipStreamBeginCapture(b->capture, hipStreamCaptureModeRelaxed) != hipSuccess) {
      ok = false;
      break;
    }
    tmx::active_recorder() = &rec[ph];
    int rc = 0;
    if (el == kSolveMixed && fl == kSolveMixed) rc = record_phase<float, float>(h, ph, rho, w, ep, fp);
    else if (el == kSolveMixed) rc = record_phase<double, float>(h, ph, rho, w, ep, fp);
    else if (fl == kSolveMixed) rc = record_phase<float, double>(h, ph, rho, w, ep, fp);
    else rc = record_phase<double, double>(h, ph, rho, w, ep, fp);
    tmx::active_recorder() = nullptr;
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->capture, &g);
    size_t nodes = 0;
    if (g) {
      if (hipGraphGetNodes(g, nullptr, &nodes) != hipSuccess) nodes = 1;
      (void)hipGraphDestroy(g);
    }
    if (rc != 0 || e != hipSuccess || nodes != 0) {
      ok = false;
      if (std::getenv("TMX_TRACE"))
        std::fprintf(stderr, "[tmx] batch: phase %d of a %dx%d design not recorded (rc %d, %s, "
                     "%zu launches outside the recorder)\n", ph, h->nx, h->ny, rc,
                     hipGetErrorString(e), nodes);
    }
  }
  h->stream = own;
  (void)hipGetLastError();
  return ok;
}

int run_phase(tmx_batch* b, int ph) {
  const unsigned char* base = b->dev_blob.get();
  for (const Merged& m : b->phase[ph]) {
    const tmx::MultiMap mm{
        reinterpret_cast<const unsigned short*>(base + m.owner),
        reinterpret_cast<const int*>(base + m.first),
        m.skip ? reinterpret_cast<const double* const*>(base + m.skip) : nullptr};
    TMX_HIP(m.fn(base + m.rec, mm, m.grid, m.lds, b->stream));
  }
  b->report.launches += (int)b->phase[ph].size();
  return 0;
}

// Every design's PCG start has been queued (with its pinned ring cleared); runs the merged
// iterations (phases it0 / it0 + 1) until every design has stopped or max_iter, as the
// single-design host loop does: one iteration queued ahead of the one waited for, none once
// every running design is predicted to converge within the next.  Then queues the end phase
// of the launched count's parity and waits for the last report.  st receives each design's
// final flag and iteration count.
int run_loop(tmx_batch* b, const std::vector<tmx_problem*>& ds, const std::vector<double>& rtol,
             int it0, int end0, int max_iter, std::vector<LoopState>& st) {
  st.assign(ds.size(), LoopState());
  auto slot = [&](size_t d, int k) {
    return static_cast<volatile double*>(ds[d]->ws.host + Workspace::kScalars + 4 * ((k - 1) & 1));
  };
  auto wait_all = [&](int k) -> int {
    for (size_t d = 0; d < ds.size(); ++d) {
      if (k > st[d].last) continue;
      volatile double* sl = slot(d, k);
      for (long spins = 1; sl[3] < (double)k; ++spins) {
        if ((spins & 1023) == 0) {
          const hipError_t q = hipStreamQuery(b->stream);
          if (q == hipSuccess && sl[3] < (double)k) {
            set_error("internal: merged PCG iteration " + std::to_string(k) + " did not report");
            return -2;
          }
          if (q != hipSuccess && q != hipErrorNotReady) TMX_HIP(q);
        }
      }
    }
    std::atomic_thread_fence(std::memory_order_acquire);
    return 0;
  };
  auto launch = [&](int k) { return run_phase(b, it0 + ((k - 1) & 1)); };
  int launched = 0, done = 0;
  bool all_near = false;
  if (max_iter >= 1) {
    TMX_TRY(launch(1));
    launched = 1;
  }
  while (done < launched) {
    if (launched == done + 1 && launched < max_iter && !all_near) {
      TMX_TRY(launch(launched + 1));
      ++launched;
    }
    TMX_TRY(wait_all(done + 1));
    ++done;
    bool all_stopped = true;
    all_near = true;
    for (size_t d = 0; d < ds.size(); ++d) {
      LoopState& s = st[d];
      if (s.bb < 0.0) {
        s.bb = ds[d]->ws.host[Workspace::kScalars + kRingBB];
        s.tol2 = rtol[d] * rtol[d] * s.bb;
        s.rr_last = s.bb;
      }
      if (s.flag != 0.0) continue;
      volatile double* sl = slot(d, done);
      s.flag = sl[0];
      if (s.flag != 0.0) {  // stopped: its final state, and its last report
        s.rr_end = sl[1];
        s.iters = (int)sl[2];
        s.last = s.iters == done ? done + 1 : done;
        continue;
      }
      all_stopped = false;
      const double rr = sl[1];
      if (s.rr_last > 0.0) s.ratio = rr / s.rr_last;
      s.rr_last = rr;
      s.near = rr * std::min(s.ratio, 1.0) <= 4.0 * s.tol2;
      all_near = all_near && s.near;
    }
    if (all_stopped) break;
    if (launched == done && done < max_iter) {
      TMX_TRY(launch(done + 1));
      ++launched;
    }
  }
  if (launched > 0) TMX_TRY(run_phase(b, end0 + ((launched - 1) & 1)));
  if (launched > done) TMX_TRY(wait_all(launched));
  for (size_t d = 0; d < ds.size(); ++d) {
    LoopState& s = st[d];
    if (s.bb < 0.0) s.bb = ds[d]->ws.host[Workspace::kScalars + kRingBB];
    if (launched == 0) {
      s.flag = 0.0;
    } else if (s.flag == 0.0) {  // (not stopped by the last report read: the final one)
      volatile double* sl = slot(d, launched);
      s.flag = sl[0];
      s.rr_end = sl[1];
      s.iters = (int)sl[2];
    }
  }
  return launched;
}

}  // namespace

namespace {

// Runs designs ids (largest first) as single-design calls on worker threads.
void run_alone(tmx_batch* b, const std::vector<int>& ids, const double* const* rho, double* J,
               double* const* w, const tmx_solve_opts* opts, tmx_solve_stats* stats,
               std::vector<int>& rc, std::vector<std::string>& msg) {
  if (ids.empty()) return;
  std::atomic<size_t> next{0};
  auto work = [&]() {
    for (;;) {
      const size_t k = next.fetch_add(1);
      if (k >= ids.size()) return;
      const int i = ids[k];
      if (hipSetDevice(b->probs[i]->device) != hipSuccess) {
        rc[i] = TMX_ERR_HIP;
        msg[i] = "hipSetDevice failed in a batch worker";
        continue;
      }
      rc[i] = tmx_objective_and_gradient(b->probs[i], rho[i], &J[i], nullptr, nullptr, w[i],
                                         nullptr, opts, stats ? &stats[i] : nullptr);
      if (rc[i] == TMX_ERR_NONFINITE) J[i] = NAN;
      else if (rc[i] != TMX_OK) msg[i] = last_error_cstr();
    }
  };
  const int nw = std::min<int>(b->workers, (int)ids.size());
  std::vector<std::thread> pool;
  for (int t = 1; t < nw; ++t) pool.emplace_back(work);
  work();
  for (auto& t : pool) t.join();
}

bool same_opts(const tmx_solve_opts* a, bool has, const tmx_solve_opts& b) {
  if (!a) return !has;
  return has && a->rtol == b.rtol && a->max_iter == b.max_iter && a->precision == b.precision &&
         a->inner_rtol == b.inner_rtol && a->max_outer == b.max_outer;
}

void reset_rings(const std::vector<tmx_problem*>& ds) {
  for (tmx_problem* h : ds) {
    volatile double* ring = h->ws.host + Workspace::kScalars;
    ring[3] = ring[7] = 0.0;
    ring[kRingBB] = std::numeric_limits<double>::quiet_NaN();
  }
}

// The merged step over designs ids.  Designs that cannot be recorded, or whose solves stop
// without converging, are appended to `alone` (their outputs are then written by a
// single-design call afterwards).
int run_merged(tmx_batch* b, const std::vector<int>& ids, const double* const* rho, double* J,
               double* const* w, const tmx_solve_opts* opts, tmx_solve_stats* stats,
               std::vector<int>& alone) {
  TMX_HIP(hipSetDevice(b->device));
  auto t0 = std::chrono::steady_clock::now();
  // ---- the plan: recorded and merged once, reused while nothing it depends on changes
  bool reuse = b->plan_valid && b->plan_designs.size() + b->plan_rejected.size() == ids.size() &&
               same_opts(opts, b->plan_has_opts, b->plan_opts);
  if (reuse) {
    size_t k = 0;
    for (int i : ids) {
      if (std::find(b->plan_rejected.begin(), b->plan_rejected.end(), i) != b->plan_rejected.end())
        continue;
      if (k >= b->plan_designs.size() || b->plan_designs[k] != b->probs[i] ||
          b->plan_versions[k] != b->probs[i]->version) {
        reuse = false;
        break;
      }
      ++k;
    }
  }
  std::vector<int> used;  // indices into probs of the merged designs, in plan order
  if (reuse) {
    for (int i : ids)
      if (std::find(b->plan_rejected.begin(), b->plan_rejected.end(), i) == b->plan_rejected.end())
        used.push_back(i);
    for (int i : b->plan_rejected) alone.push_back(i);
    b->report.record_ms = 0.0;
  } else {
    b->plan_valid = false;
    b->plan_designs.clear();
    b->plan_versions.clear();
    b->plan_rejected.clear();
    // the batch's input / output vectors of each design
    b->rho_b.resize(b->probs.size());
    b->w_b.resize(b->probs.size());
    for (int i : ids) {
      const size_t nv = nverts(b->probs[i]);
      if (b->rho_b[i].size() != nv) TMX_TRY(b->rho_b[i].alloc(nv));
      if (b->w_b[i].size() != nv) TMX_TRY(b->w_b[i].alloc(nv));
    }
    std::vector<std::array<tmx::Recorder, kPhases>> recs(ids.size());
    for (size_t k = 0; k < ids.size(); ++k) {
      const int i = ids[k];
      if (record_design(b, b->probs[i], b->rho_b[i].get(), b->w_b[i].get(), opts, recs[k])) {
        used.push_back(i);
      } else {
        b->plan_rejected.push_back(i);
        alone.push_back(i);
      }
    }
    if (const char* path = std::getenv("TMX_BATCH_DUMP")) {
      // the recorded launch sequences (kernel, blocks) of every design and phase, as JSON
      if (FILE* f = std::fopen(path, "w")) {
        std::vector<const char*> names;
        auto id = [&](const char* nm) {
          for (size_t k = 0; k < names.size(); ++k)
            if (names[k] == nm) return k;
          names.push_back(nm);
          return names.size() - 1;
        };
        std::fprintf(f, "{\"phases\": [");
        for (int ph = 0; ph < kPhases; ++ph) {
          std::fprintf(f, "%s[", ph ? ", " : "");
          bool firstd = true;
          for (size_t k = 0; k < ids.size(); ++k) {
            if (std::find(used.begin(), used.end(), ids[k]) == used.end()) continue;
            std::fprintf(f, "%s[", firstd ? "" : ", ");
            firstd = false;
            for (size_t r = 0; r < recs[k][ph].recs.size(); ++r)
              std::fprintf(f, "%s[%zu, %u]", r ? ", " : "", id(recs[k][ph].recs[r].name),
                           recs[k][ph].recs[r].nblk);
            std::fprintf(f, "]");
          }
          std::fprintf(f, "]");
        }
        std::fprintf(f, "], \"names\": [");
        for (size_t k = 0; k < names.size(); ++k) {
          std::string nm = names[k];
          for (auto& ch : nm)
            if (ch == '"' || ch == '\\') ch = '\'';
          std::fprintf(f, "%s\"%s\"", k ? ", " : "", nm.c_str());
        }
        std::fprintf(f, "]}\n");
        std::fclose(f);
      }
    }
    b->blob.clear();
    std::vector<const double*> skips;  // each merged design's PCG skip flag
    for (int i : used) skips.push_back(b->probs[i]->ws.sc() + kPcg.skip);
    for (int ph = 0; ph < kPhases; ++ph) {
      b->phase[ph].clear();
      std::vector<const tmx::Recorder*> rs;
      for (size_t k = 0; k < ids.size(); ++k)
        if (std::find(used.begin(), used.end(), ids[k]) != used.end()) rs.push_back(&recs[k][ph]);
      const bool iteration = ph == kF1It0 || ph == kF1It1 || ph == kElIt0 || ph == kElIt1 ||
                             ph == kF2It0 || ph == kF2It1;
      merge_phase(rs, iteration ? &skips : nullptr, b->phase[ph], b->blob);
    }
    for (int ph = 0; ph < kPhases; ++ph)
      if (!tables_valid(b->phase[ph], b->blob, used.size())) {
        set_error("internal: inconsistent merged launch tables");
        return TMX_ERR_INVALID;
      }
    for (int i : used) {
      b->plan_designs.push_back(b->probs[i]);
      b->plan_versions.push_back(b->probs[i]->version);
    }
    b->plan_has_opts = opts != nullptr;
    if (opts) b->plan_opts = *opts;
    // the tables go to the device once per plan (stream-ordered before the first launch)
    const size_t bytes = std::max<size_t>(b->blob.size(), 16);
    if (b->dev_blob.size() < bytes) TMX_TRY(b->dev_blob.alloc(bytes));
    if (b->pinned_size < bytes) {
      if (b->pinned) TMX_HIP(hipHostFree(b->pinned));
      b->pinned = nullptr;
      b->pinned_size = 0;
      TMX_HIP(hipHostMalloc(reinterpret_cast<void**>(&b->pinned), bytes, hipHostMallocDefault));
      b->pinned_size = bytes;
    }
    TMX_HIP(hipStreamSynchronize(b->stream));  // (the previous upload has been read)
    std::memcpy(b->pinned, b->blob.data(), b->blob.size());
    TMX_HIP(hipMemcpyAsync(b->dev_blob.get(), b->pinned, b->blob.size(), hipMemcpyHostToDevice,
                           b->stream));
    b->plan_valid = true;
    b->report.record_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  if (used.empty()) return TMX_OK;
  std::vector<tmx_problem*> ds;
  std::vector<double> rtol_f, rtol_e;
  int max_f = 0, max_e = 0;
  std::vector<SolveParams> eps(used.size()), fps(used.size());
  for (size_t k = 0; k < used.size(); ++k) {
    tmx_problem* h = b->probs[used[k]];
    ds.push_back(h);
    eps[k] = merge(h->el_prm, opts);
    fps[k] = filter_params(h, opts);
    rtol_e.push_back(eps[k].rtol);
    rtol_f.push_back(fps[k].rtol);
    max_e = std::max(max_e, eps[k].max_iter);
    max_f = std::max(max_f, fps[k].max_iter);
  }
  // after the callers' streams (each distinct one once) and anything still queued on the
  // handles' own streams
  std::vector<hipStream_t> users;
  for (tmx_problem* h : ds) {
    if (std::find(users.begin(), users.end(), h->user) == users.end()) {
      users.push_back(h->user);
      TMX_HIP(hipEventRecord(h->ev_in, h->user));
      TMX_HIP(hipStreamWaitEvent(b->stream, h->ev_in, 0));
    }
    if (hipStreamQuery(h->stream) != hipSuccess) {
      TMX_HIP(hipEventRecord(h->ev_out, h->stream));
      TMX_HIP(hipStreamWaitEvent(b->stream, h->ev_out, 0));
    }
  }
  const auto ts = std::chrono::steady_clock::now();
  // copy tasks: the callers' rho in (first half), the designs' w out (second half)
  const size_t nt = 2 * used.size();
  if (b->tasks_cap < nt) {
    if (b->tasks) TMX_HIP(hipHostFree(b->tasks));
    b->tasks = nullptr;
    b->tasks_cap = 0;
    TMX_HIP(hipHostMalloc(reinterpret_cast<void**>(&b->tasks), sizeof(CopyTask) * nt,
                          hipHostMallocDefault));
    b->tasks_cap = nt;
    TMX_TRY(b->dev_tasks.alloc(nt));
  }
  for (size_t k = 0; k < used.size(); ++k) {
    const int i = used[k];
    const size_t nv = nverts(b->probs[i]);
    b->tasks[k] = CopyTask{rho[i], b->rho_b[i].get(), nv};
    b->tasks[used.size() + k] = CopyTask{b->w_b[i].get(), w[i], nv};
  }
  TMX_HIP(hipMemcpyAsync(b->dev_tasks.get(), b->tasks, sizeof(CopyTask) * nt,
                         hipMemcpyHostToDevice, b->stream));
  TMX_HIP(copy_tasks(b->dev_tasks.get(), (int)used.size(), b->stream));
  b->report.launches += 1;
  std::vector<LoopState> s1, s2, s3;
  reset_rings(ds);
  TMX_TRY(run_phase(b, kP0));
  const int n1 = run_loop(b, ds, rtol_f, kF1It0, kF1End0, max_f, s1);
  if (n1 < 0) return n1;
  reset_rings(ds);
  TMX_TRY(run_phase(b, kP1));
  const int n2 = run_loop(b, ds, rtol_e, kElIt0, kElEnd0, max_e, s2);
  if (n2 < 0) return n2;
  reset_rings(ds);
  TMX_TRY(run_phase(b, kP2));
  const int n3 = run_loop(b, ds, rtol_f, kF2It0, kF2End0, max_f, s3);
  if (n3 < 0) return n3;
  TMX_TRY(run_phase(b, kP3));
  TMX_HIP(copy_tasks(b->dev_tasks.get() + used.size(), (int)used.size(), b->stream));
  b->report.launches += 1;
  TMX_HIP(stream_wait(b->stream));
  TMX_HIP(hipEventRecord(b->ev_done, b->stream));
  const double ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - ts).count();
  b->report.iterations[0] = n1;
  b->report.iterations[1] = n2;
  b->report.iterations[2] = n3;
  for (hipStream_t u : users) TMX_HIP(hipStreamWaitEvent(u, b->ev_done, 0));
  for (size_t k = 0; k < ds.size(); ++k) {
    tmx_problem* h = ds[k];
    const int i = used[k];
    const bool ok = s1[k].flag == 1.0 && s2[k].flag == 1.0 && s3[k].flag == 1.0 &&
                    s1[k].iters <= fps[k].max_iter && s2[k].iters <= eps[k].max_iter &&
                    s3[k].iters <= fps[k].max_iter;
    if (!ok) {
      alone.push_back(i);
      continue;
    }
    ++b->report.merged_designs;
    J[i] = h->ws.host[kScJ];
    if (stats) {
      tmx_solve_stats& o = stats[i];
      o.iterations = s2[k].iters;
      o.outer_iterations = 1;
      const double bb = s2[k].bb;
      double rel = 0.0;
      if (bb > 0.0) {
        const double rr = eps[k].verify ? h->ws.host[kScVerify] : s2[k].rr_end;
        rel = std::sqrt(std::max(rr, 0.0) / bb);
      }
      o.rel_residual = rel;
      o.setup_ms = 0.0;
      o.solve_ms = ms;
      o.levels = h->el.levels();
    }
  }
  return TMX_OK;
}

}  // namespace

extern "C" int tmx_batch_create(tmx_problem* const* problems, int count, tmx_batch** out) {
  if (!problems || count <= 0 || !out) return invalid("bad batch arguments");
  if (count > 65535) return invalid("a batch holds at most 65535 problems");
  tmx_batch* b = new (std::nothrow) tmx_batch();
  if (!b) return invalid("out of host memory");
  for (int i = 0; i < count; ++i) {
    if (!problems[i]) {
      delete b;
      return invalid("null problem in batch");
    }
    for (int j = 0; j < i; ++j)
      if (problems[j] == problems[i]) {
        delete b;
        return invalid("a problem appears twice in the batch");
      }
    b->probs.push_back(problems[i]);
    b->order.push_back(i);
  }
  std::stable_sort(b->order.begin(), b->order.end(), [&](int a, int c) {
    return nverts(b->probs[a]) > nverts(b->probs[c]);
  });
  b->device = problems[0]->device;
  int cur = 0;
  if (hipGetDevice(&cur) != hipSuccess || hipSetDevice(b->device) != hipSuccess ||
      hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&b->capture, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_done, hipEventDisableTiming) != hipSuccess) {
    (void)hipSetDevice(cur);
    delete b;
    set_error("HIP stream or event creation failed for the batch");
    return TMX_ERR_HIP;
  }
  (void)hipSetDevice(cur);
  *out = b;
  return TMX_OK;
}

extern "C" void tmx_batch_destroy(tmx_batch* b) { delete b; }

extern "C" int tmx_batch_set_workers(tmx_batch* b, int workers) {
  if (!b) return invalid("null batch");
  if (workers < 1 || workers > 64) return invalid("workers must be in [1, 64]");
  b->workers = workers;
  return TMX_OK;
}

extern "C" int tmx_batch_set_mode(tmx_batch* b, int mode) {
  if (!b) return invalid("null batch");
  if (mode != TMX_BATCH_MERGED && mode != TMX_BATCH_ALONE)
    return invalid("mode must be TMX_BATCH_MERGED or TMX_BATCH_ALONE");
  b->mode = mode;
  return TMX_OK;
}

extern "C" int tmx_batch_last_report(const tmx_batch* b, tmx_batch_report* out) {
  if (!b || !out) return invalid("null argument");
  *out = b->report;
  return TMX_OK;
}

extern "C" int tmx_batch_objective_and_gradient(tmx_batch* b, const double* const* rho,
                                                double* J, double* const* w,
                                                const tmx_solve_opts* opts,
                                                tmx_solve_stats* stats) {
  if (!b || !rho || !J || !w) return invalid("null argument");
  const int n = (int)b->probs.size();
  for (int i = 0; i < n; ++i)
    if (!rho[i] || !w[i]) return invalid("null design pointer in batch");
  b->report = tmx_batch_report{};
  std::vector<int> rc(n, TMX_OK);
  std::vector<std::string> msg(n);
  std::vector<int> merged, alone;
  for (int i : b->order) {
    if (b->mode == TMX_BATCH_MERGED && qualifies(b->probs[i], opts, b->device)) merged.push_back(i);
    else alone.push_back(i);
  }
  // designs that do not qualify run alone meanwhile, on their own streams
  std::thread side([&]() { run_alone(b, alone, rho, J, w, opts, stats, rc, msg); });
  std::vector<int> again;
  int mrc = TMX_OK;
  if (!merged.empty()) {
    int cur = 0;
    const bool restore = hipGetDevice(&cur) == hipSuccess;
    mrc = run_merged(b, merged, rho, J, w, opts, stats, again);
    if (restore) (void)hipSetDevice(cur);
  }
  side.join();
  if (mrc != TMX_OK) {
    const std::string m = last_error_cstr();
    set_error("merged batch step: " + m);
    return mrc;
  }
  // designs rejected by the recorder or not converged in the merged step: alone, from the start
  std::sort(again.begin(), again.end(), [&](int a, int c) {
    return nverts(b->probs[a]) > nverts(b->probs[c]);
  });
  run_alone(b, again, rho, J, w, opts, stats, rc, msg);
  b->report.solo_designs = (int)(alone.size() + again.size());
  for (int i = 0; i < n; ++i)
    if (rc[i] != TMX_OK && rc[i] != TMX_ERR_NONFINITE) {
      set_error("design " + std::to_string(i) + ": " + msg[i]);
      return rc[i];
    }
  return TMX_OK;
}
