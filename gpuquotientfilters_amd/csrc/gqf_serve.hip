// gqf_serve.hip -- the serving loop: a timed request stream over resident
// filters, admission against a device-memory budget, and continuous batching.
// Host C++ only (no kernels).  The scheduler knows nothing of HIP: each call
// goes through an executor -- the RSQF one at the end of this file, or one
// the caller supplies (gqf_serve_run).
//
// Workers: with coalescing, one worker per filter (its own host thread and,
// in the RSQF executor, its own stream); without, one worker for everything.
// A worker takes the request at the head of its queue and, coalescing, every
// following request of the same kind while the keys fit max_batch, then asks
// for admission: the call starts when
//     resident tables + bytes every worker holds + this call's growth
// fits the budget.  Otherwise the worker frees the holdings of idle workers,
// then waits for running calls to end; a coalesced batch that cannot fit
// even alone is cut back to its first request, and a single request that
// cannot fit alone is rejected (GQF_ERR_NOMEM) without stopping the loop.
// Because a worker's queue holds one filter's requests in arrival order and
// a batch never passes a request of another kind, every call sees the
// filter in the state sequential execution would leave it in.
#include "gqf_runtime.hpp"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <thread>
#include <vector>

namespace gqf {
namespace {

using Clock = std::chrono::steady_clock;

struct Scheduler {
  gqf_serve_config cfg{};
  gqf_serve_executor ex{};
  int nfilters = 0, nworkers = 0;
  size_t resident = 0;
  uint64_t max_batch = 0;
  std::mutex mu;
  std::condition_variable cv;                 // queue changes and call completions
  std::vector<std::deque<gqf_request*>> q;    // per worker, arrival order
  enum State : char { IDLE = 0, RUNNING = 1, TRIMMING = 2 };
  std::vector<char> state;                    // per worker
  std::vector<size_t> held;                   // bytes each worker holds
  int running = 0;   // calls in flight
  int trimming = 0;  // workers being released
  bool closing = false;
  gqf_status fatal = GQF_OK;
  char fatal_msg[512] = {0};
  Clock::time_point t0;
  int next_call = 0;
  uint64_t calls = 0;
  size_t peak = 0;

  double now_us() const { return std::chrono::duration<double, std::micro>(Clock::now() - t0).count(); }
  int worker_of(int filter) const { return cfg.coalesce ? filter : 0; }

  size_t used() const {
    size_t u = resident;
    for (size_t h : held) u += h;
    return u;
  }

  // Admission of worker w's batch (lock held).  Returns false when the batch
  // was cut back or rejected and must be formed again.
  bool admit(std::unique_lock<std::mutex>& lk, int w, std::vector<gqf_request*>& b) {
    uint64_t n = 0;
    for (gqf_request* r : b) n += (uint64_t)r->n;
    const bool coalesced = b.size() > 1;
    for (;;) {
      if (state[(size_t)w] == TRIMMING) {  // another worker is releasing our holdings
        cv.wait(lk);
        continue;
      }
      const size_t grow = ex.grow_bytes(ex.user, w, b[0]->filter, b[0]->op, n, coalesced ? 1 : 0);
      const size_t u = used();
      if (u + grow <= cfg.hbm_budget) {
        held[(size_t)w] += grow;
        peak = std::max(peak, u + grow);
        state[(size_t)w] = RUNNING;
        running++;
        return true;
      }
      // release what an idle worker holds
      int v = -1;
      for (int i = 0; i < nworkers; ++i)
        if (i != w && state[(size_t)i] == IDLE && held[(size_t)i] > 0) {
          v = i;
          break;
        }
      if (v >= 0) {
        state[(size_t)v] = TRIMMING;
        trimming++;
        lk.unlock();
        ex.trim(ex.user, v);
        const size_t h = ex.held_bytes(ex.user, v);
        lk.lock();
        held[(size_t)v] = h;
        state[(size_t)v] = IDLE;
        trimming--;
        cv.notify_all();
        continue;
      }
      if (running == 0 && trimming == 0) {
        // nothing will free up: cut the batch back, or reject its only request
        if (b.size() > 1) {
          for (size_t i = b.size(); i-- > 1;) q[(size_t)w].push_front(b[i]);
          b.resize(1);
          return false;
        }
        gqf_request* r = b[0];
        r->status = GQF_ERR_NOMEM;
        r->call = -1;
        r->start_us = r->done_us = now_us();
        b.clear();
        cv.notify_all();
        return false;
      }
      cv.wait(lk);
    }
  }

  void worker_loop(int w) {
    std::vector<gqf_request*> b;
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
      if (b.empty()) {
        cv.wait(lk, [&] { return closing || !q[(size_t)w].empty(); });
        if (q[(size_t)w].empty()) return;  // closing and drained
        std::deque<gqf_request*>& qq = q[(size_t)w];
        b.push_back(qq.front());
        qq.pop_front();
        if (cfg.coalesce) {
          uint64_t n = (uint64_t)b[0]->n;
          while (!qq.empty() && qq.front()->op == b[0]->op && qq.front()->filter == b[0]->filter &&
                 n + (uint64_t)qq.front()->n <= max_batch) {
            n += (uint64_t)qq.front()->n;
            b.push_back(qq.front());
            qq.pop_front();
          }
        }
      }
      if (!admit(lk, w, b)) continue;  // cut back to one request, or rejected
      const int call = next_call++;
      lk.unlock();
      const double ts = now_us();
      const int st = ex.run(ex.user, w, b.data(), (int)b.size());
      const double td = now_us();
      const size_t h = ex.held_bytes(ex.user, w);
      lk.lock();
      for (gqf_request* r : b) {
        r->start_us = ts;
        r->done_us = td;
        r->status = st;
        r->call = call;
      }
      if (st != GQF_OK && st != GQF_ERR_CAPACITY && st != GQF_ERR_NOMEM && fatal == GQF_OK) {
        fatal = (gqf_status)st;
        snprintf(fatal_msg, sizeof(fatal_msg), "%s", gqf_last_error());
      }
      held[(size_t)w] = h;
      state[(size_t)w] = IDLE;
      running--;
      calls++;
      b.clear();
      cv.notify_all();
    }
  }

  gqf_status run(gqf_request* reqs, int nreqs, gqf_serve_stats* stats) {
    nworkers = cfg.coalesce ? nfilters : 1;
    max_batch = cfg.max_batch > 0 ? (uint64_t)cfg.max_batch : (1ull << 24);
    q.assign((size_t)nworkers, {});
    state.assign((size_t)nworkers, IDLE);
    held.assign((size_t)nworkers, 0);
    for (int i = 0; i < nworkers; ++i) held[(size_t)i] = ex.held_bytes(ex.user, i);
    for (int i = 0; i < nreqs; ++i) {
      reqs[i].status = GQF_OK;
      reqs[i].call = -1;
      reqs[i].start_us = reqs[i].done_us = 0.0;
    }
    t0 = Clock::now();
    std::vector<std::thread> th;
    th.reserve((size_t)nworkers);
    for (int i = 0; i < nworkers; ++i) th.emplace_back([this, i] { worker_loop(i); });
    for (int i = 0; i < nreqs; ++i) {
      gqf_request& r = reqs[i];
      std::this_thread::sleep_until(t0 + std::chrono::nanoseconds((int64_t)(r.arrive_us * 1000.0)));
      std::lock_guard<std::mutex> lk(mu);
      q[(size_t)worker_of(r.filter)].push_back(&r);
      cv.notify_all();
    }
    {
      std::lock_guard<std::mutex> lk(mu);
      closing = true;
      cv.notify_all();
    }
    for (std::thread& t : th) t.join();
    if (stats) {
      gqf_serve_stats s{};
      std::vector<double> lat;
      lat.reserve((size_t)nreqs);
      double first = nreqs ? reqs[0].arrive_us : 0.0, last = first;
      for (int i = 0; i < nreqs; ++i) {
        const gqf_request& r = reqs[i];
        if (r.status == GQF_ERR_NOMEM && r.call < 0) {
          s.rejected++;
          continue;
        }
        if (r.status != GQF_OK) continue;
        s.requests++;
        s.keys += (uint64_t)r.n;
        lat.push_back(r.done_us - r.arrive_us);
        last = std::max(last, r.done_us);
      }
      s.calls = calls;
      s.span_us = last - first;
      if (!lat.empty()) {
        std::sort(lat.begin(), lat.end());
        auto pct = [&](double p) {
          size_t k = (size_t)(p * (double)lat.size() + 0.999999);
          k = k < 1 ? 1 : (k > lat.size() ? lat.size() : k);
          return lat[k - 1];
        };
        s.p50_us = pct(0.50);
        s.p99_us = pct(0.99);
        double sum = 0;
        for (double x : lat) sum += x;
        s.mean_us = sum / (double)lat.size();
      }
      s.mops = s.span_us > 0 ? (double)s.keys / s.span_us : 0.0;
      s.peak_bytes = peak;
      *stats = s;
    }
    if (fatal != GQF_OK) {
      set_error("%s", fatal_msg);
      return fatal;
    }
    return GQF_OK;
  }
};

gqf_status check_requests(int nfilters, const gqf_request* reqs, int nreqs, const gqf_serve_config* cfg,
                          bool need_out) {
  if (nfilters <= 0 || nreqs < 0 || (nreqs > 0 && !reqs) || !cfg) {
    set_error("invalid serve arguments (nfilters=%d, nreqs=%d)", nfilters, nreqs);
    return GQF_ERR_INVALID;
  }
  for (int i = 0; i < nreqs; ++i) {
    const gqf_request& r = reqs[i];
    const bool bad = r.filter < 0 || r.filter >= nfilters || r.op < GQF_OP_LOOKUP || r.op > GQF_OP_DELETE ||
                     r.n < 0 || (uint64_t)r.n >= GQF_MAX_ITEMS || (r.n > 0 && !r.d_keys) ||
                     (need_out && r.op == GQF_OP_LOOKUP && r.n > 0 && !r.d_out) || !(r.arrive_us >= 0.0) ||
                     (i > 0 && r.arrive_us < reqs[i - 1].arrive_us);
    if (bad) {
      set_error("request %d is invalid (filter, op, key count, buffers or arrival order)", i);
      return GQF_ERR_INVALID;
    }
  }
  return GQF_OK;
}

// ---------------------------------------------------------------- RSQF executor
// Worker w owns a stream (its calls' workspace lives in the stream's
// context) and, coalescing, gather buffers for keys and lookup results.
struct RsqfExec {
  const gqf_rsqf* f = nullptr;
  int sorted = 0;
  struct W {
    gqf_stream st = nullptr;
    uint32_t* keys = nullptr;
    int32_t* out = nullptr;
    size_t cap = 0;  // keys the gather buffers hold
  };
  std::vector<W> w;

  static size_t stage_bytes(size_t cap) { return cap * 8; }

  static size_t grow(void* u, int wi, int filter, int op, uint64_t n, int coalesced) {
    RsqfExec& e = *static_cast<RsqfExec*>(u);
    W& x = e.w[(size_t)wi];
    const size_t need = gqf_rsqf_workspace_bytes(e.f[filter], op, (int)n, e.sorted);
    size_t have = 0;
    (void)gqf_stream_workspace_bytes(x.st, &have);
    size_t g = need > have ? ws_alloc_bytes(need) - have : 0;
    if (coalesced && n > x.cap) g += stage_bytes((size_t)n) - stage_bytes(x.cap);
    return g;
  }
  static size_t held(void* u, int wi) {
    RsqfExec& e = *static_cast<RsqfExec*>(u);
    W& x = e.w[(size_t)wi];
    size_t have = 0;
    (void)gqf_stream_workspace_bytes(x.st, &have);
    return have + stage_bytes(x.cap);
  }
  static void trim(void* u, int wi) {
    RsqfExec& e = *static_cast<RsqfExec*>(u);
    W& x = e.w[(size_t)wi];
    (void)gqf_stream_trim(x.st);
    e.free_stage(x);
  }
  void free_stage(W& x) {
    if (x.keys) (void)hipFree(x.keys);
    if (x.out) (void)hipFree(x.out);
    x.keys = nullptr;
    x.out = nullptr;
    x.cap = 0;
  }
  static int run(void* u, int wi, gqf_request** b, int nb) {
    RsqfExec& e = *static_cast<RsqfExec*>(u);
    return (int)e.call(e.w[(size_t)wi], b, nb);
  }

  gqf_status call(W& x, gqf_request** b, int nb) {
    const gqf_rsqf& fl = f[b[0]->filter];
    const int op = b[0]->op;
    hipStream_t hs = reinterpret_cast<hipStream_t>(x.st);
    float ms = 0.f;
    if (nb == 1) {
      const gqf_request& r = *b[0];
      if (op == GQF_OP_LOOKUP) return gqf_rsqf_lookup(fl, sorted, r.n, r.d_keys, r.d_out, x.st, &ms);
      if (op == GQF_OP_INSERT) return gqf_rsqf_insert(fl, r.n, r.d_keys, nullptr, x.st, &ms);
      return gqf_rsqf_delete(fl, r.n, r.d_keys, x.st, &ms);
    }
    size_t n = 0;
    for (int i = 0; i < nb; ++i) n += (size_t)b[i]->n;
    if (n > x.cap) {
      GQF_HIP(hipStreamSynchronize(hs));
      free_stage(x);
      GQF_HIP(hipMalloc((void**)&x.keys, n * 4));
      GQF_HIP(hipMalloc((void**)&x.out, n * 4));
      x.cap = n;
    }
    size_t at = 0;
    for (int i = 0; i < nb; ++i) {
      const size_t k = (size_t)b[i]->n;
      if (k) GQF_HIP(hipMemcpyAsync(x.keys + at, b[i]->d_keys, k * 4, hipMemcpyDeviceToDevice, hs));
      at += k;
    }
    if (op == GQF_OP_INSERT) return gqf_rsqf_insert(fl, (int)n, x.keys, nullptr, x.st, &ms);
    if (op == GQF_OP_DELETE) return gqf_rsqf_delete(fl, (int)n, x.keys, x.st, &ms);
    GQF_TRY(gqf_rsqf_lookup(fl, sorted, (int)n, x.keys, x.out, x.st, &ms));
    at = 0;
    for (int i = 0; i < nb; ++i) {
      const size_t k = (size_t)b[i]->n;
      if (k) GQF_HIP(hipMemcpyAsync(b[i]->d_out, x.out + at, k * 4, hipMemcpyDeviceToDevice, hs));
      at += k;
    }
    GQF_HIP(hipStreamSynchronize(hs));
    return GQF_OK;
  }
};

}  // namespace
}  // namespace gqf

using namespace gqf;

extern "C" {

gqf_status gqf_serve_run(int nfilters, size_t resident_bytes, gqf_request* reqs, int nreqs,
                         const gqf_serve_config* cfg, const gqf_serve_executor* ex, gqf_serve_stats* stats) {
  GQF_TRY(check_requests(nfilters, reqs, nreqs, cfg, false));
  if (!ex || !ex->grow_bytes || !ex->held_bytes || !ex->trim || !ex->run) {
    set_error("executor callbacks missing");
    return GQF_ERR_INVALID;
  }
  Scheduler s;
  s.cfg = *cfg;
  s.ex = *ex;
  s.nfilters = nfilters;
  s.resident = resident_bytes;
  return s.run(reqs, nreqs, stats);
}

gqf_status gqf_rsqf_serve(gqf_rsqf* filters, int nfilters, gqf_request* reqs, int nreqs,
                          const gqf_serve_config* cfg, gqf_serve_stats* stats) {
  GQF_TRY(check_requests(nfilters, reqs, nreqs, cfg, true));
  if (!filters) return GQF_ERR_INVALID;
  size_t resident = 0;
  for (int i = 0; i < nfilters; ++i) {
    if (!filters[i].blocks || gqf_rsqf_num_blocks(filters[i].qbits) != filters[i].numBlocks) {
      set_error("filter %d is not an initialised RSQF", i);
      return GQF_ERR_INVALID;
    }
    resident += gqf_rsqf_table_bytes(filters[i].qbits, filters[i].rbits);
  }
  RsqfExec e;
  e.f = filters;
  e.sorted = cfg->sorted_lookups;
  e.w.resize(cfg->coalesce ? (size_t)nfilters : 1);
  gqf_status st = GQF_OK;
  for (RsqfExec::W& x : e.w)
    if (st == GQF_OK) st = gqf_stream_create(&x.st);
  if (st == GQF_OK) {
    gqf_serve_executor ex{&e, &RsqfExec::grow, &RsqfExec::held, &RsqfExec::trim, &RsqfExec::run};
    Scheduler s;
    s.cfg = *cfg;
    s.ex = ex;
    s.nfilters = nfilters;
    s.resident = resident;
    st = s.run(reqs, nreqs, stats);
  }
  for (RsqfExec::W& x : e.w) {
    e.free_stage(x);
    if (x.st) {
      const gqf_status d = gqf_stream_destroy(x.st);
      if (st == GQF_OK) st = d;
    }
  }
  return st;
}

}  // extern "C"
