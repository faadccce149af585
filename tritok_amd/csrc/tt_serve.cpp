// tt_serve.cpp -- request loop over whole flow cases (no reference
// counterpart: the reference runs one case per process).
//
// Cases arrive as host-side preprocessed inputs (tt_host); a pool of worker
// threads, one per slot, admits them in arrival order as device memory frees
// up (first fit against the memory budget).  A worker holds a batch of up to
// max_batch active cases of one kind (same mesh size and precision) and
// steps them together (tt_case_picard_step_batch: every LSQR iteration of a
// phase is one launch for the whole batch); between steps it retires the
// cases that finished -- copying their fields back and releasing their device
// memory -- and admits queued cases of its kind, so the batch is refilled
// continuously.  A case's results do not depend on the batch it ran in.
// Batches of different workers run concurrently on separate streams.
#include <hip/hip_runtime.h>

#include <algorithm>
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
#include <vector>

#include "tritok_bem.h"
#include "tt_internal.h"

namespace {

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// Device bytes a case will hold (the buffers tt_case_device_bytes counts),
// from the mesh sizes: macro matrices and cell systems per cell, the seven
// single-domain matrices and three inverses, and the reduced-precision copies.
int64_t estimate_bytes(const tt_case_view &v, int prec) {
  const double nb = v.mesh.nbnodes, nn = v.mesh.nnodes, nc = v.mesh.nicell;
  double b = nc * (51.0 * 27 * 8 + 51.0 * 24 + 27.0 * 27 * 3) * 8;  // macro matrices
  b += nc * (2 * 51.0 * 51 + 27.0 * 27) * 8;                          // full cell systems
  b += nc * (4 * 2604.0 + 732) * (prec == TT_PREC_FP64 ? 8 : 12);     // least-squares blocks
  if (v.iBw == 1) {
    const double el = 7 * nb * nb + 3 * nb * nn;
    b += el * 8;
    if (prec == TT_PREC_FP32) b += el * 4;
    if (prec == TT_PREC_FP16 || prec == TT_PREC_BF16) b += el * 2;
  }
  return (int64_t)(b * 1.05) + (64LL << 20);
}

struct Req {
  int64_t id = 0;
  tt_host *h = nullptr;
  int prec = 0;
  int64_t kind = 0;  // cases of one kind share batches: mesh size and precision
  tt_server_info info{};
  int64_t est = 0;
  std::vector<double> vel, vort, qvort, temp, qtemp;
  bool done = false;
};

// a case being iterated by a worker
struct Active {
  Req *r = nullptr;
  tt_case *c = nullptr;
  int its = 0;
  double change = 0.0;
};

}  // namespace

struct tt_server {
  int slots = 1, max_batch = 1, device = 0, anchor = 16, max_iter = 0;
  int64_t budget = 0, in_use = 0, next_id = 1;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Req *> queue;
  std::map<int64_t, std::unique_ptr<Req>> reqs;
  bool stopping = false;
  std::vector<std::thread> workers;
  double t0 = 0;

  // builds the case of r; on failure r is finished with the error
  tt_case *admit(Req *r) {
    tt_case_view v;
    const double *f[5];
    int rc = tt_host_case_view(r->h, &v);
    if (rc == 0) rc = tt_host_fields(r->h, &f[0], &f[1], &f[2], &f[3], &f[4]);
    tt_case *c = nullptr;
    if (rc == 0) rc = tt_case_create(&v, f[0], f[1], f[2], f[3], f[4], r->prec, 0.0, &c);
    if (rc == 0) {
      r->info.t_setup = now_s() - t0;
      r->info.device_bytes = tt_case_device_bytes(c);
      if (r->prec != TT_PREC_FP64) rc = tt_case_set_anchor_period(c, anchor);
    }
    if (rc < 0) {
      r->info.status = -1;
      std::snprintf(r->info.error, sizeof(r->info.error), "%s", tt_last_error());
      if (c) tt_case_free(c);
      return nullptr;
    }
    return c;
  }

  // copies a finished case's fields back (status 0 converged, 1 at the cap,
  // -1 failed), frees it and hands the request back
  void retire(Active &a, int status, const char *err) {
    Req *r = a.r;
    r->info.picard_iterations = a.its;
    r->info.final_change = a.change;
    if (status >= 0 && a.c) {
      tt_case_view v;
      if (tt_host_case_view(r->h, &v) == 0) {
        const size_t nn = v.mesh.nnodes, nq = v.mesh.nq;
        r->vel.resize(nn * 3), r->vort.resize(nn * 3), r->qvort.resize(nq * 3), r->temp.resize(nn),
            r->qtemp.resize(nq);
        if (tt_case_fields(a.c, r->vel.data(), r->vort.data(), r->qvort.data(), r->temp.data(), r->qtemp.data()) < 0)
          status = -1, err = tt_last_error();
      } else {
        status = -1, err = tt_last_error();
      }
    }
    r->info.status = status;
    if (status < 0) std::snprintf(r->info.error, sizeof(r->info.error), "%s", err ? err : "failed");
    if (a.c) tt_case_free(a.c);
    a.c = nullptr;
    finish(r);
  }

  void finish(Req *r) {
    r->info.t_end = now_s() - t0;
    {
      std::lock_guard<std::mutex> lk(mu);
      in_use -= r->est;
      r->done = true;
    }
    cv_work.notify_all();
    cv_done.notify_all();
  }

  // takes the first queued request (arrival order) that fits the memory
  // left and, when kind >= 0, is of that kind; nullptr if none (the caller
  // holds mu)
  Req *take(int64_t kind) {
    auto it = std::find_if(queue.begin(), queue.end(), [&](Req *q) {
      return (kind < 0 || q->kind == kind) && (in_use == 0 || in_use + q->est <= budget);
    });
    if (it == queue.end()) return nullptr;
    Req *r = *it;
    queue.erase(it);
    in_use += r->est;
    return r;
  }

  void worker() {
    (void)hipSetDevice(device);
    std::vector<Active> act;
    std::vector<tt_case *> cs;
    std::vector<double> err;
    std::vector<int32_t> nit, st;
    for (;;) {
      // admission: an idle worker waits for any request; a busy one tops its
      // batch up with queued requests of its kind
      std::vector<Req *> fresh;
      {
        std::unique_lock<std::mutex> lk(mu);
        if (act.empty()) {
          for (;;) {
            if (stopping && queue.empty()) return;
            Req *r = take(-1);
            if (r) {
              fresh.push_back(r);
              break;
            }
            cv_work.wait(lk);
          }
        }
        const int64_t kind = act.empty() ? fresh[0]->kind : act[0].r->kind;
        while ((int)(act.size() + fresh.size()) < max_batch) {
          Req *r = take(kind);
          if (!r) break;
          fresh.push_back(r);
        }
      }
      for (Req *r : fresh) {
        r->info.t_start = now_s() - t0;
        Active a;
        a.r = r;
        a.c = admit(r);
        if (!a.c) {
          finish(r);
          continue;
        }
        act.push_back(a);
      }
      if (act.empty()) continue;
      // one Picard iteration of every active case
      const int k = (int)act.size();
      cs.resize(k), err.assign((size_t)9 * k, 0.0), nit.assign((size_t)12 * k, 0), st.assign(k, 0);
      for (int i = 0; i < k; ++i) cs[i] = act[i].c;
      const int rc = tt_case_picard_step_batch(cs.data(), k, err.data(), nit.data(), st.data());
      const std::string batch_err = rc < 0 ? tt_last_error() : "";
      std::vector<Active> keep;
      for (int i = 0; i < k; ++i) {
        Active &a = act[i];
        if (rc < 0) {
          retire(a, -1, batch_err.c_str());
          continue;
        }
        ++a.its;
        a.r->info.batch_peak = std::max(a.r->info.batch_peak, k);
        for (int q = 0; q < 12; ++q) a.r->info.lsqr_iterations += nit[(size_t)12 * i + q];
        a.change = 0.0;
        for (int q = 0; q < 9; ++q) a.change = std::max(a.change, err[(size_t)9 * i + q]);
        if (st[i] < 0)
          retire(a, -1, tt_last_error());
        else if (st[i] == 0)
          retire(a, 0, nullptr);
        else if (max_iter > 0 && a.its >= max_iter)
          retire(a, 1, nullptr);
        else
          keep.push_back(a);
      }
      act.swap(keep);
    }
  }
};

extern "C" int tt_server_create_batched(int32_t slots, int32_t max_batch, int64_t mem_budget,
                                        int32_t anchor_period, int32_t max_iter, tt_server **out) {
  try {
    if (!out || slots < 1 || slots > 64) throw std::runtime_error("slots must be 1..64");
    if (max_batch < 1 || max_batch > 16) throw std::runtime_error("max_batch must be 1..16");
    auto s = std::make_unique<tt_server>();
    s->max_batch = max_batch;
    if (hipGetDevice(&s->device) != hipSuccess) throw std::runtime_error("no HIP device");
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) throw std::runtime_error("hipMemGetInfo failed");
    s->budget = mem_budget > 0 ? mem_budget : (int64_t)(0.8 * (double)fr);
    s->slots = slots;
    s->anchor = anchor_period > 0 ? anchor_period : 16;
    s->max_iter = max_iter;
    s->t0 = now_s();
    for (int k = 0; k < slots; ++k) s->workers.emplace_back([p = s.get()] { p->worker(); });
    *out = s.release();
    return 0;
  } catch (std::exception &e) {
    tt_set_error("tt_server_create: %s", e.what());
    return -1;
  }
}

extern "C" int tt_server_create(int32_t slots, int64_t mem_budget, int32_t anchor_period, int32_t max_iter,
                                tt_server **out) {
  return tt_server_create_batched(slots, 1, mem_budget, anchor_period, max_iter, out);
}

extern "C" int tt_server_submit(tt_server *s, tt_host *h, int32_t prec, int64_t *id) {
  try {
    if (!s || !h || !id) throw std::runtime_error("NULL argument");
    if (prec < TT_PREC_FP64 || prec > TT_PREC_FP16) throw std::runtime_error("unknown precision");
    tt_case_view v;
    if (tt_host_case_view(h, &v) < 0) throw std::runtime_error(tt_last_error());
    auto r = std::make_unique<Req>();
    r->h = h;
    r->prec = prec;
    r->est = estimate_bytes(v, prec);
    r->kind = (int64_t)v.mesh.nicell * 8 + prec;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      if (s->stopping) throw std::runtime_error("server is shutting down");
      r->id = s->next_id++;
      r->info.t_submit = now_s() - s->t0;
      r->info.estimated_bytes = r->est;
      *id = r->id;
      s->queue.push_back(r.get());
      s->reqs[r->id] = std::move(r);
    }
    s->cv_work.notify_all();
    return 0;
  } catch (std::exception &e) {
    tt_set_error("tt_server_submit: %s", e.what());
    return -1;
  }
}

extern "C" int tt_server_wait(tt_server *s, int64_t id, tt_server_info *info) {
  try {
    if (!s) throw std::runtime_error("NULL argument");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) throw std::runtime_error("unknown request");
    Req *r = it->second.get();
    s->cv_done.wait(lk, [&] { return r->done; });
    if (info) *info = r->info;
    return r->info.status < 0 ? -1 : 0;
  } catch (std::exception &e) {
    tt_set_error("tt_server_wait: %s", e.what());
    return -1;
  }
}

extern "C" int tt_server_fields(tt_server *s, int64_t id, double *vel, double *vort, double *qvort, double *temp,
                                double *qtemp) {
  try {
    if (!s) throw std::runtime_error("NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end() || !it->second->done) throw std::runtime_error("request not finished");
    Req *r = it->second.get();
    if (r->info.status < 0) throw std::runtime_error("request failed");
    auto cp = [](double *d, const std::vector<double> &v) {
      if (d) std::memcpy(d, v.data(), sizeof(double) * v.size());
    };
    cp(vel, r->vel), cp(vort, r->vort), cp(qvort, r->qvort), cp(temp, r->temp), cp(qtemp, r->qtemp);
    return 0;
  } catch (std::exception &e) {
    tt_set_error("tt_server_fields: %s", e.what());
    return -1;
  }
}

extern "C" int tt_server_release(tt_server *s, int64_t id) {
  if (!s) return -1;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end() || !it->second->done) {
    tt_set_error("tt_server_release: request not finished");
    return -1;
  }
  s->reqs.erase(it);
  return 0;
}

extern "C" int tt_server_load(tt_server *s, int32_t *queued, int64_t *bytes_in_use, int64_t *budget) {
  if (!s) return -1;
  std::lock_guard<std::mutex> lk(s->mu);
  if (queued) *queued = (int32_t)s->queue.size();
  if (bytes_in_use) *bytes_in_use = s->in_use;
  if (budget) *budget = s->budget;
  return 0;
}

extern "C" void tt_server_free(tt_server *s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stopping = true;
  }
  s->cv_work.notify_all();
  for (auto &t : s->workers) t.join();
  delete s;
}
