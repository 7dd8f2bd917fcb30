// cgfd_serve.cpp -- continuous-batching shot server over el_iso sessions
// (see cgfd_serve.h).  One worker thread owns every session; client threads
// only touch the request table under the server mutex.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/cgfd_el_iso.h"
#include "../../include/cgfd_serve.h"
#include "cgfd_internal.h"

namespace {

using Clock = std::chrono::steady_clock;

int sfail(int code, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  cgfd_internal_set_error(buf);
  return code;
}

std::vector<char> copy_bytes(const void *p, size_t n) {
  std::vector<char> v(n);
  if (p && n) std::memcpy(v.data(), p, n);
  return v;
}

struct Model {
  cgfd_model_desc d{};
  size_t es = 4;  // bytes per array element
  std::vector<char> metric, lam, mu, slw, mvx, mvy, A[6], B[6], D[6];
  std::vector<int> scale;
};

enum State { QUEUED, RUNNING, DONE, REJECTED };

struct Request {
  long long ticket = 0;
  int model = 0, nt = 0;
  // source tables (copied)
  int nsrc = 0, force_on = 0, moment_on = 0, max_nt = 0, max_ext = 0, sdtype = CGFD_DT_F32;
  std::vector<int> itb, ite, en;
  std::vector<long long> ei;
  std::vector<char> ec, F, M;
  // receivers
  int nrecv = 0, rdtype = CGFD_DT_F32;
  std::vector<long long> ridx;
  std::vector<char> rfrac;
  // outcome
  State state = QUEUED;
  std::string error;
  Clock::time_point t_submit, t_admit, t_done;
  std::vector<double> seismo;
};

struct Slot {
  bool busy = false;
  long long ticket = 0;
  int start = 0, nt = 0;
};

struct Group {
  int model = -1;
  cgfd_session *s = nullptr;
  int cap = 0;
  std::vector<Slot> slot;
  int clock = 0;
  bool inflight = false;
  long long bytes = 0;
  Clock::time_point last_used;
  int active() const {
    int n = 0;
    for (const Slot &q : slot) n += q.busy;
    return n;
  }
};

size_t npts(const cgfd_model_desc &d) { return (size_t)(d.ni + 6) * (d.nj + 6) * (d.nk + 6); }

}  // namespace

struct cgfd_server {
  cgfd_server_desc d{};
  long long budget = 0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  bool stop = false;
  std::vector<std::unique_ptr<Model>> models;
  std::map<long long, std::unique_ptr<Request>> reqs;
  std::deque<long long> pending;
  long long next_ticket = 1;
  cgfd_server_stats st{};
  std::vector<std::unique_ptr<Group>> groups;  // worker-owned
  std::thread worker;
  int dev = 0;

  void run();
  bool admit_one(Request &r);
  Group *make_group(int model, std::string &why);
  void retire(Group &g);
  void fail_group(Group &g, const std::string &why);
};

namespace {

long long group_bytes(const Model &m, int cap) {
  cgfd_desc dd{m.d.ni, m.d.nj, m.d.nk, m.d.storage, cap, 0};
  long long b = cgfd_footprint_bytes(&dd, m.d.max_nrecv, m.d.max_nt);
  // the footprint counts PML memory for faces of up to 11 points; add deeper ones
  const long long es = m.d.storage == CGFD_STORE_F64 ? 8 : m.d.storage == CGFD_STORE_F32 ? 4 : 2;
  const long long area[3] = {(long long)m.d.nj * m.d.nk, (long long)m.d.ni * m.d.nk, (long long)m.d.ni * m.d.nj};
  for (int f = 0; f < 6; f++)
    if (m.d.pml_nlay[f] + 1 > 11) b += (long long)cap * 5 * 9 * es * area[f / 2] * (m.d.pml_nlay[f] + 1 - 11);
  return b;
}

int upload_model(cgfd_session *s, const Model &m) {
  int r;
  if ((r = cgfd_upload_metric(s, m.d.dtype, m.metric.data()))) return r;
  if ((r = cgfd_upload_medium(s, m.d.dtype, m.lam.data(), m.mu.data(), m.slw.data()))) return r;
  if ((r = cgfd_upload_free_surface(s, m.d.free_top, m.d.dtype, m.d.free_top ? m.mvx.data() : nullptr,
                                    m.d.free_top ? m.mvy.data() : nullptr)))
    return r;
  for (int f = 0; f < 6; f++) {
    if (m.d.pml_nlay[f] <= 0) continue;
    if ((r = cgfd_upload_pml(s, f, m.d.pml_nlay[f], m.d.dtype, m.A[f].data(), m.B[f].data(), m.D[f].data())))
      return r;
  }
  if (!m.scale.empty() && (m.d.storage == CGFD_STORE_F16 || m.d.storage == CGFD_STORE_BF16))
    if ((r = cgfd_set_storage_scale(s, m.scale.data()))) return r;
  return CGFD_OK;
}

}  // namespace

// a resident group for the model, created within the budget (evicting idle
// groups of other models, least recently used first); null if it cannot fit
Group *cgfd_server::make_group(int model, std::string &why) {
  const Model &m = *models[model];
  auto used = [&]() {
    long long u = 0;
    for (auto &g : groups) u += g->bytes;
    return u;
  };
  if (group_bytes(m, 1) > budget) {
    why = "the model does not fit the HBM budget even with a single shot";
    return nullptr;
  }
  while (group_bytes(m, 1) > budget - used()) {
    Group *victim = nullptr;
    for (auto &g : groups)
      if (!g->inflight && g->active() == 0 && (!victim || g->last_used < victim->last_used)) victim = g.get();
    if (!victim) return nullptr;  // wait for memory to free up
    cgfd_destroy(victim->s);
    groups.erase(std::find_if(groups.begin(), groups.end(), [&](auto &p) { return p.get() == victim; }));
  }
  int cap = std::max(1, d.max_batch);
  const double share = d.group_share > 0 ? d.group_share : 0.25;
  const long long room = std::min(budget - used(), (long long)(share * (double)budget));
  while (cap > 1 && group_bytes(m, cap) > room) cap--;
  cgfd_desc dd{m.d.ni, m.d.nj, m.d.nk, m.d.storage, cap, 0};
  cgfd_session *s = nullptr;
  if (cgfd_create(&dd, &s) != CGFD_OK) {
    why = cgfd_last_error();
    return nullptr;
  }
  if (upload_model(s, m) != CGFD_OK) {
    why = cgfd_last_error();
    cgfd_destroy(s);
    return nullptr;
  }
  for (int q = 0; q < cap; q++) cgfd_set_shot(s, q, 0, 0);
  if (cgfd_reserve_receivers(s, m.d.max_nrecv, m.d.max_nt) != CGFD_OK) {
    why = cgfd_last_error();
    cgfd_destroy(s);
    return nullptr;
  }
  auto g = std::make_unique<Group>();
  g->model = model;
  g->s = s;
  g->cap = cap;
  g->slot.resize(cap);
  g->bytes = group_bytes(m, cap);
  g->last_used = Clock::now();
  groups.push_back(std::move(g));
  return groups.back().get();
}

// put one queued request into a free slot of its model's group
bool cgfd_server::admit_one(Request &r) {
  {
    // the request may have been rejected by cgfd_server_destroy since the
    // worker copied the queue; a stopping server admits nothing
    std::lock_guard<std::mutex> lk(mu);
    if (stop || r.state != QUEUED) return false;
  }
  Group *g = nullptr;
  for (auto &p : groups)
    if (p->model == r.model && !p->inflight && p->active() < p->cap) { g = p.get(); break; }
  if (!g) {
    bool exists = false;
    for (auto &p : groups) exists |= p->model == r.model;
    if (exists) return false;  // its group is busy or full: next boundary
    std::string why;
    g = make_group(r.model, why);
    if (!g) {
      if (!why.empty()) {
        std::lock_guard<std::mutex> lk(mu);
        r.state = REJECTED;
        r.error = why;
        r.t_done = Clock::now();
        st.rejected++;
        cv_done.notify_all();
        return true;
      }
      return false;
    }
  }
  int q = 0;
  while (g->slot[q].busy) q++;
  cgfd_session *s = g->s;
  cgfd_source_desc sd{r.nsrc, r.force_on, r.moment_on, r.max_nt, r.max_ext, r.sdtype,
                      r.itb.data(), r.ite.data(), r.en.data(), r.ei.data(), r.ec.data(),
                      r.F.empty() ? nullptr : r.F.data(), r.M.empty() ? nullptr : r.M.data()};
  int rc = cgfd_reset_shot(s, q);
  if (!rc) rc = cgfd_upload_source(s, q, r.nsrc > 0 ? &sd : nullptr);
  if (!rc) rc = cgfd_upload_receivers(s, q, r.nrecv, r.ridx.data(), r.rdtype, r.rfrac.data(), r.nt);
  if (!rc) rc = cgfd_set_shot(s, q, 1, g->clock);
  std::lock_guard<std::mutex> lk(mu);
  if (rc || r.state != QUEUED) {
    if (!rc) rc = CGFD_E_STATE;  // rejected meanwhile: free the slot again
    cgfd_set_shot(s, q, 0, 0);
    if (r.state == QUEUED) {
      r.state = REJECTED;
      r.error = cgfd_last_error();
      r.t_done = Clock::now();
      st.rejected++;
    }
    cv_done.notify_all();
    return true;
  }
  g->slot[q] = Slot{true, r.ticket, g->clock, r.nt};
  r.state = RUNNING;
  r.t_admit = Clock::now();
  st.admitted++;
  return true;
}

// end every shot of the group with an error and free its slots
void cgfd_server::fail_group(Group &g, const std::string &why) {
  std::lock_guard<std::mutex> lk(mu);
  for (int q = 0; q < g.cap; q++) {
    Slot &sl = g.slot[q];
    if (!sl.busy) continue;
    auto it = reqs.find(sl.ticket);
    if (it != reqs.end()) {
      it->second->state = REJECTED;
      it->second->error = why;
      it->second->t_done = Clock::now();
      st.rejected++;
    }
    sl = Slot{};
  }
  g.inflight = false;
  cv_done.notify_all();
}

// collect every shot of the group that has run its steps
void cgfd_server::retire(Group &g) {
  for (int q = 0; q < g.cap; q++) {
    Slot &sl = g.slot[q];
    if (!sl.busy || g.clock - sl.start < sl.nt) continue;
    Request *r = nullptr;
    {
      std::lock_guard<std::mutex> lk(mu);
      auto it = reqs.find(sl.ticket);
      if (it != reqs.end()) r = it->second.get();
    }
    std::vector<double> seis;
    int rc = CGFD_OK;
    if (r) {
      seis.resize((size_t)r->nrecv * 9 * r->nt);
      if (!seis.empty()) rc = cgfd_read_seismo(g.s, q, CGFD_DT_F64, seis.data());
    }
    cgfd_set_shot(g.s, q, 0, 0);
    sl = Slot{};
    std::lock_guard<std::mutex> lk(mu);
    if (r) {
      r->t_done = Clock::now();
      if (rc) {
        r->state = REJECTED;
        r->error = cgfd_last_error();
        st.rejected++;
      } else {
        r->seismo = std::move(seis);
        r->state = DONE;
        st.completed++;
      }
    }
    cv_done.notify_all();
  }
}

void cgfd_server::run() {
  (void)hipSetDevice(dev);
  for (;;) {
    // retire finished chunks
    bool any_inflight = false;
    for (size_t gi = 0; gi < groups.size();) {
      Group &g = *groups[gi];
      int pr = g.inflight ? cgfd_poll(g.s) : 1;
      if (g.inflight && pr == 0) {
        g.inflight = false;
        retire(g);
      } else if (g.inflight && pr < 0) {
        // the group's stream failed: its shots end with the error and the
        // group is dropped, so the worker never waits on it again
        fail_group(g, cgfd_last_error());
        cgfd_destroy(g.s);
        groups.erase(groups.begin() + (long)gi);
        continue;
      }
      any_inflight |= g.inflight;
      gi++;
    }
    // admit queued requests, oldest first, wherever a slot and memory allow
    std::vector<long long> todo;
    {
      std::unique_lock<std::mutex> lk(mu);
      if (stop && !any_inflight) break;
      if (!stop && pending.empty() && !any_inflight) {
        bool idle = true;
        for (auto &g : groups) idle &= g->active() == 0;
        if (idle) cv_work.wait_for(lk, std::chrono::milliseconds(50));
      }
      if (!stop) todo.assign(pending.begin(), pending.end());
    }
    std::vector<long long> admitted;
    for (long long t : todo) {
      Request *r;
      {
        std::lock_guard<std::mutex> lk(mu);
        r = reqs[t].get();
      }
      if (admit_one(*r)) admitted.push_back(t);
    }
    {
      std::lock_guard<std::mutex> lk(mu);
      for (long long t : admitted) {
        auto it = std::find(pending.begin(), pending.end(), t);
        if (it != pending.end()) pending.erase(it);
      }
    }
    // Launch the next chunk of idle groups with shots on board: up to the
    // first shot that finishes, at most chunk_steps while the group has free
    // slots (new arrivals wait at most that long) and 8 x chunk_steps when it
    // is full.  A group whose shots fill the device alone (kBigCells active
    // points or more) runs by itself: overlapping it with another sweep only
    // thins the L2 reuse of both.  Small groups overlap each other.  Groups
    // are visited least recently run first, and a big group that is waiting
    // for the device holds back the groups behind it, so none starves.
    constexpr double kBigCells = 16e6;
    auto big = [&](const Group &g) { return (double)npts(models[g.model]->d) * g.active() >= kBigCells; };
    bool any_run = false, big_run = false;
    for (auto &g : groups)
      if (g->inflight) {
        any_run = true;
        big_run |= big(*g);
      }
    std::vector<Group *> order;
    for (auto &g : groups)
      if (!g->inflight && g->active() > 0) order.push_back(g.get());
    std::stable_sort(order.begin(), order.end(),
                     [](const Group *a, const Group *b) { return a->last_used < b->last_used; });
    for (Group *g : order) {
      if (big_run) break;
      if (big(*g) && any_run) break;  // waits for the device; later groups wait behind it
      const Model &m = *models[g->model];
      int left = 1 << 30;
      for (const Slot &sl : g->slot)
        if (sl.busy) left = std::min(left, sl.start + sl.nt - g->clock);
      int len = (std::max(left, 1) + 7) / 8 * 8;
      len = std::min(len, g->active() < g->cap ? d.chunk_steps : 8 * d.chunk_steps);
      if (cgfd_run_async(g->s, g->clock, len, m.d.dt) != CGFD_OK) {
        fail_group(*g, cgfd_last_error());
        continue;
      }
      g->clock += len;
      g->inflight = true;
      g->last_used = Clock::now();
      any_run = true;
      big_run |= big(*g);
    }
    {
      std::lock_guard<std::mutex> lk(mu);
      long long used = 0;
      int act = 0;
      for (auto &g : groups) { used += g->bytes; act += g->active(); }
      st.groups = (int)groups.size();
      st.active_shots = act;
      st.queued = (int)pending.size();
      st.hbm_in_use = used;
    }
    std::this_thread::sleep_for(std::chrono::microseconds(50));
  }
  for (auto &g : groups) {
    cgfd_wait(g->s);
    cgfd_destroy(g->s);
  }
  groups.clear();
}

extern "C" {

int cgfd_server_create(const cgfd_server_desc *desc, cgfd_server **out) {
  if (!desc || !out) return sfail(CGFD_E_ARG, "null argument");
  *out = nullptr;
  if (desc->max_batch < 1 || desc->max_batch > 64) return sfail(CGFD_E_ARG, "max_batch %d not in 1..64", desc->max_batch);
  if (!(desc->group_share >= 0 && desc->group_share <= 1))
    return sfail(CGFD_E_ARG, "group_share %g not in [0, 1]", desc->group_share);
  if (desc->chunk_steps < 8 || desc->chunk_steps % 8)
    return sfail(CGFD_E_ARG, "chunk_steps %d must be a positive multiple of 8", desc->chunk_steps);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return sfail(CGFD_E_NODEV, "no HIP device available");
  auto *srv = new cgfd_server();
  srv->d = *desc;
  (void)hipGetDevice(&srv->dev);
  srv->budget = desc->hbm_budget;
  if (srv->budget <= 0) {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
      delete srv;
      return sfail(CGFD_E_HIP, "hipMemGetInfo failed");
    }
    srv->budget = (long long)(0.9 * (double)fr);
  }
  srv->st.hbm_budget = srv->budget;
  srv->worker = std::thread([srv]() { srv->run(); });
  *out = srv;
  return CGFD_OK;
}

void cgfd_server_destroy(cgfd_server *srv) {
  if (!srv) return;
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->stop = true;
    // queued requests will not run
    for (long long t : srv->pending) {
      Request &r = *srv->reqs[t];
      if (r.state != QUEUED) continue;
      r.state = REJECTED;
      r.error = "server shut down";
      r.t_done = Clock::now();
      srv->st.rejected++;
    }
    srv->pending.clear();
  }
  srv->cv_work.notify_all();
  if (srv->worker.joinable()) srv->worker.join();
  srv->cv_done.notify_all();
  delete srv;
}

int cgfd_server_add_model(cgfd_server *srv, const cgfd_model_desc *md, int *model_id) {
  if (!srv || !md || !model_id) return sfail(CGFD_E_ARG, "null argument");
  if (md->ni < 1 || md->nj < 1 || md->nk < 1) return sfail(CGFD_E_ARG, "bad grid");
  if (md->storage < 0 || md->storage > 3) return sfail(CGFD_E_ARG, "bad storage");
  if (!md->metric || !md->lam || !md->mu || !md->slw) return sfail(CGFD_E_ARG, "missing model arrays");
  if (md->free_top && (!md->matVx2Vz || !md->matVy2Vz)) return sfail(CGFD_E_ARG, "free surface without matrices");
  if (!(md->dt > 0)) return sfail(CGFD_E_ARG, "dt must be positive");
  if (md->max_nt < 1 || md->max_nrecv < 0) return sfail(CGFD_E_ARG, "bad request bounds");
  auto m = std::make_unique<Model>();
  m->d = *md;
  m->es = md->dtype == CGFD_DT_F64 ? 8 : 4;
  const size_t n = npts(*md);
  const size_t plane = (size_t)(md->ni + 6) * (md->nj + 6) * 9;
  m->metric = copy_bytes(md->metric, 10 * n * m->es);
  m->lam = copy_bytes(md->lam, n * m->es);
  m->mu = copy_bytes(md->mu, n * m->es);
  m->slw = copy_bytes(md->slw, n * m->es);
  if (md->free_top) {
    m->mvx = copy_bytes(md->matVx2Vz, plane * m->es);
    m->mvy = copy_bytes(md->matVy2Vz, plane * m->es);
  }
  for (int f = 0; f < 6; f++) {
    if (md->pml_nlay[f] <= 0) continue;
    if (!md->pml_A[f] || !md->pml_B[f] || !md->pml_D[f]) return sfail(CGFD_E_ARG, "face %d: null profile", f);
    const size_t k = (size_t)(md->pml_nlay[f] + 1) * m->es;
    m->A[f] = copy_bytes(md->pml_A[f], k);
    m->B[f] = copy_bytes(md->pml_B[f], k);
    m->D[f] = copy_bytes(md->pml_D[f], k);
  }
  if (md->scale_exp2) m->scale.assign(md->scale_exp2, md->scale_exp2 + 9);
  // pointers in the stored descriptor refer to the caller's arrays: clear them
  m->d.metric = m->d.lam = m->d.mu = m->d.slw = m->d.matVx2Vz = m->d.matVy2Vz = nullptr;
  for (int f = 0; f < 6; f++) m->d.pml_A[f] = m->d.pml_B[f] = m->d.pml_D[f] = nullptr;
  m->d.scale_exp2 = nullptr;
  std::lock_guard<std::mutex> lk(srv->mu);
  if (group_bytes(*m, 1) > srv->budget)
    return sfail(CGFD_E_NOMEM, "insufficient HBM: one shot of this model needs %lld bytes, the budget is %lld",
                 group_bytes(*m, 1), srv->budget);
  *model_id = (int)srv->models.size();
  srv->models.push_back(std::move(m));
  return CGFD_OK;
}

int cgfd_server_submit(cgfd_server *srv, const cgfd_request *rq, long long *ticket) {
  if (!srv || !rq || !ticket) return sfail(CGFD_E_ARG, "null argument");
  auto r = std::make_unique<Request>();
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    if (rq->model < 0 || rq->model >= (int)srv->models.size()) return sfail(CGFD_E_ARG, "unknown model %d", rq->model);
    const Model &m = *srv->models[rq->model];
    if (rq->nt_total < 1 || rq->nt_total > m.d.max_nt)
      return sfail(CGFD_E_ARG, "nt_total %d outside 1..%d of the model", rq->nt_total, m.d.max_nt);
    if (rq->nrecv < 0 || rq->nrecv > m.d.max_nrecv)
      return sfail(CGFD_E_ARG, "nrecv %d outside 0..%d of the model", rq->nrecv, m.d.max_nrecv);
  }
  if (rq->nrecv > 0 && (!rq->recv_indx8 || !rq->recv_frac3)) return sfail(CGFD_E_ARG, "receiver arrays missing");
  r->model = rq->model;
  r->nt = rq->nt_total;
  if (rq->src && rq->src->nsrc > 0) {
    const cgfd_source_desc &s = *rq->src;
    if (!s.it_begin || !s.it_end || !s.ext_num || !s.ext_indx || !s.ext_coef)
      return sfail(CGFD_E_ARG, "source arrays missing");
    const size_t es = s.dtype == CGFD_DT_F64 ? 8 : 4;
    r->nsrc = s.nsrc; r->force_on = s.force_on; r->moment_on = s.moment_on;
    r->max_nt = s.max_nt; r->max_ext = s.max_ext; r->sdtype = s.dtype;
    r->itb.assign(s.it_begin, s.it_begin + s.nsrc);
    r->ite.assign(s.it_end, s.it_end + s.nsrc);
    r->en.assign(s.ext_num, s.ext_num + s.nsrc);
    r->ei.assign(s.ext_indx, s.ext_indx + (size_t)s.nsrc * s.max_ext);
    r->ec = copy_bytes(s.ext_coef, (size_t)s.nsrc * s.max_ext * es);
    const size_t nt4 = (size_t)s.nsrc * s.max_nt * 4;
    if (s.force_on && s.F) r->F = copy_bytes(s.F, 3 * nt4 * es);
    if (s.moment_on && s.M) r->M = copy_bytes(s.M, 6 * nt4 * es);
  }
  r->nrecv = rq->nrecv;
  r->rdtype = rq->recv_dtype;
  if (rq->nrecv > 0) {
    r->ridx.assign(rq->recv_indx8, rq->recv_indx8 + (size_t)rq->nrecv * 8);
    r->rfrac = copy_bytes(rq->recv_frac3, (size_t)rq->nrecv * 3 * (rq->recv_dtype == CGFD_DT_F64 ? 8 : 4));
  }
  r->t_submit = Clock::now();
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    if (srv->stop) return sfail(CGFD_E_STATE, "server is shutting down");
    r->ticket = srv->next_ticket++;
    *ticket = r->ticket;
    srv->pending.push_back(r->ticket);
    srv->reqs[r->ticket] = std::move(r);
    srv->st.submitted++;
  }
  srv->cv_work.notify_all();
  return CGFD_OK;
}

int cgfd_server_wait(cgfd_server *srv, long long ticket, double timeout_s) {
  if (!srv) return sfail(CGFD_E_ARG, "null server");
  std::unique_lock<std::mutex> lk(srv->mu);
  auto it = srv->reqs.find(ticket);
  if (it == srv->reqs.end()) return sfail(CGFD_E_ARG, "unknown ticket %lld", ticket);
  Request *r = it->second.get();
  auto done = [&]() { return r->state == DONE || r->state == REJECTED; };
  if (timeout_s < 0) {
    srv->cv_done.wait(lk, done);
  } else if (!srv->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), done)) {
    return 1;
  }
  if (r->state == REJECTED) return sfail(CGFD_E_NOMEM, "request %lld rejected: %s", ticket, r->error.c_str());
  return CGFD_OK;
}

int cgfd_server_result(cgfd_server *srv, long long ticket, int dtype, void *seismo, double *t_admit,
                       double *t_done) {
  if (!srv) return sfail(CGFD_E_ARG, "null server");
  std::lock_guard<std::mutex> lk(srv->mu);
  auto it = srv->reqs.find(ticket);
  if (it == srv->reqs.end()) return sfail(CGFD_E_ARG, "unknown ticket %lld", ticket);
  Request &r = *it->second;
  if (r.state == QUEUED || r.state == RUNNING) return sfail(CGFD_E_STATE, "request %lld not finished", ticket);
  if (r.state == REJECTED) {
    std::string e = r.error;
    srv->reqs.erase(it);
    return sfail(CGFD_E_NOMEM, "request %lld rejected: %s", ticket, e.c_str());
  }
  if (seismo) {
    for (size_t q = 0; q < r.seismo.size(); q++) {
      if (dtype == CGFD_DT_F64) ((double *)seismo)[q] = r.seismo[q];
      else ((float *)seismo)[q] = (float)r.seismo[q];
    }
  }
  using D = std::chrono::duration<double>;
  if (t_admit) *t_admit = D(r.t_admit - r.t_submit).count();
  if (t_done) *t_done = D(r.t_done - r.t_submit).count();
  srv->reqs.erase(it);
  return CGFD_OK;
}

int cgfd_server_get_stats(cgfd_server *srv, cgfd_server_stats *out) {
  if (!srv || !out) return sfail(CGFD_E_ARG, "null argument");
  std::lock_guard<std::mutex> lk(srv->mu);
  *out = srv->st;
  return CGFD_OK;
}

}  // extern "C"
