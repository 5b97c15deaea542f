// serve.cpp - served steady solves over ccfd_batch arenas (ccfd_serve.h):
// FIFO admission within the arenas' room, step bursts, residual polling,
// retirement, and promotion to a polish arena. Host code only; the arenas
// are reached through ccfd_serve_backend.
#include "ccfd_serve.h"

#include <math.h>
#include <time.h>

#include <algorithm>
#include <deque>
#include <map>
#include <string>
#include <vector>

namespace ccfd {
int fail(int code, const char* fmt, ...);  // ccfd_gpu.hip: last-error message
}
using ccfd::fail;

namespace {

struct Queued {
  long id;
  const void* mesh;
  long n;
  const double* prim;
  double cfl;
};

struct Solve {
  long id = 0;
  const void* mesh = nullptr;
  long n = 0;
  double cfl = 0.0;
  double admitted = 0.0;
  bool haveRes0 = false;
  double res0 = 0.0;        // density residual of the first main-arena step
  long steps0 = 0;          // main-arena steps before the move
  long polishSteps = 0;
  std::vector<double> hist; // drop after each poll (promotion only)
  std::vector<double> state;  // primitive state handed to the polish arena
};

double mono_clock(void*) {
  timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

// consPrim (src/equationOfState.c:37-52)
void cons_prim(double gam1, const double* c, double* p) {
  p[0] = c[0];
  p[1] = c[1] / c[0];
  p[2] = c[2] / c[0];
  p[3] = gam1 * (c[3] - 0.5 * (c[1] * p[1] + c[2] * p[2]));
  if (p[0] < 1e-19) p[0] = 1e-10;
  if (p[3] < 1e-19) p[3] = 1e-10;
}

}  // namespace

struct ccfd_server {
  ccfd_serve_backend main{}, polish{};
  bool hasPolish = false;
  ccfd_serve_cfg cfg{};
  std::deque<Queued> queue;
  std::map<int, Solve> active;      // main arena id -> solve
  std::map<int, Solve> polishing;   // polish arena id -> solve
  std::deque<Solve> moving;         // waiting for room in the polish arena
  std::deque<ccfd_serve_result> done;
  std::map<long, std::vector<double>> kept;
  std::vector<int> ids;
  std::vector<ccfd_batch_info> info;

  double now() { return cfg.clock(cfg.clockArg); }
  int in_flight() const {
    return (int)(active.size() + polishing.size() + moving.size());
  }
};

static int retire(ccfd_server* s, const ccfd_serve_backend& be, int k,
                  Solve& r, const ccfd_batch_info& inf, double drop, bool conv,
                  long steps) {
  int rc;
  if (s->cfg.keepState) {
    std::vector<double> prim(4 * (size_t)r.n);
    if ((rc = be.get_prim(be.arena, k, prim.data()))) return rc;
    s->kept[r.id] = std::move(prim);
  }
  if ((rc = be.remove(be.arena, k))) return rc;
  ccfd_serve_result res;
  res.id = r.id;
  res.nElems = r.n;
  res.steps = steps;
  res.polishSteps = r.polishSteps;
  res.drop = drop;
  res.converged = conv ? 1 : 0;
  res.nonfinite = inf.nonfinite ? 1 : 0;
  res.admitted = r.admitted;
  res.done = s->now();
  s->done.push_back(res);
  return CCFD_OK;
}

static int poll(ccfd_server* s, const ccfd_serve_backend& be, size_t cap,
                int* n) {
  s->ids.resize(std::max<size_t>(cap, 1));
  s->info.resize(std::max<size_t>(cap, 1));
  return be.poll(be.arena, (int)s->ids.size(), s->ids.data(), s->info.data(), n);
}

// retire or promote the main arena's solves (continuous batching)
static int poll_main(ccfd_server* s) {
  int n = 0, rc;
  if ((rc = poll(s, s->main, s->active.size(), &n))) return rc;
  const ccfd_serve_cfg& c = s->cfg;
  for (int j = 0; j < n; ++j) {
    const int k = s->ids[j];
    auto it = s->active.find(k);
    if (it == s->active.end()) continue;
    Solve& r = it->second;
    const ccfd_batch_info& inf = s->info[j];
    if (!r.haveRes0) {
      r.res0 = inf.res0[0];
      r.haveRes0 = true;
    }
    const double drop = r.res0 > 0.0 ? inf.res[0] / r.res0 : 0.0;
    const bool conv = drop <= c.targetDrop;
    if (conv || inf.steps >= c.maxSteps || inf.nonfinite) {
      if ((rc = retire(s, s->main, k, r, inf, drop, conv, inf.steps))) return rc;
      s->active.erase(it);
      continue;
    }
    if (!s->hasPolish) continue;
    r.hist.push_back(drop);
    // stalled near the main arena's floor: no halving of the best drop in
    // the last stallRounds polls, within 10x of the switch level (a solve
    // that stalls far above it would not converge in the polish arena
    // either and stays on the cheaper arena until the cap)
    const size_t h = r.hist.size(), w = (size_t)c.stallRounds;
    bool stalled = false;
    if (h > w && drop <= 10.0 * c.switchDrop) {
      const double recent = *std::min_element(r.hist.end() - w, r.hist.end());
      const double before = *std::min_element(r.hist.begin(), r.hist.end() - w);
      stalled = recent > 0.5 * before;
    }
    if (drop <= c.switchDrop || stalled) {
      std::vector<double> cons(4 * (size_t)r.n);
      if ((rc = s->main.get_cons(s->main.arena, k, cons.data()))) return rc;
      r.state.resize(cons.size());
      const double gam1 = c.gamma - 1.0;
      for (long e = 0; e < r.n; ++e)
        cons_prim(gam1, cons.data() + 4 * e, r.state.data() + 4 * e);
      r.steps0 = inf.steps;
      if ((rc = s->main.remove(s->main.arena, k))) return rc;
      s->moving.push_back(std::move(r));
      s->active.erase(it);
    }
  }
  return CCFD_OK;
}

static int poll_polish(ccfd_server* s) {
  int n = 0, rc;
  if ((rc = poll(s, s->polish, s->polishing.size(), &n))) return rc;
  for (int j = 0; j < n; ++j) {
    const int k = s->ids[j];
    auto it = s->polishing.find(k);
    if (it == s->polishing.end()) continue;
    Solve& r = it->second;
    const ccfd_batch_info& inf = s->info[j];
    const double drop = r.res0 > 0.0 ? inf.res[0] / r.res0 : 0.0;
    const bool conv = drop <= s->cfg.targetDrop;
    r.polishSteps = inf.steps;
    if (conv || r.steps0 + inf.steps >= s->cfg.maxSteps || inf.nonfinite) {
      if ((rc = retire(s, s->polish, k, r, inf, drop, conv, r.steps0 + inf.steps)))
        return rc;
      s->polishing.erase(it);
    }
  }
  return CCFD_OK;
}

// ------------------------------------------------------------------ C ABI

extern "C" int ccfd_serve_create(ccfd_server** out, const ccfd_serve_backend* main,
                                 const ccfd_serve_backend* polish,
                                 const ccfd_serve_cfg* cfg) {
  if (!out || !main || !cfg) return fail(CCFD_EINVAL, "null argument");
  auto complete = [](const ccfd_serve_backend* b) {
    return b->fits && b->add && b->remove && b->step && b->poll && b->get_cons &&
           b->get_prim;
  };
  if (!complete(main) || (polish && !complete(polish)))
    return fail(CCFD_EINVAL, "backend with missing functions");
  if (cfg->stepsPerRound < 1 || cfg->maxSteps < 1 || cfg->stallRounds < 1 ||
      !(cfg->targetDrop > 0.0) || !(cfg->gamma > 1.0) || cfg->maxActive < 0)
    return fail(CCFD_EINVAL, "bad serve settings");
  ccfd_server* s = new ccfd_server;
  s->main = *main;
  s->hasPolish = polish != nullptr;
  if (polish) s->polish = *polish;
  s->cfg = *cfg;
  if (!s->cfg.clock) s->cfg.clock = mono_clock;
  *out = s;
  return CCFD_OK;
}

extern "C" int ccfd_serve_destroy(ccfd_server* s) {
  delete s;
  return CCFD_OK;
}

extern "C" int ccfd_serve_submit(ccfd_server* s, long id, const void* mesh,
                                 long nElems, const double* prim, double cfl) {
  if (!s || !mesh || nElems <= 0) return fail(CCFD_EINVAL, "bad request");
  s->queue.push_back({id, mesh, nElems, prim, cfl});
  return CCFD_OK;
}

extern "C" int ccfd_serve_round(ccfd_server* s, int* nActive, int* nPolishing,
                                int* nQueued) {
  if (!s) return fail(CCFD_EINVAL, "null server");
  int rc = CCFD_OK, f = 0;
  auto report = [&]() {
    if (nActive) *nActive = (int)s->active.size();
    if (nPolishing) *nPolishing = (int)(s->polishing.size() + s->moving.size());
    if (nQueued) *nQueued = (int)s->queue.size();
  };
  // solves handed over to the polish arena, in order
  while (!s->moving.empty()) {
    Solve& r = s->moving.front();
    if ((rc = s->polish.fits(s->polish.arena, r.mesh, &f))) return rc;
    if (!f) {
      if (s->polishing.empty()) {
        report();
        return fail(CCFD_ENOMEM, "request %ld does not fit the empty polish arena",
                    r.id);
      }
      break;
    }
    int k;
    if ((rc = s->polish.add(s->polish.arena, r.mesh, r.state.data(), r.cfl, &k)))
      return rc;
    r.state.clear();
    r.state.shrink_to_fit();
    s->polishing[k] = std::move(r);
    s->moving.pop_front();
  }
  // admit queued requests in order while they fit
  while (!s->queue.empty() &&
         (s->cfg.maxActive == 0 || s->in_flight() < s->cfg.maxActive)) {
    const Queued q = s->queue.front();
    if ((rc = s->main.fits(s->main.arena, q.mesh, &f))) return rc;
    if (!f) {
      if (s->active.empty()) {
        report();
        return fail(CCFD_ENOMEM, "request %ld (%ld cells) does not fit the empty arena",
                    q.id, q.n);
      }
      break;
    }
    int k;
    if ((rc = s->main.add(s->main.arena, q.mesh, q.prim, q.cfl, &k))) return rc;
    Solve r;
    r.id = q.id;
    r.mesh = q.mesh;
    r.n = q.n;
    r.cfl = q.cfl;
    r.admitted = s->now();
    s->active[k] = std::move(r);
    s->queue.pop_front();
  }
  const bool runMain = !s->active.empty(), runPolish = !s->polishing.empty();
  if (runMain && (rc = s->main.step(s->main.arena, s->cfg.stepsPerRound))) return rc;
  if (runPolish && (rc = s->polish.step(s->polish.arena, s->cfg.stepsPerRound)))
    return rc;
  if (runMain && (rc = poll_main(s))) return rc;
  if (runPolish && (rc = poll_polish(s))) return rc;
  report();
  return CCFD_OK;
}

extern "C" int ccfd_serve_results(ccfd_server* s, int max, ccfd_serve_result* out,
                                  int* n) {
  if (!s || !n || max < 0 || (max > 0 && !out)) return fail(CCFD_EINVAL, "bad arguments");
  int k = 0;
  while (k < max && !s->done.empty()) {
    out[k++] = s->done.front();
    s->done.pop_front();
  }
  *n = k;
  return CCFD_OK;
}

extern "C" int ccfd_serve_take_state(ccfd_server* s, long id, double* prim) {
  if (!s || !prim) return fail(CCFD_EINVAL, "null argument");
  auto it = s->kept.find(id);
  if (it == s->kept.end()) return fail(CCFD_EINVAL, "no kept state for request %ld", id);
  std::copy(it->second.begin(), it->second.end(), prim);
  s->kept.erase(it);
  return CCFD_OK;
}

// ---------------------------------------------- the backend of a ccfd_batch

namespace {
ccfd_batch* B(void* a) { return (ccfd_batch*)a; }
int b_fits(void* a, const void* m, int* f) {
  return ccfd_batch_fits(B(a), (const ccfd_mesh*)m, f);
}
int b_add(void* a, const void* m, const double* prim, double cfl, int* id) {
  ccfd_batch_case cs = {cfl, 1e300, 1e300};
  return ccfd_batch_add(B(a), (const ccfd_mesh*)m, &cs, prim, id);
}
int b_fits_raw(void* a, const void* m, int* f) {
  return ccfd_batch_fits_raw(B(a), (const ccfd_raw_mesh*)m, f);
}
int b_add_raw(void* a, const void* m, const double* prim, double cfl, int* id) {
  ccfd_batch_case cs = {cfl, 1e300, 1e300};
  return ccfd_batch_add_raw(B(a), (const ccfd_raw_mesh*)m, &cs, prim, id);
}
int b_remove(void* a, int id) { return ccfd_batch_remove(B(a), id); }
int b_step(void* a, int n) { return ccfd_batch_step(B(a), n); }
int b_poll(void* a, int maxIds, int* ids, ccfd_batch_info* info, int* n) {
  return ccfd_batch_poll(B(a), maxIds, ids, info, n);
}
int b_cons(void* a, int id, double* c) { return ccfd_batch_get_cons(B(a), id, c); }
int b_prim(void* a, int id, double* p) { return ccfd_batch_get_prim(B(a), id, p); }
}  // namespace

extern "C" int ccfd_serve_backend_batch(ccfd_batch* b, ccfd_serve_backend* out) {
  if (!b || !out) return fail(CCFD_EINVAL, "null argument");
  out->arena = b;
  out->fits = b_fits;
  out->add = b_add;
  out->remove = b_remove;
  out->step = b_step;
  out->poll = b_poll;
  out->get_cons = b_cons;
  out->get_prim = b_prim;
  return CCFD_OK;
}

extern "C" int ccfd_serve_backend_batch_raw(ccfd_batch* b, ccfd_serve_backend* out) {
  int rc = ccfd_serve_backend_batch(b, out);
  if (rc) return rc;
  out->fits = b_fits_raw;
  out->add = b_add_raw;
  return CCFD_OK;
}
