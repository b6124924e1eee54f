// bq_serve.cpp -- request loop over resident models: continuous batching of incidence
// solves on one GPU.  No counterpart in the reference, whose d3qd2_bv_solver solves one
// wave read from ipw.txt per run (d3qd2_src/d3qd2_bv_solver.c:3-21); each request here is that
// run (wave set-up, solve_bieq, EH_t at probe points) on a model kept resident.
//
// Worker slots: one host thread per slot; each owns one DQD2 per mesh (read from the mesh's
// set-up file), and the solver session of a DQD2 runs on a stream of its own, so requests in
// different slots overlap on the device (one slot's assembly, FP64 vector work, beside
// another's factorisation, MFMA work).  Requests enter the ready queue at their arrival times.
// A free slot takes, in arrival order, the first ready request it can admit: preferably one
// whose mesh and wave (wavelength, direction) are resident in that slot, whose matrices and
// factors are then reused (a polarisation or amplitude change only changes the right-hand
// side, d3qd2_src/d3qd2_setup.c:1309-1349); else the first whose device memory fits the limit
// next to what the other slots hold, releasing this slot's own idle session first and, if
// still short, the least recently used idle session of another slot.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "bq_amd.h"
#include "../bq_solver.h"

// The drop-in API's model type is C99 (complex members); here it is an opaque buffer of
// bq_sizeof_dqd2() bytes handled through the C entry points.
typedef void DQD2;
extern "C" {
void dat_read(char* fname, void* qd);
void finalize_dqd2(void* qd);
}

namespace {

using clk = std::chrono::steady_clock;

struct Mesh {
  std::string path;
  int Ne = 0, NN = 0, MN = 0;
  std::vector<int> dNe;
  double est_bytes = 0.0;  // device bytes of a resident session (estimate, then measured)
  bool measured = false;
};

struct Slot {
  std::vector<DQD2*> qd;        // per mesh, read on first use
  int mesh = -1;                // mesh of the resident session
  double lam = 0, th = 0, ph = 0;
  bool busy = false;
  double bytes = 0.0;           // device bytes of the resident session
  double last_use = 0.0;
};

struct Engine {
  std::vector<Mesh> meshes;
  std::vector<Slot> slots;
  double limit = 0.0;
  int precision = 1;
};

double now_s(clk::time_point t0) { return std::chrono::duration<double>(clk::now() - t0).count(); }

// Phase gate: at most one solve in its assembly (FP64 vector units) and one in its
// factorisation and solve (matrix units) at a time, so that concurrent requests pipeline
// instead of running the same phase side by side.
struct PhaseGate {
  std::mutex mu;
  std::condition_variable cv;
  int busy[2] = {0, 0};
};
PhaseGate g_gate;

void gate_fn(int phase, int enter)
{
  std::unique_lock<std::mutex> lk(g_gate.mu);
  if (enter) {
    g_gate.cv.wait(lk, [&] { return g_gate.busy[phase] == 0; });
    g_gate.busy[phase] = 1;
  } else {
    g_gate.busy[phase] = 0;
    g_gate.cv.notify_all();
  }
}

// resident bytes of a session of mesh m: assembled matrices (G, H of every domain, dG, dH of
// the bounded ones), the assembly work buffer and the single-object reduction's factors
double estimate_bytes(const Mesh& m, int precision)
{
  double b = 0.0, work = 0.0;
  for (int d = 0; d <= m.MN; ++d) {
    const double N = 4.0 * m.dNe[d];
    b += 32.0 * N * N + (d > 0 ? 64.0 * N * N : 0.0);
    work = std::max(work, 29.0 * N * m.dNe[d]);
  }
  const double n = m.NN;
  // six n x n factors and work matrices, and the FGMRES and preconditioner vectors
  b += work + 6.0 * (precision == 1 ? 8.0 : 16.0) * n * n + 8000.0 * n;
  return b;
}

DQD2* slot_model(Engine& E, Slot& s, int m)
{
  if ((int)s.qd.size() < (int)E.meshes.size()) s.qd.resize(E.meshes.size(), nullptr);
  if (!s.qd[m]) {
    DQD2* q = calloc(1, (size_t)bq_sizeof_dqd2());
    dat_read((char*)E.meshes[m].path.c_str(), q);
    s.qd[m] = q;
  }
  return s.qd[m];
}

void release_slot(Engine& E, Slot& s)
{
  if (s.mesh >= 0 && s.qd[s.mesh]) bq_release_solver(s.qd[s.mesh]);
  s.mesh = -1;
  s.bytes = 0.0;
}

}  // namespace

extern "C" {

void* bq_serve_create(int nmesh, const char* const* setup_dat, int nslots, double mem_limit_gb, int precision)
{
  if (nmesh <= 0 || !setup_dat || nslots <= 0 || nslots > 64 || !(mem_limit_gb > 0.0)) return nullptr;
  Engine* E = new Engine();
  E->limit = mem_limit_gb * 1073741824.0;
  E->precision = precision;
  for (int i = 0; i < nmesh; ++i) {
    Mesh m;
    m.path = setup_dat[i];
    DQD2* q = calloc(1, (size_t)bq_sizeof_dqd2());
    dat_read((char*)m.path.c_str(), q);
    m.MN = bq_model_dims(q, &m.Ne, &m.NN, nullptr);
    m.dNe.assign(m.MN + 1, 0);
    bq_model_dims(q, &m.Ne, &m.NN, m.dNe.data());
    finalize_dqd2(q);
    free(q);
    m.est_bytes = estimate_bytes(m, precision);
    E->meshes.push_back(m);
  }
  E->slots.resize(nslots);
  return E;
}

void bq_serve_destroy(void* h)
{
  Engine* E = (Engine*)h;
  if (!E) return;
  for (Slot& s : E->slots)
    for (DQD2* q : s.qd)
      if (q) {
        finalize_dqd2(q);
        free(q);
      }
  delete E;
}

int bq_serve_run(void* h, int nreq, const bq_serve_req* req, int nprobe, const double* probes, double* E_out,
                 double* H_out, bq_serve_out* out)
{
  Engine* E = (Engine*)h;
  if (!E || nreq < 0 || (nreq > 0 && (!req || !out)) || nprobe < 0 || (nprobe > 0 && (!probes || !E_out || !H_out)))
    return -2;
  for (int i = 0; i < nreq; ++i)
    if (req[i].mesh < 0 || req[i].mesh >= (int)E->meshes.size()) return -2;
  for (const Mesh& m : E->meshes)
    if (m.est_bytes > E->limit) return -4;  // a request of this mesh could never be admitted
  const bool log = getenv("BQ_SERVE_LOG") != nullptr;
  const int prev_precision = bq_solve_precision(-1);
  bq_solve_precision(E->precision);
  std::vector<int> order(nreq);
  for (int i = 0; i < nreq; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return req[a].arrival_s < req[b].arrival_s; });
  std::mutex mu;
  std::condition_variable cv;
  std::deque<int> ready;
  int released = 0, finished = 0, failed = 0;
  const clk::time_point t0 = clk::now();

  // bytes held by slots other than w
  auto held_by_others = [&](int w) {
    double b = 0.0;
    for (int i = 0; i < (int)E->slots.size(); ++i)
      if (i != w) b += E->slots[i].bytes;
    return b;
  };
  // pick a ready request for slot w (mu held); -1 if none fits now
  auto pick = [&](int w) -> int {
    Slot& s = E->slots[w];
    for (size_t q = 0; q < ready.size(); ++q) {
      const bq_serve_req& r = req[ready[q]];
      if (r.mesh == s.mesh && r.lambda0 == s.lam && r.theta == s.th && r.phi == s.ph) {
        const int id = ready[q];
        ready.erase(ready.begin() + q);
        return id;
      }
    }
    for (size_t q = 0; q < ready.size(); ++q) {
      const bq_serve_req& r = req[ready[q]];
      const double need = E->meshes[r.mesh].est_bytes;
      const double own = (r.mesh == s.mesh) ? s.bytes : 0.0;
      double others = held_by_others(w);
      if (others + std::max(need, own) > E->limit) {
        // release idle sessions of other slots, least recently used first
        std::vector<int> idle;
        for (int i = 0; i < (int)E->slots.size(); ++i)
          if (i != w && !E->slots[i].busy && E->slots[i].mesh >= 0) idle.push_back(i);
        std::sort(idle.begin(), idle.end(),
                  [&](int a, int b) { return E->slots[a].last_use < E->slots[b].last_use; });
        for (int i : idle) {
          if (others + std::max(need, own) <= E->limit) break;
          release_slot(*E, E->slots[i]);
          others = held_by_others(w);
        }
      }
      if (others + std::max(need, own) <= E->limit) {
        const int id = ready[q];
        ready.erase(ready.begin() + q);
        return id;
      }
    }
    return -1;
  };

  auto worker = [&](int w) {
    std::unique_lock<std::mutex> lk(mu);
    Slot& s = E->slots[w];
    for (;;) {
      int id = -1;
      cv.wait(lk, [&] {
        if (released == nreq && ready.empty()) return true;
        id = pick(w);
        return id >= 0;
      });
      if (id < 0) return;  // nothing left
      const bq_serve_req& r = req[id];
      s.busy = true;
      const bool same_mesh = (s.mesh == r.mesh);
      if (!same_mesh) {
        release_slot(*E, s);
        s.mesh = r.mesh;
      }
      // reserve the estimate (or the measured size) while the request runs
      s.bytes = std::max(s.bytes, E->meshes[r.mesh].est_bytes);
      lk.unlock();
      bq_serve_out o;
      memset(&o, 0, sizeof o);
      o.slot = w;
      o.t_arrival_s = r.arrival_s;
      o.t_start_s = now_s(t0);
      DQD2* qd = slot_model(*E, s, r.mesh);
      const double ex = r.pol == 1 ? 1.0 : 0.0, ey = r.pol == 1 ? 0.0 : 1.0;
      int rc = bq_set_incidence(qd, r.lambda0, r.theta, r.phi, ex, 0.0, ey, 0.0, 0.0);
      if (rc == 0) rc = bq_solve_bieq_rc(qd);
      if (rc == 0 && nprobe > 0) {
        std::vector<int> did(nprobe);
        rc = bq_EH_t_batch(qd, nprobe, probes, 1, E_out + (size_t)id * nprobe * 6, H_out + (size_t)id * nprobe * 6,
                           did.data());
      }
      if (rc == 0) {
        bq_stats st;
        if (bq_solve_stats(qd, &st) == 0) {
          o.reused = st.reused;
          o.gmres_iterations = st.gmres_iterations;
          o.backward_error = st.backward_error;
        }
      }
      o.status = rc;
      o.t_done_s = now_s(t0);
      const double used = (double)bq_solver_bytes(qd);
      lk.lock();
      out[id] = o;
      Mesh& m = E->meshes[r.mesh];
      if (rc == 0 && used > 0.0) {
        s.bytes = used;
        if (!m.measured || used > m.est_bytes) m.est_bytes = used;
        m.measured = true;
        s.lam = r.lambda0;
        s.th = r.theta;
        s.ph = r.phi;
      } else {
        release_slot(*E, s);
        s.lam = s.th = s.ph = 0.0;
      }
      if (rc != 0) ++failed;
      if (log)
        fprintf(stderr, "bq_serve: request %d (mesh %d) on slot %d: %s, start %.2f s, done %.2f s%s\n", id, r.mesh,
                w, rc ? "failed" : "ok", o.t_start_s, o.t_done_s, o.reused ? ", factors reused" : "");
      s.busy = false;
      s.last_use = o.t_done_s;
      ++finished;
      cv.notify_all();
    }
  };

  if (E->slots.size() > 1) bq_set_phase_gate(gate_fn);
  std::vector<std::thread> th;
  for (int w = 0; w < (int)E->slots.size(); ++w) th.emplace_back(worker, w);
  for (int k = 0; k < nreq; ++k) {
    const int id = order[k];
    const double wait = req[id].arrival_s - now_s(t0);
    if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
    std::lock_guard<std::mutex> lk(mu);
    ready.push_back(id);
    ++released;
    cv.notify_all();
  }
  {
    std::lock_guard<std::mutex> lk(mu);
    cv.notify_all();
  }
  for (auto& t : th) t.join();
  bq_set_phase_gate(nullptr);
  bq_solve_precision(prev_precision);
  return failed ? -5 : 0;
}

}  // extern "C"
