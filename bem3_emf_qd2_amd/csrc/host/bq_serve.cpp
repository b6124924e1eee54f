// bq_serve.cpp -- request loop over resident models: continuous batching of incidence
// solves on one GPU.  No counterpart in the reference, whose d3qd2_bv_solver solves one
// wave read from ipw.txt per run (d3qd2_src/d3qd2_bv_solver.c:3-21); each request here is that
// run (wave set-up, solve_bieq, EH_t at probe points) on a model kept resident.
//
// Resident instances: a pool of models (one DQD2 each, read from its mesh's set-up file), each
// holding the device session of the last wave it solved (matrices and factors).  Worker slots
// (one host thread each) take ready requests as they arrive; the solver session of an instance
// runs on a stream of its own, so requests in different slots overlap on the device, and a
// phase gate lets at most one solve assemble and one factorise at a time (one request's FP64
// vector work beside another's matrix work).  A free slot serves, in arrival order:
//   1. a request whose mesh and wave (wavelength, direction) an idle instance holds: its
//      matrices and factors are reused (a polarisation or amplitude change only changes the
//      right-hand side, d3qd2_src/d3qd2_setup.c:1309-1349);
//   2. else the first request whose wave no busy instance holds (those wait for it): on a new
//      instance when its session fits the memory limit next to all others (idle instances keep
//      their waves for later repeats), else on the least recently used idle instance of its
//      mesh, else on a new instance after releasing idle instances, least recently used first.
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

struct Inst {
  int mesh = -1;
  DQD2* qd = nullptr;
  bool valid = false;            // a solved wave is resident
  double lam = 0, th = 0, ph = 0;
  bool busy = false;
  double bytes = 0.0;            // device bytes of its session (reserved estimate while running)
  double last_use = 0.0;
};

struct Engine {
  std::vector<Mesh> meshes;
  std::vector<Inst*> pool;
  int nslots = 1;
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
  if (m.MN == 1) {
    // single object: six n x n factors and work matrices, the inverted 512-row diagonal blocks
    // of three factors (3 x 2 x 512 n, when n is a multiple of 512), and the FGMRES and
    // preconditioner vectors
    const double blk = (m.NN % 512 == 0) ? 3072.0 * n : 0.0;
    b += work + (6.0 * n * n + blk) * (precision == 1 ? 8.0 : 16.0) + 8000.0 * n;
  } else {
    // block elimination of order 2n: M, M3 (2n x 2n), R, L (2n x n), W (n x n) in FP64, and
    // their FP32 copies in the mixed solve
    const double n2 = 2.0 * n, fp64 = 16.0 * (2.0 * n2 * n2 + 2.0 * n2 * n + n * n);
    b += work + fp64 + (precision != 0 ? 0.5 * fp64 : 0.0) + 8000.0 * n;
  }
  return b;
}

Inst* new_inst(Engine& E, int m)
{
  Inst* I = new Inst();
  I->mesh = m;
  I->qd = calloc(1, (size_t)bq_sizeof_dqd2());
  dat_read((char*)E.meshes[m].path.c_str(), I->qd);
  E.pool.push_back(I);
  return I;
}

void drop_inst(Engine& E, Inst* I)
{
  finalize_dqd2(I->qd);  // releases its solver session
  free(I->qd);
  E.pool.erase(std::find(E.pool.begin(), E.pool.end(), I));
  delete I;
}

double held(const Engine& E)
{
  double b = 0.0;
  for (const Inst* I : E.pool) b += I->bytes;
  return b;
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
  E->nslots = nslots;
  return E;
}

void bq_serve_destroy(void* h)
{
  Engine* E = (Engine*)h;
  if (!E) return;
  while (!E->pool.empty()) drop_inst(*E, E->pool.back());
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

  auto same_wave = [](const Inst* I, const bq_serve_req& r) {
    return I->valid && I->mesh == r.mesh && I->lam == r.lambda0 && I->th == r.theta && I->ph == r.phi;
  };
  // a request and an instance for a free slot (mu held); false if nothing can start now
  auto pick = [&](int& id, Inst*& inst) -> bool {
    for (size_t q = 0; q < ready.size(); ++q)
      for (Inst* I : E->pool)
        if (!I->busy && same_wave(I, req[ready[q]])) {
          id = ready[q];
          inst = I;
          ready.erase(ready.begin() + q);
          return true;
        }
    for (size_t q = 0; q < ready.size(); ++q) {
      const bq_serve_req& r = req[ready[q]];
      bool deferred = false;
      for (const Inst* I : E->pool) deferred = deferred || (I->busy && same_wave(I, r));
      if (deferred) continue;  // its factors are in use: it reuses them when that request ends
      // a new instance while memory allows (the idle ones keep their waves for repeats), else
      // the least recently used idle instance of the mesh, else a new one after releasing idle
      // instances of any mesh, least recently used first
      const double need = E->meshes[r.mesh].est_bytes;
      Inst* best = nullptr;
      if (held(*E) + need <= E->limit) {
        best = new_inst(*E, r.mesh);
        best->bytes = need;
      }
      if (!best)
        for (Inst* I : E->pool)
          if (!I->busy && I->mesh == r.mesh && (!best || I->last_use < best->last_use)) best = I;
      if (!best) {
        std::vector<Inst*> idle;
        for (Inst* I : E->pool)
          if (!I->busy) idle.push_back(I);
        std::sort(idle.begin(), idle.end(), [](const Inst* a, const Inst* b) { return a->last_use < b->last_use; });
        size_t k = 0;
        while (held(*E) + need > E->limit && k < idle.size()) drop_inst(*E, idle[k++]);
        if (held(*E) + need > E->limit) continue;  // wait for busy instances to end
        best = new_inst(*E, r.mesh);
        best->bytes = need;
      }
      id = ready[q];
      inst = best;
      ready.erase(ready.begin() + q);
      return true;
    }
    return false;
  };

  auto worker = [&](int w) {
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
      int id = -1;
      Inst* I = nullptr;
      cv.wait(lk, [&] {
        if (released == nreq && ready.empty()) return true;
        return pick(id, I);
      });
      if (id < 0) return;  // nothing left
      const bq_serve_req& r = req[id];
      I->busy = true;
      I->bytes = std::max(I->bytes, E->meshes[r.mesh].est_bytes);
      lk.unlock();
      bq_serve_out o;
      memset(&o, 0, sizeof o);
      o.slot = w;
      o.t_arrival_s = r.arrival_s;
      o.t_start_s = now_s(t0);
      const double ex = r.pol == 1 ? 1.0 : 0.0, ey = r.pol == 1 ? 0.0 : 1.0;
      int rc = bq_set_incidence(I->qd, r.lambda0, r.theta, r.phi, ex, 0.0, ey, 0.0, 0.0);
      if (rc == 0) rc = bq_solve_bieq_rc(I->qd);
      if (rc == 0 && nprobe > 0) {
        std::vector<int> did(nprobe);
        rc = bq_EH_t_batch(I->qd, nprobe, probes, 1, E_out + (size_t)id * nprobe * 6,
                           H_out + (size_t)id * nprobe * 6, did.data());
      }
      if (rc == 0) {
        bq_stats st;
        if (bq_solve_stats(I->qd, &st) == 0) {
          o.reused = st.reused;
          o.gmres_iterations = st.gmres_iterations;
          o.backward_error = st.backward_error;
          o.n_qpgf_gl = st.reused ? 0 : st.n_qpgf_gl;
          o.asm_cut = st.asm_cut;
        }
      }
      o.status = rc;
      o.t_done_s = now_s(t0);
      std::string why = rc ? bq_last_error() : "";  // this thread's error record
      const double used = (double)bq_solver_bytes(I->qd);
      lk.lock();
      out[id] = o;
      Mesh& m = E->meshes[r.mesh];
      if (rc == 0 && used > 0.0) {
        I->bytes = used;
        if (!m.measured || used > m.est_bytes) m.est_bytes = used;
        m.measured = true;
        I->valid = true;
        I->lam = r.lambda0;
        I->th = r.theta;
        I->ph = r.phi;
      } else {
        bq_release_solver(I->qd);
        I->bytes = 0.0;
        I->valid = false;
      }
      if (rc != 0) ++failed;
      if (log)
        fprintf(stderr, "bq_serve: request %d (mesh %d) on slot %d: %s, start %.2f s, done %.2f s%s%s%s\n", id,
                r.mesh, w, rc ? "failed" : "ok", o.t_start_s, o.t_done_s, o.reused ? ", factors reused" : "",
                rc ? ": " : "", why.c_str());
      I->busy = false;
      I->last_use = o.t_done_s;
      ++finished;
      cv.notify_all();
    }
  };

  if (E->nslots > 1) bq_set_phase_gate(gate_fn);
  std::vector<std::thread> th;
  for (int w = 0; w < E->nslots; ++w) th.emplace_back(worker, w);
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
