// bq_sess.hpp -- the device solver session shared by the solver translation units
// (bq_solver.hip: set-up, assembly, block elimination; bq_reduced.hip: the single-object
// reduction and its FP64 GMRES refinement).
#pragma once
#include <vector>
#include <hip/hip_runtime.h>
#include "../bq_solver.h"
#include "../bq_plan.h"

// Factors of the single-object reduction (one object: every element has the open region as
// main domain and domain 1 as sub-domain).  All matrices n x n column-major in global
// element-node order (n = NN), complex FP64 or FP32 (prec), in buffers of FP64 size when
// cap64 is set.
struct bq_reduced {
  int prec = -1;            // -1 none, 0 complex FP64, 1 complex FP32
  bool cap64 = false;       // buffers hold complex FP64
  void* G0 = nullptr;       // LU of G0
  void* G1 = nullptr;       // LU of G1
  void* Mx = nullptr;       // M = G0^-1 H0
  void* P = nullptr;        // P = G1^-1 H1 (work; free after the Schur complement)
  void* B = nullptr;        // LU of B = P + M, then LU of S
  void* Binv = nullptr;     // B^-1
  int *ip0 = nullptr, *ip1 = nullptr, *ipB = nullptr;
  void *inv0 = nullptr, *inv1 = nullptr, *invB = nullptr;
  // inverted 512-row diagonal blocks of the G0, G1 and final (Schur) factors for the few-column
  // solves of the preconditioner and of dE/dH (null when n is not a multiple of 512)
  void *W0 = nullptr, *W1 = nullptr, *WB = nullptr;
  int* ginv[2] = {nullptr, nullptr};  // global node -> local row of domain d
  int* gidx[2] = {nullptr, nullptr};  // local row of domain d -> global node (-1: none)
  double* nn3 = nullptr;    // unit normals per global node [n][3]
  double* Uig = nullptr;    // incident Ui per global node, [3][n] complex (sub-domain side)
  double* dUig = nullptr;   // incident dUi per global node, [3][n] complex
  double* rownorm = nullptr;  // per equation row: sum |A_ij| of the unreduced operator [8 n]
  // work vectors (complex FP64) of length 8 n, and the FGMRES bases
  double *b = nullptr, *x = nullptr, *r = nullptr, *t = nullptr;
  double* loc = nullptr;    // domain-local work: 12 N complex (operands and row products)
  double* nloc = nullptr;   // normals in domain-local column order [N][3]
  double *Vb = nullptr, *Zb = nullptr;  // FGMRES bases: kRestart + 1 and kRestart vectors of 8 n
  double* part = nullptr;   // partial sums of the dot-product kernels
  void* vf = nullptr;       // FP32 / FP64 vectors for the preconditioner [8 n + 12 n] complex
  double anorm = 0.0;       // infinity norm of the unreduced operator
  size_t bytes = 0;         // device bytes held
};

struct bq_sess {
  bq_model m;
  int Ne, NN, MN, ps;
  hipStream_t s;
  hipStream_t s2 = nullptr;  // second stream: the object's factorization beside the open region's
  bool own_stream = false;
  double *E = nullptr, *ren = nullptr, *nrm = nullptr, *Ui = nullptr, *dUi = nullptr, *kn = nullptr;
  int *eni = nullptr, *md = nullptr, *sd = nullptr;
  std::vector<int> dNe;
  std::vector<std::vector<int>> sid_h;
  std::vector<int*> sid;
  std::vector<int*> rowmap;  // per domain, device: row 4(t-1)+tn -> live index or -1
  std::vector<int*> live;    // per domain, device: live index -> row
  std::vector<int> nlive;
  std::vector<double*> G, H, dG, dH, tdF;
  std::vector<double> kn_h;
  bq_qplan plan;
  bq_stats st{};
  int two_nn = 0;
  // the wave data the resident matrices were assembled for (k, vk, lattice, kn); the
  // matrices and factors stay valid while a new incidence leaves these unchanged
  bool assembled = false;
  std::vector<double> key_asm;
  std::vector<double> key_now;
  int* qerr = nullptr;       // QPGF error word of the assembly launches (device)
  double* wtab = nullptr;    // Faddeeva-series weights of the plan (device, BQ_WTAB_MAXV rows)
  std::vector<double> wtab_h;  // the same on the host, kept while its copy may be in flight
  void* work = nullptr;      // assembly work buffer
  size_t work_bytes = 0;
  bq_reduced red;
  size_t bytes = 0;          // device bytes of the matrices and work buffers
  size_t tmp_bytes = 0;      // peak of the temporaries a block-system solve allocated
};

namespace bq {
// single object: MN == 1 and every element has main domain 0 and sub-domain 1
bool single_object(const bq_sess* S);
// the reduction's factors for the resident matrices (prec 0 FP64, 1 FP32)
int reduced_factor(bq_sess* S, int prec);
// X[8 NN] complex in the reference's order from the resident factors: FP64 direct (prec 0)
// or FP32 preconditioned FP64 GMRES (prec 1)
int reduced_solve(bq_sess* S, double* Xh);
// dEH[N][6] = G_d^{-1} (H_d EH) with the reduction's factors of G_d
int reduced_deh(bq_sess* S, int d, const double* EH, double* dEH);
void reduced_free(bq_sess* S);
}  // namespace bq
