/*
 * bq_solver.h -- device solver session used by solve_bieq (internal, C ABI).
 *
 * The host flattens the model (element constants, nodes, sub-domain lists, incident data)
 * into bq_model; the session uploads it once, assembles every domain's matrices on the
 * device, solves the global system and provides the two dense products the boundary-field
 * step needs.  All complex arrays are interleaved (re, im) doubles.
 */
#ifndef BQ_SOLVER_H
#define BQ_SOLVER_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct bq_model {
  int Ne, NN, MN, ps;
  double qd[8];          /* QPD2: k, vk[3], vd1[2], vd2[2] */
  const double *kn;      /* [MN+1] complex */
  const double *E;       /* [(Ne+1)][24] element constants, bq_pack_elem layout */
  const double *ren;     /* [(Ne+1)][4][3] */
  const double *nrm;     /* [(Ne+1)][4][3] unit normals at element nodes (main side) */
  const int *eni;        /* [(Ne+1)][4] */
  const int *md, *sd;    /* [Ne+1] */
  const double *Ui, *dUi; /* [(Ne+1)][4][4] complex */
  const int *dNe;        /* [MN+1] */
  const int *const *sid; /* [MN+1][dNe+1] */
  const double *xli, *wli, *xhi, *whi; /* Gauss-Legendre 12 and 64 on [-1, 1] */
} bq_model;

typedef struct bq_stats {
  double t_assemble_ms, t_solve_ms, t_fields_ms;
  int n_gh_pairs;        /* pairs that needed the 64x64 rule, all domains */
  long long n_qpgf;      /* QPGF evaluations in the assembly (algorithmic count) */
  long long n_qpgf_gl, n_qpgf_gh, n_qpgf_sing; /* the same split by kernel phase */
  int mixed;             /* 1: the global system was solved by FP32 factors + FP64 refinement */
  int ir_iterations;     /* refinement steps taken */
  double ir_correction;  /* last correction, max |dx| / max |x| (block elimination, mixed) */
  int path;              /* 0: block elimination (two LUs of order 2 NN); 1: single-object
                            reduction (LUs of order NN) */
  int reused;            /* 1: matrices and factors of the previous solve were reused (same
                            wavelength, direction and lattice: only the polarisation or
                            amplitude changed) */
  double backward_error; /* single-object reduction: normwise backward error of the unreduced
                            system, max|b - A x| / (||A||_inf ||x||_inf + ||b||_inf) */
  double t_factor_ms;    /* of t_solve_ms: the factorisation (0 when reused) */
  int gmres_iterations;  /* FGMRES steps of the mixed-precision single-object solve */
  int qpgf_error;        /* 0, or the reference's negative QPGF code met by the assembly */
  double asm_tol;        /* assembly tolerance in force (0: the reference's accuracy) */
  double asm_cut;        /* lattice-sum cut the assembly took (terms below exp(-cut) dropped) */
} bq_stats;

typedef struct bq_sess bq_sess;

/* Element constants in the device layout (24 doubles): cr[3][4], cw[3][3] (dimension-major
 * as the reference's cr[s][d][c], cw[s][d][c]), [21] = shape (0 bilinear quad, 1 linear
 * triangle), 2 pad.  A triangle's cw is stored times 2/(3 sqrt3), which makes |W| the
 * surface Jacobian of the map from the reference triangle, as |W| already is for the
 * bilinear map (reference d3qd2_src/d3qd2_setup.c:1150-1185). */
void bq_pack_elem(double *const *cr, double *const *cw, int tri, double *dst);

bq_sess *bq_sess_create(const bq_model *m);
void bq_sess_destroy(bq_sess *s);
/* New incident wave for a resident session (same mesh and media; m's arrays as flattened for
 * the new wave).  The matrices and factors stay valid when k, the Bloch vector, the lattice
 * and every medium's wave number are unchanged (polarisation or amplitude only).  Returns 1
 * if they stay valid, 0 if the next solve reassembles, < 0 if m is not the same mesh. */
int bq_sess_update(bq_sess *s, const bq_model *m);
/* 1 if the resident matrices belong to the current wave */
int bq_sess_assembled(const bq_sess *s);
/* device bytes held by the session */
size_t bq_sess_bytes(const bq_sess *s);
int bq_sess_assemble(bq_sess *s);
/* X[8 NN] complex: the global unknowns in the reference's order (X[c NN + e]) */
int bq_sess_solve_global(bq_sess *s, double *X);
/* out[2N][4] complex = dG U' - dH U for domain d (d >= 1 or non-periodic), where
 * U'[4(s-1)+sn][i] = dU, U = U of the domain's element nodes */
int bq_sess_dudt(bq_sess *s, int d, const double *U, const double *dU, double *out);
/* dEH[N][6] = G_d^{-1} (H_d EH), EH[N][6] complex */
int bq_sess_deh(bq_sess *s, int d, const double *EH, double *dEH);
/* optional: copy a domain's G, H (row-major N x N) to host */
int bq_sess_get_GH(bq_sess *s, int d, double *G, double *H);
int bq_sess_get_dGH(bq_sess *s, int d, double *dG, double *dH, double *tdF);
const bq_stats *bq_sess_stats(const bq_sess *s);

/* solve_bieq's phase gate (bq_solve.c): fn(phase, 1) before and fn(phase, 0) after the
 * assembly (phase 0) and the factorisation + solve + boundary fields (phase 1); NULL: none */
void bq_set_phase_gate(void (*fn)(int phase, int enter));

#ifdef __cplusplus
}
#endif
#endif
