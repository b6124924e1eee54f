/*
 * bq_amd.h -- entry points of libbem3qd2 beyond the reference's own API.
 *
 * Plain C ABI: pointers and sizes, complex numbers as interleaved (re, im) doubles.
 * Every function returns 0 on success and a negative code on failure; bq_last_error()
 * then describes the failure.  None of these has a CPU fallback: without a usable gfx950
 * device they return -1.
 */
#ifndef BQ_AMD_H
#define BQ_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

const char *bq_last_error(void);

/* 0 if a gfx950 device is usable by this process. */
int bq_device_ok(void);

/* w(z) for n points; z and w are [n][2] (re, im).  Batched form of the reference's
 * Faddeeva_w (qgf2_src/Faddeeva.h:39, Faddeeva.cc:679), relative accuracy ~1e-15. */
int bq_faddeeva_w_batch(int n, const double *z, double *w);

/* erfcx(x) = exp(x^2) erfc(x) for n real points, out[n], through the table form that the
 * assembly's wavefront-uniform lattice sums use in place of the device library's erfcx
 * (piecewise polynomials in 4/(4+x), generated by csrc/dev/gen_erfcx_table.py; relative
 * accuracy ~1e-15).  The reference has no separate entry point: its Ewald sums call
 * Faddeeva::erfcx (qgf2_src/Faddeeva.cc), whose values this reproduces. */
int bq_erfcx_batch(int n, const double *x, double *out);

/* Quasi-periodic Green's function value and gradient for n relative positions r[n][3].
 * qd8 = the reference's QPD2 {k, vk[3], vd1[2], vd2[2]} as 8 doubles.
 * out[n][8] = qG, dqG/dx, dqG/dy, dqG/dz (complex).  Batched form of d3hm_qpgf_d2_v1
 * (qgf2_src/d3_qpgf_d2.h:101-112, d3_qpgf_d2.c:124). */
int bq_qpgf_v1_batch(const double *qd8, int n, const double *r, double *out);

/* Same with a choice of the reference variant it reproduces at r_z == 0, where the
 * reference's variants differ: variant 0 = d3hm_qpgf_d2_v1 / _dqG (their spectral sum
 * stops early there, d3_qpgf_d2.c:1840-1963), variant 1 = d3hm_qpgf_d2_qG (converged,
 * d3_qpgf_d2.c:1190-1335).  Elsewhere both give the converged sum.  variant 2 =
 * d3hm_qpgf_d2_v2 (d3_qpgf_d2.c:178-230): out[n][20] adds d2qG xx, yy, zz, xy, yz, zx.
 * Returns the reference's code -12 (Fourier-domain sum) or -23 (Ewald sum) when a reciprocal
 * order has beta_n == 0 exactly (qgf2_src/d3_qpgf_d2.h:113-123); out then holds finite values
 * without that order's term. */
int bq_qpgf_batch(const double *qd8, int variant, int n, const double *r, double *out);

/* bq_qpgf_batch without any device allocation (no counterpart in the reference: for callers
 * that cannot allocate, such as code under HIP graph capture or a serving loop that forbids
 * hipMalloc).  d_r [n][3] and d_out [n][8] (variant 2: [n][20]) are device buffers of the
 * caller; the work is enqueued on stream (a hipStream_t, NULL for the default stream), which
 * is synchronized before returning.  The only scratch is one error word per chunk of 2^20
 * points in a fixed 256-byte device array of the library: n is limited to 64 x 2^20 points
 * (larger n returns -2 without launching).  Every call on a device shares that array, so calls
 * that overlap on one device race.  Returns 0, -12 / -23 as bq_qpgf_batch, or < 0 on error. */
int bq_qpgf_batch_noalloc(const double *qd8, int variant, int n, const double *d_r, double *d_out, void *stream);

/* sizeof(DQD2) as compiled into the library (for callers that allocate it opaquely). */
int bq_sizeof_dqd2(void);
/* Sizes of a set-up model for callers that hold DQD2 opaquely: elements, element nodes and,
 * if dNe is not NULL, each domain's element count (dNe[MN + 1]); returns MN. */
int bq_model_dims(void *qd, int *Ne, int *NN, int *dNe);

/* Total fields at n points of a solved model in one GPU launch (batched EH_t of
 * d3qd2_src/d3qd2_field.c:271).  qd: a DQD2* after solve_bieq or dat_read; r[n][3];
 * E, H: [n][3] complex out; did[n] out (domain id, -1 if none).  type 0..3 as EH_t. */
int bq_EH_t_batch(void *qd, int n, const double *r, int type, double *E, double *H, int *did);

/* Domain ids of n points in one GPU launch (batched domain_id / q0_domain_id,
 * d3qd2_src/d3qd2_setup.c:157-245: on a lattice the point is folded into the unit cell first);
 * did[n] receives the id, -1 outside every domain.  The field entry points use it. */
int bq_domain_id_batch(void *qd, int n, const double *r, int *did);

/* Total fields from the modified potential at n points in one GPU launch (batched
 * EH_mEMP_t of d3qd2_src/d3qd2_field.c:175-190).  E, H: [n][3] complex out; ret[n] = domain id,
 * -1 outside every domain, or -did (-99 for the open region) with zero fields for a point too
 * close to the boundary.  Returns -1 if type is outside 0..3. */
int bq_EH_mEMP_t_batch(void *qd, int n, const double *r, int type, double *E, double *H, int *ret);

/* Total fields at n boundary points in one GPU launch (batched EH_t_bd of
 * d3qd2_src/d3qd2_field.c:343-391): point p is (zt[p], et[p]) on the t[p]-th element
 * (1-based) of domain did[p].  E, H: [n][3] complex out.  Returns -1 without evaluating if a
 * query names no element or type is outside 0..3. */
int bq_EH_t_bd_batch(void *qd, int n, const int *did, const int *t, const double *zt, const double *et, int type,
                     double *E, double *H);

/* Device assembly of one domain's coefficient matrices for a set-up model (DQD2* after
 * initialize_dqd2 or dat_read), copied to host buffers (NULL to skip): G, H [N][N],
 * dG, dH [2N][N] complex (rows 2 row + {zeta, eta}), tdF [N][3]; N = 4 x elements of the
 * domain.  Row/column layout of the reference's tmpG/tmpH files
 * (d3qd2_src/d3qd2_solve_bieq.c:212-234). */
int bq_assemble_domain_host(void *qd, int d, double *G, double *H, double *dG, double *dH, double *tdF);

/* Counters and phase times (ms, device clock) of the last solve_bieq in this process:
 * returns a pointer to {double t_assemble_ms, t_solve_ms, t_fields_ms; int n_gh_pairs;
 * long long n_qpgf, n_qpgf_gl, n_qpgf_gh, n_qpgf_sing; int mixed, ir_iterations;
 * double ir_correction; int path, reused; double backward_error, t_factor_ms;
 * int gmres_iterations, qpgf_error} (QPGF evaluations of the periodic domain, in total and by
 * phase: regular pairs with their distance-chosen Gauss rule, 64-point fallback rule, singular
 * rules; whether the global system took the mixed-precision path, its refinement steps and
 * the last correction max|dx| / max|x|; the solve path (0 block elimination of order 2 NN,
 * 1 single-object reduction of order NN); whether the matrices and factors were reused;
 * the single-object path's normwise backward error of the unreduced system; the
 * factorisation time; FGMRES steps; t_fields_ms = boundary E, H, dE, dH (host clock)). */
const void *bq_solve_last_stats(void);

/* Replace the incident plane wave of a set-up model (DQD2* after initialize_dqd2 or
 * dat_read) by one wave: vacuum wavelength lambda0, direction theta, phi (rad), polarisation
 * (e0x, e0y) in the wave's frame (complex, as in ipw.txt), power; the medium of the open
 * region is kept.  Recomputes what the reference derives from the wave in initialize_dqd2
 * (d3qd2_src/d3qd2_setup.c:121-128, 1309-1349): QPD2 (k, Bloch vector), every domain's wave
 * number and the incident field Ui, dUi at the element nodes.  The next solve_bieq reuses the
 * model's resident device session; when only the polarisation or the power changed it also
 * reuses the coefficient matrices and their factors.  power = 0 means amplitude 1
 * (power = n0 / 2, reference d3qd2_src/ipw.c:131).  Returns 0, or -1 for invalid arguments
 * (non-positive lambda0, negative power, zero polarisation, or on a lattice a direction with
 * cos theta = 0 or |cos theta| = 1, which the reference rejects, d3qd2_setup.c:1083-1104).
 * No counterpart in the reference, which reads the wave from ipw.txt once. */
int bq_set_incidence(void *qd, double lambda0, double theta, double phi, double e0x_re, double e0x_im,
                     double e0y_re, double e0y_im, double power);

/* solve_bieq without its progress lines and without exit(): returns 0, or a negative code
 * with the reason in bq_last_error() (the reference's QPGF codes -12 / -23 when a reciprocal
 * order has beta_n == 0, -3 for a singular system, -1 for a device failure).  For request loops
 * that must survive a failed request; solve_bieq itself prints and exits as the reference does
 * (d3qd2_src/q0_bil_dcoef_v2_cc.c:177-180). */
int bq_solve_bieq_rc(void *qd);

/* Device memory of a model's resident solver session is released by finalize_dqd2 or here
 * (no counterpart in the reference; for callers that keep many models). */
void bq_release_solver(void *qd);
/* Device bytes held by the resident solver session of qd (all sessions if qd is NULL), plus
 * the peak of the temporaries its block-system solves allocate (models with more than one
 * bounded domain): the memory a solve of the model needs at once. */
long long bq_solver_bytes(void *qd);
/* Statistics of qd's last solve_bieq (the bq_solve_last_stats layout) into out; -1 if qd
 * has no session. */
int bq_solve_stats(void *qd, void *out);

/* ---- request loop (no counterpart in the reference; its d3qd2_bv_solver solves one wave per
 * run, d3qd2_src/d3qd2_bv_solver.c:3-21) ------------------------------------------------
 * An engine keeps models resident on the GPU and serves incidence requests with continuous
 * batching: nslots worker slots (one host thread each, each model session on its own stream)
 * take ready requests as they arrive, so that work of different requests overlaps on the
 * device.  A slot prefers a request whose mesh and wave are resident in it (matrices and
 * factors reused: a polarisation change); a request is admitted when the device memory of its
 * session fits mem_limit_gb next to the sessions other slots hold (idle ones are released,
 * least recently used first).  Meshes are given as set-up files (dat_write of a model after
 * initialize_dqd2).  precision: the bq_solve_precision mode of every solve. */
typedef struct bq_serve_req {
  int mesh;                      /* index of the set-up file */
  double lambda0, theta, phi;    /* wave (um, rad) */
  int pol;                       /* 1: TM (e0x = 1), 0: TE (e0y = 1); amplitude 1 */
  double arrival_s;              /* seconds after the start of bq_serve_run */
} bq_serve_req;

typedef struct bq_serve_out {
  int status;                    /* 0, or the request's negative code */
  int slot;                      /* worker slot that served it */
  int reused;                    /* matrices and factors reused from the slot's previous request */
  int gmres_iterations;
  double t_arrival_s, t_start_s, t_done_s; /* seconds after the start of the run */
  double backward_error;
  long long n_qpgf_gl;           /* QPGF evaluations of the regular periodic pairs (0 if reused) */
  double asm_cut;                /* lattice-sum cut of its assembly */
} bq_serve_out;

void *bq_serve_create(int nmesh, const char *const *setup_dat, int nslots, double mem_limit_gb, int precision);
/* Serves nreq requests (arrivals released at req[i].arrival_s); E, H [nreq][nprobe][3] complex
 * receive EH_t (type 1) at the probes.  Returns 0, -5 if a request failed (its out.status), -4
 * if a mesh cannot fit the memory limit alone, -2 for bad arguments. */
int bq_serve_run(void *engine, int nreq, const bq_serve_req *req, int nprobe, const double *probes, double *E,
                 double *H, bq_serve_out *out);
void bq_serve_destroy(void *engine);

/* Precision of the global (open-region) linear solve (no counterpart in the reference,
 * whose solver is PARDISO / LAPACKE_zgesv in FP64, d3qd2_src/d3qd2_mkl_solver.c:18-148):
 * mode 0 = complex FP64 LU (default); mode 1 = complex FP32 LU of the same blocks used as
 * the correction solver of iterative refinement with FP64 residuals, which stops when the
 * correction is below 1e-15 of the solution or stops shrinking (the rounding floor of the
 * FP64 residual) below 1e-10, and falls back to the FP64 LU when it stalls above that;
 * mode 2 = mode 1 when the block system's order 2 NN is at least 16384, else mode 0 (smaller
 * LUs are latency bound and gain nothing from FP32).  Any other mode only queries.  Returns
 * the mode in force.  The environment variable BQ_MIXED=1 or 2 sets the initial mode.
 * Process-wide; set it between solves. */
int bq_solve_precision(int mode);

/* Accuracy of the coefficient-matrix assembly (no counterpart in the reference, whose
 * lattice sums and Gauss rules aim at full double precision): tol = 0 (default) keeps the
 * reference's accuracy (lattice sums to exp(-40) of their leading term, distance-chosen
 * Gauss orders bounded at 1e-15 of the integrand); 0 < tol < 1e-3 assembles each incidence's
 * matrices to about tol of their largest entries: the Gauss-order bound becomes tol and the
 * lattice-sum cut is tuned per incidence by evaluating the QPGF at 4096 sampled element
 * offsets with shorter sums against the full one (the smallest cut whose error stays below
 * tol / 8).  For the mixed-precision path, whose result is held to 1e-8 of the FP64 path;
 * tol < 0 only queries.  BQ_ASM_TOL sets the initial value.  Process-wide; set between
 * solves.  Returns the value in force. */
double bq_assembly_tolerance(double tol);

/* Update arithmetic of the complex FP32 factorisation (the preconditioner of the
 * mixed-precision solve; the reference's PARDISO call, d3qd2_src/d3qd2_mkl_solver.c:18-148,
 * factors in FP64 only): 1 = bf16x3 products on the matrix cores (FP32 operands split into
 * bf16 hi/lo pairs, three products per real product, FP32 accumulation; relative error ~2^-16
 * per term), 0 = FP32 products (rocBLAS cgemm).  mode < 0 only queries.  BQ_LU_UPDATE=fp32
 * sets 0 at start-up.  Returns the mode in force before the call. */
int bq_lu_update_mode(int mode);
/* Look-ahead of the dense factorisations (FP64 and FP32; no counterpart in the reference, whose
 * PARDISO call, d3qd2_src/d3qd2_mkl_solver.c:18-148, schedules its own factorisation):
 * 1 (default) = for order > 1024 the columns right of each 512-column panel are updated in
 * two parts, the next panel's first and the rest on a second stream per calling stream,
 * concurrently with the next panel's pivot steps; 0 = one stream.  The same algorithm either
 * way; the products may be tiled differently, so results agree to rounding.  mode < 0 only
 * queries.  BQ_LU_LOOKAHEAD=0 sets 0 at start-up.
 * Returns the mode in force before the call. */
int bq_lu_lookahead(int mode);
/* C -= A B, complex FP32, column-major device matrices (A m x k, B k x n, C m x n), with the
 * bf16x3 products above; the building block of the FP32 factorisation's rank-512 updates and
 * many-right-hand-side solves.  stream: a hipStream_t or NULL.  Returns 0 or -1. */
int bq_cgemm_split_dev(void *stream, int m, int n, int k, const float *A, long lda, const float *B, long ldb,
                       float *C, long ldc);
/* C = -A B on the same kernel, C not read; B may be C itself (in place: C <- -A C), since the
 * operands are packed before the product is written.  The solves apply inverted 512-row
 * diagonal blocks of the FP32 factors this way. */
int bq_cgemm_split_store_dev(void *stream, int m, int n, int k, const float *A, long lda, const float *B, long ldb,
                             float *C, long ldc);

/* Per-kernel timing (no counterpart in the reference; for benchmarks and profiling):
 * bq_kernel_timing(1) turns on HIP-event timing of every kernel launch the library makes
 * (on the stream it launches on) and clears the totals; bq_kernel_timing(0) turns it off.
 * bq_kernel_times copies up to max entries: names (48 chars each), total_ms and launch
 * count per kernel name; returns the number of entries. */
int bq_kernel_timing(int on);
int bq_kernel_times(int max, char *names, double *total_ms, long long *count);

#ifdef __cplusplus
}
#endif
#endif
