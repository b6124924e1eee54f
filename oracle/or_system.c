/*
 * or_system.c -- domain matrices, global system, solution scatter and boundary E/H.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Follows d3qd2_src/d3qd2_solve_bieq.c: create_cmatrix_domain :165-256,
 * create_matrix_csr_dac :284-421 (written dense here), tmatrix_bd_store :456-544,
 * q1_solve_eh_bv + dudt_fcm + inv_matrix_33 :931-1030, :617-660, and the right-hand side
 * of solve_deh_bv :662-791.  The dense complex solves themselves are done by the caller
 * with LAPACK (numpy/scipy), standing in for MKL PARDISO / LAPACKE_zgesv.
 */
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <complex.h>
#include "oracle.h"

static int is_tri(const or_model *m, int s) { return m->ed[4 * abs(s) + 3] == 0; }

void or_assemble_rows(const or_model *m, int did, int t0, int t1, double complex *tG,
                      double complex *tH, double complex *tdG, double complex *tdH, double *tdF)
{
  int Ne = m->dNe[did];
  size_t N = 4 * (size_t)Ne;
#pragma omp parallel for schedule(dynamic)
  for (int t = t0; t <= t1; t++) {
    int td = m->sid[did][t];
    int tl = is_tri(m, td) ? 3 : 4;
    for (int tn = 0; tn < tl; tn++) {
      size_t row = 4 * (size_t)(t - 1) + tn;
      double F = 0, dFz = 0, dFe = 0;
      for (int s = 1; s <= Ne; s++) {
        int sd = m->sid[did][s];
        double complex CC[9], dCz[9], dCe[9];
        or_pair(m, did, td, tn, sd, 2, CC, dCz, dCe);
        for (int i = 0; i < 4; i++) {
          size_t col = 4 * (size_t)(s - 1) + i;
          tG[row * N + col] = CC[i];
          tH[row * N + col] = CC[i + 4];
          if (tdG) {
            tdG[row * 2 * N + col] = dCz[i];
            tdG[row * 2 * N + N + col] = dCe[i];
            tdH[row * 2 * N + col] = dCz[i + 4];
            tdH[row * 2 * N + N + col] = dCe[i + 4];
          }
        }
        F += creal(CC[8]);
        dFz += creal(dCz[8]);
        dFe += creal(dCe[8]);
      }
      double C = (did == 0) ? 1.0 + F : F;
      tH[row * N + row] += C;
      if (tdF) {
        tdF[3 * row + 0] = C;
        tdF[3 * row + 1] = dFz;
        tdF[3 * row + 2] = dFe;
      }
    }
  }
}

static void unit_normal(const or_model *m, int s, int l, double *n)
{
  const double *w = m->wen + 12 * abs(s) + 3 * l;
  double a = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  for (int d = 0; d < 3; d++) n[d] = w[d] / a;
}

void or_global_system(const or_model *m, double complex **Gs, double complex **Hs,
                      double complex *A, double complex *B)
{
  size_t nn = m->NN, na = 8 * nn, row = 0;
  for (int did = 0; did <= m->MN; did++) {
    int Ne = m->dNe[did];
    size_t N = 4 * (size_t)Ne;
    for (int cid = 0; cid < 4; cid++) {
      for (int t = 1; t <= Ne; t++) {
        int td = m->sid[did][t];
        for (int tn = 0; tn < 4; tn++) {
          if (tn == 3 && is_tri(m, td)) continue;
          double complex *a = A + row * na, b = 0.0;
          memset(a, 0, sizeof(double complex) * na);
          size_t gr = 4 * (size_t)(t - 1) + tn;
          const double complex *gG = Gs[did] + gr * N, *gH = Hs[did] + gr * N;
          for (int s = 1; s <= Ne; s++) {
            int sd = m->sid[did][s], asd = abs(sd);
            int mdid = m->md[asd], sdid = m->sd[asd], nl = is_tri(m, sd) ? 3 : 4;
            for (int l = 0; l < nl; l++) {
              size_t e = (size_t)m->eni[4 * asd + l];
              double complex g = gG[4 * (s - 1) + l], h = gH[4 * (s - 1) + l];
              if (did == mdid) {
                a[nn * cid + e] = h;
                a[nn * (cid + 4) + e] = -g;
              } else {
                double complex k1 = m->kn[mdid], k2 = m->kn[sdid];
                double complex delt = k1 * k1 - k2 * k2, beps = k1 * k1 / (k2 * k2), sigm = 1.0 - beps;
                double vn[3];
                unit_normal(m, sd, l, vn);
                const double complex *ui = m->Ui + 16 * asd + 4 * l, *dui = m->dUi + 16 * asd + 4 * l;
                if (cid < 3) {
                  a[nn * cid + e] = h;
                  a[nn * 3 + e] = delt * vn[cid] * g;
                  a[nn * (cid + 4) + e] = g;
                  if (mdid == 0) b += -h * ui[cid] - g * dui[cid];
                } else {
                  a[nn * 0 + e] = sigm * vn[0] * g;
                  a[nn * 1 + e] = sigm * vn[1] * g;
                  a[nn * 2 + e] = sigm * vn[2] * g;
                  a[nn * 3 + e] = h;
                  a[nn * 7 + e] = beps * g;
                  if (mdid == 0) b += -sigm * g * (vn[0] * ui[0] + vn[1] * ui[1] + vn[2] * ui[2]);
                }
              }
            }
          }
          B[row] = b;
          row++;
        }
      }
    }
  }
}

void or_store_solution(const or_model *m, const double complex *X, double complex **U, double complex **dU)
{
  size_t nn = m->NN;
  for (int d = 0; d <= m->MN; d++) {
    for (int s = 1; s <= m->dNe[d]; s++) {
      int sd = m->sid[d][s], asd = abs(sd), mdid = m->md[asd], sdid = m->sd[asd];
      int nl = is_tri(m, sd) ? 3 : 4;
      double complex *u = U[d] + 16 * s, *du = dU[d] + 16 * s;
      for (int l = 0; l < nl; l++) {
        size_t e = (size_t)m->eni[4 * asd + l];
        if (d == mdid) {
          for (int i = 0; i < 4; i++) {
            u[4 * l + i] = X[nn * i + e];
            du[4 * l + i] = X[nn * (i + 4) + e];
          }
        } else {
          double complex k1 = m->kn[mdid], k2 = m->kn[sdid];
          double complex delt = k1 * k1 - k2 * k2, beps = k1 * k1 / (k2 * k2), sigm = 1.0 - beps;
          double n[3];
          unit_normal(m, asd, l, n);
          double complex tU[4], tUi[3] = {0, 0, 0};
          tU[3] = X[nn * 3 + e];
          for (int i = 0; i < 3; i++) {
            tU[i] = X[nn * i + e];
            u[4 * l + i] = tU[i];
            du[4 * l + i] = -(X[nn * (i + 4) + e] + delt * tU[3] * n[i]);
            if (mdid == 0) {
              tUi[i] = m->Ui[16 * asd + 4 * l + i];
              u[4 * l + i] += tUi[i];
              du[4 * l + i] += -m->dUi[16 * asd + 4 * l + i];
            }
          }
          u[4 * l + 3] = tU[3];
          du[4 * l + 3] = -(beps * X[nn * 7 + e] + sigm * (n[0] * tU[0] + n[1] * tU[1] + n[2] * tU[2]));
          if (mdid == 0) du[4 * l + 3] += -sigm * (n[0] * tUi[0] + n[1] * tUi[1] + n[2] * tUi[2]);
        }
      }
    }
  }
}

static void inv33(double *a, const double *vn, const double *vtz, const double *vte)
{
  double idet = 1.0 / (vn[0] * vtz[1] * vte[2] + vn[1] * vtz[2] * vte[0] + vn[2] * vtz[0] * vte[1] -
                       (vn[2] * vtz[1] * vte[0] + vn[1] * vtz[0] * vte[2] + vn[0] * vtz[2] * vte[1]));
  a[0] = vtz[1] * vte[2] - vtz[2] * vte[1];
  a[1] = -vn[1] * vte[2] + vn[2] * vte[1];
  a[2] = vn[1] * vtz[2] - vn[2] * vtz[1];
  a[3] = -vtz[0] * vte[2] + vtz[2] * vte[0];
  a[4] = vn[0] * vte[2] - vn[2] * vte[0];
  a[5] = -vn[0] * vtz[2] + vn[2] * vtz[0];
  a[6] = vtz[0] * vte[1] - vtz[1] * vte[0];
  a[7] = -vn[0] * vte[1] + vn[1] * vte[0];
  a[8] = vn[0] * vtz[1] - vn[1] * vtz[0];
  for (int i = 0; i < 9; i++) a[i] *= idet;
}

/* E = -grad(phi) + A, H = (lambda0 / 2 pi i) curl A from U, dU/dn and the tangential
 * derivatives; gradient components recovered through the inverse of [n, tz, te]. */
static void eh_from(const double *a, const double complex *U, const double complex *dUn,
                    const double complex *dUz, const double complex *dUe, double complex ch,
                    double complex *E, double complex *H)
{
  for (int i = 0; i < 3; i++) E[i] = -(dUn[3] * a[3 * i] + dUz[3] * a[3 * i + 1] + dUe[3] * a[3 * i + 2]) + U[i];
#define GR(c, i) (dUn[c] * a[3 * (i)] + dUz[c] * a[3 * (i) + 1] + dUe[c] * a[3 * (i) + 2])
  H[0] = ch * (GR(2, 1) - GR(1, 2));
  H[1] = ch * (GR(0, 2) - GR(2, 0));
  H[2] = ch * (GR(1, 0) - GR(0, 1));
#undef GR
}

void or_boundary_EH(const or_model *m, double complex **U, double complex **dU,
                    double complex **tdG, double complex **tdH, double **tdF,
                    double complex **E, double complex **H)
{
  double complex ch = m->lambda0 / (2.0 * M_PI * I);
  for (int d = 1; d <= m->MN; d++) {
    int Ne = m->dNe[d];
    size_t N = 4 * (size_t)Ne;
    for (int t = 1; t <= Ne; t++) {
      int at = m->sid[d][t], atd = abs(at);
      double sig = (at < 0) ? -1.0 : 1.0;
      for (int tn = 0; tn < 4; tn++) {
        if (tn == 3 && is_tri(m, atd)) continue;
        size_t row = 4 * (size_t)(t - 1) + tn;
        double vtz[3], vte[3], vn[3], a[9];
        or_node_tangents(m, atd, tn, vtz, vte);
        const double *w = m->wen + 12 * atd + 3 * tn;
        double aw = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
        for (int i = 0; i < 3; i++) vn[i] = sig * w[i] / aw;
        inv33(a, vn, vtz, vte);
        /* dudt_fcm */
        double complex dUz[4] = {0}, dUe[4] = {0};
        const double complex *gz = tdG[d] + row * 2 * N, *ge = gz + N;
        const double complex *hz = tdH[d] + row * 2 * N, *he = hz + N;
        for (int s = 1; s <= Ne; s++)
          for (int sn = 0; sn < 4; sn++)
            for (int i = 0; i < 4; i++) {
              size_t c = 4 * (size_t)(s - 1) + sn;
              double complex du = dU[d][16 * s + 4 * sn + i], u = U[d][16 * s + 4 * sn + i];
              dUz[i] += gz[c] * du - hz[c] * u;
              dUe[i] += ge[c] * du - he[c] * u;
            }
        const double *f = tdF[d] + 3 * row;
        const double complex *Ut = U[d] + 16 * t + 4 * tn, *dUt = dU[d] + 16 * t + 4 * tn;
        for (int i = 0; i < 4; i++) {
          dUz[i] = (dUz[i] - Ut[i] * f[1]) / f[0];
          dUe[i] = (dUe[i] - Ut[i] * f[2]) / f[0];
        }
        eh_from(a, Ut, dUt, dUz, dUe, ch, E[d] + 12 * t + 3 * tn, H[d] + 12 * t + 3 * tn);
        if (m->md[atd] == 0) {
          double complex e3[3], h3[3], dUiz[3], dUie[3], dh[3];
          const double *rn = m->ren + 12 * atd + 3 * tn;
          or_ipw_EH_dv(m, rn, vtz, e3, h3, dUiz, dh);
          or_ipw_EH_dv(m, rn, vte, e3, h3, dUie, dh);
          int s;
          for (s = 1; s <= m->dNe[0]; s++) if (m->sid[0][s] == atd) break;
          const double complex *U0 = U[0] + 16 * s + 4 * tn, *dU0 = dU[0] + 16 * s + 4 * tn;
          double complex mdn[4], z2[4], e2[4];
          for (int i = 0; i < 4; i++) {
            mdn[i] = -dU0[i];
            z2[i] = (i < 3) ? dUz[i] - dUiz[i] : dUz[i];
            e2[i] = (i < 3) ? dUe[i] - dUie[i] : dUe[i];
          }
          eh_from(a, U0, mdn, z2, e2, ch, E[0] + 12 * s + 3 * tn, H[0] + 12 * s + 3 * tn);
          /* E uses the undifferenced phi derivatives (d3qd2_solve_bieq.c:1011-1012) */
          for (int i = 0; i < 3; i++)
            E[0][12 * s + 3 * tn + i] = -(-dU0[3] * a[3 * i] + dUz[3] * a[3 * i + 1] + dUe[3] * a[3 * i + 2]) + U0[i];
        }
      }
    }
  }
}

void or_deh_rhs(const or_model *m, int d, const double complex *tH, double complex *E,
                double complex *H, double complex *B)
{
  int Ne = m->dNe[d];
  size_t N = 4 * (size_t)Ne;
  for (size_t row = 0; row < N; row++) {
    for (int i = 0; i < 6; i++) B[6 * row + i] = 0.0;
    for (int s = 1; s <= Ne; s++)
      for (int sn = 0; sn < 4; sn++) {
        double complex h = tH[row * N + 4 * (s - 1) + sn];
        for (int i = 0; i < 3; i++) {
          B[6 * row + i] += h * E[12 * s + 3 * sn + i];
          B[6 * row + 3 + i] += h * H[12 * s + 3 * sn + i];
        }
      }
  }
}
