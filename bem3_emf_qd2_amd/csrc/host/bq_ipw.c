/*
 * bq_ipw.c -- incident plane waves: input file, derived data, fields and directional
 * derivatives at points.
 *
 * Same file format and formulas as the reference (d3qd2_src/m_ipw.c:9-198,
 * d3qd2_src/ipw.c:121-256): E0 = sqrt(2 power / n), k = 2 pi n / lambda0, the (x, y)
 * polarisation rotated by theta about the axis set by phi, H = n k_hat x E.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <complex.h>
#include "bem3_emf_qd2.h"

static void die(const char *what, const char *fn)
{
  printf("read_data_mipw(), failed to read %s in %s. Exit...\n", what, fn);
  exit(1);
}

void read_data_mipw(char *fname, MIPW *obj)
{
  FILE *fp = fopen(fname, "rt");
  char buf[256];
  if (!fp) {
    printf("read_data_mipw(), failed to read %s. Exit...\n", fname);
    exit(1);
  }
  memset(obj->fn_ipw, 0, sizeof obj->fn_ipw);
  strncpy(obj->fn_ipw, fname, sizeof obj->fn_ipw - 1);
  if (!fgets(buf, 256, fp) || !fgets(buf, 256, fp)) die("the header", fname);
  int nn;
  if (fscanf(fp, "%d\n", &nn) != 1) die("the wave count", fname);
  if (!fgets(buf, 256, fp)) die("the header", fname);
  if (nn <= 0) {
    printf("read_data_mipw(), number of plane waves must be positive. Exit...\n");
    exit(1);
  }
  obj->n_ipw = nn;
  obj->bd.ipw = (Ipw *)m_alloc2(nn, sizeof(Ipw), "read_data_mipw(), ipw");
  for (int i = 0; i < nn; i++) {
    double v[12];
    for (int j = 0; j < 12; j++)
      if (fscanf(fp, "%lf", &v[j]) != 1) die("a plane-wave parameter", fname);
    Ipw *p = &obj->bd.ipw[i];
    p->lambda0 = v[0];
    p->ni = v[1];
    p->power = v[2];
    p->e0x = v[3] + I * v[4];
    p->e0y = v[5] + I * v[6];
    p->fx = v[7];
    p->fy = v[8];
    p->fz = v[9];
    p->theta = v[10];
    p->phi = v[11];
  }
  fclose(fp);
}

static void print_one(const Ipw *p)
{
  printf("-- plane wave parameter --\n");
  printf("wave length of incident beam in vacuum      : %15.14g\n", p->lambda0);
  printf("refractive index of surrounding             : %15.14g\n", p->ni);
  printf("incident beam power per unit area           : %15.14g\n", p->power);
  printf("x-poralization coefficient                  : %7.6g%+7.6gi\n", creal(p->e0x), cimag(p->e0x));
  printf("y-poralization coefficient                  : %7.6g%+7.6gi\n", creal(p->e0y), cimag(p->e0y));
  printf("x-component of translation vector           : %15.14g\n", p->fx);
  printf("y-component of translation vector           : %15.14g\n", p->fy);
  printf("z-component of translation vector           : %15.14g\n", p->fz);
  printf("rotation parameter theta               [rad]: %15.14g\n", p->theta);
  printf("rotation parameter phi                 [rad]: %15.14g\n", p->phi);
}

void print_data_mipw(MIPW *obj)
{
  for (int i = 0; i < obj->n_ipw; i++) {
    printf(" \"%s\" No.%02d ", obj->fn_ipw, i);
    print_one(&obj->bd.ipw[i]);
  }
}

void print_data_mipw_mksa(MIPW *obj)
{
  /* lengths in um and powers in mW/um^2 map to MKSA by fixed factors (com_src/osu_mksa.h:11-12) */
  for (int i = 0; i < obj->n_ipw; i++) {
    const Ipw *p = &obj->bd.ipw[i];
    printf(" \"%s\" No.%02d -- plane wave parameter --\n", obj->fn_ipw, i);
    printf("wave length of incident beam in vacuum   [m]: %15.14g\n", p->lambda0 * 1.0e-6);
    printf("refractive index of surrounding             : %15.14g\n", p->ni);
    printf("incident beam power per unit area    [W/m^2]: %15.14g\n", p->power * 1.0e-3 / 1.0e-12);
    printf("rotation parameter theta               [rad]: %15.14g\n", p->theta);
    printf("rotation parameter phi                 [rad]: %15.14g\n", p->phi);
  }
}

void check_param_mipw(MIPW *obj)
{
  obj->n_0 = obj->lambda_0 = 0.0;
  for (int i = 0; i < obj->n_ipw; i++) {
    if (i == 0) {
      obj->n_0 = obj->bd.ipw[i].ni;
      obj->lambda_0 = obj->bd.ipw[i].lambda0;
    } else {
      if (obj->n_0 != obj->bd.ipw[i].ni) { printf("check_param_mipw(),parameter error ipw ni, exit\n"); exit(1); }
      if (obj->lambda_0 != obj->bd.ipw[i].lambda0) { printf("check_param_mipw(),parameter error ipw lambda, exit\n"); exit(1); }
    }
  }
  obj->omega = 2.0 * M_PI / obj->lambda_0;
}

static void setup_ipw_one(Ipw *p)
{
  double cs = 1.0 / sqrt(pow(cabs(p->e0x), 2) + pow(cabs(p->e0y), 2));
  double complex ex = p->e0x * cs, ey = p->e0y * cs, hx = -ey * p->ni, hy = ex * p->ni;
  IpwD *d = &p->data;
  d->E0 = sqrt(p->power * 2.0 / p->ni);
  d->ki = 2.0 * M_PI * p->ni / p->lambda0;
  d->sin_t = sin(p->theta);
  d->cos_t = cos(p->theta);
  d->sin_p = sin(p->phi);
  d->cos_p = cos(p->phi);
  double st = d->sin_t, ct = d->cos_t, sp = d->sin_p, cp = d->cos_p;
  double a = sp * sp + cp * cp * ct, b = sp * cp * ct - sp * cp, c = sp * sp * ct + cp * cp;
  d->ex = ex * a + ey * b;
  d->ey = ex * b + ey * c;
  d->ez = -(ex * cp + ey * sp) * st;
  d->hx = hx * a + hy * b;
  d->hy = hx * b + hy * c;
  d->hz = -(hx * cp + hy * sp) * st;
}

void setup_mipw(MIPW *obj)
{
  for (int i = 0; i < obj->n_ipw; i++) setup_ipw_one(&obj->bd.ipw[i]);
}

void free_mipw(MIPW *obj)
{
  free(obj->bd.ipw);
  obj->bd.ipw = NULL;
  obj->n_ipw = 0;
  obj->fn_ipw[0] = '\0';
}

static double complex phase_at(const Ipw *p, const double *x, double *kh)
{
  const IpwD *d = &p->data;
  kh[0] = d->sin_t * d->cos_p;
  kh[1] = d->sin_t * d->sin_p;
  kh[2] = d->cos_t;
  double te = d->ki * ((x[0] - p->fx) * kh[0] + (x[1] - p->fy) * kh[1] + (x[2] - p->fz) * kh[2]);
  return (cos(te) + I * sin(te)) * d->E0;
}

void calc_mipw_EH(double complex *e, double complex *h, double *x, MIPW *obj)
{
  for (int i = 0; i < 3; i++) e[i] = h[i] = 0.0;
  for (int w = 0; w < obj->n_ipw; w++) {
    const Ipw *p = &obj->bd.ipw[w];
    double kh[3];
    double complex f = phase_at(p, x, kh);
    e[0] += p->data.ex * f; e[1] += p->data.ey * f; e[2] += p->data.ez * f;
    h[0] += p->data.hx * f; h[1] += p->data.hy * f; h[2] += p->data.hz * f;
  }
}

void calc_mipw_EH_dEHdv(double complex *e, double complex *h, double complex *dedv, double complex *dhdv, double *x,
                        double *v, MIPW *obj)
{
  for (int i = 0; i < 3; i++) e[i] = h[i] = dedv[i] = dhdv[i] = 0.0;
  for (int w = 0; w < obj->n_ipw; w++) {
    const Ipw *p = &obj->bd.ipw[w];
    double kh[3];
    double complex f = phase_at(p, x, kh);
    double complex g = I * p->data.ki * (kh[0] * v[0] + kh[1] * v[1] + kh[2] * v[2]);
    double complex ev[3] = {p->data.ex * f, p->data.ey * f, p->data.ez * f};
    double complex hv[3] = {p->data.hx * f, p->data.hy * f, p->data.hz * f};
    for (int i = 0; i < 3; i++) {
      e[i] += ev[i];
      h[i] += hv[i];
      dedv[i] += ev[i] * g;
      dhdv[i] += hv[i] * g;
    }
  }
}
