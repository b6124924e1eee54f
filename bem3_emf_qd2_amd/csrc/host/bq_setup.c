/*
 * bq_setup.c -- model input, set-up and result files on the host.
 *
 * Same inputs, data layout and checks as the reference's d3qd2_src/d3qd2_setup.c:
 *   read_dqd2 :3-39 (argv, 5 or 12 arguments), read_periodicity_data :557-589,
 *   read_medium_data :609-650, read_mesh_data :662-959 (Gmsh 2.2 ASCII, paired elements,
 *   open-domain swap), initialize_dqd2 :105-139, check_mipw :1075-1108,
 *   rotation_translation_obj :1110-1137, init_elem_const :1139-1222, sub-domains
 *   :1224-1307, init_boundary_data :1309-1349, domain_id / q0_domain_id :157-245 with the
 *   solid angle of :1351-1422, dat_write / dat_read :247-512 (byte layout kept, so files
 *   interchange with the reference's tools), output_node_particles :514-543.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <complex.h>
#include "bem3_emf_qd2.h"

#define MEPS 1.0e-15
#define GEPS 1.0e-15
#define MSHVER 2.2
#define ELT3 2
#define ELT4 3
#define P44_N 0.5773502691896257645091
#define P44_W 1.0
#define P49_N 0.7745966692414833770359
#define P49_W0 0.3086419753086419753086
#define P49_W1 0.4938271604938271604938
#define P49_W2 0.7901234567901234567901
#define P34_N0 0.2
#define P34_N1 0.3464101615137754587055
#define P34_W0 0.5208333333333333333333
#define P34_W1 0.5625
#define P37_N0 0.3480702390148154917985
#define P37_N1 0.2052130961576726346557
#define P37_N2 0.6028753385763033130539
#define P37_N3 0.3554395089236065568357
#define P37_W0 0.1259391805448271525957
#define P37_W1 0.1323941527885061807376
#define P37_W2 0.225

static void fail(const char *msg, const char *arg)
{
  printf("%s%s. Exit...\n", msg, arg ? arg : "");
  exit(1);
}

/* ---- Gauss-Legendre nodes (Newton on P_n, ascending; com_src/gauleg.c:3-36) ---- */
static void gauss_legendre(double *x, double *w, int n)
{
  int m = (n + 1) / 2;
  for (int i = 1; i <= m; i++) {
    double z = cos(M_PI * (i - 0.25) / (n + 0.5)), z1, pp;
    do {
      double p1 = 1.0, p2 = 0.0;
      for (int j = 1; j <= n; j++) {
        double p3 = p2;
        p2 = p1;
        p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
      }
      pp = n * (z * p1 - p2) / (z * z - 1.0);
      z1 = z;
      z = z1 - p1 / pp;
    } while (fabs(z - z1) > 3.0e-16);
    x[i - 1] = -z;
    x[n - i] = z;
    w[i - 1] = 2.0 / ((1.0 - z * z) * pp * pp);
    w[n - i] = w[i - 1];
  }
  if (n % 2 == 1) x[n / 2] = 0.0;
}

/* ---- allocation of the triple-pointer arrays ---- */
static double **alloc_d2(int n0, int n1)
{
  double **a = m_alloc2(n0, sizeof(double *), "alloc_d2");
  for (int i = 0; i < n0; i++) a[i] = m_alloc2(n1, sizeof(double), "alloc_d2");
  return a;
}
static int **alloc_i2(int n0, int n1)
{
  int **a = m_alloc2(n0, sizeof(int *), "alloc_i2");
  for (int i = 0; i < n0; i++) a[i] = m_alloc2(n1, sizeof(int), "alloc_i2");
  return a;
}
static double ***alloc_d3(int n0, int n1, int n2)
{
  double ***a = m_alloc2(n0, sizeof(double **), "alloc_d3");
  for (int i = 0; i < n0; i++) a[i] = alloc_d2(n1, n2);
  return a;
}
static double complex ***alloc_z3(int n0, int n1, int n2)
{
  double complex ***a = m_alloc2(n0, sizeof(double complex **), "alloc_z3");
  for (int i = 0; i < n0; i++) {
    a[i] = m_alloc2(n1, sizeof(double complex *), "alloc_z3");
    for (int j = 0; j < n1; j++) a[i][j] = m_alloc2(n2, sizeof(double complex), "alloc_z3");
  }
  return a;
}
static void free_2(void **a, int n0)
{
  for (int i = 0; i < n0; i++) free(a[i]);
  free(a);
}
static void free_3(void ***a, int n0, int n1)
{
  for (int i = 0; i < n0; i++) free_2(a[i], n1);
  free(a);
}

static void malloc_node(BOUD *bd) { bd->rn = alloc_d2(bd->Nn + 1, 3); }

static void malloc_elem(BOUD *bd)
{
  int n = bd->Ne + 1;
  bd->ed = alloc_i2(n, 4);
  bd->eni = alloc_i2(n, 4);
  bd->md = m_alloc2(n, sizeof(int), "malloc_elem md");
  bd->sd = m_alloc2(n, sizeof(int), "malloc_elem sd");
  bd->gd = m_alloc2(n, sizeof(int), "malloc_elem gd");
  bd->cr = alloc_d3(n, 3, 4);
  bd->cw = alloc_d3(n, 3, 3);
  bd->ren = alloc_d3(n, 4, 3);
  bd->wen = alloc_d3(n, 4, 3);
  bd->Ui = alloc_z3(n, 4, 4);
  bd->dUi = alloc_z3(n, 4, 4);
}

static void mfree_elem(BOUD *bd)
{
  int n = bd->Ne + 1;
  free_2((void **)bd->ed, n);
  free_2((void **)bd->eni, n);
  free(bd->md);
  free(bd->sd);
  free(bd->gd);
  free_3((void ***)bd->cr, n, 3);
  free_3((void ***)bd->cw, n, 3);
  free_3((void ***)bd->ren, n, 4);
  free_3((void ***)bd->wen, n, 4);
  free_3((void ***)bd->Ui, n, 4);
  free_3((void ***)bd->dUi, n, 4);
  bd->Ne = 0;
}

/* ---- readers ---- */
static void read_periodicity_data(char *fn, DQD2 *qd)
{
  FILE *fp = fopen(fn, "rt");
  char buf[256];
  if (!fp) fail("Can not open the periodicity data file ", fn);
  memset(qd->ped_fn, 0, 128);
  strncpy(qd->ped_fn, fn, 127);
  int ps;
  double a, b, c, d;
  if (!fgets(buf, 256, fp) || !fgets(buf, 256, fp)) fail("read_periodicity_data(), failed to read the header of ", fn);
  if (fscanf(fp, "%d\n", &ps) != 1) fail("read_periodicity_data(), failed to read the periodicity setting in ", fn);
  if (fscanf(fp, "%lf %lf", &a, &b) != 2) fail("read_periodicity_data(), failed to read d1 in ", fn);
  if (fscanf(fp, "%lf %lf", &c, &d) != 2) fail("read_periodicity_data(), failed to read d2 in ", fn);
  fclose(fp);
  qd->bd.ps = ps;
  qd->bd.qd.vd1[0] = a;
  qd->bd.qd.vd1[1] = b;
  qd->bd.qd.vd2[0] = c;
  qd->bd.qd.vd2[1] = d;
}

static void read_medium_data(char *fn, DQD2 *qd)
{
  FILE *fp = fopen(fn, "rt");
  char buf[256];
  if (!fp) fail("Can not open the medium data file ", fn);
  memset(qd->med_fn, 0, 128);
  strncpy(qd->med_fn, fn, 127);
  int mn;
  if (!fgets(buf, 256, fp) || !fgets(buf, 256, fp)) fail("read_medium_data(), failed to read the header of ", fn);
  if (fscanf(fp, "%d\n", &mn) != 1) fail("read_medium_data(), failed to read the number of media in ", fn);
  qd->MN = mn;
  qd->n = m_alloc2(mn + 1, sizeof(double complex), "read_medium_data(), n");
  qd->kn = m_alloc2(mn + 1, sizeof(double complex), "read_medium_data(), kn");
  if (!fgets(buf, 256, fp)) fail("read_medium_data(), failed to read the comment line of ", fn);
  for (int i = 1; i <= mn; i++) {
    double re, im;
    if (fscanf(fp, "%lf", &re) != 1 || fscanf(fp, "%lf", &im) != 1)
      fail("read_medium_data(), failed to read a refractive index in ", fn);
    qd->n[i] = re + I * im;
  }
  fclose(fp);
}

static void read_mesh_data(char *fn, DQD2 *md)
{
  FILE *fp = fopen(fn, "rt");
  char buf[256];
  if (!fp) fail("Can not open the mesh file ", fn);
  memset(md->msh_fn, 0, 128);
  strncpy(md->msh_fn, fn, 127);
  double ver;
  int fmt, prec, n;
  if (!fgets(buf, 256, fp)) fail("read_mesh_data(), failed to read $MeshFormat in ", fn);
  if (fscanf(fp, "%lf", &ver) != 1) fail("read_mesh_data(), failed to read the version in ", fn);
  if (ver < MSHVER) fail("read_mesh_data(), unsupported mesh version (2.2 or later needed) in ", fn);
  if (fscanf(fp, "%d", &fmt) != 1 || fmt != 0) fail("read_mesh_data(), mesh file is not ASCII: ", fn);
  if (fscanf(fp, "%d\n", &prec) != 1 || prec != 8) fail("read_mesh_data(), mesh data size is not 8: ", fn);
  if (!fgets(buf, 256, fp) || !fgets(buf, 256, fp)) fail("read_mesh_data(), failed to read $Nodes in ", fn);
  if (fscanf(fp, "%d\n", &n) != 1) fail("read_mesh_data(), failed to read the node count in ", fn);
  md->bd.Nn = n;
  malloc_node(&md->bd);
  for (int i = 1; i <= n; i++) {
    int id;
    if (fscanf(fp, "%d", &id) != 1) fail("read_mesh_data(), failed to read a node id in ", fn);
    if (id != i) printf("bad id %d\n", id);
    if (fscanf(fp, "%lf %lf %lf", &md->bd.rn[i][0], &md->bd.rn[i][1], &md->bd.rn[i][2]) != 3)
      fail("read_mesh_data(), failed to read a node coordinate in ", fn);
  }
  if (fscanf(fp, "%255s", buf) != 1 || fscanf(fp, "%255s", buf) != 1) fail("read_mesh_data(), failed to read $Elements in ", fn);
  int ne2;
  if (fscanf(fp, "%d", &ne2) != 1) fail("read_mesh_data(), failed to read the element count in ", fn);
  md->bd.Ne = ne2 / 2;
  malloc_elem(&md->bd);
  int nc = 0;
  for (int i = 1; i <= md->bd.Ne; i++) {
    int id, type, ntag, tag, nodes[4] = {0, 0, 0, 0}, dom[2] = {0, 0};
    for (int half = 0; half < 2; half++) {
      if (fscanf(fp, "%d %d %d", &id, &type, &ntag) != 3) fail("read_mesh_data(), failed to read an element in ", fn);
      if (id != 2 * i - 1 + half) {
        printf("read_mesh_data(), element id error: elements must come in paired order. id=%d\n", id);
        exit(1);
      }
      if (type != ELT3 && type != ELT4) {
        printf("read_mesh_data(), element type %d is not supported (2: triangle, 3: quad). Exit...\n", type);
        exit(1);
      }
      for (int j = 0; j < ntag; j++) {
        if (fscanf(fp, "%d", &tag) != 1) fail("read_mesh_data(), failed to read a tag in ", fn);
        if (j == 0) {
          if (tag == OPENDID) tag = 0;
          if (tag > md->MN) {
            printf("read_mesh_data(), domain id %d is larger than the number of media %d. Exit...\n", tag, md->MN);
            exit(1);
          }
          dom[half] = tag;
        } else if (j == 1 && half == 0) md->bd.gd[i] = tag;
      }
      int nv = (type == ELT3) ? 3 : 4;
      int v[4] = {0, 0, 0, 0};
      for (int j = 0; j < nv; j++)
        if (fscanf(fp, "%d", &v[j]) != 1) fail("read_mesh_data(), failed to read element nodes in ", fn);
      if (half == 0) {
        for (int j = 0; j < 4; j++) nodes[j] = v[j];
        for (int j = 0; j < 4; j++) md->bd.ed[i][j] = (j < nv) ? v[j] : 0;
        for (int j = 0; j < 4; j++) md->bd.eni[i][j] = (j < nv) ? nc++ : -1;
      } else {
        /* the paired element lists the same nodes in reverse order */
        int ok = (nv == 3) ? (v[0] == nodes[0] && v[1] == nodes[2] && v[2] == nodes[1])
                           : (v[0] == nodes[0] && v[1] == nodes[3] && v[2] == nodes[2] && v[3] == nodes[1]);
        if (!ok) {
          printf("read_mesh_data(), paired element %d does not list the reversed nodes. Exit...\n", id);
          exit(1);
        }
      }
    }
    md->bd.md[i] = dom[0];
    md->bd.sd[i] = dom[1];
    /* the open region is always the main domain; flip the node order with it */
    if (md->bd.sd[i] == 0) {
      md->bd.sd[i] = md->bd.md[i];
      md->bd.md[i] = 0;
      if (md->bd.ed[i][3] == 0) {
        int t = md->bd.ed[i][1];
        md->bd.ed[i][1] = md->bd.ed[i][2];
        md->bd.ed[i][2] = t;
      } else {
        int t = md->bd.ed[i][3];
        md->bd.ed[i][3] = md->bd.ed[i][1];
        md->bd.ed[i][1] = t;
      }
    }
  }
  fclose(fp);
  md->bd.NN = nc;
}

void read_dqd2(int argc, char **argv, DQD2 *qd)
{
  if (argc != 6 && argc != 13) {
    printf("usage : %s ipw_datafile periodicity_datafile medium_datafile mesh_datafile output_datafile "
           "[rvx rvy rvz theta tx ty tz]\n", argv[0]);
    exit(0);
  }
  memset(qd, 0, sizeof(DQD2));
  read_data_mipw(argv[1], &qd->pw);
  read_periodicity_data(argv[2], qd);
  read_medium_data(argv[3], qd);
  read_mesh_data(argv[4], qd);
  if (argc == 13) {
    for (int i = 0; i < 3; i++) qd->rv[i] = atof(argv[6 + i]);
    if (vabs_d(qd->rv) == 0.0) {
      printf("The vector defining rotation axis is zero-vector. Exit...\n");
      exit(1);
    }
    qd->th = atof(argv[9]);
    for (int i = 0; i < 3; i++) qd->tv[i] = atof(argv[10 + i]);
  } else {
    qd->rv[0] = 1.0;
    qd->rv[1] = qd->rv[2] = 0.0;
    qd->th = 0.0;
    qd->tv[0] = qd->tv[1] = qd->tv[2] = 0.0;
  }
}

static void print_common(DQD2 *qd)
{
  printf("-- periodicity data --\n");
  printf("periodicity data file name : %s\n", qd->ped_fn);
  printf("periodicity setting        : %17d\n", qd->bd.ps);
  printf("lattice vector d1          : (%7g, %7g)\n", qd->bd.qd.vd1[0], qd->bd.qd.vd1[1]);
  printf("lattice vector d2          : (%7g, %7g)\n\n", qd->bd.qd.vd2[0], qd->bd.qd.vd2[1]);
  printf("-- medium data --\n");
  printf("medium data file name                  : %s\n", qd->med_fn);
  for (int i = 1; i <= qd->MN; i++)
    printf("medium (domain) id %2d refractive index : %8g + %8gI\n", i, creal(qd->n[i]), cimag(qd->n[i]));
  printf("\n-- mesh data --\n");
  printf("mesh data file name    : %s\n", qd->msh_fn);
  printf("node number            : %6d\n", qd->bd.Nn);
  printf("defined element number : %6d\n\n", qd->bd.Ne * 2);
  if (qd->th != 0.0 || vabs_d(qd->tv) != 0.0) {
    printf("-- rotation and translation settings --\n");
    if (qd->th != 0.0) {
      printf("vector defining rotation axis :(% 8.7g,% 8.7g,% 8.7g)\n", qd->rv[0], qd->rv[1], qd->rv[2]);
      printf("rotation angle           [rad]: %8.7g\n", qd->th);
    }
    if (vabs_d(qd->tv) != 0.0) printf("translation vector            :(%8.7g,%8.7g,%8.7g)\n", qd->tv[0], qd->tv[1], qd->tv[2]);
    printf("\n");
  }
}

void print_dqd2(DQD2 *qd)
{
  printf("-- plane wave data --\n");
  print_data_mipw(&qd->pw);
  printf("\n");
  print_common(qd);
}

void print_dqd2_mksa(DQD2 *qd)
{
  printf("-- plane wave data --\n");
  print_data_mipw_mksa(&qd->pw);
  printf("\n");
  print_common(qd);
}

/* ---- set-up ---- */
static void check_mipw(DQD2 *qd)
{
  check_param_mipw(&qd->pw);
  if (qd->bd.ps != 1) return;
  double okx = 0, oky = 0;
  for (int i = 0; i < qd->pw.n_ipw; i++) {
    const IpwD *d = &qd->pw.bd.ipw[i].data;
    double kx = d->sin_t * d->cos_p, ky = d->sin_t * d->sin_p, kz = d->cos_t;
    if (kz == 0.0) {
      printf("check_mipw(), plane wave parameter error.\nkz must not be equal to 0. Exit...\n");
      exit(1);
    }
    if (fabs(kz) == 1.0) {
      printf("check_mipw(), plane wave parameter error.\nkz must not be equal to wave number (quasi-peridic boundary "
             "condition unsatisfied). Exit...\n");
      exit(1);
    }
    if (i == 0) {
      okx = kx;
      oky = ky;
    } else if (fabs(okx - kx) > MEPS * fabs(kx) || fabs(oky - ky) > MEPS * fabs(ky)) {
      printf("check_mipw(), plane wave parameter error.\nall of kx and ky must be same. kx=%15.14g. ky=%15.14g. Exit...\n",
             kx, ky);
      exit(1);
    }
  }
}

static void rotation_translation(DQD2 *qd)
{
  double nv[3] = {qd->rv[0], qd->rv[1], qd->rv[2]};
  vuni_d(nv);
  double st = sin(qd->th), ct = cos(qd->th), M[9];
  M[0] = ct + nv[0] * nv[0] * (1.0 - ct);
  M[1] = nv[0] * nv[1] * (1.0 - ct) - nv[2] * st;
  M[2] = nv[2] * nv[0] * (1.0 - ct) + nv[1] * st;
  M[3] = nv[0] * nv[1] * (1.0 - ct) + nv[2] * st;
  M[4] = ct + nv[1] * nv[1] * (1.0 - ct);
  M[5] = nv[1] * nv[2] * (1.0 - ct) - nv[0] * st;
  M[6] = nv[2] * nv[0] * (1.0 - ct) - nv[1] * st;
  M[7] = nv[1] * nv[2] * (1.0 - ct) + nv[0] * st;
  M[8] = ct + nv[2] * nv[2] * (1.0 - ct);
  for (int s = 1; s <= qd->bd.Nn; s++) {
    double r[3];
    for (int i = 0; i < 3; i++)
      r[i] = M[3 * i] * qd->bd.rn[s][0] + M[3 * i + 1] * qd->bd.rn[s][1] + M[3 * i + 2] * qd->bd.rn[s][2] + qd->tv[i];
    for (int i = 0; i < 3; i++) qd->bd.rn[s][i] = r[i];
  }
}

static void init_elem_const(BOUD *bd)
{
  for (int i = 1; i <= bd->Ne; i++) {
    double rc[3][4];
    for (int d = 0; d < 3; d++)
      for (int j = 0; j < 4; j++) rc[d][j] = bd->rn[bd->ed[i][j]][d];
    if (bd->ed[i][3] != 0) {
      for (int d = 0; d < 3; d++) {
        bd->cr[i][d][0] = 0.25 * (rc[d][0] + rc[d][1] + rc[d][2] + rc[d][3]);
        bd->cr[i][d][1] = 0.25 * (-rc[d][0] + rc[d][1] + rc[d][2] - rc[d][3]);
        bd->cr[i][d][2] = 0.25 * (-rc[d][0] - rc[d][1] + rc[d][2] + rc[d][3]);
        bd->cr[i][d][3] = 0.25 * (rc[d][0] - rc[d][1] + rc[d][2] - rc[d][3]);
      }
      /* W = (dr/dzeta) x (dr/deta), linear in zeta and eta */
      static const int pa[3] = {1, 2, 0}, pb[3] = {2, 0, 1};
      for (int d = 0; d < 3; d++) {
        int a = pa[d], b = pb[d];
        bd->cw[i][d][0] = 0.125 * ((rc[a][0] - rc[a][2]) * (rc[b][1] - rc[b][3]) - (rc[a][1] - rc[a][3]) * (rc[b][0] - rc[b][2]));
        bd->cw[i][d][1] = 0.125 * ((rc[a][2] - rc[a][3]) * (rc[b][0] - rc[b][1]) - (rc[a][0] - rc[a][1]) * (rc[b][2] - rc[b][3]));
        bd->cw[i][d][2] = 0.125 * ((rc[a][1] - rc[a][2]) * (rc[b][0] - rc[b][3]) - (rc[a][0] - rc[a][3]) * (rc[b][1] - rc[b][2]));
      }
    } else {
      for (int d = 0; d < 3; d++) {
        bd->cr[i][d][0] = (rc[d][0] + rc[d][1] + rc[d][2]) / 3.0;
        bd->cr[i][d][1] = (-rc[d][0] + 2.0 * rc[d][1] - rc[d][2]) / 3.0;
        bd->cr[i][d][2] = (rc[d][2] - rc[d][0]) / sqrt(3.0);
        bd->cr[i][d][3] = 0.0;
      }
      bd->cw[i][0][0] = (rc[1][1] - rc[1][0]) * (rc[2][2] - rc[2][0]) - (rc[2][1] - rc[2][0]) * (rc[1][2] - rc[1][0]);
      bd->cw[i][1][0] = (rc[2][1] - rc[2][0]) * (rc[0][2] - rc[0][0]) - (rc[0][1] - rc[0][0]) * (rc[2][2] - rc[2][0]);
      bd->cw[i][2][0] = (rc[0][1] - rc[0][0]) * (rc[1][2] - rc[1][0]) - (rc[1][1] - rc[1][0]) * (rc[0][2] - rc[0][0]);
      for (int d = 0; d < 3; d++) bd->cw[i][d][1] = bd->cw[i][d][2] = 0.0;
    }
  }
  const double z44[4] = {-P44_N, P44_N, P44_N, -P44_N}, e44[4] = {-P44_N, -P44_N, P44_N, P44_N};
  const double z49[9] = {-P49_N, P49_N, P49_N, -P49_N, 0.0, P49_N, 0.0, -P49_N, 0.0};
  const double e49[9] = {-P49_N, -P49_N, P49_N, P49_N, -P49_N, 0.0, P49_N, 0.0, 0.0};
  const double w49[9] = {P49_W0, P49_W0, P49_W0, P49_W0, P49_W1, P49_W1, P49_W1, P49_W1, P49_W2};
  const double z34[4] = {-P34_N0, 2.0 * P34_N0, -P34_N0, 0.0}, e34[4] = {-P34_N1, 0.0, P34_N1, 0.0};
  const double w34[4] = {P34_W0, P34_W0, P34_W0, -P34_W1};
  const double z37[7] = {-P37_N0, 2.0 * P37_N0, -P37_N0, P37_N1, -2.0 * P37_N1, P37_N1, 0.0};
  const double e37[7] = {-P37_N2, 0.0, P37_N2, P37_N3, 0.0, -P37_N3, 0.0};
  const double w37[7] = {P37_W0, P37_W0, P37_W0, P37_W1, P37_W1, P37_W1, P37_W2};
  for (int j = 0; j < 4; j++) {
    bd->zt_44[j] = z44[j]; bd->et_44[j] = e44[j]; bd->wt_44[j] = P44_W;
    bd->zt_34[j] = z34[j]; bd->et_34[j] = e34[j]; bd->wt_34[j] = w34[j];
  }
  for (int j = 0; j < 9; j++) { bd->zt_49[j] = z49[j]; bd->et_49[j] = e49[j]; bd->wt_49[j] = w49[j]; }
  for (int j = 0; j < 7; j++) { bd->zt_37[j] = z37[j]; bd->et_37[j] = e37[j]; bd->wt_37[j] = w37[j]; }
  gauss_legendre(bd->xli, bd->wli, GLN);
  gauss_legendre(bd->xhi, bd->whi, GHN);
}

static void malloc_sub_domain(DQD2 *md)
{
  int *Nc = m_alloc2(md->MN + 1, sizeof(int), "malloc_sub_domain");
  for (int i = 1; i <= md->bd.Ne; i++) {
    Nc[md->bd.md[i]]++;
    Nc[md->bd.sd[i]]++;
  }
  md->bd.sb = m_alloc2(md->MN + 1, sizeof(SUBD), "malloc_sub_domain sb");
  for (int i = 0; i <= md->MN; i++) {
    SUBD *s = &md->bd.sb[i];
    s->Ne = Nc[i];
    s->sid = m_alloc2(Nc[i] + 1, sizeof(int), "malloc_sub_domain sid");
    s->U = alloc_z3(Nc[i] + 1, 4, 4);
    s->dU = alloc_z3(Nc[i] + 1, 4, 4);
    s->E = alloc_z3(Nc[i] + 1, 4, 3);
    s->dE = alloc_z3(Nc[i] + 1, 4, 3);
    s->H = alloc_z3(Nc[i] + 1, 4, 3);
    s->dH = alloc_z3(Nc[i] + 1, 4, 3);
  }
  free(Nc);
}

static void mfree_sub_domain(DQD2 *md)
{
  for (int i = 0; i <= md->MN; i++) {
    SUBD *s = &md->bd.sb[i];
    int n = s->Ne + 1;
    free(s->sid);
    free_3((void ***)s->U, n, 4);
    free_3((void ***)s->dU, n, 4);
    free_3((void ***)s->E, n, 4);
    free_3((void ***)s->dE, n, 4);
    free_3((void ***)s->H, n, 4);
    free_3((void ***)s->dH, n, 4);
  }
  free(md->bd.sb);
}

static void init_sub_domain(DQD2 *md)
{
  for (int d = 0; d <= md->MN; d++) {
    int c = 1;
    for (int i = 1; i <= md->bd.Ne; i++) {
      if (md->bd.md[i] == d) md->bd.sb[d].sid[c++] = i;
      else if (md->bd.sd[i] == d) md->bd.sb[d].sid[c++] = -i;
    }
  }
}

static void init_boundary_data(DQD2 *md)
{
  BOUD *bd = &md->bd;
  for (int i = 1; i <= bd->Ne; i++) {
    int tri = (bd->ed[i][3] == 0);
    for (int j = 0; j < 4; j++) {
      double z = tri ? bd->zt_34[j] : bd->zt_44[j], e = tri ? bd->et_34[j] : bd->et_44[j], r[3], w[3];
      for (int d = 0; d < 3; d++) {
        r[d] = bd->cr[i][d][0] + bd->cr[i][d][1] * z + bd->cr[i][d][2] * e + bd->cr[i][d][3] * z * e;
        w[d] = bd->cw[i][d][0] + bd->cw[i][d][1] * z + bd->cw[i][d][2] * e;
        bd->ren[i][j][d] = r[d];
        bd->wen[i][j][d] = w[d];
      }
      vuni_d(w);
      double complex h[3], dh[3];
      calc_mipw_EH_dEHdv(bd->Ui[i][j], h, bd->dUi[i][j], dh, r, w, &md->pw);
      bd->Ui[i][j][3] = 0.0;
      bd->dUi[i][j][3] = 0.0;
    }
  }
}

/* Drops the field evaluator's device copy of qd's model and solution (bq_field.c); called
 * wherever qd's arrays are replaced, freed or rewritten. */
void bq_release_device_state(DQD2 *qd);
void bq_release_solver(void *qd); /* bq_solve.c */

/* QPD2 and the wave numbers from the (first) incident wave, as initialize_dqd2 sets them
 * (reference d3qd2_src/d3qd2_setup.c:121-128) */
static void wave_constants(DQD2 *qd)
{
  const IpwD *d = &qd->pw.bd.ipw[0].data;
  qd->bd.qd.k = d->ki;
  qd->bd.qd.vk[0] = d->ki * d->sin_t * d->cos_p;
  qd->bd.qd.vk[1] = d->ki * d->sin_t * d->sin_p;
  qd->bd.qd.vk[2] = d->ki * d->cos_t;
  if (qd->bd.qd.vk[0] == 0.0) qd->bd.qd.vk[0] = qd->bd.qd.k * MEPS;
  if (qd->bd.qd.vk[1] == 0.0) qd->bd.qd.vk[1] = qd->bd.qd.k * MEPS;
  qd->n[0] = qd->pw.n_0;
  for (int i = 0; i <= qd->MN; i++) qd->kn[i] = 2.0 * M_PI / qd->pw.lambda_0 * qd->n[i];
}

int bq_set_incidence(void *qdv, double lambda0, double theta, double phi, double e0x_re, double e0x_im,
                     double e0y_re, double e0y_im, double power)
{
  DQD2 *qd = (DQD2 *)qdv;
  if (!qd || !qd->pw.bd.ipw || qd->pw.n_ipw < 1 || !qd->bd.sb) return -1;
  if (power == 0.0) power = 0.5 * qd->pw.n_0; /* amplitude 1 (E0 = sqrt(2 power / n0)) */
  if (!(lambda0 > 0.0) || !(power > 0.0)) return -1;
  if (e0x_re == 0.0 && e0x_im == 0.0 && e0y_re == 0.0 && e0y_im == 0.0) return -1;
  const double kz = cos(theta);
  if (qd->bd.ps == 1 && (kz == 0.0 || fabs(kz) == 1.0)) return -1;
  bq_release_device_state(qd);
  Ipw *p = &qd->pw.bd.ipw[0];
  qd->pw.n_ipw = 1;
  p->lambda0 = lambda0;
  p->power = power;
  p->e0x = e0x_re + I * e0x_im;
  p->e0y = e0y_re + I * e0y_im;
  p->theta = theta;
  p->phi = phi;
  setup_mipw(&qd->pw);
  check_mipw(qd);
  wave_constants(qd);
  init_boundary_data(qd);
  return 0;
}

int bq_model_dims(void *qdv, int *Ne, int *NN, int *dNe)
{
  DQD2 *qd = (DQD2 *)qdv;
  if (Ne) *Ne = qd->bd.Ne;
  if (NN) *NN = qd->bd.NN;
  if (dNe && qd->bd.sb)
    for (int d = 0; d <= qd->MN; d++) dNe[d] = qd->bd.sb[d].Ne;
  return qd->MN;
}

void initialize_dqd2(DQD2 *qd)
{
  bq_release_device_state(qd);
  bq_release_solver(qd);
  setup_mipw(&qd->pw);
  check_mipw(qd);
  wave_constants(qd);
  if (qd->th != 0.0 || vabs_d(qd->tv) != 0.0) rotation_translation(qd);
  init_elem_const(&qd->bd);
  malloc_sub_domain(qd);
  init_sub_domain(qd);
  init_boundary_data(qd);
}

void finalize_dqd2(DQD2 *qd)
{
  bq_release_device_state(qd);
  bq_release_solver(qd);
  mfree_sub_domain(qd);
  mfree_elem(&qd->bd);
  free_2((void **)qd->bd.rn, qd->bd.Nn + 1);
  qd->bd.Nn = 0;
  free(qd->n);
  free(qd->kn);
  free_mipw(&qd->pw);
}

/* ---- domain identification by signed solid angles ---- */
static double tri_omega(double *r0, double *r1, double *r2, int *flg, double *n20o)
{
  double n01[3], n12[3], n20[3], nt[3];
  vcrs_d(n01, r0, r1);
  vcrs_d(n12, r1, r2);
  vcrs_d(n20, r2, r0);
  vcrs_d(nt, n01, n20);
  double a0 = atan2(vabs_d(nt), -vdot_d(n01, n20));
  vcrs_d(nt, n12, n01);
  double a1 = atan2(vabs_d(nt), -vdot_d(n12, n01));
  vcrs_d(nt, n20, n12);
  double a2 = atan2(vabs_d(nt), -vdot_d(n20, n12));
  double D = vdot_d(r0, n12);
  if (n20o) for (int i = 0; i < 3; i++) n20o[i] = n20[i];
  if (D > 0.0) return a0 + a1 + a2 - M_PI;
  if (D < 0.0) return -(a0 + a1 + a2 - M_PI);
  *flg = -1;
  return 0.0;
}

static double solid_angle(int quad, double r[4][3], int *flg)
{
  *flg = 0;
  int st = 0;
  for (int i = 0; i < (quad ? 4 : 3); i++) st += vuni_d(r[i]);
  if (st < 0) { *flg = -1; return 0.0; }
  double n20[3];
  double om = tri_omega(r[0], r[1], r[2], flg, n20);
  if (*flg < 0 || !quad) return om;
  double n01[3] = {-n20[0], -n20[1], -n20[2]}, n12[3], nb[3], nt[3];
  vcrs_d(n12, r[2], r[3]);
  vcrs_d(nb, r[3], r[0]);
  vcrs_d(nt, n01, nb);
  double a0 = atan2(vabs_d(nt), -vdot_d(n01, nb));
  vcrs_d(nt, n12, n01);
  double a1 = atan2(vabs_d(nt), -vdot_d(n12, n01));
  vcrs_d(nt, nb, n12);
  double a2 = atan2(vabs_d(nt), -vdot_d(nb, n12));
  double D = vdot_d(r[0], n12);
  if (D > 0.0) om += a0 + a1 + a2 - M_PI;
  else if (D < 0.0) om -= a0 + a1 + a2 - M_PI;
  else { *flg = -1; return 0.0; }
  return om;
}

static int domain_of_point(double *rt, DQD2 *qd)
{
  double *Og = m_alloc2(qd->MN + 1, sizeof(double), "domain_id");
  int res = -1;
  for (int d = 0; d <= qd->MN && res < 0; d++) {
    for (int i = 1; i <= qd->bd.sb[d].Ne; i++) {
      int s = qd->bd.sb[d].sid[i];
      if (s <= 0) continue;
      double rv[4][3];
      for (int j = 0; j < 4; j++)
        for (int k = 0; k < 3; k++) rv[j][k] = qd->bd.rn[qd->bd.ed[s][j]][k] - rt[k];
      int flg;
      double om = solid_angle(qd->bd.ed[s][3] != 0, rv, &flg);
      if (flg < 0) { res = d; break; }
      Og[d] += om;
      Og[qd->bd.sd[s]] -= om;
    }
    if (res >= 0) break;
    if (d == 0 && fabs(Og[d]) < 2.0 * M_PI) res = d;
    else if (Og[d] > 2.0 * M_PI) res = d;
  }
  free(Og);
  return res;
}

int domain_id(double *rt, DQD2 *qd) { return domain_of_point(rt, qd); }

void q0_domain_id_l(int *l1, int *l2, double *rt, DQD2 *qd)
{
  double idd = 1.0 / (qd->bd.qd.vd1[0] * qd->bd.qd.vd2[1] - qd->bd.qd.vd2[0] * qd->bd.qd.vd1[1]);
  *l1 = (int)floor((rt[0] * qd->bd.qd.vd2[1] - rt[1] * qd->bd.qd.vd2[0]) * idd + 0.5);
  *l2 = (int)floor(-(rt[0] * qd->bd.qd.vd1[1] - rt[1] * qd->bd.qd.vd1[0]) * idd + 0.5);
}

int q0_domain_id(double *ro, DQD2 *qd)
{
  int l1, l2;
  q0_domain_id_l(&l1, &l2, ro, qd);
  double rt[3] = {ro[0] - l1 * qd->bd.qd.vd1[0] - l2 * qd->bd.qd.vd2[0],
                  ro[1] - l1 * qd->bd.qd.vd1[1] - l2 * qd->bd.qd.vd2[1], ro[2]};
  return domain_of_point(rt, qd);
}

/* ---- ex.dat ---- */
static void wr(FILE *fp, const void *p, size_t sz, size_t n)
{
  if (fwrite(p, sz, n, fp) != n) {
    printf("dat_write(), write error. Exit...\n");
    exit(1);
  }
}
static void rd(FILE *fp, void *p, size_t sz, size_t n, const char *what)
{
  if (fread(p, sz, n, fp) != n) {
    printf("dat_read(), failed to read the %s. exit...\n", what);
    exit(1);
  }
}

void dat_write(char *fname, DQD2 *qd)
{
  FILE *fp = fopen(fname, "wb");
  if (!fp) {
    printf("dat_write(), Failed to create the %s file.\n", fname);
    exit(1);
  }
  BOUD *bd = &qd->bd;
  wr(fp, qd->ped_fn, 1, 128);
  wr(fp, qd->med_fn, 1, 128);
  wr(fp, qd->msh_fn, 1, 128);
  wr(fp, qd->rv, sizeof(double), 3);
  wr(fp, &qd->th, sizeof(double), 1);
  wr(fp, qd->tv, sizeof(double), 3);
  wr(fp, &qd->MN, sizeof(int), 1);
  wr(fp, qd->n, sizeof(double complex), qd->MN + 1);
  wr(fp, qd->kn, sizeof(double complex), qd->MN + 1);
  wr(fp, &qd->pw, sizeof(MIPW), 1);
  wr(fp, qd->pw.bd.ipw, sizeof(Ipw), qd->pw.n_ipw);
  wr(fp, &bd->ps, sizeof(int), 1);
  wr(fp, &bd->qd, sizeof(QPD2), 1);
  wr(fp, &bd->Nn, sizeof(int), 1);
  wr(fp, &bd->Ne, sizeof(int), 1);
  for (int i = 0; i <= bd->Nn; i++) wr(fp, bd->rn[i], sizeof(double), 3);
  for (int i = 0; i <= bd->Ne; i++) wr(fp, bd->ed[i], sizeof(int), 4);
  wr(fp, &bd->NN, sizeof(int), 1);
  for (int i = 0; i <= bd->Ne; i++) wr(fp, bd->eni[i], sizeof(int), 4);
  wr(fp, bd->md, sizeof(int), bd->Ne + 1);
  wr(fp, bd->sd, sizeof(int), bd->Ne + 1);
  wr(fp, bd->gd, sizeof(int), bd->Ne + 1);
  for (int i = 0; i <= bd->Ne; i++) for (int j = 0; j < 4; j++) wr(fp, bd->ren[i][j], sizeof(double), 3);
  for (int i = 0; i <= bd->Ne; i++) for (int j = 0; j < 4; j++) wr(fp, bd->wen[i][j], sizeof(double), 3);
  for (int i = 0; i <= bd->Ne; i++) for (int j = 0; j < 4; j++) wr(fp, bd->Ui[i][j], sizeof(double complex), 4);
  for (int i = 0; i <= bd->Ne; i++) for (int j = 0; j < 4; j++) wr(fp, bd->dUi[i][j], sizeof(double complex), 4);
  for (int d = 0; d <= qd->MN; d++) {
    SUBD *s = &bd->sb[d];
    wr(fp, &s->Ne, sizeof(int), 1);
    wr(fp, s->sid, sizeof(int), s->Ne + 1);
    for (int i = 0; i <= s->Ne; i++) for (int j = 0; j < 4; j++) wr(fp, s->U[i][j], sizeof(double complex), 4);
    for (int i = 0; i <= s->Ne; i++) for (int j = 0; j < 4; j++) wr(fp, s->dU[i][j], sizeof(double complex), 4);
    for (int i = 0; i <= s->Ne; i++) for (int j = 0; j < 4; j++) wr(fp, s->E[i][j], sizeof(double complex), 3);
    for (int i = 0; i <= s->Ne; i++) for (int j = 0; j < 4; j++) wr(fp, s->dE[i][j], sizeof(double complex), 3);
    for (int i = 0; i <= s->Ne; i++) for (int j = 0; j < 4; j++) wr(fp, s->H[i][j], sizeof(double complex), 3);
    for (int i = 0; i <= s->Ne; i++) for (int j = 0; j < 4; j++) wr(fp, s->dH[i][j], sizeof(double complex), 3);
  }
  fclose(fp);
}

void dat_read(char *fname, DQD2 *qd)
{
  bq_release_device_state(qd);
  bq_release_solver(qd);
  FILE *fp = fopen(fname, "rb");
  if (!fp) {
    printf("dat_read(), Failed to open the %s file.\n", fname);
    exit(1);
  }
  memset(qd, 0, sizeof(DQD2));
  BOUD *bd = &qd->bd;
  rd(fp, qd->ped_fn, 1, 128, "ped_fn");
  rd(fp, qd->med_fn, 1, 128, "med_fn");
  rd(fp, qd->msh_fn, 1, 128, "msh_fn");
  rd(fp, qd->rv, sizeof(double), 3, "rv");
  rd(fp, &qd->th, sizeof(double), 1, "th");
  rd(fp, qd->tv, sizeof(double), 3, "tv");
  rd(fp, &qd->MN, sizeof(int), 1, "MN");
  qd->n = m_alloc2(qd->MN + 1, sizeof(double complex), "dat_read(), n");
  qd->kn = m_alloc2(qd->MN + 1, sizeof(double complex), "dat_read(), kn");
  rd(fp, qd->n, sizeof(double complex), qd->MN + 1, "n");
  rd(fp, qd->kn, sizeof(double complex), qd->MN + 1, "kn");
  rd(fp, &qd->pw, sizeof(MIPW), 1, "MIPW");
  qd->pw.bd.ipw = m_alloc2(qd->pw.n_ipw, sizeof(Ipw), "dat_read(), ipw");
  rd(fp, qd->pw.bd.ipw, sizeof(Ipw), qd->pw.n_ipw, "ipw");
  setup_mipw(&qd->pw);
  rd(fp, &bd->ps, sizeof(int), 1, "ps");
  rd(fp, &bd->qd, sizeof(QPD2), 1, "qd");
  rd(fp, &bd->Nn, sizeof(int), 1, "Nn");
  rd(fp, &bd->Ne, sizeof(int), 1, "Ne");
  malloc_node(bd);
  malloc_elem(bd);
  for (int i = 0; i <= bd->Nn; i++) rd(fp, bd->rn[i], sizeof(double), 3, "rn");
  for (int i = 0; i <= bd->Ne; i++) rd(fp, bd->ed[i], sizeof(int), 4, "ed");
  rd(fp, &bd->NN, sizeof(int), 1, "NN");
  for (int i = 0; i <= bd->Ne; i++) rd(fp, bd->eni[i], sizeof(int), 4, "eni");
  rd(fp, bd->md, sizeof(int), bd->Ne + 1, "md");
  rd(fp, bd->sd, sizeof(int), bd->Ne + 1, "sd");
  rd(fp, bd->gd, sizeof(int), bd->Ne + 1, "gd");
  for (int i = 0; i <= bd->Ne; i++) for (int j = 0; j < 4; j++) rd(fp, bd->ren[i][j], sizeof(double), 3, "ren");
  for (int i = 0; i <= bd->Ne; i++) for (int j = 0; j < 4; j++) rd(fp, bd->wen[i][j], sizeof(double), 3, "wen");
  for (int i = 0; i <= bd->Ne; i++) for (int j = 0; j < 4; j++) rd(fp, bd->Ui[i][j], sizeof(double complex), 4, "Ui");
  for (int i = 0; i <= bd->Ne; i++) for (int j = 0; j < 4; j++) rd(fp, bd->dUi[i][j], sizeof(double complex), 4, "dUi");
  init_elem_const(bd);
  malloc_sub_domain(qd);
  for (int d = 0; d <= qd->MN; d++) {
    SUBD *s = &bd->sb[d];
    int tmp;
    rd(fp, &tmp, sizeof(int), 1, "Ne of a domain");
    rd(fp, s->sid, sizeof(int), s->Ne + 1, "sid");
    for (int i = 0; i <= s->Ne; i++) for (int j = 0; j < 4; j++) rd(fp, s->U[i][j], sizeof(double complex), 4, "U");
    for (int i = 0; i <= s->Ne; i++) for (int j = 0; j < 4; j++) rd(fp, s->dU[i][j], sizeof(double complex), 4, "dU");
    for (int i = 0; i <= s->Ne; i++) for (int j = 0; j < 4; j++) rd(fp, s->E[i][j], sizeof(double complex), 3, "E");
    for (int i = 0; i <= s->Ne; i++) for (int j = 0; j < 4; j++) rd(fp, s->dE[i][j], sizeof(double complex), 3, "dE");
    for (int i = 0; i <= s->Ne; i++) for (int j = 0; j < 4; j++) rd(fp, s->H[i][j], sizeof(double complex), 3, "H");
    for (int i = 0; i <= s->Ne; i++) for (int j = 0; j < 4; j++) rd(fp, s->dH[i][j], sizeof(double complex), 3, "dH");
  }
  fclose(fp);
}

void output_node_particles(char *fname, DQD2 *qd)
{
  char fo[512];
  size_t s1 = strlen(fname);
  if (s1 > 200) {
    printf("output_node_particles(), file name is too long. exit...\n");
    exit(1);
  }
  snprintf(fo, sizeof fo, "%s", fname);
  char *dot = strrchr(fo, '.');
  if (dot) strcpy(dot, ".particles");
  FILE *fp = fopen(fo, "wt");
  if (!fp) {
    printf("Can not open the %s file.\n", fo);
    exit(1);
  }
  fprintf(fp, "# x y z object_id\n");
  for (int i = 1; i <= qd->bd.Ne; i++)
    for (int j = 0; j < 4; j++)
      fprintf(fp, "%15.14e %15.14e %15.14e %d\n", qd->bd.ren[i][j][0], qd->bd.ren[i][j][1], qd->bd.ren[i][j][2], 0);
  fclose(fp);
}
