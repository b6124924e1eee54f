/* example2 -- electric and magnetic field intensity |E|^2, |H|^2 on the planes x = 0, y = 0 and
 * z = 0 (200 x 200 points over +-lambda0 each) from an ex.dat result, written to Ie_yz.txt,
 * Ih_yz.txt, Ie_xz.txt, Ih_xz.txt, Ie_xy.txt, Ih_xy.txt in the text layout of the reference's
 * example2.c:11-105 (gnuplot blocks: "%g %g %15.14e", a blank line after each row).
 * Each plane is one batched EH_t launch (bq_EH_t_batch) instead of 40 000 calls; the values
 * are those EH_t gives point by point.  An optional second argument sets the points per
 * side (default 200, the reference's). */
#include "bem3_emf_qd2.h"
#include "bq_amd.h"

static int plane(DQD2 *qd, int fixed, int a, int b, int max, double rang, const char *fe, const char *fh,
                 const char *head_e, const char *head_h)
{
  const double dr = rang * 2.0 / (double)(max - 1);
  const size_t n = (size_t)max * max;
  double *r = calloc(n * 3, sizeof(double));
  double complex *E = malloc(n * 3 * sizeof(double complex)), *H = malloc(n * 3 * sizeof(double complex));
  int *did = malloc(n * sizeof(int));
  if (!r || !E || !H || !did) {
    printf("example2: out of host memory\n");
    exit(1);
  }
  for (int i = 0; i < max; i++)
    for (int j = 0; j < max; j++) {
      double *p = r + 3 * ((size_t)i * max + j);
      p[fixed] = 0.0;
      p[a] = -rang + (double)i * dr;
      p[b] = -rang + (double)j * dr;
    }
  if (bq_EH_t_batch(qd, (int)n, r, 1, (double *)E, (double *)H, did) != 0) {
    printf("example2: field evaluation failed: %s\n", bq_last_error());
    exit(1);
  }
  FILE *fp1 = fopen(fe, "wt"), *fp2 = fopen(fh, "wt");
  if (!fp1 || !fp2) {
    printf("Can not open the file.\n");
    exit(1);
  }
  fprintf(fp1, "%s\n", head_e);
  fprintf(fp2, "%s\n", head_h);
  for (int i = 0; i < max; i++) {
    for (int j = 0; j < max; j++) {
      const size_t k = (size_t)i * max + j;
      const double complex *e = E + 3 * k, *h = H + 3 * k;
      double ie = creal(e[0] * conj(e[0])) + creal(e[1] * conj(e[1])) + creal(e[2] * conj(e[2]));
      double ih = creal(h[0] * conj(h[0])) + creal(h[1] * conj(h[1])) + creal(h[2] * conj(h[2]));
      fprintf(fp1, "%g %g %15.14e\n", r[3 * k + a], r[3 * k + b], ie);
      fprintf(fp2, "%g %g %15.14e\n", r[3 * k + a], r[3 * k + b], ih);
    }
    fprintf(fp1, "\n");
    fprintf(fp2, "\n");
  }
  fclose(fp1);
  fclose(fp2);
  free(r);
  free(E);
  free(H);
  free(did);
  return 0;
}

int main(int argc, char **argv)
{
  DQD2 qd;
  if (argc < 2) {
    printf("usage: %s ex.dat [points_per_side]\n", argv[0]);
    return 1;
  }
  int max = argc > 2 ? atoi(argv[2]) : 200;
  if (max < 2) {
    printf("example2: points per side must be at least 2\n");
    return 1;
  }
  dat_read(argv[1], &qd);
  print_dqd2(&qd);
  const double rang = 1.0 * qd.pw.lambda_0;
  plane(&qd, 0, 1, 2, max, rang, "Ie_yz.txt", "Ih_yz.txt", "# y z electric_field_intensity",
        "# y z magnetic_field_intensity");
  plane(&qd, 1, 0, 2, max, rang, "Ie_xz.txt", "Ih_xz.txt", "# x z electric_field_intensity",
        "# x z magnetic_field_intensity");
  plane(&qd, 2, 0, 1, max, rang, "Ie_xy.txt", "Ih_xy.txt", "# x y electric_field_intensity",
        "# x y magnetic_field_intensity");
  printf("Intensity plot is finished\n");
  finalize_dqd2(&qd);
  return 0;
}
