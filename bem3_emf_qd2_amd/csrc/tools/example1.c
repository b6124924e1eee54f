/* example1 -- total field at one point from an ex.dat result and the radiation force and
 * torque, 4-point and 9-point rules (the output of the reference's example1.c). */
#include "bem3_emf_qd2.h"

static void show(double complex *E, double complex *H)
{
  const char *n[6] = {"Ex", "Ey", "Ez", "Hx", "Hy", "Hz"};
  for (int i = 0; i < 3; i++) printf("%s = % 15.14e %+15.14e I \n", n[i], creal(E[i]), cimag(E[i]));
  for (int i = 0; i < 3; i++) printf("%s = % 15.14e %+15.14e I \n", n[3 + i], creal(H[i]), cimag(H[i]));
}

int main(int argc, char **argv)
{
  DQD2 qd;
  double complex E[3], H[3];
  double r[3] = {0.3, 0.0, 1.1};
  if (argc < 2) {
    printf("usage: %s ex.dat [x y z]\n", argv[0]);
    return 1;
  }
  if (argc == 5) for (int i = 0; i < 3; i++) r[i] = atof(argv[2 + i]);
  dat_read(argv[1], &qd);
  print_dqd2(&qd);
  EH_t(E, H, r, 0, &qd);
  printf("Total electromagnetic field at r=( % g,% g,% g )\n", r[0], r[1], r[2]);
  printf("type=%d setting ( 4 point Gauss-Legendre )\n", 0);
  show(E, H);
  EH_t(E, H, r, 1, &qd);
  printf("type=%d setting ( 9 or 7 point Gauss-Legendre )\n", 1);
  show(E, H);
  double F[3], N[3], rc[3] = {0.0, 0.0, 0.0};
  printf("\nRadiation force and torque\n");
  for (int type = 0; type < 2; type++) {
    force_FN(F, N, rc, type, &qd);
    printf("type=%d setting ( %s point Gauss-Legendre ) \n", type, type == 0 ? "4" : "9 or 7");
    printf("F = (% 15.14e,% 15.14e,% 15.14e)\n", F[0], F[1], F[2]);
    printf("N = (% 15.14e,% 15.14e,% 15.14e), center of rotation (% g,% g,% g)\n", N[0], N[1], N[2], rc[0], rc[1],
           rc[2]);
  }
  finalize_dqd2(&qd);
  return 0;
}
