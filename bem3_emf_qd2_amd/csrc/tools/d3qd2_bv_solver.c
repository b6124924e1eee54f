/* d3qd2_bv_solver -- solver executable with the reference's command line
 * (d3qd2_src/d3qd2_bv_solver.c, d3qd2_src/d3qd2_setup.c:546-555):
 *   d3qd2_bv_solver ipw.txt periodicity_data.txt medium_data.txt mesh.msh out.dat [rv th tv]
 * read -> print -> initialise -> particles -> solve (GPU) -> write ex.dat -> free. */
#include "bem3_emf_qd2.h"

int main(int argc, char **argv)
{
  DQD2 qd;
  read_dqd2(argc, argv, &qd);
  print_dqd2(&qd);
  initialize_dqd2(&qd);
  output_node_particles(argv[5], &qd);
  solve_bieq(&qd);
  printf("\noutput datafile : %s\n", argv[5]);
  dat_write(argv[5], &qd);
  finalize_dqd2(&qd);
  return 0;
}
