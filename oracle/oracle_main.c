/*
 * oracle_main.c -- command-line driver around the CPU oracle (TEST INFRASTRUCTURE).
 *
 * Mirrors the reference main (miniapps/vlp4d_mpi/openmp/Vlp4d.cpp:12-62) for
 * one rank: "-f deck.dat", the config echo of VMO/Init.cpp:92-124, "iter %d"
 * per step, nrj.out / l2norm.out in the working directory and the timer table
 * of miniapps/vlp4d_mpi/Timer.hpp:70-75.  Used as the timed CPU baseline.
 *
 * Extra flags (not in the reference): "-n K" runs only K steps of the deck;
 * "-T" prints "steptime <iter> <seconds>" after each step (wall time of that
 * onetimestep + swap, the MainLoop body).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "vlp4d_oracle.h"

static const char *labels[VO_T_COUNT] = {"total", "MainLoop", "pack", "comm", "unpack", "advec2D", "advec4D",
                                         "field", "all_reduce", "Fourier", "diag", "splinecoeff_xy",
                                         "splinecoeff_vxvy"};

static void print_config(const vo_config *c) { /* VMO/Init.cpp:92-124 */
  printf("** Definition of mesh\n");
  printf("Number of points in  x with the coarse mesh : %d\n", c->nxmax[0]);
  printf("Number of points in  y with the coarse mesh : %d\n", c->nxmax[1]);
  printf("Number of points in Vx with the coarse mesh : %d\n", c->nxmax[2]);
  printf("Number of points in Vy with the coarse mesh : %d\n", c->nxmax[3]);
  printf("\n** Defintion of the geometry of the domain\n");
  printf("Minimal value of Ex : %lf\n", c->minPhy[0]);
  printf("Maximal value of Ex : %lf\n", c->maxPhy[0]);
  printf("Minimal value of Ey : %lf\n", c->minPhy[1]);
  printf("Maximal value of Ey : %lf\n", c->maxPhy[1]);
  printf("\nMinimal value of Vx : %lf\n", c->minPhy[2]);
  printf("Maximal value of Vx   : %lf\n", c->maxPhy[2]);
  printf("Minimal value of Vy   : %lf\n", c->minPhy[3]);
  printf("Maximal value of Vy   : %lf\n", c->maxPhy[3]);
  printf("\n** Considered test cases");
  printf("\n-10- Landau Damping");
  printf("\n-11- Landau Damping 2");
  printf("\n-20- Two beam instability");
  printf("\nNumber of the chosen test case : %d\n", c->idcase);
  printf("\n** Iterations in time and diagnostics\n");
  printf("Time step : %lf\n", c->dt);
  printf("Number of total iterations : %d\n", c->nbiter);
  printf("Frequency of diagnostics : %d\n", c->ifreq);
  printf("Diagnostics of fxvx : %d\n", c->fxvx);
}

int main(int argc, char **argv) {
  const char *file = "data.dat";
  int nsteps = -1, steptimes = 0;
  for (int i = 1; i < argc; i++) {
    if ((!strcmp(argv[i], "-f") || !strcmp(argv[i], "--file")) && i + 1 < argc) file = argv[++i];
    else if (!strcmp(argv[i], "-n") && i + 1 < argc) nsteps = atoi(argv[++i]);
    else if (!strcmp(argv[i], "-T")) steptimes = 1;
  }
  printf("reading input file %s\n", file);
  vo_config cfg;
  int rc = vo_parse_deck(file, &cfg);
  if (rc == VO_E_FILE) {
    printf("import: error file not found\n");
    fflush(stdout);
    abort();
  }
  print_config(&cfg);
  if (nsteps >= 0 && nsteps < cfg.nbiter) cfg.nbiter = nsteps;
  vo_state *st = NULL;
  rc = vo_create(&cfg, &st);
  if (rc == VO_E_CASE) {
    printf("Unknown test case !\n");
    fflush(stdout);
    abort();
  }
  if (rc) {
    fprintf(stderr, "Error: invalid configuration (code %d)\n", rc);
    return 1;
  }
  while (vo_iter(st) < cfg.nbiter) {
    printf("iter %d\n", vo_iter(st));
    const double t0 = vo_timer_seconds(st, VO_T_MAINLOOP);
    rc = vo_step(st);
    if (steptimes && !rc) printf("steptime %d %.6f\n", vo_iter(st), vo_timer_seconds(st, VO_T_MAINLOOP) - t0);
    if (rc == VO_E_CFL) {
      fprintf(stderr, "Time step is too large, exiting\n");
      exit(0);
    }
  }
  vo_save(st, ".");
  for (int k = 0; k < VO_T_COUNT; k++)
    printf("%s %g [s], %d calls\n", labels[k], vo_timer_seconds(st, k), vo_timer_calls(st, k));
  vo_destroy(st);
  return 0;
}
