// vlp4d_main.cpp -- drop-in command-line driver for the vlp4d_mpi executable
// (miniapps/vlp4d_mpi/openmp/Vlp4d.cpp:12-62) on one MI355X, over the C-ABI.
//
// Same surface: "-f deck.dat" (miniapps/vlp4d_mpi/Parser.hpp:18-21), the
// config echo of VMO/Init.cpp:92-124, "iter %d" per step, nrj.out and
// l2norm.out in the working directory (VMO/Diags.cpp:78-104) and the timer
// table (miniapps/vlp4d_mpi/Timer.hpp:70-75).  The reference's exit paths are
// reproduced here and only here: a missing deck or unknown case abort(), a
// too-large time step prints its message and exits with status 0.
//
// Extra flags: "--fp32" selects FP32 storage of f and the spline
// coefficients (double interpolation arithmetic), "--fp32-acc" FP32 storage
// with float interpolation arithmetic, "-d N" the device (as the reference's
// --device).  "--lagrange" runs the Lagrange split-step solver of the non-MPI
// miniapp instead (miniapps/vlp4d/openmp/vlp4d.cpp:33-81: same deck and echo,
// its 3-column nrj.out, no l2norm.out); that executable takes the deck as its
// only argument, so a bare path is read as the deck too.
#include <sys/stat.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>

#include "../../include/vlp4d.h"

static void print_config(const vlp_config *c) {
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
  int precision = VLP_FP64, device = 0;
  bool lagrange = false;
  for (int i = 1; i < argc; i++) {
    if ((!strcmp(argv[i], "-f") || !strcmp(argv[i], "--file")) && i + 1 < argc) file = argv[++i];
    else if ((!strcmp(argv[i], "-d") || !strcmp(argv[i], "--device")) && i + 1 < argc) device = atoi(argv[++i]);
    else if (!strcmp(argv[i], "--fp32")) precision = VLP_FP32;
    else if (!strcmp(argv[i], "--fp32-acc")) precision = VLP_FP32_ACC;
    else if (!strcmp(argv[i], "--lagrange")) lagrange = true;
    else if (argv[i][0] != '-') file = argv[i];
  }
  printf("reading input file %s\n", file);
  vlp_config cfg;
  int rc = vlp_parse_deck(file, &cfg);
  if (rc == VLP_E_FILE) {
    printf("import: error file not found\n");
    fflush(stdout);
    abort();
  }
  print_config(&cfg);
  vlp_ctx *ctx = nullptr;
  rc = lagrange ? vlp_create_lagrange(&cfg, device, &ctx) : vlp_create_config(&cfg, precision, device, &ctx);
  if (rc == VLP_E_CASE) {
    printf("Unknown test case !\n");
    fflush(stdout);
    abort();
  }
  if (rc) {
    fprintf(stderr, "vlp4d: %s\n", vlp_last_error(nullptr));
    return 1;
  }
  if (!lagrange)  // the domain line of the MPI miniapp's decomposition
    printf("[0] local domain [%4u:%4u,%4u:%4u,%4u:%4u,%4u:%4u] cost %lu\n", 0u, cfg.nxmax[0] - 1, 0u, cfg.nxmax[1] - 1,
           0u, cfg.nxmax[2] - 1, 0u, cfg.nxmax[3] - 1,
           (unsigned long)cfg.nxmax[0] * cfg.nxmax[1] * cfg.nxmax[2] * cfg.nxmax[3]);
  // fxvx diagnostics: every ifreq steps (VMO/Timestep.hpp:93-95,
  // miniapps/vlp4d/openmp/timestep.hpp:49-53), and at iteration 0 for the
  // non-MPI miniapp (its vlp4d.cpp:63), into data/<miniapp>/
  const char *fx_dir = lagrange ? "data/vlp4d" : "data/vlp4d_mpi";
  auto fxvx = [&]() {
    if (!cfg.fxvx || vlp_iter(ctx) % cfg.ifreq) return;
    mkdir("data", 0755);
    mkdir(fx_dir, 0755);
    printf("print_fxvx %d\n", vlp_iter(ctx));
    if (vlp_write_fxvx(ctx, fx_dir)) fprintf(stderr, "vlp4d: %s\n", vlp_last_error(ctx));
  };
  if (lagrange) fxvx();
  while (vlp_iter(ctx) < cfg.nbiter) {
    printf("iter %d\n", vlp_iter(ctx));
    rc = vlp_step(ctx, 1);
    if (rc == VLP_E_CFL) {
      fprintf(stderr, "Time step is too large, exiting\n");
      exit(0);
    }
    if (rc) {
      fprintf(stderr, "vlp4d: %s\n", vlp_last_error(ctx));
      return 1;
    }
    fxvx();
  }
  rc = vlp_save(ctx, ".");
  if (rc) fprintf(stderr, "vlp4d: %s\n", vlp_last_error(ctx));
  for (int k = 0; k < VLP_T_COUNT; k++) {
    double s = 0;
    int calls = 0;
    vlp_timer(ctx, k, &s, &calls);
    std::cout << vlp_timer_label(k) << " " << s << " [s], " << calls << " calls" << std::endl;
  }
  vlp_destroy(ctx);
  return 0;
}
