/*
 * 1x1x1 driver around the reference's own FSM core
 * (mpi_openacc/src/phi3D_fsm.c, phi3D.c), compiled from the sources where
 * they lie under /root/reference by oracle/Makefile into oracle/_ref/.
 *
 * TEST INFRASTRUCTURE ONLY: used to pin the oracle and as the CPU baseline
 * ("kind": "reference"). Not part of the product.
 *
 * With one rank and a 1x1x1 decomposition, runFSMAndPropagate
 * (mpi_openacc/src/block_fsm.c:277-369) reduces to runFSM(pf, sn) for
 * sn = 1..8 (block_fsm.c:296-298) with no messages; the repeat-until-
 * converged loop around it is new (the reference runs one pass). Timing
 * follows GET_TIME (mpi_openacc/src/utilities.h:62-67) around the sweep
 * loop only, as block_fsm.c:189-202 does.
 */
#include "phi3D.h"
#include "phi3D_fsm.h"
#include "utilities.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static Phi make_phi(double *dist, int X, int Y, int Z, double dx, double dy,
                    double dz)
{
  Phi pf;
  pf.x = X;
  pf.y = Y;
  pf.z = Z;
  pf.dx = dx;
  pf.dy = dy;
  pf.dz = dz;
  pf.distance = dist;
  pf.F = SPEED;
  return pf;
}

/* One ordering sn (1..8) of the reference, in place on a padded array. */
int ref_run_sn(double *dist, int X, int Y, int Z, double dx, double dy,
               double dz, int sn)
{
  if (sn < 1 || sn > 8) return -1;
  Phi pf = make_phi(dist, X, Y, Z, dx, dy, dz);
  runFSM(&pf, sn);
  return 0;
}

/* Passes of sn = 1..8 until a pass changes no value by more than tol, or
 * max_passes. Returns passes; *seconds = GET_TIME span of the sweeps only. */
int ref_solve(double *dist, int X, int Y, int Z, double dx, double dy,
              double dz, double tol, int max_passes, double *seconds,
              double *last_maxdiff)
{
  Phi pf = make_phi(dist, X, Y, Z, dx, dy, dz);
  size_t n = (size_t)(X + 2) * (Y + 2) * (Z + 2);
  double *prev = (double *)malloc(n * sizeof(double));
  double t0, t1, total = 0.0, md = 0.0;
  int passes = 0;
  while (passes < max_passes) {
    memcpy(prev, dist, n * sizeof(double));
    GET_TIME(t0);
    for (int sn = 1; sn <= 8; sn++) runFSM(&pf, sn);
    GET_TIME(t1);
    total += t1 - t0;
    passes++;
    md = 0.0;
    for (size_t p = 0; p < n; p++) {
      if (prev[p] == dist[p]) continue;
      double df = fabs(prev[p] - dist[p]);
      if (!(df <= md)) md = df;
    }
    if (md <= tol) break;
  }
  free(prev);
  if (seconds) *seconds = total;
  if (last_maxdiff) *last_maxdiff = md;
  return passes;
}

/* Time `norders` single orderings (sn cycling 1..8) without a convergence
 * check: the bounded CPU-baseline sample. Returns seconds (GET_TIME). */
double ref_time_orderings(double *dist, int X, int Y, int Z, double dx,
                          double dy, double dz, int norders)
{
  Phi pf = make_phi(dist, X, Y, Z, dx, dy, dz);
  double t0, t1;
  GET_TIME(t0);
  for (int q = 0; q < norders; q++) runFSM(&pf, 1 + (q % 8));
  GET_TIME(t1);
  return t1 - t0;
}
