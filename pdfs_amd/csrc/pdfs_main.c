/*
 * PDFS CLI: drop-in for single_gpu_cuda/src/main.c:56-106.
 *
 *   PDFS <file.vti> <outputPrefix> [--passes N | --converge] [--fp32]
 *
 * With two arguments the behaviour is the reference's: read the VTI file,
 * print the dimensions, run calc_dist_field with one pass of the 8
 * orderings in FP64 reference arithmetic (phi3D.c:118), write
 * <prefix>_distfield.vti. The options are extensions: more passes, passes
 * until converged, or the FP32 column-pipeline solver.
 */
#define _POSIX_C_SOURCE 200809L
#include "pdfs.h"
#include "pdfs_vti.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

int main(int argc, char *argv[])
{
  if (argc < 3) {
    printf("Usage: %s <file.vti> <prefix>\n", argv[0]);
    exit(1);
  }
  const char *fileToRead = argv[1], *prefix = argv[2];
  int passes = 1, precision = PDFS_F64_REF;
  for (int a = 3; a < argc; a++) {
    if (!strcmp(argv[a], "--passes") && a + 1 < argc) {
      passes = atoi(argv[++a]);
    } else if (!strcmp(argv[a], "--converge")) {
      passes = 0;
    } else if (!strcmp(argv[a], "--fp32")) {
      precision = PDFS_F32;
    } else {
      printf("Usage: %s <file.vti> <prefix> [--passes N | --converge] [--fp32]\n", argv[0]);
      exit(1);
    }
  }
  pdfs_vti v;
  int rc = pdfs_vti_read(fileToRead, &v);
  if (rc == -1) {
    printf("Unable to read file \"%s\"!\n", fileToRead);
    exit(1);
  }
  if (rc != 0) {
    printf("Malformed VTI file \"%s\" (error %d)\n", fileToRead, rc);
    exit(1);
  }
  char outFilename[1024];
  snprintf(outFilename, sizeof outFilename, "%s_distfield", prefix);
  Phi pf;
  pf.x = v.x;
  pf.y = v.y;
  pf.z = v.z;
  pf.dx = v.dx;
  pf.dy = v.dy;
  pf.dz = v.dz;
  pf.location = v.location;
  pf.distance = v.distance;
  pf.F = 1;
  /* main.c:85-95 */
  printf("Dimensions:\n");
  printf("x: %d\tdx: %f\n", pf.x, pf.dx);
  printf("y: %d\tdy: %f\n", pf.y, pf.dy);
  printf("z: %d\tdz: %f\n", pf.z, pf.dz);
  printf("Size:\n");
  printf("x: %d\tdx: %f\n", pf.x, pf.dx);
  printf("y: %d\tdy: %f\n", pf.y, pf.dy);
  printf("z: %d\tdz: %f\n", pf.z, pf.dz);
  /* progress and timing lines of phi3D_fsm.cu:146,187 and :104 (the time
   * includes the device copies, as the reference's does) */
  for (int q = 1; q <= passes; q++) printf("Please wait. Sweeping...[%d/%d]\n", q, passes);
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  rc = pdfs_calc_dist_field(&pf, passes, precision);
  clock_gettime(CLOCK_MONOTONIC, &t1);
  for (int q = 1; q <= passes; q++) printf("Sweeping finished!......[%d/%d]\n", q, passes);
  if (rc == PDFS_OK)
    printf("Parallel FSM time: %f s.\n",
           (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec));
  if (rc != PDFS_OK) {
    printf("\nHIP failure: '%s'\n", pdfs_last_error());
    pdfs_vti_free(&v);
    exit(EXIT_FAILURE);
  }
  /* phi3D.c:148-157 -> file_writer.c:100-137: "<name>.vti" */
  char vtiName[1100];
  snprintf(vtiName, sizeof vtiName, "%s.vti", outFilename);
  if (pdfs_vti_write(vtiName, pf.distance, pf.x + 2, pf.y + 2, pf.z + 2, pf.dx, pf.dy, pf.dz) != 0) {
    printf("Unable to write \"%s\"\n", vtiName);
    pdfs_vti_free(&v);
    exit(1);
  }
  printf("VTI file created: %s\n", vtiName);
  pdfs_vti_free(&v);
  return 0;
}
