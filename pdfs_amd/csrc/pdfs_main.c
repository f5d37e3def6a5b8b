/*
 * PDFS CLI: drop-in for single_gpu_cuda/src/main.c:56-106 and, with
 * options, for the MPI/OpenACC build's mpi_openacc/src/main.c:71-225.
 *
 *   PDFS <file.vti> <outputPrefix> [--passes N | --converge] [--fp32]
 *   PDFS -i <file.nc> -p <outputPrefix> [--nx 1 --ny 1 --nz 1]
 *        [--passes N | --converge] [--fp32]
 *
 * With two arguments the behaviour is the reference's: read the VTI file,
 * print the dimensions, run calc_dist_field with one pass of the 8
 * orderings in FP64 reference arithmetic (phi3D.c:118), write
 * <prefix>_distfield.vti. The options are extensions: more passes, passes
 * until converged, or the FP32 column-pipeline solver.
 *
 * With -i / -p it is the MPI build on one process: read the NetCDF-4 file's
 * Flags and Distance (lib/libpdfs_nc.so, loaded only here), pad them by one
 * cell (border Flags -1, Distance +inf; process_file.c:78-120, 228-333),
 * replace unknown distances, sweep one pass, set the inside negative and
 * copy the border from the nearest interior cell (process_file.c:340-530;
 * the same steps as calc_dist_field), and write <prefix>_soln.nc. One
 * process means one block: as in the reference, the decomposition
 * --nx * --ny * --nz must equal the process count, here 1.
 */
#define _POSIX_C_SOURCE 200809L
#define _GNU_SOURCE
#include "pdfs.h"
#include "pdfs_nc.h"
#include "pdfs_vti.h"

#include <dlfcn.h>
#include <getopt.h>
#include <inttypes.h>
#include <libgen.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

typedef int (*nc_read_fn)(const char *, pdfs_nc_field *);
typedef void (*nc_free_fn)(pdfs_nc_field *);
typedef int (*nc_write_fn)(const char *, const int *, const double *, int, int, int, double,
                           double, double);

static void nc_usage(const char *prog)
{
  printf("\nUsage: %s -i <filename.nc> -p <outputPrefix> [--nx <val>] [--ny "
         "<val>] [--nz <val>]\n\n",
         prog);
}

/* lib/libpdfs_nc.so next to this program's lib/libpdfs.so */
static void *nc_open_lib(void)
{
  char exe[4096];
  ssize_t n = readlink("/proc/self/exe", exe, sizeof exe - 1);
  if (n <= 0) return NULL;
  exe[n] = 0;
  char path[4200];
  snprintf(path, sizeof path, "%s/../lib/libpdfs_nc.so", dirname(exe));
  return dlopen(path, RTLD_NOW | RTLD_LOCAL);
}

static int nc_main(int argc, char *argv[])
{
  /* checkArgs (main.c:172-225), plus the solver extensions */
  char *ipfile = NULL;
  char opfile[1024];
  size_t blk[3] = {1, 1, 1};
  int iflag = 0, pflag = 0, ok = 1, passes = 1, precision = PDFS_F64_REF;
  static struct option lopts[] = {{"input", required_argument, NULL, 'i'},
                                  {"prefix", required_argument, NULL, 'p'},
                                  {"nx", required_argument, NULL, 'x'},
                                  {"ny", required_argument, NULL, 'y'},
                                  {"nz", required_argument, NULL, 'z'},
                                  {"passes", required_argument, NULL, 'P'},
                                  {"converge", no_argument, NULL, 'C'},
                                  {"fp32", no_argument, NULL, 'F'},
                                  {0, 0, 0, 0}};
  int opt;
  while (ok && (opt = getopt_long(argc, argv, "i:p:x:y:z", lopts, NULL)) != -1) {
    switch (opt) {
      case 'i':
        ipfile = optarg;
        iflag = 1;
        break;
      case 'p':
        snprintf(opfile, sizeof opfile, "%s_soln.nc", optarg);
        pflag = 1;
        break;
      case 'x':
      case 'y':
      case 'z': {
        const size_t v = optarg ? (size_t)strtoumax(optarg, NULL, 10) : 0;
        if (v == 0) ok = 0;
        blk[opt - 'x'] = v;
        break;
      }
      case 'P':
        passes = atoi(optarg);
        break;
      case 'C':
        passes = 0;
        break;
      case 'F':
        precision = PDFS_F32;
        break;
      case '?':
        break;
      default:
        printf("Invalid option.\n");
        ok = 0;
    }
  }
  if (ok && blk[0] * blk[1] * blk[2] != 1) {
    printf("Total processes must equal total blocks in the decmposition!\n");
    ok = 0;
  }
  if (!ok || !iflag || !pflag) {
    nc_usage(argv[0]);
    return 0;
  }
  printf("\n------------------------\n");
  printf(" Input: %s\n", ipfile);
  printf("Output: %s\n", opfile);
  printf("------------------------\n");
  printf("nx: %d, ny: %d, nz: %d\n", (int)blk[0], (int)blk[1], (int)blk[2]);
  printf("------------------------\n");

  void *lib = nc_open_lib();
  nc_read_fn nc_read = lib ? (nc_read_fn)dlsym(lib, "pdfs_nc_read") : NULL;
  nc_free_fn nc_free = lib ? (nc_free_fn)dlsym(lib, "pdfs_nc_free") : NULL;
  nc_write_fn nc_write = lib ? (nc_write_fn)dlsym(lib, "pdfs_nc_write") : NULL;
  if (!nc_read || !nc_free || !nc_write) {
    printf("\nError: NetCDF-4 support (lib/libpdfs_nc.so) is not available: %s\n", dlerror());
    exit(2);
  }
  pdfs_nc_field in;
  int rc = nc_read(ipfile, &in);
  if (rc) {
    printf("\nError: %s: %s\n", ipfile,
           rc == -1 ? "cannot open the NetCDF-4 file"
                    : rc == -3 ? "out of memory" : "no Flags / Distance variables over x, y, z with delta");
    exit(2);
  }
  /* preProcessFile: padded output, interior copied, border Flags -1 and
   * Distance +inf (process_file.c:78-120, 204-333) */
  const int X = in.x + 2, Y = in.y + 2, Z = in.z + 2;
  const size_t np = (size_t)X * Y * Z;
  int *flg = (int *)malloc(np * sizeof(int));
  double *dst = (double *)malloc(np * sizeof(double));
  if (!flg || !dst) {
    printf("Error allocating memory for the Phi function.\n");
    exit(1);
  }
  for (int k = 0; k < Z; k++)
    for (int j = 0; j < Y; j++)
      for (int i = 0; i < X; i++) {
        const size_t p = ((size_t)k * Y + j) * X + i;
        if (i == 0 || j == 0 || k == 0 || i == X - 1 || j == Y - 1 || k == Z - 1) {
          flg[p] = -1;
          dst[p] = INFINITY;
        } else {
          const size_t q = ((size_t)(k - 1) * in.y + (j - 1)) * in.x + (i - 1);
          flg[p] = in.flags[q];
          dst[p] = in.distance[q];
        }
      }
  printf("BOX_X: %d, BOX_Y: %d, BOX_Z: %d\n", X, Y, Z);
  Phi pf;
  pf.x = in.x;
  pf.y = in.y;
  pf.z = in.z;
  pf.dx = in.dx;
  pf.dy = in.dy;
  pf.dz = in.dz;
  pf.location = flg;
  pf.distance = dst;
  pf.F = 1;
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  rc = pdfs_calc_dist_field(&pf, passes, precision);
  clock_gettime(CLOCK_MONOTONIC, &t1);
  if (rc != PDFS_OK) {
    printf("\nHIP failure: '%s'\n", pdfs_last_error());
    exit(EXIT_FAILURE);
  }
  printf("Parallel MPI/OpenACC Block FSM time: %f s.\n",
         (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec));
  rc = nc_write(opfile, flg, dst, X, Y, Z, in.dx, in.dy, in.dz);
  if (rc) {
    printf("\nError: %s: cannot write the NetCDF-4 file\n", opfile);
    exit(2);
  }
  nc_free(&in);
  free(flg);
  free(dst);
  return 0;
}

int main(int argc, char *argv[])
{
  if (argc >= 2 && argv[1][0] == '-') return nc_main(argc, argv);
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
