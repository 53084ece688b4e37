// Driver for the reference build (test infrastructure; runs in the build
// container only, never on the GPU box).
//
// It links the reference solver's own translation units and calls them in
// the order of the reference main program (main.cpp:99-1357) on the grid and
// surface files of one generated case, with two changes: the inner-iteration
// and physical-step counts are arguments (the convergence break is disabled
// unless a tolerance is given), and fluid densities get the seeded noise the
// cases prescribe between the initial condition and the U1 -> U1_/U1q copy.  Arrays are dumped as raw
// little-endian binaries in Array.h layout for the fixture writer; the
// reference's own writers leave IBM_minus/plus00000.dat and the PLOT3D files
// BCMdata00001.x and qBCMdata<n_steps>.q in the case directory.
//
// Usage (cwd = case directory holding BCMgrid00000.dat and BCM_STL_fine.stl):
//   ref_driver <n_steps> <n_inner> <seed> <noise_amp> <dump_dir> [dump_stages] [tol]
// With tol > 0 the inner loop ends as main.cpp:1249-1320 ends it: after each
// inner iteration E[m] = max(E[m], er[m]) and er[m] /= E[m] (m = 1..5), and
// the loop breaks when all five are below tol or at iteration n_inner (the
// reference's iteration_end_step; its tolerance is 0.02).  The inner-iteration
// count of every step is written to iters.txt.
// With dump_dir "-" nothing is written; the wall time of the time loop is
// printed as "loop_seconds <s>" (the CPU baseline of the benchmark).
#include <math.h>
#include <mpi.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>

#include "Resolution.h"

typedef double (*State)[X_size][Y_size][Z_size][Ndim];
typedef int (*Flags)[X_size][Y_size][Z_size];
typedef int (*Adj)[5][7];

// Reference entry points (signatures of the definitions in the reference tree).
void BCM_Grid(int, int, int, double*, int*, double*, double*, double*, double (*)[X_size],
              double (*)[Y_size], double (*)[Z_size], Flags, Adj, int*);
void BCM_Interface_table(int, int, int, int*, int*, int*, int*, int*, int*, int*, int*, int*,
                         int*, int*, int*, int*, int*, int*, int*, Adj, int*, int*, int*, int*,
                         int*, int*, int*, int*, int*, int*, int*, int*, int*, int*, int*);
void BCM_FWS_Interface(int, int, int, int, int, int, int, int, int (*)[MPI_Ncube], int*, int*,
                       int*, int*, int*, int*, int*, int*, int*, int*, double*, double*, double*,
                       double*, int*, int*, int*, int*, Adj, int*, int*, int*, Flags);
void BCM_Initial_condition(int, int, int, int (*)[MPI_Ncube], Flags, State, State, State);
void BCM_Immersed_boundary(int, int, int, int*, double, double, double, double, double, double,
                           double*, int*, double*, double*, double*, double (*)[X_size],
                           double (*)[Y_size], double (*)[Z_size], Flags, int*);
void BCM_Reading_IBM(int, int*, int, double*, int*, int*);
void BCM_Interface(int, int, int, int, int, int, int, int, int, int, int, int, int, int, int, int,
                   int, int, int, int, int, int, int, int, int (*)[MPI_Ncube], int*, int*, int*,
                   int*, int*, int*, int*, int*, int*, int*, int*, int*, int*, int*, double*,
                   double*, double*, double*, int*, int*, int*, int*, double*, double*, double*,
                   double*, int*, int*, int*, int*, double*, double*, double*, double*, int*,
                   int*, int*, int*, int*, int*, int*, int*, int*, Adj, int*, int*, int*, int*,
                   int*, int*, int*, int*, int*, int*, int*, int*, int*, int*, int*, State);
void BCM_X_boundary_condition(int, int, int, int*, int*, State);
void BCM_Y_boundary_condition(int, int, int*, int*, State);
void BCM_Z_boundary_condition(int, int, int*, int*, State);
void BCM_Ghostcell_minus(int, int*, double*, int*, int*, Flags, State);
void BCM_Ghostcell_plus(int, int*, double*, int*, int*, Flags, State);
void BCM_Flux_XYZ_Viscous_Runge_kutta(int, int, int, double, double, double, Flags, int*, double*,
                                      State, State, State, State, State, State, State, double*);
void BCM_Output(int, int, int, int (*)[MPI_Ncube], State, State, double*, double (*)[X_size],
                double (*)[Y_size], double (*)[Z_size]);

// Globals the reference declares extern for its default arguments.
int Ncube, N_wallcube, NBC, NBC_plus, NBC_minus, MPI_Nadj;
int NXbc_l, NXbc_u, NYbc_l, NYbc_u, NZbc_l, NZbc_u;
int Max_nei_eq, Max_nei_sb, Max_nei_bs, Ncpu_eq, Ncpu_sb, Ncpu_bs;

static double noise_unit(uint64_t seed, int cube, int i, int j, int k) {
  uint64_t key = (uint64_t)cube * 8000ull + (uint64_t)((i * 20 + j) * 20 + k);
  uint64_t z = seed * 0x9E3779B97F4A7C15ull + key + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z = z ^ (z >> 31);
  return (double)(z >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
}

static std::string g_dump;
static bool g_timing = false;  // dump_dir "-": time the loop, write nothing
static void dump(const char* name, const void* p, size_t bytes) {
  if (g_timing) return;
  std::string path = g_dump + "/" + name;
  FILE* f = fopen(path.c_str(), "wb");
  if (!f) { fprintf(stderr, "cannot write %s\n", path.c_str()); exit(3); }
  fwrite(p, 1, bytes, f);
  fclose(f);
}

static int read_count(const char* file, const char* marker) {
  FILE* f = fopen(file, "r");
  if (!f) { fprintf(stderr, "cannot open %s\n", file); exit(3); }
  char str[1024] = "";
  while (strcmp(str, marker) != 0) {
    if (fscanf(f, "%[^\n]\n", str) != 1) { fprintf(stderr, "marker missing in %s\n", file); exit(3); }
  }
  int n = 0;
  if (fscanf(f, "%d", &n) != 1) exit(3);
  fclose(f);
  return n;
}

int main(int argc, char** argv) {
  if (argc < 6) {
    fprintf(stderr, "usage: ref_driver n_steps n_inner seed noise_amp dump_dir [dump_stages] [tol]\n");
    return 2;
  }
  const int n_steps = atoi(argv[1]);
  const int n_inner = atoi(argv[2]);
  const uint64_t seed = strtoull(argv[3], nullptr, 10);
  const double noise_amp = atof(argv[4]);
  g_dump = argv[5];
  g_timing = g_dump == "-";
  const int dump_stages = argc > 6 ? atoi(argv[6]) : 0;
  const double tol = argc > 7 ? atof(argv[7]) : 0.0;
  MPI_Init(&argc, &argv);
  int myid = 0;

  const double deltaT = 0.0001;
  const double deltaTau = deltaT / 200.;
  const double e = 0.1;
  double* er = new double[8]();

  Ncube = read_count("BCMgrid00000.dat", "#_of_cubes in the computational region >>") + 1;
  N_wallcube = read_count("BCMgrid00000.dat", "#_of_wall_cubes >>") + 1;
  const int ncube = Ncube, n_wallcube = N_wallcube;
  if (ncube - 1 != MPI_Ncube) { fprintf(stderr, "MPI_Ncube mismatch\n"); return 2; }

  State U1_ = new double[ncube][X_size][Y_size][Z_size][Ndim]();
  State U1 = new double[ncube][X_size][Y_size][Z_size][Ndim]();
  State U1q = new double[ncube][X_size][Y_size][Z_size][Ndim]();
  State U1p1 = new double[ncube][X_size][Y_size][Z_size][Ndim]();
  State U1p2 = new double[ncube][X_size][Y_size][Z_size][Ndim]();
  State Residual1 = new double[ncube][X_size][Y_size][Z_size][Ndim]();
  State Rku1 = new double[ncube][X_size][Y_size][Z_size][Ndim]();
  double* cube_size = new double[ncube]();
  int* csl = new int[ncube]();
  double *Xcube = new double[ncube](), *Ycube = new double[ncube](), *Zcube = new double[ncube]();
  double(*Xcnt)[X_size] = new double[ncube][X_size]();
  double(*Ycnt)[Y_size] = new double[ncube][Y_size]();
  double(*Zcnt)[Z_size] = new double[ncube][Z_size]();
  Flags FWS = new int[ncube][X_size][Y_size][Z_size]();
  Adj adj_number = new int[ncube][5][7]();
  int* wallcube = new int[n_wallcube + 1]();
  int* lists[15];
  for (int l = 0; l < 15; ++l) lists[l] = new int[ncube + 1]();
  int nadj[15] = {0};

  BCM_Grid(myid, ncube, n_wallcube, cube_size, csl, Xcube, Ycube, Zcube, Xcnt, Ycnt, Zcnt, FWS,
           adj_number, wallcube);

  // Domain bounds and boundary-cube lists (main.cpp:108-189).
  double Xmax = -100000000.0, Xmin = 100000000.0, Ymax = -100000000.0, Ymin = 100000000.0,
         Zmax = -100000000.0, Zmin = 100000000.0;
  for (int c = 1; c < ncube; ++c) {
    Xmin = fmin(Xmin, Xcube[c]); Xmax = fmax(Xmax, Xcube[c] + cube_size[c]);
    Ymin = fmin(Ymin, Ycube[c]); Ymax = fmax(Ymax, Ycube[c] + cube_size[c]);
    Zmin = fmin(Zmin, Zcube[c]); Zmax = fmax(Zmax, Zcube[c] + cube_size[c]);
  }
  int *Xbc_u = new int[ncube + 1](), *Xbc_l = new int[ncube + 1](), *Ybc_u = new int[ncube + 1]();
  int *Ybc_l = new int[ncube + 1](), *Zbc_u = new int[ncube + 1](), *Zbc_l = new int[ncube + 1]();
  NXbc_l = NXbc_u = NYbc_l = NYbc_u = NZbc_l = NZbc_u = 0;
  const double minimum = 0.0000001;
  for (int c = 1; c < ncube; ++c) {
    if (fabs(Xcube[c] + cube_size[c] - Xmax) < minimum) Xbc_u[++NXbc_u] = c;
    if (fabs(Ycube[c] + cube_size[c] - Ymax) < minimum) Ybc_u[++NYbc_u] = c;
    if (fabs(Zcube[c] + cube_size[c] - Zmax) < minimum) Zbc_u[++NZbc_u] = c;
    if (fabs(Xcube[c] - Xmin) < minimum) Xbc_l[++NXbc_l] = c;
    if (fabs(Ycube[c] - Ymin) < minimum) Ybc_l[++NYbc_l] = c;
    if (fabs(Zcube[c] - Zmin) < minimum) Zbc_l[++NZbc_l] = c;
  }

  BCM_Interface_table(myid, ncube, n_wallcube, &nadj[0], &nadj[1], &nadj[2], &nadj[3], &nadj[4],
                      &nadj[5], &nadj[6], &nadj[7], &nadj[8], &nadj[9], &nadj[10], &nadj[11],
                      &nadj[12], &nadj[13], &nadj[14], csl, adj_number, lists[0], lists[1],
                      lists[2], lists[3], lists[4], lists[5], lists[6], lists[7], lists[8],
                      lists[9], lists[10], lists[11], lists[12], lists[13], lists[14]);

  // Single rank: every MPI neighbour count is zero; the buffers are dummies.
  int(*rank_map)[MPI_Ncube] = new int[2][MPI_Ncube]();
  int dummy_i[64] = {0};
  double dummy_d[64] = {0};
  auto fws_interface = [&]() {
    BCM_FWS_Interface(myid, ncube, 1, 0, 0, nadj[0], nadj[1], nadj[2], rank_map, dummy_i, dummy_i,
                      dummy_i, dummy_i, dummy_i, dummy_i, dummy_i, dummy_i, dummy_i, dummy_i,
                      dummy_d, dummy_d, dummy_d, dummy_d, dummy_i, dummy_i, dummy_i, csl,
                      adj_number, lists[0], lists[1], lists[2], FWS);
  };
  auto interface = [&]() {
    BCM_Interface(myid, ncube, 1, 0, 0, 0, 0, 0, 0, nadj[0], nadj[1], nadj[2], nadj[3], nadj[4],
                  nadj[5], nadj[6], nadj[7], nadj[8], nadj[9], nadj[10], nadj[11], nadj[12],
                  nadj[13], nadj[14], rank_map, dummy_i, dummy_i, dummy_i, dummy_i, dummy_i,
                  dummy_i, dummy_i, dummy_i, dummy_i, dummy_i, dummy_i, dummy_i, dummy_i, dummy_i,
                  dummy_d, dummy_d, dummy_d, dummy_d, dummy_i, dummy_i, dummy_i, dummy_i, dummy_d,
                  dummy_d, dummy_d, dummy_d, dummy_i, dummy_i, dummy_i, dummy_i, dummy_d, dummy_d,
                  dummy_d, dummy_d, dummy_i, dummy_i, dummy_i, dummy_i, dummy_i, dummy_i,
                  dummy_i, dummy_i, csl, adj_number, lists[0], lists[1], lists[2], lists[3],
                  lists[4], lists[5], lists[6], lists[7], lists[8], lists[9], lists[10],
                  lists[11], lists[12], lists[13], lists[14], U1_);
  };
  fws_interface();
  BCM_Initial_condition(myid, ncube, n_wallcube, rank_map, FWS, U1_, U1, U1q);
  if (noise_amp != 0.0)
    for (int c = 1; c < ncube; ++c)
      for (int i = 0; i < X_size; ++i)
        for (int j = 0; j < Y_size; ++j)
          for (int k = 0; k < Z_size; ++k)
            if (FWS[c][i][j][k] == 1) {
              double fac = 1.0 + noise_amp * noise_unit(seed, c, i, j, k);
              U1[c][i][j][k][0] = U1[c][i][j][k][0] * fac;
            }
  for (int c = 1; c < ncube; ++c)
    for (int i = 0; i < X_size; ++i)
      for (int j = 0; j < Y_size; ++j)
        for (int k = 0; k < Z_size; ++k)
          for (int v = 0; v < Ndim; ++v) U1_[c][i][j][k][v] = U1q[c][i][j][k][v] = U1[c][i][j][k][v];

  // Ghost-cell detection (main.cpp:694-803).
  {
    Flags tmp = new int[ncube][X_size][Y_size][Z_size];
    memcpy(tmp, FWS, sizeof(int) * (size_t)ncube * X_size * Y_size * Z_size);
    for (int c = 1; c < ncube; ++c) {
      if (csl[c] != 0) continue;
      for (int i = 1; i <= 18; ++i)
        for (int j = 1; j <= 18; ++j)
          for (int k = 1; k <= 18; ++k)
            if (FWS[c][i][j][k] == -1 &&
                (FWS[c][i + 1][j][k] == 1 || FWS[c][i - 1][j][k] == 1 || FWS[c][i][j - 1][k] == 1 ||
                 FWS[c][i][j + 1][k] == 1 || FWS[c][i][j][k - 1] == 1 || FWS[c][i][j][k + 1] == 1))
              tmp[c][i][j][k] = -2;
    }
    for (int c = 1; c < ncube; ++c) {
      if (csl[c] != 0) continue;
      for (int i = 0; i < X_size; ++i)
        for (int j = 0; j < Y_size; ++j)
          for (int k = 0; k < Z_size; ++k)
            if (tmp[c][i][j][k] == -2) FWS[c][i][j][k] = 0;
    }
    delete[] tmp;
  }

  BCM_Immersed_boundary(myid, ncube, n_wallcube, &NBC, Xmin, Xmax, Ymin, Ymax, Zmin, Zmax,
                        cube_size, csl, Xcube, Ycube, Zcube, Xcnt, Ycnt, Zcnt, FWS, wallcube);
  interface();

  NBC_minus = read_count("IBM_minus00000.dat", "Number of Boundary Cells");
  NBC_plus = read_count("IBM_plus00000.dat", "Number of Boundary Cells");
  double* weight_minus = new double[NBC_minus * 8 + 1]();
  int* GCindex_minus = new int[NBC_minus * 4 + 1]();
  int* IPsur_minus = new int[NBC_minus * 4 + 1]();
  double* weight_plus = new double[NBC_plus * 8 + 1]();
  int* GCindex_plus = new int[NBC_plus * 4 + 1]();
  int* IPsur_plus = new int[NBC_plus * 4 + 1]();
  NBC = NBC_minus;
  BCM_Reading_IBM(myid, &NBC, -1, weight_minus, GCindex_minus, IPsur_minus);
  NBC = NBC_plus;
  BCM_Reading_IBM(myid, &NBC, 1, weight_plus, GCindex_plus, IPsur_plus);
  for (int n = 1; n <= NBC_minus; ++n) {
    int* g = &GCindex_minus[(n - 1) * 4 + 1];
    FWS[g[0]][g[1]][g[2]][g[3]] = 0;
  }
  for (int n = 1; n <= NBC_plus; ++n) {
    int* g = &GCindex_plus[(n - 1) * 4 + 1];
    FWS[g[0]][g[1]][g[2]][g[3]] = 0;
  }
  interface();
  fws_interface();
  const double rho0 = 1.1842, P0 = 101300;
  for (int pass = 0; pass < 2; ++pass) {
    int n_bc = pass == 0 ? NBC_minus : NBC_plus;
    int* gci = pass == 0 ? GCindex_minus : GCindex_plus;
    for (int n = 1; n <= n_bc; ++n) {
      int* g = &gci[(n - 1) * 4 + 1];
      for (State S : {U1, U1_}) {
        S[g[0]][g[1]][g[2]][g[3]][0] = rho0;
        S[g[0]][g[1]][g[2]][g[3]][1] = 0;
        S[g[0]][g[1]][g[2]][g[3]][2] = 0;
        S[g[0]][g[1]][g[2]][g[3]][3] = 0;
        S[g[0]][g[1]][g[2]][g[3]][4] = P0 / (1.4 - 1);
      }
    }
  }

  const size_t sbytes = sizeof(double) * (size_t)ncube * X_size * Y_size * Z_size * Ndim;
  const size_t fbytes = sizeof(int) * (size_t)ncube * X_size * Y_size * Z_size;
  dump("setup_FWS.bin", FWS, fbytes);
  dump("setup_U1_.bin", U1_, sbytes);
  dump("setup_U1.bin", U1, sbytes);
  dump("setup_U1q.bin", U1q, sbytes);
  {
    int bc_counts[6] = {NXbc_l, NXbc_u, NYbc_l, NYbc_u, NZbc_l, NZbc_u};
    dump("setup_bc_counts.bin", bc_counts, sizeof bc_counts);
    dump("setup_nadj.bin", nadj, sizeof nadj);
  }

  // Time loop (main.cpp:1146-1357) with a fixed inner-iteration count.
  FILE* fer = g_timing ? fopen("/dev/null", "w") : fopen((g_dump + "/er.txt").c_str(), "w");
  FILE* fit = (tol > 0 && !g_timing) ? fopen((g_dump + "/iters.txt").c_str(), "w") : nullptr;
  const double t_loop = MPI_Wtime();
  double E[6] = {0, 0, 0, 0, 0, 0};
  int stage = 0;
  for (int step = 1; step <= n_steps; ++step) {
    for (int iteration = 1; iteration <= n_inner; ++iteration) {
      for (int RK = 1; RK < 4; ++RK) {
        BCM_X_boundary_condition(myid, NXbc_l, NXbc_u, Xbc_l, Xbc_u, U1_);
        BCM_Y_boundary_condition(NYbc_l, NYbc_u, Ybc_l, Ybc_u, U1_);
        BCM_Z_boundary_condition(NZbc_l, NZbc_u, Zbc_l, Zbc_u, U1_);
        BCM_Ghostcell_minus(myid, &NBC_minus, weight_minus, GCindex_minus, IPsur_minus, FWS, U1_);
        BCM_Ghostcell_plus(myid, &NBC_plus, weight_plus, GCindex_plus, IPsur_plus, FWS, U1_);
        interface();
        BCM_Flux_XYZ_Viscous_Runge_kutta(myid, ncube, RK, deltaT, deltaTau, e, FWS, csl,
                                         cube_size, U1_, U1, U1q, U1p1, U1p2, Rku1, Residual1,
                                         er);
        interface();
        ++stage;
        if (stage <= dump_stages) {
          char name[64];
          snprintf(name, sizeof name, "stage%d_U1_.bin", stage);
          dump(name, U1_, sbytes);
          snprintf(name, sizeof name, "stage%d_Rku1.bin", stage);
          dump(name, Rku1, sbytes);
          snprintf(name, sizeof name, "stage%d_Residual1.bin", stage);
          dump(name, Residual1, sbytes);
        }
      }
      fprintf(fer, "%d %d %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", step, iteration, er[1],
              er[2], er[3], er[4], er[5], er[6], er[7]);
      bool below = true;
      for (int m = 1; m <= 5; ++m) {
        E[m] = E[m] > er[m] ? E[m] : er[m];  // the max macro of main.cpp:14
        er[m] = er[m] / E[m];
        below = below && er[m] < tol;
      }
      if (tol > 0 && (below || iteration == n_inner)) {
        if (fit) fprintf(fit, "%d %d\n", step, iteration);
        break;
      }
    }
    for (int c = 1; c < ncube; ++c)
      for (int i = 2; i <= 17; ++i)
        for (int j = 2; j <= 17; ++j)
          for (int k = 2; k <= 17; ++k)
            for (int v = 0; v < Ndim; ++v) {
              U1q[c][i][j][k][v] = U1[c][i][j][k][v];
              U1[c][i][j][k][v] = U1_[c][i][j][k][v];
            }
  }
  fclose(fer);
  if (fit) fclose(fit);
  if (g_timing) {
    // Wall time of the time loop alone (setup excluded), for CPU baselines.
    printf("loop_seconds %.6f\n", MPI_Wtime() - t_loop);
    MPI_Finalize();
    return 0;
  }
  // PLOT3D files of the last step, written as main.cpp:1346-1351 does after
  // the time-level shift (into the case directory).
  BCM_Output(myid, ncube, n_steps, rank_map, U1_, U1q, cube_size, Xcnt, Ycnt, Zcnt);
  dump("final_U1_.bin", U1_, sbytes);
  dump("final_U1.bin", U1, sbytes);
  dump("final_U1q.bin", U1q, sbytes);
  dump("final_U1p1.bin", U1p1, sbytes);
  dump("final_U1p2.bin", U1p2, sbytes);
  dump("final_Rku1.bin", Rku1, sbytes);
  dump("final_Residual1.bin", Residual1, sbytes);
  MPI_Finalize();
  return 0;
}
