// On-disk formats of a BCM case: readers for the grid file, the ASCII STL
// surface and the IB tables; the IB-table and PLOT3D writers in the layouts
// the reference writes; the rank-decomposition files of a one-rank run; a case built from files; a restart snapshot of the
// two time levels.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "bcm_setup.h"

namespace bcm {

namespace {

struct File {
  FILE* f = nullptr;
  explicit File(const std::string& path, const char* mode) : f(fopen(path.c_str(), mode)) {}
  ~File() {
    if (f) fclose(f);
  }
};

bool skip_line(FILE* f) {
  char buf[1024];
  return fscanf(f, "%1023[^\n]\n", buf) == 1;
}

}  // namespace

// Grid file as BCM_Grid.cpp:63-261 reads it: a title line, the cube count,
// (size, corner) per cube, 4 adjacency slots per cube and direction, the cell
// count, the wall cubes, and the run-length coded flags of each wall cube in
// boustrophedon cell order (1 fluid, 0 solid; solid becomes -1).
int read_bcmgrid(const std::string& path, Case& cs, std::string& err) {
  File F(path, "r");
  if (!F.f) {
    err = "cannot open " + path;
    return 1;
  }
  FILE* f = F.f;
  char title[1024] = {0};
  int n = 0, ncell = 0, nwall = 0;
  if (fscanf(f, "%1023[^\n]\n", title) != 1 || !skip_line(f) || fscanf(f, "%d\n", &n) != 1 ||
      n < 1 || !skip_line(f)) {
    err = "grid file header unreadable";
    return 1;
  }
  cs.ncube = n + 1;
  cs.cube_size.assign(cs.ncube, 0.0);
  cs.Xcube.assign(cs.ncube, 0.0);
  cs.Ycube.assign(cs.ncube, 0.0);
  cs.Zcube.assign(cs.ncube, 0.0);
  for (int c = 1; c < cs.ncube; ++c)
    if (fscanf(f, "%lf\t%lf\t%lf\t%lf\n", &cs.cube_size[c], &cs.Xcube[c], &cs.Ycube[c],
               &cs.Zcube[c]) != 4) {
      err = "grid file cube list unreadable";
      return 1;
    }
  if (!skip_line(f)) {
    err = "grid file adjacency header missing";
    return 1;
  }
  cs.adj_number.assign((size_t)cs.ncube * 35, 0);
  for (int c = 1; c < cs.ncube; ++c)
    for (int d = 1; d <= 6; ++d) {
      int a[4];
      if (fscanf(f, "%d\t%d\t%d\t%d\n", &a[0], &a[1], &a[2], &a[3]) != 4) {
        err = "grid file adjacency unreadable";
        return 1;
      }
      for (int s = 0; s < 4; ++s) {
        if (a[s] < 0 || a[s] >= cs.ncube) {
          err = "grid file names a missing cube";
          return 1;
        }
        cs.adj_number[adj_index(c, s + 1, d)] = a[s];
      }
    }
  if (!skip_line(f) || fscanf(f, "%d\n", &ncell) != 1 || !skip_line(f) ||
      fscanf(f, "%d\n", &nwall) != 1 || !skip_line(f)) {
    err = "grid file cell/wall header unreadable";
    return 1;
  }
  if (ncell != NC * NC * NC || nwall < 0 || nwall > n) {
    err = "grid file counts do not fit 16^3-cell cubes";
    return 1;
  }
  cs.n_wallcube = nwall + 1;
  cs.wallcube.assign(cs.n_wallcube, 0);
  for (int w = 1; w < cs.n_wallcube; ++w)
    if (fscanf(f, "%d\n", &cs.wallcube[w]) != 1 || cs.wallcube[w] < 1 ||
        cs.wallcube[w] >= cs.ncube) {
      err = "grid file wall cube list unreadable";
      return 1;
    }
  if (nwall > 0 && !skip_line(f)) {
    err = "grid file flag header missing";
    return 1;
  }
  cs.fws.assign((size_t)cs.ncube * CELLS_PADDED, IFLUID);
  for (int k = 0; k < CELLS_PADDED; ++k) cs.fws[k] = 0;  // slot 0 is unused
  std::vector<int> seq(NC * NC * NC + 1);
  for (int w = 1; w < cs.n_wallcube; ++w) {
    const int c = cs.wallcube[w];
    int nch = 0, flag = 0;
    if (fscanf(f, "%d\n", &nch) != 1 || fscanf(f, "%d\n", &flag) != 1 || nch < 0) {
      err = "grid file flags unreadable";
      return 1;
    }
    if (nch == 0) {
      for (int i = ILO; i <= IHI; ++i)
        for (int j = ILO; j <= IHI; ++j)
          for (int k = ILO; k <= IHI; ++k) cs.fws[flag_index(c, i, j, k)] = flag;
      continue;
    }
    seq[1] = flag;
    int pos = 0, count = 0;
    if (fscanf(f, "%d", &pos) != 1) {
      err = "grid file flag changes unreadable";
      return 1;
    }
    for (int ii = 2; ii <= NC * NC * NC; ++ii) {
      if (ii == pos) {
        ++count;
        seq[ii] = abs(seq[ii - 1] - 1);
        if (count < nch && fscanf(f, "%d\n", &pos) != 1) {
          err = "grid file flag changes unreadable";
          return 1;
        }
      } else {
        seq[ii] = seq[ii - 1];
      }
    }
    for (int k = 1; k <= NC; ++k)
      for (int j = 1; j <= NC; ++j)
        for (int i = 1; i <= NC; ++i) {
          const int pp = j % 2, qq = k % 2;
          const int i1 = NC * (j - 1) + pp * i + (1 - pp) * (NC + 1 - i);
          const int i2 = NC * (NC - j) + pp * (NC + 1 - i) + (1 - pp) * i;
          const int ii = qq * i1 + (1 - qq) * i2 + NC * NC * (k - 1);
          cs.fws[flag_index(c, i + NB - 1, j + NB - 1, k + NB - 1)] = seq[ii];
        }
  }
  for (int c = 1; c < cs.ncube; ++c)
    for (int i = ILO; i <= IHI; ++i)
      for (int j = ILO; j <= IHI; ++j)
        for (int k = ILO; k <= IHI; ++k) {
          int& fl = cs.fws[flag_index(c, i, j, k)];
          if (fl == 0) fl = ISOLID;
        }
  return 0;
}

// ASCII STL: every "facet normal nx ny nz / outer loop / vertex x y z (x3)
// / endloop / endfacet" group is one triangle (BCM_Immersed_boundary.cpp:
// 198-275 reads the same fields in the same order).
int read_stl(const std::string& path, std::vector<Triangle>& tris, std::string& err) {
  File F(path, "r");
  if (!F.f) {
    err = "cannot open " + path;
    return 1;
  }
  tris.clear();
  char tok[1024];
  while (fscanf(F.f, "%1023s", tok) == 1) {
    if (strcmp(tok, "normal") != 0) continue;
    Triangle T;
    char s1[64], s2[64], s3[64], s4[64], s5[64];
    if (fscanf(F.f, "%lf %lf %lf %63s %63s %63s %lf %lf %lf %63s %lf %lf %lf %63s %lf %lf %lf",
               &T.n[0], &T.n[1], &T.n[2], s1, s2, s3, &T.v[0][0], &T.v[0][1], &T.v[0][2], s4,
               &T.v[1][0], &T.v[1][1], &T.v[1][2], s5, &T.v[2][0], &T.v[2][1],
               &T.v[2][2]) != 17) {
      err = "STL facet unreadable";
      return 1;
    }
    tris.push_back(T);
  }
  return 0;
}

// IB table files as BCM_Immersed_boundary.cpp:1009-1065 writes them: per
// entry the ghost cell (cube i j k), the interpolation base corner (cube i j
// k) and the 8 weights with "%f"; closed by "Number of Boundary Cells" and
// the entry count.
int write_ibm(const Case& cs, const std::string& dir) {
  for (int s = 0; s < 2; ++s) {
    const IBTable& tb = s == 0 ? cs.ib_minus : cs.ib_plus;
    const std::string path = dir + (s == 0 ? "/IBM_minus00000.dat" : "/IBM_plus00000.dat");
    File F(path, "w");
    if (!F.f) return 1;
    for (int n = 1; n <= tb.n; ++n) {
      const int* g = &tb.gcindex[(n - 1) * 4 + 1];
      const int* q = &tb.ipsur[(n - 1) * 4 + 1];
      const double* w = &tb.weight[(n - 1) * 8 + 1];
      fprintf(F.f, "%d\t%d\t%d\t%d\n", g[0], g[1], g[2], g[3]);
      fprintf(F.f, "%d\t%d\t%d\t%d\n", q[0], q[1], q[2], q[3]);
      fprintf(F.f, "%f\t%f\t%f\t%f\t%f\t%f\t%f\t%f\n", w[0], w[1], w[2], w[3], w[4], w[5], w[6],
              w[7]);
    }
    fprintf(F.f, "Number of Boundary Cells\n");
    fprintf(F.f, "%d", tb.n);
  }
  return 0;
}

// Reads one IB table file (BCM_Reading_IBM.cpp:48-75 and the entry count of
// main.cpp:866-874).
int read_ibm(const std::string& path, IBTable& tb, std::string& err) {
  File F(path, "r");
  if (!F.f) {
    err = "cannot open " + path;
    return 1;
  }
  // Entry count from the closing line.
  char line[1024];
  int n = -1;
  while (fgets(line, sizeof line, F.f))
    if (strncmp(line, "Number of Boundary Cells", 24) == 0) {
      if (fscanf(F.f, "%d", &n) != 1) n = -1;
      break;
    }
  if (n < 0) {
    err = "IB table count missing in " + path;
    return 1;
  }
  rewind(F.f);
  tb = IBTable();
  tb.n = n;
  tb.gcindex.assign((size_t)4 * n + 1, 0);
  tb.ipsur.assign((size_t)4 * n + 1, 0);
  tb.weight.assign((size_t)8 * n + 1, 0.0);
  for (int e = 1; e <= n; ++e) {
    int* g = &tb.gcindex[(e - 1) * 4 + 1];
    int* q = &tb.ipsur[(e - 1) * 4 + 1];
    double* w = &tb.weight[(e - 1) * 8 + 1];
    if (fscanf(F.f, "%d\t%d\t%d\t%d\n", &g[0], &g[1], &g[2], &g[3]) != 4 ||
        fscanf(F.f, "%d\t%d\t%d\t%d\t", &q[0], &q[1], &q[2], &q[3]) != 4 ||
        fscanf(F.f, "%lf\t%lf\t%lf\t%lf\t%lf\t%lf\t%lf\t%lf\n", &w[0], &w[1], &w[2], &w[3], &w[4],
               &w[5], &w[6], &w[7]) != 8) {
      err = "IB table entry unreadable in " + path;
      return 1;
    }
  }
  return 0;
}

// The three rank-decomposition files the reference reads at start-up, written
// for a decomposition into one rank that holds all mpi_ncube cubes:
//  - MPI_communication_table_in_turn.dat: free text up to the line
//    "Number of Cubes for connection  >>", then the count of cross-rank cube
//    connections (main.cpp:215-221 reads only that count): 0.
//  - Communication_irank.00000.dat: per transfer kind (equal size, small to
//    big, big to small) three title lines, the neighbour-rank count and then
//    one title line before each per-neighbour or per-cube list
//    (main.cpp:262-554): all counts 0, so every list is empty.
//  - rank_map.dat: a title line, then per global cube "global rank local"
//    (main.cpp:584-596); local numbers run from 1, and the last one of a rank
//    is its cube count (BCM_Output.cpp:68-77).
int write_rank_tables(int mpi_ncube, const std::string& dir) {
  {
    File F(dir + "/MPI_communication_table_in_turn.dat", "w");
    if (!F.f) return 1;
    fprintf(F.f, "Cube connections between ranks (one rank: none)\n");
    fprintf(F.f, "Number of Cubes for connection  >>\n0\n");
  }
  {
    File F(dir + "/Communication_irank.00000.dat", "w");
    if (!F.f) return 1;
    static const char* kinds[3] = {"equal size", "small to big", "big to small"};
    static const int lists[3] = {6, 7, 8};
    static const char* names[8] = {"neighbour ranks", "cubes per neighbour rank",
                                   "cubes sent",      "cubes received",
                                   "send directions", "receive directions",
                                   "adjacency slot",  "send adjacency slot"};
    for (int s = 0; s < 3; ++s) {
      fprintf(F.f, "# ------------------------------\n# %s neighbours of rank 0\n", kinds[s]);
      fprintf(F.f, "# ------------------------------\n0\n");
      for (int l = 0; l < lists[s]; ++l) fprintf(F.f, "# %s\n", names[l]);
    }
  }
  File F(dir + "/rank_map.dat", "w");
  if (!F.f) return 1;
  fprintf(F.f, "Cube(global)\tRank\tCube(local)\n");
  for (int g = 0; g < mpi_ncube; ++g) fprintf(F.f, "%d\t%d\t%d\n", g, 0, g + 1);
  return 0;
}

// Checks the rank-decomposition files present in dir against a single-rank
// run of a case of mpi_ncube cubes: no cross-rank connections, no neighbour
// ranks, and a rank map of exactly mpi_ncube cubes all on rank 0 with local
// numbers in 1..mpi_ncube.  Absent files are accepted.
int check_rank_tables(const std::string& dir, int mpi_ncube, std::string& err) {
  const std::string multi =
      " describes a decomposition over several ranks; one GPU runs a case decomposed into one "
      "rank (BCMgrid00000.dat holding every cube)";
  std::string path = dir + "/MPI_communication_table_in_turn.dat";
  if (File F(path, "r"); F.f) {
    char line[1024];
    bool found = false;
    while (fscanf(F.f, "%1023[^\n]\n", line) == 1) {
      size_t len = strlen(line);
      while (len && (line[len - 1] == ' ' || line[len - 1] == '\r')) line[--len] = 0;
      if (strcmp(line, "Number of Cubes for connection  >>") == 0) {
        found = true;
        break;
      }
    }
    int nadj = -1;
    if (!found || fscanf(F.f, "%d", &nadj) != 1 || nadj < 0) {
      err = "connection count unreadable in " + path;
      return 1;
    }
    if (nadj != 0) {
      err = path + multi;
      return 1;
    }
  }
  path = dir + "/Communication_irank.00000.dat";
  if (File F(path, "r"); F.f) {
    static const int lists[3] = {6, 7, 8};
    for (int s = 0; s < 3; ++s) {
      int ncpu = -1;
      bool ok = skip_line(F.f) && skip_line(F.f) && skip_line(F.f) && fscanf(F.f, "%d\n", &ncpu) == 1;
      if (!ok || ncpu < 0) {
        err = "neighbour-rank count unreadable in " + path;
        return 1;
      }
      if (ncpu != 0) {
        err = path + multi;
        return 1;
      }
      for (int l = 0; l < lists[s]; ++l)
        if (!skip_line(F.f)) {
          err = "title line missing in " + path;
          return 1;
        }
    }
  }
  path = dir + "/rank_map.dat";
  if (File F(path, "r"); F.f) {
    if (!skip_line(F.f)) {
      err = "title line missing in " + path;
      return 1;
    }
    int g, rank, local, rows = 0;
    while (fscanf(F.f, "%d %d %d", &g, &rank, &local) == 3) {
      if (rank != 0 || ++rows > mpi_ncube) {
        err = path + multi;
        return 1;
      }
      if (local < 1 || local > mpi_ncube) {
        err = "local cube number outside 1.." + std::to_string(mpi_ncube) + " in " + path;
        return 1;
      }
    }
    int ch;
    while ((ch = fgetc(F.f)) != EOF && (ch == ' ' || ch == '\t' || ch == '\r' || ch == '\n')) {
    }
    if (ch != EOF) {
      err = "rank map row unreadable in " + path;
      return 1;
    }
    if (rows != mpi_ncube) {
      err = path + " maps " + std::to_string(rows) + " cubes; the grid file holds " +
            std::to_string(mpi_ncube);
      return 1;
    }
  }
  return 0;
}

// A case from the reference's input files in dir: BCMgrid00000.dat and
// BCM_STL_fine.stl, with the rank-decomposition files checked where present.
// p supplies what the reference keeps in prm.h and its initial-condition
// noise: Mach, Reynolds number, reference length (p.sphere_diameter), seed
// and noise amplitude; the remaining setup
// (boundary lists, interface tables, flags, IB tables, initial state) runs as
// for a generated case.
int read_case(const std::string& dir, const CaseParams& p, Case& cs, std::string& err,
              IbDevice* dev) {
  cs = Case();
  cs.p = p;
  if (!(p.mach > 0) || !(p.reynolds > 0) || !(p.sphere_diameter > 0)) {
    err = "invalid case parameters";
    return 1;
  }
  compute_physics(p, cs.phys);
  if (read_bcmgrid(dir + "/BCMgrid00000.dat", cs, err)) return 1;
  if (check_rank_tables(dir, cs.ncube - 1, err)) return 1;
  if (read_stl(dir + "/BCM_STL_fine.stl", cs.stl, err)) return 1;
  grid_geometry(cs);
  cs.fws_grid = cs.fws;
  return finish_setup(cs, err, dev);
}

// PLOT3D multi-block grid and solution files as BCM_Output.cpp:86-301 writes
// them for one rank: BCMdata00001.x holds the block count, 18 x 18 x 18 per
// block and the cell-centre coordinates of cells 1..18 of each cube;
// qBCMdata<step>.q holds the same header, then per cube four 1.0 and the five
// conserved variables.  The reference writes the arrays with the first and
// last indices of its loops exchanged (X from k, Z from i, U[c][k][j][i]).
int write_plot3d(int ncube, const double* Xcnt, const double* Ycnt, const double* Zcnt,
                 int mpi_ncube, const double* U, int step, const std::string& dir,
                 std::string& err) {
  const int nout = NC + 2;
  const int n = ncube - 1;
  if (n < 1 || mpi_ncube < n) {
    err = "block count below the cube count";
    return 1;
  }
  std::vector<double> buf;
  buf.reserve((size_t)n * nout * nout * nout * 5 + 4 * (size_t)n);
  auto header = [&](FILE* f) {
    bool ok = fwrite(&mpi_ncube, sizeof(int), 1, f) == 1;
    const int dims[3] = {nout, nout, nout};
    for (int c = 0; c < mpi_ncube; ++c) ok = ok && fwrite(dims, sizeof(int), 3, f) == 3;
    return ok;
  };
  {
    File F(dir + "/BCMdata00001.x", "wb");
    if (!F.f) {
      err = "cannot write the PLOT3D grid file in " + dir;
      return 1;
    }
    for (int c = 1; c < ncube; ++c)
      for (int a = 0; a < 3; ++a)
        for (int i = NB - 1; i <= IHI + 1; ++i)
          for (int j = NB - 1; j <= IHI + 1; ++j)
            for (int k = NB - 1; k <= IHI + 1; ++k)
              buf.push_back(a == 0   ? Xcnt[(size_t)c * NP + k]
                            : a == 1 ? Ycnt[(size_t)c * NP + j]
                                     : Zcnt[(size_t)c * NP + i]);
    if (!header(F.f) || fwrite(buf.data(), sizeof(double), buf.size(), F.f) != buf.size()) {
      err = "short write of the PLOT3D grid file";
      return 1;
    }
  }
  buf.clear();
  char name[64];
  snprintf(name, sizeof name, "/qBCMdata%05d.q", step);
  File F(dir + name, "wb");
  if (!F.f) {
    err = "cannot write the PLOT3D solution file in " + dir;
    return 1;
  }
  for (int c = 1; c < ncube; ++c) {
    for (int m = 0; m < 4; ++m) buf.push_back(1.0);
    for (int v = 0; v < NV; ++v)
      for (int i = NB - 1; i <= IHI + 1; ++i)
        for (int j = NB - 1; j <= IHI + 1; ++j)
          for (int k = NB - 1; k <= IHI + 1; ++k) buf.push_back(U[aos_index(c, k, j, i, v)]);
  }
  if (!header(F.f) || fwrite(buf.data(), sizeof(double), buf.size(), F.f) != buf.size()) {
    err = "short write of the PLOT3D solution file";
    return 1;
  }
  return 0;
}

// Restart snapshot of everything a resumed run needs (the reference keeps
// none: its restart reader, BCM_Initial_condition.cpp:91-227, is disabled and
// its output holds U1_ only): magic, cube count, step, the five norm scales
// E1..E5, then U1_, U1 and U1q in the Array.h layout (ncube slots each).
namespace {
const char kRestartMagic[8] = {'B', 'C', 'M', 'R', 'S', 'T', '0', '1'};
}

int write_restart(const std::string& path, int ncube, long long step, const double* E5,
                  const double* U1_, const double* U1, const double* U1q, std::string& err) {
  File F(path, "wb");
  if (!F.f) {
    err = "cannot write " + path;
    return 1;
  }
  const size_t n = (size_t)ncube * CELLS_PADDED * NV;
  double e[5] = {0, 0, 0, 0, 0};
  if (E5) memcpy(e, E5, sizeof e);
  if (fwrite(kRestartMagic, 1, 8, F.f) != 8 || fwrite(&ncube, sizeof ncube, 1, F.f) != 1 ||
      fwrite(&step, sizeof step, 1, F.f) != 1 || fwrite(e, sizeof(double), 5, F.f) != 5 ||
      fwrite(U1_, sizeof(double), n, F.f) != n || fwrite(U1, sizeof(double), n, F.f) != n ||
      fwrite(U1q, sizeof(double), n, F.f) != n) {
    err = "short write of " + path;
    return 1;
  }
  return 0;
}

int read_restart(const std::string& path, int ncube, long long* step, double* E5, double* U1_,
                 double* U1, double* U1q, std::string& err) {
  File F(path, "rb");
  if (!F.f) {
    err = "cannot open " + path;
    return 1;
  }
  char magic[8];
  int nc = 0;
  long long st = 0;
  double e[5];
  if (fread(magic, 1, 8, F.f) != 8 || memcmp(magic, kRestartMagic, 8) != 0 ||
      fread(&nc, sizeof nc, 1, F.f) != 1 || fread(&st, sizeof st, 1, F.f) != 1 ||
      fread(e, sizeof(double), 5, F.f) != 5) {
    err = path + " is not a restart file";
    return 1;
  }
  if (nc != ncube) {
    err = "restart file holds " + std::to_string(nc) + " cube slots, not " + std::to_string(ncube);
    return 1;
  }
  const size_t n = (size_t)ncube * CELLS_PADDED * NV;
  if (fread(U1_, sizeof(double), n, F.f) != n || fread(U1, sizeof(double), n, F.f) != n ||
      fread(U1q, sizeof(double), n, F.f) != n) {
    err = "restart file truncated";
    return 1;
  }
  if (fgetc(F.f) != EOF) {
    err = "restart file longer than its header says";
    return 1;
  }
  if (step) *step = st;
  if (E5) memcpy(E5, e, sizeof e);
  return 0;
}

}  // namespace bcm
