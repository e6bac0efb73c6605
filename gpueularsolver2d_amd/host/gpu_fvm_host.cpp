// Host program with the reference's entry points, driving libfvm_hip.so.
//
// The reference's main (gpu_fvm.cu:27-68) runs inputMesh -> Initialization ->
// CalculationMeshGeo -> Cuda_Init -> ROE_SOLVER -> Node_Computing -> Output
// over global state.  This program keeps that sequence and those names and
// file formats; ROE_SOLVER packs the globals into the C-ABI descriptors and
// calls fvm_create / fvm_run / fvm_get_state.  Differences from the
// reference, by design:
//   - every function returns a status and main exits non-zero on failure
//     (the reference calls exit(0) on a missing file, gpu_init.cu:10-14);
//   - the final state written by Output is the solved state (the reference
//     writes its stale initial state, see the Q1 note in ROE_SOLVER);
//   - file names and scheme options come from the command line, defaulting
//     to the reference's hard-wired ones.
//
// Usage: fvm_solver [mesh.dat] [input.dat] [--fp32] [--plain] [--order 1|2]
//                   [--no-smooth] [--global-dt] [--resid-every N] [--quiet]
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "fvm_abi.h"

// Reference data model (gpu_fvm.cuh:14-118), reduced to what the host uses.
struct W {
  double density, density_U, density_V, density_E, P, T, A;
};
struct NODE {
  int NLog;
  double x, y, ROU, U, V, E, P, T, A, Mach, RoundArea;
};
struct EDGE {
  int ELog, left_cell, right_cell, farfieldid, wallid, node1, node2;
};
struct CELL {
  int Point[3];
};

// Reference globals (gpu_fvm.cu:5-12)
static int Ncell, Nnode, Nedge, MAXSTEP, WallNum, step, WallBoundNum, FarBoundNum;
static double MA, ALPHA, GAMA, PIN, TIN, CFL, EPSL, UI, VI;
static double ROUIN, AIN, VIN, PIS;
static const double R = 287.06;
static const double PI = 3.14159265358979;
static std::vector<double> CellArea, Resd;
static std::vector<CELL> cell;
static std::vector<NODE> node;
static std::vector<EDGE> WallEdge;
static std::vector<W> w;

struct Options {
  std::string mesh = "2822-3k.dat", input = "input.dat";
  fvm_scheme_opts scheme;
  int resid_every = 100;
  bool quiet = false;
};
static Options opt;

static int fail(const char* what, fvm_status s) {
  std::fprintf(stderr, "%s: %s (%s)\n", what, fvm_status_string(s), fvm_last_error());
  return (int)s;
}

// gpu_init.cu:5-47: mesh file and input.dat
int inputMesh() {
  FILE* fp = std::fopen(opt.mesh.c_str(), "r");
  if (!fp) {
    std::fprintf(stderr, "Cannot open the mesh data file %s\n", opt.mesh.c_str());
    return FVM_ERR_MESH;
  }
  bool ok = std::fscanf(fp, "%d %d", &Nnode, &Ncell) == 2 && Nnode >= 3 && Ncell >= 1;
  if (ok) {
    cell.assign(Ncell + 1, CELL{});
    node.assign(Nnode + 1, NODE{});
    for (int i = 1; ok && i <= Nnode; i++) ok = std::fscanf(fp, "%lf", &node[i].x) == 1;
    for (int i = 1; ok && i <= Nnode; i++) ok = std::fscanf(fp, "%lf", &node[i].y) == 1;
    for (int c = 1; ok && c <= Ncell; c++)
      for (int i = 0; ok && i < 3; i++) ok = std::fscanf(fp, "%d", &cell[c].Point[i]) == 1;
  }
  std::fclose(fp);
  if (!ok) {
    std::fprintf(stderr, "Mesh file %s is truncated or malformed\n", opt.mesh.c_str());
    return FVM_ERR_MESH;
  }
  fp = std::fopen(opt.input.c_str(), "r");
  if (!fp) {
    std::fprintf(stderr, "Cannot open input file %s\n", opt.input.c_str());
    return FVM_ERR_ARGUMENT;
  }
  ok = std::fscanf(fp, "%lf %lf %lf %lf %lf %lf %lf %d", &MA, &ALPHA, &GAMA, &PIN, &TIN, &CFL, &EPSL,
                   &MAXSTEP) == 8;
  std::fclose(fp);
  if (!ok || MAXSTEP < 0) {
    std::fprintf(stderr, "input file needs: MA ALPHA GAMA PIN TIN CFL EPSL MAXSTEP\n");
    return FVM_ERR_ARGUMENT;
  }
  if (!opt.quiet) std::printf("inputData finished!\n");
  return FVM_OK;
}

// gpu_init.cu:49-61
static void CalculationPTA(W& ww) {
  const double U = ww.density_U / ww.density, V = ww.density_V / ww.density;
  const double E = ww.density_E / ww.density - 0.5 * (U * U + V * V);
  ww.P = (GAMA - 1.0) * ww.density * E;
  ww.A = std::sqrt(GAMA * ww.P / ww.density);
  ww.T = (GAMA - 1.0) * E / R;
}

// gpu_init.cu:63-93: freestream constants and the initial state
int Initialization() {
  const double alpha = ALPHA * PI / 180.0;
  AIN = std::sqrt(GAMA * R * TIN);
  VIN = MA * AIN;
  ROUIN = PIN / (R * TIN);
  UI = std::cos(alpha);
  VI = std::sin(alpha);
  w.assign(Ncell + 1, W{});
  for (int c = 1; c <= Ncell; c++) {
    w[c].density = 1.0;
    w[c].density_U = std::cos(alpha);
    w[c].density_V = std::sin(alpha);
    w[c].density_E = R * TIN / (VIN * VIN) / (GAMA - 1.0) + 0.50;
    CalculationPTA(w[c]);
  }
  PIS = w[1].P;
  return FVM_OK;
}

static fvm_mesh_desc mesh_desc(std::vector<double>& x, std::vector<double>& y, std::vector<int32_t>& tri) {
  x.resize(Nnode);
  y.resize(Nnode);
  tri.resize(3 * (size_t)Ncell);
  for (int i = 1; i <= Nnode; i++) {
    x[i - 1] = node[i].x;
    y[i - 1] = node[i].y;
  }
  for (int c = 1; c <= Ncell; c++)
    for (int k = 0; k < 3; k++) tri[3 * (c - 1) + k] = cell[c].Point[k];
  return fvm_mesh_desc{Nnode, Ncell, x.data(), y.data(), tri.data()};
}

// gpu_init.cu:100-228: cell areas, edges, wall and far-field edges (the
// library's topology follows the reference's conventions exactly).
int CalculationMeshGeo() {
  std::vector<double> x, y;
  std::vector<int32_t> tri;
  fvm_mesh_desc m = mesh_desc(x, y, tri);
  int32_t counts[5];
  fvm_status s = fvm_mesh_topology(&m, counts, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                                   nullptr, nullptr, nullptr, nullptr);
  if (s != FVM_OK) return fail("CalculationMeshGeo", s);
  Nedge = counts[2];
  WallBoundNum = counts[3];
  FarBoundNum = counts[4];
  std::vector<int32_t> elog(Nedge), left(Nedge), right(Nedge), farid(Nedge), wallid(Nedge),
      n1(Nedge), n2(Nedge);
  s = fvm_mesh_topology(&m, counts, nullptr, nullptr, nullptr, elog.data(), left.data(), right.data(),
                        farid.data(), wallid.data(), n1.data(), n2.data());
  if (s != FVM_OK) return fail("CalculationMeshGeo", s);
  WallEdge.assign(WallBoundNum + 1, EDGE{});
  for (int e = 0; e < Nedge; e++)
    if (elog[e] == 2) {
      WallEdge[wallid[e]] = EDGE{2, left[e], right[e], farid[e], wallid[e], n1[e], n2[e]};
      node[n1[e]].NLog = 2;
      node[n2[e]].NLog = 2;
    }
  CellArea.assign(Ncell + 1, 0.0);
  for (int c = 1; c <= Ncell; c++) {
    const NODE &a = node[cell[c].Point[0]], &b = node[cell[c].Point[1]], &d = node[cell[c].Point[2]];
    CellArea[c] = 0.5 * ((d.y - a.y) * (b.x - a.x) - (b.y - a.y) * (d.x - a.x));
  }
  if (!opt.quiet) std::printf("Mesh data Computing's finished!!\n");
  return FVM_OK;
}

// gpu_solver.cu:2171-2473.  The iteration loop, kernels and transfers run in
// the library.  Q1: the reference's final AoS copies read the device state
// never updated by the solver (gpu_solver.cu:2464-2468); here w receives the
// solved state.
int ROE_SOLVER() {
  std::vector<double> x, y;
  std::vector<int32_t> tri;
  fvm_mesh_desc m = mesh_desc(x, y, tri);
  fvm_flow_params f{MA, ALPHA, GAMA, PIN, TIN, CFL};
  fvm_solver* s = nullptr;
  fvm_status st = fvm_create(&m, &f, &opt.scheme, &s);
  if (st != FVM_OK) return fail("fvm_create", st);
  Resd.assign(MAXSTEP + 1, 0.0);
  int32_t done = 0;
  st = fvm_run(s, MAXSTEP, EPSL, Resd.data(), &done);
  step = done;
  if (st != FVM_OK && st != FVM_ERR_NAN) {
    fvm_destroy(s);
    return fail("fvm_run", st);
  }
  if (st == FVM_ERR_NAN) std::fprintf(stderr, "warning: the residual became NaN at step %d\n", step);
  if (!opt.quiet)
    for (int i = 1; i <= step; i++) std::printf("%d,%1.20lf\n", i, Resd[i]);
  std::vector<double> rho(Ncell), ru(Ncell), rv(Ncell), re(Ncell);
  fvm_status g = fvm_get_state(s, rho.data(), ru.data(), rv.data(), re.data());
  double ms = 0;
  int64_t csu = 0;
  fvm_get_timing(s, &ms, &csu);
  double cl = 0, cd = 0;
  fvm_get_forces(s, &cl, &cd);
  fvm_destroy(s);
  if (g != FVM_OK) return fail("fvm_get_state", g);
  for (int c = 1; c <= Ncell; c++) {
    w[c].density = rho[c - 1];
    w[c].density_U = ru[c - 1];
    w[c].density_V = rv[c - 1];
    w[c].density_E = re[c - 1];
    CalculationPTA(w[c]);
  }
  std::printf("steps %d  Resd %.10e  device time %.3f ms  %.1f Mcell-stage-updates/s  Cl %.8f  Cd %.8f\n",
              step, step > 0 ? Resd[step] : 100.0, ms, ms > 0 ? csu / ms / 1e3 : 0.0, cl, cd);
  return st == FVM_ERR_NAN ? (int)FVM_ERR_NAN : (int)FVM_OK;
}

// gpu_fvm.cu:71-131: area-weighted cell-to-node averages
int Node_Computing() {
  WallNum = 0;
  for (int i = 1; i <= Nnode; i++) {
    NODE& n = node[i];
    n.RoundArea = n.P = n.ROU = n.T = n.U = n.V = n.E = n.A = 0.0;
    if (n.NLog == 2) WallNum++;
  }
  for (int c = 1; c <= Ncell; c++) {
    for (int i = 0; i < 3; i++) {
      NODE& n = node[cell[c].Point[i]];
      n.ROU += w[c].density * CellArea[c];
      n.U += w[c].density_U * CellArea[c];
      n.V += w[c].density_V * CellArea[c];
      n.E += w[c].density_E * CellArea[c];
      n.T += w[c].T * CellArea[c];
      n.P += w[c].P * CellArea[c];
      n.A += w[c].A * CellArea[c];
      n.RoundArea += CellArea[c];
    }
  }
  for (int i = 1; i <= Nnode; i++) {
    NODE& n = node[i];
    n.ROU /= n.RoundArea;
    n.U /= n.RoundArea;
    n.V /= n.RoundArea;
    n.E /= n.RoundArea;
    n.U /= n.ROU;
    n.V /= n.ROU;
    n.P /= n.RoundArea;
    n.T /= n.RoundArea;
    n.A /= n.RoundArea;
    n.Mach = std::sqrt(n.U * n.U + n.V * n.V) / n.A;
  }
  return FVM_OK;
}

// gpu_fvm.cu:133-196: Tecplot result, residual history and wall pressure
int Output() {
  FILE* fp = std::fopen("unresult.plt", "w");
  if (!fp) return FVM_ERR_ARGUMENT;
  std::fprintf(fp, "TITLE= \"Data\"\n");
  std::fprintf(fp, "VARIABLES = \"X\" \"Y\" \"Rou\" \"U\" \"V\" \"P\" \"T\" \"MA\" \n");
  std::fprintf(fp, "ZONE T=\"AirfoilMesh\"\n");
  std::fprintf(fp, "N=%d,E=%d, ZONETYPE=FETriangle\n", Nnode, Ncell);
  std::fprintf(fp, "DATAPACKING=POINT\n DT=(SINGLE SINGLE SINGLE SINGLE SINGLE SINGLE SINGLE SINGLE)\n");
  for (int i = 1; i <= Nnode; i++)
    std::fprintf(fp, "%20.10f %20.10f %20.10f %20.10f %20.10f %20.10f %20.10f %20.10f \n", node[i].x,
                 node[i].y, node[i].ROU, node[i].U, node[i].V, node[i].P, node[i].T, node[i].Mach);
  for (int c = 1; c <= Ncell; c++) {
    for (int i = 0; i < 3; i++) std::fprintf(fp, "%10d", cell[c].Point[i]);
    std::fprintf(fp, "\n");
  }
  std::fclose(fp);

  fp = std::fopen("unresidual.plt", "w");
  if (!fp) return FVM_ERR_ARGUMENT;
  std::fprintf(fp, "TITLE     = \"Dataset\"\nVARIABLES = \"Time\" \"Resid\" \n");
  std::fprintf(fp, "ZONE I=%d J=%d f=point \n", step, 1);
  for (int i = 1; i <= step; i++)
    if (i % opt.resid_every == 0) std::fprintf(fp, "%d %20.10lf\n", i, Resd[i]);
  std::fclose(fp);

  fp = std::fopen("unairfoilpress.plt", "w");
  if (!fp) return FVM_ERR_ARGUMENT;
  std::fprintf(fp, "TITLE     = \"Dataset\"\nVARIABLES = \"Position\" \"Pressure\" \n");
  std::fprintf(fp, "ZONE N=%d E=%d F=FEPOINT ET=Triangle\n", WallBoundNum * 2, WallBoundNum);
  for (int i = 1; i <= WallBoundNum; i++) {
    // -Cp as the reference writes it: -2 (p - PIN/(ROUIN VIN^2)), gpu_fvm.cu:182-185
    double pc = -2.0 * (node[WallEdge[i].node1].P - PIN / (ROUIN * VIN * VIN));
    std::fprintf(fp, "%20.10lf %20.10lf\n", node[WallEdge[i].node1].x, pc);
    pc = -2.0 * (node[WallEdge[i].node2].P - PIN / (ROUIN * VIN * VIN));
    std::fprintf(fp, "%20.10lf %20.10lf\n", node[WallEdge[i].node2].x, pc);
  }
  for (int i = 1; i <= WallBoundNum; i++) std::fprintf(fp, "%d %d %d\n", 2 * i - 1, 2 * i, 2 * i - 1);
  std::fclose(fp);
  return FVM_OK;
}

static int parse(int argc, char** argv) {
  fvm_default_opts(&opt.scheme);
  int positional = 0;
  for (int i = 1; i < argc; i++) {
    std::string a = argv[i];
    if (a == "--fp32") {
      opt.scheme.precision = FVM_FP32;
    } else if (a == "--plain") {
      opt.scheme.kernels = FVM_KERNELS_PLAIN;
    } else if (a == "--order" && i + 1 < argc) {
      opt.scheme.order = std::atoi(argv[++i]);
    } else if (a == "--no-smooth") {
      opt.scheme.smoothing = 0;
    } else if (a == "--global-dt") {
      opt.scheme.global_dt = 1;
    } else if (a == "--resid-every" && i + 1 < argc) {
      opt.resid_every = std::max(1, std::atoi(argv[++i]));
    } else if (a == "--quiet") {
      opt.quiet = true;
    } else if (!a.empty() && a[0] != '-' && positional < 2) {
      (positional++ == 0 ? opt.mesh : opt.input) = a;
    } else {
      std::fprintf(stderr, "unknown argument %s\n", a.c_str());
      return FVM_ERR_ARGUMENT;
    }
  }
  return FVM_OK;
}

int main(int argc, char** argv) {
  int st;
  if ((st = parse(argc, argv)) != FVM_OK) return st;
  if ((st = inputMesh()) != FVM_OK) return st;
  if ((st = Initialization()) != FVM_OK) return st;
  if ((st = CalculationMeshGeo()) != FVM_OK) return st;
  const int solved = ROE_SOLVER();
  if (solved != FVM_OK && solved != FVM_ERR_NAN) return solved;
  if ((st = Node_Computing()) != FVM_OK) return st;
  if ((st = Output()) != FVM_OK) return st;
  if (!opt.quiet) std::printf("All Computation's Finished!!\n");
  return solved;
}
