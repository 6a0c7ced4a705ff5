// Test driver: runs the SA flat plate through the C++ adapters of
// include/structured_gpu_adapter.hpp the way the reference's Solver::step
// drives its equation and linear solver (src/solver/solver.cpp:52-197): per
// counter calc_dt, calc_residual, then set_lhs / set_rhs / solve_and_update
// (preallocate at counter 0).  The mesh, solution and configuration are
// plain structs with the reference's member names.
//
//   adapter_driver <grid.p3d> <ni> <nj> <n_up> <mach> <re> <steps> <cfl> <out_prefix>
//
// writes <out_prefix>r0.bin (R at the initial state), <out_prefix>q.bin (q
// after the steps) and prints one line of norms per counter.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <memory>
#include <string>
#include <vector>

#include "structured_gpu_adapter.hpp"

struct Array {
  std::vector<double> v;
  double* data() { return v.data(); }
  const double* data() const { return v.data(); }
  size_t size() const { return v.size(); }
};
struct Solution {
  size_t nq = 4, ntrans = 1;
  Array q, dt, rhs;
};
struct Mesh {
  size_t ni, nj, nic, njc;
  Array xv, yv;
  std::shared_ptr<Solution> solution;
};
struct ConfigFreestream {
  double rho_inf, u_inf, v_inf, p_inf, mu_inf, pr_inf, T_inf, aoa;
};
struct ConfigSolver {
  size_t order, lhs_order;
  std::string flux;
  double dpdx, dpdy;
};
struct ConfigGeometry {
  size_t tail;
};
struct Config {
  std::shared_ptr<ConfigFreestream> freestream;
  std::shared_ptr<ConfigSolver> solver;
  std::shared_ptr<ConfigGeometry> geometry;
};

static void write_bin(const std::string& path, const Array& a) {
  std::ofstream f(path, std::ios::binary);
  f.write(reinterpret_cast<const char*>(a.data()), (std::streamsize)(a.size() * sizeof(double)));
}

int main(int argc, char** argv) {
  if (argc != 10) {
    std::fprintf(stderr, "usage: adapter_driver grid.p3d ni nj n_up mach re steps cfl out_prefix\n");
    return 2;
  }
  const int ni = std::atoi(argv[2]), nj = std::atoi(argv[3]), n_up = std::atoi(argv[4]);
  const double mach = std::atof(argv[5]), re = std::atof(argv[6]);
  const int steps = std::atoi(argv[7]);
  const double cfl = std::atof(argv[8]);
  const std::string out = argv[9];

  auto mesh = std::make_shared<Mesh>();
  mesh->ni = ni; mesh->nj = nj; mesh->nic = ni - 1; mesh->njc = nj - 1;
  mesh->xv.v.resize((size_t)ni * nj);
  mesh->yv.v.resize((size_t)ni * nj);
  if (sg_read_grid(argv[1], "p3d", ni, nj, mesh->xv.data(), mesh->yv.data()) != SG_OK) {
    std::fprintf(stderr, "grid: %s\n", sg_last_error(nullptr));
    return 3;
  }
  auto sol = std::make_shared<Solution>();
  const size_t n = mesh->nic * mesh->njc * (sol->nq + sol->ntrans);
  sol->q.v.resize(n); sol->dt.v.resize(n); sol->rhs.v.resize(n);
  mesh->solution = sol;

  auto cfg = std::make_shared<Config>();
  cfg->freestream = std::make_shared<ConfigFreestream>(
      ConfigFreestream{1.0, mach, 0.0, 1.0 / 1.4, mach / re, 0.72, 1.0 / 1.4, 0.0});
  cfg->solver = std::make_shared<ConfigSolver>(ConfigSolver{2, 1, "roe", 0.0, 0.0});
  cfg->geometry = std::make_shared<ConfigGeometry>(ConfigGeometry{1});
  // the flat plate's [[boundary]] list, in file order
  std::vector<sg_bc> bcs = {{SG_BC_SLIPWALL, SG_FACE_BOTTOM, 1, n_up, 0, 0, 0},
                            {SG_BC_WALL, SG_FACE_BOTTOM, n_up + 1, -2, 0, 0, 0},
                            {SG_BC_FREESTREAM, SG_FACE_TOP, 0, -1, 0, 0, 0},
                            {SG_BC_FREESTREAM, SG_FACE_LEFT, 0, -1, 0, 0, 0},
                            {SG_BC_OUTFLOW, SG_FACE_RIGHT, 0, -1, 0, 0, 0}};
  try {
    sg_adapter::TransportEquationGpu<Mesh, Config> transport(mesh, cfg, 3.0);
    sg_adapter::EulerEquationGpu<Mesh, Config> equation(mesh, cfg, bcs, &transport);
    sg_adapter::LinearSolverGpu<Mesh> linearsolver(mesh, equation.handle(), 6, 30, 60, 1e-3, 1, 32);
    equation.initialize();
    for (int counter = 0; counter <= steps; ++counter) {
      equation.calc_dt(cfl);
      equation.calc_residual(sol->q, sol->rhs, true);
      if (counter == 0) write_bin(out + "r0.bin", sol->rhs);
      double l2[5] = {0, 0, 0, 0, 0};
      for (size_t c = 0; c < n / 5; ++c)
        for (int k = 0; k < 5; ++k) l2[k] += sol->rhs.v[c * 5 + k] * sol->rhs.v[c * 5 + k];
      std::printf("counter %d norms %.6e %.6e %.6e %.6e %.6e\n", counter, std::sqrt(l2[0]), std::sqrt(l2[1]),
                  std::sqrt(l2[2]), std::sqrt(l2[3]), std::sqrt(l2[4]));
      if (counter == steps) break;
      if (counter == 0) linearsolver.preallocate(0);
      linearsolver.set_lhs(0, nullptr, nullptr, nullptr);
      linearsolver.set_rhs(sol->rhs.data());
      linearsolver.solve_and_update(sol->q.data(), 1.0);
    }
    write_bin(out + "q.bin", sol->q);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "adapter: %s\n", e.what());
    return 4;
  }
  return 0;
}
