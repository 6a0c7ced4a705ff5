// CPU ORACLE — test infrastructure only.  A driver of this build's own that
// links the reference's unmodified sources (compiled from /root/reference by
// oracle/Makefile into oracle/_ref/) and evaluates them on a given state, so
// that the restatement in sg_oracle.hpp can be checked against the real code.
//
//   ref_harness <config.inp> res  <state.bin> <rhs.bin>
//   ref_harness <config.inp> dt   <state.bin> <cfl> <dt.bin>
//   ref_harness <config.inp> surf <state.bin>          (writes <label>.surface)
//   ref_harness <config.inp> forces <state.bin>  (as surf, with the console at
//              info level so the reference prints its Pressure / Viscous /
//              Total Cl Cd lines, src/utils/io.cpp:249-251, to stdout)
//   ref_harness <config.inp> init <state.bin> <out.bin>  (the state q that
//              Mesh::setup leaves after EulerEquation::initialize,
//              src/model/eulerequation.cpp:262-291; state.bin is not read)
//   ref_harness <config.inp> grad <state.bin> <out.bin> (grad_{u,v,T}_{chi,eta})
//   ref_harness <config.inp> solve <state.bin> <out.bin> (Solver::solve from the
//              state with the config's explicit scheme; writes the final q)
//
// States and outputs are raw native-endian float64 in the reference layout
// q[i][j][k] (the restart format of src/utils/io.cpp:134-152).  Run it from
// the directory that holds the config and grid files.
#include <cstdio>
#include <fstream>
#include <vector>

#include "common.h"
#include "config.h"
#include "eulerequation.h"
#include "mesh.h"
#include "solution.h"
#include "solver.h"

static std::vector<double> read_bin(const char* path, size_t n) {
  std::vector<double> v(n);
  std::ifstream f(path, std::ios::binary);
  f.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(n * sizeof(double)));
  if (!f) { std::fprintf(stderr, "short read %s\n", path); std::exit(3); }
  return v;
}
static void write_bin(const char* path, const double* p, size_t n) {
  std::ofstream f(path, std::ios::binary);
  f.write(reinterpret_cast<const char*>(p), (std::streamsize)(n * sizeof(double)));
}

int main(int argc, char** argv) {
  if (argc < 4) { std::fprintf(stderr, "usage: ref_harness cfg mode state ...\n"); return 2; }
  spdlog::set_pattern("%v");
  auto logger = spdlog::stdout_logger_mt("console", true);
  logger->set_level(spdlog::level::warn);
  auto config = std::make_shared<Config<double>>(argv[1], argc, argv);
  auto mesh = std::make_shared<Mesh<double, double>>(config);
  mesh->label = "";
  mesh->setup();
  auto sol = mesh->solution;
  const size_t nic = mesh->nic, njc = mesh->njc, nv = sol->nq + sol->ntrans;
  const size_t n = nic * njc * nv;
  std::string mode = argv[2];
  if (mode == "init") {
    write_bin(argv[4], sol->q.data(), n);
    return 0;
  }
  std::vector<double> q = read_bin(argv[3], n);
  for (size_t m = 0; m < n; ++m) sol->q.data()[m] = q[m];
  if (mode == "res") {
    mesh->equation->calc_residual(sol->q.const_ref(), sol->rhs);
    write_bin(argv[4], sol->rhs.data(), n);
  } else if (mode == "dt") {
    mesh->equation->calc_dt(std::atof(argv[4]));
    write_bin(argv[5], sol->dt.data(), n);
  } else if (mode == "surf") {
    mesh->equation->calc_residual(sol->q.const_ref(), sol->rhs);
    mesh->fluid_model->primvars(sol->q.const_ref(), sol->rho, sol->u, sol->v, sol->p, sol->T);
    mesh->iomanager->write_surface();
  } else if (mode == "forces") {
    mesh->equation->calc_residual(sol->q.const_ref(), sol->rhs);
    mesh->fluid_model->primvars(sol->q.const_ref(), sol->rho, sol->u, sol->v, sol->p, sol->T);
    logger->set_level(spdlog::level::info);
    mesh->iomanager->write_surface();
  } else if (mode == "grad") {
    mesh->equation->calc_residual(sol->q.const_ref(), sol->rhs);
    auto* e = mesh->equation.get();
    std::ofstream f(argv[4], std::ios::binary);
    for (auto* a : {&e->grad_u_chi, &e->grad_v_chi, &e->grad_T_chi, &e->grad_u_eta, &e->grad_v_eta, &e->grad_T_eta})
      f.write(reinterpret_cast<const char*>(a->data()), (std::streamsize)(a->size() * sizeof(double)));
  } else if (mode == "solve") {
    // EulerEquation::initialize leaves q_tmp = q (src/model/eulerequation.cpp:282-288)
    for (size_t m = 0; m < n; ++m) sol->q_tmp.data()[m] = q[m];
    auto solver = std::make_shared<Solver<double, double>>(config);
    solver->add_mesh(mesh);
    solver->solve();
    write_bin(argv[4], sol->q.data(), n);
  } else {
    std::fprintf(stderr, "unknown mode\n");
    return 2;
  }
  return 0;
}
