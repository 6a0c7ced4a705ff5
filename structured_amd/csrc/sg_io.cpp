// Files of the reference: the grid loaders of Mesh (src/utils/mesh.cpp:133-169)
// and the restart, surface and Tecplot writers of IOManager
// (src/utils/io.cpp:49-255).  Host code; the surface writer takes its values
// from the device wall reduction (sg_wall).  Text is written with the same
// std::ostream defaults as the reference (6 significant digits, '%g' style),
// so the files read the same.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "sg_context.hpp"

namespace {

int io_fail(sg_ctx* ctx, const std::string& msg) { return sg_fail(ctx, SG_ERR_ARG, msg); }

}  // namespace

extern "C" {

int sg_read_grid(const char* path, const char* format, int ni, int nj, double* xv, double* yv) {
  if (!path || !format || !xv || !yv || ni < 2 || nj < 2) return SG_ERR_ARG;
  std::ifstream in(path);
  if (!in) return sg_fail(nullptr, SG_ERR_ARG, std::string("cannot open grid file ") + path);
  // vertex (i, j) is stored at i * nj + j (Mesh::xv[ni][nj]); both formats
  // list the vertices j-outer, i-inner
  if (!std::strcmp(format, "p3d")) {
    long nblock = 0, ni_ = 0, nj_ = 0;
    if (!(in >> nblock >> ni_ >> nj_) || nblock != 1 || ni_ != ni || nj_ != nj)
      return sg_fail(nullptr, SG_ERR_ARG, "p3d header is not '1 ni nj' for the given ni, nj");
    for (double* a : {xv, yv})
      for (int j = 0; j < nj; ++j)
        for (int i = 0; i < ni; ++i)
          if (!(in >> a[(size_t)i * nj + j])) return sg_fail(nullptr, SG_ERR_ARG, "p3d grid file truncated");
  } else if (!std::strcmp(format, "simple")) {
    for (int j = 0; j < nj; ++j)
      for (int i = 0; i < ni; ++i)
        if (!(in >> xv[(size_t)i * nj + j] >> yv[(size_t)i * nj + j]))
          return sg_fail(nullptr, SG_ERR_ARG, "simple grid file truncated");
  } else {
    return sg_fail(nullptr, SG_ERR_ARG, "grid format must be \"p3d\" or \"simple\"");
  }
  return SG_OK;
}

int sg_write_restart(sg_ctx* ctx, const char* path, const double* q_aos) {
  if (!ctx || !path || !q_aos) return SG_ERR_ARG;
  std::ofstream out(path, std::ios::binary);
  if (!out) return io_fail(ctx, std::string("cannot write ") + path);
  out.write(reinterpret_cast<const char*>(q_aos), (std::streamsize)(ctx->ncell() * ctx->nvar * sizeof(double)));
  return out ? SG_OK : io_fail(ctx, std::string("write failed: ") + path);
}

int sg_read_restart(sg_ctx* ctx, const char* path, double* q_aos) {
  if (!ctx || !path || !q_aos) return SG_ERR_ARG;
  std::ifstream in(path, std::ios::binary);
  if (!in) return io_fail(ctx, std::string("cannot open restart file ") + path);
  in.seekg(0, in.end);
  const long long size = (long long)in.tellg();
  const long long expect = (long long)(ctx->ncell() * ctx->nvar * sizeof(double));
  if (size != expect) {
    std::ostringstream m;
    m << "restart file " << path << " holds " << size << " bytes, expected " << expect << " (nic*njc*nvar doubles)";
    return io_fail(ctx, m.str());
  }
  in.seekg(0);
  in.read(reinterpret_cast<char*>(q_aos), (std::streamsize)expect);
  return in ? SG_OK : io_fail(ctx, std::string("read failed: ") + path);
}

int sg_write_surface(sg_ctx* ctx, const char* path, const double* q_aos, double* forces) {
  if (!ctx || !path || !q_aos) return SG_ERR_ARG;
  const int nb = ctx->geo.ni - 2 * ctx->opts.tail + 1;
  if (nb <= 0) return io_fail(ctx, "no surface cells for this tail");
  std::vector<double> x(nb), cp(nb), cf(nb), f(6);
  int rc = sg_wall(ctx, q_aos, x.data(), cp.data(), cf.data(), f.data());
  if (rc) return rc;
  std::ofstream out(path);
  if (!out) return io_fail(ctx, std::string("cannot write ") + path);
  for (int m = 0; m < nb; ++m) out << x[m] << " " << cp[m] << " " << cf[m] << std::endl;
  if (forces) std::copy(f.begin(), f.end(), forces);
  return out ? SG_OK : io_fail(ctx, std::string("write failed: ") + path);
}

int sg_write_tecplot(sg_ctx* ctx, const char* path, const double* q_aos) {
  if (!ctx || !path || !q_aos) return SG_ERR_ARG;
  std::ofstream out(path);
  if (!out) return io_fail(ctx, std::string("cannot write ") + path);
  const int nic = ctx->geo.nic, njc = ctx->geo.njc, nv = ctx->nvar, naux = 10;
  static const char* names[4] = {"rho", "rhou", "rhov", "rhoE"};
  out << "title = \"Solution\"" << "\n";
  out << "variables = \"x\" \"y\"";
  for (int k = 0; k < 4; ++k) out << " \"" << names[k] << "\"";
  for (int k = 4; k < nv; ++k) out << " \"psi_0\"";
  for (int k = 0; k < naux; ++k) out << " \"q_aux_" << k << "\"";
  out << "\n";
  out << "zone i=" << nic << ", j=" << njc << ", f=point\n";
  for (int j = 0; j < njc; ++j)
    for (int i = 0; i < nic; ++i) {
      const size_t c = (size_t)i * njc + j;
      out << ctx->geo.xc[c] << " " << ctx->geo.yc[c] << " ";
      for (int k = 0; k < nv; ++k) out << q_aos[c * nv + k] << " ";
      for (int k = 0; k < naux; ++k) out << 0.0 << " ";
      out << "\n";
    }
  return out ? SG_OK : io_fail(ctx, std::string("write failed: ") + path);
}

}  // extern "C"
