// Internal C++ interface between the host context and the kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "sg_device.hpp"

namespace sg {

size_t residual_partial_count(const DevGeom& g);
// Norms (optional: all of partial, counter, l2 and parity, or none): partial
// holds 2 nv per-block partial sums, counter one device word (the arrival
// count), parity a host word that alternates per fused launch (this call
// reads and flips it).  counter must be 0, partial +0.0 and parity 0 when
// they are first used together, and they then stay with the grid they were
// used for.  The residual kernel's block that arrives last finalises the L2
// norms of R, then of V R, into l2 (device): l2_count entries (0: all 2 nv;
// nv: the norms of R only).
hipError_t launch_residual(const DevGeom& g, const DevPhys& ph, const DevBc& bc, bool sa, const double* q, double* r,
                           double* partial, unsigned long long* counter, double* l2, hipStream_t s,
                           const DevClosure& clo = DevClosure(), int l2_count = 0, unsigned* parity = nullptr);
// J.v (+ (dR/dbeta).beta_dot when beta_dot is not null; v may be all zero)
hipError_t launch_jv(const DevGeom& g, const DevPhys& ph, const DevBc& bc, bool sa, const double* q, const double* v,
                     double* r, double* jv, hipStream_t s, const double* beta_dot = nullptr,
                     const DevClosure& clo = DevClosure());
hipError_t launch_dt(const DevGeom& g, double mu_inf, int nv, const double* q, const double* dsmin, double cfl,
                     double* dt, hipStream_t s);
hipError_t launch_aos_to_soa(int nic, int njc, int ldi, int nv, const double* aos, double* soa, hipStream_t s);
hipError_t launch_soa_to_aos(int nic, int njc, int ldi, int nv, const double* soa, double* aos, hipStream_t s);
hipError_t launch_wall(const DevGeom& g, const DevPhys& ph, const DevBc& bc, bool sa, const double* q, int first,
                       int nb, const double* xv0, const double* yv0, double qinf, double* cp, double* cf, double* fsum,
                       hipStream_t s);

// tangent of the wall Cf along v (dcf has nb entries)
hipError_t launch_wall_tangent(const DevGeom& g, const DevPhys& ph, const DevBc& bc, bool sa, const double* q,
                               const double* v, int first, int nb, const double* xv0, const double* yv0, double qinf,
                               double* dcf, hipStream_t s);

}  // namespace sg
