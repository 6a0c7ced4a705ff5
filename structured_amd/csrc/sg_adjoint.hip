// Adjoint products and solves are implemented with the Newton-Krylov driver
// (sg_solver.cpp), which owns the Jacobian blocks they use, and the device
// kernels in sg_linalg.hip.  This translation unit is intentionally empty.
