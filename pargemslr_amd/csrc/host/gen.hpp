// Test-problem generators and deterministic random vectors.
#pragma once

#include <cstdint>

#include "csr.hpp"

namespace pgs {

// 5-/7-point Laplacian on an nx x ny x nz grid with Dirichlet boundary,
// following CooMatrixClass::Laplacian (ParGeMSLR/SRC/matrices/coo_matrix.cpp:519-640):
// diagonal (2D ? 4 : 6) - shift, +x/+y/+z neighbours -1 - alpha, -x/-y/-z
// neighbours -1 + alpha.  Rows sorted by column.
template <typename T>
Csr<T> laplacian(int nx, int ny, int nz, T shift, double ax, double ay, double az);

// 3D first-order upwind convection-diffusion on nx^3 interior points:
// diagonal 6 + 3c, upstream neighbour (-x,-y,-z) -1 - c, downstream -1.
Csr<double> convdiff_upwind(int nx, int ny, int nz, double c);

// Counter-based generator: splitmix64 of (seed, index).
uint64_t rng_u64(uint64_t seed, uint64_t idx);
// uniform in [lo, hi)
void rng_uniform(uint64_t seed, int64_t n, double lo, double hi, double* out);
// standard normal (Box-Muller on two uniforms per index)
void rng_normal(uint64_t seed, int64_t n, double* out);

}  // namespace pgs
