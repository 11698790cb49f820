// FiniteSizeCorrection's interpolation and integration of the transition
// structure factor (algorithms/FiniteSizeCorrection.cxx:136-311), host code.
#pragma once

#include <cstdint>

namespace cc4s_amd {

struct FiniteSizeEnergy {
  double corrected = 0, uncorrected = 0;
};

// grid [NG][3] Cartesian momenta, B [3][3] reciprocal lattice vectors (rows
// Gi, Gj, Gk), S [NG]; N the fine-grid resolution (the step's
// interpolationGridSize); units of the vertex, the grid and the potential.
FiniteSizeEnergy finite_size_correction(const double *grid, int64_t NG, const double *B,
                                        const double *S, int64_t N, double vertex_unit,
                                        double grid_unit, double potential_unit);

}  // namespace cc4s_amd
