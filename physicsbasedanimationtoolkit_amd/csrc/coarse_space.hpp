// Size of the adjoint's coarse space, shared by the host set-up and the kernels.
#pragma once

namespace pbat_vbd::device {

// Coarse degrees of freedom per vertex aggregate in the adjoint's two-level
// preconditioner: three translations and three infinitesimal rotations.
constexpr int kCoarseDofs = 6;

} // namespace pbat_vbd::device
