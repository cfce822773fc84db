// Learned coarse-level initial guess: internal interface between the U-Net
// translation unit and the solver.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

struct srmg_unet;

namespace srmg_guess {

// grid extents the network was built for (unused dimensions are 1)
void extents(const srmg_unet *net, int *D, int n[3]);
// true when the network was declared for this level-1 operator
bool matches(const srmg_unet *net, int geometry, int disc, double k2, const int8_t bc[27]);
// u[g] = network prediction of the solution of L_1 u = f[g] for ngrid stacked
// level-1 grids (x fastest), queued on `stream`
void infer(srmg_unet *net, const double *f, double *u, int ngrid, hipStream_t stream);

}  // namespace srmg_guess
