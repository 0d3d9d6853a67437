// diag_host.h -- host set-up of the post-fit per-star diagnostics
// (src/photometry/psf_photometry_diagnostics.jl): the spread_model reference
// disk.  Defined in diagnostics.cpp.
#pragma once

#include "../../include/crowdphot_hip.h"

namespace cp {

// _spread_fwhm (psf_photometry_diagnostics.jl:113-123): Gaussian-equivalent
// FWHM sqrt(ea 2 ln2 / pi), ea = (sum g)^2 / sum g^2 of a unit-flux render of
// the model at (y0, x0) over its integer extent.  NaN when undefined.
double spread_fwhm_host(const cp_psf_desc* psf, double y0, double x0);

// _exp_disk_kernel_bandlimited (psf_photometry_diagnostics.jl:78-101) with
// half-width `half`: K[(i) + (j) (2 half + 1)], i, j = 0..2 half, sum 1.
void spread_kernel_host(double fwhm, int half, double* K);

}  // namespace cp
