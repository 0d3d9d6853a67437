// aperture_host.h -- exact pixel / circle overlap and the pixel-integrated
// curve of growth of a PSF model, for model_rad = :auto with ImagePSF and
// GriddedPSFModel (whose curve of growth has no closed form).  Defined in
// aperture.cpp.
#pragma once

#include "../../include/crowdphot_hip.h"

namespace cp {

// Area of the unit pixel centred at (dx, dy) inside the circle of radius r
// centred at the origin (aperture_weight(..., ExactOverlap()),
// src/apertures.jl:93-166): 0 / 1 for pixels wholly outside / inside by the
// centre-distance tests, the exact area otherwise.
double pixel_overlap(double dx, double dy, double r);

// curve_of_growth(model, radii) (src/curve_of_growth.jl:143-176) of the
// unit-flux model centred at (0, 0): flux[k] = sum over the aperture's
// bounding pixels (apertures.jl:52-65) of overlap * model.  radii strictly
// positive and increasing.  Returns false if the model is not usable.
bool model_cog_host(const cp_psf_desc* psf, int n_radii, const double* radii, double* flux);

}  // namespace cp
