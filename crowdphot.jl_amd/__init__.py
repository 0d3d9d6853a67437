"""crowdphot_amd -- MI355X (gfx950) build of CrowdPhot.jl's PSF-fitting hot path.

Host-side mirror of the reference Julia API over libcrowdphot_hip.so:
  fit_all_stars_simultaneous  (src/photometry/psf_photometry_simultaneous.jl:953)
  fit_all_stars                (src/photometry/psf_photometry_single.jl:267, sequential multi-pass)
  matched_filter, findlocalmaxima, correlate (src/detection.jl, src/correlation.jl)
  roman_crds_gridded_epsf      (src/psf/roman_crds.jl, ext/CrowdPhotASDFExt.jl)
  fit_star / fit_star_batch    (src/psf/psf_fitting.jl:229, lm_irls src/levenberg_marquardt.jl:431)
  CircularGaussianPSF, CircularMoffatPSF, ImagePSF, GriddedPSFModel
"""
from . import _abi  # noqa: F401
from .photometry import (ArgumentError, Context, LMResult, MultiPassPhotResult,  # noqa: F401
                         build_stamps, default_context, evaluate_fg, fit_all_stars,
                         fit_all_stars_simultaneous, fit_star, fit_star_batch)
from .psf import (CircularGaussianPSF, CircularMoffatPSF, GriddedPSFModel,  # noqa: F401
                  ImagePSF, moffat_epsf, pixel_response_kernel)
from .detection import (CircularGaussianPRF, GaussianPRF, MatchedFilterResult, correlate,  # noqa: F401
                        findlocalmaxima, matched_filter, render)
from . import autotune, roman_crds, serving, simulation  # noqa: F401
from .roman_crds import read_asdf, roman_crds_gridded_epsf  # noqa: F401
