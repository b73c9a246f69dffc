"""mbrec -- MI355X-native model-based MSOT operators and NNLS solvers.

Host-side mirror of the reference toolbox's operator interface
(reconstruction_model_based/models/*.m, reconstruction/*.m) over the HIP
library libmbrec.so (include/mbrec.h).
"""
from ._lib import MbrecError, lib, LIB_PATH  # noqa: F401
from .discretization import Discretization  # noqa: F401
from .probe import Probe  # noqa: F401

__all__ = ["MbrecError", "lib", "LIB_PATH", "Discretization", "Probe", "MSOTmodel",
           "define_model_for_reconstruction", "rec_nn_with_L2_reg", "rec_nn_with_L1_eye_reg", "rec_nn_with_TV_reg", "forwardMex", "transposeMex",
           "forwardDelayAndSumMex", "transposeDelayAndSumMex"]


def __getattr__(name):
    if name in ("MSOTmodel", "define_model_for_reconstruction", "DevicePlan", "travel_time_indices"):
        from . import model
        return getattr(model, name)
    if name in ("rec_nn_with_L2_reg", "rec_nn_with_L1_eye_reg", "rec_nn_with_TV_reg"):
        from . import solvers
        return getattr(solvers, name)
    if name in ("forwardMex", "transposeMex", "forwardDelayAndSumMex", "transposeDelayAndSumMex", "MexError"):
        from . import mex
        return getattr(mex, name)
    raise AttributeError(name)
