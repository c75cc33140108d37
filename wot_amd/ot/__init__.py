"""OT solvers and the OTModel driver (the wot.ot namespace of the reference)."""
from .adjoint import ot_objective  # noqa: F401
from .device import DeviceCost  # noqa: F401
from .initializer import (parse_configuration, parse_parameter_file,  # noqa: F401
                          parse_per_timepair_configuration, parse_per_timepoint_configuration)
from .optimal_transport import (compute_transport_matrix, dual, fdiv, fdivstar,  # noqa: F401
                                optimal_transport_duality_gap, primal, transport_stablev1_learnx,
                                transport_stablev2)
from .ot_model import OTModel  # noqa: F401
from .util import compute_pca  # noqa: F401
