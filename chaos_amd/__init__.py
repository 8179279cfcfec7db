"""chaos_amd: MI355X-native ensemble tangent stream (the `propagate` hot path of Chaoslab).

Hand-written HIP kernels for gfx950 behind a C ABI (include/chaoslab_hip.h); this package is the
Python host side of that ABI, mirroring the reference's `propagate` API and reducer types.
"""
from .systems import StandardMap, StandardMapK, StandardMapChain, WrappedStandardMap, HenonHeilesGauss, torus_state, hh_energy
from .systems import PairMap, pair_start, pair_frame, pair_distance, SeparationError
from .reducers import (QRSpectrum, CGS, MGS, RGS, HouseholderQR, AugmentedQR, GivensQR, Benettin, MEGNO, SALI, FLI,
                       OFLI, GALI, ResidualMonitor, RunMonitor, monitor_exponents, det_threshold, symp_threshold, sum_threshold,
                       eps_sum, eps_pair, check_residual)
from .tangent import propagate, propagate_ragged, propagate_mixed, PropagationResult, NumericalHealthError, initial_deviations
from .two_trajectory import propagate_shadow
from .reducers import RLI
from .birkhoff import WeightedPhaseAdvance, PlanarObservable, LiftedAngleObservable, circle_distance, frequency_index
from .reducers import StretchingNumbers, stretching_range, spectral_distance, spectrum_density, spectrum_cell
from .birkhoff import (WeightedBirkhoff, CosObservable, SinObservable, LiftedObservable, weighted_birkhoff,
                       birkhoff_index)

__all__ = ["propagate_mixed", "StandardMap", "StandardMapK", "StandardMapChain", "WrappedStandardMap", "HenonHeilesGauss", "hh_energy", "torus_state", "QRSpectrum", "CGS", "MGS", "RGS", "HouseholderQR",
           "AugmentedQR", "GivensQR", "Benettin", "MEGNO", "SALI", "FLI", "OFLI", "GALI", "ResidualMonitor", "RunMonitor", "monitor_exponents", "det_threshold",
           "symp_threshold", "sum_threshold", "eps_sum", "eps_pair", "check_residual", "propagate", "propagate_ragged",
           "PropagationResult", "NumericalHealthError", "initial_deviations", "propagate_shadow",
           "StretchingNumbers", "stretching_range", "spectral_distance", "spectrum_density", "spectrum_cell",
           "WeightedBirkhoff", "CosObservable", "SinObservable", "LiftedObservable", "weighted_birkhoff",
           "birkhoff_index"]
