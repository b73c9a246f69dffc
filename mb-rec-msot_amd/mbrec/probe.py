"""Probe registry (reconstruction_model_based/models/Probe.m).

Detector k sits at (r cos(theta_k), -r sin(theta_k)) with theta evenly spaced
from (180 - coverage)/2 to 180 - (180 - coverage)/2 degrees (Probe.m:110-116).
Two probes are registered:

* ``'example_system'``  the reference's made-up device (Probe.m:99-127):
  256 elements, 60 mm radius, 145 deg, 40 MHz, 2030 samples, delay 500.
* ``'arc256_r40_174'``  the benchmark device: 256 elements on a 40 mm arc over
  174 deg, 40 MHz, 2030 samples, delay 0, element 0.5 x 15 mm.

The EIR is a delta when ``include_eir`` is false (Probe.m:57-59), otherwise the
synthetic Gaussian-modulated EIR of the benchmark (fc 4 MHz, 100% bandwidth):
the reference's own EIR file is MATLAB v7.3 (HDF5) and is not used.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .irf import EIR_LENGTH, gauspuls_eir


@dataclass
class Detector:
    numOfTransducers: int
    radius: float
    angularCoverage: float
    xPositionsOfTransducers: np.ndarray
    zPositionsOfTransducers: np.ndarray
    elementWidth: float = 0.5e-3
    elementHeight: float = 15e-3
    numSubElements: int = 8
    EIR: np.ndarray = field(default=None)


@dataclass
class DAC:
    numRecordedSamplesPerTransducer: int
    delayBeforeRecording: int
    frequency: float


def arc_positions(n: int, radius: float, coverage_deg: float):
    offset = (180.0 - coverage_deg) / 2.0
    ang = (offset + np.arange(n) * (coverage_deg / (n - 1))) * np.pi / 180.0
    return radius * np.cos(ang), -radius * np.sin(ang)


def membrane_model_horizontal_example(x):
    """Probe.m:148-151: the example membrane, a horizontal line at z = -10 mm."""
    return -0.01 * np.ones_like(np.asarray(x, dtype=np.float64))


_REGISTRY = {
    # coupling medium heavy water at 1397 m/s and the horizontal example membrane (Probe.m:101-106)
    "example_system": dict(n=256, radius=0.060, coverage=145.0, samples=2030, delay=500, fs=4e7,
                           width=0.2e-3, height=20e-3, sub=8, coupling_sos=1397.0,
                           membrane=membrane_model_horizontal_example),
    "arc256_r40_174": dict(n=256, radius=0.040, coverage=174.0, samples=2030, delay=0, fs=4e7,
                           width=0.5e-3, height=15e-3, sub=8, coupling_sos=None, membrane=None),
}


@dataclass
class Coupling:
    speedOfSound: float | None = None


class Probe:
    eirLength = EIR_LENGTH

    def set_custom_membrane_model(self, custom_membrane_model):
        """Probe.m:88-94: a callable z = m(x) (vectorised over numpy arrays)."""
        if not callable(custom_membrane_model):
            raise ValueError("Custom membrane model must be function handle.")
        self.membrane = custom_membrane_model

    def __init__(self, probe_id: str, include_eir: bool = True, include_indiv_eir: bool = False,
                 num_detectors: int | None = None):
        if probe_id not in _REGISTRY:
            raise ValueError("Unknown device id.")
        cfg = dict(_REGISTRY[probe_id])
        n = num_detectors or cfg["n"]
        x, z = arc_positions(n, cfg["radius"], cfg["coverage"])
        self.probeId = probe_id
        self.includeEir = include_eir
        self.includeIndivEir = False if include_indiv_eir else False  # no individual EIRs available
        self.detector = Detector(n, cfg["radius"], cfg["coverage"], x, z, cfg["width"], cfg["height"],
                                 cfg["sub"])
        self.DAC = DAC(cfg["samples"], cfg["delay"], cfg["fs"])
        self.coupling = Coupling(cfg["coupling_sos"])
        self.membrane = cfg["membrane"]
        if include_eir:
            self.detector.EIR = gauspuls_eir(cfg["fs"])
        else:
            eir = np.zeros(EIR_LENGTH)
            eir[EIR_LENGTH // 2] = 1.0
            self.detector.EIR = eir
