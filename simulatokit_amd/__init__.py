"""simulatokit_amd — SimulatoKit's multislice STEM hot path on one AMD Instinct MI355X (gfx950).

The `.hip` backend of SimulatoKit's `Multislice.run` / `runBatch` and `STEM.scan`: hand-written HIP kernels behind
a C ABI (include/simulato_hip.h, built into simulatokit_amd/_lib/libsimulato_hip.so), with this package as the
host-side mirror of the reference's Swift API. There is no CPU implementation behind it: when the library or the
device is missing, every entry point raises.
"""
from .core import (AtomicStructure, Backend, ElectronBeam, ExecutionOptions, Grid, HipAccuracy, Normalization,
                   REGISTERED, SimulatoError, Waves)
from .detector import (STEM, AnnularDetector, CenterOfMassDetector, PixelatedDetector, ScanMeasurement, ScanWaves,
                       SegmentedDetector, WavesDetector, last_timing, set_timing)
from .multislice import (ExitPlanes, FrozenPhonons, GaussianGenerator, LobatoParameterization, Multislice, Potential,
                         PreparedPotential, ProjectedPotential, Slicing, hip_is_available)
from .probe import Aberrations, Aperture, ContrastTransferFunction, GridScan, LineScan, Probe
from .tune import TuneResult, autotune

__all__ = ["AtomicStructure", "Backend", "ElectronBeam", "ExecutionOptions", "Grid", "HipAccuracy", "Normalization",
           "REGISTERED", "SimulatoError", "Waves", "STEM", "AnnularDetector", "CenterOfMassDetector",
           "PixelatedDetector", "ScanMeasurement", "ExitPlanes", "FrozenPhonons", "GaussianGenerator",
           "LobatoParameterization", "Multislice", "Potential", "PreparedPotential", "ProjectedPotential", "Slicing",
           "hip_is_available", "Aberrations", "Aperture", "GridScan", "Probe", "last_timing", "set_timing",
           "TuneResult", "autotune", "SegmentedDetector", "LineScan", "WavesDetector", "ScanWaves",
           "ContrastTransferFunction"]
