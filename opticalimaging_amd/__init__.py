"""MI355X-native aerial-image engine with the imaging API of
gary21978/OpticalImaging.

Drop-in modules (same names and signatures as the reference):
  CalculateAerialImage      ImageData, CalculateAbbeImage, CalculateHopkinsImage
  CalculateTCCMatrix        CalculateTCCMatrix, GetTCCMatrix
  DecomposeTCC_SOCS         DecomposeTCC_SOCS
  CalculateAerialImage_SOCS CalculateAerialImage_SOCS
  CalculateCharacteristicMatrix
  Source, Projection, Scatter, Numerics (parameter objects)
All numerical work runs in hand-written HIP kernels for gfx950
(libopticalimaging_amd.so, C ABI in include/opticalimaging_amd.h).
"""
from .Numerics import Numerics
from .Source import Source, SourceData, PolarizationParameters
from .Projection import Projection
from .Scatter import Scatter
from .CalculateCharacteristicMatrix import CalculateCharacteristicMatrix
from .CalculateTCCMatrix import CalculateTCCMatrix, GetTCCMatrix
from .DecomposeTCC_SOCS import DecomposeTCC_SOCS
from .CalculateAerialImage_SOCS import CalculateAerialImage_SOCS, CalculateSOCS
from .CalculateAerialImage import ImageData, CalculateAbbeImage, CalculateHopkinsImage, CalculateOptimized
from .engine import KernelSet, socs_image

__all__ = ["Numerics", "Source", "SourceData", "PolarizationParameters", "Projection", "Scatter",
           "CalculateCharacteristicMatrix", "CalculateTCCMatrix", "GetTCCMatrix",
           "DecomposeTCC_SOCS", "CalculateAerialImage_SOCS", "ImageData", "CalculateAbbeImage",
           "CalculateHopkinsImage", "CalculateOptimized", "CalculateSOCS", "KernelSet", "socs_image"]
