"""Numerical-method configuration, attribute-compatible with the reference's
``Numerics`` (Numerics.py:1-16).

Extensions (not in the reference, all optional):
  * ``Hopkins_GridHalfWidth``: TCC pupil-frequency grid half-width N (grid
    (2N+1)^2); None keeps the reference rule ceil(2 p) (CalculateTCCMatrix.py:19-20).
  * ``TCC_Precision``: 'split_bf16' (3-term split bf16 on MFMA, fp32
    accumulation), 'fp32' or 'fp64' for the TCC contraction.
  * ``TCC_Decomposition``: how DecomposeTCC_SOCS takes the eigenpairs of a
    TCC built by CalculateTCCMatrix: 'projected' (the computed TCC on the
    range of its shifted-pupil factor, float64), 'factor' (from the factor's
    Gram matrix alone) or 'dense' (the M x M TCC itself).
"""


class Numerics:
    ImageCalculationMode: str = 'scalar'
    ImageCalculationMethod: str = 'hopkins'
    Hopkins_SettingType: str = 'order'
    Hopkins_Order: int = 100
    Hopkins_Threshold: float = 0.95
    ScatterGrid_X: int = 100
    ScatterGrid_Y: int = 100
    ScatterOrder_X: int = 15
    ScatterOrder_Y: int = 15
    Hopkins_GridHalfWidth = None
    TCC_Precision: str = 'split_bf16'
    TCC_Decomposition: str = 'projected'

    def __init__(self):
        pass
