"""MI355X (gfx950) image-domain gridder with w-stacking.

The gridding API of the reference (Core/ImageDomainGridder, Core/FFT.cs)
backed by the HIP library libp9idg.so; see p9idg.idg.
"""
from .idg import (  # noqa: F401
    FFT,
    Batch,
    IDG,
    GriddingConstants,
    P9IDGError,
    Partition,
    Partitioner,
    Plan,
    SUBGRID_DTYPE,
    Visibilities,
)
from .cycle import CYCLE_STATS_DTYPE, PSF, DeconvolutionResult, FastSerialCD, MajorCycle, PaddedConvolver  # noqa: F401
from .fitsio import FitsIO, MeasurementData  # noqa: F401
from . import _lib  # noqa: F401
from ._lib import (  # noqa: F401
    DEGRID_PLAIN,
    DEGRID_WSTACK,
    GRID_PLAIN,
    GRID_WSTACK,
    GRID_WSUM,
    SRC_RESIDUAL,
    SRC_VIS,
    STAGE_DEGRIDDER,
    STAGE_GRIDDER,
    STAGE_TO_IMAGE,
    STAGE_TO_UVW,
    STAGE_ZERO,
)
