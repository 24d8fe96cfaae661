"""numrs_amd -- the numrs FFT path (four1, fourn, realft, convlv, correl, and
twofft, cosft1, cosft2, rlft3 on top of them) as
hand-written HIP kernels for MI355X (gfx950), behind the C ABI of
include/nrfft.h.

``api`` mirrors the reference's Rust functions and processors on host numpy
arrays; ``device`` exposes the device-pointer plans used with data resident
in HBM.
"""
from .api import (  # noqa: F401
    ConvlvError,
    ConvlvProcessor,
    CorrelError,
    FFTProcessor,
    FournError,
    NrfftError,
    RealFTProcessor,
    TwoFFTProcessor,
    autocorrel,
    autocorrel_fast,
    combine_real_imag,
    complex_to_real,
    convlv,
    convlv_batch,
    convlv_batch_responses,
    correl,
    correl_batch,
    correl_normalized,
    correl_normalized_fast,
    cosft1,
    cosft1_optimized,
    cosft2,
    cosft2_simd,
    extract_real_imag,
    FourFS,
    four1,
    four1_optimized,
    fourfs,
    fourn,
    magnitude_spectrum,
    power_spectrum,
    real_to_complex,
    realft,
    realft_optimized,
    realft_sequential,
    rlft3,
    rlft3_optimized,
    twofft,
    twofft_batch,
    twofft_optimized,
)
from .device import (  # noqa: F401
    C2CPlan,
    ConvlvPlan,
    CorrelPlan,
    Cosft1Plan,
    Cosft2Plan,
    RealFTPlan,
    Rlft3Plan,
    TwofftPlan,
    UnevenPlan,
    autotune_c2c,
    autotune_uneven,
    gen_uniform,
    tune_cache_load,
    tune_cache_save,
)

__version__ = "0.1.0"
