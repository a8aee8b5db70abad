"""miscellaneous_amd -- MI355X-native (gfx950) 2D NFFT.

A from-scratch HIP implementation of the 2D NFFT path of
OrangeOwlSolutions/Miscellaneous (``NFFT/NFFT2_2D.cu``, ``NFFT/NDFT2_2D.cu``
and the MATLAB ``BP_FBP_FFBP/NFFT1_2D.m``), behind the reference's call
signatures (``include/nfft2d.h``), plus a differentiable type-2 transform
the type-2 NDFT as a split-precision MFMA contraction (``NdftPlan``) and
batches of type-2 transforms of different sizes in one launch per stage
(``Batch``), and the reference's band-limited polar interpolation on top of
the type-2 (``knabNFFT2D``), batched 1-D NFFTs (``Plan1D``, ``NFFT1_1D``,
``NFFT2_1D``), and the FP64 type-3 NUFFT (``NUFFT_2D``, the counterpart of
``NFFT/NFFT3/C/NFFT3_2D_C.cpp``).

Everything here runs through ``libnfft2d.so`` (built in-tree by
``miscellaneous_amd.build``); there is no CPU fallback.
"""
from ._lib import NFFT2DError, load_library, LIB_PATH
from .nfft import (Plan, NFFT2_2D_GPU, NFFT1_2D_GPU, NDFT2_2D_GPU, ndft_grid_info,
                   Calculate_cuFFT_plan_C2C_NFFT2_2D, Calculate_cuFFT_plan_Z2Z_NFFT2_2D, Destroy_cuFFT_plan_NFFT2_2D)
from .autograd import NFFT2DType2, nfft2d_type2
from .ndft import NdftPlan
from .batch import Batch
from .knab import knabNFFT2D, rectNFFT2D, knab_window
from .nfft1d import Plan1D, NFFT1_1D, NFFT2_1D
from .pso import FarFieldGeneration
from .nufft3 import NUFFT_2D, NUFFT_2D_host, nufft3_params

__all__ = ["NFFT2DError", "load_library", "LIB_PATH", "Plan", "NFFT2_2D_GPU", "NFFT1_2D_GPU", "NDFT2_2D_GPU",
           "ndft_grid_info", "Calculate_cuFFT_plan_C2C_NFFT2_2D", "Calculate_cuFFT_plan_Z2Z_NFFT2_2D", "Destroy_cuFFT_plan_NFFT2_2D",
           "NFFT2DType2", "nfft2d_type2", "NdftPlan", "Batch", "knabNFFT2D", "rectNFFT2D", "knab_window",
           "Plan1D", "NFFT1_1D", "NFFT2_1D", "FarFieldGeneration", "NUFFT_2D", "NUFFT_2D_host", "nufft3_params"]
