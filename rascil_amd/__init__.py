"""rascil_amd: MI355X-native (gfx950) imaging core behind RASCIL's invert_ng / predict_ng.

Public surface (mirrors the reference's names):
  invert_ng, predict_ng                      (rascil/processing_components/imaging/ng.py)
  invert_blockvisibility, predict_blockvisibility   (imaging/imaging.py)
  wgridder.ms2dirty, wgridder.dirty2ms       (the ducc0.wgridder surface ng.py calls)
  dft_skycomponent_visibility, idft_visibility_skycomponent   (imaging/dft.py)
  weight_visibility                          (imaging/weighting.py)
  apply_gaintable, create_gaintable_from_blockvisibility, GainTable
                                             (calibration/operations.py)
  wgridder.Geometry                          plan once / execute many (hipGraph capturable)
  autograd.predict                           differentiable predict (gridder = adjoint)
  BlockVisibility, DeviceBlockVisibility (columns resident in GPU memory), Image,
  PolarisationFrame, Skycomponent, Direction
The compute runs in the hipng HIP library (rascil_amd/_hipng.so); nothing falls back to
the CPU.
"""

from .calibration import GainTable, apply_gaintable, create_gaintable_from_blockvisibility
from .data_models import (BlockVisibility, DeviceBlockVisibility, Direction, Image, PolarisationFrame,
                          Skycomponent, copy_visibility, create_image, get_parameter, image_is_canonical)
from .dft import dft_skycomponent_visibility, idft_visibility_skycomponent
from .imaging import invert_blockvisibility, predict_blockvisibility
from .imaging_base import normalise_sumwt, shift_vis_to_image
from .ng import invert_ng, predict_ng
from .polarisation import convert_pol_frame
from .weighting import weight_visibility

__version__ = "0.1.0"
