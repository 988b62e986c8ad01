"""Context dispatch (rascil/processing_components/imaging/imaging.py:26-93).

"ng" (default) and "hipng" run invert_ng / predict_ng on the GPU gridder; "2d" is the
same with do_wstacking=False. Unknown contexts raise ValueError as in the reference.
"awprojection" and "wg" are other backends of the reference and are not provided here.
"""

from .ng import invert_ng, predict_ng

__all__ = ["predict_blockvisibility", "invert_blockvisibility"]


def predict_blockvisibility(vis, model, context="ng", gcfcf=None, **kwargs):
    """Predict blockvisibility from an image."""
    if context == "2d":
        return predict_ng(vis, model, do_wstacking=False, **kwargs)
    if context in ("ng", "hipng"):
        return predict_ng(vis, model, **kwargs)
    raise ValueError(f"Unknown imaging context {context}")


def invert_blockvisibility(vis, im, dopsf=False, normalise=True, context="ng", gcfcf=None,
                           **kwargs):
    """Invert blockvisibility to make an (image, sum weights) tuple."""
    if context == "2d":
        return invert_ng(vis, im, dopsf=dopsf, normalise=normalise, do_wstacking=False, **kwargs)
    if context in ("ng", "hipng"):
        return invert_ng(vis, im, dopsf=dopsf, normalise=normalise, **kwargs)
    raise ValueError(f"Unknown imaging context {context}")
