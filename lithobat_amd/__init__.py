"""MI355X-native partially coherent imaging: Abbe, SOCS and their adjoint.

Drop-in entry points with the reference's names and semantics:

* ``compute_aerial_image``, ``compute_aerial_planes``, ``normalisation_scale``,
  ``source_points`` (expose/aerial_image.py)
* ``SOCSKernels`` (expose/hopkins.py)
* ``socs_fields``, ``socs_gradient`` and the differentiable
  ``AbbeIntensity`` (opc/ilt.py)
* ``exposure_volume``, ``effective_defocus`` (develop/resist3d.py)

plus the batched device forms ``compute_aerial_batch`` and ``ImagingPlan``.
The heavy lifting is the C-ABI library ``liblithobat_amd.so`` (HIP, gfx950);
there is no CPU fallback.
"""
from .config import GridConfig, OpticsConfig, ResistConfig  # noqa: F401
from .masks import lines_and_spaces, manhattan  # noqa: F401
from .source import SourcePoints, build_source, source_points  # noqa: F401

__version__ = "0.1.0"


def __getattr__(name):
    # device-facing modules import torch and load the library lazily
    if name in ("compute_aerial_image", "compute_aerial_planes", "compute_aerial_batch",
                "normalisation_scale"):
        from . import expose
        return getattr(expose, name)
    if name in ("exposure_volume", "effective_defocus", "focus_reference_depth"):
        from . import resist3d
        return getattr(resist3d, name)
    if name == "SOCSKernels":
        from .hopkins import SOCSKernels
        return SOCSKernels
    if name in ("AbbeIntensity", "abbe_intensity", "socs_fields", "socs_gradient",
                "relaxed_resist_loss", "mask_of"):
        from . import ilt
        return getattr(ilt, name)
    if name in ("ImagingPlan", "get_plan", "clear_plan_cache"):
        from . import plan
        return getattr(plan, name)
    raise AttributeError(name)
