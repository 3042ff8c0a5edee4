"""MI355X-native nonlinear magnetostatic solve path (P1 Newton + batched PCG).

Python access to the C-ABI library ``lib/libfemm_gpu_amd.so`` and the model /
mesh utilities around it.
"""

from .model import (AXISYMMETRIC, MU0, PLANAR, Material, Model, accepted_options,  # noqa: F401
                    load_artifact, plain_options, production_options, status_name)
