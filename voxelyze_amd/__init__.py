"""voxelyze_amd -- MI355X-native explicit dynamics step of Voxelyze.

The product is the HIP library libvxg.so (C ABI in include/vxg.h, C++ facade
in include/Voxelyze.h).  This package builds it, loads it and mirrors the
reference's object API in Python (voxelyze_amd.api).
"""
from .api import (CVoxelyze, CVX_Material, CVX_Voxel, CVX_External, CVX_Link, dof,  # noqa: F401
                  X_TRANSLATE, Y_TRANSLATE, Z_TRANSLATE, X_ROTATE, Y_ROTATE, Z_ROTATE,
                  X_POS, X_NEG, Y_POS, Y_NEG, Z_POS, Z_NEG, FP64_PLAIN, MIXED)
