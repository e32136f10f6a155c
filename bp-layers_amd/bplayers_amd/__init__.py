"""MI355X (gfx950) implementation of the BP-Layers hot path (VLOGroup/bp-layers).

SAD matching-cost volumes (stereo and optical flow) feeding differentiable min-sum loopy
belief propagation on the 4-connected grid, as hand-written HIP kernels in libbpl_amd.so
behind the reference's PyTorch op and autograd signatures:

  reference extension / module                    here
  pytorch_cuda_stereo_sad_op                      bplayers_amd.pytorch_cuda_stereo_sad_op
  pytorch_cuda_flow_mp_sad_op                     bplayers_amd.pytorch_cuda_flow_mp_sad_op
  pytorch_cuda_lbp_op (ops/lbp_stereo)            bplayers_amd.pytorch_cuda_lbp_op_stereo
  pytorch_cuda_lbp_op (ops/lbp_semantic_pw)       bplayers_amd.pytorch_cuda_lbp_op_semantic_pw
  pytorch_cuda_lbp_op (ops/lbp_semantic_pw_pixel) bplayers_amd.pytorch_cuda_lbp_op_semantic_pw_pixel
  ops.* autograd Functions / nn.Modules           bplayers_amd.ops.*
"""
import importlib
import sys

from . import _native  # noqa: F401  (imports torch first, then binds the library lazily)

__version__ = "0.1.0"

_LBP_VARIANTS = {
    "stereo": "pytorch_cuda_lbp_op_stereo",
    "semantic_pw": "pytorch_cuda_lbp_op_semantic_pw",
    "semantic_pw_pixel": "pytorch_cuda_lbp_op_semantic_pw_pixel",
}


def install_reference_aliases(lbp_variant="stereo"):
    """Register this package's op modules under the reference's extension module names
    (pytorch_cuda_stereo_sad_op, pytorch_cuda_flow_mp_sad_op, pytorch_cuda_lbp_op), so that
    the reference's own Python files import them unchanged.  The reference builds three
    different modules under the one name pytorch_cuda_lbp_op; choose which one it means."""
    mods = {
        "pytorch_cuda_stereo_sad_op": "bplayers_amd.pytorch_cuda_stereo_sad_op",
        "pytorch_cuda_flow_mp_sad_op": "bplayers_amd.pytorch_cuda_flow_mp_sad_op",
        "pytorch_cuda_lbp_op": "bplayers_amd." + _LBP_VARIANTS[lbp_variant],
    }
    for alias, real in mods.items():
        sys.modules[alias] = importlib.import_module(real)
    return mods
