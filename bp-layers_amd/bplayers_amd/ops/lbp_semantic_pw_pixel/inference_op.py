"""Inference over beliefs; the same module as the stereo one (reference ops/lbp_semantic_pw_pixel/inference_op.py)."""
from ..lbp_stereo.inference_op import Inference  # noqa: F401
