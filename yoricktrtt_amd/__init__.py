"""MI355X (gfx950) footprint projectors of TRTT behind the reference C ABI.

Submodules:
  geometry    FP64 host geometry (XForm3D, frames, XOS/XSD, footprint table)
  projectors  per-view evaluators and device-resident batched projectors
  build       in-tree hipcc build of libtrtt_hip.so
"""
__all__ = ["geometry", "projectors", "build"]
__version__ = "0.1.0"
