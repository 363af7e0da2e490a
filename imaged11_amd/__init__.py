"""imaged11_amd — ImageD11's grain-scoring path as HIP kernels for MI355X.

Modules:
  cImageD11  drop-in replacements for the cImageD11 functions on the path
             (compute_xlylzl, compute_gv, score, score_and_refine,
             score_and_assign, refine_assigned, verify_rounding)
  batched    many-grain launches on device tensors (score_many,
             assign_many, sf2gv, refine_lists, ...)
  geometry   ImageD11 parameter packing (Ctransform equivalent)
  simulate   seeded synthetic workloads
"""
__version__ = "0.1.0"
