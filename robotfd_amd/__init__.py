"""MI355X-native continuous Frechet distance (RobotFD's compute_distance_parallel path).

The compute path is librfd.so: gfx950 HIP kernels behind the C ABI in
include/rfd.h.  This package is its Python mirror of the reference interface.
"""
from .frechet import (FLT_EPSILON, Context, Curve, Intersection, Point,  # noqa: F401
                      compute_distance_parallel, distance, distance_cuda)
