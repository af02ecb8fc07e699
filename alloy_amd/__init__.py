"""MI355X-native DFSPH liquid step (the `sph` solver plug-in of Alloy) on gfx950 HIP kernels."""

__version__ = "0.1.0"
