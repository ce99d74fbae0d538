"""simul_amd -- MI355X-native NOSB peridynamics solver step (drop-in for the
reference backend/fea peridynamics path).  Compute runs in libpdx.so (HIP,
gfx950) behind a C ABI; this package is the host-side mirror of the
reference's Python API."""
__version__ = "0.1.0"
