"""dgm_amd: MI355X-native explicit DG residual and Runge-Kutta engine.

A drop-in for the hot path of DGM's compressible Euler / Navier-Stokes domains
(Domain::time_derivative + TimeInt::advance), built on hand-written gfx950 HIP
kernels behind the C ABI in include/dgx.h.

    dgm_amd.dgx        ctypes binding of libdgx.so (GPUDomain)
    dgm_amd.problem    ReoProblem(args, table, comm).solve() mirror of pydgm
    dgm_amd.formats    DGM input / restart file readers and writers
    dgm_amd.mesh       periodic structured quad / hex mesh generator
    dgm_amd.ic         initial conditions
"""
__version__ = "0.1.0"
