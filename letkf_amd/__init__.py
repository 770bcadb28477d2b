"""letkf_amd -- MI355X-native LETKF local analysis.

The hot path of the reference (``letkf_core`` with ``mtx_eigen`` driven by the
per-model ``das_letkf`` loops, common/common_letkf.f90:50-243) runs as
hand-written HIP kernels for gfx950 in ``liblekf_hip.so`` behind the C ABI of
``include/lekf.h``.  This package is the Python host side over that ABI:

* :func:`letkf_core` -- per-point drop-in (same arguments as the Fortran);
* :func:`letkf_core_batch` / :func:`make_job` -- batched solve on device memory;
* :class:`GlobalLetkf`, :class:`L96Letkf`, :class:`RegionalLetkf` -- das_letkf
  handles that keep the ensemble resident on the GPU;
* :mod:`letkf_amd.synth` -- seeded synthetic generators of the benchmark configs;
* :mod:`letkf_amd.serve` -- the continuous-batching analysis server.
"""
from ._lib import LetkfError, lib  # noqa: F401
from .core import letkf_core, letkf_core_batch, make_job  # noqa: F401
from .das import GlobalLetkf, L96Letkf, RegionalLetkf, run_many  # noqa: F401

__all__ = ["letkf_core", "letkf_core_batch", "make_job", "GlobalLetkf", "L96Letkf",
           "RegionalLetkf", "run_many", "LetkfError", "lib"]
