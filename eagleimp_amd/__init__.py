"""MI355X (gfx950) engine for EagleImp's phasing pre-stage.

For each target sample: the K best conditioning reference haplotypes by
Hamming distance at the target's homozygous sites, then the target's
condensed PBWT (forward and backward compressed columns with cumulative
zero counts, plus count0 per row), bit-exact to the reference CPU path.

  include/eimp.h          the C ABI (drop-in beside the reference's GPUEngine)
  eagleimp_amd/csrc       HIP kernels + engine, generator
  eagleimp_amd/engine.py  ctypes binding used by tests / bench / serving
"""
from .engine import Engine, EimpError, device_count, pbwt_variants  # noqa: F401

__all__ = ["Engine", "EimpError", "device_count", "pbwt_variants"]
