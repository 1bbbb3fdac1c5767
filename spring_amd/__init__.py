"""MI355X-native three/four-phonon scattering-rate path (SPRING's Kappa / LineWidth hot path).

Host side (numpy, FP64): crystal + force-constant model generation, the Gamma-centred
q-mesh and its irreducible wedge, harmonic phonons, and the driver that mirrors
src/Kappa/Src/main.f90.  Device side: libspring_gpu.so (hand-written gfx950 HIP kernels
behind a C ABI, declared in include/spring_gpu.h), loaded by spring_amd.lib.
"""
__all__ = ["constants", "model", "mesh", "phonons", "lib"]
