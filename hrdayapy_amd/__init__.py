"""MI355X-native hot path of HrdayaPy: TTP06 reaction, anisotropic 19-point
monodomain diffusion, Purkinje cable with PMJ coupling and the torso Poisson
forward solve, as hand-written gfx950 HIP kernels behind a C ABI
(include/hrdaya_hip.h), with the reference's hrdayapy.simulation /
hrdayapy.anisotropic_simulation / hrdayapy.ecg call surface on top."""
__version__ = "0.1.0"
