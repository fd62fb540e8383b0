"""MI355X-native Mi3 inference engine: drop-in host package for the Mi3
inverse-Potts inference round (mcmcGPU / node_manager / NewtonSteps / Mi3
command line), computing on hand-written gfx950 HIP kernels in libmi3hip.so."""
__version__ = '0.1.0'
