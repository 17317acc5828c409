"""MI355X-native BioFVM reaction-diffusion backend.

physicore_amd.bfvm   ctypes binding of libbfvm_hip.so (the C ABI in include/bfvm_hip.h)
physicore_amd.build  in-tree build recipe of the native libraries
physicore_amd/host   C++20 hip_solver behind the reference's biofvm::solver interface
"""
__version__ = "0.1.0"
