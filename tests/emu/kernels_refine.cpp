// Test harness: the mixed-precision solver kernel sources, compiled for the CPU emulator.
#include "refine_kernels.hip"
