// Test harness: the step-norms kernel sources, compiled for the CPU emulator.
#include "norms_kernels.hip"
