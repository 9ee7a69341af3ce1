// Test harness: the single-simulation kernel sources, compiled for the CPU emulator.
#include "step_kernels.hip"
