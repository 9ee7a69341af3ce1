// Test harness: the batched kernel sources, compiled for the CPU emulator.
#include "batch_kernels.hip"
