// Test harness: the proxy-geometry kernel source, compiled for the CPU emulator.
#include "proxy_kernels.hip"
