// Sweep kernels and launchers for 1024-point waves.
#include "sweeps_tpl.hpp"

MSE_INSTANTIATE_SWEEPS(1024)
