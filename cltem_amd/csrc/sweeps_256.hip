// Sweep kernels and launchers for 256-point waves.
#include "sweeps_tpl.hpp"

MSE_INSTANTIATE_SWEEPS(256)
