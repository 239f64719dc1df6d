// Sweep kernels and launchers for 512-point waves.
#include "sweeps_tpl.hpp"

MSE_INSTANTIATE_SWEEPS(512)
