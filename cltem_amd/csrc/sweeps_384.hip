// Sweep kernels and launchers for 384-point waves.
#include "sweeps_tpl.hpp"

MSE_INSTANTIATE_SWEEPS(384)
