// Sweep kernels and launchers for 768-point waves.
#include "sweeps_tpl.hpp"

MSE_INSTANTIATE_SWEEPS(768)
