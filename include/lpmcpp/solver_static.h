// solver_static.h -- LPM-CPP's include/solver_static.h for drivers built against the GPU adapter;
// its classes are those of lpm_gpu.hpp (see lpm.h in this directory).
#pragma once
#include "lpm.h"
