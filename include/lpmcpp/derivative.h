// derivative.h -- LPM-CPP's include/derivative.h for drivers built against the GPU adapter;
// its classes are those of lpm_gpu.hpp (see lpm.h in this directory).
#pragma once
#include "lpm.h"
