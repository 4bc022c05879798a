// fw_baseline_cpu.h -- drop-in for include/fw_baseline_cpu.h:6.
#pragma once

#include "constants.h"

// Sequential k-i-j Floyd-Warshall on the host; tileSize is ignored.
void fwBaselineCPU(WeightType* D, int n, int tileSize);
