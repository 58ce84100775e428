// The step launchers and kernels for float32, radius 1 (space order 2).
#define CW_T float
#define CW_R 1
#include "steps.inc"
