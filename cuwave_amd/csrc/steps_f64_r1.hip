// The step launchers and kernels for float64, radius 1 (space order 2).
#define CW_T double
#define CW_R 1
#include "steps.inc"
