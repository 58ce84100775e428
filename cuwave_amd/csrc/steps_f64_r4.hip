// The step launchers and kernels for float64, radius 4 (space order 8).
#define CW_T double
#define CW_R 4
#include "steps.inc"
