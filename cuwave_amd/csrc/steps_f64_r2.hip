// The step launchers and kernels for float64, radius 2 (space order 4).
#define CW_T double
#define CW_R 2
#include "steps.inc"
