// The step launchers and kernels for float64, radius 3 (space order 6).
#define CW_T double
#define CW_R 3
#include "steps.inc"
