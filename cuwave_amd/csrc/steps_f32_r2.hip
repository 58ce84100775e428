// The step launchers and kernels for float32, radius 2 (space order 4).
#define CW_T float
#define CW_R 2
#include "steps.inc"
