// The step launchers and kernels for float32, radius 4 (space order 8).
#define CW_T float
#define CW_R 4
#include "steps.inc"
