// The step launchers and kernels for float32, radius 3 (space order 6).
#define CW_T float
#define CW_R 3
#include "steps.inc"
