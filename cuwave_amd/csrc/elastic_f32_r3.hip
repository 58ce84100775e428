// elastic step and gradient launchers for float at stencil radius 3
#define CW_T float
#define CW_R 3
#include "elastic_steps.inc"
