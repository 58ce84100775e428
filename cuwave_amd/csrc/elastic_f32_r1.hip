// elastic step and gradient launchers for float at stencil radius 1
#define CW_T float
#define CW_R 1
#include "elastic_steps.inc"
