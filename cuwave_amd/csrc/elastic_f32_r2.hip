// elastic step and gradient launchers for float at stencil radius 2
#define CW_T float
#define CW_R 2
#include "elastic_steps.inc"
