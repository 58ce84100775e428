// elastic step and gradient launchers for float at stencil radius 4
#define CW_T float
#define CW_R 4
#include "elastic_steps.inc"
