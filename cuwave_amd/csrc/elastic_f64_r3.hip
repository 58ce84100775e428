// elastic step and gradient launchers for double at stencil radius 3
#define CW_T double
#define CW_R 3
#include "elastic_steps.inc"
