// elastic step and gradient launchers for double at stencil radius 2
#define CW_T double
#define CW_R 2
#include "elastic_steps.inc"
