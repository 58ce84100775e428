// elastic step and gradient launchers for double at stencil radius 1
#define CW_T double
#define CW_R 1
#include "elastic_steps.inc"
