// elastic step and gradient launchers for double at stencil radius 4
#define CW_T double
#define CW_R 4
#include "elastic_steps.inc"
