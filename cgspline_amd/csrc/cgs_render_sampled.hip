// The sampled (Monte Carlo) render kernel, kRender<true, true>, for every
// scene (a variant without the wide-hash and surface-cull paths measured
// slower on the 1M-patch frame).  It is compiled in a translation unit of its
// own from the same source as the plain render: in one code object with the
// plain render kernels it raised their register spills from 74 to 697 VGPRs.
// Only this kernel and its launcher (launchRenderSampled) are emitted here.
#define CGS_SAMPLED_TU 1
#include "cgs_trace.hip"
