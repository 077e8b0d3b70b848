// Device-side data of the StratRace2D plain-MPPI engine.  The batch layout
// is the quadrotor engine's (dr_kernels.h: InstDev, InstDyn, RTile, WTile)
// with 2-D controls: nominal float[sum_i T_i*2] ([t][d]), noise float4 rows
// [k][t] holding (eps0, eps1, 0, 0) so that the shared weighted-average
// kernel reads them unchanged; one world state (dr_sr_state) per instance.
#pragma once

#include "dr_kernels.h"

namespace dr {

struct SrInstDev
{
    int theta;     // opponent model followed by the rollouts and the world
    float sigma;   // control noise standard deviation (both dimensions)
    float fill[2]; // value of the freed last step after the shift
    int adapt;
    int _pad[3];
};

} // namespace dr
