// Host build of the DroneRace raceline projection of the device header
// (dr_drace.h fast_project: full scan, or the greedy search from the previous
// progress) as a small shared library, so that a CPU test can compare it with
// the oracle's restatement.  No device code is called.
#include <hip/hip_runtime.h>

#include "../../drone-race_amd/csrc/dr_drace.h"

extern "C" float hd_drace_project(const float* tp, float x, float y, float prev_s)
{
    const float pos[2] = {x, y};
    return dr::drace::fast_project(tp, pos, prev_s);
}
