// Host-side helpers shared by the MPPI and PR-MPPI engines: the error
// string behind dr_last_error(), device buffers, HIP status checks and the
// validation of the ABI's track descriptor.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "../../include/dr_api.h"

namespace dr {

// Records msg as the calling thread's dr_last_error() and returns code.
int host_fail(int code, const std::string& msg);

// DR_OK, or DR_EINVAL with a message naming track idx.
int check_track(const dr_track& t, int idx);
// Drone states must hold a unit attitude quaternion (|q|^2 within 1e-2 of 1)
// and finite values: the attitude recurrence renormalises to first order.
int check_drones(const dr_drone_state* s, int n);

// DR_OK, or DR_EINVAL naming the first invalid field of a StratRace2D environment.
int check_sr_env(const dr_sr_env& env);
int check_ho_env(const dr_ho_env& env);
int check_drace_env(const dr_drace_env& env);

template <class T>
struct DevBuf
{
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release()
    {
        if (p)
            (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    hipError_t alloc(size_t count)
    {
        release();
        n = count;
        if (count == 0)
            return hipSuccess;
        return hipMalloc((void**)&p, count * sizeof(T));
    }
};

} // namespace dr

#define HIPCHK(call)                                                                                                                           \
    do                                                                                                                                         \
    {                                                                                                                                          \
        hipError_t _e = (call);                                                                                                                \
        if (_e != hipSuccess)                                                                                                                  \
            return ::dr::host_fail(DR_EHIP, std::string(#call) + ": " + hipGetErrorString(_e));                                              \
    } while (0)
