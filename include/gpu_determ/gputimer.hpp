// Drop-in for the reference's gputimer.hpp (gputimer.hpp:10-46): wall time
// between two device-synchronised points, in milliseconds.
#ifndef GPU_DETERM_GPUTIMER_HPP
#define GPU_DETERM_GPUTIMER_HPP

#include <chrono>

#include "../gdt.h"

class GpuTimer
{
public:
    void start()
    {
        gdt_device_synchronize();
        t0_ = std::chrono::steady_clock::now();
    }
    void stop()
    {
        gdt_device_synchronize();
        t1_ = std::chrono::steady_clock::now();
    }
    float elapsed() { return std::chrono::duration<float, std::milli>(t1_ - t0_).count(); }

private:
    std::chrono::steady_clock::time_point t0_, t1_;
};

#endif
