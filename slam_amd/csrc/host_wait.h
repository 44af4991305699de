// host_wait.h -- the host's wait for results it reads back on every frame
// (the tracked pose and statistics, the surfel count).
//
// A blocking stream synchronisation sleeps until the runtime's completion
// interrupt: in the frame loop each such readback left the GPU idle for
// 30-40 us before the next launch.  Polling the stream returns within about
// a microsecond of the last command, at the cost of one busy host thread for
// the wait's duration.  The destination of the readback should be pinned
// host memory, so that the copy itself is asynchronous.
#pragma once

#include <hip/hip_runtime.h>

namespace slam {

inline hipError_t poll_stream(hipStream_t s)
{
    for (;;) {
        const hipError_t e = hipStreamQuery(s);
        if (e != hipErrorNotReady) return e;
        __builtin_ia32_pause();
    }
}

}  // namespace slam
