// staging.hpp — small host-to-device tables (plan arrays, task lists) for
// fused launches, passed through pinned host memory so the copy is
// asynchronous and stream-ordered with the launch that reads it.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace xg {

// A pinned host buffer and a device buffer of the same size.  Use:
//   char* h = s.host(bytes, stream);   // waits until the previous upload has left
//   ... fill h ...
//   char* d = s.upload(bytes, stream); // async copy, then launch kernels reading d
// The device buffer may be read by kernels enqueued on `stream` after the
// upload; it is reused (overwritten) by the next upload on the same stream,
// which is ordered after those kernels.
class Staging {
  public:
    Staging() = default;
    Staging(const Staging&) = delete;
    Staging& operator=(const Staging&) = delete;
    ~Staging();

    char* host(size_t bytes, hipStream_t stream);
    char* upload(size_t bytes, hipStream_t stream);

  private:
    char* h_ = nullptr;
    char* d_ = nullptr;
    size_t cap_ = 0;
    hipEvent_t copied_ = nullptr;
    bool pending_ = false;
};

// The staging buffer of (current device, stream).
Staging& staging_for(hipStream_t stream);

}  // namespace xg
