// Internal interface of the fake HIP device (tests/fakehip) between the
// runtime (fake_runtime.cpp), the fake kernels (fake_kernels.cpp) and the
// stress driver (stress.cpp).  Test infrastructure only.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <string>
#include <vector>

namespace fakehip {

// Runs `fn` on stream `s` in order (or records it when `s` is capturing).
void enqueue(hipStream_t s, std::function<void()> fn);

// Whether [p, p + bytes) lies inside one live device allocation (hipMalloc);
// *why explains a failure.
bool live(const void *p, size_t bytes, std::string *why);
// Bytes from p to the end of the live device allocation holding p (0: none).
uint64_t bytes_from(const void *p);
// Sequence number of the latest host-to-device copy or memset that wrote any
// byte of [p, p + bytes) (0: never written by a copy).  Sequence numbers grow
// in the order operations execute on the fake device.
uint64_t last_write(const void *p, size_t bytes);

// Records a violation (kept for the driver's report; the run goes on).
void error(const std::string &msg);
std::vector<std::string> errors();

// Called with (base, size) of every device allocation as it is freed, so the
// fake kernels can drop state they keep per buffer.
void on_free(std::function<void(const void *, size_t)> fn);

// Operations sleep a random 0..max_us microseconds before they run, so that
// work on different streams interleaves as on a real device.
void set_max_delay_us(int max_us);

struct Stats {
  uint64_t ops = 0, kernels = 0, copies = 0, graph_launches = 0, frees = 0, mallocs = 0;
  uint64_t live_allocs = 0, streams = 0;
};
Stats stats();
void count_kernel();

}  // namespace fakehip
