// lc_host.h -- host helpers shared by the engine's translation units.
#pragma once
#include <string>

// L1 distances on the host (kind 0: xl1, 1: xl11, 2: xl12); *st = status.
double lcx_host_l1(int kind, double q, double spin, int *st);

namespace lch {
// Message for lc_thread_error() of the calling thread.
void set_thread_error(const std::string &s);
}  // namespace lch
