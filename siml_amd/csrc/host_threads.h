// host_threads.h -- default number of host worker threads of the library.
#pragma once

#include <algorithm>
#include <cstdlib>
#include <thread>

namespace siml {

// OMP_NUM_THREADS when set (the process's CPU share), else the hardware
// threads; at most cap.
inline int default_threads(int cap) {
    const char* env = std::getenv("OMP_NUM_THREADS");
    int n = env ? std::atoi(env) : 0;
    if (n <= 0) n = (int)std::max(1u, std::thread::hardware_concurrency());
    return std::min(n, cap);
}

}  // namespace siml
