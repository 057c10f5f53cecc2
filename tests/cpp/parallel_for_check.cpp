// The setup's thread pool (parallel_for in csr.hpp) hands a worker's
// exception to the caller: a std::bad_alloc thrown inside a body must reach a
// catch on the calling thread (where the setup turns it into PGS_ERROR_OOM)
// instead of terminating the process, and the remaining chunks must stop.
//
// usage: parallel_for_check   (exit 0 when every case behaves)
#include <atomic>
#include <cstdio>
#include <new>
#include <stdexcept>

#include "../../pargemslr_amd/csrc/host/csr.hpp"

int main() {
  int fails = 0;
  // 1. bad_alloc from one body on a worker thread reaches the caller
  for (int threads : {1, 4, 16}) {
    std::atomic<int> ran{0};
    bool caught = false;
    try {
      pgs::parallel_for(100000, [&](int i) {
        ran.fetch_add(1);
        if (i == 777) throw std::bad_alloc();
      }, threads);
    } catch (const std::bad_alloc&) {
      caught = true;
    }
    const bool ok = caught && ran.load() < 100000;
    std::printf("bad_alloc with %d threads: caught %d, bodies run %d %s\n", threads, (int)caught, ran.load(),
                ok ? "ok" : "FAIL");
    fails += ok ? 0 : 1;
  }
  // 2. the first exception's type is kept (a runtime_error stays one)
  try {
    pgs::parallel_for(1000, [](int i) {
      if (i % 100 == 3) throw std::runtime_error("body failed");
    }, 8);
    std::printf("runtime_error: not thrown FAIL\n");
    ++fails;
  } catch (const std::runtime_error& e) {
    std::printf("runtime_error: caught \"%s\" ok\n", e.what());
  }
  // 3. no exception: every index runs exactly once
  std::atomic<long> sum{0};
  pgs::parallel_for(50000, [&](int i) { sum.fetch_add(i); }, 8);
  const bool all = sum.load() == 50000L * 49999L / 2;
  std::printf("no exception: sum %ld %s\n", sum.load(), all ? "ok" : "FAIL");
  fails += all ? 0 : 1;
  return fails ? 1 : 0;
}
