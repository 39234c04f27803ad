// for_block_chunks' error path: a chunk that throws on a worker thread is
// rethrown on the caller after every thread has been joined, and a run
// without errors covers every block once.
#include <atomic>
#include <cstdio>
#include <stdexcept>

#include "../../treegeneralizedharmonic.jl_amd/csrc/tgh_forest.hpp"

int main() {
  std::atomic<int> seen{0};
  const int n = tgh::for_block_chunks(4096, [&](int, int b0, int b1) { seen += b1 - b0; });
  if (seen != 4096) return std::printf("covered %d of 4096 blocks\n", seen.load()), 1;
  for (int bad = 0; bad < n; ++bad) {
    std::atomic<int> done{0};
    bool caught = false;
    try {
      tgh::for_block_chunks(4096, [&](int c, int, int) {
        if (c == bad) throw std::runtime_error("chunk failed");
        ++done;
      });
    } catch (const std::runtime_error& e) {
      caught = std::string(e.what()) == "chunk failed";
    }
    if (!caught) return std::printf("chunk %d's exception was not rethrown\n", bad), 1;
    if (done != n - 1) return std::printf("chunk %d: %d other chunks finished, want %d\n", bad, done.load(), n - 1), 1;
  }
  std::printf("ok %d chunks\n", n);
  return 0;
}
