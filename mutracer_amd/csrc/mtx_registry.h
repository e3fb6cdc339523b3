// Handle registry: uint64 handles (monotonic, never reused) to resident objects.
#pragma once
#include <atomic>
#include <memory>
#include <mutex>
#include <unordered_map>

#include "mtx_scene.h"

namespace mtx {

struct Context;
struct Server;

struct Registry {
  std::mutex m;
  std::atomic<uint64_t> nextHandle{1};
  std::unordered_map<uint64_t, std::shared_ptr<Blas>> blas;
  std::unordered_map<uint64_t, std::shared_ptr<Accel>> accel;
  std::unordered_map<uint64_t, std::shared_ptr<Context>> contexts;
  std::unordered_map<uint64_t, std::shared_ptr<Server>> servers;
  uint64_t next();
  std::shared_ptr<Blas> get_blas(uint64_t h) {
    std::lock_guard<std::mutex> g(m);
    auto it = blas.find(h);
    if (it == blas.end()) throw Error(MTX_INVALID_HANDLE, "unknown BLAS handle");
    return it->second;
  }
  std::shared_ptr<Accel> get_accel(uint64_t h) {
    std::lock_guard<std::mutex> g(m);
    auto it = accel.find(h);
    if (it == accel.end()) throw Error(MTX_INVALID_HANDLE, "unknown acceleration handle");
    return it->second;
  }
};

Registry &registry();
uint32_t trace_grid();
int32_t *spill_for(hipStream_t s, uint32_t blocksNeeded, uint32_t blockThreads);

}  // namespace mtx
