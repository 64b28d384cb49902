// Internal declarations shared by the csmi translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/csmi.h"

namespace csmi {

// An error carrying a cs_status; converted to a return code at the C boundary.
struct Error : std::runtime_error {
  cs_status status;
  Error(cs_status s, const std::string &msg) : std::runtime_error(msg), status(s) {}
};

void set_last_error(const std::string &msg);
[[noreturn]] void fail(cs_status status, const std::string &msg);
void check_hip(hipError_t e, const char *what);

#define CS_HIP(call) ::csmi::check_hip((call), #call)

// Runs `body`, mapping exceptions to a cs_status and the thread's last error message.
template <class F>
cs_status guard(F &&body) {
  try {
    body();
    return CS_SUCCESS;
  } catch (const Error &e) {
    set_last_error(e.what());
    return e.status;
  } catch (const std::bad_alloc &) {
    set_last_error("host allocation failed");
    return CS_OUT_OF_HOST_MEMORY;
  } catch (const std::exception &e) {
    set_last_error(e.what());
    return CS_INVALID_OPERATION;
  }
}

struct Slab;

// One arena block: [ptr, ptr+size) inside `slab`.
struct Block {
  char *ptr = nullptr;
  size_t size = 0;
  Slab *slab = nullptr;
};

struct Slab {
  char *base = nullptr;
  size_t size = 0;
  std::map<size_t, size_t> free_ranges;  // offset -> length, coalesced
  size_t in_use = 0;
};

// An event recorded on a stream (nullptr: the null stream) when a block was released.
struct StreamEvent {
  hipStream_t stream;
  hipEvent_t event;
};

// A freed block waiting for the streams that may still read or write it.
struct PendingFree {
  Block block;
  std::vector<StreamEvent> events;
};

// Deterministic pooled arena over hipMalloc slabs. Allocation is first fit by
// (slab creation order, offset); frees coalesce. A freed block becomes reusable once the
// work queued on every stream of the device at release time has finished. Every allocation
// first hands back all pending frees in release order, so an identical sequence of calls
// yields identical addresses. How the allocation waits for them depends on the caller:
//   * without a stream, the host waits for the release events (the block is free for any
//     stream when the call returns);
//   * for a stream (stream-ordered), nothing waits on the host: the stream is made to wait
//     on the device for the release events of the other streams (its own work is ordered
//     already), so the block may only be used on that stream, or after it.
class Arena {
 public:
  explicit Arena(int ordinal) : ordinal_(ordinal) {}
  ~Arena();
  Block allocate(size_t bytes);
  Block allocate_on(size_t bytes, hipStream_t stream);
  void release(Block b, std::vector<StreamEvent> events);
  // Waits for every pending free and hands its block back; the release events are destroyed.
  void settle() {
    drain_pending(true);
    reap_events(true);
  }
  void trim();
  size_t reserved() const { return reserved_; }
  size_t in_use() const { return in_use_; }
  size_t slab_count() const { return slabs_.size(); }

 private:
  bool try_allocate(size_t bytes, Block *out);
  Block take(size_t bytes);
  void drain_pending(bool wait);
  // Hands back every pending free now; `stream` waits on the device for every release event
  // of another stream that has not completed (kept in retired_ until it has).
  void hand_back_ordered(hipStream_t stream);
  // Destroys retired events that have completed (all of them, waiting, with `wait`).
  void reap_events(bool wait);
  void give_back(const Block &b);
  int ordinal_;
  std::vector<std::unique_ptr<Slab>> slabs_;
  std::deque<PendingFree> pending_;
  std::deque<StreamEvent> retired_;  // release events of handed-back blocks, until complete
  size_t reserved_ = 0;
  size_t in_use_ = 0;
};

// Pinned host staging ring used by host<->device copies.
// Pinned host slots a transfer moves through chunk by chunk: the host copy of one chunk
// overlaps the DMA of the previous ones.
struct Staging {
  static constexpr size_t kChunk = 4u << 20;  // bytes per slot
  static constexpr int kSlots = 16;
  char *host = nullptr;                        // kSlots * kChunk pinned bytes
  hipEvent_t done[kSlots] = {};
  bool busy[kSlots] = {};
  int next = 0;
};

}  // namespace csmi

struct cs_device {
  int ordinal = 0;
  hipDeviceProp_t props{};
  std::string arch;
  std::recursive_mutex mu;
  std::unique_ptr<csmi::Arena> arena;
  std::set<cs_stream *> streams;
  csmi::Staging staging;
  std::mutex staging_mu;
  // Structural cache of compiled fused kernels: canonical key bytes -> kernel.
  std::unordered_map<std::string, cs_fused *> fused_cache;
  int open_count = 0;
};

struct cs_stream {
  cs_device *device = nullptr;
  hipStream_t stream = nullptr;
  // Scratch owned by this stream (reduction partials, GEMM operand copies).
  // Work on one stream is ordered, so calls on it may share the scratch.
  void *scratch = nullptr;
  size_t scratch_bytes = 0;
  // Smaller scratch buffers this one replaced. A graph captured on the stream before the
  // scratch grew still addresses them, so they are freed only with the stream.
  std::vector<void *> retired_scratch;
};

struct cs_event {
  std::atomic<int> refs{1};
  hipEvent_t event = nullptr;
  int ordinal = 0;
};

struct cs_buffer {
  std::atomic<int> refs{1};
  cs_device *device = nullptr;
  csmi::Block block;
  size_t bytes = 0;
};

struct cs_program {
  std::atomic<int> refs{1};
  cs_device *device = nullptr;
  hipModule_t module = nullptr;
  std::string source;
};

struct cs_kernel {
  cs_program *program = nullptr;
  hipFunction_t function = nullptr;
  std::string name;
};

struct cs_batch;
struct cs_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace csmi {
// Stream-owned scratch of at least `bytes`, zero-filled when (re)allocated. Growing it
// synchronizes the stream, so the first call of a given size must not be under capture.
// The buffer it replaces stays allocated until the stream is destroyed.
void *stream_scratch(cs_stream *stream, size_t bytes);
// FP64 block partials of a sum in the stream's scratch (n of them), and the fixed-order FP64
// combine that turns n partials into *out (FP32, nullable) and *out64 (nullable).
double *sum_partials(cs_stream *stream, size_t n);
void launch_finish_sum(cs_stream *stream, const double *partials, int64_t n, float *out, double *out64);
// Events recorded now on every live stream of the device (for stream-ordered frees).
std::vector<StreamEvent> record_on_all_streams(cs_device *device);
// Makes `device` current on the calling thread.
void set_device(cs_device *device);
// hiprtc build of `source` for the device's arch; throws Error with the log on failure.
hipModule_t build_module(cs_device *device, const std::string &source,
                         const std::vector<std::string> &options, std::string *log);
// Compile-only check (no device needed); returns true on success and fills *log.
bool compile_check(const std::string &source, const std::vector<std::string> &options,
                   const std::string &arch, std::string *log, size_t *code_size);
// Compiles into the on-disk code-object cache without a device; true if cached or built.
bool precompile(const std::string &source, const std::vector<std::string> &options, const std::string &arch,
                std::string *log);
// Directory of the on-disk code-object cache ("" when disabled).
std::string kernel_cache_dir();
// Options every HIP build in this library uses.
std::vector<std::string> base_compile_options(const std::string &arch);
}  // namespace csmi
