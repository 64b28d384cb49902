// Device runtime: devices, streams, events, the pooled arena, pinned staging, hiprtc
// programs and graphs. Replaces the OpenCL module (OpenCL.scala:1139-1433) and the
// Memory module (Memory.scala:13-335) for the Tensor API.
#include <dlfcn.h>
#include <hip/hiprtc.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <fstream>
#include <condition_variable>
#include <cstdio>
#include <functional>
#include <thread>

#include "internal.h"

namespace csmi {

namespace {
thread_local std::string g_last_error;
std::mutex g_devices_mu;
std::map<int, cs_device *> g_devices;
constexpr size_t kAlign = 256;
constexpr size_t kMinSlab = 256u << 20;
constexpr size_t kSlabGranule = 2u << 20;
size_t round_up(size_t v, size_t a) { return (v + a - 1) / a * a; }
}  // namespace

void set_last_error(const std::string &msg) { g_last_error = msg; }

void fail(cs_status status, const std::string &msg) { throw Error(status, msg); }

void check_hip(hipError_t e, const char *what) {
  if (e == hipSuccess) return;
  cs_status s = CS_OUT_OF_RESOURCES;
  if (e == hipErrorOutOfMemory) s = CS_MEM_OBJECT_ALLOCATION_FAILURE;
  else if (e == hipErrorInvalidValue) s = CS_INVALID_VALUE;
  else if (e == hipErrorInvalidDevice || e == hipErrorNoDevice) s = CS_DEVICE_NOT_FOUND;
  else if (e == hipErrorNotFound) s = CS_INVALID_KERNEL_NAME;
  else if (e == hipErrorInvalidHandle) s = CS_INVALID_OPERATION;
  (void)hipGetLastError();  // clear the sticky non-fatal error
  fail(s, std::string(what) + ": " + hipGetErrorString(e));
}

void set_device(cs_device *device) { CS_HIP(hipSetDevice(device->ordinal)); }

// ------------------------------------------------------------------------ arena
Arena::~Arena() {
  // Slabs are returned to HIP at process exit by the driver; freeing them here could race
  // with static destruction of the HIP runtime.
}

bool Arena::try_allocate(size_t bytes, Block *out) {
  for (auto &slab : slabs_) {
    for (auto it = slab->free_ranges.begin(); it != slab->free_ranges.end(); ++it) {
      if (it->second >= bytes) {
        size_t off = it->first, len = it->second;
        slab->free_ranges.erase(it);
        if (len > bytes) slab->free_ranges.emplace(off + bytes, len - bytes);
        slab->in_use += bytes;
        in_use_ += bytes;
        out->ptr = slab->base + off;
        out->size = bytes;
        out->slab = slab.get();
        return true;
      }
    }
  }
  return false;
}

void Arena::give_back(const Block &b) {
  Slab *slab = b.slab;
  size_t off = static_cast<size_t>(b.ptr - slab->base), len = b.size;
  auto next = slab->free_ranges.lower_bound(off);
  if (next != slab->free_ranges.end() && off + len == next->first) {
    len += next->second;
    next = slab->free_ranges.erase(next);
  }
  if (next != slab->free_ranges.begin()) {
    auto prev = std::prev(next);
    if (prev->first + prev->second == off) {
      off = prev->first;
      len += prev->second;
      slab->free_ranges.erase(prev);
    }
  }
  slab->free_ranges.emplace(off, len);
  slab->in_use -= b.size;
  in_use_ -= b.size;
}

void Arena::drain_pending(bool wait) {
  while (!pending_.empty()) {
    PendingFree &p = pending_.front();
    bool done = true;
    for (const StreamEvent &se : p.events) {
      if (wait) {
        CS_HIP(hipEventSynchronize(se.event));
      } else if (hipEventQuery(se.event) != hipSuccess) {
        (void)hipGetLastError();
        done = false;
        break;
      }
    }
    if (!done) return;
    for (const StreamEvent &se : p.events) (void)hipEventDestroy(se.event);
    give_back(p.block);
    pending_.pop_front();
  }
}

void Arena::hand_back_ordered(hipStream_t stream) {
  // Every pending block goes back to the free lists now; its release events join the
  // retired ones. A block handed back here may go to any later stream-ordered allocation, so
  // the stream waits for every retired event that has not completed, not only for those of
  // the blocks it takes (its own events need no wait: its work is ordered).
  while (!pending_.empty()) {
    PendingFree &p = pending_.front();
    for (const StreamEvent &se : p.events) retired_.push_back(se);
    give_back(p.block);
    pending_.pop_front();
  }
  reap_events(false);
  for (const StreamEvent &se : retired_) {
    if (se.stream == stream) continue;
    if (hipEventQuery(se.event) == hipSuccess) continue;
    (void)hipGetLastError();
    CS_HIP(hipStreamWaitEvent(stream, se.event, 0));
  }
}

void Arena::reap_events(bool wait) {
  while (!retired_.empty()) {
    hipEvent_t ev = retired_.front().event;
    if (wait) {
      CS_HIP(hipEventSynchronize(ev));
    } else if (hipEventQuery(ev) != hipSuccess) {
      (void)hipGetLastError();
      return;
    }
    (void)hipEventDestroy(ev);
    retired_.pop_front();
  }
}

Block Arena::allocate(size_t bytes) {
  // Deterministic: every pending free is settled before the search, and the host waits for
  // the release events of blocks handed back by stream-ordered allocations, so that every
  // free block is free for any stream.
  drain_pending(true);
  reap_events(true);
  return take(bytes);
}

Block Arena::allocate_on(size_t bytes, hipStream_t stream) {
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  CS_HIP(hipStreamIsCapturing(stream, &cap));
  if (cap != hipStreamCaptureStatusNone) {
    // a capture cannot wait on events recorded outside it: the host waits instead
    drain_pending(true);
    reap_events(true);
  } else {
    hand_back_ordered(stream);
  }
  return take(bytes);
}

Block Arena::take(size_t bytes) {
  if (bytes == 0) bytes = 1;
  bytes = round_up(bytes, kAlign);
  Block b;
  if (try_allocate(bytes, &b)) return b;
  size_t slab_bytes = std::max(kMinSlab, round_up(bytes, kSlabGranule));
  void *p = nullptr;
  CS_HIP(hipSetDevice(ordinal_));  // the slab belongs to this arena's device, whatever is current
  hipError_t e = hipMalloc(&p, slab_bytes);
  if (e != hipSuccess && slab_bytes > bytes) {
    (void)hipGetLastError();
    slab_bytes = round_up(bytes, kSlabGranule);
    e = hipMalloc(&p, slab_bytes);
  }
  if (e != hipSuccess) {
    (void)hipGetLastError();
    fail(CS_MEM_OBJECT_ALLOCATION_FAILURE,
         "hipMalloc of a " + std::to_string(slab_bytes) + "-byte arena slab failed for a " +
             std::to_string(bytes) + "-byte buffer");
  }
  auto slab = std::make_unique<Slab>();
  slab->base = static_cast<char *>(p);
  slab->size = slab_bytes;
  slab->free_ranges.emplace(0, slab_bytes);
  reserved_ += slab_bytes;
  slabs_.push_back(std::move(slab));
  if (!try_allocate(bytes, &b)) fail(CS_MEM_OBJECT_ALLOCATION_FAILURE, "arena inconsistency");
  return b;
}

void Arena::release(Block b, std::vector<StreamEvent> events) {
  pending_.push_back(PendingFree{b, std::move(events)});
  drain_pending(false);
  reap_events(false);
}

void Arena::trim() {
  settle();
  CS_HIP(hipSetDevice(ordinal_));
  for (auto it = slabs_.begin(); it != slabs_.end();) {
    if ((*it)->in_use == 0) {
      (void)hipFree((*it)->base);
      reserved_ -= (*it)->size;
      it = slabs_.erase(it);
    } else {
      ++it;
    }
  }
}

std::vector<StreamEvent> record_on_all_streams(cs_device *device) {
  std::vector<StreamEvent> evs;
  auto record = [&](hipStream_t s) {
    hipEvent_t ev;
    CS_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    CS_HIP(hipEventRecord(ev, s));
    evs.push_back(StreamEvent{s, ev});
  };
  // A stream that is capturing a graph is skipped: an event recorded there would become a
  // node of the graph, and capture starts on an idle stream (cs_graph_begin_capture waits
  // for it), so none of its pending work can still use the block.
  bool capturing = false;
  std::vector<hipStream_t> live;
  for (cs_stream *s : device->streams) {
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    CS_HIP(hipStreamIsCapturing(s->stream, &cap));
    if (cap == hipStreamCaptureStatusNone) live.push_back(s->stream);
    else capturing = true;
  }
  // The null stream covers raw launches made outside library streams; it is left alone
  // while a capture is open.
  if (!capturing) record(nullptr);
  for (hipStream_t s : live) record(s);
  return evs;
}

void *stream_scratch(cs_stream *stream, size_t bytes) {
  if (stream->scratch_bytes >= bytes) return stream->scratch;
  set_device(stream->device);
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  CS_HIP(hipStreamIsCapturing(stream->stream, &cap));
  if (cap != hipStreamCaptureStatusNone)
    fail(CS_INVALID_OPERATION, "stream scratch must grow outside graph capture; run the call once first");
  CS_HIP(hipStreamSynchronize(stream->stream));
  // kept, not freed: graphs captured earlier on this stream may still address it
  if (stream->scratch) stream->retired_scratch.push_back(stream->scratch);
  stream->scratch = nullptr;
  stream->scratch_bytes = 0;
  size_t n = round_up(bytes, 1u << 20);
  CS_HIP(hipMalloc(&stream->scratch, n));
  CS_HIP(hipMemsetAsync(stream->scratch, 0, n, stream->stream));
  CS_HIP(hipStreamSynchronize(stream->stream));
  stream->scratch_bytes = n;
  return stream->scratch;
}

// ------------------------------------------------------------------------ hiprtc
std::vector<std::string> base_compile_options(const std::string &arch) {
  return {"--offload-arch=" + arch, "-O3", "-std=c++17",
          // Packed FP32 instructions are disabled for every HIP build in this library.
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
}

static bool rtc_compile(const std::string &source, const std::vector<std::string> &options,
                        std::string *log, std::vector<char> *code) {
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, source.c_str(), "csmi_kernel.hip", 0, nullptr, nullptr) !=
      HIPRTC_SUCCESS)
    fail(CS_OUT_OF_HOST_MEMORY, "hiprtcCreateProgram failed");
  std::vector<const char *> opts;
  for (auto &o : options) opts.push_back(o.c_str());
  hiprtcResult r = hiprtcCompileProgram(prog, static_cast<int>(opts.size()), opts.data());
  size_t log_size = 0;
  hiprtcGetProgramLogSize(prog, &log_size);
  std::string l(log_size, '\0');
  if (log_size) hiprtcGetProgramLog(prog, &l[0]);
  while (!l.empty() && l.back() == '\0') l.pop_back();
  if (log) *log = l;
  bool ok = r == HIPRTC_SUCCESS;
  if (ok && code) {
    size_t n = 0;
    hiprtcGetCodeSize(prog, &n);
    code->resize(n);
    hiprtcGetCode(prog, code->data());
  }
  hiprtcDestroyProgram(&prog);
  return ok;
}

// ------------------------------------------------------------------------ code-object cache
// Compiled code objects are kept on disk, keyed by a hash of (arch, options, source), in
// $CSMI_KERNEL_CACHE or, by default, the `kcache` directory next to the library's `lib`
// directory. Generated kernels are deterministic, so a cache filled on one machine (for
// example while building) serves another with the same toolchain.
std::string kernel_cache_dir() {
  const char *env = std::getenv("CSMI_KERNEL_CACHE");
  if (env) return env;  // empty disables the cache
  Dl_info info;
  if (dladdr(reinterpret_cast<void *>(&kernel_cache_dir), &info) && info.dli_fname) {
    std::string so = info.dli_fname;
    size_t slash = so.rfind('/');
    std::string dir = slash == std::string::npos ? "." : so.substr(0, slash);
    return dir + "/../kcache";
  }
  return "";
}

static std::string cache_path(const std::string &source, const std::vector<std::string> &options) {
  std::string dir = kernel_cache_dir();
  if (dir.empty()) return "";
  std::string k;
  for (auto &o : options) k += o + "\n";
  k += "\x01" + source;
  uint64_t h = 1469598103934665603ull, h2 = 0x9e3779b97f4a7c15ull;
  for (unsigned char c : k) {
    h ^= c;
    h *= 1099511628211ull;
    h2 = (h2 ^ c) * 0x100000001b3ull + (h2 >> 29);
  }
  char name[64];
  std::snprintf(name, sizeof name, "%016llx%016llx.co", static_cast<unsigned long long>(h),
                static_cast<unsigned long long>(h2));
  return dir + "/" + name;
}

static bool cache_load(const std::string &path, std::vector<char> *code) {
  if (path.empty()) return false;
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  code->assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
  return !code->empty();
}

static void cache_store(const std::string &path, const std::vector<char> &code) {
  if (path.empty()) return;
  std::string dir = path.substr(0, path.rfind('/'));
  ::mkdir(dir.c_str(), 0755);
  // unique per writer (process and call), so that threads compiling the same kernel at once
  // never write one temporary file together; the rename publishes a complete file
  static std::atomic<unsigned> writes{0};
  std::string tmp = path + ".tmp" + std::to_string(::getpid()) + "." + std::to_string(writes.fetch_add(1));
  {
    std::ofstream f(tmp, std::ios::binary);
    if (!f) return;
    f.write(code.data(), static_cast<std::streamsize>(code.size()));
    if (!f) return;
  }
  std::rename(tmp.c_str(), path.c_str());
}

bool precompile(const std::string &source, const std::vector<std::string> &options, const std::string &arch,
                std::string *log) {
  std::vector<std::string> all = base_compile_options(arch);
  all.insert(all.end(), options.begin(), options.end());
  std::string path = cache_path(source, all);
  std::vector<char> code;
  if (cache_load(path, &code)) return true;
  if (!rtc_compile(source, all, log, &code)) return false;
  cache_store(path, code);
  return true;
}

bool compile_check(const std::string &source, const std::vector<std::string> &options,
                   const std::string &arch, std::string *log, size_t *code_size) {
  std::vector<std::string> all = base_compile_options(arch);
  all.insert(all.end(), options.begin(), options.end());
  std::vector<char> code;
  bool ok = rtc_compile(source, all, log, &code);
  if (code_size) *code_size = code.size();
  return ok;
}

hipModule_t build_module(cs_device *device, const std::string &source,
                         const std::vector<std::string> &options, std::string *log) {
  std::vector<std::string> all = base_compile_options(device->arch);
  all.insert(all.end(), options.begin(), options.end());
  std::vector<char> code;
  std::string l;
  std::string path = cache_path(source, all);
  if (!cache_load(path, &code)) {
    if (!rtc_compile(source, all, &l, &code)) {
      if (log) *log = l;
      fail(CS_BUILD_PROGRAM_FAILURE, "hiprtc build failed:\n" + l);
    }
    cache_store(path, code);
  }
  if (log) *log = l;
  set_device(device);
  hipModule_t m;
  CS_HIP(hipModuleLoadData(&m, code.data()));
  return m;
}

// ------------------------------------------------------------------------ completion thread
// Completions are delivered on one library thread, never on HIP's callback thread, so the
// callbacks may take locks that HIP calls hold.
namespace {
struct Completion {
  hipEvent_t event;
  std::function<void()> fn;
};
class CompletionThread {
 public:
  void post(hipEvent_t ev, std::function<void()> fn) {
    std::lock_guard<std::mutex> l(mu_);
    if (!started_) {
      started_ = true;
      std::thread([this] { loop(); }).detach();
    }
    queue_.push_back({ev, std::move(fn)});
    cv_.notify_one();
  }

 private:
  void loop() {
    for (;;) {
      Completion c;
      {
        std::unique_lock<std::mutex> l(mu_);
        cv_.wait(l, [this] { return !queue_.empty(); });
        c = std::move(queue_.front());
        queue_.pop_front();
      }
      (void)hipEventSynchronize(c.event);
      (void)hipEventDestroy(c.event);
      c.fn();
    }
  }
  std::mutex mu_;
  std::condition_variable cv_;
  std::deque<Completion> queue_;
  bool started_ = false;
};
CompletionThread &completions() {
  static CompletionThread *t = new CompletionThread();  // intentionally leaked
  return *t;
}
}  // namespace

}  // namespace csmi

using namespace csmi;

// ============================================================================ C ABI
extern "C" {

size_t cs_last_error(char *buf, size_t buf_size) {
  const std::string &m = g_last_error;
  if (buf && buf_size) {
    size_t n = std::min(buf_size - 1, m.size());
    std::memcpy(buf, m.data(), n);
    buf[n] = '\0';
  }
  return m.size();
}

const char *cs_status_string(cs_status status) {
  switch (status) {
    case CS_SUCCESS: return "CS_SUCCESS";
    case CS_DEVICE_NOT_FOUND: return "CS_DEVICE_NOT_FOUND";
    case CS_MEM_OBJECT_ALLOCATION_FAILURE: return "CS_MEM_OBJECT_ALLOCATION_FAILURE";
    case CS_OUT_OF_RESOURCES: return "CS_OUT_OF_RESOURCES";
    case CS_OUT_OF_HOST_MEMORY: return "CS_OUT_OF_HOST_MEMORY";
    case CS_BUILD_PROGRAM_FAILURE: return "CS_BUILD_PROGRAM_FAILURE";
    case CS_INVALID_VALUE: return "CS_INVALID_VALUE";
    case CS_INVALID_DEVICE: return "CS_INVALID_DEVICE";
    case CS_INVALID_MEM_OBJECT: return "CS_INVALID_MEM_OBJECT";
    case CS_INVALID_PROGRAM: return "CS_INVALID_PROGRAM";
    case CS_INVALID_KERNEL_NAME: return "CS_INVALID_KERNEL_NAME";
    case CS_INVALID_KERNEL_ARGS: return "CS_INVALID_KERNEL_ARGS";
    case CS_INVALID_WORK_GROUP_SIZE: return "CS_INVALID_WORK_GROUP_SIZE";
    case CS_INVALID_EVENT: return "CS_INVALID_EVENT";
    case CS_INVALID_OPERATION: return "CS_INVALID_OPERATION";
    case CS_INVALID_BUFFER_SIZE: return "CS_INVALID_BUFFER_SIZE";
    default: return "CS_UNKNOWN_STATUS";
  }
}

const char *cs_version(void) { return "csmi 0.1 gfx950"; }

void cs_free(void *p) { std::free(p); }

cs_status cs_device_count(int *count) {
  return guard([&] {
    if (!count) fail(CS_INVALID_VALUE, "count is null");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      n = 0;
    }
    *count = n;
  });
}

cs_status cs_device_open(int ordinal, cs_device **device) {
  return guard([&] {
    if (!device) fail(CS_INVALID_VALUE, "device is null");
    std::lock_guard<std::mutex> l(g_devices_mu);
    auto it = g_devices.find(ordinal);
    if (it != g_devices.end()) {
      it->second->open_count++;
      *device = it->second;
      return;
    }
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || ordinal < 0 || ordinal >= n) {
      (void)hipGetLastError();
      fail(CS_DEVICE_NOT_FOUND, "no HIP device with ordinal " + std::to_string(ordinal) +
                                    " (" + std::to_string(n) + " visible)");
    }
    std::unique_ptr<cs_device> d(new cs_device());
    d->ordinal = ordinal;
    CS_HIP(hipSetDevice(ordinal));
    CS_HIP(hipGetDeviceProperties(&d->props, ordinal));
    std::string arch = d->props.gcnArchName;
    d->arch = arch.substr(0, arch.find(':'));
    d->arena = std::make_unique<Arena>(ordinal);
    d->open_count = 1;
    g_devices[ordinal] = d.get();
    *device = d.release();
  });
}

cs_status cs_device_close(cs_device *device) {
  return guard([&] {
    if (!device) fail(CS_INVALID_DEVICE, "device is null");
    std::lock_guard<std::mutex> l(g_devices_mu);
    // The device object lives until process exit so that buffers released after the
    // last close still find their arena.
    if (device->open_count > 0) device->open_count--;
  });
}

cs_status cs_device_info(cs_device *device, int key, int64_t *value) {
  return guard([&] {
    if (!device) fail(CS_INVALID_DEVICE, "device is null");
    if (!value) fail(CS_INVALID_VALUE, "value is null");
    const hipDeviceProp_t &p = device->props;
    std::lock_guard<std::recursive_mutex> l(device->mu);
    switch (key) {
      case CS_INFO_DEVICE_TYPE: *value = 4; break;  // CL_DEVICE_TYPE_GPU
      case CS_INFO_COMPUTE_UNITS: *value = p.multiProcessorCount; break;
      case CS_INFO_MAX_WORK_GROUP_SIZE: *value = p.maxThreadsPerBlock; break;
      case CS_INFO_WAVEFRONT_SIZE: *value = p.warpSize; break;
      case CS_INFO_LDS_BYTES: *value = static_cast<int64_t>(p.sharedMemPerBlock); break;
      case CS_INFO_GLOBAL_MEM_BYTES: *value = static_cast<int64_t>(p.totalGlobalMem); break;
      case CS_INFO_L2_BYTES: *value = p.l2CacheSize; break;
      case CS_INFO_CLOCK_KHZ: *value = p.clockRate; break;
      case CS_INFO_ARENA_RESERVED_BYTES: *value = static_cast<int64_t>(device->arena->reserved()); break;
      case CS_INFO_ARENA_IN_USE_BYTES: *value = static_cast<int64_t>(device->arena->in_use()); break;
      case CS_INFO_ARENA_SLABS: *value = static_cast<int64_t>(device->arena->slab_count()); break;
      case CS_INFO_FUSED_CACHE_ENTRIES: *value = static_cast<int64_t>(device->fused_cache.size()); break;
      default: fail(CS_INVALID_VALUE, "unknown cs_info_key " + std::to_string(key));
    }
  });
}

cs_status cs_device_name(cs_device *device, const char **name) {
  return guard([&] {
    if (!device) fail(CS_INVALID_DEVICE, "device is null");
    if (!name) fail(CS_INVALID_VALUE, "name is null");
    *name = device->arch.c_str();
  });
}

// ---------------------------------------------------------------- streams
cs_status cs_stream_create(cs_device *device, cs_stream **stream) {
  return guard([&] {
    if (!device) fail(CS_INVALID_DEVICE, "device is null");
    if (!stream) fail(CS_INVALID_VALUE, "stream is null");
    set_device(device);
    hipStream_t hs = nullptr;
    // Default priority on purpose: mixed-priority streams are not used anywhere here.
    CS_HIP(hipStreamCreateWithFlags(&hs, hipStreamNonBlocking));
    auto *s = new cs_stream();
    s->device = device;
    s->stream = hs;
    std::lock_guard<std::recursive_mutex> l(device->mu);
    device->streams.insert(s);
    *stream = s;
  });
}

cs_status cs_stream_destroy(cs_stream *stream) {
  return guard([&] {
    if (!stream) fail(CS_INVALID_VALUE, "stream is null");
    set_device(stream->device);
    CS_HIP(hipStreamSynchronize(stream->stream));
    {
      std::lock_guard<std::recursive_mutex> l(stream->device->mu);
      stream->device->streams.erase(stream);
      // Pending frees hold events recorded on this stream; they are waited for and destroyed
      // while the stream still exists, so that no event outlives the stream it was recorded on.
      stream->device->arena->settle();
    }
    CS_HIP(hipStreamDestroy(stream->stream));
    if (stream->scratch) CS_HIP(hipFree(stream->scratch));
    for (void *p : stream->retired_scratch) CS_HIP(hipFree(p));
    delete stream;
  });
}

cs_status cs_stream_synchronize(cs_stream *stream) {
  return guard([&] {
    if (!stream) fail(CS_INVALID_VALUE, "stream is null");
    CS_HIP(hipStreamSynchronize(stream->stream));
  });
}

cs_status cs_stream_device(cs_stream *stream, cs_device **device) {
  return guard([&] {
    if (!stream || !device) fail(CS_INVALID_VALUE, "null argument");
    *device = stream->device;
  });
}

cs_status cs_stream_native(cs_stream *stream, void **hip_stream) {
  return guard([&] {
    if (!stream || !hip_stream) fail(CS_INVALID_VALUE, "null argument");
    *hip_stream = reinterpret_cast<void *>(stream->stream);
  });
}

// ---------------------------------------------------------------- events
cs_status cs_event_record(cs_stream *stream, cs_event **event) {
  return guard([&] {
    if (!stream || !event) fail(CS_INVALID_VALUE, "null argument");
    set_device(stream->device);
    hipEvent_t h = nullptr;
    CS_HIP(hipEventCreate(&h));
    hipError_t r = hipEventRecord(h, stream->stream);
    if (r != hipSuccess) {
      (void)hipEventDestroy(h);
      CS_HIP(r);
    }
    auto *e = new cs_event();
    e->ordinal = stream->device->ordinal;
    e->event = h;
    *event = e;
  });
}

cs_status cs_event_retain(cs_event *event) {
  return guard([&] {
    if (!event) fail(CS_INVALID_EVENT, "event is null");
    event->refs++;
  });
}

cs_status cs_event_release(cs_event *event) {
  return guard([&] {
    if (!event) fail(CS_INVALID_EVENT, "event is null");
    if (--event->refs == 0) {
      (void)hipEventDestroy(event->event);
      delete event;
    }
  });
}

cs_status cs_event_wait(cs_event *event) {
  return guard([&] {
    if (!event) fail(CS_INVALID_EVENT, "event is null");
    CS_HIP(hipEventSynchronize(event->event));
  });
}

cs_status cs_event_query(cs_event *event, int *complete) {
  return guard([&] {
    if (!event || !complete) fail(CS_INVALID_VALUE, "null argument");
    hipError_t e = hipEventQuery(event->event);
    if (e == hipErrorNotReady) {
      (void)hipGetLastError();
      *complete = 0;
      return;
    }
    CS_HIP(e);
    *complete = 1;
  });
}

cs_status cs_stream_wait_event(cs_stream *stream, cs_event *event) {
  return guard([&] {
    if (!stream || !event) fail(CS_INVALID_VALUE, "null argument");
    CS_HIP(hipStreamWaitEvent(stream->stream, event->event, 0));
  });
}

cs_status cs_event_elapsed_ms(cs_event *start, cs_event *end, float *ms) {
  return guard([&] {
    if (!start || !end || !ms) fail(CS_INVALID_VALUE, "null argument");
    CS_HIP(hipEventElapsedTime(ms, start->event, end->event));
  });
}

cs_status cs_event_on_complete(cs_event *event, void (*callback)(void *), void *user) {
  return guard([&] {
    if (!event || !callback) fail(CS_INVALID_VALUE, "null argument");
    // A private event recorded after `event` keeps the completion independent of the
    // caller's reference to `event`.
    hipEvent_t mine;
    CS_HIP(hipEventCreateWithFlags(&mine, hipEventDisableTiming));
    hipStream_t s;
    CS_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    CS_HIP(hipStreamWaitEvent(s, event->event, 0));
    CS_HIP(hipEventRecord(mine, s));
    completions().post(mine, [callback, user, s] {
      (void)hipStreamDestroy(s);
      callback(user);
    });
  });
}

// ---------------------------------------------------------------- buffers
cs_status cs_buffer_alloc(cs_device *device, size_t bytes, cs_buffer **buffer) {
  return guard([&] {
    if (!device) fail(CS_INVALID_DEVICE, "device is null");
    if (!buffer) fail(CS_INVALID_VALUE, "buffer is null");
    if (bytes > device->props.totalGlobalMem)
      fail(CS_INVALID_BUFFER_SIZE, "buffer of " + std::to_string(bytes) +
                                       " bytes exceeds device memory");
    set_device(device);
    std::lock_guard<std::recursive_mutex> l(device->mu);
    std::unique_ptr<cs_buffer> b(new cs_buffer());
    b->device = device;
    b->block = device->arena->allocate(bytes);
    b->bytes = bytes;
    *buffer = b.release();
  });
}

cs_status cs_buffer_alloc_on(cs_stream *stream, size_t bytes, cs_buffer **buffer) {
  return guard([&] {
    if (!stream) fail(CS_INVALID_VALUE, "stream is null");
    if (!buffer) fail(CS_INVALID_VALUE, "buffer is null");
    cs_device *device = stream->device;
    if (bytes > device->props.totalGlobalMem)
      fail(CS_INVALID_BUFFER_SIZE, "buffer of " + std::to_string(bytes) +
                                       " bytes exceeds device memory");
    set_device(device);
    std::lock_guard<std::recursive_mutex> l(device->mu);
    std::unique_ptr<cs_buffer> b(new cs_buffer());
    b->device = device;
    b->block = device->arena->allocate_on(bytes, stream->stream);
    b->bytes = bytes;
    *buffer = b.release();
  });
}

cs_status cs_buffer_retain(cs_buffer *buffer) {
  return guard([&] {
    if (!buffer) fail(CS_INVALID_MEM_OBJECT, "buffer is null");
    if (buffer->refs.load() <= 0) fail(CS_INVALID_MEM_OBJECT, "buffer already released");
    buffer->refs++;
  });
}

cs_status cs_buffer_release(cs_buffer *buffer) {
  return guard([&] {
    if (!buffer) fail(CS_INVALID_MEM_OBJECT, "buffer is null");
    int left = --buffer->refs;
    if (left < 0) fail(CS_INVALID_MEM_OBJECT, "buffer released more often than retained");
    if (left == 0) {
      cs_device *d = buffer->device;
      set_device(d);
      std::lock_guard<std::recursive_mutex> l(d->mu);
      d->arena->release(buffer->block, record_on_all_streams(d));
      delete buffer;
    }
  });
}

cs_status cs_buffer_device_ptr(cs_buffer *buffer, void **device_ptr) {
  return guard([&] {
    if (!buffer) fail(CS_INVALID_MEM_OBJECT, "buffer is null");
    if (!device_ptr) fail(CS_INVALID_VALUE, "device_ptr is null");
    *device_ptr = buffer->block.ptr;
  });
}

cs_status cs_buffer_size(cs_buffer *buffer, size_t *bytes) {
  return guard([&] {
    if (!buffer) fail(CS_INVALID_MEM_OBJECT, "buffer is null");
    if (!bytes) fail(CS_INVALID_VALUE, "bytes is null");
    *bytes = buffer->bytes;
  });
}

// Host copies into and out of the pinned slots on two threads: one thread copies at about
// 33 GB/s, the DMA moves about 56 GB/s. The caller copies the first half of a chunk while a
// helper thread copies the second. One helper serves every device (callers take turns); it is
// created on first use in each process (a forked child has no copy of the parent's thread) and
// never destroyed, so it may outlive static destruction at exit.
class CopyHelper {
 public:
  static CopyHelper &get() {
    static std::mutex *m = new std::mutex();
    static CopyHelper *h = nullptr;
    static pid_t owner = 0;
    std::lock_guard<std::mutex> l(*m);
    if (!h || owner != getpid()) {
      h = new CopyHelper();
      owner = getpid();
    }
    return *h;
  }
  void copy(char *dst, const char *src, size_t n) {
    constexpr size_t kMinSplit = 256u << 10;
    if (n < kMinSplit) {
      std::memcpy(dst, src, n);
      return;
    }
    std::lock_guard<std::mutex> turn(call_mu_);
    const size_t half = (n / 2) & ~size_t(4095);
    {
      std::lock_guard<std::mutex> l(mu_);
      dst_ = dst + half;
      src_ = src + half;
      n_ = n - half;
      pending_ = true;
    }
    cv_.notify_one();
    std::memcpy(dst, src, half);
    std::unique_lock<std::mutex> l(mu_);
    done_cv_.wait(l, [&] { return !pending_; });
  }

 private:
  CopyHelper() { std::thread([this] { run(); }).detach(); }
  void run() {
    std::unique_lock<std::mutex> l(mu_);
    for (;;) {
      cv_.wait(l, [&] { return pending_; });
      char *d = dst_;
      const char *s = src_;
      const size_t n = n_;
      l.unlock();
      std::memcpy(d, s, n);
      l.lock();
      pending_ = false;
      done_cv_.notify_one();
    }
  }
  std::mutex call_mu_, mu_;
  std::condition_variable cv_, done_cv_;
  char *dst_ = nullptr;
  const char *src_ = nullptr;
  size_t n_ = 0;
  bool pending_ = false;
};

static void staged_htod(cs_stream *stream, char *dst, const char *src, size_t bytes, bool blocking) {
  cs_device *d = stream->device;
  set_device(d);
  std::lock_guard<std::mutex> l(d->staging_mu);
  Staging &st = d->staging;
  if (!st.host) {
    CS_HIP(hipHostMalloc(reinterpret_cast<void **>(&st.host), Staging::kChunk * Staging::kSlots,
                         hipHostMallocDefault));
    for (auto &ev : st.done) CS_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  }
  size_t done = 0;
  while (done < bytes) {
    int slot = st.next;
    st.next = (st.next + 1) % Staging::kSlots;
    if (st.busy[slot]) CS_HIP(hipEventSynchronize(st.done[slot]));
    size_t n = std::min(Staging::kChunk, bytes - done);
    char *h = st.host + static_cast<size_t>(slot) * Staging::kChunk;
    CopyHelper::get().copy(h, src + done, n);
    CS_HIP(hipMemcpyAsync(dst + done, h, n, hipMemcpyHostToDevice, stream->stream));
    CS_HIP(hipEventRecord(st.done[slot], stream->stream));
    st.busy[slot] = true;
    done += n;
  }
  if (blocking) CS_HIP(hipStreamSynchronize(stream->stream));
}

static void staged_dtoh(cs_stream *stream, char *dst, const char *src, size_t bytes) {
  cs_device *d = stream->device;
  set_device(d);
  std::lock_guard<std::mutex> l(d->staging_mu);
  Staging &st = d->staging;
  if (!st.host) {
    CS_HIP(hipHostMalloc(reinterpret_cast<void **>(&st.host), Staging::kChunk * Staging::kSlots,
                         hipHostMallocDefault));
    for (auto &ev : st.done) CS_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  }
  for (int s = 0; s < Staging::kSlots; ++s)
    if (st.busy[s]) {
      CS_HIP(hipEventSynchronize(st.done[s]));
      st.busy[s] = false;
    }
  // Every slot in flight: chunk c goes to slot c % kSlots; once the host has drained chunk c,
  // its slot takes chunk c + kSlots.
  const size_t chunks = (bytes + Staging::kChunk - 1) / Staging::kChunk;
  auto issue = [&](size_t c) {
    const int slot = static_cast<int>(c % Staging::kSlots);
    const size_t off = c * Staging::kChunk, n = std::min(Staging::kChunk, bytes - off);
    CS_HIP(hipMemcpyAsync(st.host + slot * Staging::kChunk, src + off, n, hipMemcpyDeviceToHost,
                          stream->stream));
    CS_HIP(hipEventRecord(st.done[slot], stream->stream));
  };
  size_t issued = 0;
  for (; issued < chunks && issued < static_cast<size_t>(Staging::kSlots); ++issued) issue(issued);
  for (size_t c = 0; c < chunks; ++c) {
    const int slot = static_cast<int>(c % Staging::kSlots);
    CS_HIP(hipEventSynchronize(st.done[slot]));
    const size_t off = c * Staging::kChunk, n = std::min(Staging::kChunk, bytes - off);
    CopyHelper::get().copy(dst + off, st.host + slot * Staging::kChunk, n);
    if (issued < chunks) issue(issued++);  // slot (c + kSlots) % kSlots == slot
  }
}

cs_status cs_buffer_write(cs_stream *stream, cs_buffer *buffer, size_t offset, const void *host,
                          size_t bytes, int blocking) {
  return guard([&] {
    if (!stream || !buffer) fail(CS_INVALID_VALUE, "null stream or buffer");
    if (bytes && !host) fail(CS_INVALID_VALUE, "host is null");
    if (offset > buffer->bytes || bytes > buffer->bytes - offset)
      fail(CS_INVALID_VALUE, "write out of the buffer's bounds");
    staged_htod(stream, buffer->block.ptr + offset, static_cast<const char *>(host), bytes,
                blocking != 0);
  });
}

cs_status cs_buffer_read(cs_stream *stream, cs_buffer *buffer, size_t offset, void *host, size_t bytes) {
  return guard([&] {
    if (!stream || !buffer) fail(CS_INVALID_VALUE, "null stream or buffer");
    if (bytes && !host) fail(CS_INVALID_VALUE, "host is null");
    if (offset > buffer->bytes || bytes > buffer->bytes - offset)
      fail(CS_INVALID_VALUE, "read out of the buffer's bounds");
    staged_dtoh(stream, static_cast<char *>(host), buffer->block.ptr + offset, bytes);
  });
}

cs_status cs_memcpy_htod(cs_stream *stream, void *device_dst, const void *host_src, size_t bytes) {
  return guard([&] {
    if (!stream || (bytes && (!device_dst || !host_src))) fail(CS_INVALID_VALUE, "null argument");
    staged_htod(stream, static_cast<char *>(device_dst), static_cast<const char *>(host_src), bytes, true);
  });
}

cs_status cs_memcpy_dtoh(cs_stream *stream, void *host_dst, const void *device_src, size_t bytes) {
  return guard([&] {
    if (!stream || (bytes && (!device_src || !host_dst))) fail(CS_INVALID_VALUE, "null argument");
    staged_dtoh(stream, static_cast<char *>(host_dst), static_cast<const char *>(device_src), bytes);
  });
}

cs_status cs_memcpy_dtod(cs_stream *stream, void *device_dst, const void *device_src, size_t bytes) {
  return guard([&] {
    if (!stream || (bytes && (!device_src || !device_dst))) fail(CS_INVALID_VALUE, "null argument");
    set_device(stream->device);
    CS_HIP(hipMemcpyAsync(device_dst, device_src, bytes, hipMemcpyDeviceToDevice, stream->stream));
  });
}

cs_status cs_arena_trim(cs_device *device) {
  return guard([&] {
    if (!device) fail(CS_INVALID_DEVICE, "device is null");
    set_device(device);
    std::lock_guard<std::recursive_mutex> l(device->mu);
    device->arena->trim();
  });
}

// ---------------------------------------------------------------- programs
cs_status cs_program_build(cs_device *device, const char *source, const char *options,
                           cs_program **program, char **log) {
  if (log) *log = nullptr;
  std::string l;
  cs_status s = guard([&] {
    if (!device) fail(CS_INVALID_DEVICE, "device is null");
    if (!source || !program) fail(CS_INVALID_VALUE, "null argument");
    std::vector<std::string> opts;
    if (options) {
      std::string o(options), cur;
      for (char c : o) {
        if (c == ' ') {
          if (!cur.empty()) opts.push_back(cur);
          cur.clear();
        } else {
          cur += c;
        }
      }
      if (!cur.empty()) opts.push_back(cur);
    }
    hipModule_t m = build_module(device, source, opts, &l);
    auto *p = new cs_program();
    p->device = device;
    p->module = m;
    p->source = source;
    *program = p;
  });
  if (log && !l.empty()) {
    *log = static_cast<char *>(std::malloc(l.size() + 1));
    if (*log) std::memcpy(*log, l.c_str(), l.size() + 1);
  }
  return s;
}

cs_status cs_program_release(cs_program *program) {
  return guard([&] {
    if (!program) fail(CS_INVALID_PROGRAM, "program is null");
    if (--program->refs == 0) {
      set_device(program->device);
      (void)hipModuleUnload(program->module);
      delete program;
    }
  });
}

cs_status cs_kernel_max_block(cs_kernel *kernel, int *max_block) {
  return guard([&] {
    if (!kernel) fail(CS_INVALID_VALUE, "kernel is null");
    if (!max_block) fail(CS_INVALID_VALUE, "max_block is null");
    set_device(kernel->program->device);
    int v = 0;
    CS_HIP(hipFuncGetAttribute(&v, HIP_FUNC_ATTRIBUTE_MAX_THREADS_PER_BLOCK, kernel->function));
    *max_block = v;
  });
}

cs_status cs_program_get_kernel(cs_program *program, const char *name, cs_kernel **kernel) {
  return guard([&] {
    if (!program) fail(CS_INVALID_PROGRAM, "program is null");
    if (!name || !kernel) fail(CS_INVALID_VALUE, "null argument");
    hipFunction_t f;
    hipError_t e = hipModuleGetFunction(&f, program->module, name);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      fail(CS_INVALID_KERNEL_NAME, std::string("no kernel named ") + name);
    }
    auto *k = new cs_kernel();
    program->refs++;
    k->program = program;
    k->function = f;
    k->name = name;
    *kernel = k;
  });
}

cs_status cs_kernel_release(cs_kernel *kernel) {
  return guard([&] {
    if (!kernel) fail(CS_INVALID_VALUE, "kernel is null");
    cs_program_release(kernel->program);
    delete kernel;
  });
}

cs_status cs_launch(cs_stream *stream, cs_kernel *kernel, const uint32_t grid[3],
                    const uint32_t block[3], uint32_t dynamic_lds_bytes, void **args, int nargs) {
  return guard([&] {
    if (!stream || !kernel || !grid || !block) fail(CS_INVALID_VALUE, "null argument");
    if (nargs > 0 && !args) fail(CS_INVALID_KERNEL_ARGS, "args is null");
    uint64_t threads = uint64_t(block[0]) * block[1] * block[2];
    if (threads == 0 || threads > 1024)
      fail(CS_INVALID_WORK_GROUP_SIZE, "workgroup of " + std::to_string(threads) + " threads");
    if (uint64_t(grid[0]) * grid[1] * grid[2] == 0) return;
    set_device(stream->device);
    CS_HIP(hipModuleLaunchKernel(kernel->function, grid[0], grid[1], grid[2], block[0], block[1],
                                 block[2], dynamic_lds_bytes, stream->stream, args, nullptr));
  });
}

// ---------------------------------------------------------------- graphs
cs_status cs_graph_begin_capture(cs_stream *stream) {
  return guard([&] {
    if (!stream) fail(CS_INVALID_VALUE, "stream is null");
    set_device(stream->device);
    // Capture starts on an idle stream, so buffers released during the capture are not in
    // use by work queued on it before (their release events skip capturing streams).
    CS_HIP(hipStreamSynchronize(stream->stream));
    CS_HIP(hipStreamBeginCapture(stream->stream, hipStreamCaptureModeThreadLocal));
  });
}

cs_status cs_graph_end_capture(cs_stream *stream, cs_graph **graph) {
  return guard([&] {
    if (!stream || !graph) fail(CS_INVALID_VALUE, "null argument");
    hipGraph_t h = nullptr;
    CS_HIP(hipStreamEndCapture(stream->stream, &h));
    hipGraphExec_t x = nullptr;
    hipError_t r = hipGraphInstantiate(&x, h, nullptr, nullptr, 0);
    if (r != hipSuccess) {
      (void)hipGraphDestroy(h);
      CS_HIP(r);
    }
    // Uploaded now, so that the first launch costs what later ones do.
    r = hipGraphUpload(x, stream->stream);
    if (r != hipSuccess) {
      (void)hipGraphExecDestroy(x);
      (void)hipGraphDestroy(h);
      CS_HIP(r);
    }
    auto *g = new cs_graph();
    g->graph = h;
    g->exec = x;
    *graph = g;
  });
}

cs_status cs_graph_launch(cs_stream *stream, cs_graph *graph) {
  return guard([&] {
    if (!stream || !graph) fail(CS_INVALID_VALUE, "null argument");
    CS_HIP(hipGraphLaunch(graph->exec, stream->stream));
  });
}

cs_status cs_graph_release(cs_graph *graph) {
  return guard([&] {
    if (!graph) fail(CS_INVALID_VALUE, "graph is null");
    (void)hipGraphExecDestroy(graph->exec);
    (void)hipGraphDestroy(graph->graph);
    delete graph;
  });
}

}  // extern "C"
