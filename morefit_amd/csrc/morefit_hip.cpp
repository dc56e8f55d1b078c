// morefit_hip.cpp -- libmorefit_hip.so: the C ABI declared in include/morefit_hip.h.
//
// Owns the HIP device/stream, device memory, hipRTC compilation (gfx950) of generated
// likelihood modules with an in-memory and on-disk code-object cache, kernel launches,
// hipGraph capture, HIP-event timing and an HBM copy-ceiling kernel.  No torch types and no
// C++ types cross the boundary.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <dlfcn.h>
#include <sys/stat.h>
#include <unistd.h>

#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "morefit_hip.h"
#include "kernel_runtime_src.h"  // MF_ABI_SRC, MF_RUNTIME_SRC (generated by the Makefile)

namespace {

thread_local std::string g_last_error;
std::mutex g_fault_mu;
std::map<std::string, int> g_faults;

int set_error(int code, const char* fmt, ...) {
  char buf[2048];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

bool injected(const char* entry) {
  std::lock_guard<std::mutex> lk(g_fault_mu);
  auto it = g_faults.find(entry);
  if (it == g_faults.end() || it->second <= 0) return false;
  it->second--;
  g_last_error = std::string("injected failure in ") + entry;
  return true;
}

#define MF_INJECT(name) \
  if (injected(name)) return MF_ERR_INJECTED;
#define MF_HIP(call, what)                                                                          \
  do {                                                                                              \
    hipError_t e_ = (call);                                                                         \
    if (e_ != hipSuccess) return set_error(MF_ERR_HIP, "%s: %s (%s)", what, hipGetErrorString(e_), #call); \
  } while (0)

uint64_t fnv1a(const std::string& s, uint64_t h = 1469598103934665603ULL) {
  for (unsigned char c : s) {
    h ^= c;
    h *= 1099511628211ULL;
  }
  return h;
}

std::string library_dir() {
  Dl_info info;
  if (dladdr((void*)&mf_last_error, &info) && info.dli_fname) {
    std::string p = info.dli_fname;
    size_t k = p.rfind('/');
    if (k != std::string::npos) return p.substr(0, k);
  }
  return ".";
}

std::string cache_dir() {
  const char* env = getenv("MOREFIT_KCACHE");
  if (env && *env) return env;
  return library_dir() + "/../kcache";
}

const char* kArch = "gfx950";
const char* kMarker = "//@MF_RUNTIME@";

std::string assemble(const char* generated) {
  std::string g(generated);
  size_t m = g.find(kMarker);
  std::string prelude = m == std::string::npos ? std::string() : g.substr(0, m);
  std::string body = m == std::string::npos ? g : g.substr(m + strlen(kMarker));
  return prelude + "\n#define MF_KERNEL_SIDE 1\n" + MF_ABI_SRC + "\n" + MF_RUNTIME_SRC + "\n" + body;
}

std::vector<std::string> compile_options(const char* extra) {
  std::vector<std::string> o = {std::string("--offload-arch=") + kArch, "-O3", "-std=c++17",
                                "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
  if (extra) {
    std::istringstream is(extra);
    std::string t;
    while (is >> t) o.push_back(t);
  }
  return o;
}

int rtc_compile(const std::string& src, const std::vector<std::string>& opts, std::vector<char>& code, double* ms) {
  auto t0 = std::chrono::steady_clock::now();
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "morefit_module.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS)
    return set_error(MF_ERR_COMPILE, "hiprtcCreateProgram failed");
  std::vector<const char*> co;
  for (auto& s : opts) co.push_back(s.c_str());
  hiprtcResult r = hiprtcCompileProgram(prog, (int)co.size(), co.data());
  if (r != HIPRTC_SUCCESS) {
    size_t n = 0;
    hiprtcGetProgramLogSize(prog, &n);
    std::string log(n + 1, '\0');
    hiprtcGetProgramLog(prog, &log[0]);
    hiprtcDestroyProgram(&prog);
    return set_error(MF_ERR_COMPILE, "hipRTC compilation failed:\n%s", log.c_str());
  }
  size_t n = 0;
  hiprtcGetCodeSize(prog, &n);
  code.resize(n);
  hiprtcGetCode(prog, code.data());
  hiprtcDestroyProgram(&prog);
  if (ms) *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return MF_OK;
}

bool read_file(const std::string& path, std::vector<char>& out) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  out.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
  return !out.empty();
}

void write_file_atomic(const std::string& path, const std::vector<char>& data) {
  std::string dir = path.substr(0, path.rfind('/'));
  mkdir(dir.c_str(), 0755);
  std::string tmp = path + ".tmp" + std::to_string(getpid());
  {
    std::ofstream f(tmp, std::ios::binary);
    if (!f) return;
    f.write(data.data(), (std::streamsize)data.size());
  }
  rename(tmp.c_str(), path.c_str());
}

// code object for (source, options): memory -> disk -> hipRTC
int get_code(const char* generated, const char* options, std::vector<char>& code, uint64_t* hash_out, double* compile_ms) {
  std::string src = assemble(generated);
  auto opts = compile_options(options);
  std::string key = src;
  for (auto& o : opts) key += "\n" + o;
  uint64_t h = fnv1a(key);
  if (hash_out) *hash_out = h;
  char name[64];
  snprintf(name, sizeof(name), "/%016llx.co", (unsigned long long)h);
  std::string path = cache_dir() + name;
  if (compile_ms) *compile_ms = 0.0;
  if (read_file(path, code)) return MF_OK;
  int rc = rtc_compile(src, opts, code, compile_ms);
  if (rc != MF_OK) return rc;
  write_file_atomic(path, code);
  return MF_OK;
}

// built-in kernels (compiled into this library for gfx950)
__global__ void k_fill_zero(unsigned int* p, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = 0u;
}
__global__ void k_fill_zero_bytes(unsigned char* p, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}
typedef float mf_v4f __attribute__((ext_vector_type(4)));
__global__ void __launch_bounds__(256) k_copy(const mf_v4f* __restrict__ src, mf_v4f* __restrict__ dst, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

}  // namespace

struct mf_module;

struct mf_device {
  int ordinal = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  mf_device_props props{};
  std::unordered_map<uint64_t, mf_module*> modules;
  std::mutex mu;
};

struct mf_module {
  mf_device* dev = nullptr;
  hipModule_t mod = nullptr;
  std::map<std::string, hipFunction_t> fns;
  uint64_t hash = 0;
  int refs = 0;
  double compile_ms = 0.0;
};

struct mf_graph {
  mf_device* dev = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

extern "C" {

const char* mf_last_error(void) { return g_last_error.c_str(); }

int mf_debug_fail_next(const char* entry_point, int count) {
  if (!entry_point) return set_error(MF_ERR_INVALID, "mf_debug_fail_next: null entry point");
  std::lock_guard<std::mutex> lk(g_fault_mu);
  g_faults[entry_point] = count;
  return MF_OK;
}

int mf_device_open(int ordinal, mf_device** out) {
  MF_INJECT("mf_device_open");
  if (!out) return set_error(MF_ERR_INVALID, "mf_device_open: null output");
  int n = 0;
  MF_HIP(hipGetDeviceCount(&n), "mf_device_open");
  if (ordinal < 0 || ordinal >= n) return set_error(MF_ERR_INVALID, "mf_device_open: device %d of %d", ordinal, n);
  MF_HIP(hipSetDevice(ordinal), "mf_device_open");
  hipDeviceProp_t p;
  MF_HIP(hipGetDeviceProperties(&p, ordinal), "mf_device_open");
  if (strncmp(p.gcnArchName, kArch, strlen(kArch)) != 0)
    return set_error(MF_ERR_INVALID, "mf_device_open: device %d is %s, this library is built for %s", ordinal, p.gcnArchName, kArch);
  auto* d = new mf_device();
  d->ordinal = ordinal;
  if (hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&d->ev0) != hipSuccess ||
      hipEventCreate(&d->ev1) != hipSuccess) {
    delete d;
    return set_error(MF_ERR_HIP, "mf_device_open: stream/event creation failed");
  }
  snprintf(d->props.name, sizeof(d->props.name), "%s", p.name);
  snprintf(d->props.arch, sizeof(d->props.arch), "%s", p.gcnArchName);
  d->props.compute_units = p.multiProcessorCount;
  d->props.wavefront_size = p.warpSize;
  d->props.total_mem = p.totalGlobalMem;
  d->props.clock_khz = p.clockRate;
  d->props.l2_bytes = p.l2CacheSize;
  *out = d;
  return MF_OK;
}

int mf_device_close(mf_device* dev) {
  if (!dev) return MF_OK;
  (void)hipSetDevice(dev->ordinal);
  (void)hipStreamSynchronize(dev->stream);
  for (auto& kv : dev->modules) {
    (void)hipModuleUnload(kv.second->mod);
    delete kv.second;
  }
  (void)hipEventDestroy(dev->ev0);
  (void)hipEventDestroy(dev->ev1);
  (void)hipStreamDestroy(dev->stream);
  delete dev;
  return MF_OK;
}

int mf_device_query(mf_device* dev, mf_device_props* props) {
  if (!dev || !props) return set_error(MF_ERR_INVALID, "mf_device_query: null argument");
  size_t fr = 0, tot = 0;
  MF_HIP(hipSetDevice(dev->ordinal), "mf_device_query");
  MF_HIP(hipMemGetInfo(&fr, &tot), "mf_device_query");
  dev->props.free_mem = fr;
  *props = dev->props;
  return MF_OK;
}

int mf_malloc(mf_device* dev, size_t bytes, void** dptr) {
  MF_INJECT("mf_malloc");
  if (!dev || !dptr) return set_error(MF_ERR_INVALID, "mf_malloc: null argument");
  *dptr = nullptr;
  if (bytes == 0) return MF_OK;
  hipError_t e = hipMalloc(dptr, bytes);
  if (e != hipSuccess) return set_error(MF_ERR_NOMEM, "mf_malloc(%zu): %s", bytes, hipGetErrorString(e));
  return MF_OK;
}

int mf_free(mf_device* dev, void* dptr) {
  if (!dev) return set_error(MF_ERR_INVALID, "mf_free: null device");
  if (!dptr) return MF_OK;
  MF_HIP(hipFree(dptr), "mf_free");
  return MF_OK;
}

int mf_host_alloc(mf_device* dev, size_t bytes, void** hptr) {
  MF_INJECT("mf_host_alloc");
  if (!dev || !hptr) return set_error(MF_ERR_INVALID, "mf_host_alloc: null argument");
  *hptr = nullptr;
  if (bytes == 0) return MF_OK;
  hipError_t e = hipHostMalloc(hptr, bytes, hipHostMallocDefault);
  if (e != hipSuccess) return set_error(MF_ERR_NOMEM, "mf_host_alloc(%zu): %s", bytes, hipGetErrorString(e));
  return MF_OK;
}

int mf_host_free(mf_device* dev, void* hptr) {
  if (!dev) return set_error(MF_ERR_INVALID, "mf_host_free: null device");
  if (!hptr) return MF_OK;
  MF_HIP(hipHostFree(hptr), "mf_host_free");
  return MF_OK;
}

int mf_memcpy_h2d(mf_device* dev, void* dst, const void* src, size_t bytes) {
  MF_INJECT("mf_memcpy_h2d");
  if (!dev || (!dst && bytes) || (!src && bytes)) return set_error(MF_ERR_INVALID, "mf_memcpy_h2d: null argument");
  if (!bytes) return MF_OK;
  MF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, dev->stream), "mf_memcpy_h2d");
  return MF_OK;
}

int mf_memcpy_d2h(mf_device* dev, void* dst, const void* src, size_t bytes) {
  MF_INJECT("mf_memcpy_d2h");
  if (!dev || (!dst && bytes) || (!src && bytes)) return set_error(MF_ERR_INVALID, "mf_memcpy_d2h: null argument");
  if (!bytes) return MF_OK;
  MF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, dev->stream), "mf_memcpy_d2h");
  return MF_OK;
}

int mf_memcpy_d2d(mf_device* dev, void* dst, const void* src, size_t bytes) {
  if (!dev || (!dst && bytes) || (!src && bytes)) return set_error(MF_ERR_INVALID, "mf_memcpy_d2d: null argument");
  if (!bytes) return MF_OK;
  MF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, dev->stream), "mf_memcpy_d2d");
  return MF_OK;
}

int mf_fill_zero(mf_device* dev, void* dptr, size_t bytes) {
  if (!dev || (!dptr && bytes)) return set_error(MF_ERR_INVALID, "mf_fill_zero: null argument");
  if (!bytes) return MF_OK;
  if ((bytes % 4) == 0 && ((uintptr_t)dptr % 4) == 0) {
    size_t n = bytes / 4;
    int grid = (int)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill_zero, dim3(grid), dim3(256), 0, dev->stream, (unsigned int*)dptr, n);
  } else {
    int grid = (int)((bytes + 255) / 256);
    hipLaunchKernelGGL(k_fill_zero_bytes, dim3(grid), dim3(256), 0, dev->stream, (unsigned char*)dptr, bytes);
  }
  MF_HIP(hipGetLastError(), "mf_fill_zero");
  return MF_OK;
}

int mf_sync(mf_device* dev) {
  if (!dev) return set_error(MF_ERR_INVALID, "mf_sync: null device");
  MF_HIP(hipStreamSynchronize(dev->stream), "mf_sync");
  return MF_OK;
}

int mf_module_compile_only(const char* generated_src, const char* options, size_t* code_bytes) {
  MF_INJECT("mf_module_compile_only");
  if (!generated_src) return set_error(MF_ERR_INVALID, "mf_module_compile_only: null source");
  std::vector<char> code;
  int rc = get_code(generated_src, options, code, nullptr, nullptr);
  if (rc == MF_OK && code_bytes) *code_bytes = code.size();
  return rc;
}

int mf_module_load(mf_device* dev, const char* generated_src, const char* options, mf_module** out) {
  MF_INJECT("mf_module_load");
  if (!dev || !generated_src || !out) return set_error(MF_ERR_INVALID, "mf_module_load: null argument");
  std::vector<char> code;
  uint64_t h = 0;
  double ms = 0.0;
  {
    // memory cache first (same source + options on the same device)
    std::string src = assemble(generated_src);
    auto opts = compile_options(options);
    std::string key = src;
    for (auto& o : opts) key += "\n" + o;
    h = fnv1a(key);
    std::lock_guard<std::mutex> lk(dev->mu);
    auto it = dev->modules.find(h);
    if (it != dev->modules.end()) {
      it->second->refs++;
      *out = it->second;
      return MF_OK;
    }
  }
  int rc = get_code(generated_src, options, code, &h, &ms);
  if (rc != MF_OK) return rc;
  MF_HIP(hipSetDevice(dev->ordinal), "mf_module_load");
  hipModule_t hm;
  MF_HIP(hipModuleLoadData(&hm, code.data()), "mf_module_load");
  auto* m = new mf_module();
  m->dev = dev;
  m->mod = hm;
  m->hash = h;
  m->refs = 1;
  m->compile_ms = ms;
  std::lock_guard<std::mutex> lk(dev->mu);
  dev->modules[h] = m;
  *out = m;
  return MF_OK;
}

int mf_module_unload(mf_module* mod) {
  if (!mod) return MF_OK;
  mf_device* dev = mod->dev;
  std::lock_guard<std::mutex> lk(dev->mu);
  if (--mod->refs > 0) return MF_OK;
  dev->modules.erase(mod->hash);
  (void)hipModuleUnload(mod->mod);
  delete mod;
  return MF_OK;
}

double mf_module_compile_ms(mf_module* mod) { return mod ? mod->compile_ms : 0.0; }

static bool mod_has(mf_module* mod, const char* kernel, hipFunction_t* fn) {
  auto it = mod->fns.find(kernel);
  if (it != mod->fns.end()) {
    *fn = it->second;
    return it->second != nullptr;
  }
  hipError_t e = hipModuleGetFunction(fn, mod->mod, kernel);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    mod->fns[kernel] = nullptr;  // remember the absence
    return false;
  }
  mod->fns[kernel] = *fn;
  return true;
}

static int get_fn(mf_module* mod, const char* kernel, hipFunction_t* fn) {
  auto it = mod->fns.find(kernel);
  if (it != mod->fns.end() && it->second) {
    *fn = it->second;
    return MF_OK;
  }
  hipError_t e = hipModuleGetFunction(fn, mod->mod, kernel);
  if (e != hipSuccess) return set_error(MF_ERR_NOKERNEL, "kernel %s not in module: %s", kernel, hipGetErrorString(e));
  mod->fns[kernel] = *fn;
  return MF_OK;
}

int mf_module_has_kernel(mf_module* mod, const char* kernel) {
  if (!mod || !kernel) return 0;
  hipFunction_t fn;
  return get_fn(mod, kernel, &fn) == MF_OK ? 1 : 0;
}

int mf_launch_eval(mf_module* mod, const char* kernel, const mf_eval_args* args, int grid, int block) {
  MF_INJECT("mf_launch_eval");
  if (!mod || !kernel || !args) return set_error(MF_ERR_INVALID, "mf_launch_eval: null argument");
  if (grid <= 0 || block <= 0 || block % 64 != 0 || block > 1024)
    return set_error(MF_ERR_INVALID, "mf_launch_eval: bad launch shape %d x %d", grid, block);
  if (args->nchunks < 0 || args->nseg < 0 || !args->segs || !args->chunks || !args->params || !args->data)
    return set_error(MF_ERR_INVALID, "mf_launch_eval: incomplete argument block");
  if (args->nchunks == 0) return MF_OK;
  hipFunction_t fn;
  int rc = get_fn(mod, kernel, &fn);
  if (rc != MF_OK) return rc;
  mf_eval_args a = *args;
  void* kp[] = {(void*)&a};
  MF_HIP(hipModuleLaunchKernel(fn, grid, 1, 1, block, 1, 1, 0, mod->dev->stream, kp, nullptr), "mf_launch_eval");
  std::string fin = std::string(kernel) + "_final";
  hipFunction_t ff;
  if (args->nseg > 0 && mod_has(mod, fin.c_str(), &ff))
    MF_HIP(hipModuleLaunchKernel(ff, args->nseg, 1, 1, 64, 1, 1, 0, mod->dev->stream, kp, nullptr), "mf_launch_eval final");
  return MF_OK;
}

int mf_launch_prep(mf_module* mod, const double* theta, void* params, const int* seg_list, int nseg) {
  MF_INJECT("mf_launch_prep");
  if (!mod || !theta || !params || nseg < 0) return set_error(MF_ERR_INVALID, "mf_launch_prep: bad argument");
  if (nseg == 0) return MF_OK;
  hipFunction_t fn;
  int rc = get_fn(mod, "mf_prep", &fn);
  if (rc != MF_OK) return rc;
  void* kp[] = {(void*)&theta, (void*)&params, (void*)&seg_list, (void*)&nseg};
  int grid = (nseg + 63) / 64;
  MF_HIP(hipModuleLaunchKernel(fn, grid, 1, 1, 64, 1, 1, 0, mod->dev->stream, kp, nullptr), "mf_launch_prep");
  return MF_OK;
}

int mf_launch_generate(mf_module* mod, const mf_gen_args* args, int grid, int block) {
  MF_INJECT("mf_launch_generate");
  if (!mod || !args || grid <= 0 || block <= 0) return set_error(MF_ERR_INVALID, "mf_launch_generate: bad argument");
  if (args->nstreams == 0) return MF_OK;
  if ((long long)grid * block < args->nstreams) return set_error(MF_ERR_INVALID, "mf_launch_generate: grid covers fewer threads than streams");
  hipFunction_t fn;
  int rc = get_fn(mod, "mf_generate", &fn);
  if (rc != MF_OK) return rc;
  mf_gen_args a = *args;
  void* kp[] = {(void*)&a};
  MF_HIP(hipModuleLaunchKernel(fn, grid, 1, 1, block, 1, 1, 0, mod->dev->stream, kp, nullptr), "mf_launch_generate");
  return MF_OK;
}

int mf_graph_begin(mf_device* dev) {
  if (!dev) return set_error(MF_ERR_INVALID, "mf_graph_begin: null device");
  MF_HIP(hipStreamBeginCapture(dev->stream, hipStreamCaptureModeThreadLocal), "mf_graph_begin");
  return MF_OK;
}

int mf_graph_end(mf_device* dev, mf_graph** out) {
  if (!dev || !out) return set_error(MF_ERR_INVALID, "mf_graph_end: null argument");
  hipGraph_t g;
  MF_HIP(hipStreamEndCapture(dev->stream, &g), "mf_graph_end");
  hipGraphExec_t ex;
  hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(g);
    return set_error(MF_ERR_HIP, "mf_graph_end: instantiate failed: %s", hipGetErrorString(e));
  }
  auto* r = new mf_graph();
  r->dev = dev;
  r->graph = g;
  r->exec = ex;
  *out = r;
  return MF_OK;
}

int mf_graph_launch(mf_graph* g) {
  MF_INJECT("mf_graph_launch");
  if (!g) return set_error(MF_ERR_INVALID, "mf_graph_launch: null graph");
  MF_HIP(hipGraphLaunch(g->exec, g->dev->stream), "mf_graph_launch");
  return MF_OK;
}

int mf_graph_destroy(mf_graph* g) {
  if (!g) return MF_OK;
  (void)hipGraphExecDestroy(g->exec);
  (void)hipGraphDestroy(g->graph);
  delete g;
  return MF_OK;
}

int mf_timer_begin(mf_device* dev) {
  if (!dev) return set_error(MF_ERR_INVALID, "mf_timer_begin: null device");
  MF_HIP(hipEventRecord(dev->ev0, dev->stream), "mf_timer_begin");
  return MF_OK;
}

int mf_timer_end(mf_device* dev, float* ms) {
  if (!dev || !ms) return set_error(MF_ERR_INVALID, "mf_timer_end: null argument");
  MF_HIP(hipEventRecord(dev->ev1, dev->stream), "mf_timer_end");
  MF_HIP(hipEventSynchronize(dev->ev1), "mf_timer_end");
  MF_HIP(hipEventElapsedTime(ms, dev->ev0, dev->ev1), "mf_timer_end");
  return MF_OK;
}

int mf_bench_copy(mf_device* dev, size_t bytes, int iters, double* gbps) {
  if (!dev || !gbps || bytes < 16 || iters <= 0) return set_error(MF_ERR_INVALID, "mf_bench_copy: bad argument");
  size_t n = bytes / 16;
  void *a = nullptr, *b = nullptr;
  int rc = mf_malloc(dev, n * 16, &a);
  if (rc) return rc;
  rc = mf_malloc(dev, n * 16, &b);
  if (rc) {
    mf_free(dev, a);
    return rc;
  }
  mf_fill_zero(dev, a, n * 16);
  int grid = dev->props.compute_units * 16;
  float best = 1e30f;
  for (int it = 0; it < iters + 1; it++) {
    mf_timer_begin(dev);
    hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, dev->stream, (const mf_v4f*)a, (mf_v4f*)b, n);
    float ms = 0;
    rc = mf_timer_end(dev, &ms);
    if (rc) break;
    if (it > 0 && ms < best) best = ms;  // first launch is a warm-up
  }
  mf_free(dev, a);
  mf_free(dev, b);
  if (rc) return rc;
  *gbps = 2.0 * (double)(n * 16) / (best * 1e-3) / 1e9;
  return MF_OK;
}

}  // extern "C"
