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

// ---------------------------------------------------------------- code-object metadata
// A gfx950 code object is an ELF file whose SHT_NOTE section holds an NT_AMDGPU_METADATA
// (type 32, owner "AMDGPU") note: a MessagePack map with "amdhsa.kernels", one map per
// kernel carrying .name, .vgpr_count, .agpr_count, ... .  Only what that needs is read here.
struct MpVal {
  enum Kind { NIL, BOOL, INT, FLOAT, STR, BIN, ARR, MAP, EXT } kind = NIL;
  long long i = 0;
  std::string s;
  std::vector<MpVal> items;  // array items; for a map: key, value, key, value, ...
  const MpVal* get(const char* key) const {
    if (kind != MAP) return nullptr;
    for (size_t k = 0; k + 1 < items.size(); k += 2)
      if (items[k].kind == STR && items[k].s == key) return &items[k + 1];
    return nullptr;
  }
};

struct MpReader {
  const unsigned char *p, *e;
  bool ok = true;
  uint64_t be(int n) {
    if (e - p < n) {
      ok = false;
      return 0;
    }
    uint64_t v = 0;
    for (int k = 0; k < n; k++) v = (v << 8) | p[k];
    p += n;
    return v;
  }
  void bytes(size_t n, std::string* out) {
    if ((size_t)(e - p) < n) {
      ok = false;
      return;
    }
    if (out) out->assign((const char*)p, n);
    p += n;
  }
  void seq(MpVal& v, uint64_t n, bool map, int depth) {
    v.kind = map ? MpVal::MAP : MpVal::ARR;
    uint64_t m = map ? 2 * n : n;
    if (m > (uint64_t)(e - p)) {  // every element takes at least one byte
      ok = false;
      return;
    }
    v.items.resize(m);
    for (uint64_t k = 0; k < m && ok; k++) value(v.items[k], depth + 1);
  }
  void value(MpVal& v, int depth = 0) {
    if (depth > 64 || p >= e) {
      ok = false;
      return;
    }
    unsigned c = *p++;
    if (c <= 0x7f) { v.kind = MpVal::INT; v.i = c; }
    else if (c >= 0xe0) { v.kind = MpVal::INT; v.i = (int)c - 256; }
    else if ((c & 0xf0) == 0x80) seq(v, c & 0x0f, true, depth);
    else if ((c & 0xf0) == 0x90) seq(v, c & 0x0f, false, depth);
    else if ((c & 0xe0) == 0xa0) { v.kind = MpVal::STR; bytes(c & 0x1f, &v.s); }
    else switch (c) {
      case 0xc0: v.kind = MpVal::NIL; break;
      case 0xc2: case 0xc3: v.kind = MpVal::BOOL; v.i = c == 0xc3; break;
      case 0xc4: case 0xc5: case 0xc6: v.kind = MpVal::BIN; bytes(be(1 << (c - 0xc4)), nullptr); break;
      case 0xc7: case 0xc8: case 0xc9: { v.kind = MpVal::EXT; uint64_t n = be(1 << (c - 0xc7)); be(1); bytes(n, nullptr); break; }
      case 0xca: v.kind = MpVal::FLOAT; be(4); break;
      case 0xcb: v.kind = MpVal::FLOAT; be(8); break;
      case 0xcc: case 0xcd: case 0xce: case 0xcf: v.kind = MpVal::INT; v.i = (long long)be(1 << (c - 0xcc)); break;
      case 0xd0: v.kind = MpVal::INT; v.i = (int8_t)be(1); break;
      case 0xd1: v.kind = MpVal::INT; v.i = (int16_t)be(2); break;
      case 0xd2: v.kind = MpVal::INT; v.i = (int32_t)be(4); break;
      case 0xd3: v.kind = MpVal::INT; v.i = (long long)be(8); break;
      case 0xd4: case 0xd5: case 0xd6: case 0xd7: case 0xd8: v.kind = MpVal::EXT; be(1); bytes((size_t)1 << (c - 0xd4), nullptr); break;
      case 0xd9: case 0xda: case 0xdb: v.kind = MpVal::STR; bytes(be(1 << (c - 0xd9)), &v.s); break;
      case 0xdc: case 0xdd: seq(v, be(c == 0xdc ? 2 : 4), false, depth); break;
      case 0xde: case 0xdf: seq(v, be(c == 0xde ? 2 : 4), true, depth); break;
      default: ok = false;
    }
  }
};

template <typename T>
bool rd_le(const std::vector<char>& b, size_t off, T& v) {
  if (off + sizeof(T) > b.size()) return false;
  std::memcpy(&v, b.data() + off, sizeof(T));
  return true;
}

// metadata note of a code object (an ELF64 image, or a clang offload bundle holding one)
bool code_metadata(const std::vector<char>& code, MpVal& meta) {
  static const char kBundle[] = "__CLANG_OFFLOAD_BUNDLE__";
  if (code.size() > 32 && std::memcmp(code.data(), kBundle, 24) == 0) {
    uint64_t n = 0;
    rd_le(code, 24, n);
    size_t q = 32;
    for (uint64_t k = 0; k < n && k < 64; k++) {
      uint64_t off = 0, size = 0, idlen = 0;
      if (!rd_le(code, q, off) || !rd_le(code, q + 8, size) || !rd_le(code, q + 16, idlen)) return false;
      if (q + 24 + idlen > code.size()) return false;
      std::string id(code.data() + q + 24, idlen);
      q += 24 + idlen;
      if (id.find(kArch) != std::string::npos && off + size <= code.size()) {
        std::vector<char> inner(code.begin() + off, code.begin() + off + size);
        return code_metadata(inner, meta);
      }
    }
    return false;
  }
  if (code.size() < 64 || std::memcmp(code.data(), "\x7f" "ELF", 4) != 0 || code[4] != 2 || code[5] != 1) return false;
  uint64_t shoff = 0;
  uint16_t shentsize = 0, shnum = 0;
  if (!rd_le(code, 0x28, shoff) || !rd_le(code, 0x3a, shentsize) || !rd_le(code, 0x3c, shnum)) return false;
  for (uint16_t k = 0; k < shnum; k++) {
    size_t sh = shoff + (size_t)k * shentsize;
    uint32_t type = 0;
    uint64_t off = 0, size = 0;
    if (!rd_le(code, sh + 4, type) || !rd_le(code, sh + 0x18, off) || !rd_le(code, sh + 0x20, size)) return false;
    if (type != 7 || off + size > code.size()) continue;  // SHT_NOTE
    size_t q = off, end = off + size;
    while (q + 12 <= end) {
      uint32_t namesz = 0, descsz = 0, ntype = 0;
      rd_le(code, q, namesz);
      rd_le(code, q + 4, descsz);
      rd_le(code, q + 8, ntype);
      size_t name = q + 12, desc = name + ((namesz + 3u) & ~3u);
      if (desc + descsz > end) break;
      if (ntype == 32 && namesz >= 6 && std::memcmp(code.data() + name, "AMDGPU", 6) == 0) {
        MpReader r{(const unsigned char*)code.data() + desc, (const unsigned char*)code.data() + desc + descsz};
        r.value(meta);
        return r.ok && meta.kind == MpVal::MAP;
      }
      q = desc + ((descsz + 3u) & ~3u);
    }
  }
  return false;
}

int code_resources(const std::vector<char>& code, std::vector<mf_kernel_resources>& out) {
  MpVal meta;
  if (!code_metadata(code, meta)) return set_error(MF_ERR_COMPILE, "code object: AMDGPU metadata note not found or malformed");
  const MpVal* ks = meta.get("amdhsa.kernels");
  if (!ks || ks->kind != MpVal::ARR) return set_error(MF_ERR_COMPILE, "code object: no amdhsa.kernels in metadata");
  out.clear();
  for (const MpVal& k : ks->items) {
    mf_kernel_resources r{};
    auto num = [&](const char* key) {
      const MpVal* v = k.get(key);
      return v && v->kind == MpVal::INT ? (int)v->i : 0;
    };
    const MpVal* nm = k.get(".name");
    if (nm && nm->kind == MpVal::STR) snprintf(r.name, sizeof(r.name), "%s", nm->s.c_str());
    r.vgpr_count = num(".vgpr_count");
    r.agpr_count = num(".agpr_count");
    r.sgpr_count = num(".sgpr_count");
    r.vgpr_spill_count = num(".vgpr_spill_count");
    r.sgpr_spill_count = num(".sgpr_spill_count");
    r.private_segment_bytes = num(".private_segment_fixed_size");
    r.group_segment_bytes = num(".group_segment_fixed_size");
    out.push_back(r);
  }
  return MF_OK;
}

// built-in kernels (compiled into this library for gfx950)
// Per-chunk partials of mf_launch_eval_noalloc: one array per GPU, for the life of the process.
// Host code never uses the name itself (on the host it is only a stand-in for the symbol):
// mf_device_open takes each device's address of it with hipGetSymbolAddress, and launches
// pass that address (mf_device::scratch).
static __device__ __align__(256) double scratch[MF_NOALLOC_SCRATCH_BYTES / sizeof(double)];

__global__ void k_fill_zero(unsigned int* p, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = 0u;
}
__global__ void k_fill_zero_bytes(unsigned char* p, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}
// generator stream states by jump-ahead: state = J^k base, J^(2^i) given as column sets
__global__ void __launch_bounds__(256) k_seed_streams(const unsigned int* __restrict__ powers, int levels,
                                                      const mf_gen_stream* __restrict__ streams, int nstreams, int eps,
                                                      const unsigned int* __restrict__ seg_base, unsigned int* __restrict__ states) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nstreams) return;
  const mf_gen_stream st = streams[g];
  unsigned long long k = (unsigned long long)(st.begin / eps);
  unsigned int s[4];
  for (int w = 0; w < 4; w++) s[w] = seg_base[4 * st.seg + w];
  for (int i = 0; i < levels && (k >> i); i++) {
    if (!((k >> i) & 1ull)) continue;
    const unsigned int* cols = powers + (size_t)i * 512;
    unsigned int a0 = 0, a1 = 0, a2 = 0, a3 = 0;
    for (int b = 0; b < 128; b++) {
      const unsigned int m = 0u - ((s[b >> 5] >> (b & 31)) & 1u);
      a0 ^= cols[4 * b] & m;
      a1 ^= cols[4 * b + 1] & m;
      a2 ^= cols[4 * b + 2] & m;
      a3 ^= cols[4 * b + 3] & m;
    }
    s[0] = a0;
    s[1] = a1;
    s[2] = a2;
    s[3] = a3;
  }
  for (int w = 0; w < 4; w++) states[4 * (size_t)g + w] = s[w];
}

// generator stream table from per-slot counts: stream g belongs to the slot l whose range
// [first[l], first[l + 1]) holds g (binary search over the nslots + 1 prefix sums); it is the
// (g - first[l])-th stream of that slot, events [j * eps, min(n, (j + 1) * eps))
__global__ void __launch_bounds__(256) k_expand_streams(const long long* __restrict__ first, const int* __restrict__ slot,
                                                        const long long* __restrict__ n, int nslots, int eps,
                                                        mf_gen_stream* __restrict__ out, int nstreams) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nstreams) return;
  int lo = 0, hi = nslots;  // invariant: first[lo] <= g < first[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (first[mid] <= g) lo = mid;
    else hi = mid;
  }
  const long long j = g - first[lo];
  const long long b = j * eps, e = b + eps < n[lo] ? b + eps : n[lo];
  out[g] = mf_gen_stream{slot[lo], (int)b, (int)e, 0};
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
  double* scratch = nullptr;  // this device's address of `scratch` (mf_launch_eval_noalloc)
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
  std::map<std::string, int> nacc;  // evaluation kernel -> FP64 partials per chunk (its <name>_nacc)
  int nmom = 0;                      // data moments per segment its outputs read (mf_nmom)
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
  // the device's own copy of the no-allocation scratch array, with this device current
  {
    void* sp = nullptr;
    hipError_t e = hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !sp) {
      delete d;
      return set_error(MF_ERR_HIP, "mf_device_open: hipGetSymbolAddress(scratch) on device %d: %s", ordinal, hipGetErrorString(e));
    }
    d->scratch = (double*)sp;
  }
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

int mf_host_alloc_mapped(mf_device* dev, size_t bytes, void** hptr, void** dptr) {
  MF_INJECT("mf_host_alloc_mapped");
  if (!dev || !hptr || !dptr) return set_error(MF_ERR_INVALID, "mf_host_alloc_mapped: null argument");
  *hptr = *dptr = nullptr;
  if (bytes == 0) return MF_OK;
  MF_HIP(hipSetDevice(dev->ordinal), "mf_host_alloc_mapped");
  hipError_t e = hipHostMalloc(hptr, bytes, hipHostMallocMapped | hipHostMallocCoherent);
  if (e != hipSuccess) return set_error(MF_ERR_NOMEM, "mf_host_alloc_mapped(%zu): %s", bytes, hipGetErrorString(e));
  e = hipHostGetDevicePointer(dptr, *hptr, 0);
  if (e != hipSuccess) {
    (void)hipHostFree(*hptr);
    *hptr = nullptr;
    return set_error(MF_ERR_HIP, "mf_host_alloc_mapped: hipHostGetDevicePointer: %s", hipGetErrorString(e));
  }
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

int mf_query(mf_device* dev, int* idle) {
  if (!dev || !idle) return set_error(MF_ERR_INVALID, "mf_query: null argument");
  const hipError_t e = hipStreamQuery(dev->stream);
  if (e == hipErrorNotReady) {
    *idle = 0;
    return MF_OK;
  }
  MF_HIP(e, "mf_query");
  *idle = 1;
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

int mf_module_resources(const char* generated_src, const char* options, mf_kernel_resources* out, int cap, int* count) {
  MF_INJECT("mf_module_resources");
  if (!generated_src || !count || (cap > 0 && !out)) return set_error(MF_ERR_INVALID, "mf_module_resources: null argument");
  std::vector<char> code;
  int rc = get_code(generated_src, options, code, nullptr, nullptr);
  if (rc != MF_OK) return rc;
  std::vector<mf_kernel_resources> res;
  rc = code_resources(code, res);
  if (rc != MF_OK) return rc;
  *count = (int)res.size();
  for (int k = 0; k < cap && k < (int)res.size(); k++) out[k] = res[k];
  return MF_OK;
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
  std::vector<mf_kernel_resources> res;
  rc = code_resources(code, res);
  if (rc != MF_OK) return rc;
  for (const auto& r : res)
    if (r.agpr_count > 0)
      return set_error(MF_ERR_COMPILE,
                       "mf_module_load: kernel %s allocates %d AGPRs (%d VGPRs); modules whose kernels need more than the "
                       "256 architectural VGPRs are refused -- lower that kernel's unroll",
                       r.name, r.agpr_count, r.vgpr_count);
  MF_HIP(hipSetDevice(dev->ordinal), "mf_module_load");
  hipModule_t hm;
  MF_HIP(hipModuleLoadData(&hm, code.data()), "mf_module_load");
  auto* m = new mf_module();
  m->dev = dev;
  m->mod = hm;
  m->hash = h;
  m->refs = 1;
  m->compile_ms = ms;
  // partials per chunk of each evaluation kernel (the module's <kernel>_nacc integers)
  for (const auto& r : res) {
    const std::string sym = std::string(r.name) + "_nacc";
    hipDeviceptr_t dp = nullptr;
    size_t bytes = 0;
    if (hipModuleGetGlobal(&dp, &bytes, hm, sym.c_str()) != hipSuccess || bytes != sizeof(int)) {
      (void)hipGetLastError();
      continue;
    }
    int v = 0;
    if (hipMemcpyDtoH(&v, dp, sizeof(int)) != hipSuccess) {
      (void)hipModuleUnload(hm);
      delete m;
      return set_error(MF_ERR_HIP, "mf_module_load: reading %s failed", sym.c_str());
    }
    m->nacc[r.name] = v;
  }
  {
    hipDeviceptr_t dp = nullptr;
    size_t bytes = 0;
    if (hipModuleGetGlobal(&dp, &bytes, hm, "mf_nmom") == hipSuccess && bytes == sizeof(int)) {
      if (hipMemcpyDtoH(&m->nmom, dp, sizeof(int)) != hipSuccess) {
        (void)hipModuleUnload(hm);
        delete m;
        return set_error(MF_ERR_HIP, "mf_module_load: reading mf_nmom failed");
      }
    } else {
      (void)hipGetLastError();
    }
  }
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

int mf_module_occupancy(mf_module* mod, const char* kernel, int block, int* blocks_per_cu) {
  if (!mod || !kernel || !blocks_per_cu || block <= 0) return set_error(MF_ERR_INVALID, "mf_module_occupancy: bad argument");
  hipFunction_t fn;
  int rc = get_fn(mod, kernel, &fn);
  if (rc != MF_OK) return rc;
  int nb = 0;
  MF_HIP(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, block, 0), "mf_module_occupancy");
  *blocks_per_cu = nb;
  return MF_OK;
}

int mf_module_has_kernel(mf_module* mod, const char* kernel) {
  if (!mod || !kernel) return 0;
  hipFunction_t fn;
  return get_fn(mod, kernel, &fn) == MF_OK ? 1 : 0;
}

static int launch_eval(mf_module* mod, const char* kernel, const mf_eval_args* args, int grid, int block);

int mf_launch_eval(mf_module* mod, const char* kernel, const mf_eval_args* args, int grid, int block) {
  MF_INJECT("mf_launch_eval");
  return launch_eval(mod, kernel, args, grid, block);
}

int mf_launch_eval_noalloc(mf_module* mod, const char* kernel, const mf_eval_args* args, int grid, int block) {
  MF_INJECT("mf_launch_eval_noalloc");
  if (!mod || !kernel || !args) return set_error(MF_ERR_INVALID, "mf_launch_eval_noalloc: null argument");
  if (args->seg_done)
    return set_error(MF_ERR_INVALID, "mf_launch_eval_noalloc: in-kernel segment reduction (seg_done) is not supported; the "
                                     "final kernel reduces the segments");
  auto it = mod->nacc.find(kernel);
  if (it == mod->nacc.end())
    return set_error(MF_ERR_NOKERNEL, "mf_launch_eval_noalloc: %s is not an evaluation kernel of this module", kernel);
  const long long need = (long long)(args->nchunks > 0 ? args->nchunks : 0) * it->second * (long long)sizeof(double);
  if (need > (long long)MF_NOALLOC_SCRATCH_BYTES)
    return set_error(MF_ERR_INVALID,
                     "mf_launch_eval_noalloc: %d chunks x %d partials need %lld bytes of scratch; the maximum is %lld "
                     "(use fewer, larger chunks, or mf_launch_eval with a partials buffer)",
                     args->nchunks, it->second, need, (long long)MF_NOALLOC_SCRATCH_BYTES);
  mf_eval_args a = *args;
  a.partials = mod->dev->scratch;
  return launch_eval(mod, kernel, &a, grid, block);
}

int mf_launch_check(mf_module* mod, const char* kernel, const mf_eval_args* args, int grid, int block) {
  MF_INJECT("mf_launch_check");
  if (!mod || !kernel || !args) return set_error(MF_ERR_INVALID, "mf_launch_check: null argument");
  if (grid <= 0 || block <= 0 || block % 64 != 0 || block > 1024)
    return set_error(MF_ERR_INVALID, "mf_launch_check: bad launch shape %d x %d", grid, block);
  if (args->nchunks < 0 || args->nseg < 0 || !args->segs || !args->chunks || !args->params || !args->data || !args->store)
    return set_error(MF_ERR_INVALID, "mf_launch_check: incomplete argument block");
  if (args->nchunks == 0) return MF_OK;
  hipFunction_t fn;
  int rc = get_fn(mod, kernel, &fn);
  if (rc != MF_OK) return rc;
  mf_eval_args a = *args;
  void* kp[] = {(void*)&a};
  MF_HIP(hipModuleLaunchKernel(fn, grid, 1, 1, block, 1, 1, 0, mod->dev->stream, kp, nullptr), "mf_launch_check");
  return MF_OK;
}

static int launch_eval(mf_module* mod, const char* kernel, const mf_eval_args* args, int grid, int block) {
  if (!mod || !kernel || !args) return set_error(MF_ERR_INVALID, "mf_launch_eval: null argument");
  if (grid <= 0 || block <= 0 || block % 64 != 0 || block > 1024)
    return set_error(MF_ERR_INVALID, "mf_launch_eval: bad launch shape %d x %d", grid, block);
  if (args->nchunks < 0 || args->nseg < 0 || !args->segs || !args->chunks || !args->params || !args->data || !args->partials)
    return set_error(MF_ERR_INVALID, "mf_launch_eval: incomplete argument block");
  if (mod->nmom > 0 && !args->moments && mod->nacc.count(kernel) && std::strcmp(kernel, "mf_eval_mom") != 0)
    return set_error(MF_ERR_INVALID,
                     "mf_launch_eval: the outputs of %s read the segments' data moments (mf_eval_args.moments, %d per segment); "
                     "compute them with mf_eval_mom first",
                     kernel, mod->nmom);
  std::string fin = std::string(kernel) + "_final";
  hipFunction_t ff = nullptr;
  const bool final = args->nseg > 0 && !args->seg_done && !(args->flags & MF_EVAL_SKIP_FINAL) && mod_has(mod, fin.c_str(), &ff);
  if (args->done_flag && (args->nseg != 1 || args->nchunks == 0 || !(final || (args->nchunks == 1 && grid == 1))))
    return set_error(MF_ERR_INVALID,
                     "mf_launch_eval: a done flag needs one segment, reduced by the final kernel or held in a single chunk");
  if (args->nchunks == 0) return MF_OK;
  hipFunction_t fn;
  int rc = get_fn(mod, kernel, &fn);
  if (rc != MF_OK) return rc;
  mf_eval_args a = *args;
  void* kp[] = {(void*)&a};
  MF_HIP(hipModuleLaunchKernel(fn, grid, 1, 1, block, 1, 1, 0, mod->dev->stream, kp, nullptr), "mf_launch_eval");
  if (final)
    MF_HIP(hipModuleLaunchKernel(ff, args->nseg, 1, 1, MF_FINAL_THREADS, 1, 1, 0, mod->dev->stream, kp, nullptr), "mf_launch_eval final");
  return MF_OK;
}

int mf_launch_prep(mf_module* mod, const double* theta, void* params, double* hoist64, const int* seg_list, int nseg,
                   const mf_segment* seg_src, mf_segment* seg_dst) {
  MF_INJECT("mf_launch_prep");
  if (!mod || !theta || !params || nseg < 0 || (seg_src && !seg_dst))
    return set_error(MF_ERR_INVALID, "mf_launch_prep: bad argument");
  if (nseg == 0) return MF_OK;
  hipFunction_t fn;
  int rc = get_fn(mod, "mf_prep", &fn);
  if (rc != MF_OK) return rc;
  void* kp[] = {(void*)&theta, (void*)&params, (void*)&hoist64, (void*)&seg_list, (void*)&nseg, (void*)&seg_src, (void*)&seg_dst};
  MF_HIP(hipModuleLaunchKernel(fn, MF_PREP_SPLIT, (nseg + 63) / 64, 1, 64, 1, 1, 0, mod->dev->stream, kp, nullptr), "mf_launch_prep");
  return MF_OK;
}

int mf_launch_prep_inline(mf_module* mod, const double* theta, int ntheta, int nseg, void* params, double* hoist64,
                          const mf_segment* seg_src, mf_segment* seg_dst) {
  MF_INJECT("mf_launch_prep_inline");
  if (!mod || (!theta && ntheta > 0) || !params || nseg < 0 || ntheta < 0 || (seg_src && !seg_dst))
    return set_error(MF_ERR_INVALID, "mf_launch_prep_inline: bad argument");
  if ((long long)nseg * ntheta > MF_PREP_INLINE_THETA)
    return set_error(MF_ERR_INVALID, "mf_launch_prep_inline: %d x %d parameters exceed the %d passed by value", nseg, ntheta,
                     MF_PREP_INLINE_THETA);
  if (nseg == 0) return MF_OK;
  hipFunction_t fn;
  int rc = get_fn(mod, "mf_prep_inline", &fn);
  if (rc != MF_OK) return rc;
  mf_prep_inline_args g{};
  std::memcpy(g.theta, theta, sizeof(double) * (size_t)nseg * ntheta);
  g.params = params;
  g.hoist64 = hoist64;
  g.seg_src = seg_src;
  g.seg_dst = seg_dst;
  g.nseg = nseg;
  void* kp[] = {(void*)&g};
  MF_HIP(hipModuleLaunchKernel(fn, MF_PREP_SPLIT, (nseg + 63) / 64, 1, 64, 1, 1, 0, mod->dev->stream, kp, nullptr), "mf_launch_prep_inline");
  return MF_OK;
}

int mf_launch_seed_streams(mf_device* dev, const unsigned int* powers, int levels, const mf_gen_stream* streams, int nstreams,
                           int events_per_stream, const unsigned int* seg_base, unsigned int* states) {
  MF_INJECT("mf_launch_seed_streams");
  if (!dev || nstreams < 0 || events_per_stream <= 0 || levels < 0 || levels > 63 ||
      (nstreams > 0 && (!powers || !streams || !seg_base || !states)))
    return set_error(MF_ERR_INVALID, "mf_launch_seed_streams: bad argument");
  if (nstreams == 0) return MF_OK;
  MF_HIP(hipSetDevice(dev->ordinal), "mf_launch_seed_streams");
  hipLaunchKernelGGL(k_seed_streams, dim3((nstreams + 255) / 256), dim3(256), 0, dev->stream, powers, levels, streams, nstreams,
                     events_per_stream, seg_base, states);
  MF_HIP(hipGetLastError(), "mf_launch_seed_streams");
  return MF_OK;
}

int mf_launch_expand_streams(mf_device* dev, const long long* first, const int* slot, const long long* n, int nslots,
                             int events_per_stream, mf_gen_stream* streams, int nstreams) {
  MF_INJECT("mf_launch_expand_streams");
  if (!dev || nslots < 0 || nstreams < 0 || events_per_stream <= 0 ||
      (nstreams > 0 && (nslots == 0 || !first || !slot || !n || !streams)))
    return set_error(MF_ERR_INVALID, "mf_launch_expand_streams: bad argument");
  if (nstreams == 0) return MF_OK;
  MF_HIP(hipSetDevice(dev->ordinal), "mf_launch_expand_streams");
  hipLaunchKernelGGL(k_expand_streams, dim3((nstreams + 255) / 256), dim3(256), 0, dev->stream, first, slot, n, nslots,
                     events_per_stream, streams, nstreams);
  MF_HIP(hipGetLastError(), "mf_launch_expand_streams");
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
