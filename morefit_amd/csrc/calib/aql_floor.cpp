// aql_floor.cpp -- calibration: the host round trip of kernel dispatches written as AQL packets
// straight into an HSA queue of our own (no HIP runtime in between), to compare with
// launch_floor's hipModuleLaunchKernel figures: one dispatch whose completion flag is polled, and
// two dependent dispatches (barrier bit) then the flag.  usage: aql_floor CODE_OBJECT (the raw
// gfx950 code object of aql_kernels.hip).
#include <hsa/hsa.h>
#include <hsa/hsa_ext_amd.h>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <vector>
#define HC(x)                                                        \
  do {                                                               \
    hsa_status_t st = (x);                                           \
    if (st != HSA_STATUS_SUCCESS) {                                  \
      const char* m = "";                                            \
      hsa_status_string(st, &m);                                     \
      std::printf("HSA error %s at line %d\n", m, __LINE__);         \
      return 1;                                                      \
    }                                                                \
  } while (0)

static hsa_agent_t gpu{0};
static hsa_region_t kernarg_region{0};
static hsa_status_t find_gpu(hsa_agent_t a, void*) {
  hsa_device_type_t t;
  hsa_agent_get_info(a, HSA_AGENT_INFO_DEVICE, &t);
  if (t == HSA_DEVICE_TYPE_GPU && gpu.handle == 0) gpu = a;
  return HSA_STATUS_SUCCESS;
}
static hsa_status_t find_kernarg(hsa_region_t r, void*) {
  hsa_region_segment_t seg;
  hsa_region_get_info(r, HSA_REGION_INFO_SEGMENT, &seg);
  if (seg != HSA_REGION_SEGMENT_GLOBAL) return HSA_STATUS_SUCCESS;
  uint32_t flags;
  hsa_region_get_info(r, HSA_REGION_INFO_GLOBAL_FLAGS, &flags);
  if ((flags & HSA_REGION_GLOBAL_FLAG_KERNARG) && kernarg_region.handle == 0) kernarg_region = r;
  return HSA_STATUS_SUCCESS;
}
struct Kern {
  uint64_t object;
  uint32_t kernarg_size, group_size, private_size;
};

int main(int argc, char** argv) {
  if (argc < 2) {
    std::printf("usage: %s CODE_OBJECT\n", argv[0]);
    return 2;
  }
  HC(hsa_init());
  HC(hsa_iterate_agents(find_gpu, nullptr));
  if (!gpu.handle) { std::printf("no GPU agent\n"); return 1; }
  HC(hsa_agent_iterate_regions(gpu, find_kernarg, nullptr));
  if (!kernarg_region.handle) { std::printf("no kernel-argument region\n"); return 1; }
  std::ifstream f(argv[1], std::ios::binary);
  std::vector<char> co((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
  if (co.empty()) {
    std::printf("cannot read %s\n", argv[1]);
    return 1;
  }
  hsa_code_object_reader_t rd;
  HC(hsa_code_object_reader_create_from_memory(co.data(), co.size(), &rd));
  hsa_executable_t exe;
  HC(hsa_executable_create_alt(HSA_PROFILE_FULL, HSA_DEFAULT_FLOAT_ROUNDING_MODE_DEFAULT, nullptr, &exe));
  HC(hsa_executable_load_agent_code_object(exe, gpu, rd, nullptr, nullptr));
  HC(hsa_executable_freeze(exe, nullptr));
  Kern kf{}, kn{};
  for (int i = 0; i < 2; i++) {
    hsa_executable_symbol_t sym;
    HC(hsa_executable_get_symbol_by_name(exe, i ? "k_nop.kd" : "k_flag.kd", &gpu, &sym));
    Kern& k = i ? kn : kf;
    HC(hsa_executable_symbol_get_info(sym, HSA_EXECUTABLE_SYMBOL_INFO_KERNEL_OBJECT, &k.object));
    HC(hsa_executable_symbol_get_info(sym, HSA_EXECUTABLE_SYMBOL_INFO_KERNEL_KERNARG_SEGMENT_SIZE, &k.kernarg_size));
    HC(hsa_executable_symbol_get_info(sym, HSA_EXECUTABLE_SYMBOL_INFO_KERNEL_GROUP_SEGMENT_SIZE, &k.group_size));
    HC(hsa_executable_symbol_get_info(sym, HSA_EXECUTABLE_SYMBOL_INFO_KERNEL_PRIVATE_SEGMENT_SIZE, &k.private_size));
    std::printf("kernel %d: kernarg %u group %u private %u\n", i, k.kernarg_size, k.group_size, k.private_size);
  }
  hsa_queue_t* q;
  HC(hsa_queue_create(gpu, 256, HSA_QUEUE_TYPE_SINGLE, nullptr, nullptr, UINT32_MAX, UINT32_MAX, &q));
  // kernarg buffers (one per packet slot in flight) and the flag, all in the kernarg region
  char* ka;
  HC(hsa_memory_allocate(kernarg_region, 2 * 512, (void**)&ka));
  int* flag;
  HC(hsa_memory_allocate(kernarg_region, 64, (void**)&flag));
  *flag = 0;
  std::memset(ka, 0, 2 * 512);
  auto submit = [&](const Kern& k, const void* args, size_t nargs, int slot) {
    char* kp = ka + 512 * slot;
    std::memcpy(kp, args, nargs);
    const uint64_t idx = hsa_queue_add_write_index_relaxed(q, 1);
    while (idx - hsa_queue_load_read_index_scacquire(q) >= q->size) {
    }
    hsa_kernel_dispatch_packet_t* p = (hsa_kernel_dispatch_packet_t*)q->base_address + (idx & (q->size - 1));
    std::memset(((char*)p) + 4, 0, sizeof(*p) - 4);
    p->workgroup_size_x = 64;
    p->workgroup_size_y = 1;
    p->workgroup_size_z = 1;
    p->grid_size_x = 64;
    p->grid_size_y = 1;
    p->grid_size_z = 1;
    p->private_segment_size = k.private_size;
    p->group_segment_size = k.group_size;
    p->kernel_object = k.object;
    p->kernarg_address = kp;
    p->completion_signal.handle = 0;
    const uint16_t header = (HSA_PACKET_TYPE_KERNEL_DISPATCH << HSA_PACKET_HEADER_TYPE) | (1 << HSA_PACKET_HEADER_BARRIER) |
                            (HSA_FENCE_SCOPE_SYSTEM << HSA_PACKET_HEADER_SCACQUIRE_FENCE_SCOPE) |
                            (HSA_FENCE_SCOPE_SYSTEM << HSA_PACKET_HEADER_SCRELEASE_FENCE_SCOPE);
    const uint16_t setup = 1 << HSA_KERNEL_DISPATCH_PACKET_SETUP_DIMENSIONS;
    __atomic_store_n((uint32_t*)p, (uint32_t)header | ((uint32_t)setup << 16), __ATOMIC_RELEASE);
    hsa_signal_store_screlease(q->doorbell_signal, idx);
  };
  auto poll = [&](int v) {
    for (long spin = 0; spin < 2000000000L; spin++)
      if (__atomic_load_n(flag, __ATOMIC_ACQUIRE) == v) return true;
    return false;
  };
  struct { int* f; int v; int pad; } fa{flag, 0, 0};
  struct { int* p; } na{nullptr};
  int v = 0;
  const int it = 3000;
  for (int pass = 0; pass < 2; pass++) {
    auto t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < it; i++) {
      fa.v = ++v;
      submit(kf, &fa, sizeof(fa), 0);
      if (!poll(v)) return std::printf("flag not set\n"), 1;
    }
    auto t1 = std::chrono::steady_clock::now();
    double a = std::chrono::duration<double, std::micro>(t1 - t0).count() / it;
    t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < it; i++) {
      fa.v = ++v;
      submit(kn, &na, sizeof(na), 1);
      submit(kf, &fa, sizeof(fa), 0);
      if (!poll(v)) return std::printf("flag not set\n"), 1;
    }
    t1 = std::chrono::steady_clock::now();
    double b = std::chrono::duration<double, std::micro>(t1 - t0).count() / it;
    std::printf("AQL: dispatch+poll %.2f us, two dispatches+poll %.2f us\n", a, b);
  }
  HC(hsa_queue_destroy(q));
  return 0;
}
