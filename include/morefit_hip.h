/*
 * morefit_hip.h -- C ABI of libmorefit_hip.so, the MI355X (gfx950) compute backend of the
 * MoreFit likelihood hot path.
 *
 * This is the layer the reference's ComputeBlock implementations talk to their device
 * through (reference compute.hh:110-149; OpenCL calls in compute_opencl.hh:166-995).
 * Plain C: opaque handles, plain pointers and sizes, int status (0 = MF_OK) plus
 * mf_last_error().  Device memory is owned by the library (mf_malloc/mf_free), host memory
 * by the caller.  One device handle owns one HIP stream (default priority); everything
 * enqueued through a handle is ordered on that stream.  A handle is not thread-safe.
 *
 * Kernels are generated per density by the C++ host (graph.hh emit_kernel) and compiled
 * here with hipRTC for gfx950; mf_module_load prepends the kernel runtime (the structs
 * below, the fused per-event accumulation, the wave64/LDS block reduction and the
 * deterministic per-segment final reduction) to the generated source.
 *
 * Entry points and the reference interface each one replaces:
 *   mf_device_open/close            OpenCLBackend ctor: platform/device/context/queue (compute_opencl.hh:30-206)
 *   mf_malloc/mf_free               OpenCLBlock::Setup*Buffer / dtor (compute_opencl.hh:258-350)
 *   mf_host_alloc_mapped            no counterpart (zero-copy staging of parameters and totals)
 *   mf_memcpy_h2d                   CopyToInputBuffer / CopyToParameterBuffer / CopyToSeedBuffer (compute_opencl.hh:351-400)
 *   mf_memcpy_d2h                   CopyFromOutputBuffer / CopyFromKahanBuffer (compute_opencl.hh:401-424, 977-986)
 *   mf_module_load                  MakeComputeKernel / MakeGenerateKernel build step (compute_opencl.hh:425-804)
 *   mf_module_resources             no counterpart (register budget of each compiled kernel)
 *   mf_module_occupancy             no counterpart (resident work-groups per CU, for balanced chunk tables)
 *   mf_launch_eval                  SubmitKernel: per-event kernel + kahan_<name> passes (compute_opencl.hh:880-995)
 *   mf_launch_eval_noalloc          no counterpart: mf_launch_eval with the library's own static scratch for the
 *                                   partials, for callers that cannot allocate
 *   mf_launch_check                 the NaN check of the host sum (fitter.hh:1197-1204), as a per-event device pass
 *   mf_launch_prep                  host eval() of the hoisted parameter terms (fitter.hh:1063-1072)
 *   mf_launch_prep_inline           the same with theta passed by value (no counterpart: latency of one call)
 *   mf_launch_generate              SubmitKernel of gen_kernel (compute_opencl.hh:540-575)
 *   mf_launch_seed_streams          host seeding of per-work-item generator states (compute_opencl.hh:522-539), on the device
 *   mf_launch_expand_streams        no counterpart (the per-stream event ranges built on the device from per-slot counts)
 *   mf_sync                         Finish (clFinish)
 *   mf_query                        no counterpart (a non-blocking stream check that lets the runtime retire finished launches)
 *   mf_graph_*                      no counterpart (per-iteration launch sequence as a hipGraph)
 *   mf_timer_*, mf_bench_copy       no counterpart (measurement)
 */
#ifndef MOREFIT_HIP_H
#define MOREFIT_HIP_H

#ifndef MF_KERNEL_SIDE
#include <stddef.h>
#include <stdint.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- status codes ---------------- */
#define MF_OK 0
#define MF_ERR_INVALID 1     /* bad argument */
#define MF_ERR_HIP 2         /* HIP runtime error */
#define MF_ERR_COMPILE 3     /* hipRTC compilation failed (log in mf_last_error) */
#define MF_ERR_NOMEM 4       /* allocation failed or over budget */
#define MF_ERR_NOKERNEL 5    /* kernel not found in module */
#define MF_ERR_INJECTED 6    /* failure injected by mf_debug_fail_next */

/* ---------------- structures shared with the device (same layout on both sides) ----------------
 * A segment is one dataset: the single fit's events, one toy, or one served request.
 * Its observables are columns of `stride` elements starting at data_off.  The segment's
 * kernel parameters (floating parameters, then hoisted terms, in the reference's
 * parameter-buffer order, fitter.hh:1063-1072) start at params[param_off].  Its events
 * are split into chunks [begin, end) listed contiguously at chunks[chunk_begin ..
 * chunk_end); each chunk's sums go to partials[chunk * nout ...] and the segment's totals
 * (FP64) to results[out_off ...].  mode 0 skips the segment in a launch.  mf_launch_eval runs
 * the per-chunk kernel <kernel> and then, when the module has it, <kernel>_final (one
 * 256-thread work-group per segment, fixed-order reduction of the partials of segments with
 * more than one chunk; a one-chunk segment's totals come from the per-chunk kernel). */
/* Event indices inside a segment are 32-bit (mf_segment.n, mf_chunk.begin/end, the kernels'
 * event index, which runs up to end + unroll x block): a segment holds at most
 * MF_MAX_SEGMENT_EVENTS events (8.6 GB of FP32 mass events).  Hosts refuse larger segments;
 * a larger dataset is split into several segments whose totals add up. */
#define MF_MAX_SEGMENT_EVENTS 2146435072LL /* 2^31 - 2^20 */

typedef struct mf_segment {
  long long data_off;
  long long stride;
  int n;
  int param_off;
  int chunk_begin;
  int chunk_end;
  int out_off;
  int mode;
} mf_segment;

typedef struct mf_chunk {
  int seg;
  int begin;
  int end;
  int pad;
} mf_chunk;

/* argument block of every generated evaluation kernel (mf_eval_<kind>) */
/* capacity of the parameters passed by value in mf_eval_args: 128 32-bit words (128 FP32 or 64
 * FP64 kernel parameters) and 64 FP64 hoisted terms; a model with more takes the prep kernel.
 * The size of the argument block (1.1 KB) was measured: a one-fit call of the N = 10^7 FP32
 * gradient (two launches) took 23.6 us with it, 27.8 and 28.0 us with 0.6 and 0.9 KB blocks,
 * 24.2-24.9 us with 2.1 KB (and 23.3 us with the 0.1 KB block before the parameters moved in) */
#define MF_INLINE_PARAM_WORDS 128
#define MF_INLINE_HOIST 64

typedef struct mf_eval_args {
  const void* data;           /* T* SoA event columns */
  const mf_segment* segs;
  const mf_chunk* chunks;
  int nchunks;
  int nseg;
  const void* params;         /* T* kernel parameters */
  double* partials;           /* [nchunks][nout] */
  double* results;            /* [..][nout] per segment at out_off */
  const double* hoist64;      /* FP64 hoisted terms [segment][nhoist] from mf_prep: totals of outputs
                                 with a scale (functor scale(k) >= 0) are multiplied by
                                 hoist64[seg * nhoist + scale(k)]; NULL: no scaling */
  void* store;                /* T* per-event outputs (store kernels only) or NULL */
  long long store_stride;     /* column stride of `store` */
  int flags;                  /* MF_EVAL_* bits */
  int pad;
  unsigned int* seg_done;     /* [nseg] chunk counters, zero between launches (the kernel that
                                 completes a segment sets its counter back to zero).  Given, the
                                 work-group that finishes a segment's last chunk reduces its totals
                                 and no final kernel is launched; NULL: the `_final` kernel does. */
  int* done_flag;             /* mapped host memory or NULL: once every total is written, the final
                                 kernel stores done_value there (after a system-scope fence), so the
                                 host can poll it instead of waiting on the stream.  Only with one
                                 segment whose totals the final kernel reduces (nseg == 1, neither
                                 MF_EVAL_SKIP_FINAL nor seg_done), or with one segment of a single
                                 chunk launched as one work-group (the per-chunk kernel signals). */
  int done_value;
  int inline_params;          /* 1: one-segment call whose kernel parameters are pinline (MF_T values,
                                 [theta, hoisted]) and whose FP64 hoisted terms are hinline, read from
                                 this argument block (kernarg segment) instead of params / hoist64 */
  unsigned int pinline[MF_INLINE_PARAM_WORDS];
  double hinline[MF_INLINE_HOIST];
  const double* moments;      /* [nseg][MF_NMOM] data moments of each segment (sums of observable-only
                                 terms, computed once per dataset by the module's mf_eval_mom) that the
                                 outputs' combination reads, or NULL when the module has none */
} mf_eval_args;

/* mf_eval_args.flags: launch only the per-chunk kernel.  Correct when no active segment has
 * more than one chunk (the per-chunk kernel then writes the totals itself); otherwise the
 * totals are left unreduced (used to time the per-chunk kernel alone). */
#define MF_EVAL_SKIP_FINAL 1

/* argument block of mf_prep_inline (mf_launch_prep_inline) */
#define MF_PREP_INLINE_THETA 64
typedef struct mf_prep_inline_args {
  double theta[MF_PREP_INLINE_THETA]; /* [segment][ntheta] */
  void* params;
  double* hoist64;
  const mf_segment* seg_src;          /* NULL: the device segment table is left as it is */
  mf_segment* seg_dst;
  int nseg;
  int pad;
} mf_prep_inline_args;

/* prep kernels (mf_prep, mf_prep_inline): work-groups of 64 segments along y, and MF_PREP_SPLIT
 * along x that share a segment's hoisted terms (term k in work-group k % MF_PREP_SPLIT) */
#define MF_PREP_SPLIT 16

/* work-group size of the `_final` kernels (one work-group per segment) */
#define MF_FINAL_THREADS 1024

/* argument block of the generated generator kernel (mf_generate):
 * stream k owns events [begin, end) of segment seg and the xoshiro128++ state
 * states[4k..4k+3], advanced in place. */
typedef struct mf_gen_stream {
  int seg;
  int begin;
  int end;
  int pad;
} mf_gen_stream;

typedef struct mf_gen_args {
  void* data;                 /* T* SoA output columns (segment layout as above) */
  const mf_segment* segs;
  const mf_gen_stream* streams;
  int nstreams;
  int pad;
  unsigned int* states;
  const void* params;         /* T* kernel parameters of the likelihood kernel */
  const double* maxprob;      /* per segment */
  unsigned long long* proposals; /* per segment proposal counters (may be NULL) */
} mf_gen_args;

#ifndef MF_KERNEL_SIDE

typedef struct mf_device mf_device;
typedef struct mf_module mf_module;
typedef struct mf_graph mf_graph;

typedef struct mf_device_props {
  char name[64];
  char arch[32];
  int compute_units;
  int wavefront_size;
  size_t total_mem;
  size_t free_mem;
  int clock_khz;
  int l2_bytes;
} mf_device_props;

const char* mf_last_error(void);

/* device */
int mf_device_open(int ordinal, mf_device** out);
int mf_device_close(mf_device* dev);
int mf_device_query(mf_device* dev, mf_device_props* props);

/* memory (device memory is library-owned; host memory from mf_host_alloc is pinned) */
int mf_malloc(mf_device* dev, size_t bytes, void** dptr);
int mf_free(mf_device* dev, void* dptr);
int mf_host_alloc(mf_device* dev, size_t bytes, void** hptr);
/* coherent pinned host memory that kernels read and write in place: *hptr for the host, *dptr
 * for kernel arguments (staged parameters in, totals out, with no copy in between); freed
 * with mf_host_free */
int mf_host_alloc_mapped(mf_device* dev, size_t bytes, void** hptr, void** dptr);
int mf_host_free(mf_device* dev, void* hptr);
int mf_memcpy_h2d(mf_device* dev, void* dst, const void* src, size_t bytes);   /* async, stream-ordered */
int mf_memcpy_d2h(mf_device* dev, void* dst, const void* src, size_t bytes);   /* async, stream-ordered */
int mf_memcpy_d2d(mf_device* dev, void* dst, const void* src, size_t bytes);
int mf_fill_zero(mf_device* dev, void* dptr, size_t bytes);                    /* kernel fill, graph-safe */
int mf_sync(mf_device* dev);
/* *idle = 1 when all work queued on the device's stream has finished, else 0; does not wait.
 * Callers that learn of completion another way (a polled flag) call it now and then so that
 * the runtime retires its finished launches */
int mf_query(mf_device* dev, int* idle);

/* modules: generated source (+ the kernel runtime) compiled by hipRTC for gfx950 and cached
 * in memory and on disk by a hash of source and options.
 *
 * mf_module_load refuses (MF_ERR_COMPILE) a module in which any kernel allocates AGPRs.
 * When a kernel's register demand exceeds the 256 architectural VGPRs, the ROCm 7.2 hipRTC
 * compiler keeps loop-carried values in AGPRs and was found to place the VGPR->AGPR copies
 * inside a divergent branch (under a narrowed exec mask) while reloading them for every lane
 * at the loop latch, so lanes outside the branch read unset registers -- in the event loop
 * these were load addresses.  Hosts choose the unroll of each evaluation kernel with
 * mf_module_resources so that no kernel needs AGPRs. */
int mf_module_load(mf_device* dev, const char* generated_src, const char* options, mf_module** out);
int mf_module_compile_only(const char* generated_src, const char* options, size_t* code_bytes);

/* resources of one compiled kernel, read from the code object's AMDGPU metadata note */
typedef struct mf_kernel_resources {
  char name[64];
  int vgpr_count;             /* architectural VGPRs per lane */
  int agpr_count;             /* accumulation VGPRs per lane */
  int sgpr_count;
  int vgpr_spill_count;
  int sgpr_spill_count;
  int private_segment_bytes;  /* scratch per work-item */
  int group_segment_bytes;    /* LDS per work-group */
  int pad;
} mf_kernel_resources;

/* compile (or fetch from the cache) without a device and list the resources of every kernel
 * in the module: up to `cap` entries in out[], the number of kernels in *count */
int mf_module_resources(const char* generated_src, const char* options, mf_kernel_resources* out, int cap, int* count);
int mf_module_unload(mf_module* mod);
int mf_module_has_kernel(mf_module* mod, const char* kernel);
/* work-groups of `block` threads of `kernel` that one compute unit holds at once */
int mf_module_occupancy(mf_module* mod, const char* kernel, int block, int* blocks_per_cu);
double mf_module_compile_ms(mf_module* mod);  /* 0 when served from a cache */

/* launches (on the device stream) */
int mf_launch_eval(mf_module* mod, const char* kernel, const mf_eval_args* args, int grid, int block);

/* mf_launch_eval that allocates nothing: no counterpart in the reference.  For callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  args->partials is ignored: the per-chunk partials go to a static
 * array of MF_NOALLOC_SCRATCH_BYTES that the library holds on each GPU it runs on until the
 * process ends; a launch needing more (nchunks x the kernel's partials per chunk x 8 bytes) is
 * refused with MF_ERR_INVALID before anything is enqueued.  Nothing in the scratch has to start
 * at zero (each partial is written before it is read), so no clearing is enqueued; in-kernel
 * segment reduction (args->seg_done) is refused, the final kernel reduces.  All calls on one
 * GPU share its array: launches that overlap on one GPU (other streams, other threads) race
 * and must be serialised by the caller.  Inputs, outputs and tables are the caller's, as for
 * mf_launch_eval. */
#define MF_NOALLOC_SCRATCH_BYTES (4u << 20)
int mf_launch_eval_noalloc(mf_module* mod, const char* kernel, const mf_eval_args* args, int grid, int block);
/* a diagnostic kernel of the module (MF_CHECK_KERNEL: its functor's per-event terms over the
 * active segments of args): args->store points at unsigned int [args->nseg][3] that the caller
 * sets to {0, 0xffffffff, 0} per segment; the kernel adds the events whose terms are not all
 * finite, lowers the second word to the smallest such event index, and adds the events with an
 * observable outside the density's range (or not a number; modules generated without ranges
 * leave that word alone).  args->partials and args->results are not used. */
int mf_launch_check(mf_module* mod, const char* kernel, const mf_eval_args* args, int grid, int block);
/* kernel parameters of segments seg_list[0..nseg) (all 0..nseg-1 when seg_list is NULL) from
 * their FP64 theta: params[s * (ntheta + nhoist) ...] in the module's type and, when hoist64 is
 * set, the hoisted terms in FP64 at hoist64[s * nhoist ...].  With seg_src set, each of those
 * segments' table entries is also copied seg_src[s] -> seg_dst[s] (a mapped host table staged
 * into device memory in the same launch). */
int mf_launch_prep(mf_module* mod, const double* theta, void* params, double* hoist64, const int* seg_list, int nseg,
                   const mf_segment* seg_src, mf_segment* seg_dst);
/* the same for segments 0..nseg-1 with theta (host memory, nseg * ntheta values, ntheta the
 * module's floating parameters) copied into the kernel arguments at the call: the kernel reads
 * nothing from host memory.  nseg * ntheta <= MF_PREP_INLINE_THETA.  Not for graph capture
 * (a captured launch keeps the theta of the capture). */
int mf_launch_prep_inline(mf_module* mod, const double* theta, int ntheta, int nseg, void* params, double* hoist64,
                          const mf_segment* seg_src, mf_segment* seg_dst);
int mf_launch_generate(mf_module* mod, const mf_gen_args* args, int grid, int block);
/* xoshiro128++ states of generator streams, all device pointers: stream g (of nstreams)
 * gets J^k applied to seg_base[4 * streams[g].seg ...], k = streams[g].begin /
 * events_per_stream, J the jump() map; powers holds J^(2^i), i < levels, as 128 columns of 4
 * words each (the layout of random.hh xoshiro_jump_powers).  One thread per stream. */
int mf_launch_seed_streams(mf_device* dev, const unsigned int* powers, int levels, const mf_gen_stream* streams, int nstreams,
                           int events_per_stream, const unsigned int* seg_base, unsigned int* states);
/* the stream table of a generation on the device, from per-slot counts (all device pointers):
 * slot l (of nslots) owns streams [first[l], first[l + 1]) (first[nslots] == nstreams), its
 * j-th stream covering events [j * events_per_stream, min(n[l], (j + 1) * events_per_stream))
 * of segment slot[l].  One thread per stream. */
int mf_launch_expand_streams(mf_device* dev, const long long* first, const int* slot, const long long* n, int nslots,
                             int events_per_stream, mf_gen_stream* streams, int nstreams);

/* hipGraph capture of everything enqueued on the device stream between begin and end */
int mf_graph_begin(mf_device* dev);
int mf_graph_end(mf_device* dev, mf_graph** out);
int mf_graph_launch(mf_graph* g);
int mf_graph_destroy(mf_graph* g);

/* timing with HIP events on the device stream */
int mf_timer_begin(mf_device* dev);
int mf_timer_end(mf_device* dev, float* ms);

/* HBM copy ceiling: a streaming read+write kernel over `bytes`, best of `iters` (GB/s of read+write) */
int mf_bench_copy(mf_device* dev, size_t bytes, int iters, double* gbps);

/* failure injection for host-side error-path tests: the next `count` calls of the named
 * entry point ("mf_malloc", "mf_module_load", "mf_launch_eval", ...) fail with MF_ERR_INJECTED
 * before touching the device. */
int mf_debug_fail_next(const char* entry_point, int count);

#endif /* MF_KERNEL_SIDE */

#ifdef __cplusplus
}
#endif

#endif
