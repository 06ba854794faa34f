// CPU emulation of the HIP subset used by ggcat_amd/csrc (test harness only).
// Kernels run one workgroup at a time, each thread on its own std::thread,
// __syncthreads on a std::barrier and wavefront collectives on per-wave
// barriers.  Never linked into the product library.
#pragma once
#include <algorithm>
#include <atomic>
#include <barrier>
#include <chrono>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <thread>
#include <vector>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static
#define __align__(x) alignas(x)

struct dim3 {
    unsigned x = 1, y = 1, z = 1;
    dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct uint4 { unsigned x, y, z, w; };
inline uint4 make_uint4(unsigned a, unsigned b, unsigned c, unsigned d) { return uint4{a, b, c, d}; }
struct ulonglong2 { unsigned long long x, y; };

namespace emu {
extern thread_local dim3 threadIdx;
extern thread_local dim3 blockIdx;
extern dim3 blockDim, gridDim;
extern std::barrier<>* bar;
extern std::barrier<>* wbar[16];
extern std::atomic<unsigned long long> wmask[16];
extern unsigned long long wval[16];
extern uint8_t* lds;
}  // namespace emu
using emu::blockDim;
using emu::blockIdx;
using emu::gridDim;
using emu::threadIdx;

inline void __syncthreads() { emu::bar->arrive_and_wait(); }
inline void __threadfence() { std::atomic_thread_fence(std::memory_order_seq_cst); }
inline void __threadfence_block() { std::atomic_thread_fence(std::memory_order_seq_cst); }
template <class T> inline T atomicAdd(T* p, T v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
template <class T> inline T atomicOr(T* p, T v) { return __atomic_fetch_or(p, v, __ATOMIC_SEQ_CST); }
template <class T> inline T atomicMax(T* p, T v) {
    T cur = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (cur < v && !__atomic_compare_exchange_n(p, &cur, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
    }
    return cur;
}
template <class T> inline T atomicMin(T* p, T v) {
    T cur = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (cur > v && !__atomic_compare_exchange_n(p, &cur, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
    }
    return cur;
}
template <class T> inline T atomicExch(T* p, T v) { return __atomic_exchange_n(p, v, __ATOMIC_SEQ_CST); }
template <class T> inline T atomicCAS(T* p, T c, T v) {
    __atomic_compare_exchange_n(p, &c, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST);
    return c;
}
#define __HIP_MEMORY_SCOPE_AGENT 0
#define __HIP_MEMORY_SCOPE_WORKGROUP 0
template <class T> inline T __hip_atomic_load(T* p, int, int) { return __atomic_load_n(p, __ATOMIC_SEQ_CST); }
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
inline int __popc(unsigned x) { return __builtin_popcount(x); }
inline int __ffs(unsigned x) { return __builtin_ffs((int)x); }
inline int __ffsll(unsigned long long x) { return __builtin_ffsll((long long)x); }
inline unsigned __brev(unsigned x) {
    x = ((x >> 1) & 0x55555555u) | ((x & 0x55555555u) << 1);
    x = ((x >> 2) & 0x33333333u) | ((x & 0x33333333u) << 2);
    x = ((x >> 4) & 0x0F0F0F0Fu) | ((x & 0x0F0F0F0Fu) << 4);
    x = ((x >> 8) & 0x00FF00FFu) | ((x & 0x00FF00FFu) << 8);
    return (x >> 16) | (x << 16);
}
inline int __clz(unsigned x) { return x ? __builtin_clz(x) : 32; }
inline unsigned long long __brevll(unsigned long long x) {
    return ((unsigned long long)__brev((unsigned)x) << 32) | __brev((unsigned)(x >> 32));
}
inline unsigned long long __ballot(int pred) {
    unsigned w = threadIdx.x / 64, l = threadIdx.x % 64;
    emu::wbar[w]->arrive_and_wait();
    if (pred) emu::wmask[w].fetch_or(1ull << l);
    emu::wbar[w]->arrive_and_wait();
    unsigned long long r = emu::wmask[w].load();
    emu::wbar[w]->arrive_and_wait();
    if (l == 0) emu::wmask[w].store(0);
    emu::wbar[w]->arrive_and_wait();
    return r;
}
template <class T> inline T __shfl(T v, int src) {
    unsigned w = threadIdx.x / 64, l = threadIdx.x % 64;
    emu::wbar[w]->arrive_and_wait();
    if (l == (unsigned)src) std::memcpy(&emu::wval[w], &v, sizeof(T));
    emu::wbar[w]->arrive_and_wait();
    T r;
    std::memcpy(&r, &emu::wval[w], sizeof(T));
    emu::wbar[w]->arrive_and_wait();
    return r;
}
// DPP move; only wave_shl:1 (0x130: each lane reads lane + 1, the last lane
// gets `old`) is emulated
inline int __builtin_amdgcn_update_dpp(int old, int src, int ctrl, int, int, bool) {
    unsigned w = threadIdx.x / 64, l = threadIdx.x % 64;
    static int slots[16][64];
    if (ctrl != 0x130 && ctrl != 0x138) std::abort();  // wave_shl:1 / wave_shr:1
    emu::wbar[w]->arrive_and_wait();
    slots[w][l] = src;
    emu::wbar[w]->arrive_and_wait();
    const int r = ctrl == 0x130 ? ((l + 1 < 64) ? slots[w][l + 1] : old) : (l >= 1 ? slots[w][l - 1] : old);
    emu::wbar[w]->arrive_and_wait();
    return r;
}
template <class T> inline T __shfl_down(T v, unsigned d) {
    unsigned w = threadIdx.x / 64, l = threadIdx.x % 64;
    static unsigned long long slots[16][64];
    emu::wbar[w]->arrive_and_wait();
    std::memcpy(&slots[w][l], &v, sizeof(T));
    emu::wbar[w]->arrive_and_wait();
    T r = v;
    if (l + d < 64) std::memcpy(&r, &slots[w][l + d], sizeof(T));
    emu::wbar[w]->arrive_and_wait();
    return r;
}
template <class T> inline T __shfl_up(T v, unsigned d) {
    static thread_local std::vector<unsigned long long>* dummy = nullptr;
    (void)dummy;
    unsigned w = threadIdx.x / 64, l = threadIdx.x % 64;
    static unsigned long long slots[16][64];
    emu::wbar[w]->arrive_and_wait();
    std::memcpy(&slots[w][l], &v, sizeof(T));
    emu::wbar[w]->arrive_and_wait();
    T r = v;
    if (l >= d) std::memcpy(&r, &slots[w][l - d], sizeof(T));
    emu::wbar[w]->arrive_and_wait();
    return r;
}

// ------------------------------------------------------------- runtime ---
typedef int hipError_t;
constexpr hipError_t hipSuccess = 0;
typedef void* hipStream_t;
typedef struct { std::chrono::steady_clock::time_point t; }* hipEvent_t;
enum hipMemcpyKind { hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice };
constexpr unsigned hipStreamNonBlocking = 1;
struct hipDeviceProp_t { int multiProcessorCount = 4; };
inline hipError_t hipSetDevice(int) { return hipSuccess; }
inline hipError_t hipGetLastError() { return hipSuccess; }
inline const char* hipGetErrorString(hipError_t) { return "emulation error"; }
inline hipError_t hipGetDeviceProperties(hipDeviceProp_t* p, int) { p->multiProcessorCount = 4; return hipSuccess; }
inline hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) { *s = (void*)1; return hipSuccess; }
inline hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipEventCreate(hipEvent_t* e) { *e = new std::remove_pointer_t<hipEvent_t>(); return hipSuccess; }
inline hipError_t hipEventDestroy(hipEvent_t e) { delete e; return hipSuccess; }
inline hipError_t hipEventRecord(hipEvent_t e, hipStream_t) { e->t = std::chrono::steady_clock::now(); return hipSuccess; }
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
// every call runs to completion in order, so a stream never has to wait
inline hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
// one resident workgroup per CU (workgroups run one at a time here)
inline hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int* n, const void*, int, size_t) {
    *n = 1;
    return hipSuccess;
}
constexpr hipError_t hipErrorNotReady = 600;
constexpr hipError_t hipErrorInvalidSymbol = 13;
// device symbols are host variables here: their address is the symbol's own
#define HIP_SYMBOL(x) (&(x))
template <class T> inline hipError_t hipGetSymbolAddress(void** p, T* sym) { *p = (void*)sym; return hipSuccess; }
// HIP graphs are not emulated: capture fails, so callers keep direct launches.
typedef void* hipGraph_t;
typedef void* hipGraphExec_t;
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal, hipStreamCaptureModeThreadLocal, hipStreamCaptureModeRelaxed };
inline hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) { return 801; }
inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t*) { return 801; }
inline hipError_t hipGraphInstantiate(hipGraphExec_t*, hipGraph_t, void*, void*, size_t) { return 801; }
inline hipError_t hipGraphDestroy(hipGraph_t) { return hipSuccess; }
inline hipError_t hipGraphExecDestroy(hipGraphExec_t) { return hipSuccess; }
inline hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return 801; }
constexpr unsigned hipEventDisableTiming = 2;
inline hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
inline hipError_t hipEventQuery(hipEvent_t) { return hipSuccess; }
struct hipFuncAttributes {
    size_t sharedSizeBytes;
};
inline hipError_t hipFuncGetAttributes(hipFuncAttributes* a, const void*) {
    a->sharedSizeBytes = 0;
    return hipSuccess;
}
inline hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b) {
    *ms = std::chrono::duration<float, std::milli>(b->t - a->t).count();
    return hipSuccess;
}
inline hipError_t hipMalloc(void** p, size_t n) { *p = std::calloc(1, n + 64); return *p ? hipSuccess : 1; }
template <class T>
inline hipError_t hipMalloc(T** p, size_t n) { return hipMalloc(reinterpret_cast<void**>(p), n); }
constexpr unsigned hipHostMallocDefault = 0;
template <class T>
inline hipError_t hipHostMalloc(T** p, size_t n, unsigned) { return hipMalloc(reinterpret_cast<void**>(p), n); }
inline hipError_t hipHostFree(void* p) { std::free(p); return hipSuccess; }
inline hipError_t hipFree(void* p) { std::free(p); return hipSuccess; }
inline hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) { if (n) std::memcpy(d, s, n); return hipSuccess; }
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind k, hipStream_t) { return hipMemcpy(d, s, n, k); }
inline hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t) { std::memset(d, v, n); return hipSuccess; }

namespace emu {
// Runs the workgroups of the grid in turn with one team of threads (one
// std::thread per workitem); the team meets on a barrier between workgroups.
template <class F>
void launch(dim3 grid, dim3 block, size_t shmem, F&& body) {
    static std::vector<uint8_t> shm;
    if (shm.size() < shmem + 256) shm.assign(shmem + 256, 0);
    lds = shm.data();
    gridDim = grid;
    blockDim = block;
    const unsigned nt = block.x;
    std::barrier<> wg(nt);
    bar = &wg;
    std::vector<std::barrier<>*> wb;
    for (unsigned w = 0; w < (nt + 63) / 64; w++) {
        wb.push_back(new std::barrier<>(std::min(64u, nt - w * 64)));
        wbar[w] = wb.back();
    }
    std::vector<std::thread> th;
    th.reserve(nt);
    for (unsigned t = 0; t < nt; t++)
        th.emplace_back([&, t] {
            threadIdx.x = t;
            for (unsigned b = 0; b < grid.x; b++) {
                blockIdx.x = b;
                body();
                wg.arrive_and_wait();
            }
        });
    for (auto& x : th) x.join();
    for (auto* p : wb) delete p;
}
}  // namespace emu
#define hipLaunchKernelGGL(K, G, B, S, STR, ...) emu::launch(dim3(G), dim3(B), (S), [&] { K(__VA_ARGS__); })
#define HIP_EMU_DEFINE_GLOBALS                                      \
    namespace emu {                                                 \
    thread_local dim3 threadIdx;                                    \
    thread_local dim3 blockIdx;                                     \
    dim3 blockDim, gridDim;                                         \
    std::barrier<>* bar;                                            \
    std::barrier<>* wbar[16];                                       \
    std::atomic<unsigned long long> wmask[16];                      \
    unsigned long long wval[16];                                    \
    uint8_t* lds;                                                   \
    }
