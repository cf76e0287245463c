// Host build of the library's kernels, for sanitizers: this header stands in
// for <hip/hip_runtime.h> when solver.hip is compiled as ordinary C++
// (clang++ -x c++ -I tests/hostemu/include), so that AddressSanitizer,
// UndefinedBehaviorSanitizer (array bounds of the kernels' private arrays
// included) and pattern-initialised locals can watch the kernel bodies run
// on the CPU.  Test infrastructure only; the library itself is built for
// gfx950 by hipcc.
//
// Execution model: every launch runs at once, one workgroup after another;
// the workgroup's threads are fibers that meet at __syncthreads and at each
// lane exchange (__shfl_xor, and the DPP quad permutations of
// __builtin_amdgcn_mov_dpp), so kernels run with the same synchronisation
// they have on the GPU; wave priority (__builtin_amdgcn_s_setprio) is a
// no-op.  A thread that returns takes no part in later
// barriers, as an exited wave leaves s_barrier.
// Streams are all the same in-order queue; graph capture records the
// launches and copies and a graph launch replays them; events read the host
// clock.  Device memory is host memory from aligned_alloc, so ASan sees every
// buffer's bounds.
#pragma once

#include <math.h>
#include <ucontext.h>
#if defined(__SANITIZE_ADDRESS__) || defined(__has_feature)
#if defined(__SANITIZE_ADDRESS__) || __has_feature(address_sanitizer)
#include <sanitizer/common_interface_defs.h>
#endif
#endif

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <type_traits>
#include <vector>

#define __global__
#define __device__
#define __host__
#define __shared__ static
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __align__(n) __attribute__((aligned(n)))
// launch-shape attributes of the gfx950 kernels mean nothing on the host
#define amdgpu_flat_work_group_size(...) unused
#define amdgpu_waves_per_eu(...) unused

// gfx950 instructions used by the kernels, as their exact host counterparts
#define __builtin_amdgcn_rsq(x) (1.0 / std::sqrt((double)(x)))
#define __builtin_amdgcn_rcp(x) (1.0 / (double)(x))
#define __builtin_amdgcn_exp2f(x) exp2f(x)
#define __builtin_amdgcn_logf(x) log2f(x)  // the instruction is log base 2

struct dim3 {
	unsigned x, y, z;
	constexpr dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct alignas(8) int2 {
	int x, y;
};
struct alignas(16) int4 {
	int x, y, z, w;
};
inline int4 make_int4(int x, int y, int z, int w) { return int4{x, y, z, w}; }
inline int2 make_int2(int x, int y) { return int2{x, y}; }

// the device library's integer min and max
template <typename A, typename B>
inline std::common_type_t<A, B> min(A a, B b)
{
	return b < a ? b : a;
}
template <typename A, typename B>
inline std::common_type_t<A, B> max(A a, B b)
{
	return a < b ? b : a;
}

inline dim3 threadIdx, blockIdx, blockDim, gridDim;  // the running fiber's (set by the scheduler)

namespace hostemu {

constexpr int MAX_THREADS = 1024;
constexpr size_t FIBER_STACK = 256 << 10;

// One workgroup at a time, its threads as fibers (ucontext) on the calling
// host thread, resumed round-robin: each pass over the fibers runs every
// live one up to its next barrier (or its end), so a pass is one barrier
// phase.  The sanitizer is told of every stack switch.
struct Fiber {
	ucontext_t ctx;
	char* stack = nullptr;
	dim3 tid;
	int orcall = 0;
	bool done = false;
	void* fake = nullptr;
};

struct Group {
	std::vector<Fiber> fibers;
	Fiber* cur = nullptr;
	ucontext_t sched;
	const void* sched_bottom = nullptr;
	size_t sched_size = 0;
	const std::function<void()>* body = nullptr;
	alignas(16) unsigned char lane[MAX_THREADS][8];  // __shfl_xor exchange, one slot per thread
	int orv[3] = {0, 0, 0};                           // __syncthreads_or, used round-robin
};
inline Group& g_group = *new Group;  // never destroyed

inline void switch_start(void** save, const void* bottom, size_t size)
{
#if defined(__SANITIZE_ADDRESS__) || __has_feature(address_sanitizer)
	__sanitizer_start_switch_fiber(save, bottom, size);
#else
	(void)save, (void)bottom, (void)size;
#endif
}
inline void switch_finish(void* save, const void** bottom, size_t* size)
{
#if defined(__SANITIZE_ADDRESS__) || __has_feature(address_sanitizer)
	__sanitizer_finish_switch_fiber(save, bottom, size);
#else
	(void)save, (void)bottom, (void)size;
#endif
}

// a thread of the workgroup waits for the others: back to the scheduler
inline void yield()
{
	Group& G = g_group;
	Fiber* f = G.cur;
	switch_start(&f->fake, G.sched_bottom, G.sched_size);
	swapcontext(&f->ctx, &G.sched);
	switch_finish(f->fake, &G.sched_bottom, &G.sched_size);
}

inline void fiber_main()
{
	Group& G = g_group;
	switch_finish(nullptr, &G.sched_bottom, &G.sched_size);
	(*G.body)();
	G.cur->done = true;
	switch_start(nullptr, G.sched_bottom, G.sched_size);  // this fiber's stack is finished with
	setcontext(&G.sched);
}

inline void run_group(dim3 g, dim3 blk, dim3 bid, const std::function<void()>& body)
{
	int n = (int)(blk.x * blk.y * blk.z);
	if (n < 1 || n > MAX_THREADS) std::abort();
	Group& G = g_group;
	if ((int)G.fibers.size() < n) G.fibers.resize(n);
	G.body = &body;
	G.orv[0] = G.orv[1] = G.orv[2] = 0;
	blockIdx = bid;
	blockDim = blk;
	gridDim = g;
	for (int t = 0; t < n; t++) {
		Fiber& f = G.fibers[t];
		if (!f.stack) f.stack = static_cast<char*>(std::malloc(FIBER_STACK));
		f.tid = dim3(t % blk.x, (t / blk.x) % blk.y, t / (blk.x * blk.y));
		f.orcall = 0;
		f.done = false;
		getcontext(&f.ctx);
		f.ctx.uc_stack.ss_sp = f.stack;
		f.ctx.uc_stack.ss_size = FIBER_STACK;
		f.ctx.uc_link = nullptr;
		makecontext(&f.ctx, fiber_main, 0);
	}
	for (int live = n; live > 0;) {
		live = 0;
		for (int t = 0; t < n; t++) {
			Fiber& f = G.fibers[t];
			if (f.done) continue;
			G.cur = &f;
			threadIdx = f.tid;
			void* save = nullptr;
			switch_start(&save, f.stack, FIBER_STACK);
			swapcontext(&G.sched, &f.ctx);
			switch_finish(save, nullptr, nullptr);
			if (!f.done) live++;
		}
	}
	G.cur = nullptr;
}

inline void run_grid(dim3 g, dim3 blk, const std::function<void()>& body)
{
	for (unsigned z = 0; z < g.z; z++)
		for (unsigned y = 0; y < g.y; y++)
			for (unsigned x = 0; x < g.x; x++) run_group(g, blk, dim3(x, y, z), body);
}

inline int flat_tid() { return (int)(threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z)); }

}  // namespace hostemu

inline void __syncthreads() { hostemu::yield(); }

inline int __syncthreads_or(int v)
{
	auto& G = hostemu::g_group;
	int k = G.cur->orcall++ % 3;
	// the slot after this one was last read two calls ago, in a phase that
	// every thread has left: clear it for the next call
	G.orv[(k + 1) % 3] = 0;
	if (v) G.orv[k] = 1;
	hostemu::yield();
	int r = G.orv[k];
	hostemu::yield();
	return r;
}

template <typename T>
inline T __shfl_xor(T v, int mask, int = 64)
{
	static_assert(sizeof(T) <= 8, "lane exchange of up to 8 bytes");
	auto& G = hostemu::g_group;
	int t = hostemu::flat_tid();
	int src = (t & ~63) | ((t & 63) ^ mask);  // within the 64-lane wave
	std::memcpy(G.lane[t], &v, sizeof(T));
	hostemu::yield();
	T r;
	std::memcpy(&r, G.lane[src], sizeof(T));
	hostemu::yield();
	return r;
}

// DPP quad permutation (dpp_ctrl < 0x100: lane 4k + j reads lane 4k +
// perm[j], perm[j] = bits 2j..2j+1 of the control) of a 32-bit value; the
// row and bank masks must be full, as in the kernels
inline int hostemu_mov_dpp(int v, int ctrl, int row_mask, int bank_mask)
{
	if (ctrl < 0 || ctrl >= 0x100 || row_mask != 0xF || bank_mask != 0xF) {
		fprintf(stderr, "hostemu: only full-mask DPP quad permutations are emulated\n");
		abort();
	}
	auto& G = hostemu::g_group;
	int t = hostemu::flat_tid();
	int src = (t & ~3) | ((ctrl >> (2 * (t & 3))) & 3);
	std::memcpy(G.lane[t], &v, sizeof(int));
	hostemu::yield();
	int r;
	std::memcpy(&r, G.lane[src], sizeof(int));
	hostemu::yield();
	return r;
}
#define __builtin_amdgcn_s_setprio(p) ((void)(p))
#define __builtin_amdgcn_mov_dpp(v, ctrl, row_mask, bank_mask, bound_ctrl) \
	hostemu_mov_dpp((v), (ctrl), (row_mask), (bank_mask))

template <typename T, typename U>
inline T atomicAdd(T* p, U v)
{
	T old = *p;
	*p = old + (T)v;
	return old;
}
template <typename T, typename U>
inline T atomicOr(T* p, U v)
{
	T old = *p;
	*p = old | (T)v;
	return old;
}

// ---- runtime -------------------------------------------------------------

enum hipError_t {
	hipSuccess = 0,
	hipErrorInvalidValue = 1,
	hipErrorOutOfMemory = 2,
	hipErrorNotSupported = 801,
	hipErrorUnknown = 999,
};
enum hipMemcpyKind {
	hipMemcpyHostToHost = 0,
	hipMemcpyHostToDevice = 1,
	hipMemcpyDeviceToHost = 2,
	hipMemcpyDeviceToDevice = 3,
	hipMemcpyDefault = 4,
};
enum hipStreamCaptureMode {
	hipStreamCaptureModeGlobal = 0,
	hipStreamCaptureModeThreadLocal = 1,
	hipStreamCaptureModeRelaxed = 2,
};
constexpr unsigned hipStreamNonBlocking = 1;
constexpr unsigned hipHostMallocDefault = 0;

namespace hostemu {
struct Stream;
struct Graph {
	std::vector<std::function<void()>> ops;
	std::vector<Stream*> joined;  // streams that joined the capture by waiting on its events
};
struct Stream {
	Graph* capture = nullptr;
};
struct Event {
	std::chrono::steady_clock::time_point t;
	Graph* capture = nullptr;  // recorded into this capture
};
inline Stream g_null_stream;
inline Stream* stream_of(Stream* s) { return s ? s : &g_null_stream; }
// runs op now, or appends it to the graph the stream is capturing
inline void submit(Stream* s, std::function<void()> op)
{
	Stream* q = stream_of(s);
	if (q->capture) q->capture->ops.push_back(std::move(op));
	else op();
}
// 256-byte aligned, exactly the size asked for (ASan's bounds are the buffer's)
inline void* dev_alloc(size_t bytes)
{
	void* p = nullptr;
	return posix_memalign(&p, 256, std::max<size_t>(bytes, 1)) == 0 ? p : nullptr;
}
}  // namespace hostemu

typedef hostemu::Stream* hipStream_t;
typedef hostemu::Event* hipEvent_t;
typedef hostemu::Graph* hipGraph_t;
typedef hostemu::Graph* hipGraphExec_t;
typedef void* hipGraphNode_t;

inline const char* hipGetErrorString(hipError_t) { return "host emulation error"; }
inline hipError_t hipGetLastError() { return hipSuccess; }
inline hipError_t hipGetDevice(int* d)
{
	*d = 0;
	return hipSuccess;
}
inline hipError_t hipGetDeviceCount(int* n)
{
	*n = 1;
	return hipSuccess;
}
inline hipError_t hipMemGetInfo(size_t* fr, size_t* tot)
{
	*fr = (size_t)1 << 30;
	*tot = (size_t)2 << 30;
	return hipSuccess;
}
inline hipError_t hipMalloc(void** p, size_t bytes)
{
	*p = hostemu::dev_alloc(bytes);
	return *p ? hipSuccess : hipErrorOutOfMemory;
}
template <typename T>
inline hipError_t hipMalloc(T** p, size_t bytes)
{
	return hipMalloc(reinterpret_cast<void**>(p), bytes);
}
inline hipError_t hipFree(void* p)
{
	std::free(p);
	return hipSuccess;
}
inline hipError_t hipHostMalloc(void** p, size_t bytes, unsigned)
{
	*p = hostemu::dev_alloc(bytes);
	return *p ? hipSuccess : hipErrorOutOfMemory;
}
template <typename T>
inline hipError_t hipHostMalloc(T** p, size_t bytes, unsigned f)
{
	return hipHostMalloc(reinterpret_cast<void**>(p), bytes, f);
}
inline hipError_t hipHostFree(void* p)
{
	std::free(p);
	return hipSuccess;
}
inline hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind)
{
	if (n) std::memcpy(d, s, n);
	return hipSuccess;
}
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind, hipStream_t st = nullptr)
{
	hostemu::submit(st, [=] {
		if (n) std::memcpy(d, s, n);
	});
	return hipSuccess;
}
inline hipError_t hipStreamCreate(hipStream_t* s)
{
	*s = new hostemu::Stream;
	return hipSuccess;
}
inline hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) { return hipStreamCreate(s); }
inline hipError_t hipStreamDestroy(hipStream_t s)
{
	delete s;
	return hipSuccess;
}
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipDeviceSynchronize() { return hipSuccess; }
inline hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode)
{
	hostemu::Stream* q = hostemu::stream_of(s);
	if (q->capture) return hipErrorInvalidValue;
	q->capture = new hostemu::Graph;
	return hipSuccess;
}
inline hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* g)
{
	hostemu::Stream* q = hostemu::stream_of(s);
	if (!q->capture) return hipErrorInvalidValue;
	*g = q->capture;
	for (auto* j : q->capture->joined) j->capture = nullptr;
	q->capture->joined.clear();
	q->capture = nullptr;
	return hipSuccess;
}
inline hipError_t hipGraphInstantiate(hipGraphExec_t* x, hipGraph_t g, hipGraphNode_t*, char*, size_t)
{
	*x = new hostemu::Graph(*g);
	return hipSuccess;
}
inline hipError_t hipGraphLaunch(hipGraphExec_t x, hipStream_t s)
{
	for (auto& op : x->ops) hostemu::submit(s, op);
	return hipSuccess;
}
inline hipError_t hipGraphDestroy(hipGraph_t g)
{
	delete g;
	return hipSuccess;
}
inline hipError_t hipGraphExecDestroy(hipGraphExec_t x)
{
	delete x;
	return hipSuccess;
}
inline hipError_t hipEventCreate(hipEvent_t* e)
{
	*e = new hostemu::Event;
	return hipSuccess;
}
inline hipError_t hipEventDestroy(hipEvent_t e)
{
	delete e;
	return hipSuccess;
}
inline hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
constexpr unsigned hipEventDisableTiming = 2;
// one in-order queue: an event is complete once recorded, except that a
// stream waiting on an event recorded into a capture joins that capture
// (its later work is recorded into the same graph, in enqueue order)
inline hipError_t hipEventRecord(hipEvent_t e, hipStream_t s = nullptr)
{
	e->capture = hostemu::stream_of(s)->capture;
	hostemu::submit(s, [=] { e->t = std::chrono::steady_clock::now(); });
	return hipSuccess;
}
inline hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned)
{
	hostemu::Stream* q = hostemu::stream_of(s);
	if (e->capture && q->capture != e->capture) {
		q->capture = e->capture;
		e->capture->joined.push_back(q);
	}
	return hipSuccess;
}
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b)
{
	*ms = std::chrono::duration<float, std::milli>(b->t - a->t).count();
	return hipSuccess;
}
inline hipError_t hipGetSymbolAddress(void** p, const void* sym)
{
	*p = const_cast<void*>(sym);
	return hipSuccess;
}
#define HIP_SYMBOL(x) (&(x))

namespace hostemu {
// the kernel's arguments are copied when the launch is submitted (a graph
// replays the values seen at capture, as on the device)
template <typename K, typename... A>
void launch(dim3 g, dim3 b, size_t, hipStream_t s, K k, A... a)
{
	submit(s, [=] { run_grid(g, b, [=] { k(a...); }); });
}
}  // namespace hostemu

#define hipLaunchKernelGGL(K, G, B, SH, ST, ...) hostemu::launch(G, B, SH, ST, K, ##__VA_ARGS__)
