// launch.h -- kernel launches that can be recorded and grouped across
// independent solves (C++ side; the C interface is in engine.h).
//
// A launch either goes to the GPU at once or, while the calling thread
// records (bbjd_rec_begin), is appended to the current job's list.
// bbjd_rec_flush then launches the lists of all jobs: the launches of one
// job in their order, and launches of the same kernel from different jobs
// merged into one grouped launch.  A grouped launch runs each job's blocks
// with the grid, arguments and block indices that job's own launch would
// have had (the kernel body sees a virtual block index, VB), so every job
// computes exactly what it computes alone; only the number of launches
// changes.  Launches of different jobs must not share writable memory.
//
// A groupable kernel is written as a body taking VB first, plus a __global__
// wrapper that passes the hardware indices (VB_HW); launchers call
// bbjd::gl<body, block size>(kernel, grid, dynamic LDS, stream, args...).
// Any other launch or copy on a recording path goes through bbjd::sl
// (BBJD_SL), which records it as a closure that keeps its place in the
// job's order.
#ifndef BBJD_LAUNCH_H
#define BBJD_LAUNCH_H

#include <hip/hip_runtime.h>
#include <atomic>
#include <functional>
#include <new>
#include <type_traits>
#include <tuple>
#include <utility>
#include <vector>

// the block index and grid size a kernel body sees; unsigned like the
// hardware's, so a body called with VB_HW compiles to the kernel it was
struct VB { unsigned bx, by, gx, gy; };
#define VB_HW VB{blockIdx.x, blockIdx.y, gridDim.x, gridDim.y}

namespace bbjd {

constexpr int kGroupMax = 16;                 // jobs per grouped launch

// Per-job arguments of a grouped launch, passed by value in the kernel
// arguments: job j owns flat blocks [start[j], start[j + 1]) laid out as its
// own gx x gy grid.
template <typename... P> struct Group {
    int nj;
    int start[kGroupMax + 1];
    int gx[kGroupMax];
    int gy[kGroupMax];
    std::tuple<P...> a[kGroupMax];
};

template <auto BODY, int BS, typename... P>
__global__ __launch_bounds__(BS) void k_group(Group<P...> g)
{
    int j = 0;
    while (j + 1 < g.nj && (int)blockIdx.x >= g.start[j + 1]) j++;
    const unsigned l = blockIdx.x - (unsigned)g.start[j], gx = (unsigned)g.gx[j];
    const VB vb{l % gx, l / gx, gx, (unsigned)g.gy[j]};
    std::apply([&](P... p) { BODY(vb, p...); }, g.a[j]);
}

struct Op;
typedef void (*GroupFn)(Op *const *ops, int n, hipStream_t s, const unsigned char *arena);

// One recorded launch: a groupable kernel (fn != nullptr; ops with the same
// fn are merged) or a closure run in order on its own.
struct Op {
    GroupFn fn = nullptr;
    void *kern = nullptr;          // the job's own kernel (one-job launch)
    dim3 grid;
    size_t lds = 0;
    size_t args = 0;               // offset of its std::tuple<P...> in the arena
    std::function<void(hipStream_t)> solo;
};

struct Recorder {
    std::vector<std::vector<Op>> jobs;
    std::vector<unsigned char> arena;   // argument tuples of the recorded launches
    int cur = -1;
    // counters read by other threads (bbjd_rec_counts) while this one runs
    std::atomic<long> launches{0};   // launches made by flushes so far
    std::atomic<long> recorded{0};   // launches recorded so far

    // room for an object of the given size and alignment (<= 16); arena
    // offsets stay valid when it grows
    size_t alloc(size_t size, size_t align)
    {
        const size_t off = (arena.size() + align - 1) / align * align;
        arena.resize(off + size);
        return off;
    }
};

// the recorder of the calling thread while it records, else null
Recorder *&tl_rec();

template <auto BODY, int BS, typename... P>
void group_launch(Op *const *ops, int n, hipStream_t s, const unsigned char *arena)
{
    typedef void (*K)(P...);
    typedef std::tuple<P...> T;
    if (n == 1) {
        const Op &o = *ops[0];
        const K kern = reinterpret_cast<K>(o.kern);
        std::apply([&](P... p) { hipLaunchKernelGGL(kern, o.grid, dim3(BS), o.lds, s, p...); },
                   *reinterpret_cast<const T *>(arena + o.args));
        return;
    }
    static_assert(sizeof(Group<P...>) <= 4096, "grouped kernel arguments exceed 4 KiB");
    Group<P...> g;
    g.nj = n;
    int tot = 0;
    size_t lds = 0;
    for (int j = 0; j < n; j++) {
        const Op &o = *ops[j];
        g.start[j] = tot;
        g.gx[j] = (int)o.grid.x;
        g.gy[j] = (int)o.grid.y;
        g.a[j] = *reinterpret_cast<const T *>(arena + o.args);
        tot += (int)(o.grid.x * o.grid.y);
        if (o.lds > lds) lds = o.lds;
    }
    g.start[n] = tot;
    hipLaunchKernelGGL((k_group<BODY, BS, P...>), dim3((unsigned)tot), dim3(BS), lds, s, g);
}

// groupable launch of kern (whose body is BODY) with block size BS
template <auto BODY, int BS, typename... P, typename... A>
inline void gl(void (*kern)(P...), dim3 grid, size_t lds, hipStream_t s, A... a)
{
    Recorder *r = tl_rec();
    if (__builtin_expect(r == nullptr, 1)) {
        hipLaunchKernelGGL(kern, grid, dim3(BS), lds, s, static_cast<P>(a)...);
        return;
    }
    typedef std::tuple<P...> T;
    static_assert(std::is_trivially_destructible<T>::value && alignof(T) <= 16, "kernel arguments");
    Op o;
    o.fn = &group_launch<BODY, BS, P...>;
    o.kern = reinterpret_cast<void *>(kern);
    o.grid = grid;
    o.lds = lds;
    o.args = r->alloc(sizeof(T), alignof(T));
    new (r->arena.data() + o.args) T(static_cast<P>(a)...);
    r->jobs[r->cur].push_back(std::move(o));
    r->recorded.fetch_add(1, std::memory_order_relaxed);
}

// any other launch or copy: run now, or recorded as a closure
template <typename F> inline void sl(hipStream_t s, F &&f)
{
    Recorder *r = tl_rec();
    if (__builtin_expect(r == nullptr, 1)) { f(s); return; }
    Op o;
    o.solo = std::function<void(hipStream_t)>(std::forward<F>(f));
    r->jobs[r->cur].push_back(std::move(o));
    r->recorded.fetch_add(1, std::memory_order_relaxed);
}

}  // namespace bbjd

#define BBJD_SL(S, KERNEL, GRID, BLOCK, LDS, ...)                                                                  \
    bbjd::sl(S, [=](hipStream_t s_) { hipLaunchKernelGGL(KERNEL, GRID, BLOCK, LDS, s_, __VA_ARGS__); })

#endif
