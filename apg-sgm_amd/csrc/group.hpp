// Launches that span several stereo pairs (batched form).
//
// A kernel of the fused schedule is written as a device body taking its block
// index as an argument; its one-pair __global__ entry passes blockIdx, and
// group_kernel runs the same body for many pairs in one grid:
//
//   * each pair's blocks start at a multiple of kGroupChunk in the flattened
//     1-D grid, in its own grid's x-fastest order, and the blocks past a
//     pair's count (at most kGroupChunk - 1 per pair) exit at once;
//   * block b reads record b / kGroupChunk: where the chunk starts in its
//     pair's grid, that grid's shape, and the pair's kernel arguments as one
//     Pack<...>, exactly the values its one-pair launch passes.  One record
//     per chunk keeps the block's argument loads one scalar load away from
//     its index (a table of pairs indexed through a chunk map would be two).
//
// The pairs of one group launch share the kernel instance (template
// arguments) and the block size; the dynamic LDS size is the largest any of
// them asks for.  Every pair computes exactly what its one-pair launch does.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <type_traits>
#include <vector>

namespace apgk {

constexpr uint32_t kGroupChunkLog2 = 5;
constexpr uint32_t kGroupChunk = 1u << kGroupChunkLog2;

struct GroupChunk {
    uint32_t base;    // the chunk's first block in its pair's grid
    uint32_t count;   // blocks of the pair's grid, gx * gy * gz
    uint32_t gx, gxy;
};

// A flat record of kernel arguments: head, then the rest.
template <typename... T>
struct Pack;
template <>
struct Pack<> {};
template <typename H, typename... T>
struct Pack<H, T...> {
    H h;
    Pack<T...> t;
};

inline Pack<> make_pack() { return {}; }
template <typename H, typename... T>
Pack<H, T...> make_pack(H h, T... t) {
    Pack<H, T...> p;
    p.h = h;
    p.t = make_pack(t...);
    return p;
}

// A pointer field read as a pointer of the global address space, as clang
// passes a kernel's pointer arguments: its loads and stores are then
// global_* instructions, not flat_* (which also wait on the LDS counter).
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Wincompatible-pointer-types-discards-qualifiers"
template <typename T>
__device__ __forceinline__ T* global_ptr(T* const& field) {
    typedef T __attribute__((address_space(1))) GT;
    GT* const g = *reinterpret_cast<GT* const*>(&field);
    return (T*)g;
}
#pragma clang diagnostic pop

// Structs whose arrays a body indexes at run time (PlaneSet::p,
// AuxPair::img): passed by reference, read in place as a one-pair kernel
// reads its argument segment (a private copy of an array indexed at run time
// goes to scratch memory).
template <typename T>
struct GroupByRef : std::false_type {};

// Every other field is read here, at kernel entry, into registers (pointers
// marked global), as a one-pair kernel's arguments are; a body then never
// reloads them after its own stores.
template <typename Fn, typename... Done>
__device__ __forceinline__ void pack_run(const uint3 b, const Pack<>&, const Done&... d) {
    Fn::run(b, d...);
}
template <typename Fn, typename H, typename... T, typename... Done>
__device__ __forceinline__ void pack_run(const uint3 b, const Pack<H, T...>& p, const Done&... d) {
    if constexpr (std::is_pointer<H>::value) {
        const H v = global_ptr(p.h);
        pack_run<Fn>(b, p.t, d..., v);
    } else if constexpr (GroupByRef<H>::value) {
        pack_run<Fn>(b, p.t, d..., p.h);
    } else {
        const H v = p.h;
        pack_run<Fn>(b, p.t, d..., v);
    }
}

__device__ __forceinline__ uint3 block3() { return make_uint3(blockIdx.x, blockIdx.y, blockIdx.z); }

template <typename P>
struct GroupRecord {
    GroupChunk c;
    P args;
};

// The records are read through the constant address space, as kernel
// arguments are: the compiler then knows no store of the body changes them
// and keeps their fields in scalar registers instead of reloading them.
template <typename P>
using ConstRecord = const __attribute__((address_space(4))) GroupRecord<P>;

template <typename Fn, typename P>
__global__ __launch_bounds__(Fn::kThreads) void group_kernel(ConstRecord<P>* __restrict__ recs) {
    ConstRecord<P>* rc = recs + (blockIdx.x >> kGroupChunkLog2);
    const uint32_t l = rc->c.base + (blockIdx.x & (kGroupChunk - 1));
    if (l >= rc->c.count) return;   // the pair's padding to a whole chunk
    uint3 bid;
    bid.z = l / rc->c.gxy;
    const uint32_t q = l - bid.z * rc->c.gxy;
    bid.y = q / rc->c.gx;
    bid.x = q - bid.y * rc->c.gx;
    const GroupRecord<P>& r = *(const GroupRecord<P>*)rc;
    pack_run<Fn>(bid, r.args);
}

// One pair's launch of one kernel instance, recorded instead of launched.
struct GroupRec {
    const void* kernel;   // group_kernel<Fn, P>
    int threads;
    size_t shm;
    dim3 grid;
    std::vector<uint8_t> args;   // the P record
    size_t rec_bytes;            // sizeof(GroupRecord<P>)
    size_t args_at;              // offsetof(GroupRecord<P>, args)
};

// Collects the launches of one stage of several pairs.  A launcher given a
// sink records the launches it supports here and enqueues the others as
// usual.
struct LaunchSink {
    std::vector<GroupRec> recs;
    template <typename Fn, typename... A>
    hipError_t add(dim3 grid, size_t shm, A... a) {
        using P = Pack<A...>;
        const P p = make_pack(a...);
        GroupRec r;
        r.kernel = (const void*)group_kernel<Fn, P>;
        r.threads = Fn::kThreads;
        r.shm = shm;
        r.grid = grid;
        r.args.assign((const uint8_t*)&p, (const uint8_t*)&p + sizeof p);
        r.rec_bytes = sizeof(GroupRecord<P>);
        r.args_at = offsetof(GroupRecord<P>, args);
        recs.push_back(std::move(r));
        return hipSuccess;
    }
};

}  // namespace apgk
