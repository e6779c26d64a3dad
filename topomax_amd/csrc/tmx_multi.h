// Merged launches: one launch of a kernel covers the same step of many independent problems
// (the designs of a batch).
//
// A host launcher called while a Recorder is active on the calling thread does not launch: it
// appends the kernel's merged form, its grid for this problem, its dynamic LDS and its
// arguments to the recorder.  The batch driver (tmx_batch.hip) records each design's launch
// sequence for one phase of the objective step, merges the sequences of all designs into one
// launch per kernel instance and position, and replays those: each block looks up its design
// in a block -> design table and runs the kernel body with its block index inside that design's
// own grid.  The body is the code the single-problem kernel runs, so every design gets the same
// arithmetic, in the same order, as in a launch of its own.
#pragma once

#include <hip/hip_runtime.h>

#include <cstring>
#include <tuple>
#include <vector>

namespace tmx {

// A kernel body's view of its block: index and count within its own problem's grid.
struct BlkId {
  unsigned id, n;
};

__device__ __forceinline__ BlkId hw_blk() { return BlkId{blockIdx.x, gridDim.x}; }

// Block -> design table of one merged launch: block b belongs to design owner[b] (an index into
// the launch's argument records), whose blocks are first[d] .. first[d + 1] - 1.  With skip,
// the blocks of design d do nothing while *skip[d] is nonzero (a device flag that only kernels
// of other launches write).
struct MultiMap {
  const unsigned short* owner;
  const int* first;
  const double* const* skip;
};

// Pointers read from an argument record in memory are generic pointers to the compiler; this
// gives them the global address space, so the kernel's loads and stores through them compile
// to global (not flat) memory instructions, as they do for pointers passed as kernel arguments.
// (The host pass of the compiler only type-checks device code and sees the plain pointer.)
#if defined(__HIP_DEVICE_COMPILE__)
template <typename P>
__device__ __forceinline__ __attribute__((address_space(1))) P* glob(P* p) {
  return (__attribute__((address_space(1))) P*)p;
}
#else
template <typename P>
__device__ __forceinline__ P* glob(P* p) {
  return p;
}
#endif

template <int NT, typename Tup, auto Body>
__global__ __launch_bounds__(NT) void k_multi(const Tup* __restrict__ recs, MultiMap m) {
  const unsigned d = m.owner[blockIdx.x];
  if (m.skip && *m.skip[d] != 0.0) return;  // (uniform across the block)
  const int f0 = m.first[d], f1 = m.first[d + 1];
  const Tup& r = recs[d];
  const BlkId b{blockIdx.x - (unsigned)f0, (unsigned)(f1 - f0)};
  std::apply([&](const auto&... a) { Body(b, a...); }, r);
}

using MultiFn = hipError_t (*)(const void* recs, MultiMap m, unsigned grid, size_t lds,
                               hipStream_t s);

template <int NT, typename Tup, auto Body>
hipError_t multi_launch(const void* recs, MultiMap m, unsigned grid, size_t lds, hipStream_t s) {
  if (lds > 64 * 1024) {  // (dynamic LDS beyond 64 KiB needs the kernel's attribute raised)
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_multi<NT, Tup, Body>),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  k_multi<NT, Tup, Body><<<grid, NT, lds, s>>>(static_cast<const Tup*>(recs), m);
  return hipGetLastError();
}

// the merged kernel's name with its template arguments (diagnostics)
template <int NT, typename Tup, auto Body>
const char* multi_name() {
  return __PRETTY_FUNCTION__;
}

// One recorded launch: the merged kernel, this problem's grid and LDS, and where its argument
// record sits in the recorder's arena.
struct LaunchRec {
  MultiFn fn;
  const char* name;
  unsigned nblk;
  unsigned threads;
  size_t lds;
  size_t off, size;
};

class Recorder {
 public:
  std::vector<LaunchRec> recs;
  std::vector<unsigned char> arena;
  void clear() {
    recs.clear();
    arena.clear();
  }
  template <int NT, auto Body, typename... A>
  void add(unsigned nblk, size_t lds, const std::tuple<A...>& args) {
    using Tup = std::tuple<A...>;
    static_assert(alignof(Tup) <= 16, "argument records are 16-byte aligned");
    const size_t off = (arena.size() + 15) & ~size_t(15);
    arena.resize(off + sizeof(Tup));
    std::memcpy(arena.data() + off, static_cast<const void*>(&args), sizeof(Tup));
    recs.push_back(LaunchRec{&multi_launch<NT, Tup, Body>, multi_name<NT, Tup, Body>(), nblk,
                             (unsigned)NT, lds, off, sizeof(Tup)});
  }
};

// The recorder that launchers on this thread append to instead of launching (null: launch).
inline Recorder*& active_recorder() {
  static thread_local Recorder* r = nullptr;
  return r;
}

// Records the launch when a recorder is active, else launches KERNEL on the stream.  BODY is
// the kernel's body (first parameter BlkId, then the kernel's own parameters).
#define TMX_LAUNCH_OR_RECORD(NT, BODY, NB, LDS, S, KERNEL, ...)                            \
  do {                                                                                     \
    if (::tmx::Recorder* rec_ = ::tmx::active_recorder())                                  \
      rec_->add<NT, BODY>((unsigned)(NB), (LDS), std::make_tuple(__VA_ARGS__));            \
    else                                                                                   \
      (void)(KERNEL<<<(NB), (NT), (LDS), (S)>>>(__VA_ARGS__));                             \
  } while (0)

}  // namespace tmx
