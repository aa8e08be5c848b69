// Batched FRSZ2: many independent arrays of uneven length, mixed float32 /
// float64, bit widths and directions, in one persistent launch captured in a
// hipGraph.
//
// The jobs' tiles form one global index space (job j owns tiles
// [tile0_j, tile0_j + ntiles_j)).  A persistent grid of workgroups claims
// tiles from a device counter (`chunk` tiles per atomic claim; chunk 0 =
// static round-robin, no counter), reads the tile's job from the plan's tile
// table (or, without one, finds it by binary search over tile0), and runs
// the same tile code as the plain kernels for the job's class (direction,
// dtype, storage width U, packed or not).  Jobs the tile path cannot take (B
// not a power of two or larger than a tile, unaligned pointers) run as plain
// launches of the generic kernels inside the same graph.  The counter is
// reset by a one-thread kernel captured at the head of the graph.
//
// Tile shapes are per class and come in three settings (narrow, balanced,
// wide; see class_vpt).  The autotuner times candidate (workgroups, tile
// shape, chunk) settings on the plan's own jobs and keeps the fastest.
//
// No counterpart in the reference: its kernels take one array per launch
// (include/frsz2.hpp:1258-1288).
#include "frsz2_capi.h"
#include "frsz2_device.hpp"
#include "frsz2_internal.hpp"
#include "tile.hpp"
#include "batch_internal.hpp"
#include "tuning_defaults.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace frsz2 {

using tile::tplan;
using tile::WG;

namespace {

// class = op * 14 + dtype_slot for compress / decompress, dtype_slot 0..13:
//   float32: U8, U16, U32 x {aligned, packed} -> 0..5
//   float64: U8, U16, U32, U64 x {aligned, packed} -> 6..13
// gather: 28 + (float32 U8, U16, U32 -> 0..2; float64 U8..U64 -> 3..6)
__host__ __device__ constexpr int
class_of(int op, int dtype, int bits)
{
  const int U = storage_bits(bits);
  const int ui = U == 8 ? 0 : U == 16 ? 1 : U == 32 ? 2 : 3;
  if (op == FRSZ2_OP_GATHER) {
    return 28 + (dtype == FRSZ2_FLOAT32 ? ui : 3 + ui);
  }
  const int packed = bits < U ? 1 : 0;
  const int slot = dtype == FRSZ2_FLOAT32 ? ui * 2 + packed : 6 + ui * 2 + packed;
  return op * 14 + slot;
}

// indices per gather tile (256 threads x 4)
constexpr int GATHER_TILE = WG * 4;

// random-access decode of indices [t*GATHER_TILE, +GATHER_TILE) of a gather job
template<typename T, int U>
__device__ __forceinline__ void
gather_tile(const bjob& J, std::uint64_t t)
{
  const layout L{ J.n, J.B, J.nb, J.R, J.CB, J.bits, U };
  const auto* stream = static_cast<const std::uint8_t*>(J.in);
  T* out = static_cast<T*>(J.out);
  const std::uint64_t i0 = t * GATHER_TILE + static_cast<std::uint64_t>(threadIdx.x) * 4;
  std::uint64_t idx[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    idx[v] = i0 + v < J.m ? J.idx[i0 + v] : ~std::uint64_t{ 0 };
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    if (i0 + v < J.m) {
      out[i0 + v] = idx[v] < J.n ? decode_element<T, U, true>(stream, L, idx[v]) : T(0);
    }
  }
}

// (op, dtype, U index, packed) of a class
__host__ __device__ constexpr int cls_op(int cls) { return cls >= 28 ? 2 : cls / 14; }
__host__ __device__ constexpr int cls_dtype(int cls) { return cls >= 28 ? (cls < 31 ? 0 : 1) : ((cls % 14) < 6 ? 0 : 1); }
__host__ __device__ constexpr int cls_ui(int cls)
{
  return cls >= 28 ? (cls < 31 ? cls - 28 : cls - 31) : ((cls % 14) < 6 ? (cls % 14) / 2 : ((cls % 14) - 6) / 2);
}
__host__ __device__ constexpr int cls_packed(int cls) { return cls >= 28 ? 0 : (cls % 14) % 2; }
__host__ __device__ constexpr int cls_default(int cls)
{
  return default_variant(cls_op(cls) == 2 ? 1 : cls_op(cls), cls_dtype(cls), cls_ui(cls), cls_packed(cls));
}

// Per-class tile shape, three settings:
//   0 narrow:   4 values per thread for every class (smallest tiles, finest
//               load balance, fewest registers), float64 decompress staged
//   1 balanced: up to 8 values per thread, float64 decompress through the
//               LDS output stage
//   2 wide:     the shapes measured fastest for the plain kernels (up to 16
//               values per thread, staged decompress stores)
//   3 grouped:  not one persistent kernel but one launch per class present,
//               each with that class's default plain-kernel variant and its
//               own register allocation (launch_grouped)
// In one kernel every class shares the register allocation of the widest,
// so wider shapes trade occupancy for bytes in flight; the autotuner picks.
__host__ __device__ constexpr int
class_vpt(int cls, int shape)
{
  if (shape == 3) {
    return venc_vpt(cls_default(cls)); // grouped: the plain kernels' default variant
  }
  const int op = cls / 14;
  const int slot = cls % 14;
  const bool f32 = slot < 6;
  const int ui = f32 ? slot / 2 : (slot - 6) / 2;
  const bool packed = (slot % 2) == 1;
  if (shape == 0 || op == 2) {
    return 4;
  }
  if (op == 0) {
    return f32 ? ((ui == 2 && !packed) ? 4 : 8) : 4;
  }
  if (shape == 1) {
    return f32 ? 8 : 4;
  }
  if (f32) {
    return packed ? 8 : (ui == 2 ? 8 : 16);
  }
  return packed ? 4 : 8;
}

// threads per workgroup of a class's tiles: 64 where the grouped shape runs
// a plain-kernel default that uses one-wave workgroups, else 256
__host__ __device__ constexpr int
class_wgs(int cls, int shape)
{
  return shape == 3 && cls_op(cls) != 2 && venc_small(cls_default(cls)) != 0 ? 64 : WG;
}

__host__ __device__ constexpr bool
class_stage(int cls, int shape)
{
  if (shape == 3) {
    return cls_op(cls) == 1 && venc_stage(cls_default(cls)) != 0;
  }
  const int slot = cls % 14;
  const bool f32 = slot < 6;
  const bool packed = (slot % 2) == 1;
  if (cls / 14 != 1 || cls >= 28 || packed) {
    return false;
  }
  // bits == U: the strided decode loads one code per lane per step, which
  // pays off for 4-byte codes (float32, U 32) and not for 1-2-byte ones
  return !(f32 && slot / 2 == 2);
}

// LDS dwords of the batch kernel: the largest payload image or output stage
// of any class (a tile uses one or the other)
template<int SHAPE>
constexpr int
batch_lds_dwords()
{
  return SHAPE == 0 ? tile::lds_dwords<4>() : tile::lds_dwords<8>();
}
static_assert(tile::stage_dwords<double, 4>() <= tile::lds_dwords<4>(), "stage fits");
static_assert(tile::stage_dwords<float, 16>() <= tile::lds_dwords<8>(), "stage fits");
static_assert(tile::stage_dwords<double, 8>() <= tile::lds_dwords<8>(), "stage fits");

// OPS: bit mask of the directions a kernel instance handles (1 compress,
// 2 decompress, 4 gather); the code (and register allocation) of the other
// classes is left out.
template<int SHAPE, int OPS>
__device__ __forceinline__ void
run_tile(const bjob& J, std::uint64_t t, std::uint32_t* lds, std::uint32_t* wred)
{
  const tplan p{ J.n, J.nb, J.R, J.CB, J.bits, J.lb < 0 ? 0 : J.lb, 0 };
  auto* o8 = static_cast<std::uint8_t*>(J.out);
  const auto* i8 = static_cast<const std::uint8_t*>(J.in);
#define FRSZ2_C(K, T, U, PK)                                                                                 \
  case K:                                                                                                    \
    if constexpr ((OPS & 1) != 0)                                                                            \
    tile::compress_tile<T, U, class_vpt(K, SHAPE), PK, true>(static_cast<const typename fmt<T>::uw*>(J.in),   \
                                                            o8, p, t, lds, wred);                            \
    if constexpr ((OPS & 1) != 0)                                                                            \
    if (t == 0)                                                                                              \
      tile::zero_pad(o8, J.nb, U, J.R, static_cast<int>(threadIdx.x), WG);                                   \
    break
#define FRSZ2_D(K, T, U, PK)                                                                                 \
  case K:                                                                                                    \
    if constexpr ((OPS & 2) != 0)                                                                            \
    tile::decompress_tile<T, U, class_vpt(K, SHAPE), PK, true, class_stage(K, SHAPE)>(                         \
      i8, static_cast<typename fmt<T>::uw*>(J.out), p, t, lds, lds);                                         \
    break
  switch (J.cls) {
    FRSZ2_C(0, float, 8, false);
    FRSZ2_C(1, float, 8, true);
    FRSZ2_C(2, float, 16, false);
    FRSZ2_C(3, float, 16, true);
    FRSZ2_C(4, float, 32, false);
    FRSZ2_C(5, float, 32, true);
    FRSZ2_C(6, double, 8, false);
    FRSZ2_C(7, double, 8, true);
    FRSZ2_C(8, double, 16, false);
    FRSZ2_C(9, double, 16, true);
    FRSZ2_C(10, double, 32, false);
    FRSZ2_C(11, double, 32, true);
    FRSZ2_C(12, double, 64, false);
    FRSZ2_C(13, double, 64, true);
    FRSZ2_D(14, float, 8, false);
    FRSZ2_D(15, float, 8, true);
    FRSZ2_D(16, float, 16, false);
    FRSZ2_D(17, float, 16, true);
    FRSZ2_D(18, float, 32, false);
    FRSZ2_D(19, float, 32, true);
    FRSZ2_D(20, double, 8, false);
    FRSZ2_D(21, double, 8, true);
    FRSZ2_D(22, double, 16, false);
    FRSZ2_D(23, double, 16, true);
    FRSZ2_D(24, double, 32, false);
    FRSZ2_D(25, double, 32, true);
    FRSZ2_D(26, double, 64, false);
    FRSZ2_D(27, double, 64, true);
    case 28:
      if constexpr ((OPS & 4) != 0)
        gather_tile<float, 8>(J, t);
      break;
    case 29:
      if constexpr ((OPS & 4) != 0)
        gather_tile<float, 16>(J, t);
      break;
    case 30:
      if constexpr ((OPS & 4) != 0)
        gather_tile<float, 32>(J, t);
      break;
    case 31:
      if constexpr ((OPS & 4) != 0)
        gather_tile<double, 8>(J, t);
      break;
    case 32:
      if constexpr ((OPS & 4) != 0)
        gather_tile<double, 16>(J, t);
      break;
    case 33:
      if constexpr ((OPS & 4) != 0)
        gather_tile<double, 32>(J, t);
      break;
    case 34:
      if constexpr ((OPS & 4) != 0)
        gather_tile<double, 64>(J, t);
      break;
    default:
      break;
  }
#undef FRSZ2_C
#undef FRSZ2_D
}

// job owning global tile g, searching from hint (jobs sorted by tile0)
__device__ __forceinline__ int
find_job(const bjob* __restrict__ jobs, int njobs, std::uint64_t g, int hint)
{
  if (g >= jobs[hint].tile0 && (hint + 1 == njobs || g < jobs[hint + 1].tile0)) {
    return hint;
  }
  int lo = 0;
  int hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].tile0 <= g) {
      lo = mid;
    } else {
      hi = mid - 1;
    }
  }
  return lo;
}

// The persistent loop: tiles claimed `chunk` at a time from *counter, or
// (chunk == 0) taken round-robin by workgroup index.  A tile's job comes from
// jtab (descriptor index per tile) when given, else from a search.
template<int SHAPE, int OPS>
__device__ __forceinline__ void
batch_loop(const bjob* __restrict__ jobs, int njobs, std::uint64_t total_tiles, unsigned long long* counter,
           int chunk, const std::uint32_t* __restrict__ jtab)
{
  __shared__ std::uint32_t lds[batch_lds_dwords<SHAPE>()];
  __shared__ std::uint32_t wred[WG / tile::WAVE];
  __shared__ unsigned long long claim;
  if (chunk == 0 && jtab) {
    // Round-robin with the job lookups run ahead: the descriptor of the next
    // tile and the table entry of the one after it are loaded before the
    // current tile runs, so their latency passes under its work instead of
    // between tiles.  Every value loaded here is workgroup-uniform.
    // (Past the last tile the lookups read the last tile's entry, so every
    // load is unconditional and in bounds.)
    const std::uint64_t G = gridDim.x;
    const std::uint64_t last = total_tiles - 1;
    std::uint64_t g = blockIdx.x;
    if (g >= total_tiles) {
      return;
    }
    bjob J = jobs[jtab[g]];
    std::uint32_t jn = jtab[g + G < last ? g + G : last];
    for (;;) {
      const std::uint64_t gn = g + G;
      const bjob Jn = jobs[jn];
      const std::uint32_t jnn = jtab[gn + G < last ? gn + G : last];
      run_tile<SHAPE, OPS>(J, g - J.tile0, lds, wred);
      if (gn >= total_tiles) {
        break;
      }
      g = gn;
      J = Jn;
      jn = jnn;
    }
    return;
  }
  int j = 0;
  std::uint64_t next = blockIdx.x;
  for (;;) {
    std::uint64_t g0;
    std::uint64_t g1;
    if (chunk > 0) {
      if (threadIdx.x == 0) {
        claim = atomicAdd(counter, static_cast<unsigned long long>(chunk));
      }
      __syncthreads();
      g0 = claim;
      __syncthreads();
      g1 = g0 + static_cast<std::uint64_t>(chunk);
    } else {
      g0 = next;
      g1 = g0 + 1;
      next += gridDim.x;
    }
    if (g0 >= total_tiles) {
      break;
    }
    if (g1 > total_tiles) {
      g1 = total_tiles;
    }
    for (std::uint64_t g = g0; g < g1; ++g) {
      j = __builtin_amdgcn_readfirstlane(jtab ? static_cast<int>(jtab[g]) : find_job(jobs, njobs, g, j));
      const bjob J = jobs[j];
      run_tile<SHAPE, OPS>(J, g - J.tile0, lds, wred);
    }
  }
}

template<int SHAPE, int OPS>
__global__ __launch_bounds__(WG) void k_batch(const bjob* __restrict__ jobs,
                                              int njobs,
                                              std::uint64_t total_tiles,
                                              unsigned long long* counter,
                                              int chunk,
                                              const std::uint32_t* __restrict__ jtab)
{
  batch_loop<SHAPE, OPS>(jobs, njobs, total_tiles, counter, chunk, jtab);
}

// Up to INLINE_JOBS descriptors passed in the kernel arguments: a launch of
// a few jobs needs no host-to-device copy of its descriptors before it.
struct job_pack
{
  bjob j[INLINE_JOBS];
};

template<int SHAPE, int OPS>
__global__ __launch_bounds__(WG) void k_batch_inline(job_pack pack, int njobs, std::uint64_t total_tiles)
{
  batch_loop<SHAPE, OPS>(pack.j, njobs, total_tiles, nullptr, 0, nullptr);
}

__global__ void
k_reset_counter(unsigned long long* counter)
{
  *counter = 0ull;
}

// Grouped mode: one workgroup per tile of the jobs of one class (jobs[0..njobs)
// are that class's descriptors, tiles tile_base + blockIdx.x).
template<typename T, int U, int VPT, bool PACKED, bool NT, bool STAGE, int OP, int WGS>
__device__ __forceinline__ void
grouped_tile(const bjob* __restrict__ jobs, int njobs, std::uint64_t tile_base, const std::uint32_t* __restrict__ jtab)
{
  __shared__ std::uint32_t lds[PACKED ? tile::lds_dwords_u<U, VPT, WGS>() : 1];
  __shared__ std::uint32_t stage[STAGE ? tile::stage_dwords<T, VPT, WGS>() : 1];
  __shared__ std::uint32_t wred[WGS / tile::WAVE];
  const std::uint64_t g = tile_base + blockIdx.x;
  // the plan's tile table gives the job in one load; without one, a binary
  // search over the class's descriptors
  const int j = jtab ? static_cast<int>(jtab[g]) : __builtin_amdgcn_readfirstlane(find_job(jobs, njobs, g, 0));
  const bjob J = jobs[j];
  const std::uint64_t t = g - J.tile0;
  if constexpr (OP == 2) {
    gather_tile<T, U>(J, t);
  } else {
    const tplan p{ J.n, J.nb, J.R, J.CB, J.bits, J.lb < 0 ? 0 : J.lb, 0 };
    if constexpr (OP == 0) {
      auto* o8 = static_cast<std::uint8_t*>(J.out);
      tile::compress_tile<T, U, VPT, PACKED, NT, WGS>(static_cast<const typename fmt<T>::uw*>(J.in), o8, p, t, lds,
                                                      wred);
      if (t == 0) {
        tile::zero_pad(o8, J.nb, U, J.R, static_cast<int>(threadIdx.x), WGS);
      }
    } else {
      tile::decompress_tile<T, U, VPT, PACKED, NT, STAGE, WGS>(static_cast<const std::uint8_t*>(J.in),
                                                               static_cast<typename fmt<T>::uw*>(J.out), p, t, lds,
                                                               stage);
    }
  }
}

template<typename T, int U, int VPT, bool PACKED, bool NT, bool STAGE, int OP, int WGS>
__global__ __launch_bounds__(WGS) void k_grouped(const bjob* __restrict__ jobs,
                                                 int njobs,
                                                 std::uint64_t tile_base,
                                                 const std::uint32_t* __restrict__ jtab)
{
  grouped_tile<T, U, VPT, PACKED, NT, STAGE, OP, WGS>(jobs, njobs, tile_base, jtab);
}

// the class's descriptors (at most INLINE_JOBS) in the kernel arguments
template<typename T, int U, int VPT, bool PACKED, bool NT, bool STAGE, int OP, int WGS>
__global__ __launch_bounds__(WGS) void k_grouped_inline(job_pack pack, int njobs, std::uint64_t tile_base)
{
  grouped_tile<T, U, VPT, PACKED, NT, STAGE, OP, WGS>(pack.j, njobs, tile_base, nullptr);
}

// h: host copy of the class's descriptors when they go in the kernel
// arguments (nd <= INLINE_JOBS), else null
template<int CLS>
void
launch_class(const bjob* d, const bjob* h, int nd, std::uint64_t t0, std::uint64_t nt, const std::uint32_t* jtab,
             hipStream_t s)
{
  using T = std::conditional_t<cls_dtype(CLS) == 0, float, double>;
  constexpr int U = 8 << cls_ui(CLS);
  constexpr int V = cls_default(CLS);
  constexpr int OP = cls_op(CLS);
  constexpr int VPT = OP == 2 ? 4 : venc_vpt(V);
  constexpr bool PACKED = cls_packed(CLS) != 0;
  constexpr bool NT = venc_nt(V) != 0;
  constexpr bool STAGE = OP == 1 && venc_stage(V) != 0;
  constexpr int WGS = class_wgs(CLS, 3);
  if (h != nullptr) {
    job_pack pack{};
    std::copy(h, h + nd, pack.j);
    hipLaunchKernelGGL((k_grouped_inline<T, U, VPT, PACKED, NT, STAGE, OP, WGS>), dim3(static_cast<unsigned>(nt)),
                       dim3(WGS), 0, s, pack, nd, t0);
  } else {
    hipLaunchKernelGGL((k_grouped<T, U, VPT, PACKED, NT, STAGE, OP, WGS>), dim3(static_cast<unsigned>(nt)),
                       dim3(WGS), 0, s, d, nd, t0, jtab);
  }
}

int
ilog2_exact(std::uint64_t B)
{
  if (B == 0 || (B & (B - 1)) != 0) {
    return -1;
  }
  int l = 0;
  while ((std::uint64_t{ 1 } << l) < B) {
    ++l;
  }
  return l;
}

bool
aligned16(const void* p)
{
  return (reinterpret_cast<std::uintptr_t>(p) & 15u) == 0;
}

// grouped mode: one k_grouped launch per run of same-class descriptors
// (descriptors are built in class order)
// the grouped kernel of class cls over descriptors d[0..nd) (h: their host
// copy when they go in the kernel arguments), tiles [t0, t0 + nt)
hipError_t
launch_class_of(int cls, const bjob* d, const bjob* h, int nd, std::uint64_t t0, std::uint64_t nt,
                const std::uint32_t* jtab, hipStream_t s)
{
#define FRSZ2_G(K)                                                                                           \
  case K:                                                                                                    \
    launch_class<K>(d, h, nd, t0, nt, jtab, s);                                                              \
    break
  switch (cls) {
    FRSZ2_G(0); FRSZ2_G(1); FRSZ2_G(2); FRSZ2_G(3); FRSZ2_G(4); FRSZ2_G(5); FRSZ2_G(6);
    FRSZ2_G(7); FRSZ2_G(8); FRSZ2_G(9); FRSZ2_G(10); FRSZ2_G(11); FRSZ2_G(12); FRSZ2_G(13);
    FRSZ2_G(14); FRSZ2_G(15); FRSZ2_G(16); FRSZ2_G(17); FRSZ2_G(18); FRSZ2_G(19); FRSZ2_G(20);
    FRSZ2_G(21); FRSZ2_G(22); FRSZ2_G(23); FRSZ2_G(24); FRSZ2_G(25); FRSZ2_G(26); FRSZ2_G(27);
    FRSZ2_G(28); FRSZ2_G(29); FRSZ2_G(30); FRSZ2_G(31); FRSZ2_G(32); FRSZ2_G(33); FRSZ2_G(34);
    default:
      break;
  }
#undef FRSZ2_G
  return hipGetLastError();
}

hipError_t
launch_grouped(const bjob* d_desc, const bjob* h_desc, int nd, std::uint64_t total_tiles, const std::uint32_t* jtab,
               bool inl, hipStream_t s)
{
  int d0 = 0;
  while (d0 < nd) {
    const int cls = h_desc[d0].cls;
    int d1 = d0 + 1;
    while (d1 < nd && h_desc[d1].cls == cls) {
      ++d1;
    }
    const std::uint64_t t0 = h_desc[d0].tile0;
    const std::uint64_t t1 = d1 < nd ? h_desc[d1].tile0 : total_tiles;
    if (t1 > t0) {
      const hipError_t e =
        launch_class_of(cls, d_desc + d0, inl ? h_desc + d0 : nullptr, d1 - d0, t0, t1 - t0, jtab, s);
      if (e != hipSuccess) {
        return e;
      }
    }
    d0 = d1;
  }
  return hipSuccess;
}

} // namespace

} // namespace frsz2

namespace frsz2 {

int
validate_job(const frsz2_job& j, int idx)
{
  const std::string tag = "job " + std::to_string(idx) + ": ";
  if (j.op != FRSZ2_OP_COMPRESS && j.op != FRSZ2_OP_DECOMPRESS && j.op != FRSZ2_OP_GATHER) {
    return set_error(FRSZ2_EARG, tag + "op must be compress, decompress or gather");
  }
  int rc = check_params(j.dtype, j.bits, j.block);
  if (rc) {
    return rc;
  }
  const layout L = layout::make(j.bits, j.block, j.n);
  if (j.op == FRSZ2_OP_COMPRESS && j.out_bytes < L.total_bytes()) {
    return set_error(FRSZ2_ECAPACITY, tag + "output capacity below stream size");
  }
  if (j.op != FRSZ2_OP_COMPRESS && j.in_bytes < L.total_bytes()) {
    return set_error(FRSZ2_ECAPACITY, tag + "input stream below stream size");
  }
  const std::uintptr_t vs = j.dtype == FRSZ2_FLOAT32 ? 4 : 8;
  // decompress writes n values, gather m values: out_bytes must hold them
  const std::uint64_t outv = j.op == FRSZ2_OP_DECOMPRESS ? j.n : j.op == FRSZ2_OP_GATHER ? j.m : 0;
  if (j.op != FRSZ2_OP_COMPRESS && j.out_bytes / vs < outv) {
    return set_error(FRSZ2_ECAPACITY, tag + "output capacity " + std::to_string(j.out_bytes) + " bytes below " +
                                        std::to_string(outv) + " values");
  }
  const void* vals = j.op == FRSZ2_OP_COMPRESS ? j.in : j.out;
  const void* strm = j.op == FRSZ2_OP_COMPRESS ? j.out : j.in;
  const std::uint64_t nvals = j.op == FRSZ2_OP_GATHER ? j.m : j.n;
  if (!strm || (nvals > 0 && !vals) || (reinterpret_cast<std::uintptr_t>(vals) & (vs - 1)) ||
      (reinterpret_cast<std::uintptr_t>(strm) & 7u) ||
      (j.op == FRSZ2_OP_GATHER && j.m > 0 && (!j.idx || (reinterpret_cast<std::uintptr_t>(j.idx) & 7u)))) {
    return set_error(FRSZ2_EARG, tag + "null or misaligned pointer");
  }
  return FRSZ2_OK;
}

// Descriptor of job j for the tile path at tile_shape, its tiles starting at
// global tile tile0: the job's tile count, or -1 when the tile path cannot
// take it (B not a power of two or larger than the class's tile, pointers
// off 16-byte alignment), in which case it runs as a plain launch.
std::int64_t
describe_job(const frsz2_job& j, int tile_shape, std::uint64_t tile0, bjob* d)
{
  const int cls = class_of(j.op, j.dtype, j.bits);
  const layout L = layout::make(j.bits, j.block, j.n);
  const int lb = ilog2_exact(j.block);
  std::uint64_t nt;
  if (j.op == FRSZ2_OP_GATHER) {
    nt = (j.m + GATHER_TILE - 1) / GATHER_TILE;
  } else {
    const std::uint64_t TILE = static_cast<std::uint64_t>(class_wgs(cls, tile_shape)) * class_vpt(cls, tile_shape);
    if (lb < 0 || j.block > TILE || !aligned16(j.in) || !aligned16(j.out)) {
      return -1;
    }
    nt = (j.n + TILE - 1) / TILE;
    if (j.op == FRSZ2_OP_COMPRESS && nt == 0) {
      nt = 1; // the pad of an empty stream is still written
    }
  }
  *d = bjob{};
  d->in = j.in;
  d->out = j.out;
  d->idx = j.idx;
  d->n = j.n;
  d->nb = L.nb;
  d->R = L.R;
  d->tile0 = tile0;
  d->m = j.m;
  d->B = j.block;
  d->CB = static_cast<std::uint32_t>(L.CB);
  d->bits = j.bits;
  d->lb = lb;
  d->cls = cls;
  return static_cast<std::int64_t>(nt);
}

std::uint64_t
build_descriptors(const frsz2_job* jobs, std::size_t n, int tile_shape, std::vector<bjob>& desc,
                  std::vector<char>& tiled)
{
  tiled.assign(n, 0);
  std::uint64_t tiles = 0;
  // jobs of one class get consecutive tiles, so a workgroup's successive
  // claims mostly run the same code path of the kernel
  std::vector<std::size_t> order(n);
  for (std::size_t i = 0; i < n; ++i) {
    order[i] = i;
  }
  std::stable_sort(order.begin(), order.end(), [&](std::size_t a, std::size_t b) {
    return class_of(jobs[a].op, jobs[a].dtype, jobs[a].bits) < class_of(jobs[b].op, jobs[b].dtype, jobs[b].bits);
  });
  for (const std::size_t i : order) {
    bjob d;
    const std::int64_t nt = describe_job(jobs[i], tile_shape, tiles, &d);
    if (nt < 0) {
      continue;
    }
    tiled[i] = 1;
    if (nt == 0) {
      continue;
    }
    desc.push_back(d);
    tiles += static_cast<std::uint64_t>(nt);
  }
  return tiles;
}

std::vector<std::uint32_t>
tile_table(const std::vector<bjob>& desc, std::uint64_t total_tiles, bool within_class)
{
  std::vector<std::uint32_t> tab(total_tiles, 0u);
  const std::size_t nd = desc.size();
  std::size_t d0 = 0; // first descriptor of the current class (0 for global indices)
  for (std::size_t d = 0; d < nd; ++d) {
    if (within_class && desc[d].cls != desc[d0].cls) {
      d0 = d;
    }
    const std::uint64_t t1 = d + 1 < nd ? desc[d + 1].tile0 : total_tiles;
    for (std::uint64_t t = desc[d].tile0; t < t1; ++t) {
      tab[t] = static_cast<std::uint32_t>(d - d0);
    }
  }
  return tab;
}

static std::uint64_t
persistent_grid(const batch_shape& shape, std::uint64_t total_tiles)
{
  const std::uint64_t grid = static_cast<std::uint64_t>(shape.workgroups);
  return grid < total_tiles ? grid : total_tiles;
}

bool
descriptors_inline(const batch_shape& shape, int ndesc, std::uint64_t total_tiles)
{
  if (ndesc <= 0 || ndesc > INLINE_JOBS || total_tiles == 0) {
    return false;
  }
  return shape.tile_shape == 3 || persistent_grid(shape, total_tiles) == total_tiles;
}

hipError_t
launch_persistent(const bjob* d_desc, const bjob* h_desc, int ndesc, std::uint64_t total_tiles,
                  unsigned long long* d_counter, const batch_shape& shape, hipStream_t s, const std::uint32_t* d_jtab)
{
  if (ndesc == 0 || total_tiles == 0) {
    return hipSuccess;
  }
  if (shape.tile_shape == 3) {
    return launch_grouped(d_desc, h_desc, ndesc, total_tiles, d_jtab,
                          h_desc != nullptr && descriptors_inline(shape, ndesc, total_tiles), s);
  }
  const std::uint64_t grid = persistent_grid(shape, total_tiles);
  // a grid that covers every tile takes one tile per workgroup: no claims,
  // so no counter to reset
  const int chunk = grid == total_tiles ? 0 : shape.tiles_per_claim;
  const bool inl = h_desc != nullptr && descriptors_inline(shape, ndesc, total_tiles);
  if (chunk > 0) {
    hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(1), 0, s, d_counter);
  }
  const dim3 gd(static_cast<unsigned>(grid));
  job_pack pack{};
  if (inl) {
    std::copy(h_desc, h_desc + ndesc, pack.j);
  }
  // one kernel instance per (tile shape, set of directions present)
  const int ops = shape.ops == 0 ? 7 : shape.ops == 1 ? 1 : (shape.ops & 1) != 0 ? 7 : 6;
#define FRSZ2_KB(S, O)                                                                                       \
  if (inl) {                                                                                                 \
    hipLaunchKernelGGL((k_batch_inline<S, O>), gd, dim3(WG), 0, s, pack, ndesc, total_tiles);                \
  } else {                                                                                                   \
    hipLaunchKernelGGL((k_batch<S, O>), gd, dim3(WG), 0, s, d_desc, ndesc, total_tiles, d_counter, chunk,    \
                       d_jtab);                                                                              \
  }
  switch (shape.tile_shape * 8 + ops) {
    case 0 * 8 + 1: FRSZ2_KB(0, 1); break;
    case 0 * 8 + 6: FRSZ2_KB(0, 6); break;
    case 0 * 8 + 7: FRSZ2_KB(0, 7); break;
    case 1 * 8 + 1: FRSZ2_KB(1, 1); break;
    case 1 * 8 + 6: FRSZ2_KB(1, 6); break;
    case 1 * 8 + 7: FRSZ2_KB(1, 7); break;
    case 2 * 8 + 1: FRSZ2_KB(2, 1); break;
    case 2 * 8 + 6: FRSZ2_KB(2, 6); break;
    default: FRSZ2_KB(2, 7); break;
  }
#undef FRSZ2_KB
  return hipGetLastError();
}

int
run_plain(const frsz2_job& j, hipStream_t s)
{
  switch (j.op) {
    case FRSZ2_OP_COMPRESS:
      return compress_device(j.dtype, j.bits, j.block, j.in, j.n, j.out, j.out_bytes, s);
    case FRSZ2_OP_DECOMPRESS:
      return decompress_device(j.dtype, j.bits, j.block, j.in, j.in_bytes, j.n, j.out, s);
    default:
      return gather_device(j.dtype, j.bits, j.block, j.in, j.in_bytes, j.n, j.idx, j.m, j.out, s);
  }
}

int
device_cus()
{
  int dev = 0;
  int cus = 0;
  if (hipGetDevice(&dev) == hipSuccess &&
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cus > 0) {
    return cus;
  }
  return 256;
}

} // namespace frsz2

struct frsz2_batch_plan
{
  std::vector<frsz2_job> jobs;        // as given
  std::vector<char> tiled;            // job index -> 1 if it takes the persistent kernel
  std::vector<frsz2::bjob> host_desc; // descriptors for cfg.tile_shape
  frsz2::bjob* d_desc = nullptr;
  std::uint32_t* d_jtab = nullptr; // tile -> descriptor index (within its class for tile shape 3)
  unsigned long long* d_counter = nullptr;
  int ndesc = 0;
  std::uint64_t total_tiles = 0;
  frsz2_batch_config cfg{};
  int cus = 256;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int ops = 0; // directions present among the jobs (1 compress, 2 decompress, 4 gather)
};

namespace frsz2 {
namespace {

// (re)build and upload the descriptors for cfg.tile_shape
int
upload_descriptors(frsz2_batch_plan* P)
{
  P->host_desc.clear();
  P->total_tiles = build_descriptors(P->jobs.data(), P->jobs.size(), P->cfg.tile_shape, P->host_desc, P->tiled);
  if (P->d_desc) {
    (void)hipFree(P->d_desc);
    P->d_desc = nullptr;
  }
  if (P->d_jtab) {
    (void)hipFree(P->d_jtab);
    P->d_jtab = nullptr;
  }
  P->ndesc = static_cast<int>(P->host_desc.size());
  if (P->ndesc > 0) {
    hipError_t e = hipMalloc(&P->d_desc, P->host_desc.size() * sizeof(bjob));
    if (e != hipSuccess) {
      return hip_status(e, "hipMalloc");
    }
    e = hipMemcpy(P->d_desc, P->host_desc.data(), P->host_desc.size() * sizeof(bjob), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      return hip_status(e, "hipMemcpy");
    }
    if (P->total_tiles > 0) {
      // the tile's job by table (4 bytes per tile of at least 1 KiB of data)
      // instead of a search over thousands of jobs
      const std::vector<std::uint32_t> tab = tile_table(P->host_desc, P->total_tiles, P->cfg.tile_shape == 3);
      e = hipMalloc(&P->d_jtab, tab.size() * sizeof(std::uint32_t));
      if (e != hipSuccess) {
        return hip_status(e, "hipMalloc");
      }
      e = hipMemcpy(P->d_jtab, tab.data(), tab.size() * sizeof(std::uint32_t), hipMemcpyHostToDevice);
      if (e != hipSuccess) {
        return hip_status(e, "hipMemcpy");
      }
    }
  }
  return FRSZ2_OK;
}

// enqueue the whole batch on s (also used under capture)
int
enqueue(frsz2_batch_plan* P, hipStream_t s)
{
  if (P->cfg.plain_launches) {
    for (const frsz2_job& j : P->jobs) {
      const int rc = run_plain(j, s);
      if (rc) {
        return rc;
      }
    }
    return FRSZ2_OK;
  }
  const batch_shape shape{ P->cfg.tile_shape, P->cfg.workgroups, P->cfg.tiles_per_claim, P->ops };
  const hipError_t e =
    launch_persistent(P->d_desc, P->host_desc.data(), P->ndesc, P->total_tiles, P->d_counter, shape, s, P->d_jtab);
  if (e != hipSuccess) {
    return hip_status(e, "batch launch");
  }
  // jobs outside the tile path: plain launches on the same stream
  for (std::size_t i = 0; i < P->jobs.size(); ++i) {
    if (!P->tiled[i]) {
      const int rc = run_plain(P->jobs[i], s);
      if (rc) {
        return rc;
      }
    }
  }
  return FRSZ2_OK;
}

void
drop_graph(frsz2_batch_plan* P)
{
  if (P->exec) {
    (void)hipGraphExecDestroy(P->exec);
    P->exec = nullptr;
  }
  if (P->graph) {
    (void)hipGraphDestroy(P->graph);
    P->graph = nullptr;
  }
}

int
capture_graph(frsz2_batch_plan* P)
{
  drop_graph(P);
  if (!P->cfg.use_graph) {
    return FRSZ2_OK;
  }
  hipStream_t cs = nullptr;
  hipError_t e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
  if (e != hipSuccess) {
    return hip_status(e, "hipStreamCreate");
  }
  e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  int rc = e == hipSuccess ? enqueue(P, cs) : hip_status(e, "hipStreamBeginCapture");
  hipGraph_t g = nullptr;
  const hipError_t e2 = hipStreamEndCapture(cs, &g);
  if (rc == FRSZ2_OK && e2 != hipSuccess) {
    rc = hip_status(e2, "hipStreamEndCapture");
  }
  if (rc == FRSZ2_OK) {
    P->graph = g;
    e = hipGraphInstantiate(&P->exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      rc = hip_status(e, "hipGraphInstantiate");
    }
  } else if (g) {
    (void)hipGraphDestroy(g);
  }
  (void)hipStreamDestroy(cs);
  return rc;
}

int
apply_config(frsz2_batch_plan* P, const frsz2_batch_config& c)
{
  frsz2_batch_config n = c;
  if (n.tile_shape < 0 || n.tile_shape > 3) {
    n.tile_shape = 3;
  }
  if (n.workgroups <= 0) {
    n.workgroups = 4 * P->cus;
  }
  if (n.tiles_per_claim < 0) {
    n.tiles_per_claim = 0;
  }
  const bool rebuild = n.tile_shape != P->cfg.tile_shape || P->host_desc.empty();
  P->cfg = n;
  if (rebuild) {
    const int rc = upload_descriptors(P);
    if (rc) {
      return rc;
    }
  }
  return capture_graph(P);
}

void
destroy(frsz2_batch_plan* P)
{
  drop_graph(P);
  if (P->d_desc) {
    (void)hipFree(P->d_desc);
  }
  if (P->d_jtab) {
    (void)hipFree(P->d_jtab);
  }
  if (P->d_counter) {
    (void)hipFree(P->d_counter);
  }
  delete P;
}

} // namespace
} // namespace frsz2

extern "C" {

int
frsz2_batch_create(const frsz2_job* jobs, size_t njobs, const frsz2_batch_config* cfg, frsz2_batch_plan** plan)
{
  using namespace frsz2;
  clear_error();
  if (!plan || (njobs > 0 && !jobs)) {
    return set_error(FRSZ2_EARG, "null pointer");
  }
  *plan = nullptr;
  for (size_t i = 0; i < njobs; ++i) {
    const int rc = validate_job(jobs[i], static_cast<int>(i));
    if (rc) {
      return rc;
    }
  }
  auto* P = new frsz2_batch_plan();
  P->jobs.assign(jobs, jobs + njobs);
  P->cus = device_cus();
  for (size_t i = 0; i < njobs; ++i) {
    P->ops |= 1 << jobs[i].op;
  }
  hipError_t e = hipMalloc(&P->d_counter, sizeof(unsigned long long));
  if (e != hipSuccess) {
    destroy(P);
    return hip_status(e, "hipMalloc");
  }
  frsz2_batch_config c{};
  if (cfg) {
    c = *cfg;
  } else {
    c.workgroups = 0;
    c.tile_shape = 3;
    c.tiles_per_claim = 4;
    c.use_graph = 1;
    c.plain_launches = 0;
  }
  P->cfg.tile_shape = -1; // force a descriptor build
  const int rc = apply_config(P, c);
  if (rc) {
    destroy(P);
    return rc;
  }
  *plan = P;
  return FRSZ2_OK;
}

int
frsz2_batch_launch(frsz2_batch_plan* plan, frsz2_stream_t stream)
{
  using namespace frsz2;
  clear_error();
  if (!plan) {
    return set_error(FRSZ2_EARG, "null plan");
  }
  const auto s = static_cast<hipStream_t>(stream);
  if (plan->exec) {
    return hip_status(hipGraphLaunch(plan->exec, s), "hipGraphLaunch");
  }
  return enqueue(plan, s);
}

int
frsz2_batch_get_config(const frsz2_batch_plan* plan, frsz2_batch_config* cfg)
{
  frsz2::clear_error();
  if (!plan || !cfg) {
    return frsz2::set_error(FRSZ2_EARG, "null pointer");
  }
  *cfg = plan->cfg;
  return FRSZ2_OK;
}

int
frsz2_batch_set_config(frsz2_batch_plan* plan, const frsz2_batch_config* cfg)
{
  frsz2::clear_error();
  if (!plan || !cfg) {
    return frsz2::set_error(FRSZ2_EARG, "null pointer");
  }
  return frsz2::apply_config(plan, *cfg);
}

int
frsz2_batch_autotune(frsz2_batch_plan* plan, frsz2_stream_t stream, int reps, frsz2_batch_config* chosen)
{
  using namespace frsz2;
  clear_error();
  if (!plan) {
    return set_error(FRSZ2_EARG, "null plan");
  }
  if (reps <= 0) {
    reps = 3;
  }
  const auto s = static_cast<hipStream_t>(stream);
  const int use_graph = plan->cfg.use_graph;
  hipEvent_t e0 = nullptr;
  hipEvent_t e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    return set_error(FRSZ2_EHIP, "hipEventCreate failed");
  }
  frsz2_batch_config best = plan->cfg;
  float best_ms = -1.0f;
  const int shapes[4] = { 0, 1, 2, 3 };
  const int wg_per_cu[4] = { 2, 4, 8, 16 };
  const int chunks[4] = { 0, 1, 4, 16 };
  int rc = FRSZ2_OK;
  for (int vi = 0; vi < 4 && rc == FRSZ2_OK; ++vi) {
    for (int wi = 0; wi < 4 && rc == FRSZ2_OK; ++wi) {
      for (int ci = 0; ci < 4 && rc == FRSZ2_OK; ++ci) {
        if (shapes[vi] == 3 && (wi > 0 || ci > 0)) {
          continue; // grouped launches have no grid or claim setting
        }
        frsz2_batch_config c{};
        c.tile_shape = shapes[vi];
        c.workgroups = wg_per_cu[wi] * plan->cus;
        c.tiles_per_claim = chunks[ci];
        c.use_graph = 0;
        c.plain_launches = 0;
        rc = apply_config(plan, c);
        if (rc) {
          break;
        }
        rc = enqueue(plan, s); // warm-up
        if (rc) {
          break;
        }
        (void)hipEventRecord(e0, s);
        for (int r = 0; r < reps && rc == FRSZ2_OK; ++r) {
          rc = enqueue(plan, s);
        }
        (void)hipEventRecord(e1, s);
        if (hipEventSynchronize(e1) != hipSuccess) {
          rc = set_error(FRSZ2_EHIP, "hipEventSynchronize failed");
          break;
        }
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (best_ms < 0.0f || ms < best_ms) {
          best_ms = ms;
          best = c;
        }
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  best.use_graph = use_graph;
  const int rc2 = apply_config(plan, best);
  if (rc == FRSZ2_OK) {
    rc = rc2;
  }
  if (chosen) {
    *chosen = plan->cfg;
  }
  return rc;
}

int
frsz2_batch_destroy(frsz2_batch_plan* plan)
{
  frsz2::clear_error();
  if (plan) {
    frsz2::destroy(plan);
  }
  return FRSZ2_OK;
}

} // extern "C"

// ---------------------------------------------------------------- allocation-free batched launch
//
// frsz2_batch_run_noalloc runs a batch the way a plan with the grouped shape
// does (one launch per kernel class present, one workgroup per tile, the
// plain kernels' variants), without allocating: the descriptors of the
// tiled jobs live in `scratch`, a file-scope device array of
// NOALLOC_SCRATCH_BYTES that every call on a device shares; a workgroup
// finds its tile's job by binary search over its class's descriptors (no
// tile table).  The descriptors are written into scratch by small
// stream-ordered kernels that take them in their kernel arguments, so no
// host staging buffer is involved.  Jobs the tiles cannot take run as plain
// launches, as in a plan.

namespace frsz2 {
namespace {

constexpr std::size_t NOALLOC_SCRATCH_BYTES = std::size_t{ 1 } << 20;
constexpr std::size_t NOALLOC_MAX_DESC = NOALLOC_SCRATCH_BYTES / sizeof(bjob); // 10922

// descriptors per store kernel: 40 x 96 bytes of kernel arguments
constexpr int STORE_JOBS = 40;
struct store_pack
{
  bjob j[STORE_JOBS];
};
static_assert(sizeof(store_pack) + 16 <= 4096, "store kernel arguments fit the kernel argument space");

// The device copy of the array is only reachable through the address
// hipGetSymbolAddress returns for the current device; its name in host code
// is not a device address.
static __device__ __align__(256) std::uint8_t scratch[NOALLOC_SCRATCH_BYTES];

// dst[i] = pack.j[i] for i < count (one thread per descriptor)
__global__ __launch_bounds__(64) void k_store_descriptors(store_pack pack, int count, bjob* __restrict__ dst)
{
  const int i = static_cast<int>(threadIdx.x);
  if (i < count) {
    dst[i] = pack.j[i];
  }
}

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch[MAX_DEVICES] = {};

// device address of the current device's copy of `scratch`, looked up once
// per device with that device current
int
scratch_address(bjob** out)
{
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    return hip_status(e, "hipGetDevice");
  }
  if (dev < 0 || dev >= MAX_DEVICES) {
    return set_error(FRSZ2_EARG, "device " + std::to_string(dev) + " beyond the allocation-free path's " +
                                   std::to_string(MAX_DEVICES) + " devices");
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      return hip_status(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
    }
    g_scratch[dev] = p;
  }
  *out = static_cast<bjob*>(g_scratch[dev]);
  return FRSZ2_OK;
}

constexpr int NUM_CLASSES = 35;

} // namespace
} // namespace frsz2

extern "C" {

int
frsz2_batch_run_noalloc(const frsz2_job* jobs, size_t njobs, frsz2_stream_t stream)
{
  using namespace frsz2;
  clear_error();
  if (njobs > 0 && !jobs) {
    return set_error(FRSZ2_EARG, "null jobs");
  }
  for (size_t i = 0; i < njobs; ++i) {
    const int rc = validate_job(jobs[i], static_cast<int>(i));
    if (rc) {
      return rc;
    }
  }
  // descriptors needed: one per job the tiles take that has tiles
  std::size_t ndesc = 0;
  for (size_t i = 0; i < njobs; ++i) {
    bjob d;
    if (describe_job(jobs[i], 3, 0, &d) > 0) {
      ++ndesc;
    }
  }
  if (ndesc > NOALLOC_MAX_DESC) {
    return set_error(FRSZ2_ECAPACITY, "the allocation-free batch takes at most " + std::to_string(NOALLOC_MAX_DESC) +
                                        " tiled jobs (" + std::to_string(NOALLOC_SCRATCH_BYTES) +
                                        " bytes of descriptor scratch); this batch has " + std::to_string(ndesc));
  }
  const auto s = static_cast<hipStream_t>(stream);
  bjob* d_scratch = nullptr;
  if (ndesc > 0) {
    const int rc = scratch_address(&d_scratch);
    if (rc) {
      return rc;
    }
  }
  // class by class, in job order within a class (the order a plan uses):
  // store the class's descriptors, then launch its grouped kernel over them
  std::size_t pos = 0;     // descriptors stored so far
  std::uint64_t tiles = 0; // global tile numbering, as in a plan
  store_pack pack{};
  for (int cls = 0; cls < NUM_CLASSES; ++cls) {
    const std::size_t d0 = pos;
    const std::uint64_t t0 = tiles;
    int fill = 0;
    std::size_t flushed = pos;
    for (size_t i = 0; i < njobs; ++i) {
      const frsz2_job& j = jobs[i];
      if (class_of(j.op, j.dtype, j.bits) != cls) {
        continue;
      }
      bjob d;
      const std::int64_t nt = describe_job(j, 3, tiles, &d);
      if (nt <= 0) {
        continue;
      }
      pack.j[fill++] = d;
      ++pos;
      tiles += static_cast<std::uint64_t>(nt);
      if (fill == STORE_JOBS) {
        hipLaunchKernelGGL(k_store_descriptors, dim3(1), dim3(64), 0, s, pack, fill, d_scratch + flushed);
        flushed += static_cast<std::size_t>(fill);
        fill = 0;
      }
    }
    if (fill > 0) {
      hipLaunchKernelGGL(k_store_descriptors, dim3(1), dim3(64), 0, s, pack, fill, d_scratch + flushed);
    }
    if (pos > d0) {
      hipError_t e = hipGetLastError();
      if (e == hipSuccess) {
        e = launch_class_of(cls, d_scratch + d0, nullptr, static_cast<int>(pos - d0), t0, tiles - t0, nullptr, s);
      }
      if (e != hipSuccess) {
        return hip_status(e, "allocation-free batch launch");
      }
    }
  }
  // jobs outside the tile path: plain launches on the same stream
  for (size_t i = 0; i < njobs; ++i) {
    bjob d;
    if (describe_job(jobs[i], 3, 0, &d) < 0) {
      const int rc = run_plain(jobs[i], s);
      if (rc) {
        return rc;
      }
    }
  }
  return FRSZ2_OK;
}

} // extern "C"
