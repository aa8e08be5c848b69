// Batched FRSZ2: many independent arrays of uneven length, mixed float32 /
// float64, bit widths and directions, in one captured hipGraph launch.
//
// The jobs' tiles form one global index space, ordered by kernel class
// (direction, dtype, storage width U, packed or not; job j owns tiles
// [tile0_j, tile0_j + ntiles_j)).  The graph holds one grouped kernel per
// class present, each with the plain kernels' built-in variant for that
// class and its own register allocation, running the same tile code as the
// plain kernels; a workgroup reads its tile's job from the plan's tile table
// (or, without one, finds it by binary search over tile0).  Jobs the tile
// path cannot take (unaligned pointers, blocks larger than a tile) run as
// plain launches inside the same graph.
//
// The tile schedule is chosen by the autotuner on the plan's own jobs: one
// workgroup per tile or a persistent per-class grid of what the device holds
// at once, each in workgroup order or with one contiguous range of a
// class's tiles per XCD.
//
// No counterpart in the reference: its kernels take one array per launch
// (include/frsz2.hpp:1258-1288).
#include "frsz2_capi.h"
#include "frsz2_device.hpp"
#include "frsz2_internal.hpp"
#include "frsz2_tile.hpp"
#include "batch_internal.hpp"
#include "frsz2_tuning.hpp"

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

// Values per thread, workgroup size and staged decompress stores of a
// class's tiles: the plain kernels' built-in variant for the class
// (include/frsz2_tuning.hpp); gathers take 4 indices per thread.
__host__ __device__ constexpr int
class_vpt(int cls)
{
  return cls_op(cls) == 2 ? 4 : venc_vpt(cls_default(cls));
}

__host__ __device__ constexpr int
class_wgs(int cls)
{
  return cls_op(cls) != 2 && venc_small(cls_default(cls)) != 0 ? 64 : WG;
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

// Up to INLINE_JOBS descriptors passed in the kernel arguments: a launch of
// a few jobs needs no host-to-device copy of its descriptors before it.
struct job_pack
{
  bjob j[INLINE_JOBS];
};

// Grouped mode: global tile g of the jobs of one class (jobs[0..njobs) are
// that class's descriptors), with the workgroup's LDS arrays.
template<typename T, int U, int VPT, bool PACKED, bool NT, bool STAGE, int OP, int WGS>
__device__ __forceinline__ void
grouped_tile_at(const bjob* __restrict__ jobs, int njobs, std::uint64_t g, const std::uint32_t* __restrict__ jtab,
                std::uint32_t* lds, std::uint32_t* stage, std::uint32_t* wred)
{
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

// the class's tile of virtual index v < ntiles: v itself (workgroup order)
// or, with xcd, the XCD-contiguous order (workgroups are dealt round-robin to
// the 8 XCDs; each XCD then takes one contiguous range of the class's tiles)
__device__ __forceinline__ std::uint64_t
class_tile(std::uint64_t v, std::uint64_t ntiles, int xcd)
{
  return xcd ? tile::xcd_tile(static_cast<std::uint32_t>(v), static_cast<std::uint32_t>(ntiles)) : v;
}

// one workgroup per tile: virtual tile blockIdx.x of the class's ntiles
template<typename T, int U, int VPT, bool PACKED, bool NT, bool STAGE, int OP, int WGS>
__device__ __forceinline__ void
grouped_tile(const bjob* __restrict__ jobs, int njobs, std::uint64_t tile_base, std::uint64_t ntiles, int xcd,
             const std::uint32_t* __restrict__ jtab)
{
  __shared__ std::uint32_t lds[PACKED ? tile::lds_dwords_u<U, VPT, WGS>() : 1];
  __shared__ std::uint32_t stage[STAGE ? tile::stage_dwords<T, VPT, WGS>() : 1];
  __shared__ std::uint32_t wred[WGS / tile::WAVE];
  grouped_tile_at<T, U, VPT, PACKED, NT, STAGE, OP, WGS>(jobs, njobs, tile_base + class_tile(blockIdx.x, ntiles, xcd),
                                                         jtab, lds, stage, wred);
}

template<typename T, int U, int VPT, bool PACKED, bool NT, bool STAGE, int OP, int WGS>
__global__ __launch_bounds__(WGS) void k_grouped(const bjob* __restrict__ jobs,
                                                 int njobs,
                                                 std::uint64_t tile_base,
                                                 std::uint64_t ntiles,
                                                 int xcd,
                                                 const std::uint32_t* __restrict__ jtab)
{
  grouped_tile<T, U, VPT, PACKED, NT, STAGE, OP, WGS>(jobs, njobs, tile_base, ntiles, xcd, jtab);
}

// Persistent: a grid of at most what the device holds at once strides over
// the class's ntiles tiles: workgroup w takes virtual tiles w, w + grid, ...
// (with xcd the grid is a multiple of 8, so a workgroup's virtual tiles all
// map into its own XCD's range).  Every tile function ends with the barrier
// that frees its LDS for the next tile.
template<typename T, int U, int VPT, bool PACKED, bool NT, bool STAGE, int OP, int WGS>
__global__ __launch_bounds__(WGS) void k_grouped_persistent(const bjob* __restrict__ jobs,
                                                            int njobs,
                                                            std::uint64_t tile_base,
                                                            std::uint64_t ntiles,
                                                            int xcd,
                                                            const std::uint32_t* __restrict__ jtab)
{
  __shared__ std::uint32_t lds[PACKED ? tile::lds_dwords_u<U, VPT, WGS>() : 1];
  __shared__ std::uint32_t stage[STAGE ? tile::stage_dwords<T, VPT, WGS>() : 1];
  __shared__ std::uint32_t wred[WGS / tile::WAVE];
  for (std::uint64_t v = blockIdx.x; v < ntiles; v += gridDim.x) {
    grouped_tile_at<T, U, VPT, PACKED, NT, STAGE, OP, WGS>(jobs, njobs, tile_base + class_tile(v, ntiles, xcd), jtab,
                                                           lds, stage, wred);
  }
}

// the class's descriptors (at most INLINE_JOBS) in the kernel arguments
template<typename T, int U, int VPT, bool PACKED, bool NT, bool STAGE, int OP, int WGS>
__global__ __launch_bounds__(WGS) void k_grouped_inline(job_pack pack,
                                                        int njobs,
                                                        std::uint64_t tile_base,
                                                        std::uint64_t ntiles,
                                                        int xcd)
{
  grouped_tile<T, U, VPT, PACKED, NT, STAGE, OP, WGS>(pack.j, njobs, tile_base, ntiles, xcd, nullptr);
}

// Tiles per launch of one workgroup per tile: 2^22 tiles of at most 256
// threads stay below the 2^32 work-items of one dispatch (the classes of the
// stated 4096-array workload hold 3.8-8.4 million tiles, so most of them
// take two launches).
constexpr std::uint64_t MAX_LAUNCH_TILES = std::uint64_t{ 1 } << 22;

// One workgroup per tile of class CLS in the tile shape VPT x WGS (the
// class's non-temporal hint and staged stores), descriptors d[0..nd) in
// device memory, tiles [t0, t0 + nt), at most MAX_LAUNCH_TILES tiles per
// launch (a dispatch counts its work-items in 32 bits).
template<int CLS, int VPT, int WGS>
void
launch_tiles(const bjob* d, int nd, std::uint64_t t0, std::uint64_t nt, const std::uint32_t* jtab, hipStream_t s,
             int xcd)
{
  using T = std::conditional_t<cls_dtype(CLS) == 0, float, double>;
  constexpr int U = 8 << cls_ui(CLS);
  constexpr int V = cls_default(CLS);
  constexpr int OP = cls_op(CLS);
  constexpr bool PACKED = cls_packed(CLS) != 0;
  constexpr bool NT = venc_nt(V) != 0;
  constexpr bool STAGE = OP == 1 && venc_stage(V) != 0;
  for (std::uint64_t c0 = 0; c0 < nt; c0 += MAX_LAUNCH_TILES) {
    const std::uint64_t cn = nt - c0 < MAX_LAUNCH_TILES ? nt - c0 : MAX_LAUNCH_TILES;
    hipLaunchKernelGGL((k_grouped<T, U, VPT, PACKED, NT, STAGE, OP, WGS>), dim3(static_cast<unsigned>(cn)), dim3(WGS),
                       0, s, d, nd, t0 + c0, cn, xcd, jtab);
  }
}

// the same in the tile shape vpt x wgs chosen at run time (a shape
// validate_shapes accepts for the class)
template<int CLS>
void
launch_tiles_shape(int vpt, int wgs, const bjob* d, int nd, std::uint64_t t0, std::uint64_t nt,
                   const std::uint32_t* jtab, hipStream_t s, int xcd)
{
  if constexpr (cls_op(CLS) != 2) {
    constexpr bool two = cls_dtype(CLS) == 1 && cls_ui(CLS) >= 1; // 2 values per thread: one 16-byte float64 chunk
#define FRSZ2_TS(V, W) launch_tiles<CLS, V, W>(d, nd, t0, nt, jtab, s, xcd)
    if (wgs == 64) {
      switch (vpt) {
        case 2:
          if constexpr (two) {
            FRSZ2_TS(2, 64);
          }
          break;
        case 4: FRSZ2_TS(4, 64); break;
        case 8: FRSZ2_TS(8, 64); break;
        default: FRSZ2_TS(16, 64); break;
      }
    } else {
      switch (vpt) {
        case 2:
          if constexpr (two) {
            FRSZ2_TS(2, WG);
          }
          break;
        case 4: FRSZ2_TS(4, WG); break;
        case 8: FRSZ2_TS(8, WG); break;
        default: FRSZ2_TS(16, WG); break;
      }
    }
#undef FRSZ2_TS
  }
}

// The grouped kernel of class CLS over its descriptors d[0..nd), tiles
// [t0, t0 + nt).  h: host copy of the descriptors when they go in the kernel
// arguments (nd <= INLINE_JOBS, one workgroup per tile, built-in shape),
// else null.  vpt x wgs: the tile shape of the one-workgroup-per-tile
// schedule (the persistent schedule runs the built-in shape).
template<int CLS>
void
launch_class(const bjob* d, const bjob* h, int nd, std::uint64_t t0, std::uint64_t nt, const std::uint32_t* jtab,
             hipStream_t s, const batch_schedule& sc, int vpt, int wgs)
{
  if (!sc.persistent && (vpt != class_vpt(CLS) || wgs != class_wgs(CLS))) {
    launch_tiles_shape<CLS>(vpt, wgs, d, nd, t0, nt, jtab, s, sc.xcd_order ? 1 : 0);
    return;
  }
  using T = std::conditional_t<cls_dtype(CLS) == 0, float, double>;
  constexpr int U = 8 << cls_ui(CLS);
  constexpr int V = cls_default(CLS);
  constexpr int OP = cls_op(CLS);
  constexpr int VPT = class_vpt(CLS);
  constexpr bool PACKED = cls_packed(CLS) != 0;
  constexpr bool NT = venc_nt(V) != 0;
  constexpr bool STAGE = OP == 1 && venc_stage(V) != 0;
  constexpr int WGS = class_wgs(CLS);
  const int xcd = sc.xcd_order ? 1 : 0;
  if (sc.persistent) {
    const auto kern = k_grouped_persistent<T, U, VPT, PACKED, NT, STAGE, OP, WGS>;
    // workgroups of this kernel resident per CU (registers, LDS, waves), once per process
    static const int resident = [kern]() {
      int b = 0;
      return hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, kern, WGS, 0) == hipSuccess && b > 0 ? b : 4;
    }();
    const int per_cu = sc.workgroups > 0 && sc.workgroups < resident ? sc.workgroups : resident;
    const std::uint64_t cap = static_cast<std::uint64_t>(per_cu) * static_cast<std::uint64_t>(device_cus());
    std::uint64_t grid = nt < cap ? nt : cap;
    if (xcd && grid >= 8) {
      grid = grid / 8 * 8;
    }
    hipLaunchKernelGGL(kern, dim3(static_cast<unsigned>(grid)), dim3(WGS), 0, s, d, nd, t0, nt, xcd, jtab);
  } else {
    // one workgroup per tile, at most MAX_LAUNCH_TILES tiles per launch (a
    // dispatch counts its work-items in 32 bits)
    job_pack pack{};
    if (h != nullptr) {
      std::copy(h, h + nd, pack.j);
    }
    for (std::uint64_t c0 = 0; c0 < nt; c0 += MAX_LAUNCH_TILES) {
      const std::uint64_t cn = nt - c0 < MAX_LAUNCH_TILES ? nt - c0 : MAX_LAUNCH_TILES;
      if (h != nullptr) {
        hipLaunchKernelGGL((k_grouped_inline<T, U, VPT, PACKED, NT, STAGE, OP, WGS>), dim3(static_cast<unsigned>(cn)),
                           dim3(WGS), 0, s, pack, nd, t0 + c0, cn, xcd);
      } else {
        hipLaunchKernelGGL((k_grouped<T, U, VPT, PACKED, NT, STAGE, OP, WGS>), dim3(static_cast<unsigned>(cn)),
                           dim3(WGS), 0, s, d, nd, t0 + c0, cn, xcd, jtab);
      }
    }
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

// the grouped kernel of class cls over descriptors d[0..nd) (h: their host
// copy when they go in the kernel arguments), tiles [t0, t0 + nt), in the
// class's tile shape under sh
hipError_t
launch_class_of(int cls, const bjob* d, const bjob* h, int nd, std::uint64_t t0, std::uint64_t nt,
                const std::uint32_t* jtab, hipStream_t s, const batch_schedule& sc, const class_shapes* sh = nullptr)
{
  int vpt = 0;
  int wgs = 0;
  class_shape(cls, sh, vpt, wgs);
#define FRSZ2_G(K)                                                                                           \
  case K:                                                                                                    \
    launch_class<K>(d, h, nd, t0, nt, jtab, s, sc, vpt, wgs);                                                \
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

// one grouped launch per run of same-class descriptors (descriptors are
// built in class order)
hipError_t
launch_grouped(const bjob* d_desc, const bjob* h_desc, int nd, std::uint64_t total_tiles, const std::uint32_t* jtab,
               bool inl, hipStream_t s, const batch_schedule& sc, const class_shapes* sh)
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
        launch_class_of(cls, d_desc + d0, inl ? h_desc + d0 : nullptr, d1 - d0, t0, t1 - t0, jtab, s, sc, sh);
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
job_class(const frsz2_job& j)
{
  return class_of(j.op, j.dtype, j.bits);
}

void
class_shape(int cls, const class_shapes* sh, int& vpt, int& wgs)
{
  vpt = class_vpt(cls);
  wgs = class_wgs(cls);
  if (sh != nullptr && cls >= 0 && cls < FRSZ2_BATCH_CLASSES) {
    if (sh->vpt[cls] != 0) {
      vpt = sh->vpt[cls];
    }
    if (sh->wgs[cls] != 0) {
      wgs = sh->wgs[cls];
    }
  }
}

int
validate_shapes(const class_shapes& sh)
{
  for (int c = 0; c < FRSZ2_BATCH_CLASSES; ++c) {
    const int v = sh.vpt[c];
    const int w = sh.wgs[c];
    const bool two = cls_dtype(c) == 1 && cls_ui(c) >= 1;
    if (!(v == 0 || v == 4 || v == 8 || v == 16 || (v == 2 && two))) {
      return set_error(FRSZ2_EARG, "class " + std::to_string(c) + ": values per thread " + std::to_string(v) +
                                     (two ? " (0, 2, 4, 8 or 16)" : " (0, 4, 8 or 16)"));
    }
    if (!(w == 0 || w == 64 || w == 256)) {
      return set_error(FRSZ2_EARG,
                       "class " + std::to_string(c) + ": workgroup size " + std::to_string(w) + " (0, 64 or 256)");
    }
  }
  return FRSZ2_OK;
}

namespace {

// true when every class runs its built-in shape under sh
bool
builtin_shapes(const class_shapes* sh)
{
  if (sh == nullptr) {
    return true;
  }
  for (int c = 0; c < FRSZ2_BATCH_CLASSES; ++c) {
    int v = 0;
    int w = 0;
    class_shape(c, sh, v, w);
    if (v != class_vpt(c) || w != class_wgs(c)) {
      return false;
    }
  }
  return true;
}

} // namespace

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

// Descriptor of job j for the tile path, its tiles starting at global tile
// tile0: the job's tile count, or -1 when the tile path cannot take it (B
// not a power of two or larger than the class's tile, pointers off 16-byte
// alignment), in which case it runs as a plain launch.
std::int64_t
describe_job(const frsz2_job& j, std::uint64_t tile0, bjob* d, const class_shapes* sh = nullptr)
{
  const int cls = class_of(j.op, j.dtype, j.bits);
  const layout L = layout::make(j.bits, j.block, j.n);
  const int lb = ilog2_exact(j.block);
  std::uint64_t nt;
  if (j.op == FRSZ2_OP_GATHER) {
    nt = (j.m + GATHER_TILE - 1) / GATHER_TILE;
  } else {
    int vpt = 0;
    int wgs = 0;
    class_shape(cls, sh, vpt, wgs);
    const std::uint64_t TILE = static_cast<std::uint64_t>(wgs) * static_cast<std::uint64_t>(vpt);
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
build_descriptors(const frsz2_job* jobs, std::size_t n, std::vector<bjob>& desc, std::vector<char>& tiled,
                  const class_shapes* sh)
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
    const std::int64_t nt = describe_job(jobs[i], tiles, &d, sh);
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

bool
descriptors_inline(const batch_schedule& sc, int ndesc, std::uint64_t total_tiles, const class_shapes* sh)
{
  // the persistent kernels read device descriptors; the inline kernels exist
  // in the built-in shapes only
  return ndesc > 0 && ndesc <= INLINE_JOBS && total_tiles > 0 && !sc.persistent && builtin_shapes(sh);
}

hipError_t
launch_batch(const bjob* d_desc, const bjob* h_desc, int ndesc, std::uint64_t total_tiles, const batch_schedule& sc,
             hipStream_t s, const std::uint32_t* d_jtab, const class_shapes* sh)
{
  if (ndesc == 0 || total_tiles == 0) {
    return hipSuccess;
  }
  return launch_grouped(d_desc, h_desc, ndesc, total_tiles, d_jtab,
                        h_desc != nullptr && descriptors_inline(sc, ndesc, total_tiles, sh), s, sc, sh);
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
  std::vector<char> tiled;            // job index -> 1 if it takes the grouped tile kernels
  std::vector<frsz2::bjob> host_desc; // descriptors, in class order
  frsz2::bjob* d_desc = nullptr;
  std::uint32_t* d_jtab = nullptr; // tile -> descriptor index within its class
  int ndesc = 0;
  std::uint64_t total_tiles = 0;
  bool uploaded = false;       // descriptors built for `shapes`
  frsz2::class_shapes shapes{}; // the tile shapes the descriptors were built with
  frsz2_batch_config cfg{};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace frsz2 {
namespace {

// Device copies of a descriptor set and its tile table (within-class
// indices), built on the host with shapes sh.
struct device_descriptors
{
  std::vector<bjob> host;
  std::vector<char> tiled;
  bjob* d_desc = nullptr;
  std::uint32_t* d_jtab = nullptr;
  std::uint64_t tiles = 0;

  void release()
  {
    if (d_desc) {
      (void)hipFree(d_desc);
      d_desc = nullptr;
    }
    if (d_jtab) {
      (void)hipFree(d_jtab);
      d_jtab = nullptr;
    }
  }

  int build(const frsz2_job* jobs, std::size_t n, const class_shapes* sh)
  {
    release();
    host.clear();
    tiles = build_descriptors(jobs, n, host, tiled, sh);
    if (host.empty()) {
      return FRSZ2_OK;
    }
    hipError_t e = hipMalloc(&d_desc, host.size() * sizeof(bjob));
    if (e != hipSuccess) {
      return hip_status(e, "hipMalloc");
    }
    e = hipMemcpy(d_desc, host.data(), host.size() * sizeof(bjob), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      return hip_status(e, "hipMemcpy");
    }
    if (tiles > 0) {
      // the tile's job by table (4 bytes per tile of at least 1 KiB of data)
      // instead of a search over thousands of jobs
      const std::vector<std::uint32_t> tab = tile_table(host, tiles, true);
      e = hipMalloc(&d_jtab, tab.size() * sizeof(std::uint32_t));
      if (e != hipSuccess) {
        return hip_status(e, "hipMalloc");
      }
      e = hipMemcpy(d_jtab, tab.data(), tab.size() * sizeof(std::uint32_t), hipMemcpyHostToDevice);
      if (e != hipSuccess) {
        return hip_status(e, "hipMemcpy");
      }
    }
    return FRSZ2_OK;
  }
};

// The tile shapes a configuration runs: its own under one workgroup per
// tile, the built-in ones under the persistent schedule.
class_shapes
effective_shapes(const frsz2_batch_config& c)
{
  class_shapes sh{};
  if (c.schedule == FRSZ2_SCHEDULE_TILES) {
    std::copy(c.values_per_thread, c.values_per_thread + FRSZ2_BATCH_CLASSES, sh.vpt);
    std::copy(c.workgroup_size, c.workgroup_size + FRSZ2_BATCH_CLASSES, sh.wgs);
  }
  return sh;
}

bool
same_shapes(const class_shapes& a, const class_shapes& b)
{
  for (int c = 0; c < FRSZ2_BATCH_CLASSES; ++c) {
    int va = 0, wa = 0, vb = 0, wb = 0;
    class_shape(c, &a, va, wa);
    class_shape(c, &b, vb, wb);
    if (va != vb || wa != wb) {
      return false;
    }
  }
  return true;
}

// build and upload the descriptors and the tile table for the tile shapes sh
int
upload_descriptors(frsz2_batch_plan* P, const class_shapes& sh)
{
  device_descriptors dd;
  dd.d_desc = P->d_desc;
  dd.d_jtab = P->d_jtab;
  P->d_desc = nullptr;
  P->d_jtab = nullptr;
  P->uploaded = false;
  const int rc = dd.build(P->jobs.data(), P->jobs.size(), &sh);
  if (rc) {
    dd.release();
    return rc;
  }
  P->host_desc = std::move(dd.host);
  P->tiled = std::move(dd.tiled);
  P->d_desc = dd.d_desc;
  P->d_jtab = dd.d_jtab;
  P->total_tiles = dd.tiles;
  P->ndesc = static_cast<int>(P->host_desc.size());
  P->shapes = sh;
  P->uploaded = true;
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
  if (!P->uploaded) {
    // a configuration whose descriptors could not be built left none
    return set_error(FRSZ2_EHIP, "batch plan has no descriptors (its last configuration failed)");
  }
  const batch_schedule sc{ P->cfg.schedule, P->cfg.workgroups, P->cfg.xcd_order };
  const hipError_t e =
    launch_batch(P->d_desc, P->host_desc.data(), P->ndesc, P->total_tiles, sc, s, P->d_jtab, &P->shapes);
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

// Takes configuration c (checked first: a bad tile shape leaves the plan as
// it was), rebuilds the descriptors when its tile shapes differ from those
// they were built with, and re-captures the graph.
int
apply_config(frsz2_batch_plan* P, const frsz2_batch_config& c)
{
  frsz2_batch_config n = c;
  if (n.schedule != FRSZ2_SCHEDULE_PERSISTENT) {
    n.schedule = FRSZ2_SCHEDULE_TILES;
  }
  if (n.workgroups < 0) {
    n.workgroups = 0;
  }
  n.xcd_order = n.xcd_order ? 1 : 0;
  class_shapes given{};
  std::copy(n.values_per_thread, n.values_per_thread + FRSZ2_BATCH_CLASSES, given.vpt);
  std::copy(n.workgroup_size, n.workgroup_size + FRSZ2_BATCH_CLASSES, given.wgs);
  int rc = validate_shapes(given);
  if (rc) {
    return rc;
  }
  P->cfg = n;
  const class_shapes sh = effective_shapes(n);
  if (!P->uploaded || !same_shapes(sh, P->shapes)) {
    drop_graph(P);
    rc = upload_descriptors(P, sh);
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
  int rc = FRSZ2_OK;
  frsz2_batch_config c{};
  if (cfg) {
    c = *cfg;
  } else {
    c.schedule = FRSZ2_SCHEDULE_TILES;
    c.workgroups = 0;
    c.xcd_order = 0;
    c.use_graph = 1;
    c.plain_launches = 0;
  }
  rc = apply_config(P, c);
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
  // time of reps runs of `run` after one warm-up, in ms
  auto timed = [&](auto&& run, float& ms) -> int {
    int rc = run();
    if (rc) {
      return rc;
    }
    (void)hipEventRecord(e0, s);
    for (int r = 0; r < reps && rc == FRSZ2_OK; ++r) {
      rc = run();
    }
    (void)hipEventRecord(e1, s);
    if (hipEventSynchronize(e1) != hipSuccess) {
      return set_error(FRSZ2_EHIP, "hipEventSynchronize failed");
    }
    (void)hipEventElapsedTime(&ms, e0, e1);
    return rc;
  };
  // 1. the tile schedules, built-in tile shapes: one workgroup per tile or a
  // persistent grid, each in workgroup order or XCD order
  frsz2_batch_config best{};
  float best_ms = -1.0f;
  int rc = FRSZ2_OK;
  for (int sched = 0; sched < 2 && rc == FRSZ2_OK; ++sched) {
    for (int xcd = 0; xcd < 2 && rc == FRSZ2_OK; ++xcd) {
      frsz2_batch_config c{};
      c.schedule = sched;
      c.workgroups = 0;
      c.xcd_order = xcd;
      c.use_graph = 0;
      c.plain_launches = 0;
      rc = apply_config(plan, c);
      if (rc) {
        break;
      }
      float ms = 0.0f;
      rc = timed([&] { return enqueue(plan, s); }, ms);
      if (rc == FRSZ2_OK && (best_ms < 0.0f || ms < best_ms)) {
        best_ms = ms;
        best = c;
      }
    }
  }
  // 2. one workgroup per tile: each class's tile shape, on that class's jobs
  // alone (the jobs the built-in shapes tile; a shape is a candidate when
  // its tile holds all of their blocks)
  if (rc == FRSZ2_OK && best.schedule == FRSZ2_SCHEDULE_TILES) {
    rc = apply_config(plan, best);
    const std::vector<char> tiled = plan->tiled; // under the built-in shapes
    const batch_schedule sc{ 0, 0, best.xcd_order };
    for (int cls = 0; cls < FRSZ2_BATCH_CLASSES && rc == FRSZ2_OK; ++cls) {
      std::vector<frsz2_job> sub;
      std::uint64_t maxB = 0;
      for (std::size_t i = 0; i < plan->jobs.size(); ++i) {
        if (tiled[i] && job_class(plan->jobs[i]) == cls) {
          sub.push_back(plan->jobs[i]);
          maxB = plan->jobs[i].block > maxB ? plan->jobs[i].block : maxB;
        }
      }
      if (sub.empty()) {
        continue;
      }
      const bool two = cls_dtype(cls) == 1 && cls_ui(cls) >= 1;
      float cls_best = -1.0f;
      for (int vpt : { 2, 4, 8, 16 }) {
        for (int wgs : { 64, 256 }) {
          if ((vpt == 2 && !two) || static_cast<std::uint64_t>(vpt) * static_cast<std::uint64_t>(wgs) < maxB) {
            continue;
          }
          class_shapes sh{};
          sh.vpt[cls] = static_cast<std::uint8_t>(vpt);
          sh.wgs[cls] = static_cast<std::uint16_t>(wgs);
          device_descriptors dd;
          rc = dd.build(sub.data(), sub.size(), &sh);
          float ms = 0.0f;
          if (rc == FRSZ2_OK) {
            rc = timed(
              [&] {
                const hipError_t e = launch_batch(dd.d_desc, dd.host.data(), static_cast<int>(dd.host.size()), dd.tiles,
                                                  sc, s, dd.d_jtab, &sh);
                return e == hipSuccess ? FRSZ2_OK : hip_status(e, "batch launch");
              },
              ms);
          }
          dd.release();
          if (rc) {
            break;
          }
          if (cls_best < 0.0f || ms < cls_best) {
            cls_best = ms;
            int bv = 0;
            int bw = 0;
            class_shape(cls, nullptr, bv, bw);
            best.values_per_thread[cls] = static_cast<unsigned char>(vpt == bv ? 0 : vpt);
            best.workgroup_size[cls] = static_cast<unsigned short>(wgs == bw ? 0 : wgs);
          }
        }
        if (rc) {
          break;
        }
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (best_ms < 0.0f) {
    best = plan->cfg; // no schedule could be timed: keep the configuration
  }
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
// frsz2_batch_run_noalloc runs a batch the way a plan with the default
// schedule does (one launch per kernel class present, one workgroup per tile,
// the plain kernels' variants), without allocating: the descriptors of the
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
    if (describe_job(jobs[i], 0, &d) > 0) {
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
      const std::int64_t nt = describe_job(j, tiles, &d);
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
        e = launch_class_of(cls, d_scratch + d0, nullptr, static_cast<int>(pos - d0), t0, tiles - t0, nullptr, s,
                            batch_schedule{ FRSZ2_SCHEDULE_TILES, 0, 0 });
      }
      if (e != hipSuccess) {
        return hip_status(e, "allocation-free batch launch");
      }
    }
  }
  // jobs outside the tile path: plain launches on the same stream
  for (size_t i = 0; i < njobs; ++i) {
    bjob d;
    if (describe_job(jobs[i], 0, &d) < 0) {
      const int rc = run_plain(jobs[i], s);
      if (rc) {
        return rc;
      }
    }
  }
  return FRSZ2_OK;
}

} // extern "C"
