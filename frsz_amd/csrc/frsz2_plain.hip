// Plain single-array FRSZ2 kernels for gfx950 and their C-ABI entry points.
//
//   compress    k_compress_tile<T,U,VPT,PACKED>   one 256-thread workgroup per
//               tile of 256*VPT values (power-of-two B <= tile), XCD-aware
//               tile order; for any other B and for pointers that are not
//               16-byte aligned, k_compress_block (one thread per block, B up
//               to 16) or k_block_exponents (block maxima) then
//               k_encode_words (one thread per payload word).
//   decompress  k_decompress_tile<...> / k_decompress_elem (one thread per value)
//   gather      k_gather (one thread per index)
//
// The arithmetic follows include/frsz2.hpp of robertu94/frsz; the value-level
// rules live in include/frsz2_device.hpp and the tile-level ones in include/frsz2_tile.hpp.
#include "frsz2_capi.h"
#include "frsz2_device.hpp"
#include "frsz2_internal.hpp"
#include "gtile.hpp"
#include "frsz2_tile.hpp"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace frsz2 {

using tile::tplan;
using tile::WG;

// ---------------------------------------------------------------- kernels

// WGS: threads per workgroup, 256 or 64 (one wave; the tile is then a
// quarter as large and a workgroup's resources free as soon as its one wave
// ends)
// VEC = false: values aligned to their size only (loaded one by one)
template<typename T, int U, int VPT, bool PACKED, bool NT, int WGS, bool VEC = true>
__global__ __launch_bounds__(WGS) void k_compress_tile(const typename fmt<T>::uw* __restrict__ in,
                                                       std::uint8_t* __restrict__ out,
                                                       tplan p,
                                                       std::uint32_t ntiles,
                                                       int linear)
{
  __shared__ std::uint32_t lds[PACKED ? tile::lds_dwords_u<U, VPT, WGS>() : 1];
  __shared__ std::uint32_t wred[WGS / tile::WAVE];
  const std::uint32_t t = linear ? blockIdx.x : tile::xcd_tile(blockIdx.x, ntiles);
  tile::compress_tile<T, U, VPT, PACKED, NT, WGS, VEC>(in, out, p, t, lds, wred);
  if (t == 0) {
    tile::zero_plan_pad<U>(out, p, static_cast<int>(threadIdx.x), WGS);
  }
}

template<typename T, int U, int VPT, bool PACKED, bool NT, bool STAGE, int WGS>
__global__ __launch_bounds__(WGS) void k_decompress_tile(const std::uint8_t* __restrict__ in,
                                                         typename fmt<T>::uw* __restrict__ out,
                                                         tplan p,
                                                         std::uint32_t ntiles,
                                                         int linear)
{
  // one array, the payload image first: it then starts at LDS address 0 and
  // the code extraction's addresses need no base (a multiple of 4 dwords, so
  // the output stage after it is 16-byte aligned)
  constexpr int LD = PACKED ? tile::lds_dwords_u<U, VPT, WGS>() : 0;
  constexpr int SD = STAGE ? tile::stage_dwords<T, VPT, WGS>() : 0;
  static_assert(LD % 4 == 0, "stage alignment");
  __shared__ __attribute__((aligned(16))) std::uint32_t smem[LD + SD > 0 ? LD + SD : 4];
  const std::uint32_t t = linear ? blockIdx.x : tile::xcd_tile(blockIdx.x, ntiles);
  tile::decompress_tile<T, U, VPT, PACKED, NT, STAGE, WGS>(in, out, p, t, smem, smem + LD);
}

// Generic compress, blocks of up to GENERIC_SMALL_B values: one thread per
// exponent block (encode_block: each payload word is built from the codes
// that overlap it, all CB bytes of the block written).  A thread's values
// are consecutive, so a wave's loads stay within a few cache lines.  The
// first 256 threads clear the zpad exponent-pad bytes after slot nb.
constexpr std::uint64_t GENERIC_SMALL_B = 16;

template<typename T, int U>
__global__ __launch_bounds__(WG) void k_compress_block(const T* __restrict__ in,
                                                       std::uint8_t* __restrict__ out,
                                                       layout L,
                                                       std::uint32_t zpad)
{
  const std::uint64_t gid0 = static_cast<std::uint64_t>(blockIdx.x) * WG + threadIdx.x;
  if (gid0 < 256) {
    std::uint8_t* q = out + L.nb * static_cast<std::uint64_t>(U / 8);
    for (std::uint32_t i = static_cast<std::uint32_t>(gid0); i < zpad; i += 256) {
      q[i] = 0;
    }
  }
  for (std::uint64_t gid = gid0; gid < L.nb; gid += static_cast<std::uint64_t>(gridDim.x) * WG) {
    const std::uint64_t first = gid * L.B;
    const std::uint64_t len = (L.n - first) < L.B ? (L.n - first) : L.B;
    encode_block<T, U>(out, L, gid, in + first, len);
  }
}

// Generic compress, larger blocks (any B, any alignment), in two passes that
// read the input with consecutive lanes on consecutive values (one thread
// per block would read a block's values serially, lanes B values apart).
//
// Pass 1: the exponent of every block (compress_cpu_block's block maximum,
// include/frsz2.hpp:1121-1128).  A group of G = 2^lg lanes (G divides the
// wave) takes one block: lane l reads values l, l + G, ... of it, and the
// key maximum is a butterfly over the group.  The first 256 threads clear the
// zpad exponent-pad bytes after slot nb.
template<typename T, int U>
__global__ __launch_bounds__(WG) void k_block_exponents(const T* __restrict__ in,
                                                        std::uint8_t* __restrict__ out,
                                                        layout L,
                                                        std::uint32_t zpad,
                                                        int lg)
{
  using uw = typename fmt<T>::uw;
  const std::uint64_t gid0 = static_cast<std::uint64_t>(blockIdx.x) * WG + threadIdx.x;
  if (gid0 < 256) {
    std::uint8_t* q = out + L.nb * static_cast<std::uint64_t>(U / 8);
    for (std::uint32_t i = static_cast<std::uint32_t>(gid0); i < zpad; i += 256) {
      q[i] = 0;
    }
  }
  const int G = 1 << lg;
  const std::uint64_t total = L.nb << lg;
  // grid-stride over workgroup-sized steps: the loop condition is uniform
  // across the workgroup, so every lane of a wave takes part in each exchange
  for (std::uint64_t base = static_cast<std::uint64_t>(blockIdx.x) * WG; base < total;
       base += static_cast<std::uint64_t>(gridDim.x) * WG) {
    const std::uint64_t gid = base + threadIdx.x;
    const std::uint64_t k = gid >> lg;
    const std::uint64_t lane = gid & static_cast<std::uint64_t>(G - 1);
    std::uint32_t m = 0;
    if (k < L.nb) {
      const std::uint64_t first = k * L.B;
      const std::uint64_t count = (L.n - first) < L.B ? (L.n - first) : L.B;
      for (std::uint64_t j = lane; j < count; j += static_cast<std::uint64_t>(G)) {
        const std::uint32_t key = max_key<T>(__builtin_bit_cast(uw, in[first + j]));
        m = m > key ? m : key;
      }
    }
    for (int o = G >> 1; o > 0; o >>= 1) {
      const std::uint32_t other = static_cast<std::uint32_t>(__shfl_xor(static_cast<int>(m), o, G));
      m = m > other ? m : other;
    }
    if (k < L.nb && lane == 0) {
      store_block_exponent<U>(out, k, block_exponent_from_key<T, U>(m));
    }
  }
}

// Pass 2: one thread per payload word, built from the codes that overlap it
// (as encode_block does: codes past the block's count are zero), with the
// block exponents of pass 1.  wpb = payload words per block.  Offsets within
// a block fit 32 bits (B <= 2^24, bits <= 64), and so does the word index
// when the stream has fewer than 2^32 payload words (narrow = true), which
// keeps the divisions 32-bit.
// payload word gid (of block gid / wpb) of the generic two-pass compress
template<typename T, int U>
__device__ __forceinline__ void
encode_word(const T* __restrict__ in, std::uint8_t* __restrict__ out, const layout& L, std::uint32_t wpb, bool narrow,
            std::uint64_t gid)
{
  using uw = typename fmt<T>::uw;
  using wt = typename word<U>::type;
  std::uint64_t k;
  std::uint32_t w;
  if (narrow) {
    const std::uint32_t g32 = static_cast<std::uint32_t>(gid);
    k = g32 / wpb;
    w = g32 - static_cast<std::uint32_t>(k) * wpb;
  } else {
    k = gid / wpb;
    w = static_cast<std::uint32_t>(gid - k * wpb);
  }
  const std::uint64_t first = k * L.B;
  const std::uint32_t count = static_cast<std::uint32_t>((L.n - first) < L.B ? (L.n - first) : L.B);
  const int E = load_block_exponent<U>(out, k);
  const std::uint32_t bits = static_cast<std::uint32_t>(L.bits);
  const T* v = in + first;
  std::uint64_t acc = 0;
  if (bits == U) {
    // one code per word
    if (w < count) {
      acc = static_cast<std::uint64_t>(encode_value<T>(__builtin_bit_cast(uw, v[w]), E, L.bits));
    }
  } else {
    const std::uint32_t lo = w * U;
    const std::uint32_t j0 = lo / bits;
    std::uint32_t j1 = (lo + U + bits - 1) / bits;
    j1 = j1 < count ? j1 : count;
    for (std::uint32_t j = j0; j < j1; ++j) {
      const std::uint64_t c = static_cast<std::uint64_t>(encode_value<T>(__builtin_bit_cast(uw, v[j]), E, L.bits));
      const int sh = static_cast<int>(j * bits) - static_cast<int>(lo);
      acc |= sh >= 0 ? (c << sh) : (c >> (-sh));
    }
  }
  reinterpret_cast<wt*>(out + L.R + k * L.CB)[w] = static_cast<wt>(acc);
}

template<typename T, int U>
__global__ __launch_bounds__(WG) void k_encode_words(const T* __restrict__ in,
                                                     std::uint8_t* __restrict__ out,
                                                     layout L,
                                                     std::uint32_t wpb,
                                                     bool narrow)
{
  const std::uint64_t words = L.nb * wpb;
  for (std::uint64_t gid = static_cast<std::uint64_t>(blockIdx.x) * WG + threadIdx.x; gid < words;
       gid += static_cast<std::uint64_t>(gridDim.x) * WG) {
    encode_word<T, U>(in, out, L, wpb, narrow, gid);
  }
}

// Generic-B tiles (gtile.hpp): one workgroup per tile of TB whole blocks;
// `align` is the alignment of the payload (16, 8, 4 or 1 bytes)
template<typename T, int U, int VPT, bool PACKED>
__global__ __launch_bounds__(gtile::WG) void k_compress_gtile(const T* __restrict__ in,
                                                              std::uint8_t* __restrict__ out,
                                                              gtile::gplan p,
                                                              int align)
{
  extern __shared__ std::uint32_t gsmem[];
  gtile::compress_gtile<T, U, VPT, PACKED, true>(in, out, p, blockIdx.x, gsmem, align);
}

template<typename T, int U, int VPT, bool PACKED>
__global__ __launch_bounds__(gtile::WG) void k_decompress_gtile(const std::uint8_t* __restrict__ in,
                                                                T* __restrict__ out,
                                                                gtile::gplan p,
                                                                int align)
{
  extern __shared__ std::uint32_t gsmem[];
  gtile::decompress_gtile<T, U, VPT, PACKED, true>(in, out, p, blockIdx.x, gsmem, align);
}

template<typename T, int U>
__global__ __launch_bounds__(WG) void k_decompress_elem(const std::uint8_t* __restrict__ in,
                                                        T* __restrict__ out,
                                                        layout L)
{
  for (std::uint64_t i = static_cast<std::uint64_t>(blockIdx.x) * WG + threadIdx.x; i < L.n;
       i += static_cast<std::uint64_t>(gridDim.x) * WG) {
    out[i] = decode_element<T, U>(in, L, i);
  }
}

template<typename T, int U>
__global__ __launch_bounds__(WG) void k_gather(const std::uint8_t* __restrict__ in,
                                               layout L,
                                               const std::uint64_t* __restrict__ idx,
                                               std::uint64_t m,
                                               T* __restrict__ out)
{
  for (std::uint64_t i = static_cast<std::uint64_t>(blockIdx.x) * WG + threadIdx.x; i < m;
       i += static_cast<std::uint64_t>(gridDim.x) * WG) {
    const std::uint64_t j = idx[i];
    out[i] = j < L.n ? decode_element<T, U, true>(in, L, j) : T(0);
  }
}

// ---------------------------------------------------------------- launchers

namespace {

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

inline unsigned
grid_for(std::uint64_t work)
{
  // the one-thread-per-item kernels stride over the grid: at most 2^22
  // workgroups (2^30 work-items, within the 32-bit work-item count of a
  // dispatch whatever the array's length)
  const std::uint64_t g = (work + WG - 1) / WG;
  return static_cast<unsigned>(g < (std::uint64_t{ 1 } << 22) ? g : (std::uint64_t{ 1 } << 22));
}

template<typename T, int U, int VPT, bool NT, int WGS, bool VEC = true>
hipError_t
launch_compress_tiled(const void* in, void* out, const layout& L, std::uint32_t zpad, int lb, int linear,
                      hipStream_t s)
{
  constexpr std::uint64_t TILE = static_cast<std::uint64_t>(WGS) * VPT;
  tplan p{ L.n, L.nb, L.R, static_cast<std::uint32_t>(L.CB), L.bits, lb, zpad };
  std::uint64_t nt = (L.n + TILE - 1) / TILE;
  if (nt == 0) {
    nt = 1;
  }
  const auto* x = static_cast<const typename fmt<T>::uw*>(in);
  auto* o = static_cast<std::uint8_t*>(out);
  if (L.bits == U) {
    hipLaunchKernelGGL((k_compress_tile<T, U, VPT, false, NT, WGS, VEC>), dim3(nt), dim3(WGS), 0, s, x, o, p,
                       static_cast<std::uint32_t>(nt), linear);
  } else {
    hipLaunchKernelGGL((k_compress_tile<T, U, VPT, true, NT, WGS, VEC>), dim3(nt), dim3(WGS), 0, s, x, o, p,
                       static_cast<std::uint32_t>(nt), linear);
  }
  return hipGetLastError();
}

template<typename T, int U, int VPT, bool NT, bool STAGE, int WGS>
hipError_t
launch_decompress_tiled(const void* in, void* out, const layout& L, int lb, int linear, hipStream_t s)
{
  constexpr std::uint64_t TILE = static_cast<std::uint64_t>(WGS) * VPT;
  tplan p{ L.n, L.nb, L.R, static_cast<std::uint32_t>(L.CB), L.bits, lb, 0u };
  const std::uint64_t nt = (L.n + TILE - 1) / TILE;
  if (nt == 0) {
    return hipSuccess;
  }
  const auto* x = static_cast<const std::uint8_t*>(in);
  auto* o = static_cast<typename fmt<T>::uw*>(out);
  if (L.bits == U) {
    hipLaunchKernelGGL((k_decompress_tile<T, U, VPT, false, NT, STAGE, WGS>), dim3(nt), dim3(WGS), 0, s, x, o, p,
                       static_cast<std::uint32_t>(nt), linear);
  } else {
    hipLaunchKernelGGL((k_decompress_tile<T, U, VPT, true, NT, STAGE, WGS>), dim3(nt), dim3(WGS), 0, s, x, o, p,
                       static_cast<std::uint32_t>(nt), linear);
  }
  return hipGetLastError();
}

int
alignment_of(const void* p)
{
  const std::uintptr_t a = reinterpret_cast<std::uintptr_t>(p);
  return (a & 15u) == 0 ? 16 : (a & 7u) == 0 ? 8 : (a & 3u) == 0 ? 4 : 1;
}

std::uint32_t
gcd32(std::uint32_t a, std::uint32_t b)
{
  while (b != 0) {
    const std::uint32_t r = a % b;
    a = b;
    b = r;
  }
  return a;
}

// Generic-tile plan of a run, or false when the tiles cannot take it (no
// values, B = 1, or a block whose aligned tile would exceed GT_MAX_VALUES).  TB is
// the largest multiple of 16 / gcd(CB, 16) blocks within `target` values
// (gtile::tile_values; at least that multiple), so every tile's payload is a
// whole number of 16-byte chunks.
bool
make_gplan(const layout& L, std::uint32_t zpad, gtile::gplan* g, std::uint32_t target)
{
  if (L.n == 0 || L.B < 2 || L.B > static_cast<std::uint64_t>(gtile::GT_MAX_VALUES)) {
    return false;
  }
  const std::uint32_t B = static_cast<std::uint32_t>(L.B);
  const std::uint32_t CB = static_cast<std::uint32_t>(L.CB);
  const std::uint32_t step = 16u / gcd32(CB, 16u);
  std::uint32_t TB = target / B / step * step;
  if (TB == 0) {
    TB = step;
  }
  if (static_cast<std::uint64_t>(TB) * B > static_cast<std::uint64_t>(gtile::GT_MAX_VALUES)) {
    return false;
  }
  *g = gtile::gplan{};
  g->n = L.n;
  g->nb = L.nb;
  g->R = L.R;
  g->B = B;
  g->CB = CB;
  g->TB = TB;
  g->V = TB * B;
  g->magic = static_cast<std::uint32_t>(((std::uint64_t{ 1 } << 32) + B - 1) / B);
  // lanes per block for the block maxima: as many as keep all of the tile's
  // blocks in one pass of the workgroup (at most 16, at most B)
  g->glg = 0;
  while (g->glg < 4 && (2u << g->glg) <= B && TB * (2u << g->glg) <= static_cast<std::uint32_t>(gtile::WG)) {
    ++g->glg;
  }
  g->bits = L.bits;
  g->zpad = zpad;
  return true;
}

template<typename T, int U, int VPT, bool PACKED>
hipError_t
launch_gtile_v(bool compress, const void* in, void* out, const gtile::gplan& g, hipStream_t s)
{
  const std::uint64_t tiles = (g.n + g.V - 1) / g.V;
  const std::size_t lds =
    static_cast<std::size_t>(compress ? gtile::compress_lds_dwords(g) : gtile::decompress_lds_dwords(g)) * 4u;
  const dim3 grid(static_cast<unsigned>(tiles));
  if (compress) {
    const int align = alignment_of(static_cast<const std::uint8_t*>(out) + g.R);
    hipLaunchKernelGGL((k_compress_gtile<T, U, VPT, PACKED>), grid, dim3(gtile::WG), lds, s,
                       static_cast<const T*>(in), static_cast<std::uint8_t*>(out), g, align);
  } else {
    const int align = alignment_of(static_cast<const std::uint8_t*>(in) + g.R);
    hipLaunchKernelGGL((k_decompress_gtile<T, U, VPT, PACKED>), grid, dim3(gtile::WG), lds, s,
                       static_cast<const std::uint8_t*>(in), static_cast<T*>(out), g, align);
  }
  return hipGetLastError();
}

// tiles of more than TARGET_VALUES values take 16 values per thread
template<typename T, int U>
hipError_t
launch_gtile(bool compress, const void* in, void* out, const gtile::gplan& g, hipStream_t s)
{
  constexpr int V8 = gtile::TARGET_VALUES / gtile::WG;
  constexpr int V16 = gtile::GT_MAX_VALUES / gtile::WG;
  const bool wide = g.V > static_cast<std::uint32_t>(gtile::TARGET_VALUES);
  const bool packed = g.bits < U;
  if (wide) {
    return packed ? launch_gtile_v<T, U, V16, true>(compress, in, out, g, s)
                  : launch_gtile_v<T, U, V16, false>(compress, in, out, g, s);
  }
  return packed ? launch_gtile_v<T, U, V8, true>(compress, in, out, g, s)
                : launch_gtile_v<T, U, V8, false>(compress, in, out, g, s);
}

template<typename T, int U>
hipError_t
launch_compress_generic(const void* in, void* out, const layout& L, std::uint32_t zpad, hipStream_t s)
{
  gtile::gplan g;
  if (make_gplan(L, zpad, &g, gtile::tile_values<T>(true, L.bits))) {
    return launch_gtile<T, U>(true, in, out, g, s);
  }
  if (L.B <= GENERIC_SMALL_B) {
    const std::uint64_t threads = L.nb < 256 ? 256 : L.nb;
    hipLaunchKernelGGL((k_compress_block<T, U>), dim3(grid_for(threads)), dim3(WG), 0, s, static_cast<const T*>(in),
                       static_cast<std::uint8_t*>(out), L, zpad);
    return hipGetLastError();
  }
  // G lanes per block: the power of two at or above B, at most a wave
  int lg = 0;
  while (lg < 6 && (std::uint64_t{ 1 } << lg) < L.B) {
    ++lg;
  }
  const std::uint64_t t1 = L.nb << lg;
  hipLaunchKernelGGL((k_block_exponents<T, U>), dim3(grid_for(t1 < 256 ? 256 : t1)), dim3(WG), 0, s,
                     static_cast<const T*>(in), static_cast<std::uint8_t*>(out), L, zpad, lg);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess || L.nb == 0) {
    return e;
  }
  const std::uint64_t wpb = L.CB / static_cast<std::uint64_t>(U / 8);
  const std::uint64_t words = L.nb * wpb;
  const bool narrow = words < (std::uint64_t{ 1 } << 32);
  hipLaunchKernelGGL((k_encode_words<T, U>), dim3(grid_for(words)), dim3(WG), 0, s, static_cast<const T*>(in),
                     static_cast<std::uint8_t*>(out), L, static_cast<std::uint32_t>(wpb), narrow);
  return hipGetLastError();
}

template<typename T, int U>
hipError_t
launch_decompress_generic(const void* in, void* out, const layout& L, hipStream_t s)
{
  if (L.n == 0) {
    return hipSuccess;
  }
  gtile::gplan g;
  if (make_gplan(L, 0, &g, gtile::tile_values<T>(false, L.bits))) {
    return launch_gtile<T, U>(false, in, out, g, s);
  }
  hipLaunchKernelGGL((k_decompress_elem<T, U>), dim3(grid_for(L.n)), dim3(WG), 0, s,
                     static_cast<const std::uint8_t*>(in), static_cast<T*>(out), L);
  return hipGetLastError();
}

template<typename T, int U>
hipError_t
launch_gather(const void* in, const layout& L, const std::uint64_t* idx, std::uint64_t m, void* out, hipStream_t s)
{
  if (m == 0) {
    return hipSuccess;
  }
  hipLaunchKernelGGL((k_gather<T, U>), dim3(grid_for(m)), dim3(WG), 0, s, static_cast<const std::uint8_t*>(in), L,
                     idx, m, static_cast<T*>(out));
  return hipGetLastError();
}

// Values per thread actually launched: the tuned one, raised to the smallest
// supported one whose tile (WGS x VPT values) holds a whole exponent block.
// 2 values per thread exists for float64 with U >= 16 only (one 16-byte load
// per lane).  0 = no tile holds the block (the generic kernels).
template<typename T, int U>
int
effective_vpt(int vpt, std::uint64_t B, int wgs)
{
  if (vpt == 2 && !(sizeof(T) == 8 && U >= 16)) {
    vpt = 4;
  }
  while (vpt <= 16 && static_cast<std::uint64_t>(wgs) * vpt < B) {
    vpt *= 2;
  }
  return vpt <= 16 ? vpt : 0;
}

// Tiled launch with the variant of the tuning table (values per thread,
// non-temporal hint, tile order, staged decompress stores, workgroup size).
// A block too large for the tuned tile grows the tile: more values per
// thread, and a 64-thread tile becomes the 256-thread one when no 64-thread
// tile holds the block.  prefer_wide (decompress): a 64-thread tile becomes
// the 256-thread one as soon as that needs fewer values per thread (float64
// bits 32 decompress at B = 1024: 64 x 16 ran at 5.2 TB/s, 256 x 8 at 6.3;
// compress measured better the other way, float32 bits 32 at B = 512: 6.5
// against 6.0).  B beyond the largest tile, or unaligned pointers, take the
// generic kernels.  Workgroups of 64 threads are built with the
// non-temporal hint only.
template<typename T, int U>
int
tiled_vpt(bool tiled, int vpt, std::uint64_t B, bool& small, bool prefer_wide)
{
  if (!tiled) {
    return 0;
  }
  int v = effective_vpt<T, U>(vpt, B, small ? 64 : WG);
  if (small) {
    const int wide = effective_vpt<T, U>(vpt, B, WG);
    if (v == 0 || (prefer_wide && wide != 0 && wide < v)) {
      small = false;
      v = wide;
    }
  }
  return v;
}

template<typename T, int U>
hipError_t
run_compress(const void* in, void* out, const layout& L, std::uint32_t zpad, bool tiled, bool elementwise, int lb,
             hipStream_t s)
{
  if (tiled && elementwise) {
    // values aligned to their size only: the 4096-value tile that loads them
    // one by one (it holds any power-of-two block up to 4096)
    if (L.B <= static_cast<std::uint64_t>(WG) * 16) {
      return launch_compress_tiled<T, U, 16, true, WG, false>(in, out, L, zpad, lb, 0, s);
    }
    return launch_compress_generic<T, U>(in, out, L, zpad, s);
  }
  const variant v = get_variant(OP_COMPRESS, sizeof(T) == 4 ? FRSZ2_FLOAT32 : FRSZ2_FLOAT64, L.bits);
  bool small = v.small != 0;
  int vpt = tiled_vpt<T, U>(tiled, v.vpt, L.B, small, false);
  if (vpt > 0) {
#define FRSZ2_CL(V, N, W) return launch_compress_tiled<T, U, V, N, W>(in, out, L, zpad, lb, v.linear, s)
    if (small) {
      if constexpr (sizeof(T) == 8 && U >= 16) {
        if (vpt == 2) {
          FRSZ2_CL(2, true, 64);
        }
      }
      vpt = vpt < 4 ? 4 : vpt;
      switch (vpt) {
        case 4: FRSZ2_CL(4, true, 64);
        case 8: FRSZ2_CL(8, true, 64);
        default: FRSZ2_CL(16, true, 64);
      }
    }
    if constexpr (sizeof(T) == 8 && U >= 16) {
      if (vpt == 2) {
        if (v.nt) {
          FRSZ2_CL(2, true, WG);
        }
        FRSZ2_CL(2, false, WG);
      }
    }
    switch (vpt * 2 + v.nt) {
      case 8: FRSZ2_CL(4, false, WG);
      case 9: FRSZ2_CL(4, true, WG);
      case 16: FRSZ2_CL(8, false, WG);
      case 17: FRSZ2_CL(8, true, WG);
      case 32: FRSZ2_CL(16, false, WG);
      default: FRSZ2_CL(16, true, WG);
    }
#undef FRSZ2_CL
  }
  return launch_compress_generic<T, U>(in, out, L, zpad, s);
}

template<typename T, int U>
hipError_t
run_decompress(const void* in, void* out, const layout& L, bool tiled, bool strided, int lb, hipStream_t s)
{
  const variant v = get_variant(OP_DECOMPRESS, sizeof(T) == 4 ? FRSZ2_FLOAT32 : FRSZ2_FLOAT64, L.bits);
  bool small = v.small != 0;
  int vpt = tiled_vpt<T, U>(tiled, v.vpt, L.B, small, true);
  // strided: the decode that stores single values (any value alignment).
  // A float64 tile grown to 16 values per thread to hold its block takes it
  // too: the staged decode at 16 float64 values per thread (128 bytes of
  // output per lane) measured 4.1-5.2 TB/s at B 512-4096, the strided one
  // 6.0-6.1.
  const int stage = strided || (sizeof(T) == 8 && vpt == 16 && v.vpt < 16) ? 0 : v.stage;
  if (vpt > 0) {
#define FRSZ2_DL(V, N, S, W) return launch_decompress_tiled<T, U, V, N, S, W>(in, out, L, lb, v.linear, s)
    if (small) {
      if constexpr (sizeof(T) == 8 && U >= 16) {
        if (vpt == 2) {
          if (stage) {
            FRSZ2_DL(2, true, true, 64);
          }
          FRSZ2_DL(2, true, false, 64);
        }
      }
      vpt = vpt < 4 ? 4 : vpt;
      switch (vpt * 2 + stage) {
        case 8: FRSZ2_DL(4, true, false, 64);
        case 9: FRSZ2_DL(4, true, true, 64);
        case 16: FRSZ2_DL(8, true, false, 64);
        case 17: FRSZ2_DL(8, true, true, 64);
        case 32: FRSZ2_DL(16, true, false, 64);
        default: FRSZ2_DL(16, true, true, 64);
      }
    }
    if constexpr (sizeof(T) == 8 && U >= 16) {
      if (vpt == 2) {
        switch (stage * 2 + v.nt) {
          case 0: FRSZ2_DL(2, false, false, WG);
          case 1: FRSZ2_DL(2, true, false, WG);
          case 2: FRSZ2_DL(2, false, true, WG);
          default: FRSZ2_DL(2, true, true, WG);
        }
      }
    }
    switch (vpt * 4 + stage * 2 + v.nt) {
      case 16: FRSZ2_DL(4, false, false, WG);
      case 17: FRSZ2_DL(4, true, false, WG);
      case 18: FRSZ2_DL(4, false, true, WG);
      case 19: FRSZ2_DL(4, true, true, WG);
      case 32: FRSZ2_DL(8, false, false, WG);
      case 33: FRSZ2_DL(8, true, false, WG);
      case 34: FRSZ2_DL(8, false, true, WG);
      case 35: FRSZ2_DL(8, true, true, WG);
      case 64: FRSZ2_DL(16, false, false, WG);
      case 65: FRSZ2_DL(16, true, false, WG);
      case 66: FRSZ2_DL(16, false, true, WG);
      default: FRSZ2_DL(16, true, true, WG);
    }
#undef FRSZ2_DL
  }
  return launch_decompress_generic<T, U>(in, out, L, s);
}

template<template<typename, int> class Op, typename... A>
hipError_t
by_type(int dtype, int U, A&&... a)
{
  if (dtype == FRSZ2_FLOAT32) {
    switch (U) {
      case 8:
        return Op<float, 8>::run(a...);
      case 16:
        return Op<float, 16>::run(a...);
      default:
        return Op<float, 32>::run(a...);
    }
  }
  switch (U) {
    case 8:
      return Op<double, 8>::run(a...);
    case 16:
      return Op<double, 16>::run(a...);
    case 32:
      return Op<double, 32>::run(a...);
    default:
      return Op<double, 64>::run(a...);
  }
}

template<typename T, int U>
struct compress_op
{
  static hipError_t run(const void* in, void* out, const layout& L, std::uint32_t zpad, bool tiled,
                        bool elementwise, int lb, hipStream_t s)
  {
    return run_compress<T, U>(in, out, L, zpad, tiled, elementwise, lb, s);
  }
};

template<typename T, int U>
struct decompress_op
{
  static hipError_t run(const void* in, void* out, const layout& L, bool tiled, bool strided, int lb, hipStream_t s)
  {
    return run_decompress<T, U>(in, out, L, tiled, strided, lb, s);
  }
};

template<typename T, int U>
struct gather_op
{
  static hipError_t run(const void* in, const layout& L, const std::uint64_t* idx, std::uint64_t m, void* out,
                        hipStream_t s)
  {
    return launch_gather<T, U>(in, L, idx, m, out, s);
  }
};

inline bool
aligned(const void* p, std::uintptr_t a)
{
  return (reinterpret_cast<std::uintptr_t>(p) & (a - 1)) == 0;
}

} // namespace

// ---------------------------------------------------------------- internal API

int
check_params(int dtype, int bits, std::uint64_t block)
{
  if (dtype != FRSZ2_FLOAT32 && dtype != FRSZ2_FLOAT64) {
    return set_error(FRSZ2_EDTYPE, "unsupported dtype");
  }
  const int W = dtype == FRSZ2_FLOAT32 ? 32 : 64;
  if (bits < 4 || bits > W) {
    return set_error(FRSZ2_EPARAM, "Unsupported number of bits: " + std::to_string(bits));
  }
  if (block == 0 || block > FRSZ2_MAX_BLOCK) {
    return set_error(FRSZ2_EPARAM, "Unsupported exponent block size: " + std::to_string(block));
  }
  return FRSZ2_OK;
}

int
hip_status(hipError_t e, const char* what)
{
  if (e != hipSuccess) {
    return set_error(FRSZ2_EHIP, std::string(what) + ": " + hipGetErrorString(e));
  }
  return FRSZ2_OK;
}

run_plan
make_run(int bits, std::uint64_t block, std::uint64_t n, std::uint64_t first, std::uint64_t count)
{
  const layout G = layout::make(bits, block, n);
  const std::uint64_t es = static_cast<std::uint64_t>(G.U / 8);
  run_plan r{};
  r.L = G;
  r.L.n = count;
  r.L.nb = count / block + (count % block != 0);
  r.exp_off = first * es;
  r.pay_off = G.R + first * G.CB;
  r.L.R = r.pay_off - r.exp_off;
  r.zpad = first + r.L.nb == G.nb ? static_cast<std::uint32_t>(G.R - G.nb * es) : 0u;
  return r;
}

int
check_run(int dtype, int bits, std::uint64_t block, std::uint64_t n, std::uint64_t first, std::uint64_t count)
{
  int rc = check_params(dtype, bits, block);
  if (rc) {
    return rc;
  }
  const std::uint64_t nb = n / block + (n % block != 0);
  const std::uint64_t room = first >= nb ? 0 : n - first * block;
  if (first > nb || count > room) {
    return set_error(FRSZ2_EARG, "blocks from " + std::to_string(first) + " holding " + std::to_string(count) +
                                   " values exceed the " + std::to_string(n) + "-value array");
  }
  return FRSZ2_OK;
}

int
compress_run(int dtype, const void* d_values, void* d_out, const layout& L, std::uint32_t zpad, hipStream_t s)
{
  if (L.n == 0 && zpad == 0) {
    return FRSZ2_OK;
  }
  const int lb = ilog2_exact(L.B);
  // the tiles write the payload with 16-byte stores; values off 16-byte
  // alignment are loaded one by one
  const bool tiled = lb >= 0 && aligned(static_cast<const std::uint8_t*>(d_out) + L.R, 16);
  const bool elementwise = !aligned(d_values, 16);
  return hip_status(by_type<compress_op>(dtype, L.U, d_values, d_out, L, zpad, tiled, elementwise, lb, s),
                    "compress launch");
}

int
decompress_run(int dtype, const void* d_in, void* d_values, const layout& L, hipStream_t s)
{
  if (L.n == 0) {
    return FRSZ2_OK;
  }
  const int lb = ilog2_exact(L.B);
  const bool vals16 = aligned(d_values, 16);
  const bool pay16 = aligned(static_cast<const std::uint8_t*>(d_in) + L.R, 16);
  // Values off 16-byte alignment, or an unpacked payload off it, still take
  // the tiles through the strided decode: it stores single values and reads
  // an unpacked payload word by word.  A packed payload is staged with
  // 16-byte loads, so it needs the alignment.
  const bool packed = L.bits < L.U;
  const bool tiled = lb >= 0 && (pay16 || !packed);
  const bool strided = !(vals16 && pay16);
  return hip_status(by_type<decompress_op>(dtype, L.U, d_in, d_values, L, tiled, strided, lb, s),
                    "decompress launch");
}

// Blocks per launch of a long run: a whole multiple of 16 blocks (each
// launch's payload then keeps the 16-byte alignment of the whole run's, and
// its values the alignment of the run's values) holding at most 2^30 values,
// so no launch comes near the 2^32 work-items a dispatch can count.
std::uint64_t
launch_blocks(std::uint64_t block)
{
  const std::uint64_t b = (std::uint64_t{ 1 } << 30) / block / 16 * 16;
  return b > 16 ? b : 16;
}

int
compress_blocks_device(int dtype, int bits, std::uint64_t block, std::uint64_t n, std::uint64_t first,
                       const void* d_values, std::uint64_t count, void* d_stream, std::uint64_t stream_bytes,
                       hipStream_t s)
{
  int rc = check_run(dtype, bits, block, n, first, count);
  if (rc) {
    return rc;
  }
  const std::uint64_t total = layout::make(bits, block, n).total_bytes();
  if (stream_bytes < total) {
    return set_error(FRSZ2_ECAPACITY,
                     "output capacity " + std::to_string(stream_bytes) + " < stream size " + std::to_string(total));
  }
  const std::uintptr_t vs = dtype == FRSZ2_FLOAT32 ? 4 : 8;
  if (!d_stream || (count > 0 && !d_values) || !aligned(d_values, vs) || !aligned(d_stream, 8)) {
    return set_error(FRSZ2_EARG, "null or misaligned pointer");
  }
  const std::uint64_t step = launch_blocks(block);
  std::uint64_t k = first;
  std::uint64_t left = count;
  do {
    const std::uint64_t take = left < step * block ? left : step * block;
    const run_plan r = make_run(bits, block, n, k, take);
    rc = compress_run(dtype, static_cast<const std::uint8_t*>(d_values) + (k - first) * block * vs,
                      static_cast<std::uint8_t*>(d_stream) + r.exp_off, r.L, r.zpad, s);
    if (rc) {
      return rc;
    }
    k += step;
    left -= take;
  } while (left > 0);
  return FRSZ2_OK;
}

int
decompress_blocks_device(int dtype, int bits, std::uint64_t block, std::uint64_t n, std::uint64_t first,
                         const void* d_stream, std::uint64_t stream_bytes, std::uint64_t count, void* d_values,
                         hipStream_t s)
{
  int rc = check_run(dtype, bits, block, n, first, count);
  if (rc) {
    return rc;
  }
  const std::uint64_t total = layout::make(bits, block, n).total_bytes();
  if (stream_bytes < total) {
    return set_error(FRSZ2_ECAPACITY,
                     "input stream of " + std::to_string(stream_bytes) + " bytes < stream size " + std::to_string(total));
  }
  const std::uintptr_t vs = dtype == FRSZ2_FLOAT32 ? 4 : 8;
  if (!d_stream || (count > 0 && !d_values) || !aligned(d_stream, 8) || !aligned(d_values, vs)) {
    return set_error(FRSZ2_EARG, "null or misaligned pointer");
  }
  const std::uint64_t step = launch_blocks(block);
  std::uint64_t k = first;
  std::uint64_t left = count;
  do {
    const std::uint64_t take = left < step * block ? left : step * block;
    const run_plan r = make_run(bits, block, n, k, take);
    rc = decompress_run(dtype, static_cast<const std::uint8_t*>(d_stream) + r.exp_off,
                        static_cast<std::uint8_t*>(d_values) + (k - first) * block * vs, r.L, s);
    if (rc) {
      return rc;
    }
    k += step;
    left -= take;
  } while (left > 0);
  return FRSZ2_OK;
}

int
compress_device(int dtype, int bits, std::uint64_t block, const void* d_in, std::uint64_t n, void* d_out,
                std::uint64_t out_capacity, hipStream_t s)
{
  return compress_blocks_device(dtype, bits, block, n, 0, d_in, n, d_out, out_capacity, s);
}

int
decompress_device(int dtype, int bits, std::uint64_t block, const void* d_in, std::uint64_t in_bytes,
                  std::uint64_t n, void* d_out, hipStream_t s)
{
  return decompress_blocks_device(dtype, bits, block, n, 0, d_in, in_bytes, n, d_out, s);
}

int
gather_device(int dtype, int bits, std::uint64_t block, const void* d_in, std::uint64_t in_bytes, std::uint64_t n,
              const std::uint64_t* d_idx, std::uint64_t m, void* d_out, hipStream_t s)
{
  int rc = check_params(dtype, bits, block);
  if (rc) {
    return rc;
  }
  const layout L = layout::make(bits, block, n);
  if (in_bytes < L.total_bytes()) {
    return set_error(FRSZ2_ECAPACITY, "input stream of " + std::to_string(in_bytes) + " bytes < stream size " +
                                        std::to_string(L.total_bytes()));
  }
  const std::uintptr_t vs = dtype == FRSZ2_FLOAT32 ? 4 : 8;
  if (m > 0 && (!d_in || !d_idx || !d_out || !aligned(d_in, 8) || !aligned(d_out, vs) || !aligned(d_idx, 8))) {
    return set_error(FRSZ2_EARG, "null or misaligned pointer");
  }
  return hip_status(by_type<gather_op>(dtype, L.U, d_in, L, d_idx, m, d_out, s), "gather launch");
}

} // namespace frsz2

// ---------------------------------------------------------------- C ABI

extern "C" {

const char*
frsz2_version(void)
{
  return "frsz2-mi355x 0.1.0 (stream format of frsz2 plugin 0.0.2)";
}

int
frsz2_check(int dtype, int bits, uint64_t block)
{
  frsz2::clear_error();
  return frsz2::check_params(dtype, bits, block);
}

int
frsz2_compressed_size(int dtype, int bits, uint64_t block, uint64_t n, uint64_t* bytes)
{
  frsz2::clear_error();
  const int rc = frsz2::check_params(dtype, bits, block);
  if (rc) {
    return rc;
  }
  if (!bytes) {
    return frsz2::set_error(FRSZ2_EARG, "null output pointer");
  }
  *bytes = frsz2::layout::make(bits, block, n).total_bytes();
  return FRSZ2_OK;
}

int
frsz2_compress(int dtype, int bits, uint64_t block, const void* d_in, uint64_t n, void* d_out,
               uint64_t out_capacity, frsz2_stream_t stream)
{
  frsz2::clear_error();
  return frsz2::compress_device(dtype, bits, block, d_in, n, d_out, out_capacity,
                                static_cast<hipStream_t>(stream));
}

int
frsz2_decompress(int dtype, int bits, uint64_t block, const void* d_in, uint64_t in_bytes, uint64_t n,
                 void* d_out, frsz2_stream_t stream)
{
  frsz2::clear_error();
  return frsz2::decompress_device(dtype, bits, block, d_in, in_bytes, n, d_out, static_cast<hipStream_t>(stream));
}

int
frsz2_gather(int dtype, int bits, uint64_t block, const void* d_in, uint64_t in_bytes, uint64_t n,
             const uint64_t* d_idx, uint64_t m, void* d_out, frsz2_stream_t stream)
{
  frsz2::clear_error();
  return frsz2::gather_device(dtype, bits, block, d_in, in_bytes, n, d_idx, m, d_out,
                              static_cast<hipStream_t>(stream));
}

int
frsz2_compress_blocks(int dtype, int bits, uint64_t block, uint64_t n, uint64_t first_block, const void* d_values,
                      uint64_t count, void* d_stream, uint64_t stream_bytes, frsz2_stream_t stream)
{
  frsz2::clear_error();
  return frsz2::compress_blocks_device(dtype, bits, block, n, first_block, d_values, count, d_stream, stream_bytes,
                                       static_cast<hipStream_t>(stream));
}

int
frsz2_decompress_blocks(int dtype, int bits, uint64_t block, uint64_t n, uint64_t first_block,
                        const void* d_stream, uint64_t stream_bytes, uint64_t count, void* d_values,
                        frsz2_stream_t stream)
{
  frsz2::clear_error();
  return frsz2::decompress_blocks_device(dtype, bits, block, n, first_block, d_stream, stream_bytes, count,
                                         d_values, static_cast<hipStream_t>(stream));
}

} // extern "C"
