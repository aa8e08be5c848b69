// Tiles of whole exponent blocks for any block size B (gfx950): the path for
// block sizes that are not powers of two, and for payloads off 16-byte
// alignment, where the power-of-two tiles (include/frsz2_tile.hpp) do not apply.
//
// A tile is TB whole exponent blocks, V = TB*B values (at most GT_MAX_VALUES),
// with TB chosen on the host so that TB*CB is a multiple of 16 bytes: every
// tile's payload then starts at the same alignment as the stream's payload.
// Thread t handles tile values t, t + WG, ... (consecutive lanes on
// consecutive values, so every wave load / store is one contiguous span).
// The block of tile value i is i / B, by one multiply-high with a magic
// number (exact for i < 2^16, 2 <= B <= 2^16).  Lanes past the tile's values
// (V is often below VPT*WG; the last tile is short) take the tile's last
// value for their loads and arithmetic, so the per-value code has no
// branches, and only their LDS atomics and stores are masked off.
//
// compress:   load -> block maxima (LDS atomics for B < 6; otherwise keys
//             into LDS and one block per group of up to 16 lanes, a DPP
//             butterfly) -> block exponents ->
//             codes OR-ed into an LDS image of the tile's payload at their bit
//             offsets -> exponents and the image written out (the widest
//             stores the payload alignment allows).  Same bytes as
//             compress_cpu_block (include/frsz2.hpp:1108-1179).
// decompress: the tile's payload staged into LDS with wide loads, its block
//             exponents into LDS, then per value: extract the code, decode,
//             store.  Same values as decompress_cpu_block (:1217-1240).
#pragma once

#include "frsz2_device.hpp"
#include "frsz2_tile.hpp"

#include <hip/hip_runtime.h>

#include <cstdint>

namespace frsz2 {
namespace gtile {

constexpr int WG = 256;
constexpr int GT_MAX_VALUES = 4096;
// tiles of up to TARGET_VALUES values take VPT = TARGET_VALUES / WG values
// per thread, larger ones (TB rounded up for alignment) GT_MAX_VALUES / WG
constexpr int TARGET_VALUES = 2048;

// Values a run's tiles aim for.  Per-tile work (clearing and writing out the
// payload image, the barriers, the block-maximum passes) and the bytes one
// workgroup keeps in flight favour large tiles, until the register and LDS
// footprint of a tile moving much more than about 20 KB costs more than it
// saves.  Measured on 10^8-value float32 and 5*10^7-value float64 data at
// B 5-100: float32 below 32 bits, and float64 compress up to 16 bits, run
// 4096-value tiles (float32 bits 8, B 9: compress 4.8 -> 5.6 TB/s,
// decompress 4.3 -> 6.0); float32 at 32 bits and the other float64 cases
// were faster with 2048, and 8192-value float32 tiles 3-30% slower than
// 4096.
template<typename T>
constexpr std::uint32_t
tile_values(bool compress, int bits)
{
  return (sizeof(T) == 4 ? bits < 32 : compress && bits <= 16) ? static_cast<std::uint32_t>(GT_MAX_VALUES)
                                                                : static_cast<std::uint32_t>(TARGET_VALUES);
}

// per-run parameters (host-computed)
struct gplan
{
  std::uint64_t n;      // values of the run
  std::uint64_t nb;     // exponent blocks of the run
  std::uint64_t R;      // payload offset from the run's exponent base
  std::uint32_t B;      // block size
  std::uint32_t CB;     // payload bytes per block
  std::uint32_t TB;     // blocks per tile
  std::uint32_t V;      // values per tile (TB * B)
  std::uint32_t magic;  // ceil(2^32 / B), B >= 2
  int glg;              // log2 of the lanes that reduce one block's maximum (2^glg <= min(16, B))
  int bits;
  std::uint32_t zpad;   // exponent-pad bytes after slot nb to clear (tile 0)
};

// block of tile value i (i < 2^16): one multiply-high, exact for every
// B >= 2 (for powers of two the magic number is exactly 2^32 / B)
__device__ __forceinline__ std::uint32_t
block_of(std::uint32_t i, const gplan& p)
{
  return __umulhi(i, p.magic);
}

// LDS dwords of a tile's payload image, plus slack for the reads and ORs
// that touch the dwords after a code's last one
__host__ __device__ inline std::uint32_t
image_dwords(const gplan& p)
{
  return p.TB * p.CB / 4 + 4;
}

// copy nbytes between global memory at g (aligned like the stream payload:
// `align` = 16, 8 or 4 bytes, or 1) and the LDS image, all threads of the
// workgroup taking part
template<bool TO_LDS, bool NT>
__device__ __forceinline__ void
move_payload(std::uint8_t* g, std::uint32_t* lds, std::uint32_t nbytes, int align)
{
  using tile::u32x4;
  typedef std::uint32_t u32x2 __attribute__((ext_vector_type(2)));
  const std::uint32_t t = threadIdx.x;
  std::uint32_t done = 0;
  if (align >= 16) {
    u32x4* g4 = reinterpret_cast<u32x4*>(g);
    u32x4* l4 = reinterpret_cast<u32x4*>(lds);
    for (std::uint32_t c = t; c < nbytes / 16; c += WG) {
      if constexpr (TO_LDS) {
        l4[c] = tile::ld<NT>(g4 + c);
      } else {
        tile::st<NT>(l4[c], g4 + c);
      }
    }
    done = nbytes & ~15u;
  } else if (align >= 8) {
    u32x2* g2 = reinterpret_cast<u32x2*>(g);
    u32x2* l2 = reinterpret_cast<u32x2*>(lds);
    for (std::uint32_t c = t; c < nbytes / 8; c += WG) {
      if constexpr (TO_LDS) {
        l2[c] = tile::ld<NT>(g2 + c);
      } else {
        tile::st<NT>(l2[c], g2 + c);
      }
    }
    done = nbytes & ~7u;
  } else if (align >= 4) {
    std::uint32_t* g1 = reinterpret_cast<std::uint32_t*>(g);
    for (std::uint32_t c = t; c < nbytes / 4; c += WG) {
      if constexpr (TO_LDS) {
        lds[c] = g1[c];
      } else {
        g1[c] = lds[c];
      }
    }
    done = nbytes & ~3u;
  }
  std::uint8_t* lb = reinterpret_cast<std::uint8_t*>(lds);
  for (std::uint32_t o = done + t; o < nbytes; o += WG) {
    if constexpr (TO_LDS) {
      lb[o] = g[o];
    } else {
      g[o] = lb[o];
    }
  }
}

// LDS of a decompress tile: the payload image, then (from a 16-byte aligned
// dword offset) the tile's TB block exponents as they lie in the stream,
// TB U/8 <= 2 TB dwords
__host__ __device__ inline std::uint32_t
exponent_offset(const gplan& p)
{
  return (image_dwords(p) + 3u) & ~3u;
}

__host__ __device__ inline std::uint32_t
decompress_lds_dwords(const gplan& p)
{
  return exponent_offset(p) + 2u * p.TB;
}

// largest of 16, 8, 4 and 1 that divides the address
__device__ __forceinline__ int
alignment_at(const void* a)
{
  const std::uintptr_t v = reinterpret_cast<std::uintptr_t>(a);
  return (v & 15u) == 0 ? 16 : (v & 7u) == 0 ? 8 : (v & 3u) == 0 ? 4 : 1;
}

// LDS dwords of a compress tile: payload image, TB block keys, V 16-bit value keys
__host__ __device__ inline std::uint32_t
compress_lds_dwords(const gplan& p)
{
  return image_dwords(p) + p.TB + (p.V + 1) / 2;
}

// Compress tile `tile` of a run: exponents at out, payload at out + p.R.
// LDS: compress_lds_dwords(p).  PACKED: bits < U (codes OR-ed into the
// image at their bit offsets); otherwise the image is an array of U-bit
// words, word i holding value i's code (CB = B*U/8 then, no block padding).
template<typename T, int U, int VPT, bool PACKED, bool NT>
__device__ __forceinline__ void
compress_gtile(const T* __restrict__ in, std::uint8_t* __restrict__ out, const gplan& p, std::uint64_t tile,
               std::uint32_t* smem, int align)
{
  using uw = typename fmt<T>::uw;
  using wt = typename word<U>::type;
  const std::uint32_t t = threadIdx.x;
  std::uint32_t* img = smem;
  const std::uint32_t idw = image_dwords(p);
  std::uint32_t* keys = smem + idw;
  const std::uint64_t k0 = tile * p.TB;
  const std::uint64_t base = k0 * p.B;
  const std::uint64_t left = p.n - base;
  const std::uint32_t cnt = left < p.V ? static_cast<std::uint32_t>(left) : p.V;
  const std::uint32_t nblk = block_of(cnt - 1, p) + 1;
  for (std::uint32_t i = t; i < idw; i += WG) {
    img[i] = 0;
  }
  const T* src = in + base;
  uw x[VPT];
  std::uint32_t ic[VPT]; // the value index a lane works on (clamped to the tile's last)
#pragma unroll
  for (int r = 0; r < VPT; ++r) {
    const std::uint32_t i = t + static_cast<std::uint32_t>(r * WG);
    ic[r] = i < cnt ? i : cnt - 1;
    x[r] = tile::ld<NT>(reinterpret_cast<const uw*>(src) + ic[r]);
  }
  // Block maxima into keys.  A wave's 64 values are consecutive.  For B < 6
  // its lanes spread over many keys, and each lane atomically max-es its key
  // in.  For larger B, where many lanes would hit one address and
  // same-address LDS atomics serialize (measured: B = 7 ran 2-9% faster on
  // the key array below, B = 5 9-12% slower), every value's key goes to an LDS
  // array (16 bits each, consecutive lanes on consecutive entries), and
  // aligned groups of G = 2^glg lanes (G <= 16, at most B, chosen so that
  // one pass of the workgroup covers the tile's blocks where it can) take
  // one block each: lane l of a group reads entries l, l + G, ... of its
  // block, and a DPP butterfly within the group gives the block maximum.
  if (p.B < 6) {
    for (std::uint32_t i = t; i < nblk; i += WG) {
      keys[i] = 0;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < VPT; ++r) {
      if (t + static_cast<std::uint32_t>(r * WG) < cnt) {
        atomicMax(&keys[block_of(ic[r], p)], max_key<T>(x[r]));
      }
    }
  } else {
    std::uint16_t* kv = reinterpret_cast<std::uint16_t*>(keys + p.TB);
#pragma unroll
    for (int r = 0; r < VPT; ++r) {
      if (t + static_cast<std::uint32_t>(r * WG) < cnt) {
        kv[ic[r]] = static_cast<std::uint16_t>(max_key<T>(x[r]));
      }
    }
    __syncthreads();
    const int G = 1 << p.glg;
    const std::uint32_t l = t & static_cast<std::uint32_t>(G - 1);
    const std::uint32_t groups = static_cast<std::uint32_t>(WG) >> p.glg;
    for (std::uint32_t b0 = 0; b0 < nblk; b0 += groups) {
      const std::uint32_t b = b0 + (t >> p.glg);
      std::uint32_t m = 0;
      if (b < nblk) {
        const std::uint32_t first = b * p.B;
        const std::uint32_t end = first + p.B < cnt ? first + p.B : cnt;
        for (std::uint32_t j = first + l; j < end; j += static_cast<std::uint32_t>(G)) {
          const std::uint32_t v = kv[j];
          m = m > v ? m : v;
        }
      }
      m = tile::group_max(m, G, nullptr); // G <= 16: DPP steps within the group only
      if (l == 0 && b < nblk) {
        keys[b] = m;
      }
    }
  }
  __syncthreads();
  const bool narrow = p.bits <= fmt<T>::SB + 1;
  const std::uint32_t bits = static_cast<std::uint32_t>(p.bits);
  // bit offset of tile value i in block k = i / B: i*bits plus the padding
  // bits of the k blocks before it, k*(CB*8 - B*bits)
  const std::uint32_t padbits = p.CB * 8u - p.B * bits;
#pragma unroll
  for (int r = 0; r < VPT; ++r) {
    const std::uint32_t i = ic[r];
    const bool own = t + static_cast<std::uint32_t>(r * WG) < cnt;
    const std::uint32_t k = block_of(i, p);
    const int E = block_exponent_from_key<T, U>(keys[k]);
    const uw code = narrow ? encode_value_short<T>(x[r], E, p.bits) : encode_value<T>(x[r], E, p.bits);
    if constexpr (PACKED) {
      // a zero code leaves the zeroed image as it is; skipping it saves the
      // LDS atomics (codes below the block's top bits are often zero)
      if (own && code != 0) {
        tile::or_code<U>(img, i * bits + k * padbits, static_cast<std::uint64_t>(code), p.bits);
      }
    } else if (own) {
      reinterpret_cast<wt*>(img)[i] = static_cast<wt>(code);
    }
  }
  for (std::uint32_t b = t; b < nblk; b += WG) {
    store_block_exponent<U>(out, k0 + b, block_exponent_from_key<T, U>(keys[b]));
  }
  __syncthreads();
  move_payload<false, NT>(out + p.R + k0 * p.CB, img, nblk * p.CB, align);
  if (tile == 0) {
    tile::zero_plan_pad<U>(out, tile::tplan{ p.n, p.nb, p.R, p.CB, p.bits, 0, p.zpad }, static_cast<int>(t), WG);
  }
}

// Decode of unpacked codes (bits = U <= 32) four consecutive values at a
// time: chunk c of thread t holds tile values 4 (c WG + t) .. + 3, so one
// LDS read of 4 U/8 bytes gives a chunk's codes and a wave's chunk stores
// are one contiguous span (16-byte stores when the tile's output is 16-byte
// aligned).  For bits = U a block's payload is its B codes with no padding
// (CB = B U / 8), so code i is word i of the image, and V = TB B is a
// multiple of 4 (TB CB is a multiple of 16 bytes), so every chunk that
// starts below cnt reads inside the image.  Positions at or past cnt take
// cnt - 1 for their exponent and are not stored.
template<typename T, int U, int NC, bool NT>
__device__ __forceinline__ void
decode_chunks(const std::uint32_t* img, const std::uint8_t* exb, T* __restrict__ o, std::uint32_t cnt,
              const gplan& p)
{
  static_assert(U <= 32, "chunks of four codes are read as at most 16 bytes");
  using uw = typename fmt<T>::uw;
  const std::uint32_t t = threadIdx.x;
  const std::uint32_t last = cnt - 1;
  uw code[NC][4];
  int E[NC][4];
  int emin = 0x7fffffff;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const std::uint32_t c0 = 4u * (t + static_cast<std::uint32_t>(c * WG));
    const std::uint32_t i0 = c0 <= last ? c0 : (last & ~3u);
    if constexpr (U == 8) {
      const std::uint32_t w = img[i0 / 4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        code[c][q] = static_cast<uw>((w >> (8 * q)) & 0xFFu);
      }
    } else if constexpr (U == 16) {
      const std::uint32_t w0 = img[i0 / 2];
      const std::uint32_t w1 = img[i0 / 2 + 1];
      code[c][0] = static_cast<uw>(w0 & 0xFFFFu);
      code[c][1] = static_cast<uw>(w0 >> 16);
      code[c][2] = static_cast<uw>(w1 & 0xFFFFu);
      code[c][3] = static_cast<uw>(w1 >> 16);
    } else {
      const tile::u32x4 w = *reinterpret_cast<const tile::u32x4*>(img + i0);
      code[c][0] = static_cast<uw>(w.x);
      code[c][1] = static_cast<uw>(w.y);
      code[c][2] = static_cast<uw>(w.z);
      code[c][3] = static_cast<uw>(w.w);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const std::uint32_t i = i0 + static_cast<std::uint32_t>(q);
      E[c][q] = load_block_exponent<U>(exb, block_of(i <= last ? i : last, p));
      emin = E[c][q] < emin ? E[c][q] : emin;
    }
  }
  // decode_value_short needs bits <= SB + 1 and E + bias >= 0 (every block a
  // compress writes); checked over all of the thread's values
  const bool fast = p.bits <= fmt<T>::SB + 1 && emin + fmt<T>::BIAS >= 0;
  const bool vec = (reinterpret_cast<std::uintptr_t>(o) & 15u) == 0;
  uw* ow = reinterpret_cast<uw*>(o);
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const std::uint32_t c0 = 4u * (t + static_cast<std::uint32_t>(c * WG));
    if (c0 > last) {
      continue;
    }
    uw y[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      y[q] = fast ? decode_value_short<T>(code[c][q], E[c][q], p.bits) : decode_value<T>(code[c][q], E[c][q], p.bits);
    }
    if (vec && c0 + 3 <= last) {
      tile::u32x4* d = reinterpret_cast<tile::u32x4*>(ow + c0);
      if constexpr (sizeof(T) == 4) {
        const tile::u32x4 v = { static_cast<std::uint32_t>(y[0]), static_cast<std::uint32_t>(y[1]),
                                static_cast<std::uint32_t>(y[2]), static_cast<std::uint32_t>(y[3]) };
        tile::st<NT>(v, d);
      } else {
        const tile::u32x4 v0 = { static_cast<std::uint32_t>(y[0]), static_cast<std::uint32_t>(y[0] >> 32),
                                 static_cast<std::uint32_t>(y[1]), static_cast<std::uint32_t>(y[1] >> 32) };
        const tile::u32x4 v1 = { static_cast<std::uint32_t>(y[2]), static_cast<std::uint32_t>(y[2] >> 32),
                                 static_cast<std::uint32_t>(y[3]), static_cast<std::uint32_t>(y[3] >> 32) };
        tile::st<NT>(v0, d);
        tile::st<NT>(v1, d + 1);
      }
    } else {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (c0 + static_cast<std::uint32_t>(q) <= last) {
          tile::st<NT>(y[q], ow + c0 + q);
        }
      }
    }
  }
}

// Decompress tile `tile` of a run (exponents at in, payload at in + p.R)
// into out.  LDS: decompress_lds_dwords(p) (payload image, then the tile's
// block exponents).
// PACKED: bits < U; otherwise word i of the image is value i's code.
template<typename T, int U, int VPT, bool PACKED, bool NT>
__device__ __forceinline__ void
decompress_gtile(const std::uint8_t* __restrict__ in, T* __restrict__ out, const gplan& p, std::uint64_t tile,
                 std::uint32_t* smem, int align)
{
  using uw = typename fmt<T>::uw;
  using wt = typename word<U>::type;
  const std::uint32_t t = threadIdx.x;
  std::uint32_t* img = smem;
  const std::uint8_t* exb = reinterpret_cast<const std::uint8_t*>(smem + exponent_offset(p));
  const std::uint64_t k0 = tile * p.TB;
  const std::uint64_t base = k0 * p.B;
  const std::uint64_t left = p.n - base;
  const std::uint32_t cnt = left < p.V ? static_cast<std::uint32_t>(left) : p.V;
  const std::uint32_t nblk = block_of(cnt - 1, p) + 1;
  // The tile's block exponents into LDS.  More blocks than threads (B < 9
  // at the target tile size): the contiguous run of exponents moves with the
  // widest loads its alignment allows, instead of several rounds of one load
  // per block (float32 bits 8, B = 3: 4.2 -> 5.2 TB/s).  Fewer: one load per
  // block and thread, which measured faster there (B = 12 and 100: 5-20%).
  std::uint8_t* eg = const_cast<std::uint8_t*>(in) + k0 * static_cast<std::uint64_t>(U / 8);
  std::uint8_t* el = reinterpret_cast<std::uint8_t*>(smem + exponent_offset(p));
  if (nblk > static_cast<std::uint32_t>(WG)) {
    move_payload<true, false>(eg, smem + exponent_offset(p), nblk * static_cast<std::uint32_t>(U / 8), alignment_at(eg));
  } else {
    for (std::uint32_t b = t; b < nblk; b += WG) {
      store_block_exponent<U>(el, b, load_block_exponent<U>(in, k0 + b));
    }
  }
  move_payload<true, NT>(const_cast<std::uint8_t*>(in) + p.R + k0 * p.CB, img, nblk * p.CB, align);
  __syncthreads();
  if constexpr (!PACKED && U <= 32 && VPT % 4 == 0) {
    decode_chunks<T, U, VPT / 4, NT>(img, exb, out + base, cnt, p);
    return;
  }
  const std::uint32_t bits = static_cast<std::uint32_t>(p.bits);
  const std::uint32_t padbits = p.CB * 8u - p.B * bits;
  uw code[VPT];
  int E[VPT];
  std::uint32_t ic[VPT];
  int emin = 0x7fffffff;
#pragma unroll
  for (int r = 0; r < VPT; ++r) {
    const std::uint32_t i0 = t + static_cast<std::uint32_t>(r * WG);
    const std::uint32_t i = i0 < cnt ? i0 : cnt - 1;
    ic[r] = i;
    const std::uint32_t k = block_of(i, p);
    if constexpr (PACKED) {
      code[r] = static_cast<uw>(tile::extract_code<U>(img, i * bits + k * padbits));
    } else {
      code[r] = static_cast<uw>(reinterpret_cast<const wt*>(img)[i]);
    }
    E[r] = load_block_exponent<U>(exb, k);
    emin = E[r] < emin ? E[r] : emin;
  }
  // decode_value_short needs bits <= SB + 1 and E + bias >= 0 (every block a
  // compress writes); checked over all of the thread's values
  const bool fast = p.bits <= fmt<T>::SB + 1 && emin + fmt<T>::BIAS >= 0;
  uw* o = reinterpret_cast<uw*>(out + base);
  if (fast) {
#pragma unroll
    for (int r = 0; r < VPT; ++r) {
      const uw y = decode_value_short<T>(code[r], E[r], p.bits);
      if (t + static_cast<std::uint32_t>(r * WG) < cnt) {
        tile::st<NT>(y, o + ic[r]);
      }
    }
  } else {
#pragma unroll
    for (int r = 0; r < VPT; ++r) {
      const uw y = decode_value<T>(code[r], E[r], p.bits);
      if (t + static_cast<std::uint32_t>(r * WG) < cnt) {
        tile::st<NT>(y, o + ic[r]);
      }
    }
  }
}

} // namespace gtile
} // namespace frsz2
