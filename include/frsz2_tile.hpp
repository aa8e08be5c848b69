// Tile-level FRSZ2 compress / decompress for gfx950, shared by the library's
// plain and batch kernels and by the kernel templates of the drop-in header
// (frsz2.hpp), which compile it into the caller's code.
//
// A tile is WGS*VPT consecutive values (WGS = 256 threads or one 64-lane
// wave, VPT values per thread), i.e. TB = tile/B whole exponent blocks; B is
// a power of two no larger than the tile.  A thread's values are 16-byte
// chunks, lane-interleaved, loaded and stored with 128-bit accesses.
//
// compress:  per-value max key (biased exponent, NaN -> 0) -> thread max ->
//            xor-butterfly max over the B/VPT lanes of one block (plus an LDS
//            step across waves when a block spans more than one wave) ->
//            block exponent -> per-value code.
//   bits == U: codes are U-bit words, payload word i = value i; stored
//            straight from registers.
//   bits <  U: codes are OR-ed into an LDS image of the tile's payload
//            (ds_or_b32 at their bit offsets), which is then written out with
//            128-bit stores.  For power-of-two B the tile payload TB*CB is a
//            multiple of 16 bytes and starts at a 16-byte aligned offset.
// decompress: the reverse; for bits < U the tile payload is staged into LDS
//            with 128-bit loads and codes are extracted with v_alignbit.
#pragma once

#include "frsz2_device.hpp"

#include <hip/hip_runtime.h>

#include <cstdint>

namespace frsz2 {
namespace tile {

constexpr int WG = 256;
constexpr int WAVE = 64;

// Per-array parameters of the tile path.
struct tplan
{
  std::uint64_t n;
  std::uint64_t nb;
  std::uint64_t R;
  std::uint32_t CB;
  int bits;
  int lb; // B = 1 << lb
  // exponent-pad bytes to clear after exponent slot nb (by the first tile):
  // the whole pad for a stream, 0 for a run of blocks that does not end at
  // the stream's last block
  std::uint32_t zpad;
};

typedef std::uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef std::uint32_t u32x2 __attribute__((ext_vector_type(2)));

// LDS dwords a packed tile needs (payload image plus 3 dwords of slack for
// the funnel reads / ORs that touch the dword after a code's last one)
template<int VPT>
constexpr int lds_dwords()
{
  return WG * VPT * 2 + 4;
}

// the same for codes of at most U bits (a tile's payload is at most
// WG*VPT*U/8 bytes)
template<int U, int VPT, int WGS = WG>
constexpr int lds_dwords_u()
{
  return WGS * VPT * U / 32 + 4;
}

// 128/64/32-bit global access, optionally with the non-temporal hint
template<bool NT, typename V>
__device__ __forceinline__ V
ld(const V* p)
{
  if constexpr (NT) {
    return __builtin_nontemporal_load(p);
  } else {
    return *p;
  }
}

template<bool NT, typename V>
__device__ __forceinline__ void
st(V v, V* p)
{
  if constexpr (NT) {
    __builtin_nontemporal_store(v, p);
  } else {
    *p = v;
  }
}

template<typename T, int VPT, bool NT>
__device__ __forceinline__ void
load_values(const typename fmt<T>::uw* __restrict__ p, typename fmt<T>::uw (&x)[VPT])
{
  constexpr int NV = VPT * static_cast<int>(sizeof(T)) / 16;
  const u32x4* q = reinterpret_cast<const u32x4*>(p);
#pragma unroll
  for (int c = 0; c < NV; ++c) {
    const u32x4 d = ld<NT>(q + c);
    if (sizeof(T) == 4) {
      x[4 * c + 0] = d.x;
      x[4 * c + 1] = d.y;
      x[4 * c + 2] = d.z;
      x[4 * c + 3] = d.w;
    } else {
      x[2 * c + 0] = static_cast<typename fmt<T>::uw>(d.x) | (static_cast<std::uint64_t>(d.y) << 32);
      x[2 * c + 1] = static_cast<typename fmt<T>::uw>(d.z) | (static_cast<std::uint64_t>(d.w) << 32);
    }
  }
}

template<typename T, int VPT, bool NT>
__device__ __forceinline__ void
store_values(typename fmt<T>::uw* __restrict__ p, const typename fmt<T>::uw (&y)[VPT])
{
  constexpr int NV = VPT * static_cast<int>(sizeof(T)) / 16;
  u32x4* q = reinterpret_cast<u32x4*>(p);
#pragma unroll
  for (int c = 0; c < NV; ++c) {
    u32x4 d;
    if (sizeof(T) == 4) {
      d.x = static_cast<std::uint32_t>(y[4 * c + 0]);
      d.y = static_cast<std::uint32_t>(y[4 * c + 1]);
      d.z = static_cast<std::uint32_t>(y[4 * c + 2]);
      d.w = static_cast<std::uint32_t>(y[4 * c + 3]);
    } else {
      d.x = static_cast<std::uint32_t>(y[2 * c + 0]);
      d.y = static_cast<std::uint32_t>(static_cast<std::uint64_t>(y[2 * c + 0]) >> 32);
      d.z = static_cast<std::uint32_t>(y[2 * c + 1]);
      d.w = static_cast<std::uint32_t>(static_cast<std::uint64_t>(y[2 * c + 1]) >> 32);
    }
    st<NT>(d, q + c);
  }
}

// VPT U-bit codes <-> VPT*U/32 dwords (little-endian word order)
template<int U, int VPT, typename UW>
__device__ __forceinline__ void
pack_words(const UW (&code)[VPT], std::uint32_t (&w)[VPT * U / 32 > 0 ? VPT * U / 32 : 1])
{
  if constexpr (U == 64) {
#pragma unroll
    for (int v = 0; v < VPT; ++v) {
      w[2 * v] = static_cast<std::uint32_t>(code[v]);
      w[2 * v + 1] = static_cast<std::uint32_t>(static_cast<std::uint64_t>(code[v]) >> 32);
    }
  } else {
    constexpr int PER = 32 / U;
#pragma unroll
    for (int d = 0; d < VPT / PER; ++d) {
      std::uint32_t acc = 0;
#pragma unroll
      for (int i = 0; i < PER; ++i) {
        acc |= static_cast<std::uint32_t>(code[d * PER + i]) << (i * U);
      }
      w[d] = acc;
    }
  }
}

template<int U, int VPT, typename UW>
__device__ __forceinline__ void
unpack_words(const std::uint32_t (&w)[VPT * U / 32 > 0 ? VPT * U / 32 : 1], UW (&code)[VPT])
{
  if constexpr (U == 64) {
#pragma unroll
    for (int v = 0; v < VPT; ++v) {
      code[v] = static_cast<UW>(w[2 * v]) | static_cast<UW>(static_cast<std::uint64_t>(w[2 * v + 1]) << 32);
    }
  } else {
    constexpr int PER = 32 / U;
    constexpr std::uint32_t M = U == 32 ? 0xffffffffu : ((1u << (U % 32)) - 1u);
#pragma unroll
    for (int d = 0; d < VPT / PER; ++d) {
#pragma unroll
      for (int i = 0; i < PER; ++i) {
        code[d * PER + i] = static_cast<UW>((w[d] >> (i * U)) & M);
      }
    }
  }
}

template<bool NT, int N>
__device__ __forceinline__ void
store_dwords(std::uint8_t* __restrict__ dst, const std::uint32_t (&w)[N])
{
  if constexpr (N % 4 == 0) {
    u32x4* q = reinterpret_cast<u32x4*>(dst);
#pragma unroll
    for (int c = 0; c < N / 4; ++c) {
      u32x4 d;
      d.x = w[4 * c];
      d.y = w[4 * c + 1];
      d.z = w[4 * c + 2];
      d.w = w[4 * c + 3];
      st<NT>(d, q + c);
    }
  } else if constexpr (N == 2) {
    typedef std::uint32_t u32x2 __attribute__((ext_vector_type(2)));
    u32x2 d;
    d.x = w[0];
    d.y = w[1];
    st<NT>(d, reinterpret_cast<u32x2*>(dst));
  } else {
    st<NT>(w[0], reinterpret_cast<std::uint32_t*>(dst));
  }
}

template<bool NT, int N>
__device__ __forceinline__ void
load_dwords(const std::uint8_t* __restrict__ src, std::uint32_t (&w)[N])
{
  if constexpr (N % 4 == 0) {
    const u32x4* q = reinterpret_cast<const u32x4*>(src);
#pragma unroll
    for (int c = 0; c < N / 4; ++c) {
      const u32x4 d = ld<NT>(q + c);
      w[4 * c] = d.x;
      w[4 * c + 1] = d.y;
      w[4 * c + 2] = d.z;
      w[4 * c + 3] = d.w;
    }
  } else if constexpr (N == 2) {
    typedef std::uint32_t u32x2 __attribute__((ext_vector_type(2)));
    const u32x2 d = ld<NT>(reinterpret_cast<const u32x2*>(src));
    w[0] = d.x;
    w[1] = d.y;
  } else {
    w[0] = ld<NT>(reinterpret_cast<const std::uint32_t*>(src));
  }
}

template<int U>
__device__ __forceinline__ void
store_exponent(std::uint8_t* __restrict__ out, std::uint64_t k, int e)
{
  store_block_exponent<U>(out, k, e);
}

// max of v and the value of the lane that DPP control CTRL selects (a lane of
// the same 16-lane row)
template<int CTRL>
__device__ __forceinline__ std::uint32_t
dpp_max(std::uint32_t v)
{
  const std::uint32_t o =
    static_cast<std::uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), CTRL, 0xF, 0xF, false));
  return v > o ? v : o;
}

// max over aligned groups of G lanes (G a power of two up to the wave).
// Within a row the steps are DPP moves fused into the max: quad_perm
// [1,0,3,2] and [2,3,0,1] (each lane then holds its quad's max), row half
// mirror (lane i <- 7-i: the other quad of the 8-lane group) and row mirror
// (lane i <- 15-i: the other half of the row).  Groups of 32 and 64 lanes
// finish with cross-row shuffles.
template<int G>
__device__ __forceinline__ std::uint32_t
lane_max(std::uint32_t v)
{
  if constexpr (G >= 2) {
    v = dpp_max<0xB1>(v);
  }
  if constexpr (G >= 4) {
    v = dpp_max<0x4E>(v);
  }
  if constexpr (G >= 8) {
    v = dpp_max<0x141>(v);
  }
  if constexpr (G >= 16) {
    v = dpp_max<0x140>(v);
  }
  if constexpr (G >= 32) {
    const std::uint32_t o = static_cast<std::uint32_t>(__shfl_xor(static_cast<int>(v), 16));
    v = v > o ? v : o;
  }
  if constexpr (G >= 64) {
    const std::uint32_t o = static_cast<std::uint32_t>(__shfl_xor(static_cast<int>(v), 32));
    v = v > o ? v : o;
  }
  return v;
}

// max over the G lanes of an aligned lane group (G a power of two); for
// G > 64 the wave maxima are combined through wred (one dword per wave).
__device__ __forceinline__ std::uint32_t
group_max(std::uint32_t v, int G, std::uint32_t* wred)
{
  // G is uniform: one branch picks a butterfly with compile-time lane masks
  switch (G < WAVE ? G : WAVE) {
    case 1:
      break;
    case 2:
      v = lane_max<2>(v);
      break;
    case 4:
      v = lane_max<4>(v);
      break;
    case 8:
      v = lane_max<8>(v);
      break;
    case 16:
      v = lane_max<16>(v);
      break;
    case 32:
      v = lane_max<32>(v);
      break;
    default:
      v = lane_max<64>(v);
      break;
  }
  if (G > WAVE) {
    const int wave = threadIdx.x / WAVE;
    if ((threadIdx.x & (WAVE - 1)) == 0) {
      wred[wave] = v;
    }
    __syncthreads();
    const int wpg = G / WAVE;
    const int w0 = wave & ~(wpg - 1);
    v = wred[w0];
    for (int i = 1; i < wpg; ++i) {
      const std::uint32_t o = wred[w0 + i];
      v = v > o ? v : o;
    }
    __syncthreads();
  }
  return v;
}

// Store the low `bytes` bytes (1, 2, 4 or 8; p aligned to min(bytes, 4))
// of w into LDS.
__device__ __forceinline__ void
store_block_bytes(std::uint8_t* p, std::uint64_t w, std::uint32_t bytes)
{
  switch (bytes) {
    case 1:
      *p = static_cast<std::uint8_t>(w);
      break;
    case 2:
      *reinterpret_cast<std::uint16_t*>(p) = static_cast<std::uint16_t>(w);
      break;
    case 4:
      *reinterpret_cast<std::uint32_t*>(p) = static_cast<std::uint32_t>(w);
      break;
    default:
      reinterpret_cast<std::uint32_t*>(p)[0] = static_cast<std::uint32_t>(w);
      reinterpret_cast<std::uint32_t*>(p)[1] = static_cast<std::uint32_t>(w >> 32);
      break;
  }
}

// OR a code into the LDS bit image at bit position pos
template<int U>
__device__ __forceinline__ void
or_code(std::uint32_t* lds, std::uint32_t pos, std::uint64_t code, int bits)
{
  const std::uint32_t q = pos >> 5;
  const int o = static_cast<int>(pos & 31u);
  if constexpr (U <= 32) {
    const std::uint64_t s = code << o;
    atomicOr(lds + q, static_cast<std::uint32_t>(s));
    if (o + bits > 32) {
      atomicOr(lds + q + 1, static_cast<std::uint32_t>(s >> 32));
    }
  } else {
    const std::uint64_t s = code << o;
    const std::uint32_t lo = static_cast<std::uint32_t>(s);
    if (lo) {
      atomicOr(lds + q, lo);
    }
    if (o + bits > 32) {
      atomicOr(lds + q + 1, static_cast<std::uint32_t>(s >> 32));
    }
    if (o + bits > 64) {
      atomicOr(lds + q + 2, static_cast<std::uint32_t>(code >> (64 - o)));
    }
  }
}

// OR the VC codes of one chunk into the LDS bit image when the chunk lies in
// one exponent block: its codes are then consecutive in the bitstream, from
// bit pos0 on.  They are assembled into one 128-bit value (VC * bits <= 128;
// codes are below 2^bits, and bits < U, so pairs of float32 codes fit one
// 64-bit word), shifted by pos0 mod 32 and OR-ed in as the 1..5 dwords they
// touch: one LDS atomic per touched dword instead of one or two per code.
template<int VC, typename UW>
__device__ __forceinline__ void
or_chunk(std::uint32_t* lds, std::uint32_t pos0, const UW (&code)[VC], int bits)
{
  std::uint64_t lo = 0;
  std::uint64_t hi = 0;
  if constexpr (VC == 4) {
    // float32 chunk, packed codes (bits <= 31): two codes fit one 64-bit word
    const std::uint64_t p01 = static_cast<std::uint64_t>(code[0]) | (static_cast<std::uint64_t>(code[1]) << bits);
    const std::uint64_t p23 = static_cast<std::uint64_t>(code[2]) | (static_cast<std::uint64_t>(code[3]) << bits);
    const int s2 = 2 * bits;
    lo = p01 | (p23 << s2);
    hi = p23 >> (64 - s2);
  } else if constexpr (VC == 2) {
    // float64 chunk, packed codes (bits <= 63)
    lo = static_cast<std::uint64_t>(code[0]) | (static_cast<std::uint64_t>(code[1]) << bits);
    hi = static_cast<std::uint64_t>(code[1]) >> (64 - bits);
  } else {
#pragma unroll
    for (int v = 0; v < VC; ++v) {
      const std::uint64_t c = static_cast<std::uint64_t>(code[v]);
      const int sh = v * bits;
      if (sh == 0) {
        lo |= c;
      } else if (sh < 64) {
        lo |= c << sh;
        hi |= c >> (64 - sh);
      } else {
        hi |= c << (sh - 64);
      }
    }
  }
  const std::uint32_t q = pos0 >> 5;
  const int o = static_cast<int>(pos0 & 31u);
  // (hi:lo) << o as three 64-bit words; (x >> 1) >> (63 - o) is x >> (64 - o), and 0 for o == 0
  const std::uint64_t w0 = lo << o;
  const std::uint64_t w1 = (hi << o) | ((lo >> 1) >> (63 - o));
  const std::uint64_t w2 = (hi >> 1) >> (63 - o);
  const int nd = (o + VC * bits + 31) >> 5;
  atomicOr(lds + q, static_cast<std::uint32_t>(w0));
  if (nd > 1) {
    atomicOr(lds + q + 1, static_cast<std::uint32_t>(w0 >> 32));
  }
  if (nd > 2) {
    atomicOr(lds + q + 2, static_cast<std::uint32_t>(w1));
  }
  if (nd > 3) {
    atomicOr(lds + q + 3, static_cast<std::uint32_t>(w1 >> 32));
  }
  if (nd > 4) {
    atomicOr(lds + q + 4, static_cast<std::uint32_t>(w2));
  }
}

// The U bits at bit offset pos of the LDS payload image: a code in the low
// `bits` bits, with the bits above it left in (the decodes ignore them).
template<int U>
__device__ __forceinline__ std::uint64_t
extract_code(const std::uint32_t* lds, std::uint32_t pos)
{
  const std::uint32_t q = pos >> 5;
  const std::uint32_t o = pos & 31u;
  const std::uint32_t lo = __builtin_amdgcn_alignbit(lds[q + 1], lds[q], o);
  if constexpr (U <= 32) {
    return lo;
  } else {
    const std::uint32_t hi = __builtin_amdgcn_alignbit(lds[q + 2], lds[q + 1], o);
    return static_cast<std::uint64_t>(lo) | (static_cast<std::uint64_t>(hi) << 32);
  }
}

// Block exponents of a thread's NCH chunks of VC values (chunk c of thread t
// holds tile values (c*WGS + t)*VC .. +VC), per value in E.  The lane that
// holds a block's first value stores the block's exponent.
template<typename T, int U, int VC, int NCH, int WGS>
__device__ __forceinline__ void
chunk_exponents(const std::uint32_t (&key)[NCH][VC],
                int (&E)[NCH][VC],
                std::uint64_t base,
                const tplan& p,
                std::uint8_t* __restrict__ out,
                std::uint32_t* wred)
{
  const int B = 1 << p.lb;
  const int t = static_cast<int>(threadIdx.x);
  if (B >= WGS * VC) {
    // a block covers R whole chunk rows of the workgroup
    const int R = B / (WGS * VC);
#pragma unroll
    for (int r0 = 0; r0 < NCH; ++r0) {
      if ((r0 & (R - 1)) == 0) {
        std::uint32_t m = 0;
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
          if (c >= r0 && c < r0 + R) {
#pragma unroll
            for (int v = 0; v < VC; ++v) {
              m = m > key[c][v] ? m : key[c][v];
            }
          }
        }
        m = group_max(m, WGS, wred);
        const int e = block_exponent_from_key<T, U>(m);
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
          if (c >= r0 && c < r0 + R) {
#pragma unroll
            for (int v = 0; v < VC; ++v) {
              E[c][v] = e;
            }
          }
        }
        const std::uint64_t k = (base + static_cast<std::uint64_t>(r0) * WGS * VC) >> p.lb;
        if (t == 0 && k < p.nb) {
          store_exponent<U>(out, k, e);
        }
      }
    }
  } else if (B >= VC) {
    // a block covers B/VC lanes of one chunk row
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
      std::uint32_t m = key[c][0];
#pragma unroll
      for (int v = 1; v < VC; ++v) {
        m = m > key[c][v] ? m : key[c][v];
      }
      m = group_max(m, B / VC, wred);
      const int e = block_exponent_from_key<T, U>(m);
#pragma unroll
      for (int v = 0; v < VC; ++v) {
        E[c][v] = e;
      }
      const std::uint64_t i0 = base + static_cast<std::uint64_t>(c * WGS + t) * VC;
      if ((i0 & static_cast<std::uint64_t>(B - 1)) == 0 && (i0 >> p.lb) < p.nb) {
        store_exponent<U>(out, i0 >> p.lb, e);
      }
    }
  } else {
    // B < VC: whole blocks inside a chunk; butterfly max over runs of B
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
      std::uint32_t m[VC];
#pragma unroll
      for (int v = 0; v < VC; ++v) {
        m[v] = key[c][v];
      }
#pragma unroll
      for (int s = 1; s < VC; s <<= 1) {
        if (s < B) {
          std::uint32_t q[VC];
#pragma unroll
          for (int v = 0; v < VC; ++v) {
            q[v] = m[v] > m[v ^ s] ? m[v] : m[v ^ s];
          }
#pragma unroll
          for (int v = 0; v < VC; ++v) {
            m[v] = q[v];
          }
        }
      }
      const std::uint64_t i0 = base + static_cast<std::uint64_t>(c * WGS + t) * VC;
#pragma unroll
      for (int v = 0; v < VC; ++v) {
        E[c][v] = block_exponent_from_key<T, U>(m[v]);
        if ((v & (B - 1)) == 0 && ((i0 + v) >> p.lb) < p.nb) {
          store_exponent<U>(out, (i0 + v) >> p.lb, E[c][v]);
        }
      }
    }
  }
}

// Compress one tile.  A thread's VPT values are NCH chunks of VC = 16 bytes
// of input (4 float32 / 2 float64), chunk c of thread t holding tile values
// [(c*WGS + t)*VC, +VC), so every load instruction of a wave reads 1 KiB of
// consecutive input and every code store writes consecutive bytes.  VEC =
// false loads the values one by one (input aligned to its element size
// only; the payload must still be 16-byte aligned).
template<typename T, int U, int VPT, bool PACKED, bool NT, int WGS = WG, bool VEC = true>
__device__ __forceinline__ void
compress_tile(const typename fmt<T>::uw* __restrict__ in,
              std::uint8_t* __restrict__ out,
              const tplan& p,
              std::uint64_t tile,
              std::uint32_t* lds,
              std::uint32_t* wred)
{
  using uw = typename fmt<T>::uw;
  constexpr int TILE = WGS * VPT;
  constexpr int VC = 16 / static_cast<int>(sizeof(T));
  constexpr int NCH = VPT / VC;
  static_assert(VPT % VC == 0, "a thread holds whole 16-byte chunks");
  const int t = static_cast<int>(threadIdx.x);
  const std::uint64_t base = tile * TILE;

  uw x[NCH][VC];
#pragma unroll
  for (int c = 0; c < NCH; ++c) {
    const std::uint64_t i0 = base + static_cast<std::uint64_t>(c * WGS + t) * VC;
    if (VEC && i0 + VC <= p.n) {
      load_values<T, VC, NT>(in + i0, x[c]);
    } else if (i0 + VC <= p.n) {
#pragma unroll
      for (int v = 0; v < VC; ++v) {
        x[c][v] = ld<NT>(in + i0 + v);
      }
    } else {
#pragma unroll
      for (int v = 0; v < VC; ++v) {
        x[c][v] = (i0 + v < p.n) ? in[i0 + v] : uw{ 0 };
      }
    }
  }
  std::uint32_t key[NCH][VC];
#pragma unroll
  for (int c = 0; c < NCH; ++c) {
#pragma unroll
    for (int v = 0; v < VC; ++v) {
      key[c][v] = max_key<T>(x[c][v]);
    }
  }
  int E[NCH][VC];
  chunk_exponents<T, U, VC, NCH, WGS>(key, E, base, p, out, wred);

  // narrow widths with one exponent per chunk: encode_value_short, with the
  // per-block terms computed once per chunk (both flags are uniform)
  const bool short_codes = p.bits <= fmt<T>::SB + 1 && (1 << p.lb) >= VC;
  uw code[NCH][VC];
#pragma unroll
  for (int c = 0; c < NCH; ++c) {
    if (short_codes) {
      const int e = E[c][0];
#pragma unroll
      for (int v = 0; v < VC; ++v) {
        code[c][v] = encode_value_short<T>(x[c][v], e, p.bits);
      }
    } else {
#pragma unroll
      for (int v = 0; v < VC; ++v) {
        code[c][v] = encode_value<T>(x[c][v], E[c][v], p.bits);
      }
    }
  }

  std::uint8_t* pay = out + p.R;
  if (!PACKED) {
    // bits == U: payload word i holds value i; words [n, nb*B) get zero codes
    using wt = typename word<U>::type;
    const std::uint64_t limit = p.nb << p.lb;
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
      const std::uint64_t i0 = base + static_cast<std::uint64_t>(c * WGS + t) * VC;
      if (VC * U >= 32 && i0 + VC <= limit) {
        std::uint32_t w[VC * U / 32 > 0 ? VC * U / 32 : 1];
        pack_words<U, VC>(code[c], w);
        store_dwords<NT>(pay + i0 * (U / 8), w);
      } else {
#pragma unroll
        for (int v = 0; v < VC; ++v) {
          if (i0 + v < limit) {
            reinterpret_cast<wt*>(pay)[i0 + v] = static_cast<wt>(code[c][v]);
          }
        }
      }
    }
  } else {
    const int B = 1 << p.lb;
    const std::uint64_t k0 = base >> p.lb;
    const std::uint32_t TB = static_cast<std::uint32_t>(TILE >> p.lb);
    const std::uint32_t tile_bytes = TB * p.CB;
    u32x4* l4 = reinterpret_cast<u32x4*>(lds);
    for (std::uint32_t c = t; c < tile_bytes / 16; c += WGS) {
      l4[c] = u32x4{ 0u, 0u, 0u, 0u };
    }
    __syncthreads();
    if (B >= VC) {
      // each chunk lies in one block.  The bit offset of tile value i is
      // f(i) = (i >> lb)*CB*8 + (i & (B-1))*bits, and for power-of-two B,
      // f(c*WGS*VC + t*VC) = f(c*WGS*VC) + f(t*VC): a uniform term per chunk
      const std::uint32_t cb8 = p.CB * 8u;
      const std::uint32_t bm = static_cast<std::uint32_t>(B - 1);
      const std::uint32_t bits = static_cast<std::uint32_t>(p.bits);
      const std::uint32_t tv = static_cast<std::uint32_t>(t * VC);
      const std::uint32_t P = (tv >> p.lb) * cb8 + (tv & bm) * bits;
#pragma unroll
      for (int c = 0; c < NCH; ++c) {
        const std::uint32_t cw = static_cast<std::uint32_t>(c * WGS * VC);
        const std::uint32_t sc = __builtin_amdgcn_readfirstlane((cw >> p.lb) * cb8 + (cw & bm) * bits);
        or_chunk<VC>(lds, P + sc, code[c], p.bits);
      }
    } else if (B <= 2) {
      // B < VC (B = 1 or 2): the chunk holds whole blocks, and a block's CB
      // bytes (ceil(B*bits/U) words of U bits: 1, 2, 4 or 8 bytes, so
      // aligned to their size) belong to this lane alone.  Its codes,
      // assembled LSB-first, are stored into the image, with no atomics.
      std::uint8_t* img = reinterpret_cast<std::uint8_t*>(lds);
      const std::uint32_t cb = p.CB;
#pragma unroll
      for (int c = 0; c < NCH; ++c) {
        const std::uint32_t local0 = static_cast<std::uint32_t>((c * WGS + t) * VC);
#pragma unroll
        for (int v = 0; v < VC; v += 2) {
          std::uint64_t w0 = static_cast<std::uint64_t>(code[c][v]);
          std::uint64_t w1 = VC > 1 ? static_cast<std::uint64_t>(code[c][v + 1]) : 0;
          std::uint32_t k = local0 + static_cast<std::uint32_t>(v);
          if (B == 2) {
            // one block of two codes (2 * bits <= 2 * (U - 1) <= 62 for float32)
            store_block_bytes(img + (k >> 1) * cb, w0 | (w1 << p.bits), cb);
          } else {
            store_block_bytes(img + k * cb, w0, cb);
            store_block_bytes(img + (k + 1) * cb, w1, cb);
          }
        }
      }
    } else {
#pragma unroll
      for (int c = 0; c < NCH; ++c) {
#pragma unroll
        for (int v = 0; v < VC; ++v) {
          if (code[c][v] != 0) {
            const std::uint32_t local = static_cast<std::uint32_t>((c * WGS + t) * VC + v);
            const std::uint32_t kl = local >> p.lb;
            const std::uint32_t j = local & static_cast<std::uint32_t>(B - 1);
            const std::uint32_t pos = kl * p.CB * 8u + j * static_cast<std::uint32_t>(p.bits);
            or_code<U>(lds, pos, static_cast<std::uint64_t>(code[c][v]), p.bits);
          }
        }
      }
    }
    __syncthreads();
    const std::uint64_t left = p.nb - k0;
    const std::uint32_t blocks = left < TB ? static_cast<std::uint32_t>(left) : TB;
    const std::uint32_t nbytes = blocks * p.CB;
    std::uint8_t* dst = pay + k0 * p.CB;
    u32x4* d4 = reinterpret_cast<u32x4*>(dst);
    for (std::uint32_t c = t; c < nbytes / 16; c += WGS) {
      st<NT>(l4[c], d4 + c);
    }
    const std::uint32_t tail = nbytes & 15u;
    if (static_cast<std::uint32_t>(t) < tail) {
      const std::uint32_t o = nbytes - tail + t;
      dst[o] = reinterpret_cast<const std::uint8_t*>(lds)[o];
    }
    __syncthreads();
  }
}

// Values of a chunk of VC codes (bits of a code above `bits` are ignored).
// `uniform`: the chunk lies in one exponent block (E[0] holds for all of it,
// so the per-block terms are computed once); `narrow`: bits <= SB + 1.  Both
// flags are the same for every thread of the launch; together with
// E + bias >= 0 (true for every block a compress writes; a stream from
// elsewhere may hold any exponent) they take decode_value_short.
template<typename T, int VC>
__device__ __forceinline__ void
decode_chunk(const typename fmt<T>::uw (&code)[VC],
             const int (&E)[VC],
             bool uniform,
             bool narrow,
             int bits,
             typename fmt<T>::uw (&y)[VC])
{
  if (narrow && uniform && E[0] + fmt<T>::BIAS >= 0) {
    const int e = E[0];
#pragma unroll
    for (int v = 0; v < VC; ++v) {
      y[v] = decode_value_short<T>(code[v], e, bits);
    }
    return;
  }
  // several blocks in the chunk (B < VC): the short form per value when
  // every one of their exponents allows it
  int emin = E[0];
#pragma unroll
  for (int v = 1; v < VC; ++v) {
    emin = E[v] < emin ? E[v] : emin;
  }
  if (narrow && emin + fmt<T>::BIAS >= 0) {
#pragma unroll
    for (int v = 0; v < VC; ++v) {
      y[v] = decode_value_short<T>(code[v], E[v], bits);
    }
  } else {
#pragma unroll
    for (int v = 0; v < VC; ++v) {
      y[v] = decode_value<T>(code[v], E[v], bits);
    }
  }
}

// NB-byte code load / value store helpers of the chunked decompress
template<bool NT, int NB>
__device__ __forceinline__ void
load_bytes(const std::uint8_t* __restrict__ src, std::uint32_t (&w)[NB >= 4 ? NB / 4 : 1])
{
  if constexpr (NB == 2) {
    w[0] = ld<NT>(reinterpret_cast<const std::uint16_t*>(src));
  } else {
    load_dwords<NT>(src, w);
  }
}

// Decompress one tile.  A thread's VPT values are split into NCH chunks of
// VC consecutive values, chunk c of thread t holding values
// [(c*WG + t)*VC, +VC) of the tile, so each load / store instruction of the
// workgroup covers one contiguous span.  All loads of a thread are issued
// before any value is decoded.
//   direct (!STAGE): a chunk is 16 bytes of output (4 float32 / 2 float64),
//     stored straight from registers; its codes are VC*U/8 bytes (bits == U)
//     or extracted from the LDS image of the tile's payload (bits < U).
//   STAGE: a chunk is 16 bytes of codes (bits == U; VC = 128/U, at most VPT)
//     or all VPT values of the thread (bits < U); the decoded tile goes
//     through LDS and out with 16 bytes per lane per store instruction.
//     This keeps the code loads 128-bit wide when a chunk's output is 32+
//     bytes per lane, without stores that half-fill their cache lines.
// LDS dwords of the output staging buffer of a decompress tile
template<typename T, int VPT, int WGS = WG>
constexpr int
stage_dwords()
{
  return WGS * VPT * static_cast<int>(sizeof(T)) / 4;
}

// Decompress one tile, strided: thread t decodes tile values k*WGS + t
// (k < VPT).  Packed codes (bits < U) are read from an LDS image of the
// tile's payload; codes of bits == U are loaded from HBM directly (word i of
// the payload is value i).  The bit offset of tile value i in the
// tile's payload is f(i) = (i >> lb)*CB*8 + (i & (B-1))*bits, and for
// power-of-two B, f(k*WGS + t) = f(k*WGS) + f(t) (B >= WGS: k*WGS mod B + t
// < B; B < WGS: k*WGS is a multiple of B), so a value's offset is one add of
// a workgroup-uniform term.  Each wave instruction reads consecutive dwords
// of the LDS image (no bank conflicts, whatever bits is) and stores 64
// consecutive values, so no output staging is needed.  A value's exponent
// is loaded for it (lanes of one block read one address).
template<typename T, int U, int VPT, bool PACKED, bool NT, int WGS>
__device__ __forceinline__ void
decompress_tile_strided(const std::uint8_t* __restrict__ in,
                        typename fmt<T>::uw* __restrict__ out,
                        const tplan& p,
                        std::uint64_t tile,
                        std::uint32_t* lds)
{
  using uw = typename fmt<T>::uw;
  constexpr int TILE = WGS * VPT;
  const std::uint32_t t = threadIdx.x;
  const std::uint64_t base = tile * TILE;
  const std::uint32_t bm = (1u << p.lb) - 1u;
  const std::uint32_t cb8 = p.CB * 8u;
  const std::uint32_t bits = static_cast<std::uint32_t>(p.bits);

  const std::uint64_t k0 = base >> p.lb;
  const bool full = base + TILE <= p.n;

  // exponents first: their loads are in flight with the payload's.  Value
  // k*WGS + t is in tile block (k*WGS >> lb) + (t >> lb) (power-of-two B), a
  // uniform part and a lane part.
  using et = typename signed_word<U>::type;
  const std::uint32_t tl = t >> p.lb;
  const et* e0 = reinterpret_cast<const et*>(in) + k0;
  int E[VPT];
  if (full) {
#pragma unroll
    for (int k = 0; k < VPT; ++k) {
      const et* ek = e0 + (static_cast<std::uint32_t>(k * WGS) >> p.lb);
      E[k] = ek[tl];
    }
  } else {
#pragma unroll
    for (int k = 0; k < VPT; ++k) {
      const std::uint32_t kw = static_cast<std::uint32_t>(k * WGS);
      E[k] = base + kw + t < p.n ? static_cast<int>(e0[(kw >> p.lb) + tl]) : 0;
    }
  }

  uw code[VPT];
  if constexpr (PACKED) {
    const std::uint32_t TB = static_cast<std::uint32_t>(TILE >> p.lb);
    const std::uint64_t left = p.nb - k0;
    const std::uint32_t blocks = left < TB ? static_cast<std::uint32_t>(left) : TB;
    const std::uint32_t nbytes = blocks * p.CB;
    const std::uint8_t* src = in + p.R + k0 * p.CB;
    const u32x4* s4 = reinterpret_cast<const u32x4*>(src);
    u32x4* l4 = reinterpret_cast<u32x4*>(lds);
    for (std::uint32_t c = t; c < nbytes / 16; c += WGS) {
      l4[c] = ld<NT>(s4 + c);
    }
    const std::uint32_t tail = nbytes & 15u;
    if (t < tail) {
      const std::uint32_t o = nbytes - tail + t;
      reinterpret_cast<std::uint8_t*>(lds)[o] = src[o];
    }
    __syncthreads();

    const std::uint32_t P = (t >> p.lb) * cb8 + (t & bm) * bits;
#pragma unroll
    for (int k = 0; k < VPT; ++k) {
      const std::uint32_t kw = static_cast<std::uint32_t>(k * WGS);
      // uniform; kept whole so the per-value work is the one add
      const std::uint32_t sk = __builtin_amdgcn_readfirstlane((kw >> p.lb) * cb8 + (kw & bm) * bits);
      code[k] = static_cast<uw>(extract_code<U>(lds, P + sk));
    }
  } else {
    // bits == U: payload word i holds value i (words up to nb*B exist)
    using wt = typename word<U>::type;
    const wt* w = reinterpret_cast<const wt*>(in + p.R) + base;
    const std::uint64_t limit = p.nb << p.lb;
#pragma unroll
    for (int k = 0; k < VPT; ++k) {
      const std::uint32_t kw = static_cast<std::uint32_t>(k * WGS);
      code[k] = full || base + kw + t < limit ? static_cast<uw>(ld<NT>(w + kw + t)) : uw{ 0 };
    }
  }
  // decode_value_short needs bits <= SB + 1 and E + bias >= 0 (every block a
  // compress writes); the check is per thread, over all of its values
  int emin = E[0];
#pragma unroll
  for (int k = 1; k < VPT; ++k) {
    emin = E[k] < emin ? E[k] : emin;
  }
  uw y[VPT];
  if (p.bits <= fmt<T>::SB + 1 && emin + fmt<T>::BIAS >= 0) {
#pragma unroll
    for (int k = 0; k < VPT; ++k) {
      y[k] = decode_value_short<T>(code[k], E[k], p.bits);
    }
  } else {
#pragma unroll
    for (int k = 0; k < VPT; ++k) {
      y[k] = decode_value<T>(code[k], E[k], p.bits);
    }
  }
  if (full) {
    uw* ob = out + base;
#pragma unroll
    for (int k = 0; k < VPT; ++k) {
      st<NT>(y[k], ob + static_cast<std::uint32_t>(k * WGS) + t);
    }
  } else {
#pragma unroll
    for (int k = 0; k < VPT; ++k) {
      const std::uint64_t i = base + static_cast<std::uint32_t>(k * WGS) + t;
      if (i < p.n) {
        out[i] = y[k];
      }
    }
  }
  if constexpr (PACKED) {
    __syncthreads(); // the LDS image is reused by the next tile of a persistent workgroup
  }
}

// Block exponents of one chunk of VC values that holds NBK = VC >> LB whole
// exponent blocks (B = 2^LB < VC): their NBK slots are contiguous in the
// stream and aligned to their size (the chunk starts at a multiple of VC
// values), so they come in with one load of NBK*U/8 <= 16 bytes instead of
// one load per value.  E[v] is the exponent of value v's block.
template<int U, int VC, int LB>
__device__ __forceinline__ void
chunk_exponents_vec(const std::uint8_t* __restrict__ in, std::uint64_t k0, int (&E)[VC])
{
  constexpr int NBK = VC >> LB;
  constexpr int BYTES = NBK * (U / 8);
  static_assert(BYTES >= 1 && BYTES <= 16, "one load of at most 16 bytes");
  const std::uint8_t* p = in + k0 * static_cast<std::uint64_t>(U / 8);
  std::uint32_t w[4] = { 0, 0, 0, 0 };
  if constexpr (BYTES == 16) {
    const u32x4 d = *reinterpret_cast<const u32x4*>(p);
    w[0] = d.x;
    w[1] = d.y;
    w[2] = d.z;
    w[3] = d.w;
  } else if constexpr (BYTES == 8) {
    const u32x2 d = *reinterpret_cast<const u32x2*>(p);
    w[0] = d.x;
    w[1] = d.y;
  } else if constexpr (BYTES == 4) {
    w[0] = *reinterpret_cast<const std::uint32_t*>(p);
  } else if constexpr (BYTES == 2) {
    w[0] = *reinterpret_cast<const std::uint16_t*>(p);
  } else {
    w[0] = *p;
  }
#pragma unroll
  for (int v = 0; v < VC; ++v) {
    constexpr int UB = U < 32 ? U : 32; // a 64-bit slot's exponent is in its low dword
    const int bit = (v >> LB) * U;
    const std::uint32_t d = w[bit >> 5] >> (bit & 31);
    if constexpr (UB == 32) {
      E[v] = static_cast<int>(d);
    } else {
      E[v] = static_cast<int>(d << (32 - UB)) >> (32 - UB); // sign-extend the U-bit slot
    }
  }
}

template<typename T, int U, int VPT, bool PACKED, bool NT, int WGS>
__device__ __forceinline__ void
decompress_tile_chunked(const std::uint8_t* __restrict__ in,
                        typename fmt<T>::uw* __restrict__ out,
                        const tplan& p,
                        std::uint64_t tile,
                        std::uint32_t* lds,
                        std::uint32_t* stage)
{
  using uw = typename fmt<T>::uw;
  constexpr int TILE = WGS * VPT;
  // values per chunk: 16 bytes of codes (or all of a thread's values when the
  // codes come from the LDS image), restaged through LDS for whole-line stores
  constexpr int VC = PACKED ? VPT : (VPT < 128 / U ? VPT : 128 / U);
  constexpr int NCH = VPT / VC;
  constexpr int NB = VC * U / 8; // code bytes per chunk
  static_assert(VC * static_cast<int>(sizeof(T)) % 16 == 0, "a chunk is whole 16-byte stores");
  const int t = static_cast<int>(threadIdx.x);
  const std::uint64_t base = tile * TILE;
  const int B = 1 << p.lb;
  const std::uint8_t* pay = in + p.R;

  // block exponents first: they do not depend on the payload, so their loads
  // are in flight together with the payload loads (and, for packed codes,
  // before the barrier that publishes the staged payload)
  int E[NCH][VC];
#pragma unroll
  for (int c = 0; c < NCH; ++c) {
    const std::uint64_t i0 = base + static_cast<std::uint32_t>((c * WGS + t) * VC);
    if (B >= VC) {
      const int e = i0 < p.n ? load_block_exponent<U>(in, i0 >> p.lb) : 0;
#pragma unroll
      for (int v = 0; v < VC; ++v) {
        E[c][v] = e;
      }
    } else {
      // B < VC: whole blocks in the chunk; one vector load of their
      // exponents when the chunk is full and the stream base 16-byte aligned
      // (a run of blocks may start anywhere)
      bool done = false;
      if (i0 + VC <= p.n && (reinterpret_cast<std::uintptr_t>(in) & 15u) == 0) {
        const std::uint64_t k0 = i0 >> p.lb;
        done = true;
        switch (p.lb) {
          case 0:
            if constexpr (VC * (U / 8) <= 16) {
              chunk_exponents_vec<U, VC, 0>(in, k0, E[c]);
            } else {
              done = false;
            }
            break;
          case 1:
            if constexpr (VC >= 2 && (VC >> 1) * (U / 8) <= 16) {
              chunk_exponents_vec<U, VC, 1>(in, k0, E[c]);
            } else {
              done = false;
            }
            break;
          case 2:
            if constexpr (VC >= 4 && (VC >> 2) * (U / 8) <= 16) {
              chunk_exponents_vec<U, VC, 2>(in, k0, E[c]);
            } else {
              done = false;
            }
            break;
          case 3:
            if constexpr (VC >= 8 && (VC >> 3) * (U / 8) <= 16) {
              chunk_exponents_vec<U, VC, 3>(in, k0, E[c]);
            } else {
              done = false;
            }
            break;
          default:
            done = false;
            break;
        }
      }
      if (!done) {
#pragma unroll
        for (int v = 0; v < VC; ++v) {
          E[c][v] = (i0 + v < p.n) ? load_block_exponent<U>(in, (i0 + v) >> p.lb) : 0;
        }
      }
    }
  }

  if (PACKED) {
    const std::uint64_t k0 = base >> p.lb;
    const std::uint32_t TB = static_cast<std::uint32_t>(TILE >> p.lb);
    const std::uint64_t left = p.nb - k0;
    const std::uint32_t blocks = left < TB ? static_cast<std::uint32_t>(left) : TB;
    const std::uint32_t nbytes = blocks * p.CB;
    const std::uint8_t* src = pay + k0 * p.CB;
    const u32x4* s4 = reinterpret_cast<const u32x4*>(src);
    u32x4* l4 = reinterpret_cast<u32x4*>(lds);
    for (std::uint32_t c = t; c < nbytes / 16; c += WGS) {
      l4[c] = ld<NT>(s4 + c);
    }
    const std::uint32_t tail = nbytes & 15u;
    if (static_cast<std::uint32_t>(t) < tail) {
      const std::uint32_t o = nbytes - tail + t;
      reinterpret_cast<std::uint8_t*>(lds)[o] = src[o];
    }
    __syncthreads();
  }

  const std::uint64_t limit = p.nb << p.lb; // payload words
  const bool uniform = B >= VC;              // one exponent per chunk
  const bool narrow = p.bits <= fmt<T>::SB + 1;
  uw code[NCH][VC];
#pragma unroll
  for (int c = 0; c < NCH; ++c) {
    const std::uint32_t local0 = static_cast<std::uint32_t>((c * WGS + t) * VC);
    const std::uint64_t i0 = base + local0;
    if (!PACKED) {
      if (i0 + VC <= limit) {
        std::uint32_t w[NB >= 4 ? NB / 4 : 1];
        load_bytes<NT, NB>(pay + i0 * (U / 8), w);
        if constexpr (NB == 2) {
#pragma unroll
          for (int v = 0; v < VC; ++v) {
            code[c][v] = static_cast<uw>((w[0] >> (8 * v)) & 0xffu);
          }
        } else {
          unpack_words<U, VC>(w, code[c]);
        }
      } else {
        using wt = typename word<U>::type;
#pragma unroll
        for (int v = 0; v < VC; ++v) {
          code[c][v] = (i0 + v < limit) ? static_cast<uw>(reinterpret_cast<const wt*>(pay)[i0 + v]) : uw{ 0 };
        }
      }
    } else if (uniform) {
      // one block: the chunk's codes are consecutive in the bitstream
      std::uint32_t pos = (local0 >> p.lb) * p.CB * 8u +
                          (local0 & static_cast<std::uint32_t>(B - 1)) * static_cast<std::uint32_t>(p.bits);
#pragma unroll
      for (int v = 0; v < VC; ++v) {
        code[c][v] = static_cast<uw>(extract_code<U>(lds, pos));
        pos += static_cast<std::uint32_t>(p.bits);
      }
    } else {
#pragma unroll
      for (int v = 0; v < VC; ++v) {
        const std::uint32_t local = local0 + v;
        const std::uint32_t kl = local >> p.lb;
        const std::uint32_t j = local & static_cast<std::uint32_t>(B - 1);
        const std::uint32_t pos = kl * p.CB * 8u + j * static_cast<std::uint32_t>(p.bits);
        code[c][v] = static_cast<uw>(extract_code<U>(lds, pos));
      }
    }
  }
  uw* sv = reinterpret_cast<uw*>(stage);
#pragma unroll
  for (int c = 0; c < NCH; ++c) {
    const std::uint32_t local0 = static_cast<std::uint32_t>((c * WGS + t) * VC);
    uw y[VC];
    decode_chunk<T, VC>(code[c], E[c], uniform, narrow, p.bits, y);
    store_values<T, VC, false>(sv + local0, y);
  }
  __syncthreads();
  constexpr int CPV = 16 / static_cast<int>(sizeof(T));
  constexpr int NQ = TILE / CPV; // 16-byte chunks of the tile
  const std::uint64_t left = p.n - base;
  const std::uint32_t nvals = left < static_cast<std::uint64_t>(TILE) ? static_cast<std::uint32_t>(left) : TILE;
  const u32x4* s4 = reinterpret_cast<const u32x4*>(stage);
  u32x4* o4 = reinterpret_cast<u32x4*>(out + base);
#pragma unroll
  for (int q0 = 0; q0 < NQ; q0 += WGS) {
    const std::uint32_t q = static_cast<std::uint32_t>(q0 + t);
    if ((q + 1) * CPV <= nvals) {
      st<NT>(s4[q], o4 + q);
    } else if (q * CPV < nvals) {
      for (std::uint32_t v = q * CPV; v < nvals; ++v) {
        out[base + v] = sv[v];
      }
    }
  }
  __syncthreads(); // the stage and the payload image are reused by the next tile of a persistent workgroup
}

// Decompress one tile: staged, the chunked decode with LDS-staged stores;
// otherwise the strided decode.
template<typename T, int U, int VPT, bool PACKED, bool NT, bool STAGE, int WGS = WG>
__device__ __forceinline__ void
decompress_tile(const std::uint8_t* __restrict__ in,
                typename fmt<T>::uw* __restrict__ out,
                const tplan& p,
                std::uint64_t tile,
                std::uint32_t* lds,
                std::uint32_t* stage)
{
  if constexpr (STAGE) {
    decompress_tile_chunked<T, U, VPT, PACKED, NT, WGS>(in, out, p, tile, lds, stage);
  } else {
    decompress_tile_strided<T, U, VPT, PACKED, NT, WGS>(in, out, p, tile, lds);
  }
}

// Zero the exponent pad [nb*U/8, R) of a stream (the reference never writes
// it; a canonical stream has zeros there).
__device__ __forceinline__ void
zero_pad(std::uint8_t* __restrict__ out, std::uint64_t nb, int U, std::uint64_t R, int t, int nthreads)
{
  const std::uint64_t x = nb * static_cast<std::uint64_t>(U / 8);
  for (std::uint64_t i = x + t; i < R; i += nthreads) {
    out[i] = 0;
  }
}

// the exponent pad of a tplan: zpad bytes after exponent slot nb
template<int U>
__device__ __forceinline__ void
zero_plan_pad(std::uint8_t* __restrict__ out, const tplan& p, int t, int nthreads)
{
  std::uint8_t* q = out + p.nb * static_cast<std::uint64_t>(U / 8);
  for (std::uint32_t i = static_cast<std::uint32_t>(t); i < p.zpad; i += static_cast<std::uint32_t>(nthreads)) {
    q[i] = 0;
  }
}

// logical tile of workgroup `bid` such that each of the 8 XCDs (which receive
// workgroups round-robin) processes one contiguous range of tiles
__device__ __forceinline__ std::uint32_t
xcd_tile(std::uint32_t bid, std::uint32_t ntiles)
{
  const std::uint32_t xcd = bid & 7u;
  const std::uint32_t local = bid >> 3;
  const std::uint32_t q = ntiles >> 3;
  const std::uint32_t r = ntiles & 7u;
  return xcd < r ? xcd * (q + 1) + local : r * (q + 1) + (xcd - r) * q + local;
}

} // namespace tile
} // namespace frsz2
