// Drop-in C++ API of the FRSZ2 codec, running on MI355X through the C ABI of
// libfrsz2_hip.so (include/frsz2_capi.h).
//
// Mirrors the public surface of robertu94/frsz include/frsz2.hpp:
//   frsz::frsz2_compressor<bits, B, T>   non-owning view over a stream buffer
//                                        (include/frsz2.hpp:504-794) with the
//                                        same constants, size functions and
//                                        method names
//   frsz::int_list_t, frsz::dispatch_frsz2_compression / _decompression
//                                        runtime (bits, B) dispatch with the
//                                        same std::runtime_error messages
//                                        (include/frsz2.hpp:1292-1347)
// The *_cpu_* methods keep their host-memory meaning (buffers in host
// memory) but do the work with the HIP kernels: inputs are staged to the GPU
// and results copied back.  The *_gpu methods take device pointers and a
// stream and run asynchronously.  Every method reports a failure of the
// library by throwing std::runtime_error with the library's message.
//
// compress_cpu_block / decompress_cpu_block keep their per-block meaning;
// compress_gpu_blocks / decompress_gpu_blocks do the same for a run of
// blocks in device memory.  decompress_cpu_element decodes its one value in
// place from the host stream (two or three loads, the same integer code the
// kernels run); decompress_cpu_elements sends a batch of indices to the GPU.
//
// The public static helpers of the reference view are kept:
// compressed_to_uint_fp / uint_fp_to_compressed and write_shift_value /
// retrieve_shift_value (include/frsz2.hpp:550-644).
//
// Compiled by hipcc, the header also offers the reference's device-side API:
//   frsz::compress_gpu<C, block_size><<<grid, block_size>>>(d_values, view)
//   frsz::decompress_gpu<C><<<grid, threads>>>(d_output, view)
//                                        the free kernels of
//                                        include/frsz2.hpp:1258-1288, correct
//                                        for any 1-D grid (grid-stride loops)
//   view.compress_gpu_function<block_size>(global_value_idx, value)
//                                        the cooperative per-value compress
//                                        of include/frsz2.hpp:916-1094
//   view.decompress_gpu_element(idx)     include/frsz2.hpp:1096-1104
// and, beyond the reference, compress_gpu_block / decompress_gpu_block, the
// per-block compress and decode by one thread.
#ifndef FRSZ2_MI355X_HPP
#define FRSZ2_MI355X_HPP

#include "frsz2_capi.h"
#include "frsz2_device.hpp"
#if defined(__HIPCC__)
#include "frsz2_tile.hpp"
#include "frsz2_tuning.hpp"
#endif

#include <algorithm>
#include <climits>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <type_traits>

namespace frsz {

namespace detail {

template<std::size_t N>
using storage_t = std::conditional_t<
  (N <= 8),
  std::uint8_t,
  std::conditional_t<(N <= 16), std::uint16_t, std::conditional_t<(N <= 32), std::uint32_t, std::uint64_t>>>;

template<class T>
using scaled_t = std::conditional_t<sizeof(T) == 8, std::uint64_t, std::uint32_t>;

template<class T>
constexpr int
dtype_code()
{
  static_assert(std::is_same<T, float>::value || std::is_same<T, double>::value, "float or double only");
  return std::is_same<T, float>::value ? FRSZ2_FLOAT32 : FRSZ2_FLOAT64;
}

inline void
throw_if(int rc)
{
  if (rc != FRSZ2_OK) {
    throw std::runtime_error(frsz2_last_error());
  }
}

constexpr std::size_t
ceildiv(std::size_t a, std::size_t b)
{
  return a / b + (a % b != 0);
}

#if defined(__HIPCC__)
// The library's tile kernels as the kernel templates run them
// (frsz2_tile.hpp): tiles of WGS threads in the measured variant for C's (dtype, storage
// width, packed) (frsz2_tuning.hpp), its values per thread raised until a
// tile holds an exponent block.  ok: C's block size is a power of two that a
// WGS x 16-value tile holds.
template<class C, int OP, int WGS>
struct tile_shape
{
  using T = typename C::fp_type;
  static constexpr int U = C::uint_compressed_size_bit;
  static constexpr int B = static_cast<int>(C::max_exp_block_size);
  static constexpr bool packed = C::bits_per_value < U;
  static constexpr int v =
    frsz2::default_variant(OP, sizeof(T) == 4 ? 0 : 1, U == 8 ? 0 : U == 16 ? 1 : U == 32 ? 2 : 3, packed ? 1 : 0);
  // 2 values per thread exist for float64 with U >= 16 only (one 16-byte chunk)
  static constexpr int vpt0 = frsz2::venc_vpt(v) == 2 && !(sizeof(T) == 8 && U >= 16) ? 4 : frsz2::venc_vpt(v);
  static constexpr int vpt = WGS * vpt0 >= B       ? vpt0
                             : 2 * WGS * vpt0 >= B ? 2 * vpt0
                             : 4 * WGS * vpt0 >= B ? 4 * vpt0
                                                   : 8 * vpt0;
  static constexpr bool ok = (B & (B - 1)) == 0 && vpt <= 16 && WGS * vpt >= B;
  static constexpr int lb = B <= 1 ? 0 : B <= 2 ? 1 : B <= 4 ? 2 : B <= 8 ? 3 : B <= 16 ? 4 : B <= 32 ? 5
                            : B <= 64 ? 6 : B <= 128 ? 7 : B <= 256 ? 8 : B <= 512 ? 9 : B <= 1024 ? 10
                            : B <= 2048 ? 11 : 12;
  // decompress: the staged decode unless a float64 tile had to grow to 16
  // values per thread (the strided decode is faster there)
  static constexpr bool stage =
    OP == 1 && frsz2::venc_stage(v) != 0 && !(sizeof(T) == 8 && vpt == 16 && vpt0 < 16);
  static constexpr int tile = WGS * vpt;
  // LDS dwords: compress, the payload image; decompress, the payload image
  // and the output stage
  static constexpr int lds_c = packed ? frsz2::tile::lds_dwords_u<U, (vpt <= 16 ? vpt : 16), WGS>() : 1;
  static constexpr int lds_d_image = packed ? frsz2::tile::lds_dwords_u<U, (vpt <= 16 ? vpt : 16), WGS>() : 0;
  static constexpr int lds_d = lds_d_image + (stage ? frsz2::tile::stage_dwords<T, (vpt <= 16 ? vpt : 16), WGS>() : 0);

  static __device__ frsz2::tile::tplan plan(std::size_t n)
  {
    const frsz2::layout L = frsz2::layout::make(C::bits_per_value, C::max_exp_block_size, n);
    return frsz2::tile::tplan{ L.n,
                               L.nb,
                               L.R,
                               static_cast<std::uint32_t>(L.CB),
                               L.bits,
                               lb,
                               static_cast<std::uint32_t>(L.R - L.nb * static_cast<std::uint64_t>(U / 8)) };
  }
};

// Tile compress of the whole array by the first WGS threads of every
// workgroup, grid-stride over the tiles (lds: tile_shape::lds_c dwords,
// wred: WGS / 64 dwords).  Tile 0 also clears the exponent pad.
template<class C, int WGS>
__device__ void
compress_tiles(const typename C::fp_type* __restrict__ data, std::uint8_t* __restrict__ stream, std::size_t n,
               std::uint32_t* lds, std::uint32_t* wred)
{
  using TS = tile_shape<C, 0, WGS>;
  using uw = typename frsz2::fmt<typename C::fp_type>::uw;
  const frsz2::tile::tplan p = TS::plan(n);
  const std::uint64_t tiles = n == 0 ? 1 : (n + TS::tile - 1) / TS::tile; // an empty stream still has its pad
  for (std::uint64_t t = blockIdx.x; t < tiles; t += gridDim.x) {
    frsz2::tile::compress_tile<typename C::fp_type, C::uint_compressed_size_bit, TS::vpt, TS::packed, true, WGS>(
      reinterpret_cast<const uw*>(data), stream, p, t, lds, wred);
    if (t == 0) {
      frsz2::tile::zero_plan_pad<C::uint_compressed_size_bit>(stream, p, static_cast<int>(threadIdx.x), WGS);
    }
  }
}

// Tile decompress of the whole stream by the first WGS threads of every
// workgroup, grid-stride over the tiles (smem: tile_shape::lds_d dwords,
// 16-byte aligned).
template<class C, int WGS>
__device__ void
decompress_tiles(const std::uint8_t* __restrict__ stream, typename C::fp_type* __restrict__ output, std::size_t n,
                 std::uint32_t* smem)
{
  using TS = tile_shape<C, 1, WGS>;
  using uw = typename frsz2::fmt<typename C::fp_type>::uw;
  const frsz2::tile::tplan p = TS::plan(n);
  const std::uint64_t tiles = (n + TS::tile - 1) / TS::tile;
  for (std::uint64_t t = blockIdx.x; t < tiles; t += gridDim.x) {
    frsz2::tile::decompress_tile<typename C::fp_type, C::uint_compressed_size_bit, TS::vpt, TS::packed, true,
                                 TS::stage, WGS>(stream, reinterpret_cast<uw*>(output), p, t, smem,
                                                 smem + TS::lds_d_image);
  }
}

inline __device__ bool
aligned16(const void* p)
{
  return (reinterpret_cast<std::uintptr_t>(p) & 15u) == 0;
}

// Launch shape of the cooperative compress of one workgroup of block_size
// threads: it compresses `seg` exponent blocks per step (block_size / B, at
// least one), through an LDS image of their payload and one key per block.
template<class C, int block_size>
struct wg_shape
{
  static constexpr std::size_t B = C::max_exp_block_size;
  static constexpr std::size_t seg = block_size >= static_cast<int>(B) ? block_size / B : 1;
  static constexpr std::size_t image_dwords = (seg * C::compressed_block_size_byte + 3) / 4;
  // the LDS path is used while the image fits comfortably in a workgroup's LDS
  static constexpr bool lds = image_dwords * 4 + seg * 4 <= 48 * 1024;
};

// One step of the cooperative compress: exponent blocks [k0, k0 + seg) of
// the stream at `stream` (layout L), whose value i (array index) has the bit
// pattern value(i).  Every thread of the 1-D workgroup calls it with the same
// k0 (k0 < L.nb).  Block maxima are atomicMax-ed into keys (one per block),
// codes are OR-ed into an LDS image of the blocks' payload at their bit
// offsets, and the image is written out whole: every payload byte of the
// blocks (codes past n are zero), as compress_cpu_block writes them
// (include/frsz2.hpp:1108-1179).
template<class C, int block_size, class Load>
__device__ void
wg_compress_step(std::uint8_t* __restrict__ stream,
                 const frsz2::layout& L,
                 std::uint64_t k0,
                 Load value,
                 std::uint32_t* keys,
                 std::uint32_t* image)
{
  using T = typename C::fp_type;
  using S = wg_shape<C, block_size>;
  constexpr int U = C::uint_compressed_size_bit;
  constexpr std::uint32_t B = static_cast<std::uint32_t>(S::B);
  constexpr std::uint32_t CB = static_cast<std::uint32_t>(C::compressed_block_size_byte);
  const std::uint32_t t = threadIdx.x;
  const std::uint32_t nt = blockDim.x;
  const std::uint64_t first = k0 * B;
  const std::uint64_t left = L.n - first;
  const std::uint32_t cnt = left < S::seg * B ? static_cast<std::uint32_t>(left) : static_cast<std::uint32_t>(S::seg * B);
  const std::uint32_t nblk = (cnt + B - 1) / B;
  const std::uint32_t nbytes = nblk * CB;
  for (std::uint32_t i = t; i < nblk; i += nt) {
    keys[i] = 0;
  }
  for (std::uint32_t i = t; i < (nbytes + 3) / 4; i += nt) {
    image[i] = 0;
  }
  __syncthreads();
  for (std::uint32_t i = t; i < cnt; i += nt) {
    atomicMax(&keys[i / B], frsz2::max_key<T>(value(first + i)));
  }
  __syncthreads();
  for (std::uint32_t b = t; b < nblk; b += nt) {
    frsz2::store_block_exponent<U>(stream, k0 + b, frsz2::block_exponent_from_key<T, U>(keys[b]));
  }
  const std::uint32_t bits = static_cast<std::uint32_t>(C::bits_per_value);
  for (std::uint32_t i = t; i < cnt; i += nt) {
    const std::uint32_t b = i / B;
    const int E = frsz2::block_exponent_from_key<T, U>(keys[b]);
    const std::uint64_t code = static_cast<std::uint64_t>(frsz2::encode_value<T>(value(first + i), E, C::bits_per_value));
    if (code != 0) {
      // the code's bits [0, bits) go to image bits [pos, pos + bits)
      const std::uint32_t pos = b * CB * 8u + (i - b * B) * bits;
      const std::uint32_t q = pos >> 5;
      const std::uint32_t o = pos & 31u;
      const std::uint64_t lo = code << o;
      atomicOr(&image[q], static_cast<std::uint32_t>(lo));
      if (o + bits > 32) {
        atomicOr(&image[q + 1], static_cast<std::uint32_t>(lo >> 32));
      }
      if (o + bits > 64) {
        atomicOr(&image[q + 2], static_cast<std::uint32_t>(code >> (64 - o)));
      }
    }
  }
  __syncthreads();
  std::uint8_t* dst = stream + L.R + k0 * CB;
  if (CB % 4 == 0 && (reinterpret_cast<std::uintptr_t>(dst) & 3u) == 0) {
    std::uint32_t* d4 = reinterpret_cast<std::uint32_t*>(dst);
    for (std::uint32_t i = t; i < nbytes / 4; i += nt) {
      d4[i] = image[i];
    }
  } else {
    const std::uint8_t* src = reinterpret_cast<const std::uint8_t*>(image);
    for (std::uint32_t i = t; i < nbytes; i += nt) {
      dst[i] = src[i];
    }
  }
  __syncthreads(); // keys and image are reused by the next step
}
#endif

} // namespace detail

template<int bits_per_value_, int max_exp_block_size_, class FpType_>
struct frsz2_compressor
{
  constexpr static auto bits_per_value = bits_per_value_;
  constexpr static auto max_exp_block_size = max_exp_block_size_;
  using fp_type = FpType_;
  using exp_type = std::make_signed_t<detail::storage_t<bits_per_value>>;
  using uint_compressed_type = detail::storage_t<bits_per_value>;
  using uint_fp_type = detail::scaled_t<fp_type>;

  static_assert(bits_per_value <= sizeof(fp_type) * CHAR_BIT,
                "The number of bits per compressed value must be smaller (or equal to) the size of the "
                "original value type!");
  static_assert(0 < max_exp_block_size, "max_exp_block_size must be positive!");
  static_assert(4 <= bits_per_value, "Minimum bits per value is 4 (so at most 2 values share the same byte).");
  static_assert(bits_per_value <= 64, "maximum number for bits_per_value is 64.");

  static constexpr int uint_compressed_size_bit = sizeof(uint_compressed_type) * CHAR_BIT;
  static constexpr int fp_bits = sizeof(fp_type) * CHAR_BIT;
  static constexpr int compressed_block_size_element_count =
    static_cast<int>(detail::ceildiv(static_cast<std::size_t>(max_exp_block_size) * bits_per_value,
                                     sizeof(uint_compressed_type) * CHAR_BIT));
  static constexpr int compressed_block_size_byte =
    compressed_block_size_element_count * static_cast<int>(sizeof(uint_compressed_type));

  // include/frsz2.hpp:534-540: exponent array rounded up to the next multiple
  // of 256 bytes (always 1..256 bytes of pad)
  static constexpr FRSZ2_HD std::size_t compute_exponent_block_size_byte(std::size_t number_elements)
  {
    constexpr std::size_t block_alignment{ 256 };
    const std::size_t exponent_element_size =
      detail::ceildiv(number_elements, max_exp_block_size) * sizeof(exp_type);
    return exponent_element_size + block_alignment - exponent_element_size % block_alignment;
  }

  // include/frsz2.hpp:542-548
  static constexpr FRSZ2_HD std::size_t compute_compressed_memory_size_byte(std::size_t number_elements)
  {
    return compute_exponent_block_size_byte(number_elements) +
           detail::ceildiv(number_elements, max_exp_block_size) * compressed_block_size_byte;
  }

  // ---- code <-> word helpers (public static members of the reference view,
  // include/frsz2.hpp:550-644).  A code is the top bits_per_value bits of a
  // W-bit word (W = bits of fp_type); in the payload it is stored as the low
  // bits_per_value bits of a U-bit word (U = uint_compressed_size_bit), at a
  // bit offset within that word, continuing into the next word when it does
  // not fit.

  static constexpr FRSZ2_HD uint_fp_type compressed_to_uint_fp(const uint_compressed_type& code)
  {
    return static_cast<uint_fp_type>(static_cast<uint_fp_type>(code) << (fp_bits - bits_per_value));
  }

  static constexpr FRSZ2_HD uint_compressed_type uint_fp_to_compressed(const uint_fp_type& word)
  {
    return static_cast<uint_compressed_type>(word >> (fp_bits - bits_per_value));
  }

  // Split the code of `input` at bit_offset of a word: first_val gets the
  // bits that fall in the word, second_val those that spill into the next one
  // (0 when none do).  With bits_per_value == U the code fills the word:
  // first_val = the code, and bit_offset and second_val are left alone.
  static constexpr FRSZ2_HD void write_shift_value(const int bit_offset,
                                                   const uint_fp_type& input,
                                                   uint_compressed_type& first_val,
                                                   uint_compressed_type& second_val)
  {
    const uint_compressed_type code = uint_fp_to_compressed(input);
    if (uint_compressed_size_bit == bits_per_value) {
      first_val = code;
      return;
    }
    first_val = static_cast<uint_compressed_type>(static_cast<std::uint64_t>(code) << bit_offset);
    second_val = bit_offset + bits_per_value > uint_compressed_size_bit
                   ? static_cast<uint_compressed_type>(static_cast<std::uint64_t>(code) >>
                                                       (uint_compressed_size_bit - bit_offset))
                   : uint_compressed_type{};
  }

  // The code at bit_offset of first_val (continued in second_val when it
  // spills over), as a W-bit word; bits of the words above the code drop out.
  static constexpr FRSZ2_HD uint_fp_type retrieve_shift_value(const int bit_offset,
                                                              const uint_compressed_type& first_val,
                                                              const uint_compressed_type& second_val)
  {
    if (uint_compressed_size_bit == bits_per_value) {
      return compressed_to_uint_fp(first_val);
    }
    std::uint64_t r = static_cast<std::uint64_t>(first_val) >> bit_offset;
    if (bit_offset + bits_per_value > uint_compressed_size_bit) {
      r |= static_cast<std::uint64_t>(second_val) << (uint_compressed_size_bit - bit_offset);
    }
    return compressed_to_uint_fp(static_cast<uint_compressed_type>(r));
  }

  // The same from a pointer to the first word; the second word is read only
  // when the code spills into it.
  static constexpr FRSZ2_HD uint_fp_type retrieve_shift_value(const int bit_offset, const uint_compressed_type* values)
  {
    return uint_compressed_size_bit == bits_per_value || bit_offset + bits_per_value <= uint_compressed_size_bit
             ? retrieve_shift_value(bit_offset, values[0], values[0])
             : retrieve_shift_value(bit_offset, values[0], values[1]);
  }

  FRSZ2_HD frsz2_compressor(std::uint8_t* data, std::size_t total_elements)
    : data_{ data }
    , exponents_{ reinterpret_cast<exp_type*>(data) }
    , compressed_{ reinterpret_cast<uint_compressed_type*>(data + compute_exponent_block_size_byte(total_elements)) }
    , total_elements_{ total_elements }
  {
  }

  FRSZ2_HD uint_compressed_type* get_compressed_data() { return compressed_; }
  FRSZ2_HD const uint_compressed_type* get_compressed_data() const { return compressed_; }
  FRSZ2_HD const uint_compressed_type* get_const_compressed_data() const { return compressed_; }
  FRSZ2_HD exp_type* get_exponent_data() { return exponents_; }
  FRSZ2_HD const exp_type* get_exponent_data() const { return exponents_; }
  FRSZ2_HD const exp_type* get_const_exponent_data() const { return exponents_; }
  FRSZ2_HD std::size_t get_total_elements() const { return total_elements_; }

  // ---- host memory (the view's buffer and the values are host memory)

  // compress_cpu_impl, include/frsz2.hpp:1181-1193
  int compress_cpu_impl(const fp_type* data)
  {
    detail::throw_if(frsz2_compress_host(detail::dtype_code<fp_type>(), bits_per_value, max_exp_block_size, data,
                                         total_elements_, data_, stream_bytes()));
    return 0;
  }

  // decompress_cpu_impl, include/frsz2.hpp:1242-1249
  int decompress_cpu_impl(fp_type* output) const
  {
    detail::throw_if(frsz2_decompress_host(detail::dtype_code<fp_type>(), bits_per_value, max_exp_block_size,
                                           data_, stream_bytes(), total_elements_, output));
    return 0;
  }

  // decompress_cpu_element, include/frsz2.hpp:1199-1214: one value of the
  // host stream, decoded in place (its block exponent and the one or two
  // payload words holding its code) by the integer decode the kernels use.
  // An index >= n yields 0 instead of reading past the stream.
  fp_type decompress_cpu_element(std::size_t idx) const
  {
    if (idx >= total_elements_) {
      return fp_type{};
    }
    const frsz2::layout L = frsz2::layout::make(bits_per_value, max_exp_block_size, total_elements_);
    return frsz2::decode_element<fp_type, uint_compressed_size_bit>(data_, L, idx);
  }

  // decompress_gpu_value, include/frsz2.hpp:665-700: the value of one code
  // (its low bits_per_value bits; higher bits are ignored) in an exponent
  // block of the given exponent, decoded as the host path decodes it (a zero
  // code gives +-2^(exponent - 31), include/frsz2.hpp:1199-1214).  Callable
  // on the host and in kernels, for every bits_per_value.
  static FRSZ2_HD fp_type decompress_gpu_value(const exp_type exponent, const uint_compressed_type extracted_value)
  {
    using uw = typename frsz2::fmt<fp_type>::uw;
    return __builtin_bit_cast(
      fp_type, frsz2::decode_value<fp_type>(static_cast<uw>(extracted_value), static_cast<int>(exponent), bits_per_value));
  }

  // decompress_gpu_element_impl, include/frsz2.hpp:646-775: value idx given
  // the payload words (block k starts at word k *
  // compressed_block_size_element_count, as get_compressed_data() lays them
  // out) and the exponent slots (get_exponent_data()).  Reads exponent slot
  // idx / B and the one or two words holding the code, nothing past the
  // block.  Callable on the host and in kernels.
  template<typename IndexType>
  static FRSZ2_HD fp_type decompress_gpu_element_impl(const uint_compressed_type* __restrict__ compressed,
                                                      const exp_type* __restrict__ exponents, const IndexType idx)
  {
    const std::uint64_t i = static_cast<std::uint64_t>(idx);
    const std::uint64_t k = i / static_cast<std::uint64_t>(max_exp_block_size);
    const std::uint64_t j = i - k * static_cast<std::uint64_t>(max_exp_block_size);
    const auto* blk = reinterpret_cast<const std::uint8_t*>(
      compressed + k * static_cast<std::uint64_t>(compressed_block_size_element_count));
    const std::uint64_t code = frsz2::load_code<uint_compressed_size_bit>(blk, j, bits_per_value);
    return decompress_gpu_value(exponents[k], static_cast<uint_compressed_type>(code));
  }

  // compress_cpu_block, include/frsz2.hpp:1108-1179: block exp_block_id from
  // its first exp_block_elements values (the rest of the block's codes are
  // written as zero)
  void compress_cpu_block(std::size_t exp_block_id, const fp_type* data,
                          std::size_t exp_block_elements = max_exp_block_size)
  {
    detail::throw_if(frsz2_compress_blocks_host(detail::dtype_code<fp_type>(), bits_per_value, max_exp_block_size,
                                                total_elements_, exp_block_id, data, exp_block_elements, data_,
                                                stream_bytes()));
  }

  // decompress_cpu_block, include/frsz2.hpp:1217-1240
  void decompress_cpu_block(std::size_t exp_block_id, fp_type* output) const
  {
    const std::size_t first = exp_block_id * max_exp_block_size;
    const std::size_t count =
      first < total_elements_ ? std::min<std::size_t>(max_exp_block_size, total_elements_ - first) : 0;
    detail::throw_if(frsz2_decompress_blocks_host(detail::dtype_code<fp_type>(), bits_per_value, max_exp_block_size,
                                                  total_elements_, exp_block_id, data_, stream_bytes(), count,
                                                  output));
  }

  // random-access decode of many indices (host memory)
  void decompress_cpu_elements(const std::uint64_t* idx, std::size_t m, fp_type* out) const
  {
    detail::throw_if(frsz2_gather_host(detail::dtype_code<fp_type>(), bits_per_value, max_exp_block_size, data_,
                                       stream_bytes(), total_elements_, idx, m, out));
  }

  // ---- device memory (the view's buffer and the values are device memory)

  // replaces the compress_gpu kernel, include/frsz2.hpp:1258-1274
  int compress_gpu(const fp_type* d_data, frsz2_stream_t stream = nullptr)
  {
    detail::throw_if(frsz2_compress(detail::dtype_code<fp_type>(), bits_per_value, max_exp_block_size, d_data,
                                    total_elements_, data_, stream_bytes(), stream));
    return 0;
  }

  // replaces the decompress_gpu kernel, include/frsz2.hpp:1276-1288
  int decompress_gpu(fp_type* d_output, frsz2_stream_t stream = nullptr) const
  {
    detail::throw_if(frsz2_decompress(detail::dtype_code<fp_type>(), bits_per_value, max_exp_block_size, data_,
                                      stream_bytes(), total_elements_, d_output, stream));
    return 0;
  }

  void gather_gpu(const std::uint64_t* d_idx, std::size_t m, fp_type* d_out, frsz2_stream_t stream = nullptr) const
  {
    detail::throw_if(frsz2_gather(detail::dtype_code<fp_type>(), bits_per_value, max_exp_block_size, data_,
                                  stream_bytes(), total_elements_, d_idx, m, d_out, stream));
  }

  // a run of blocks in device memory: values [first_block*B, +count)
  void compress_gpu_blocks(std::size_t first_block, const fp_type* d_values, std::size_t count,
                           frsz2_stream_t stream = nullptr)
  {
    detail::throw_if(frsz2_compress_blocks(detail::dtype_code<fp_type>(), bits_per_value, max_exp_block_size,
                                           total_elements_, first_block, d_values, count, data_, stream_bytes(),
                                           stream));
  }

  void decompress_gpu_blocks(std::size_t first_block, std::size_t count, fp_type* d_values,
                             frsz2_stream_t stream = nullptr) const
  {
    detail::throw_if(frsz2_decompress_blocks(detail::dtype_code<fp_type>(), bits_per_value, max_exp_block_size,
                                             total_elements_, first_block, data_, stream_bytes(), count, d_values,
                                             stream));
  }

#if defined(__HIPCC__)
  // decompress_gpu_element, include/frsz2.hpp:1096-1104: decode one value
  // inside a kernel (the view is passed to the kernel by value)
  __device__ fp_type decompress_gpu_element(std::size_t idx) const
  {
    const frsz2::layout L = frsz2::layout::make(bits_per_value, max_exp_block_size, total_elements_);
    return frsz2::decode_element<fp_type, uint_compressed_size_bit>(data_, L, idx);
  }

  // compress_cpu_block inside a kernel, by one thread: block exp_block_id
  // from its first exp_block_elements values (exponent slot and all payload
  // bytes of the block; a kernel that compresses every block also clears the
  // exponent pad with clear_exponent_pad)
  __device__ void compress_gpu_block(std::size_t exp_block_id, const fp_type* values,
                                     std::size_t exp_block_elements = max_exp_block_size)
  {
    const frsz2::layout L = frsz2::layout::make(bits_per_value, max_exp_block_size, total_elements_);
    frsz2::encode_block<fp_type, uint_compressed_size_bit>(data_, L, exp_block_id, values, exp_block_elements);
  }

  // decompress_cpu_block inside a kernel, by one thread
  __device__ void decompress_gpu_block(std::size_t exp_block_id, fp_type* output) const
  {
    const frsz2::layout L = frsz2::layout::make(bits_per_value, max_exp_block_size, total_elements_);
    const std::size_t first = exp_block_id * max_exp_block_size;
    const std::size_t count = total_elements_ - first < max_exp_block_size ? total_elements_ - first
                                                                           : max_exp_block_size;
    frsz2::decode_block<fp_type, uint_compressed_size_bit>(data_, L, exp_block_id, output, count);
  }

  // general_compress_gpu_function, include/frsz2.hpp:797-914: exponent
  // block exponent_block_idx compressed by one 1-D workgroup of at least B
  // threads (B <= 1024), thread t passing value t of the block; every thread
  // calls it with the same exponent_block_idx.  Threads at or past B, and
  // values at or past n, are ignored.  Writes the block's exponent slot and
  // all of its payload bytes in the stream format, the bytes
  // compress_cpu_block writes.  (The reference's version writes the exponent
  // in front of the block's payload, a layout its decoders do not read.)
  __device__ void general_compress_gpu_function(const std::size_t exponent_block_idx, const fp_type fp_input_value)
  {
    static_assert(max_exp_block_size <= 1024, "one workgroup holds the exponent block");
    using shape = detail::wg_shape<frsz2_compressor, max_exp_block_size>;
    using uw = typename frsz2::fmt<fp_type>::uw;
    __shared__ uw vals[max_exp_block_size];
    __shared__ std::uint32_t keys[shape::seg];
    __shared__ std::uint32_t image[shape::image_dwords];
    if (threadIdx.x < max_exp_block_size) {
      vals[threadIdx.x] = __builtin_bit_cast(uw, fp_input_value);
    }
    __syncthreads();
    const frsz2::layout L = frsz2::layout::make(bits_per_value, max_exp_block_size, total_elements_);
    if (exponent_block_idx < L.nb) {
      const std::size_t first = exponent_block_idx * max_exp_block_size;
      detail::wg_compress_step<frsz2_compressor, max_exp_block_size>(
        data_, L, exponent_block_idx, [&](std::size_t i) { return vals[i - first]; }, keys, image);
    }
    __syncthreads(); // vals is reused by the workgroup's next call
  }

  // zero the exponent pad (byte i of it, for i = lane, lane + lanes, ...)
  __device__ void clear_exponent_pad(unsigned lane, unsigned lanes)
  {
    const std::size_t nb = detail::ceildiv(total_elements_, max_exp_block_size);
    const std::size_t end = compute_exponent_block_size_byte(total_elements_);
    for (std::size_t i = nb * sizeof(exp_type) + lane; i < end; i += lanes) {
      data_[i] = 0;
    }
  }

  // compress_gpu_function<block_size>(global_value_idx, value),
  // include/frsz2.hpp:916-1094: every thread of a 1-D workgroup of
  // block_size threads calls it together, thread t with value
  // global_value_idx = base + t of the array (base a multiple of block_size;
  // values at or past n are ignored).  The workgroup then writes the
  // exponent slots and payload blocks of its block_size / B exponent blocks.
  // block_size must be a multiple of B.  It does not write the exponent pad:
  // a kernel that compresses every block also calls clear_exponent_pad.
  template<int block_size>
  __device__ void compress_gpu_function(const std::size_t global_value_idx, const fp_type fp_input_value)
  {
    static_assert(block_size % max_exp_block_size == 0, "block_size must be a multiple of the exponent block size");
    using shape = detail::wg_shape<frsz2_compressor, block_size>;
    using uw = typename frsz2::fmt<fp_type>::uw;
    __shared__ uw vals[block_size];
    __shared__ std::uint32_t keys[shape::seg];
    __shared__ std::uint32_t image[shape::image_dwords];
    const std::size_t base = global_value_idx - threadIdx.x;
    vals[threadIdx.x] = __builtin_bit_cast(uw, fp_input_value);
    __syncthreads();
    const frsz2::layout L = frsz2::layout::make(bits_per_value, max_exp_block_size, total_elements_);
    if (base < total_elements_) {
      detail::wg_compress_step<frsz2_compressor, block_size>(
        data_, L, base / max_exp_block_size, [&](std::size_t i) { return vals[i - base]; }, keys, image);
    }
    __syncthreads(); // vals is reused by the workgroup's next call
  }
#endif

private:
  std::uint64_t stream_bytes() const { return compute_compressed_memory_size_byte(total_elements_); }

  std::uint8_t* const data_;
  exp_type* const exponents_;
  uint_compressed_type* const compressed_;
  const std::size_t total_elements_;
};

#if defined(__HIPCC__)
// compress_gpu<C, block_size>, include/frsz2.hpp:1258-1274: compress the
// view's n values from d_data into its (device) stream.  Correct for any 1-D
// grid and any blockDim.x <= block_size: workgroups step through groups of
// max(1, block_size / B) exponent blocks, grid-stride, each group compressed
// cooperatively through LDS; workgroup 0 also clears the exponent pad.  The
// stream bytes equal compress_cpu_impl's.  (Blocks whose payload image would
// not fit the workgroup's LDS are compressed by one thread per block.)
template<typename Frsz2Compressor, int block_size>
__global__ __launch_bounds__(block_size) void compress_gpu(const typename Frsz2Compressor::fp_type* __restrict__ data,
                                                           Frsz2Compressor compressor)
{
  using C = Frsz2Compressor;
  using S = detail::wg_shape<C, block_size>;
  using uw = typename frsz2::fmt<typename C::fp_type>::uw;
  std::uint8_t* stream = reinterpret_cast<std::uint8_t*>(compressor.get_exponent_data());
  const std::size_t n = compressor.get_total_elements();
  // the library's tile compress, grid-stride over the tiles, with 16-byte
  // aligned values and stream: the first wave (64 threads) of every
  // workgroup compresses one tile at a time, and the workgroup's other waves
  // have no work and end here (a barrier waits only for the waves that are
  // still running).  One-wave tiles measured faster under the reference's
  // launch shapes than 128- or 256-thread tiles (a workgroup's spare waves
  // then end at once, and more workgroups hold a tile).
  using TS = detail::tile_shape<C, 0, 64>;
  if constexpr (block_size >= 64 && TS::ok) {
    if (blockDim.x >= 64 && detail::aligned16(data) && detail::aligned16(stream)) {
      if (threadIdx.x >= 64) {
        return;
      }
      __shared__ __attribute__((aligned(16))) std::uint32_t tlds[TS::lds_c];
      __shared__ std::uint32_t twred[1];
      detail::compress_tiles<C, 64>(data, stream, n, tlds, twred);
      return;
    }
  }
  if (blockIdx.x == 0) {
    compressor.clear_exponent_pad(threadIdx.x, blockDim.x);
  }
  const frsz2::layout L = frsz2::layout::make(C::bits_per_value, C::max_exp_block_size, n);
  if constexpr (S::lds) {
    __shared__ std::uint32_t keys[S::seg];
    __shared__ std::uint32_t image[S::image_dwords > 0 ? S::image_dwords : 1];
    const std::uint64_t groups = (L.nb + S::seg - 1) / S::seg;
    for (std::uint64_t g = blockIdx.x; g < groups; g += gridDim.x) {
      detail::wg_compress_step<C, block_size>(
        stream, L, g * S::seg, [&](std::size_t i) { return __builtin_bit_cast(uw, data[i]); }, keys, image);
    }
  } else {
    const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * blockDim.x;
    for (std::uint64_t k = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; k < L.nb; k += stride) {
      const std::size_t first = k * C::max_exp_block_size;
      compressor.compress_gpu_block(k, data + first, n - first < C::max_exp_block_size ? n - first : C::max_exp_block_size);
    }
  }
}

// decompress_gpu<C>, include/frsz2.hpp:1276-1288: value i of the view's
// stream into output[i], grid-stride, for any 1-D grid.
template<typename Frsz2Compressor>
__global__ void
decompress_gpu(typename Frsz2Compressor::fp_type* const __restrict__ output, Frsz2Compressor compressor)
{
  using C = Frsz2Compressor;
  const std::size_t n = compressor.get_total_elements();
  // the library's tile decompress, grid-stride over the tiles, with 16-byte
  // aligned stream and output: the first wave of every workgroup decodes
  // one tile at a time; the other waves end here
  using TS = detail::tile_shape<C, 1, 64>;
  if constexpr (TS::ok) {
    const std::uint8_t* stream = reinterpret_cast<const std::uint8_t*>(compressor.get_const_exponent_data());
    if (blockDim.x >= 64 && detail::aligned16(stream) && detail::aligned16(output)) {
      if (threadIdx.x >= 64) {
        return;
      }
      __shared__ __attribute__((aligned(16))) std::uint32_t tsmem[TS::lds_d > 0 ? TS::lds_d : 4];
      detail::decompress_tiles<C, 64>(stream, output, n, tsmem);
      return;
    }
  }
  const std::size_t stride = static_cast<std::size_t>(gridDim.x) * blockDim.x;
  for (std::size_t i = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    output[i] = compressor.decompress_gpu_element(i);
  }
}
#endif

template<int... values>
struct int_list_t
{};

namespace detail {

template<int... Vs>
constexpr bool
in_list(int v, int_list_t<Vs...>)
{
  bool found = false;
  using expand = int[];
  (void)expand{ 0, (found = found || (v == Vs), 0)... };
  return found;
}

// include/frsz2.hpp:1296-1344: bits are checked first, then the block size,
// each against the caller's list, with the reference's messages
template<class BitList, class ExpList>
void
check_lists(BitList bl, int bits, ExpList el, int block)
{
  if (!in_list(bits, bl)) {
    throw std::runtime_error("Unsupported number of bits: " + std::to_string(bits));
  }
  if (!in_list(block, el)) {
    throw std::runtime_error("Unsupported exponent block size: " + std::to_string(block));
  }
}

} // namespace detail

// dispatch_frsz2_compression<T>(bit list, bits, block list, B, data, n, out):
// host memory in and out, like the reference (include/frsz2.hpp:1346)
template<class FpType, int... bit_values, int... exp_block_sizes>
void
dispatch_frsz2_compression(int_list_t<bit_values...> bl,
                           int bits,
                           int_list_t<exp_block_sizes...> el,
                           int block,
                           FpType* data,
                           std::size_t total_elements,
                           std::uint8_t* compressed)
{
  detail::check_lists(bl, bits, el, block);
  std::uint64_t bytes = 0;
  detail::throw_if(frsz2_compressed_size(detail::dtype_code<FpType>(), bits, block, total_elements, &bytes));
  detail::throw_if(
    frsz2_compress_host(detail::dtype_code<FpType>(), bits, block, data, total_elements, compressed, bytes));
}

// include/frsz2.hpp:1347
template<class FpType, int... bit_values, int... exp_block_sizes>
void
dispatch_frsz2_decompression(int_list_t<bit_values...> bl,
                             int bits,
                             int_list_t<exp_block_sizes...> el,
                             int block,
                             FpType* data,
                             std::size_t total_elements,
                             std::uint8_t* compressed)
{
  detail::check_lists(bl, bits, el, block);
  std::uint64_t bytes = 0;
  detail::throw_if(frsz2_compressed_size(detail::dtype_code<FpType>(), bits, block, total_elements, &bytes));
  detail::throw_if(
    frsz2_decompress_host(detail::dtype_code<FpType>(), bits, block, compressed, bytes, total_elements, data));
}

} // namespace frsz

#endif
