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
// blocks in device memory.
//
// Compiled by hipcc, the view also offers, for use inside user kernels (the
// view passed by value): decompress_gpu_element(idx)
// (include/frsz2.hpp:1096-1104), and compress_gpu_block /
// decompress_gpu_block, the per-block compress and decode by one thread.
#ifndef FRSZ2_MI355X_HPP
#define FRSZ2_MI355X_HPP

#include "frsz2_capi.h"

#include <algorithm>
#include <climits>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>

#if defined(__HIPCC__)
#include "frsz2_device.hpp"
#endif

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
  static constexpr int compressed_block_size_element_count =
    static_cast<int>(detail::ceildiv(static_cast<std::size_t>(max_exp_block_size) * bits_per_value,
                                     sizeof(uint_compressed_type) * CHAR_BIT));
  static constexpr int compressed_block_size_byte =
    compressed_block_size_element_count * static_cast<int>(sizeof(uint_compressed_type));

  // include/frsz2.hpp:534-540: exponent array rounded up to the next multiple
  // of 256 bytes (always 1..256 bytes of pad)
  static constexpr std::size_t compute_exponent_block_size_byte(std::size_t number_elements)
  {
    constexpr std::size_t block_alignment{ 256 };
    const std::size_t exponent_element_size =
      detail::ceildiv(number_elements, max_exp_block_size) * sizeof(exp_type);
    return exponent_element_size + block_alignment - exponent_element_size % block_alignment;
  }

  // include/frsz2.hpp:542-548
  static constexpr std::size_t compute_compressed_memory_size_byte(std::size_t number_elements)
  {
    return compute_exponent_block_size_byte(number_elements) +
           detail::ceildiv(number_elements, max_exp_block_size) * compressed_block_size_byte;
  }

  frsz2_compressor(std::uint8_t* data, std::size_t total_elements)
    : data_{ data }
    , exponents_{ reinterpret_cast<exp_type*>(data) }
    , compressed_{ reinterpret_cast<uint_compressed_type*>(data + compute_exponent_block_size_byte(total_elements)) }
    , total_elements_{ total_elements }
  {
  }

  uint_compressed_type* get_compressed_data() { return compressed_; }
  const uint_compressed_type* get_compressed_data() const { return compressed_; }
  const uint_compressed_type* get_const_compressed_data() const { return compressed_; }
  exp_type* get_exponent_data() { return exponents_; }
  const exp_type* get_exponent_data() const { return exponents_; }
  const exp_type* get_const_exponent_data() const { return exponents_; }
  std::size_t get_total_elements() const { return total_elements_; }

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

  // decompress_cpu_element, include/frsz2.hpp:1199-1214
  fp_type decompress_cpu_element(std::size_t idx) const
  {
    fp_type out{};
    const std::uint64_t i = idx;
    detail::throw_if(frsz2_gather_host(detail::dtype_code<fp_type>(), bits_per_value, max_exp_block_size, data_,
                                       stream_bytes(), total_elements_, &i, 1, &out));
    return out;
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

  // zero the exponent pad (byte i of it, for i = lane, lane + lanes, ...)
  __device__ void clear_exponent_pad(unsigned lane, unsigned lanes)
  {
    const std::size_t nb = detail::ceildiv(total_elements_, max_exp_block_size);
    const std::size_t end = compute_exponent_block_size_byte(total_elements_);
    for (std::size_t i = nb * sizeof(exp_type) + lane; i < end; i += lanes) {
      data_[i] = 0;
    }
  }
#endif

private:
  std::uint64_t stream_bytes() const { return compute_compressed_memory_size_byte(total_elements_); }

  std::uint8_t* const data_;
  exp_type* const exponents_;
  uint_compressed_type* const compressed_;
  const std::size_t total_elements_;
};

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
