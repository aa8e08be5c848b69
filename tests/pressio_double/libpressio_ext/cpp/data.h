// Test double of libpressio_ext/cpp/data.h: a typed, possibly owning,
// N-d buffer in host memory.
#pragma once

#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <functional>
#include <memory>
#include <numeric>
#include <vector>

enum pressio_dtype
{
  pressio_double_dtype,
  pressio_float_dtype,
  pressio_uint8_dtype,
  pressio_uint16_dtype,
  pressio_uint32_dtype,
  pressio_uint64_dtype,
  pressio_int8_dtype,
  pressio_int16_dtype,
  pressio_int32_dtype,
  pressio_int64_dtype,
  pressio_byte_dtype,
  pressio_bool_dtype
};

inline std::size_t
pressio_dtype_size(pressio_dtype d)
{
  switch (d) {
    case pressio_double_dtype:
    case pressio_uint64_dtype:
    case pressio_int64_dtype:
      return 8;
    case pressio_float_dtype:
    case pressio_uint32_dtype:
    case pressio_int32_dtype:
      return 4;
    case pressio_uint16_dtype:
    case pressio_int16_dtype:
      return 2;
    default:
      return 1;
  }
}

struct pressio_data
{
  pressio_data() = default;

  // a zero-filled buffer of its own
  static pressio_data owning(pressio_dtype dtype, std::vector<std::size_t> const& dims)
  {
    pressio_data d(dtype, dims);
    d.storage_ = std::make_shared<std::vector<std::uint8_t>>(d.size_in_bytes());
    d.ptr_ = d.storage_->data();
    return d;
  }

  // a view of the caller's buffer
  static pressio_data nonowning(pressio_dtype dtype, void* data, std::vector<std::size_t> const& dims)
  {
    pressio_data d(dtype, dims);
    d.ptr_ = data;
    return d;
  }

  static pressio_data empty(pressio_dtype dtype, std::vector<std::size_t> const& dims)
  {
    return pressio_data(dtype, dims);
  }

  void* data() const { return ptr_; }
  pressio_dtype dtype() const { return dtype_; }
  std::vector<std::size_t> const& dimensions() const { return dims_; }
  std::size_t num_dimensions() const { return dims_.size(); }
  std::size_t num_elements() const
  {
    return dims_.empty() ? 0
                         : std::accumulate(dims_.begin(), dims_.end(), std::size_t{ 1 }, std::multiplies<>());
  }
  std::size_t size_in_bytes() const { return num_elements() * pressio_dtype_size(dtype_); }
  bool has_data() const { return ptr_ != nullptr; }

private:
  pressio_data(pressio_dtype dtype, std::vector<std::size_t> dims)
    : dtype_(dtype)
    , dims_(std::move(dims))
  {
  }

  pressio_dtype dtype_ = pressio_byte_dtype;
  std::vector<std::size_t> dims_;
  std::shared_ptr<std::vector<std::uint8_t>> storage_;
  void* ptr_ = nullptr;
};
