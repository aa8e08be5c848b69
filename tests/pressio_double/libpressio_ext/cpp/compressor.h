// Test double of libpressio_ext/cpp/compressor.h: the compressor plugin
// base class.  The public calls clear the error state and forward to the
// plugin's *_impl overrides, as LibPressio's do.
#pragma once

#include "data.h"
#include "options.h"

#include <memory>
#include <string>

class libpressio_compressor_plugin
{
public:
  virtual ~libpressio_compressor_plugin() = default;

  int compress(const pressio_data* input, pressio_data* output)
  {
    clear_error();
    return compress_impl(input, output);
  }
  int decompress(const pressio_data* input, pressio_data* output)
  {
    clear_error();
    return decompress_impl(input, output);
  }
  pressio_options get_options() const { return get_options_impl(); }
  int set_options(pressio_options const& o)
  {
    clear_error();
    return set_options_impl(o);
  }
  pressio_options get_configuration() const { return get_configuration_impl(); }
  pressio_options get_documentation() const { return get_documentation_impl(); }
  pressio_options get_metrics_results() const { return get_metrics_results_impl(); }

  int error_code() const { return error_code_; }
  const char* error_msg() const { return error_msg_.c_str(); }

  virtual int major_version() const = 0;
  virtual int minor_version() const = 0;
  virtual int patch_version() const = 0;
  virtual const char* version() const = 0;
  virtual const char* prefix() const = 0;
  virtual std::shared_ptr<libpressio_compressor_plugin> clone() = 0;

protected:
  virtual pressio_options get_options_impl() const = 0;
  virtual pressio_options get_configuration_impl() const = 0;
  virtual pressio_options get_documentation_impl() const = 0;
  virtual int set_options_impl(pressio_options const&) = 0;
  virtual int compress_impl(const pressio_data* input, pressio_data* output) = 0;
  virtual int decompress_impl(const pressio_data* input, pressio_data* output) = 0;
  virtual pressio_options get_metrics_results_impl() const { return {}; }

  int set_error(int code, std::string const& msg)
  {
    error_code_ = code;
    error_msg_ = msg;
    return code;
  }

  template<class T>
  void set(pressio_options& o, std::string const& key, T const& v) const
  {
    pressio_options_set(o, key, v);
  }

  template<class T>
  int get(pressio_options const& o, std::string const& key, T* v) const
  {
    return pressio_options_get(o, key, v);
  }

private:
  void clear_error()
  {
    error_code_ = 0;
    error_msg_.clear();
  }

  int error_code_ = 0;
  std::string error_msg_;
};
