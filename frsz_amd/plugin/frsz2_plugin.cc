// The "frsz2" LibPressio compressor plugin, running on MI355X through the C
// ABI of libfrsz2_hip.so (include/frsz2_capi.h).
//
// Drop-in for the plugin of robertu94/frsz src/frsz2.cc:25-208:
//   options      frsz2:bits, frsz2:max_exp_block_size (uint64, defaults 32 /
//                32, src/frsz2.cc:28-35, :58-64, :201-202)
//   configuration pressio:thread_safe = multiple, pressio:stability =
//                experimental (:37-43)
//   supported    float bits 4..32, double bits 4..64, block sizes 2, 4, ...,
//                1024 (:83-134); anything else fails with error 2 and the
//                runtime dispatch's message (include/frsz2.hpp:1300, :1326),
//                bits checked first; a dtype other than float / double fails
//                with error 1 "unsupported dtype" (:164, :181)
//   compress     replaces *output with a new byte buffer holding the stream
//                (:152-169); the input is host memory, flattened
//   decompress   dtype and element count come from *output (:171-186)
//   version      0.0.2, prefix "frsz2", clone copies the options (:188-199)
//   registration under "frsz2" (:206-208)
//
// Deliberate differences: the output buffer has the true stream size (the
// reference's output_size, :136-150, is 1..256 bytes short and the
// reference codec writes past it); a compressed input shorter than the
// stream fails with the library's error 3 instead of being read past its
// end; an error leaves *output unchanged.  Library failures other than the
// plugin's own checks report the library's code (3..6) and message.
//
// Builds against an installed LibPressio (libpressio_ext/cpp/*.h,
// std_compat) and links libfrsz2_hip.so.
#include "frsz2_capi.h"
#include "libpressio_ext/cpp/compressor.h"
#include "libpressio_ext/cpp/data.h"
#include "libpressio_ext/cpp/options.h"
#include "libpressio_ext/cpp/pressio.h"
#include "std_compat/memory.h"

#include <cstdint>
#include <string>

namespace libpressio {
namespace frsz2_ns {

class frsz2_compressor_plugin : public libpressio_compressor_plugin
{
public:
  pressio_options get_options_impl() const override
  {
    pressio_options o;
    set(o, "frsz2:bits", bits);
    set(o, "frsz2:max_exp_block_size", max_exp_block_size);
    return o;
  }

  pressio_options get_configuration_impl() const override
  {
    pressio_options o;
    set(o, "pressio:thread_safe", static_cast<std::int32_t>(pressio_thread_safety_multiple));
    set(o, "pressio:stability", "experimental");
    return o;
  }

  pressio_options get_documentation_impl() const override
  {
    pressio_options o;
    set(o, "pressio:description",
        "FRSZ2 fixed-rate compressor: values of a block share its largest exponent, each value keeps the top "
        "frsz2:bits bits of its aligned significand (MI355X kernels)");
    set(o, "frsz2:bits", "bits per value: 4..32 for float, 4..64 for double");
    set(o, "frsz2:max_exp_block_size",
        "values per shared exponent; one of 2, 4, 8, 16, 32, 64, 128, 256, 512, 1024");
    return o;
  }

  int set_options_impl(pressio_options const& o) override
  {
    get(o, "frsz2:bits", &bits);
    get(o, "frsz2:max_exp_block_size", &max_exp_block_size);
    return 0;
  }

  int compress_impl(const pressio_data* input, pressio_data* output) override
  {
    int dtype = 0;
    if (const int rc = check(input->dtype(), &dtype)) {
      return rc;
    }
    const std::uint64_t n = input->num_elements();
    std::uint64_t bytes = 0;
    if (const int rc = frsz2_compressed_size(dtype, as_int(bits), as_int(max_exp_block_size), n, &bytes)) {
      return set_error(rc, frsz2_last_error());
    }
    pressio_data stream = pressio_data::owning(pressio_byte_dtype, { static_cast<std::size_t>(bytes) });
    if (const int rc = frsz2_compress_host(dtype, as_int(bits), as_int(max_exp_block_size), input->data(), n,
                                           stream.data(), bytes)) {
      return set_error(rc, frsz2_last_error());
    }
    *output = std::move(stream);
    return 0;
  }

  int decompress_impl(const pressio_data* input, pressio_data* output) override
  {
    int dtype = 0;
    if (const int rc = check(output->dtype(), &dtype)) {
      return rc;
    }
    if (const int rc = frsz2_decompress_host(dtype, as_int(bits), as_int(max_exp_block_size), input->data(),
                                             input->size_in_bytes(), output->num_elements(), output->data())) {
      return set_error(rc, frsz2_last_error());
    }
    return 0;
  }

  // bytes a compress of n values writes: the stream's true size
  std::uint64_t output_size(pressio_dtype dtype, std::uint64_t n) const
  {
    std::uint64_t bytes = 0;
    const int code = dtype == pressio_float_dtype ? FRSZ2_FLOAT32 : FRSZ2_FLOAT64;
    return frsz2_compressed_size(code, as_int(bits), as_int(max_exp_block_size), n, &bytes) == 0 ? bytes : 0;
  }

  int major_version() const override { return 0; }
  int minor_version() const override { return 0; }
  int patch_version() const override { return 2; }
  const char* version() const override { return "0.0.2"; }
  const char* prefix() const override { return "frsz2"; }

  pressio_options get_metrics_results_impl() const override { return {}; }

  std::shared_ptr<libpressio_compressor_plugin> clone() override
  {
    return compat::make_unique<frsz2_compressor_plugin>(*this);
  }

  std::uint64_t bits = 32;
  std::uint64_t max_exp_block_size = 32;

private:
  // the options reach the dispatch as int (src/frsz2.cc:91-92): the low 32
  // bits, two's complement
  static int as_int(std::uint64_t v) { return static_cast<int>(static_cast<std::uint32_t>(v)); }

  // dtype, then bits, then block size, each against the plugin's lists
  int check(pressio_dtype d, int* code)
  {
    if (d != pressio_float_dtype && d != pressio_double_dtype) {
      return set_error(1, "unsupported dtype");
    }
    *code = d == pressio_float_dtype ? FRSZ2_FLOAT32 : FRSZ2_FLOAT64;
    const int b = as_int(bits);
    const int max_bits = d == pressio_float_dtype ? 32 : 64;
    if (b < 4 || b > max_bits) {
      return set_error(2, "Unsupported number of bits: " + std::to_string(b));
    }
    const int B = as_int(max_exp_block_size);
    if (B < 2 || B > 1024 || (B & (B - 1)) != 0) {
      return set_error(2, "Unsupported exponent block size: " + std::to_string(B));
    }
    return 0;
  }
};

static pressio_register frsz2_plugin_registration(compressor_plugins(), "frsz2", []() {
  return compat::make_unique<frsz2_compressor_plugin>();
});

} // namespace frsz2_ns
} // namespace libpressio

// Link anchor: clients call it before looking the plugin up, so that a
// static link keeps this translation unit (include/frsz.h:1, src/frsz.cc:11-13).
extern "C" void
register_frsz()
{
}
