// Thin C entry points over the reference's own CPU codec, compiled from the
// reference header where it lies (robertu94/frsz include/frsz2.hpp, found via
// -I at build time; never copied into this repository).
//
// TEST INFRASTRUCTURE ONLY: the build recipe is oracle/Makefile, its output
// goes to oracle/_ref/ (git-ignored).  tests/golden/make_golden.py uses it to
// write golden vectors, tests use it to pin the C restatement, and bench.py
// times it as the "reference" CPU baseline.  The product never loads it.
//
// REF_DTYPE selects the float type of this translation unit and
// REF_BITS_LO..REF_BITS_HI the bit widths it instantiates, so the template
// instantiations can be split over several objects.
#include "frsz2.hpp"

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>

#ifndef REF_DTYPE
#error "REF_DTYPE must be float or double"
#endif
#ifndef REF_TAG
#error "REF_TAG must name the translation unit"
#endif

#define REF_CAT2(a, b) a##b
#define REF_CAT(a, b) REF_CAT2(a, b)
#define REF_FN(name) REF_CAT(REF_CAT(name, _), REF_TAG)

namespace {

using T = REF_DTYPE;

template<class C>
struct tag
{
  using type = C;
};

// exponent block sizes instantiated: the plugin's list (src/frsz2.cc:86) plus
// the odd sizes the reference's GPU test uses (test/test_frsz_gpu.cu:399-408)
using block_list = frsz::int_list_t<1, 2, 3, 4, 5, 7, 8, 9, 16, 32, 64, 128, 256, 512, 1024>;

template<int b, class F>
bool
with_block(int, F&, frsz::int_list_t<>)
{
  return false;
}

template<int b, class F, int B0, int... Bs>
bool
with_block(int B, F& f, frsz::int_list_t<B0, Bs...>)
{
  if (B == B0) {
    f(tag<frsz::frsz2_compressor<b, B0, T>>{});
    return true;
  }
  return with_block<b>(B, f, frsz::int_list_t<Bs...>{});
}

template<class F, int... Is>
bool
with_bits(int bits, int B, F& f, std::integer_sequence<int, Is...>)
{
  bool found = false;
  using expand = int[];
  (void)expand{ 0, (bits == REF_BITS_LO + Is ? (found = with_block<REF_BITS_LO + Is>(B, f, block_list{}), 0) : 0)... };
  return found;
}

template<class F>
bool
dispatch(int bits, int B, F&& f)
{
  return with_bits(bits, B, f, std::make_integer_sequence<int, REF_BITS_HI - REF_BITS_LO + 1>{});
}

} // namespace

extern "C" {

// 0 ok, 2 unsupported (bits, B) in this translation unit
int
REF_FN(ref_compressed_size)(int bits, int B, std::uint64_t n, std::uint64_t* out)
{
  return dispatch(bits, B, [&](auto t) {
           using C = typename decltype(t)::type;
           *out = C::compute_compressed_memory_size_byte(n);
         })
           ? 0
           : 2;
}

int
REF_FN(ref_compress)(int bits, int B, const void* in, std::uint64_t n, std::uint8_t* out)
{
  return dispatch(bits, B, [&](auto t) {
           using C = typename decltype(t)::type;
           C c(out, n);
           c.compress_cpu_impl(static_cast<const T*>(in));
         })
           ? 0
           : 2;
}

// compress_cpu_block(k, vals, len) on a view over n values
int
REF_FN(ref_compress_block)(int bits, int B, std::uint64_t n, std::uint64_t k, const void* vals, std::uint64_t len,
                           std::uint8_t* out)
{
  return dispatch(bits, B, [&](auto t) {
           using C = typename decltype(t)::type;
           C c(out, n);
           c.compress_cpu_block(k, static_cast<const T*>(vals), len);
         })
           ? 0
           : 2;
}

// decompress_cpu_block(k, vals) on a view over n values
int
REF_FN(ref_decompress_block)(int bits, int B, std::uint64_t n, std::uint64_t k, const std::uint8_t* in, void* vals)
{
  return dispatch(bits, B, [&](auto t) {
           using C = typename decltype(t)::type;
           const C c(const_cast<std::uint8_t*>(in), n);
           c.decompress_cpu_block(k, static_cast<T*>(vals));
         })
           ? 0
           : 2;
}

int
REF_FN(ref_decompress)(int bits, int B, const std::uint8_t* in, std::uint64_t n, void* out)
{
  return dispatch(bits, B, [&](auto t) {
           using C = typename decltype(t)::type;
           const C c(const_cast<std::uint8_t*>(in), n);
           c.decompress_cpu_impl(static_cast<T*>(out));
         })
           ? 0
           : 2;
}

int
REF_FN(ref_gather)(int bits,
                   int B,
                   const std::uint8_t* in,
                   std::uint64_t n,
                   const std::uint64_t* idx,
                   std::uint64_t m,
                   void* out)
{
  return dispatch(bits, B, [&](auto t) {
           using C = typename decltype(t)::type;
           const C c(const_cast<std::uint8_t*>(in), n);
           T* o = static_cast<T*>(out);
           for (std::uint64_t i = 0; i < m; ++i) {
             o[i] = c.decompress_cpu_element(idx[i]);
           }
         })
           ? 0
           : 2;
}

// single-value helpers of include/frsz2.hpp:312-399
std::int32_t
REF_FN(ref_exponent)(T x)
{
  return frsz::detail::fp::exponent(x);
}

std::uint64_t
REF_FN(ref_significand)(T x)
{
  return frsz::detail::fp::significand(x);
}

int
REF_FN(ref_sign)(T x)
{
  return frsz::detail::fp::sign(x) ? 1 : 0;
}

std::uint64_t
REF_FN(ref_floating_to_fixed)(T x, std::int16_t e)
{
  return frsz::detail::fp::floating_to_fixed(x, e);
}

T
REF_FN(ref_fixed_to_floating)(std::uint64_t fixed, std::int16_t e)
{
  return frsz::detail::fp::fixed_to_floating<T>(static_cast<frsz::detail::scaled_t<T>>(fixed), e);
}

#ifdef REF_WITH_DISPATCH
// runtime dispatch of include/frsz2.hpp:1296-1347 with the plugin's block
// list (src/frsz2.cc:86) and a short bits list; writes the exception text
// (if any) to msg.
int
REF_FN(ref_dispatch_compress)(int bits, int B, const void* in, std::uint64_t n, std::uint8_t* out, char* msg,
                              int msglen)
{
  using exp_list = frsz::int_list_t<2, 4, 8, 16, 32, 64, 128, 256, 512, 1024>;
  try {
    frsz::dispatch_frsz2_compression<T>(frsz::int_list_t<4, 8, 12, 16, 21, 32>{},
                                        bits,
                                        exp_list{},
                                        B,
                                        const_cast<T*>(static_cast<const T*>(in)),
                                        static_cast<std::size_t>(n),
                                        out);
  } catch (std::exception const& ex) {
    std::strncpy(msg, ex.what(), static_cast<std::size_t>(msglen - 1));
    msg[msglen - 1] = '\0';
    return 2;
  }
  return 0;
}
#endif

} // extern "C"
