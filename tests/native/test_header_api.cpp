// Exercises the drop-in C++ header (include/frsz2.hpp) the way the
// reference's GPU test drives its own header (test/test_frsz_gpu.cu:316-410):
// the 9-value repeating pattern, double n = 2049 and float n = 2160, with the
// same (bits, B) list, compress + decompress, lossless round trip.  Also the
// device-pointer methods, the per-block methods (host, device runs and in
// kernel), the in-kernel decompress_gpu_element accessor and the runtime
// dispatch with its error messages.
//
// Usage: test_header_api <outdir>
//   writes <outdir>/<float|double>_b<bits>_B<B>.bin (the stream bytes) for
//   pytest to compare with the golden vectors; prints "FAIL ..." lines and
//   exits non-zero on any failed check.
#include "frsz2.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

int g_fail = 0;

#define CHECK(cond, ...)                                                                                     \
  do {                                                                                                       \
    if (!(cond)) {                                                                                           \
      std::printf("FAIL %s:%d: ", __FILE__, __LINE__);                                                       \
      std::printf(__VA_ARGS__);                                                                              \
      std::printf("\n");                                                                                     \
      ++g_fail;                                                                                              \
    }                                                                                                        \
  } while (0)

#define HIP_OK(expr) CHECK((expr) == hipSuccess, "%s", #expr)

template<class C>
__global__ void
element_kernel(C view, std::size_t n, typename C::fp_type* out)
{
  const std::size_t i = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) {
    out[i] = view.decompress_gpu_element(i);
  }
}

// in-kernel per-block compress / decode: one thread per exponent block
template<class C>
__global__ void
block_compress_kernel(C view, const typename C::fp_type* in, std::size_t n)
{
  constexpr std::size_t B = C::max_exp_block_size;
  const std::size_t k = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (blockIdx.x == 0) {
    view.clear_exponent_pad(threadIdx.x, blockDim.x);
  }
  if (k * B < n) {
    view.compress_gpu_block(k, in + k * B, n - k * B < B ? n - k * B : B);
  }
}

template<class C>
__global__ void
block_decode_kernel(C view, typename C::fp_type* out, std::size_t n)
{
  constexpr std::size_t B = C::max_exp_block_size;
  const std::size_t k = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k * B < n) {
    view.decompress_gpu_block(k, out + k * B);
  }
}

// a user kernel built on the cooperative per-value compress: workgroups of
// BS threads step through the array, each thread passing its value
template<class C, int BS>
__global__ __launch_bounds__(BS) void function_kernel(C view, const typename C::fp_type* in, std::size_t n)
{
  if (blockIdx.x == 0) {
    view.clear_exponent_pad(threadIdx.x, blockDim.x);
  }
  const std::size_t end = (n + BS - 1) / BS * BS;
  for (std::size_t base = static_cast<std::size_t>(blockIdx.x) * BS; base < end;
       base += static_cast<std::size_t>(gridDim.x) * BS) {
    const std::size_t i = base + threadIdx.x;
    view.template compress_gpu_function<BS>(i, i < n ? in[i] : typename C::fp_type{});
  }
}

// general_compress_gpu_function: one workgroup per exponent block, thread t
// passing value t of the block; workgroup 0 also clears the exponent pad
template<class C>
__global__ void
general_kernel(C view, const typename C::fp_type* in, std::size_t n)
{
  constexpr std::size_t B = C::max_exp_block_size;
  if (blockIdx.x == 0) {
    view.clear_exponent_pad(threadIdx.x, blockDim.x);
  }
  const std::size_t k = blockIdx.x;
  const std::size_t i = k * B + threadIdx.x;
  view.general_compress_gpu_function(k, threadIdx.x < B && i < n ? in[i] : typename C::fp_type{});
}

// the static decoders in a kernel: value i from the view's payload words and
// exponent slots
template<class C>
__global__ void
impl_kernel(C view, std::size_t n, typename C::fp_type* out)
{
  const std::size_t i = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) {
    out[i] = C::decompress_gpu_element_impl(view.get_const_compressed_data(), view.get_const_exponent_data(), i);
  }
}

// the reference header's free kernels, launched as the reference's GPU test
// launches them (test/test_frsz_gpu.cu:325-341, 369-370) and with other
// grids (grid-stride loops, fewer threads than block_size), into buffers of
// other contents
template<class C>
void
check_free_kernels(C dev_view, const typename C::fp_type* d_in, typename C::fp_type* d_out, std::uint8_t* d_stream,
                   std::size_t n, const std::vector<std::uint8_t>& stream, const std::vector<typename C::fp_type>& out,
                   const char* tname)
{
  using T = typename C::fp_type;
  constexpr int B = C::max_exp_block_size;
  constexpr int bits = C::bits_per_value;
  const std::size_t size = stream.size();
  std::vector<std::uint8_t> dstream(size);
  std::vector<T> dout(n);
  constexpr int comp_threads = 4 * B;
  constexpr int blocks_per_tb = 512 / B > 1 ? 512 / B : 1;
  constexpr int decomp_threads = blocks_per_tb * B;
  struct grid
  {
    unsigned g, t;
  };
  const grid cgrids[] = { { static_cast<unsigned>((n + comp_threads - 1) / comp_threads), comp_threads },
                          { 3, comp_threads },
                          { 1, comp_threads / 2 > 0 ? comp_threads / 2 : 1 } };
  for (const grid& g : cgrids) {
    HIP_OK(hipMemset(d_stream, 0xFF, size));
    frsz::compress_gpu<C, comp_threads><<<dim3(g.g), dim3(g.t)>>>(d_in, dev_view);
    HIP_OK(hipDeviceSynchronize());
    HIP_OK(hipMemcpy(dstream.data(), d_stream, size, hipMemcpyDeviceToHost));
    CHECK(dstream == stream, "%s b%d B%d: compress_gpu<C,%d><<<%u,%u>>> stream differs", tname, bits, B, comp_threads,
          g.g, g.t);
  }
  const grid dgrids[] = { { static_cast<unsigned>((n + decomp_threads - 1) / decomp_threads),
                            static_cast<unsigned>(decomp_threads) },
                          { 2, 96 } };
  for (const grid& g : dgrids) {
    HIP_OK(hipMemset(d_out, 0x7F, n * sizeof(T)));
    frsz::decompress_gpu<C><<<dim3(g.g), dim3(g.t)>>>(d_out, dev_view);
    HIP_OK(hipDeviceSynchronize());
    HIP_OK(hipMemcpy(dout.data(), d_out, n * sizeof(T), hipMemcpyDeviceToHost));
    CHECK(std::memcmp(dout.data(), out.data(), n * sizeof(T)) == 0, "%s b%d B%d: decompress_gpu<<<%u,%u>>> differs",
          tname, bits, B, g.g, g.t);
  }
  // general_compress_gpu_function: one workgroup of B threads (or B rounded
  // up to a whole wave) per exponent block
  if constexpr (B <= 1024) {
    const unsigned nb = static_cast<unsigned>((n + B - 1) / B);
    for (unsigned t : { static_cast<unsigned>(B), static_cast<unsigned>((B + 63) / 64 * 64) }) {
      HIP_OK(hipMemset(d_stream, 0x5A, size));
      general_kernel<C><<<dim3(nb), dim3(t)>>>(dev_view, d_in, n);
      HIP_OK(hipDeviceSynchronize());
      HIP_OK(hipMemcpy(dstream.data(), d_stream, size, hipMemcpyDeviceToHost));
      CHECK(dstream == stream, "%s b%d B%d: general_compress_gpu_function (%u threads) stream differs", tname, bits,
            B, t);
    }
  }
  // decompress_gpu_element_impl in a kernel
  HIP_OK(hipMemcpy(d_stream, stream.data(), size, hipMemcpyHostToDevice));
  HIP_OK(hipMemset(d_out, 0x7F, n * sizeof(T)));
  impl_kernel<C><<<dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256)>>>(dev_view, n, d_out);
  HIP_OK(hipDeviceSynchronize());
  HIP_OK(hipMemcpy(dout.data(), d_out, n * sizeof(T), hipMemcpyDeviceToHost));
  CHECK(std::memcmp(dout.data(), out.data(), n * sizeof(T)) == 0, "%s b%d B%d: decompress_gpu_element_impl differs",
        tname, bits, B);
  // compress_gpu_function in a user kernel, workgroups of 4B threads
  for (unsigned g : { static_cast<unsigned>((n + comp_threads - 1) / comp_threads), 2u }) {
    HIP_OK(hipMemset(d_stream, 0xC3, size));
    function_kernel<C, comp_threads><<<dim3(g), dim3(comp_threads)>>>(dev_view, d_in, n);
    HIP_OK(hipDeviceSynchronize());
    HIP_OK(hipMemcpy(dstream.data(), d_stream, size, hipMemcpyDeviceToHost));
    CHECK(dstream == stream, "%s b%d B%d: compress_gpu_function (grid %u) stream differs", tname, bits, B, g);
  }
}

template<int bits, int B, class T>
void
run_case(const std::vector<T>& in, const std::string& outdir)
{
  using C = frsz::frsz2_compressor<bits, B, T>;
  const std::size_t n = in.size();
  const std::size_t size = C::compute_compressed_memory_size_byte(n);
  const char* tname = sizeof(T) == 4 ? "float" : "double";

  // host path
  std::vector<std::uint8_t> stream(size, 0xFF);
  C host_view(stream.data(), n);
  CHECK(host_view.compress_cpu_impl(in.data()) == 0, "compress_cpu_impl");
  std::vector<T> out(n, T(-1));
  CHECK(host_view.decompress_cpu_impl(out.data()) == 0, "decompress_cpu_impl");
  CHECK(std::memcmp(out.data(), in.data(), n * sizeof(T)) == 0, "%s b%d B%d: round trip not lossless", tname, bits,
        B);
  for (std::size_t i : { std::size_t{ 0 }, n / 3, n - 1 }) {
    const T e = host_view.decompress_cpu_element(i);
    CHECK(std::memcmp(&e, &out[i], sizeof(T)) == 0, "decompress_cpu_element(%zu)", i);
  }
  // per-block methods, every block into a buffer of other contents
  {
    const std::size_t nb = (n + B - 1) / B;
    std::vector<std::uint8_t> s2(size, 0xA5);
    C v2(s2.data(), n);
    for (std::size_t k = 0; k < nb; ++k) {
      v2.compress_cpu_block(k, in.data() + k * B, std::min<std::size_t>(B, n - k * B));
    }
    CHECK(s2 == stream, "%s b%d B%d: compress_cpu_block stream differs", tname, bits, B);
    std::vector<T> o2(n, T(-1));
    for (std::size_t k = 0; k < nb; ++k) {
      v2.decompress_cpu_block(k, o2.data() + k * B);
    }
    CHECK(std::memcmp(o2.data(), out.data(), n * sizeof(T)) == 0, "%s b%d B%d: decompress_cpu_block differs", tname,
          bits, B);
  }
  const std::string path = outdir + "/" + tname + "_b" + std::to_string(bits) + "_B" + std::to_string(B) + ".bin";
  if (FILE* f = std::fopen(path.c_str(), "wb")) {
    std::fwrite(stream.data(), 1, size, f);
    std::fclose(f);
  }

  // device path: same bytes, same values
  T* d_in = nullptr;
  T* d_out = nullptr;
  std::uint8_t* d_stream = nullptr;
  HIP_OK(hipMalloc(&d_in, n * sizeof(T)));
  HIP_OK(hipMalloc(&d_out, n * sizeof(T)));
  HIP_OK(hipMalloc(&d_stream, size));
  HIP_OK(hipMemcpy(d_in, in.data(), n * sizeof(T), hipMemcpyHostToDevice));
  C dev_view(d_stream, n);
  dev_view.compress_gpu(d_in);
  std::vector<std::uint8_t> dstream(size);
  HIP_OK(hipMemcpy(dstream.data(), d_stream, size, hipMemcpyDeviceToHost));
  CHECK(dstream == stream, "%s b%d B%d: device stream differs from host-API stream", tname, bits, B);
  dev_view.decompress_gpu(d_out);
  std::vector<T> dout(n);
  HIP_OK(hipMemcpy(dout.data(), d_out, n * sizeof(T), hipMemcpyDeviceToHost));
  CHECK(std::memcmp(dout.data(), out.data(), n * sizeof(T)) == 0, "decompress_gpu differs");
  // in-kernel accessor
  HIP_OK(hipMemset(d_out, 0, n * sizeof(T)));
  element_kernel<C><<<dim3((n + 255) / 256), dim3(256)>>>(dev_view, n, d_out);
  HIP_OK(hipDeviceSynchronize());
  HIP_OK(hipMemcpy(dout.data(), d_out, n * sizeof(T), hipMemcpyDeviceToHost));
  CHECK(std::memcmp(dout.data(), out.data(), n * sizeof(T)) == 0, "decompress_gpu_element differs");
  check_free_kernels<C>(dev_view, d_in, d_out, d_stream, n, stream, out, tname);
  // in-kernel per-block compress and decode
  {
    const std::size_t nb = (n + B - 1) / B;
    HIP_OK(hipMemset(d_stream, 0x3C, size));
    block_compress_kernel<C><<<dim3((nb + 63) / 64), dim3(64)>>>(dev_view, d_in, n);
    HIP_OK(hipDeviceSynchronize());
    HIP_OK(hipMemcpy(dstream.data(), d_stream, size, hipMemcpyDeviceToHost));
    CHECK(dstream == stream, "%s b%d B%d: compress_gpu_block stream differs", tname, bits, B);
    HIP_OK(hipMemset(d_out, 0, n * sizeof(T)));
    block_decode_kernel<C><<<dim3((nb + 63) / 64), dim3(64)>>>(dev_view, d_out, n);
    HIP_OK(hipDeviceSynchronize());
    HIP_OK(hipMemcpy(dout.data(), d_out, n * sizeof(T), hipMemcpyDeviceToHost));
    CHECK(std::memcmp(dout.data(), out.data(), n * sizeof(T)) == 0, "decompress_gpu_block differs");
    // two runs of blocks through the device-memory run methods
    const std::size_t half = nb / 2;
    HIP_OK(hipMemset(d_stream, 0x77, size));
    dev_view.compress_gpu_blocks(half, d_in + half * B, n - half * B);
    dev_view.compress_gpu_blocks(0, d_in, half * B);
    HIP_OK(hipMemcpy(dstream.data(), d_stream, size, hipMemcpyDeviceToHost));
    CHECK(dstream == stream, "%s b%d B%d: compress_gpu_blocks stream differs", tname, bits, B);
    HIP_OK(hipMemset(d_out, 0, n * sizeof(T)));
    dev_view.decompress_gpu_blocks(0, half * B, d_out);
    dev_view.decompress_gpu_blocks(half, n - half * B, d_out + half * B);
    HIP_OK(hipMemcpy(dout.data(), d_out, n * sizeof(T), hipMemcpyDeviceToHost));
    CHECK(std::memcmp(dout.data(), out.data(), n * sizeof(T)) == 0, "decompress_gpu_blocks differs");
  }
  HIP_OK(hipFree(d_in));
  HIP_OK(hipFree(d_out));
  HIP_OK(hipFree(d_stream));
}

// beyond the reference's list: bits = W with B = 1 (lossless,
// test/test_frsz_gpu.cu:414-441), a block larger than the workgroup, and
// non-power-of-two blocks, on the free kernels against the host-API stream
template<int bits, int B, class T>
void
free_kernels_only(const std::vector<T>& in)
{
  using C = frsz::frsz2_compressor<bits, B, T>;
  const std::size_t n = in.size();
  const std::size_t size = C::compute_compressed_memory_size_byte(n);
  const char* tname = sizeof(T) == 4 ? "float" : "double";
  std::vector<std::uint8_t> stream(size, 0);
  C host_view(stream.data(), n);
  host_view.compress_cpu_impl(in.data());
  std::vector<T> out(n);
  host_view.decompress_cpu_impl(out.data());
  T* d_in = nullptr;
  T* d_out = nullptr;
  std::uint8_t* d_stream = nullptr;
  HIP_OK(hipMalloc(&d_in, n * sizeof(T)));
  HIP_OK(hipMalloc(&d_out, n * sizeof(T)));
  HIP_OK(hipMalloc(&d_stream, size));
  HIP_OK(hipMemcpy(d_in, in.data(), n * sizeof(T), hipMemcpyHostToDevice));
  C dev_view(d_stream, n);
  if constexpr (4 * B <= 1024) {
    check_free_kernels<C>(dev_view, d_in, d_out, d_stream, n, stream, out, tname);
  } else {
    // a block of more values than the workgroup has threads
    std::vector<std::uint8_t> dstream(size);
    HIP_OK(hipMemset(d_stream, 0xFF, size));
    frsz::compress_gpu<C, 256><<<dim3(5), dim3(256)>>>(d_in, dev_view);
    HIP_OK(hipDeviceSynchronize());
    HIP_OK(hipMemcpy(dstream.data(), d_stream, size, hipMemcpyDeviceToHost));
    CHECK(dstream == stream, "%s b%d B%d: compress_gpu<C,256> stream differs", tname, bits, B);
    std::vector<T> dout(n);
    frsz::decompress_gpu<C><<<dim3(7), dim3(256)>>>(d_out, dev_view);
    HIP_OK(hipDeviceSynchronize());
    HIP_OK(hipMemcpy(dout.data(), d_out, n * sizeof(T), hipMemcpyDeviceToHost));
    CHECK(std::memcmp(dout.data(), out.data(), n * sizeof(T)) == 0, "%s b%d B%d: decompress_gpu differs", tname,
          bits, B);
  }
  if (bits == static_cast<int>(sizeof(T) * 8) && B == 1) {
    CHECK(std::memcmp(out.data(), in.data(), n * sizeof(T)) == 0, "%s: bits = W, B = 1 not lossless", tname);
  }
  HIP_OK(hipFree(d_in));
  HIP_OK(hipFree(d_out));
  HIP_OK(hipFree(d_stream));
}

void
check_wide_cases()
{
  // random finite values of both signs over a wide exponent range, a few
  // zeros and subnormals
  std::vector<float> f(100003);
  std::vector<double> d(50021);
  std::uint64_t st = 0x9E3779B97F4A7C15ull;
  auto next = [&st]() {
    st ^= st << 13;
    st ^= st >> 7;
    st ^= st << 17;
    return st;
  };
  for (auto& v : f) {
    const std::uint32_t r = static_cast<std::uint32_t>(next());
    std::uint32_t u = (r & 0x807fffffu) | ((60u + (r >> 8) % 136u) << 23);
    if (r % 97 == 0) {
      u &= 0x807fffffu; // subnormal or zero
    }
    std::memcpy(&v, &u, 4);
  }
  for (auto& v : d) {
    const std::uint64_t r = next();
    std::uint64_t u = (r & 0x800fffffffffffffull) | ((900ull + (r >> 20) % 250ull) << 52);
    if (r % 89 == 0) {
      u &= 0x800fffffffffffffull;
    }
    std::memcpy(&v, &u, 8);
  }
  free_kernels_only<32, 1, float>(f);
  free_kernels_only<64, 1, double>(d);
  free_kernels_only<12, 3, float>(f);
  free_kernels_only<21, 7, float>(f);
  free_kernels_only<40, 100, double>(d);
  free_kernels_only<16, 256, float>(f);
  free_kernels_only<24, 1024, float>(f);
  free_kernels_only<53, 2048, double>(d);
  // power-of-two blocks the kernel templates run on the library's tiles
  // (the first wave of each workgroup, grid-stride over many tiles)
  free_kernels_only<16, 32, float>(f);
  free_kernels_only<12, 32, float>(f);
  free_kernels_only<8, 16, float>(f);
  free_kernels_only<9, 2, float>(f);
  free_kernels_only<21, 64, float>(f);
  free_kernels_only<32, 32, double>(d);
  free_kernels_only<12, 4, double>(d);
  free_kernels_only<40, 128, double>(d);
}

void
check_dispatch()
{
  std::vector<float> x(100, 1.5f);
  using C = frsz::frsz2_compressor<16, 32, float>;
  std::vector<std::uint8_t> s(C::compute_compressed_memory_size_byte(x.size()));
  using bits = frsz::int_list_t<8, 16, 32>;
  using exps = frsz::int_list_t<2, 4, 8, 16, 32, 64, 128, 256, 512, 1024>;
  frsz::dispatch_frsz2_compression<float>(bits{}, 16, exps{}, 32, x.data(), x.size(), s.data());
  std::vector<float> y(x.size());
  frsz::dispatch_frsz2_decompression<float>(bits{}, 16, exps{}, 32, y.data(), y.size(), s.data());
  CHECK(y == x, "dispatch round trip");
  std::string msg;
  try {
    frsz::dispatch_frsz2_compression<float>(bits{}, 12, exps{}, 32, x.data(), x.size(), s.data());
  } catch (std::runtime_error const& e) {
    msg = e.what();
  }
  CHECK(msg == "Unsupported number of bits: 12", "bits message: %s", msg.c_str());
  msg.clear();
  try {
    frsz::dispatch_frsz2_compression<float>(bits{}, 16, exps{}, 3, x.data(), x.size(), s.data());
  } catch (std::runtime_error const& e) {
    msg = e.what();
  }
  CHECK(msg == "Unsupported exponent block size: 3", "block message: %s", msg.c_str());
}

} // namespace

int
main(int argc, char** argv)
{
  const std::string outdir = argc > 1 ? argv[1] : ".";
  const double pat[9] = { 1., 2., 3., 4., 0.25, -0.25, -0.125, 1 / 32., 0.125 };
  std::vector<double> vd(2049);
  for (std::size_t i = 0; i < vd.size(); ++i) {
    vd[i] = pat[i % 9];
  }
  std::vector<float> vf(2049 + 111);
  for (std::size_t i = 0; i < vf.size(); ++i) {
    vf[i] = static_cast<float>(pat[i % 9]);
  }
  run_case<32, 32, double>(vd, outdir);
  run_case<16, 32, double>(vd, outdir);
  run_case<16, 8, double>(vd, outdir);
  run_case<15, 8, double>(vd, outdir);
  run_case<9, 8, double>(vd, outdir);
  run_case<9, 4, double>(vd, outdir);
  run_case<9, 5, double>(vd, outdir);
  run_case<16, 8, float>(vf, outdir);
  run_case<15, 8, float>(vf, outdir);
  run_case<9, 4, float>(vf, outdir);
  run_case<9, 5, float>(vf, outdir);
  check_wide_cases();
  check_dispatch();
  std::printf("%s (%d failures)\n", g_fail ? "FAILED" : "OK", g_fail);
  return g_fail ? 1 : 0;
}
