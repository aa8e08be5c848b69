// HBM streaming ceilings of the device: what plain coalesced 16-byte loads
// and stores reach for the read:write mixes the codec's kernels have.
//
//   read-only (1:0), write-only (0:1), copy (1:1),
//   2:1 (compress of float64 at 32 bits reads 8 bytes per 4 written),
//   1:2 (the matching decompress),
//   1:4 (decompress of float32 at 8 bits: 1 byte of code per 4 written),
//   4:1 (the matching compress).
//
// Every workgroup streams one tile: it reads R units and writes W units of
// blockDim.x * UNR 16-byte vectors, each lane loading all of its vectors
// before it stores any.  The tiles cover the buffers exactly (the sizes are
// multiples of the largest tile), so no lane is out of bounds.  Read-only
// tiles store only when a lane's XOR of its vectors equals a key the fill
// pattern never produces, into a sink of blockDim.x vectors.  Variants:
// plain or non-temporal loads and stores (NT), and workgroup order or one
// contiguous range of tiles per XCD (workgroup b runs on XCD b mod 8), the
// two choices the codec's tile kernels make.
//
// Vectors per lane per unit (UNR): 1, 2 or 4.
//
// Usage: stream_ceiling [--mib N]   (N MiB per side, default 1024)
// Prints one JSON line per (pattern, workgroup size, UNR, NT, order) with
// the best of three rounds of ten launches, event-timed, and exits non-zero
// if a copy's or a mix's bytes are wrong.
//
// Random-access ceiling: stream_ceiling --random M [--values N]
// The memory pattern of a random-access decode of M uniform indices into a
// float32 stream of N values (default 100000037) at bits 16, B 32, without
// the decode: per index an 8-byte index load (consecutive), a 2-byte load
// of a random code in the 2 N-byte payload and of its block's 2-byte
// exponent in the N/16-byte exponent array, and a 4-byte store of their
// XOR (consecutive).  Variants: both loads, code load only, exponent load
// only; plain or non-temporal code loads.  One JSON line per variant (best
// of three rounds of ten launches); the stored words are checked against
// the host's.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#define HIP_CHECK(expr)                                                                                     \
  do {                                                                                                      \
    hipError_t e_ = (expr);                                                                                 \
    if (e_ != hipSuccess) {                                                                                 \
      std::fprintf(stderr, "%s failed: %s\n", #expr, hipGetErrorString(e_));                               \
      std::exit(2);                                                                                         \
    }                                                                                                       \
  } while (0)

namespace {

__device__ __forceinline__ uint4
xor4(uint4 a, uint4 b)
{
  return make_uint4(a.x ^ b.x, a.y ^ b.y, a.z ^ b.z, a.w ^ b.w);
}

__device__ __forceinline__ uint4
pattern(std::uint64_t i)
{
  const auto lo = static_cast<std::uint32_t>(i);
  return make_uint4(lo * 0x9E3779B1u, lo, ~lo, lo ^ 0x55AA55AAu);
}

__global__ void
k_fill(uint4* p, std::uint64_t count)
{
  for (std::uint64_t i = blockIdx.x * std::uint64_t(blockDim.x) + threadIdx.x; i < count;
       i += std::uint64_t(gridDim.x) * blockDim.x)
    p[i] = pattern(i);
}

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

template<bool NT>
__device__ __forceinline__ uint4
load(const uint4* p)
{
  if constexpr (NT) {
    const u32x4 v = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(p));
    return make_uint4(v.x, v.y, v.z, v.w);
  } else {
    return *p;
  }
}

template<bool NT>
__device__ __forceinline__ void
store(uint4 v, uint4* p)
{
  if constexpr (NT) {
    const u32x4 w = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(w, reinterpret_cast<u32x4*>(p));
  } else {
    *p = v;
  }
}

// tile of workgroup b: with xcd order, the tiles are split into 8 contiguous
// ranges (the first ntiles % 8 one longer) and XCD b % 8 takes range b % 8
__device__ __forceinline__ std::uint32_t
tile_of(std::uint32_t b, std::uint32_t ntiles, bool xcd)
{
  if (!xcd)
    return b;
  const std::uint32_t x = b & 7u, local = b >> 3, q = ntiles >> 3, r = ntiles & 7u;
  return x < r ? x * (q + 1) + local : r * (q + 1) + (x - r) * q + local;
}

// R units read and W units written per tile; a unit is blockDim.x * UNR vectors
template<int R, int W, int UNR, bool NT>
__global__ void
k_stream(const uint4* __restrict__ src, uint4* __restrict__ dst, uint4* __restrict__ sink, uint4 key, int xcd)
{
  const std::uint64_t unit = std::uint64_t(blockDim.x) * UNR;
  const std::uint64_t t = tile_of(blockIdx.x, gridDim.x, xcd != 0);
  const unsigned lane = threadIdx.x;
  uint4 v[(R > 0 ? R : 1) * UNR];
  if constexpr (R > 0) {
    const uint4* s = src + t * R * unit;
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int u = 0; u < UNR; ++u)
        v[r * UNR + u] = load<NT>(s + r * unit + u * blockDim.x + lane);
  } else {
#pragma unroll
    for (int u = 0; u < UNR; ++u)
      v[u] = make_uint4(static_cast<unsigned>(t), lane, u, 0x5A5A5A5Au);
  }
  if constexpr (W == 0) {
    uint4 acc = v[0];
#pragma unroll
    for (int k = 1; k < R * UNR; ++k)
      acc = xor4(acc, v[k]);
    if (acc.x == key.x && acc.y == key.y && acc.z == key.z && acc.w == key.w)
      sink[lane] = acc;
  } else {
    uint4* d = dst + t * W * unit;
#pragma unroll
    for (int w = 0; w < W; ++w)
#pragma unroll
      for (int u = 0; u < UNR; ++u) {
        uint4 o = v[u];
#pragma unroll
        for (int r = 1; r < R; ++r)
          o = xor4(o, v[r * UNR + u]);
        if (w > 0)
          o = xor4(o, make_uint4(w, 2u * w, 3u * w, 4u * w));
        store<NT>(o, d + w * unit + u * blockDim.x + lane);
      }
  }
}

struct result
{
  double us;
  double gbs;
};

// ---- random-access ceiling

// index i of the random pattern: a 64-bit mix of i, reduced to [0, n)
__host__ __device__ inline std::uint64_t
random_index(std::uint64_t i, std::uint64_t n)
{
  std::uint64_t h = (i + 1) * 0x9E3779B97F4A7C15ull;
  h ^= h >> 31;
  h *= 0xBF58476D1CE4E5B9ull;
  h ^= h >> 29;
  return h % n;
}

__global__ void
k_fill_indices(std::uint64_t* idx, std::uint64_t m, std::uint64_t n)
{
  for (std::uint64_t i = blockIdx.x * std::uint64_t(blockDim.x) + threadIdx.x; i < m;
       i += std::uint64_t(gridDim.x) * blockDim.x)
    idx[i] = random_index(i, n);
}

// 16-bit word k of the filled payload / exponent arrays (k_fill's pattern)
__host__ __device__ inline std::uint16_t
pattern_half(std::uint64_t k)
{
  const auto lo = static_cast<std::uint32_t>(k / 8);
  const std::uint32_t w[4] = { lo * 0x9E3779B1u, lo, ~lo, lo ^ 0x55AA55AAu };
  const std::uint32_t d = w[(k / 2) % 4];
  return static_cast<std::uint16_t>(k % 2 ? d >> 16 : d);
}

// CODE / EXP: which of the two random loads run; NTC: non-temporal code loads
template<bool CODE, bool EXP, bool NTC>
__global__ __launch_bounds__(256) void
k_random(const std::uint64_t* __restrict__ idx, std::uint64_t m, const std::uint16_t* __restrict__ code,
         const std::uint16_t* __restrict__ exps, std::uint32_t* __restrict__ out)
{
  const std::uint64_t i = blockIdx.x * std::uint64_t(256) + threadIdx.x;
  if (i >= m)
    return;
  const std::uint64_t j = idx[i];
  std::uint32_t v = 0;
  if constexpr (CODE)
    v = NTC ? __builtin_nontemporal_load(code + j) : code[j];
  if constexpr (EXP)
    v ^= static_cast<std::uint32_t>(exps[j >> 5]) << 16;
  out[i] = v;
}

template<bool CODE, bool EXP, bool NTC>
int
run_random(const char* name, const std::uint64_t* idx, std::uint64_t m, std::uint64_t n, const std::uint16_t* code,
           const std::uint16_t* exps, std::uint32_t* out, const std::vector<std::uint64_t>& hidx, hipEvent_t e0,
           hipEvent_t e1)
{
  const unsigned grid = static_cast<unsigned>((m + 255) / 256);
  auto launch = [&] {
    hipLaunchKernelGGL((k_random<CODE, EXP, NTC>), dim3(grid), dim3(256), 0, nullptr, idx, m, code, exps, out);
  };
  for (int i = 0; i < 2; ++i)
    launch();
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipDeviceSynchronize());
  float best = 0.f;
  for (int round = 0; round < 3; ++round) {
    HIP_CHECK(hipEventRecord(e0, nullptr));
    for (int i = 0; i < 10; ++i)
      launch();
    HIP_CHECK(hipEventRecord(e1, nullptr));
    HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    if (round == 0 || ms < best)
      best = ms;
  }
  // check the first and last 4096 stored words
  int fails = 0;
  const std::uint64_t k = m < 4096 ? m : 4096;
  std::vector<std::uint32_t> got(k);
  for (int part = 0; part < 2; ++part) {
    const std::uint64_t first = part == 0 ? 0 : m - k;
    HIP_CHECK(hipMemcpy(got.data(), out + first, k * 4, hipMemcpyDeviceToHost));
    for (std::uint64_t q = 0; q < k; ++q) {
      const std::uint64_t j = hidx[part * k + q];
      std::uint32_t want = 0;
      if (CODE)
        want = pattern_half(j);
      if (EXP)
        want ^= static_cast<std::uint32_t>(pattern_half(j >> 5)) << 16;
      if (got[q] != want) {
        ++fails;
        break;
      }
    }
  }
  const double us = best / 10.0 * 1e3;
  std::printf("{\"pattern\": \"random\", \"loads\": \"%s\", \"nt_code\": %d, \"indices\": %llu, \"values\": %llu, "
              "\"us\": %.1f, \"gindex_s\": %.2f, \"ok\": %d}\n",
              name, int(NTC), static_cast<unsigned long long>(m), static_cast<unsigned long long>(n), us,
              double(m) / us * 1e-3, fails == 0);
  std::fflush(stdout);
  return fails;
}

int
random_ceiling(std::uint64_t m, std::uint64_t n)
{
  const std::uint64_t nexp = (n + 31) / 32;
  // payload: n codes; exponents: nexp words; both padded to whole vectors
  const std::uint64_t pvec = (2 * n + 15) / 16;
  const std::uint64_t evec = (2 * nexp + 15) / 16;
  uint4 *code = nullptr, *exps = nullptr;
  std::uint64_t* idx = nullptr;
  std::uint32_t* out = nullptr;
  HIP_CHECK(hipMalloc(&code, pvec * 16));
  HIP_CHECK(hipMalloc(&exps, evec * 16));
  HIP_CHECK(hipMalloc(&idx, m * 8));
  HIP_CHECK(hipMalloc(&out, m * 4));
  hipLaunchKernelGGL(k_fill, dim3(4096), dim3(256), 0, nullptr, code, pvec);
  hipLaunchKernelGGL(k_fill, dim3(4096), dim3(256), 0, nullptr, exps, evec);
  hipLaunchKernelGGL(k_fill_indices, dim3(4096), dim3(256), 0, nullptr, idx, m, n);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipDeviceSynchronize());
  const std::uint64_t k = m < 4096 ? m : 4096;
  std::vector<std::uint64_t> hidx(2 * k);
  for (std::uint64_t q = 0; q < k; ++q) {
    hidx[q] = random_index(q, n);
    hidx[k + q] = random_index(m - k + q, n);
  }
  hipEvent_t e0, e1;
  HIP_CHECK(hipEventCreate(&e0));
  HIP_CHECK(hipEventCreate(&e1));
  const auto* c16 = reinterpret_cast<const std::uint16_t*>(code);
  const auto* x16 = reinterpret_cast<const std::uint16_t*>(exps);
  int fails = 0;
  fails += run_random<true, true, false>("code+exponent", idx, m, n, c16, x16, out, hidx, e0, e1);
  fails += run_random<true, true, true>("code+exponent", idx, m, n, c16, x16, out, hidx, e0, e1);
  fails += run_random<true, false, false>("code", idx, m, n, c16, x16, out, hidx, e0, e1);
  fails += run_random<true, false, true>("code", idx, m, n, c16, x16, out, hidx, e0, e1);
  fails += run_random<false, true, false>("exponent", idx, m, n, c16, x16, out, hidx, e0, e1);
  HIP_CHECK(hipFree(code));
  HIP_CHECK(hipFree(exps));
  HIP_CHECK(hipFree(idx));
  HIP_CHECK(hipFree(out));
  if (fails)
    std::printf("FAIL %d random-access checks\n", fails);
  return fails ? 1 : 0;
}

template<int R, int W, int UNR, bool NT>
result
run(const uint4* src, uint4* dst, uint4* sink, std::uint64_t side_bytes, int wg, int xcd, hipEvent_t e0,
    hipEvent_t e1)
{
  const std::uint64_t unit_bytes = std::uint64_t(wg) * UNR * 16;
  const int m = R > W ? R : W;
  const std::uint64_t tiles = side_bytes / (unit_bytes * m);
  const uint4 key = make_uint4(0xDEADBEEFu, 0x01234567u, 0x89ABCDEFu, 0xFEDCBA98u);
  auto launch = [&] {
    hipLaunchKernelGGL((k_stream<R, W, UNR, NT>), dim3(static_cast<unsigned>(tiles)), dim3(wg), 0, nullptr, src,
                       dst, sink, key, xcd);
  };
  for (int i = 0; i < 2; ++i)
    launch();
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipDeviceSynchronize());
  float best = 0.f;
  for (int round = 0; round < 3; ++round) {
    HIP_CHECK(hipEventRecord(e0, nullptr));
    for (int i = 0; i < 10; ++i)
      launch();
    HIP_CHECK(hipEventRecord(e1, nullptr));
    HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    if (round == 0 || ms < best)
      best = ms;
  }
  const double us = best / 10.0 * 1e3;
  const double bytes = double(tiles) * unit_bytes * (R + W);
  return {us, bytes / (us * 1e-6) / 1e9};
}

// host copy of the vectors [first, first + count) of the fill pattern
std::vector<uint4>
host_pattern(std::uint64_t first, std::uint64_t count)
{
  std::vector<uint4> h(count);
  for (std::uint64_t k = 0; k < count; ++k) {
    const auto lo = static_cast<std::uint32_t>(first + k);
    h[k] = make_uint4(lo * 0x9E3779B1u, lo, ~lo, lo ^ 0x55AA55AAu);
  }
  return h;
}

bool
same(const uint4& a, const uint4& b)
{
  return a.x == b.x && a.y == b.y && a.z == b.z && a.w == b.w;
}

// checks the first `count` vectors of dst after a pattern (R, W) ran with
// workgroup size wg and UNR (the first tile holds them when count <= unit)
bool
check_tile(const uint4* dst, int R, int W, int wg, int unr)
{
  const std::uint64_t unit = std::uint64_t(wg) * unr;
  std::vector<uint4> got(unit * W);
  HIP_CHECK(hipMemcpy(got.data(), dst, got.size() * 16, hipMemcpyDeviceToHost));
  std::vector<std::vector<uint4>> s;
  for (int r = 0; r < (R > 0 ? R : 1); ++r)
    s.push_back(host_pattern(r * unit, unit));
  for (int w = 0; w < W; ++w)
    for (std::uint64_t k = 0; k < unit; ++k) {
      uint4 want;
      if (R == 0) {
        const unsigned lane = static_cast<unsigned>(k % wg), u = static_cast<unsigned>(k / wg);
        want = make_uint4(0u, lane, u, 0x5A5A5A5Au);
      } else {
        want = s[0][k];
        for (int r = 1; r < R; ++r)
          want = make_uint4(want.x ^ s[r][k].x, want.y ^ s[r][k].y, want.z ^ s[r][k].z, want.w ^ s[r][k].w);
      }
      const unsigned c = static_cast<unsigned>(w);
      want = make_uint4(want.x ^ c, want.y ^ (2u * c), want.z ^ (3u * c), want.w ^ (4u * c));
      if (!same(got[w * unit + k], want)) {
        std::printf("FAIL pattern %d:%d wg %d unr %d vector %llu\n", R, W, wg, unr,
                    static_cast<unsigned long long>(w * unit + k));
        return false;
      }
    }
  return true;
}

} // namespace

int
main(int argc, char** argv)
{
  std::uint64_t mib = 1024;
  std::uint64_t random_m = 0;
  std::uint64_t random_n = 100000037;
  for (int i = 1; i < argc; ++i) {
    if (!std::strcmp(argv[i], "--mib") && i + 1 < argc)
      mib = std::strtoull(argv[++i], nullptr, 10);
    else if (!std::strcmp(argv[i], "--random") && i + 1 < argc)
      random_m = std::strtoull(argv[++i], nullptr, 10);
    else if (!std::strcmp(argv[i], "--values") && i + 1 < argc)
      random_n = std::strtoull(argv[++i], nullptr, 10);
  }
  if (random_m > 0) {
    // one launch of one thread per index: at most 2^32 work-items
    if (random_m > (std::uint64_t{ 1 } << 31) || random_n < 32 || random_n > (std::uint64_t{ 1 } << 34)) {
      std::fprintf(stderr, "--random must be in 1..2^31 and --values in 32..2^34\n");
      return 2;
    }
    return random_ceiling(random_m, random_n);
  }
  // the largest tile is 4 units of 256 * 4 vectors (64 KiB): sizes in MiB divide evenly
  if (mib < 1 || mib > 16384) {
    std::fprintf(stderr, "--mib must be in 1..16384\n");
    return 2;
  }
  const std::uint64_t side = mib << 20;
  uint4 *src = nullptr, *dst = nullptr, *sink = nullptr;
  HIP_CHECK(hipMalloc(&src, side));
  HIP_CHECK(hipMalloc(&dst, side));
  HIP_CHECK(hipMalloc(&sink, 256 * 16));
  hipLaunchKernelGGL(k_fill, dim3(4096), dim3(256), 0, nullptr, src, side / 16);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipDeviceSynchronize());
  hipEvent_t e0, e1;
  HIP_CHECK(hipEventCreate(&e0));
  HIP_CHECK(hipEventCreate(&e1));
  int fails = 0;
  auto report = [&](const char* name, int R, int W, int wg, int unr, int nt, int xcd, result r) {
    std::printf("{\"pattern\": \"%s\", \"read\": %d, \"write\": %d, \"workgroup\": %d, \"vectors_per_lane\": %d, "
                "\"nt\": %d, \"xcd_order\": %d, \"side_mib\": %llu, \"us\": %.1f, \"gbs\": %.1f}\n",
                name, R, W, wg, unr, nt, xcd, static_cast<unsigned long long>(mib), r.us, r.gbs);
    std::fflush(stdout);
    if (W > 0 && !check_tile(dst, R, W, wg, unr))
      ++fails;
  };
#define CEIL_RUN(name, R, W, UNR, NT)                                                                      \
  report(name, R, W, wg, UNR, NT, xcd, run<R, W, UNR, NT>(src, dst, sink, side, wg, xcd, e0, e1))
  for (int wg : {256, 64})
    for (int xcd : {0, 1}) {
      CEIL_RUN("read-only", 1, 0, 4, false);
      CEIL_RUN("read-only", 1, 0, 2, false);
      CEIL_RUN("read-only", 1, 0, 1, false);
      CEIL_RUN("read-only", 1, 0, 4, true);
      CEIL_RUN("read-only", 1, 0, 2, true);
      CEIL_RUN("read-only", 1, 0, 1, true);
      CEIL_RUN("write-only", 0, 1, 4, false);
      CEIL_RUN("write-only", 0, 1, 2, false);
      CEIL_RUN("write-only", 0, 1, 1, false);
      CEIL_RUN("write-only", 0, 1, 4, true);
      CEIL_RUN("write-only", 0, 1, 2, true);
      CEIL_RUN("write-only", 0, 1, 1, true);
      CEIL_RUN("copy 1:1", 1, 1, 4, false);
      CEIL_RUN("copy 1:1", 1, 1, 2, false);
      CEIL_RUN("copy 1:1", 1, 1, 1, false);
      CEIL_RUN("copy 1:1", 1, 1, 4, true);
      CEIL_RUN("copy 1:1", 1, 1, 2, true);
      CEIL_RUN("copy 1:1", 1, 1, 1, true);
      CEIL_RUN("2:1", 2, 1, 4, false);
      CEIL_RUN("2:1", 2, 1, 2, false);
      CEIL_RUN("2:1", 2, 1, 1, false);
      CEIL_RUN("2:1", 2, 1, 4, true);
      CEIL_RUN("2:1", 2, 1, 2, true);
      CEIL_RUN("2:1", 2, 1, 1, true);
      CEIL_RUN("1:2", 1, 2, 4, false);
      CEIL_RUN("1:2", 1, 2, 2, false);
      CEIL_RUN("1:2", 1, 2, 1, false);
      CEIL_RUN("1:2", 1, 2, 4, true);
      CEIL_RUN("1:2", 1, 2, 2, true);
      CEIL_RUN("1:2", 1, 2, 1, true);
      CEIL_RUN("1:4", 1, 4, 4, false);
      CEIL_RUN("1:4", 1, 4, 2, false);
      CEIL_RUN("1:4", 1, 4, 1, false);
      CEIL_RUN("1:4", 1, 4, 4, true);
      CEIL_RUN("1:4", 1, 4, 2, true);
      CEIL_RUN("1:4", 1, 4, 1, true);
      CEIL_RUN("4:1", 4, 1, 4, false);
      CEIL_RUN("4:1", 4, 1, 2, false);
      CEIL_RUN("4:1", 4, 1, 1, false);
      CEIL_RUN("4:1", 4, 1, 4, true);
      CEIL_RUN("4:1", 4, 1, 2, true);
      CEIL_RUN("4:1", 4, 1, 1, true);
    }
#undef CEIL_RUN
  HIP_CHECK(hipFree(src));
  HIP_CHECK(hipFree(dst));
  HIP_CHECK(hipFree(sink));
  if (fails)
    std::printf("FAIL %d pattern checks\n", fails);
  return fails ? 1 : 0;
}
