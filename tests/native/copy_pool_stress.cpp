// Stress test of the host copy pool (frsz_amd/csrc/copy_pool.hpp), built by
// tests/test_copy_pool.py with -fsanitize=thread: several pools used from
// several threads at once, copies of every size class, pools created and
// destroyed while others run.  Exits non-zero on a wrong byte.
#include "copy_pool.hpp"

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <thread>
#include <vector>

namespace {

int
run_pool(int seed, int copies)
{
  std::mt19937_64 rng(static_cast<std::uint64_t>(seed));
  frsz2::copy_pool pool(3 + seed % 5);
  const std::size_t max_bytes = (std::size_t{ 6 } << 20) + 12345;
  std::vector<unsigned char> src(max_bytes), dst(max_bytes);
  for (std::size_t i = 0; i < max_bytes; ++i) {
    src[i] = static_cast<unsigned char>(rng());
  }
  for (int k = 0; k < copies; ++k) {
    const std::size_t sizes[] = { 0, 1, 4095, frsz2::PARALLEL_COPY_MIN - 1, frsz2::PARALLEL_COPY_MIN,
                                  frsz2::PARALLEL_COPY_MIN + 4097, max_bytes };
    const std::size_t n = sizes[rng() % 7];
    const std::size_t off = rng() % 64;
    const std::size_t len = n > off ? n - off : 0;
    std::memset(dst.data(), 0xEE, max_bytes);
    pool.copy(dst.data() + off, src.data() + off, len);
    if (std::memcmp(dst.data() + off, src.data() + off, len) != 0) {
      std::printf("FAIL seed %d copy %d bytes %zu\n", seed, k, len);
      return 1;
    }
    if (off > 0 && dst[off - 1] != 0xEE) {
      std::printf("FAIL seed %d copy %d wrote before the destination\n", seed, k);
      return 1;
    }
    if (off + len < max_bytes && dst[off + len] != 0xEE) {
      std::printf("FAIL seed %d copy %d wrote past the destination\n", seed, k);
      return 1;
    }
  }
  return 0;
}

} // namespace

int
main()
{
  std::vector<int> rc(6, 0);
  std::vector<std::thread> ts;
  for (int i = 0; i < 6; ++i) {
    ts.emplace_back([i, &rc] { rc[i] = run_pool(i, 12); });
  }
  for (auto& t : ts) {
    t.join();
  }
  for (int r : rc) {
    if (r) {
      return 1;
    }
  }
  std::printf("ok\n");
  return 0;
}
