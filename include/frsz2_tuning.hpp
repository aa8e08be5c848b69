// Default kernel variants of the tiled kernels, measured on MI355X (1 GiB
// float64 field; 10^8-value float32 and 5*10^7-value float64 arrays of the
// edge-case generator), shared by the tuning table of the plain kernels, the
// grouped batch kernels and the drop-in header's kernel templates.
#pragma once

namespace frsz2 {

// encoded variant: values per thread, staged decompress (the chunked decode
// with LDS-staged stores; off: the strided decode), the non-temporal hint,
// linear (workgroup-order) tiles, 64-thread workgroups
constexpr int
venc(int vpt, int stage, int nt, int linear = 0, int small = 0)
{
  return vpt * 16 + small * 8 + linear * 4 + stage * 2 + nt;
}
constexpr int venc_vpt(int v) { return v / 16; }
constexpr int venc_small(int v) { return (v / 8) % 2; }
constexpr int venc_linear(int v) { return (v / 4) % 2; }
constexpr int venc_stage(int v) { return (v / 2) % 2; }
constexpr int venc_nt(int v) { return v % 2; }

// op 0 compress / 1 decompress; dtype 0 float32 / 1 float64; ui = index of
// U in {8, 16, 32, 64}; packed = bits < U
constexpr int
default_variant(int op, int dtype, int ui, int packed)
{
  if (op == 0) {
    if (dtype == 0) {
      if (packed) {
        return ui == 2 ? venc(16, 0, 1, 1) : venc(16, 0, 1, 1, 1);
      }
      return ui == 0 ? venc(8, 0, 1, 0, 1) : ui == 1 ? venc(8, 0, 1, 1, 1) : venc(4, 0, 1, 1, 1);
    }
    if (packed) {
      return ui == 2 ? venc(8, 0, 1, 1, 1) : venc(8, 0, 1, 1, 0);
    }
    return ui <= 2 ? venc(4, 0, 1, 0, 1) : venc(2, 0, 1, 1, 1);
  }
  if (dtype == 0) {
    if (packed) {
      return ui == 0 ? venc(16, 1, 1, 1) : ui == 1 ? venc(8, 0, 1, 1) : venc(16, 0, 1, 1, 1);
    }
    return ui == 2 ? venc(8, 0, 1, 1) : venc(16, 1, 1, 1, 1);
  }
  if (packed) {
    return ui <= 1 ? venc(8, 1, 1, 1) : ui == 2 ? venc(4, 0, 1, 1) : venc(2, 0, 1, 1, 1);
  }
  return ui == 2 ? venc(8, 1, 1, 1, 1) : venc(8, 1, 1, 1);
}

} // namespace frsz2
