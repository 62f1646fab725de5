// fastdiv.h -- division by a run-time invariant divisor as a multiply-high, an
// add and a shift, for the index arithmetic of the kernels (a hardware-free
// 32-bit division costs ~18 VALU instructions per lane on CDNA).
//
// Method (Granlund & Montgomery, "Division by invariant integers using
// multiplication", PLDI 1994, round-up variant): for 1 <= d <= 2^31 take
// l = ceil(log2 d) and the 33-bit multiplier 2^32 + m with
// m = floor(2^32 (2^l - d) / d) + 1; then for every 0 <= n < 2^31
//   n / d = (mulhi(m, n) + n) >> l,
// where mulhi(m, n) + n < 2n stays below 2^32.  Plain C++ (constexpr, hence
// callable from device code), so the host test compiles this same header.
#pragma once

namespace qamd {

struct FastDiv {
  unsigned d;  // divisor
  unsigned m;  // low 32 bits of the multiplier
  unsigned l;  // shift
};

// 1 <= d <= 2^31
constexpr FastDiv make_fastdiv(unsigned d)
{
  unsigned l = 0;
  while ((1ull << l) < d) l++;
  const unsigned m = (unsigned)((1ull << 32) * ((1ull << l) - d) / d + 1);
  return FastDiv{d, m, l};
}

// n / f.d for 0 <= n < 2^31
constexpr unsigned fdiv(unsigned n, FastDiv f)
{
  return ((unsigned)(((unsigned long long)f.m * n) >> 32) + n) >> f.l;
}

// n % f.d for 0 <= n < 2^31
constexpr unsigned frem(unsigned n, FastDiv f) { return n - fdiv(n, f) * f.d; }

}  // namespace qamd
