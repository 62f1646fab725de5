// Checks the kernels' multiply-shift division (qphix-codegen_amd/csrc/fastdiv.h)
// against the built-in / and % on the host: every divisor 1..2048 with every
// numerator below 20000, the top 2000 numerators below 2^31 and random ones,
// and large divisors (lattice volumes, powers of two, 2^31 - 1, 2^31) with
// random numerators and the top of the range.  Exit status 0 when all agree.
#include "fastdiv.h"

#include <cstdio>
#include <random>

int main()
{
  using qamd::FastDiv;
  std::mt19937_64 rng(20261203);
  long bad = 0, checked = 0;
  auto check = [&](unsigned d, unsigned n) {
    const FastDiv f = qamd::make_fastdiv(d);
    checked++;
    if (qamd::fdiv(n, f) != n / d || qamd::frem(n, f) != n % d) {
      if (bad++ < 5) std::printf("mismatch: d=%u n=%u\n", d, n);
    }
  };
  const unsigned top = 0x7fffffffu;  // largest numerator the kernels pass
  for (unsigned d = 1; d <= 2048; d++) {
    for (unsigned n = 0; n < 20000; n++) check(d, n);
    for (unsigned n = top - 2000; n <= top - 1; n++) check(d, n);
    check(d, top);
    for (int k = 0; k < 500; k++) check(d, (unsigned)(rng() & top));
  }
  const unsigned big[] = {5308416u, 1048576u, 663552u, 65537u, 1000003u, 1u << 30, top, 1u << 31};
  for (unsigned d : big) {
    for (int k = 0; k < 200000; k++) check(d, (unsigned)(rng() & top));
    for (unsigned n = top - 20000; n <= top - 1; n++) check(d, n);
    check(d, top);
  }
  std::printf("fastdiv: %ld checked, %ld mismatches\n", checked, bad);
  return bad != 0;
}
