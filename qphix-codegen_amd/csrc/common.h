// common.h -- device-side building blocks shared by every kernel of the
// library: storage traits, the native site-fastest field layout, complex
// arithmetic and checkerboard site decoding.
//
// Native layout (one checkerboard, vcb sites): a field with N reals per
// site is stored as vcb-long rows of W-element vectors,
//     element k of site i  at  ((k / W) * vcb + i) * W + (k % W),
// W = the widest of 16/8/4 bytes that divides the field's unit (a spinor,
// one link, one clover block), so one load instruction of a 64-lane
// wavefront reads 64 consecutive sites' chunk: 1 KiB at 16 B per lane.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>
#include <stdint.h>

#include "../../include/qphix_amd.h"
#include "fastdiv.h"

namespace qamd {

// operator families: plain Wilson, clover, twisted mass, twisted-mass clover
enum { OP_WILSON = 0, OP_CLOVER = 1, OP_TM = 2, OP_TMCLOVER = 3 };

// ---- storage traits ------------------------------------------------------
template <typename S> struct Traits;
template <> struct Traits<double> {
  using R = double;
  static constexpr int prec = QPHIX_AMD_F64;
};
template <> struct Traits<float> {
  using R = float;
  static constexpr int prec = QPHIX_AMD_F32;
};
template <> struct Traits<__half> {
  using R = float;
  static constexpr int prec = QPHIX_AMD_F16;
};
template <> struct Traits<__hip_bfloat16> {
  using R = float;
  static constexpr int prec = QPHIX_AMD_BF16;
};

__device__ __forceinline__ double to_real(double v) { return v; }
__device__ __forceinline__ float to_real(float v) { return v; }
__device__ __forceinline__ float to_real(__half v) { return __half2float(v); }
__device__ __forceinline__ float to_real(__hip_bfloat16 v) { return __bfloat162float(v); }

template <typename S> __device__ __forceinline__ S from_real(double v);
template <> __device__ __forceinline__ double from_real<double>(double v) { return v; }
template <> __device__ __forceinline__ float from_real<float>(double v) { return (float)v; }
template <> __device__ __forceinline__ __half from_real<__half>(double v) { return __float2half((float)v); }
template <> __device__ __forceinline__ __hip_bfloat16 from_real<__hip_bfloat16>(double v) { return __float2bfloat16((float)v); }

template <typename S> __device__ __forceinline__ S from_realf(float v);
template <> __device__ __forceinline__ double from_realf<double>(float v) { return v; }
template <> __device__ __forceinline__ float from_realf<float>(float v) { return v; }
// FP16 stores round the FP32 result to FP16: the value is made opaque first,
// so the compiler cannot fuse the last multiply-add with the conversion into
// one rounding straight to FP16 (v_fma_mixlo_f16).  The reference's
// low-precision kernels compute in FP32 and then convert
// (codegen/inst_sp_vec16.cc:129-135), and a fused form would round some
// elements differently, depending on which outputs the compiler fuses.
template <> __device__ __forceinline__ __half from_realf<__half>(float v)
{
  asm("" : "+v"(v));
  return __float2half(v);
}
template <> __device__ __forceinline__ __hip_bfloat16 from_realf<__hip_bfloat16>(float v) { return __float2bfloat16(v); }

template <typename S> __device__ __forceinline__ S to_storage(double v) { return from_real<S>(v); }
template <typename S> __device__ __forceinline__ S to_storage(float v) { return from_realf<S>(v); }

// Vector width (elements) for a unit of N reals of storage type S.
template <typename S> constexpr int pick_w(int N)
{
  return (N % (16 / (int)sizeof(S)) == 0) ? 16 / (int)sizeof(S)
         : (N % (8 / (int)sizeof(S)) == 0) ? 8 / (int)sizeof(S)
                                            : 4 / (int)sizeof(S);
}

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
// Unit widths of each field.  The two links of one direction mu (back and
// forward) form one unit, as do the two packed clover blocks of a site, so
// that FP16/BF16 12-real links (48 B per pair) and packed clover (144 B)
// load as 16-byte vectors.
template <typename S> constexpr int spinor_w() { return pick_w<S>(24); }
template <typename S, int RECON> constexpr int gauge_w() { return pick_w<S>(2 * RECON); }
template <typename S> constexpr int clover_w() { return pick_w<S>(72); }
template <typename S> constexpr int fclover_w() { return pick_w<S>(144); }

template <int BYTES> struct RawVec;
template <> struct RawVec<16> { using T = u32x4; };
template <> struct RawVec<8> { using T = u32x2; };
template <> struct RawVec<4> { using T = unsigned int; };

// Cache policy of a field access.  Fields read once per launch (links,
// clover blocks) and the output are streamed with non-temporal accesses so
// that L2 keeps the neighbour spinors, which every site reads 8 times.
#ifndef QAMD_NT_STREAM
#define QAMD_NT_STREAM 1
#endif

// Load reals [K0, K0+N) of site `site` of a field whose unit width is W.
// K0 and N are multiples of W.
template <typename S, int W, int K0, int N, typename R, bool NT = false>
__device__ __forceinline__ void load_reals(const S *__restrict__ base, int vcb,
                                           int site, R *__restrict__ out)
{
  static_assert(K0 % W == 0 && N % W == 0, "unit misaligned");
  using V = typename RawVec<W * sizeof(S)>::T;
  const V *vb = reinterpret_cast<const V *>(base);
#pragma unroll
  for (int c = 0; c < N / W; c++) {
    const V *a = vb + (unsigned)((K0 / W + c) * vcb + site);
    V v = (NT && QAMD_NT_STREAM) ? __builtin_nontemporal_load(a) : *a;
    S tmp[W];
    __builtin_memcpy(tmp, &v, sizeof(V));
#pragma unroll
    for (int e = 0; e < W; e++) out[c * W + e] = (R)to_real(tmp[e]);
  }
}

template <typename S, int W, int K0, int N, typename R, bool NT = false>
__device__ __forceinline__ void store_reals(S *__restrict__ base, int vcb,
                                            int site, const R *__restrict__ in)
{
  static_assert(K0 % W == 0 && N % W == 0, "unit misaligned");
  using V = typename RawVec<W * sizeof(S)>::T;
  V *vb = reinterpret_cast<V *>(base);
#pragma unroll
  for (int c = 0; c < N / W; c++) {
    S tmp[W];
#pragma unroll
    for (int e = 0; e < W; e++) tmp[e] = to_storage<S>(in[c * W + e]);
    V v;
    __builtin_memcpy(&v, tmp, sizeof(V));
    V *a = vb + (unsigned)((K0 / W + c) * vcb + site);
    if (NT && QAMD_NT_STREAM)
      __builtin_nontemporal_store(v, a);
    else
      *a = v;
  }
}

// ---- complex arithmetic ---------------------------------------------------
// A fused multiply-add in the arithmetic type itself (never a wider one).
__device__ __forceinline__ float fmx(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmx(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <typename R> struct cx {
  R re, im;
};
template <typename R> __device__ __forceinline__ cx<R> mk(R a, R b) { return cx<R>{a, b}; }
template <typename R> __device__ __forceinline__ cx<R> operator+(cx<R> a, cx<R> b) { return {a.re + b.re, a.im + b.im}; }
template <typename R> __device__ __forceinline__ cx<R> operator-(cx<R> a, cx<R> b) { return {a.re - b.re, a.im - b.im}; }
template <typename R> __device__ __forceinline__ cx<R> operator*(cx<R> a, cx<R> b)
{
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename R> __device__ __forceinline__ cx<R> operator*(R s, cx<R> a) { return {s * a.re, s * a.im}; }
template <typename R> __device__ __forceinline__ cx<R> conj(cx<R> a) { return {a.re, -a.im}; }
// i*a
template <typename R> __device__ __forceinline__ cx<R> timesI(cx<R> a) { return {-a.im, a.re}; }
// conj(a) * b
template <typename R> __device__ __forceinline__ cx<R> cmul_conj(cx<R> a, cx<R> b)
{
  return {a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re};
}
// acc + a*b and acc + conj(a)*b as four fused multiply-adds (sums of complex
// products otherwise compile to a multiply and an add per term)
template <typename R> __device__ __forceinline__ cx<R> cmac(cx<R> acc, cx<R> a, cx<R> b)
{
  R re = fma(a.re, b.re, acc.re);
  re = fma(-a.im, b.im, re);
  R im = fma(a.re, b.im, acc.im);
  im = fma(a.im, b.re, im);
  return {re, im};
}
template <typename R> __device__ __forceinline__ cx<R> cmac_conj(cx<R> acc, cx<R> a, cx<R> b)
{
  R re = fma(a.re, b.re, acc.re);
  re = fma(a.im, b.im, re);
  R im = fma(a.re, b.im, acc.im);
  im = fma(-a.im, b.re, im);
  return {re, im};
}
// row . vec of three complex entries: u0 h0 + u1 h1 + u2 h2
template <typename R>
__device__ __forceinline__ cx<R> cdot3(cx<R> u0, cx<R> u1, cx<R> u2, cx<R> h0, cx<R> h1, cx<R> h2)
{
  cx<R> r = {u0.re * h0.re, u0.re * h0.im};
  r.re = fma(-u0.im, h0.im, r.re);
  r.im = fma(u0.im, h0.re, r.im);
  return cmac(cmac(r, u1, h1), u2, h2);
}
// The same with the first product of each chain taken as fma(a, b, nz):
// with nz = -0.0 that is exactly the product (signed zeros included), and
// when nz is opaque to the compiler the product stays a fused multiply-add,
// which can read an FP16 operand without a separate conversion.
template <typename R>
__device__ __forceinline__ cx<R> cdot3(cx<R> u0, cx<R> u1, cx<R> u2, cx<R> h0, cx<R> h1, cx<R> h2, R nz)
{
  cx<R> r = {fma(u0.re, h0.re, nz), fma(u0.re, h0.im, nz)};
  r.re = fma(-u0.im, h0.im, r.re);
  r.im = fma(u0.im, h0.re, r.im);
  return cmac(cmac(r, u1, h1), u2, h2);
}
// conj(p q - r s) (the third row of an SU(3) matrix from its first two),
// with the reference's rounding sequence (codegen/dslash_common.cc:
// 701-718, Conj_CrossProd): the real part starts from p.re q.re, the
// imaginary part from r.re s.im, one product rounded and three fused
// multiply-adds each.
template <typename R> __device__ __forceinline__ cx<R> conj_diff(cx<R> p, cx<R> q, cx<R> r, cx<R> s, R nz)
{
  R re = fma(p.re, q.re, nz);
  re = fma(-p.im, q.im, re);
  re = fma(-r.re, s.re, re);
  re = fma(r.im, s.im, re);
  R im = fma(r.re, s.im, nz);
  im = fma(r.im, s.re, im);
  im = fma(-p.re, q.im, im);
  im = fma(-p.im, q.re, im);
  return {re, im};
}
template <typename R> __device__ __forceinline__ cx<R> conj_diff(cx<R> p, cx<R> q, cx<R> r, cx<R> s)
{
  R re = p.re * q.re;
  re = fma(-p.im, q.im, re);
  re = fma(-r.re, s.re, re);
  re = fma(r.im, s.im, re);
  R im = r.re * s.im;
  im = fma(r.im, s.re, im);
  im = fma(-p.re, q.im, im);
  im = fma(-p.im, q.re, im);
  return {re, im};
}
// a b with one rounded product and one fused multiply-add per part, the
// order the reference's complex multiplies use (codegen/dslash_common.cc:
// 1188-1232, 1491-1524): re = fma(-a.im, b.im, a.re b.re),
// im = fma(a.im, b.re, a.re b.im).
template <typename R> __device__ __forceinline__ cx<R> cmul_ord(cx<R> a, cx<R> b)
{
  R re = a.re * b.re;
  R im = a.re * b.im;
  return {fma(-a.im, b.im, re), fma(a.im, b.re, im)};
}

// ---- checkerboard geometry ------------------------------------------------
// cb index i = ((t*Lz + z)*Ly + y)*Lxh + xh,  x = 2 xh + ((cb+y+z+t) & 1).
struct Geom {
  int Lx, Ly, Lz, Lt, Lxh;
  int vcb;         // sites per checkerboard
  int sy, sz, st;  // cb-index strides of y, z, t
  FastDiv fxh, fy, fz;  // division by Lxh, Ly, Lz
};

struct SiteCoord {
  int x, y, z, t, xh;
};

// 32-bit decode (every supported lattice has vcb < 2^31 sites).
__device__ __forceinline__ SiteCoord decode(const Geom &g, int i, int cb)
{
  SiteCoord c;
  unsigned r = (unsigned)i;
  unsigned q = fdiv(r, g.fxh);
  c.xh = (int)(r - q * (unsigned)g.Lxh);
  r = q;
  q = fdiv(r, g.fy);
  c.y = (int)(r - q * (unsigned)g.Ly);
  r = q;
  q = fdiv(r, g.fz);
  c.z = (int)(r - q * (unsigned)g.Lz);
  c.t = (int)q;
  c.x = 2 * c.xh + ((cb + c.y + c.z + c.t) & 1);
  return c;
}

// Lexicographic full-lattice index of a cb site.
__device__ __forceinline__ long lex_index(const Geom &g, const SiteCoord &c)
{
  return (((long)c.t * g.Lz + c.z) * g.Ly + c.y) * g.Lx + c.x;
}

// XCD-aware work-group remap: the dispatcher hands work-group b to XCD
// b % 8, so logical block (b % 8) * per_xcd + b / 8 gives each XCD one
// contiguous slab of the lattice (neighbour re-reads then hit its own L2).
// The grid is padded to a multiple of 8; padded blocks return.
constexpr int kNumXcd = 8;
__device__ __forceinline__ int xcd_block()
{
  int b = (int)blockIdx.x;
  int per = (int)gridDim.x / kNumXcd;
  return (b % kNumXcd) * per + b / kNumXcd;
}

inline int xcd_grid(long nblocks) { return (int)((nblocks + kNumXcd - 1) / kNumXcd * kNumXcd); }

// Stencil kernels on the lexicographic site order, with `sub` work items
// (fastest) per chunk of blockDim.x consecutive sites.  When Lz is a
// multiple of 8 the XCDs sweep t together: work-group b runs on XCD
// x = b % 8 (the dispatch order), which owns the z-planes [x zs, x zs + zs),
// zs = Lz / 8, and walks them t-slice by t-slice, so a t-neighbour was
// loaded one step earlier by some XCD (one t-slice of data stays in the
// shared last-level cache) and a z-neighbour by the same XCD (its L2).
// Otherwise each XCD sweeps one contiguous range.  Returns the thread's
// lexicographic site, or -1 past the end; item: the sub item.
// bs sites per work-group (blockDim.x by default), this thread's site the
// local-th of them.
__device__ __forceinline__ long sweep_site(const Geom &g, int sub, int &item, int bs, int local)
{
  const long lxy = (long)g.Lx * g.Ly, V = lxy * g.Lz * g.Lt;
  if (g.Lz % kNumXcd) {
    const long b = xcd_block();
    item = (int)(b % sub);
    const long y = (b / sub) * bs + local;
    return y < V ? y : -1;
  }
  const int x = (int)(blockIdx.x % kNumXcd);
  long j = blockIdx.x / kNumXcd;
  item = (int)(j % sub);
  j /= sub;
  const long slab = (long)(g.Lz / kNumXcd) * lxy;  // sites of one XCD per t
  const long nch = (slab + bs - 1) / bs;
  const long t = j / nch, off = (j % nch) * bs + local;
  if (t >= g.Lt || off >= slab) return -1;
  return (t * g.Lz + (long)x * (g.Lz / kNumXcd)) * lxy + off;
}

__device__ __forceinline__ long sweep_site(const Geom &g, int sub, int &item)
{
  return sweep_site(g, sub, item, blockDim.x, threadIdx.x);
}

inline int sweep_grid(const Geom &g, int sub, int bs)
{
  const long lxy = (long)g.Lx * g.Ly, V = lxy * g.Lz * g.Lt;
  if (g.Lz % kNumXcd) return xcd_grid((long)sub * ((V + bs - 1) / bs));
  const long slab = (long)(g.Lz / kNumXcd) * lxy;
  return (int)((long)kNumXcd * g.Lt * ((slab + bs - 1) / bs) * sub);
}

// Stencil kernels on the lexicographic site order visited in (x, y)
// cross-sections: a work-group is `sub` work items (fastest, a direction for
// example) of one chunk of bs sites of a cross-section of Lx * yb sites (yb
// rows, at most kSectionSites sites).  With zparts = 8 (Lz a multiple of 8)
// work-group b runs on XCD b % 8 (the dispatch order), which owns the
// z-planes [k zs, (k + 1) zs); each XCD walks, for one y-block after the
// other, its z-planes (fastest) and then t, so that a site's z- and
// t-neighbours were read a few cross-sections earlier and are still in that
// XCD's L2 (whole t-slabs per XCD put a t-neighbour a whole slab of
// traffic back).
constexpr int kSectionSites = 512;
struct CrossSweep {
  int zparts, zs;  // z-parts (8 or 1) and z-planes per part
  int yb;          // rows of a cross-section (a divisor of Ly)
  int cs, ncx;     // sites of a cross-section and its chunks of bs sites
  int bs, sub;
};

inline CrossSweep cross_sweep(const Geom &g, int bs, int sub)
{
  CrossSweep w;
  w.zparts = g.Lz % kNumXcd ? 1 : kNumXcd;
  w.zs = g.Lz / w.zparts;
  w.yb = 1;
  for (int d = 1; d <= g.Ly; d++)
    if (g.Ly % d == 0 && (long)g.Lx * d <= kSectionSites) w.yb = d;
  w.cs = g.Lx * w.yb;
  w.ncx = (w.cs + bs - 1) / bs;
  w.bs = bs;
  w.sub = sub;
  return w;
}

inline long cross_sweep_blocks(const Geom &g, const CrossSweep &w)
{
  return (long)w.zparts * w.sub * w.ncx * w.zs * g.Lt * (g.Ly / w.yb);
}

// This thread's site (the local-th of its work-group's chunk): its
// coordinates in c and its lexicographic index, or -1 past the end of a
// partial chunk; item: the work item.
__device__ __forceinline__ long cross_sweep_site(const Geom &g, const CrossSweep &w, int local, int &item, int c[4])
{
  long j = blockIdx.x;
  int part = 0;
  if (w.zparts == kNumXcd) {
    part = (int)(j % kNumXcd);
    j /= kNumXcd;
  }
  item = (int)(j % w.sub);
  j /= w.sub;
  const int ch = (int)(j % w.ncx);
  j /= w.ncx;
  const int zl = (int)(j % w.zs);
  j /= w.zs;
  const int t = (int)(j % g.Lt);
  const int yblk = (int)(j / g.Lt);
  const int o = ch * w.bs + local;
  if (o >= w.cs) return -1;
  c[0] = o % g.Lx;
  c[1] = yblk * w.yb + o / g.Lx;
  c[2] = part * w.zs + zl;
  c[3] = t;
  return (((long)c[3] * g.Lz + c[2]) * g.Ly + c[1]) * g.Lx + c[0];
}

}  // namespace qamd
