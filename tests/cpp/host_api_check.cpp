// C++ host program over include/qphix_amd.hpp (test infrastructure).
//
//   host_api_check cpu
//     context-only checks, no device work: geometry queries, the traversal
//     choice, and the exceptions the classes throw on bad arguments.
//   host_api_check gpu <in.bin> <out.bin>
//     reads a lattice written by tests/test_gpu_cpp_host.py, runs every
//     operator family through the C++ classes on the GPU (FP64 and FP32,
//     12-real links, both isigns), writes the unpacked results to out.bin,
//     then solves M^dag M x = b with the even-odd operator (FP64 CG and
//     FP16 mixed CG) and prints one line "cg ..." with both results.
//
// in.bin: int32 dims[4], then float64 arrays, lexicographic: U [V][4][3][3][2],
// psi, chi [V][4][3][2], clov, inv [V][2][36], and full blocks
// [V][2][6][6][2] for (isign +1: B = inverse, F = direct), (isign -1: B, F).
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "qphix_amd.hpp"

namespace {

int fail(const char *msg)
{
  std::fprintf(stderr, "host_api_check: %s\n", msg);
  return 1;
}

#define HIPCHECK(x)                                              \
  do {                                                           \
    hipError_t e_ = (x);                                         \
    if (e_ != hipSuccess) {                                      \
      std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_)); \
      std::exit(1);                                              \
    }                                                            \
  } while (0)

struct Dev {
  void *p = nullptr;
  explicit Dev(size_t bytes) { HIPCHECK(hipMalloc(&p, bytes)); }
  ~Dev() { hipFree(p); }
  Dev(const Dev &) = delete;
  Dev &operator=(const Dev &) = delete;
  double *d() const { return static_cast<double *>(p); }
};

int cpu_mode()
{
  using namespace qphix_amd;
  Geometry g(16, 16, 8, 8, -1.0);
  if (g.sites_cb() != 16 * 16 * 8 * 8 / 2) return fail("sites_cb");
  if (g.field_bytes(QPHIX_AMD_FIELD_SPINOR, QPHIX_AMD_F32) != (size_t)g.sites_cb() * 24 * 4)
    return fail("field_bytes");
  if (g.traversal() != 1) return fail("automatic traversal (Lt <= Lz)");
  g.set_traversal(0);
  if (g.traversal() != 0) return fail("set_traversal");
  bool threw = false;
  try {
    g.set_traversal(7);
  } catch (const Error &e) {
    threw = e.code() == QPHIX_AMD_EINVAL;
  }
  if (!threw) return fail("set_traversal(7) did not throw EINVAL");
  threw = false;
  try {
    Geometry bad(15, 16, 8, 8);
  } catch (const Error &e) {
    threw = e.code() == QPHIX_AMD_EGEOM;
  }
  if (!threw) return fail("odd extent did not throw EGEOM");
  // argument checks happen before any device work
  Dslash D(g, QPHIX_AMD_F32, 12);
  threw = false;
  try {
    D.dslash(nullptr, nullptr, nullptr, 2, 0);
  } catch (const Error &e) {
    threw = e.code() == QPHIX_AMD_EINVAL;
  }
  if (!threw) return fail("isign 2 did not throw EINVAL");
  ClovDslash C(g, QPHIX_AMD_F16, 11);
  threw = false;
  try {
    C.dslash(nullptr, nullptr, nullptr, nullptr, 1, 0);
  } catch (const Error &e) {
    threw = e.code() == QPHIX_AMD_EINVAL;
  }
  if (!threw) return fail("recon 11 did not throw EINVAL");
  // the face, differentiation and link wrappers reach their C entry points
  // (bad arguments are rejected there, before any device work)
  if (g.face_sites(0) != 8 * 8 * 16 / 2 || g.face_sites(3) != 16 * 16 * 8 / 2 || g.face_sites(4) != 0)
    return fail("face_sites");
  double dummy[4];
  TMDslash T(g, QPHIX_AMD_F32, 12, 0.1, 0.1);
  TMClovDslash TC(g, 7, 12);
  const struct {
    const char *what;
    std::function<void()> call;
  } bad[] = {
      {"face_pack dim 4", [&] { g.face_pack(QPHIX_AMD_F32, 1, 0, 4, 0, dummy, dummy + 2); }},
      {"dslash_adjoint cb 2", [&] { g.dslash_adjoint(QPHIX_AMD_F32, 12, 1, 2, dummy, dummy, dummy); }},
      {"dslash_force null force", [&] { g.dslash_force(QPHIX_AMD_F64, 1, 0, nullptr, dummy, dummy, dummy); }},
      {"clover_force null force", [&] { g.clover_force(nullptr, dummy, dummy, dummy, 1.0); }},
      {"reunitarize null links", [&] { g.reunitarize(nullptr); }},
      {"face_finish_achimbdpsi isign 0", [&] { D.face_finish_achimbdpsi(dummy, dummy, dummy, 1.0, 0, 0, 0, 0); }},
      {"clov face_finish_dslash recon 11", [&] { C.face_finish_dslash(dummy, dummy, dummy, dummy, 1.0, 1, 0, 0, 0); }},
      {"tm face_finish_dslash dim 5", [&] { T.face_finish_dslash(dummy, dummy, dummy, -1.0, 1, 0, 5, 0); }},
      {"tm_clov face_finish_dslash prec 7", [&] { TC.face_finish_dslash(dummy, dummy, dummy, dummy, 1.0, 1, 0, 0, 0); }},
  };
  for (const auto &b : bad) {
    threw = false;
    try {
      b.call();
    } catch (const Error &e) {
      threw = e.code() == QPHIX_AMD_EINVAL;
    }
    if (!threw) return fail(b.what);
  }
  std::printf("cpu ok\n");
  return 0;
}

int gpu_mode(const char *in_path, const char *out_path)
{
  using namespace qphix_amd;
  FILE *f = std::fopen(in_path, "rb");
  if (!f) return fail("cannot open input");
  int dims[4];
  if (std::fread(dims, sizeof(int), 4, f) != 4) return fail("short input");
  const long V = (long)dims[0] * dims[1] * dims[2] * dims[3];
  const long nU = V * 72, nS = V * 24, nC = V * 72, nF = V * 144;
  std::vector<double> h(nU + 2 * nS + 2 * nC + 4 * nF);
  if (std::fread(h.data(), sizeof(double), h.size(), f) != h.size()) return fail("short input");
  std::fclose(f);
  const double *hU = h.data(), *hpsi = hU + nU, *hchi = hpsi + nS, *hclov = hchi + nS, *hinv = hclov + nC;
  const double *hBp = hinv + nC, *hFp = hBp + nF, *hBm = hFp + nF, *hFm = hBm + nF;

  Dev U(nU * 8), psi(nS * 8), chi(nS * 8), clov(nC * 8), inv(nC * 8), Bp(nF * 8), Fp(nF * 8), Bm(nF * 8),
      Fm(nF * 8), outlex(nS * 8);
  HIPCHECK(hipMemcpy(U.p, hU, nU * 8, hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(psi.p, hpsi, nS * 8, hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(chi.p, hchi, nS * 8, hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(clov.p, hclov, nC * 8, hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(inv.p, hinv, nC * 8, hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(Bp.p, hBp, nF * 8, hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(Fp.p, hFp, nF * 8, hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(Bm.p, hBm, nF * 8, hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(Fm.p, hFm, nF * 8, hipMemcpyHostToDevice));

  Geometry geom(dims[0], dims[1], dims[2], dims[3], -1.0);
  const double aniso_s = 0.9, aniso_t = 1.1, alpha = 4.2, beta = 0.25, rho = 0.05, mass = 0.2, tmu = 0.1;
  FILE *o = std::fopen(out_path, "wb");
  if (!o) return fail("cannot open output");
  std::vector<double> host(nS);
  const int precs[2] = {QPHIX_AMD_F64, QPHIX_AMD_F32};
  for (int pi = 0; pi < 2; pi++) {
    const int prec = precs[pi];
    Dslash W(geom, prec, 12, aniso_s, aniso_t);
    ClovDslash C(geom, prec, 12, aniso_s, aniso_t);
    TMDslash T(geom, prec, 12, mass, tmu, aniso_s, aniso_t);
    TMClovDslash TC(geom, prec, 12, aniso_s, aniso_t);
    for (int isign = 1; isign >= -1; isign -= 2) {
      const int cb = isign > 0 ? 0 : 1;
      Dev g(geom.field_bytes(QPHIX_AMD_FIELD_GAUGE, prec, 12)), p(geom.field_bytes(QPHIX_AMD_FIELD_SPINOR, prec)),
          c(geom.field_bytes(QPHIX_AMD_FIELD_SPINOR, prec)), r(geom.field_bytes(QPHIX_AMD_FIELD_SPINOR, prec)),
          A(geom.field_bytes(QPHIX_AMD_FIELD_CLOVER, prec)), Ai(geom.field_bytes(QPHIX_AMD_FIELD_CLOVER, prec)),
          B(geom.field_bytes(QPHIX_AMD_FIELD_FULLCLOVER, prec)), F(geom.field_bytes(QPHIX_AMD_FIELD_FULLCLOVER, prec));
      geom.pack_gauge(g.p, prec, 12, U.d(), cb);
      geom.pack_spinor(p.p, prec, psi.d(), 1 - cb);
      geom.pack_spinor(c.p, prec, chi.d(), cb);
      geom.pack_clover(A.p, prec, clov.d(), cb);
      geom.pack_clover(Ai.p, prec, inv.d(), cb);
      geom.pack_full_clover(B.p, prec, (isign > 0 ? Bp : Bm).d(), cb);
      geom.pack_full_clover(F.p, prec, (isign > 0 ? Fp : Fm).d(), cb);
      // the same classes on a QPhiX host's FourSpinorBlock arrays in place
      // (veclen 4 soalen 2 in FP64, 8 and 4 in FP32): the same bits
      const qphix_amd_qphix_layout lay = {prec, prec == QPHIX_AMD_F64 ? 4 : 8, prec == QPHIX_AMD_F64 ? 2 : 4, 1, 0, 0};
      const size_t qb = geom.qphix_field_bytes(lay, QPHIX_AMD_FIELD_SPINOR);
      const size_t nb = geom.field_bytes(QPHIX_AMD_FIELD_SPINOR, prec);
      Dev pq(qb), cq(qb), rq(qb), r2(nb);
      geom.spinor_to_qphix(lay, pq.p, p.p, prec, 1 - cb);
      geom.spinor_to_qphix(lay, cq.p, c.p, prec, cb);
      std::vector<char> n1(nb), n2(nb);
      for (int kind = 0; kind < 4; kind++) {
        for (int achi = 0; achi < 2; achi++) {
          for (int inplace = 0; inplace < 2; inplace++) {
            void *res = inplace ? rq.p : r.p;
            const void *in = inplace ? pq.p : p.p, *ch = inplace ? cq.p : c.p;
            Dslash &base = kind == 0 ? W : kind == 1 ? static_cast<Dslash &>(C)
                                       : kind == 2 ? static_cast<Dslash &>(T) : static_cast<Dslash &>(TC);
            if (inplace)
              base.use_qphix_arrays(lay);
            else
              base.use_native_spinors();
            switch (kind) {
              case 0:
                if (achi) W.dslashAChiMinusBDPsi(res, in, ch, g.p, alpha, beta, isign, cb);
                else W.dslash(res, in, g.p, isign, cb);
                break;
              case 1:
                if (achi) C.dslashAChiMinusBDPsi(res, in, ch, g.p, A.p, beta, isign, cb, rho);
                else C.dslash(res, in, g.p, Ai.p, isign, cb);
                break;
              case 2:
                if (achi) T.dslashAChiMinusBDPsi(res, in, ch, g.p, alpha, beta, isign, cb);
                else T.dslash(res, in, g.p, isign, cb);
                break;
              default:
                if (achi) TC.dslashAChiMinusBDPsi(res, in, ch, g.p, F.p, beta, isign, cb, rho);
                else TC.dslash(res, in, g.p, B.p, isign, cb);
            }
            base.use_native_spinors();
          }
          geom.spinor_from_qphix(lay, r2.p, prec, rq.p, cb);
          HIPCHECK(hipMemcpy(n1.data(), r.p, nb, hipMemcpyDeviceToHost));
          HIPCHECK(hipMemcpy(n2.data(), r2.p, nb, hipMemcpyDeviceToHost));
          if (std::memcmp(n1.data(), n2.data(), nb) != 0) return fail("in-place class call differs from native");
          HIPCHECK(hipMemset(outlex.p, 0, nS * 8));
          geom.unpack_spinor(outlex.d(), r.p, prec, cb);
          HIPCHECK(hipMemcpy(host.data(), outlex.p, nS * 8, hipMemcpyDeviceToHost));
          std::fwrite(host.data(), sizeof(double), nS, o);
        }
      }
    }
  }
  std::fclose(o);

  // even-odd Wilson-clover operator from the links, FP64 CG and FP16 mixed CG
  Dev A(nC * 8);
  geom.clover_from_links(A.d(), U.d(), 1.0 / (2 * 0.135), 1.0);
  EvenOddOperator op(geom, U.d(), A.d(), 0.0, QPHIX_AMD_F16, 12);
  const size_t sb = geom.field_bytes(QPHIX_AMD_FIELD_SPINOR, QPHIX_AMD_F64);
  Dev b(sb), x64(sb), x16(sb);
  geom.pack_spinor(b.p, QPHIX_AMD_F64, psi.d(), 1);
  HIPCHECK(hipMemset(x64.p, 0, sb));
  HIPCHECK(hipMemset(x16.p, 0, sb));
  qphix_amd_cg_result r64 = op.cg(x64.p, b.p, 1e-10, 2000, false);
  qphix_amd_cg_result r16 = op.cg(x16.p, b.p, 1e-10, 2000, true);
  std::vector<double> a64(sb / 8), a16(sb / 8);
  HIPCHECK(hipMemcpy(a64.data(), x64.p, sb, hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(a16.data(), x16.p, sb, hipMemcpyDeviceToHost));
  double num = 0, den = 0;
  for (size_t i = 0; i < a64.size(); i++) {
    num += (a16[i] - a64[i]) * (a16[i] - a64[i]);
    den += a64[i] * a64[i];
  }
  std::printf("cg f64_iters %d f64_rsd %.3e f16_iters %d f16_rsd %.3e f16_reliable %d rel_diff %.3e\n", r64.iters,
              r64.rsd_final, r16.iters, r16.rsd_final, r16.reliable_updates, std::sqrt(num / den));

  // pseudofermion force at the FP64 solution: hopping term alone, then with
  // the clover term (both projected to su(3)), written to <out>.force
  Dev fh(nU * 8), fc(nU * 8);
  HIPCHECK(hipMemset(fh.p, 0, nU * 8));
  HIPCHECK(hipMemset(fc.p, 0, nU * 8));
  op.force(static_cast<double *>(fh.p), x64.p, U.d(), 0.5);
  op.force_clover(static_cast<double *>(fc.p), x64.p, U.d(), 1.0, 0.5);
  std::vector<double> hf(2 * nU);
  HIPCHECK(hipMemcpy(hf.data(), fh.p, nU * 8, hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(hf.data() + nU, fc.p, nU * 8, hipMemcpyDeviceToHost));
  const std::string fpath = std::string(out_path) + ".force";
  FILE *of = std::fopen(fpath.c_str(), "wb");
  if (!of) return fail("cannot open force output");
  std::fwrite(hf.data(), sizeof(double), hf.size(), of);
  std::fclose(of);
  std::printf("force written\n");
  return 0;
}

// A lattice split in two along t, emulated on one GPU through the C++
// classes: each half has its own Geometry (t marked split with set_comm),
// its links handed over in a QPhiX layout with the back links of its
// boundary sites from across the cut (gauge halo), runs every family's
// operator (the interior part), packs both t faces, and completes them from
// the other half's faces through the family's face_finish_* wrapper.
// Input: int dims[4] of a half, int n (doubles of a half's QPhiX gauge
// array: FP64, veclen = soalen = 1, 18 reals), then per half the gauge array
// and lexicographic FP64 psi, chi, A^-1, A, B (the (A + i mu g5)^-1 blocks)
// and F (the A + i mu g5 blocks), isign +1, output checkerboard 0.
// Output: per (family, dslash | achimbdpsi), per half, the lexicographic
// result (checkerboard 0 sites set).
int split_mode(const char *in_path, const char *out_path)
{
  using namespace qphix_amd;
  FILE *f = std::fopen(in_path, "rb");
  if (!f) return fail("cannot open input");
  int dims[4], ng = 0;
  if (std::fread(dims, sizeof(int), 4, f) != 4 || std::fread(&ng, sizeof(int), 1, f) != 1)
    return fail("short input");
  const long V = (long)dims[0] * dims[1] * dims[2] * dims[3];
  const long nS = V * 24, nC = V * 72, nF = V * 144;
  const long per = ng + 2 * nS + 2 * nC + 2 * nF;
  std::vector<double> h(2 * per);
  if (std::fread(h.data(), sizeof(double), h.size(), f) != h.size()) return fail("short input");
  std::fclose(f);

  const double aniso_s = 0.9, aniso_t = 1.1, alpha = 4.2, beta = 0.25, rho = 0.05, mass = 0.2, tmu = 0.1;
  const double t_boundary = -1.0;
  const int prec = QPHIX_AMD_F64, recon = 18, isign = 1, cb = 0, T = 3;
  const qphix_amd_qphix_layout lay = {QPHIX_AMD_F64, 1, 1, 0, 0, 0};
  const bool comm[4] = {false, false, false, true};
  struct Half {
    Geometry geom;
    Dev g, psi, chi, Ai, A, B, F, res, outlex, face0, face1;
    Half(const int *d, long nS)
        : geom(d[0], d[1], d[2], d[3], -1.0),
          g(geom.field_bytes(QPHIX_AMD_FIELD_GAUGE, QPHIX_AMD_F64, 18)),
          psi(geom.field_bytes(QPHIX_AMD_FIELD_SPINOR, QPHIX_AMD_F64)),
          chi(geom.field_bytes(QPHIX_AMD_FIELD_SPINOR, QPHIX_AMD_F64)),
          Ai(geom.field_bytes(QPHIX_AMD_FIELD_CLOVER, QPHIX_AMD_F64)),
          A(geom.field_bytes(QPHIX_AMD_FIELD_CLOVER, QPHIX_AMD_F64)),
          B(geom.field_bytes(QPHIX_AMD_FIELD_FULLCLOVER, QPHIX_AMD_F64)),
          F(geom.field_bytes(QPHIX_AMD_FIELD_FULLCLOVER, QPHIX_AMD_F64)),
          res(geom.field_bytes(QPHIX_AMD_FIELD_SPINOR, QPHIX_AMD_F64)),
          outlex(nS * 8),
          face0(geom.face_sites(3) * 12 * 8),
          face1(geom.face_sites(3) * 12 * 8)
    {
    }
    void *face(int fb) const { return fb ? face1.p : face0.p; }
  };
  std::vector<std::unique_ptr<Half>> half;
  for (int k = 0; k < 2; k++) {
    half.emplace_back(new Half(dims, nS));
    Half &H = *half[k];
    H.geom.set_comm(comm);
    const double *src = h.data() + k * per;
    Dev qg(ng * 8), lex(nF * 8);
    HIPCHECK(hipMemcpy(qg.p, src, ng * 8, hipMemcpyHostToDevice));
    if (H.geom.qphix_field_bytes(lay, QPHIX_AMD_FIELD_GAUGE) != (size_t)ng * 8) return fail("gauge array size");
    H.geom.gauge_from_qphix(lay, H.g.p, prec, recon, qg.p, cb);
    src += ng;
    struct {
      Dev *dst;
      long n;
      int kind;  // 0 spinor cb 1, 1 spinor cb 0, 2 clover, 3 full clover
    } fields[] = {{&H.psi, nS, 0}, {&H.chi, nS, 1}, {&H.Ai, nC, 2}, {&H.A, nC, 2}, {&H.B, nF, 3}, {&H.F, nF, 3}};
    for (const auto &fd : fields) {
      HIPCHECK(hipMemcpy(lex.p, src, fd.n * 8, hipMemcpyHostToDevice));
      if (fd.kind == 0) H.geom.pack_spinor(fd.dst->p, prec, lex.d(), 1 - cb);
      if (fd.kind == 1) H.geom.pack_spinor(fd.dst->p, prec, lex.d(), cb);
      if (fd.kind == 2) H.geom.pack_clover(fd.dst->p, prec, lex.d(), cb);
      if (fd.kind == 3) H.geom.pack_full_clover(fd.dst->p, prec, lex.d(), cb);
      src += fd.n;
    }
  }
  // the same split again with the spinors kept in QPhiX FourSpinorBlock
  // arrays (FP64, veclen = soalen = 2): the classes in their in-place mode,
  // faces packed from and finished into those arrays (must give the same
  // bits as the native run)
  const qphix_amd_qphix_layout lq = {QPHIX_AMD_F64, 2, 2, 0, 0, 0};
  std::vector<std::unique_ptr<Dev>> qarr;  // per half: psi, chi, res
  for (int k = 0; k < 2; k++) {
    Half &H = *half[k];
    const size_t qb = H.geom.qphix_field_bytes(lq, QPHIX_AMD_FIELD_SPINOR);
    for (int a = 0; a < 3; a++) qarr.emplace_back(new Dev(qb));
    H.geom.spinor_to_qphix(lq, qarr[3 * k]->p, H.psi.p, prec, 1 - cb);
    H.geom.spinor_to_qphix(lq, qarr[3 * k + 1]->p, H.chi.p, prec, cb);
  }
  const size_t nb = half[0]->geom.field_bytes(QPHIX_AMD_FIELD_SPINOR, prec);
  Dev back(nb);
  std::vector<char> hn(nb), hq(nb);

  // one (family, kernel) on both halves: interior, faces, finishes
  auto run = [&](int kind, int achi, bool inplace) {
    for (int k = 0; k < 2; k++) {
      Half &H = *half[k];
      void *res = inplace ? qarr[3 * k + 2]->p : H.res.p;
      const void *psi = inplace ? qarr[3 * k]->p : H.psi.p, *chi = inplace ? qarr[3 * k + 1]->p : H.chi.p;
      Dslash W(H.geom, prec, recon, aniso_s, aniso_t);
      ClovDslash C(H.geom, prec, recon, aniso_s, aniso_t);
      TMDslash Tm(H.geom, prec, recon, mass, tmu, aniso_s, aniso_t);
      TMClovDslash TC(H.geom, prec, recon, aniso_s, aniso_t);
      if (inplace) {
        W.use_qphix_arrays(lq);
        C.use_qphix_arrays(lq);
        Tm.use_qphix_arrays(lq);
        TC.use_qphix_arrays(lq);
      }
      switch (kind) {  // the interior part
        case 0:
          if (achi) W.dslashAChiMinusBDPsi(res, psi, chi, H.g.p, alpha, beta, isign, cb);
          else W.dslash(res, psi, H.g.p, isign, cb);
          break;
        case 1:
          if (achi) C.dslashAChiMinusBDPsi(res, psi, chi, H.g.p, H.A.p, beta, isign, cb, rho);
          else C.dslash(res, psi, H.g.p, H.Ai.p, isign, cb);
          break;
        case 2:
          if (achi) Tm.dslashAChiMinusBDPsi(res, psi, chi, H.g.p, alpha, beta, isign, cb);
          else Tm.dslash(res, psi, H.g.p, isign, cb);
          break;
        default:
          if (achi) TC.dslashAChiMinusBDPsi(res, psi, chi, H.g.p, H.F.p, beta, isign, cb, rho);
          else TC.dslash(res, psi, H.g.p, H.B.p, isign, cb);
      }
      for (int fb = 0; fb < 2; fb++) {
        if (inplace)
          H.geom.face_pack(lq, isign, cb, T, fb, H.face(fb), psi);
        else
          H.geom.face_pack(prec, isign, cb, T, fb, H.face(fb), psi);
      }
    }
    for (int k = 0; k < 2; k++) {
      Half &H = *half[k];
      void *res = inplace ? qarr[3 * k + 2]->p : H.res.p;
      Dslash W(H.geom, prec, recon, aniso_s, aniso_t);
      ClovDslash C(H.geom, prec, recon, aniso_s, aniso_t);
      TMDslash Tm(H.geom, prec, recon, mass, tmu, aniso_s, aniso_t);
      TMClovDslash TC(H.geom, prec, recon, aniso_s, aniso_t);
      if (inplace) {
        W.use_qphix_arrays(lq);
        C.use_qphix_arrays(lq);
        Tm.use_qphix_arrays(lq);
        TC.use_qphix_arrays(lq);
      }
      for (int fb = 0; fb < 2; fb++) {
        // fb 0 completes t = 0 from the half behind (its forward face),
        // fb 1 t = Lt - 1 from the half ahead (its backward face)
        const Half &nb = *half[1 - k];
        const void *buf = nb.face(1 - fb);
        const bool edge = fb == 0 ? k == 0 : k == 1;  // the global time boundary
        const double coeff = aniso_t * (edge ? t_boundary : 1.0);
        if (achi) {
          const double b = beta * coeff;
          switch (kind) {
            case 0: W.face_finish_achimbdpsi(res, buf, H.g.p, b, isign, cb, T, fb); break;
            case 1: C.face_finish_achimbdpsi(res, buf, H.g.p, b, isign, cb, T, fb); break;
            case 2: Tm.face_finish_achimbdpsi(res, buf, H.g.p, b, isign, cb, T, fb); break;
            default: TC.face_finish_achimbdpsi(res, buf, H.g.p, b, isign, cb, T, fb);
          }
        } else {
          switch (kind) {
            case 0: W.face_finish_dslash(res, buf, H.g.p, -coeff, isign, cb, T, fb); break;
            case 1: C.face_finish_dslash(res, buf, H.g.p, H.Ai.p, coeff, isign, cb, T, fb); break;
            case 2: Tm.face_finish_dslash(res, buf, H.g.p, -coeff, isign, cb, T, fb); break;
            default: TC.face_finish_dslash(res, buf, H.g.p, H.B.p, coeff, isign, cb, T, fb);
          }
        }
      }
    }
  };

  FILE *o = std::fopen(out_path, "wb");
  if (!o) return fail("cannot open output");
  std::vector<double> host(nS);
  for (int kind = 0; kind < 4; kind++)
    for (int achi = 0; achi < 2; achi++) {
      run(kind, achi, false);
      for (int k = 0; k < 2; k++) {
        Half &H = *half[k];
        HIPCHECK(hipMemset(H.outlex.p, 0, nS * 8));
        H.geom.unpack_spinor(H.outlex.d(), H.res.p, prec, cb);
        HIPCHECK(hipMemcpy(host.data(), H.outlex.p, nS * 8, hipMemcpyDeviceToHost));
        std::fwrite(host.data(), sizeof(double), nS, o);
      }
      run(kind, achi, true);
      for (int k = 0; k < 2; k++) {
        Half &H = *half[k];
        H.geom.spinor_from_qphix(lq, back.p, prec, qarr[3 * k + 2]->p, cb);
        HIPCHECK(hipMemcpy(hn.data(), H.res.p, nb, hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(hq.data(), back.p, nb, hipMemcpyDeviceToHost));
        if (std::memcmp(hn.data(), hq.data(), nb) != 0) return fail("in-place split differs from native split");
      }
    }
  std::fclose(o);
  std::printf("split ok (in place on QPhiX arrays: same bits)\n");
  return 0;
}

}  // namespace

int main(int argc, char **argv)
{
  try {
    if (argc >= 2 && std::strcmp(argv[1], "cpu") == 0) return cpu_mode();
    if (argc >= 4 && std::strcmp(argv[1], "gpu") == 0) return gpu_mode(argv[2], argv[3]);
    if (argc >= 4 && std::strcmp(argv[1], "split") == 0) return split_mode(argv[2], argv[3]);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "host_api_check: %s\n", e.what());
    return 1;
  }
  return fail("usage: host_api_check cpu | gpu <in.bin> <out.bin> | split <in.bin> <out.bin>");
}
