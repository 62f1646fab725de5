/* qphix_amd.hpp -- C++ host classes over the C ABI of qphix_amd.h, shaped
   like the QPhiX classes that call the reference's generated kernels:
   Dslash, ClovDslash, TMDslash and TMClovDslash with dslash() and
   dslashAChiMinusBDPsi(), plus a geometry that owns the context and the
   even-odd operator / CG on top.

   The reference declares one function template per body kernel,
   specialised on <FT, veclen, soalen, compress12> and called once per SIMD
   block by QPhiX (codegen/jinja/dslash_decl.h.j2:6-115,
   clov_dslash_decl.h.j2:6-113, tm_dslash_decl.h.j2:6-109,
   tm_clov_dslash_decl.h.j2:6-113).  Here the storage type and 12/18-real
   links are run-time arguments, one call covers a checkerboard, and the
   fields are device pointers in the native layout (convert QPhiX block
   arrays with Geometry::*_from_qphix).  Every call is asynchronous on the
   stream given to the constructor (nullptr: the default stream).

   Header only, C++11, no HIP headers needed: link with -lqphix_amd.
   Errors from the library are thrown as qphix_amd::Error (the reference
   prints and exits on a bad direction, codegen/jinja/dslash_spec.h.j2:28-31;
   the C ABI returns a status instead).                                    */
#ifndef QPHIX_AMD_HPP
#define QPHIX_AMD_HPP

#include <stdexcept>
#include <string>

#include "qphix_amd.h"

namespace qphix_amd {

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string &what)
      : std::runtime_error(what + ": " + qphix_amd_error_string(code)), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

inline void check(int rc, const char *what)
{
  if (rc != QPHIX_AMD_OK) throw Error(rc, what);
}

/* call C entry point fn and throw its status, named, if it is not OK */
#define QPHIX_AMD_CALL(fn, ...) ::qphix_amd::check(fn(__VA_ARGS__), #fn)

/* Storage precision of a QPhiX floating type (float16 storage: pass
   QPHIX_AMD_F16 explicitly). */
template <typename FT> struct Prec;
template <> struct Prec<double> { static const int value = QPHIX_AMD_F64; };
template <> struct Prec<float> { static const int value = QPHIX_AMD_F32; };

/* Lattice geometry and its context (QPhiX Geometry; antiperiodic time is
   t_boundary = -1).  Owns the context. */
class Geometry {
 public:
  Geometry(int Lx, int Ly, int Lz, int Lt, double t_boundary = -1.0) : ctx_(nullptr)
  {
    QPHIX_AMD_CALL(qphix_amd_ctx_create, &ctx_, Lx, Ly, Lz, Lt, t_boundary);
  }
  ~Geometry()
  {
    if (ctx_) qphix_amd_ctx_destroy(ctx_);
  }
  Geometry(const Geometry &) = delete;
  Geometry &operator=(const Geometry &) = delete;
  Geometry(Geometry &&o) noexcept : ctx_(o.ctx_) { o.ctx_ = nullptr; }

  qphix_amd_ctx *ctx() const { return ctx_; }
  long sites_cb() const { return qphix_amd_sites_cb(ctx_); }
  /* bytes of one checkerboard of a native field */
  size_t field_bytes(int field, int prec, int recon = 18) const
  {
    return qphix_amd_field_bytes(ctx_, field, prec, recon);
  }

  /* twisted boundary phases: tbc[mu] != 0 multiplies forward hops in mu by
     phases[mu] (re, im) and backward hops by its conjugate */
  void set_tbc(const bool tbc[4], const double phases[4][2])
  {
    int t[4];
    double p[8];
    for (int m = 0; m < 4; m++) {
      t[m] = tbc[m] ? 1 : 0;
      p[2 * m] = phases[m][0];
      p[2 * m + 1] = phases[m][1];
    }
    QPHIX_AMD_CALL(qphix_amd_set_tbc, ctx_, t, p);
  }
  /* split dimensions of a lattice spread over several GPUs */
  void set_comm(const bool comm[4])
  {
    int c[4];
    for (int m = 0; m < 4; m++) c[m] = comm[m] ? 1 : 0;
    QPHIX_AMD_CALL(qphix_amd_set_comm, ctx_, c);
  }
  void set_traversal(int order) { QPHIX_AMD_CALL(qphix_amd_set_traversal, ctx_, order); }
  int traversal() const { return qphix_amd_traversal(ctx_); }

  /* device lexicographic FP64 <-> native (see qphix_amd.h for the layouts) */
  void pack_spinor(void *dst, int prec, const double *lex, int cb, void *stream = nullptr) const
  {
    QPHIX_AMD_CALL(qphix_amd_pack_spinor, ctx_, dst, prec, lex, cb, stream);
  }
  void unpack_spinor(double *lex, const void *src, int prec, int cb, void *stream = nullptr) const
  {
    QPHIX_AMD_CALL(qphix_amd_unpack_spinor, ctx_, lex, src, prec, cb, stream);
  }
  void pack_gauge(void *dst, int prec, int recon, const double *u_lex, int cb, void *stream = nullptr) const
  {
    QPHIX_AMD_CALL(qphix_amd_pack_gauge, ctx_, dst, prec, recon, u_lex, cb, stream);
  }
  void pack_clover(void *dst, int prec, const double *clov_lex, int cb, void *stream = nullptr) const
  {
    QPHIX_AMD_CALL(qphix_amd_pack_clover, ctx_, dst, prec, clov_lex, cb, stream);
  }
  void pack_full_clover(void *dst, int prec, const double *fclov_lex, int cb, void *stream = nullptr) const
  {
    QPHIX_AMD_CALL(qphix_amd_pack_full_clover, ctx_, dst, prec, fclov_lex, cb, stream);
  }

  /* clover term from links and its inverse / twisted-mass blocks (device
     lexicographic FP64) */
  void clover_from_links(double *clov_lex, const double *u_lex, double alpha, double csw,
                         void *stream = nullptr) const
  {
    QPHIX_AMD_CALL(qphix_amd_clover_from_links, ctx_, clov_lex, u_lex, alpha, csw, stream);
  }
  void clover_invert(double *inv_lex, const double *clov_lex, void *stream = nullptr) const
  {
    QPHIX_AMD_CALL(qphix_amd_clover_invert, ctx_, inv_lex, clov_lex, stream);
  }
  void tmclover_blocks(double *out_lex, const double *clov_lex, double mu, int sign, bool invert,
                       void *stream = nullptr) const
  {
    QPHIX_AMD_CALL(qphix_amd_tmclover_blocks, ctx_, out_lex, clov_lex, mu, sign, invert ? 1 : 0, stream);
  }

  /* device QPhiX block arrays -> native */
  void spinor_from_qphix(const qphix_amd_qphix_layout &lay, void *dst, int prec, const void *src, int cb,
                         void *stream = nullptr) const
  {
    QPHIX_AMD_CALL(qphix_amd_spinor_from_qphix, ctx_, &lay, dst, prec, src, cb, stream);
  }
  void spinor_to_qphix(const qphix_amd_qphix_layout &lay, void *dst, const void *src, int prec, int cb,
                       void *stream = nullptr) const
  {
    QPHIX_AMD_CALL(qphix_amd_spinor_to_qphix, ctx_, &lay, dst, src, prec, cb, stream);
  }
  void gauge_from_qphix(const qphix_amd_qphix_layout &lay, void *dst, int prec, int recon, const void *src,
                        int cb, void *stream = nullptr) const
  {
    QPHIX_AMD_CALL(qphix_amd_gauge_from_qphix, ctx_, &lay, dst, prec, recon, src, cb, stream);
  }
  void clover_from_qphix(const qphix_amd_qphix_layout &lay, void *dst, int prec, const void *src, int cb,
                         void *stream = nullptr) const
  {
    QPHIX_AMD_CALL(qphix_amd_clover_from_qphix, ctx_, &lay, dst, prec, src, cb, stream);
  }
  void full_clover_from_qphix(const qphix_amd_qphix_layout &lay, void *dst, int prec, const void *src, int cb,
                              void *stream = nullptr) const
  {
    QPHIX_AMD_CALL(qphix_amd_full_clover_from_qphix, ctx_, &lay, dst, prec, src, cb, stream);
  }
  /* bytes of a QPhiX block array of this lattice in layout lay */
  size_t qphix_field_bytes(const qphix_amd_qphix_layout &lay, int field) const
  {
    return qphix_amd_qphix_field_bytes(ctx_, &lay, field);
  }

  /* faces of a lattice split over several GPUs (set_comm): half-spinors
     per face, and the projected half-spinors of one face of `in`; each
     operator class completes a face with face_finish_dslash or
     face_finish_achimbdpsi */
  long face_sites(int dim) const { return qphix_amd_face_sites(ctx_, dim); }
  void face_pack(int prec, int isign, int cb, int dim, int fb, void *buf, const void *in,
                 void *stream = nullptr) const
  {
    QPHIX_AMD_CALL(qphix_amd_face_pack, ctx_, prec, isign, cb, dim, fb, buf, in, stream);
  }
  /* the same from a QPhiX FourSpinorBlock array `in` of layout lay, read in
     place */
  void face_pack(const qphix_amd_qphix_layout &lay, int isign, int cb, int dim, int fb, void *buf, const void *in,
                 void *stream = nullptr) const
  {
    QPHIX_AMD_CALL(qphix_amd_qphix_face_pack, ctx_, &lay, isign, cb, dim, fb, buf, in, stream);
  }

  /* differentiation of L = Re<chi, D psi> (no QPhiX counterpart): the
     adjoint D^dagger chi on checkerboard 1-cb, and the link gradient
     (project: su(3) force) accumulated into lexicographic FP64 force_lex */
  void dslash_adjoint(int prec, int recon, int isign, int cb, void *grad_psi, const void *chi,
                      const void *gauge, double coeff_s = 1.0, double coeff_t_f = 1.0, double coeff_t_b = 1.0,
                      void *stream = nullptr) const
  {
    QPHIX_AMD_CALL(qphix_amd_dslash_adjoint, ctx_, prec, recon, isign, cb, grad_psi, chi, gauge, coeff_s,
                   coeff_t_f, coeff_t_b, stream);
  }
  void dslash_force(int prec, int isign, int cb, double *force_lex, const void *chi, const void *psi,
                    const double *u_lex, double scale = 1.0, bool project = true, double coeff_s = 1.0,
                    double coeff_t_f = 1.0, double coeff_t_b = 1.0, void *stream = nullptr) const
  {
    QPHIX_AMD_CALL(qphix_amd_dslash_force, ctx_, prec, isign, cb, force_lex, chi, psi, u_lex, coeff_s, coeff_t_f,
                   coeff_t_b, scale, project ? 1 : 0, stream);
  }
  /* link gradient of Re sum_x w(x) chi^dag A psi, A = clover_from_links(.., csw),
     w = scale_even / scale_odd by site parity (lexicographic FP64 fields) */
  void clover_force(double *force_lex, const double *chi_lex, const double *psi_lex, const double *u_lex,
                    double csw, double scale_even = 1.0, double scale_odd = 1.0, bool project = true,
                    void *stream = nullptr) const
  {
    QPHIX_AMD_CALL(qphix_amd_clover_force, ctx_, force_lex, chi_lex, psi_lex, u_lex, csw, scale_even, scale_odd,
                   project ? 1 : 0, stream);
  }
  /* project lexicographic FP64 links back onto SU(3), in place */
  void reunitarize(double *u_lex, void *stream = nullptr) const
  {
    QPHIX_AMD_CALL(qphix_amd_reunitarize, ctx_, u_lex, stream);
  }

 private:
  qphix_amd_ctx *ctx_;
};

/* Wilson hopping term (QPhiX Dslash):
     dslash:               res = D psi                 (beta 1 per hop)
     dslashAChiMinusBDPsi: res = alpha chi - beta D psi
   with the anisotropy coefficients on the spatial / time hops. */
class Dslash {
 public:
  Dslash(const Geometry &geom, int prec, int recon, double aniso_coeff_s = 1.0, double aniso_coeff_t = 1.0,
         void *stream = nullptr)
      : ctx_(geom.ctx()), prec_(prec), recon_(recon), cs_(aniso_coeff_s), ct_(aniso_coeff_t), stream_(stream)
  {
  }

  /* Operate on a QPhiX host's FourSpinorBlock arrays in place: after this
     call res, psi and chi of every method (and res of the face finishes)
     are arrays of the layout lay (FT = the class's precision), with the
     optional per-site accumulate mask (one byte per slot of the padded
     array); links and clover blocks stay native fields.
     use_native_spinors() switches back. */
  void use_qphix_arrays(const qphix_amd_qphix_layout &lay, const unsigned char *accumulate = nullptr)
  {
    if (lay.ft != prec_) throw Error(QPHIX_AMD_EINVAL, "use_qphix_arrays: layout FT is not the class's precision");
    lay_ = lay;
    q_ = true;
    acc_ = accumulate;
  }
  void use_native_spinors()
  {
    q_ = false;
    acc_ = nullptr;
  }

  void dslash(void *res, const void *psi, const void *u, int isign, int cb) const
  {
    if (q_)
      QPHIX_AMD_CALL(qphix_amd_qphix_dslash, ctx_, &lay_, recon_, isign, cb, res, psi, u, acc_, cs_, ct_, ct_,
                     stream_);
    else
      QPHIX_AMD_CALL(qphix_amd_dslash, ctx_, prec_, recon_, isign, cb, res, psi, u, cs_, ct_, ct_, stream_);
  }
  void dslashAChiMinusBDPsi(void *res, const void *psi, const void *chi, const void *u, double alpha,
                            double beta, int isign, int cb) const
  {
    if (q_)
      QPHIX_AMD_CALL(qphix_amd_qphix_dslash_achimbdpsi, ctx_, &lay_, recon_, isign, cb, res, psi, chi, u, acc_,
                     alpha, beta * cs_, beta * ct_, beta * ct_, stream_);
    else
      QPHIX_AMD_CALL(qphix_amd_dslash_achimbdpsi, ctx_, prec_, recon_, isign, cb, res, psi, chi, u, alpha,
                     beta * cs_, beta * ct_, beta * ct_, stream_);
  }
  /* complete the sites of face (dim, fb) from a received face buffer:
     after dslash (beta = -coeff; each family has its own), and after
     dslashAChiMinusBDPsi (beta = +coeff, the same for every family) */
  void face_finish_dslash(void *res, const void *buf, const void *u, double beta, int isign, int cb, int dim,
                          int fb) const
  {
    face_finish_achimbdpsi(res, buf, u, beta, isign, cb, dim, fb);
  }
  void face_finish_achimbdpsi(void *res, const void *buf, const void *u, double beta, int isign, int cb, int dim,
                              int fb) const
  {
    if (q_)
      QPHIX_AMD_CALL(qphix_amd_qphix_face_finish, ctx_, &lay_, recon_, isign, cb, dim, fb, res, buf, u, beta,
                     stream_);
    else
      QPHIX_AMD_CALL(qphix_amd_face_finish, ctx_, prec_, recon_, isign, cb, dim, fb, res, buf, u, beta, stream_);
  }

 protected:
  qphix_amd_ctx *ctx_;
  int prec_, recon_;
  double cs_, ct_;
  void *stream_;
  bool q_ = false;  // spinors are QPhiX arrays of lay_
  qphix_amd_qphix_layout lay_{};
  const unsigned char *acc_ = nullptr;
};

/* Wilson-clover (QPhiX ClovDslash):
     dslash:               res = A^-1 D psi            (invclov packed blocks)
     dslashAChiMinusBDPsi: res = A chi - beta D psi (+ i rho g5 chi) */
class ClovDslash : public Dslash {
 public:
  using Dslash::Dslash;

  void dslash(void *res, const void *psi, const void *u, const void *invclov, int isign, int cb) const
  {
    if (q_)
      QPHIX_AMD_CALL(qphix_amd_qphix_clov_dslash, ctx_, &lay_, recon_, isign, cb, res, psi, u, invclov, acc_, cs_,
                     ct_, ct_, stream_);
    else
      QPHIX_AMD_CALL(qphix_amd_clov_dslash, ctx_, prec_, recon_, isign, cb, res, psi, u, invclov, cs_, ct_, ct_,
                     stream_);
  }
  void dslashAChiMinusBDPsi(void *res, const void *psi, const void *chi, const void *u, const void *clov,
                            double beta, int isign, int cb, double prec_mass_rho = 0.0) const
  {
    if (q_)
      QPHIX_AMD_CALL(qphix_amd_qphix_clov_dslash_achimbdpsi, ctx_, &lay_, recon_, isign, cb, res, psi, chi, u,
                     clov, acc_, beta * cs_, beta * ct_, beta * ct_, prec_mass_rho, stream_);
    else
      QPHIX_AMD_CALL(qphix_amd_clov_dslash_achimbdpsi, ctx_, prec_, recon_, isign, cb, res, psi, chi, u, clov,
                     beta * cs_, beta * ct_, beta * ct_, prec_mass_rho, stream_);
  }
  /* face completion of dslash (beta = +coeff) */
  void face_finish_dslash(void *res, const void *buf, const void *u, const void *invclov, double beta, int isign,
                          int cb, int dim, int fb) const
  {
    if (q_)
      QPHIX_AMD_CALL(qphix_amd_qphix_clov_face_finish, ctx_, &lay_, recon_, isign, cb, dim, fb, res, buf, u,
                     invclov, beta, stream_);
    else
      QPHIX_AMD_CALL(qphix_amd_clov_face_finish, ctx_, prec_, recon_, isign, cb, dim, fb, res, buf, u, invclov,
                     beta, stream_);
  }
};

/* Twisted-mass Wilson (QPhiX TMDslash), alpha = 4 + mass.  The kernels take
   mu = twisted_mass / alpha and muinv = alpha / (alpha^2 + twisted_mass^2)
   (codegen/dslash_common.cc:1256-1266):
     dslash:               res = muinv (1 - isign i mu g5) D psi
     dslashAChiMinusBDPsi: res = alpha (1 + isign i mu' g5) chi - beta D psi,
                           mu' = twisted_mass / alpha of the call */
class TMDslash : public Dslash {
 public:
  TMDslash(const Geometry &geom, int prec, int recon, double mass, double twisted_mass,
           double aniso_coeff_s = 1.0, double aniso_coeff_t = 1.0, void *stream = nullptr)
      : Dslash(geom, prec, recon, aniso_coeff_s, aniso_coeff_t, stream),
        alpha_(4.0 + mass),
        mu_phys_(twisted_mass)
  {
  }

  void dslash(void *res, const void *psi, const void *u, int isign, int cb) const
  {
    const double mu = mu_phys_ / alpha_, muinv = alpha_ / (alpha_ * alpha_ + mu_phys_ * mu_phys_);
    if (q_)
      QPHIX_AMD_CALL(qphix_amd_qphix_tm_dslash, ctx_, &lay_, recon_, isign, cb, res, psi, u, acc_, cs_, ct_, ct_,
                     mu, muinv, stream_);
    else
      QPHIX_AMD_CALL(qphix_amd_tm_dslash, ctx_, prec_, recon_, isign, cb, res, psi, u, cs_, ct_, ct_, mu, muinv,
                     stream_);
  }
  void dslashAChiMinusBDPsi(void *res, const void *psi, const void *chi, const void *u, double alpha,
                            double beta, int isign, int cb) const
  {
    if (q_)
      QPHIX_AMD_CALL(qphix_amd_qphix_tm_dslash_achimbdpsi, ctx_, &lay_, recon_, isign, cb, res, psi, chi, u, acc_,
                     alpha, beta * cs_, beta * ct_, beta * ct_, mu_phys_ / alpha, stream_);
    else
      QPHIX_AMD_CALL(qphix_amd_tm_dslash_achimbdpsi, ctx_, prec_, recon_, isign, cb, res, psi, chi, u, alpha,
                     beta * cs_, beta * ct_, beta * ct_, mu_phys_ / alpha, stream_);
  }
  /* face completion of dslash (beta = -coeff), with the twisted epilogue */
  void face_finish_dslash(void *res, const void *buf, const void *u, double beta, int isign, int cb, int dim,
                          int fb) const
  {
    const double mu = mu_phys_ / alpha_, muinv = alpha_ / (alpha_ * alpha_ + mu_phys_ * mu_phys_);
    if (q_)
      QPHIX_AMD_CALL(qphix_amd_qphix_tm_face_finish, ctx_, &lay_, recon_, isign, cb, dim, fb, res, buf, u, beta,
                     mu, muinv, stream_);
    else
      QPHIX_AMD_CALL(qphix_amd_tm_face_finish, ctx_, prec_, recon_, isign, cb, dim, fb, res, buf, u, beta, mu,
                     muinv, stream_);
  }

 private:
  double alpha_, mu_phys_;
};

/* Twisted-mass clover (QPhiX TMClovDslash), full complex 6x6 blocks:
     dslash:               res = (A + i isign mu g5)^-1 D psi  (fclov: the inverse blocks)
     dslashAChiMinusBDPsi: res = (A + i isign mu g5) chi - beta D psi (+ i rho g5 chi) */
class TMClovDslash : public Dslash {
 public:
  using Dslash::Dslash;

  void dslash(void *res, const void *psi, const void *u, const void *invfclov, int isign, int cb) const
  {
    if (q_)
      QPHIX_AMD_CALL(qphix_amd_qphix_tm_clov_dslash, ctx_, &lay_, recon_, isign, cb, res, psi, u, invfclov, acc_,
                     cs_, ct_, ct_, stream_);
    else
      QPHIX_AMD_CALL(qphix_amd_tm_clov_dslash, ctx_, prec_, recon_, isign, cb, res, psi, u, invfclov, cs_, ct_,
                     ct_, stream_);
  }
  void dslashAChiMinusBDPsi(void *res, const void *psi, const void *chi, const void *u, const void *fclov,
                            double beta, int isign, int cb, double prec_mass_rho = 0.0) const
  {
    if (q_)
      QPHIX_AMD_CALL(qphix_amd_qphix_tm_clov_dslash_achimbdpsi, ctx_, &lay_, recon_, isign, cb, res, psi, chi, u,
                     fclov, acc_, beta * cs_, beta * ct_, beta * ct_, prec_mass_rho, stream_);
    else
      QPHIX_AMD_CALL(qphix_amd_tm_clov_dslash_achimbdpsi, ctx_, prec_, recon_, isign, cb, res, psi, chi, u, fclov,
                     beta * cs_, beta * ct_, beta * ct_, prec_mass_rho, stream_);
  }
  /* face completion of dslash (beta = +coeff) */
  void face_finish_dslash(void *res, const void *buf, const void *u, const void *invfclov, double beta, int isign,
                          int cb, int dim, int fb) const
  {
    if (q_)
      QPHIX_AMD_CALL(qphix_amd_qphix_tm_clov_face_finish, ctx_, &lay_, recon_, isign, cb, dim, fb, res, buf, u,
                     invfclov, beta, stream_);
    else
      QPHIX_AMD_CALL(qphix_amd_tm_clov_face_finish, ctx_, prec_, recon_, isign, cb, dim, fb, res, buf, u, invfclov,
                     beta, stream_);
  }
};

/* Even-odd clover / twisted-mass-clover operator M = F_oo - 1/4 D_oe
   F_ee^-1 D_eo with its mixed-precision CG (the solver QPhiX builds on its
   operator classes).  Owns device memory; u_lex / clov_lex are device
   lexicographic FP64 arrays. */
class EvenOddOperator {
 public:
  EvenOddOperator(const Geometry &geom, const double *u_lex, const double *clov_lex, double mu, int prec_sloppy,
                  int recon_sloppy, void *stream = nullptr)
      : op_(nullptr), stream_(stream)
  {
    QPHIX_AMD_CALL(qphix_amd_eo_op_create, &op_, geom.ctx(), u_lex, clov_lex, mu, prec_sloppy, recon_sloppy, stream);
  }
  ~EvenOddOperator()
  {
    if (op_) qphix_amd_eo_op_destroy(op_);
  }
  EvenOddOperator(const EvenOddOperator &) = delete;
  EvenOddOperator &operator=(const EvenOddOperator &) = delete;

  qphix_amd_eo_op *get() const { return op_; }
  /* out = M in (dagger false) or M^dagger in; prec F64 or the sloppy one */
  void apply(void *out, const void *in, int prec, bool dagger = false) const
  {
    QPHIX_AMD_CALL(qphix_amd_eo_op_apply, op_, prec, dagger ? 1 : 0, out, in, stream_);
  }
  /* solve M^dag M x = b (FP64 native spinors on cb 1); throws on failure,
     including QPHIX_AMD_ENOCONV when max_iters is reached */
  qphix_amd_cg_result cg(void *x, const void *b, double rsd_target, int max_iters, bool mixed = true,
                         double delta = 0.0) const
  {
    qphix_amd_cg_params p;
    p.rsd_target = rsd_target;
    p.max_iters = max_iters;
    p.delta = delta;
    p.mixed = mixed ? 1 : 0;
    qphix_amd_cg_result r;
    QPHIX_AMD_CALL(qphix_amd_cg_solve, op_, x, b, &p, &r, stream_);
    return r;
  }
  /* accumulate scale * dS/dU of S = phi^dag (M^dag M)^-1 phi at
     x = (M^dag M)^-1 phi (FP64 native, cb 1) into force_lex (lexicographic
     FP64, 4 links x 18 reals per site): the hopping term; project: su(3)
     force, else the link gradient */
  void force(double *force_lex, const void *x, const double *u_lex, double scale = 1.0,
             bool project = true) const
  {
    QPHIX_AMD_CALL(qphix_amd_eo_op_force, op_, force_lex, x, u_lex, scale, project ? 1 : 0, stream_);
  }
  /* the same with the clover term, for an operator built on
     clover_from_links(.., csw) of the same links */
  void force_clover(double *force_lex, const void *x, const double *u_lex, double csw, double scale = 1.0,
                    bool project = true) const
  {
    QPHIX_AMD_CALL(qphix_amd_eo_op_force_clover, op_, force_lex, x, u_lex, csw, scale, project ? 1 : 0, stream_);
  }

 private:
  qphix_amd_eo_op *op_;
  void *stream_;
};

}  // namespace qphix_amd

#undef QPHIX_AMD_CALL

#endif /* QPHIX_AMD_HPP */
