/*
 * qphix_oracle.c -- FP64 CPU restatement of the Wilson-type Dirac operators
 * that qphix-codegen emits as SIMD kernel bodies.
 *
 * THIS FILE IS TEST INFRASTRUCTURE.  It is the checker for the HIP library:
 * only tests/, the smoke check and the CPU-baseline leg of bench.py load it.
 * No product code path links or calls it.
 *
 * The restatement is written independently of the reference's half-spinor
 * tables: every hop applies the full 4x4 spin matrix (1 +/- gamma_mu) of the
 * DeGrand-Rossi basis to U*psi.  That basis is the one the reference's
 * projector/reconstruct tables encode (codegen/dslash_common.cc:748-890).
 *
 * Field layouts (all double, lexicographic full lattice,
 *   site = ((t*Lz + z)*Ly + y)*Lx + x, parity(site) = (x+y+z+t) & 1):
 *   spinor      [V][4 spin][3 colour][2 reim]
 *   gauge       [V][4 mu][3 row][3 col][2 reim]       U_mu(x), row-major
 *   clover      [V][2 block][36]  = 6 real diagonals, then 15 complex
 *               lower-triangle entries, od(r,c) = r(r-1)/2 + c for r > c
 *               (codegen/include/decl_common.h:15-32,
 *                codegen/dslash_common.cc:1098-1186)
 *   full clover [V][2 block][6 row][6 col][2 reim]
 *               (codegen/include/decl_common.h:56-61,
 *                codegen/dslash_common.cc:1188-1232)
 * Block b of a clover acts on spins 2b, 2b+1 with combined index
 * k = 3*(spin - 2b) + colour.
 *
 * Operators restated (per output site x of checkerboard cb; s = isign):
 *   hop(x) = sum_mu  bb_mu(x) (1 + s g_mu) conj(ph_mu) U_mu(x-mu)^+ psi(x-mu)
 *                  + bf_mu(x) (1 - s g_mu)      ph_mu  U_mu(x)    psi(x+mu)
 *     (codegen/dslash_common.cc:1729-1826, codegen/dslash.cc:200-323;
 *      bb/bf = coeff_s for x,y,z and coeff_t_b/coeff_t_f for t,
 *      codegen/dslash.cc:79-86; the antiperiodic-in-t sign is the caller's
 *      boundary factor on coeff_t_{b,f}; ph_mu = tbc phase when tbc[mu],
 *      codegen/dslash_common.cc:1491-1524)
 *   wilson          out = hop
 *   clover          out = A^-1 hop            (A^-1 passed as "clover")
 *   tm              out = muinv (1 - s i mu g5) hop
 *                                             (codegen/dslash_common.cc:1234-1359)
 *   tm_clover       out = B hop               (B full 6x6 blocks)
 *   wilson achi     out = alpha chi - hop
 *   clover achi     out = A chi + s i rho g5 chi - hop
 *                                             (codegen/dslash_common.cc:1830-1932)
 *   tm achi         out = alpha (1 + s i mu g5) chi - hop
 *                                             (codegen/dslash_common.cc:1364-1489)
 *   tm_clover achi  out = B chi + s i rho g5 chi - hop
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;

enum { KIND_WILSON = 0, KIND_CLOVER = 1, KIND_TM = 2, KIND_TMCLOVER = 3 };

typedef struct {
  int L[4];
  long V;
} lat_t;

static void lat_init(lat_t *g, int Lx, int Ly, int Lz, int Lt)
{
  g->L[0] = Lx;
  g->L[1] = Ly;
  g->L[2] = Lz;
  g->L[3] = Lt;
  g->V = (long)Lx * Ly * Lz * Lt;
}

static void site_coords(const lat_t *g, long s, int c[4])
{
  c[0] = (int)(s % g->L[0]);
  s /= g->L[0];
  c[1] = (int)(s % g->L[1]);
  s /= g->L[1];
  c[2] = (int)(s % g->L[2]);
  c[3] = (int)(s / g->L[2]);
}

static long site_index(const lat_t *g, const int c[4])
{
  return ((long)(c[3] * g->L[2] + c[2]) * g->L[1] + c[1]) * g->L[0] + c[0];
}

static long neighbour(const lat_t *g, long s, int mu, int dir)
{
  int c[4];
  site_coords(g, s, c);
  c[mu] = (c[mu] + dir + g->L[mu]) % g->L[mu];
  return site_index(g, c);
}

/* DeGrand-Rossi gamma matrices; gamma_5 = g_x g_y g_z g_t = diag(1,1,-1,-1). */
static cplx gamma_mat[4][4][4];
static int gamma_ready = 0;

static void init_gamma(void)
{
  if (gamma_ready) return;
  memset(gamma_mat, 0, sizeof(gamma_mat));
  /* g_x */
  gamma_mat[0][0][3] = I;
  gamma_mat[0][1][2] = I;
  gamma_mat[0][2][1] = -I;
  gamma_mat[0][3][0] = -I;
  /* g_y */
  gamma_mat[1][0][3] = -1;
  gamma_mat[1][1][2] = 1;
  gamma_mat[1][2][1] = 1;
  gamma_mat[1][3][0] = -1;
  /* g_z */
  gamma_mat[2][0][2] = I;
  gamma_mat[2][1][3] = -I;
  gamma_mat[2][2][0] = -I;
  gamma_mat[2][3][1] = I;
  /* g_t */
  gamma_mat[3][0][2] = 1;
  gamma_mat[3][1][3] = 1;
  gamma_mat[3][2][0] = 1;
  gamma_mat[3][3][1] = 1;
  gamma_ready = 1;
}

static inline cplx ld(const double *p) { return p[0] + I * p[1]; }
static inline void st(double *p, cplx v)
{
  p[0] = creal(v);
  p[1] = cimag(v);
}

/* U_mu(x) element (i,j) */
static inline cplx link(const double *U, long s, int mu, int i, int j)
{
  return ld(U + ((s * 4 + mu) * 9 + i * 3 + j) * 2);
}

static inline cplx spin(const double *psi, long s, int sp, int c)
{
  return ld(psi + ((s * 4 + sp) * 3 + c) * 2);
}

/* Hopping term at one site s, as in the header comment. */
static void hop_site(cplx acc[4][3], long s, const double *psi, const double *U,
                     const lat_t *g, int isign, double coeff_s,
                     double coeff_t_f, double coeff_t_b, double t_boundary,
                     const int *tbc, const double *tbc_phases)
{
  int c[4];
  site_coords(g, s, c);
  for (int sp = 0; sp < 4; sp++)
    for (int i = 0; i < 3; i++) acc[sp][i] = 0;
  for (int mu = 0; mu < 4; mu++) {
    cplx ph = 1.0;
    if (tbc && tbc[mu]) ph = tbc_phases[2 * mu] + I * tbc_phases[2 * mu + 1];
    for (int fb = 0; fb < 2; fb++) {
      /* fb = 0: back hop from x - mu with U^+, (1 + s g_mu);
         fb = 1: forward hop from x + mu with U, (1 - s g_mu). */
      long n = neighbour(g, s, mu, fb ? +1 : -1);
      double beta;
      if (mu < 3) {
        beta = coeff_s;
      } else if (fb) {
        beta = coeff_t_f * (c[3] == g->L[3] - 1 ? t_boundary : 1.0);
      } else {
        beta = coeff_t_b * (c[3] == 0 ? t_boundary : 1.0);
      }
      cplx w[4][3];
      for (int sp = 0; sp < 4; sp++) {
        for (int i = 0; i < 3; i++) {
          cplx v = 0;
          for (int j = 0; j < 3; j++) {
            if (fb)
              v += link(U, s, mu, i, j) * spin(psi, n, sp, j);
            else
              v += conj(link(U, n, mu, j, i)) * spin(psi, n, sp, j);
          }
          w[sp][i] = v * (fb ? ph : conj(ph));
        }
      }
      double sg = fb ? -(double)isign : (double)isign;
      for (int sp = 0; sp < 4; sp++) {
        for (int i = 0; i < 3; i++) {
          cplx v = w[sp][i];
          for (int k = 0; k < 4; k++) v += sg * gamma_mat[mu][sp][k] * w[k][i];
          acc[sp][i] += beta * v;
        }
      }
    }
  }
}

void oracle_hop(double *out, const double *psi, const double *U, int Lx,
                int Ly, int Lz, int Lt, int isign, int cb, double coeff_s,
                double coeff_t_f, double coeff_t_b, double t_boundary,
                const int *tbc, const double *tbc_phases)
{
  lat_t g;
  lat_init(&g, Lx, Ly, Lz, Lt);
  init_gamma();
  for (long s = 0; s < g.V; s++) {
    int c[4];
    site_coords(&g, s, c);
    if (((c[0] + c[1] + c[2] + c[3]) & 1) != cb) continue;
    cplx acc[4][3];
    hop_site(acc, s, psi, U, &g, isign, coeff_s, coeff_t_f, coeff_t_b,
             t_boundary, tbc, tbc_phases);
    for (int sp = 0; sp < 4; sp++)
      for (int i = 0; i < 3; i++) st(out + ((s * 4 + sp) * 3 + i) * 2, acc[sp][i]);
  }
}

/* Expand a packed Hermitian clover block into a dense 6x6 matrix. */
static void clover_block_dense(const double *blk, cplx A[6][6])
{
  for (int r = 0; r < 6; r++) {
    A[r][r] = blk[r];
    for (int c = 0; c < r; c++) {
      int od = r * (r - 1) / 2 + c;
      cplx v = blk[6 + 2 * od] + I * blk[6 + 2 * od + 1];
      A[r][c] = v;
      A[c][r] = conj(v);
    }
  }
}

static void full_block_dense(const double *blk, cplx A[6][6])
{
  for (int r = 0; r < 6; r++)
    for (int c = 0; c < 6; c++) A[r][c] = ld(blk + (r * 6 + c) * 2);
}

/* r = A v at site s, A packed Hermitian (full = 0) or full complex. */
static void blocks_site(cplx r[4][3], const cplx v[4][3], const double *clov,
                        int full, long s)
{
  int per_site = full ? 144 : 72;
  int per_block = full ? 72 : 36;
  for (int b = 0; b < 2; b++) {
    cplx A[6][6];
    const double *blk = clov + s * per_site + b * per_block;
    if (full)
      full_block_dense(blk, A);
    else
      clover_block_dense(blk, A);
    for (int i = 0; i < 6; i++) {
      cplx acc = 0;
      for (int k = 0; k < 6; k++) acc += A[i][k] * v[2 * b + k / 3][k % 3];
      r[2 * b + i / 3][i % 3] = acc;
    }
  }
}

void oracle_clover_apply(double *out, const double *in, const double *clov,
                         int full, int Lx, int Ly, int Lz, int Lt, int cb)
{
  lat_t g;
  lat_init(&g, Lx, Ly, Lz, Lt);
  for (long s = 0; s < g.V; s++) {
    int c[4];
    site_coords(&g, s, c);
    if (((c[0] + c[1] + c[2] + c[3]) & 1) != cb) continue;
    cplx v[4][3], r[4][3];
    for (int sp = 0; sp < 4; sp++)
      for (int i = 0; i < 3; i++) v[sp][i] = spin(in, s, sp, i);
    blocks_site(r, v, clov, full, s);
    for (int sp = 0; sp < 4; sp++)
      for (int i = 0; i < 3; i++) st(out + ((s * 4 + sp) * 3 + i) * 2, r[sp][i]);
  }
}

/* One output site of any operator family (see the header comment). */
static void dslash_site(cplx res[4][3], long s, const double *psi,
                        const double *chi, const double *U, const double *clov,
                        int kind, int achi, const lat_t *g, int isign,
                        double coeff_s, double coeff_t_f, double coeff_t_b,
                        double t_boundary, const int *tbc,
                        const double *tbc_phases, double alpha, double mu,
                        double muinv, double rho)
{
  cplx hop[4][3];
  hop_site(hop, s, psi, U, g, isign, coeff_s, coeff_t_f, coeff_t_b, t_boundary,
           tbc, tbc_phases);
  double sg = (double)isign;
  if (!achi) {
    if (kind == KIND_WILSON) {
      memcpy(res, hop, sizeof(hop));
    } else if (kind == KIND_CLOVER || kind == KIND_TMCLOVER) {
      blocks_site(res, (const cplx(*)[3])hop, clov, kind == KIND_TMCLOVER, s);
    } else { /* tm: muinv (1 - s i mu g5) hop */
      for (int sp = 0; sp < 4; sp++)
        for (int i = 0; i < 3; i++) {
          double g5 = sp < 2 ? 1.0 : -1.0;
          res[sp][i] = muinv * (1.0 - sg * I * mu * g5) * hop[sp][i];
        }
    }
    return;
  }
  cplx x[4][3];
  for (int sp = 0; sp < 4; sp++)
    for (int i = 0; i < 3; i++) x[sp][i] = spin(chi, s, sp, i);
  if (kind == KIND_WILSON) {
    for (int sp = 0; sp < 4; sp++)
      for (int i = 0; i < 3; i++) res[sp][i] = alpha * x[sp][i];
  } else if (kind == KIND_TM) {
    for (int sp = 0; sp < 4; sp++)
      for (int i = 0; i < 3; i++) {
        double g5 = sp < 2 ? 1.0 : -1.0;
        res[sp][i] = alpha * (1.0 + sg * I * mu * g5) * x[sp][i];
      }
  } else {
    blocks_site(res, (const cplx(*)[3])x, clov, kind == KIND_TMCLOVER, s);
    for (int sp = 0; sp < 4; sp++)
      for (int i = 0; i < 3; i++) {
        double g5 = sp < 2 ? 1.0 : -1.0;
        res[sp][i] += sg * I * rho * g5 * x[sp][i];
      }
  }
  for (int sp = 0; sp < 4; sp++)
    for (int i = 0; i < 3; i++) res[sp][i] -= hop[sp][i];
}

/* Main operator entry: every (kind, achimbdpsi) family of the reference.
   sites == NULL: all sites of parity cb; otherwise only the nsites listed
   (lexicographic indices, which must have parity cb). */
void oracle_dslash(double *out, const double *psi, const double *chi,
                   const double *U, const double *clov, int kind, int achi,
                   int Lx, int Ly, int Lz, int Lt, int isign, int cb,
                   double coeff_s, double coeff_t_f, double coeff_t_b,
                   double t_boundary, const int *tbc, const double *tbc_phases,
                   double alpha, double mu, double muinv, double rho)
{
  lat_t g;
  lat_init(&g, Lx, Ly, Lz, Lt);
  init_gamma();
  for (long s = 0; s < g.V; s++) {
    int c[4];
    site_coords(&g, s, c);
    if (((c[0] + c[1] + c[2] + c[3]) & 1) != cb) continue;
    cplx r[4][3];
    dslash_site(r, s, psi, chi, U, clov, kind, achi, &g, isign, coeff_s,
                coeff_t_f, coeff_t_b, t_boundary, tbc, tbc_phases, alpha, mu,
                muinv, rho);
    for (int sp = 0; sp < 4; sp++)
      for (int i = 0; i < 3; i++) st(out + ((s * 4 + sp) * 3 + i) * 2, r[sp][i]);
  }
}

/* Same operator at the listed lexicographic sites only; out is
   [nsites][4][3][2]. */
void oracle_dslash_sites(double *out, const long *sites, long nsites,
                         const double *psi, const double *chi, const double *U,
                         const double *clov, int kind, int achi, int Lx, int Ly,
                         int Lz, int Lt, int isign, double coeff_s,
                         double coeff_t_f, double coeff_t_b, double t_boundary,
                         const int *tbc, const double *tbc_phases, double alpha,
                         double mu, double muinv, double rho)
{
  lat_t g;
  lat_init(&g, Lx, Ly, Lz, Lt);
  init_gamma();
  for (long k = 0; k < nsites; k++) {
    cplx r[4][3];
    dslash_site(r, sites[k], psi, chi, U, clov, kind, achi, &g, isign, coeff_s,
                coeff_t_f, coeff_t_b, t_boundary, tbc, tbc_phases, alpha, mu,
                muinv, rho);
    for (int sp = 0; sp < 4; sp++)
      for (int i = 0; i < 3; i++) st(out + ((k * 4 + sp) * 3 + i) * 2, r[sp][i]);
  }
}

/* ------------------------------------------------------------------------
 * Clover term from links.  No counterpart in the reference (its kernels
 * take A or A^-1 as input); the convention is fixed here:
 *   A(x) = alpha + (csw/2) sum_{mu<nu} sigma_{mu nu} F_{mu nu}(x),
 *   sigma_{mu nu} = (i/2)[g_mu, g_nu],
 *   F_{mu nu} = traceless part of (Q_{mu nu} - Q_{mu nu}^+)/(8i),
 *   Q_{mu nu} = sum of the four plaquette leaves based at x.
 * sigma_{mu nu} commutes with g5 in this basis, so A splits into the two
 * 6x6 Hermitian blocks of the packed clover layout.
 * ---------------------------------------------------------------------- */
typedef struct {
  cplx m[3][3];
} su3;

static su3 get_link(const double *U, long s, int mu)
{
  su3 r;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) r.m[i][j] = link(U, s, mu, i, j);
  return r;
}

static su3 mul(su3 a, su3 b)
{
  su3 r;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      cplx v = 0;
      for (int k = 0; k < 3; k++) v += a.m[i][k] * b.m[k][j];
      r.m[i][j] = v;
    }
  return r;
}

static su3 dag(su3 a)
{
  su3 r;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) r.m[i][j] = conj(a.m[j][i]);
  return r;
}

static long shift2(const lat_t *g, long s, int mu, int dmu, int nu, int dnu)
{
  long t = s;
  if (dmu) t = neighbour(g, t, mu, dmu);
  if (dnu) t = neighbour(g, t, nu, dnu);
  return t;
}

/* Field strength F_{mu nu}(x), traceless Hermitian. */
static void field_strength(const double *U, const lat_t *g, long s, int mu,
                           int nu, cplx F[3][3])
{
  long xpm = shift2(g, s, mu, 1, nu, 0), xpn = shift2(g, s, mu, 0, nu, 1);
  long xmm = shift2(g, s, mu, -1, nu, 0), xmn = shift2(g, s, mu, 0, nu, -1);
  long xmmpn = shift2(g, s, mu, -1, nu, 1), xmmmn = shift2(g, s, mu, -1, nu, -1);
  long xpmmn = shift2(g, s, mu, 1, nu, -1);
  su3 P1 = mul(mul(get_link(U, s, mu), get_link(U, xpm, nu)),
               mul(dag(get_link(U, xpn, mu)), dag(get_link(U, s, nu))));
  su3 P2 = mul(mul(get_link(U, s, nu), dag(get_link(U, xmmpn, mu))),
               mul(dag(get_link(U, xmm, nu)), get_link(U, xmm, mu)));
  su3 P3 = mul(mul(dag(get_link(U, xmm, mu)), dag(get_link(U, xmmmn, nu))),
               mul(get_link(U, xmmmn, mu), get_link(U, xmn, nu)));
  su3 P4 = mul(mul(dag(get_link(U, xmn, nu)), get_link(U, xmn, mu)),
               mul(get_link(U, xpmmn, nu), dag(get_link(U, s, mu))));
  cplx Q[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      Q[i][j] = P1.m[i][j] + P2.m[i][j] + P3.m[i][j] + P4.m[i][j];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) F[i][j] = (Q[i][j] - conj(Q[j][i])) / (8.0 * I);
  cplx tr = (F[0][0] + F[1][1] + F[2][2]) / 3.0;
  for (int i = 0; i < 3; i++) F[i][i] -= tr;
}

/* A(x) at site s: alpha + (csw/2) sum_{mu<nu} sigma_{mu nu} F_{mu nu}(x),
   sigma_{mu nu} = (i/2)(g_mu g_nu - g_nu g_mu), stored packed at blk. */
static void clover_site(double *blk0, const double *U, const lat_t *g, long s, double alpha,
                        double csw)
{
  cplx sigma[4][4][4][4];
  for (int mu = 0; mu < 4; mu++)
    for (int nu = 0; nu < 4; nu++)
      for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++) {
          cplx v = 0;
          for (int k = 0; k < 4; k++)
            v += gamma_mat[mu][a][k] * gamma_mat[nu][k][b] -
                 gamma_mat[nu][a][k] * gamma_mat[mu][k][b];
          sigma[mu][nu][a][b] = 0.5 * I * v;
        }
  cplx A[2][6][6] = {{{0}}};
  for (int b = 0; b < 2; b++)
    for (int k = 0; k < 6; k++) A[b][k][k] = alpha;
  for (int mu = 0; mu < 4; mu++)
    for (int nu = mu + 1; nu < 4; nu++) {
      cplx F[3][3];
      field_strength(U, g, s, mu, nu, F);
      for (int b = 0; b < 2; b++)
        for (int r = 0; r < 6; r++)
          for (int c = 0; c < 6; c++) {
            int s1 = 2 * b + r / 3, s2 = 2 * b + c / 3;
            A[b][r][c] += 0.5 * csw * sigma[mu][nu][s1][s2] * F[r % 3][c % 3];
          }
    }
  for (int b = 0; b < 2; b++) {
    double *blk = blk0 + b * 36;
    for (int r = 0; r < 6; r++) {
      blk[r] = creal(A[b][r][r]);
      for (int c = 0; c < r; c++) {
        int od = r * (r - 1) / 2 + c;
        blk[6 + 2 * od] = creal(A[b][r][c]);
        blk[6 + 2 * od + 1] = cimag(A[b][r][c]);
      }
    }
  }
}

void oracle_clover_from_links(double *clov, const double *U, int Lx, int Ly,
                              int Lz, int Lt, double alpha, double csw)
{
  lat_t g;
  lat_init(&g, Lx, Ly, Lz, Lt);
  init_gamma();
  for (long s = 0; s < g.V; s++) clover_site(clov + s * 72, U, &g, s, alpha, csw);
}

/* The same at the listed lexicographic sites only: clov[n][72]. */
void oracle_clover_from_links_sites(double *clov, const long *sites, long nsites, const double *U,
                                    int Lx, int Ly, int Lz, int Lt, double alpha, double csw)
{
  lat_t g;
  lat_init(&g, Lx, Ly, Lz, Lt);
  init_gamma();
  for (long k = 0; k < nsites; k++) clover_site(clov + k * 72, U, &g, sites[k], alpha, csw);
}

/* Dense 6x6 complex inverse by Gauss-Jordan with partial pivoting. */
static int invert6(cplx A[6][6], cplx R[6][6])
{
  cplx M[6][12];
  for (int i = 0; i < 6; i++)
    for (int j = 0; j < 6; j++) {
      M[i][j] = A[i][j];
      M[i][6 + j] = (i == j) ? 1.0 : 0.0;
    }
  for (int col = 0; col < 6; col++) {
    int piv = col;
    for (int r = col + 1; r < 6; r++)
      if (cabs(M[r][col]) > cabs(M[piv][col])) piv = r;
    if (cabs(M[piv][col]) == 0.0) return 1;
    if (piv != col)
      for (int j = 0; j < 12; j++) {
        cplx t = M[col][j];
        M[col][j] = M[piv][j];
        M[piv][j] = t;
      }
    cplx inv = 1.0 / M[col][col];
    for (int j = 0; j < 12; j++) M[col][j] *= inv;
    for (int r = 0; r < 6; r++) {
      if (r == col) continue;
      cplx f = M[r][col];
      for (int j = 0; j < 12; j++) M[r][j] -= f * M[col][j];
    }
  }
  for (int i = 0; i < 6; i++)
    for (int j = 0; j < 6; j++) R[i][j] = M[i][6 + j];
  return 0;
}

/* Packed Hermitian blocks -> packed Hermitian inverse blocks. */
int oracle_clover_invert(double *out, const double *clov, long nsites)
{
  for (long s = 0; s < nsites; s++)
    for (int b = 0; b < 2; b++) {
      cplx A[6][6], R[6][6];
      clover_block_dense(clov + s * 72 + b * 36, A);
      if (invert6(A, R)) return 1;
      double *blk = out + s * 72 + b * 36;
      for (int r = 0; r < 6; r++) {
        blk[r] = creal(R[r][r]);
        for (int c = 0; c < r; c++) {
          int od = r * (r - 1) / 2 + c;
          blk[6 + 2 * od] = creal(R[r][c]);
          blk[6 + 2 * od + 1] = cimag(R[r][c]);
        }
      }
    }
  return 0;
}

/* Full complex blocks of A + i*sign*mu*g5 (invert = 0) or its inverse
   (invert = 1), from packed Hermitian A.  g5 = +1 on block 0, -1 on block 1. */
int oracle_tmclover_blocks(double *out, const double *clov, long nsites,
                           double mu, int sign, int invert)
{
  for (long s = 0; s < nsites; s++)
    for (int b = 0; b < 2; b++) {
      cplx A[6][6], R[6][6];
      clover_block_dense(clov + s * 72 + b * 36, A);
      double g5 = b == 0 ? 1.0 : -1.0;
      for (int k = 0; k < 6; k++) A[k][k] += I * (double)sign * mu * g5;
      if (invert) {
        if (invert6(A, R)) return 1;
      } else {
        memcpy(R, A, sizeof(R));
      }
      double *blk = out + s * 144 + b * 72;
      for (int r = 0; r < 6; r++)
        for (int c = 0; c < 6; c++) st(blk + (r * 6 + c) * 2, R[r][c]);
    }
  return 0;
}

/* ------------------------------------------------------------------------
 * Gauge-link gradient of L = Re<chi, hop psi> (chi on cb, psi on 1-cb).
 * No counterpart in the reference.  For every hop of every output site x,
 * the hop term  beta ph_fb chi(x)^+ (1 -/+ s g_mu) V psi(n)  with
 * V = U_mu(x) (forward) or U_mu(x-mu)^+ (backward) contributes to
 * Re tr(U G) of the link it uses; G is accumulated per link here by
 * expanding each term entry by entry.  Output per link [V][4][3][3][2]:
 *   project = 1: TA(U G) (traceless anti-Hermitian, d/de L(e^{ieT}U) =
 *                Re tr(i T F)),  project = 0: G^dagger.
 * ---------------------------------------------------------------------- */
/* G of link (y, mu): the one hop term that uses it.  y on cb: the forward
   hop at x = y (n = y + mu); y on 1-cb: the backward hop at x = y + mu
   (n = y).  Gl[i*3+j] is G[i][j] with L = sum_links Re tr(U G). */
static void force_link(cplx Gl[9], const double *chi, const double *psi, const lat_t *g, long y,
                       int mu, int isign, int cb, double coeff_s, double coeff_t_f,
                       double coeff_t_b, double t_boundary, const int *tbc,
                       const double *tbc_phases)
{
  int cy[4];
  site_coords(g, y, cy);
  const int fb = ((cy[0] + cy[1] + cy[2] + cy[3]) & 1) == cb;
  const long x = fb ? y : neighbour(g, y, mu, +1);
  const long n = fb ? neighbour(g, y, mu, +1) : y;
  int c[4];
  site_coords(g, x, c);
  cplx ph = 1.0;
  if (tbc && tbc[mu]) ph = tbc_phases[2 * mu] + I * tbc_phases[2 * mu + 1];
  double beta;
  if (mu < 3)
    beta = coeff_s;
  else if (fb)
    beta = coeff_t_f * (c[3] == g->L[3] - 1 ? t_boundary : 1.0);
  else
    beta = coeff_t_b * (c[3] == 0 ? t_boundary : 1.0);
  double sg = fb ? -(double)isign : (double)isign;
  /* eta = (1 + sg g_mu)^+ chi(x) = (1 + sg g_mu) chi(x) (g Hermitian) */
  cplx eta[4][3];
  for (int sp = 0; sp < 4; sp++)
    for (int i = 0; i < 3; i++) {
      cplx v = spin(chi, x, sp, i);
      for (int k = 0; k < 4; k++) v += sg * gamma_mat[mu][sp][k] * spin(chi, x, k, i);
      eta[sp][i] = v;
    }
  /* term = beta * sum_s sum_ij conj(eta_si) V_ij psi(n)_sj with
     V = ph U_mu(x) (fb = 1) or conj(ph) U_mu(n)^+ (fb = 0).
     fb = 1: term = tr(U_mu(x) Gx), Gx[j][i] += beta ph psi_j conj(eta_i)
     fb = 0: V_ij = conj(ph) conj(U_mu(n)_ji):
             Re term = Re tr(U_mu(n) Gn), Gn[i][j] += beta ph eta_i conj(psi_j) */
  for (int k = 0; k < 9; k++) Gl[k] = 0;
  for (int sp = 0; sp < 4; sp++)
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        if (fb)
          Gl[j * 3 + i] += beta * ph * spin(psi, n, sp, j) * conj(eta[sp][i]);
        else
          Gl[i * 3 + j] += beta * ph * eta[sp][i] * conj(spin(psi, n, sp, j));
      }
}

/* Output of one link from its G: G^dagger, or TA(U G). */
static void force_out(double *out, const cplx Gl[9], const double *U, long s, int mu, int project)
{
  if (!project) {
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) st(out + (i * 3 + j) * 2, conj(Gl[j * 3 + i]));
    return;
  }
  cplx W[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      cplx v = 0;
      for (int k = 0; k < 3; k++) v += link(U, s, mu, i, k) * Gl[k * 3 + j];
      W[i][j] = v;
    }
  cplx tr = 0;
  for (int i = 0; i < 3; i++) tr += W[i][i] - conj(W[i][i]);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      cplx v = 0.5 * (W[i][j] - conj(W[j][i]));
      if (i == j) v -= tr / 6.0;
      st(out + (i * 3 + j) * 2, v);
    }
}

void oracle_force(double *F, const double *chi, const double *psi, const double *U,
                  int Lx, int Ly, int Lz, int Lt, int isign, int cb, double coeff_s,
                  double coeff_t_f, double coeff_t_b, double t_boundary,
                  const int *tbc, const double *tbc_phases, int project)
{
  lat_t g;
  lat_init(&g, Lx, Ly, Lz, Lt);
  init_gamma();
  for (long l = 0; l < g.V * 4; l++) {
    cplx Gl[9];
    force_link(Gl, chi, psi, &g, l / 4, (int)(l % 4), isign, cb, coeff_s, coeff_t_f, coeff_t_b,
               t_boundary, tbc, tbc_phases);
    force_out(F + l * 18, Gl, U, l / 4, (int)(l % 4), project);
  }
}

/* The same for the listed links only (link = 4 * site + mu): F[n][18]. */
void oracle_force_links(double *F, const long *links, long nlinks, const double *chi,
                        const double *psi, const double *U, int Lx, int Ly, int Lz, int Lt,
                        int isign, int cb, double coeff_s, double coeff_t_f, double coeff_t_b,
                        double t_boundary, const int *tbc, const double *tbc_phases, int project)
{
  lat_t g;
  lat_init(&g, Lx, Ly, Lz, Lt);
  init_gamma();
  for (long k = 0; k < nlinks; k++) {
    const long l = links[k];
    cplx Gl[9];
    force_link(Gl, chi, psi, &g, l / 4, (int)(l % 4), isign, cb, coeff_s, coeff_t_f, coeff_t_b,
               t_boundary, tbc, tbc_phases);
    force_out(F + k * 18, Gl, U, l / 4, (int)(l % 4), project);
  }
}
