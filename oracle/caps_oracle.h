/* CPU oracle for the force-and-step path of CAPS — TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference algorithms, used by tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg as the checker. Nothing in caps_amd/ links, loads or calls it.
 *
 * Each function names the reference file:line it restates and is checked independently (tests/test_oracle_*.py): against
 * the reference build in oracle/_ref (list build and prune bit-exact, the cluster kernel within nbnxm_check.cpp's bars,
 * the Fortran PME, ewald_beta, good sizes, grid choice, the rng.hpp Normal/Gamma/UniformReal streams bit-exact), closed
 * forms (Madelung constant, DFT, the tail integral), finite differences (pair, bonded and reciprocal forces, the virial
 * as the strain derivative, the MLP gradients), invariants (every pair within r_c listed once, momentum, temperature,
 * KG bond lengths and angles) and published known-answer values (mt19937_64's 10000th output).
 * Draw orders: orc_kremer_grest is bit-exact against the reference's builder (kremer_grest.cpp:15-83, built alone by
 * this directory's Makefile; tests/golden/kg.npz); orc_initial_velocities consumes the reference's own normal stream
 * (tests/golden/rng.npz) in the order of dynamics.cpp:231-247 (that function links the whole core and is not built).
 *
 * Conventions: a cell is double[12] = origin(3), a(3), b(3), c(3), fully periodic; positions are 3N doubles (Å);
 * forces 3N doubles (kcal/mol/Å); energies kcal/mol; the virial tensor w[6] = xx yy zz xy xz yz (Σ r_a f_b, symmetrised).
 */
#ifndef CAPS_ORACLE_H
#define CAPS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_PAD 0xffffffffu

/* ---- cluster pair list (core/src/nbnxm.cpp:22-326, 328-390) ---------------------------------------------------- */
typedef struct orc_list orc_list;

/* Builds the half list of 4x4 cluster pairs at radius rlist. excl: CSR (n+1 offsets, sorted partners). p14: CSR of
 * the excluded partners whose Coulomb keeps a fraction, or NULL; with keep_excluded != 0 the entries carry the excluded
 * lanes (ex, ex14) for the kernel's Ewald corrections, as List::build with p14 (nbnxm.cpp:229-253).
 * Returns NULL on error (message in orc_last_error). */
orc_list* orc_list_build(const double* x, int32_t n, const double cell[12], double rlist, const int32_t* excl_off,
                         const int32_t* excl_idx, const int32_t* p14_off, const int32_t* p14_idx, int32_t keep_excluded);
void orc_list_free(orc_list* L);
/* sizes: n atoms, clusters, slots (4 per cluster), entries, image codes, columns nu x nv */
void orc_list_sizes(const orc_list* L, int64_t out[7]);
/* copies of the list arrays; any pointer may be NULL */
void orc_list_get(const orc_list* L, uint32_t* perm, uint32_t* iperm, int64_t* first, uint32_t* cj, uint16_t* code,
                  uint16_t* mask, uint16_t* ex, uint16_t* ex14, double* org, double* off, int32_t* ioff, int32_t* qint,
                  double* qvec, float* bx, float* by, float* bz);
/* List::positions (nbnxm.cpp:378-390): float slot coordinates relative to the cluster origins */
void orc_list_positions(const orc_list* L, const double* x, float* X, float* Y, float* Z);
/* List::prune (nbnxm.cpp:328-363) against the slot coordinates X, Y, Z: the pruned masks per outer entry
 * (mask, ex, ex14 each ANDed with the lanes within `radius`) */
void orc_list_prune(const orc_list* L, const float* X, const float* Y, const float* Z, double radius, uint16_t* pmask,
                    uint16_t* pex, uint16_t* pex14);

/* ---- plain pair terms by a direct FP64 sum over atom pairs and images (the pin of the cluster kernel) ----------
 * Follows the double atom-list semantics the reference's nbnxm_check.cpp:78-119 compares the kernel against, with the
 * pair forms of field.cpp:1378-1487 (12-6 / 9-6 LJ) and 1700-1790 (Ewald real space / DSF):
 *   lj: 1 = 12-6, 2 = 9-6; c12, c6: nt x nt pair coefficients (12-6: 4εσ¹², 4εσ⁶; 9-6: 2εσ⁹, 3εσ⁶)
 *   coul: 0 none, 1 Ewald real space erfc(κr)/r, 2 damped shifted force (dsf_e0, dsf_f0)
 *   shift: LJ shifted to zero at rc; f14: Coulomb fraction a 1-4 pair keeps (excluded pairs' corrections included
 *   when with_excl != 0, at their nearest image, within rc)
 * out e[2] = {vdw, coulomb}; f (3N, added to); w (6, added to). */
int orc_pairs_direct(const double* x, int32_t n, const double cell[12], const int32_t* type, int32_t nt, const double* c12,
                     const double* c6, int32_t lj, const double* q, double kcoul, int32_t coul, double kappa, double dsf_e0,
                     double dsf_f0, int32_t shift, double rc, const int32_t* excl_off, const int32_t* excl_idx,
                     const int32_t* p14_off, const int32_t* p14_idx, double f14, int32_t with_excl, double* e, double* f,
                     double* w);

/* ---- bonded terms (core/src/field.cpp:1980-2003 harmonic bond / angle, 2150-2180 FENE + WCA) --------------------
 * bonds: ij (2 per), kr (k, r0 per); angles: ijk (3 per), kt (k_θ, θ0 in radians); fene: ij, p (K, R0, ε, σ per).
 * e[3] = {bond, angle, virial}; f added to; w added to. Returns -1 when a FENE bond is stretched to R0 or past. */
int orc_bonded(const double* x, int32_t n, const double cell[12], int32_t nb, const int32_t* bij, const double* bkr,
               int32_t na, const int32_t* aijk, const double* akt, int32_t nf, const int32_t* fij, const double* fp,
               double* e, double* f, double* w);

/* LJ tail correction (field.cpp:2501-2537), 12-6 (lj = 1) or 9-6 (lj = 2): eps/sig6 per type pair (nt x nt), counts
 * per type; out[2] = {energy, virial} */
void orc_tail(int32_t nt, const double* eps, const double* sig6, const double* count, int32_t lj, double rc, double volume,
              double out[2]);

/* ---- k-space (core/src/kspace.cpp:35-52; core/fortran/caps_kspace.f90) ------------------------------------------ */
double orc_ewald_beta(double cutoff, double rtol);                       /* kspace.cpp:35-43 */
int32_t orc_good_size(int32_t n);                                        /* caps_kspace.f90:23-36 */
void orc_pme_grid(const double cell[12], double spacing, int32_t order, int32_t k[3]);   /* kspace.cpp:45-52 */
/* B-spline weights of order p at offset w (caps_kspace.f90:568-598) */
void orc_bspline(double w, int32_t p, double* theta, double* dtheta);
/* |b(m)|² per axis (caps_kspace.f90:601-620) */
void orc_bspline_moduli(int32_t k, int32_t p, double* bsp);
/* 1-D complex DFT, X(m) = Σ x(k) exp(sign 2πi mk/n), out of place, any n (mixed radix) */
void orc_fft1d(int32_t n, int32_t sign, const double* in_re, const double* in_im, double* out_re, double* out_im);
/* in-place 3-D transform of data[i + k1 (j + k2 k)] */
void orc_fft3d(int32_t k1, int32_t k2, int32_t k3, int32_t sign, double* re, double* im);
/* SPME reciprocal part (caps_kspace.f90:625-715 caps_f_pme): energy returned, forces added to f, vir[6] */
double orc_pme(const double* x, const double* q, int32_t n, const double cell[12], const int32_t k[3], int32_t order,
               double beta, double kcoul, double* f, double* vir);
/* plain Ewald reciprocal sum over |m_d| <= kmax (caps_kspace.f90:718-764) */
double orc_ewald_recip(const double* x, const double* q, int32_t n, const double cell[12], const int32_t kmax[3],
                       double beta, double kcoul, double* f, double* vir);
/* the influence function with its virial weights at grid index (a, b, c) (caps_kspace.f90:518-564) */
void orc_pme_green(int32_t a, int32_t b, int32_t c, const int32_t k[3], const double cell[12], int32_t order, double beta,
                   double kcoul, double* g, double w[6]);

/* ---- random numbers (core/include/caps/rng.hpp) ------------------------------------------------------------------ */
typedef struct {
  uint64_t mt[312];
  int32_t mti;
  int32_t have_spare;
  double spare;
} orc_rng;
void orc_rng_seed(orc_rng* r, uint64_t seed);
uint64_t orc_rng_next(orc_rng* r);                 /* std::mt19937_64 */
double orc_rng_uniform01(orc_rng* r);              /* rng.hpp:26-33 */
double orc_rng_normal(orc_rng* r);                 /* rng.hpp:70-89, Marsaglia polar, spare kept */
/* rng.hpp:120-147, Marsaglia–Tsang; the distribution owns its own normal generator (spare) as caps::Gamma does */
double orc_rng_gamma(orc_rng* r, double shape, double scale, int32_t* g_have, double* g_spare);

/* ---- dynamics (core/src/dynamics.cpp) ---------------------------------------------------------------------------- */
double orc_kinetic_energy(const double* v, const double* m, int32_t n);    /* dynamics.cpp:65-69 */
/* initial Maxwell–Boltzmann velocities without centre-of-mass motion, rescaled to T (dynamics.cpp:236-256) */
void orc_initial_velocities(orc_rng* r, const double* m, int32_t n, double temperature, double* v);
/* Bussi's new kinetic energy factor α from K, the draws r1 and Σ R_i² (dynamics.cpp:466-477) */
double orc_bussi_alpha(double k, double kt_target, double ndof, double c, double r1, double sum);
/* Kremer–Grest chains (kremer_grest.cpp:15-83): x (3·chains·beads), returns the box length */
double orc_kremer_grest(int32_t chains, int32_t beads, double density, double k_theta, double bond, uint64_t seed, double* x);

/* ---- learned r-RESPA (no reference counterpart): FP64 restatement of the MLP and its descriptors ---------------- */
/* 96-128-128-3 SiLU MLP: w1 [128][96], b1[128], w2 [128][128], b2[128], w3 [3][128], b3[3]; x [n][96] -> y [n][3] */
void orc_mlp_forward(const double* w1, const double* b1, const double* w2, const double* b2, const double* w3,
                     const double* b3, const double* x, int32_t n, double* y);
/* gradients of L = ½ Σ (y − t)² / n over the batch; g* accumulate (set to zero first) */
double orc_mlp_backward(const double* w1, const double* b1, const double* w2, const double* b2, const double* w3,
                        const double* b3, const double* x, const double* t, int32_t n, double* gw1, double* gb1,
                        double* gw2, double* gb2, double* gw3, double* gb3);

/* the 96 descriptors per atom (24 hat bins on 0 … 8 Å × {q, q d̂_x, q d̂_y, q d̂_z}, cosine cut-off): out [n][96] */
void orc_descriptors(const double* x, const double* q, int32_t n, const double cell[12], double* out);

const char* orc_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
