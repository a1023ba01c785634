/*
 * librint-amd: C ABI of the MI355X (gfx950) two-electron path.
 *
 * Part 1 mirrors the reference's C boundary symbol for symbol (reference
 * src/p2c.rs), so a caller that binds librint.so through ctypes/cgo/FFI loads
 * this library unchanged.  Conventions kept from the reference:
 *   - atm/bas/env are libcint arrays given as (pointer, element count); they
 *     are copied on entry (reference src/p2c.rs:28-46);
 *   - returned double* buffers are owned by the library and released with
 *     free_c(ptr, exact_len) (reference src/p2c.rs:48-75);
 *   - failures return NULL (arrays) or NaN (scalars) with the reason on
 *     stderr (reference src/p2c.rs:166-174, 222-229).  Invalid coord/typec
 *     codes also return NULL instead of exiting the process (the reference
 *     calls exit(1), src/scf.rs:25-28).
 *
 * Part 2 adds what a GPU path needs: a persistent context (basis staged in
 * HBM, Schwarz bounds and quartet bins computed once), the fused Fock build
 * on host or device buffers, batched shell-quartet blocks and counters.
 */
#ifndef LIBRINT_AMD_H
#define LIBRINT_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------ reference boundary ---- */

/* reference src/p2c.rs:67-75 */
void free_c(double *ptr, size_t len);

/* n*n one-electron matrix, R[mu*n+nu]; coord 0 cartesian (n = nao), 1 real
 * spherical (n = sum of 2l+1, transform of src/cart2sph.rs:88380-88530);
 * typec 0 overlap, 1 kinetic, 2 nuclear.  Computed on the GPU, l <= 4.
 * reference src/p2c.rs:85-100, src/scf.rs:160-213 */
double *int1e_c(int *atm, size_t atm_l, int *bas, size_t bas_l, double *env, size_t env_l, int coord, int typec);

/* n^4 ERI tensor R[((i*n+j)*n+k)*n+l] = (ij|kl), computed on the GPU (l <= 4);
 * coord 0 cartesian (n = nao), 1 real spherical.  reference src/p2c.rs:108-123,
 * src/scf.rs:215-311 */
double *int2e_c(int *atm, size_t atm_l, int *bas, size_t bas_l, double *env, size_t env_l, int coord);

/* Converged RHF density (nao*nao) or NULL; reference SCF protocol (P0 = I,
 * Roothaan, |dP|_F/2 <= conv).  reference src/p2c.rs:152-177, src/scf.rs:743-799 */
double *density_c(int *atm, size_t atm_l, int *bas, size_t bas_l, double *env, size_t env_l, size_t nelec,
                  int imax, double conv);

/* E[P] = tr(P(T+V)) + 1/2 tr(P G[P]) + E_nuc.  reference src/p2c.rs:185-202, src/scf.rs:827-956 */
double energy_c(int *atm, size_t atm_l, int *bas, size_t bas_l, double *env, size_t env_l, double *P, size_t P_l);

/* SCF energy or NaN.  reference src/p2c.rs:209-230, src/scf.rs:958-969 */
double scf_c(int *atm, size_t atm_l, int *bas, size_t bas_l, double *env, size_t env_l, size_t nelec, int imax,
             double conv);

/* Basis-parameter gradients of the reference, each the derivative with
 * respect to env[s1..s2) (s1, s2 from the split rule of reference
 * src/utils.rs:152-175: every primitive exponent and contraction
 * coefficient), P the density they are evaluated at; release with free_c.
 * Both the two-electron part (adjoint Rys kernels) and the one-electron
 * derivatives (S, T, V) run on the GPU; shells with l <= 4, the reference's
 * domain (reference python/librint/dscf.py:11; NULL and an error message
 * above it).
 *   dS_u          nao*nao*(s2-s1), [(a*nao+b)*L + k] = dS_ab/denv[s1+k]
 *                 reference src/p2c.rs:130-144, src/dscf.rs:96-150
 *   grad_c        dHcore + dR            reference src/p2c.rs:240-258, src/dscf.rs:547-557
 *   dS_c          sum_ab (PFP)_ab dS_ab  reference src/p2c.rs:267-285, src/dscf.rs:354-368
 *   dHcore_c      sum_ab P_ab d(T+V)_ab  reference src/p2c.rs:294-312, src/dscf.rs:320-336
 *   dR_c          d/denv 1/2 tr(P G[P])  reference src/p2c.rs:321-339, src/dscf.rs:370-524
 *   danalytical_c dHcore + dR - dS/2     reference src/p2c.rs:348-366, src/dscf.rs:526-537
 *   denergy_c     same as danalytical_c  reference src/p2c.rs:376-394, src/dscf.rs:559-565 */
double *dS_u(int *atm, size_t atm_l, int *bas, size_t bas_l, double *env, size_t env_l);
double *grad_c(int *atm, size_t atm_l, int *bas, size_t bas_l, double *env, size_t env_l, double *P, size_t P_l);
double *dS_c(int *atm, size_t atm_l, int *bas, size_t bas_l, double *env, size_t env_l, double *P, size_t P_l);
double *dHcore_c(int *atm, size_t atm_l, int *bas, size_t bas_l, double *env, size_t env_l, double *P, size_t P_l);
double *dR_c(int *atm, size_t atm_l, int *bas, size_t bas_l, double *env, size_t env_l, double *P, size_t P_l);
double *danalytical_c(int *atm, size_t atm_l, int *bas, size_t bas_l, double *env, size_t env_l, double *P,
                      size_t P_l);
double *denergy_c(int *atm, size_t atm_l, int *bas, size_t bas_l, double *env, size_t env_l, double *P, size_t P_l);

/* --------------------------------------------------------- extensions ---- */

/* Stateless direct two-electron Fock matrix G = J - K/2 (n*n, release
 * with free_c), the GPU counterpart of integral2e_fock (reference
 * src/scf.rs:316-410).  coord 0: cartesian, P and G nao*nao; coord 1: real
 * spherical, P and G nsph*nsph (the cartesian build between the spherical
 * transforms, G_sph = C G_cart[C^T P C] C^T).  schwarz_tol: quartets with
 * Q_ij*Q_kl <= tol are skipped (0 keeps every quartet).  precision: 0 = FP64. */
double *fock_c(int *atm, size_t atm_l, int *bas, size_t bas_l, double *env, size_t env_l, double *P, size_t P_l,
               int coord, int precision, double schwarz_tol);

typedef struct librint_ctx librint_ctx;

/* Stage a basis on the current GPU; NULL on failure. */
librint_ctx *librint_ctx_create(const int *atm, size_t atm_l, const int *bas, size_t bas_l, const double *env,
                                size_t env_l);
void librint_ctx_destroy(librint_ctx *ctx);
int librint_ctx_nao(const librint_ctx *ctx);
/* Set the Schwarz threshold (default 1e-14); returns 0 on success. */
int librint_ctx_set_schwarz(librint_ctx *ctx, double tol);
/* Restrict the context's Fock builds to one device's share of the work when
 * nshards devices split it (no counterpart in the reference, which runs on one
 * host; for a multi-device caller with one context per device).  The groups
 * of quartet tasks of every class launch are dealt heaviest first, each to the
 * share with the least primitive-quartet cost so far (ties to the lowest
 * index: every device computes the same deal), and this context keeps share
 * `shard`.  librint_fock* then return that share's G; the G's of the nshards
 * shares sum to the whole build (one all-reduce).  nshards = 1 (default): the
 * whole build.  The SCF and gradient calls refuse a context with nshards > 1.
 * Returns 0 on success, -1 for shard outside [0, nshards). */
int librint_ctx_set_shard(librint_ctx *ctx, int nshards, int shard);

/* Precision of the Fock build.  mode 0: FP64 everywhere (default).
 * mode 1: mixed -- every build bounds each 64-ket task by the largest
 * Q_ab*Q_cd*max(|D| over the six density blocks a quartet reads) of its kets,
 * from that build's density; tasks below tau_mix run their Rys recurrences in
 * FP32 (the Rys table too; J/K accumulation stays FP64).  tau_skip > 0
 * additionally drops, per build, the tasks below tau_skip (density
 * screening; also honoured in mode 0).  dmax is ignored (kept for ABI
 * compatibility).  Returns 0 on success. */
int librint_ctx_set_precision(librint_ctx *ctx, int mode, double tau_mix, double tau_skip, double dmax);
/* Quartets and primitive quartets of the last build by task flag:
 * out[0..2] quartets FP64/FP32/skipped, out[3..5] primitive quartets. */
int librint_ctx_flag_counts(librint_ctx *ctx, double *out6);

/* G = J - K/2 from host buffers (nao*nao each).  Returns 0 on success. */
int librint_fock(librint_ctx *ctx, const double *P, double *G);
/* Same on device buffers, enqueued on `stream` (a hipStream_t, NULL = default).
 * Asynchronous: returns after enqueueing.  Returns 0 on success. */
int librint_fock_device(librint_ctx *ctx, const double *P_dev, double *G_dev, void *stream);

/* librint_fock_device without any device allocation (no counterpart in the
 * reference; for callers that cannot allocate: code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The build's
 * scratch (the nao*nao accumulator, the shell-block density maxima when
 * density screening or mixed precision is on, the work counters and the
 * block scratch of f/g classes) lives in a static 64 MiB device buffer of the
 * library, one per GPU, held until the process ends; the caller passes P and
 * G as for librint_fock_device.  One stream (`stream`) and no Coulomb slabs:
 * the ket-pair Coulomb blocks are added with atomics.  Returns -1 with an
 * error message when the build needs more scratch than the buffer holds
 * (about nao <= 2800 for an s/p/d basis without density screening).  Every
 * call on a GPU shares that GPU's buffer, so calls that overlap on one GPU
 * (several streams or threads) race; calls in sequence on one stream are
 * safe.  The context must be set up (librint_ctx_create and its set_*
 * calls) beforehand: those allocate. */
int librint_fock_device_noalloc(librint_ctx *ctx, const double *P_dev, double *G_dev, void *stream);

/* Blocks (ij|kl) of nq shell quartets (4 ints each, indices into bas), each
 * in cint2e_cart layout (i fastest), concatenated into out.  Returns the
 * number of doubles written, or (size_t)-1 on failure.  With out == NULL only
 * the size is returned. */
size_t librint_int2e_shells(librint_ctx *ctx, const int *quartets, int nq, double *out);

/* Schwarz bound sqrt(max|(ij|ij)|) per shell pair i >= j, index i*(i+1)/2+j
 * (nbas*(nbas+1)/2 doubles). Returns 0 on success. */
int librint_schwarz(librint_ctx *ctx, double *out);

/* Counters of the current quartet plan, out[0..10]:
 * quartets, primitive quartets, contracted integrals, algorithmic flops,
 * candidate quartets, kernel launches, device ms of the last build, Schwarz tol,
 * device ms of the last adjoint (gradient) kernel launches, densities the last
 * SCF on the context took from the purified projector instead of eigh,
 * streams the last build's class launches ran on.
 * Returns the number of values written. */
int librint_ctx_stats(const librint_ctx *ctx, double *out, int n);

/* Per angular-class launch of the current plan, 14 values each:
 * la, lb, lc, ld, quartets, primitive quartets, algorithmic flops, lanes
 * launched (64 x tasks), device ms of the launch in the last host-side build
 * (librint_fock), task count, mixed (1: the launch picks FP32 recurrences per
 * task in mixed precision), kernel of the last build (0 jk_kernel,
 * 1 jk_group_kernel, 2 jk_team_kernel, 3 hl_kernel for classes with an f or
 * g shell, -1 none yet), bra density rows staged in LDS (0/1), device ms of
 * the Coulomb-slab reductions inside the launch's time (launches whose slabs
 * run in chunks reduce each chunk right after it).
 * Returns the number of launches (out may be NULL to query it); writes at
 * most n launches. */
int librint_ctx_class_stats(const librint_ctx *ctx, double *out, int n);

/* RHF energy gradient at a converged density P (host, nao*nao):
 *   gnuc[3*natm] = dE/dR  (core H incl. the point-charge positions, the
 *                  two-electron adjoint, -1/2 tr(PFP dS), nuclear repulsion)
 *   denv[env_l]  = dE/d(env entry) for every primitive exponent and
 *                  contraction coefficient (dHcore + dR - dS/2)
 * parts (optional): [gH | gR | gS | gN] (3*natm each) then [dH | dR | dS]
 * (env_l each).  Quartets are screened with the context's Schwarz threshold.
 * Shells with l <= 4 (returns -1 otherwise).
 * No counterpart in the reference (its gradients are basis-parameter only). */
int librint_ctx_gradient(librint_ctx *ctx, const double *P, double *gnuc, double *denv, double *parts);

/* Nuclear-coordinate RHF gradient (3*natm, release with free_c) at a
 * converged density; stateless form of librint_ctx_gradient without
 * screening.  No counterpart in the reference. */
double *grad_nuc_c(int *atm, size_t atm_l, int *bas, size_t bas_l, double *env, size_t env_l, double *P, size_t P_l);

/* RHF on the context.  guess: 0 identity (reference), 1 core Hamiltonian,
 * 2 caller density in P_inout.  diis: 0/1.  On return P_inout (nao*nao,
 * may be NULL) holds the final density and *iters the cycle count.
 * Returns the energy, NaN on failure.  The reference mode (guess 0, no
 * DIIS) diagonalises F in every cycle as the reference does; with DIIS and
 * guess 1 or 2 each density is the occupied-space projector of F by
 * trace-correcting purification (matrix products only), falling back to
 * the eigensolver where that does not converge to a projector of trace
 * nelec/2 (LIBRINT_SCF_EIGH=1: always the eigensolver). */
double librint_scf(librint_ctx *ctx, int nelec, int imax, double conv, int guess, int diis, double *P_inout,
                   int *iters);

/* librint_scf with an energy criterion as well: converged when
 * |P - P_old|_F / 2 <= conv and (conv_e <= 0 or |E_k - E_(k-1)| <= conv_e),
 * E_k the energy of the density entering iteration k.  No counterpart in the
 * reference. */
double librint_scf_ex(librint_ctx *ctx, int nelec, int imax, double conv, double conv_e, int guess, int diis,
                      double *P_inout, int *iters);

/* ------------------------------------------------ learned SCF guess ---- */
/* Kernels of the density-guess MLP (BF16 weights/activations as raw uint16
 * patterns, FP32 accumulation).  All array arguments are device pointers;
 * stream NULL = the default stream.  Return 0, or -1 with librint_last_error.
 * No counterpart in the reference (its SCF starts from P = I).
 *   gemm:     C = act(op(A) op(B) + bias [+ C]); op(A)(m,k) = transA ? A[k*lda+m]
 *             : A[m*lda+k], op(B)(k,n) = transB ? B[n*ldb+k] : B[k*ldb+n];
 *             act 0 none, 1 ReLU; C (FP32) and/or Cb (BF16) may be NULL (not both)
 *   relu_bwd: dpre = dz * (h > 0) as FP32 and/or BF16
 *   colsum:   out[c] (+)= sum_r x[r*ld + c]
 *   mse:      dy = 2 mask (y - t) / scale on the first cols columns (0 beyond),
 *             *loss += sum mask (y - t)^2 / scale
 *   adam:     Adam step (bias-corrected, step >= 1) on FP32 w, refreshing wb
 *   to_bf16:  y = bf16(x), round to nearest even */
int librint_mlp_gemm(int M, int N, int K, const uint16_t *A, int lda, int transA, const uint16_t *B, int ldb,
                     int transB, float *C, int ldc, const float *bias, int act, int accumulate, uint16_t *Cb, int ldcb,
                     void *stream);
int librint_mlp_relu_bwd(size_t n, const float *dz, const float *h, float *dpre, uint16_t *dpre_b, void *stream);
int librint_mlp_colsum(int rows, int cols, const float *x, int ld, float *out, int accumulate, void *stream);
int librint_mlp_mse(int rows, int cols, const float *y, int ldy, const float *t, const float *mask, int ldt,
                    float scale, float *dy, uint16_t *dy_b, double *loss, void *stream);
int librint_mlp_adam(size_t n, float *w, const float *g, float *m, float *v, float lr, float b1, float b2, float eps,
                     int step, uint16_t *wb, void *stream);
int librint_mlp_to_bf16(size_t n, const float *x, uint16_t *y, void *stream);

/* Last error message of this thread ("" if none). */
const char *librint_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
