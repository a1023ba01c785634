/*
 * ORACLE -- test infrastructure only.  Nothing in the product links or calls
 * this file; tests, the smoke check and the CPU-baseline leg of the benchmark
 * load it (as oracle/_build/librefcpu.so) to check or time the HIP path.
 *
 * A plain-C restatement of the reference's two-electron path and the RHF
 * driver around it:
 *
 *   rc_boys            Boys function           reference src/fmt.rs:63-105
 *   rc_rys_roots       Rys nodes/weights       reference src/rys_roots.rs:3036-3064 (output convention)
 *   rc_int2e_shell     contracted (ij|kl)      reference src/eri.rs:98-399,409-692
 *   rc_int2e           nao^4 tensor            reference src/scf.rs:215-311
 *   rc_fock2e          G = J - K/2 direct      reference src/scf.rs:316-410
 *   rc_calc_F          F from stored tensor    reference src/scf.rs:651-670
 *   rc_int1e           S, T, V                 reference src/scf.rs:160-213, src/g1e.rs, src/cint1e.rs
 *   rc_energyfast      E[P]                    reference src/scf.rs:827-956
 *   rc_density/rc_scf  Roothaan SCF from P=I   reference src/scf.rs:572-799, 958-969
 *   rc_schwarz, rc_fock2e_screened  the same G from canonical Schwarz-screened
 *                      quartets, 8-fold symmetry, OpenMP (a faster CPU comparator)
 *
 * Numerical choices of this restatement (it does not reuse any table of the
 * reference):
 *   - Boys: power series below a turnover, erf + upward recursion above,
 *     in long double.
 *   - Rys rule: discretised Stieltjes procedure on a 56-point Gauss-Legendre
 *     discretisation of exp(-x t^2) dt, then a Jacobi eigen-solve of the
 *     n x n Jacobi matrix, all in long double (Gautschi, "Orthogonal
 *     Polynomials: Computation and Approximation", 2004, sec. 2.2.3).
 *   - Eigen-solver of the SCF: cyclic Jacobi rotations in double.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define ATM_SLOTS 6
#define BAS_SLOTS 8
#define LMAX 4
#define NFMAX 15
#define NRMAX 10
#define GSIZE_MAX ((LMAX + 1) * (LMAX + 1) * (2 * LMAX + 1) * (2 * LMAX + 1))

typedef long double ld;

/* ------------------------------------------------------------------ Boys -- */
/* F_m(x) = int_0^1 t^{2m} exp(-x t^2) dt for m = 0..mmax.
 * Reference: gamma_inc_like, src/fmt.rs:63-105 (series below a turnover point,
 * erf with upward recursion above). */
void rc_boys(int mmax, double x, double *out)
{
    ld f[64];
    ld X = x;
    if (x < 1e-300 && x >= 0) {
        for (int m = 0; m <= mmax; m++) out[m] = 1.0 / (2 * m + 1);
        return;
    }
    if (x < 40.0 + mmax) {
        /* F_M by series  e^{-x} sum_k (2x)^k / ((2M+1)(2M+3)...(2M+2k+1)),
         * then downward recursion F_m = (2x F_{m+1} + e^{-x}) / (2m+1). */
        ld e = expl(-X);
        ld term = 1.0L / (2 * mmax + 1), sum = term;
        for (int k = 1; k < 2000; k++) {
            term *= 2 * X / (2 * mmax + 2 * k + 1);
            sum += term;
            if (term < sum * 1e-22L) break;
        }
        f[mmax] = e * sum;
        for (int m = mmax - 1; m >= 0; m--) f[m] = (2 * X * f[m + 1] + e) / (2 * m + 1);
    } else {
        ld e = expl(-X);
        ld sx = sqrtl(X);
        f[0] = 0.886226925452758013649083741671L * erfl(sx) / sx; /* sqrt(pi)/2 */
        for (int m = 0; m < mmax; m++) f[m + 1] = ((2 * m + 1) * f[m] - e) / (2 * X);
    }
    for (int m = 0; m <= mmax; m++) out[m] = (double)f[m];
}

/* ------------------------------------------------------------ Rys rule -- */
#define NGL 56
static ld gl_x[NGL], gl_w[NGL];
static int gl_ready = 0;

static void gl_init(void)
{
    if (gl_ready) return;
    const ld PI = 3.141592653589793238462643383279502884L;
    for (int i = 0; i < NGL; i++) {
        ld z = cosl(PI * (i + 0.75L) / (NGL + 0.5L)), dp = 0;
        for (int it = 0; it < 100; it++) {
            ld p0 = 1, p1 = z;
            for (int k = 2; k <= NGL; k++) {
                ld p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = NGL * (z * p1 - p0) / (z * z - 1);
            ld dz = p1 / dp;
            z -= dz;
            if (fabsl(dz) < 1e-21L) break;
        }
        {
            ld p0 = 1, p1 = z;
            for (int k = 2; k <= NGL; k++) {
                ld p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = NGL * (z * p1 - p0) / (z * z - 1);
        }
        gl_x[i] = z;
        gl_w[i] = 2 / ((1 - z * z) * dp * dp);
    }
    gl_ready = 1;
}

/* symmetric eigen-decomposition of a small matrix by cyclic Jacobi (long double) */
static void jacobi_ld(int n, ld *A, ld *V)
{
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) V[i * n + j] = (i == j);
    for (int sweep = 0; sweep < 100; sweep++) {
        ld off = 0, dia = 0;
        for (int i = 0; i < n; i++) {
            dia += A[i * n + i] * A[i * n + i];
            for (int j = i + 1; j < n; j++) off += A[i * n + j] * A[i * n + j];
        }
        if (off <= 1e-40L * dia) break;
        for (int p = 0; p < n; p++)
            for (int q = p + 1; q < n; q++) {
                ld apq = A[p * n + q];
                if (fabsl(apq) < 1e-300L) continue;
                ld theta = (A[q * n + q] - A[p * n + p]) / (2 * apq);
                ld t = (theta >= 0 ? 1 : -1) / (fabsl(theta) + sqrtl(theta * theta + 1));
                ld c = 1 / sqrtl(t * t + 1), s = t * c;
                for (int k = 0; k < n; k++) {
                    ld akp = A[k * n + p], akq = A[k * n + q];
                    A[k * n + p] = c * akp - s * akq;
                    A[k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; k++) {
                    ld apk = A[p * n + k], aqk = A[q * n + k];
                    A[p * n + k] = c * apk - s * aqk;
                    A[q * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; k++) {
                    ld vkp = V[k * n + p], vkq = V[k * n + q];
                    V[k * n + p] = c * vkp - s * vkq;
                    V[k * n + q] = s * vkp + c * vkq;
                }
            }
    }
}

/* n-point Rys rule for exp(-x t^2) on t in [0,1]: nodes v = t^2, returned as
 * u = v/(1-v) with plain weights w, sum_r w_r = F_0(x).  The u convention is
 * the one of CINTrys_roots (reference src/rys_roots.rs:3036-3064). */
void rc_rys_roots_v(int n, double x, double *v_out, double *w_out)
{
    gl_init();
    ld X = x;
    ld T = 1.0L;
    if (X > 90.0L) T = sqrtl(90.0L / X); /* exp(-x t^2) < e^-90 beyond T */
    ld v[NGL], W[NGL], p0[NGL], p1[NGL];
    for (int i = 0; i < NGL; i++) {
        ld t = 0.5L * T * (gl_x[i] + 1);
        v[i] = t * t;
        W[i] = 0.5L * T * gl_w[i] * expl(-X * t * t);
    }
    ld alpha[NRMAX], beta[NRMAX];
    ld nrm_prev = 0, nrm = 0;
    for (int i = 0; i < NGL; i++) {
        p0[i] = 0;
        p1[i] = 1;
        nrm += W[i];
    }
    beta[0] = nrm;
    for (int k = 0; k < n; k++) {
        ld a = 0;
        for (int i = 0; i < NGL; i++) a += W[i] * v[i] * p1[i] * p1[i];
        alpha[k] = a / nrm;
        if (k + 1 == n) break;
        nrm_prev = nrm;
        nrm = 0;
        for (int i = 0; i < NGL; i++) {
            ld p2 = (v[i] - alpha[k]) * p1[i] - (k > 0 ? beta[k] : 0) * p0[i];
            p0[i] = p1[i];
            p1[i] = p2;
            nrm += W[i] * p2 * p2;
        }
        beta[k + 1] = nrm / nrm_prev;
        /* rescale p_k and p_{k+1} by one common factor: the recurrence is
         * linear and alpha/beta are ratios, so only the magnitudes change */
        ld s = 1 / sqrtl(nrm);
        for (int i = 0; i < NGL; i++) {
            p0[i] *= s;
            p1[i] *= s;
        }
        nrm = 1;
    }
    ld J[NRMAX * NRMAX], V[NRMAX * NRMAX];
    memset(J, 0, sizeof(J));
    for (int i = 0; i < n; i++) {
        J[i * n + i] = alpha[i];
        if (i + 1 < n) J[i * n + i + 1] = J[(i + 1) * n + i] = sqrtl(beta[i + 1]);
    }
    jacobi_ld(n, J, V);
    /* sort ascending */
    int idx[NRMAX];
    for (int i = 0; i < n; i++) idx[i] = i;
    for (int i = 0; i < n; i++)
        for (int j = i + 1; j < n; j++)
            if (J[idx[j] * n + idx[j]] < J[idx[i] * n + idx[i]]) {
                int t = idx[i];
                idx[i] = idx[j];
                idx[j] = t;
            }
    for (int r = 0; r < n; r++) {
        int c = idx[r];
        v_out[r] = (double)J[c * n + c];
        w_out[r] = (double)(beta[0] * V[0 * n + c] * V[0 * n + c]);
    }
}

void rc_rys_roots(int n, double x, double *u, double *w)
{
    double v[NRMAX];
    rc_rys_roots_v(n, x, v, w);
    for (int r = 0; r < n; r++) u[r] = v[r] / (1 - v[r]);
}

/* ------------------------------------------------------------ basis -- */
static double common_fac_sp(int l)
{ /* reference src/g1e.rs:776-782 */
    if (l == 0) return 0.282094791773878143;
    if (l == 1) return 0.488602511902919921;
    return 1.0;
}

static void cart_comp(int l, int *nx, int *ny, int *nz)
{ /* x-major descending, reference src/cint_bas.rs:234-255 */
    int n = 0;
    for (int lx = l; lx >= 0; lx--)
        for (int ly = l - lx; ly >= 0; ly--) {
            nx[n] = lx;
            ny[n] = ly;
            nz[n] = l - lx - ly;
            n++;
        }
}

static int ncart(int l) { return (l + 1) * (l + 2) / 2; }

int rc_nao(const int *bas, int nbas, int coord)
{
    int n = 0;
    for (int s = 0; s < nbas; s++) {
        int l = bas[BAS_SLOTS * s + 1], nctr = bas[BAS_SLOTS * s + 3];
        n += (coord == 0 ? ncart(l) : 2 * l + 1) * nctr;
    }
    return n;
}

static void ao_offsets(const int *bas, int nbas, int *off)
{
    off[0] = 0;
    for (int s = 0; s < nbas; s++)
        off[s + 1] = off[s] + ncart(bas[BAS_SLOTS * s + 1]) * bas[BAS_SLOTS * s + 3];
}

/* ----------------------------------------------------------- 2e ERI -- */
typedef struct {
    int l[4], np[4], nctr[4], pa[4], pc[4];
    double r[4][3];
    int nf[4], nmax, mmax, nroots, dk, dl, dj;
    double common_factor, expcutoff, rirj[3], rkrl[3], rr_ij, rr_kl, log_rr_ij, log_rr_kl;
} quartet_t;

static void quartet_ctx(quartet_t *q, const int *shls, const int *atm, const int *bas, const double *env)
{ /* reference src/eri.rs:98-227 */
    for (int a = 0; a < 4; a++) {
        const int *b = bas + BAS_SLOTS * shls[a];
        q->l[a] = b[1];
        q->np[a] = b[2];
        q->nctr[a] = b[3];
        q->pa[a] = b[5];
        q->pc[a] = b[6];
        int pr = atm[ATM_SLOTS * b[0] + 1];
        for (int d = 0; d < 3; d++) q->r[a][d] = env[pr + d];
        q->nf[a] = ncart(b[1]);
    }
    q->nmax = q->l[0] + q->l[1];
    q->mmax = q->l[2] + q->l[3];
    q->nroots = (q->nmax + q->mmax) / 2 + 1;
    q->dk = q->l[0] + 1;
    q->dl = q->dk * (q->l[2] + 1);
    q->dj = q->dl * (q->mmax + 1);
    q->common_factor = M_PI * M_PI * M_PI * 2.0 / 1.772453850905516 * common_fac_sp(q->l[0]) *
                       common_fac_sp(q->l[1]) * common_fac_sp(q->l[2]) * common_fac_sp(q->l[3]);
    q->expcutoff = env[0] == 0.0 ? 60.0 : (env[0] > 40.0 ? env[0] : 40.0) + 1.0;
    q->rr_ij = q->rr_kl = 0;
    for (int d = 0; d < 3; d++) {
        q->rirj[d] = q->r[0][d] - q->r[1][d];
        q->rkrl[d] = q->r[2][d] - q->r[3][d];
        q->rr_ij += q->rirj[d] * q->rirj[d];
        q->rr_kl += q->rkrl[d] * q->rkrl[d];
    }
    double aij_last = env[q->pa[0] + q->np[0] - 1] + env[q->pa[1] + q->np[1] - 1];
    q->log_rr_ij = 1.7 - 1.5 * log(aij_last);
    if (q->nmax > 0) q->log_rr_ij += q->nmax * log(sqrt(q->rr_ij) + 1.0);
    double akl_last = env[q->pa[2] + q->np[2] - 1] + env[q->pa[3] + q->np[3] - 1];
    q->log_rr_kl = 1.7 - 1.5 * log(akl_last);
    if (q->mmax > 0) q->log_rr_kl += q->mmax * log(sqrt(q->rr_kl) + 1.0);
}

/* One Rys root of a primitive quartet: 2D VRR building momentum on j and l,
 * then HRR j->i, l->k.  reference src/eri.rs:263-399 */
static void prim_g_root(double *gx, double *gy, double *gz, const quartet_t *q, double aij, double akl,
                        double a0, double fac1, double u_r, double w_r, const double *rij,
                        const double *rkl, const double *rijrkl)
{
    const int nmax = q->nmax, mmax = q->mmax, dk = q->dk, dl = q->dl, dj = q->dj;
    const double a1 = aij * akl;
    const double u2 = a0 * u_r;
    const double tmp4 = 0.5 / (u2 * (aij + akl) + a1);
    const double tmp5 = u2 * tmp4;
    const double tmp1 = 2.0 * tmp5;
    const double tmp2 = tmp1 * akl, tmp3 = tmp1 * aij;
    const double b00 = tmp5, b10 = tmp5 + tmp4 * akl, b01 = tmp5 + tmp4 * aij;
    double c00[3], c0p[3];
    for (int d = 0; d < 3; d++) {
        c00[d] = rij[d] - q->r[1][d] - tmp2 * rijrkl[d];
        c0p[d] = rkl[d] - q->r[3][d] + tmp3 * rijrkl[d];
    }
    gx[0] = 1.0;
    gy[0] = 1.0;
    gz[0] = w_r * fac1;
    double *gs[3] = {gx, gy, gz};
    for (int ax = 0; ax < 3; ax++) {
        double *g = gs[ax];
        double cn = c00[ax], cm = c0p[ax];
        if (nmax > 0) {
            double s0 = g[0], s1 = cn * s0;
            g[dj] = s1;
            for (int n = 1; n < nmax; n++) {
                double s2 = cn * s1 + n * b10 * s0;
                g[(n + 1) * dj] = s2;
                s0 = s1;
                s1 = s2;
            }
        }
        if (mmax > 0) {
            double s0 = g[0], s1 = cm * s0;
            g[dl] = s1;
            for (int m = 1; m < mmax; m++) {
                double s2 = cm * s1 + m * b01 * s0;
                g[(m + 1) * dl] = s2;
                s0 = s1;
                s1 = s2;
            }
            if (nmax > 0) {
                s0 = g[dj];
                s1 = cm * s0 + b00 * g[0];
                g[dj + dl] = s1;
                for (int m = 1; m < mmax; m++) {
                    double s2 = cm * s1 + m * b01 * s0 + b00 * g[m * dl];
                    g[dj + (m + 1) * dl] = s2;
                    s0 = s1;
                    s1 = s2;
                }
            }
        }
        if (nmax > 1) {
            for (int m = 1; m <= mmax; m++) {
                int base = m * dl;
                double s0 = g[base], s1 = g[base + dj];
                for (int n = 1; n < nmax; n++) {
                    double s2 = cn * s1 + n * b10 * s0 + m * b00 * g[base + n * dj - dl];
                    g[base + (n + 1) * dj] = s2;
                    s0 = s1;
                    s1 = s2;
                }
            }
        }
    }
    if (q->l[0] > 0 || q->l[2] > 0) {
        for (int ax = 0; ax < 3; ax++) {
            double *g = gs[ax];
            double rx = -q->rirj[ax];
            for (int i = 1; i <= q->l[0]; i++)
                for (int j = 0; j <= nmax - i; j++)
                    for (int l = 0; l <= mmax; l++) {
                        int p = j * dj + l * dl + i;
                        g[p] = rx * g[p - 1] + g[p - 1 + dj];
                    }
            rx = -q->rkrl[ax];
            for (int j = 0; j <= q->l[1]; j++)
                for (int k = 1; k <= q->l[2]; k++)
                    for (int l = 0; l <= mmax - k; l++) {
                        int p = j * dj + l * dl + k * dk;
                        for (int n = p; n < p + dk; n++) g[n] = rx * g[n - dk] + g[n - dk + dl];
                    }
        }
    }
}

/* Contracted shell quartet for one (ci,cj,ck,cl) contraction combination;
 * accumulates into out[fi + nfi*(fj + nfj*(fk + nfk*fl))].
 * reference src/eri.rs:460-561 (segmented) and 571-692 (general) */
static void eri_one_ctr(double *out, const quartet_t *q, const double *env, const int *ctr)
{
    static __thread double gx[GSIZE_MAX], gy[GSIZE_MAX], gz[GSIZE_MAX];
    int inx[NFMAX], iny[NFMAX], inz[NFMAX], jnx[NFMAX], jny[NFMAX], jnz[NFMAX];
    int knx[NFMAX], kny[NFMAX], knz[NFMAX], lnx[NFMAX], lny[NFMAX], lnz[NFMAX];
    cart_comp(q->l[0], inx, iny, inz);
    cart_comp(q->l[1], jnx, jny, jnz);
    cart_comp(q->l[2], knx, kny, knz);
    cart_comp(q->l[3], lnx, lny, lnz);
    const int nfi = q->nf[0], nfj = q->nf[1], nfk = q->nf[2], nfl = q->nf[3];
    const int dk = q->dk, dl = q->dl, dj = q->dj;
    /* log of the largest |coefficient| of a primitive over its contractions */
    double lmc[4][32];
    for (int a = 0; a < 4; a++)
        for (int p = 0; p < q->np[a]; p++) {
            double m = 0;
            for (int c = 0; c < q->nctr[a]; c++) {
                double v = fabs(env[q->pc[a] + c * q->np[a] + p]);
                if (v > m) m = v;
            }
            lmc[a][p] = log(m);
        }
    const double *ci = env + q->pc[0] + ctr[0] * q->np[0];
    const double *cj = env + q->pc[1] + ctr[1] * q->np[1];
    const double *ck = env + q->pc[2] + ctr[2] * q->np[2];
    const double *cl = env + q->pc[3] + ctr[3] * q->np[3];
    for (int lp = 0; lp < q->np[3]; lp++) {
        double al = env[q->pa[3] + lp];
        for (int kp = 0; kp < q->np[2]; kp++) {
            double ak = env[q->pa[2] + kp];
            double akl = ak + al;
            double ekl_val = q->rr_kl * ak * al / akl;
            double ccekl = ekl_val - q->log_rr_kl - lmc[2][kp] - lmc[3][lp];
            if (ccekl > q->expcutoff) continue;
            double rkl[3];
            for (int d = 0; d < 3; d++) rkl[d] = (ak * q->r[2][d] + al * q->r[3][d]) / akl;
            double eijcutoff = q->expcutoff - ccekl;
            double ekl = exp(-ekl_val);
            for (int jp = 0; jp < q->np[1]; jp++) {
                double aj = env[q->pa[1] + jp];
                for (int ip = 0; ip < q->np[0]; ip++) {
                    double ai = env[q->pa[0] + ip];
                    double aij = ai + aj;
                    double eij_val = q->rr_ij * ai * aj / aij;
                    double cceij = eij_val - q->log_rr_ij - lmc[0][ip] - lmc[1][jp];
                    if (cceij > eijcutoff) continue;
                    double wj = aj / aij;
                    double rij[3];
                    for (int d = 0; d < 3; d++) rij[d] = q->r[0][d] + wj * (q->r[1][d] - q->r[0][d]);
                    double fac = q->common_factor * cl[lp] * ck[kp] * cj[jp] * ci[ip] * exp(-eij_val) * ekl;
                    double a1 = aij * akl, a0 = a1 / (aij + akl);
                    double fac1 = sqrt(a0 / (a1 * a1 * a1)) * fac;
                    double rijrkl[3], rr = 0;
                    for (int d = 0; d < 3; d++) {
                        rijrkl[d] = rij[d] - rkl[d];
                        rr += rijrkl[d] * rijrkl[d];
                    }
                    double u[NRMAX], w[NRMAX];
                    rc_rys_roots(q->nroots, a0 * rr, u, w);
                    for (int r = 0; r < q->nroots; r++) {
                        prim_g_root(gx, gy, gz, q, aij, akl, a0, fac1, u[r], w[r], rij, rkl, rijrkl);
                        for (int jf = 0; jf < nfj; jf++)
                            for (int lf = 0; lf < nfl; lf++)
                                for (int kf = 0; kf < nfk; kf++)
                                    for (int f = 0; f < nfi; f++) {
                                        int ix = jnx[jf] * dj + lnx[lf] * dl + knx[kf] * dk + inx[f];
                                        int iy = jny[jf] * dj + lny[lf] * dl + kny[kf] * dk + iny[f];
                                        int iz = jnz[jf] * dj + lnz[lf] * dl + knz[kf] * dk + inz[f];
                                        out[f + nfi * (jf + nfj * (kf + nfk * lf))] += gx[ix] * gy[iy] * gz[iz];
                                    }
                    }
                }
            }
        }
    }
}

/* (ij|kl) block in cint2e_cart's layout out[I + ni*(J + nj*(K + nk*L))] with
 * I = fi + nfi*ci (reference src/cint2e.rs:3128-3152 contract; out is zeroed). */
int rc_int2e_shell(double *out, const int *shls, const int *atm, int natm, const int *bas, int nbas,
                   const double *env)
{
    (void)natm;
    (void)nbas;
    quartet_t q;
    quartet_ctx(&q, shls, atm, bas, env);
    if (q.l[0] > LMAX || q.l[1] > LMAX || q.l[2] > LMAX || q.l[3] > LMAX) return -1;
    const int nfi = q.nf[0], nfj = q.nf[1], nfk = q.nf[2], nfl = q.nf[3];
    const int ni = nfi * q.nctr[0], nj = nfj * q.nctr[1], nk = nfk * q.nctr[2], nl = nfl * q.nctr[3];
    memset(out, 0, sizeof(double) * ni * nj * nk * nl);
    int nf = nfi * nfj * nfk * nfl;
    double *tmp = (double *)malloc(sizeof(double) * nf);
    int ctr[4];
    for (ctr[3] = 0; ctr[3] < q.nctr[3]; ctr[3]++)
        for (ctr[2] = 0; ctr[2] < q.nctr[2]; ctr[2]++)
            for (ctr[1] = 0; ctr[1] < q.nctr[1]; ctr[1]++)
                for (ctr[0] = 0; ctr[0] < q.nctr[0]; ctr[0]++) {
                    memset(tmp, 0, sizeof(double) * nf);
                    eri_one_ctr(tmp, &q, env, ctr);
                    for (int lf = 0; lf < nfl; lf++)
                        for (int kf = 0; kf < nfk; kf++)
                            for (int jf = 0; jf < nfj; jf++)
                                for (int f = 0; f < nfi; f++) {
                                    int I = f + nfi * ctr[0], J = jf + nfj * ctr[1];
                                    int K = kf + nfk * ctr[2], L = lf + nfl * ctr[3];
                                    out[I + ni * (J + nj * (K + nk * L))] =
                                        tmp[f + nfi * (jf + nfj * (kf + nfk * lf))];
                                }
                }
    free(tmp);
    return 1;
}

static void quartet_dims(const int *bas, const int *shls, int *d)
{
    for (int a = 0; a < 4; a++) d[a] = ncart(bas[BAS_SLOTS * shls[a] + 1]) * bas[BAS_SLOTS * shls[a] + 3];
}

/* full tensor R[((mu*n+nu)*n+sig)*n+lam] = (mu nu|sig lam); reference src/scf.rs:215-311 */
void rc_int2e(double *R, const int *atm, int natm, const int *bas, int nbas, const double *env)
{
    int *off = (int *)malloc(sizeof(int) * (nbas + 1));
    ao_offsets(bas, nbas, off);
    size_t n = off[nbas];
    double *buf = (double *)malloc(sizeof(double) * 50625);
    int shls[4], d[4];
    for (shls[0] = 0; shls[0] < nbas; shls[0]++)
        for (shls[1] = 0; shls[1] < nbas; shls[1]++)
            for (shls[2] = 0; shls[2] < nbas; shls[2]++)
                for (shls[3] = 0; shls[3] < nbas; shls[3]++) {
                    rc_int2e_shell(buf, shls, atm, natm, bas, nbas, env);
                    quartet_dims(bas, shls, d);
                    int c = 0;
                    for (int l = 0; l < d[3]; l++)
                        for (int k = 0; k < d[2]; k++)
                            for (int j = 0; j < d[1]; j++)
                                for (int i = 0; i < d[0]; i++) {
                                    size_t mu = off[shls[0]] + i, nu = off[shls[1]] + j;
                                    size_t sg = off[shls[2]] + k, la = off[shls[3]] + l;
                                    R[((mu * n + nu) * n + sg) * n + la] = buf[c++];
                                }
                }
    free(buf);
    free(off);
}

/* Direct G = J - K/2 over all nbas^4 quartets; reference src/scf.rs:316-410:
 *   G[mu,nu] += P[lam,sig] v ;  G[mu,lam] -= 0.5 P[nu,sig] v  for v = (mu nu|sig lam) */
void rc_fock2e(double *G, const double *P, const int *atm, int natm, const int *bas, int nbas,
               const double *env)
{
    int *off = (int *)malloc(sizeof(int) * (nbas + 1));
    ao_offsets(bas, nbas, off);
    size_t n = off[nbas];
    memset(G, 0, sizeof(double) * n * n);
    double *buf = (double *)malloc(sizeof(double) * 50625);
    int shls[4], d[4];
    for (shls[0] = 0; shls[0] < nbas; shls[0]++)
        for (shls[1] = 0; shls[1] < nbas; shls[1]++)
            for (shls[2] = 0; shls[2] < nbas; shls[2]++)
                for (shls[3] = 0; shls[3] < nbas; shls[3]++) {
                    rc_int2e_shell(buf, shls, atm, natm, bas, nbas, env);
                    quartet_dims(bas, shls, d);
                    int c = 0;
                    for (int l = 0; l < d[3]; l++)
                        for (int k = 0; k < d[2]; k++)
                            for (int j = 0; j < d[1]; j++)
                                for (int i = 0; i < d[0]; i++) {
                                    size_t mu = off[shls[0]] + i, nu = off[shls[1]] + j;
                                    size_t sg = off[shls[2]] + k, la = off[shls[3]] + l;
                                    double v = buf[c++];
                                    G[mu * n + nu] += P[la * n + sg] * v;
                                    G[mu * n + la] += -0.5 * P[nu * n + sg] * v;
                                }
                }
    free(buf);
    free(off);
}

/* F = H + G from the stored tensor; reference src/scf.rs:651-670 */
void rc_calc_F(int n, const double *P, const double *two, const double *H, double *F)
{
    size_t N = n;
    for (size_t mu = 0; mu < N; mu++)
        for (size_t nu = 0; nu < N; nu++) {
            double g = 0;
            for (size_t la = 0; la < N; la++)
                for (size_t sg = 0; sg < N; sg++)
                    g += P[la * N + sg] * (two[((mu * N + nu) * N + sg) * N + la] -
                                           0.5 * two[((mu * N + la) * N + sg) * N + nu]);
            F[mu * N + nu] = g + H[mu * N + nu];
        }
}

/* ----------------------------------------------------------- 1e ints -- */
/* 1D overlap table s[i][j] for powers i on A, j on B (Obara-Saika). */
static void ovlp1d(int imax, int jmax, double PA, double PB, double p, double s00, double s[8][10])
{
    double h = 0.5 / p;
    for (int i = 0; i <= imax; i++)
        for (int j = 0; j <= jmax; j++) {
            double v;
            if (i == 0 && j == 0)
                v = s00;
            else if (i > 0) {
                v = PA * s[i - 1][j];
                if (i > 1) v += (i - 1) * h * s[i - 2][j];
                if (j > 0) v += j * h * s[i - 1][j - 1];
            } else {
                v = PB * s[i][j - 1];
                if (j > 1) v += (j - 1) * h * s[i][j - 2];
            }
            s[i][j] = v;
        }
}

/* S (typ 0), T (typ 1) or V (typ 2) block for a shell pair, one contraction
 * combination, out[fi + nfi*fj] (i fastest, as cint1e_*_cart). */
static void int1e_one_ctr(double *out, int typ, int ish, int jsh, int ci, int cj, const int *atm, int natm,
                          const int *bas, const double *env)
{
    const int *bi = bas + BAS_SLOTS * ish, *bj = bas + BAS_SLOTS * jsh;
    int li = bi[1], lj = bj[1];
    int nfi = ncart(li), nfj = ncart(lj);
    int inx[NFMAX], iny[NFMAX], inz[NFMAX], jnx[NFMAX], jny[NFMAX], jnz[NFMAX];
    cart_comp(li, inx, iny, inz);
    cart_comp(lj, jnx, jny, jnz);
    const double *A = env + atm[ATM_SLOTS * bi[0] + 1], *B = env + atm[ATM_SLOTS * bj[0] + 1];
    double AB2 = 0, AB[3];
    for (int d = 0; d < 3; d++) {
        AB[d] = A[d] - B[d];
        AB2 += AB[d] * AB[d];
    }
    double cf = common_fac_sp(li) * common_fac_sp(lj);
    memset(out, 0, sizeof(double) * nfi * nfj);
    for (int jp = 0; jp < bj[2]; jp++)
        for (int ip = 0; ip < bi[2]; ip++) {
            double a = env[bi[5] + ip], b = env[bj[5] + jp];
            double ca = env[bi[6] + ci * bi[2] + ip], cb = env[bj[6] + cj * bj[2] + jp];
            double p = a + b, mu = a * b / p;
            double P[3];
            for (int d = 0; d < 3; d++) P[d] = (a * A[d] + b * B[d]) / p;
            double pref = ca * cb * cf * exp(-mu * AB2);
            if (typ <= 1) {
                double s[3][8][10];
                for (int d = 0; d < 3; d++)
                    ovlp1d(li, lj + 2, P[d] - A[d], P[d] - B[d], p, sqrt(M_PI / p), s[d]);
                for (int jf = 0; jf < nfj; jf++)
                    for (int f = 0; f < nfi; f++) {
                        int ix[3] = {inx[f], iny[f], inz[f]}, jx[3] = {jnx[jf], jny[jf], jnz[jf]};
                        double val;
                        if (typ == 0) {
                            val = s[0][ix[0]][jx[0]] * s[1][ix[1]][jx[1]] * s[2][ix[2]][jx[2]];
                        } else {
                            double t[3], sv[3];
                            for (int d = 0; d < 3; d++) {
                                int i = ix[d], j = jx[d];
                                sv[d] = s[d][i][j];
                                double tt = -2.0 * b * (2 * j + 1) * s[d][i][j] + 4.0 * b * b * s[d][i][j + 2];
                                if (j >= 2) tt += j * (j - 1) * s[d][i][j - 2];
                                t[d] = -0.5 * tt;
                            }
                            val = t[0] * sv[1] * sv[2] + sv[0] * t[1] * sv[2] + sv[0] * sv[1] * t[2];
                        }
                        out[f + nfi * jf] += pref * val;
                    }
            } else {
                int nr = (li + lj) / 2 + 1;
                for (int c = 0; c < natm; c++) {
                    const double *C = env + atm[ATM_SLOTS * c + 1];
                    double Z = fabs((double)atm[ATM_SLOTS * c]);
                    double PC2 = 0;
                    for (int d = 0; d < 3; d++) PC2 += (P[d] - C[d]) * (P[d] - C[d]);
                    double v[NRMAX], w[NRMAX];
                    rc_rys_roots_v(nr, p * PC2, v, w);
                    double fac = -Z * 2.0 * M_PI / p * pref;
                    for (int r = 0; r < nr; r++) {
                        /* 1D: I(i,0) by VRR, then HRR onto B */
                        double I[3][8][8];
                        for (int d = 0; d < 3; d++) {
                            double c00 = (P[d] - A[d]) - v[r] * (P[d] - C[d]);
                            double b10 = (1 - v[r]) * 0.5 / p;
                            double g[8];
                            g[0] = 1;
                            if (li + lj > 0) g[1] = c00;
                            for (int i = 1; i < li + lj; i++) g[i + 1] = c00 * g[i] + i * b10 * g[i - 1];
                            for (int i = 0; i <= li + lj; i++) I[d][i][0] = g[i];
                            for (int j = 1; j <= lj; j++)
                                for (int i = 0; i <= li + lj - j; i++)
                                    I[d][i][j] = I[d][i + 1][j - 1] + AB[d] * I[d][i][j - 1];
                        }
                        for (int jf = 0; jf < nfj; jf++)
                            for (int f = 0; f < nfi; f++)
                                out[f + nfi * jf] += fac * w[r] * I[0][inx[f]][jnx[jf]] *
                                                     I[1][iny[f]][jny[jf]] * I[2][inz[f]][jnz[jf]];
                    }
                }
            }
        }
}

/* nao x nao one-electron matrix, R[mu*n+nu]; typ 0 overlap, 1 kinetic, 2 nuclear.
 * reference src/scf.rs:160-213 (integral1e), cartesian only. */
void rc_int1e(double *R, int typ, const int *atm, int natm, const int *bas, int nbas, const double *env)
{
    int *off = (int *)malloc(sizeof(int) * (nbas + 1));
    ao_offsets(bas, nbas, off);
    int n = off[nbas];
    double buf[NFMAX * NFMAX];
    for (int i = 0; i < nbas; i++)
        for (int j = 0; j < nbas; j++) {
            int nfi = ncart(bas[BAS_SLOTS * i + 1]), nfj = ncart(bas[BAS_SLOTS * j + 1]);
            for (int ci = 0; ci < bas[BAS_SLOTS * i + 3]; ci++)
                for (int cj = 0; cj < bas[BAS_SLOTS * j + 3]; cj++) {
                    int1e_one_ctr(buf, typ, i, j, ci, cj, atm, natm, bas, env);
                    for (int jf = 0; jf < nfj; jf++)
                        for (int f = 0; f < nfi; f++)
                            R[(size_t)(off[i] + f + nfi * ci) * n + off[j] + jf + nfj * cj] = buf[f + nfi * jf];
                }
        }
    free(off);
}

/* --------------------------------------------------------------- SCF -- */
static void jacobi_eigh(int n, const double *Ain, double *eig, double *U)
{ /* A = U diag(eig) U^T, eig ascending, columns of U (row-major U[i*n+k]) */
    double *A = (double *)malloc(sizeof(double) * n * n);
    memcpy(A, Ain, sizeof(double) * n * n);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) U[i * n + j] = (i == j);
    for (int sweep = 0; sweep < 200; sweep++) {
        double off = 0, tot = 0;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                tot += A[i * n + j] * A[i * n + j];
                if (i != j) off += A[i * n + j] * A[i * n + j];
            }
        if (off <= 1e-30 * tot) break;
        for (int p = 0; p < n; p++)
            for (int q = p + 1; q < n; q++) {
                double apq = A[p * n + q];
                if (apq == 0) continue;
                double theta = (A[q * n + q] - A[p * n + p]) / (2 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1));
                double c = 1 / sqrt(t * t + 1), s = t * c;
                for (int k = 0; k < n; k++) {
                    double akp = A[k * n + p], akq = A[k * n + q];
                    A[k * n + p] = c * akp - s * akq;
                    A[k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; k++) {
                    double apk = A[p * n + k], aqk = A[q * n + k];
                    A[p * n + k] = c * apk - s * aqk;
                    A[q * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; k++) {
                    double ukp = U[k * n + p], ukq = U[k * n + q];
                    U[k * n + p] = c * ukp - s * ukq;
                    U[k * n + q] = s * ukp + c * ukq;
                }
            }
    }
    for (int i = 0; i < n; i++) eig[i] = A[i * n + i];
    /* sort ascending with columns */
    for (int i = 0; i < n; i++) {
        int m = i;
        for (int j = i + 1; j < n; j++)
            if (eig[j] < eig[m]) m = j;
        if (m != i) {
            double t = eig[i];
            eig[i] = eig[m];
            eig[m] = t;
            for (int k = 0; k < n; k++) {
                t = U[k * n + i];
                U[k * n + i] = U[k * n + m];
                U[k * n + m] = t;
            }
        }
    }
    free(A);
}

static void matmul(int n, const double *A, const double *B, double *C)
{
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            double s = 0;
            for (int k = 0; k < n; k++) s += A[i * n + k] * B[k * n + j];
            C[i * n + j] = s;
        }
}

static double nuclear_repulsion(const int *atm, int natm, const double *env)
{ /* reference src/scf.rs:946-953 */
    double E = 0;
    for (int i = 0; i < natm; i++)
        for (int j = 0; j < i; j++) {
            const double *a = env + atm[ATM_SLOTS * i + 1], *b = env + atm[ATM_SLOTS * j + 1];
            double r = sqrt((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) + (a[2] - b[2]) * (a[2] - b[2]));
            E += (double)(atm[ATM_SLOTS * i] * atm[ATM_SLOTS * j]) / r;
        }
    return E;
}

/* E[P] = sum P (T+V) + 1/2 sum (P P - 1/2 P P)(mu nu|sig lam) + E_nuc, evaluated
 * as tr(P H) + 1/2 tr(P G) with G from rc_fock2e; reference src/scf.rs:827-956 */
double rc_energyfast(const double *P, const int *atm, int natm, const int *bas, int nbas, const double *env)
{
    int n = rc_nao(bas, nbas, 0);
    double *T = (double *)malloc(sizeof(double) * n * n), *V = (double *)malloc(sizeof(double) * n * n);
    double *G = (double *)malloc(sizeof(double) * n * n);
    rc_int1e(T, 1, atm, natm, bas, nbas, env);
    rc_int1e(V, 2, atm, natm, bas, nbas, env);
    rc_fock2e(G, P, atm, natm, bas, nbas, env);
    double E = 0;
    for (int i = 0; i < n * n; i++) E += P[i] * (T[i] + V[i]) + 0.5 * P[i] * G[i];
    free(T);
    free(V);
    free(G);
    return E + nuclear_repulsion(atm, natm, env);
}

/* Roothaan SCF from P = I; reference src/scf.rs:743-799.  Uses the direct
 * Fock build when store_tensor == 0 (same G).  Returns 0 on success and fills
 * P; nonzero on failure (1 no convergence, 2 invalid density, 3 singular S). */
int rc_density(double *P, int *iters, const int *atm, int natm, const int *bas, int nbas, const double *env,
               int nelec, int imax, double conv, int store_tensor)
{
    int n = rc_nao(bas, nbas, 0);
    size_t nn = (size_t)n * n;
    double *S = malloc(sizeof(double) * nn), *H = malloc(sizeof(double) * nn), *V = malloc(sizeof(double) * nn);
    double *X = malloc(sizeof(double) * nn), *U = malloc(sizeof(double) * nn), *eig = malloc(sizeof(double) * n);
    double *F = malloc(sizeof(double) * nn), *Fp = malloc(sizeof(double) * nn), *tmp = malloc(sizeof(double) * nn);
    double *C = malloc(sizeof(double) * nn), *Pold = malloc(sizeof(double) * nn);
    double *two = NULL;
    int rc = 0;
    rc_int1e(S, 0, atm, natm, bas, nbas, env);
    rc_int1e(H, 1, atm, natm, bas, nbas, env);
    rc_int1e(V, 2, atm, natm, bas, nbas, env);
    for (size_t i = 0; i < nn; i++) H[i] += V[i];
    if (store_tensor) {
        two = malloc(sizeof(double) * nn * nn);
        rc_int2e(two, atm, natm, bas, nbas, env);
    }
    jacobi_eigh(n, S, eig, U);
    if (eig[0] < 1e-8) {
        rc = 3;
        goto done;
    }
    for (int i = 0; i < n; i++)
        for (int k = 0; k < n; k++) X[i * n + k] = U[i * n + k] / sqrt(eig[k]);
    for (size_t i = 0; i < nn; i++) P[i] = 0;
    for (int i = 0; i < n; i++) P[i * n + i] = 1.0;
    double delta = 1.0;
    int it = 0;
    while (delta > conv && it < imax) {
        memcpy(Pold, P, sizeof(double) * nn);
        if (store_tensor)
            rc_calc_F(n, P, two, H, F);
        else {
            rc_fock2e(F, P, atm, natm, bas, nbas, env);
            for (size_t i = 0; i < nn; i++) F[i] += H[i];
        }
        /* F' = X^T F X */
        double *Xt = U; /* reuse */
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) Xt[i * n + j] = X[j * n + i];
        matmul(n, Xt, F, tmp);
        matmul(n, tmp, X, Fp);
        jacobi_eigh(n, Fp, eig, tmp);
        matmul(n, X, tmp, C);
        for (int mu = 0; mu < n; mu++)
            for (int nu = 0; nu < n; nu++) {
                double s = 0;
                for (int i = 0; i < nelec / 2; i++) s += 2.0 * C[mu * n + i] * C[nu * n + i];
                P[mu * n + nu] = s;
            }
        delta = 0;
        for (size_t i = 0; i < nn; i++) delta += (P[i] - Pold[i]) * (P[i] - Pold[i]);
        delta = sqrt(delta) / 2.0;
        it++;
    }
    if (iters) *iters = it;
    if (delta > conv) {
        rc = 1;
        goto done;
    }
    { /* reference src/scf.rs:572-600 */
        double tr = 0;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) tr += P[i * n + j] * S[j * n + i];
        if (fabs(tr - nelec) > 1e-6) {
            rc = 2;
            goto done;
        }
        matmul(n, P, S, tmp);
        matmul(n, tmp, P, F);
        double idem = 0;
        for (size_t i = 0; i < nn; i++) idem = fmax(idem, fabs(F[i] - 2.0 * P[i]));
        if (idem > 1e-6) rc = 2;
    }
done:
    free(S); free(H); free(V); free(X); free(U); free(eig); free(F); free(Fp); free(tmp); free(C); free(Pold);
    free(two);
    return rc;
}

double rc_scf(const int *atm, int natm, const int *bas, int nbas, const double *env, int nelec, int imax,
              double conv, int store_tensor)
{
    int n = rc_nao(bas, nbas, 0);
    double *P = malloc(sizeof(double) * n * n);
    int it;
    int rc = rc_density(P, &it, atm, natm, bas, nbas, env, nelec, imax, conv, store_tensor);
    double E = rc ? NAN : rc_energyfast(P, atm, natm, bas, nbas, env);
    free(P);
    return E;
}

/* Fock build restricted to a caller-chosen list of shell quartets (each of the
 * nbas^4 ordered quartets listed visits its slots exactly as rc_fock2e does);
 * used to time a bounded sample of the reference loop. */
void rc_fock2e_sample(double *G, const double *P, const int *quartets, long nq, const int *atm, int natm,
                      const int *bas, int nbas, const double *env)
{
    int *off = (int *)malloc(sizeof(int) * (nbas + 1));
    ao_offsets(bas, nbas, off);
    size_t n = off[nbas];
    double *buf = (double *)malloc(sizeof(double) * 50625);
    int d[4];
    for (long t = 0; t < nq; t++) {
        const int *shls = quartets + 4 * t;
        rc_int2e_shell(buf, shls, atm, natm, bas, nbas, env);
        quartet_dims(bas, shls, d);
        int c = 0;
        for (int l = 0; l < d[3]; l++)
            for (int k = 0; k < d[2]; k++)
                for (int j = 0; j < d[1]; j++)
                    for (int i = 0; i < d[0]; i++) {
                        size_t mu = off[shls[0]] + i, nu = off[shls[1]] + j;
                        size_t sg = off[shls[2]] + k, la = off[shls[3]] + l;
                        double v = buf[c++];
                        G[mu * n + nu] += P[la * n + sg] * v;
                        G[mu * n + la] += -0.5 * P[nu * n + sg] * v;
                    }
    }
    free(buf);
    free(off);
}

/* ------------------------------------------- screened symmetric G -- */
/* A second CPU comparator for the Fock build, with the shortcuts a tuned CPU
 * code takes that the reference loop (rc_fock2e, every ordered quartet) does
 * not: 8-fold permutational symmetry, Schwarz screening and threads.  Its
 * quartet set is the one the HIP path evaluates, so on the same inputs the two
 * agree to rounding; against rc_fock2e it differs by the screened quartets. */

/* Q[ij] = sqrt(max |(ij|ij)|) over the block, ij = i(i+1)/2 + j, i >= j. */
void rc_schwarz(double *Q, const int *atm, int natm, const int *bas, int nbas, const double *env, int nthreads)
{
    double u[NRMAX], w[NRMAX];
    rc_rys_roots(1, 1.0, u, w); /* fills the static quadrature before the threads start */
    const long npair = (long)nbas * (nbas + 1) / 2;
#pragma omp parallel num_threads(nthreads)
    {
        double *buf = (double *)malloc(sizeof(double) * 50625);
#pragma omp for schedule(dynamic, 16)
        for (long ij = 0; ij < npair; ij++) {
            int i = (int)((sqrt(8.0 * (double)ij + 1.0) - 1.0) / 2.0);
            while ((long)i * (i + 1) / 2 > ij) i--;
            while ((long)(i + 1) * (i + 2) / 2 <= ij) i++;
            const int j = (int)(ij - (long)i * (i + 1) / 2);
            int shls[4] = {i, j, i, j}, d[4];
            rc_int2e_shell(buf, shls, atm, natm, bas, nbas, env);
            quartet_dims(bas, shls, d);
            const long n = (long)d[0] * d[1] * d[2] * d[3];
            double m = 0.0;
            for (long t = 0; t < n; t++) m = fmax(m, fabs(buf[t]));
            Q[ij] = sqrt(m);
        }
        free(buf);
    }
}

/* G = J - K/2 (the convention of rc_fock2e, symmetric P) from the canonical
 * quartets i >= j, k >= l, ij >= kl with Q[ij] Q[kl] > tau, each evaluated
 * once and added to its eight images with weight
 * 2^-([i=j] + [k=l] + [ij=kl]).  Only bra pairs ij with ij % stride == phase
 * are visited (stride 1: the whole build; larger strides time a bounded
 * sample).  Threads keep private G's, summed at the end.  Returns the number
 * of quartets evaluated. */
long rc_fock2e_screened(double *G, const double *P, const double *Q, double tau, int stride, int phase,
                        const int *atm, int natm, const int *bas, int nbas, const double *env, int nthreads)
{
    double u[NRMAX], w[NRMAX];
    rc_rys_roots(1, 1.0, u, w);
    int *off = (int *)malloc(sizeof(int) * (nbas + 1));
    ao_offsets(bas, nbas, off);
    const size_t n = off[nbas];
    memset(G, 0, sizeof(double) * n * n);
    const long npair = (long)nbas * (nbas + 1) / 2;
    int *pi = (int *)malloc(sizeof(int) * npair), *pj = (int *)malloc(sizeof(int) * npair);
    for (int i = 0, c = 0; i < nbas; i++)
        for (int j = 0; j <= i; j++, c++) {
            pi[c] = i;
            pj[c] = j;
        }
    long count = 0;
#pragma omp parallel num_threads(nthreads) reduction(+ : count)
    {
        double *buf = (double *)malloc(sizeof(double) * 50625);
        double *Gt = (double *)calloc(n * n, sizeof(double));
#pragma omp for schedule(dynamic, 1)
        for (long ij = phase; ij < npair; ij += stride) {
            if (Q[ij] == 0.0) continue;
            for (long kl = 0; kl <= ij; kl++) {
                if (!(Q[ij] * Q[kl] > tau)) continue;
                int shls[4] = {pi[ij], pj[ij], pi[kl], pj[kl]}, d[4];
                rc_int2e_shell(buf, shls, atm, natm, bas, nbas, env);
                quartet_dims(bas, shls, d);
                double s = 1.0;
                if (shls[0] == shls[1]) s *= 0.5;
                if (shls[2] == shls[3]) s *= 0.5;
                if (ij == kl) s *= 0.5;
                count++;
                int c = 0;
                for (int l = 0; l < d[3]; l++)
                    for (int k = 0; k < d[2]; k++)
                        for (int j = 0; j < d[1]; j++)
                            for (int i = 0; i < d[0]; i++) {
                                const size_t a = off[shls[0]] + i, b = off[shls[1]] + j;
                                const size_t cc = off[shls[2]] + k, dd = off[shls[3]] + l;
                                const double v = s * buf[c++];
                                /* Coulomb: the four images with bra (a,b) or (b,a) add P_cd, ... */
                                const double jab = 2.0 * v * P[cc * n + dd], jcd = 2.0 * v * P[a * n + b];
                                Gt[a * n + b] += jab;
                                Gt[b * n + a] += jab;
                                Gt[cc * n + dd] += jcd;
                                Gt[dd * n + cc] += jcd;
                                /* exchange: G[p][s] -= P[q][r] v / 2 for every image (p q|r s) */
                                const double h = 0.5 * v;
                                Gt[a * n + dd] -= h * P[b * n + cc];
                                Gt[b * n + dd] -= h * P[a * n + cc];
                                Gt[a * n + cc] -= h * P[b * n + dd];
                                Gt[b * n + cc] -= h * P[a * n + dd];
                                Gt[cc * n + b] -= h * P[dd * n + a];
                                Gt[dd * n + b] -= h * P[cc * n + a];
                                Gt[cc * n + a] -= h * P[dd * n + b];
                                Gt[dd * n + a] -= h * P[cc * n + b];
                            }
            }
        }
#pragma omp critical
        for (size_t t = 0; t < n * n; t++) G[t] += Gt[t];
        free(Gt);
        free(buf);
    }
    free(pi);
    free(pj);
    free(off);
    return count;
}
