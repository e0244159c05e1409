/*
 * ed_oracle.c -- FP64 CPU restatement of the zero-temperature ED impurity
 * solver of yw-choi/DMFT-ED-ZT, used ONLY as the checker for the HIP path.
 *
 * TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library.  The product path
 * (libedgpu.so) never links, calls or falls back to it.
 *
 * Every function follows the cited reference file:line (paths relative to the
 * reference tree).  Conventions kept from the reference:
 *   - mode (bit) p = (spin-1)*nsite + site-1; sites 1..norb are impurity
 *     orbitals, norb+1..nsite are bath sites (ed_basis.f90:153-156);
 *   - a Fock state is up + 2^nsite * dn (ed_basis.f90:135);
 *   - global row (1-based) = (rank_dn-1)*nup + rank_up, up index fastest
 *     (ed_basis.f90:132-135,149-150); C arrays here are 0-based;
 *   - em(norb,2), ek(nbath,2), vmk(norb,nbath,2) column-major FP64.
 *
 * What pins each function (closed forms, brute-force enumeration, the dense
 * Jordan-Wigner construction in oracle/dense.py, LAPACK, golden fixtures) is
 * exercised by tests/test_oracle.py.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    int norb, nbath, nsite;
    double U, Up, Jex, Jp, mu;
} orc_model;

typedef struct {
    int nsite, ne_up, ne_dn;
    int nup, ndn;
    long ntot;
    int *up, *dn;           /* bit patterns, increasing order              */
    int *idx_up, *idx_dn;   /* pattern -> 1-based rank, 0 if not in sector */
} orc_basis;

#define EM(iorb, is) em[((is) - 1) * p->norb + (iorb) - 1]
#define EK(j, is) ek[((is) - 1) * p->nbath + (j) - 1]
#define VMK(iorb, j, is) vmk[(((is) - 1) * p->nbath + (j) - 1) * p->norb + (iorb) - 1]
#define BTEST(s, b) ((((uint32_t)(s)) >> (b)) & 1u)

/* numeric_utils.f90:181-222 (ifact, iP, iCom) in 64-bit to avoid overflow */
long orc_icom(int n, int m)
{
    if (m < 0 || m > n) return 0;
    if (n == m || m == 0) return 1;
    int k = (m > n / 2) ? n - m : m;
    long r = 1;
    for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
    return r;
}

/* ed_basis.f90:42-114 generate_basis: enumerate [minrange,maxrange], keep
 * popcount == N in increasing integer order.  Single rank: nloc == ntot. */
int orc_generate_basis(int nsite, int ne_up, int ne_dn, orc_basis *b)
{
    memset(b, 0, sizeof(*b));
    if (nsite < 1 || nsite > 15 || ne_up < 0 || ne_dn < 0 || ne_up > nsite || ne_dn > nsite)
        return 3;
    b->nsite = nsite;
    b->ne_up = ne_up;
    b->ne_dn = ne_dn;
    b->nup = (int)orc_icom(nsite, ne_up);
    b->ndn = (int)orc_icom(nsite, ne_dn);
    b->ntot = (long)b->nup * b->ndn;
    b->up = (int *)malloc(sizeof(int) * b->nup);
    b->dn = (int *)malloc(sizeof(int) * b->ndn);
    b->idx_up = (int *)calloc((size_t)1 << nsite, sizeof(int));
    b->idx_dn = (int *)calloc((size_t)1 << nsite, sizeof(int));
    int nud[2] = {ne_up, ne_dn};
    for (int ispin = 1; ispin <= 2; ++ispin) {
        int minrange = 0, maxrange = 0;
        for (int i = 1; i <= nud[ispin - 1]; ++i) {
            minrange += 1 << (i - 1);
            maxrange += 1 << (nsite - i);
        }
        int counts = 0;
        for (int i = minrange; i <= maxrange; ++i) {
            int nbit = 0;
            for (int j = 0; j < nsite; ++j) nbit += BTEST(i, j);
            if (nbit == nud[ispin - 1]) {
                ++counts;
                if (ispin == 1) { b->up[counts - 1] = i; b->idx_up[i] = counts; }
                else            { b->dn[counts - 1] = i; b->idx_dn[i] = counts; }
            }
        }
    }
    return 0;
}

void orc_free_basis(orc_basis *b)
{
    free(b->up); free(b->dn); free(b->idx_up); free(b->idx_dn);
    memset(b, 0, sizeof(*b));
}

/* ed_basis.f90:122-136 ed_basis_get (1-based row -> Fock state) */
static inline long basis_get(const orc_basis *b, long idx)
{
    long iup = (idx - 1) % b->nup + 1;
    long idown = (idx - 1) / b->nup + 1;
    return (long)b->up[iup - 1] + ((long)b->dn[idown - 1] << b->nsite);
}

/* ed_basis.f90:139-151 ed_basis_idx (Fock state -> 1-based row) */
static inline long basis_idx(const orc_basis *b, long s)
{
    long su = s & ((1L << b->nsite) - 1);
    long sd = s >> b->nsite;
    return (long)(b->idx_dn[sd] - 1) * b->nup + b->idx_up[su];
}

long orc_basis_get(const orc_basis *b, long idx) { return basis_get(b, idx); }
long orc_basis_idx(const orc_basis *b, long s) { return basis_idx(b, s); }

/* impurity_hamiltonian.f90:251-260 sgn: (-1)^{sum_{k=i}^{j-1} n_k} */
static inline int sgn(long state, int i, int j)
{
    int s = 1;
    for (int k = i; k <= j - 1; ++k)
        if ((state >> (k - 1)) & 1L) s = -s;
    return s;
}

/* impurity_hamiltonian.f90:264-276 sgn2: (-1)^{sum_{k=i}^{j-1} n_k,up + n_k,dn} */
static inline int sgn2(long state, int nsite, int i, int j)
{
    int s = 1;
    for (int k = i; k <= j - 1; ++k) {
        if ((state >> (k - 1)) & 1L) s = -s;
        if ((state >> (nsite + k - 1)) & 1L) s = -s;
    }
    return s;
}

/* impurity_hamiltonian.f90:77-247 hx: y = H x (py=0) or y += H x (py=1).
 * x_all is x itself (single rank: the allgatherv at :95-96 is the identity). */
void orc_hx(const orc_model *p, const double *em, const double *ek, const double *vmk,
            const orc_basis *b, const double *x, double *y, int py)
{
    const int norb = p->norb, nbath = p->nbath, nsite = p->nsite;
    const double U = p->U, Up = p->Up, Jex = p->Jex, Jp = p->Jp, mu = p->mu;
#pragma omp parallel for schedule(static)
    for (long bb = 1; bb <= b->ntot; ++bb) {
        long bra = basis_get(b, bb);
        double rowsum = 0.0;
        int ni[2], nj[2];
        for (int iorb = 1; iorb <= norb; ++iorb) {
            ni[0] = (int)((bra >> (iorb - 1)) & 1L);
            ni[1] = (int)((bra >> (nsite + iorb - 1)) & 1L);
            /* :119-121 onsite + intra-orbital U */
            rowsum += ((EM(iorb, 1) - mu) * ni[0] + (EM(iorb, 2) - mu) * ni[1]
                       + U * ni[0] * ni[1]) * x[bb - 1];
            /* :126-141 inter-orbital density-density */
            for (int jorb = iorb + 1; jorb <= norb; ++jorb) {
                nj[0] = (int)((bra >> (jorb - 1)) & 1L);
                nj[1] = (int)((bra >> (nsite + jorb - 1)) & 1L);
                rowsum += ((Up - Jex) * (ni[0] * nj[0] + ni[1] * nj[1])
                           + Up * (ni[0] * nj[1] + ni[1] * nj[0])) * x[bb - 1];
            }
            /* :146-174 hybridization */
            for (int ispin = 1; ispin <= 2; ++ispin) {
                for (int jbath = 1; jbath <= nbath; ++jbath) {
                    double v = VMK(iorb, jbath, ispin);
                    if (fabs(v) < 1.0e-16) continue;
                    int isite = (ispin - 1) * nsite + iorb;
                    int jsite = (ispin - 1) * nsite + norb + jbath;
                    int jb = (int)((bra >> (jsite - 1)) & 1L);
                    if (ni[ispin - 1] == 0 && jb) {
                        long ket = (bra & ~(1L << (jsite - 1))) | (1L << (isite - 1));
                        long a = basis_idx(b, ket);
                        rowsum += sgn(bra, isite, jsite) * v * x[a - 1];
                    } else if (ni[ispin - 1] != 0 && !jb) {
                        long ket = (bra & ~(1L << (isite - 1))) | (1L << (jsite - 1));
                        long a = basis_idx(b, ket);
                        rowsum -= sgn(bra, isite, jsite) * v * x[a - 1];
                    }
                }
            }
            /* :179-225 spin-flip and pair-hopping */
            for (int jorb = 1; jorb <= norb; ++jorb) {
                if (jorb == iorb) continue;
                nj[0] = (int)((bra >> (jorb - 1)) & 1L);
                nj[1] = (int)((bra >> (nsite + jorb - 1)) & 1L);
                int imin = iorb < jorb ? iorb : jorb;
                int imax = iorb < jorb ? jorb : iorb;
                if (ni[0] == 0 && nj[0] != 0 && nj[1] == 0 && ni[1] != 0) {
                    long ket = bra & ~(1L << (nsite + iorb - 1));
                    ket |= 1L << (nsite + jorb - 1);
                    ket &= ~(1L << (jorb - 1));
                    ket |= 1L << (iorb - 1);
                    long a = basis_idx(b, ket);
                    rowsum -= sgn2(bra, nsite, imin, imax) * Jp * x[a - 1];
                }
                if (ni[0] == 0 && ni[1] == 0 && nj[0] != 0 && nj[1] != 0) {
                    long ket = bra & ~(1L << (nsite + jorb - 1));
                    ket |= 1L << (nsite + iorb - 1);
                    ket &= ~(1L << (jorb - 1));
                    ket |= 1L << (iorb - 1);
                    long a = basis_idx(b, ket);
                    rowsum += sgn2(bra, nsite, imin, imax) * Jp * x[a - 1];
                }
            }
        }
        /* :232-239 bath levels */
        for (int jbath = 1; jbath <= nbath; ++jbath) {
            if ((bra >> (norb + jbath - 1)) & 1L) rowsum += EK(jbath, 1) * x[bb - 1];
            if ((bra >> (nsite + norb + jbath - 1)) & 1L) rowsum += EK(jbath, 2) * x[bb - 1];
        }
        if (py) y[bb - 1] += rowsum;
        else    y[bb - 1] = rowsum;
    }
}

/* numeric_utils.f90:159-179 mpi_dot_product / mpi_norm (single rank) */
/* Deterministic for a given thread count: fixed contiguous chunks, partial
 * sums combined in chunk order. */
static double dot(const double *a, const double *b, long n)
{
    enum { MAXCH = 256 };
    double part[MAXCH];
    int nch = 1;
#ifdef _OPENMP
    nch = omp_get_max_threads();
    if (nch > MAXCH) nch = MAXCH;
#endif
    long chunk = (n + nch - 1) / nch;
#pragma omp parallel for schedule(static, 1)
    for (int c = 0; c < nch; ++c) {
        double s = 0.0;
        long lo = (long)c * chunk, hi = lo + chunk < n ? lo + chunk : n;
        for (long i = lo; i < hi; ++i) s += a[i] * b[i];
        part[c] = s;
    }
    double s = 0.0;
    for (int c = 0; c < nch; ++c) s += part[c];
    return s;
}
double orc_dot(const double *a, const double *b, long n) { return dot(a, b, n); }

/* lanczos_solver.f90:92-173 lanczos_iteration.  a, b have maxnstep entries,
 * zero-padded beyond nstep (:116-117).  The breakdown test b(k)<1e-16 comes
 * after the divide at :154, as in the reference. */
void orc_lanczos_iteration(const orc_model *p, const double *em, const double *ek,
                           const double *vmk, const orc_basis *bs, const double *v_init,
                           int maxnstep, int *nstep, double *a, double *b)
{
    long n = bs->ntot;
    double *v = (double *)malloc(sizeof(double) * n);
    double *w = (double *)malloc(sizeof(double) * n);
    for (int i = 0; i < maxnstep; ++i) a[i] = b[i] = 0.0;
    double norm_v = sqrt(dot(v_init, v_init, n));
    for (long i = 0; i < n; ++i) w[i] = v_init[i] / norm_v;
    orc_hx(p, em, ek, vmk, bs, w, v, 0);
    a[0] = dot(w, v, n);
    b[0] = 0.0;
    if (maxnstep == 1) { *nstep = 1; free(v); free(w); return; }
    for (long i = 0; i < n; ++i) v[i] -= a[0] * w[i];
    b[1] = sqrt(dot(v, v, n));
    int k = 2;
    for (;;) {
        double bk = b[k - 1];
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n; ++i) {
            double t = w[i];
            w[i] = v[i] / bk;
            v[i] = -bk * t;
        }
        orc_hx(p, em, ek, vmk, bs, w, v, 1);
        a[k - 1] = dot(w, v, n);
        if (k >= maxnstep || b[k - 1] < 1.0e-16) { *nstep = k; break; }
        double ak = a[k - 1];
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n; ++i) v[i] -= ak * w[i];
        b[k] = sqrt(dot(v, v, n));
        ++k;
    }
    free(v); free(w);
}

/* Symmetric tridiagonal eigensystem by implicit QL with Wilkinson shifts
 * (textbook algorithm), standing in for MKL dstevx('V','A') at
 * lanczos_solver.f90:175-211.  d[n] diagonal, e[n] with e[i] coupling i,i+1
 * (e[n-1] unused).  On return w[] ascending, z column-major n x n with
 * column j the eigenvector of w[j].  Returns 0 or the failing index+1. */
int orc_tridiag_eig(int n, const double *d_in, const double *e_in, double *w, double *z)
{
    double *d = (double *)malloc(sizeof(double) * n);
    double *e = (double *)malloc(sizeof(double) * (n + 1));
    for (int i = 0; i < n; ++i) { d[i] = d_in[i]; e[i] = (i < n - 1) ? e_in[i] : 0.0; }
    e[n] = 0.0;
    for (int i = 0; i < n * n; ++i) z[i] = 0.0;
    for (int i = 0; i < n; ++i) z[i * n + i] = 1.0;
    int info = 0;
    for (int l = 0; l < n && !info; ++l) {
        int iter = 0, m;
        do {
            for (m = l; m < n - 1; ++m) {
                double dd = fabs(d[m]) + fabs(d[m + 1]);
                if (fabs(e[m]) <= 1.0e-300 || fabs(e[m]) <= 2.220446049250313e-16 * dd) break;
            }
            if (m != l) {
                if (iter++ == 60) { info = l + 1; break; }
                double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
                double r = hypot(g, 1.0);
                g = d[m] - d[l] + e[l] / (g + (g >= 0 ? fabs(r) : -fabs(r)));
                double s = 1.0, c = 1.0, pp = 0.0;
                int i;
                for (i = m - 1; i >= l; --i) {
                    double f = s * e[i], bb = c * e[i];
                    r = hypot(f, g);
                    e[i + 1] = r;
                    if (r == 0.0) { d[i + 1] -= pp; e[m] = 0.0; break; }
                    s = f / r;
                    c = g / r;
                    g = d[i + 1] - pp;
                    r = (d[i] - g) * s + 2.0 * c * bb;
                    pp = s * r;
                    d[i + 1] = g + pp;
                    g = c * r - bb;
                    for (int k = 0; k < n; ++k) {
                        double zf = z[(i + 1) * n + k];
                        z[(i + 1) * n + k] = s * z[i * n + k] + c * zf;
                        z[i * n + k] = c * z[i * n + k] - s * zf;
                    }
                }
                if (r == 0.0 && i >= l) continue;
                d[l] -= pp;
                e[l] = g;
                e[m] = 0.0;
            }
        } while (m != l);
    }
    /* selection sort ascending, carrying vectors */
    for (int i = 0; i < n - 1; ++i) {
        int kmin = i;
        for (int j = i + 1; j < n; ++j) if (d[j] < d[kmin]) kmin = j;
        if (kmin != i) {
            double t = d[i]; d[i] = d[kmin]; d[kmin] = t;
            for (int k = 0; k < n; ++k) {
                double tz = z[i * n + k]; z[i * n + k] = z[kmin * n + k]; z[kmin * n + k] = tz;
            }
        }
    }
    for (int i = 0; i < n; ++i) w[i] = d[i];
    free(d); free(e);
    return info;
}

/* lanczos_solver.f90:213-300 lanczos_ground_state: rerun the recurrence from
 * v_init, accumulate eigvec = sum_k coeff(k) w_k, normalize, residual. */
void orc_lanczos_ground_state(const orc_model *p, const double *em, const double *ek,
                              const double *vmk, const orc_basis *bs, const double *v_init,
                              int nstep, const double *a, const double *b, const double *coeff,
                              double eigval, double *eigvec, double *residual)
{
    long n = bs->ntot;
    double *v = (double *)malloc(sizeof(double) * n);
    double *w = (double *)malloc(sizeof(double) * n);
    double norm_v = sqrt(dot(v_init, v_init, n));
    for (long i = 0; i < n; ++i) { w[i] = v_init[i] / norm_v; eigvec[i] = coeff[0] * w[i]; }
    orc_hx(p, em, ek, vmk, bs, w, v, 0);
    if (nstep == 1) {
        /* the reference returns here without normalizing (:259-262) */
        *residual = 0.0;
        free(v); free(w);
        return;
    }
    for (long i = 0; i < n; ++i) v[i] -= a[0] * w[i];
    int k = 2;
    for (;;) {
        double bk = b[k - 1], ck = coeff[k - 1];
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n; ++i) {
            double t = w[i];
            w[i] = v[i] / bk;
            v[i] = -bk * t;
            eigvec[i] += ck * w[i];
        }
        orc_hx(p, em, ek, vmk, bs, w, v, 1);
        if (k >= nstep || b[k - 1] < 1.0e-16) break;
        double ak = a[k - 1];
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n; ++i) v[i] -= ak * w[i];
        ++k;
    }
    double nrm = sqrt(dot(eigvec, eigvec, n));
    for (long i = 0; i < n; ++i) eigvec[i] /= nrm;
    orc_hx(p, em, ek, vmk, bs, eigvec, v, 0);
    for (long i = 0; i < n; ++i) v[i] -= eigval * eigvec[i];
    *residual = dot(v, v, n);
    free(v); free(w);
}

/* Counter-based U[0,1) start vector: splitmix64 of (seed, index).  Stands in
 * for the time-seeded random_number at lanczos_solver.f90:42-46 so that two
 * implementations can start from the same vector. */
static inline uint64_t splitmix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
void orc_start_vector(uint64_t seed, long n, double *v)
{
    for (long i = 0; i < n; ++i)
        v[i] = (double)(splitmix64(seed * 0xD1B54A32D192ED03ull + (uint64_t)i) >> 11) * 0x1.0p-53;
}

/* lanczos_solver.f90:20-90 diag_lanczos with a caller-supplied v_init. */
int orc_diag_lanczos(const orc_model *p, const double *em, const double *ek, const double *vmk,
                     const orc_basis *bs, const double *v_init, int maxnstep, double *eigval,
                     double *eigvec, int *nstep_out, double *residual)
{
    double *a = (double *)malloc(sizeof(double) * maxnstep);
    double *b = (double *)malloc(sizeof(double) * maxnstep);
    int nstep;
    orc_lanczos_iteration(p, em, ek, vmk, bs, v_init, maxnstep, &nstep, a, b);
    double *ev = (double *)malloc(sizeof(double) * nstep);
    double *z = (double *)malloc(sizeof(double) * nstep * nstep);
    int info = orc_tridiag_eig(nstep, a, b + 1, ev, z);
    if (info == 0) {
        *eigval = ev[0];
        orc_lanczos_ground_state(p, em, ek, vmk, bs, v_init, nstep, a, b, z, ev[0], eigvec, residual);
    }
    *nstep_out = nstep;
    free(a); free(b); free(ev); free(z);
    return info ? 4 : 0;
}

/* ed_operator.f90:13-76 apply_c: vec_out = c^+ vec (pm=1) or c vec (pm=2),
 * output-indexed, sign over all modes below isite of the output state. */
void orc_apply_c(const orc_basis *bin, const double *vec, int pm, int iorb, int ispin,
                 const orc_basis *bout, double *vec_out)
{
    int nsite = bin->nsite;
    int isite = (ispin - 1) * nsite + iorb;
#pragma omp parallel for schedule(static)
    for (long i = 1; i <= bout->ntot; ++i) {
        long basis_i = basis_get(bout, i);
        int sgnsum = 0;
        for (int k = 1; k <= isite - 1; ++k) sgnsum += (int)((basis_i >> (k - 1)) & 1L);
        int sgntot = (sgnsum & 1) ? -1 : 1;
        long basis_j;
        if (pm == 1) {
            if (!((basis_i >> (isite - 1)) & 1L)) sgntot = 0;
            basis_j = basis_i & ~(1L << (isite - 1));
        } else {
            if ((basis_i >> (isite - 1)) & 1L) sgntot = 0;
            basis_j = basis_i | (1L << (isite - 1));
        }
        vec_out[i - 1] = 0.0;
        if (sgntot == 0) continue;
        long j = basis_idx(bin, basis_j);
        vec_out[i - 1] += vec[j - 1] * sgntot;
    }
}

/* numeric_utils.f90:11-23 continued_fraction_p and :32-44 _m (complex z). */
void orc_continued_fraction(double zr, double zi, int n, const double *a, const double *b,
                            int minus, double *fr, double *fi)
{
    double s = minus ? 1.0 : -1.0; /* z - a (plus) or z + a (minus) */
    double _Complex z = zr + zi * I;
    double _Complex f = b[n - 1] * b[n - 1] / (z + s * a[n - 1]);
    for (int i = n - 1; i >= 2; --i) f = b[i - 1] * b[i - 1] / (z + s * a[i - 1] - f);
    /* the last level is applied after the loop even when n == 1 */
    f = b[0] * b[0] / (z + s * a[0] - f);
    *fr = creal(f);
    *fi = cimag(f);
}

/* dmft_green.f90:177-260 green_diag for one (iorb, ispin): G+ from the c^+
 * chain in N+1, G- from the c chain in N-1.  g holds nw complex values
 * (interleaved re, im); ap..bm have maxnstep entries each.  Returns the two
 * chain lengths through nstep_pm[2]. */
void orc_green_diag(const orc_model *p, const double *em, const double *ek, const double *vmk,
                    const orc_basis *bs, const double *gsvec, double e0, int iorb, int ispin,
                    int maxnstep, int nw, const double *omega, double *g, double *ap,
                    double *bp, double *am, double *bm, int *nstep_pm)
{
    orc_basis bo;
    for (int iw = 0; iw < nw; ++iw) g[2 * iw] = g[2 * iw + 1] = 0.0;
    for (int pm = 1; pm <= 2; ++pm) {
        int dup = (ispin == 1) ? (pm == 1 ? 1 : -1) : 0;
        int ddn = (ispin == 2) ? (pm == 1 ? 1 : -1) : 0;
        double *aa = (pm == 1) ? ap : am, *bb = (pm == 1) ? bp : bm;
        for (int i = 0; i < maxnstep; ++i) aa[i] = bb[i] = 0.0;
        if (orc_generate_basis(bs->nsite, bs->ne_up + dup, bs->ne_dn + ddn, &bo) != 0) {
            nstep_pm[pm - 1] = 0;
            continue;
        }
        double *vi = (double *)malloc(sizeof(double) * bo.ntot);
        orc_apply_c(bs, gsvec, pm, iorb, ispin, &bo, vi);
        int nstep;
        orc_lanczos_iteration(p, em, ek, vmk, &bo, vi, maxnstep, &nstep, aa, bb);
        bb[0] = sqrt(dot(vi, vi, bo.ntot));
        nstep_pm[pm - 1] = nstep;
        for (int iw = 0; iw < nw; ++iw) {
            double fr, fi;
            if (pm == 1) orc_continued_fraction(e0, omega[iw], nstep, aa, bb, 0, &fr, &fi);
            else         orc_continued_fraction(-e0, omega[iw], nstep, aa, bb, 1, &fr, &fi);
            g[2 * iw] += fr;
            g[2 * iw + 1] += fi;
        }
        free(vi);
        orc_free_basis(&bo);
    }
}

/* post_process.f90:63-112 occupancy n(iorb,ispin) and double occupancy. */
void orc_observables(const orc_model *p, const orc_basis *bs, const double *gsvec,
                     double *occ /* norb x 2 col-major */, double *docc /* norb */)
{
    int norb = p->norb, nsite = p->nsite;
    for (int ispin = 1; ispin <= 2; ++ispin)
        for (int iorb = 1; iorb <= norb; ++iorb) {
            double s = 0.0;
            for (long i = 1; i <= bs->ntot; ++i) {
                long ib = basis_get(bs, i);
                if ((ib >> ((ispin - 1) * nsite + iorb - 1)) & 1L) s += gsvec[i - 1] * gsvec[i - 1];
            }
            occ[(ispin - 1) * norb + iorb - 1] = s;
        }
    for (int iorb = 1; iorb <= norb; ++iorb) {
        double s = 0.0;
        for (long i = 1; i <= bs->ntot; ++i) {
            long ib = basis_get(bs, i);
            if (((ib >> (iorb - 1)) & 1L) && ((ib >> (nsite + iorb - 1)) & 1L))
                s += gsvec[i - 1] * gsvec[i - 1];
        }
        docc[iorb - 1] = s;
    }
}

int orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}
