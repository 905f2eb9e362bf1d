/*
 * snb_oracle.c -- CPU restatement of SPONGE's nonbonded force path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this file; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, and only as the checker.
 *
 * Two layers, both plain C (gcc, -ffp-contract=off, no fast-math):
 *
 *  FP64 layer (orc_*_f64): the reference algorithm evaluated in double
 *  precision from the same float inputs.  It is the accuracy anchor for the
 *  5e-5 relative-RMS force bound and the 1e-6 energy/virial bound.
 *
 *  FP32 twin (orc_*_twin): the float operation sequence the HIP kernels
 *  execute, written independently here, with every per-pair / per-point term
 *  rounded to int64 fixed point (scale 2^32) before any summation.  Because
 *  integer sums are associative, the GPU accumulators must equal these
 *  bit-for-bit whatever the summation order on the device.
 *
 * Reference citations are paths inside yuhaosimba/SPONGE2.0 (file:line).
 *
 * Parity pins (what checks each function; see tests/test_oracle_*.py):
 *   orc_get_beta            beta the reference computes at rc 8 and 12 + erfc(beta rc)/rc = tol
 *   orc_fft_size            reference grid sizes for the fixture boxes + 2/3/5/7-smoothness
 *   orc_bspline_moduli_f64  closed form |sum M4(k+1) e^{2 pi i m k/K}|^-2 at m=0, K/2
 *   orc_nl_build            brute-force set definition itself; invariants (no dup, j>i, d<rc+skin)
 *   orc_pair_f64            two-atom closed forms; GROMACS fixture LJ_sr within the reference's test tolerance
 *   orc_pme_f64             exact Ewald sum (orc_ewald_f64) on small boxes; NaCl Madelung constant;
 *                           GROMACS fixture coulomb_sr+coulomb_recip within the reference's tolerance
 *   orc_fft3d_f64           naive DFT on small grids
 *   orc_excl_f64            closed form erf; E_excl + E_direct(no cutoff) = bare Coulomb
 *   orc_*_twin              the matching orc_*_f64 within float tolerance
 *   orc_ewald_f64           textbook Madelung constant of rock salt (1.747565)
 *   orc_nb14_f64            GROMACS fixture coulomb14 / lj14 terms; central-difference gradient
 *   orc_bond_f64            GROMACS fixture bond term (iterations 0-2) within the reference's
 *                           0.5 kcal/mol; closed form E = k (r - r0)^2, |f| = 2 k |r - r0|;
 *                           central-difference gradient; virial trace = sum f . dr
 *   orc_stress_twin         trace of the kinetic part = 2 E_kin (E_kin summed independently);
 *                           closed-form off-diagonal layout; potential part = numpy sum
 *   orc_pme_f64 (virial)    central differences of E_rec under box scaling: each diagonal
 *                           entry is +1/2 dE_rec/dlambda (the reference's sign and factor)
 *   orc_angle_f64           GROMACS fixture angle + Urey-Bradley term within the reference's
 *                           0.5 kcal/mol; with the bonds and the path, the fixture's per-atom
 *                           GROMACS forces on the water atoms; central-difference gradient
 *   orc_philox4x32_10       Random123 known-answer vectors
 *   orc_langevin_*          moments and a KS test of the normals; free-particle temperature
 *   orc_settle_*            closed-form TIP3P frame; restored geometry, centre of mass and
 *                           momentum after an unconstrained move
 *   orc_settle_pair_*       restored pair distance, zero momentum change, correction along the
 *                           remembered bond; a bond stretched along itself comes back exactly
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>

/* the erfcx polynomial table the kernels use (data shared with the library) */
#include "../sponge2.0_amd/csrc/snb_erfcx_table.h"
#endif

#define ORC_EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------------ */
/* constants                                                                 */
/* ------------------------------------------------------------------------ */
#define ORC_PI_D 3.14159265358979323846
/* PM_force.cpp:15-17 and Lennard_Jones_force.h:7 */
#define ORC_PI_F 3.1415926f
#define ORC_FIXED_SCALE 4294967296.0 /* 2^32 */

/* ------------------------------------------------------------------------ */
/* small helpers                                                             */
/* ------------------------------------------------------------------------ */
static inline float f_from_bits(uint32_t u)
{
    float f;
    memcpy(&f, &u, 4);
    return f;
}
static inline uint32_t bits_from_f(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}

/* round(f * 2^32) to nearest-even, exact for |f| < 2^19, via the 1.5*2^52
 * magic constant: the low mantissa bits of the sum hold the integer. */
static inline int64_t to_fixed(float f)
{
    double d = fma((double)f, ORC_FIXED_SCALE, 6755399441055744.0);
    uint64_t u;
    memcpy(&u, &d, 8);
    return (int64_t)u - (int64_t)0x4338000000000000LL;
}

/* ------------------------------------------------------------------------ */
/* FP32 math written from published methods (twin of the kernels' math)     */
/* ------------------------------------------------------------------------ */
/* 1/sqrt(x) for normal x > 0: integer seed 0x5f375a86 - (bits >> 1)
 * (Lomont, "Fast inverse square root", 2003) refined by three Newton steps
 * y <- y (1.5 - (x/2) y^2); only IEEE mul/fma, so it is reproducible. */
static inline float twin_rsqrtf(float x)
{
    float y = f_from_bits(0x5f375a86u - (bits_from_f(x) >> 1));
    float h = 0.5f * x;
    y = y * fmaf(-(h * y), y, 1.5f);
    y = y * fmaf(-(h * y), y, 1.5f);
    y = y * fmaf(-(h * y), y, 1.5f);
    return y;
}

/* FP64 twin pieces: exp by Cody-Waite (33-bit ln2 head) + degree-13 Taylor;
 * erfc(x) = e^{-x^2} erfcx(x) with erfcx from the kernels' table of
 * piecewise Chebyshev-interpolation polynomials (degree 12 on half-unit
 * intervals of [0, 6), Horner in t = (x - centre)/0.25; 0 from x = 6 on), and
 * g = 2/sqrt(pi) x e^{-x^2}.  Pinned against scipy's erfc in
 * tests/test_oracle_pins.py. */
static inline double twin_exp_d(double y)
{
    static const double inv_fact[14] = {1.0, 1.0, 0.5, 1.0 / 6.0, 1.0 / 24.0, 1.0 / 120.0,
                                        1.0 / 720.0, 1.0 / 5040.0, 1.0 / 40320.0,
                                        1.0 / 362880.0, 1.0 / 3628800.0, 1.0 / 39916800.0,
                                        1.0 / 479001600.0, 1.0 / 6227020800.0};
    y = fmax(y, -700.0);
    double n = rint(y * 1.4426950408889634);
    double r = fma(-n, 0.6931471806019545, y);
    r = fma(-n, -4.2009150726810846e-11, r);
    double p = inv_fact[13];
    for (int k = 12; k >= 0; k--) p = fma(p, r, inv_fact[k]);
    int64_t ni = (int64_t)n;
    uint64_t b = (uint64_t)(ni + 1023) << 52;
    double s;
    memcpy(&s, &b, 8);
    return p * s;
}

static inline double twin_erfc_parts_d(double x, double* g_out)
{
    double e = twin_exp_d(-(x * x));
    *g_out = 1.1283791670955126 * x * e;
    if (!(x < 6.0)) return 0.0;
    int k = (int)(x * 2.0);
    if (k < 0) k = 0;
    double t = (x - ((double)k + 0.5) * 0.5) * 4.0;
    const double* c = snb_erfcx_coef[k];
    double p = c[SNB_ERFCX_DEG];
    for (int j = SNB_ERFCX_DEG - 1; j >= 0; j--) p = fma(p, t, c[j]);
    return e * p;
}

/* Screened-Coulomb force shape h(x) = erfc(x) + 2/sqrt(pi) x e^{-x^2}
 * (the factor of qq/r^3 in Lennard_Jones_force.cpp:87-95, x = beta r) as the
 * kernels evaluate it: TWIN_COUL_N cubic Hermite pieces over
 * [0, beta rc (1 + 2^-7)], node k at k / scale with scale the float of
 * N / X; the piece coefficients come from h and h' = -4/sqrt(pi) x^2 e^{-x^2}
 * at both nodes in double, each rounded once to float.  Pinned against
 * scipy's erfc in tests/test_oracle_pins.py. */
#define TWIN_COUL_N 256
static float twin_coul_table(float beta, float cutoff, float* tab)
{
    double X = (double)beta * (double)cutoff * (1.0 + 1.0 / 128.0);
    if (!(X > 0.0))
    {
        for (int k = 0; k < TWIN_COUL_N; k++)
        {
            tab[4 * k] = 1.0f;
            tab[4 * k + 1] = tab[4 * k + 2] = tab[4 * k + 3] = 0.0f;
        }
        return 0.0f;
    }
    float scale = (float)((double)TWIN_COUL_N / X);
    double step = 1.0 / (double)scale;
    double hprev = 0.0, dprev = 0.0;
    for (int k = 0; k <= TWIN_COUL_N; k++)
    {
        double x = (double)k * step;
        double g;
        double hk = twin_erfc_parts_d(x, &g) + g;  /* erfc(x) + 2/sqrt(pi) x e^{-x^2} */
        double dk = ((-2.2567583341910252 * x) * x) * twin_exp_d(-(x * x)) * step;
        if (k > 0)
        {
            float* c = tab + 4 * (k - 1);
            c[0] = (float)hprev;
            c[1] = (float)dprev;
            c[2] = (float)((3.0 * (hk - hprev) - 2.0 * dprev) - dk);
            c[3] = (float)((2.0 * (hprev - hk) + dprev) + dk);
        }
        hprev = hk;
        dprev = dk;
    }
    return scale;
}

static inline float twin_coul_h(float x, float scale, const float* tab)
{
    float u = x * scale;
    int k = (int)u;
    if (k > TWIN_COUL_N - 1) k = TWIN_COUL_N - 1;
    float t = u - (float)k;
    const float* c = tab + 4 * k;
    return fmaf(fmaf(fmaf(c[3], t, c[2]), t, c[1]), t, c[0]);
}

static inline int64_t to_fixed_d(double d)
{
    double v = fma(d, ORC_FIXED_SCALE, 6755399441055744.0);
    uint64_t u;
    memcpy(&u, &v, 8);
    return (int64_t)u - (int64_t)0x4338000000000000LL;
}

/* ------------------------------------------------------------------------ */
/* periodic displacement (utils/vector.hpp:316-327), float and double        */
/* cell/rcell are LTMatrix3 {a11, a21, a22, a31, a32, a33}                   */
/* ------------------------------------------------------------------------ */
static inline void pbc_disp_f(const float* a, const float* b, const float* cell,
                              const float* rcell, float* d)
{
    float dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
    float sx = (dx * rcell[0] + dy * rcell[1]) + dz * rcell[3];
    float sy = dy * rcell[2] + dz * rcell[4];
    float sz = dz * rcell[5];
    float kx = floorf(sx + 0.5f), ky = floorf(sy + 0.5f), kz = floorf(sz + 0.5f);
    d[0] = dx - ((kx * cell[0] + ky * cell[1]) + kz * cell[3]);
    d[1] = dy - (ky * cell[2] + kz * cell[4]);
    d[2] = dz - kz * cell[5];
}

static inline void pbc_disp_d(const float* a, const float* b, const float* cell,
                              const float* rcell, double* d)
{
    double dx = (double)a[0] - b[0], dy = (double)a[1] - b[1],
           dz = (double)a[2] - b[2];
    double sx = dx * rcell[0] + dy * rcell[1] + dz * rcell[3];
    double sy = dy * rcell[2] + dz * rcell[4];
    double sz = dz * rcell[5];
    double kx = floor(sx + 0.5), ky = floor(sy + 0.5), kz = floor(sz + 0.5);
    d[0] = dx - (kx * cell[0] + ky * cell[1] + kz * cell[3]);
    d[1] = dy - (ky * cell[2] + kz * cell[4]);
    d[2] = dz - kz * cell[5];
}

/* Lennard_Jones_force.h:18-27: triangular type-pair index */
static inline int lj_pair_index(int a, int b)
{
    int hi = a > b ? a : b, lo = a > b ? b : a;
    return ((hi * (hi + 1)) >> 1) + lo;
}

/* neighbor_list.cpp:618-637: the reference tests membership only when j lies
 * between the first and the last entry of i's exclusion list. */
/* OpenMP helpers: thread count (ORC threads = OMP_NUM_THREADS when built with
 * -fopenmp, else 1) and the fixed-order sum of per-thread integer arrays. */
static inline int orc_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

static inline int orc_thread_id(void)
{
#ifdef _OPENMP
    return omp_get_thread_num();
#else
    return 0;
#endif
}

static void reduce_parts(const int64_t* part, int nt, size_t len, int64_t* out)
{
#pragma omp parallel for num_threads(nt) schedule(static)
    for (size_t k = 0; k < len; k++)
    {
        int64_t s = 0;
        for (int t = 0; t < nt; t++) s += part[(size_t)t * len + k];
        out[k] += s;
    }
}

ORC_EXPORT int orc_num_threads(void) { return orc_threads(); }

static int is_excluded(int i, int j, const int* excl_start, const int* excl_num,
                       const int* excl_list)
{
    int n = excl_num[i];
    if (n <= 0) return 0;
    int s = excl_start[i];
    int lo = excl_list[s], hi = excl_list[s + n - 1];
    if (j < lo || j > hi) return 0;
    for (int k = s; k < s + n; k++)
        if (excl_list[k] == j) return 1;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* host precompute                                                           */
/* ------------------------------------------------------------------------ */
/* PM_force.cpp:244-272: doubling then 50 bisection steps, in float. */
ORC_EXPORT float orc_get_beta(float cutoff, float tolerance)
{
    float beta = 0.0f, low, high, tempf;
    int ilow, ihigh;
    high = 1.0f;
    ihigh = 1;
    while (1)
    {
        tempf = erfcf(high * cutoff) / cutoff;
        if (tempf <= tolerance) break;
        high *= 2;
        ihigh++;
    }
    ihigh += 50;
    low = 0.0f;
    for (ilow = 1; ilow < ihigh; ilow++)
    {
        beta = (low + high) / 2;
        tempf = erfcf(beta * cutoff) / cutoff;
        if (tempf >= tolerance)
            low = beta;
        else
            high = beta;
    }
    return beta;
}

/* common.cpp:521-557 */
ORC_EXPORT int orc_check_2357(int number)
{
    static const int primes[4] = {2, 3, 5, 7};
    if (number <= 0) return 0;
    for (int p = 0; p < 4; p++)
        while (number % primes[p] == 0) number /= primes[p];
    return number == 1;
}

/* common.cpp:639-659 */
ORC_EXPORT int orc_fft_size(float length)
{
    int tempi = (int)ceil(length + 3) >> 2 << 2;
    if (tempi >= 60 && tempi <= 68)
        tempi = 64;
    else if (tempi >= 120 && tempi <= 136)
        tempi = 128;
    else if (tempi >= 240 && tempi <= 272)
        tempi = 256;
    else if (tempi >= 480 && tempi <= 544)
        tempi = 512;
    else if (tempi >= 960 && tempi <= 1088)
        tempi = 1024;
    while (!orc_check_2357(tempi)) tempi += 4;
    return tempi;
}

/* PM_force.cpp:27-62 (order 4): b(m) = |e^{2 pi i 3 m/K} /
 * sum_{k=0}^{2} M4(k+1) e^{2 pi i m k/K}|^2 with M4(1)=1/6, M4(2)=2/3,
 * M4(3)=1/6 (the cardinal cubic B-spline at the integers). */
ORC_EXPORT void orc_bspline_moduli_f64(int K, double* out)
{
    static const double M4[3] = {1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0};
    for (int m = 0; m < K; m++)
    {
        double re = 0, im = 0;
        for (int k = 0; k < 3; k++)
        {
            double a = 2.0 * ORC_PI_D * m * k / K;
            re += M4[k] * cos(a);
            im += M4[k] * sin(a);
        }
        out[m] = 1.0 / (re * re + im * im);
    }
}

/* PM_force.cpp:544-612: BC over the half spectrum kz <= fftz/2, index
 * kx*ffty*(fftz/2+1) + ky*(fftz/2+1) + kz; m = MultiplyTranspose(k', rcell).
 * As the kernels do, kz times rcell's last entry is a float product;
 * bc_all_double takes every product in double (the FP64 evaluation's). */
static void bc_build(int fx, int fy, int fz, const float* rcell, double beta, double volume,
                     double* bc, int all_double);

ORC_EXPORT void orc_bc_f64(int fx, int fy, int fz, const float* rcell,
                           double beta, double volume, double* bc)
{
    bc_build(fx, fy, fz, rcell, beta, volume, bc, 0);
}

static void bc_build(int fx, int fy, int fz, const float* rcell, double beta, double volume,
                     double* bc, int all_double)
{
    double* b1 = (double*)malloc(sizeof(double) * fx);
    double* b2 = (double*)malloc(sizeof(double) * fy);
    double* b3 = (double*)malloc(sizeof(double) * fz);
    orc_bspline_moduli_f64(fx, b1);
    orc_bspline_moduli_f64(fy, b2);
    orc_bspline_moduli_f64(fz, b3);
    int fc = fz / 2 + 1;
    double mpre = ORC_PI_D * ORC_PI_D / (beta * beta);
    for (int kx = 0; kx < fx; kx++)
    {
        double kxr = kx > fx / 2 ? kx - fx : kx;
        for (int ky = 0; ky < fy; ky++)
        {
            double kyr = ky > fy / 2 ? ky - fy : ky;
            for (int kz = 0; kz < fc; kz++)
            {
                double mx = kxr * rcell[0];
                double my = kxr * rcell[1] + kyr * rcell[2];
                double mz = all_double
                                ? kxr * rcell[3] + kyr * rcell[4] + (double)kz * rcell[5]
                                : kxr * rcell[3] + kyr * rcell[4] + kz * rcell[5];
                double msq = mx * mx + my * my + mz * mz;
                size_t idx = ((size_t)kx * fy + ky) * fc + kz;
                if (kx + ky + kz == 0)
                    bc[idx] = 0.0;
                else
                    bc[idx] = exp(-mpre * msq) / (ORC_PI_D * msq * volume) *
                              b1[kx] * b2[ky] * b3[kz];
            }
        }
    }
    free(b1);
    free(b2);
    free(b3);
}

/* ------------------------------------------------------------------------ */
/* neighbor list (neighbor_list.cpp:206-275, :495-591, :594-665)            */
/* ------------------------------------------------------------------------ */
/* The reference's half list as a set: j in row i iff j > i, the float
 * displacement of Get_Periodic_Displacement(crd_i, crd_j) has
 * (x*x + y*y) + z*z < (0.5(rc+skin))^2 * 4, and j is not excluded from i.
 * Rows come out sorted ascending.  Brute force O(N^2).
 * Returns 1 if any row overflowed max_nb (the row is then truncated). */
ORC_EXPORT int orc_nl_build(int n, const float* crd, const float* cell,
                            const float* rcell, float cutoff, float skin,
                            const int* excl_start, const int* excl_num,
                            const int* excl_list, int max_nb, int* counts,
                            int* rows)
{
    float g = 0.5f * (cutoff + skin);
    float cs2 = g * g * 4.0f;
    int overflow = 0;
    for (int i = 0; i < n; i++)
    {
        int c = 0;
        for (int j = i + 1; j < n; j++)
        {
            float d[3];
            pbc_disp_f(crd + 3 * i, crd + 3 * j, cell, rcell, d);
            float d2 = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2];
            if (!(d2 < cs2)) continue;
            if (excl_num && is_excluded(i, j, excl_start, excl_num, excl_list))
                continue;
            if (c < max_nb)
                rows[(size_t)i * max_nb + c] = j;
            else
                overflow = 1;
            c++;
        }
        counts[i] = c < max_nb ? c : max_nb;
    }
    return overflow;
}

/* ------------------------------------------------------------------------ */
/* LJ + direct Coulomb (Lennard_Jones_force.cpp:43-136, Lennard_Jones_force.h:43-73) */
/* ------------------------------------------------------------------------ */
/* FP64: per pair (i row atom, j entry), dr = pbc(r_j - r_i); if r < rc:
 *   F/r = (B - A r^-6) r^-8 - qq r^-3 (2 beta r/sqrt(pi) e^{-beta^2 r^2} + erfc(beta r))
 *   f = (F/r) dr; frc[i] += f; frc[j] -= f;   virial[i] -= f (x) dr (symmetrised)
 *   e_lj[i] += A/12 r^-12 - B/6 r^-6;   e_c[i] += qq erfc(beta r)/r
 * A, B are the in-memory tables (file values x12 and x6). */
ORC_EXPORT void orc_pair_f64(int n, const float* crd, const float* charge,
                             const int* type, const float* lj_a,
                             const float* lj_b, const float* cell,
                             const float* rcell, float cutoff, float beta,
                             const int* counts, const int* rows, int max_nb,
                             double* frc, double* e_lj, double* e_c,
                             double* virial)
{
    double b = beta, rc = cutoff;
    for (int i = 0; i < n; i++)
    {
        for (int k = 0; k < counts[i]; k++)
        {
            int j = rows[(size_t)i * max_nb + k];
            double d[3];
            pbc_disp_d(crd + 3 * j, crd + 3 * i, cell, rcell, d);
            double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            double r = sqrt(r2);
            if (!(r < rc)) continue;
            int p = lj_pair_index(type[i], type[j]);
            double A = lj_a[p], B = lj_b[p];
            double ri2 = 1.0 / r2, ri6 = ri2 * ri2 * ri2;
            double qq = (double)charge[i] * charge[j];
            double x = b * r;
            double ex = exp(-x * x);
            double fc = qq * ri2 / r * (2.0 / sqrt(ORC_PI_D) * x * ex + erfc(x));
            double fabs_ = (B - A * ri6) * ri6 * ri2 - fc;
            double f[3] = {fabs_ * d[0], fabs_ * d[1], fabs_ * d[2]};
            for (int c = 0; c < 3; c++)
            {
                frc[3 * i + c] += f[c];
                frc[3 * j + c] -= f[c];
            }
            if (e_lj) e_lj[i] += (A / 12.0 * ri6 - B / 6.0) * ri6;
            if (e_c) e_c[i] += qq * erfc(x) / r;
            if (virial)
            {
                double* v = virial + 6 * i;
                v[0] -= f[0] * d[0];
                v[1] -= f[0] * d[1] + f[1] * d[0];
                v[2] -= f[1] * d[1];
                v[3] -= f[0] * d[2] + f[2] * d[0];
                v[4] -= f[1] * d[2] + f[2] * d[1];
                v[5] -= f[2] * d[2];
            }
        }
    }
}

/* FP32 twin of the pair kernel.  Accumulators (+=) are int64 at scale 2^32:
 * frc[3n] (SoA x|y|z), e_lj[n], e_c[n], vir[6n] (SoA by component). */
ORC_EXPORT void orc_pair_twin(int n, const float* crd, const float* charge,
                              const int* type, const float* lj_a,
                              const float* lj_b, const float* cell,
                              const float* rcell, float cutoff, float beta,
                              const int* counts, const int* rows, int max_nb,
                              int64_t* frc, int64_t* e_lj, int64_t* e_c,
                              int64_t* vir)
{
    float rc2 = cutoff * cutoff;
    float ctab[4 * TWIN_COUL_N];
    float cscale = twin_coul_table(beta, cutoff, ctab);
    /* rows in parallel; the i- and j-side force terms go to per-thread
     * arrays (integer sums, so the result does not depend on the split);
     * energies and virials belong to row i alone */
    int nt = orc_threads();
    int64_t* part = (int64_t*)calloc((size_t)nt * 3 * n + 1, sizeof(int64_t));
#pragma omp parallel num_threads(nt)
    {
    int64_t* f = part + (size_t)orc_thread_id() * 3 * n;
#pragma omp for schedule(dynamic, 64)
    for (int i = 0; i < n; i++)
    {
        const float* ri = crd + 3 * i;
        float qi = charge[i];
        int ti = type[i];
        for (int k = 0; k < counts[i]; k++)
        {
            int j = rows[(size_t)i * max_nb + k];
            float d[3];
            pbc_disp_f(crd + 3 * j, ri, cell, rcell, d);
            float r2 = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2];
            if (!(r2 < rc2)) continue;
            int p = lj_pair_index(ti, type[j]);
            float A = lj_a[p], B = lj_b[p];
            float rinv = twin_rsqrtf(r2);
            float r = r2 * rinv;
            float rinv2 = rinv * rinv;
            float rinv6 = (rinv2 * rinv2) * rinv2;
            float flj = ((B - A * rinv6) * rinv6) * rinv2;
            float x = beta * r;
            float qq = qi * charge[j];
            float fc = (qq * (rinv2 * rinv)) * twin_coul_h(x, cscale, ctab);
            float fabs_ = flj - fc;
            float fx = fabs_ * d[0], fy = fabs_ * d[1], fz = fabs_ * d[2];
            int64_t ix = to_fixed(fx), iy = to_fixed(fy), iz = to_fixed(fz);
            f[i] += ix;
            f[n + i] += iy;
            f[2 * n + i] += iz;
            f[j] -= ix;
            f[n + j] -= iy;
            f[2 * n + j] -= iz;
            if (e_lj || e_c)
            {
                /* energies in FP64 from the float displacement */
                double ddx = d[0], ddy = d[1], ddz = d[2];
                double r2d = (ddx * ddx + ddy * ddy) + ddz * ddz;
                double rd = sqrt(r2d);
                double rid = 1.0 / rd;
                double ri2 = rid * rid;
                double ri6 = (ri2 * ri2) * ri2;
                if (e_lj)
                    e_lj[i] += to_fixed_d((((double)A * (1.0 / 12.0)) * ri6 -
                                           (double)B * (1.0 / 6.0)) * ri6);
                double gd;
                double erfcd = twin_erfc_parts_d((double)beta * rd, &gd);
                if (e_c)
                    e_c[i] += to_fixed_d((((double)qi * (double)charge[j]) * erfcd) * rid);
            }
            if (vir)
            {
                vir[0 * (size_t)n + i] += to_fixed(-(fx * d[0]));
                vir[1 * (size_t)n + i] += to_fixed(-(fx * d[1] + fy * d[0]));
                vir[2 * (size_t)n + i] += to_fixed(-(fy * d[1]));
                vir[3 * (size_t)n + i] += to_fixed(-(fx * d[2] + fz * d[0]));
                vir[4 * (size_t)n + i] += to_fixed(-(fy * d[2] + fz * d[1]));
                vir[5 * (size_t)n + i] += to_fixed(-(fz * d[2]));
            }
        }
    }
    }
    reduce_parts(part, nt, 3 * (size_t)n, frc);
    free(part);
}

/* ------------------------------------------------------------------------ */
/* 1-4 pairs (nb14/nb14.cpp:7-100)                                          */
/* ------------------------------------------------------------------------ */
/* Per pair k: a = atom_a[k], b = atom_b[k], dr = pbc(r_b - r_a); charges
 * below 1e-10 in magnitude become 1e-10;
 *   f = -A r^-14 + B r^-8 - cf qa qb r^-3,  force f dr on a and -f dr on b;
 *   E_cf = cf qa qb / r,  E_lj = 0.08333333 A r^-12 - 0.1666666 B r^-6
 *   (the reference's constants; A, B carry the x12 / x6 of the LJ files);
 *   atom a takes E_cf + E_lj and the virial -f dr (x) dr.
 * The FP64 layer evaluates this in double from the float inputs.  e_pair
 * (2m: LJ | Coulomb) and e_atom may be NULL. */
static inline float nb14_charge(float q)
{
    return (double)fabsf(q) < 1e-10 ? (float)1e-10 : q;
}

ORC_EXPORT void orc_nb14_f64(int n, const float* crd, const float* charge, const float* cell,
                             const float* rcell, int m, const int* atom_a, const int* atom_b,
                             const float* A, const float* B, const float* cf, double* frc,
                             double* e_pair, double* e_atom, double* vir)
{
    for (int k = 0; k < m; k++)
    {
        int a = atom_a[k], b = atom_b[k];
        double d[3];
        pbc_disp_d(crd + 3 * b, crd + 3 * a, cell, rcell, d);
        double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        double r = sqrt(r2);
        double qa = nb14_charge(charge[a]), qb = nb14_charge(charge[b]);
        double ir2 = 1.0 / r2;
        double ir6 = ir2 * ir2 * ir2;
        double f = -(double)A[k] * ir6 * ir6 * ir2 + (double)B[k] * ir6 * ir2 -
                   (double)cf[k] * qa * qb * ir2 / r;
        for (int c = 0; c < 3; c++)
        {
            frc[3 * a + c] += f * d[c];
            frc[3 * b + c] -= f * d[c];
        }
        double ecf = (double)cf[k] * qa * qb / r;
        double elj = 0.08333333 * (double)A[k] * ir6 * ir6 - 0.1666666 * (double)B[k] * ir6;
        if (e_pair)
        {
            e_pair[k] = elj;
            e_pair[m + k] = ecf;
        }
        if (e_atom) e_atom[a] += ecf + elj;
        if (vir)
        {
            double* v = vir + 6 * a;
            double fx = f * d[0], fy = f * d[1], fz = f * d[2];
            v[0] -= fx * d[0];
            v[1] -= fx * d[1] + fy * d[0];
            v[2] -= fy * d[1];
            v[3] -= fx * d[2] + fz * d[0];
            v[4] -= fy * d[2] + fz * d[1];
            v[5] -= fz * d[2];
        }
    }
}

/* Float twin of the 1-4 kernel: the float force in the reference's operation
 * order (1/r^2 and 1/r as double quotients rounded to float, 1/r from the
 * correctly rounded sqrtf), every term rounded to int64 at scale 2^32;
 * energies in double from the float displacement.  Accumulators (+=): frc[3n]
 * SoA, e_pair[2m] (LJ | Coulomb), e_atom[n], vir[6n] SoA; the last three may
 * be NULL. */
ORC_EXPORT void orc_nb14_twin(int n, const float* crd, const float* charge, const float* cell,
                              const float* rcell, int m, const int* atom_a, const int* atom_b,
                              const float* A, const float* B, const float* cf, int64_t* frc,
                              int64_t* e_pair, int64_t* e_atom, int64_t* vir)
{
    for (int k = 0; k < m; k++)
    {
        int a = atom_a[k], b = atom_b[k];
        float d[3];
        pbc_disp_f(crd + 3 * b, crd + 3 * a, cell, rcell, d);
        float dr2 = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2];
        float dr_2 = (float)(1.0 / (double)dr2);
        float dr_4 = dr_2 * dr_2;
        float dr_8 = dr_4 * dr_4;
        float dr_14 = (dr_8 * dr_4) * dr_2;
        float dr_1 = (float)(1.0 / (double)sqrtf(dr2));
        float qa = nb14_charge(charge[a]), qb = nb14_charge(charge[b]);
        float fcf = ((-((cf[k] * dr_2) * dr_1)) * qa) * qb;
        float flj = (-A[k]) * dr_14 + B[k] * dr_8;
        float f = flj + fcf;
        float fv[3] = {f * d[0], f * d[1], f * d[2]};
        for (int c = 0; c < 3; c++)
        {
            int64_t q = to_fixed(fv[c]);
            frc[(size_t)c * n + a] += q;
            frc[(size_t)c * n + b] -= q;
        }
        if (e_pair || e_atom)
        {
            double dx = d[0], dy = d[1], dz = d[2];
            double r2d = (dx * dx + dy * dy) + dz * dz;
            double d1 = 1.0 / sqrt(r2d);
            double d2 = 1.0 / r2d;
            double d4 = d2 * d2, d8 = d4 * d4;
            int64_t pc = to_fixed_d((((double)cf[k] * d1) * (double)qa) * (double)qb);
            int64_t pl = to_fixed_d(((0.08333333 * (double)A[k]) * d4) * d8 -
                                    ((0.1666666 * (double)B[k]) * d4) * d2);
            if (e_pair)
            {
                e_pair[k] += pl;
                e_pair[m + k] += pc;
            }
            if (e_atom) e_atom[a] += pl + pc;
        }
        if (vir)
        {
            vir[0 * (size_t)n + a] += to_fixed(-(fv[0] * d[0]));
            vir[1 * (size_t)n + a] += to_fixed(-(fv[0] * d[1] + fv[1] * d[0]));
            vir[2 * (size_t)n + a] += to_fixed(-(fv[1] * d[1]));
            vir[3 * (size_t)n + a] += to_fixed(-(fv[0] * d[2] + fv[2] * d[0]));
            vir[4 * (size_t)n + a] += to_fixed(-(fv[1] * d[2] + fv[2] * d[1]));
            vir[5 * (size_t)n + a] += to_fixed(-(fv[2] * d[2]));
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Harmonic bonds (bond/bond.cpp:15-64)                                     */
/* ------------------------------------------------------------------------ */
/* dr = pbc(r_a - r_b), E = k (r - r0)^2, force f = -2 k (r - r0) dr / r on a,
 * -f on b; atom a takes the energy and the virial f (x) dr.  FP64 layer:
 * e_pair[m], e_atom[n], vir[n][6] may be NULL. */
ORC_EXPORT void orc_bond_f64(int n, const float* crd, const float* cell, const float* rcell,
                             int m, const int* atom_a, const int* atom_b, const float* k,
                             const float* r0, double* frc, double* e_pair, double* e_atom,
                             double* vir)
{
    for (int q = 0; q < m; q++)
    {
        int a = atom_a[q], b = atom_b[q];
        double d[3];
        pbc_disp_d(crd + 3 * a, crd + 3 * b, cell, rcell, d);
        double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        double x = r - (double)r0[q];
        double s = -2.0 * (double)k[q] * x / r;
        for (int c = 0; c < 3; c++)
        {
            frc[3 * a + c] += s * d[c];
            frc[3 * b + c] -= s * d[c];
        }
        double e = (double)k[q] * x * x;
        if (e_pair) e_pair[q] = e;
        if (e_atom) e_atom[a] += e;
        if (vir)
        {
            double* v = vir + 6 * a;
            double f[3] = {s * d[0], s * d[1], s * d[2]};
            v[0] += f[0] * d[0];
            v[1] += f[0] * d[1] + f[1] * d[0];
            v[2] += f[1] * d[1];
            v[3] += f[0] * d[2] + f[2] * d[0];
            v[4] += f[1] * d[2] + f[2] * d[1];
            v[5] += f[2] * d[2];
        }
    }
}

/* Float twin of the bond kernel: the reference's float order
 * (r = sqrtf(r^2), s = -(2 (r - r0) k) / r), terms rounded to int64 at 2^32;
 * energies in double from the float displacement.  frc[3n], e_pair[m],
 * e_atom[n], vir[6n] (+=; the last three may be NULL). */
ORC_EXPORT void orc_bond_twin(int n, const float* crd, const float* cell, const float* rcell,
                              int m, const int* atom_a, const int* atom_b, const float* k,
                              const float* r0, int64_t* frc, int64_t* e_pair, int64_t* e_atom,
                              int64_t* vir)
{
    for (int q = 0; q < m; q++)
    {
        int a = atom_a[q], b = atom_b[q];
        float d[3];
        pbc_disp_f(crd + 3 * a, crd + 3 * b, cell, rcell, d);
        float r = sqrtf((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
        float x = r - r0[q];
        float g = (2.0f * x) * k[q];
        float s = (-g) / r;
        float f[3] = {s * d[0], s * d[1], s * d[2]};
        for (int c = 0; c < 3; c++)
        {
            int64_t v = to_fixed(f[c]);
            frc[(size_t)c * n + a] += v;
            frc[(size_t)c * n + b] -= v;
        }
        if (e_pair || e_atom)
        {
            double dx = d[0], dy = d[1], dz = d[2];
            double dd = sqrt((dx * dx + dy * dy) + dz * dz) - (double)r0[q];
            int64_t pe = to_fixed_d(((double)k[q] * dd) * dd);
            if (e_pair) e_pair[q] += pe;
            if (e_atom) e_atom[a] += pe;
        }
        if (vir)
        {
            vir[0 * (size_t)n + a] += to_fixed(f[0] * d[0]);
            vir[1 * (size_t)n + a] += to_fixed(f[0] * d[1] + f[1] * d[0]);
            vir[2 * (size_t)n + a] += to_fixed(f[1] * d[1]);
            vir[3 * (size_t)n + a] += to_fixed(f[0] * d[2] + f[2] * d[0]);
            vir[4 * (size_t)n + a] += to_fixed(f[1] * d[2] + f[2] * d[1]);
            vir[5 * (size_t)n + a] += to_fixed(f[2] * d[2]);
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Harmonic angles (angle/angle.cpp:6-83), the angle half of UREY_BRADLEY     */
/* (angle/Urey_Bradley_force.cpp:101-115; its bond half is orc_bond_f64)     */
/* ------------------------------------------------------------------------ */
/* FP64 only: a test aid that lets the checker assemble the total force of a
 * fixture whose reference forces include angle terms.  With
 * dij = pbc(r_a - r_b), dkj = pbc(r_c - r_b), cos t = dij.dkj / (|dij||dkj|)
 * clamped to +-0.999999, E = k (t - t0)^2; g = -2 k (t - t0) / sin t;
 * f_a = g cos t / |dij|^2 dij - g / (|dij||dkj|) dkj, f_c likewise,
 * f_b = -f_a - f_c; energy to atom a. */
ORC_EXPORT void orc_angle_f64(int n, const float* crd, const float* cell, const float* rcell,
                              int m, const int* atom_a, const int* atom_b, const int* atom_c,
                              const float* k, const float* theta0, double* frc, double* e_pair)
{
    for (int q = 0; q < m; q++)
    {
        int a = atom_a[q], b = atom_b[q], c = atom_c[q];
        double dij[3], dkj[3];
        pbc_disp_d(crd + 3 * a, crd + 3 * b, cell, rcell, dij);
        pbc_disp_d(crd + 3 * c, crd + 3 * b, cell, rcell, dkj);
        double rij2 = 1.0 / (dij[0] * dij[0] + dij[1] * dij[1] + dij[2] * dij[2]);
        double rkj2 = 1.0 / (dkj[0] * dkj[0] + dkj[1] * dkj[1] + dkj[2] * dkj[2]);
        double rr = sqrt(rij2 * rkj2);
        double ct = (dij[0] * dkj[0] + dij[1] * dkj[1] + dij[2] * dkj[2]) * rr;
        if (ct > 0.999999) ct = 0.999999;
        if (ct < -0.999999) ct = -0.999999;
        double th = acos(ct);
        double dth = th - (double)theta0[q];
        double g = -2.0 * (double)k[q] * dth / sin(th);
        double cross = g * rr, self = g * ct;
        for (int d = 0; d < 3; d++)
        {
            double fa = self * rij2 * dij[d] - cross * dkj[d];
            double fc = self * rkj2 * dkj[d] - cross * dij[d];
            frc[3 * a + d] += fa;
            frc[3 * c + d] += fc;
            frc[3 * b + d] -= fa + fc;
        }
        if (e_pair) e_pair[q] = (double)k[q] * dth * dth;
    }
}

/* ------------------------------------------------------------------------ */
/* Stress assembly (MD_core/sys.hpp:14-85, :493-524; MD_core.cpp:948-975)    */
/* ------------------------------------------------------------------------ */
static int64_t stress_fixed(double x) { return (int64_t)rint(x * ORC_FIXED_SCALE); }

/* Float twin of the library's pressure assembly: sums of the atom virials
 * (float entries) and of m v (x) v (FP64 products, off-diagonals 2 m v_a v_b)
 * rounded per term to int64 at 2^32; then, in float,
 *   vt += sum_pot;  stress += inv_v * vt;  stress += inv_v * sum_kin.
 * atom_virial[6n] (a11 a21 a22 a31 a32 a33 per atom) or NULL (no potential
 * part); vel[3n]/mass[n] or NULL (no kinetic part); vt[6], stress[6] in/out.
 * sums[12] (may be NULL) receives the two int64 sums. */
ORC_EXPORT void orc_stress_twin(int n, const float* atom_virial, const float* vel,
                                const float* mass, float inv_v, float* vt, float* stress,
                                int64_t* sums)
{
    int64_t w[12] = {0};
    for (int i = 0; i < n; i++)
    {
        if (atom_virial)
            for (int k = 0; k < 6; k++) w[k] += stress_fixed((double)atom_virial[6 * i + k]);
        if (vel && mass)
        {
            double m = mass[i], x = vel[3 * i], y = vel[3 * i + 1], z = vel[3 * i + 2];
            w[6] += stress_fixed(m * (x * x));
            w[7] += stress_fixed(m * (2.0 * (x * y)));
            w[8] += stress_fixed(m * (y * y));
            w[9] += stress_fixed(m * (2.0 * (x * z)));
            w[10] += stress_fixed(m * (2.0 * (y * z)));
            w[11] += stress_fixed(m * (z * z));
        }
    }
    if (atom_virial)
        for (int k = 0; k < 6; k++)
        {
            vt[k] += (float)((double)w[k] * (1.0 / ORC_FIXED_SCALE));
            stress[k] += inv_v * vt[k];
        }
    if (vel && mass)
        for (int k = 0; k < 6; k++)
            stress[k] += inv_v * (float)((double)w[6 + k] * (1.0 / ORC_FIXED_SCALE));
    if (sums) memcpy(sums, w, sizeof(w));
}

/* ------------------------------------------------------------------------ */
/* PME exclusion correction (PM_force.cpp:1743-1814), factor 1 (one rank)   */
/* ------------------------------------------------------------------------ */
/* For j in excl(i): dr = pbc(r_j - r_i), x = beta r,
 *   f = qq (erf(x) - 2x/sqrt(pi) e^{-x^2}) / r^3 * dr  (force on i, -f on j)
 *   e[i] -= qq erf(x)/r;   virial[i] -= f (x) dr                            */
ORC_EXPORT void orc_excl_f64(int n, const float* crd, const float* charge,
                             const float* cell, const float* rcell, float beta,
                             const int* excl_start, const int* excl_num,
                             const int* excl_list, double* frc, double* e,
                             double* virial)
{
    double b = beta;
    for (int i = 0; i < n; i++)
    {
        for (int k = excl_start[i]; k < excl_start[i] + excl_num[i]; k++)
        {
            int j = excl_list[k];
            double d[3];
            pbc_disp_d(crd + 3 * j, crd + 3 * i, cell, rcell, d);
            double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            double r = sqrt(r2);
            double x = b * r;
            double qq = (double)charge[i] * charge[j];
            double h = erf(x) - 2.0 / sqrt(ORC_PI_D) * x * exp(-x * x);
            double fa = qq * h / (r2 * r);
            double f[3] = {fa * d[0], fa * d[1], fa * d[2]};
            for (int c = 0; c < 3; c++)
            {
                frc[3 * i + c] += f[c];
                frc[3 * j + c] -= f[c];
            }
            if (e) e[i] -= qq * erf(x) / r;
            if (virial)
            {
                double* v = virial + 6 * i;
                v[0] -= f[0] * d[0];
                v[1] -= f[0] * d[1] + f[1] * d[0];
                v[2] -= f[1] * d[1];
                v[3] -= f[0] * d[2] + f[2] * d[0];
                v[4] -= f[1] * d[2] + f[2] * d[1];
                v[5] -= f[2] * d[2];
            }
        }
    }
}

ORC_EXPORT void orc_excl_twin(int n, const float* crd, const float* charge,
                              const float* cell, const float* rcell, float beta,
                              const int* excl_start, const int* excl_num,
                              const int* excl_list, int64_t* frc, int64_t* e,
                              int64_t* vir)
{
    int nt = orc_threads();
    int64_t* part = (int64_t*)calloc((size_t)nt * 3 * n + 1, sizeof(int64_t));
#pragma omp parallel num_threads(nt)
    {
    int64_t* f = part + (size_t)orc_thread_id() * 3 * n;
#pragma omp for schedule(dynamic, 256)
    for (int i = 0; i < n; i++)
    {
        float qi = charge[i];
        for (int k = excl_start[i]; k < excl_start[i] + excl_num[i]; k++)
        {
            int j = excl_list[k];
            float d[3];
            pbc_disp_f(crd + 3 * j, crd + 3 * i, cell, rcell, d);
            /* FP64 from the float displacement */
            double ddx = d[0], ddy = d[1], ddz = d[2];
            double r2d = (ddx * ddx + ddy * ddy) + ddz * ddz;
            double rd = sqrt(r2d);
            double rid = 1.0 / rd;
            double gd;
            double erfd = 1.0 - twin_erfc_parts_d((double)beta * rd, &gd);
            double hd = erfd - gd;
            double qq = (double)qi * (double)charge[j];
            double fa = (qq * hd) * ((rid * rid) * rid);
            double fx = fa * ddx, fy = fa * ddy, fz = fa * ddz;
            int64_t ix = to_fixed_d(fx), iy = to_fixed_d(fy), iz = to_fixed_d(fz);
            f[i] += ix;
            f[n + i] += iy;
            f[2 * n + i] += iz;
            f[j] -= ix;
            f[n + j] -= iy;
            f[2 * n + j] -= iz;
            if (e) e[i] += to_fixed_d(-((qq * erfd) * rid));
            if (vir)
            {
                vir[0 * (size_t)n + i] += to_fixed_d(-(fx * ddx));
                vir[1 * (size_t)n + i] += to_fixed_d(-(fx * ddy + fy * ddx));
                vir[2 * (size_t)n + i] += to_fixed_d(-(fy * ddy));
                vir[3 * (size_t)n + i] += to_fixed_d(-(fx * ddz + fz * ddx));
                vir[4 * (size_t)n + i] += to_fixed_d(-(fy * ddz + fz * ddy));
                vir[5 * (size_t)n + i] += to_fixed_d(-(fz * ddz));
            }
        }
    }
    }
    reduce_parts(part, nt, 3 * (size_t)n, frc);
    free(part);
}

/* ------------------------------------------------------------------------ */
/* FFT (double): recursive mixed-radix Cooley-Tukey, any length              */
/* ------------------------------------------------------------------------ */
typedef struct
{
    double re, im;
} cplx;

/* out[k] = sum_j in[j*stride] exp(sign 2 pi i j k / n); root[t] holds
 * exp(sign 2 pi i t / n0) for the top-level length n0 (n divides n0). */
static void fft_rec(const cplx* in, int stride, cplx* out, int n, int n0,
                    const cplx* root, cplx* scratch)
{
    if (n == 1)
    {
        out[0] = in[0];
        return;
    }
    int p = 2;
    while (n % p) p++;
    int m = n / p;
    int step = n0 / n;
    /* sub-transforms of the p decimated sequences, written to scratch */
    for (int r = 0; r < p; r++)
        fft_rec(in + r * stride, stride * p, scratch + r * m, m, n0, root,
                out /* out used as scratch by the child */);
    for (int k = 0; k < n; k++)
    {
        double sr = 0, si = 0;
        for (int r = 0; r < p; r++)
        {
            cplx v = scratch[r * m + (k % m)];
            cplx w = root[((size_t)r * k % n) * step];
            sr += v.re * w.re - v.im * w.im;
            si += v.re * w.im + v.im * w.re;
        }
        out[k].re = sr;
        out[k].im = si;
    }
}

static void fft1d(cplx* data, int n, int stride, const cplx* root, cplx* work)
{
    cplx* in = work;
    cplx* out = work + n;
    cplx* scratch = work + 2 * n;
    for (int j = 0; j < n; j++) in[j] = data[(size_t)j * stride];
    fft_rec(in, 1, out, n, n, root, scratch);
    for (int j = 0; j < n; j++) data[(size_t)j * stride] = out[j];
}

static cplx* make_roots(int n, int sign)
{
    cplx* r = (cplx*)malloc(sizeof(cplx) * n);
    for (int t = 0; t < n; t++)
    {
        double a = sign * 2.0 * ORC_PI_D * (double)t / n;
        r[t].re = cos(a);
        r[t].im = sin(a);
    }
    return r;
}

/* In-place complex 3-D DFT of a [n0][n1][n2] array (last index fastest),
 * exponent sign `sign` (-1 forward as in FFTW), unnormalised. */
ORC_EXPORT void orc_fft3d_f64(double* data_ri, int n0, int n1, int n2, int sign)
{
    cplx* a = (cplx*)data_ri;
    int nmax = n0 > n1 ? n0 : n1;
    nmax = nmax > n2 ? nmax : n2;
    int nt = orc_threads();
    cplx* works = (cplx*)malloc(sizeof(cplx) * 3 * nmax * (size_t)nt);
    /* each line is transformed by one thread in the same operation order */
    cplx* r2 = make_roots(n2, sign);
#pragma omp parallel for num_threads(nt) schedule(static)
    for (int ij = 0; ij < n0 * n1; ij++)
        fft1d(a + (size_t)ij * n2, n2, 1, r2, works + (size_t)orc_thread_id() * 3 * nmax);
    free(r2);
    cplx* r1 = make_roots(n1, sign);
#pragma omp parallel for num_threads(nt) schedule(static)
    for (int ik = 0; ik < n0 * n2; ik++)
    {
        int i = ik / n2, k = ik - i * n2;
        fft1d(a + (size_t)i * n1 * n2 + k, n1, n2, r1, works + (size_t)orc_thread_id() * 3 * nmax);
    }
    free(r1);
    cplx* r0 = make_roots(n0, sign);
#pragma omp parallel for num_threads(nt) schedule(static)
    for (int jk = 0; jk < n1 * n2; jk++)
        fft1d(a + (size_t)jk, n0, n1 * n2, r0, works + (size_t)orc_thread_id() * 3 * nmax);
    free(r0);
    free(works);
}

/* Naive O(N^2) DFT of the same array, to pin orc_fft3d_f64 on small grids. */
ORC_EXPORT void orc_dft3d_naive_f64(const double* in_ri, double* out_ri, int n0,
                                    int n1, int n2, int sign)
{
    const cplx* in = (const cplx*)in_ri;
    cplx* out = (cplx*)out_ri;
    for (int k0 = 0; k0 < n0; k0++)
        for (int k1 = 0; k1 < n1; k1++)
            for (int k2 = 0; k2 < n2; k2++)
            {
                double sr = 0, si = 0;
                for (int x0 = 0; x0 < n0; x0++)
                    for (int x1 = 0; x1 < n1; x1++)
                        for (int x2 = 0; x2 < n2; x2++)
                        {
                            double a = sign * 2.0 * ORC_PI_D *
                                       ((double)k0 * x0 / n0 + (double)k1 * x1 / n1 +
                                        (double)k2 * x2 / n2);
                            cplx v = in[((size_t)x0 * n1 + x1) * n2 + x2];
                            sr += v.re * cos(a) - v.im * sin(a);
                            si += v.re * sin(a) + v.im * cos(a);
                        }
                out[((size_t)k0 * n1 + k1) * n2 + k2].re = sr;
                out[((size_t)k0 * n1 + k1) * n2 + k2].im = si;
            }
}

/* ------------------------------------------------------------------------ */
/* PME reciprocal (PM_force.cpp:697-916, :1021-1165)                         */
/* ------------------------------------------------------------------------ */
/* B-spline coefficient tables, PM_force.cpp:18-24 */
static const double BS_MA[4] = {1.0 / 6.0, -0.5, 0.5, -1.0 / 6.0};
static const double BS_MB[4] = {0, 0.5, -1, 0.5};
static const double BS_MC[4] = {0, 0.5, 0, -0.5};
static const double BS_MD[4] = {0, 1.0 / 6.0, 4.0 / 6.0, 1.0 / 6.0};
static const double BS_DMA[4] = {0.5, -1.5, 1.5, -0.5};
static const double BS_DMB[4] = {0, 1, -2, 1};
static const double BS_DMC[4] = {0, 0.5, 0, -0.5};

/* PM_force.cpp:706-737: base grid point u and fraction fr, in float exactly
 * as the reference (fractional coordinate crd*rcell wrapped to [0,1)). */
static void atom_near_f(const float* r, const float* rcell, int fx, int fy,
                        int fz, int* u, float* fr)
{
    float f[3];
    f[0] = (r[0] * rcell[0] + r[1] * rcell[1]) + r[2] * rcell[3];
    f[1] = r[1] * rcell[2] + r[2] * rcell[4];
    f[2] = r[2] * rcell[5];
    for (int c = 0; c < 3; c++) f[c] = f[c] - floorf(f[c]);
    if (!isfinite(f[0]) || !isfinite(f[1]) || !isfinite(f[2]))
        f[0] = f[1] = f[2] = 0.0f;
    int dims[3] = {fx, fy, fz};
    for (int c = 0; c < 3; c++)
    {
        float s = f[c] * (float)dims[c];
        int t = (int)s;
        t = t < 0 ? 0 : (t < dims[c] ? t : dims[c] - 1);
        float q = s - (float)t;
        fr[c] = q - floorf(q);
        u[c] = t;
    }
}

/* the same point and fractions with every step in double (the FP64
 * evaluation of the library: snb_f64_force) */
static void atom_near_d(const float* r, const float* rcell, int fx, int fy, int fz, int* u,
                        double* fr)
{
    double f[3];
    f[0] = ((double)r[0] * rcell[0] + (double)r[1] * rcell[1]) + (double)r[2] * rcell[3];
    f[1] = (double)r[1] * rcell[2] + (double)r[2] * rcell[4];
    f[2] = (double)r[2] * rcell[5];
    for (int c = 0; c < 3; c++) f[c] = f[c] - floor(f[c]);
    if (!isfinite(f[0]) || !isfinite(f[1]) || !isfinite(f[2])) f[0] = f[1] = f[2] = 0.0;
    int dims[3] = {fx, fy, fz};
    for (int c = 0; c < 3; c++)
    {
        double s = f[c] * (double)dims[c];
        int t = (int)s;
        t = t < 0 ? 0 : (t < dims[c] ? t : dims[c] - 1);
        fr[c] = s - (double)t;
        u[c] = t;
    }
}

static inline int wrapi(int a, int n)
{
    a %= n;
    return a < 0 ? a + n : a;
}

/* FP64 reciprocal pipeline.  Outputs: frc += update_interval * F (double[3n]),
 * *e_rec = 0.5 sum Q phi, vir[6] = the reference's PME_Sum_Virial result
 * (-sum_half w |FQ|^2 VirialBC with w = 0.5 on kz = 0 and kz = fftz/2),
 * and optionally the real-space grids Q and phi (double[Nall]). */
static void pme_f64_impl(int n, const float* crd, const float* charge, const float* rcell,
                         int fx, int fy, int fz, float beta, double volume, double* frc,
                         double* e_rec, double* vir, double* q_out, double* phi_out,
                         int dfrac);

ORC_EXPORT void orc_pme_f64(int n, const float* crd, const float* charge,
                            const float* rcell, int fx, int fy, int fz,
                            float beta, double volume, double* frc,
                            double* e_rec, double* vir, double* q_out,
                            double* phi_out)
{
    pme_f64_impl(n, crd, charge, rcell, fx, fy, fz, beta, volume, frc, e_rec, vir, q_out,
                 phi_out, 0);
}

/* As orc_pme_f64 with every step in double: the atoms' grid points and
 * spline fractions (instead of the reference's float PME_Atom_Near) and the
 * wave vectors of the influence function and the virial (the kernels take
 * kz times rcell's last entry as a float product). */
ORC_EXPORT void orc_pme_f64_dfrac(int n, const float* crd, const float* charge,
                                  const float* rcell, int fx, int fy, int fz, float beta,
                                  double volume, double* frc, double* e_rec, double* vir)
{
    pme_f64_impl(n, crd, charge, rcell, fx, fy, fz, beta, volume, frc, e_rec, vir, NULL, NULL,
                 1);
}

static void pme_f64_impl(int n, const float* crd, const float* charge, const float* rcell,
                         int fx, int fy, int fz, float beta, double volume, double* frc,
                         double* e_rec, double* vir, double* q_out, double* phi_out,
                         int dfrac)
{
    size_t nall = (size_t)fx * fy * fz;
    cplx* grid = (cplx*)calloc(nall, sizeof(cplx));
    int* u = (int*)malloc(sizeof(int) * 3 * n);
    double* frd = (double*)malloc(sizeof(double) * 3 * n);
    /* spread into per-thread grids, summed in thread order afterwards */
    int nt = orc_threads();
    double* tg = (double*)calloc((size_t)nt * nall, sizeof(double));
#pragma omp parallel for num_threads(nt) schedule(static)
    for (int i = 0; i < n; i++)
    {
        double* gq = tg + (size_t)orc_thread_id() * nall;
        if (dfrac)
            atom_near_d(crd + 3 * i, rcell, fx, fy, fz, u + 3 * i, frd + 3 * i);
        else
        {
            float frf[3];
            atom_near_f(crd + 3 * i, rcell, fx, fy, fz, u + 3 * i, frf);
            for (int c = 0; c < 3; c++) frd[3 * i + c] = frf[c];
        }
        double w[3][4];
        for (int c = 0; c < 3; c++)
        {
            double t = frd[3 * i + c];
            for (int k = 0; k < 4; k++)
                w[c][k] = ((BS_MA[k] * t + BS_MB[k]) * t + BS_MC[k]) * t + BS_MD[k];
        }
        for (int kx = 0; kx < 4; kx++)
            for (int ky = 0; ky < 4; ky++)
                for (int kz = 0; kz < 4; kz++)
                {
                    size_t idx = ((size_t)wrapi(u[3 * i] - kx, fx) * fy +
                                  wrapi(u[3 * i + 1] - ky, fy)) * fz +
                                 wrapi(u[3 * i + 2] - kz, fz);
                    gq[idx] += charge[i] * w[0][kx] * w[1][ky] * w[2][kz];
                }
    }
#pragma omp parallel for num_threads(nt) schedule(static)
    for (size_t p = 0; p < nall; p++)
    {
        double acc = 0.0;
        for (int t = 0; t < nt; t++) acc += tg[(size_t)t * nall + p];
        grid[p].re = acc;
    }
    free(tg);
    double* q = (double*)malloc(sizeof(double) * nall);
    for (size_t p = 0; p < nall; p++) q[p] = grid[p].re;
    orc_fft3d_f64((double*)grid, fx, fy, fz, -1);
    /* convolution over the full spectrum with BC from the half spectrum */
    int fc = fz / 2 + 1;
    double* bc = (double*)malloc(sizeof(double) * (size_t)fx * fy * fc);
    bc_build(fx, fy, fz, rcell, beta, volume, bc, dfrac);
    double mpre = ORC_PI_D * ORC_PI_D / ((double)beta * beta);
    double v0 = 0, v1 = 0, v2 = 0, v3 = 0, v4 = 0, v5 = 0;
#pragma omp parallel for num_threads(nt) schedule(static) \
    reduction(+ : v0, v1, v2, v3, v4, v5)
    for (int kx = 0; kx < fx; kx++)
        for (int ky = 0; ky < fy; ky++)
            for (int kz = 0; kz < fz; kz++)
            {
                size_t idx = ((size_t)kx * fy + ky) * fz + kz;
                int hz = kz, hx = kx, hy = ky;
                if (kz >= fc)
                { /* conjugate partner in the half spectrum */
                    hz = fz - kz;
                    hx = (fx - kx) % fx;
                    hy = (fy - ky) % fy;
                }
                double b = bc[((size_t)hx * fy + hy) * fc + hz];
                if (kz < fc && b != 0.0)
                {
                    /* PM_force.cpp:1021-1065 virial over the half spectrum */
                    double w = (kz == 0 || kz == fc - 1) ? 0.5 : 1.0;
                    double a2 = grid[idx].re * grid[idx].re + grid[idx].im * grid[idx].im;
                    double kxr = kx > fx / 2 ? kx - fx : kx;
                    double kyr = ky > fy / 2 ? ky - fy : ky;
                    double mx = kxr * rcell[0];
                    double my = kxr * rcell[1] + kyr * rcell[2];
                    double mz = dfrac
                                    ? kxr * rcell[3] + kyr * rcell[4] + (double)kz * rcell[5]
                                    : kxr * rcell[3] + kyr * rcell[4] + kz * rcell[5];
                    double msq = mx * mx + my * my + mz * mz;
                    double cf = 2.0 / msq * (1.0 + mpre * msq);
                    double hb = 0.5 * b * w * a2;
                    v0 -= hb * (1.0 - cf * mx * mx);
                    v1 -= hb * (0.0 - cf * my * mx);
                    v2 -= hb * (1.0 - cf * my * my);
                    v3 -= hb * (0.0 - cf * mz * mx);
                    v4 -= hb * (0.0 - cf * mz * my);
                    v5 -= hb * (1.0 - cf * mz * mz);
                }
                grid[idx].re *= b;
                grid[idx].im *= b;
            }
    double v[6] = {v0, v1, v2, v3, v4, v5};
    orc_fft3d_f64((double*)grid, fx, fy, fz, +1);
    double e = 0;
    for (size_t p = 0; p < nall; p++) e += q[p] * grid[p].re;
    *e_rec = 0.5 * e;
    if (vir)
        for (int c = 0; c < 6; c++) vir[c] = v[c];
    if (q_out) memcpy(q_out, q, sizeof(double) * nall);
    if (phi_out)
        for (size_t p = 0; p < nall; p++) phi_out[p] = grid[p].re;
    /* gather (PM_force.cpp:826-890) */
#pragma omp parallel for num_threads(nt) schedule(static)
    for (int i = 0; i < n; i++)
    {
        double w[3][4], dw[3][4];
        for (int c = 0; c < 3; c++)
        {
            double t = frd[3 * i + c];
            for (int k = 0; k < 4; k++)
            {
                w[c][k] = ((BS_MA[k] * t + BS_MB[k]) * t + BS_MC[k]) * t + BS_MD[k];
                dw[c][k] = (BS_DMA[k] * t + BS_DMB[k]) * t + BS_DMC[k];
            }
        }
        double s[3] = {0, 0, 0};
        for (int kx = 0; kx < 4; kx++)
            for (int ky = 0; ky < 4; ky++)
                for (int kz = 0; kz < 4; kz++)
                {
                    size_t idx = ((size_t)wrapi(u[3 * i] - kx, fx) * fy +
                                  wrapi(u[3 * i + 1] - ky, fy)) * fz +
                                 wrapi(u[3 * i + 2] - kz, fz);
                    double ph = grid[idx].re;
                    s[0] += ph * dw[0][kx] * w[1][ky] * w[2][kz];
                    s[1] += ph * w[0][kx] * dw[1][ky] * w[2][kz];
                    s[2] += ph * w[0][kx] * w[1][ky] * dw[2][kz];
                }
        double t0 = s[0] * fx, t1 = s[1] * fy, t2 = s[2] * fz;
        double q_ = -(double)charge[i];
        frc[3 * i + 0] += q_ * (t0 * rcell[0]);
        frc[3 * i + 1] += q_ * (t0 * rcell[1] + t1 * rcell[2]);
        frc[3 * i + 2] += q_ * (t0 * rcell[3] + t1 * rcell[4] + t2 * rcell[5]);
    }
    free(bc);
    free(q);
    free(grid);
    free(u);
    free(frd);
}

/* FP32 twin of the spread: the Q grid as int64 fixed point (sum of
 * round(((q*wx)*wy)*wz * 2^32) over atoms), weights evaluated in float with
 * the reference's op order (PM_force.cpp:776-805). */
static inline float bs_w_f(int k, float t)
{
    static const float MA[4] = {1.0f / 6.0f, -0.5f, 0.5f, -1.0f / 6.0f};
    static const float MB[4] = {0.0f, 0.5f, -1.0f, 0.5f};
    static const float MC[4] = {0.0f, 0.5f, 0.0f, -0.5f};
    static const float MD[4] = {0.0f, 1.0f / 6.0f, 4.0f / 6.0f, 1.0f / 6.0f};
    float t2 = t * t;
    return ((MA[k] * t) * t2 + MB[k] * t2) + MC[k] * t + MD[k];
}
static inline float bs_dw_f(int k, float t)
{
    static const float DMA[4] = {0.5f, -1.5f, 1.5f, -0.5f};
    static const float DMB[4] = {0.0f, 1.0f, -2.0f, 1.0f};
    static const float DMC[4] = {0.0f, 0.5f, 0.0f, -0.5f};
    float t2 = t * t;
    return (DMA[k] * t2 + DMB[k] * t) + DMC[k];
}

ORC_EXPORT void orc_pme_spread_twin(int n, const float* crd, const float* charge,
                                    const float* rcell, int fx, int fy, int fz,
                                    int64_t* qgrid)
{
    for (int i = 0; i < n; i++)
    {
        int u[3];
        float fr[3];
        atom_near_f(crd + 3 * i, rcell, fx, fy, fz, u, fr);
        float w[3][4];
        for (int c = 0; c < 3; c++)
            for (int k = 0; k < 4; k++) w[c][k] = bs_w_f(k, fr[c]);
        for (int kx = 0; kx < 4; kx++)
            for (int ky = 0; ky < 4; ky++)
                for (int kz = 0; kz < 4; kz++)
                {
                    size_t idx = ((size_t)wrapi(u[0] - kx, fx) * fy +
                                  wrapi(u[1] - ky, fy)) * fz + wrapi(u[2] - kz, fz);
                    float v = ((charge[i] * w[0][kx]) * w[1][ky]) * w[2][kz];
                    qgrid[idx] += to_fixed(v);
                }
    }
}

/* FP32 twin of the gather: given the float potential grid phi (the C2R
 * output), per atom sum in stencil order kx, ky, kz (outer to inner)
 *   s_x = fma(phi, (dwx*wy)*wz, s_x) etc., then t = s * fft dims,
 *   F = -q * MultiplyTranspose(t, rcell) * update_interval, to fixed. */
ORC_EXPORT void orc_pme_gather_twin(int n, const float* crd, const float* charge,
                                    const float* rcell, int fx, int fy, int fz,
                                    const float* phi, float update_interval,
                                    int64_t* frc)
{
    for (int i = 0; i < n; i++)
    {
        int u[3];
        float fr[3];
        atom_near_f(crd + 3 * i, rcell, fx, fy, fz, u, fr);
        float w[3][4], dw[3][4];
        for (int c = 0; c < 3; c++)
            for (int k = 0; k < 4; k++)
            {
                w[c][k] = bs_w_f(k, fr[c]);
                dw[c][k] = bs_dw_f(k, fr[c]);
            }
        float sx = 0.0f, sy = 0.0f, sz = 0.0f;
        for (int kx = 0; kx < 4; kx++)
            for (int ky = 0; ky < 4; ky++)
                for (int kz = 0; kz < 4; kz++)
                {
                    size_t idx = ((size_t)wrapi(u[0] - kx, fx) * fy +
                                  wrapi(u[1] - ky, fy)) * fz + wrapi(u[2] - kz, fz);
                    float ph = phi[idx];
                    sx = fmaf(ph, (dw[0][kx] * w[1][ky]) * w[2][kz], sx);
                    sy = fmaf(ph, (w[0][kx] * dw[1][ky]) * w[2][kz], sy);
                    sz = fmaf(ph, (w[0][kx] * w[1][ky]) * dw[2][kz], sz);
                }
        float tx = sx * (float)fx, ty = sy * (float)fy, tz = sz * (float)fz;
        float mq = -charge[i] * update_interval;
        float gx = tx * rcell[0];
        float gy = tx * rcell[1] + ty * rcell[2];
        float gz = (tx * rcell[3] + ty * rcell[4]) + tz * rcell[5];
        frc[i] += to_fixed(mq * gx);
        frc[n + i] += to_fixed(mq * gy);
        frc[2 * n + i] += to_fixed(mq * gz);
    }
}

/* ------------------------------------------------------------------------ */
/* exact Ewald sum (validation of the PME layer on small boxes)              */
/* ------------------------------------------------------------------------ */
/* Orthorhombic box L[3]; E = E_dir + E_rec + E_self with
 * E_dir = sum_{i<j, images n, |r|<rcut} qq erfc(beta r)/r (no exclusions),
 * E_rec = 1/(2 pi V) sum_{m != 0, |m|<=kmax} exp(-pi^2 m^2/beta^2)/m^2 |S(m)|^2,
 * E_self = -beta/sqrt(pi) sum q^2.  Forces likewise.  All double. */
ORC_EXPORT void orc_ewald_f64(int n, const double* crd, const double* q,
                              const double* L, double beta, double rcut,
                              int kmax, double* frc, double* e_dir,
                              double* e_rec, double* e_self)
{
    double V = L[0] * L[1] * L[2];
    double ed = 0, er = 0, es = 0;
    int nim[3];
    for (int c = 0; c < 3; c++) nim[c] = (int)ceil(rcut / L[c]) + 1;
    for (int i = 0; i < n; i++)
    {
        es -= beta / sqrt(ORC_PI_D) * q[i] * q[i];
        for (int j = 0; j < n; j++)
            for (int a = -nim[0]; a <= nim[0]; a++)
                for (int b = -nim[1]; b <= nim[1]; b++)
                    for (int c = -nim[2]; c <= nim[2]; c++)
                    {
                        if (i == j && a == 0 && b == 0 && c == 0) continue;
                        double d[3] = {crd[3 * j] - crd[3 * i] + a * L[0],
                                       crd[3 * j + 1] - crd[3 * i + 1] + b * L[1],
                                       crd[3 * j + 2] - crd[3 * i + 2] + c * L[2]};
                        double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                        double r = sqrt(r2);
                        if (r >= rcut) continue;
                        double x = beta * r;
                        ed += 0.5 * q[i] * q[j] * erfc(x) / r;
                        double fa = -q[i] * q[j] *
                                    (2.0 / sqrt(ORC_PI_D) * x * exp(-x * x) + erfc(x)) /
                                    (r2 * r);
                        for (int t = 0; t < 3; t++) frc[3 * i + t] += fa * d[t];
                    }
    }
    for (int a = -kmax; a <= kmax; a++)
        for (int b = -kmax; b <= kmax; b++)
            for (int c = -kmax; c <= kmax; c++)
            {
                if (a == 0 && b == 0 && c == 0) continue;
                double m[3] = {a / L[0], b / L[1], c / L[2]};
                double m2 = m[0] * m[0] + m[1] * m[1] + m[2] * m[2];
                double pref = exp(-ORC_PI_D * ORC_PI_D * m2 / (beta * beta)) /
                              (2.0 * ORC_PI_D * V * m2);
                double sr = 0, si = 0;
                for (int i = 0; i < n; i++)
                {
                    double ph = 2.0 * ORC_PI_D *
                                (m[0] * crd[3 * i] + m[1] * crd[3 * i + 1] +
                                 m[2] * crd[3 * i + 2]);
                    sr += q[i] * cos(ph);
                    si += q[i] * sin(ph);
                }
                er += pref * (sr * sr + si * si);
                for (int i = 0; i < n; i++)
                {
                    double ph = 2.0 * ORC_PI_D *
                                (m[0] * crd[3 * i] + m[1] * crd[3 * i + 1] +
                                 m[2] * crd[3 * i + 2]);
                    /* F_i = -dE/dr_i = pref * 2 * q_i * 2 pi m (sin ph Sr - cos ph Si) */
                    double g = pref * 2.0 * q[i] * 2.0 * ORC_PI_D *
                               (sin(ph) * sr - cos(ph) * si);
                    for (int t = 0; t < 3; t++) frc[3 * i + t] += g * m[t];
                }
            }
    *e_dir = ed;
    *e_rec = er;
    *e_self = es;
}

/* ------------------------------------------------------------------------ */
/* LJ long-range correction (Lennard_Jones_force.cpp:146-184, :246-272)     */
/* ------------------------------------------------------------------------ */
/* sum over all ordered pairs (i, j), i == j included, of B_ij, via a type
 * histogram; returns the double total (the reference rounds it to float). */
ORC_EXPORT double orc_lj_c6_total(int n, const int* type, int ntypes,
                                  const float* lj_b)
{
    double* hist = (double*)calloc(ntypes, sizeof(double));
    for (int i = 0; i < n; i++) hist[type[i]] += 1.0;
    double s = 0;
    for (int a = 0; a < ntypes; a++)
        for (int b = 0; b < ntypes; b++)
            s += hist[a] * hist[b] * lj_b[lj_pair_index(a, b)];
    free(hist);
    return s;
}

/* exposes the twin's scalar math for unit tests */
ORC_EXPORT void orc_twin_erfc_parts(double x, double* erfc_out, double* g_out)
{
    *erfc_out = twin_erfc_parts_d(x, g_out);
}
ORC_EXPORT double orc_twin_exp_d(double y) { return twin_exp_d(y); }
ORC_EXPORT float orc_twin_rsqrtf(float x) { return twin_rsqrtf(x); }
/* h(x) of the pair kernels' Coulomb table for (beta, cutoff) */
ORC_EXPORT float orc_twin_coul_h(float x, float beta, float cutoff)
{
    float tab[4 * TWIN_COUL_N];
    float scale = twin_coul_table(beta, cutoff, tab);
    return twin_coul_h(x, scale, tab);
}
ORC_EXPORT int64_t orc_to_fixed(float f) { return to_fixed(f); }

/* ------------------------------------------------------------------------ */
/* Integration: leap-frog (MD_core/nve.hpp:3-48), middle-scheme Langevin     */
/* (thermostat/Middle_Langevin_MD.cpp:5-66) and SETTLE (constrain/settle.cpp) */
/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC 2011): ten rounds with the
 * published multipliers 0xD2511F53 / 0xCD9E8D57 and Weyl increments
 * 0x9E3779B9 / 0xBB67AE85.  Pinned by the Random123 known-answer vectors in
 * tests/test_oracle_pins.py. */
static void twin_philox(uint32_t c[4], uint32_t k0, uint32_t k1)
{
    for (int r = 0; r < 10; r++)
    {
        uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        c[1] = (uint32_t)p1;
        c[3] = (uint32_t)p0;
        c[0] = n0;
        c[2] = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

static inline float twin_uniform01(uint32_t x)
{
    return (float)(x >> 8) * 5.9604644775390625e-08f + 2.98023223876953125e-08f;
}

/* ln u, u in (0, 1]: u = m 2^e, m in [sqrt(1/2), sqrt(2)); ln m = 2 atanh(s),
 * s = (m - 1)/(m + 1), odd series to s^9 */
static inline float twin_log_unit(float u)
{
    uint32_t b = bits_from_f(u);
    int e = (int)(b >> 23) - 127;
    float m = f_from_bits((b & 0x007FFFFFu) | 0x3F800000u);
    if (m > 1.41421356f)
    {
        m = m * 0.5f;
        e += 1;
    }
    float s = (m - 1.0f) / (m + 1.0f);
    float s2 = s * s;
    float p = 1.0f / 9.0f;
    p = fmaf(p, s2, 1.0f / 7.0f);
    p = fmaf(p, s2, 0.2f);
    p = fmaf(p, s2, 1.0f / 3.0f);
    p = fmaf(p, s2, 1.0f);
    return fmaf((float)e, 0.693147182f, (2.0f * s) * p);
}

/* sin, cos of 2 pi u: reduction to +-pi/4 by quarter turns, Taylor to a^9 */
static inline void twin_sincos_2pi(float u, float* so, float* co)
{
    float x = u - rintf(u);
    float jf = rintf(4.0f * x);
    float a = (x - 0.25f * jf) * 6.28318548f;
    float a2 = a * a;
    float sp = 1.0f / 362880.0f;
    sp = fmaf(sp, a2, -1.0f / 5040.0f);
    sp = fmaf(sp, a2, 1.0f / 120.0f);
    sp = fmaf(sp, a2, -1.0f / 6.0f);
    float sa = fmaf(a * a2, sp, a);
    float cp = 1.0f / 40320.0f;
    cp = fmaf(cp, a2, -1.0f / 720.0f);
    cp = fmaf(cp, a2, 1.0f / 24.0f);
    cp = fmaf(cp, a2, -0.5f);
    float ca = fmaf(a2, cp, 1.0f);
    int j = (int)jf & 3;
    float s = sa, c = ca;
    if (j == 1) { s = ca; c = -sa; }
    else if (j == 2) { s = -sa; c = -ca; }
    else if (j == 3) { s = -ca; c = sa; }
    *so = s;
    *co = c;
}

static inline void twin_box_muller(uint32_t x0, uint32_t x1, float* n0, float* n1)
{
    float r = sqrtf(-2.0f * twin_log_unit(twin_uniform01(x0)));
    float s, c;
    twin_sincos_2pi(twin_uniform01(x1), &s, &c);
    *n0 = r * c;
    *n1 = r * s;
}

#define TWIN_LANGEVIN_KEY1 0x53504F4Eu

ORC_EXPORT void orc_philox4x32_10(const uint32_t* ctr, uint32_t k0, uint32_t k1, uint32_t* out)
{
    uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    twin_philox(c, k0, k1);
    for (int k = 0; k < 4; k++) out[k] = c[k];
}

/* normal numbers the thermostat draws: for atom i at stream position c, the
 * four Box-Muller outputs of Philox({i, c, 0, 0}, {seed, key1}) */
ORC_EXPORT void orc_langevin_normals(int n, uint32_t seed, const uint32_t* counter, float* out4)
{
    for (int i = 0; i < n; i++)
    {
        uint32_t c[4] = {(uint32_t)i, counter[i], 0u, 0u};
        twin_philox(c, seed, TWIN_LANGEVIN_KEY1);
        twin_box_muller(c[0], c[1], out4 + 4 * i, out4 + 4 * i + 1);
        twin_box_muller(c[2], c[3], out4 + 4 * i + 2, out4 + 4 * i + 3);
    }
}

static inline void twin_cap_speed(float* v, float vmax)
{
    float v2 = (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2];
    float f = v2 > 0.0f ? fminf(1.0f, vmax * twin_rsqrtf(v2)) : 1.0f;
    for (int k = 0; k < 3; k++) v[k] = f * v[k];
}

/* One middle-scheme Langevin step (counter[] advanced), or with seed < 0 a
 * leap-frog step with inv_mass and dt only. */
ORC_EXPORT void orc_langevin_twin(int n, const float* inv_mass, const float* sqrt_mass,
                                  uint32_t* counter, uint32_t seed, const float* frc, float* vel,
                                  float* crd, float dt, float half_dt, float exp_gamma,
                                  float vmax)
{
    for (int i = 0; i < n; i++)
    {
        uint32_t c[4] = {(uint32_t)i, counter[i], 0u, 0u};
        counter[i] = c[1] + 1u;
        twin_philox(c, seed, TWIN_LANGEVIN_KEY1);
        float r[4];
        twin_box_muller(c[0], c[1], r, r + 1);
        twin_box_muller(c[2], c[3], r + 2, r + 3);
        float v[3], x1[3];
        for (int k = 0; k < 3; k++)
        {
            float a = inv_mass[i] * frc[3 * i + k];
            v[k] = vel[3 * i + k] + dt * a;
            x1[k] = crd[3 * i + k] + half_dt * v[k];
            v[k] = exp_gamma * v[k] + sqrt_mass[i] * r[k];
        }
        if (vmax > 0.0f) twin_cap_speed(v, vmax);
        for (int k = 0; k < 3; k++)
        {
            vel[3 * i + k] = v[k];
            crd[3 * i + k] = x1[k] + half_dt * v[k];
        }
    }
}

ORC_EXPORT void orc_leap_frog_twin(int n, const float* inv_mass, const float* frc, float* vel,
                                   float* crd, float dt, float vmax)
{
    for (int i = 0; i < n; i++)
    {
        float v[3];
        for (int k = 0; k < 3; k++) v[k] = vel[3 * i + k] + dt * (inv_mass[i] * frc[3 * i + k]);
        if (vmax > 0.0f) twin_cap_speed(v, vmax);
        for (int k = 0; k < 3; k++)
        {
            vel[3 * i + k] = v[k];
            crd[3 * i + k] = crd[3 * i + k] + dt * v[k];
        }
    }
}

/* Get_Rabcde_From_SSS (settle.cpp:50-98), float */
ORC_EXPORT void orc_settle_frame(float rAB, float rAC, float rBC, float mA, float mB, float mC,
                                 float* out)
{
    float mt = (mA + mB) + mC;
    float Ax = 0.0f, Ay = 0.0f, Bx = -rAB, By = 0.0f;
    float ct = ((rBC * rBC - rAC * rAC) - rAB * rAB) / ((2.0f * rAC) * rAB);
    float Cx = rAC * ct;
    float st = sqrtf(1.0f - ct * ct);
    float Cy = rAC * st;
    float Ox = (Bx * mB + Cx * mC) / mt;
    float Oy = (Cy * mC) / mt;
    Ax -= Ox; Ay -= Oy; Bx -= Ox; By -= Oy; Cx -= Ox; Cy -= Oy;
    ct = 1.0f / sqrtf(1.0f + ((Ax * Ax) / Ay) / Ay);
    st = (ct * Ax) / Ay;
    float ra = Ax * st + Ay * ct;
    float rc = Bx * ct - By * st;
    float rb = Bx * st + By * ct;
    float rd = Cx * ct - Cy * st;
    float re = Cx * st + Cy * ct;
    if (ra < 0.0f) { ra = -ra; rb = -rb; re = -re; }
    out[0] = ra; out[1] = rb; out[2] = rc; out[3] = rd; out[4] = re;
}

static inline void v3sub(const float* a, const float* b, float* o)
{
    o[0] = a[0] - b[0]; o[1] = a[1] - b[1]; o[2] = a[2] - b[2];
}
static inline float v3dot(const float* a, const float* b)
{
    return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}
static inline void v3cross(const float* a, const float* b, float* o)
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
static inline void v3unit(float* a)
{
    float s = twin_rsqrtf(v3dot(a, a));
    a[0] = s * a[0]; a[1] = s * a[1]; a[2] = s * a[2];
}

/* remember: ba[3m], ca[3m] = B - A, C - A (minimum image) */
ORC_EXPORT void orc_settle_remember(int m, const int* tri, const float* crd, const float* cell,
                                    const float* rcell, float* ba, float* ca)
{
    for (int t = 0; t < m; t++)
    {
        const int A = tri[3 * t], B = tri[3 * t + 1], C = tri[3 * t + 2];
        pbc_disp_f(crd + 3 * B, crd + 3 * A, cell, rcell, ba + 3 * t);
        pbc_disp_f(crd + 3 * C, crd + 3 * A, cell, rcell, ca + 3 * t);
    }
}

/* Float twin of the SETTLE kernel (settle.cpp:101-253): crd, vel updated;
 * vir6 (+=, int64 at 2^32) gets the per-triangle constraint virial in FP64
 * when vir6 != NULL. */
ORC_EXPORT void orc_settle_twin(int m, const int* tri, const float* geo, const float* mass,
                                float* crd, const float* cell, const float* rcell,
                                const float* ba, const float* ca, float dt, float v_factor,
                                float x_factor, float* vel, int64_t* vir6)
{
    for (int t = 0; t < m; t++)
    {
        const int A = tri[3 * t], B = tri[3 * t + 1], C = tri[3 * t + 2];
        const float ra = geo[5 * t], rb = geo[5 * t + 1], rc = geo[5 * t + 2];
        const float rd = geo[5 * t + 3], re = geo[5 * t + 4];
        float rA[3] = {crd[3 * A], crd[3 * A + 1], crd[3 * A + 2]}, rB[3], rC[3];
        pbc_disp_f(crd + 3 * B, rA, cell, rcell, rB);
        pbc_disp_f(crd + 3 * C, rA, cell, rcell, rC);
        const float mA = mass[A], mB = mass[B], mC = mass[C];
        const float imt = 1.0f / ((mA + mB) + mC);
        float rO[3], rA1[3], rB1[3], rC1[3];
        for (int k = 0; k < 3; k++)
        {
            rO[k] = imt * (mB * rB[k] + mC * rC[k]) + rA[k];
            rA1[k] = rA[k] - rO[k];
            rB1[k] = rB[k] + rA1[k];
            rC1[k] = rC[k] + rA1[k];
        }
        const float* rB0 = ba + 3 * t;
        const float* rC0 = ca + 3 * t;
        float ez[3], ex[3], ey[3];
        v3cross(rB0, rC0, ez);
        v3cross(rA1, ez, ex);
        v3cross(ez, ex, ey);
        v3unit(ex);
        v3unit(ey);
        v3unit(ez);
        float b0x = v3dot(ex, rB0), b0y = v3dot(ey, rB0);
        float c0x = v3dot(ex, rC0), c0y = v3dot(ey, rC0);
        float a1z = v3dot(ez, rA1);
        float b1x = v3dot(ex, rB1), b1y = v3dot(ey, rB1), b1z = v3dot(ez, rB1);
        float c1x = v3dot(ex, rC1), c1y = v3dot(ey, rC1), c1z = v3dot(ez, rC1);
        float sinphi = a1z / ra;
        float cosphi = sqrtf(1.0f - sinphi * sinphi);
        float sinpsi = ((b1z - c1z) - (rb - re) * sinphi) / ((rd - rc) * cosphi);
        float cospsi = sqrtf(1.0f - sinpsi * sinpsi);
        float a2y = ra * cosphi;
        float b2x = rc * cospsi, b2y = rb * cosphi + (rc * sinpsi) * sinphi;
        float c2x = rd * cospsi, c2y = re * cosphi + (rd * sinpsi) * sinphi;
        float alpha = ((b2x * b0x + c2x * c0x) + b2y * b0y) + c2y * c0y;
        float beta = (((-b2x) * b0y - c2x * c0y) + b2y * b0x) + c2y * c0x;
        float gamma = ((b1y * b0x - b1x * b0y) + c1y * c0x) - c1x * c0y;
        float ab = alpha * alpha + beta * beta;
        float sintheta = (alpha * gamma - beta * sqrtf(ab - gamma * gamma)) / ab;
        float costheta = sqrtf(1.0f - sintheta * sintheta);
        float p3[3][3] = {{-a2y * sintheta, a2y * costheta, a1z},
                          {b2x * costheta - b2y * sintheta, b2x * sintheta + b2y * costheta, b1z},
                          {c2x * costheta - c2y * sintheta, c2x * sintheta + c2y * costheta, c1z}};
        float r3[3][3];
        for (int q = 0; q < 3; q++)
            for (int k = 0; k < 3; k++)
                r3[q][k] = (p3[q][0] * ex[k] + p3[q][1] * ey[k]) + p3[q][2] * ez[k];
        const float kv = (v_factor / dt) / x_factor;
        const float* r1[3] = {rA1, rB1, rC1};
        const int at[3] = {A, B, C};
        float d[3][3];
        for (int q = 0; q < 3; q++)
            for (int k = 0; k < 3; k++)
            {
                d[q][k] = r3[q][k] - r1[q][k];
                vel[3 * at[q] + k] = vel[3 * at[q] + k] + kv * d[q][k];
            }
        if (vir6)
        {
            const float kf = ((1.0f / dt) / dt) / x_factor;
            float fB[3], fC[3];
            for (int k = 0; k < 3; k++)
            {
                fB[k] = (kf * mB) * d[1][k];
                fC[k] = (kf * mC) * d[2][k];
            }
            double v[6] = {
                (double)fB[0] * rB0[0] + (double)fC[0] * rC0[0],
                ((double)fB[0] * rB0[1] + (double)fB[1] * rB0[0]) +
                    ((double)fC[0] * rC0[1] + (double)fC[1] * rC0[0]),
                (double)fB[1] * rB0[1] + (double)fC[1] * rC0[1],
                ((double)fB[0] * rB0[2] + (double)fB[2] * rB0[0]) +
                    ((double)fC[0] * rC0[2] + (double)fC[2] * rC0[0]),
                ((double)fB[1] * rB0[2] + (double)fB[2] * rB0[1]) +
                    ((double)fC[1] * rC0[2] + (double)fC[2] * rC0[1]),
                (double)fB[2] * rB0[2] + (double)fC[2] * rC0[2]};
            for (int k = 0; k < 6; k++) vir6[k] += to_fixed_d(v[k]);
        }
        for (int q = 0; q < 3; q++)
            for (int k = 0; k < 3; k++) crd[3 * at[q] + k] = r3[q][k] + rO[k];
    }
}

/* remember_pair_AB (settle.cpp:25-44): ab[3m] = j - i (minimum image) */
ORC_EXPORT void orc_settle_pair_remember(int m, const int* pairs, const float* crd,
                                         const float* cell, const float* rcell, float* ab)
{
    for (int t = 0; t < m; t++)
        pbc_disp_f(crd + 3 * pairs[2 * t + 1], crd + 3 * pairs[2 * t], cell, rcell, ab + 3 * t);
}

/* Float twin of settle_pair (settle.cpp:258-307): per pair the displacement
 * k r2 (r2 the remembered j - i) with |r1 + k r2| = r0 for this step's
 * j - i = r1, split by the inverse masses (constrain_k = 1/(m_i + m_j)); the
 * velocities change by (v_factor / dt / x_factor) times the displacement;
 * vir6 (+=, int64 at 2^32) gets the constraint force (kv m_j / dt / v_factor
 * times j's displacement) times r2 in FP64 when vir6 != NULL. */
ORC_EXPORT void orc_settle_pair_twin(int m, const int* pairs, const float* constant_r,
                                     const float* constrain_k, const float* mass, float* crd,
                                     const float* cell, const float* rcell, const float* ab,
                                     float dt, float v_factor, float x_factor, float* vel,
                                     int64_t* vir6)
{
    for (int t = 0; t < m; t++)
    {
        const int i = pairs[2 * t], j = pairs[2 * t + 1];
        float r1[3];
        pbc_disp_f(crd + 3 * j, crd + 3 * i, cell, rcell, r1);
        const float* r2 = ab + 3 * t;
        const float mA = mass[i], mB = mass[j];
        const float r0r0 = constant_r[t] * constant_r[t];
        const float r1r1 = v3dot(r1, r1), r1r2 = v3dot(r1, r2), r2r2 = v3dot(r2, r2);
        const float k = (sqrtf((r1r2 * r1r2 - r1r1 * r2r2) + r2r2 * r0r0) - r1r2) / r2r2;
        const float ci = (-mB) * constrain_k[t], cj = mA * constrain_k[t];
        const float kv = (v_factor / dt) / x_factor;
        float dj[3];
        for (int q = 0; q < 3; q++)
        {
            const float kr2 = k * r2[q];
            const float di = ci * kr2;
            dj[q] = cj * kr2;
            crd[3 * i + q] = crd[3 * i + q] + di;
            crd[3 * j + q] = crd[3 * j + q] + dj[q];
            vel[3 * i + q] = vel[3 * i + q] + kv * di;
            vel[3 * j + q] = vel[3 * j + q] + kv * dj[q];
        }
        if (vir6)
        {
            const float kf = ((kv * mB) / dt) / v_factor;
            const float f[3] = {kf * dj[0], kf * dj[1], kf * dj[2]};
            const double v[6] = {(double)f[0] * r2[0],
                                 (double)f[0] * r2[1] + (double)f[1] * r2[0],
                                 (double)f[1] * r2[1],
                                 (double)f[0] * r2[2] + (double)f[2] * r2[0],
                                 (double)f[1] * r2[2] + (double)f[2] * r2[1],
                                 (double)f[2] * r2[2]};
            for (int q = 0; q < 6; q++) vir6[q] += to_fixed_d(v[q]);
        }
    }
}
