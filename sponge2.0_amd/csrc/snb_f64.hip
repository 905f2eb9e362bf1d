// snb_f64.hip -- an all-FP64 evaluation of the whole force path and an FP64
// NVE leap-frog + SETTLE step, for precision studies of the FP32 / int64 path.
//
// Same modules and inputs as the FP32 path (the context's LJ tables, 1-4
// pairs, bonds, exclusions, PME grid, beta and B-spline moduli, SETTLE
// triangles), with double coordinates, velocities and forces:
//  - pairs over the context's reference-layout half list (rebuilt from a float
//    copy of the coordinates with the context's skin and refresh policy), each
//    pair in double: minimum image, LJ, erfc-screened Coulomb (the device
//    library's erfc / exp);
//  - exclusion correction, 1-4 pairs, harmonic bonds in double;
//  - PME reciprocal in double: order-4 B-spline spread, a full complex 3-D
//    transform as three passes of direct DFTs along each axis (n <= a few
//    hundred points per line, twiddles e^{-2 pi i m/n} tabulated in double on
//    the host), the influence function on the full grid, energy and virial
//    from the spectrum, inverse transform, gather;
//  - forces summed in int64 fixed point at 2^40 (resolution 9e-13 kcal/mol/A,
//    order-independent, so a run is reproducible), energies and the virial at
//    2^32 (as the FP32 path's energies).
// Nothing here is on the FP32 path's hot loop; it is the FP64 reference a
// trajectory of the FP32 path is compared with.
#include <vector>

#include "snb_internal.h"

#define F64_SCALE 1099511627776.0           // 2^40
#define F64_INV 9.094947017729282e-13       // 2^-40

enum
{
    F64_E_LJ = 0,
    F64_E_DIRECT,
    F64_E_CORR,
    F64_E_NB14_LJ,
    F64_E_NB14_EE,
    F64_E_BOND,
    F64_Q2,  // sum q^2
    F64_Q1,  // sum q
    F64_VIR, // 6 slots: a11 a21 a22 a31 a32 a33
    F64_NSUM = F64_VIR + 6
};

struct SnbF64
{
    int n = 0;
    int64_t* fx = nullptr;      // 3n, SoA, scale 2^40
    int64_t* sums = nullptr;    // F64_NSUM, scale 2^32
    snb_vector* crd32 = nullptr;
    // PME
    int fx_ = 0, fy = 0, fz = 0;
    int64_t* qfix = nullptr;    // nall, scale 2^40
    double2* ga = nullptr;      // nall
    double2* gb = nullptr;      // nall
    double* bc = nullptr;       // nall (full grid)
    double2* tw = nullptr;      // fx + fy + fz: e^{-2 pi i m/n} per axis
    double* partials = nullptr; // conv blocks x 8
    int conv_blocks = 0;
    float bc_rcell[6] = {0, 0, 0, 0, 0, 0};
    // SETTLE
    double* geo = nullptr;      // 5 per triangle
    double* last = nullptr;     // 6 per triangle: previous B - A, C - A
};

static __device__ __forceinline__ int64_t fix40(double v)
{
    return (int64_t)__builtin_rint(v * F64_SCALE);
}

static __device__ __forceinline__ int64_t fix32(double v)
{
    return (int64_t)__builtin_rint(v * SNB_FIXED_SCALE);
}

static __device__ __forceinline__ void add_i64(int64_t* p, int64_t v)
{
    if (v) atomicAdd((unsigned long long*)p, (unsigned long long)v);
}

static __device__ __forceinline__ int64_t wsum(int64_t v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

struct BoxD
{
    double c[6], rc[6];
};

static BoxD box_d(snb_ltmatrix3 cell, snb_ltmatrix3 rcell)
{
    BoxD b;
    const float* c = &cell.a11;
    const float* r = &rcell.a11;
    for (int k = 0; k < 6; k++)
    {
        b.c[k] = c[k];
        b.rc[k] = r[k];
    }
    return b;
}

// d = r_j - r_i to the nearest image (row vector times the lower-triangular
// cell, as the FP32 path)
static __device__ __forceinline__ void disp_d(const double* rj, const double* ri,
                                              const BoxD& b, double* d)
{
    double dx = rj[0] - ri[0], dy = rj[1] - ri[1], dz = rj[2] - ri[2];
    double sx = (dx * b.rc[0] + dy * b.rc[1]) + dz * b.rc[3];
    double sy = dy * b.rc[2] + dz * b.rc[4];
    double sz = dz * b.rc[5];
    double kx = floor(sx + 0.5), ky = floor(sy + 0.5), kz = floor(sz + 0.5);
    d[0] = dx - ((kx * b.c[0] + ky * b.c[1]) + kz * b.c[3]);
    d[1] = dy - (ky * b.c[2] + kz * b.c[4]);
    d[2] = dz - kz * b.c[5];
}

// -F (x) d of one pair into six fixed-point lane sums
static __device__ __forceinline__ void vir_add(int64_t* v, const double* f, const double* d)
{
    v[0] += fix32(-(f[0] * d[0]));
    v[1] += fix32(-(f[0] * d[1] + f[1] * d[0]));
    v[2] += fix32(-(f[1] * d[1]));
    v[3] += fix32(-(f[0] * d[2] + f[2] * d[0]));
    v[4] += fix32(-(f[1] * d[2] + f[2] * d[1]));
    v[5] += fix32(-(f[2] * d[2]));
}

static __device__ __forceinline__ void vir_flush(int64_t* v, int64_t* sums)
{
    for (int c = 0; c < 6; c++)
    {
        int64_t s = wsum(v[c]);
        if ((threadIdx.x & 63) == 0) add_i64(sums + F64_VIR + c, s);
    }
}

__global__ void k_f64_to_f32(int n, const double* __restrict__ crd, snb_vector* __restrict__ out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = {(float)crd[3 * i], (float)crd[3 * i + 1], (float)crd[3 * i + 2]};
}

// LJ + erfc Coulomb over the half list: one wavefront per row
template <bool E, bool V>
__global__ void __launch_bounds__(256) k_f64_pair(int n, const snb_atom_group* __restrict__ nl,
                                                  const double* __restrict__ crd,
                                                  const float* __restrict__ charge,
                                                  const int* __restrict__ type,
                                                  const float* __restrict__ lj_a,
                                                  const float* __restrict__ lj_b, BoxD box,
                                                  double rc, double beta,
                                                  int64_t* __restrict__ fx,
                                                  int64_t* __restrict__ sums)
{
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= n) return;
    const snb_atom_group g = nl[i];
    const double ri[3] = {crd[3 * i], crd[3 * i + 1], crd[3 * i + 2]};
    const double qi = charge[i];
    const int ti = type[i];
    int64_t sx = 0, sy = 0, sz = 0, elj = 0, ec = 0;
    int64_t v[6] = {0, 0, 0, 0, 0, 0};
    const double tsp = 1.1283791670955126;  // 2/sqrt(pi)
    for (int k = lane; k < g.atom_numbers; k += 64)
    {
        const int j = g.atom_serial[k];
        const double rj[3] = {crd[3 * j], crd[3 * j + 1], crd[3 * j + 2]};
        double d[3];
        disp_d(rj, ri, box, d);
        const double r2 = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2];
        const double r = sqrt(r2);
        if (!(r < rc)) continue;
        const int p = snb_lj_pair_index(ti, type[j]);
        const double A = lj_a[p], B = lj_b[p];
        const double ri2 = 1.0 / r2, ri6 = (ri2 * ri2) * ri2;
        const double qq = qi * (double)charge[j];
        const double x = beta * r;
        const double ex = exp(-x * x);
        const double erfcx_ = erfc(x);
        const double fc = (qq * ri2 / r) * (tsp * x * ex + erfcx_);
        const double fa = (B - A * ri6) * ri6 * ri2 - fc;
        const double f[3] = {fa * d[0], fa * d[1], fa * d[2]};
        const int64_t ix = fix40(f[0]), iy = fix40(f[1]), iz = fix40(f[2]);
        sx += ix;
        sy += iy;
        sz += iz;
        add_i64(fx + j, -ix);
        add_i64(fx + n + j, -iy);
        add_i64(fx + 2 * n + j, -iz);
        if (E)
        {
            elj += fix32((A * (1.0 / 12.0) * ri6 - B * (1.0 / 6.0)) * ri6);
            ec += fix32(qq * erfcx_ / r);
        }
        if (V) vir_add(v, f, d);
    }
    sx = wsum(sx);
    sy = wsum(sy);
    sz = wsum(sz);
    if (lane == 0)
    {
        add_i64(fx + i, sx);
        add_i64(fx + n + i, sy);
        add_i64(fx + 2 * n + i, sz);
    }
    if (E)
    {
        elj = wsum(elj);
        ec = wsum(ec);
        if (lane == 0)
        {
            add_i64(sums + F64_E_LJ, elj);
            add_i64(sums + F64_E_DIRECT, ec);
        }
    }
    if (V) vir_flush(v, sums);
}

// Ewald exclusion correction: minus the erf part of every excluded pair
template <bool E, bool V>
__global__ void __launch_bounds__(256) k_f64_excl(int n, const int* __restrict__ ex_start,
                                                  const int* __restrict__ ex_list,
                                                  const int* __restrict__ ex_num,
                                                  const double* __restrict__ crd,
                                                  const float* __restrict__ charge, BoxD box,
                                                  double beta, int64_t* __restrict__ fx,
                                                  int64_t* __restrict__ sums)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int64_t e = 0;
    int64_t v[6] = {0, 0, 0, 0, 0, 0};
    if (i < n)
    {
        const double ri[3] = {crd[3 * i], crd[3 * i + 1], crd[3 * i + 2]};
        const double qi = charge[i];
        const int s0 = ex_start[i], cnt = ex_num[i];
        int64_t sx = 0, sy = 0, sz = 0;
        for (int k = s0; k < s0 + cnt; k++)
        {
            const int j = ex_list[k];
            const double rj[3] = {crd[3 * j], crd[3 * j + 1], crd[3 * j + 2]};
            double d[3];
            disp_d(rj, ri, box, d);
            const double r = sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
            const double x = beta * r;
            const double er = erf(x);
            const double qq = qi * (double)charge[j];
            const double h = er - 1.1283791670955126 * x * exp(-x * x);
            const double fa = qq * h / (r * r * r);
            const double f[3] = {fa * d[0], fa * d[1], fa * d[2]};
            const int64_t ix = fix40(f[0]), iy = fix40(f[1]), iz = fix40(f[2]);
            sx += ix;
            sy += iy;
            sz += iz;
            add_i64(fx + j, -ix);
            add_i64(fx + n + j, -iy);
            add_i64(fx + 2 * n + j, -iz);
            if (E) e += fix32(-(qq * er / r));
            if (V) vir_add(v, f, d);
        }
        add_i64(fx + i, sx);
        add_i64(fx + n + i, sy);
        add_i64(fx + 2 * n + i, sz);
    }
    if (E)
    {
        e = wsum(e);
        if ((threadIdx.x & 63) == 0) add_i64(sums + F64_E_CORR, e);
    }
    if (V) vir_flush(v, sums);
}

// scaled 1-4 pairs: f on a = (-A r^-14 + B r^-8 - cf qa qb r^-3) d, d = r_b - r_a
template <bool E, bool V>
__global__ void __launch_bounds__(256) k_f64_nb14(int m, int n, const int* __restrict__ pa,
                                                  const int* __restrict__ pb,
                                                  const float* __restrict__ A14,
                                                  const float* __restrict__ B14,
                                                  const float* __restrict__ cf14,
                                                  const double* __restrict__ crd,
                                                  const float* __restrict__ charge, BoxD box,
                                                  int64_t* __restrict__ fx,
                                                  int64_t* __restrict__ sums)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    int64_t elj = 0, eee = 0;
    int64_t v[6] = {0, 0, 0, 0, 0, 0};
    if (t < m)
    {
        const int a = pa[t], b = pb[t];
        const double ra[3] = {crd[3 * a], crd[3 * a + 1], crd[3 * a + 2]};
        const double rb[3] = {crd[3 * b], crd[3 * b + 1], crd[3 * b + 2]};
        double d[3];
        disp_d(rb, ra, box, d);
        const double r2 = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2];
        const double r = sqrt(r2);
        const double ri2 = 1.0 / r2, ri6 = ri2 * ri2 * ri2;
        const double A = A14[t], B = B14[t];
        const double qq = (double)cf14[t] * (double)charge[a] * (double)charge[b];
        const double fa = -A * ri6 * ri6 * ri2 + B * ri6 * ri2 - qq * ri2 / r;
        const double f[3] = {fa * d[0], fa * d[1], fa * d[2]};
        for (int c = 0; c < 3; c++)
        {
            const int64_t q = fix40(f[c]);
            add_i64(fx + (size_t)c * n + a, q);
            add_i64(fx + (size_t)c * n + b, -q);
        }
        if (E)
        {
            // the reference's constants of the 1-4 energy (nb14.cpp:86-87), so
            // that the FP32 path is compared on the same definition
            elj = fix32((0.08333333 * A * ri6 - 0.1666666 * B) * ri6);
            eee = fix32(qq / r);
        }
        if (V) vir_add(v, f, d);
    }
    if (E)
    {
        elj = wsum(elj);
        eee = wsum(eee);
        if ((threadIdx.x & 63) == 0)
        {
            add_i64(sums + F64_E_NB14_LJ, elj);
            add_i64(sums + F64_E_NB14_EE, eee);
        }
    }
    if (V) vir_flush(v, sums);
}

// harmonic bonds E = k (r - r0)^2: f on a = 2 k (r - r0) / r d, d = r_b - r_a
template <bool E, bool V>
__global__ void __launch_bounds__(256) k_f64_bond(int m, int n, const int* __restrict__ pa,
                                                  const int* __restrict__ pb,
                                                  const float* __restrict__ kk,
                                                  const float* __restrict__ r0,
                                                  const double* __restrict__ crd, BoxD box,
                                                  int64_t* __restrict__ fx,
                                                  int64_t* __restrict__ sums)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    int64_t eb = 0;
    int64_t v[6] = {0, 0, 0, 0, 0, 0};
    if (t < m)
    {
        const int a = pa[t], b = pb[t];
        const double ra[3] = {crd[3 * a], crd[3 * a + 1], crd[3 * a + 2]};
        const double rb[3] = {crd[3 * b], crd[3 * b + 1], crd[3 * b + 2]};
        double d[3];
        disp_d(rb, ra, box, d);
        const double r = sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
        const double dr = r - (double)r0[t];
        const double fa = 2.0 * (double)kk[t] * dr / r;
        const double f[3] = {fa * d[0], fa * d[1], fa * d[2]};
        for (int c = 0; c < 3; c++)
        {
            const int64_t q = fix40(f[c]);
            add_i64(fx + (size_t)c * n + a, q);
            add_i64(fx + (size_t)c * n + b, -q);
        }
        if (E) eb = fix32((double)kk[t] * dr * dr);
        if (V) vir_add(v, f, d);
    }
    if (E)
    {
        eb = wsum(eb);
        if ((threadIdx.x & 63) == 0) add_i64(sums + F64_E_BOND, eb);
    }
    if (V) vir_flush(v, sums);
}

// ---------------------------------------------------------------------------
// PME in double
// ---------------------------------------------------------------------------
static __device__ __forceinline__ void spline_d(double t, double* w, double* dw)
{
    const double t2 = t * t, t3 = t2 * t;
    w[0] = t3 / 6.0;
    w[1] = (-3.0 * t3 + 3.0 * t2 + 3.0 * t + 1.0) / 6.0;
    w[2] = (3.0 * t3 - 6.0 * t2 + 4.0) / 6.0;
    w[3] = (1.0 - t) * (1.0 - t) * (1.0 - t) / 6.0;
    dw[0] = 0.5 * t2;
    dw[1] = (-3.0 * t2 + 2.0 * t + 1.0) * 0.5;
    dw[2] = (3.0 * t2 - 4.0 * t) * 0.5;
    dw[3] = -0.5 * (1.0 - t) * (1.0 - t);
}

// base grid point u and fraction t per axis (points u - k get weight w[k])
static __device__ __forceinline__ void near_d(const double* r, const double* rc, int fx, int fy,
                                              int fz, int* u, double* t)
{
    double f[3] = {(r[0] * rc[0] + r[1] * rc[1]) + r[2] * rc[3], r[1] * rc[2] + r[2] * rc[4],
                   r[2] * rc[5]};
    const int K[3] = {fx, fy, fz};
    for (int a = 0; a < 3; a++)
    {
        double g = f[a] - floor(f[a]);
        if (!isfinite(g)) g = 0.0;
        double s = g * K[a];
        int ui = (int)s;
        ui = ui < 0 ? 0 : (ui < K[a] ? ui : K[a] - 1);
        u[a] = ui;
        t[a] = s - ui;
    }
}

static __device__ __forceinline__ int wrapk(int g, int K) { return g < 0 ? g + K : g; }

__global__ void k_f64_spread(int n, const double* __restrict__ crd,
                             const float* __restrict__ charge, BoxD box, int fx, int fy, int fz,
                             int64_t* __restrict__ q)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int u[3];
    double t[3], wx[4], wy[4], wz[4], dd[4];
    near_d(crd + 3 * i, box.rc, fx, fy, fz, u, t);
    spline_d(t[0], wx, dd);
    spline_d(t[1], wy, dd);
    spline_d(t[2], wz, dd);
    const double qi = charge[i];
    for (int a = 0; a < 4; a++)
    {
        const int gx = wrapk(u[0] - a, fx);
        for (int b = 0; b < 4; b++)
        {
            const int gy = wrapk(u[1] - b, fy);
            const double wab = qi * wx[a] * wy[b];
            for (int c = 0; c < 4; c++)
            {
                const int gz = wrapk(u[2] - c, fz);
                add_i64(q + ((size_t)gx * fy + gy) * fz + gz, fix40(wab * wz[c]));
            }
        }
    }
}

__global__ void k_f64_grid_in(size_t nall, const int64_t* __restrict__ q, double2* __restrict__ g)
{
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nall) g[t] = make_double2((double)q[t] * F64_INV, 0.0);
}

// One pass of direct DFTs along one axis of a [fx][fy][fz] complex grid:
// out[.., k, ..] = sum_j in[.., j, ..] e^{sign 2 pi i j k / n}
__global__ void k_f64_dft(size_t nall, int n, size_t stride, const double2* __restrict__ tw,
                          double sign, const double2* __restrict__ in,
                          double2* __restrict__ out)
{
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nall) return;
    const size_t hi = t / (stride * n), lo = t % stride;
    const int k = (int)((t / stride) % n);
    const double2* line = in + hi * stride * n + lo;
    double re = 0.0, im = 0.0;
    int m = 0;
    for (int j = 0; j < n; j++)
    {
        const double2 a = line[(size_t)j * stride];
        const double2 w = tw[m];  // e^{-2 pi i m / n}
        const double wi = sign < 0 ? w.y : -w.y;
        re += a.x * w.x - a.y * wi;
        im += a.x * wi + a.y * w.x;
        m += k;
        if (m >= n) m -= n;
    }
    out[t] = make_double2(re, im);
}

// influence function on the full grid (as the FP32 path's half-spectrum one)
__global__ void k_f64_bc(int fx, int fy, int fz, const double* __restrict__ bmod, BoxD box,
                         double mpre, double volume, double* __restrict__ bc)
{
    const int kx = blockIdx.x / fy, ky = blockIdx.x % fy;
    const double kxr = kx > fx / 2 ? kx - fx : kx;
    const double kyr = ky > fy / 2 ? ky - fy : ky;
    const double bxy = bmod[kx] * bmod[fx + ky];
    for (int kz = threadIdx.x; kz < fz; kz += blockDim.x)
    {
        const double kzr = kz > fz / 2 ? kz - fz : kz;
        const double mx = kxr * box.rc[0];
        const double my = kxr * box.rc[1] + kyr * box.rc[2];
        const double mz = kxr * box.rc[3] + kyr * box.rc[4] + kzr * box.rc[5];
        const double msq = mx * mx + my * my + mz * mz;
        const size_t idx = ((size_t)kx * fy + ky) * fz + kz;
        bc[idx] = (kx == 0 && ky == 0 && kz == 0)
                      ? 0.0
                      : exp(-mpre * msq) / (M_PI * msq * volume) * bxy * bmod[fx + fy + kz];
    }
}

// spectrum *= BC; with EV, the energy 1/2 sum BC |F|^2 and the virial
// -1/4 sum BC |F|^2 (delta - 2 (1 + pi^2 m^2 / beta^2) m m / m^2) (the FP32
// path's half-spectrum sums written over the full grid) into block partials
template <bool EV>
__global__ void __launch_bounds__(256) k_f64_conv(int fx, int fy, int fz, double2* __restrict__ g,
                                                  const double* __restrict__ bc, BoxD box,
                                                  double mpre, double* __restrict__ partials)
{
    const size_t nall = (size_t)fx * fy * fz;
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < nall;
         t += (size_t)gridDim.x * blockDim.x)
    {
        const double b = bc[t];
        double2 f = g[t];
        if (EV && b != 0.0)
        {
            const int kz = (int)(t % fz), ky = (int)((t / fz) % fy), kx = (int)(t / ((size_t)fy * fz));
            const double kxr = kx > fx / 2 ? kx - fx : kx;
            const double kyr = ky > fy / 2 ? ky - fy : ky;
            const double kzr = kz > fz / 2 ? kz - fz : kz;
            const double mx = kxr * box.rc[0];
            const double my = kxr * box.rc[1] + kyr * box.rc[2];
            const double mz = kxr * box.rc[3] + kyr * box.rc[4] + kzr * box.rc[5];
            const double msq = mx * mx + my * my + mz * mz;
            const double a2 = f.x * f.x + f.y * f.y;
            acc[0] += 0.5 * b * a2;
            const double cf = 2.0 / msq * (1.0 + mpre * msq);
            const double hb = 0.25 * b * a2;
            acc[1] -= hb * (1.0 - cf * mx * mx);
            acc[2] -= hb * (0.0 - cf * my * mx);
            acc[3] -= hb * (1.0 - cf * my * my);
            acc[4] -= hb * (0.0 - cf * mz * mx);
            acc[5] -= hb * (0.0 - cf * mz * my);
            acc[6] -= hb * (1.0 - cf * mz * mz);
        }
        g[t] = make_double2(f.x * b, f.y * b);
    }
    if (EV)
    {
        __shared__ double red[7][4];
        const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
        for (int c = 0; c < 7; c++)
        {
            double v = acc[c];
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
            if (lane == 0) red[c][wid] = v;
        }
        __syncthreads();
        if (threadIdx.x < 7)
        {
            double s = 0;
            for (int q = 0; q < 4; q++) s += red[threadIdx.x][q];
            partials[(size_t)blockIdx.x * 8 + threadIdx.x] = s;
        }
    }
}

__global__ void k_f64_gather(int n, const double* __restrict__ crd,
                             const float* __restrict__ charge, BoxD box, int fx, int fy, int fz,
                             const double2* __restrict__ phi, int64_t* __restrict__ frc)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int u[3];
    double t[3], wx[4], wy[4], wz[4], dx[4], dy[4], dz[4];
    near_d(crd + 3 * i, box.rc, fx, fy, fz, u, t);
    spline_d(t[0], wx, dx);
    spline_d(t[1], wy, dy);
    spline_d(t[2], wz, dz);
    double sx = 0, sy = 0, sz = 0;
    for (int a = 0; a < 4; a++)
    {
        const int gx = wrapk(u[0] - a, fx);
        for (int b = 0; b < 4; b++)
        {
            const int gy = wrapk(u[1] - b, fy);
            for (int c = 0; c < 4; c++)
            {
                const int gz = wrapk(u[2] - c, fz);
                const double ph = phi[((size_t)gx * fy + gy) * fz + gz].x;
                sx += ph * dx[a] * wy[b] * wz[c];
                sy += ph * wx[a] * dy[b] * wz[c];
                sz += ph * wx[a] * wy[b] * dz[c];
            }
        }
    }
    const double tx = sx * fx, ty = sy * fy, tz = sz * fz;
    const double mq = -(double)charge[i];
    const double gx = tx * box.rc[0];
    const double gy = tx * box.rc[1] + ty * box.rc[2];
    const double gz = (tx * box.rc[3] + ty * box.rc[4]) + tz * box.rc[5];
    add_i64(frc + i, fix40(mq * gx));
    add_i64(frc + n + i, fix40(mq * gy));
    add_i64(frc + 2 * n + i, fix40(mq * gz));
}

__global__ void k_f64_charge_sums(int n, const float* __restrict__ q, int64_t* __restrict__ sums)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int64_t s2 = 0, s1 = 0;
    if (i < n)
    {
        const double v = q[i];
        s2 = fix32(v * v);
        s1 = fix32(v);
    }
    s2 = wsum(s2);
    s1 = wsum(s1);
    if ((threadIdx.x & 63) == 0)
    {
        add_i64(sums + F64_Q2, s2);
        add_i64(sums + F64_Q1, s1);
    }
}

__global__ void k_f64_flush(int n, int64_t* __restrict__ fx, double* __restrict__ frc)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int c = 0; c < 3; c++)
    {
        frc[3 * i + c] += (double)fx[(size_t)c * n + i] * F64_INV;
        fx[(size_t)c * n + i] = 0;
    }
}

__global__ void k_f64_fill(size_t m, int64_t* __restrict__ p)
{
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < m) p[t] = 0;
}

// ---------------------------------------------------------------------------
// FP64 NVE step: remember, leap-frog, SETTLE
// ---------------------------------------------------------------------------
struct D3
{
    double x, y, z;
};
static __device__ __forceinline__ D3 d3(double x, double y, double z) { return {x, y, z}; }
static __device__ __forceinline__ D3 operator+(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
static __device__ __forceinline__ D3 operator-(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static __device__ __forceinline__ D3 operator*(double s, D3 a) { return {s * a.x, s * a.y, s * a.z}; }
static __device__ __forceinline__ double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static __device__ __forceinline__ D3 cross(D3 a, D3 b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
static __device__ __forceinline__ D3 unit(D3 a) { return (1.0 / sqrt(dot(a, a))) * a; }
static __device__ __forceinline__ D3 ld3(const double* p, int i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
static __device__ __forceinline__ void st3(double* p, int i, D3 v)
{
    p[3 * i] = v.x;
    p[3 * i + 1] = v.y;
    p[3 * i + 2] = v.z;
}
static __device__ __forceinline__ D3 disp3(D3 rj, D3 ri, const BoxD& b)
{
    double a[3] = {rj.x, rj.y, rj.z}, c[3] = {ri.x, ri.y, ri.z}, d[3];
    disp_d(a, c, b, d);
    return {d[0], d[1], d[2]};
}

__global__ void k_f64_remember(int m, const int4* __restrict__ tri, const double* __restrict__ crd,
                               BoxD box, double* __restrict__ last)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m) return;
    const int4 a = tri[t];
    const D3 rA = ld3(crd, a.x);
    const D3 ba = disp3(ld3(crd, a.y), rA, box), ca = disp3(ld3(crd, a.z), rA, box);
    st3(last, 2 * t, ba);
    st3(last, 2 * t + 1, ca);
}

// v += f / m dt; x += v dt (NVE_iteration::Leap_Frog, MD_core/nve.hpp:3-48)
__global__ void k_f64_leap_frog(int n, const double* __restrict__ inv_mass,
                                const double* __restrict__ frc, double* __restrict__ vel,
                                double* __restrict__ crd, double dt)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int c = 0; c < 3; c++)
    {
        const double v = vel[3 * i + c] + frc[3 * i + c] * inv_mass[i] * dt;
        vel[3 * i + c] = v;
        crd[3 * i + c] += v * dt;
    }
}

// SETTLE (Miyamoto & Kollman, J. Comput. Chem. 13, 952 (1992)) in double, the
// same canonical-frame steps as the FP32 path, with v_factor = x_factor = 1
__global__ void k_f64_settle(int m, const int4* __restrict__ tri, const double* __restrict__ geo,
                             const double* __restrict__ mass, const double* __restrict__ last,
                             BoxD box, double dt, double* __restrict__ crd,
                             double* __restrict__ vel)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m) return;
    const int4 id = tri[t];
    const double ra = geo[5 * t], rb = geo[5 * t + 1], rc = geo[5 * t + 2];
    const double rd = geo[5 * t + 3], re = geo[5 * t + 4];
    const double mA = mass[id.x], mB = mass[id.y], mC = mass[id.z];
    const D3 rB0 = ld3(last, 2 * t), rC0 = ld3(last, 2 * t + 1);
    const D3 rA = ld3(crd, id.x);
    const D3 rB = disp3(ld3(crd, id.y), rA, box);
    const D3 rC = disp3(ld3(crd, id.z), rA, box);
    const double inv_mt = 1.0 / (mA + mB + mC);
    const D3 rO = (inv_mt * (mB * rB + mC * rC)) + rA;
    const D3 rA1 = rA - rO, rB1 = rB + rA1, rC1 = rC + rA1;
    D3 ez = cross(rB0, rC0);
    D3 ex = cross(rA1, ez);
    D3 ey = cross(ez, ex);
    ex = unit(ex);
    ey = unit(ey);
    ez = unit(ez);
    const double b0x = dot(ex, rB0), b0y = dot(ey, rB0);
    const double c0x = dot(ex, rC0), c0y = dot(ey, rC0);
    const double a1z = dot(ez, rA1);
    const double b1x = dot(ex, rB1), b1y = dot(ey, rB1), b1z = dot(ez, rB1);
    const double c1x = dot(ex, rC1), c1y = dot(ey, rC1), c1z = dot(ez, rC1);
    const double sinphi = a1z / ra;
    const double cosphi = sqrt(1.0 - sinphi * sinphi);
    const double sinpsi = ((b1z - c1z) - (rb - re) * sinphi) / ((rd - rc) * cosphi);
    const double cospsi = sqrt(1.0 - sinpsi * sinpsi);
    const double a2y = ra * cosphi;
    const double b2x = rc * cospsi, b2y = rb * cosphi + rc * sinpsi * sinphi;
    const double c2x = rd * cospsi, c2y = re * cosphi + rd * sinpsi * sinphi;
    const double alpha = b2x * b0x + c2x * c0x + b2y * b0y + c2y * c0y;
    const double beta = -b2x * b0y - c2x * c0y + b2y * b0x + c2y * c0x;
    const double gamma = b1y * b0x - b1x * b0y + c1y * c0x - c1x * c0y;
    const double ab = alpha * alpha + beta * beta;
    const double sintheta = (alpha * gamma - beta * sqrt(ab - gamma * gamma)) / ab;
    const double costheta = sqrt(1.0 - sintheta * sintheta);
    const double a3x = -a2y * sintheta, a3y = a2y * costheta, a3z = a1z;
    const double b3x = b2x * costheta - b2y * sintheta, b3y = b2x * sintheta + b2y * costheta;
    const double c3x = c2x * costheta - c2y * sintheta, c3y = c2x * sintheta + c2y * costheta;
    const D3 rA3 = (a3x * ex + a3y * ey) + a3z * ez;
    const D3 rB3 = (b3x * ex + b3y * ey) + b1z * ez;
    const D3 rC3 = (c3x * ex + c3y * ey) + c1z * ez;
    const double kv = 1.0 / dt;
    st3(vel, id.x, ld3(vel, id.x) + kv * (rA3 - rA1));
    st3(vel, id.y, ld3(vel, id.y) + kv * (rB3 - rB1));
    st3(vel, id.z, ld3(vel, id.z) + kv * (rC3 - rC1));
    // positions about the old A image, as the FP32 path writes them
    st3(crd, id.x, rA3 + rO);
    st3(crd, id.y, rB3 + rO);
    st3(crd, id.z, rC3 + rO);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
void snb_f64_free(snb_context* c)
{
    SnbF64* F = c->f64;
    if (!F) return;
    void* bufs[] = {F->fx, F->sums, F->crd32, F->qfix, F->ga, F->gb, F->bc, F->tw, F->partials,
                    F->geo, F->last};
    for (void* p : bufs)
        if (p) (void)hipFree(p);
    delete F;
    c->f64 = nullptr;
}

template <class T>
static int alloc_n(T** p, size_t n)
{
    if (hipMalloc((void**)p, sizeof(T) * (n ? n : 1)) != hipSuccess)
    {
        *p = nullptr;
        return SNB_ERR_OUT_OF_MEMORY;
    }
    return SNB_OK;
}

static int f64_setup(snb_context* c)
{
    if (c->f64) return SNB_OK;
    SnbF64* F = new SnbF64();
    c->f64 = F;
    F->n = c->n;
    int rc = alloc_n(&F->fx, 3 * (size_t)c->n);
    if (!rc) rc = alloc_n(&F->sums, F64_NSUM);
    if (!rc) rc = alloc_n(&F->crd32, (size_t)c->n);
    if (!rc && c->pme.init)
    {
        F->fx_ = c->pme.fx;
        F->fy = c->pme.fy;
        F->fz = c->pme.fz;
        size_t nall = (size_t)F->fx_ * F->fy * F->fz;
        rc = alloc_n(&F->qfix, nall);
        if (!rc) rc = alloc_n(&F->ga, nall);
        if (!rc) rc = alloc_n(&F->gb, nall);
        if (!rc) rc = alloc_n(&F->bc, nall);
        F->conv_blocks = 1024;
        if (!rc) rc = alloc_n(&F->partials, 8 * (size_t)F->conv_blocks);
        if (!rc)
        {
            std::vector<double2> h;
            for (int K : {F->fx_, F->fy, F->fz})
                for (int m = 0; m < K; m++)
                {
                    const double a = -2.0 * M_PI * m / K;
                    h.push_back(make_double2(cos(a), sin(a)));
                }
            rc = alloc_n(&F->tw, h.size());
            if (!rc && hipMemcpy(F->tw, h.data(), sizeof(double2) * h.size(),
                                 hipMemcpyHostToDevice) != hipSuccess)
                rc = SNB_ERR_HIP;
        }
    }
    if (!rc && c->settle.init)
    {
        const int m = c->settle.count;
        rc = alloc_n(&F->last, 6 * (size_t)m);
        if (!rc) rc = alloc_n(&F->geo, 5 * (size_t)m);
        if (!rc && hipMemcpy(F->geo, c->settle.geo_d.data(), sizeof(double) * 5 * (size_t)m,
                             hipMemcpyHostToDevice) != hipSuccess)
            rc = SNB_ERR_HIP;
    }
    if (!rc)
    {
        hipLaunchKernelGGL(k_f64_fill, dim3((3 * (size_t)c->n + 255) / 256 + 1), dim3(256), 0,
                           c->stream, 3 * (size_t)c->n, F->fx);
        if (hipGetLastError() != hipSuccess) rc = SNB_ERR_HIP;
    }
    if (rc) snb_f64_free(c);
    return rc;
}

static int f64_pme(snb_context* c, const double* crd, const float* charge, const BoxD& box,
                   bool ev, double* e_rec, double* vir)
{
    SnbF64* F = c->f64;
    SnbPme& P = c->pme;
    hipStream_t s = c->stream;
    const int n = c->n;
    const int fx = F->fx_, fy = F->fy, fz = F->fz;
    const size_t nall = (size_t)fx * fy * fz;
    const unsigned gb = (unsigned)((nall + 255) / 256);
    const double beta = P.beta;
    const double mpre = M_PI * M_PI / (beta * beta);
    bool same_box = true;
    for (int k = 0; k < 6; k++) same_box = same_box && (F->bc_rcell[k] == (float)box.rc[k]);
    if (!same_box)
    {
        hipLaunchKernelGGL(k_f64_bc, dim3(fx * fy), dim3(64), 0, s, fx, fy, fz, P.bmod, box, mpre,
                           P.volume, F->bc);
        for (int k = 0; k < 6; k++) F->bc_rcell[k] = (float)box.rc[k];
    }
    hipLaunchKernelGGL(k_f64_fill, dim3(gb), dim3(256), 0, s, nall, F->qfix);
    if (n > 0)
        hipLaunchKernelGGL(k_f64_spread, dim3((n + 255) / 256), dim3(256), 0, s, n, crd, charge,
                           box, fx, fy, fz, F->qfix);
    hipLaunchKernelGGL(k_f64_grid_in, dim3(gb), dim3(256), 0, s, nall, F->qfix, F->ga);
    const double2* twx = F->tw;
    const double2* twy = F->tw + fx;
    const double2* twz = F->tw + fx + fy;
    // forward: z, y, x
    hipLaunchKernelGGL(k_f64_dft, dim3(gb), dim3(256), 0, s, nall, fz, (size_t)1, twz, -1.0,
                       F->ga, F->gb);
    hipLaunchKernelGGL(k_f64_dft, dim3(gb), dim3(256), 0, s, nall, fy, (size_t)fz, twy, -1.0,
                       F->gb, F->ga);
    hipLaunchKernelGGL(k_f64_dft, dim3(gb), dim3(256), 0, s, nall, fx, (size_t)fy * fz, twx,
                       -1.0, F->ga, F->gb);
    if (ev)
        hipLaunchKernelGGL(k_f64_conv<true>, dim3(F->conv_blocks), dim3(256), 0, s, fx, fy, fz,
                           F->gb, F->bc, box, mpre, F->partials);
    else
        hipLaunchKernelGGL(k_f64_conv<false>, dim3(F->conv_blocks), dim3(256), 0, s, fx, fy, fz,
                           F->gb, F->bc, box, mpre, F->partials);
    // inverse: x, y, z
    hipLaunchKernelGGL(k_f64_dft, dim3(gb), dim3(256), 0, s, nall, fx, (size_t)fy * fz, twx, 1.0,
                       F->gb, F->ga);
    hipLaunchKernelGGL(k_f64_dft, dim3(gb), dim3(256), 0, s, nall, fy, (size_t)fz, twy, 1.0,
                       F->ga, F->gb);
    hipLaunchKernelGGL(k_f64_dft, dim3(gb), dim3(256), 0, s, nall, fz, (size_t)1, twz, 1.0,
                       F->gb, F->ga);
    if (n > 0)
        hipLaunchKernelGGL(k_f64_gather, dim3((n + 255) / 256), dim3(256), 0, s, n, crd, charge,
                           box, fx, fy, fz, F->ga, F->fx);
    SNB_CHECK(hipGetLastError());
    if (ev)
    {
        std::vector<double> h(8 * (size_t)F->conv_blocks);
        SNB_CHECK(hipMemcpyAsync(h.data(), F->partials, sizeof(double) * h.size(),
                                 hipMemcpyDeviceToHost, s));
        SNB_CHECK(hipStreamSynchronize(s));
        double acc[7] = {0, 0, 0, 0, 0, 0, 0};
        for (int b = 0; b < F->conv_blocks; b++)
            for (int k = 0; k < 7; k++) acc[k] += h[8 * (size_t)b + k];
        *e_rec = acc[0];
        for (int k = 0; k < 6; k++) vir[k] = acc[1 + k];
    }
    return SNB_OK;
}

extern "C" int snb_f64_force(snb_context* c, const double* d_crd, const float* d_charge,
                             snb_ltmatrix3 cell, snb_ltmatrix3 rcell, int step, int forced_nl,
                             double* d_frc, int need_energy, int need_virial, double* h_terms,
                             double* h_virial)
{
    if (!c) return SNB_ERR_ARGUMENT;
    if (!c->nl.init || !c->lj.init || !c->pme.init) return SNB_ERR_NOT_INITIALIZED;
    if (c->n > 0 && (!d_crd || !d_charge || !d_frc)) return SNB_ERR_ARGUMENT;
    if ((need_energy && !h_terms) || (need_virial && !h_virial)) return SNB_ERR_ARGUMENT;
    int dev_prev = -1;
    (void)hipGetDevice(&dev_prev);
    if (dev_prev != c->device) (void)hipSetDevice(c->device);
    int rc = f64_setup(c);
    SnbF64* F = c->f64;
    const int n = c->n;
    hipStream_t s = c->stream;
    const BoxD box = box_d(cell, rcell);
    const bool E = need_energy != 0, V = need_virial != 0;
    if (!rc && n > 0)
    {
        // the reference-layout list, rebuilt from the float copy
        if (!c->nl.export_mode)
        {
            c->nl.export_mode = 1;
            forced_nl = 1;
        }
        hipLaunchKernelGGL(k_f64_to_f32, dim3((n + 255) / 256), dim3(256), 0, s, n, d_crd,
                           F->crd32);
        rc = snb_nl_update(c, nullptr, n, 0, F->crd32, cell, rcell, step, forced_nl, nullptr,
                           nullptr, nullptr);
    }
    if (!rc)
    {
        hipLaunchKernelGGL(k_f64_fill, dim3(1), dim3(64), 0, s, (size_t)F64_NSUM, F->sums);
        if (n > 0)
        {
            const unsigned gw = (unsigned)((n + 3) / 4), ga = (unsigned)((n + 255) / 256);
            const double rcut = c->lj.cutoff;
            const double beta = c->pme.beta;
#define F64_EV(KERNEL, GRID, ...)                                                          \
    do                                                                                     \
    {                                                                                      \
        if (E && V)                                                                        \
            hipLaunchKernelGGL((KERNEL<true, true>), GRID, dim3(256), 0, s, __VA_ARGS__);   \
        else if (E)                                                                        \
            hipLaunchKernelGGL((KERNEL<true, false>), GRID, dim3(256), 0, s, __VA_ARGS__);  \
        else if (V)                                                                        \
            hipLaunchKernelGGL((KERNEL<false, true>), GRID, dim3(256), 0, s, __VA_ARGS__);  \
        else                                                                               \
            hipLaunchKernelGGL((KERNEL<false, false>), GRID, dim3(256), 0, s, __VA_ARGS__); \
    } while (0)
            F64_EV(k_f64_pair, dim3(gw), n, c->nl.dnl, d_crd, d_charge, c->type, c->lj.a,
                   c->lj.b, box, rcut, beta, F->fx, F->sums);
            if (c->has_excl)
                F64_EV(k_f64_excl, dim3(ga), n, c->ex_start, c->ex_list, c->ex_num, d_crd,
                       d_charge, box, beta, F->fx, F->sums);
            if (c->nb14.init && c->nb14.count > 0)
                F64_EV(k_f64_nb14, dim3((c->nb14.count + 255) / 256), c->nb14.count, n,
                       c->nb14.a, c->nb14.b, c->nb14.A, c->nb14.B, c->nb14.cf, d_crd, d_charge,
                       box, F->fx, F->sums);
            if (c->bond.init && c->bond.count > 0)
                F64_EV(k_f64_bond, dim3((c->bond.count + 255) / 256), c->bond.count, n,
                       c->bond.a, c->bond.b, c->bond.k, c->bond.r0, d_crd, box, F->fx, F->sums);
#undef F64_EV
            if (E)
                hipLaunchKernelGGL(k_f64_charge_sums, dim3(ga), dim3(256), 0, s, n, d_charge,
                                   F->sums);
            if (hipGetLastError() != hipSuccess) rc = SNB_ERR_HIP;
        }
    }
    double e_rec = 0.0, vrec[6] = {0, 0, 0, 0, 0, 0};
    if (!rc) rc = f64_pme(c, d_crd, d_charge, box, E || V, &e_rec, vrec);
    if (!rc && n > 0)
    {
        hipLaunchKernelGGL(k_f64_flush, dim3((n + 255) / 256), dim3(256), 0, s, n, F->fx, d_frc);
        if (hipGetLastError() != hipSuccess) rc = SNB_ERR_HIP;
    }
    if (!rc && (E || V))
    {
        int64_t h[F64_NSUM];
        if (hipMemcpyAsync(h, F->sums, sizeof(h), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            rc = SNB_ERR_HIP;
        if (!rc && E)
        {
            const double beta = c->pme.beta;
            const double q2 = snb_fixed_to_double(h[F64_Q2]), q1 = snb_fixed_to_double(h[F64_Q1]);
            h_terms[0] = snb_fixed_to_double(h[F64_E_LJ]);
            // Lennard_Jones_force.cpp:246-272 in double: -2/3 pi C6_total / rc^3 / 6 / V
            const double rcut = c->lj.cutoff;
            h_terms[1] = c->lj.c6_total * (-2.0 / 3.0 * M_PI / (rcut * rcut * rcut) / 6.0) /
                         c->pme.volume;
            h_terms[2] = snb_fixed_to_double(h[F64_E_DIRECT]);
            h_terms[3] = e_rec;
            h_terms[4] = -beta / sqrt(M_PI) * q2 - 0.5 * M_PI / (beta * beta * c->pme.volume) * q1 * q1;
            h_terms[5] = snb_fixed_to_double(h[F64_E_CORR]);
            h_terms[6] = snb_fixed_to_double(h[F64_E_NB14_LJ]);
            h_terms[7] = snb_fixed_to_double(h[F64_E_NB14_EE]);
            h_terms[8] = snb_fixed_to_double(h[F64_E_BOND]);
        }
        if (!rc && V)
            for (int k = 0; k < 6; k++)
                h_virial[k] = snb_fixed_to_double(h[F64_VIR + k]) + vrec[k];
    }
    if (dev_prev >= 0 && dev_prev != c->device) (void)hipSetDevice(dev_prev);
    return rc;
}

extern "C" int snb_f64_nve_iterate(snb_context* c, const double* d_mass,
                                   const double* d_inverse_mass, const double* d_frc,
                                   double* d_vel, double* d_crd, snb_ltmatrix3 cell,
                                   snb_ltmatrix3 rcell, double dt)
{
    if (!c) return SNB_ERR_ARGUMENT;
    if (c->n > 0 && (!d_mass || !d_inverse_mass || !d_frc || !d_vel || !d_crd))
        return SNB_ERR_ARGUMENT;
    if (c->settle.npair > 0) return SNB_ERR_UNSUPPORTED;  // rigid triangles only
    int dev_prev = -1;
    (void)hipGetDevice(&dev_prev);
    if (dev_prev != c->device) (void)hipSetDevice(c->device);
    int rc = f64_setup(c);
    SnbF64* F = c->f64;
    hipStream_t s = c->stream;
    const BoxD box = box_d(cell, rcell);
    const int n = c->n, m = c->settle.init ? c->settle.count : 0;
    if (!rc && n > 0)
    {
        if (m > 0)
            hipLaunchKernelGGL(k_f64_remember, dim3((m + 255) / 256), dim3(256), 0, s, m,
                               c->settle.tri, d_crd, box, F->last);
        hipLaunchKernelGGL(k_f64_leap_frog, dim3((n + 255) / 256), dim3(256), 0, s, n,
                           d_inverse_mass, d_frc, d_vel, d_crd, dt);
        if (m > 0)
            hipLaunchKernelGGL(k_f64_settle, dim3((m + 255) / 256), dim3(256), 0, s, m,
                               c->settle.tri, F->geo, d_mass, F->last, box, dt, d_crd, d_vel);
        if (hipGetLastError() != hipSuccess) rc = SNB_ERR_HIP;
    }
    if (dev_prev >= 0 && dev_prev != c->device) (void)hipSetDevice(dev_prev);
    return rc;
}
