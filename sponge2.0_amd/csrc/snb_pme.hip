// snb_pme.hip -- smooth PME reciprocal space on gfx950 (stands in for
// Particle_Mesh::PME_Reciprocal_Force_With_Energy_And_Virial,
// PM_force.cpp:1067-1165).
//
// Pipeline per call (all on the context stream, no host synchronisation):
//   prep    per atom: base grid point u, spline fraction, brick id, counting
//           sort of atoms by brick (PME_Atom_Near, PM_force.cpp:697-759);
//   spread  one workgroup per grid brick: the atoms of the brick and of its
//           +1 neighbours deposit their 4x4x4 B-spline stencils into an int64
//           fixed-point LDS tile, which is written once (no global atomics,
//           result independent of atom order; PME_Q_Spread :762-811);
//   FFT     LDS-resident line transforms (radix 4/2/3/5/7 Stockham): z R2C
//           and y forward line passes (small LDS, they run beside the
//           direct-space kernel); x forward + influence-function product +
//           energy/virial accumulation + x inverse in one kernel; then per
//           x-plane y inverse + z C2R in one kernel (hipFFT R2C/C2R +
//           PME_BCFQ + PME_Sum_Virial, :1100-1122, :814-823, :1021-1065).
//           Planes too large for LDS take separate y and z line passes;
//   gather  one thread per atom sums the 64 stencil points of the potential
//           grid in a fixed order (PME_Final :826-890) and adds the force in
//           fixed point.
#include <math.h>

#include <vector>

#include "snb_internal.h"

// ---------------------------------------------------------------------------
// FFT plans and line transforms
// ---------------------------------------------------------------------------
int snb_fft_make_plan(int n, SnbFftPlan* plan)
{
    if (n < 1 || n > SNB_FFT_MAX_N) return SNB_ERR_UNSUPPORTED;
    plan->n = n;
    plan->nstages = 0;
    int m = n;
    const int order[5] = {4, 2, 3, 5, 7};
    for (int k = 0; k < 5; k++)
        while (m % order[k] == 0)
        {
            if (plan->nstages >= SNB_FFT_MAX_STAGES) return SNB_ERR_UNSUPPORTED;
            plan->radix[plan->nstages++] = order[k];
            m /= order[k];
        }
    if (m != 1) return SNB_ERR_UNSUPPORTED;
    return SNB_OK;
}

__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// multiply by s*i, s = +-1
__device__ __forceinline__ float2 cmul_si(float2 a, float s) { return make_float2(-s * a.y, s * a.x); }

// In-register DFT of length R with exponent sign s (X_m = sum v_r e^{s 2 pi i r m / R}).
template <int R>
__device__ __forceinline__ void small_dft(float2* v, float s);

template <>
__device__ __forceinline__ void small_dft<2>(float2* v, float s)
{
    float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

template <>
__device__ __forceinline__ void small_dft<4>(float2* v, float s)
{
    float2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    float2 t2 = cadd(v[1], v[3]), t3 = cmul_si(csub(v[1], v[3]), s);
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3);
    v[3] = csub(t1, t3);
}

template <>
__device__ __forceinline__ void small_dft<3>(float2* v, float s)
{
    const float c = -0.5f, sn = 0.86602540378443865f;
    float2 t1 = cadd(v[1], v[2]), t2 = csub(v[1], v[2]);
    float2 m = make_float2(v[0].x + c * t1.x, v[0].y + c * t1.y);
    float2 r = cmul_si(make_float2(sn * t2.x, sn * t2.y), s);
    v[0] = cadd(v[0], t1);
    v[1] = cadd(m, r);
    v[2] = csub(m, r);
}

template <>
__device__ __forceinline__ void small_dft<5>(float2* v, float s)
{
    const float c1 = 0.30901699437494742f, c2 = -0.80901699437494742f;
    const float s1 = 0.95105651629515357f, s2 = 0.58778525229247313f;
    float2 t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
    float2 t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
    float2 a0 = v[0];
    float2 m1 = make_float2(a0.x + c1 * t1.x + c2 * t2.x, a0.y + c1 * t1.y + c2 * t2.y);
    float2 m2 = make_float2(a0.x + c2 * t1.x + c1 * t2.x, a0.y + c2 * t1.y + c1 * t2.y);
    float2 r1 = cmul_si(make_float2(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y), s);
    float2 r2 = cmul_si(make_float2(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y), s);
    v[0] = cadd(a0, cadd(t1, t2));
    v[1] = cadd(m1, r1);
    v[4] = csub(m1, r1);
    v[2] = cadd(m2, r2);
    v[3] = csub(m2, r2);
}

template <>
__device__ __forceinline__ void small_dft<7>(float2* v, float s)
{
    const float C[7] = {1.0f, 0.62348980185873353f, -0.22252093395631440f,
                        -0.90096886790241913f, -0.90096886790241913f,
                        -0.22252093395631440f, 0.62348980185873353f};
    const float S[7] = {0.0f, 0.78183148246802981f, 0.97492791218182361f,
                        0.43388373911755812f, -0.43388373911755812f,
                        -0.97492791218182361f, -0.78183148246802981f};
    float2 out[7];
    for (int m = 0; m < 7; m++)
    {
        float2 acc = v[0];
        for (int r = 1; r < 7; r++)
        {
            int q = (r * m) % 7;
            float2 w = make_float2(C[q], s * S[q]);
            acc = cadd(acc, cmul(v[r], w));
        }
        out[m] = acc;
    }
    for (int m = 0; m < 7; m++) v[m] = out[m];
}

// One Stockham stage for `lines` lines of length n stored contiguously
// (stride n) in src, result in dst.  tw = e^{-2 pi i k/n} table in LDS.
template <int R>
__device__ void stockham_stage(const float2* src, float2* dst, int n, int lines, int ns,
                               const float2* tw, float s)
{
    const int nr = n / R;
    const int total = lines * nr;
    for (int t = threadIdx.x; t < total; t += blockDim.x)
    {
        int line = t / nr, j = t - line * nr;
        const float2* in = src + line * n;
        float2* out = dst + line * n;
        int k = j % ns;
        float2 v[R];
#pragma unroll
        for (int r = 0; r < R; r++) v[r] = in[j + r * nr];
        if (ns > 1)
        {
            int step = n / (ns * R);  // twiddle w^{r k}, w = e^{s 2 pi i/(ns R)}
#pragma unroll
            for (int r = 1; r < R; r++)
            {
                float2 w = tw[(r * k * step) % n];
                w.y *= -s;  // table holds e^{-i..}; s = -1 keeps it, s = +1 conjugates
                v[r] = cmul(v[r], w);
            }
        }
        small_dft<R>(v, s);
        int idx = (j / ns) * ns * R + k;
#pragma unroll
        for (int r = 0; r < R; r++) out[idx + r * ns] = v[r];
    }
}

// Transforms `lines` contiguous lines in a (ping-pong with b); returns the
// buffer holding the result.
__device__ float2* fft_lines(float2* a, float2* b, const SnbFftPlan& plan, int lines,
                             const float2* tw, float s)
{
    int ns = 1;
    float2* src = a;
    float2* dst = b;
    for (int st = 0; st < plan.nstages; st++)
    {
        int R = plan.radix[st];
        switch (R)
        {
            case 4: stockham_stage<4>(src, dst, plan.n, lines, ns, tw, s); break;
            case 2: stockham_stage<2>(src, dst, plan.n, lines, ns, tw, s); break;
            case 3: stockham_stage<3>(src, dst, plan.n, lines, ns, tw, s); break;
            case 5: stockham_stage<5>(src, dst, plan.n, lines, ns, tw, s); break;
            default: stockham_stage<7>(src, dst, plan.n, lines, ns, tw, s); break;
        }
        __syncthreads();
        ns *= R;
        float2* t = src;
        src = dst;
        dst = t;
    }
    return src;
}

__device__ void load_twiddles(float2* dst, const float2* src, int n)
{
    for (int k = threadIdx.x; k < n; k += blockDim.x) dst[k] = src[k];
}

static int lines_per_block(int n)
{
    // two ping-pong buffers of lines*n float2 within ~32 KB of LDS, small
    // enough to share a compute unit with the direct-space kernel
    int l = 32768 / (16 * n);
    if (l > 64) l = 64;
    if (l < 1) l = 1;
    return l;
}

// z pass forward: real rows [nrow][fz] -> half spectra [nrow][fc]
__device__ __forceinline__ void k_fft_z_r2c_body(const float* __restrict__ real,
                                                   float2* __restrict__ spec, int nrow,
                                                   SnbFftPlan plan, int fc,
                                                   const float2* __restrict__ twg, int L,
    unsigned bid, unsigned nbid)
{
    extern __shared__ float2 smem[];
    int n = plan.n;
    float2* tw = smem;
    float2* a = smem + n;
    float2* b = a + L * n;
    load_twiddles(tw, twg, n);
    int row0 = bid * L;
    int lines = min(L, nrow - row0);
    for (int t = threadIdx.x; t < lines * n; t += blockDim.x)
        a[t] = make_float2(real[(size_t)row0 * n + t], 0.0f);
    __syncthreads();
    float2* res = fft_lines(a, b, plan, lines, tw, -1.0f);
    for (int t = threadIdx.x; t < lines * fc; t += blockDim.x)
    {
        int l = t / fc, k = t - l * fc;
        spec[(size_t)(row0 + l) * fc + k] = res[l * n + k];
    }
}

__global__ void __launch_bounds__(256) k_fft_z_r2c(const float* __restrict__ real,
                                                   float2* __restrict__ spec, int nrow,
                                                   SnbFftPlan plan, int fc,
                                                   const float2* __restrict__ twg, int L)
{
    k_fft_z_r2c_body(real, spec, nrow, plan, fc, twg, L, blockIdx.x, gridDim.x);
}

struct FftZr2cArgs
{
    const float* real;
    float2* spec;
    int nrow;
    SnbFftPlan plan;
    int fc;
    const float2* twg;
    int L;
};

__global__ void __launch_bounds__(256) k_fft_z_r2c_batch(SnbSegTab t)
{
    unsigned bid, nbid;
    const FftZr2cArgs& a = snb_seg<FftZr2cArgs>(t, bid, nbid);
    k_fft_z_r2c_body(a.real, a.spec, a.nrow, a.plan, a.fc, a.twg, a.L, bid, nbid);
}

// z pass inverse: half spectra [nrow][fc] -> real rows [nrow][fz]
__device__ __forceinline__ void k_fft_z_c2r_body(const float2* __restrict__ spec,
                                                   float* __restrict__ real, int nrow,
                                                   SnbFftPlan plan, int fc,
                                                   const float2* __restrict__ twg, int L,
    unsigned bid, unsigned nbid)
{
    extern __shared__ float2 smem[];
    int n = plan.n;
    float2* tw = smem;
    float2* a = smem + n;
    float2* b = a + L * n;
    load_twiddles(tw, twg, n);
    int row0 = bid * L;
    int lines = min(L, nrow - row0);
    for (int t = threadIdx.x; t < lines * n; t += blockDim.x)
    {
        int l = t / n, k = t - l * n;
        float2 v;
        if (k < fc)
            v = spec[(size_t)(row0 + l) * fc + k];
        else
        {
            v = spec[(size_t)(row0 + l) * fc + (n - k)];
            v.y = -v.y;
        }
        a[t] = v;
    }
    __syncthreads();
    float2* res = fft_lines(a, b, plan, lines, tw, 1.0f);
    for (int t = threadIdx.x; t < lines * n; t += blockDim.x)
        real[(size_t)row0 * n + t] = res[t].x;
}

__global__ void __launch_bounds__(256) k_fft_z_c2r(const float2* __restrict__ spec,
                                                   float* __restrict__ real, int nrow,
                                                   SnbFftPlan plan, int fc,
                                                   const float2* __restrict__ twg, int L)
{
    k_fft_z_c2r_body(spec, real, nrow, plan, fc, twg, L, blockIdx.x, gridDim.x);
}

struct FftZc2rArgs
{
    const float2* spec;
    float* real;
    int nrow;
    SnbFftPlan plan;
    int fc;
    const float2* twg;
    int L;
};

__global__ void __launch_bounds__(256) k_fft_z_c2r_batch(SnbSegTab t)
{
    unsigned bid, nbid;
    const FftZc2rArgs& a = snb_seg<FftZc2rArgs>(t, bid, nbid);
    k_fft_z_c2r_body(a.spec, a.real, a.nrow, a.plan, a.fc, a.twg, a.L, bid, nbid);
}

// y pass over spec[fx][fy][fc]: block = (x, chunk of W consecutive kz).
__device__ __forceinline__ void k_fft_y_body(float2* __restrict__ spec, int fx, int fy,
                                               int fc, SnbFftPlan plan,
                                               const float2* __restrict__ twg, int W, float s,
    unsigned bid, unsigned nbid)
{
    extern __shared__ float2 smem[];
    int n = plan.n;  // == fy
    float2* tw = smem;
    float2* a = smem + n;
    float2* b = a + W * n;
    load_twiddles(tw, twg, n);
    int nchunk = (fc + W - 1) / W;
    int x = bid / nchunk;
    int k0 = (bid - x * nchunk) * W;
    int w = min(W, fc - k0);
    for (int t = threadIdx.x; t < n * w; t += blockDim.x)
    {
        int y = t / w, l = t - y * w;
        a[l * n + y] = spec[((size_t)x * fy + y) * fc + k0 + l];
    }
    __syncthreads();
    float2* res = fft_lines(a, b, plan, w, tw, s);
    for (int t = threadIdx.x; t < n * w; t += blockDim.x)
    {
        int y = t / w, l = t - y * w;
        spec[((size_t)x * fy + y) * fc + k0 + l] = res[l * n + y];
    }
}

__global__ void __launch_bounds__(256) k_fft_y(float2* __restrict__ spec, int fx, int fy,
                                               int fc, SnbFftPlan plan,
                                               const float2* __restrict__ twg, int W, float s)
{
    k_fft_y_body(spec, fx, fy, fc, plan, twg, W, s, blockIdx.x, gridDim.x);
}

struct FftYArgs
{
    float2* spec;
    int fx;
    int fy;
    int fc;
    SnbFftPlan plan;
    const float2* twg;
    int W;
    float s;
};

__global__ void __launch_bounds__(256) k_fft_y_batch(SnbSegTab t)
{
    unsigned bid, nbid;
    const FftYArgs& a = snb_seg<FftYArgs>(t, bid, nbid);
    k_fft_y_body(a.spec, a.fx, a.fy, a.fc, a.plan, a.twg, a.W, a.s, bid, nbid);
}

// x pass with the convolution: block = (y, chunk of W consecutive kz).
// Forward FFT along x, then per element (kx, y, kz):
//   energy  += 1/2 * wfull * BC |F|^2        (wfull = 1 on kz = 0 and the
//                                               Nyquist plane, else 2)
//   virial  -= wref * 1/2 BC |F|^2 (delta - 2 (1 + pi^2 m^2/beta^2) m m/m^2)
//             (wref = 0.5 on kz = 0 and kz = fz/2, else 1: PM_force.cpp:1033-1037)
//   F *= BC
// then inverse FFT along x.  Per-block partial sums go to partials[block][8].
template <bool NEED_EV>
__device__ __forceinline__ void k_fft_x_conv_body(
    float2* __restrict__ spec, const float* __restrict__ bc, int fx, int fy, int fz, int fc,
    SnbFftPlan plan, const float2* __restrict__ twg, int W, double r11, double r21,
    double r22, double r31, double r32, double r33, double mpre, double* __restrict__ partials,
    unsigned bid, unsigned nbid)
{
    extern __shared__ float2 smem[];
    int n = plan.n;  // == fx
    float2* tw = smem;
    float2* a = smem + n;
    float2* b = a + W * n;
    load_twiddles(tw, twg, n);
    int nchunk = (fc + W - 1) / W;
    int y = bid / nchunk;
    int k0 = (bid - y * nchunk) * W;
    int w = min(W, fc - k0);
    for (int t = threadIdx.x; t < n * w; t += blockDim.x)
    {
        int x = t / w, l = t - x * w;
        a[l * n + x] = spec[((size_t)x * fy + y) * fc + k0 + l];
    }
    __syncthreads();
    float2* res = fft_lines(a, b, plan, w, tw, -1.0f);
    double e = 0, v0 = 0, v1 = 0, v2 = 0, v3 = 0, v4 = 0, v5 = 0;
    for (int t = threadIdx.x; t < n * w; t += blockDim.x)
    {
        int l = t / n, kx = t - l * n;
        int kz = k0 + l;
        float bcv = bc[((size_t)kx * fy + y) * fc + kz];
        float2 f = res[t];
        if (NEED_EV && bcv != 0.0f)
        {
            double a2 = (double)f.x * f.x + (double)f.y * f.y;
            double wfull = (kz == 0 || (2 * kz == fz)) ? 1.0 : 2.0;
            e += 0.5 * wfull * (double)bcv * a2;
            double wref = (kz == 0 || kz == fc - 1) ? 0.5 : 1.0;
            double kxr = kx > fx / 2 ? kx - fx : kx;
            double kyr = y > fy / 2 ? y - fy : y;
            double mx = kxr * r11;
            double my = kxr * r21 + kyr * r22;
            double mz = kxr * r31 + kyr * r32 + kz * r33;
            double msq = mx * mx + my * my + mz * mz;
            double cf = 2.0 / msq * (1.0 + mpre * msq);
            double hb = wref * 0.5 * (double)bcv * a2;
            v0 -= hb * (1.0 - cf * mx * mx);
            v1 -= hb * (0.0 - cf * my * mx);
            v2 -= hb * (1.0 - cf * my * my);
            v3 -= hb * (0.0 - cf * mz * mx);
            v4 -= hb * (0.0 - cf * mz * my);
            v5 -= hb * (1.0 - cf * mz * mz);
        }
        res[t] = make_float2(f.x * bcv, f.y * bcv);
    }
    __syncthreads();
    float2* other = (res == a) ? b : a;
    float2* res2 = fft_lines(res, other, plan, w, tw, 1.0f);
    for (int t = threadIdx.x; t < n * w; t += blockDim.x)
    {
        int x = t / w, l = t - x * w;
        spec[((size_t)x * fy + y) * fc + k0 + l] = res2[l * n + x];
    }
    if (NEED_EV)
    {
        __shared__ double red[8][4];
        double vals[7] = {e, v0, v1, v2, v3, v4, v5};
        int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
        for (int c = 0; c < 7; c++)
        {
            double v = vals[c];
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
            if (lane == 0) red[c][wid] = v;
        }
        __syncthreads();
        if (threadIdx.x < 7)
        {
            double s = 0;
            for (int q = 0; q < (int)(blockDim.x >> 6); q++) s += red[threadIdx.x][q];
            partials[(size_t)bid * 8 + threadIdx.x] = s;
        }
    }
}

template <bool NEED_EV>
__global__ void __launch_bounds__(256) k_fft_x_conv(
    float2* __restrict__ spec, const float* __restrict__ bc, int fx, int fy, int fz, int fc,
    SnbFftPlan plan, const float2* __restrict__ twg, int W, double r11, double r21,
    double r22, double r31, double r32, double r33, double mpre, double* __restrict__ partials)
{
    k_fft_x_conv_body<NEED_EV>(
        spec, bc, fx, fy, fz, fc, plan, twg, W, r11, r21, r22, r31, r32, r33, mpre, partials,
        blockIdx.x, gridDim.x);
}

struct FftXConvArgs
{
    float2* spec;
    const float* bc;
    int fx;
    int fy;
    int fz;
    int fc;
    SnbFftPlan plan;
    const float2* twg;
    int W;
    double r11;
    double r21;
    double r22;
    double r31;
    double r32;
    double r33;
    double mpre;
    double* partials;
};

template <bool NEED_EV>
__global__ void __launch_bounds__(256) k_fft_x_conv_batch(SnbSegTab t)
{
    unsigned bid, nbid;
    const FftXConvArgs& a = snb_seg<FftXConvArgs>(t, bid, nbid);
    k_fft_x_conv_body<NEED_EV>(
        a.spec, a.bc, a.fx, a.fy, a.fz, a.fc, a.plan, a.twg, a.W, a.r11, a.r21, a.r22, a.r31, a.r32,
        a.r33, a.mpre, a.partials, bid, nbid);
}

// ---------------------------------------------------------------------------
// Fused inverse plane pass: one workgroup per x-plane does the y and z
// transforms in LDS.  Two real z-rows travel as one complex row (row 2p as the
// real part, row 2p+1 as the imaginary part), which halves the z work.
// LDS: tw_z[fz] | tw_y[fy] | A[cap] | B[cap], cap = max(ceil(fy/2)*fz, fc*fy).
// ---------------------------------------------------------------------------
__host__ __device__ static size_t plane_cap(int fy, int fz, int fc)
{
    size_t a = (size_t)((fy + 1) / 2) * fz, b = (size_t)fc * fy;
    return a > b ? a : b;
}

static size_t plane_lds(int fy, int fz, int fc)
{
    return sizeof(float2) * ((size_t)fz + fy + 2 * plane_cap(fy, fz, fc));
}

// spec[x][fy][fc] -> real[x][fy][fz] (y inverse then z C2R), exponent +1
__device__ __forceinline__ void k_fft_plane_inv_body(const float2* __restrict__ spec,
                                                        float* __restrict__ real, int fy,
                                                        int fz, int fc, SnbFftPlan plan_z,
                                                        SnbFftPlan plan_y,
                                                        const float2* __restrict__ twzg,
                                                        const float2* __restrict__ twyg,
    unsigned bid, unsigned nbid)
{
    extern __shared__ float2 smem[];
    const int cap = (int)plane_cap(fy, fz, fc);
    float2* twz = smem;
    float2* twy = twz + fz;
    float2* A = twy + fy;
    float2* B = A + cap;
    load_twiddles(twz, twzg, fz);
    load_twiddles(twy, twyg, fy);
    const int x = bid;
    const int np = (fy + 1) / 2;
    const float2* in = spec + (size_t)x * fy * fc;
    for (int t = threadIdx.x; t < fy * fc; t += blockDim.x)
    {
        int y = t / fc, k = t - y * fc;
        A[k * fy + y] = in[t];
    }
    __syncthreads();
    float2* res = fft_lines(A, B, plan_y, fc, twy, 1.0f);
    float2* O = (res == A) ? B : A;
    // Hermitian extension of rows 2p and 2p+1 along z, packed as Ya + i Yb.
    // The imaginary parts at kz = 0 and at the Nyquist plane are dropped,
    // which is what taking the real part of a C2R output does.
    const bool nyq = (fz % 2) == 0;
    for (int t = threadIdx.x; t < np * fz; t += blockDim.x)
    {
        int p = t / fz, k = t - p * fz;
        int kk = k < fc ? k : fz - k;
        float2 a = res[kk * fy + 2 * p];
        float2 b = (2 * p + 1 < fy) ? res[kk * fy + 2 * p + 1] : make_float2(0.0f, 0.0f);
        if (k >= fc)
        {
            a.y = -a.y;
            b.y = -b.y;
        }
        if (k == 0 || (nyq && 2 * k == fz))
        {
            a.y = 0.0f;
            b.y = 0.0f;
        }
        O[t] = make_float2(a.x - b.y, a.y + b.x);
    }
    __syncthreads();
    float2* res2 = fft_lines(O, res, plan_z, np, twz, 1.0f);
    float* plane = real + (size_t)x * fy * fz;
    for (int t = threadIdx.x; t < np * fz; t += blockDim.x)
    {
        int p = t / fz, z = t - p * fz;
        float2 v = res2[t];
        plane[(size_t)(2 * p) * fz + z] = v.x;
        if (2 * p + 1 < fy) plane[(size_t)(2 * p + 1) * fz + z] = v.y;
    }
}

__global__ void __launch_bounds__(1024) k_fft_plane_inv(const float2* __restrict__ spec,
                                                        float* __restrict__ real, int fy,
                                                        int fz, int fc, SnbFftPlan plan_z,
                                                        SnbFftPlan plan_y,
                                                        const float2* __restrict__ twzg,
                                                        const float2* __restrict__ twyg)
{
    k_fft_plane_inv_body(spec, real, fy, fz, fc, plan_z, plan_y, twzg, twyg, blockIdx.x, gridDim.x);
}

struct PlaneInvArgs
{
    const float2* spec;
    float* real;
    int fy;
    int fz;
    int fc;
    SnbFftPlan plan_z;
    SnbFftPlan plan_y;
    const float2* twzg;
    const float2* twyg;
};

__global__ void __launch_bounds__(1024) k_fft_plane_inv_batch(SnbSegTab t)
{
    unsigned bid, nbid;
    const PlaneInvArgs& a = snb_seg<PlaneInvArgs>(t, bid, nbid);
    k_fft_plane_inv_body(
        a.spec, a.real, a.fy, a.fz, a.fc, a.plan_z, a.plan_y, a.twzg, a.twyg, bid, nbid);
}

// Fixed-order reduction of the per-block partials into dscal[0..6] and the
// self + neutralising energy into dscal[7] (PM_force.cpp:1146-1159).
__global__ void k_pme_finalize(const double* __restrict__ partials, int nblocks,
                               double* __restrict__ dscal, const int64_t* __restrict__ scal,
                               double beta, double nfactor, int want_self)
{
    // wavefront c sums component c of the per-workgroup partials in a fixed
    // order (lane strides, then a butterfly); wavefront 7 the self term
    const int lane = threadIdx.x & 63, c = threadIdx.x >> 6;
    if (c < 7 && nblocks > 0)
    {
        double s = 0;
        for (int k = lane; k < nblocks; k += 64) s += partials[(size_t)k * 8 + c];
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
        if (lane == 0) dscal[c] = s;
    }
    if (c == 7 && lane == 0 && want_self)
    {
        double q2 = snb_fixed_to_double(scal[SNB_S_Q2]);
        double q1 = snb_fixed_to_double(scal[SNB_S_Q]);
        dscal[7] = -beta / sqrt(3.14159265358979323846) * q2 + nfactor * q1 * q1;
    }
}

// ---------------------------------------------------------------------------
// atoms -> bricks, spread, gather
// ---------------------------------------------------------------------------
// PME_Atom_Near (PM_force.cpp:706-737) in the reference's float operations.
__device__ __forceinline__ void atom_near(float4 r, const float* rc, int fx, int fy, int fz,
                                          int& ux, int& uy, int& uz, float& frx, float& fry,
                                          float& frz)
{
    float f0 = (r.x * rc[0] + r.y * rc[1]) + r.z * rc[3];
    float f1 = r.y * rc[2] + r.z * rc[4];
    float f2 = r.z * rc[5];
    f0 = f0 - floorf(f0);
    f1 = f1 - floorf(f1);
    f2 = f2 - floorf(f2);
    if (!(__builtin_isfinite(f0) && __builtin_isfinite(f1) && __builtin_isfinite(f2)))
        f0 = f1 = f2 = 0.0f;
    float s;
    float q;
    s = f0 * (float)fx;
    ux = (int)s;
    ux = ux < 0 ? 0 : (ux < fx ? ux : fx - 1);
    q = s - (float)ux;
    frx = q - floorf(q);
    s = f1 * (float)fy;
    uy = (int)s;
    uy = uy < 0 ? 0 : (uy < fy ? uy : fy - 1);
    q = s - (float)uy;
    fry = q - floorf(q);
    s = f2 * (float)fz;
    uz = (int)s;
    uz = uz < 0 ? 0 : (uz < fz ? uz : fz - 1);
    q = s - (float)uz;
    frz = q - floorf(q);
}

struct SnbRc
{
    float v[6];
};

// Per atom: base grid point, spline fractions and brick (PME_Atom_Near), and
// its slot in the brick's list.  Atoms are visited in the brick order of the
// previous call (any permutation of 0..n-1 is correct; that one keeps each
// workgroup's atoms in few bricks, so the brick counters see few global
// atomics).  With LDS_HIST the workgroup counts its atoms per brick in LDS,
// lists the bricks it touched, and takes one global counter add per listed
// brick (a single round trip); otherwise one add per distinct brick of each
// wavefront.
#define SNB_PREP_HIST_MAX 8192
template <bool LDS_HIST>
__device__ __forceinline__ void k_pme_prep_body(int n, const snb_vector* __restrict__ crd,
                                                  const float* __restrict__ charge, SnbRc rc,
                                                  int fx, int fy, int fz, int bx, int by, int bz,
                                                  int nby, int nbz, int nbrick,
                                                  const int* __restrict__ visit,
                                                  int4* __restrict__ atom_u,
                                                  float4* __restrict__ atom_fr,
                                                  int* __restrict__ slot, int* __restrict__ cnt,
    unsigned bid, unsigned nbid)
{
    __shared__ int hist[LDS_HIST ? SNB_PREP_HIST_MAX : 1];
    __shared__ int touched[LDS_HIST ? 256 : 1];
    __shared__ int ntouched;
    const int v = bid * blockDim.x + threadIdx.x;
    const int i = v < n ? visit[v] : -1;
    if (LDS_HIST)
    {
        for (int b = threadIdx.x; b < nbrick; b += blockDim.x) hist[b] = 0;
        if (threadIdx.x == 0) ntouched = 0;
        __syncthreads();
    }
    int br = -1;
    if (i >= 0)
    {
        const snb_vector ri = crd[i];
        float4 r = make_float4(ri.x, ri.y, ri.z, charge[i]);
        int ux, uy, uz;
        float a, b, c;
        atom_near(r, rc.v, fx, fy, fz, ux, uy, uz, a, b, c);
        br = ((ux / bx) * nby + uy / by) * nbz + uz / bz;
        atom_u[i] = make_int4(ux, uy, uz, br);
        atom_fr[i] = make_float4(a, b, c, r.w);
    }
    if (LDS_HIST)
    {
        int local = 0;
        if (br >= 0)
        {
            local = atomicAdd(&hist[br], 1);
            if (local == 0) touched[atomicAdd(&ntouched, 1)] = br;
        }
        __syncthreads();
        if ((int)threadIdx.x < ntouched)
        {
            const int b = touched[threadIdx.x];
            hist[b] = atomicAdd(&cnt[b], hist[b]);
        }
        __syncthreads();
        if (br >= 0) slot[i] = hist[br] + local;
        return;
    }
    if (i < 0) return;
    const int lane = threadIdx.x & 63;
    const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    unsigned long long todo = __ballot(1);
    while (todo)
    {
        const int leader = __ffsll((long long)todo) - 1;
        const int b0 = __shfl(br, leader);
        const unsigned long long same = __ballot(br == b0) & todo;
        int base = 0;
        if (lane == leader) base = atomicAdd(&cnt[b0], __popcll(same));
        base = __shfl(base, leader);
        if (br == b0 && ((todo >> lane) & 1ull)) slot[i] = base + __popcll(same & below);
        todo &= ~same;
    }
}

template <bool LDS_HIST>
__global__ void __launch_bounds__(256) k_pme_prep(int n, const snb_vector* __restrict__ crd,
                                                  const float* __restrict__ charge, SnbRc rc,
                                                  int fx, int fy, int fz, int bx, int by, int bz,
                                                  int nby, int nbz, int nbrick,
                                                  const int* __restrict__ visit,
                                                  int4* __restrict__ atom_u,
                                                  float4* __restrict__ atom_fr,
                                                  int* __restrict__ slot, int* __restrict__ cnt)
{
    k_pme_prep_body<LDS_HIST>(
        n, crd, charge, rc, fx, fy, fz, bx, by, bz, nby, nbz, nbrick, visit, atom_u, atom_fr, slot,
        cnt, blockIdx.x, gridDim.x);
}

struct PmePrepArgs
{
    int n;
    const snb_vector* crd;
    const float* charge;
    SnbRc rc;
    int fx;
    int fy;
    int fz;
    int bx;
    int by;
    int bz;
    int nby;
    int nbz;
    int nbrick;
    const int* visit;
    int4* atom_u;
    float4* atom_fr;
    int* slot;
    int* cnt;
};

template <bool LDS_HIST>
__global__ void __launch_bounds__(256) k_pme_prep_batch(SnbSegTab t)
{
    unsigned bid, nbid;
    const PmePrepArgs& a = snb_seg<PmePrepArgs>(t, bid, nbid);
    k_pme_prep_body<LDS_HIST>(
        a.n, a.crd, a.charge, a.rc, a.fx, a.fy, a.fz, a.bx, a.by, a.bz, a.nby, a.nbz, a.nbrick,
        a.visit, a.atom_u, a.atom_fr, a.slot, a.cnt, bid, nbid);
}

__device__ __forceinline__ void k_pme_scatter_body(int n, const int4* __restrict__ atom_u, const int* __restrict__ slot,
                              const int* __restrict__ start, int* __restrict__ brick_atoms,
    unsigned bid, unsigned nbid)
{
    int i = bid * blockDim.x + threadIdx.x;
    if (i >= n) return;
    brick_atoms[start[atom_u[i].w] + slot[i]] = i;
}

__global__ void k_pme_scatter(int n, const int4* __restrict__ atom_u, const int* __restrict__ slot,
                              const int* __restrict__ start, int* __restrict__ brick_atoms)
{
    k_pme_scatter_body(n, atom_u, slot, start, brick_atoms, blockIdx.x, gridDim.x);
}

struct PmeScatterArgs
{
    int n;
    const int4* atom_u;
    const int* slot;
    const int* start;
    int* brick_atoms;
};

__global__ void k_pme_scatter_batch(SnbSegTab t)
{
    unsigned bid, nbid;
    const PmeScatterArgs& a = snb_seg<PmeScatterArgs>(t, bid, nbid);
    k_pme_scatter_body(a.n, a.atom_u, a.slot, a.start, a.brick_atoms, bid, nbid);
}

__device__ __forceinline__ int wrap_axis(int v, int n) { return v < 0 ? v + n : v; }

// The device dispatches workgroups to its 8 XCDs round-robin by index, and
// each XCD has its own L2.  A launch of xcd_grid(nblocks) workgroups maps
// onto blocks 0 .. nblocks - 1 so that XCD x takes the contiguous range
// [x * chunk, (x + 1) * chunk): neighbouring bricks (and atoms in brick
// order), which read the same grid points and atom records, share one L2.
// -1 for the padding workgroups.
#define SNB_XCDS 8
__device__ __forceinline__ int xcd_block(int nblocks, unsigned bid)
{
    const int chunk = (nblocks + SNB_XCDS - 1) / SNB_XCDS;
    const int b = (int)(bid % SNB_XCDS) * chunk + (int)(bid / SNB_XCDS);
    return b < nblocks ? b : -1;
}
static inline unsigned xcd_grid(int nblocks)
{
    return (unsigned)(SNB_XCDS * ((nblocks + SNB_XCDS - 1) / SNB_XCDS));
}

// One workgroup per brick of bx*by*bz grid points.  An atom with base point u
// touches points u-3..u on each axis, so a brick receives from atoms whose
// base lies in the brick itself or in the next brick on each axis (bricks are
// >= 4 points wide): 8 source bricks.
// Bits k (0..3) for which stencil point u - k of an atom at offset d = (u - o)
// mod n from a brick origin o falls in the brick [o, o + b): max(0, d - b + 1)
// <= k <= min(3, d); all four when the axis is one brick (it wraps inside).
__device__ __forceinline__ int stencil_mask(int d, int b, bool single)
{
    if (single) return 15;
    int k0 = d - b + 1;
    k0 = k0 < 0 ? 0 : (k0 > 4 ? 4 : k0);
    int k1 = d < 3 ? d : 3;
    return ((2 << k1) - 1) & ~((1 << k0) - 1);
}

// The workgroup takes the candidate atoms of the 8 source bricks in chunks of
// 256 (one per thread).  A candidate whose stencil reaches the brick gets an
// LDS record: its 12 spline weights (x weights times q 2^32, the fixed-point
// scale, which is exact) and its offsets and masks.  Then each thread owns a
// stencil column (atom, kx, ky) of 16 listed atoms at a time and adds its up
// to 4 z points to the int64 LDS tile: ((q wx) wy) wz per point, the
// reference's product order, rounded to fixed point.  For an atom at offset
// d, stencil point u - k lies in the brick when 0 <= d - k < b: only that k
// range is visited, so every stencil point is added by exactly one brick.
#define SNB_SPREAD_LIST 256
__device__ __forceinline__ void k_pme_spread_body(
    const int4* __restrict__ atom_u, const float4* __restrict__ atom_fr,
    const int* __restrict__ start, const int* __restrict__ brick_atoms, int fx, int fy, int fz,
    int bx, int by, int bz, int nbx, int nby, int nbz, float* __restrict__ q,
    unsigned bid, unsigned nbid)
{
    extern __shared__ long long acc[];
    const int npt = bx * by * bz;
    float* lw = (float*)(acc + npt);                // [SNB_SPREAD_LIST][12]
    int* lmeta = (int*)(lw + SNB_SPREAD_LIST * 12);  // [SNB_SPREAD_LIST]
    __shared__ int src_first[8], src_pref[9], nlist;
    const int t = threadIdx.x;
    const int B = xcd_block(nbx * nby * nbz, bid);
    if (B < 0) return;
    for (int p = t; p < npt; p += blockDim.x) acc[p] = 0;
    const int Bz = B % nbz, By = (B / nbz) % nby, Bx = B / (nbz * nby);
    const int ox = Bx * bx, oy = By * by, oz = Bz * bz;
    if (t < 8)
    {
        const int sx = t >> 2, sy = (t >> 1) & 1, sz = t & 1;
        int f = 0, c = 0;
        if (!((sx && nbx == 1) || (sy && nby == 1) || (sz && nbz == 1)))
        {
            int sb = (((Bx + sx) % nbx) * nby + (By + sy) % nby) * nbz + (Bz + sz) % nbz;
            f = start[sb];
            c = start[sb + 1] - f;
        }
        src_first[t] = f;
        src_pref[t + 1] = c;
    }
    __syncthreads();
    if (t == 0)
    {
        src_pref[0] = 0;
        for (int k = 1; k <= 8; k++) src_pref[k] += src_pref[k - 1];
        nlist = 0;
    }
    __syncthreads();
    const int ncand = src_pref[8];
    const int kx = (t >> 2) & 3, ky = t & 3;
    for (int c0 = 0; c0 < ncand; c0 += blockDim.x)
    {
        const int ci = c0 + t;
        if (ci < ncand)
        {
            int sidx = 0;
            while (sidx < 7 && src_pref[sidx + 1] <= ci) sidx++;
            const int i = brick_atoms[src_first[sidx] + (ci - src_pref[sidx])];
            const int4 u = atom_u[i];
            int dx = u.x - ox, dy = u.y - oy, dz = u.z - oz;
            dx += dx < 0 ? fx : 0;
            dy += dy < 0 ? fy : 0;
            dz += dz < 0 ? fz : 0;
            const int mx = stencil_mask(dx, bx, nbx == 1);
            const int my = stencil_mask(dy, by, nby == 1);
            const int mz = stencil_mask(dz, bz, nbz == 1);
            if (mx != 0 && my != 0 && mz != 0)
            {
                // dx < bx + 3 <= 19 (or < fx <= 16 on a one-brick axis): 5 bits
                const float4 fr = atom_fr[i];
                const int slot = atomicAdd(&nlist, 1);
                const float q2 = fr.w * 4294967296.0f;
                float* w = lw + slot * 12;
#pragma unroll
                for (int k = 0; k < 4; k++)
                {
                    w[k] = q2 * snb_bspline_w(k, fr.x);
                    w[4 + k] = snb_bspline_w(k, fr.y);
                    w[8 + k] = snb_bspline_w(k, fr.z);
                }
                lmeta[slot] = dx | (dy << 5) | (dz << 10) | (mx << 15) | (my << 19) | (mz << 23);
            }
        }
        __syncthreads();
        const int nl = nlist;
        for (int a = t >> 4; a < nl; a += 16)
        {
            const int meta = lmeta[a];
            const int mx = (meta >> 15) & 15, my = (meta >> 19) & 15;
            if (!((mx >> kx) & 1) || !((my >> ky) & 1)) continue;
            const int dx = meta & 31, dy = (meta >> 5) & 31, dz = (meta >> 10) & 31;
            const int mz = (meta >> 23) & 15;
            const float* w = lw + a * 12;
            const float qxy = w[kx] * w[4 + ky];
            int lx = dx - kx, ly = dy - ky;
            lx += lx < 0 ? fx : 0;
            ly += ly < 0 ? fy : 0;
            const int row = (lx * by + ly) * bz;
#pragma unroll
            for (int kz = 0; kz < 4; kz++)
            {
                if (!((mz >> kz) & 1)) continue;
                int lz = dz - kz;
                lz += lz < 0 ? fz : 0;
                atomicAdd((unsigned long long*)&acc[row + lz],
                          (unsigned long long)snb_fixed_from_scaled(qxy * w[8 + kz]));
            }
        }
        __syncthreads();
        if (t == 0) nlist = 0;
        __syncthreads();
    }
    for (int p = t; p < npt; p += blockDim.x)
    {
        int lz = p % bz, ly = (p / bz) % by, lx = p / (bz * by);
        q[((size_t)(ox + lx) * fy + (oy + ly)) * fz + (oz + lz)] = snb_fixed_to_float(acc[p]);
    }
}

__global__ void __launch_bounds__(256) k_pme_spread(
    const int4* __restrict__ atom_u, const float4* __restrict__ atom_fr,
    const int* __restrict__ start, const int* __restrict__ brick_atoms, int fx, int fy, int fz,
    int bx, int by, int bz, int nbx, int nby, int nbz, float* __restrict__ q)
{
    k_pme_spread_body(
        atom_u, atom_fr, start, brick_atoms, fx, fy, fz, bx, by, bz, nbx, nby, nbz, q, blockIdx.x,
        gridDim.x);
}

struct PmeSpreadArgs
{
    const int4* atom_u;
    const float4* atom_fr;
    const int* start;
    const int* brick_atoms;
    int fx;
    int fy;
    int fz;
    int bx;
    int by;
    int bz;
    int nbx;
    int nby;
    int nbz;
    float* q;
};

__global__ void __launch_bounds__(256) k_pme_spread_batch(SnbSegTab t)
{
    unsigned bid, nbid;
    const PmeSpreadArgs& a = snb_seg<PmeSpreadArgs>(t, bid, nbid);
    k_pme_spread_body(
        a.atom_u, a.atom_fr, a.start, a.brick_atoms, a.fx, a.fy, a.fz, a.bx, a.by, a.bz, a.nbx,
        a.nby, a.nbz, a.q, bid, nbid);
}

// PME_Final + device_add_force (PM_force.cpp:826-890, :1011-1019): one thread
// per atom, stencil order kx, ky, kz (outer to inner), s = fma(phi, g, s).
// Atoms are visited in brick order (the lists of this call), so the lanes of a
// wavefront read neighbouring grid points.
__device__ __forceinline__ void k_pme_gather_body(int n, const int* __restrict__ visit,
                                                    const int4* __restrict__ atom_u,
                                                    const float4* __restrict__ atom_fr,
                                                    const float* __restrict__ phi, int fx,
                                                    int fy, int fz, SnbRc rc, float upd,
                                                    int64_t* __restrict__ frc,
    unsigned bid, unsigned nbid)
{
    const int blk = xcd_block((n + (int)blockDim.x - 1) / (int)blockDim.x, bid);
    if (blk < 0) return;
    const int v = blk * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const int i = visit[v];
    int4 u = atom_u[i];
    float4 fr = atom_fr[i];
    float wx[4], wy[4], wz[4], dx[4], dy[4], dz[4];
#pragma unroll
    for (int k = 0; k < 4; k++)
    {
        wx[k] = snb_bspline_w(k, fr.x);
        wy[k] = snb_bspline_w(k, fr.y);
        wz[k] = snb_bspline_w(k, fr.z);
        dx[k] = snb_bspline_dw(k, fr.x);
        dy[k] = snb_bspline_dw(k, fr.y);
        dz[k] = snb_bspline_dw(k, fr.z);
    }
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
#pragma unroll
    for (int kx = 0; kx < 4; kx++)
    {
        int gx = wrap_axis(u.x - kx, fx);
#pragma unroll
        for (int ky = 0; ky < 4; ky++)
        {
            int gy = wrap_axis(u.y - ky, fy);
            const float* rowp = phi + ((size_t)gx * fy + gy) * fz;
#pragma unroll
            for (int kz = 0; kz < 4; kz++)
            {
                int gz = wrap_axis(u.z - kz, fz);
                float ph = rowp[gz];
                sx = __builtin_fmaf(ph, (dx[kx] * wy[ky]) * wz[kz], sx);
                sy = __builtin_fmaf(ph, (wx[kx] * dy[ky]) * wz[kz], sy);
                sz = __builtin_fmaf(ph, (wx[kx] * wy[ky]) * dz[kz], sz);
            }
        }
    }
    float tx = sx * (float)fx, ty = sy * (float)fy, tz = sz * (float)fz;
    float mq = -fr.w * upd;
    float gx = tx * rc.v[0];
    float gy = tx * rc.v[1] + ty * rc.v[2];
    float gz = (tx * rc.v[3] + ty * rc.v[4]) + tz * rc.v[5];
    frc[i] += snb_to_fixed(mq * gx);
    frc[n + i] += snb_to_fixed(mq * gy);
    frc[2 * n + i] += snb_to_fixed(mq * gz);
}

__global__ void __launch_bounds__(256) k_pme_gather(int n, const int* __restrict__ visit,
                                                    const int4* __restrict__ atom_u,
                                                    const float4* __restrict__ atom_fr,
                                                    const float* __restrict__ phi, int fx,
                                                    int fy, int fz, SnbRc rc, float upd,
                                                    int64_t* __restrict__ frc)
{
    k_pme_gather_body(
        n, visit, atom_u, atom_fr, phi, fx, fy, fz, rc, upd, frc, blockIdx.x, gridDim.x);
}

struct PmeGatherArgs
{
    int n;
    const int* visit;
    const int4* atom_u;
    const float4* atom_fr;
    const float* phi;
    int fx;
    int fy;
    int fz;
    SnbRc rc;
    float upd;
    int64_t* frc;
};

__global__ void __launch_bounds__(256) k_pme_gather_batch(SnbSegTab t)
{
    unsigned bid, nbid;
    const PmeGatherArgs& a = snb_seg<PmeGatherArgs>(t, bid, nbid);
    k_pme_gather_body(
        a.n, a.visit, a.atom_u, a.atom_fr, a.phi, a.fx, a.fy, a.fz, a.rc, a.upd, a.frc, bid, nbid);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static double bspline_modulus(int m, int K)
{
    // PM_force.cpp:27-62, order 4: |sum_{k=0}^{2} M4(k+1) e^{2 pi i m k/K}|^-2
    const double M4[3] = {1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0};
    double re = 0, im = 0;
    for (int k = 0; k < 3; k++)
    {
        double a = 2.0 * M_PI * m * k / K;
        re += M4[k] * cos(a);
        im += M4[k] * sin(a);
    }
    return 1.0 / (re * re + im * im);
}

static int brick_dim(int n)
{
    // a divisor of n in [4, 16], nearest to 8; the whole axis if n <= 16
    if (n <= 16) return n;
    int best = 0;
    for (int d = 4; d <= 16; d++)
        if (n % d == 0 && (best == 0 || abs(d - 8) < abs(best - 8))) best = d;
    return best;
}

static int upload_twiddles(int n, float2** dst)
{
    std::vector<float2> h(n);
    for (int k = 0; k < n; k++)
    {
        double a = -2.0 * M_PI * k / n;
        h[k] = make_float2((float)cos(a), (float)sin(a));
    }
    SNB_CHECK(hipMalloc(dst, sizeof(float2) * n));
    SNB_CHECK(hipMemcpy(*dst, h.data(), sizeof(float2) * n, hipMemcpyHostToDevice));
    return SNB_OK;
}

// lines per workgroup for the y and x passes (a chunk of consecutive kz)
static int fft_w(int fc, int n)
{
    int w = lines_per_block(n);
    if (w > 16) w = 16;
    return fc < w ? fc : w;
}

// Influence function BC[kx][ky][kz] over the half spectrum (PM_force.cpp:544-612,
// recomputed for a new box by up_box_bc, :1168-1237):
//   BC = exp(-pi^2 m^2 / beta^2) / (pi m^2 V) * b1[kx] b2[ky] b3[kz],  BC[0] = 0,
// with m = k' rcell^T (k' the signed index), in FP64, stored as float.
// One block per (kx, ky) line, threads over kz.
__global__ void k_pme_bc(int fx, int fy, int fc, const double* __restrict__ bmod, SnbBox box,
                         double mpre, double volume, float* __restrict__ bc)
{
    const int kx = blockIdx.x / fy, ky = blockIdx.x % fy;
    const double kxr = kx > fx / 2 ? kx - fx : kx;
    const double kyr = ky > fy / 2 ? ky - fy : ky;
    const double bxy = bmod[kx] * bmod[fx + ky];
    for (int kz = threadIdx.x; kz < fc; kz += blockDim.x)
    {
        const double mx = kxr * box.rc[0];
        const double my = kxr * box.rc[1] + kyr * box.rc[2];
        const double mz = kxr * box.rc[3] + kyr * box.rc[4] + kz * box.rc[5];
        const double msq = mx * mx + my * my + mz * mz;
        const size_t idx = ((size_t)kx * fy + ky) * fc + kz;
        bc[idx] = (kx + ky + kz == 0)
                      ? 0.0f
                      : (float)(exp(-mpre * msq) / (M_PI * msq * volume) * bxy *
                                bmod[fx + fy + kz]);
    }
}

int snb_pme_compute_bc(snb_context* ctx, const SnbBox& box)
{
    SnbPme& P = ctx->pme;
    const double beta = P.beta;
    const double mpre = M_PI * M_PI / (beta * beta);
    SNB_LAUNCH(k_pme_bc, dim3(P.fx * P.fy), dim3(64), 0,
               ctx->stream, P.fx, P.fy, P.fc, P.bmod, box, mpre, P.volume, P.bc);
    SNB_CHECK(hipGetLastError());
    for (int c = 0; c < 6; c++) P.rcell[c] = box.rc[c];
    return SNB_OK;
}

int snb_pme_setup_tables(snb_context* ctx, const SnbBox& box)
{
    SnbPme& P = ctx->pme;
    int rc;
    if ((rc = snb_fft_make_plan(P.fx, &P.plan_x))) return rc;
    if ((rc = snb_fft_make_plan(P.fy, &P.plan_y))) return rc;
    if ((rc = snb_fft_make_plan(P.fz, &P.plan_z))) return rc;
    P.fc = P.fz / 2 + 1;
    P.bx = brick_dim(P.fx);
    P.by = brick_dim(P.fy);
    P.bz = brick_dim(P.fz);
    if (!P.bx || !P.by || !P.bz) return SNB_ERR_UNSUPPORTED;
    if ((P.bx < 4 && P.bx != P.fx) || (P.by < 4 && P.by != P.fy) || (P.bz < 4 && P.bz != P.fz))
        return SNB_ERR_UNSUPPORTED;
    P.nbx = P.fx / P.bx;
    P.nby = P.fy / P.by;
    P.nbz = P.fz / P.bz;
    P.nbrick = P.nbx * P.nby * P.nbz;
    size_t nall = (size_t)P.fx * P.fy * P.fz;
    size_t nfft = (size_t)P.fx * P.fy * P.fc;
    int n = ctx->n;
    if ((rc = upload_twiddles(P.fx, &P.tw_x))) return rc;
    if ((rc = upload_twiddles(P.fy, &P.tw_y))) return rc;
    if ((rc = upload_twiddles(P.fz, &P.tw_z))) return rc;
    SNB_CHECK(hipMalloc(&P.bc, sizeof(float) * nfft));
    SNB_CHECK(hipMalloc(&P.q, sizeof(float) * nall));
    SNB_CHECK(hipMalloc(&P.phi, sizeof(float) * nall));
    SNB_CHECK(hipMalloc(&P.spec, sizeof(float2) * nfft));
    SNB_CHECK(hipMalloc(&P.atom_u, sizeof(int4) * (n > 0 ? n : 1)));
    SNB_CHECK(hipMalloc(&P.atom_fr, sizeof(float4) * (n > 0 ? n : 1)));
    SNB_CHECK(hipMalloc(&P.atom_slot, sizeof(int) * (n > 0 ? n : 1)));
    SNB_CHECK(hipMalloc(&P.brick_atoms, sizeof(int) * (n > 0 ? n : 1)));
    {
        // the prep kernel visits atoms in the previous call's brick order;
        // the first call visits them in index order
        std::vector<int> iota(n > 0 ? n : 1);
        for (size_t k = 0; k < iota.size(); k++) iota[k] = (int)k;
        SNB_CHECK(hipMemcpy(P.brick_atoms, iota.data(), sizeof(int) * iota.size(),
                            hipMemcpyHostToDevice));
    }
    SNB_CHECK(hipMalloc(&P.brick_count, sizeof(int) * (P.nbrick + 1)));
    SNB_CHECK(hipMemset(P.brick_count, 0, sizeof(int) * (P.nbrick + 1)));
    SNB_CHECK(hipMalloc(&P.brick_start, sizeof(int) * (P.nbrick + 1)));
    int W = fft_w(P.fc, P.fx);
    P.conv_blocks = P.fy * ((P.fc + W - 1) / W);
    SNB_CHECK(hipMalloc(&P.partials, sizeof(double) * 8 * P.conv_blocks));
    // B-spline moduli |b(k)|^2 per axis (PM_force.cpp:40-62), then BC on the device
    std::vector<double> bm(P.fx + P.fy + P.fz);
    for (int k = 0; k < P.fx; k++) bm[k] = bspline_modulus(k, P.fx);
    for (int k = 0; k < P.fy; k++) bm[P.fx + k] = bspline_modulus(k, P.fy);
    for (int k = 0; k < P.fz; k++) bm[P.fx + P.fy + k] = bspline_modulus(k, P.fz);
    SNB_CHECK(hipMalloc(&P.bmod, sizeof(double) * bm.size()));
    SNB_CHECK(hipMemcpy(P.bmod, bm.data(), sizeof(double) * bm.size(), hipMemcpyHostToDevice));
    return snb_pme_compute_bc(ctx, box);
}

static size_t lds_lines(int n, int L) { return sizeof(float2) * ((size_t)n + 2 * (size_t)L * n); }

// Whether one x-plane of the grid fits the fused plane kernels' LDS; grants
// the dynamic LDS on first use.
static int plane_ok(snb_context* ctx, bool* ok)
{
    SnbPme& P = ctx->pme;
    size_t lds = plane_lds(P.fy, P.fz, P.fc);
    *ok = lds <= (size_t)ctx->max_lds;
    if (*ok && P.plane_lds_set != lds)
    {
        SNB_CHECK(hipFuncSetAttribute((const void*)k_fft_plane_inv,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        P.plane_lds_set = lds;
    }
    return SNB_OK;
}

// y inverse + z C2R: spec[fx][fy][fc] -> real[fx][fy][fz]
static int fft_yz_inverse(snb_context* ctx, float2* cplx, float* real)
{
    SnbPme& P = ctx->pme;
    hipStream_t s = ctx->stream;
    bool ok;
    int rc = plane_ok(ctx, &ok);
    if (rc) return rc;
    if (ok)
    {
        SNB_LAUNCH_B(k_fft_plane_inv, k_fft_plane_inv_batch, PlaneInvArgs, dim3(P.fx), dim3(1024),
                     plane_lds(P.fy, P.fz, P.fc), s, cplx, real, P.fy, P.fz, P.fc, P.plan_z,
                     P.plan_y, P.tw_z, P.tw_y);
        SNB_CHECK(hipGetLastError());
        return SNB_OK;
    }
    int W = fft_w(P.fc, P.fy);
    int nchunk = (P.fc + W - 1) / W;
    SNB_LAUNCH_B(k_fft_y, k_fft_y_batch, FftYArgs, dim3(P.fx * nchunk), dim3(256),
                 lds_lines(P.fy, W), s, cplx, P.fx, P.fy, P.fc, P.plan_y, P.tw_y, W, 1.0f);
    int nrow = P.fx * P.fy;
    int Lz = lines_per_block(P.fz);
    SNB_LAUNCH_B(k_fft_z_c2r, k_fft_z_c2r_batch, FftZc2rArgs, dim3((nrow + Lz - 1) / Lz),
                 dim3(256), lds_lines(P.fz, Lz), s, cplx, real, nrow, P.plan_z, P.fc, P.tw_z, Lz);
    SNB_CHECK(hipGetLastError());
    return SNB_OK;
}

int snb_launch_fft_r2c(snb_context* ctx, const float* real, float2* cplx)
{
    SnbPme& P = ctx->pme;
    hipStream_t s = ctx->stream;
    // the forward transform runs while the direct-space kernel holds most of
    // the LDS: line passes with small LDS share compute units with it
    int nrow = P.fx * P.fy;
    int Lz = lines_per_block(P.fz);
    SNB_LAUNCH_B(k_fft_z_r2c, k_fft_z_r2c_batch, FftZr2cArgs, dim3((nrow + Lz - 1) / Lz), dim3(256),
                 lds_lines(P.fz, Lz), s, real, cplx, nrow, P.plan_z, P.fc, P.tw_z, Lz);
    int W = fft_w(P.fc, P.fy);
    int nchunk = (P.fc + W - 1) / W;
    SNB_LAUNCH_B(k_fft_y, k_fft_y_batch, FftYArgs, dim3(P.fx * nchunk), dim3(256),
                 lds_lines(P.fy, W), s, cplx, P.fx, P.fy, P.fc, P.plan_y, P.tw_y, W, -1.0f);
    SNB_CHECK(hipGetLastError());
    return SNB_OK;
}

// x-direction plain transform (no convolution), used by the FFT wrappers.
__global__ void __launch_bounds__(256) k_fft_x(float2* __restrict__ spec, int fx, int fy, int fc,
                                               SnbFftPlan plan, const float2* __restrict__ twg,
                                               int W, float s)
{
    extern __shared__ float2 smem[];
    int n = plan.n;
    float2* tw = smem;
    float2* a = smem + n;
    float2* b = a + W * n;
    load_twiddles(tw, twg, n);
    int nchunk = (fc + W - 1) / W;
    int y = blockIdx.x / nchunk;
    int k0 = (blockIdx.x - y * nchunk) * W;
    int w = min(W, fc - k0);
    for (int t = threadIdx.x; t < n * w; t += blockDim.x)
    {
        int x = t / w, l = t - x * w;
        a[l * n + x] = spec[((size_t)x * fy + y) * fc + k0 + l];
    }
    __syncthreads();
    float2* res = fft_lines(a, b, plan, w, tw, s);
    for (int t = threadIdx.x; t < n * w; t += blockDim.x)
    {
        int x = t / w, l = t - x * w;
        spec[((size_t)x * fy + y) * fc + k0 + l] = res[l * n + x];
    }
}

static int fft_x_plain(snb_context* ctx, float2* cplx, float sign)
{
    SnbPme& P = ctx->pme;
    int W = fft_w(P.fc, P.fx);
    int nchunk = (P.fc + W - 1) / W;
    SNB_LAUNCH(k_fft_x, dim3(P.fy * nchunk), dim3(256), lds_lines(P.fx, W),
               ctx->stream, cplx, P.fx, P.fy, P.fc, P.plan_x, P.tw_x, W, sign);
    SNB_CHECK(hipGetLastError());
    return SNB_OK;
}

int snb_launch_fft_c2r(snb_context* ctx, const float2* cplx_in, float* real)
{
    SnbPme& P = ctx->pme;
    hipStream_t s = ctx->stream;
    size_t nfft = (size_t)P.fx * P.fy * P.fc;
    SNB_CHECK(hipMemcpyAsync(P.spec, cplx_in, sizeof(float2) * nfft, hipMemcpyDeviceToDevice, s));
    int rc = fft_x_plain(ctx, P.spec, 1.0f);
    if (rc) return rc;
    return fft_yz_inverse(ctx, P.spec, real);
}

int snb_fft_forward_full(snb_context* ctx, const float* real, float2* cplx)
{
    int rc = snb_launch_fft_r2c(ctx, real, cplx);
    if (rc) return rc;
    return fft_x_plain(ctx, cplx, -1.0f);
}

int snb_launch_pme_recip(snb_context* ctx, const snb_vector* crd, const float* charge,
                         const SnbBox& box, int need_energy, int need_virial, int with_force)
{
    SnbPme& P = ctx->pme;
    hipStream_t s = ctx->stream;
    int n = ctx->n;
    int rc;
    SnbRc r6;
    for (int c = 0; c < 6; c++) r6.v[c] = box.rc[c];
    snb_prof_begin(ctx, SNB_PROF_PME);
    if (n > 0)
    {
        // brick counts are zero on entry (cleared at set-up and by the scan
        // of the previous call)
        if (P.nbrick <= SNB_PREP_HIST_MAX)
            SNB_LAUNCH_B(k_pme_prep<true>, k_pme_prep_batch<true>, PmePrepArgs,
                         dim3((n + 255) / 256), dim3(256), 0, s, n, crd, charge, r6, P.fx, P.fy,
                         P.fz, P.bx, P.by, P.bz, P.nby, P.nbz, P.nbrick, P.brick_atoms, P.atom_u,
                         P.atom_fr, P.atom_slot, P.brick_count);
        else
            SNB_LAUNCH_B(k_pme_prep<false>, k_pme_prep_batch<false>, PmePrepArgs,
                         dim3((n + 255) / 256), dim3(256), 0, s, n, crd, charge, r6, P.fx, P.fy,
                         P.fz, P.bx, P.by, P.bz, P.nby, P.nbz, P.nbrick, P.brick_atoms, P.atom_u,
                         P.atom_fr, P.atom_slot, P.brick_count);
        if ((rc = snb_launch_exclusive_scan_clear(s, P.brick_count, P.brick_start, P.nbrick)))
            return rc;
        SNB_LAUNCH_B(k_pme_scatter, k_pme_scatter_batch, PmeScatterArgs, dim3((n + 255) / 256),
                     dim3(256), 0, s, n, P.atom_u, P.atom_slot, P.brick_start, P.brick_atoms);
    }
    else
    {
        if ((rc = snb_launch_fill_int(s, P.brick_start, P.nbrick + 1, 0))) return rc;
    }
    size_t lds = sizeof(long long) * (size_t)P.bx * P.by * P.bz + 13 * sizeof(float) * SNB_SPREAD_LIST;
    SNB_LAUNCH_B(k_pme_spread, k_pme_spread_batch, PmeSpreadArgs, dim3(xcd_grid(P.nbrick)),
                 dim3(256), lds, s, P.atom_u, P.atom_fr, P.brick_start, P.brick_atoms, P.fx, P.fy,
                 P.fz, P.bx, P.by, P.bz, P.nbx, P.nby, P.nbz, P.q);
    SNB_CHECK(hipGetLastError());
    if ((rc = snb_launch_fft_r2c(ctx, P.q, P.spec))) return rc;
    int W = fft_w(P.fc, P.fx);
    int nchunk = (P.fc + W - 1) / W;
    double mpre = M_PI * M_PI / ((double)P.beta * P.beta);
    bool ev = need_energy || need_virial;
    if (ev)
        SNB_LAUNCH_B((k_fft_x_conv<true>), (k_fft_x_conv_batch<true>), FftXConvArgs,
                     dim3(P.fy * nchunk), dim3(256), lds_lines(P.fx, W), s, P.spec, P.bc, P.fx,
                     P.fy, P.fz, P.fc, P.plan_x, P.tw_x, W, (double)box.rc[0], (double)box.rc[1],
                     (double)box.rc[2], (double)box.rc[3], (double)box.rc[4], (double)box.rc[5],
                     mpre, P.partials);
    else
        SNB_LAUNCH_B((k_fft_x_conv<false>), (k_fft_x_conv_batch<false>), FftXConvArgs,
                     dim3(P.fy * nchunk), dim3(256), lds_lines(P.fx, W), s, P.spec, P.bc, P.fx,
                     P.fy, P.fz, P.fc, P.plan_x, P.tw_x, W, (double)box.rc[0], (double)box.rc[1],
                     (double)box.rc[2], (double)box.rc[3], (double)box.rc[4], (double)box.rc[5],
                     mpre, P.partials);
    SNB_CHECK(hipGetLastError());
    if ((rc = fft_yz_inverse(ctx, P.spec, P.phi))) return rc;
    if (n > 0 && with_force)
        SNB_LAUNCH_B(k_pme_gather, k_pme_gather_batch, PmeGatherArgs,
                     dim3(xcd_grid((n + 255) / 256)), dim3(256), 0, s, n, P.brick_atoms, P.atom_u,
                     P.atom_fr, P.phi, P.fx, P.fy, P.fz, r6, (float)P.update_interval, ctx->fx_pme);
    snb_prof_end(ctx, SNB_PROF_PME);
    SNB_CHECK(hipGetLastError());
    if (ev)
    {
        double nfactor = -0.5 * M_PI / ((double)P.beta * P.beta * P.volume);
        SNB_LAUNCH(k_pme_finalize, dim3(1), dim3(512), 0, s, P.partials, P.conv_blocks,
                   ctx->dscal, ctx->scal, (double)P.beta, nfactor, need_energy ? 1 : 0);
        SNB_CHECK(hipGetLastError());
    }
    return SNB_OK;
}

int snb_launch_pme_self(snb_context* ctx, const float* charge)
{
    SnbPme& P = ctx->pme;
    int rc = snb_launch_charge_sums(ctx, charge);
    if (rc) return rc;
    double nfactor = -0.5 * M_PI / ((double)P.beta * P.beta * P.volume);
    SNB_LAUNCH(k_pme_finalize, dim3(1), dim3(512), 0,
               ctx->stream, P.partials, 0, ctx->dscal, ctx->scal, (double)P.beta, nfactor, 1);
    SNB_CHECK(hipGetLastError());
    return SNB_OK;
}

// d_potential[0] += E_rec + E_self (PME_Add_Energy_To_Potential, PM_force.cpp:1004-1009)
__global__ void k_pme_add_potential(const double* __restrict__ dscal, float* pot)
{
    pot[0] += (float)(dscal[SNB_D_E_REC] + dscal[SNB_D_E_SELF]);
}

int snb_launch_pme_energy_to_potential(snb_context* ctx, float* potential)
{
    SNB_LAUNCH(k_pme_add_potential, dim3(1), dim3(1), 0, ctx->stream, ctx->dscal, potential);
    SNB_CHECK(hipGetLastError());
    return SNB_OK;
}

__global__ void k_pme_add_virial(const double* __restrict__ dscal, snb_ltmatrix3* vir)
{
    vir[0].a11 += (float)dscal[SNB_D_VIR0 + 0];
    vir[0].a21 += (float)dscal[SNB_D_VIR0 + 1];
    vir[0].a22 += (float)dscal[SNB_D_VIR0 + 2];
    vir[0].a31 += (float)dscal[SNB_D_VIR0 + 3];
    vir[0].a32 += (float)dscal[SNB_D_VIR0 + 4];
    vir[0].a33 += (float)dscal[SNB_D_VIR0 + 5];
}

int snb_launch_pme_virial_to(snb_context* ctx, snb_ltmatrix3* vir)
{
    SNB_LAUNCH(k_pme_add_virial, dim3(1), dim3(1), 0, ctx->stream, ctx->dscal, vir);
    SNB_CHECK(hipGetLastError());
    return SNB_OK;
}
