// snb_common.hip -- per-atom packing, fills, scans and the fixed-point
// boundary conversions of libsponge_nb.
#include "snb_internal.h"

// Copy_Crd_And_Charge_To_New_Crd (Lennard_Jones_force.cpp:17-27): packs the
// caller's AoS VECTOR coordinates and charges into one 16-byte record per
// atom, so a pair gather is a single float4 load.  Also raises the sticky
// non-finite flag.
__device__ __forceinline__ void pack_atoms_body(int n, const snb_vector* __restrict__ crd,
                                                const float* __restrict__ charge,
                                                float4* __restrict__ out,
                                                int* __restrict__ status, unsigned bid)
{
    int i = bid * blockDim.x + threadIdx.x;
    if (i >= n) return;
    snb_vector r = crd[i];
    out[i] = make_float4(r.x, r.y, r.z, charge[i]);
    if (!(__builtin_isfinite(r.x) && __builtin_isfinite(r.y) && __builtin_isfinite(r.z)))
        atomicOr(status, SNB_FLAG_NONFINITE_CRD);
}

__global__ void k_pack_atoms(int n, const snb_vector* __restrict__ crd,
                             const float* __restrict__ charge, float4* __restrict__ out,
                             int* __restrict__ status)
{
    pack_atoms_body(n, crd, charge, out, status, blockIdx.x);
}

struct PackAtomsArgs
{
    int n;
    const snb_vector* crd;
    const float* charge;
    float4* out;
    int* status;
};

__global__ void k_pack_atoms_batch(SnbSegTab t)
{
    unsigned bid, nbid;
    const PackAtomsArgs& a = snb_seg<PackAtomsArgs>(t, bid, nbid);
    pack_atoms_body(a.n, a.crd, a.charge, a.out, a.status, bid);
}

int snb_launch_pack_atoms(snb_context* ctx, const snb_vector* crd, const float* charge)
{
    int n = ctx->n;
    if (n <= 0) return SNB_OK;
    SNB_LAUNCH_B(k_pack_atoms, k_pack_atoms_batch, PackAtomsArgs, dim3((n + 255) / 256), dim3(256),
                 0, ctx->stream, n, crd, charge, ctx->atom4, ctx->status);
    SNB_CHECK(hipGetLastError());
    return SNB_OK;
}

__global__ void k_fill_int(int* p, int n, int v)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill_i64(int64_t* p, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

int snb_launch_fill_int(hipStream_t s, int* p, int n, int v)
{
    if (n <= 0) return SNB_OK;
    SNB_LAUNCH(k_fill_int, dim3((n + 255) / 256), dim3(256), 0, s, p, n, v);
    SNB_CHECK(hipGetLastError());
    return SNB_OK;
}

__global__ void k_add_i64(const int64_t* __restrict__ a, const int64_t* __restrict__ b,
                          int64_t* __restrict__ out, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = a[i] + b[i];
}

int snb_launch_add_i64(hipStream_t s, const int64_t* a, const int64_t* b, int64_t* out, size_t n)
{
    if (n == 0) return SNB_OK;
    SNB_LAUNCH(k_add_i64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, b, out, n);
    SNB_CHECK(hipGetLastError());
    return SNB_OK;
}

int snb_launch_fill_i64(hipStream_t s, int64_t* p, size_t n)
{
    if (n == 0) return SNB_OK;
    SNB_LAUNCH(k_fill_i64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n);
    SNB_CHECK(hipGetLastError());
    return SNB_OK;
}

// Exclusive scan of n ints by one 1024-thread workgroup: each thread scans a
// contiguous chunk serially, the chunk totals are scanned in LDS.  Used for
// cell and brick offsets (n <= a few 1e5).  CLEAR zeroes the input after use
// (counters that must start the next call at zero).
template <bool CLEAR>
__device__ __forceinline__ void exclusive_scan_body(int* __restrict__ in, int* __restrict__ out,
                                                    int n)
{
    __shared__ int part[1024];
    int t = threadIdx.x;
    int chunk = (n + 1023) / 1024;
    int b = t * chunk, e = min(b + chunk, n);
    int s = 0;
    for (int i = b; i < e; i++) s += in[i];
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1)
    {
        int v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int run = t > 0 ? part[t - 1] : 0;
    for (int i = b; i < e; i++)
    {
        int v = in[i];
        out[i] = run;
        run += v;
        if (CLEAR) in[i] = 0;
    }
    if (t == 1023) out[n] = part[1023];
}

template <bool CLEAR>
__global__ void __launch_bounds__(1024) k_exclusive_scan(int* __restrict__ in,
                                                         int* __restrict__ out, int n)
{
    exclusive_scan_body<CLEAR>(in, out, n);
}

struct ScanArgs
{
    int* in;
    int* out;
    int n;
};

// one workgroup per context
template <bool CLEAR>
__global__ void __launch_bounds__(1024) k_exclusive_scan_batch(SnbSegTab t)
{
    unsigned bid, nbid;
    const ScanArgs& a = snb_seg<ScanArgs>(t, bid, nbid);
    exclusive_scan_body<CLEAR>(a.in, a.out, a.n);
}

int snb_launch_exclusive_scan(hipStream_t s, const int* in, int* out, int n)
{
    SNB_LAUNCH_B(k_exclusive_scan<false>, k_exclusive_scan_batch<false>, ScanArgs, dim3(1),
                 dim3(1024), 0, s, (int*)in, out, n);
    SNB_CHECK(hipGetLastError());
    return SNB_OK;
}

int snb_launch_exclusive_scan_clear(hipStream_t s, int* in, int* out, int n)
{
    SNB_LAUNCH_B(k_exclusive_scan<true>, k_exclusive_scan_batch<true>, ScanArgs, dim3(1),
                 dim3(1024), 0, s, in, out, n);
    SNB_CHECK(hipGetLastError());
    return SNB_OK;
}

// frc[i] += fixed[i] * 2^-32 and clears the accumulator.
__global__ void k_flush_forces(int n, int64_t* __restrict__ fx, int64_t* __restrict__ fx_pme,
                               snb_vector* __restrict__ frc)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    snb_vector f = frc[i];
    f.x += snb_fixed_to_float(fx[i] + fx_pme[i]);
    f.y += snb_fixed_to_float(fx[n + i] + fx_pme[n + i]);
    f.z += snb_fixed_to_float(fx[2 * n + i] + fx_pme[2 * n + i]);
    frc[i] = f;
    fx[i] = 0;
    fx[n + i] = 0;
    fx[2 * n + i] = 0;
    fx_pme[i] = 0;
    fx_pme[n + i] = 0;
    fx_pme[2 * n + i] = 0;
}

int snb_launch_flush_forces(snb_context* ctx, snb_vector* frc)
{
    int n = ctx->n;
    if (n <= 0) return SNB_OK;
    int rc = snb_join_pme(ctx);
    if (rc) return rc;
    SNB_LAUNCH(k_flush_forces, dim3((n + 255) / 256), dim3(256), 0,
               ctx->stream, n, ctx->fx, ctx->fx_pme, frc);
    SNB_CHECK(hipGetLastError());
    return SNB_OK;
}

// atom_energy[i] += sum of the selected per-atom terms; direct[i] = the
// direct Coulomb term (the reference memsets and refills it each energy step,
// Lennard_Jones_force.cpp:376-382).  which_mask bit k selects e_atom[k].
__global__ void k_flush_atom_energy(int n, const int64_t* __restrict__ e, float* atom_energy,
                                    float* direct, int which_mask)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int64_t s = 0;
    for (int k = 0; k < 5; k++)  // LJ | direct | exclusion correction | 1-4 | bond
        if (which_mask & (1 << k)) s += e[(size_t)k * n + i];
    if (atom_energy) atom_energy[i] += snb_fixed_to_float(s);
    if (direct) direct[i] = snb_fixed_to_float(e[(size_t)n + i]);
}

int snb_launch_flush_atom_energy(snb_context* ctx, float* atom_energy, float* direct,
                                 int which_mask)
{
    int n = ctx->n;
    if (n <= 0 || (!atom_energy && !direct)) return SNB_OK;
    SNB_LAUNCH(k_flush_atom_energy, dim3((n + 255) / 256), dim3(256), 0,
               ctx->stream, n, ctx->e_atom, atom_energy, direct, which_mask);
    SNB_CHECK(hipGetLastError());
    return SNB_OK;
}

// atom_virial[i] += fixed virial, accumulator cleared.
__global__ void k_flush_virial(int n, int64_t* __restrict__ v, snb_ltmatrix3* __restrict__ out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    snb_ltmatrix3 m = out[i];
    m.a11 += snb_fixed_to_float(v[i]);
    m.a21 += snb_fixed_to_float(v[(size_t)n + i]);
    m.a22 += snb_fixed_to_float(v[(size_t)2 * n + i]);
    m.a31 += snb_fixed_to_float(v[(size_t)3 * n + i]);
    m.a32 += snb_fixed_to_float(v[(size_t)4 * n + i]);
    m.a33 += snb_fixed_to_float(v[(size_t)5 * n + i]);
    out[i] = m;
    for (int c = 0; c < 6; c++) v[(size_t)c * n + i] = 0;
}

int snb_launch_flush_virial(snb_context* ctx, snb_ltmatrix3* atom_virial)
{
    int n = ctx->n;
    if (n <= 0 || !atom_virial) return SNB_OK;
    SNB_LAUNCH(k_flush_virial, dim3((n + 255) / 256), dim3(256), 0,
               ctx->stream, n, ctx->vir_atom, atom_virial);
    SNB_CHECK(hipGetLastError());
    return SNB_OK;
}

// Exact sum of an int64 array into scal[slot] (integer adds commute, so the
// atomics leave a deterministic result).
__global__ void k_sum_fixed(const int64_t* __restrict__ a, int n, int64_t* __restrict__ dst)
{
    int64_t s = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        s += a[i];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if ((threadIdx.x & 63) == 0 && s != 0)
        atomicAdd((unsigned long long*)dst, (unsigned long long)s);
}

int snb_launch_sum_fixed(snb_context* ctx, const int64_t* arr, int n, int slot)
{
    SNB_LAUNCH(k_fill_i64, dim3(1), dim3(64), 0, ctx->stream, ctx->scal + slot, (size_t)1);
    if (n <= 0) return SNB_OK;
    int blocks = (n + 1023) / 1024;
    if (blocks > 256) blocks = 256;
    SNB_LAUNCH(k_sum_fixed, dim3(blocks), dim3(256), 0, ctx->stream, arr, n, ctx->scal + slot);
    SNB_CHECK(hipGetLastError());
    return SNB_OK;
}

// sum q^2 and sum q in fixed point (charge_square_kernel + Sum_Of_List,
// PM_force.cpp:280-289, :1146-1156).
__global__ void k_charge_sums(int n, const float* __restrict__ q, int64_t* __restrict__ dst)
{
    int64_t s2 = 0, s1 = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    {
        double v = q[i];
        s2 += snb_to_fixed_d(v * v);  // exact product of two floats
        s1 += snb_to_fixed_d(v);
    }
    for (int off = 32; off > 0; off >>= 1)
    {
        s2 += __shfl_xor(s2, off);
        s1 += __shfl_xor(s1, off);
    }
    // one add per workgroup and word
    __shared__ int64_t part[2][16];
    const int wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
    if ((threadIdx.x & 63) == 0)
    {
        part[0][wave] = s2;
        part[1][wave] = s1;
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        for (int w = 1; w < nwave; w++)
        {
            s2 += part[0][w];
            s1 += part[1][w];
        }
        atomicAdd((unsigned long long*)(dst + SNB_S_Q2), (unsigned long long)s2);
        atomicAdd((unsigned long long*)(dst + SNB_S_Q), (unsigned long long)s1);
    }
}

int snb_launch_charge_sums(snb_context* ctx, const float* charge)
{
    SNB_LAUNCH(k_fill_i64, dim3(1), dim3(64), 0, ctx->stream, ctx->scal + SNB_S_Q2, (size_t)2);
    int n = ctx->n;
    if (n <= 0) return SNB_OK;
    int blocks = (n + 1023) / 1024;
    if (blocks > 256) blocks = 256;
    SNB_LAUNCH(k_charge_sums, dim3(blocks), dim3(256), 0, ctx->stream, n, charge, ctx->scal);
    SNB_CHECK(hipGetLastError());
    return SNB_OK;
}

__global__ void k_add_scalar(float* dst, float v) { dst[0] += v; }

int snb_launch_add_scalar(hipStream_t s, float* dst, float v)
{
    SNB_LAUNCH(k_add_scalar, dim3(1), dim3(1), 0, s, dst, v);
    SNB_CHECK(hipGetLastError());
    return SNB_OK;
}

// Long_Range_Virial_Correction and device_add (Lennard_Jones_force.cpp:312-318, :38-41).
__global__ void k_lj_long_range(snb_ltmatrix3* vir, float* pot, float vir_add, float pot_add)
{
    if (vir)
    {
        vir[0].a11 += vir_add;
        vir[0].a22 += vir_add;
        vir[0].a33 += vir_add;
    }
    if (pot) pot[0] += pot_add;
}

int snb_launch_lj_long_range(hipStream_t s, snb_ltmatrix3* vir, float* pot, float vir_add,
                             float pot_add)
{
    SNB_LAUNCH(k_lj_long_range, dim3(1), dim3(1), 0, s, vir, pot, vir_add, pot_add);
    SNB_CHECK(hipGetLastError());
    return SNB_OK;
}
