// snb_pair.hip -- LJ + erfc-screened Coulomb over the half neighbor list, and
// the PME exclusion correction.
//
// Pair kernel (stands in for Lennard_Jones_And_Direct_Coulomb_Device,
// Lennard_Jones_force.cpp:43-136): one wavefront per row i of the reference
// layout half list; 64 lanes stride the row (coalesced 256-byte reads of the
// padded int table), gather j's packed {x, y, z, q} record and LJ type, and
// evaluate the pair in float with the library's own exp/erfc/rsqrt.  Every
// per-pair term is rounded to int64 fixed point before it is added, so the
// i-sums (wave reduction) and the j-scatter (64-bit integer atomics) give the
// same bits in any order.  Energies and virials of a pair go to row atom i,
// as in the reference; they are written once per row with plain stores.
#include "snb_internal.h"

__device__ __forceinline__ void pbc_disp(float& dx, float& dy, float& dz, const SnbBox& box)
{
    float sx = (dx * box.rc[0] + dy * box.rc[1]) + dz * box.rc[3];
    float sy = dy * box.rc[2] + dz * box.rc[4];
    float sz = dz * box.rc[5];
    float kx = floorf(sx + 0.5f), ky = floorf(sy + 0.5f), kz = floorf(sz + 0.5f);
    dx = dx - ((kx * box.c[0] + ky * box.c[1]) + kz * box.c[3]);
    dy = dy - (ky * box.c[2] + kz * box.c[4]);
    dz = dz - kz * box.c[5];
}

// Element i of an array addressed with a 32-bit byte offset: the load can use
// the scalar-base + unsigned vector-offset form (the offset is computed in 32
// bits, which holds for every array of this library).
template <class T>
__device__ __forceinline__ T ld32(const T* __restrict__ base, unsigned i)
{
    return *(const T*)((const char*)base + i * (unsigned)sizeof(T));
}

__device__ __forceinline__ int64_t wave_sum(int64_t v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

__device__ __forceinline__ void atomic_add_i64(int64_t* p, int64_t v)
{
    atomicAdd((unsigned long long*)p, (unsigned long long)v);
}

// Force factor of one pair (Lennard_Jones_force.cpp:73-97): LJ plus the
// erfc-screened Coulomb term, float arithmetic of the library's own math,
// the Coulomb shape from the Hermite table (ctab, SNB_COUL_N float4 pieces).
// A, B and qq may carry a common power-of-two scale; the result then carries
// it exactly (every operation is a product or difference of scaled terms).
__device__ __forceinline__ float pair_force_factor(float r2, float A, float B, float qq,
                                                   float beta, float cscale,
                                                   const float4* ctab)
{
    float rinv = snb_rsqrtf(r2);
    float r = r2 * rinv;
    float rinv2 = rinv * rinv;
    float rinv6 = (rinv2 * rinv2) * rinv2;
    float flj = ((B - A * rinv6) * rinv6) * rinv2;
    float t;
    int k = snb_coul_piece(beta * r, cscale, &t);
    float4 c = ctab[k];
    float fc = (qq * (rinv2 * rinv)) * snb_coul_eval(t, c.x, c.y, c.z, c.w);
    return flj - fc;
}

template <bool NEED_ENERGY, bool NEED_VIRIAL>
__device__ __forceinline__ void k_lj_direct_body(
    int n, const snb_atom_group* __restrict__ nl, const float4* __restrict__ atoms,
    const int* __restrict__ type, const float* __restrict__ lj_a,
    const float* __restrict__ lj_b, const float4* __restrict__ coul_tab, float cscale,
    SnbBox box, float rc2, float beta,
    int64_t* __restrict__ fx, int64_t* __restrict__ e_atom, int64_t* __restrict__ vir,
    unsigned bid, unsigned nbid)
{
    const int lane = threadIdx.x & 63;
    const int waves = nbid * 4;
    for (int i = bid * 4 + (threadIdx.x >> 6); i < n; i += waves)
    {
        const snb_atom_group g = nl[i];
        const float4 ai = atoms[i];
        const int ti = type[i];
        int64_t sx = 0, sy = 0, sz = 0, elj = 0, ec = 0;
        int64_t v0 = 0, v1 = 0, v2 = 0, v3 = 0, v4 = 0, v5 = 0;
        for (int k = lane; k < g.atom_numbers; k += 64)
        {
            const int j = g.atom_serial[k];
            const float4 aj = atoms[j];
            // Get_Periodic_Displacement(r_j, r_i) (Lennard_Jones_force.cpp:72)
            float dx = aj.x - ai.x, dy = aj.y - ai.y, dz = aj.z - ai.z;
            pbc_disp(dx, dy, dz, box);
            float r2 = (dx * dx + dy * dy) + dz * dz;
            if (!(r2 < rc2)) continue;
            const int p = snb_lj_pair_index(ti, type[j]);
            const float A = lj_a[p], B = lj_b[p];
            const float fa = pair_force_factor(r2, A, B, ai.w * aj.w, beta, cscale, coul_tab);
            float fxv = fa * dx, fyv = fa * dy, fzv = fa * dz;
            int64_t ix = snb_to_fixed(fxv), iy = snb_to_fixed(fyv), iz = snb_to_fixed(fzv);
            sx += ix;
            sy += iy;
            sz += iz;
            atomic_add_i64(fx + j, -ix);
            atomic_add_i64(fx + n + j, -iy);
            atomic_add_i64(fx + 2 * n + j, -iz);
            if (NEED_ENERGY)
            {
                // energies in FP64 from the (exact) float displacement
                double ddx = dx, ddy = dy, ddz = dz;
                double r2d = (ddx * ddx + ddy * ddy) + ddz * ddz;
                double rd = __builtin_sqrt(r2d);
                double rid = 1.0 / rd;
                double ri2 = rid * rid;
                double ri6 = (ri2 * ri2) * ri2;
                elj += snb_to_fixed_d((((double)A * (1.0 / 12.0)) * ri6 - (double)B * (1.0 / 6.0)) *
                                      ri6);
                double gd;
                const double erfcd = snb_erfc_parts_d((double)beta * rd, &gd);
                ec += snb_to_fixed_d((((double)ai.w * (double)aj.w) * erfcd) * rid);
            }
            if (NEED_VIRIAL)
            {
                v0 += snb_to_fixed(-(fxv * dx));
                v1 += snb_to_fixed(-(fxv * dy + fyv * dx));
                v2 += snb_to_fixed(-(fyv * dy));
                v3 += snb_to_fixed(-(fxv * dz + fzv * dx));
                v4 += snb_to_fixed(-(fyv * dz + fzv * dy));
                v5 += snb_to_fixed(-(fzv * dz));
            }
        }
        sx = wave_sum(sx);
        sy = wave_sum(sy);
        sz = wave_sum(sz);
        if (NEED_ENERGY)
        {
            elj = wave_sum(elj);
            ec = wave_sum(ec);
        }
        if (NEED_VIRIAL)
        {
            v0 = wave_sum(v0); v1 = wave_sum(v1); v2 = wave_sum(v2);
            v3 = wave_sum(v3); v4 = wave_sum(v4); v5 = wave_sum(v5);
        }
        if (lane == 0)
        {
            atomic_add_i64(fx + i, sx);
            atomic_add_i64(fx + n + i, sy);
            atomic_add_i64(fx + 2 * n + i, sz);
            if (NEED_ENERGY)
            {
                e_atom[i] += elj;
                e_atom[(size_t)n + i] += ec;
            }
            if (NEED_VIRIAL)
            {
                vir[i] += v0;
                vir[(size_t)n + i] += v1;
                vir[(size_t)2 * n + i] += v2;
                vir[(size_t)3 * n + i] += v3;
                vir[(size_t)4 * n + i] += v4;
                vir[(size_t)5 * n + i] += v5;
            }
        }
    }
}

template <bool NEED_ENERGY, bool NEED_VIRIAL>
__global__ void __launch_bounds__(256) k_lj_direct(
    int n, const snb_atom_group* __restrict__ nl, const float4* __restrict__ atoms,
    const int* __restrict__ type, const float* __restrict__ lj_a,
    const float* __restrict__ lj_b, const float4* __restrict__ coul_tab, float cscale,
    SnbBox box, float rc2, float beta,
    int64_t* __restrict__ fx, int64_t* __restrict__ e_atom, int64_t* __restrict__ vir)
{
    k_lj_direct_body<NEED_ENERGY, NEED_VIRIAL>(
        n, nl, atoms, type, lj_a, lj_b, coul_tab, cscale, box, rc2, beta, fx, e_atom, vir,
        blockIdx.x, gridDim.x);
}

struct LjDirectArgs
{
    int n;
    const snb_atom_group* nl;
    const float4* atoms;
    const int* type;
    const float* lj_a;
    const float* lj_b;
    const float4* coul_tab;
    float cscale;
    SnbBox box;
    float rc2;
    float beta;
    int64_t* fx;
    int64_t* e_atom;
    int64_t* vir;
};

template <bool NEED_ENERGY, bool NEED_VIRIAL>
__global__ void __launch_bounds__(256) k_lj_direct_batch(SnbSegTab t)
{
    unsigned bid, nbid;
    const LjDirectArgs& a = snb_seg<LjDirectArgs>(t, bid, nbid);
    k_lj_direct_body<NEED_ENERGY, NEED_VIRIAL>(
        a.n, a.nl, a.atoms, a.type, a.lj_a, a.lj_b, a.coul_tab, a.cscale, a.box, a.rc2, a.beta,
        a.fx, a.e_atom, a.vir, bid, nbid);
}

// Force-only steps of the row kernel for small systems: every atom's force
// accumulators in LDS (3 x n int64, n <= SNB_LJ_LDS_MAX), each workgroup a
// run of SNB_LJ_LDS_ROWS consecutive rows, and one global add per nonzero
// sum at the end.  The same integer sums as the global atomics of
// k_lj_direct, with about a quarter of the global atomics: where several
// small systems share the device (a step batch), the atomics' throughput,
// not their latency, bounds the row kernel.
#define SNB_LJ_LDS_MAX 2560
#define SNB_LJ_LDS_ROWS 64
__device__ __forceinline__ void k_lj_lds_body(
    int n, const snb_atom_group* __restrict__ nl, const float4* __restrict__ atoms,
    const int* __restrict__ type, const float* __restrict__ lj_a,
    const float* __restrict__ lj_b, const float4* __restrict__ coul_tab, float cscale,
    SnbBox box, float rc2, float beta, int64_t* __restrict__ fx, unsigned bid, unsigned nbid)
{
    extern __shared__ long long acc[];  // [3][n]
    for (int t = threadIdx.x; t < 3 * n; t += blockDim.x) acc[t] = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
    const int r0 = (int)bid * SNB_LJ_LDS_ROWS;
    const int r1 = r0 + SNB_LJ_LDS_ROWS < n ? r0 + SNB_LJ_LDS_ROWS : n;
    for (int i = r0 + wave; i < r1; i += nwave)
    {
        const snb_atom_group g = nl[i];
        const float4 ai = atoms[i];
        const int ti = type[i];
        int64_t sx = 0, sy = 0, sz = 0;
        for (int k = lane; k < g.atom_numbers; k += 64)
        {
            const int j = g.atom_serial[k];
            const float4 aj = atoms[j];
            float dx = aj.x - ai.x, dy = aj.y - ai.y, dz = aj.z - ai.z;
            pbc_disp(dx, dy, dz, box);
            float r2 = (dx * dx + dy * dy) + dz * dz;
            if (!(r2 < rc2)) continue;
            const int p = snb_lj_pair_index(ti, type[j]);
            const float fa =
                pair_force_factor(r2, lj_a[p], lj_b[p], ai.w * aj.w, beta, cscale, coul_tab);
            const int64_t ix = snb_to_fixed(fa * dx), iy = snb_to_fixed(fa * dy),
                          iz = snb_to_fixed(fa * dz);
            sx += ix;
            sy += iy;
            sz += iz;
            atomicAdd((unsigned long long*)&acc[j], (unsigned long long)(-ix));
            atomicAdd((unsigned long long*)&acc[n + j], (unsigned long long)(-iy));
            atomicAdd((unsigned long long*)&acc[2 * n + j], (unsigned long long)(-iz));
        }
        sx = wave_sum(sx);
        sy = wave_sum(sy);
        sz = wave_sum(sz);
        if (lane == 0)
        {
            atomicAdd((unsigned long long*)&acc[i], (unsigned long long)sx);
            atomicAdd((unsigned long long*)&acc[n + i], (unsigned long long)sy);
            atomicAdd((unsigned long long*)&acc[2 * n + i], (unsigned long long)sz);
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < n; t += blockDim.x)
    {
        const long long vx = acc[t], vy = acc[n + t], vz = acc[2 * n + t];
        if (vx) atomic_add_i64(fx + t, vx);
        if (vy) atomic_add_i64(fx + n + t, vy);
        if (vz) atomic_add_i64(fx + 2 * n + t, vz);
    }
}

__global__ void __launch_bounds__(1024) k_lj_lds(
    int n, const snb_atom_group* __restrict__ nl, const float4* __restrict__ atoms,
    const int* __restrict__ type, const float* __restrict__ lj_a,
    const float* __restrict__ lj_b, const float4* __restrict__ coul_tab, float cscale,
    SnbBox box, float rc2, float beta, int64_t* __restrict__ fx)
{
    k_lj_lds_body(n, nl, atoms, type, lj_a, lj_b, coul_tab, cscale, box, rc2, beta, fx,
                  blockIdx.x, gridDim.x);
}

struct LjLdsArgs
{
    int n;
    const snb_atom_group* nl;
    const float4* atoms;
    const int* type;
    const float* lj_a;
    const float* lj_b;
    const float4* coul_tab;
    float cscale;
    SnbBox box;
    float rc2;
    float beta;
    int64_t* fx;
};

__global__ void __launch_bounds__(1024) k_lj_lds_batch(SnbSegTab t)
{
    unsigned bid, nbid;
    const LjLdsArgs& a = snb_seg<LjLdsArgs>(t, bid, nbid);
    k_lj_lds_body(a.n, a.nl, a.atoms, a.type, a.lj_a, a.lj_b, a.coul_tab, a.cscale, a.box,
                  a.rc2, a.beta, a.fx, bid, nbid);
}

int snb_launch_lj_direct(snb_context* ctx, const snb_atom_group* nl, const SnbBox& box,
                         float beta, int need_energy, int need_virial)
{
    int n = ctx->n;
    if (n <= 0) return SNB_OK;
    float rc2 = ctx->lj.cutoff * ctx->lj.cutoff;
    // LDS accumulators: on request, or by default inside a recorded step batch
    const int lds_mode = ctx->lj.row_lds;
    if (!need_energy && !need_virial && n <= SNB_LJ_LDS_MAX &&
        (lds_mode > 0 || (lds_mode < 0 && snb_rec)))
    {
        snb_prof_begin(ctx, SNB_PROF_PAIR);
        SNB_LAUNCH_B(k_lj_lds, k_lj_lds_batch, LjLdsArgs,
                     dim3((n + SNB_LJ_LDS_ROWS - 1) / SNB_LJ_LDS_ROWS), dim3(1024),
                     sizeof(long long) * 3 * (size_t)n, ctx->stream, n, nl, ctx->atom4,
                     ctx->type, ctx->lj.a, ctx->lj.b, ctx->lj.coul, ctx->lj.coul_scale, box, rc2,
                     beta, ctx->fx);
        snb_prof_end(ctx, SNB_PROF_PAIR);
        SNB_CHECK(hipGetLastError());
        return SNB_OK;
    }
    int blocks = (n + 3) / 4;
    if (blocks > 65535 * 4) blocks = 65535 * 4;
    snb_prof_begin(ctx, SNB_PROF_PAIR);
#define SNB_LJ_LAUNCH(E, V)                                                                   \
    SNB_LAUNCH_B((k_lj_direct<E, V>), (k_lj_direct_batch<E, V>), LjDirectArgs, dim3(blocks),   \
                 dim3(256), 0, ctx->stream, n, nl, ctx->atom4, ctx->type, ctx->lj.a, ctx->lj.b, \
                 ctx->lj.coul, ctx->lj.coul_scale, box, rc2, beta, ctx->fx, ctx->e_atom,        \
                 ctx->vir_atom)
    if (need_energy && need_virial)
        SNB_LJ_LAUNCH(true, true);
    else if (need_energy)
        SNB_LJ_LAUNCH(true, false);
    else if (need_virial)
        SNB_LJ_LAUNCH(false, true);
    else
        SNB_LJ_LAUNCH(false, false);
#undef SNB_LJ_LAUNCH
    snb_prof_end(ctx, SNB_PROF_PAIR);
    SNB_CHECK(hipGetLastError());
    return SNB_OK;
}

// MPI_PME_Excluded_Force_With_Atom_Energy_Correction (PM_force.cpp:1743-1814)
// with exclude_factor 1: for j in excl(i), f = qq (erf(x) - 2x/sqrt(pi)
// e^{-x^2}) r^-3 dr on i and -f on j, energy -qq erf(x)/r and virial to i,
// evaluated in FP64.  One thread per atom i; i-sums in registers, j-terms by
// integer atomics (a few per atom).
template <bool NEED_ENERGY, bool NEED_VIRIAL>
__device__ __forceinline__ void k_excl_body(int n, const snb_vector* __restrict__ crd,
                                              const float* __restrict__ charge,
                                              SnbBox box, float beta,
                                              const int* __restrict__ ex_start,
                                              const int* __restrict__ ex_list,
                                              const int* __restrict__ ex_num,
                                              int64_t* __restrict__ fx,
                                              int64_t* __restrict__ e_atom,
                                              int64_t* __restrict__ vir,
    unsigned bid, unsigned nbid)
{
    int i = bid * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int cnt = ex_num[i];
    if (cnt <= 0) return;
    int s0 = ex_start[i];
    const snb_vector ri = crd[i];
    const float4 ai = make_float4(ri.x, ri.y, ri.z, charge[i]);
    int64_t sx = 0, sy = 0, sz = 0, e = 0;
    int64_t v0 = 0, v1 = 0, v2 = 0, v3 = 0, v4 = 0, v5 = 0;
    for (int k = s0; k < s0 + cnt; k++)
    {
        int j = ex_list[k];
        const snb_vector rj = crd[j];
        const float4 aj = make_float4(rj.x, rj.y, rj.z, charge[j]);
        float dx = aj.x - ai.x, dy = aj.y - ai.y, dz = aj.z - ai.z;
        pbc_disp(dx, dy, dz, box);
        // FP64 from the float displacement: the erf part of an excluded pair is
        // a small difference of large terms in float arithmetic
        double ddx = dx, ddy = dy, ddz = dz;
        double r2d = (ddx * ddx + ddy * ddy) + ddz * ddz;
        double rd = __builtin_sqrt(r2d);
        double rid = 1.0 / rd;
        double gd;
        const double erfd = 1.0 - snb_erfc_parts_d((double)beta * rd, &gd);
        const double hd = erfd - gd;
        double qq = (double)ai.w * (double)aj.w;
        double fa = (qq * hd) * ((rid * rid) * rid);
        double fxv = fa * ddx, fyv = fa * ddy, fzv = fa * ddz;
        int64_t ix = snb_to_fixed_d(fxv), iy = snb_to_fixed_d(fyv), iz = snb_to_fixed_d(fzv);
        sx += ix;
        sy += iy;
        sz += iz;
        atomic_add_i64(fx + j, -ix);
        atomic_add_i64(fx + n + j, -iy);
        atomic_add_i64(fx + 2 * n + j, -iz);
        if (NEED_ENERGY) e += snb_to_fixed_d(-((qq * erfd) * rid));
        if (NEED_VIRIAL)
        {
            v0 += snb_to_fixed_d(-(fxv * ddx));
            v1 += snb_to_fixed_d(-(fxv * ddy + fyv * ddx));
            v2 += snb_to_fixed_d(-(fyv * ddy));
            v3 += snb_to_fixed_d(-(fxv * ddz + fzv * ddx));
            v4 += snb_to_fixed_d(-(fyv * ddz + fzv * ddy));
            v5 += snb_to_fixed_d(-(fzv * ddz));
        }
    }
    atomic_add_i64(fx + i, sx);
    atomic_add_i64(fx + n + i, sy);
    atomic_add_i64(fx + 2 * n + i, sz);
    if (NEED_ENERGY) e_atom[(size_t)2 * n + i] += e;
    if (NEED_VIRIAL)
    {
        vir[i] += v0;
        vir[(size_t)n + i] += v1;
        vir[(size_t)2 * n + i] += v2;
        vir[(size_t)3 * n + i] += v3;
        vir[(size_t)4 * n + i] += v4;
        vir[(size_t)5 * n + i] += v5;
    }
}

template <bool NEED_ENERGY, bool NEED_VIRIAL>
__global__ void __launch_bounds__(256) k_excl(int n, const snb_vector* __restrict__ crd,
                                              const float* __restrict__ charge,
                                              SnbBox box, float beta,
                                              const int* __restrict__ ex_start,
                                              const int* __restrict__ ex_list,
                                              const int* __restrict__ ex_num,
                                              int64_t* __restrict__ fx,
                                              int64_t* __restrict__ e_atom,
                                              int64_t* __restrict__ vir)
{
    k_excl_body<NEED_ENERGY, NEED_VIRIAL>(
        n, crd, charge, box, beta, ex_start, ex_list, ex_num, fx, e_atom, vir, blockIdx.x,
        gridDim.x);
}

struct ExclArgs
{
    int n;
    const snb_vector* crd;
    const float* charge;
    SnbBox box;
    float beta;
    const int* ex_start;
    const int* ex_list;
    const int* ex_num;
    int64_t* fx;
    int64_t* e_atom;
    int64_t* vir;
};

template <bool NEED_ENERGY, bool NEED_VIRIAL>
__global__ void __launch_bounds__(256) k_excl_batch(SnbSegTab t)
{
    unsigned bid, nbid;
    const ExclArgs& a = snb_seg<ExclArgs>(t, bid, nbid);
    k_excl_body<NEED_ENERGY, NEED_VIRIAL>(
        a.n, a.crd, a.charge, a.box, a.beta, a.ex_start, a.ex_list, a.ex_num, a.fx, a.e_atom, a.vir,
        bid, nbid);
}

int snb_launch_excl_to(hipStream_t s, int n, const snb_vector* crd, const float* charge,
                       const SnbBox& box, float beta, const int* ex_start, const int* ex_list,
                       const int* ex_num, int need_energy, int need_virial, int64_t* fx,
                       int64_t* e_atom, int64_t* vir)
{
    if (n <= 0) return SNB_OK;
    dim3 g((n + 255) / 256);
#define SNB_EX_LAUNCH(E, V)                                                                  \
    SNB_LAUNCH_B((k_excl<E, V>), (k_excl_batch<E, V>), ExclArgs, g, dim3(256), 0, s, n, crd, \
                 charge, box, beta, ex_start, ex_list, ex_num, fx, e_atom, vir)
    if (need_energy && need_virial)
        SNB_EX_LAUNCH(true, true);
    else if (need_energy)
        SNB_EX_LAUNCH(true, false);
    else if (need_virial)
        SNB_EX_LAUNCH(false, true);
    else
        SNB_EX_LAUNCH(false, false);
#undef SNB_EX_LAUNCH
    SNB_CHECK(hipGetLastError());
    return SNB_OK;
}

int snb_launch_excl(snb_context* ctx, const snb_vector* crd, const float* charge,
                    const SnbBox& box, float beta,
                    const int* ex_start, const int* ex_list, const int* ex_num,
                    int need_energy, int need_virial, int64_t* fx)
{
    if (ctx->n <= 0) return SNB_OK;
    snb_prof_begin(ctx, SNB_PROF_EXCL);
    int rc = snb_launch_excl_to(ctx->stream, ctx->n, crd, charge, box, beta, ex_start, ex_list,
                                ex_num, need_energy, need_virial, fx, ctx->e_atom,
                                ctx->vir_atom);
    snb_prof_end(ctx, SNB_PROF_EXCL);
    return rc;
}

// Scaled 1-4 LJ and plain Coulomb of the 1-4 pairs (stands in for
// Dihedral_14_LJ_CF_Force_With_Atom_Energy_And_Virial_Device, nb14.cpp:7-100).
// One thread per pair.  The force is float, in the reference's operation
// order: dr = pbc(r_b - r_a), 1/r^2 and 1/r as a double quotient rounded to
// float, charges below 1e-10 in magnitude raised to 1e-10, f = -A r^-14 +
// B r^-8 - cf qa qb r^-3; +f dr on a and -f dr on b, rounded to fixed point.
// Energies in FP64 from the float displacement with the reference's
// constants (cf qa qb / r and 0.08333333 A r^-12 - 0.1666666 B r^-6), kept
// per pair and added to atom a; the virial -f dr (x) dr goes to atom a.
// 1/r uses sqrtf (correctly rounded) where the reference takes norm3df.
template <bool NEED_ENERGY, bool NEED_VIRIAL>
__global__ void __launch_bounds__(256) k_nb14(int m, int n, const int* __restrict__ a14,
                                              const int* __restrict__ b14,
                                              const float* __restrict__ A14,
                                              const float* __restrict__ B14,
                                              const float* __restrict__ cf14,
                                              const snb_vector* __restrict__ crd,
                                              const float* __restrict__ charge, SnbBox box,
                                              int64_t* __restrict__ fx,
                                              int64_t* __restrict__ e_nb14,
                                              int64_t* __restrict__ pair_e,
                                              int64_t* __restrict__ vir)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    const int i = a14[k], j = b14[k];
    const snb_vector ri = crd[i], rj = crd[j];
    float dx = rj.x - ri.x, dy = rj.y - ri.y, dz = rj.z - ri.z;
    pbc_disp(dx, dy, dz, box);
    const float dr2 = (dx * dx + dy * dy) + dz * dz;
    const float dr_2 = (float)(1.0 / (double)dr2);
    const float dr_4 = dr_2 * dr_2;
    const float dr_8 = dr_4 * dr_4;
    const float dr_14 = (dr_8 * dr_4) * dr_2;
    const float dr_1 = (float)(1.0 / (double)sqrtf(dr2));
    float qi = charge[i], qj = charge[j];
    if ((double)fabsf(qi) < 1e-10) qi = (float)1e-10;
    if ((double)fabsf(qj) < 1e-10) qj = (float)1e-10;
    const float A = A14[k], B = B14[k], cf = cf14[k];
    const float fcf = ((-((cf * dr_2) * dr_1)) * qi) * qj;
    const float flj = (-A) * dr_14 + B * dr_8;
    const float f = flj + fcf;
    const float fxv = f * dx, fyv = f * dy, fzv = f * dz;
    const int64_t ix = snb_to_fixed(fxv), iy = snb_to_fixed(fyv), iz = snb_to_fixed(fzv);
    atomic_add_i64(fx + i, ix);
    atomic_add_i64(fx + n + i, iy);
    atomic_add_i64(fx + 2 * n + i, iz);
    atomic_add_i64(fx + j, -ix);
    atomic_add_i64(fx + n + j, -iy);
    atomic_add_i64(fx + 2 * n + j, -iz);
    if (NEED_ENERGY)
    {
        const double ddx = dx, ddy = dy, ddz = dz;
        const double r2d = (ddx * ddx + ddy * ddy) + ddz * ddz;
        const double d1 = 1.0 / __builtin_sqrt(r2d);
        const double d2 = 1.0 / r2d;
        const double d4 = d2 * d2, d8 = d4 * d4;
        const int64_t pc = snb_to_fixed_d((((double)cf * d1) * (double)qi) * (double)qj);
        const int64_t pl = snb_to_fixed_d(((0.08333333 * (double)A) * d4) * d8 -
                                          ((0.1666666 * (double)B) * d4) * d2);
        pair_e[k] = pl;
        pair_e[m + k] = pc;
        atomic_add_i64(e_nb14 + i, pl + pc);
    }
    if (NEED_VIRIAL)
    {
        atomic_add_i64(vir + i, snb_to_fixed(-(fxv * dx)));
        atomic_add_i64(vir + (size_t)n + i, snb_to_fixed(-(fxv * dy + fyv * dx)));
        atomic_add_i64(vir + (size_t)2 * n + i, snb_to_fixed(-(fyv * dy)));
        atomic_add_i64(vir + (size_t)3 * n + i, snb_to_fixed(-(fxv * dz + fzv * dx)));
        atomic_add_i64(vir + (size_t)4 * n + i, snb_to_fixed(-(fyv * dz + fzv * dy)));
        atomic_add_i64(vir + (size_t)5 * n + i, snb_to_fixed(-(fzv * dz)));
    }
}

int snb_launch_nb14(snb_context* ctx, const snb_vector* crd, const float* charge,
                    const SnbBox& box, int need_energy, int need_virial, int64_t* fx)
{
    const SnbNb14& M = ctx->nb14;
    const int m = M.count, n = ctx->n;
    if (m <= 0) return SNB_OK;
    dim3 g((m + 255) / 256);
#define SNB_NB14_LAUNCH(E, V)                                                                \
    SNB_LAUNCH((k_nb14<E, V>), g, dim3(256), 0, ctx->stream, m, n, M.a, M.b, M.A, M.B, \
                       M.cf, crd, charge, box, fx, ctx->e_atom + 3 * (size_t)n, M.pair_e,      \
                       ctx->vir_atom)
    if (need_energy && need_virial)
        SNB_NB14_LAUNCH(true, true);
    else if (need_energy)
        SNB_NB14_LAUNCH(true, false);
    else if (need_virial)
        SNB_NB14_LAUNCH(false, true);
    else
        SNB_NB14_LAUNCH(false, false);
#undef SNB_NB14_LAUNCH
    SNB_CHECK(hipGetLastError());
    return SNB_OK;
}

// Harmonic bonds (stands in for Bond_Force_With_Atom_Energy_And_Virial_Device,
// bond/bond.cpp:15-64): dr = pbc(r_a - r_b), r = |dr| (sqrtf, where the
// reference calls norm3df), f = -(2 k (r - r0)) / r * dr on a and -f on b in
// the reference's float order, fixed point; the energy k (r - r0)^2 in FP64
// from the float displacement, kept per bond and added to atom a; the virial
// f (x) dr to atom a.
template <bool NEED_ENERGY, bool NEED_VIRIAL>
__global__ void __launch_bounds__(256) k_bond(int m, int n, const int* __restrict__ ba,
                                              const int* __restrict__ bb,
                                              const float* __restrict__ bk,
                                              const float* __restrict__ br0,
                                              const snb_vector* __restrict__ crd, SnbBox box,
                                              int64_t* __restrict__ fx,
                                              int64_t* __restrict__ e_bond,
                                              int64_t* __restrict__ pair_e,
                                              int64_t* __restrict__ vir)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    const int i = ba[k], j = bb[k];
    const snb_vector ri = crd[i], rj = crd[j];
    float dx = ri.x - rj.x, dy = ri.y - rj.y, dz = ri.z - rj.z;
    pbc_disp(dx, dy, dz, box);
    const float r = sqrtf((dx * dx + dy * dy) + dz * dz);
    const float kk = bk[k], r0 = br0[k];
    const float d = r - r0;
    const float g = (2.0f * d) * kk;
    const float s = (-g) / r;
    const float fxv = s * dx, fyv = s * dy, fzv = s * dz;
    const int64_t ix = snb_to_fixed(fxv), iy = snb_to_fixed(fyv), iz = snb_to_fixed(fzv);
    atomic_add_i64(fx + i, ix);
    atomic_add_i64(fx + n + i, iy);
    atomic_add_i64(fx + 2 * n + i, iz);
    atomic_add_i64(fx + j, -ix);
    atomic_add_i64(fx + n + j, -iy);
    atomic_add_i64(fx + 2 * n + j, -iz);
    if (NEED_ENERGY)
    {
        const double ddx = dx, ddy = dy, ddz = dz;
        const double dd = __builtin_sqrt((ddx * ddx + ddy * ddy) + ddz * ddz) - (double)r0;
        const int64_t pe = snb_to_fixed_d(((double)kk * dd) * dd);
        pair_e[k] = pe;
        atomic_add_i64(e_bond + i, pe);
    }
    if (NEED_VIRIAL)
    {
        atomic_add_i64(vir + i, snb_to_fixed(fxv * dx));
        atomic_add_i64(vir + (size_t)n + i, snb_to_fixed(fxv * dy + fyv * dx));
        atomic_add_i64(vir + (size_t)2 * n + i, snb_to_fixed(fyv * dy));
        atomic_add_i64(vir + (size_t)3 * n + i, snb_to_fixed(fxv * dz + fzv * dx));
        atomic_add_i64(vir + (size_t)4 * n + i, snb_to_fixed(fyv * dz + fzv * dy));
        atomic_add_i64(vir + (size_t)5 * n + i, snb_to_fixed(fzv * dz));
    }
}

int snb_launch_bond(snb_context* ctx, const snb_vector* crd, const SnbBox& box, int need_energy,
                    int need_virial, int64_t* fx)
{
    const SnbBond& M = ctx->bond;
    const int m = M.count, n = ctx->n;
    if (m <= 0) return SNB_OK;
    dim3 g((m + 255) / 256);
#define SNB_BOND_LAUNCH(E, V)                                                               \
    SNB_LAUNCH((k_bond<E, V>), g, dim3(256), 0, ctx->stream, m, n, M.a, M.b, M.k,    \
                       M.r0, crd, box, fx, ctx->e_atom + 4 * (size_t)n, M.pair_e,            \
                       ctx->vir_atom)
    if (need_energy && need_virial)
        SNB_BOND_LAUNCH(true, true);
    else if (need_energy)
        SNB_BOND_LAUNCH(true, false);
    else if (need_virial)
        SNB_BOND_LAUNCH(false, true);
    else
        SNB_BOND_LAUNCH(false, false);
#undef SNB_BOND_LAUNCH
    SNB_CHECK(hipGetLastError());
    return SNB_OK;
}

// ---------------------------------------------------------------------------
// windowed path
// ---------------------------------------------------------------------------
// Packs coordinates and charges in sorted (cell) order: satom[s] = (crd, q) of
// atom order[s]; raises the non-finite flag.
__device__ __forceinline__ void k_pack_sorted_body(int n, const int* __restrict__ order,
                              const snb_vector* __restrict__ crd, const float* __restrict__ charge,
                              float4* __restrict__ satom, int* __restrict__ status,
    unsigned bid, unsigned nbid)
{
    int s = bid * blockDim.x + threadIdx.x;
    if (s >= n) return;
    int id = order[s];
    snb_vector r = crd[id];
    satom[s] = make_float4(r.x, r.y, r.z, charge[id]);
    if (!(__builtin_isfinite(r.x) && __builtin_isfinite(r.y) && __builtin_isfinite(r.z)))
        atomicOr(status, SNB_FLAG_NONFINITE_CRD);
}

__global__ void k_pack_sorted(int n, const int* __restrict__ order,
                              const snb_vector* __restrict__ crd, const float* __restrict__ charge,
                              float4* __restrict__ satom, int* __restrict__ status)
{
    k_pack_sorted_body(n, order, crd, charge, satom, status, blockIdx.x, gridDim.x);
}

struct PackSortedArgs
{
    int n;
    const int* order;
    const snb_vector* crd;
    const float* charge;
    float4* satom;
    int* status;
};

__global__ void k_pack_sorted_batch(SnbSegTab t)
{
    unsigned bid, nbid;
    const PackSortedArgs& a = snb_seg<PackSortedArgs>(t, bid, nbid);
    k_pack_sorted_body(a.n, a.order, a.crd, a.charge, a.satom, a.status, bid, nbid);
}

// One workgroup per block of a x b columns x S z-cells (snb_window.h).  The
// workgroup keeps int64 force accumulators for every atom of its window in
// LDS; a wavefront evaluates one row at a time (64 lanes over its entries,
// two entries per lane in flight so the gathers of one overlap the other),
// adds -f to the j slot (ds_add_u64) and, after a wave reduction, +sum f to
// the row's own slot.  At the end the accumulators go to the stage buffer;
// k_stage_gather sums each atom's stage entries.  Windows larger than the LDS
// capacity fall back to integer atomics on the global accumulators.
// Per-atom energies and virials go, as in the reference, to the lower atom id
// of the pair: row-owner sums when the owner is the lower id, otherwise
// integer atomics on the j atom's entry of the sorted-order arrays (the j
// atoms of one row lie in a few neighbouring cells, so a wavefront's adds
// fall on few cache lines); k_stage_gather moves the sorted-order sums to the
// id-ordered arrays.  With ref_attr == 0 every pair's terms stay with the row
// owner.

// Periodic displacement.  ORTHO (diagonal cell) drops the off-diagonal terms
// and takes k = rint(s) for floor(s + 0.5).  For every pair that can reach
// the cutoff this is the reference's arithmetic bit for bit: s lies within
// rc / L < 1/2 of an integer there, so both give the same k.  The two differ
// only for s within an ulp of a half-integer, i.e. for pairs half a box
// length apart, which are beyond the cutoff either way.
template <bool ORTHO>
__device__ __forceinline__ void win_disp(float& dx, float& dy, float& dz, const SnbBox& box)
{
    if (ORTHO)
    {
        const float kx = rintf(dx * box.rc[0]);
        const float ky = rintf(dy * box.rc[2]);
        const float kz = rintf(dz * box.rc[5]);
        dx = dx - kx * box.c[0];
        dy = dy - ky * box.c[2];
        dz = dz - kz * box.c[5];
    }
    else
        pbc_disp(dx, dy, dz, box);
}

struct WinRowSums
{
    int64_t sx, sy, sz, elj, ec, v0, v1, v2, v3, v4, v5;
};

// Everything one in-cutoff pair contributes: f (force on the row atom i) to
// the register sums, the j share to the j accumulator, the row owner's share
// of energy/virial to the register sums.  A2, B2 and qi2 carry the 2^32
// fixed-point scale (exact power-of-two scaling), so f comes out scaled and
// its conversion is one double add.  The LDS accumulators hold the negated
// force (j receives +f, the row atom -sum f at the row's end), which saves
// the 64-bit negation per pair; the global arrays of the overflow mode hold
// the force itself.
template <bool NEED_ENERGY, bool NEED_VIRIAL>
__device__ __forceinline__ void win_pair(WinRowSums& S, float dx, float dy, float dz, float r2,
                                         float A2, float B2, float qi2, float qj, float beta,
                                         float cscale, const float4* ctab,
                                         int slot, int sj, bool ov, int wcap, long long* acc,
                                         int n, const int* __restrict__ order,
                                         int64_t* __restrict__ fx, int ref_attr, int id_i,
                                         int64_t* __restrict__ e_sorted,
                                         int64_t* __restrict__ v_sorted)
{
    const float fa2 = pair_force_factor(r2, A2, B2, qi2 * qj, beta, cscale, ctab);
    const float fx2 = fa2 * dx, fy2 = fa2 * dy, fz2 = fa2 * dz;
    int64_t ix = snb_fixed_from_scaled(fx2), iy = snb_fixed_from_scaled(fy2),
            iz = snb_fixed_from_scaled(fz2);
    S.sx += ix;
    S.sy += iy;
    S.sz += iz;
    if (!ov)
    {
        atomicAdd((unsigned long long*)&acc[slot], (unsigned long long)ix);
        atomicAdd((unsigned long long*)&acc[wcap + slot], (unsigned long long)iy);
        atomicAdd((unsigned long long*)&acc[2 * wcap + slot], (unsigned long long)iz);
    }
    else
    {
        int idj = order[sj];
        atomic_add_i64(fx + idj, -ix);
        atomic_add_i64(fx + n + idj, -iy);
        atomic_add_i64(fx + 2 * n + idj, -iz);
    }
    if (NEED_ENERGY || NEED_VIRIAL)
    {
        // the lower id takes the pair's terms
        const bool to_j = ref_attr && order[sj] < id_i;
        if (NEED_ENERGY)
        {
            // unscaled parameters (2^-32 scaling is exact)
            const float A = A2 * 2.3283064365386963e-10f, B = B2 * 2.3283064365386963e-10f;
            const float qi = qi2 * 2.3283064365386963e-10f;
            double ddx = dx, ddy = dy, ddz = dz;
            double r2d = (ddx * ddx + ddy * ddy) + ddz * ddz;
            double rd = __builtin_sqrt(r2d);
            double rid = 1.0 / rd;
            double ri2 = rid * rid;
            double ri6 = (ri2 * ri2) * ri2;
            int64_t pl = snb_to_fixed_d((((double)A * (1.0 / 12.0)) * ri6 -
                                         (double)B * (1.0 / 6.0)) * ri6);
            double gd;
            const double erfcd = snb_erfc_parts_d((double)beta * rd, &gd);
            int64_t pc = snb_to_fixed_d((((double)qi * (double)qj) * erfcd) * rid);
            if (to_j)
            {
                atomic_add_i64(e_sorted + sj, pl);
                atomic_add_i64(e_sorted + n + sj, pc);
            }
            else
            {
                S.elj += pl;
                S.ec += pc;
            }
        }
        if (NEED_VIRIAL)
        {
            const float fxv = fx2 * 2.3283064365386963e-10f, fyv = fy2 * 2.3283064365386963e-10f,
                        fzv = fz2 * 2.3283064365386963e-10f;
            int64_t p0 = snb_to_fixed(-(fxv * dx));
            int64_t p1 = snb_to_fixed(-(fxv * dy + fyv * dx));
            int64_t p2 = snb_to_fixed(-(fyv * dy));
            int64_t p3 = snb_to_fixed(-(fxv * dz + fzv * dx));
            int64_t p4 = snb_to_fixed(-(fyv * dz + fzv * dy));
            int64_t p5 = snb_to_fixed(-(fzv * dz));
            if (to_j)
            {
                atomic_add_i64(v_sorted + sj, p0);
                atomic_add_i64(v_sorted + (size_t)n + sj, p1);
                atomic_add_i64(v_sorted + (size_t)2 * n + sj, p2);
                atomic_add_i64(v_sorted + (size_t)3 * n + sj, p3);
                atomic_add_i64(v_sorted + (size_t)4 * n + sj, p4);
                atomic_add_i64(v_sorted + (size_t)5 * n + sj, p5);
            }
            else
            {
                S.v0 += p0; S.v1 += p1; S.v2 += p2; S.v3 += p3; S.v4 += p4; S.v5 += p5;
            }
        }
    }
}

// window slot -> sorted index: the LDS map, or a search of the range table
// when the window did not fit
__device__ __forceinline__ int win_slot_sorted(int slot, bool ov, const int* slot_s,
                                               const int* rfirst, const int* rpref, int nrange)
{
    if (!ov) return slot_s[slot];
    int lo = 0, hi = nrange - 1;
    while (lo < hi)
    {
        int mid = (lo + hi + 1) >> 1;
        if (rpref[mid] <= slot)
            lo = mid;
        else
            hi = mid - 1;
    }
    return rfirst[lo] + (slot - rpref[lo]);
}

template <bool NEED_ENERGY, bool NEED_VIRIAL, bool ORTHO, bool STREAM>
__device__ __forceinline__ void k_pair_win_body(
    int n, SnbWin W, const int* __restrict__ start, const int* __restrict__ sorted_cell,
    const int* __restrict__ row_count, const int* __restrict__ row_split,
    const uint16_t* __restrict__ nl16, int max_nb,
    const float4* __restrict__ satom, const int* __restrict__ stype,
    const int* __restrict__ order, const float2* __restrict__ ab2, int ntypes,
    const float4* __restrict__ coul_tab, float cscale, SnbBox box, float rc2, float beta,
    const int* __restrict__ stage_base, long long* __restrict__ stage, size_t stage_cap,
    int64_t* __restrict__ fx, int64_t* __restrict__ e_sorted, int64_t* __restrict__ v_sorted,
    int ref_attr, const int* __restrict__ perm, int jdirect,
    unsigned bid, unsigned nbid)
{
    extern __shared__ long long smem[];
    float4* ctab = (float4*)smem;                   // [SNB_COUL_N]
    long long* acc = smem + 2 * SNB_COUL_N;         // [3][wcap]
    int* slot_s = (int*)(acc + 3 * (size_t)W.wcap);  // [wcap]
    int* rfirst = slot_s + W.wcap;                  // [128]
    int* rpref = rfirst + SNB_WIN_MAX_RANGES;       // [129]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nwave = blockDim.x >> 6;
    const int w = perm[bid];  // heaviest blocks first
    if (tid < SNB_COUL_N) ctab[tid] = coul_tab[tid];
    int x0, y0, z0;
    snb_win_block(W, w, x0, y0, z0);
    // range table: sizes, then a wave-0 scan
    int rsz = 0;
    if (tid < SNB_WIN_MAX_RANGES)
    {
        int f = 0;
        if (tid < W.nrange)
        {
            int fc, nc;
            snb_win_range_cells(W, x0, y0, z0, tid, fc, nc);
            f = start[fc];
            rsz = nc > 0 ? start[fc + nc] - f : 0;
        }
        rfirst[tid] = f;
        rpref[tid + 1] = rsz;
    }
    __syncthreads();
    if (wave == 0)
    {
        int a0 = rpref[lane + 1], a1 = rpref[lane + 65];
        int i0 = a0, i1 = a1;
        for (int off = 1; off < 64; off <<= 1)
        {
            int u0 = __shfl_up(i0, off), u1 = __shfl_up(i1, off);
            if (lane >= off)
            {
                i0 += u0;
                i1 += u1;
            }
        }
        int t0 = __shfl(i0, 63);
        rpref[lane + 1] = i0;
        rpref[lane + 65] = t0 + i1;
        if (lane == 0) rpref[0] = 0;
    }
    __syncthreads();
    const int total = rpref[SNB_WIN_MAX_RANGES];
    const bool ov = total > W.wcap;
    if (!ov)
    {
        for (int k = 0; k < W.nrange; k++)
        {
            int b0 = rpref[k], cnt = rpref[k + 1] - b0, f = rfirst[k];
            for (int t = tid; t < cnt; t += blockDim.x) slot_s[b0 + t] = f + t;
        }
        for (int t = tid; t < 3 * W.wcap; t += blockDim.x) acc[t] = 0;
    }
    __syncthreads();
    if constexpr (STREAM)
    {
        // STREAM (force-only steps of large systems): one software pipeline
        // per wavefront over the 128-entry steps (two entries per lane) of
        // all its rows, across row ends, so that a row starts with its
        // records already fetched.  Per
        // iteration: the LJ parameters of step t (its types arrived during
        // step t - 1), the records of step t + 1 (slot map, then position and
        // type; its list entries arrived during step t - 1), the list entries
        // of step t + 2, the metadata of the wavefront's next row, then the
        // pairs of step t.  Records and entries alternate between two
        // register sets (the loop body is written out twice), so no value
        // is copied while its load is in flight; every load is issued on
        // every iteration (clamped addresses) so that their count is fixed.
        const int ncol = W.a * W.b;
        struct RowM
        {
            int si, col, cnt, split, type, cell;
            float4 ai;
        };
        struct Rec
        {
            int sla, slb, tja, tjb;
            unsigned sja, sjb;
            float4 aja, ajb;
        };
        int cur_col = 0, cr0 = 0, cr1 = 0, cur_si = 0;
        auto col_range = [&](int col) {
            const int ua = col / W.b, vb = col - ua * W.b;
            cr0 = cr1 = 0;
            if (x0 + ua >= W.nx || y0 + vb >= W.ny) return;
            const int colbase = ((x0 + ua) * W.ny + (y0 + vb)) * W.nz;
            const int z1 = z0 + W.s < W.nz ? z0 + W.s : W.nz;
            cr0 = start[colbase + z0];
            cr1 = start[colbase + z1];
        };
        // cur_si: this wavefront's next row (-1 after its last)
        auto seek = [&]() {
            while (cur_si >= cr1)
            {
                if (++cur_col >= ncol)
                {
                    cur_si = -1;
                    return;
                }
                col_range(cur_col);
                cur_si = cr0 + wave;
            }
        };
        col_range(0);
        cur_si = cr0 + wave;
        seek();
        // metadata of row cur_si (of row 0 when there is none: discarded)
        auto load_row = [&]() -> RowM {
            const int si = cur_si >= 0 ? cur_si : 0;
            RowM r;
            r.si = cur_si;
            r.col = cur_col;
            r.cnt = row_count[si];
            r.split = row_split[si];
            r.type = stype[si];
            r.cell = sorted_cell[si];
            r.ai = satom[si];
            return r;
        };
        // the loaded metadata, in scalar registers (one value per wavefront)
        auto uniform = [](const RowM& r) -> RowM {
            RowM u;
            u.si = __builtin_amdgcn_readfirstlane(r.si);
            u.col = __builtin_amdgcn_readfirstlane(r.col);
            u.cnt = __builtin_amdgcn_readfirstlane(r.cnt);
            u.split = __builtin_amdgcn_readfirstlane(r.split);
            u.type = __builtin_amdgcn_readfirstlane(r.type);
            u.cell = __builtin_amdgcn_readfirstlane(r.cell);
            u.ai.x = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(r.ai.x)));
            u.ai.y = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(r.ai.y)));
            u.ai.z = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(r.ai.z)));
            u.ai.w = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(r.ai.w)));
            return u;
        };
        auto entry_of = [&](const RowM& r, int k) -> int {
            const int si = r.si >= 0 ? r.si : 0;
            const int kk = k < r.cnt ? k : 0;
            const int e = (int)ld32(nl16, (unsigned)si * (unsigned)max_nb +
                                              (unsigned)snb_row_entry(kk, r.split, max_nb));
            return r.si >= 0 && k < r.cnt ? e : -1;
        };
        auto records = [&](Rec& q, int sla, int slb, int si) {
            si = si >= 0 ? si : 0;
            q.sla = sla;
            q.slb = slb;
            q.sja = sla >= 0 ? (unsigned)win_slot_sorted(sla, ov, slot_s, rfirst, rpref, W.nrange)
                             : (unsigned)si;
            q.sjb = slb >= 0 ? (unsigned)win_slot_sorted(slb, ov, slot_s, rfirst, rpref, W.nrange)
                             : (unsigned)si;
            q.aja = ld32(satom, q.sja);
            q.ajb = ld32(satom, q.sjb);
            q.tja = ld32(stype, q.sja);
            q.tjb = ld32(stype, q.sjb);
        };
        WinRowSums S = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        // the units of the compute, record and entry stages, the next row
        RowM cRow = uniform(load_row());
        if (cur_si >= 0)
        {
            cur_si += nwave;
            seek();
        }
        RowM R = load_row();
        int cK = 0;
        Rec r0, r1;
        int e0a = entry_of(cRow, lane), e0b = entry_of(cRow, lane + 64);
        records(r0, e0a, e0b, cRow.si);
        RowM nRow = cRow, eRow = cRow;
        int nK = 0, eK = 0;
        // the unit after E (R, the metadata of row cur_si, is loaded again
        // after each call)
        auto advance = [&]() {
            if (eRow.si < 0) return;
            if (eK + 128 < eRow.cnt)
                eK += 128;
            else
            {
                eRow = uniform(R);
                eK = 0;
                if (cur_si >= 0)
                {
                    cur_si += nwave;
                    seek();
                }
            }
        };
        advance();
        R = load_row();
        nRow = eRow;
        nK = eK;
        int e1a = entry_of(eRow, eK + lane), e1b = entry_of(eRow, eK + 64 + lane);
        int e2a = -1, e2b = -1;
        // one iteration: compute with cur, fill nxt from (ea, eb), load the
        // entries of the unit after E into (fa, fb)
        auto step = [&](Rec& cur, Rec& nxt, int ea, int eb, int& fa, int& fb) {
            const unsigned trow = (unsigned)(cRow.type * ntypes);
            const float2 ABa = ld32(ab2, trow + (unsigned)cur.tja);
            const float2 ABb = ld32(ab2, trow + (unsigned)cur.tjb);
            records(nxt, ea, eb, nRow.si);
            advance();
            fa = entry_of(eRow, eK + lane);
            fb = entry_of(eRow, eK + 64 + lane);
            R = load_row();
            if (cRow.si >= 0)
            {
                const float4 ai = cRow.ai;
                const float qi2 = ai.w * 4294967296.0f;  // 2^32 fixed-point scale
                float dxa = cur.aja.x - ai.x, dya = cur.aja.y - ai.y, dza = cur.aja.z - ai.z;
                float dxb = cur.ajb.x - ai.x, dyb = cur.ajb.y - ai.y, dzb = cur.ajb.z - ai.z;
                win_disp<ORTHO>(dxa, dya, dza, box);
                win_disp<ORTHO>(dxb, dyb, dzb, box);
                const float r2a = (dxa * dxa + dya * dya) + dza * dza;
                const float r2b = (dxb * dxb + dyb * dyb) + dzb * dzb;
                // keep the parameter gathers ahead of the branches (an empty
                // statement that reads the registers; nothing is emitted)
                asm volatile("" ::"v"(ABa.x), "v"(ABa.y), "v"(ABb.x), "v"(ABb.y));
                if (cur.sla >= 0 && r2a < rc2)
                    win_pair<false, false>(S, dxa, dya, dza, r2a, ABa.x, ABa.y, qi2, cur.aja.w,
                                           beta, cscale, ctab, cur.sla, cur.sja, ov, W.wcap, acc,
                                           n, order, fx, 0, 0, e_sorted, v_sorted);
                if (cur.slb >= 0 && r2b < rc2)
                    win_pair<false, false>(S, dxb, dyb, dzb, r2b, ABb.x, ABb.y, qi2, cur.ajb.w,
                                           beta, cscale, ctab, cur.slb, cur.sjb, ov, W.wcap, acc,
                                           n, order, fx, 0, 0, e_sorted, v_sorted);
                if (cK + 128 >= cRow.cnt)
                {
                    // the end of the row: its own force
                    const int64_t sx = wave_sum(S.sx), sy = wave_sum(S.sy), sz = wave_sum(S.sz);
                    if (lane == 0)
                    {
                        if (!ov)
                        {
                            const int ua = cRow.col / W.b, vb = cRow.col - ua * W.b;
                            const int cz = cRow.cell % W.nz;
                            const int k = snb_win_range_of(W, z0, ua, vb + 2, cz - z0 + 2);
                            const int slot_i = rpref[k] + (cRow.si - rfirst[k]);
                            // negated-force convention of the LDS accumulators
                            atomicAdd((unsigned long long*)&acc[slot_i], (unsigned long long)(-sx));
                            atomicAdd((unsigned long long*)&acc[W.wcap + slot_i],
                                      (unsigned long long)(-sy));
                            atomicAdd((unsigned long long*)&acc[2 * W.wcap + slot_i],
                                      (unsigned long long)(-sz));
                        }
                        else
                        {
                            const int idi = order[cRow.si];
                            atomic_add_i64(fx + idi, sx);
                            atomic_add_i64(fx + n + idi, sy);
                            atomic_add_i64(fx + 2 * n + idi, sz);
                        }
                    }
                    S.sx = S.sy = S.sz = 0;
                }
            }
            cRow = nRow;
            cK = nK;
            nRow = eRow;
            nK = eK;
        };
        while (true)
        {
            step(r0, r1, e1a, e1b, e2a, e2b);
            if (cRow.si < 0) break;
            step(r1, r0, e2a, e2b, e1a, e1b);
            if (cRow.si < 0) break;
        }
    }
    else
    {
        for (int col = 0; col < W.a * W.b; col++)
        {
            int ua = col / W.b, vb = col - ua * W.b;
            if (x0 + ua >= W.nx || y0 + vb >= W.ny) continue;
            int colbase = ((x0 + ua) * W.ny + (y0 + vb)) * W.nz;
            int z1 = z0 + W.s < W.nz ? z0 + W.s : W.nz;
            int r0 = start[colbase + z0], r1 = start[colbase + z1];
            for (int si = r0 + wave; si < r1; si += nwave)
            {
                const int cnt = row_count[si], split = row_split[si];
                const unsigned row0 = (unsigned)si * (unsigned)max_nb;  // n * max_nb < 2^31
                const float4 ai = satom[si];
                const float qi2 = ai.w * 4294967296.0f;  // 2^32 fixed-point scale
                const unsigned trow = (unsigned)(stype[si] * ntypes);
                const int id_i = (NEED_ENERGY || NEED_VIRIAL) && ref_attr ? order[si] : 0;
                WinRowSums S = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
                // Software pipeline over 128-entry steps (two entries per lane):
                // step k uses atoms whose records were fetched during step k - 1,
                // fetches the records of step k + 1 (slot map, then position and
                // type) and the list entries of step k + 2.  32-bit byte offsets:
                // the loads take the scalar-base + vector-offset form.  Energy /
                // virial variants fetch each step's records when it starts: their
                // FP64 energies and row sums leave no registers for a second set
                // of records, and a pipelined version spills.
                constexpr bool PIPE = !(NEED_ENERGY || NEED_VIRIAL);
                auto entry = [&](int k) -> int {
                    return k < cnt ? (int)ld32(nl16, row0 + (unsigned)snb_row_entry(k, split, max_nb))
                                   : -1;
                };
                auto sorted_of = [&](int sl) -> unsigned {
                    return sl >= 0 ? (unsigned)win_slot_sorted(sl, ov, slot_s, rfirst, rpref, W.nrange)
                                   : (unsigned)si;
                };
                int sla = 0, slb = 0, tja = 0, tjb = 0, nxa = 0, nxb = 0;
                unsigned sja = 0, sjb = 0;
                float4 aja = {0, 0, 0, 0}, ajb = {0, 0, 0, 0};
                auto records = [&](int k) {
                    sla = entry(k + lane);
                    slb = entry(k + 64 + lane);
                    sja = sorted_of(sla);
                    sjb = sorted_of(slb);
                    aja = ld32(satom, sja);
                    ajb = ld32(satom, sjb);
                    tja = ld32(stype, sja);
                    tjb = ld32(stype, sjb);
                };
                if (PIPE)
                {
                    records(0);
                    nxa = entry(lane + 128);
                    nxb = entry(lane + 192);
                }
                for (int k0 = 0; k0 < cnt; k0 += 128)
                {
                    if (!PIPE) records(k0);
                    const float2 ABa = ld32(ab2, trow + (unsigned)tja);
                    const float2 ABb = ld32(ab2, trow + (unsigned)tjb);
                    // records of the next step
                    int sla_n = 0, slb_n = 0, tja_n = 0, tjb_n = 0;
                    unsigned sja_n = 0, sjb_n = 0;
                    float4 aja_n = {0, 0, 0, 0}, ajb_n = {0, 0, 0, 0};
                    if (PIPE)
                    {
                        sla_n = nxa;
                        slb_n = nxb;
                        sja_n = sorted_of(sla_n);
                        sjb_n = sorted_of(slb_n);
                        aja_n = ld32(satom, sja_n);
                        ajb_n = ld32(satom, sjb_n);
                        tja_n = ld32(stype, sja_n);
                        tjb_n = ld32(stype, sjb_n);
                        nxa = entry(k0 + 256 + lane);
                        nxb = entry(k0 + 320 + lane);
                    }
                    float dxa = aja.x - ai.x, dya = aja.y - ai.y, dza = aja.z - ai.z;
                    float dxb = ajb.x - ai.x, dyb = ajb.y - ai.y, dzb = ajb.z - ai.z;
                    win_disp<ORTHO>(dxa, dya, dza, box);
                    win_disp<ORTHO>(dxb, dyb, dzb, box);
                    const float r2a = (dxa * dxa + dya * dya) + dza * dza;
                    const float r2b = (dxb * dxb + dyb * dyb) + dzb * dzb;
                    // keep the parameter gathers ahead of the branches (the
                    // compiler would otherwise sink them into the branches, where
                    // their latency is not hidden behind the distance test); an
                    // empty statement that reads the registers, nothing is emitted
                    asm volatile("" ::"v"(ABa.x), "v"(ABa.y), "v"(ABb.x), "v"(ABb.y));
                    if (sla >= 0 && r2a < rc2)
                        win_pair<NEED_ENERGY, NEED_VIRIAL>(S, dxa, dya, dza, r2a, ABa.x, ABa.y, qi2,
                                                           aja.w, beta, cscale, ctab, sla, sja, ov,
                                                           W.wcap, acc, n, order, fx, ref_attr, id_i,
                                                           e_sorted, v_sorted);
                    if (slb >= 0 && r2b < rc2)
                        win_pair<NEED_ENERGY, NEED_VIRIAL>(S, dxb, dyb, dzb, r2b, ABb.x, ABb.y, qi2,
                                                           ajb.w, beta, cscale, ctab, slb, sjb, ov,
                                                           W.wcap, acc, n, order, fx, ref_attr, id_i,
                                                           e_sorted, v_sorted);
                    if (PIPE)
                    {
                        sla = sla_n;
                        slb = slb_n;
                        sja = sja_n;
                        sjb = sjb_n;
                        aja = aja_n;
                        ajb = ajb_n;
                        tja = tja_n;
                        tjb = tjb_n;
                    }
                }
                int64_t sx = wave_sum(S.sx), sy = wave_sum(S.sy), sz = wave_sum(S.sz);
                int64_t elj = 0, ec = 0, v0 = 0, v1 = 0, v2 = 0, v3 = 0, v4 = 0, v5 = 0;
                if (NEED_ENERGY)
                {
                    elj = wave_sum(S.elj);
                    ec = wave_sum(S.ec);
                }
                if (NEED_VIRIAL)
                {
                    v0 = wave_sum(S.v0); v1 = wave_sum(S.v1); v2 = wave_sum(S.v2);
                    v3 = wave_sum(S.v3); v4 = wave_sum(S.v4); v5 = wave_sum(S.v5);
                }
                if (lane == 0)
                {
                    if (!ov)
                    {
                        int cz = sorted_cell[si] % W.nz;
                        int k = snb_win_range_of(W, z0, ua, vb + 2, cz - z0 + 2);
                        int slot_i = rpref[k] + (si - rfirst[k]);
                        // negated-force convention of the LDS accumulators
                        atomicAdd((unsigned long long*)&acc[slot_i], (unsigned long long)(-sx));
                        atomicAdd((unsigned long long*)&acc[W.wcap + slot_i],
                                  (unsigned long long)(-sy));
                        atomicAdd((unsigned long long*)&acc[2 * W.wcap + slot_i],
                                  (unsigned long long)(-sz));
                    }
                    else
                    {
                        int idi = order[si];
                        atomic_add_i64(fx + idi, sx);
                        atomic_add_i64(fx + n + idi, sy);
                        atomic_add_i64(fx + 2 * n + idi, sz);
                    }
                    // other rows' lower-id terms may land on si as well
                    // (ref_attr: the arrays were cleared before the launch)
                    if (NEED_ENERGY)
                    {
                        if (ref_attr)
                        {
                            atomic_add_i64(e_sorted + si, elj);
                            atomic_add_i64(e_sorted + (size_t)n + si, ec);
                        }
                        else
                        {
                            e_sorted[si] = elj;
                            e_sorted[(size_t)n + si] = ec;
                        }
                    }
                    if (NEED_VIRIAL)
                    {
                        const int64_t vv[6] = {v0, v1, v2, v3, v4, v5};
                        for (int c = 0; c < 6; c++)
                        {
                            if (ref_attr)
                                atomic_add_i64(v_sorted + (size_t)c * n + si, vv[c]);
                            else
                                v_sorted[(size_t)c * n + si] = vv[c];
                        }
                    }
                }
            }
        }
    }
    __syncthreads();
    if (!NEED_ENERGY && !NEED_VIRIAL && jdirect)
    {
        // force-only steps: the window's sums go straight to the id-ordered
        // accumulators (integer sums, so the order of the adds does not
        // matter); slots that received nothing are skipped
        if (!ov)
            for (int t = tid; t < total; t += blockDim.x)
            {
                const long long vx = acc[t], vy = acc[W.wcap + t], vz = acc[2 * W.wcap + t];
                if ((vx | vy | vz) == 0) continue;
                const int id = order[slot_s[t]];
                if (vx) atomic_add_i64(fx + id, -vx);
                if (vy) atomic_add_i64(fx + n + id, -vy);
                if (vz) atomic_add_i64(fx + 2 * n + id, -vz);
            }
        return;
    }
    const int base = stage_base[w];
    for (int t = tid; t < total; t += blockDim.x)
    {
        if ((size_t)(base + t) >= stage_cap) break;  // cannot happen: stage holds n x invmax
        long long* dst = stage + 3 * (size_t)(base + t);
        dst[0] = ov ? 0 : -acc[t];
        dst[1] = ov ? 0 : -acc[W.wcap + t];
        dst[2] = ov ? 0 : -acc[2 * W.wcap + t];
    }
}

template <bool NEED_ENERGY, bool NEED_VIRIAL, bool ORTHO, bool STREAM>
__global__ void __launch_bounds__(1024) k_pair_win(
    int n, SnbWin W, const int* __restrict__ start, const int* __restrict__ sorted_cell,
    const int* __restrict__ row_count, const int* __restrict__ row_split,
    const uint16_t* __restrict__ nl16, int max_nb,
    const float4* __restrict__ satom, const int* __restrict__ stype,
    const int* __restrict__ order, const float2* __restrict__ ab2, int ntypes,
    const float4* __restrict__ coul_tab, float cscale, SnbBox box, float rc2, float beta,
    const int* __restrict__ stage_base, long long* __restrict__ stage, size_t stage_cap,
    int64_t* __restrict__ fx, int64_t* __restrict__ e_sorted, int64_t* __restrict__ v_sorted,
    int ref_attr, const int* __restrict__ perm, int jdirect)
{
    k_pair_win_body<NEED_ENERGY, NEED_VIRIAL, ORTHO, STREAM>(
        n, W, start, sorted_cell, row_count, row_split, nl16, max_nb, satom, stype, order, ab2,
        ntypes, coul_tab, cscale, box, rc2, beta, stage_base, stage, stage_cap, fx, e_sorted,
        v_sorted, ref_attr, perm, jdirect, blockIdx.x, gridDim.x);
}

struct PairWinArgs
{
    int n;
    SnbWin W;
    const int* start;
    const int* sorted_cell;
    const int* row_count;
    const int* row_split;
    const uint16_t* nl16;
    int max_nb;
    const float4* satom;
    const int* stype;
    const int* order;
    const float2* ab2;
    int ntypes;
    const float4* coul_tab;
    float cscale;
    SnbBox box;
    float rc2;
    float beta;
    const int* stage_base;
    long long* stage;
    size_t stage_cap;
    int64_t* fx;
    int64_t* e_sorted;
    int64_t* v_sorted;
    int ref_attr;
    const int* perm;
    int jdirect;
};

template <bool NEED_ENERGY, bool NEED_VIRIAL, bool ORTHO, bool STREAM>
__global__ void __launch_bounds__(1024) k_pair_win_batch(SnbSegTab t)
{
    unsigned bid, nbid;
    const PairWinArgs& a = snb_seg<PairWinArgs>(t, bid, nbid);
    k_pair_win_body<NEED_ENERGY, NEED_VIRIAL, ORTHO, STREAM>(
        a.n, a.W, a.start, a.sorted_cell, a.row_count, a.row_split, a.nl16, a.max_nb, a.satom,
        a.stype, a.order, a.ab2, a.ntypes, a.coul_tab, a.cscale, a.box, a.rc2, a.beta, a.stage_base,
        a.stage, a.stage_cap, a.fx, a.e_sorted, a.v_sorted, a.ref_attr, a.perm, a.jdirect, bid,
        nbid);
}

// Per sorted atom: sum of its stage entries into the id-ordered accumulators
// (plus the row-owner energies / virials on those steps).
__device__ __forceinline__ void k_stage_gather_body(int n, const int* __restrict__ inv_count,
                               const int* __restrict__ inv_list, int invmax,
                               const long long* __restrict__ stage, size_t stage_cap,
                               const int* __restrict__ order,
                               int64_t* __restrict__ fx, int need_energy, int need_virial,
                               const int64_t* __restrict__ e_sorted,
                               const int64_t* __restrict__ v_sorted, int64_t* __restrict__ e_atom,
                               int64_t* __restrict__ vir,
    unsigned bid, unsigned nbid)
{
    int s = bid * blockDim.x + threadIdx.x;
    if (s >= n) return;
    int id = order[s];
    int m = inv_count[s];
    if (m > invmax) m = invmax;
    int64_t sx = 0, sy = 0, sz = 0;
    const int* lst = inv_list + (size_t)s * invmax;
    for (int k = 0; k < m; k++)
    {
        if ((size_t)lst[k] >= stage_cap) continue;
        const long long* p = stage + 3 * (size_t)lst[k];
        sx += p[0];
        sy += p[1];
        sz += p[2];
    }
    fx[id] += sx;
    fx[n + id] += sy;
    fx[2 * n + id] += sz;
    if (need_energy)
    {
        e_atom[id] += e_sorted[s];
        e_atom[(size_t)n + id] += e_sorted[(size_t)n + s];
    }
    if (need_virial)
        for (int c = 0; c < 6; c++) vir[(size_t)c * n + id] += v_sorted[(size_t)c * n + s];
}

__global__ void k_stage_gather(int n, const int* __restrict__ inv_count,
                               const int* __restrict__ inv_list, int invmax,
                               const long long* __restrict__ stage, size_t stage_cap,
                               const int* __restrict__ order,
                               int64_t* __restrict__ fx, int need_energy, int need_virial,
                               const int64_t* __restrict__ e_sorted,
                               const int64_t* __restrict__ v_sorted, int64_t* __restrict__ e_atom,
                               int64_t* __restrict__ vir)
{
    k_stage_gather_body(
        n, inv_count, inv_list, invmax, stage, stage_cap, order, fx, need_energy, need_virial,
        e_sorted, v_sorted, e_atom, vir, blockIdx.x, gridDim.x);
}

struct StageGatherArgs
{
    int n;
    const int* inv_count;
    const int* inv_list;
    int invmax;
    const long long* stage;
    size_t stage_cap;
    const int* order;
    int64_t* fx;
    int need_energy;
    int need_virial;
    const int64_t* e_sorted;
    const int64_t* v_sorted;
    int64_t* e_atom;
    int64_t* vir;
};

__global__ void k_stage_gather_batch(SnbSegTab t)
{
    unsigned bid, nbid;
    const StageGatherArgs& a = snb_seg<StageGatherArgs>(t, bid, nbid);
    k_stage_gather_body(
        a.n, a.inv_count, a.inv_list, a.invmax, a.stage, a.stage_cap, a.order, a.fx, a.need_energy,
        a.need_virial, a.e_sorted, a.v_sorted, a.e_atom, a.vir, bid, nbid);
}

int snb_launch_lj_window(snb_context* ctx, const snb_vector* crd, const float* charge,
                         const SnbBox& box, float beta, int need_energy, int need_virial)
{
    int n = ctx->n;
    if (n <= 0) return SNB_OK;
    SnbNl& L = ctx->nl;
    const SnbWin& W = L.win;
    hipStream_t s = ctx->stream;
    SNB_LAUNCH_B(k_pack_sorted, k_pack_sorted_batch, PackSortedArgs, dim3((n + 255) / 256),
                 dim3(256), 0, s, n, L.order, crd, charge, ctx->satom, ctx->status);
    float rc2 = ctx->lj.cutoff * ctx->lj.cutoff;
    size_t lds = 28 * (size_t)W.wcap + SNB_WIN_FIXED_LDS;
    if (L.win_lds_set != lds)
    {
        // dynamic LDS above 64 KiB is granted per kernel
        const void* fns[10] = {(const void*)k_pair_win<true, true, true, false>,
                               (const void*)k_pair_win<true, false, true, false>,
                               (const void*)k_pair_win<false, true, true, false>,
                               (const void*)k_pair_win<false, false, true, false>,
                               (const void*)k_pair_win<false, false, true, true>,
                               (const void*)k_pair_win<true, true, false, false>,
                               (const void*)k_pair_win<true, false, false, false>,
                               (const void*)k_pair_win<false, true, false, false>,
                               (const void*)k_pair_win<false, false, false, false>,
                               (const void*)k_pair_win<false, false, false, true>};
        for (const void* f : fns)
            SNB_CHECK(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        L.win_lds_set = lds;
    }
    // diagonal cell and rcell: the displacement skips the zero terms
    const bool ortho = box.c[1] == 0.0f && box.c[3] == 0.0f && box.c[4] == 0.0f &&
                       box.rc[1] == 0.0f && box.rc[3] == 0.0f && box.rc[4] == 0.0f;
    // Force-only steps: the stage (about 430 bytes per atom at the preferred
    // shapes) and its gather stay in the 256 MB last-level cache for
    // mid-size systems, where they beat the blocks' atomic adds.  Past 128 MiB
    // of stage the atomic adds are cheaper than the stage's HBM round trip;
    // below 64k atoms the gather has too few workgroups to hide its chains
    // of dependent loads (48 us at 2,100 atoms, against 17 us for the pair
    // kernel), so the atomic adds win there too.
    int jdirect = ctx->lj.j_direct;
    if (jdirect < 0)
        jdirect = (n < 65536 || (size_t)n * 432 > ((size_t)128 << 20)) ? 1 : 0;
    if (need_energy || need_virial) jdirect = 0;
    const int ref_attr = ctx->lj.owner_attribution ? 0 : 1;
    int rc;
    if (ref_attr && need_energy && (rc = snb_launch_fill_i64(s, ctx->e_sorted, 2 * (size_t)n)))
        return rc;
    if (ref_attr && need_virial && (rc = snb_launch_fill_i64(s, ctx->v_sorted, 6 * (size_t)n)))
        return rc;
    // Force-only steps of large systems run the pipeline across row ends
    // (STREAM): at 1M atoms the pair kernel takes 2.15 instead of 2.34 ms.
    // At 93k atoms it is also faster alone (160 vs 170 us), but the step is
    // not: it holds more registers (87 vs 76 per lane), and the PME
    // workgroups that run beside it get fewer slots (0.2729 vs 0.2697 ms per
    // step).
    const int stream_mode = ctx->lj.row_stream;
    const bool stream = !need_energy && !need_virial &&
                        (stream_mode > 0 || (stream_mode < 0 && n >= (1 << 18)));
    snb_prof_begin(ctx, SNB_PROF_PAIR);
    // recorded step batches: the work of a workgroup is its rows
    if (snb_rec) snb_rec->weight_next = (double)n / W.nblock;
#define SNB_WIN_LAUNCH(E, V, O, ST)                                                             \
    SNB_LAUNCH_B((k_pair_win<E, V, O, ST>), (k_pair_win_batch<E, V, O, ST>), PairWinArgs,       \
                 dim3(W.nblock), dim3(1024), lds, s, n, W, L.cell_start, L.sorted_cell,        \
                 L.row_count, L.row_split, L.nl16, L.max_nb, ctx->satom, ctx->stype, L.order,  \
                 ctx->lj.ab2, ctx->lj.ntypes, ctx->lj.coul, ctx->lj.coul_scale, box, rc2, beta, \
                 L.stage_base, L.stage, (size_t)n * L.invmax, ctx->fx, ctx->e_sorted,           \
                 ctx->v_sorted, ref_attr, L.block_perm, jdirect)
#define SNB_WIN_LAUNCH_EV(O)                     \
    if (need_energy && need_virial)              \
        SNB_WIN_LAUNCH(true, true, O, false);    \
    else if (need_energy)                        \
        SNB_WIN_LAUNCH(true, false, O, false);   \
    else if (need_virial)                        \
        SNB_WIN_LAUNCH(false, true, O, false);   \
    else if (stream)                             \
        SNB_WIN_LAUNCH(false, false, O, true);   \
    else                                         \
        SNB_WIN_LAUNCH(false, false, O, false);
    if (ortho)
    {
        SNB_WIN_LAUNCH_EV(true)
    }
    else
    {
        SNB_WIN_LAUNCH_EV(false)
    }
#undef SNB_WIN_LAUNCH_EV
#undef SNB_WIN_LAUNCH
    snb_prof_end(ctx, SNB_PROF_PAIR);
    SNB_CHECK(hipGetLastError());
    if (jdirect) return SNB_OK;
    SNB_LAUNCH_B(k_stage_gather, k_stage_gather_batch, StageGatherArgs, dim3((n + 255) / 256),
                 dim3(256), 0, s, n, L.inv_count, L.inv_list, L.invmax, L.stage,
                 (size_t)n * L.invmax, L.order, ctx->fx, need_energy, need_virial, ctx->e_sorted,
                 ctx->v_sorted, ctx->e_atom, ctx->vir_atom);
    SNB_CHECK(hipGetLastError());
    return SNB_OK;
}
