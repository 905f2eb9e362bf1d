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

__device__ __forceinline__ int64_t wave_sum(int64_t v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

__device__ __forceinline__ void atomic_add_i64(int64_t* p, int64_t v)
{
    atomicAdd((unsigned long long*)p, (unsigned long long)v);
}

template <bool NEED_ENERGY, bool NEED_VIRIAL>
__global__ void __launch_bounds__(256) k_lj_direct(
    int n, const snb_atom_group* __restrict__ nl, const float4* __restrict__ atoms,
    const int* __restrict__ type, const float* __restrict__ lj_a,
    const float* __restrict__ lj_b, SnbBox box, float rc2, float beta,
    int64_t* __restrict__ fx, int64_t* __restrict__ e_atom, int64_t* __restrict__ vir)
{
    const int lane = threadIdx.x & 63;
    const int waves = gridDim.x * 4;
    for (int i = blockIdx.x * 4 + (threadIdx.x >> 6); i < n; i += waves)
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
            float rinv = snb_rsqrtf(r2);
            float r = r2 * rinv;
            float rinv2 = rinv * rinv;
            float rinv6 = (rinv2 * rinv2) * rinv2;
            float flj = ((B - A * rinv6) * rinv6) * rinv2;
            float x = beta * r;
            float ex = snb_expf(-(x * x));
            float erfcv = snb_erfc_core(x, ex);
            float qq = ai.w * aj.w;
            float fc = (qq * (rinv2 * rinv)) * ((x * SNB_TWO_OVER_SQRT_PI) * ex + erfcv);
            float fa = flj - fc;
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
                double erfd, hd;
                snb_erf_parts_d((double)beta * rd, &erfd, &hd);
                ec += snb_to_fixed_d((((double)ai.w * (double)aj.w) * (1.0 - erfd)) * rid);
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

int snb_launch_lj_direct(snb_context* ctx, const snb_atom_group* nl, const SnbBox& box,
                         float beta, int need_energy, int need_virial)
{
    int n = ctx->n;
    if (n <= 0) return SNB_OK;
    float rc2 = ctx->lj.cutoff * ctx->lj.cutoff;
    int blocks = (n + 3) / 4;
    if (blocks > 65535 * 4) blocks = 65535 * 4;
    snb_prof_begin(ctx, SNB_PROF_PAIR);
#define SNB_LJ_LAUNCH(E, V)                                                                   \
    hipLaunchKernelGGL((k_lj_direct<E, V>), dim3(blocks), dim3(256), 0, ctx->stream, n, nl,    \
                       ctx->atom4, ctx->type, ctx->lj.a, ctx->lj.b, box, rc2, beta, ctx->fx, \
                       ctx->e_atom, ctx->vir_atom)
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
__global__ void __launch_bounds__(256) k_excl(int n, const float4* __restrict__ atoms,
                                              SnbBox box, float beta,
                                              const int* __restrict__ ex_start,
                                              const int* __restrict__ ex_list,
                                              const int* __restrict__ ex_num,
                                              int64_t* __restrict__ fx,
                                              int64_t* __restrict__ e_atom,
                                              int64_t* __restrict__ vir)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int cnt = ex_num[i];
    if (cnt <= 0) return;
    int s0 = ex_start[i];
    float4 ai = atoms[i];
    int64_t sx = 0, sy = 0, sz = 0, e = 0;
    int64_t v0 = 0, v1 = 0, v2 = 0, v3 = 0, v4 = 0, v5 = 0;
    for (int k = s0; k < s0 + cnt; k++)
    {
        int j = ex_list[k];
        float4 aj = atoms[j];
        float dx = aj.x - ai.x, dy = aj.y - ai.y, dz = aj.z - ai.z;
        pbc_disp(dx, dy, dz, box);
        // FP64 from the float displacement: the erf part of an excluded pair is
        // a small difference of large terms in float arithmetic
        double ddx = dx, ddy = dy, ddz = dz;
        double r2d = (ddx * ddx + ddy * ddy) + ddz * ddz;
        double rd = __builtin_sqrt(r2d);
        double rid = 1.0 / rd;
        double erfd, hd;
        snb_erf_parts_d((double)beta * rd, &erfd, &hd);
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

int snb_launch_excl(snb_context* ctx, const float4* atoms, const SnbBox& box, float beta,
                    const int* ex_start, const int* ex_list, const int* ex_num,
                    int need_energy, int need_virial)
{
    int n = ctx->n;
    if (n <= 0) return SNB_OK;
    dim3 g((n + 255) / 256);
    snb_prof_begin(ctx, SNB_PROF_EXCL);
#define SNB_EX_LAUNCH(E, V)                                                                  \
    hipLaunchKernelGGL((k_excl<E, V>), g, dim3(256), 0, ctx->stream, n, atoms, box, beta,     \
                       ex_start, ex_list, ex_num, ctx->fx, ctx->e_atom, ctx->vir_atom)
    if (need_energy && need_virial)
        SNB_EX_LAUNCH(true, true);
    else if (need_energy)
        SNB_EX_LAUNCH(true, false);
    else if (need_virial)
        SNB_EX_LAUNCH(false, true);
    else
        SNB_EX_LAUNCH(false, false);
#undef SNB_EX_LAUNCH
    snb_prof_end(ctx, SNB_PROF_EXCL);
    SNB_CHECK(hipGetLastError());
    return SNB_OK;
}
