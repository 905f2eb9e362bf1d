// snb_md.hip -- integration on the device: leap-frog (stands in for
// MD_INFORMATION::NVE_iteration::Leap_Frog, MD_core/nve.hpp:3-48), the
// middle-scheme Langevin leap-frog (MIDDLE_Langevin_INFORMATION,
// thermostat/Middle_Langevin_MD.cpp:5-66, :305-331) and SETTLE for rigid
// triangles (constrain/settle.cpp:50-215, :573-697).
//
// The float updates keep the reference's operation order; there is no
// contraction (-ffp-contract=off), so the CPU twin reproduces every bit.  The
// Langevin noise comes from the library's own counter-based generator
// (Philox4x32-10 with Box-Muller, snb_math.h): atom i at its c-th step draws
// from counter {i, c, 0, 0}, with one counter per atom kept on the device, so
// a replayed graph still advances the stream.
#include "snb_internal.h"

// |v| capped at vmax (Make_Vector_Not_Exceed_Value, utils/vector.hpp:207-212),
// with 1/|v| from the library's rsqrt
__device__ __forceinline__ void cap_speed(float& vx, float& vy, float& vz, float vmax)
{
    const float v2 = (vx * vx + vy * vy) + vz * vz;
    const float f = v2 > 0.0f ? fminf(1.0f, vmax * snb_rsqrtf(v2)) : 1.0f;
    vx = f * vx;
    vy = f * vy;
    vz = f * vz;
}

// NVE leap-frog: v += dt inv_m f, x += dt v.
__global__ void __launch_bounds__(256) k_leap_frog(int n, const float* __restrict__ inv_mass,
                                                   const snb_vector* __restrict__ frc,
                                                   snb_vector* __restrict__ vel,
                                                   snb_vector* __restrict__ crd, float dt,
                                                   float vmax)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float im = inv_mass[i];
    const snb_vector f = frc[i];
    snb_vector v = vel[i], x = crd[i];
    v.x = v.x + dt * (im * f.x);
    v.y = v.y + dt * (im * f.y);
    v.z = v.z + dt * (im * f.z);
    if (vmax > 0.0f) cap_speed(v.x, v.y, v.z, vmax);
    x.x = x.x + dt * v.x;
    x.y = x.y + dt * v.y;
    x.z = x.z + dt * v.z;
    vel[i] = v;
    crd[i] = x;
}

// Middle scheme (Liu Jian): v1 = v + dt inv_m f; x1 = x + dt/2 v1;
// v2 = e^{-gamma dt} v1 + sqrt(kT (1 - e^{-2 gamma dt}) / m) R; x2 = x1 + dt/2 v2.
__device__ __forceinline__ void langevin_update(int i, const float* __restrict__ inv_mass,
                                                const float* __restrict__ sqrt_mass,
                                                uint32_t* __restrict__ counter, uint32_t seed,
                                                snb_vector f, snb_vector& v, snb_vector& x,
                                                float dt, float half_dt, float exp_gamma,
                                                float vmax)
{
    uint32_t c[4] = {(uint32_t)i, counter[i], 0u, 0u};
    counter[i] = c[1] + 1u;
    snb_philox4x32_10(c, seed, SNB_LANGEVIN_KEY1);
    float r0, r1, r2, r3;
    snb_box_muller(c[0], c[1], &r0, &r1);
    snb_box_muller(c[2], c[3], &r2, &r3);
    const float im = inv_mass[i], sm = sqrt_mass[i];
    float ax = im * f.x, ay = im * f.y, az = im * f.z;
    float vx = v.x + dt * ax, vy = v.y + dt * ay, vz = v.z + dt * az;
    const float xx = x.x + half_dt * vx, xy = x.y + half_dt * vy, xz = x.z + half_dt * vz;
    vx = exp_gamma * vx + sm * r0;
    vy = exp_gamma * vy + sm * r1;
    vz = exp_gamma * vz + sm * r2;
    if (vmax > 0.0f) cap_speed(vx, vy, vz, vmax);
    v.x = vx;
    v.y = vy;
    v.z = vz;
    x.x = xx + half_dt * vx;
    x.y = xy + half_dt * vy;
    x.z = xz + half_dt * vz;
}

__global__ void __launch_bounds__(256) k_langevin(int n, const float* __restrict__ inv_mass,
                                                  const float* __restrict__ sqrt_mass,
                                                  uint32_t* __restrict__ counter, uint32_t seed,
                                                  const snb_vector* __restrict__ frc,
                                                  snb_vector* __restrict__ vel,
                                                  snb_vector* __restrict__ crd, float dt,
                                                  float half_dt, float exp_gamma, float vmax)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    snb_vector v = vel[i], x = crd[i];
    langevin_update(i, inv_mass, sqrt_mass, counter, seed, frc[i], v, x, dt, half_dt, exp_gamma,
                    vmax);
    vel[i] = v;
    crd[i] = x;
}

// ---------------------------------------------------------------------------
// SETTLE
// ---------------------------------------------------------------------------
__device__ __forceinline__ snb_vector vsub(snb_vector a, snb_vector b)
{
    return {a.x - b.x, a.y - b.y, a.z - b.z};
}
__device__ __forceinline__ snb_vector vadd(snb_vector a, snb_vector b)
{
    return {a.x + b.x, a.y + b.y, a.z + b.z};
}
__device__ __forceinline__ snb_vector vscale(float s, snb_vector a)
{
    return {s * a.x, s * a.y, s * a.z};
}
__device__ __forceinline__ float vdot(snb_vector a, snb_vector b)
{
    return (a.x * b.x + a.y * b.y) + a.z * b.z;
}
__device__ __forceinline__ snb_vector vcross(snb_vector a, snb_vector b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ snb_vector vunit(snb_vector a)
{
    return vscale(snb_rsqrtf(vdot(a, a)), a);
}

__device__ __forceinline__ snb_vector settle_disp(snb_vector a, snb_vector b, const SnbBox& box)
{
    // Get_Periodic_Displacement(a, b): a - b, minimum image
    float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    float sx = (dx * box.rc[0] + dy * box.rc[1]) + dz * box.rc[3];
    float sy = dy * box.rc[2] + dz * box.rc[4];
    float sz = dz * box.rc[5];
    float kx = floorf(sx + 0.5f), ky = floorf(sy + 0.5f), kz = floorf(sz + 0.5f);
    dx = dx - ((kx * box.c[0] + ky * box.c[1]) + kz * box.c[3]);
    dy = dy - (ky * box.c[2] + kz * box.c[4]);
    dz = dz - kz * box.c[5];
    return {dx, dy, dz};
}

// remember_triangle_BA_CA (settle.cpp:3-23)
__global__ void __launch_bounds__(256) k_settle_remember(int m, const int4* __restrict__ tri,
                                                         const snb_vector* __restrict__ crd,
                                                         SnbBox box,
                                                         snb_vector* __restrict__ last_ba,
                                                         snb_vector* __restrict__ last_ca)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m) return;
    const int4 a = tri[t];
    const snb_vector rA = crd[a.x];
    last_ba[t] = settle_disp(crd[a.y], rA, box);
    last_ca[t] = settle_disp(crd[a.z], rA, box);
}

// remember_pair_AB (settle.cpp:25-44): j - i (minimum image) before the update
__global__ void __launch_bounds__(256) k_settle_pair_remember(int m, const int2* __restrict__ pairs,
                                                              const snb_vector* __restrict__ crd,
                                                              SnbBox box,
                                                              snb_vector* __restrict__ last_ab)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m) return;
    const int2 p = pairs[t];
    last_ab[t] = settle_disp(crd[p.y], crd[p.x], box);
}

// One rigid triangle per thread (settle_triangle + SETTLE_DO_TRIANGLE,
// settle.cpp:101-253, after Miyamoto & Kollman, J. Comput. Chem. 13, 952
// (1992)).  rB0, rC0: the previous step's B - A and C - A; rA1, rB1, rC1:
// this step's positions about the centre of mass.  The triangle's canonical
// frame {ra, rb, rc, rd, re} puts A at (0, ra), B at (rc, rb), C at (rd, re)
// with the centre of mass at the origin.  The per-triangle virial (FP64, to
// fixed point) is summed into vir6 when need_virial.
// The triangle update: positions pA, pB, pC (this step, unconstrained) and
// velocities vA, vB, vC in, constrained positions and corrected velocities
// out; w receives the fixed-point virial when need_virial.
__device__ __forceinline__ void settle_core(const float* __restrict__ g, float mA, float mB,
                                            float mC, snb_vector& pA, snb_vector& pB,
                                            snb_vector& pC, snb_vector& vA, snb_vector& vB,
                                            snb_vector& vC, snb_vector rB0, snb_vector rC0,
                                            const SnbBox& box, float dt, float v_factor,
                                            float x_factor, int need_virial, int64_t* w)
{
    {
        const float ra = g[0], rb = g[1], rc = g[2];
        const float rd = g[3], re = g[4];
        snb_vector rA = pA;
        snb_vector rB = settle_disp(pB, rA, box);
        snb_vector rC = settle_disp(pC, rA, box);
        // centre of mass rO = (mB rB + mC rC)/(mA + mB + mC) + rA
        const float inv_mt = 1.0f / ((mA + mB) + mC);
        const snb_vector rO = vadd(vscale(inv_mt, vadd(vscale(mB, rB), vscale(mC, rC))), rA);
        const snb_vector rA1 = vsub(rA, rO);
        const snb_vector rB1 = vadd(rB, rA1);
        const snb_vector rC1 = vadd(rC, rA1);
        // step 0: frame from the previous plane and this step's A
        snb_vector ez = vcross(rB0, rC0);
        snb_vector ex = vcross(rA1, ez);
        snb_vector ey = vcross(ez, ex);
        ex = vunit(ex);
        ey = vunit(ey);
        ez = vunit(ez);
        // step 1: projections
        const float b0x = vdot(ex, rB0), b0y = vdot(ey, rB0);
        const float c0x = vdot(ex, rC0), c0y = vdot(ey, rC0);
        const float a1z = vdot(ez, rA1);
        const float b1x = vdot(ex, rB1), b1y = vdot(ey, rB1), b1z = vdot(ez, rB1);
        const float c1x = vdot(ex, rC1), c1y = vdot(ey, rC1), c1z = vdot(ez, rC1);
        // step 2: rotations psi (about y) and phi (about x)
        const float sinphi = a1z / ra;
        const float cosphi = sqrtf(1.0f - sinphi * sinphi);
        const float sinpsi = ((b1z - c1z) - (rb - re) * sinphi) / ((rd - rc) * cosphi);
        const float cospsi = sqrtf(1.0f - sinpsi * sinpsi);
        const float a2y = ra * cosphi;
        const float b2x = rc * cospsi, b2y = rb * cosphi + (rc * sinpsi) * sinphi;
        const float c2x = rd * cospsi, c2y = re * cosphi + (rd * sinpsi) * sinphi;
        // step 3: alpha, beta, gamma
        const float alpha = ((b2x * b0x + c2x * c0x) + b2y * b0y) + c2y * c0y;
        const float beta = (((-b2x) * b0y - c2x * c0y) + b2y * b0x) + c2y * c0x;
        const float gamma = ((b1y * b0x - b1x * b0y) + c1y * c0x) - c1x * c0y;
        // step 4: rotation theta about z
        const float ab = alpha * alpha + beta * beta;
        const float sintheta = (alpha * gamma - beta * sqrtf(ab - gamma * gamma)) / ab;
        const float costheta = sqrtf(1.0f - sintheta * sintheta);
        const float a3x = -a2y * sintheta, a3y = a2y * costheta, a3z = a1z;
        const float b3x = b2x * costheta - b2y * sintheta;
        const float b3y = b2x * sintheta + b2y * costheta, b3z = b1z;
        const float c3x = c2x * costheta - c2y * sintheta;
        const float c3y = c2x * sintheta + c2y * costheta, c3z = c1z;
        // step 5: back to the lab frame
        const snb_vector rA3 = {(a3x * ex.x + a3y * ey.x) + a3z * ez.x,
                                (a3x * ex.y + a3y * ey.y) + a3z * ez.y,
                                (a3x * ex.z + a3y * ey.z) + a3z * ez.z};
        const snb_vector rB3 = {(b3x * ex.x + b3y * ey.x) + b3z * ez.x,
                                (b3x * ex.y + b3y * ey.y) + b3z * ez.y,
                                (b3x * ex.z + b3y * ey.z) + b3z * ez.z};
        const snb_vector rC3 = {(c3x * ex.x + c3y * ey.x) + c3z * ez.x,
                                (c3x * ex.y + c3y * ey.y) + c3z * ez.y,
                                (c3x * ex.z + c3y * ey.z) + c3z * ez.z};
        // step 6: velocity change and virial of the constraint forces
        const float kv = (v_factor / dt) / x_factor;
        const snb_vector dA = vsub(rA3, rA1), dB = vsub(rB3, rB1), dC = vsub(rC3, rC1);
        vA = vadd(vA, vscale(kv, dA));
        vB = vadd(vB, vscale(kv, dB));
        vC = vadd(vC, vscale(kv, dC));
        if (need_virial)
        {
            const float kf = ((1.0f / dt) / dt) / x_factor;
            const snb_vector fB = vscale(kf * mB, dB), fC = vscale(kf * mC, dC);
            // Get_Virial_From_Force_Dis(fB, rB0) + (fC, rC0), in FP64
            const double v[6] = {
                (double)fB.x * rB0.x + (double)fC.x * rC0.x,
                ((double)fB.x * rB0.y + (double)fB.y * rB0.x) +
                    ((double)fC.x * rC0.y + (double)fC.y * rC0.x),
                (double)fB.y * rB0.y + (double)fC.y * rC0.y,
                ((double)fB.x * rB0.z + (double)fB.z * rB0.x) +
                    ((double)fC.x * rC0.z + (double)fC.z * rC0.x),
                ((double)fB.y * rB0.z + (double)fB.z * rB0.y) +
                    ((double)fC.y * rC0.z + (double)fC.z * rC0.y),
                (double)fB.z * rB0.z + (double)fC.z * rC0.z};
            for (int k = 0; k < 6; k++) w[k] = snb_to_fixed_d(v[k]);
        }
        pA = vadd(rA3, rO);
        pB = vadd(rB3, rO);
        pC = vadd(rC3, rO);
    }
}

// sums each wavefront's triangle virials into vir6 (int64, order-free)
__device__ __forceinline__ void settle_virial_sum(const int64_t* w, int64_t* __restrict__ vir6)
{
    for (int k = 0; k < 6; k++)
    {
        int64_t s = w[k];
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
        if ((threadIdx.x & 63) == 0 && s != 0)
            atomicAdd((unsigned long long*)&vir6[k], (unsigned long long)s);
    }
}

// One rigid pair per thread (settle_pair, settle.cpp:258-307): with r1 this
// step's unconstrained j - i and r2 the previous step's, the displacement
// along r2 that restores |r1 + k r2| = r0, shared by the two atoms in inverse
// proportion to their masses; the velocities take the same change over the
// step, and the constraint force times r2 is the virial (in FP64, summed into
// vir6 when need_virial).  Float operations in the reference's order.
__global__ void __launch_bounds__(256) k_settle_pair(int m, const int2* __restrict__ pairs,
                                                     const float2* __restrict__ rk,
                                                     const float* __restrict__ mass,
                                                     snb_vector* __restrict__ crd, SnbBox box,
                                                     const snb_vector* __restrict__ last_ab,
                                                     float dt, float v_factor, float x_factor,
                                                     snb_vector* __restrict__ vel,
                                                     int need_virial, int64_t* __restrict__ vir6)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    int64_t w[6] = {0, 0, 0, 0, 0, 0};
    if (t < m)
    {
        const int2 p = pairs[t];
        const float2 c = rk[t];
        const snb_vector r1 = settle_disp(crd[p.y], crd[p.x], box);
        const snb_vector r2 = last_ab[t];
        const float mA = mass[p.x], mB = mass[p.y];
        const float r0r0 = c.x * c.x, r1r1 = vdot(r1, r1), r1r2 = vdot(r1, r2),
                    r2r2 = vdot(r2, r2);
        const float k = (sqrtf((r1r2 * r1r2 - r1r1 * r2r2) + r2r2 * r0r0) - r1r2) / r2r2;
        const snb_vector kr2 = vscale(k, r2);
        const snb_vector di = vscale((-mB) * c.y, kr2);
        const snb_vector dj = vscale(mA * c.y, kr2);
        crd[p.x] = vadd(crd[p.x], di);
        crd[p.y] = vadd(crd[p.y], dj);
        const float kv = (v_factor / dt) / x_factor;
        vel[p.x] = vadd(vel[p.x], vscale(kv, di));
        vel[p.y] = vadd(vel[p.y], vscale(kv, dj));
        if (need_virial)
        {
            // Get_Virial_From_Force_Dis(f, r2), f = (kv mB / dt / v_factor) dj
            const snb_vector f = vscale(((kv * mB) / dt) / v_factor, dj);
            const double v[6] = {(double)f.x * r2.x,
                                 (double)f.x * r2.y + (double)f.y * r2.x,
                                 (double)f.y * r2.y,
                                 (double)f.x * r2.z + (double)f.z * r2.x,
                                 (double)f.y * r2.z + (double)f.z * r2.y,
                                 (double)f.z * r2.z};
            for (int q = 0; q < 6; q++) w[q] = snb_to_fixed_d(v[q]);
        }
    }
    if (need_virial) settle_virial_sum(w, vir6);
}

__global__ void __launch_bounds__(256) k_settle(int m, const int4* __restrict__ tri,
                                                const float* __restrict__ geo,
                                                const float* __restrict__ mass,
                                                snb_vector* __restrict__ crd, SnbBox box,
                                                const snb_vector* __restrict__ last_ba,
                                                const snb_vector* __restrict__ last_ca,
                                                float dt, float v_factor, float x_factor,
                                                snb_vector* __restrict__ vel, int need_virial,
                                                int64_t* __restrict__ vir6)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    int64_t w[6] = {0, 0, 0, 0, 0, 0};
    if (t < m)
    {
        const int4 a = tri[t];
        snb_vector pA = crd[a.x], pB = crd[a.y], pC = crd[a.z];
        snb_vector vA = vel[a.x], vB = vel[a.y], vC = vel[a.z];
        settle_core(geo + 5 * t, mass[a.x], mass[a.y], mass[a.z], pA, pB, pC, vA, vB, vC,
                    last_ba[t], last_ca[t], box, dt, v_factor, x_factor, need_virial, w);
        vel[a.x] = vA;
        vel[a.y] = vB;
        vel[a.z] = vC;
        crd[a.x] = pA;
        crd[a.y] = pB;
        crd[a.z] = pC;
    }
    if (need_virial) settle_virial_sum(w, vir6);
}

// Fused iteration: per rigid triangle, SETTLE's remember (B - A, C - A before
// the update), the middle Langevin update of its three atoms and SETTLE; per
// free atom (in no triangle), the Langevin update alone.  Forces come from
// the fixed-point accumulators (direct + reciprocal, converted as a flush into
// a zeroed force array converts them) and are cleared.  Thread t < ntri owns
// triangle t, the next nfree threads the free atoms.
__device__ __forceinline__ void k_md_fused_body(
    int ntri, const int4* __restrict__ tri, const float* __restrict__ geo,
    const float* __restrict__ mass, int nfree, const int* __restrict__ free_atoms, int n,
    int64_t* __restrict__ fx, int64_t* __restrict__ fx_pme, const float* __restrict__ inv_mass,
    const float* __restrict__ sqrt_mass, uint32_t* __restrict__ counter, uint32_t seed,
    snb_vector* __restrict__ vel, snb_vector* __restrict__ crd, SnbBox box, float dt,
    float half_dt, float exp_gamma, float vmax, float v_factor, float x_factor, int need_virial,
    int64_t* __restrict__ vir6,
    unsigned bid, unsigned nbid)
{
    const int t = bid * blockDim.x + threadIdx.x;
    int64_t w[6] = {0, 0, 0, 0, 0, 0};
    auto force_of = [&](int i) {
        snb_vector f;
        f.x = 0.0f + snb_fixed_to_float(fx[i] + fx_pme[i]);
        f.y = 0.0f + snb_fixed_to_float(fx[n + i] + fx_pme[n + i]);
        f.z = 0.0f + snb_fixed_to_float(fx[2 * n + i] + fx_pme[2 * n + i]);
        fx[i] = 0;
        fx[n + i] = 0;
        fx[2 * n + i] = 0;
        fx_pme[i] = 0;
        fx_pme[n + i] = 0;
        fx_pme[2 * n + i] = 0;
        return f;
    };
    if (t < ntri)
    {
        const int4 a = tri[t];
        snb_vector pA = crd[a.x], pB = crd[a.y], pC = crd[a.z];
        const snb_vector rB0 = settle_disp(pB, pA, box), rC0 = settle_disp(pC, pA, box);
        snb_vector vA = vel[a.x], vB = vel[a.y], vC = vel[a.z];
        langevin_update(a.x, inv_mass, sqrt_mass, counter, seed, force_of(a.x), vA, pA, dt,
                        half_dt, exp_gamma, vmax);
        langevin_update(a.y, inv_mass, sqrt_mass, counter, seed, force_of(a.y), vB, pB, dt,
                        half_dt, exp_gamma, vmax);
        langevin_update(a.z, inv_mass, sqrt_mass, counter, seed, force_of(a.z), vC, pC, dt,
                        half_dt, exp_gamma, vmax);
        settle_core(geo + 5 * t, mass[a.x], mass[a.y], mass[a.z], pA, pB, pC, vA, vB, vC, rB0,
                    rC0, box, dt, v_factor, x_factor, need_virial, w);
        vel[a.x] = vA;
        vel[a.y] = vB;
        vel[a.z] = vC;
        crd[a.x] = pA;
        crd[a.y] = pB;
        crd[a.z] = pC;
    }
    else if (t - ntri < nfree)
    {
        const int i = free_atoms ? free_atoms[t - ntri] : t - ntri;
        snb_vector v = vel[i], x = crd[i];
        langevin_update(i, inv_mass, sqrt_mass, counter, seed, force_of(i), v, x, dt, half_dt,
                        exp_gamma, vmax);
        vel[i] = v;
        crd[i] = x;
    }
    if (need_virial) settle_virial_sum(w, vir6);
}

__global__ void __launch_bounds__(256) k_md_fused(
    int ntri, const int4* __restrict__ tri, const float* __restrict__ geo,
    const float* __restrict__ mass, int nfree, const int* __restrict__ free_atoms, int n,
    int64_t* __restrict__ fx, int64_t* __restrict__ fx_pme, const float* __restrict__ inv_mass,
    const float* __restrict__ sqrt_mass, uint32_t* __restrict__ counter, uint32_t seed,
    snb_vector* __restrict__ vel, snb_vector* __restrict__ crd, SnbBox box, float dt,
    float half_dt, float exp_gamma, float vmax, float v_factor, float x_factor, int need_virial,
    int64_t* __restrict__ vir6)
{
    k_md_fused_body(
        ntri, tri, geo, mass, nfree, free_atoms, n, fx, fx_pme, inv_mass, sqrt_mass, counter, seed,
        vel, crd, box, dt, half_dt, exp_gamma, vmax, v_factor, x_factor, need_virial, vir6,
        blockIdx.x, gridDim.x);
}

struct MdFusedArgs
{
    int ntri;
    const int4* tri;
    const float* geo;
    const float* mass;
    int nfree;
    const int* free_atoms;
    int n;
    int64_t* fx;
    int64_t* fx_pme;
    const float* inv_mass;
    const float* sqrt_mass;
    uint32_t* counter;
    uint32_t seed;
    snb_vector* vel;
    snb_vector* crd;
    SnbBox box;
    float dt;
    float half_dt;
    float exp_gamma;
    float vmax;
    float v_factor;
    float x_factor;
    int need_virial;
    int64_t* vir6;
};

__global__ void __launch_bounds__(256) k_md_fused_batch(SnbSegTab t)
{
    unsigned bid, nbid;
    const MdFusedArgs& a = snb_seg<MdFusedArgs>(t, bid, nbid);
    k_md_fused_body(
        a.ntri, a.tri, a.geo, a.mass, a.nfree, a.free_atoms, a.n, a.fx, a.fx_pme, a.inv_mass,
        a.sqrt_mass, a.counter, a.seed, a.vel, a.crd, a.box, a.dt, a.half_dt, a.exp_gamma, a.vmax,
        a.v_factor, a.x_factor, a.need_virial, a.vir6, bid, nbid);
}

// stress += (1/V) * summed virial (Sum_Virial_Tensor_To_Stress, settle.cpp:307-340);
// clears the sums
__global__ void k_settle_stress(int64_t* __restrict__ vir6, float inv_volume,
                                snb_ltmatrix3* __restrict__ stress)
{
    if (threadIdx.x != 0) return;
    snb_ltmatrix3 s = stress[0];
    s.a11 += inv_volume * snb_fixed_to_float(vir6[0]);
    s.a21 += inv_volume * snb_fixed_to_float(vir6[1]);
    s.a22 += inv_volume * snb_fixed_to_float(vir6[2]);
    s.a31 += inv_volume * snb_fixed_to_float(vir6[3]);
    s.a32 += inv_volume * snb_fixed_to_float(vir6[4]);
    s.a33 += inv_volume * snb_fixed_to_float(vir6[5]);
    stress[0] = s;
    for (int k = 0; k < 6; k++) vir6[k] = 0;
}

int snb_launch_leap_frog(snb_context* ctx, const float* inv_mass, const snb_vector* frc,
                         snb_vector* vel, snb_vector* crd, float dt, float vmax)
{
    SnbRecScope mark(0, true);  // writes coordinates (recorded step batches)
    const int n = ctx->n;
    if (n <= 0) return SNB_OK;
    SNB_LAUNCH(k_leap_frog, dim3((n + 255) / 256), dim3(256), 0,
               ctx->stream, n, inv_mass, frc, vel, crd, dt, vmax);
    SNB_CHECK(hipGetLastError());
    return SNB_OK;
}

int snb_launch_langevin(snb_context* ctx, const snb_vector* frc, snb_vector* vel,
                        snb_vector* crd)
{
    SnbRecScope mark(0, true);  // writes coordinates (recorded step batches)
    const SnbLangevin& T = ctx->langevin;
    const int n = ctx->n;
    if (n <= 0) return SNB_OK;
    SNB_LAUNCH(k_langevin, dim3((n + 255) / 256), dim3(256), 0,
               ctx->stream, n, T.inv_mass, T.sqrt_mass, T.counter, T.seed, frc, vel, crd, T.dt, T.half_dt, T.exp_gamma, T.max_velocity);
    SNB_CHECK(hipGetLastError());
    return SNB_OK;
}

int snb_launch_settle_remember(snb_context* ctx, const snb_vector* crd, const SnbBox& box)
{
    const SnbSettle& S = ctx->settle;
    if (S.npair > 0)
        SNB_LAUNCH(k_settle_pair_remember, dim3((S.npair + 255) / 256), dim3(256), 0, ctx->stream,
                   S.npair, S.pairs, crd, box, S.last_ab);
    if (S.count > 0)
        SNB_LAUNCH(k_settle_remember, dim3((S.count + 255) / 256), dim3(256), 0,
                   ctx->stream, S.count, S.tri, crd, box, S.last_ba, S.last_ca);
    SNB_CHECK(hipGetLastError());
    return SNB_OK;
}

int snb_launch_settle(snb_context* ctx, const float* mass, snb_vector* crd, const SnbBox& box,
                      snb_vector* vel, int need_pressure, snb_ltmatrix3* stress, float dt,
                      float v_factor, float x_factor)
{
    SnbRecScope mark(0, true);  // writes coordinates (recorded step batches)
    const SnbSettle& S = ctx->settle;
    const int want = need_pressure && stress ? 1 : 0;
    // rcell a11 a22 a33 is 1/V for the lower-triangular cell
    const float inv_volume = (box.rc[0] * box.rc[2]) * box.rc[5];
    // pairs, then triangles (SETTLE::Do_SETTLE, settle.cpp:660-700); they share no atom
    if (S.npair > 0)
    {
        SNB_LAUNCH(k_settle_pair, dim3((S.npair + 255) / 256), dim3(256), 0, ctx->stream,
                   S.npair, S.pairs, S.pair_rk, mass, crd, box, S.last_ab, dt, v_factor, x_factor,
                   vel, want, S.pair_vir6);
        if (want)
            SNB_LAUNCH(k_settle_stress, dim3(1), dim3(64), 0, ctx->stream, S.pair_vir6,
                       inv_volume, stress);
    }
    if (S.count > 0)
    {
        SNB_LAUNCH(k_settle, dim3((S.count + 255) / 256), dim3(256), 0, ctx->stream, S.count,
                   S.tri, S.geo, mass, crd, box, S.last_ba, S.last_ca, dt, v_factor, x_factor,
                   vel, want, S.vir6);
        if (want)
            SNB_LAUNCH(k_settle_stress, dim3(1), dim3(64), 0, ctx->stream, S.vir6, inv_volume,
                       stress);
    }
    SNB_CHECK(hipGetLastError());
    return SNB_OK;
}

// ---------------------------------------------------------------------------
// pressure assembly
// ---------------------------------------------------------------------------
// Rounds a double to int64 at 2^32 (nearest, ties to even); unlike the magic-
// constant conversion it covers |x| up to 2^31, which a system virial folded
// into one atom's entry can approach.
__device__ __forceinline__ int64_t stress_fixed(double x)
{
    return (int64_t)__builtin_rint(x * SNB_FIXED_SCALE);
}

// Sums over atoms of the atom virials (Get_Stress_Potential_Contribution,
// MD_core/sys.hpp:50-78) and of m v (x) v in the Get_Virial_From_Force_Dis
// layout, off-diagonals 2 m v_a v_b (Get_Stress_Kinetic_Contribution, :14-48).
// Each entry or FP64 product is rounded to int64 at 2^32 before any addition,
// so the sums are exact and independent of the order: acc[0..5] potential,
// acc[6..11] kinetic.
template <bool POT, bool KIN>
__global__ void __launch_bounds__(256) k_stress_sums(int n,
                                                     const snb_ltmatrix3* __restrict__ atom_virial,
                                                     const snb_vector* __restrict__ vel,
                                                     const float* __restrict__ mass,
                                                     int64_t* __restrict__ acc)
{
    int64_t w[12];
    for (int k = 0; k < 12; k++) w[k] = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    {
        if (POT)
        {
            const snb_ltmatrix3 v = atom_virial[i];
            w[0] += stress_fixed(v.a11);
            w[1] += stress_fixed(v.a21);
            w[2] += stress_fixed(v.a22);
            w[3] += stress_fixed(v.a31);
            w[4] += stress_fixed(v.a32);
            w[5] += stress_fixed(v.a33);
        }
        if (KIN)
        {
            const snb_vector u = vel[i];
            const double m = mass[i], x = u.x, y = u.y, z = u.z;
            w[6] += stress_fixed(m * (x * x));
            w[7] += stress_fixed(m * (2.0 * (x * y)));
            w[8] += stress_fixed(m * (y * y));
            w[9] += stress_fixed(m * (2.0 * (x * z)));
            w[10] += stress_fixed(m * (2.0 * (y * z)));
            w[11] += stress_fixed(m * (z * z));
        }
    }
    __shared__ int64_t part[4][12];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int k = POT ? 0 : 6; k < (KIN ? 12 : 6); k++)
    {
        int64_t s = w[k];
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
        if (lane == 0) part[wave][k] = s;
    }
    __syncthreads();
    const int k = threadIdx.x;
    if (k < 12 && (POT || k >= 6) && (KIN || k < 6))
    {
        const int64_t s = part[0][k] + part[1][k] + part[2][k] + part[3][k];
        if (s != 0) atomicAdd((unsigned long long*)&acc[k], (unsigned long long)s);
    }
}

// virial_tensor += potential sums; stress += inv_volume * virial_tensor
// (Get_Stress_From_virial, sys.hpp:80-85); stress += inv_volume * kinetic
// sums; clears the sums.
template <bool POT, bool KIN>
__global__ void k_stress_finish(int64_t* __restrict__ acc, float inv_volume,
                                snb_ltmatrix3* __restrict__ virial_tensor,
                                snb_ltmatrix3* __restrict__ stress)
{
    if (threadIdx.x != 0) return;
    snb_ltmatrix3 s = stress[0];
    if (POT)
    {
        snb_ltmatrix3 t = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        if (virial_tensor) t = virial_tensor[0];
        t.a11 += snb_fixed_to_float(acc[0]);
        t.a21 += snb_fixed_to_float(acc[1]);
        t.a22 += snb_fixed_to_float(acc[2]);
        t.a31 += snb_fixed_to_float(acc[3]);
        t.a32 += snb_fixed_to_float(acc[4]);
        t.a33 += snb_fixed_to_float(acc[5]);
        if (virial_tensor) virial_tensor[0] = t;
        s.a11 += inv_volume * t.a11;
        s.a21 += inv_volume * t.a21;
        s.a22 += inv_volume * t.a22;
        s.a31 += inv_volume * t.a31;
        s.a32 += inv_volume * t.a32;
        s.a33 += inv_volume * t.a33;
    }
    if (KIN)
    {
        s.a11 += inv_volume * snb_fixed_to_float(acc[6]);
        s.a21 += inv_volume * snb_fixed_to_float(acc[7]);
        s.a22 += inv_volume * snb_fixed_to_float(acc[8]);
        s.a31 += inv_volume * snb_fixed_to_float(acc[9]);
        s.a32 += inv_volume * snb_fixed_to_float(acc[10]);
        s.a33 += inv_volume * snb_fixed_to_float(acc[11]);
    }
    stress[0] = s;
    for (int k = 0; k < 12; k++) acc[k] = 0;
}

int snb_launch_stress(snb_context* ctx, int n, int pot, int kin, const snb_ltmatrix3* atom_virial,
                      const snb_vector* vel, const float* mass, float inv_volume,
                      snb_ltmatrix3* virial_tensor, snb_ltmatrix3* stress)
{
    if (!pot && !kin) return SNB_OK;
    int64_t* acc = ctx->scal + SNB_S_STRESS0;
    if (n > 0)
    {
        const int blocks = (n + 255) / 256 < 512 ? (n + 255) / 256 : 512;
#define SNB_STRESS_SUMS(P, K)                                                                     \
    SNB_LAUNCH((k_stress_sums<P, K>), dim3(blocks), dim3(256), 0, ctx->stream, n,      \
                       atom_virial, vel, mass, acc)
        if (pot && kin)
            SNB_STRESS_SUMS(true, true);
        else if (pot)
            SNB_STRESS_SUMS(true, false);
        else
            SNB_STRESS_SUMS(false, true);
#undef SNB_STRESS_SUMS
        SNB_CHECK(hipGetLastError());
    }
    if (pot && kin)
        SNB_LAUNCH((k_stress_finish<true, true>), dim3(1), dim3(64), 0,
                   ctx->stream, acc, inv_volume, virial_tensor, stress);
    else if (pot)
        SNB_LAUNCH((k_stress_finish<true, false>), dim3(1), dim3(64), 0,
                   ctx->stream, acc, inv_volume, virial_tensor, stress);
    else
        SNB_LAUNCH((k_stress_finish<false, true>), dim3(1), dim3(64), 0,
                   ctx->stream, acc, inv_volume, virial_tensor, stress);
    SNB_CHECK(hipGetLastError());
    return SNB_OK;
}

int snb_launch_md_fused(snb_context* ctx, const float* mass, snb_vector* crd, const SnbBox& box,
                        snb_vector* vel, int need_pressure, snb_ltmatrix3* stress)
{
    SnbRecScope mark(0, true);  // writes coordinates (recorded step batches)
    const SnbLangevin& T = ctx->langevin;
    const SnbSettle& S = ctx->settle;
    const int n = ctx->n;
    if (n <= 0) return SNB_OK;
    int rc = snb_join_pme(ctx);
    if (rc) return rc;
    const int ntri = S.init ? S.count : 0;
    const int nfree = S.init ? S.nfree : n;
    const int* free_atoms = S.init ? S.free_atoms : nullptr;
    const int want = need_pressure && stress && ntri > 0 ? 1 : 0;
    const int total = ntri + nfree;
    // constrain.v_factor / x_factor of the middle scheme (main.cpp:650-656)
    const float v_factor = T.exp_gamma, x_factor = 0.5f * T.exp_gamma + 0.5f;
    if (total > 0)
    {
        SNB_LAUNCH_B(k_md_fused, k_md_fused_batch, MdFusedArgs, dim3((total + 255) / 256),
                     dim3(256), 0,
                     ctx->stream, ntri, S.tri, S.geo, mass, nfree, free_atoms, n, ctx->fx, ctx->fx_pme, T.inv_mass, T.sqrt_mass, T.counter, T.seed, vel, crd, box, T.dt, T.half_dt, T.exp_gamma, T.max_velocity, v_factor, x_factor, want, S.vir6);
        SNB_CHECK(hipGetLastError());
    }
    if (want)
    {
        const float inv_volume = (box.rc[0] * box.rc[2]) * box.rc[5];
        SNB_LAUNCH(k_settle_stress, dim3(1), dim3(64), 0, ctx->stream, S.vir6, inv_volume, stress);
        SNB_CHECK(hipGetLastError());
    }
    return SNB_OK;
}
