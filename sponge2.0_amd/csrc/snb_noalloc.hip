// snb_noalloc.hip -- the exclusion correction without a context and without
// any device allocation, for callers that cannot allocate (code under HIP
// graph capture, serving loops, harnesses that forbid hipMalloc).
//
// The fixed-point accumulators that a context keeps in its own buffers live
// here in one file-scope device array, `scratch`, of SNB_NOALLOC_MAX_ATOMS x
// 10 int64 words (10 MiB): forces [3][n], the correction energy [n] and the
// atom virial [6][n].  Every device that calls the entry point holds its own
// copy until the process ends.  The host never uses the array's name as a
// device address: on a device's first call the address comes from
// hipGetSymbolAddress and is kept per device.  Each call clears the words it
// uses (fill kernels, no memset nodes), so calls on one device share and
// reuse the array; calls that overlap on one device race.
#include <mutex>

#include "snb_internal.h"

#define SNB_NOALLOC_MAX_DEVICES 64
#define SNB_NOALLOC_WORDS (10 * (size_t)SNB_NOALLOC_MAX_ATOMS)

static __device__ __align__(256) int64_t scratch[SNB_NOALLOC_WORDS];

static std::mutex g_scratch_mutex;
static int64_t* g_scratch_ptr[SNB_NOALLOC_MAX_DEVICES];

// Device address of `scratch` on the current device (looked up once).
static int scratch_address(int64_t** out)
{
    int dev = -1;
    SNB_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= SNB_NOALLOC_MAX_DEVICES) return SNB_ERR_UNSUPPORTED;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev])
    {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return SNB_ERR_HIP;
        g_scratch_ptr[dev] = (int64_t*)p;
    }
    *out = g_scratch_ptr[dev];
    return SNB_OK;
}

// frc[i] += fixed forces, energy[i] += correction energy, virial[i] += fixed
// virial (float conversion as the context path's flush kernels).
__global__ void k_noalloc_flush(int n, const int64_t* __restrict__ fx,
                                const int64_t* __restrict__ e, const int64_t* __restrict__ v,
                                snb_vector* __restrict__ frc, float* __restrict__ atom_energy,
                                snb_ltmatrix3* __restrict__ atom_virial)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (frc)
    {
        snb_vector f = frc[i];
        f.x += snb_fixed_to_float(fx[i]);
        f.y += snb_fixed_to_float(fx[n + i]);
        f.z += snb_fixed_to_float(fx[2 * n + i]);
        frc[i] = f;
    }
    if (atom_energy) atom_energy[i] += snb_fixed_to_float(e[i]);
    if (atom_virial)
    {
        snb_ltmatrix3 m = atom_virial[i];
        m.a11 += snb_fixed_to_float(v[i]);
        m.a21 += snb_fixed_to_float(v[(size_t)n + i]);
        m.a22 += snb_fixed_to_float(v[(size_t)2 * n + i]);
        m.a31 += snb_fixed_to_float(v[(size_t)3 * n + i]);
        m.a32 += snb_fixed_to_float(v[(size_t)4 * n + i]);
        m.a33 += snb_fixed_to_float(v[(size_t)5 * n + i]);
        atom_virial[i] = m;
    }
}

extern "C" int snb_pme_excluded_force_noalloc(
    int atom_numbers, const snb_vector* d_crd, snb_ltmatrix3 cell, snb_ltmatrix3 rcell,
    const float* d_charge, float beta, const int* d_excluded_list_start,
    const int* d_excluded_list, const int* d_excluded_numbers, snb_vector* d_frc,
    int need_energy, float* d_atom_energy, int need_virial, snb_ltmatrix3* d_atom_virial,
    void* hip_stream)
{
    if (atom_numbers < 0 || !(beta > 0.0f)) return SNB_ERR_ARGUMENT;
    if (atom_numbers > SNB_NOALLOC_MAX_ATOMS) return SNB_ERR_TOO_LARGE;
    if (atom_numbers == 0) return SNB_OK;
    if (!d_crd || !d_charge || !d_excluded_list_start || !d_excluded_list ||
        !d_excluded_numbers)
        return SNB_ERR_ARGUMENT;
    int64_t* base = nullptr;
    int rc = scratch_address(&base);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)hip_stream;
    const int n = atom_numbers;
    const bool want_e = need_energy && d_atom_energy;
    const bool want_v = need_virial && d_atom_virial;
    // scratch: fx [0, 3n), energy [3n, 4n), virial [4n, 10n); only the words
    // this call reads are cleared
    int64_t* fx = base;
    int64_t* e = base + 3 * (size_t)n;
    int64_t* v = base + 4 * (size_t)n;
    if ((rc = snb_launch_fill_i64(s, fx, (want_e ? 4 : 3) * (size_t)n))) return rc;
    if (want_v && (rc = snb_launch_fill_i64(s, v, 6 * (size_t)n))) return rc;
    // k_excl adds the correction energy of atom i to e_atom[2n + i]
    rc = snb_launch_excl_to(s, n, d_crd, d_charge, snb_box(cell, rcell), beta,
                            d_excluded_list_start, d_excluded_list, d_excluded_numbers,
                            want_e ? 1 : 0, want_v ? 1 : 0, fx, e - 2 * (size_t)n, v);
    if (rc) return rc;
    hipLaunchKernelGGL(k_noalloc_flush, dim3((n + 255) / 256), dim3(256), 0, s, n, fx, e, v,
                       d_frc, want_e ? d_atom_energy : nullptr,
                       want_v ? d_atom_virial : nullptr);
    SNB_CHECK(hipGetLastError());
    return SNB_OK;
}
