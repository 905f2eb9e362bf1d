// snb_api.hip -- the extern "C" entry points of libsponge_nb (include/sponge_nb.h).
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "snb_internal.h"

namespace
{
template <class T>
int dalloc(T** p, size_t count)
{
    if (count == 0) count = 1;
    if (hipMalloc((void**)p, sizeof(T) * count) != hipSuccess) return SNB_ERR_OUT_OF_MEMORY;
    return SNB_OK;
}

template <class T>
void dfree(T*& p)
{
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}


// PM_force.cpp:244-272: doubling, then 50 bisection steps, in float with the
// host libm erfc.
float get_beta(float cutoff, float tolerance)
{
    float beta = 0.0f, low, high = 1.0f, tempf;
    int ihigh = 1;
    while (1)
    {
        tempf = erfcf(high * cutoff) / cutoff;
        if (tempf <= tolerance) break;
        high *= 2;
        ihigh++;
        if (ihigh > 64) break;
    }
    ihigh += 50;
    low = 0.0f;
    for (int ilow = 1; ilow < ihigh; ilow++)
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

// common.cpp:521-557, :639-659
bool smooth_2357(int v)
{
    if (v <= 0) return false;
    const int p[4] = {2, 3, 5, 7};
    for (int k = 0; k < 4; k++)
        while (v % p[k] == 0) v /= p[k];
    return v == 1;
}

int fft_size(float length)
{
    int t = (int)ceil(length + 3) >> 2 << 2;
    if (t >= 60 && t <= 68)
        t = 64;
    else if (t >= 120 && t <= 136)
        t = 128;
    else if (t >= 240 && t <= 272)
        t = 256;
    else if (t >= 480 && t <= 544)
        t = 512;
    else if (t >= 960 && t <= 1088)
        t = 1024;
    while (!smooth_2357(t)) t += 4;
    return t;
}

bool orthorhombic(snb_ltmatrix3 c)
{
    return c.a21 == 0.0f && c.a31 == 0.0f && c.a32 == 0.0f;
}

void free_window(SnbNl& L);

void free_nl(SnbNl& L)
{
    dfree(L.nbr_cells);
    dfree(L.nbr_count);
    L.triclinic = false;
    dfree(L.cell_count);
    dfree(L.cell_start);
    dfree(L.atom_cell);
    dfree(L.atom_slot);
    dfree(L.sorted_id);
    dfree(L.sorted_crd);
    dfree(L.serial);
    dfree(L.dnl);
    dfree(L.need);
    dfree(L.old_crd);
    dfree(L.order);
    dfree(L.sorted_cell);
    free_window(L);
    L.init = false;
}

void free_window(SnbNl& L)
{
    dfree(L.nl16);
    dfree(L.row_count);
    dfree(L.row_split);
    dfree(L.win_size);
    dfree(L.stage_base);
    dfree(L.stage);
    dfree(L.inv_count);
    dfree(L.inv_list);
    dfree(L.pub_count);
    dfree(L.ex_rng);
    dfree(L.block_work);
    dfree(L.block_perm);
    L.windowed = false;
}

// Chooses the block of the windowed pair path (snb_window.h): the largest
// owned block whose expected window fills at most 75% of the LDS capacity and
// that still gives at least min_blocks blocks (two per compute unit, so the
// heaviest-first launch order can even out dense and sparse regions); when no
// block gives that many, the largest that fits.  Needs >= 5 cells on every
// axis (unique forward displacements).
bool choose_window_pass(int nx, int ny, int nz, double atoms_per_cell, int max_lds,
                        int min_blocks, SnbWin* W);
bool choose_window(int nx, int ny, int nz, double atoms_per_cell, int max_lds, int min_blocks,
                   SnbWin* W)
{
    return choose_window_pass(nx, ny, nz, atoms_per_cell, max_lds, min_blocks, W) ||
           choose_window_pass(nx, ny, nz, atoms_per_cell, max_lds, 1, W);
}

// LDS slots of the largest window one workgroup can hold: 3 x int64
// accumulator + an int slot map per slot, plus the fixed part
static int window_slots_max(int max_lds)
{
    int wcap = (max_lds - SNB_WIN_FIXED_LDS) / 28;
    return wcap > 65535 ? 65535 : wcap;
}

// Fills W for blocks of a x b x s cells when the grid allows that shape and
// its expected window (atoms_per_cell per cell) fits `fill` of the LDS.
static bool window_shape(int nx, int ny, int nz, int a, int b, int s, double atoms_per_cell,
                         int max_lds, double fill, SnbWin* W)
{
    if (nx < 5 || ny < 5 || nz < 5 || a < 1 || b < 1 || s < 1) return false;
    if (a > nx - 2 || b > ny - 4 || s > nz - 4) return false;
    if (2 * (a + 2) * (b + 4) > SNB_WIN_MAX_RANGES) return false;
    const int wcap = window_slots_max(max_lds);
    double expect = (double)(a + 2) * (b + 4) * (s + 4) * atoms_per_cell;
    if (expect > fill * wcap) return false;
    W->nx = nx;
    W->ny = ny;
    W->nz = nz;
    W->a = a;
    W->b = b;
    W->s = s;
    W->nbx = (nx + a - 1) / a;
    W->nby = (ny + b - 1) / b;
    W->nbz = (nz + s - 1) / s;
    W->nblock = W->nbx * W->nby * W->nbz;
    W->wu = a + 2;
    W->wv = b + 4;
    W->wz = s + 4;
    W->nrange = 2 * W->wu * W->wv;
    // request LDS for 1.3x the mean window (the densest windows of the
    // test systems reach 1.15x; a larger one takes the overflow path) so
    // that small kernels of the reciprocal stream can share the CU
    int want = (int)(1.3 * expect) + 64;
    W->wcap = want < wcap ? want : wcap;
    return true;
}

bool choose_window_pass(int nx, int ny, int nz, double atoms_per_cell, int max_lds,
                        int min_blocks, SnbWin* W)
{
    static const int cand[][3] = {{2, 2, 8}, {2, 2, 6}, {2, 2, 4}, {2, 2, 3}, {2, 2, 2},
                                  {1, 2, 4}, {1, 1, 4}, {1, 1, 2}, {1, 1, 1}};
    for (const auto& c : cand)
    {
        int a = c[0], b = c[1], s = c[2];
        long nblk = (long)((nx + a - 1) / a) * ((ny + b - 1) / b) * ((nz + s - 1) / s);
        if (nblk < min_blocks) continue;
        if (window_shape(nx, ny, nz, a, b, s, atoms_per_cell, max_lds, 0.75, W)) return true;
    }
    return false;
}

// blocks whose window can contain one cell, per axis (+1 when a partial last
// block makes origins closer than the block size across the wrap)
int window_cover(int extent, int size, int n)
{
    return (extent + size - 1) / size + (n % size ? 1 : 0);
}

void free_nb14(SnbNb14& M)
{
    dfree(M.a);
    dfree(M.b);
    dfree(M.A);
    dfree(M.B);
    dfree(M.cf);
    dfree(M.pair_e);
    M.count = 0;
    M.init = false;
}

void free_bond(SnbBond& M)
{
    dfree(M.a);
    dfree(M.b);
    dfree(M.k);
    dfree(M.r0);
    dfree(M.pair_e);
    M.count = 0;
    M.init = false;
}

void free_langevin(SnbLangevin& T)
{
    dfree(T.inv_mass);
    dfree(T.sqrt_mass);
    dfree(T.counter);
    T.init = false;
}

// the triangles' state (snb_settle_initial)
void free_settle_triangles(SnbSettle& S)
{
    dfree(S.tri);
    dfree(S.geo);
    dfree(S.last_ba);
    dfree(S.last_ca);
    dfree(S.vir6);
    dfree(S.free_atoms);
    S.geo_d.clear();
    S.count = 0;
    S.nfree = 0;
    S.disjoint = false;
    S.init = false;
}

// the pairs' state (snb_settle_pair_initial)
void free_settle_pairs(SnbSettle& S)
{
    dfree(S.pairs);
    dfree(S.pair_rk);
    dfree(S.last_ab);
    dfree(S.pair_vir6);
    S.npair = 0;
}

void free_settle(SnbSettle& S)
{
    free_settle_triangles(S);
    free_settle_pairs(S);
}

// Get_Rabcde_From_SSS (settle.cpp:50-98): the canonical frame of a triangle
// with sides AB, AC, BC and masses mA, mB, mC (host; float for the FP32 path,
// double for the FP64 one).
template <class T>
static void settle_frame(T rAB, T rAC, T rBC, T mA, T mB, T mC, T* out)
{
    const T mt = (mA + mB) + mC;
    T Ax = 0, Ay = 0, Bx = -rAB, By = 0;
    T ct = ((rBC * rBC - rAC * rAC) - rAB * rAB) / ((T(2) * rAC) * rAB);
    T Cx = rAC * ct;
    T st = std::sqrt(T(1) - ct * ct);
    T Cy = rAC * st;
    const T Ox = (Bx * mB + Cx * mC) / mt;
    const T Oy = (Cy * mC) / mt;
    Ax -= Ox;
    Ay -= Oy;
    Bx -= Ox;
    By -= Oy;
    Cx -= Ox;
    Cy -= Oy;
    ct = T(1) / std::sqrt(T(1) + ((Ax * Ax) / Ay) / Ay);
    st = (ct * Ax) / Ay;
    T ra = Ax * st + Ay * ct;
    T rc = Bx * ct - By * st;
    T rb = Bx * st + By * ct;
    T rd = Cx * ct - Cy * st;
    T re = Cx * st + Cy * ct;
    if (ra < T(0))
    {
        ra = -ra;
        rb = -rb;
        re = -re;
    }
    out[0] = ra;
    out[1] = rb;
    out[2] = rc;
    out[3] = rd;
    out[4] = re;
}

void free_pme(SnbPme& P)
{
    dfree(P.tw_x);
    dfree(P.tw_y);
    dfree(P.tw_z);
    dfree(P.bc);
    dfree(P.bmod);
    dfree(P.q);
    dfree(P.phi);
    dfree(P.spec);
    dfree(P.atom_u);
    dfree(P.atom_fr);
    dfree(P.brick_count);
    dfree(P.brick_start);
    dfree(P.atom_slot);
    dfree(P.brick_atoms);
    dfree(P.partials);
    P.init = false;
}

const int* ex_or(const int* given, const int* own) { return given ? given : own; }

// resolves the recorded event pairs of one group into done_ms / done_n
int prof_collect(snb_context* c, int g)
{
    SnbProf& P = c->prof;
    for (int k = 0; k < P.count[g]; k++)
    {
        hipEvent_t* e = P.ev + ((size_t)g * SNB_PROF_CAP + k) * 2;
        float ms = 0.0f;
        SNB_CHECK(hipEventElapsedTime(&ms, e[0], e[1]));
        P.done_ms[g] += ms;
        P.done_n[g] += 1;
    }
    P.count[g] = 0;
    return SNB_OK;
}
}  // namespace

// Makes the context stream wait for the reciprocal pipeline queued on
// pme_stream (ev_pme), if one is pending.
int snb_join_pme(snb_context* c)
{
    if (!c->pme_pending) return SNB_OK;
    SNB_CHECK(hipStreamWaitEvent(c->stream, c->ev_pme, 0));
    c->pme_pending = false;
    return SNB_OK;
}

// Force-only work of the fused mode (no float output) that the reciprocal
// stream takes over: the exclusion correction, the 1-4 pairs and the bonds
// read only coordinates, charges and module tables, so they may run beside
// the direct-space kernels, on pme_stream ahead of the reciprocal pipeline,
// adding into the reciprocal accumulator (fx_pme).  That balances the two
// streams of a step; the integer sums are the same either way.
static bool side_ok(const snb_context* c, int need_energy, int need_virial, const void* d_frc)
{
    return c->overlap && c->pme_stream && !need_energy && !need_virial && !d_frc;
}

// Points the context at pme_stream after the step's inputs are final (the
// event of the last NL update, or one recorded now); returns the home stream.
static int side_begin(snb_context* c, hipStream_t* home)
{
    *home = c->stream;
    if (!c->inputs_marked)
    {
        SNB_CHECK(hipEventRecord(c->ev_inputs, c->stream));
        c->inputs_marked = true;  // the reciprocal call of this step waits on the same event
        c->pme_waits_inputs = false;
    }
    // one wait per record: later work on pme_stream is ordered after it
    if (!c->pme_waits_inputs)
    {
        SNB_CHECK(hipStreamWaitEvent(c->pme_stream, c->ev_inputs, 0));
        c->pme_waits_inputs = true;
    }
    c->stream = c->pme_stream;
    return SNB_OK;
}

static int side_end(snb_context* c, hipStream_t home, int rc)
{
    c->stream = home;
    if (!rc && hipEventRecord(c->ev_pme, c->pme_stream) != hipSuccess) rc = SNB_ERR_HIP;
    if (!rc) c->pme_pending = true;
    return rc;
}

void snb_prof_begin(snb_context* c, int g)
{
    SnbProf& P = c->prof;
    if (!P.on) return;
    if (P.count[g] >= SNB_PROF_CAP)
    {
        (void)hipStreamSynchronize(c->stream);
        (void)prof_collect(c, g);
    }
    (void)hipEventRecord(P.ev[((size_t)g * SNB_PROF_CAP + P.count[g]) * 2], c->stream);
}

void snb_prof_end(snb_context* c, int g)
{
    SnbProf& P = c->prof;
    if (!P.on) return;
    (void)hipEventRecord(P.ev[((size_t)g * SNB_PROF_CAP + P.count[g]) * 2 + 1], c->stream);
    P.count[g] += 1;
}

extern "C" {

int snb_abi_version(void) { return 1; }

const char* snb_status_string(int code)
{
    switch (code)
    {
        case SNB_OK: return "ok";
        case SNB_ERR_ARGUMENT: return "invalid argument";
        case SNB_ERR_HIP: return "HIP runtime error";
        case SNB_ERR_BOX_TOO_SMALL: return "the box is too small for the neighbor-list cell grid";
        case SNB_ERR_UNSUPPORTED: return "unsupported configuration";
        case SNB_ERR_NOT_INITIALIZED: return "module not initialised";
        case SNB_ERR_OUT_OF_MEMORY: return "device allocation failed";
        case SNB_ERR_DEVICE_STATUS: return "a device status flag is set";
        case SNB_ERR_TOO_LARGE:
            return "atom_numbers is above SNB_NOALLOC_MAX_ATOMS, the capacity of the "
                   "allocation-free entry point's device array";
        default: return "unknown status";
    }
}

int snb_context_create(snb_context** out, int device, void* hip_stream, int atom_numbers)
{
    if (!out || atom_numbers < 0) return SNB_ERR_ARGUMENT;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
        return SNB_ERR_HIP;
    DeviceGuard g(device);
    snb_context* c = new snb_context();
    c->device = device;
    {
        int lds = 0;
        if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) ==
                hipSuccess &&
            lds > 0)
            c->max_lds = lds;
        int ncu = 0;
        if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) ==
                hipSuccess &&
            ncu > 0)
            c->num_cu = ncu;
    }
    c->stream = (hipStream_t)hip_stream;
    c->n = atom_numbers;
    int n = atom_numbers;
    int rc = SNB_OK;
    if (!rc) rc = dalloc(&c->atom4, n);
    if (!rc) rc = dalloc(&c->type, n);
    if (!rc) rc = dalloc(&c->status, 1);
    if (!rc) rc = dalloc(&c->fx, 3 * (size_t)n);
    if (!rc) rc = dalloc(&c->fx_pme, 3 * (size_t)n);
    if (!rc && hipStreamCreateWithFlags(&c->pme_stream, hipStreamNonBlocking) != hipSuccess)
        rc = SNB_ERR_HIP;
    if (!rc && hipEventCreateWithFlags(&c->ev_inputs, hipEventDisableTiming) != hipSuccess)
        rc = SNB_ERR_HIP;
    if (!rc && hipEventCreateWithFlags(&c->ev_pme, hipEventDisableTiming) != hipSuccess)
        rc = SNB_ERR_HIP;
    if (!rc) rc = dalloc(&c->e_atom, 5 * (size_t)n);
    if (!rc) rc = dalloc(&c->vir_atom, 6 * (size_t)n);
    if (!rc) rc = dalloc(&c->satom, n);
    if (!rc) rc = dalloc(&c->stype, n);
    if (!rc) rc = dalloc(&c->e_sorted, 2 * (size_t)n);
    if (!rc) rc = dalloc(&c->v_sorted, 6 * (size_t)n);
    if (!rc) rc = dalloc(&c->scal, SNB_S_COUNT);
    if (!rc) rc = dalloc(&c->dscal, SNB_D_COUNT);
    if (rc)
    {
        snb_context_destroy(c);
        return rc;
    }
    hipStream_t s = c->stream;
    snb_launch_fill_int(s, c->status, 1, 0);
    snb_launch_fill_int(s, c->type, n, 0);
    snb_launch_fill_i64(s, c->fx, 3 * (size_t)n);
    snb_launch_fill_i64(s, c->fx_pme, 3 * (size_t)n);
    snb_launch_fill_i64(s, c->e_atom, 5 * (size_t)n);
    snb_launch_fill_i64(s, c->vir_atom, 6 * (size_t)n);
    snb_launch_fill_i64(s, c->scal, SNB_S_COUNT);
    snb_launch_fill_i64(s, (int64_t*)c->dscal, SNB_D_COUNT);
    if (hipStreamSynchronize(s) != hipSuccess)
    {
        snb_context_destroy(c);
        return SNB_ERR_HIP;
    }
    *out = c;
    return SNB_OK;
}

int snb_context_destroy(snb_context* c)
{
    if (!c) return SNB_OK;
    DeviceGuard g(c->device);
    (void)hipStreamSynchronize(c->stream);
    if (c->pme_stream) (void)hipStreamSynchronize(c->pme_stream);
    if (c->ev_inputs) (void)hipEventDestroy(c->ev_inputs);
    if (c->ev_pme) (void)hipEventDestroy(c->ev_pme);
    if (c->pme_stream) (void)hipStreamDestroy(c->pme_stream);
    dfree(c->fx_pme);
    dfree(c->atom4);
    dfree(c->type);
    dfree(c->status);
    dfree(c->fx);
    dfree(c->e_atom);
    free_nb14(c->nb14);
    free_bond(c->bond);
    free_langevin(c->langevin);
    free_settle(c->settle);
    snb_f64_free(c);
    dfree(c->vir_atom);
    dfree(c->satom);
    dfree(c->stype);
    dfree(c->e_sorted);
    dfree(c->v_sorted);
    dfree(c->scal);
    dfree(c->dscal);
    dfree(c->ex_start);
    dfree(c->ex_list);
    dfree(c->ex_num);
    free_nl(c->nl);
    free_pme(c->pme);
    if (c->prof.ev)
    {
        for (size_t k = 0; k < (size_t)SNB_PROF_GROUPS * SNB_PROF_CAP * 2; k++)
            (void)hipEventDestroy(c->prof.ev[k]);
        free(c->prof.ev);
    }
    dfree(c->lj.a);
    dfree(c->lj.b);
    dfree(c->lj.ab2);
    dfree(c->lj.coul);
    delete c;
    return SNB_OK;
}

int snb_set_stream(snb_context* c, void* hip_stream)
{
    if (!c) return SNB_ERR_ARGUMENT;
    c->stream = (hipStream_t)hip_stream;
    c->inputs_marked = false;  // an inputs event recorded on the old stream is not reused
    c->pme_waits_inputs = false;
    return SNB_OK;
}

int snb_set_pme_overlap(snb_context* c, int enable)
{
    if (!c) return SNB_ERR_ARGUMENT;
    DeviceGuard g(c->device);
    int rc = snb_join_pme(c);
    if (rc) return rc;
    c->overlap = enable != 0;
    return SNB_OK;
}

int snb_synchronize(snb_context* c)
{
    if (!c) return SNB_ERR_ARGUMENT;
    DeviceGuard g(c->device);
    int rc = snb_join_pme(c);
    if (rc) return rc;
    SNB_CHECK(hipStreamSynchronize(c->stream));
    return SNB_OK;
}

int snb_read_status(snb_context* c, int clear, int* flags)
{
    if (!c || !flags) return SNB_ERR_ARGUMENT;
    DeviceGuard g(c->device);
    int rc = snb_join_pme(c);
    if (rc) return rc;
    SNB_CHECK(hipStreamSynchronize(c->stream));
    SNB_CHECK(hipMemcpy(flags, c->status, sizeof(int), hipMemcpyDeviceToHost));
    if (clear)
    {
        snb_launch_fill_int(c->stream, c->status, 1, 0);
        SNB_CHECK(hipStreamSynchronize(c->stream));
    }
    return SNB_OK;
}

int snb_set_exclusions(snb_context* c, const int* h_start, const int* h_list, const int* h_num,
                       int total)
{
    if (!c || total < 0 || (c->n > 0 && (!h_start || !h_num)) || (total > 0 && !h_list))
        return SNB_ERR_ARGUMENT;
    // validate: every listed partner is an atom; per-atom ranges inside the list
    for (int i = 0; i < c->n; i++)
    {
        if (h_num[i] < 0 || h_start[i] < 0 || h_start[i] + h_num[i] > total)
            return SNB_ERR_ARGUMENT;
        for (int k = h_start[i]; k < h_start[i] + h_num[i]; k++)
            if (h_list[k] < 0 || h_list[k] >= c->n) return SNB_ERR_ARGUMENT;
    }
    DeviceGuard g(c->device);
    dfree(c->ex_start);
    dfree(c->ex_list);
    dfree(c->ex_num);
    int rc = dalloc(&c->ex_start, c->n);
    if (!rc) rc = dalloc(&c->ex_list, total);
    if (!rc) rc = dalloc(&c->ex_num, c->n);
    if (rc) return rc;
    if (c->n > 0)
    {
        SNB_CHECK(hipMemcpy(c->ex_start, h_start, sizeof(int) * c->n, hipMemcpyHostToDevice));
        SNB_CHECK(hipMemcpy(c->ex_num, h_num, sizeof(int) * c->n, hipMemcpyHostToDevice));
    }
    if (total > 0)
        SNB_CHECK(hipMemcpy(c->ex_list, h_list, sizeof(int) * total, hipMemcpyHostToDevice));
    c->ex_total = total;
    c->has_excl = true;
    return SNB_OK;
}

// ----------------------------------------------------------------- neighbor list
static bool window_list_fits(const snb_context* c)
{
    return (double)c->n * c->nl.max_nb * sizeof(uint16_t) < 4294967296.0;
}

// Allocates the windowed path's buffers for the shape in c->nl.win.
static int setup_window(snb_context* c)
{
    SnbNl& L = c->nl;
    const SnbWin& W = L.win;
    const int n = c->n;
    int rc = SNB_OK;
    L.invmax = window_cover(W.wu, W.a, W.nx) * window_cover(W.wv, W.b, W.ny) *
               window_cover(W.wz, W.s, W.nz);
    if (!rc) rc = dalloc(&L.nl16, (size_t)n * L.max_nb);
    if (!rc) rc = dalloc(&L.row_count, n);
    if (!rc) rc = dalloc(&L.row_split, n);
    if (!rc) rc = dalloc(&L.win_size, W.nblock + 1);
    if (!rc) rc = dalloc(&L.stage_base, W.nblock + 1);
    // every atom sits in at most invmax windows
    if (!rc) rc = dalloc(&L.stage, 3 * (size_t)n * L.invmax);
    if (!rc) rc = dalloc(&L.inv_count, n);
    if (!rc) rc = dalloc(&L.inv_list, (size_t)n * L.invmax);
    if (!rc) rc = dalloc(&L.pub_count, n);
    if (!rc) rc = dalloc(&L.ex_rng, n);
    if (!rc) rc = dalloc(&L.block_work, W.nblock);
    if (!rc) rc = dalloc(&L.block_perm, W.nblock);
    if (rc)
    {
        free_window(L);
        return rc;
    }
    L.windowed = true;
    snb_launch_fill_int(c->stream, L.row_count, n, 0);
    snb_launch_fill_int(c->stream, L.row_split, n, 0);
    float near = L.cutoff + 0.25f * L.skin;
    L.near2 = near * near;
    return SNB_OK;
}

// Find_Neighor_Grids_Device (neighbor_list.cpp:5-48) for a triclinic cell:
// cell j is a neighbour of cell i when some periodic image (sx, sy, sz in
// {-1, 0, 1}) puts j within [lx, 2] x [ly, 2] x [-2, 2] cells of i, an image
// moving x by sx nx + sy txy + sz txz cells, y by sy ny + sz tyz and z by
// sz nz, with the tilts in whole cells, (int)(a21 / grid_length) etc., and
// lx = -3 (ly = -3) where the image moves x (y) by a tilt.  Each cell's list
// in ascending cell order, as the reference's search over all cells finds it.
static int triclinic_neighbour_cells(SnbNl& L, const snb_ltmatrix3& cell)
{
    const int nx = L.nx, ny = L.ny, nz = L.nz, ncell = L.ncell;
    const int txy = (int)(cell.a21 / L.grid_length), txz = (int)(cell.a31 / L.grid_length),
              tyz = (int)(cell.a32 / L.grid_length);
    std::vector<int> cells((size_t)ncell * SNB_NL_MAX_NBR), count(ncell);
    std::vector<int> mark(ncell, -1), found;
    for (int ci = 0; ci < ncell; ci++)
    {
        const int ix = ci / (ny * nz), iy = (ci / nz) % ny, iz = ci % nz;
        found.clear();
        for (int sx = -1; sx <= 1; sx++)
            for (int sy = -1; sy <= 1; sy++)
                for (int sz = -1; sz <= 1; sz++)
                {
                    const int lx = sy * cell.a21 + sz * cell.a31 == 0.0f ? -2 : -3;
                    const int ly = sz * cell.a32 == 0.0f ? -2 : -3;
                    // delta = i + shift - j in [low, 2] per axis
                    const int bx = ix + sx * nx + sy * txy + sz * txz;
                    const int by = iy + sy * ny + sz * tyz;
                    const int bz = iz + sz * nz;
                    for (int jx = bx - 2; jx <= bx - lx; jx++)
                        for (int jy = by - 2; jy <= by - ly; jy++)
                            for (int jz = bz - 2; jz <= bz + 2; jz++)
                            {
                                if (jx < 0 || jx >= nx || jy < 0 || jy >= ny || jz < 0 || jz >= nz)
                                    continue;
                                const int cj = (jx * ny + jy) * nz + jz;
                                if (mark[cj] == ci) continue;
                                mark[cj] = ci;
                                found.push_back(cj);
                            }
                }
        if (found.size() > SNB_NL_MAX_NBR) return SNB_ERR_UNSUPPORTED;  // too skewed a cell
        std::sort(found.begin(), found.end());
        count[ci] = (int)found.size();
        std::copy(found.begin(), found.end(), cells.begin() + (size_t)ci * SNB_NL_MAX_NBR);
    }
    int rc = dalloc(&L.nbr_cells, cells.size());
    if (!rc) rc = dalloc(&L.nbr_count, (size_t)ncell);
    if (rc) return rc;
    SNB_CHECK(hipMemcpy(L.nbr_cells, cells.data(), sizeof(int) * cells.size(),
                        hipMemcpyHostToDevice));
    SNB_CHECK(hipMemcpy(L.nbr_count, count.data(), sizeof(int) * ncell, hipMemcpyHostToDevice));
    return SNB_OK;
}

int snb_nl_initial(snb_context* c, float cutoff, float skin, snb_ltmatrix3 cell,
                   snb_ltmatrix3 rcell, int max_neighbor_numbers, int max_atom_in_grid_numbers,
                   int refresh_interval, float skin_permit)
{
    (void)rcell;
    if (!c || !(cutoff > 0.0f) || !(skin >= 0.0f) || max_neighbor_numbers <= 0 ||
        max_atom_in_grid_numbers <= 0)
        return SNB_ERR_ARGUMENT;
    DeviceGuard g(c->device);
    SnbNl& L = c->nl;
    free_nl(L);
    L.triclinic = !orthorhombic(cell);
    L.cutoff = cutoff;
    L.skin = skin;
    // neighbor_list.cpp:801-803, :705 -- the same float expressions
    L.grid_length = (cutoff + skin) * 0.5f;
    L.cs2 = L.grid_length * L.grid_length * 4.0f;
    L.max_nb = max_neighbor_numbers;
    L.max_in_cell = max_atom_in_grid_numbers;
    L.refresh_interval = refresh_interval;
    L.skin_permit = skin_permit;
    // GRIDS::Initial (neighbor_list.cpp:58-60)
    L.nx = (int)floorf(cell.a11 / L.grid_length);
    L.ny = (int)floorf(cell.a22 / L.grid_length);
    L.nz = (int)floorf(cell.a33 / L.grid_length);
    if (L.nx <= 2 || L.ny <= 2 || L.nz <= 2) return SNB_ERR_BOX_TOO_SMALL;
    L.ncell = L.nx * L.ny * L.nz;
    int n = c->n;
    int rc = SNB_OK;
    if (!rc) rc = dalloc(&L.cell_count, L.ncell + 1);
    if (!rc) rc = dalloc(&L.cell_start, L.ncell + 1);
    if (!rc) rc = dalloc(&L.atom_cell, n);
    if (!rc) rc = dalloc(&L.atom_slot, n);
    if (!rc) rc = dalloc(&L.sorted_id, n);
    if (!rc) rc = dalloc(&L.sorted_crd, n);
    if (!rc) rc = dalloc(&L.serial, (size_t)n * L.max_nb);
    if (!rc) rc = dalloc(&L.dnl, n);
    if (!rc) rc = dalloc(&L.need, 1);
    if (!rc) rc = dalloc(&L.old_crd, n);
    if (!rc) rc = dalloc(&L.order, n);
    if (!rc) rc = dalloc(&L.sorted_cell, n);
    if (rc)
    {
        free_nl(L);
        return rc;
    }
    if (L.triclinic && (rc = triclinic_neighbour_cells(L, cell)))
    {
        free_nl(L);
        return rc;
    }
    // the windowed kernels address the list with 32-bit byte offsets; they
    // take orthorhombic cells
    L.windowed = !L.triclinic && window_list_fits(c) && max_atom_in_grid_numbers <= SNB_NLC_MAXA &&
                 choose_window(L.nx, L.ny, L.nz, (double)n / L.ncell, c->max_lds, 2 * c->num_cu,
                               &L.win);
    if (L.windowed && (rc = setup_window(c)))
    {
        free_nl(L);
        return rc;
    }
    // ATOM_GROUP rows point into the padded table (neighbor_list.cpp:764-771)
    std::vector<snb_atom_group> h(n > 0 ? n : 1);
    for (int i = 0; i < n; i++)
    {
        h[i].atom_numbers = 0;
        h[i].ghost_numbers = 0;
        h[i].atom_serial = L.serial + (size_t)i * L.max_nb;
    }
    if (n > 0) SNB_CHECK(hipMemcpy(L.dnl, h.data(), sizeof(snb_atom_group) * n, hipMemcpyHostToDevice));
    snb_launch_fill_int(c->stream, L.need, 1, 1);
    SNB_CHECK(hipStreamSynchronize(c->stream));
    L.init = true;
    return SNB_OK;
}

int snb_nl_update(snb_context* c, const int* d_atom_local, int local_atom_numbers,
                  int ghost_numbers, const snb_vector* d_crd, snb_ltmatrix3 cell,
                  snb_ltmatrix3 rcell, int step, int update, const int* d_ex_start,
                  const int* d_ex_list, const int* d_ex_num)
{
    if (!c) return SNB_ERR_ARGUMENT;
    if (!c->nl.init) return SNB_ERR_NOT_INITIALIZED;
    if (ghost_numbers != 0 || d_atom_local) return SNB_ERR_UNSUPPORTED;
    if (local_atom_numbers != c->n || (c->n > 0 && !d_crd)) return SNB_ERR_ARGUMENT;
    // a triclinic cell needs the neighbour-cell table made at snb_nl_initial
    if (!orthorhombic(cell) && !c->nl.triclinic) return SNB_ERR_UNSUPPORTED;
    // cells keep the width of the initial box (grid_length) and the last one on
    // each axis takes the remainder; a box that shrank below n cells of that
    // width would leave a thinner cell and pairs beyond the +-2-cell stencil:
    // re-initialise then (the reference does, main.cpp:880-884)
    if (cell.a11 < c->nl.nx * c->nl.grid_length || cell.a22 < c->nl.ny * c->nl.grid_length ||
        cell.a33 < c->nl.nz * c->nl.grid_length)
        return SNB_ERR_BOX_TOO_SMALL;
    if (c->n == 0) return SNB_OK;
    DeviceGuard g(c->device);
    SnbBox box = snb_box(cell, rcell);
    const int* es = ex_or(d_ex_start, c->ex_start);
    const int* el = ex_or(d_ex_list, c->ex_list);
    const int* en = ex_or(d_ex_num, c->ex_num);
    if (!d_ex_num && !c->has_excl) es = el = en = nullptr;
    // the step's coordinates are final here: the reciprocal pipeline of this
    // step may start from this point of the stream
    if (c->overlap)
    {
        SNB_CHECK(hipEventRecord(c->ev_inputs, c->stream));
        c->inputs_marked = true;
        c->pme_waits_inputs = false;
    }
    int forced = (update || !c->nl.built) ? 1 : 0;
    // With a fixed refresh interval the host knows the decision
    // (neighbor_list.cpp:843-856): a step that does not rebuild launches
    // nothing.  The displacement criterion (interval 0) is decided on the
    // device and gates the kernels there.
    if (!forced && c->nl.refresh_interval > 0 && (step + 1) % c->nl.refresh_interval != 0)
        return SNB_OK;
    int rc = snb_launch_nl_update(c, d_crd, box, forced, step, es, el, en);
    if (!rc) c->nl.built = true;
    return rc;
}

int snb_nl_device_list(snb_context* c, snb_atom_group** d_nl, int** d_serial, int* max_nb)
{
    if (!c) return SNB_ERR_ARGUMENT;
    if (!c->nl.init) return SNB_ERR_NOT_INITIALIZED;
    if (d_nl) *d_nl = c->nl.dnl;
    if (d_serial) *d_serial = c->nl.serial;
    if (max_nb) *max_nb = c->nl.max_nb;
    return SNB_OK;
}

__global__ void k_nl_counts(int n, const snb_atom_group* __restrict__ nl, int* __restrict__ out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = nl[i].atom_numbers;
}

int snb_nl_copy_rows(snb_context* c, int* d_counts, int* d_rows)
{
    if (!c || !d_counts || !d_rows) return SNB_ERR_ARGUMENT;
    if (!c->nl.init) return SNB_ERR_NOT_INITIALIZED;
    if (c->n == 0) return SNB_OK;
    DeviceGuard g(c->device);
    hipLaunchKernelGGL(k_nl_counts, dim3((c->n + 255) / 256), dim3(256), 0, c->stream, c->n,
                       c->nl.dnl, d_counts);
    SNB_CHECK(hipGetLastError());
    SNB_CHECK(hipMemcpyAsync(d_rows, c->nl.serial, sizeof(int) * (size_t)c->n * c->nl.max_nb,
                             hipMemcpyDeviceToDevice, c->stream));
    return SNB_OK;
}

int snb_lj_set_owner_attribution(snb_context* c, int enable)
{
    if (!c) return SNB_ERR_ARGUMENT;
    c->lj.owner_attribution = enable != 0;
    return SNB_OK;
}

int snb_nl_set_export(snb_context* c, int enable)
{
    if (!c) return SNB_ERR_ARGUMENT;
    if (enable < 0 || enable > 2) return SNB_ERR_ARGUMENT;
    c->nl.export_mode = enable;
    return SNB_OK;
}

int snb_nl_set_window_capacity(snb_context* c, int slots)
{
    if (!c || slots < 64) return SNB_ERR_ARGUMENT;
    if (!c->nl.init) return SNB_ERR_NOT_INITIALIZED;
    if (!c->nl.windowed) return SNB_OK;
    int most = (c->max_lds - SNB_WIN_FIXED_LDS) / 28;
    if (most > 65535) most = 65535;
    DeviceGuard g(c->device);
    int rc = snb_join_pme(c);
    if (rc) return rc;
    SNB_CHECK(hipStreamSynchronize(c->stream));  // no pair launch in flight uses the old size
    c->nl.win.wcap = slots < most ? slots : most;
    return SNB_OK;
}

int snb_nl_set_window_shape(snb_context* c, int a, int b, int s)
{
    if (!c) return SNB_ERR_ARGUMENT;
    if (snb_rec) return SNB_ERR_UNSUPPORTED;  // changes buffers: not inside a recording
    SnbNl& L = c->nl;
    if (!L.init) return SNB_ERR_NOT_INITIALIZED;
    const bool plain = a == 0 && b == 0 && s == 0;
    SnbWin W;
    if (!plain && (!window_list_fits(c) ||
                   !window_shape(L.nx, L.ny, L.nz, a, b, s, (double)c->n / L.ncell, c->max_lds,
                                 1.0, &W)))
        return SNB_ERR_ARGUMENT;
    // the shape in use: nothing to change
    if (plain ? !L.windowed : (L.windowed && L.win.a == W.a && L.win.b == W.b && L.win.s == W.s))
        return SNB_OK;
    DeviceGuard g(c->device);
    int rc = snb_join_pme(c);
    if (rc) return rc;
    SNB_CHECK(hipStreamSynchronize(c->stream));  // no launch in flight uses the old buffers
    free_window(L);
    L.built = false;  // the next snb_nl_update rebuilds for the new path
    if (plain) return SNB_OK;
    L.win = W;
    return setup_window(c);
}

int snb_lj_set_direct_j(snb_context* c, int mode)
{
    if (!c) return SNB_ERR_ARGUMENT;
    if (mode < -1 || mode > 1) return SNB_ERR_ARGUMENT;
    c->lj.j_direct = mode;
    return SNB_OK;
}

int snb_lj_set_row_stream(snb_context* c, int mode)
{
    if (!c) return SNB_ERR_ARGUMENT;
    if (mode < -1 || mode > 1) return SNB_ERR_ARGUMENT;
    c->lj.row_stream = mode;
    return SNB_OK;
}

int snb_lj_set_row_lds(snb_context* c, int mode)
{
    if (!c) return SNB_ERR_ARGUMENT;
    if (mode < -1 || mode > 1) return SNB_ERR_ARGUMENT;
    c->lj.row_lds = mode;
    return SNB_OK;
}

int snb_nl_window_info(snb_context* c, int* windowed, int* a, int* b, int* s, int* wcap,
                       int* nblock)
{
    if (!c) return SNB_ERR_ARGUMENT;
    if (!c->nl.init) return SNB_ERR_NOT_INITIALIZED;
    const SnbWin& W = c->nl.win;
    bool on = c->nl.windowed;
    if (windowed) *windowed = on ? 1 : 0;
    if (a) *a = on ? W.a : 0;
    if (b) *b = on ? W.b : 0;
    if (s) *s = on ? W.s : 0;
    if (wcap) *wcap = on ? W.wcap : 0;
    if (nblock) *nblock = on ? W.nblock : 0;
    return SNB_OK;
}

int snb_nl_copy_internal(snb_context* c, int* d_counts, int* d_rows)
{
    if (!c || !d_counts || !d_rows) return SNB_ERR_ARGUMENT;
    if (!c->nl.init) return SNB_ERR_NOT_INITIALIZED;
    DeviceGuard g(c->device);
    if (!c->nl.windowed) return snb_nl_copy_rows(c, d_counts, d_rows);
    return snb_launch_nl_copy_internal(c, d_counts, d_rows);
}

int snb_nl_grid_shape(snb_context* c, int* nx, int* ny, int* nz)
{
    if (!c) return SNB_ERR_ARGUMENT;
    if (!c->nl.init) return SNB_ERR_NOT_INITIALIZED;
    if (nx) *nx = c->nl.nx;
    if (ny) *ny = c->nl.ny;
    if (nz) *nz = c->nl.nz;
    return SNB_OK;
}

// ----------------------------------------------------------------------- LJ
int snb_lj_initial(snb_context* c, int ntypes, const float* h_a, const float* h_b,
                   const int* h_type, float cutoff)
{
    if (!c || ntypes <= 0 || !h_a || !h_b || (c->n > 0 && !h_type) || !(cutoff > 0.0f))
        return SNB_ERR_ARGUMENT;
    for (int i = 0; i < c->n; i++)
        if (h_type[i] < 0 || h_type[i] >= ntypes) return SNB_ERR_ARGUMENT;
    DeviceGuard g(c->device);
    SnbLj& J = c->lj;
    dfree(J.a);
    dfree(J.b);
    dfree(J.ab2);
    size_t npair = (size_t)ntypes * (ntypes + 1) / 2;
    int rc = dalloc(&J.a, npair);
    if (!rc) rc = dalloc(&J.b, npair);
    if (!rc) rc = dalloc(&J.ab2, (size_t)ntypes * ntypes);
    if (!rc && !J.coul) rc = dalloc(&J.coul, SNB_COUL_N);
    if (rc) return rc;
    SNB_CHECK(hipMemcpy(J.a, h_a, sizeof(float) * npair, hipMemcpyHostToDevice));
    SNB_CHECK(hipMemcpy(J.b, h_b, sizeof(float) * npair, hipMemcpyHostToDevice));
    // square table for the windowed kernel, both parameters times 2^32 (an
    // exact power-of-two scaling)
    std::vector<float2> ab2((size_t)ntypes * ntypes);
    for (int a = 0; a < ntypes; a++)
        for (int b = 0; b < ntypes; b++)
        {
            int p = snb_lj_pair_index(a, b);
            ab2[(size_t)a * ntypes + b] =
                make_float2(h_a[p] * 4294967296.0f, h_b[p] * 4294967296.0f);
        }
    SNB_CHECK(hipMemcpy(J.ab2, ab2.data(), sizeof(float2) * ab2.size(), hipMemcpyHostToDevice));
    J.coul_beta = -1.0f;
    if (c->n > 0) SNB_CHECK(hipMemcpy(c->type, h_type, sizeof(int) * c->n, hipMemcpyHostToDevice));
    J.ntypes = ntypes;
    J.cutoff = cutoff;
    // Total_C6_Get (Lennard_Jones_force.cpp:146-184) = sum over all ordered
    // atom pairs, i == j included, of B_ij; from the type histogram.
    std::vector<double> hist(ntypes, 0.0);
    for (int i = 0; i < c->n; i++) hist[h_type[i]] += 1.0;
    double tot = 0.0;
    for (int a = 0; a < ntypes; a++)
    {
        if (hist[a] == 0.0) continue;
        for (int b = 0; b < ntypes; b++)
        {
            if (hist[b] == 0.0) continue;
            int hi = a > b ? a : b, lo = a > b ? b : a;
            tot += hist[a] * hist[b] * (double)h_b[(size_t)hi * (hi + 1) / 2 + lo];
        }
    }
    J.c6_total = tot;
    // Lennard_Jones_force.cpp:267-268 in float
    float f = (float)tot;
    f *= -2.0f / 3.0f * 3.1415926535897932f / cutoff / cutoff / cutoff / 6.0f;
    J.long_range_factor = f;
    J.init = true;
    return SNB_OK;
}

// The Coulomb force-shape table for beta and the LJ cutoff, rebuilt when
// either changed.  The copy is ordered on the context stream after every
// kernel that may still read the old table.
static int coulomb_table(snb_context* c, float beta)
{
    SnbLj& J = c->lj;
    if (J.coul_beta == beta && J.coul_cutoff == J.cutoff) return SNB_OK;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    SNB_CHECK(hipStreamIsCapturing(c->stream, &cs));
    if (cs != hipStreamCaptureStatusNone) return SNB_ERR_UNSUPPORTED;  // not inside a capture
    SNB_CHECK(hipStreamSynchronize(c->stream));  // the previous host copy is consumed
    J.coul_host.assign(4 * SNB_COUL_N, 0.0f);
    J.coul_scale = snb_coul_table(beta, J.cutoff, J.coul_host.data());
    SNB_CHECK(hipMemcpyAsync(J.coul, J.coul_host.data(), sizeof(float) * 4 * SNB_COUL_N,
                             hipMemcpyHostToDevice, c->stream));
    J.coul_beta = beta;
    J.coul_cutoff = J.cutoff;
    return SNB_OK;
}

int snb_lj_pme_direct_force(snb_context* c, int atom_numbers, int local_atom_numbers,
                            int solvent_numbers, int ghost_numbers, const snb_vector* d_crd,
                            const float* d_charge, snb_vector* d_frc, snb_ltmatrix3 cell,
                            snb_ltmatrix3 rcell, const snb_atom_group* d_nl, float pme_beta,
                            int need_atom_energy, float* d_atom_energy, int need_virial,
                            snb_ltmatrix3* d_atom_virial, float* d_atom_direct_pme_energy)
{
    if (!c) return SNB_ERR_ARGUMENT;
    if (!c->lj.init) return SNB_ERR_NOT_INITIALIZED;
    if (solvent_numbers != 0 || ghost_numbers != 0) return SNB_ERR_UNSUPPORTED;
    if (atom_numbers != c->n || local_atom_numbers != c->n) return SNB_ERR_ARGUMENT;
    if (c->n == 0) return SNB_OK;
    if (!d_crd || !d_charge || !d_nl) return SNB_ERR_ARGUMENT;
    // beta rc above 4.9 means an Ewald tolerance below 1e-11 (erfc(4.9) = 4e-12)
    if (!(pme_beta >= 0.0f) || pme_beta * c->lj.cutoff > 4.9f) return SNB_ERR_ARGUMENT;
    DeviceGuard g(c->device);
    SnbBox box = snb_box(cell, rcell);
    int rc;
    if ((rc = coulomb_table(c, pme_beta))) return rc;
    if (need_atom_energy)
    {
        // the per-atom LJ and direct arrays restart each energy call
        // (Lennard_Jones_force.cpp:376-382)
        if ((rc = snb_launch_fill_i64(c->stream, c->e_atom, 2 * (size_t)c->n))) return rc;
    }
    // The library's own list goes through the windowed kernel (same pair set,
    // internal row layout); any other list through the row kernel.
    if (c->nl.init && c->nl.windowed && c->nl.built && d_nl == c->nl.dnl)
        rc = snb_launch_lj_window(c, d_crd, d_charge, box, pme_beta, need_atom_energy,
                                  need_virial);
    else
    {
        if ((rc = snb_launch_pack_atoms(c, d_crd, d_charge))) return rc;
        rc = snb_launch_lj_direct(c, d_nl, box, pme_beta, need_atom_energy, need_virial);
    }
    if (rc) return rc;
    if (d_frc && (rc = snb_launch_flush_forces(c, d_frc))) return rc;
    if (need_atom_energy)
    {
        if ((rc = snb_launch_sum_fixed(c, c->e_atom, c->n, SNB_S_E_LJ))) return rc;
        if ((rc = snb_launch_sum_fixed(c, c->e_atom + c->n, c->n, SNB_S_E_DIRECT))) return rc;
        if ((rc = snb_launch_flush_atom_energy(c, d_atom_energy, d_atom_direct_pme_energy, 3)))
            return rc;
    }
    if (need_virial && d_atom_virial && (rc = snb_launch_flush_virial(c, d_atom_virial)))
        return rc;
    return SNB_OK;
}

int snb_lj_long_range_correction(snb_context* c, int need_pressure, snb_ltmatrix3* d_virial,
                                 int need_potential, float* d_potential, float volume)
{
    if (!c) return SNB_ERR_ARGUMENT;
    if (!c->lj.init) return SNB_ERR_NOT_INITIALIZED;
    if (!(volume > 0.0f)) return SNB_ERR_ARGUMENT;
    DeviceGuard g(c->device);
    float f = c->lj.long_range_factor;
    c->lj.last_long_energy = need_potential ? (double)(f / volume) : c->lj.last_long_energy;
    snb_ltmatrix3* v = need_pressure ? d_virial : nullptr;
    float* p = need_potential ? d_potential : nullptr;
    if (!v && !p) return SNB_OK;
    return snb_launch_lj_long_range(c->stream, v, p, 2 * f / volume, f / volume);
}

int snb_lj_energy(snb_context* c, double* lj_short, double* lj_long)
{
    if (!c) return SNB_ERR_ARGUMENT;
    if (!c->lj.init) return SNB_ERR_NOT_INITIALIZED;
    DeviceGuard g(c->device);
    SNB_CHECK(hipStreamSynchronize(c->stream));
    int64_t v = 0;
    SNB_CHECK(hipMemcpy(&v, c->scal + SNB_S_E_LJ, sizeof(int64_t), hipMemcpyDeviceToHost));
    if (lj_short) *lj_short = snb_fixed_to_double(v);
    if (lj_long) *lj_long = c->lj.last_long_energy;
    return SNB_OK;
}

// --------------------------------------------------------------------- 1-4
int snb_nb14_initial(snb_context* c, int nb14_numbers, const int* h_atom_a, const int* h_atom_b,
                     const float* h_A, const float* h_B, const float* h_cf_scale_factor)
{
    if (!c || nb14_numbers < 0) return SNB_ERR_ARGUMENT;
    if (nb14_numbers > 0 && (!h_atom_a || !h_atom_b || !h_A || !h_B || !h_cf_scale_factor))
        return SNB_ERR_ARGUMENT;
    for (int k = 0; k < nb14_numbers; k++)
        if (h_atom_a[k] < 0 || h_atom_a[k] >= c->n || h_atom_b[k] < 0 || h_atom_b[k] >= c->n)
            return SNB_ERR_ARGUMENT;
    DeviceGuard g(c->device);
    SnbNb14& M = c->nb14;
    free_nb14(M);
    M.count = nb14_numbers;
    if (nb14_numbers > 0)
    {
        size_t m = (size_t)nb14_numbers;
        int rc = dalloc(&M.a, m);
        if (!rc) rc = dalloc(&M.b, m);
        if (!rc) rc = dalloc(&M.A, m);
        if (!rc) rc = dalloc(&M.B, m);
        if (!rc) rc = dalloc(&M.cf, m);
        if (!rc) rc = dalloc(&M.pair_e, 2 * m);
        if (rc)
        {
            free_nb14(M);
            return rc;
        }
        SNB_CHECK(hipMemcpy(M.a, h_atom_a, sizeof(int) * m, hipMemcpyHostToDevice));
        SNB_CHECK(hipMemcpy(M.b, h_atom_b, sizeof(int) * m, hipMemcpyHostToDevice));
        SNB_CHECK(hipMemcpy(M.A, h_A, sizeof(float) * m, hipMemcpyHostToDevice));
        SNB_CHECK(hipMemcpy(M.B, h_B, sizeof(float) * m, hipMemcpyHostToDevice));
        SNB_CHECK(hipMemcpy(M.cf, h_cf_scale_factor, sizeof(float) * m, hipMemcpyHostToDevice));
        SNB_CHECK(hipMemset(M.pair_e, 0, sizeof(int64_t) * 2 * m));
    }
    // no pairs: not initialised, and the force call does nothing, as in the
    // reference (nb14.cpp:165-168, :288)
    M.init = nb14_numbers > 0;
    return SNB_OK;
}

int snb_nb14_force(snb_context* c, const snb_vector* d_crd, const float* d_charge,
                   snb_ltmatrix3 cell, snb_ltmatrix3 rcell, snb_vector* d_frc,
                   int need_atom_energy, float* d_atom_energy, int need_virial,
                   snb_ltmatrix3* d_atom_virial)
{
    if (!c) return SNB_ERR_ARGUMENT;
    if (!c->nb14.init) return SNB_OK;
    if (!d_crd || !d_charge) return SNB_ERR_ARGUMENT;
    DeviceGuard g(c->device);
    SnbBox box = snb_box(cell, rcell);
    int rc;
    // the module's per-atom energy restarts each energy call
    if (need_atom_energy &&
        (rc = snb_launch_fill_i64(c->stream, c->e_atom + 3 * (size_t)c->n, c->n)))
        return rc;
    if (side_ok(c, need_atom_energy, need_virial, d_frc))
    {
        hipStream_t home;
        if ((rc = side_begin(c, &home))) return rc;
        return side_end(c, home, snb_launch_nb14(c, d_crd, d_charge, box, 0, 0, c->fx_pme));
    }
    if ((rc = snb_launch_nb14(c, d_crd, d_charge, box, need_atom_energy, need_virial, c->fx)))
        return rc;
    if (d_frc && (rc = snb_launch_flush_forces(c, d_frc))) return rc;
    if (need_atom_energy)
    {
        const int m = c->nb14.count;
        if ((rc = snb_launch_sum_fixed(c, c->nb14.pair_e, m, SNB_S_E_NB14_LJ))) return rc;
        if ((rc = snb_launch_sum_fixed(c, c->nb14.pair_e + m, m, SNB_S_E_NB14_EE))) return rc;
        if ((rc = snb_launch_flush_atom_energy(c, d_atom_energy, nullptr, 8))) return rc;
    }
    if (need_virial && d_atom_virial && (rc = snb_launch_flush_virial(c, d_atom_virial)))
        return rc;
    return SNB_OK;
}

int snb_nb14_energy(snb_context* c, double* lj14, double* ee14)
{
    if (!c) return SNB_ERR_ARGUMENT;
    if (!c->nb14.init)
    {
        if (lj14) *lj14 = 0.0;
        if (ee14) *ee14 = 0.0;
        return SNB_OK;
    }
    DeviceGuard g(c->device);
    SNB_CHECK(hipStreamSynchronize(c->stream));
    int64_t v[2] = {0, 0};
    SNB_CHECK(hipMemcpy(v, c->scal + SNB_S_E_NB14_LJ, sizeof(v), hipMemcpyDeviceToHost));
    if (lj14) *lj14 = snb_fixed_to_double(v[0]);
    if (ee14) *ee14 = snb_fixed_to_double(v[1]);
    return SNB_OK;
}

// -------------------------------------------------------------------- bonds
int snb_bond_initial(snb_context* c, int bond_numbers, const int* h_atom_a, const int* h_atom_b,
                     const float* h_k, const float* h_r0)
{
    if (!c || bond_numbers < 0) return SNB_ERR_ARGUMENT;
    if (bond_numbers > 0 && (!h_atom_a || !h_atom_b || !h_k || !h_r0)) return SNB_ERR_ARGUMENT;
    for (int k = 0; k < bond_numbers; k++)
        if (h_atom_a[k] < 0 || h_atom_a[k] >= c->n || h_atom_b[k] < 0 || h_atom_b[k] >= c->n)
            return SNB_ERR_ARGUMENT;
    DeviceGuard g(c->device);
    SnbBond& M = c->bond;
    free_bond(M);
    if (bond_numbers == 0) return SNB_OK;  // not initialised, as in the reference
    const size_t m = (size_t)bond_numbers;
    int rc = dalloc(&M.a, m);
    if (!rc) rc = dalloc(&M.b, m);
    if (!rc) rc = dalloc(&M.k, m);
    if (!rc) rc = dalloc(&M.r0, m);
    if (!rc) rc = dalloc(&M.pair_e, m);
    if (rc)
    {
        free_bond(M);
        return rc;
    }
    SNB_CHECK(hipMemcpy(M.a, h_atom_a, sizeof(int) * m, hipMemcpyHostToDevice));
    SNB_CHECK(hipMemcpy(M.b, h_atom_b, sizeof(int) * m, hipMemcpyHostToDevice));
    SNB_CHECK(hipMemcpy(M.k, h_k, sizeof(float) * m, hipMemcpyHostToDevice));
    SNB_CHECK(hipMemcpy(M.r0, h_r0, sizeof(float) * m, hipMemcpyHostToDevice));
    SNB_CHECK(hipMemset(M.pair_e, 0, sizeof(int64_t) * m));
    M.count = bond_numbers;
    M.init = true;
    return SNB_OK;
}

int snb_bond_force(snb_context* c, const snb_vector* d_crd, snb_ltmatrix3 cell,
                   snb_ltmatrix3 rcell, snb_vector* d_frc, int need_atom_energy,
                   float* d_atom_energy, int need_virial, snb_ltmatrix3* d_atom_virial)
{
    if (!c) return SNB_ERR_ARGUMENT;
    if (!c->bond.init) return SNB_OK;
    if (!d_crd) return SNB_ERR_ARGUMENT;
    DeviceGuard g(c->device);
    int rc;
    if (need_atom_energy &&
        (rc = snb_launch_fill_i64(c->stream, c->e_atom + 4 * (size_t)c->n, c->n)))
        return rc;
    if (side_ok(c, need_atom_energy, need_virial, d_frc))
    {
        hipStream_t home;
        if ((rc = side_begin(c, &home))) return rc;
        return side_end(c, home, snb_launch_bond(c, d_crd, snb_box(cell, rcell), 0, 0, c->fx_pme));
    }
    if ((rc = snb_launch_bond(c, d_crd, snb_box(cell, rcell), need_atom_energy, need_virial,
                              c->fx)))
        return rc;
    if (d_frc && (rc = snb_launch_flush_forces(c, d_frc))) return rc;
    if (need_atom_energy)
    {
        if ((rc = snb_launch_sum_fixed(c, c->bond.pair_e, c->bond.count, SNB_S_E_BOND))) return rc;
        if ((rc = snb_launch_flush_atom_energy(c, d_atom_energy, nullptr, 16))) return rc;
    }
    if (need_virial && d_atom_virial && (rc = snb_launch_flush_virial(c, d_atom_virial)))
        return rc;
    return SNB_OK;
}

int snb_bond_energy(snb_context* c, double* energy)
{
    if (!c) return SNB_ERR_ARGUMENT;
    if (!c->bond.init)
    {
        if (energy) *energy = 0.0;
        return SNB_OK;
    }
    DeviceGuard g(c->device);
    SNB_CHECK(hipStreamSynchronize(c->stream));
    int64_t v = 0;
    SNB_CHECK(hipMemcpy(&v, c->scal + SNB_S_E_BOND, sizeof(v), hipMemcpyDeviceToHost));
    if (energy) *energy = snb_fixed_to_double(v);
    return SNB_OK;
}

// -------------------------------------------------------------- integration
int snb_leap_frog(snb_context* c, const float* d_inverse_mass, const snb_vector* d_frc,
                  snb_vector* d_vel, snb_vector* d_crd, float dt, float max_velocity)
{
    if (!c) return SNB_ERR_ARGUMENT;
    if (c->n > 0 && (!d_inverse_mass || !d_frc || !d_vel || !d_crd)) return SNB_ERR_ARGUMENT;
    DeviceGuard g(c->device);
    return snb_launch_leap_frog(c, d_inverse_mass, d_frc, d_vel, d_crd, dt, max_velocity);
}

int snb_langevin_initial(snb_context* c, const float* h_mass, float target_temperature,
                         float tau, int seed, float dt, float max_velocity)
{
    if (!c || (c->n > 0 && !h_mass) || !(target_temperature >= 0.0f) || !(tau > 0.0f) ||
        !(dt > 0.0f))
        return SNB_ERR_ARGUMENT;
    DeviceGuard g(c->device);
    SnbLangevin& T = c->langevin;
    free_langevin(T);
    const int n = c->n;
    const size_t m = (size_t)(n > 0 ? n : 1);
    int rc = dalloc(&T.inv_mass, m);
    if (!rc) rc = dalloc(&T.sqrt_mass, m);
    if (!rc) rc = dalloc(&T.counter, m);
    if (rc)
    {
        free_langevin(T);
        return rc;
    }
    // Middle_Langevin_MD.cpp:172-245: friction 1/tau per ps, internal time
    // unit 1/20.455 ps, kB 0.00198716 kcal/mol/K
    const float time_unit = 20.455f, kb = 0.00198716f;
    const float dti = dt * time_unit;
    const float gamma = (1.0f / tau) / time_unit;
    const float eg = expf(-gamma * dti);
    const float sg = sqrtf((float)((1.0 - (double)(eg * eg)) * (double)target_temperature *
                                   (double)kb));
    std::vector<float> im(m, 0.0f), sm(m, 0.0f);
    for (int i = 0; i < n; i++)
    {
        if (h_mass[i] != 0.0f)
        {
            im[i] = 1.0f / h_mass[i];
            sm[i] = sg * sqrtf((float)(1.0 / (double)h_mass[i]));
        }
    }
    SNB_CHECK(hipMemcpy(T.inv_mass, im.data(), sizeof(float) * m, hipMemcpyHostToDevice));
    SNB_CHECK(hipMemcpy(T.sqrt_mass, sm.data(), sizeof(float) * m, hipMemcpyHostToDevice));
    SNB_CHECK(hipMemset(T.counter, 0, sizeof(uint32_t) * m));
    T.seed = (uint32_t)seed;
    T.dt = dti;
    T.half_dt = 0.5f * dti;
    T.exp_gamma = eg;
    T.max_velocity = max_velocity > 0.0f ? max_velocity : 0.0f;
    T.init = true;
    return SNB_OK;
}

int snb_langevin_iterate(snb_context* c, const snb_vector* d_frc, snb_vector* d_vel,
                         snb_vector* d_acc, snb_vector* d_crd)
{
    (void)d_acc;
    if (!c) return SNB_ERR_ARGUMENT;
    if (!c->langevin.init) return SNB_ERR_NOT_INITIALIZED;
    if (c->n > 0 && (!d_frc || !d_vel || !d_crd)) return SNB_ERR_ARGUMENT;
    DeviceGuard g(c->device);
    return snb_launch_langevin(c, d_frc, d_vel, d_crd);
}

int snb_langevin_parameters(snb_context* c, float* dt_internal, float* exp_gamma,
                            float* v_factor, float* x_factor)
{
    if (!c) return SNB_ERR_ARGUMENT;
    if (!c->langevin.init) return SNB_ERR_NOT_INITIALIZED;
    const SnbLangevin& T = c->langevin;
    if (dt_internal) *dt_internal = T.dt;
    if (exp_gamma) *exp_gamma = T.exp_gamma;
    // constrain.v_factor / x_factor as main.cpp:650-652 sets them
    if (v_factor) *v_factor = T.exp_gamma;
    if (x_factor) *x_factor = 0.5f * T.exp_gamma + 0.5f;
    return SNB_OK;
}

int snb_settle_initial(snb_context* c, int triangle_numbers, const int* h_atom_A,
                       const int* h_atom_B, const int* h_atom_C, const float* h_rAB,
                       const float* h_rAC, const float* h_rBC, const float* h_mass)
{
    if (!c || triangle_numbers < 0) return SNB_ERR_ARGUMENT;
    if (triangle_numbers > 0 &&
        (!h_atom_A || !h_atom_B || !h_atom_C || !h_rAB || !h_rAC || !h_rBC || !h_mass))
        return SNB_ERR_ARGUMENT;
    for (int t = 0; t < triangle_numbers; t++)
    {
        const int a = h_atom_A[t], b = h_atom_B[t], d = h_atom_C[t];
        if (a < 0 || a >= c->n || b < 0 || b >= c->n || d < 0 || d >= c->n || a == b ||
            a == d || b == d)
            return SNB_ERR_ARGUMENT;
        if (!(h_rAB[t] > 0.0f) || !(h_rAC[t] > 0.0f) || !(h_rBC[t] > 0.0f)) return SNB_ERR_ARGUMENT;
    }
    DeviceGuard g(c->device);
    snb_f64_free(c);  // the FP64 step's copy of the frames is made again on next use
    SnbSettle& S = c->settle;
    free_settle_triangles(S);
    if (triangle_numbers == 0) return SNB_OK;  // not initialised: Do_SETTLE is a no-op
    const size_t m = (size_t)triangle_numbers;
    int rc = dalloc(&S.tri, m);
    if (!rc) rc = dalloc(&S.geo, 5 * m);
    if (!rc) rc = dalloc(&S.last_ba, m);
    if (!rc) rc = dalloc(&S.last_ca, m);
    if (!rc) rc = dalloc(&S.vir6, (size_t)6);
    if (rc)
    {
        free_settle_triangles(S);
        return rc;
    }
    std::vector<int4> tri(m);
    std::vector<float> geo(5 * m);
    S.geo_d.assign(5 * m, 0.0);
    for (size_t t = 0; t < m; t++)
    {
        tri[t] = make_int4(h_atom_A[t], h_atom_B[t], h_atom_C[t], 0);
        settle_frame<float>(h_rAB[t], h_rAC[t], h_rBC[t], h_mass[h_atom_A[t]],
                            h_mass[h_atom_B[t]], h_mass[h_atom_C[t]], geo.data() + 5 * t);
        settle_frame<double>(h_rAB[t], h_rAC[t], h_rBC[t], h_mass[h_atom_A[t]],
                             h_mass[h_atom_B[t]], h_mass[h_atom_C[t]], S.geo_d.data() + 5 * t);
    }
    SNB_CHECK(hipMemcpy(S.tri, tri.data(), sizeof(int4) * m, hipMemcpyHostToDevice));
    SNB_CHECK(hipMemcpy(S.geo, geo.data(), sizeof(float) * 5 * m, hipMemcpyHostToDevice));
    SNB_CHECK(hipMemset(S.vir6, 0, sizeof(int64_t) * 6));
    // atoms outside every triangle, and whether triangles share atoms
    {
        std::vector<int> uses(c->n > 0 ? c->n : 1, 0);
        for (size_t t = 0; t < m; t++)
        {
            uses[h_atom_A[t]]++;
            uses[h_atom_B[t]]++;
            uses[h_atom_C[t]]++;
        }
        std::vector<int> fr;
        S.disjoint = true;
        for (int i = 0; i < c->n; i++)
        {
            if (uses[i] == 0) fr.push_back(i);
            if (uses[i] > 1) S.disjoint = false;
        }
        S.nfree = (int)fr.size();
        if ((rc = dalloc(&S.free_atoms, fr.size() > 0 ? fr.size() : (size_t)1)))
        {
            free_settle_triangles(S);
            return rc;
        }
        if (!fr.empty())
            SNB_CHECK(hipMemcpy(S.free_atoms, fr.data(), sizeof(int) * fr.size(),
                                hipMemcpyHostToDevice));
    }
    S.count = triangle_numbers;
    S.init = true;
    return SNB_OK;
}

int snb_settle_pair_initial(snb_context* c, int pair_numbers, const int* h_atom_i,
                            const int* h_atom_j, const float* h_constant_r, const float* h_mass)
{
    if (!c || pair_numbers < 0) return SNB_ERR_ARGUMENT;
    if (pair_numbers > 0 && (!h_atom_i || !h_atom_j || !h_constant_r || !h_mass))
        return SNB_ERR_ARGUMENT;
    std::vector<char> used(c->n > 0 ? c->n : 1, 0);
    for (int t = 0; t < pair_numbers; t++)
    {
        const int i = h_atom_i[t], j = h_atom_j[t];
        if (i < 0 || i >= c->n || j < 0 || j >= c->n || i == j || used[i] || used[j])
            return SNB_ERR_ARGUMENT;  // the reference's rigid pairs share no atom
        if (!(h_constant_r[t] > 0.0f) || !(h_mass[i] + h_mass[j] > 0.0f)) return SNB_ERR_ARGUMENT;
        used[i] = used[j] = 1;
    }
    DeviceGuard g(c->device);
    SnbSettle& S = c->settle;
    free_settle_pairs(S);
    if (pair_numbers == 0) return SNB_OK;
    const size_t m = (size_t)pair_numbers;
    int rc = dalloc(&S.pairs, m);
    if (!rc) rc = dalloc(&S.pair_rk, m);
    if (!rc) rc = dalloc(&S.last_ab, m);
    if (!rc) rc = dalloc(&S.pair_vir6, (size_t)6);
    if (rc)
    {
        free_settle_pairs(S);
        return rc;
    }
    std::vector<int2> pr(m);
    std::vector<float2> rk(m);
    for (size_t t = 0; t < m; t++)
    {
        pr[t] = make_int2(h_atom_i[t], h_atom_j[t]);
        // constrain_k = 1 / (m_i + m_j) in float (settle.cpp:494-495)
        rk[t] = make_float2(h_constant_r[t], 1.0f / (h_mass[h_atom_i[t]] + h_mass[h_atom_j[t]]));
    }
    SNB_CHECK(hipMemcpy(S.pairs, pr.data(), sizeof(int2) * m, hipMemcpyHostToDevice));
    SNB_CHECK(hipMemcpy(S.pair_rk, rk.data(), sizeof(float2) * m, hipMemcpyHostToDevice));
    SNB_CHECK(hipMemset(S.pair_vir6, 0, sizeof(int64_t) * 6));
    S.npair = pair_numbers;
    return SNB_OK;
}

int snb_settle_remember(snb_context* c, const snb_vector* d_crd, snb_ltmatrix3 cell,
                        snb_ltmatrix3 rcell)
{
    if (!c) return SNB_ERR_ARGUMENT;
    if (!c->settle.init && c->settle.npair == 0) return SNB_OK;
    if (!d_crd) return SNB_ERR_ARGUMENT;
    DeviceGuard g(c->device);
    return snb_launch_settle_remember(c, d_crd, snb_box(cell, rcell));
}

int snb_settle_do(snb_context* c, const float* d_mass, snb_vector* d_crd, snb_ltmatrix3 cell,
                  snb_ltmatrix3 rcell, snb_vector* d_vel, int need_pressure,
                  snb_ltmatrix3* d_stress, float dt, float v_factor, float x_factor)
{
    if (!c) return SNB_ERR_ARGUMENT;
    if (!c->settle.init && c->settle.npair == 0) return SNB_OK;
    if (!d_mass || !d_crd || !d_vel || !(dt > 0.0f) || !(x_factor > 0.0f) ||
        (c->settle.npair > 0 && !(v_factor > 0.0f)))
        return SNB_ERR_ARGUMENT;
    DeviceGuard g(c->device);
    return snb_launch_settle(c, d_mass, d_crd, snb_box(cell, rcell), d_vel, need_pressure,
                             d_stress, dt, v_factor, x_factor);
}

// ----------------------------------------------------------------- pressure
// 1 / V with V = cell.a11 cell.a22 cell.a33 in float (Get_Volume, sys.hpp:438-443)
static float inverse_volume(const snb_ltmatrix3& cell)
{
    const float volume = (cell.a11 * cell.a22) * cell.a33;
    return 1.0f / volume;
}

static int stress_call(snb_context* c, int atom_numbers, int pot, int kin,
                       const snb_ltmatrix3* d_atom_virial, const snb_vector* d_vel,
                       const float* d_mass, snb_ltmatrix3 cell, snb_ltmatrix3* d_virial_tensor,
                       snb_ltmatrix3* d_stress)
{
    if (!c || atom_numbers < 0 || !d_stress) return SNB_ERR_ARGUMENT;
    if (atom_numbers > 0 && ((pot && !d_atom_virial) || (kin && (!d_vel || !d_mass))))
        return SNB_ERR_ARGUMENT;
    if (!(cell.a11 > 0.0f && cell.a22 > 0.0f && cell.a33 > 0.0f)) return SNB_ERR_ARGUMENT;
    DeviceGuard g(c->device);
    // the reciprocal virial may still be in flight on the PME stream
    int rc = snb_join_pme(c);
    if (rc) return rc;
    return snb_launch_stress(c, atom_numbers, pot, kin, d_atom_virial, d_vel, d_mass,
                             inverse_volume(cell), d_virial_tensor, d_stress);
}

int snb_get_potential_to_stress(snb_context* c, int atom_numbers,
                                const snb_ltmatrix3* d_atom_virial, snb_ltmatrix3 cell,
                                snb_ltmatrix3* d_virial_tensor, snb_ltmatrix3* d_stress)
{
    return stress_call(c, atom_numbers, 1, 0, d_atom_virial, nullptr, nullptr, cell,
                       d_virial_tensor, d_stress);
}

int snb_get_kinetic_to_stress(snb_context* c, int atom_numbers, const snb_vector* d_vel,
                              const float* d_mass, snb_ltmatrix3 cell, snb_ltmatrix3* d_stress)
{
    return stress_call(c, atom_numbers, 0, 1, nullptr, d_vel, d_mass, cell, nullptr, d_stress);
}

int snb_get_pressure(snb_context* c, int atom_numbers, const snb_vector* d_vel,
                     const float* d_mass, const snb_ltmatrix3* d_atom_virial, snb_ltmatrix3 cell,
                     snb_ltmatrix3* d_virial_tensor, snb_ltmatrix3* d_stress,
                     snb_ltmatrix3* h_stress, float* h_pressure)
{
    int rc = stress_call(c, atom_numbers, 1, 1, d_atom_virial, d_vel, d_mass, cell,
                         d_virial_tensor, d_stress);
    if (rc || (!h_stress && !h_pressure)) return rc;
    DeviceGuard g(c->device);
    snb_ltmatrix3 s;
    SNB_CHECK(hipStreamSynchronize(c->stream));
    SNB_CHECK(hipMemcpy(&s, d_stress, sizeof(s), hipMemcpyDeviceToHost));
    if (h_stress) *h_stress = s;
    if (h_pressure) *h_pressure = ((s.a11 + s.a22) + s.a33) / 3.0f;
    return SNB_OK;
}

int snb_md_iterate_fused(snb_context* c, const float* d_mass, snb_vector* d_crd,
                         snb_ltmatrix3 cell, snb_ltmatrix3 rcell, snb_vector* d_vel,
                         int need_pressure, snb_ltmatrix3* d_stress)
{
    if (!c) return SNB_ERR_ARGUMENT;
    if (!c->langevin.init) return SNB_ERR_NOT_INITIALIZED;
    if (c->n > 0 && (!d_crd || !d_vel || (c->settle.init && !d_mass))) return SNB_ERR_ARGUMENT;
    if (c->settle.init && !c->settle.disjoint) return SNB_ERR_UNSUPPORTED;
    if (c->settle.npair > 0) return SNB_ERR_UNSUPPORTED;  // pairs: the separate calls
    DeviceGuard g(c->device);
    return snb_launch_md_fused(c, d_mass, d_crd, snb_box(cell, rcell), d_vel, need_pressure,
                               d_stress);
}

// ---------------------------------------------------------------------- PME
int snb_pme_initial(snb_context* c, snb_ltmatrix3 cell, snb_ltmatrix3 rcell,
                    snb_vector box_length, float cutoff,
                    int no_direct_interaction_virtual_atom_numbers, float tolerance, int fftx,
                    int ffty, int fftz, float grid_spacing, int update_interval)
{
    if (!c || !(cutoff > 0.0f) || !(tolerance > 0.0f) || update_interval <= 0 ||
        no_direct_interaction_virtual_atom_numbers < 0)
        return SNB_ERR_ARGUMENT;
    if (no_direct_interaction_virtual_atom_numbers > 0) return SNB_ERR_UNSUPPORTED;
    if (grid_spacing <= 0.0f) grid_spacing = 1.0f;
    DeviceGuard g(c->device);
    snb_f64_free(c);  // FP64 grids follow the PME grid
    SnbPme& P = c->pme;
    free_pme(P);
    P.cutoff = cutoff;
    P.tolerance = tolerance;
    P.update_interval = update_interval;
    P.fx = fftx > 0 ? fftx : fft_size(box_length.x / grid_spacing);
    P.fy = ffty > 0 ? ffty : fft_size(box_length.y / grid_spacing);
    P.fz = fftz > 0 ? fftz : fft_size(box_length.z / grid_spacing);
    P.beta = get_beta(cutoff, tolerance);
    // PM_force.cpp:382, :409 (volume from the cell diagonal)
    P.volume = (double)(cell.a11 * cell.a22 * cell.a33);
    P.neutralizing_factor = (float)(-0.5 * M_PI / ((double)P.beta * P.beta * P.volume));
    SnbBox box = snb_box(cell, rcell);
    int rc = snb_pme_setup_tables(c, box);
    if (rc)
    {
        free_pme(P);
        return rc;
    }
    SNB_CHECK(hipStreamSynchronize(c->stream));
    P.init = true;
    return SNB_OK;
}

int snb_pme_update_box(snb_context* c, snb_ltmatrix3 cell, snb_ltmatrix3 rcell, snb_ltmatrix3 g,
                       float dt)
{
    (void)g;  // g and dt move the domain-decomposition corners only (PM_force.cpp:1261-1262)
    (void)dt;
    if (!c) return SNB_ERR_ARGUMENT;
    if (snb_rec) return SNB_ERR_UNSUPPORTED;  // changes module state: not inside a recording
    if (!c->pme.init) return SNB_ERR_NOT_INITIALIZED;
    if (!(cell.a11 > 0.0f && cell.a22 > 0.0f && cell.a33 > 0.0f)) return SNB_ERR_ARGUMENT;
    DeviceGuard guard(c->device);
    SnbPme& P = c->pme;
    // the previous step's pipeline may still read BC on the PME stream
    int rc = snb_join_pme(c);
    if (rc) return rc;
    // PM_force.cpp:1254-1256 (volume from the cell diagonal, float)
    P.volume = (double)(cell.a11 * cell.a22 * cell.a33);
    P.neutralizing_factor = (float)(-0.5 * M_PI / ((double)P.beta * P.beta * P.volume));
    if ((rc = snb_pme_compute_bc(c, snb_box(cell, rcell)))) return rc;
    // a pipeline of this step that was already cleared to start must also
    // see the new table
    if (c->overlap && c->inputs_marked)
    {
        SNB_CHECK(hipEventRecord(c->ev_inputs, c->stream));
        c->pme_waits_inputs = false;
    }
    return SNB_OK;
}

int snb_pme_parameters(snb_context* c, float* beta, int* fx, int* fy, int* fz)
{
    if (!c) return SNB_ERR_ARGUMENT;
    if (!c->pme.init) return SNB_ERR_NOT_INITIALIZED;
    if (beta) *beta = c->pme.beta;
    if (fx) *fx = c->pme.fx;
    if (fy) *fy = c->pme.fy;
    if (fz) *fz = c->pme.fz;
    return SNB_OK;
}

int snb_pme_excluded_force(snb_context* c, int atom_numbers, const int* d_atom_local,
                           const int* d_atom_local_id, const snb_vector* d_crd,
                           snb_ltmatrix3 cell, snb_ltmatrix3 rcell, const float* d_charge,
                           const int* d_ex_start, const int* d_ex_list, const int* d_ex_num,
                           snb_vector* d_frc, int need_energy, float* d_atom_energy,
                           int need_virial, snb_ltmatrix3* d_atom_virial)
{
    if (!c) return SNB_ERR_ARGUMENT;
    if (!c->pme.init) return SNB_ERR_NOT_INITIALIZED;
    if (d_atom_local || d_atom_local_id) return SNB_ERR_UNSUPPORTED;
    if (atom_numbers != c->n) return SNB_ERR_ARGUMENT;
    if (c->n == 0) return SNB_OK;
    if (!d_crd || !d_charge) return SNB_ERR_ARGUMENT;
    const int* es = ex_or(d_ex_start, c->ex_start);
    const int* el = ex_or(d_ex_list, c->ex_list);
    const int* en = ex_or(d_ex_num, c->ex_num);
    if (!d_ex_num && !c->has_excl) return SNB_ERR_NOT_INITIALIZED;
    DeviceGuard g(c->device);
    SnbBox box = snb_box(cell, rcell);
    int rc;
    if (side_ok(c, need_energy, need_virial, d_frc))
    {
        hipStream_t home;
        if ((rc = side_begin(c, &home))) return rc;
        return side_end(c, home,
                        snb_launch_excl(c, d_crd, d_charge, box, c->pme.beta, es, el, en, 0, 0,
                                        c->fx_pme));
    }
    if (need_energy && (rc = snb_launch_fill_i64(c->stream, c->e_atom + 2 * (size_t)c->n, c->n)))
        return rc;
    rc = snb_launch_excl(c, d_crd, d_charge, box, c->pme.beta, es, el, en, need_energy,
                         need_virial, c->fx);
    if (rc) return rc;
    if (d_frc && (rc = snb_launch_flush_forces(c, d_frc))) return rc;
    if (need_energy)
    {
        if ((rc = snb_launch_sum_fixed(c, c->e_atom + 2 * (size_t)c->n, c->n, SNB_S_E_CORR)))
            return rc;
        if ((rc = snb_launch_flush_atom_energy(c, d_atom_energy, nullptr, 4))) return rc;
    }
    if (need_virial && d_atom_virial && (rc = snb_launch_flush_virial(c, d_atom_virial)))
        return rc;
    return SNB_OK;
}

int snb_pme_reciprocal_force(snb_context* c, const snb_vector* d_crd, snb_ltmatrix3 cell,
                             snb_ltmatrix3 rcell, const float* d_charge, snb_vector* d_frc,
                             int need_virial, int need_energy, snb_ltmatrix3* d_virial,
                             float* d_potential, int step)
{
    if (!c) return SNB_ERR_ARGUMENT;
    if (!c->pme.init) return SNB_ERR_NOT_INITIALIZED;
    if (c->n > 0 && (!d_crd || !d_charge)) return SNB_ERR_ARGUMENT;
    SnbPme& P = c->pme;
    for (int k = 0; k < 6; k++)
        if ((&rcell.a11)[k] != P.rcell[k]) return SNB_ERR_ARGUMENT;  // box changed: snb_pme_update_box
    DeviceGuard g(c->device);
    SnbBox box = snb_box(cell, rcell);
    int rc;
    // PM_force.cpp:1079-1163: the reciprocal force (scaled by update_interval)
    // and virial exist only on steps with step % update_interval == 0.  The
    // energy of any other step is that of the grids of the last update step,
    // plus the self term of the current charges.  So with an interval above 1
    // every update step evaluates the reciprocal energy (it is a by-product of
    // the convolution pass) and keeps it for the steps in between.
    bool do_force = (step % P.update_interval) == 0;
    if (!do_force)
    {
        need_virial = 0;
        if (need_energy)
        {
            // after the last update step's pipeline, which wrote the kept energy
            if ((rc = snb_join_pme(c))) return rc;
            if ((rc = snb_launch_pme_self(c, d_charge))) return rc;
            P.have_energy = true;
        }
    }
    else
    {
        int keep_energy = need_energy || P.update_interval > 1;
        // the pipeline reads only crd/charge and module buffers and writes
        // fx_pme / dscal / the Q slots, so it may run beside the direct-space
        // work: on pme_stream, after the step's inputs are final
        hipStream_t home = c->stream;
        if (c->overlap)
        {
            // ev_inputs is recorded after every earlier use of the reciprocal
            // accumulator on the context stream (a flush comes before the next
            // step's NL update), so the pipeline may start once it fires
            if (!c->inputs_marked)
            {
                SNB_CHECK(hipEventRecord(c->ev_inputs, home));
                c->pme_waits_inputs = false;
            }
            if (!c->pme_waits_inputs) SNB_CHECK(hipStreamWaitEvent(c->pme_stream, c->ev_inputs, 0));
            c->pme_waits_inputs = false;  // the next step records the event again
            c->inputs_marked = false;
            c->stream = c->pme_stream;
        }
        {
            // recorded step batches: the reciprocal lane
            SnbRecScope lane(1, false);
            rc = keep_energy ? snb_launch_charge_sums(c, d_charge) : SNB_OK;
            if (!rc)
                rc = snb_launch_pme_recip(c, d_crd, d_charge, box, keep_energy, need_virial, 1);
        }
        if (c->overlap)
        {
            c->stream = home;
            if (!rc && hipEventRecord(c->ev_pme, c->pme_stream) != hipSuccess) rc = SNB_ERR_HIP;
            if (!rc) c->pme_pending = true;
        }
        if (rc) return rc;
        P.have_energy = keep_energy != 0;
    }
    if (d_frc && (rc = snb_launch_flush_forces(c, d_frc))) return rc;
    if ((d_potential && need_energy) || (d_virial && need_virial))
        if ((rc = snb_join_pme(c))) return rc;
    if (need_energy && d_potential && (rc = snb_launch_pme_energy_to_potential(c, d_potential)))
        return rc;
    if (need_virial && d_virial && (rc = snb_launch_pme_virial_to(c, d_virial))) return rc;
    return SNB_OK;
}

int snb_pme_energy(snb_context* c, double* direct, double* reciprocal, double* self,
                   double* correction)
{
    if (!c) return SNB_ERR_ARGUMENT;
    DeviceGuard g(c->device);
    int rc = snb_join_pme(c);
    if (rc) return rc;
    SNB_CHECK(hipStreamSynchronize(c->stream));
    int64_t s[SNB_S_COUNT];
    double d[SNB_D_COUNT];
    SNB_CHECK(hipMemcpy(s, c->scal, sizeof(s), hipMemcpyDeviceToHost));
    SNB_CHECK(hipMemcpy(d, c->dscal, sizeof(d), hipMemcpyDeviceToHost));
    if (direct) *direct = snb_fixed_to_double(s[SNB_S_E_DIRECT]);
    if (correction) *correction = snb_fixed_to_double(s[SNB_S_E_CORR]);
    if (reciprocal) *reciprocal = d[SNB_D_E_REC];
    if (self) *self = d[SNB_D_E_SELF];
    return SNB_OK;
}

int snb_pme_reciprocal_virial(snb_context* c, double* v6)
{
    if (!c || !v6) return SNB_ERR_ARGUMENT;
    DeviceGuard g(c->device);
    int rc = snb_join_pme(c);
    if (rc) return rc;
    SNB_CHECK(hipStreamSynchronize(c->stream));
    double d[SNB_D_COUNT];
    SNB_CHECK(hipMemcpy(d, c->dscal, sizeof(d), hipMemcpyDeviceToHost));
    for (int k = 0; k < 6; k++) v6[k] = d[SNB_D_VIR0 + k];
    return SNB_OK;
}

// ------------------------------------------------------- fixed-point helpers
int snb_flush_forces(snb_context* c, snb_vector* d_frc)
{
    if (!c || (!d_frc && c->n > 0)) return SNB_ERR_ARGUMENT;
    DeviceGuard g(c->device);
    return snb_launch_flush_forces(c, d_frc);
}

int snb_flush_atom_energy(snb_context* c, float* d_atom_energy, float* d_direct)
{
    if (!c) return SNB_ERR_ARGUMENT;
    DeviceGuard g(c->device);
    return snb_launch_flush_atom_energy(c, d_atom_energy, d_direct, 31);
}

int snb_flush_atom_virial(snb_context* c, snb_ltmatrix3* d_atom_virial)
{
    if (!c || (!d_atom_virial && c->n > 0)) return SNB_ERR_ARGUMENT;
    DeviceGuard g(c->device);
    return snb_launch_flush_virial(c, d_atom_virial);
}

int snb_copy_fixed_forces(snb_context* c, int64_t* d_out)
{
    if (!c || !d_out) return SNB_ERR_ARGUMENT;
    DeviceGuard g(c->device);
    int rc = snb_join_pme(c);
    if (rc) return rc;
    return snb_launch_add_i64(c->stream, c->fx, c->fx_pme, d_out, 3 * (size_t)c->n);
}

int snb_clear_fixed_forces(snb_context* c)
{
    if (!c) return SNB_ERR_ARGUMENT;
    DeviceGuard g(c->device);
    int rc = snb_join_pme(c);
    if (rc) return rc;
    if ((rc = snb_launch_fill_i64(c->stream, c->fx_pme, 3 * (size_t)c->n))) return rc;
    return snb_launch_fill_i64(c->stream, c->fx, 3 * (size_t)c->n);
}

int snb_copy_fixed_atom_terms(snb_context* c, int which, int64_t* d_out)
{
    if (!c || !d_out || which < 0 || which > 5) return SNB_ERR_ARGUMENT;
    DeviceGuard g(c->device);
    const int64_t* src = which == 3 ? c->vir_atom
                                    : c->e_atom + (size_t)(which >= 4 ? which - 1 : which) * c->n;
    size_t cnt = which == 3 ? 6 * (size_t)c->n : (size_t)c->n;
    SNB_CHECK(hipMemcpyAsync(d_out, src, sizeof(int64_t) * cnt, hipMemcpyDeviceToDevice,
                             c->stream));
    return SNB_OK;
}

int snb_copy_pme_grid(snb_context* c, int which, float* d_out)
{
    if (!c || !d_out || which < 0 || which > 1) return SNB_ERR_ARGUMENT;
    if (!c->pme.init) return SNB_ERR_NOT_INITIALIZED;
    DeviceGuard g(c->device);
    int rc = snb_join_pme(c);
    if (rc) return rc;
    size_t nall = (size_t)c->pme.fx * c->pme.fy * c->pme.fz;
    SNB_CHECK(hipMemcpyAsync(d_out, which == 0 ? c->pme.q : c->pme.phi, sizeof(float) * nall,
                             hipMemcpyDeviceToDevice, c->stream));
    return SNB_OK;
}

int snb_fft_r2c(snb_context* c, const float* d_real, float* d_complex)
{
    if (!c || !d_real || !d_complex) return SNB_ERR_ARGUMENT;
    if (!c->pme.init) return SNB_ERR_NOT_INITIALIZED;
    DeviceGuard g(c->device);
    int rc = snb_join_pme(c);
    if (rc) return rc;
    return snb_fft_forward_full(c, d_real, (float2*)d_complex);
}

int snb_fft_c2r(snb_context* c, const float* d_complex, float* d_real)
{
    if (!c || !d_real || !d_complex) return SNB_ERR_ARGUMENT;
    if (!c->pme.init) return SNB_ERR_NOT_INITIALIZED;
    DeviceGuard g(c->device);
    int rc = snb_join_pme(c);
    if (rc) return rc;
    return snb_launch_fft_c2r(c, (const float2*)d_complex, d_real);
}

int snb_profile_enable(snb_context* c, int enable)
{
    if (!c) return SNB_ERR_ARGUMENT;
    DeviceGuard g(c->device);
    SnbProf& P = c->prof;
    if (enable && !P.ev)
    {
        size_t m = (size_t)SNB_PROF_GROUPS * SNB_PROF_CAP * 2;
        P.ev = (hipEvent_t*)calloc(m, sizeof(hipEvent_t));
        if (!P.ev) return SNB_ERR_OUT_OF_MEMORY;
        for (size_t k = 0; k < m; k++) SNB_CHECK(hipEventCreate(&P.ev[k]));
    }
    P.on = enable != 0;
    return SNB_OK;
}

int snb_profile_read(snb_context* c, int group, int reset, double* total_ms, int* launches)
{
    if (!c || group < 0 || group >= SNB_PROF_GROUPS) return SNB_ERR_ARGUMENT;
    DeviceGuard g(c->device);
    int rc0 = snb_join_pme(c);
    if (rc0) return rc0;
    SNB_CHECK(hipStreamSynchronize(c->stream));
    SnbProf& P = c->prof;
    if (P.ev)
    {
        int rc = prof_collect(c, group);
        if (rc) return rc;
    }
    if (total_ms) *total_ms = P.done_ms[group];
    if (launches) *launches = P.done_n[group];
    if (reset)
    {
        P.done_ms[group] = 0.0;
        P.done_n[group] = 0;
    }
    return SNB_OK;
}

}  // extern "C"
