// snb_internal.h -- context layout and kernel launchers of libsponge_nb.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <functional>
#include <tuple>
#include <vector>

#include "../../include/sponge_nb.h"
#include "snb_math.h"
#include "snb_window.h"

// Scalar accumulator slots (int64 fixed point, device, zeroed per call that
// produces them).
enum SnbScalarSlot
{
    SNB_S_E_LJ = 0,      // sum of per-atom LJ energy
    SNB_S_E_DIRECT = 1,  // sum of per-atom direct Coulomb energy
    SNB_S_E_CORR = 2,    // sum of per-atom exclusion-correction energy
    SNB_S_Q2 = 3,        // sum q^2
    SNB_S_Q = 4,         // sum q
    SNB_S_E_NB14_LJ = 5,  // sum of 1-4 pair LJ energies
    SNB_S_E_NB14_EE = 6,  // sum of 1-4 pair Coulomb energies
    SNB_S_E_BOND = 7,     // sum of bond energies
    SNB_S_STRESS0 = 8,    // 12 words: atom-virial sums (6), then m v (x) v sums (6)
    SNB_S_COUNT = 20
};

// Double scalar slots written by the PME reduction.
enum SnbDoubleSlot
{
    SNB_D_E_REC = 0,
    SNB_D_VIR0 = 1,  // 6 slots: reciprocal virial a11..a33
    SNB_D_E_SELF = 7,
    SNB_D_COUNT = 8
};

// Atoms per cell the windowed list builder keeps row counts for in LDS.  The
// windowed path is taken only when max_atom_in_grid_numbers is at most this
// (the reference's default is 150), so its overflow flag keeps the meaning
// "a cell held more than max_atom_in_grid_numbers atoms".
#define SNB_NLC_MAXA 256
// neighbour cells per cell of a triclinic grid (MAX_GRID_NEIGHBORS, neighbor_list.cpp:3)
#define SNB_NL_MAX_NBR 192
#define SNB_FFT_MAX_STAGES 16
#define SNB_FFT_MAX_N 1024

struct SnbFftPlan
{
    int n;
    int nstages;
    int radix[SNB_FFT_MAX_STAGES];
};

struct SnbNl
{
    bool init = false;
    bool built = false;  // the first update always rebuilds (main.cpp:571-575)
    float cutoff = 10.0f, skin = 2.0f, grid_length = 6.0f, cs2 = 144.0f;
    float skin_permit = 0.5f;
    int refresh_interval = 0;
    int max_nb = 1200, max_in_cell = 150;
    int nx = 0, ny = 0, nz = 0, ncell = 0;
    int* cell_count = nullptr;   // ncell + 1
    int* cell_start = nullptr;   // ncell + 1 (exclusive scan)
    int* atom_cell = nullptr;    // n
    int* atom_slot = nullptr;    // n
    int* sorted_id = nullptr;    // n (scatter order)
    float4* sorted_crd = nullptr;  // n: x, y, z, bits(atom id), sorted by cell then id
    int* order = nullptr;        // n: atom id at sorted position s
    int* sorted_cell = nullptr;  // n: cell of sorted position s
    int* serial = nullptr;       // n * max_nb
    snb_atom_group* dnl = nullptr;  // n
    int* need = nullptr;         // rebuild flag (device)
    snb_vector* old_crd = nullptr;
    // triclinic cell: each cell's neighbour cells from a table (the stencil
    // is not a box of cells), SNB_NL_MAX_NBR per cell
    bool triclinic = false;
    int* nbr_cells = nullptr;  // ncell * SNB_NL_MAX_NBR
    int* nbr_count = nullptr;  // ncell
    // windowed pair path (every axis >= 5 cells)
    bool windowed = false;
    int export_mode = 1;  // reference-layout list at rebuild: 0 no, 1 yes, 2 yes with sorted rows
    SnbWin win;
    uint16_t* nl16 = nullptr;    // n * max_nb window slots, rows by sorted position
    int* row_count = nullptr;    // n: entries of each row
    int* row_split = nullptr;    // n: near pairs of each row (front part, snb_row_entry)
    float near2 = 0.0f;          // (cutoff + skin/4)^2: near/far split of the rows
    int* win_size = nullptr;     // nblock + 1
    int* stage_base = nullptr;   // nblock + 1 (exclusive scan of win_size)
    long long* stage = nullptr;  // nblock * wcap * 3
    int* inv_count = nullptr;    // n
    int* inv_list = nullptr;     // n * invmax: stage slots of each sorted atom
    int invmax = 0;
    int* pub_count = nullptr;    // n (export scratch)
    int2* ex_rng = nullptr;      // n: first/last exclusion partner per atom id
    int* block_work = nullptr;   // nblock: list entries owned by each block
    int* block_perm = nullptr;   // nblock: launch order, heaviest block first
    size_t win_lds_set = 0;      // dynamic LDS granted to the pair kernel
};

struct SnbLj
{
    bool init = false;
    int ntypes = 0;
    float cutoff = 10.0f;
    float* a = nullptr;  // pair tables
    float* b = nullptr;
    // (A, B) of type pair (ti, tj) at ti * ntypes + tj, scaled by 2^32
    float2* ab2 = nullptr;
    // Coulomb force shape table (SNB_COUL_N pieces) for coul_beta and the cutoff
    float4* coul = nullptr;
    float coul_scale = 0.0f, coul_beta = -1.0f, coul_cutoff = -1.0f;
    std::vector<float> coul_host;
    double c6_total = 0.0;
    float long_range_factor = 0.0f;
    double last_long_energy = 0.0;
    // per-atom energy/virial of a pair: false = the lower atom id (the
    // reference's row atom), true = the windowed list's row owner (cheaper)
    bool owner_attribution = false;
    // force-only pair loop: -1 by size, 0 pipelined within each row, 1
    // pipelined across row ends
    int row_stream = -1;
    // window forces of force-only steps: -1 by size, 0 through the stage and
    // its per-atom gather, 1 by atomic adds at the end of each block
    int j_direct = -1;
    // force-only row kernel of small systems with LDS accumulators: -1 inside
    // recorded step batches only, 0 never, 1 always
    int row_lds = -1;
};

// 1-4 pairs (NON_BOND_14): per pair the two atoms, A (x12 form), B (x6 form)
// and the Coulomb scale, resolved by the caller as the reference's loader does
struct SnbNb14
{
    bool init = false;
    int count = 0;
    int* a = nullptr;
    int* b = nullptr;
    float* A = nullptr;
    float* B = nullptr;
    float* cf = nullptr;
    int64_t* pair_e = nullptr;  // 2 * count, fixed point: LJ | Coulomb of each pair
};

// Harmonic bonds (BOND): atoms, k (E = k (r - r0)^2) and r0 per bond
struct SnbBond
{
    bool init = false;
    int count = 0;
    int* a = nullptr;
    int* b = nullptr;
    float* k = nullptr;
    float* r0 = nullptr;
    int64_t* pair_e = nullptr;  // count, fixed point
};

// Middle-scheme Langevin thermostat (MIDDLE_Langevin_INFORMATION); times in
// the internal unit (ps * 20.455)
#define SNB_LANGEVIN_KEY1 0x53504F4Eu  // second Philox key word ("SPON")
struct SnbLangevin
{
    bool init = false;
    float* inv_mass = nullptr;   // 1/m (0 for m = 0)
    float* sqrt_mass = nullptr;  // sqrt(kT (1 - e^{-2 gamma dt}) / m) (0 for m = 0)
    uint32_t* counter = nullptr; // per-atom random-stream position
    uint32_t seed = 0;
    float dt = 0.0f, half_dt = 0.0f, exp_gamma = 1.0f, max_velocity = 0.0f;
};

// SETTLE rigid triangles: atoms (A, B, C) and the canonical frame
// {ra, rb, rc, rd, re} per triangle
struct SnbSettle
{
    bool init = false;
    int count = 0;
    int4* tri = nullptr;
    float* geo = nullptr;           // 5 per triangle
    snb_vector* last_ba = nullptr;  // previous step's B - A
    snb_vector* last_ca = nullptr;  // previous step's C - A
    int64_t* vir6 = nullptr;        // summed constraint virial, fixed point
    int* free_atoms = nullptr;      // atoms in no triangle (the fused iteration's list)
    std::vector<double> geo_d;      // the canonical frames in double (FP64 step)
    int nfree = 0;
    bool disjoint = false;          // no atom in two triangles (needed by the fused iteration)
    // rigid two-atom pairs (CONSTRAIN_PAIR)
    int npair = 0;
    int2* pairs = nullptr;          // atom i, atom j
    float2* pair_rk = nullptr;      // constant_r, constrain_k = 1 / (m_i + m_j)
    snb_vector* last_ab = nullptr;  // previous step's j - i
    int64_t* pair_vir6 = nullptr;   // summed pair constraint virial, fixed point
};

struct SnbPme
{
    bool init = false;
    int fx = 0, fy = 0, fz = 0, fc = 0;
    int bx = 0, by = 0, bz = 0, nbx = 0, nby = 0, nbz = 0, nbrick = 0;
    float beta = 0.0f, tolerance = 1e-5f, cutoff = 10.0f;
    int update_interval = 1;
    double volume = 0.0;
    float neutralizing_factor = 0.0f;
    SnbFftPlan plan_x, plan_y, plan_z;
    float2* tw_x = nullptr;  // e^{-2 pi i k/n}
    float2* tw_y = nullptr;
    float2* tw_z = nullptr;
    float* bc = nullptr;      // [fx][fy][fc]
    double* bmod = nullptr;   // B-spline moduli: fx, then fy, then fz entries
    float* q = nullptr;       // [fx][fy][fz]
    float* phi = nullptr;     // [fx][fy][fz]
    float2* spec = nullptr;   // [fx][fy][fc]
    int4* atom_u = nullptr;   // ux, uy, uz, brick
    float4* atom_fr = nullptr;  // frx, fry, frz, q
    int* brick_count = nullptr;  // nbrick + 1 (zero between calls)
    int* brick_start = nullptr;  // nbrick + 1
    int* atom_slot = nullptr;
    int* brick_atoms = nullptr;
    double* partials = nullptr;  // conv blocks * 8
    int conv_blocks = 0;
    float rcell[6];
    bool have_energy = false;
    size_t plane_lds_set = 0;  // dynamic LDS granted to the fused plane FFT kernels
};

#define SNB_PROF_GROUPS 4
#define SNB_PROF_CAP 4096

// Event pairs recorded around kernel groups while profiling is enabled.
struct SnbProf
{
    bool on = false;
    hipEvent_t* ev = nullptr;  // SNB_PROF_GROUPS * SNB_PROF_CAP * 2
    int count[SNB_PROF_GROUPS] = {0, 0, 0, 0};
    double done_ms[SNB_PROF_GROUPS] = {0, 0, 0, 0};
    int done_n[SNB_PROF_GROUPS] = {0, 0, 0, 0};
};

struct SnbBox
{
    float c[6];   // cell a11 a21 a22 a31 a32 a33
    float rc[6];  // rcell
};

struct snb_context
{
    int device = 0;
    hipStream_t stream = nullptr;
    int n = 0;
    float4* atom4 = nullptr;  // x, y, z, q
    int* type = nullptr;
    float4* satom = nullptr;     // x, y, z, q in sorted (cell) order
    int* stype = nullptr;        // LJ type in sorted order
    int64_t* e_sorted = nullptr; // 2n: LJ | direct energy of row owners, sorted order
    int64_t* v_sorted = nullptr; // 6n: virial of row owners, sorted order
    int max_lds = 65536;         // dynamic LDS bytes per workgroup the device allows
    int num_cu = 256;            // compute units of the device
    int* status = nullptr;
    int64_t* fx = nullptr;       // 3n, SoA: direct-space and exclusion forces
    int64_t* fx_pme = nullptr;   // 3n, SoA: reciprocal forces (own accumulator: the PME
                                 // pipeline may run concurrently on pme_stream)
    int64_t* e_atom = nullptr;   // 5n: LJ | direct | correction | 1-4 | bond
    int64_t* vir_atom = nullptr; // 6n, SoA
    int64_t* scal = nullptr;     // SNB_S_COUNT
    double* dscal = nullptr;     // SNB_D_COUNT
    int* scan_tmp = nullptr;     // scratch for scans
    int scan_tmp_len = 0;
    // exclusions
    int* ex_start = nullptr;
    int* ex_list = nullptr;
    int* ex_num = nullptr;
    int ex_total = 0;
    bool has_excl = false;
    // Reciprocal pipeline on a second stream: it waits for ev_inputs (recorded
    // on `stream` when the step's coordinates are final -- at the NL update,
    // the first call of a step -- or at the PME call itself) and signals
    // ev_pme; everything that reads PME results on `stream` waits for ev_pme
    // first (snb_join_pme).
    bool overlap = true;
    hipStream_t pme_stream = nullptr;
    hipEvent_t ev_inputs = nullptr;
    hipEvent_t ev_pme = nullptr;
    bool inputs_marked = false;
    bool pme_waits_inputs = false;  // pme_stream already waits for the latest ev_inputs record
    bool pme_pending = false;
    SnbNl nl;
    SnbLj lj;
    SnbNb14 nb14;
    SnbBond bond;
    SnbLangevin langevin;
    SnbSettle settle;
    SnbPme pme;
    SnbProf prof;
    struct SnbF64* f64 = nullptr;  // FP64 evaluation state (snb_f64.hip), made on first use
};

// frees the FP64 evaluation state (snb_f64.hip)
void snb_f64_free(snb_context* ctx);

// Makes `stream` wait for the reciprocal pipeline launched last (no-op when
// none is pending).
int snb_join_pme(snb_context* ctx);

// Profiling brackets (no-ops unless snb_profile_enable was called).
void snb_prof_begin(snb_context* ctx, int group);
void snb_prof_end(snb_context* ctx, int group);

// ---------------------------------------------------------------------------
// launchers (defined in the .hip files)
// ---------------------------------------------------------------------------
#define SNB_CHECK(x)                              \
    do                                            \
    {                                             \
        hipError_t e_ = (x);                      \
        if (e_ != hipSuccess) return SNB_ERR_HIP; \
    } while (0)

// Makes `dev` current for the lifetime of the guard.
struct DeviceGuard
{
    int prev = -1;
    explicit DeviceGuard(int dev)
    {
        if (hipGetDevice(&prev) == hipSuccess && prev != dev) (void)hipSetDevice(dev);
    }
    ~DeviceGuard()
    {
        int cur;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};

// ---------------------------------------------------------------------------
// Launch recording (snb_batch.hip).  While a context's steps are recorded,
// the launchers below store each kernel launch instead of making it: `solo`
// makes it exactly as the single-context path would, on a given stream, with
// the arguments evaluated at record time.  A kernel that also has a batched
// form carries that form and its argument block, so the same launch of
// several contexts can be issued as one dispatch over a SnbSegTab.
// ---------------------------------------------------------------------------
// Lanes: the reciprocal pipeline (lane 1) may run beside the direct-space
// work (lane 0), as on a context's second stream.  It reads the coordinates
// and charges and writes only its own buffers and accumulator, so it waits
// for the last launch that writes coordinates (`writes_inputs`), and a lane-0
// launch that follows reciprocal launches (`join`: the iteration reads the
// reciprocal forces) waits for them.
struct SnbRecOp
{
    std::function<void(hipStream_t)> solo;
    const void* bfn = nullptr;  // batched kernel, or nullptr
    std::vector<char> args;     // its argument block for this context
    unsigned grid = 0, block = 0;
    size_t lds = 0;
    int lane = 0;
    bool join = false, writes_inputs = false;
    double weight = 0.0;  // work per workgroup (a batched dispatch starts the heaviest first)
};
struct SnbRecorder
{
    std::vector<SnbRecOp> ops;
    int lane = 0;                // lane of the launches being recorded
    bool writes_inputs = false;  // they write coordinates
    int prev_lane = 0;
    double weight = 0.0;         // the context's atom count
    double weight_next = -1.0;   // set by a launcher for its next launch
};
// the recorder of the step being recorded on this thread, else nullptr
extern thread_local SnbRecorder* snb_rec;

static inline void snb_rec_push(SnbRecOp&& op)
{
    op.lane = snb_rec->lane;
    op.join = op.lane == 0 && snb_rec->prev_lane == 1;
    op.writes_inputs = snb_rec->writes_inputs;
    op.weight = snb_rec->weight_next >= 0.0 ? snb_rec->weight_next : snb_rec->weight;
    snb_rec->weight_next = -1.0;
    snb_rec->prev_lane = op.lane;
    snb_rec->ops.push_back(std::move(op));
}

// Sets the recorder's lane / coordinate-writer mark for the scope.
struct SnbRecScope
{
    int lane0 = 0;
    bool w0 = false;
    SnbRecScope(int lane, bool writes_inputs)
    {
        if (!snb_rec) return;
        lane0 = snb_rec->lane;
        w0 = snb_rec->writes_inputs;
        snb_rec->lane = lane;
        snb_rec->writes_inputs = writes_inputs;
    }
    ~SnbRecScope()
    {
        if (!snb_rec) return;
        snb_rec->lane = lane0;
        snb_rec->writes_inputs = w0;
    }
};

// Argument table of a batched launch over nseg contexts: workgroup b belongs
// to the context s with blk0[s] <= b < blk0[s + 1] and is its workgroup
// b - blk0[s] of blk0[s + 1] - blk0[s]; that context's argument block is
// args + s * stride.
struct SnbSegTab
{
    const char* args;
    const int* blk0;
    int nseg, stride;
};

template <class A>
__device__ __forceinline__ const A& snb_seg(const SnbSegTab& t, unsigned& bid, unsigned& nbid)
{
    const int b = (int)blockIdx.x;
    int s = 0;
    while (s + 1 < t.nseg && t.blk0[s + 1] <= b) s++;
    bid = (unsigned)(b - t.blk0[s]);
    nbid = (unsigned)(t.blk0[s + 1] - t.blk0[s]);
    return *(const A*)(t.args + (size_t)s * t.stride);
}

// A launch of a kernel without a batched form.
#define SNB_LAUNCH(K, G, B, LDS, S, ...)                                                  \
    do                                                                                    \
    {                                                                                     \
        if (snb_rec)                                                                      \
        {                                                                                 \
            const dim3 g_ = (G), b_ = (B);                                                \
            const size_t l_ = (LDS);                                                      \
            auto t_ = std::make_tuple(__VA_ARGS__);                                       \
            SnbRecOp op_;                                                                 \
            op_.solo = [=](hipStream_t s_) {                                              \
                std::apply([&](auto... a_) { hipLaunchKernelGGL(K, g_, b_, l_, s_, a_...); }, \
                           t_);                                                           \
            };                                                                            \
            snb_rec_push(std::move(op_));                                                 \
        }                                                                                 \
        else                                                                              \
            hipLaunchKernelGGL(K, G, B, LDS, S, __VA_ARGS__);                             \
    } while (0)

// A launch of kernel K whose batched form KB takes a SnbSegTab of argument
// blocks A, aggregate-initialised from the same arguments (1-D grids only).
#define SNB_LAUNCH_B(K, KB, A, G, B, LDS, S, ...)                                         \
    do                                                                                    \
    {                                                                                     \
        if (snb_rec)                                                                      \
        {                                                                                 \
            const dim3 g_ = (G), b_ = (B);                                                \
            const size_t l_ = (LDS);                                                      \
            auto t_ = std::make_tuple(__VA_ARGS__);                                       \
            const A a_{__VA_ARGS__};                                                      \
            SnbRecOp op_;                                                                 \
            op_.solo = [=](hipStream_t s_) {                                              \
                std::apply([&](auto... a_) { hipLaunchKernelGGL(K, g_, b_, l_, s_, a_...); }, \
                           t_);                                                           \
            };                                                                            \
            if (g_.y == 1 && g_.z == 1 && b_.y == 1 && b_.z == 1)                         \
            {                                                                             \
                op_.bfn = (const void*)(KB);                                              \
                op_.args.assign((const char*)&a_, (const char*)&a_ + sizeof(A));          \
                op_.grid = g_.x;                                                          \
                op_.block = b_.x;                                                         \
                op_.lds = l_;                                                             \
            }                                                                             \
            snb_rec_push(std::move(op_));                                                 \
        }                                                                                 \
        else                                                                              \
            hipLaunchKernelGGL(K, G, B, LDS, S, __VA_ARGS__);                             \
    } while (0)


static inline SnbBox snb_box(snb_ltmatrix3 cell, snb_ltmatrix3 rcell)
{
    SnbBox b;
    b.c[0] = cell.a11; b.c[1] = cell.a21; b.c[2] = cell.a22;
    b.c[3] = cell.a31; b.c[4] = cell.a32; b.c[5] = cell.a33;
    b.rc[0] = rcell.a11; b.rc[1] = rcell.a21; b.rc[2] = rcell.a22;
    b.rc[3] = rcell.a31; b.rc[4] = rcell.a32; b.rc[5] = rcell.a33;
    return b;
}

// common (snb_common.hip)
int snb_launch_pack_atoms(snb_context* ctx, const snb_vector* crd, const float* charge);
int snb_launch_fill_int(hipStream_t s, int* p, int n, int v);
int snb_launch_fill_i64(hipStream_t s, int64_t* p, size_t n);
int snb_launch_add_i64(hipStream_t s, const int64_t* a, const int64_t* b, int64_t* out, size_t n);
int snb_launch_exclusive_scan(hipStream_t s, const int* in, int* out, int n);
int snb_launch_exclusive_scan_clear(hipStream_t s, int* in, int* out, int n);
int snb_launch_flush_forces(snb_context* ctx, snb_vector* frc);  // joins the PME stream
int snb_launch_flush_atom_energy(snb_context* ctx, float* atom_energy, float* direct_energy,
                                 int which_mask);
int snb_launch_flush_virial(snb_context* ctx, snb_ltmatrix3* atom_virial);
int snb_launch_sum_fixed(snb_context* ctx, const int64_t* arr, int n, int slot);
int snb_launch_charge_sums(snb_context* ctx, const float* charge);
int snb_launch_add_scalar(hipStream_t s, float* dst, float v);
int snb_launch_lj_long_range(hipStream_t s, snb_ltmatrix3* vir, float* pot, float vir_add,
                             float pot_add);

// neighbor list (snb_nl.hip)
int snb_launch_nl_update(snb_context* ctx, const snb_vector* crd, const SnbBox& box, int forced,
                         int step, const int* ex_start, const int* ex_list, const int* ex_num);

// windowed path (snb_nlwin.hip, snb_pair.hip)
int snb_launch_nl_window(snb_context* ctx, const SnbBox& box, const int* ex_start,
                         const int* ex_list, const int* ex_num);
int snb_launch_nl_copy_internal(snb_context* ctx, int* counts, int* rows);
int snb_launch_lj_window(snb_context* ctx, const snb_vector* crd, const float* charge,
                         const SnbBox& box, float beta, int need_energy, int need_virial);

// pair kernel (snb_pair.hip)
int snb_launch_lj_direct(snb_context* ctx, const snb_atom_group* nl, const SnbBox& box,
                         float beta, int need_energy, int need_virial);
// k_excl on any stream and accumulators: fx [3][n], e_atom with the
// correction energy in its third slot (e_atom[2n + i]), vir [6][n]
int snb_launch_excl_to(hipStream_t s, int n, const snb_vector* crd, const float* charge,
                       const SnbBox& box, float beta, const int* ex_start, const int* ex_list,
                       const int* ex_num, int need_energy, int need_virial, int64_t* fx,
                       int64_t* e_atom, int64_t* vir);
int snb_launch_excl(snb_context* ctx, const snb_vector* crd, const float* charge,
                    const SnbBox& box, float beta,
                    const int* ex_start, const int* ex_list, const int* ex_num,
                    int need_energy, int need_virial, int64_t* fx);
int snb_launch_nb14(snb_context* ctx, const snb_vector* crd, const float* charge,
                    const SnbBox& box, int need_energy, int need_virial, int64_t* fx);
int snb_launch_bond(snb_context* ctx, const snb_vector* crd, const SnbBox& box, int need_energy,
                    int need_virial, int64_t* fx);
int snb_launch_leap_frog(snb_context* ctx, const float* inv_mass, const snb_vector* frc,
                         snb_vector* vel, snb_vector* crd, float dt, float vmax);
int snb_launch_langevin(snb_context* ctx, const snb_vector* frc, snb_vector* vel,
                        snb_vector* crd);
int snb_launch_settle_remember(snb_context* ctx, const snb_vector* crd, const SnbBox& box);
// Fused remember + middle Langevin + SETTLE with the forces taken from the
// fixed-point accumulators (cleared).
int snb_launch_md_fused(snb_context* ctx, const float* mass, snb_vector* crd, const SnbBox& box,
                        snb_vector* vel, int need_pressure, snb_ltmatrix3* stress);
// Pressure assembly: pot adds the atom-virial sum (into virial_tensor, when
// given, and stress), kin the m v (x) v sum (into stress).
int snb_launch_stress(snb_context* ctx, int n, int pot, int kin, const snb_ltmatrix3* atom_virial,
                      const snb_vector* vel, const float* mass, float inv_volume,
                      snb_ltmatrix3* virial_tensor, snb_ltmatrix3* stress);
int snb_launch_settle(snb_context* ctx, const float* mass, snb_vector* crd, const SnbBox& box,
                      snb_vector* vel, int need_pressure, snb_ltmatrix3* stress, float dt,
                      float v_factor, float x_factor);

// PME (snb_pme.hip)
int snb_pme_setup_tables(snb_context* ctx, const SnbBox& box);
// (Re)computes the influence function for the box on the context stream and
// records its rcell (up_box_bc, PM_force.cpp:1168-1237).
int snb_pme_compute_bc(snb_context* ctx, const SnbBox& box);
int snb_launch_pme_recip(snb_context* ctx, const snb_vector* crd, const float* charge,
                         const SnbBox& box, int need_energy, int need_virial, int with_force);
// self + neutralising energy alone from the current charges (no grid work)
int snb_launch_pme_self(snb_context* ctx, const float* charge);
int snb_fft_forward_full(snb_context* ctx, const float* real, float2* cplx);
int snb_launch_pme_energy_to_potential(snb_context* ctx, float* potential);
int snb_launch_pme_virial_to(snb_context* ctx, snb_ltmatrix3* vir);
int snb_fft_make_plan(int n, SnbFftPlan* plan);
int snb_launch_fft_r2c(snb_context* ctx, const float* real, float2* cplx);
int snb_launch_fft_c2r(snb_context* ctx, const float2* cplx, float* real);
