// snb_nlwin.hip -- neighbor-list build for the windowed pair path.
//
// After the cell sort (snb_nl.hip), one workgroup per cell builds the rows of
// its atoms: the candidates of the forward half-stencil (13 columns, runs of
// up to five z-cells, see snb_window.h) are staged in LDS, tested 64 at a time
// and the survivors are stored as 16-bit slots of the owning block's window.  The same pair set
// as the reference's half list results -- every unordered pair within
// rc + skin once, minus exclusions decided on the lower atom id as the
// reference does (neighbor_list.cpp:594-665) -- only the row that holds a
// pair differs.  When the reference-layout list is wanted (the public d_nl),
// it is derived from these rows (export) and each row sorted.
#include "snb_internal.h"

// Range table of the window of block (x0, y0, z0) into LDS: first sorted index
// and exclusive slot prefix of each range; one wavefront, two ranges per lane.
// Caller synchronises before reading.
__device__ void win_table_wave(const SnbWin& W, const int* __restrict__ start, int x0, int y0,
                               int z0, int lane, int* first, int* pref)
{
    int sz[2], fs[2];
    for (int h = 0; h < 2; h++)
    {
        int k = lane + 64 * h;
        sz[h] = 0;
        fs[h] = 0;
        if (k < W.nrange)
        {
            int fc, nc;
            snb_win_range_cells(W, x0, y0, z0, k, fc, nc);
            fs[h] = start[fc];
            sz[h] = nc > 0 ? start[fc + nc] - fs[h] : 0;
        }
    }
    // exclusive prefix over k = lane, then k = lane + 64
    int tot0 = sz[0], inc0 = sz[0], inc1 = sz[1];
    for (int off = 1; off < 64; off <<= 1)
    {
        int a = __shfl_up(inc0, off);
        int b = __shfl_up(inc1, off);
        if (lane >= off)
        {
            inc0 += a;
            inc1 += b;
        }
    }
    tot0 = __shfl(inc0, 63);
    first[lane] = fs[0];
    first[lane + 64] = fs[1];
    pref[lane] = inc0 - sz[0];
    pref[lane + 64] = tot0 + inc1 - sz[1];
    if (lane == 63) pref[SNB_WIN_MAX_RANGES] = tot0 + inc1;
}

__device__ __forceinline__ bool excluded_pair(int id_a, int id_b, const int* __restrict__ ex_start,
                                              const int* __restrict__ ex_list,
                                              const int* __restrict__ ex_num)
{
    // the reference deletes hi from row lo when hi is in lo's list, testing
    // membership only between the list's first and last entries
    if (!ex_num) return false;
    int lo = id_a < id_b ? id_a : id_b, hi = id_a < id_b ? id_b : id_a;
    int en = ex_num[lo];
    if (en <= 0) return false;
    int es = ex_start[lo];
    if (hi < ex_list[es] || hi > ex_list[es + en - 1]) return false;
    for (int t = es; t < es + en; t++)
        if (ex_list[t] == hi) return true;
    return false;
}

__device__ __forceinline__ void win_block_of_cell(const SnbWin& W, int c, int& cx, int& cy,
                                                  int& cz, int& x0, int& y0, int& z0)
{
    cz = c % W.nz;
    cy = (c / W.nz) % W.ny;
    cx = c / (W.nz * W.ny);
    x0 = (cx / W.a) * W.a;
    y0 = (cy / W.b) * W.b;
    z0 = (cz / W.s) * W.s;
}

// Candidate capacity of one LDS chunk (the atoms per cell whose row counts
// live in LDS, SNB_NLC_MAXA, are in snb_internal.h).
// 1408 candidates (22 groups, 37 KB of LDS) hold a mean stencil of ~1,360
// atoms at 6 A cells; four workgroups then fit a CU beside the reciprocal
// pipeline's (1792 measured slower per step, though not alone)
#define SNB_NLC_CAP 1408
#define SNB_NLC_MAXRUN 63

// One 512-thread workgroup per cell.  The candidates of the cell's forward
// stencil -- 13 columns of up to 5 cells, one run per cell, at most 63 runs
// -- are staged in LDS once (in chunks of SNB_NLC_CAP) with their window
// slots, and every atom of the cell tests them from LDS: wave w builds the
// rows of atoms w, w + 8, ...  Candidate order is column, z (unwrapped), sorted
// index, so a row lists its entries in that order; the first run is the cell
// itself, where only larger sorted indices count.  For a diagonal box each
// atom first marks the stencil cells whose box lies within the list radius
// of its (wrapped) position, with a margin over float rounding, and skips the
// groups of 64 candidates that lie in unmarked cells only.
//
// The distance is taken as q_i - q_j for the row atom i.  The reference takes
// it from the higher id to the lower; the two agree bit for bit (negation is
// exact) except where floorf(s + 0.5) rounds differently at |s| = k + 1/2,
// i.e. where the minimum image is half a box away in that dimension.  With
// >= 5 cells of >= (rc + skin)/2 per axis that is >= 1.25 (rc + skin), so
// such a pair fails the test either way.
template <bool ORTHO>
__device__ __forceinline__ void k_nl_build_cell_body(
    const int* __restrict__ need, SnbWin W, const int* __restrict__ start,
    const float4* __restrict__ sorted_crd, SnbBox box, float cs2,
    const int* __restrict__ ex_start, const int* __restrict__ ex_list,
    const int* __restrict__ ex_num, const int2* __restrict__ ex_rng, int max_nb, float near2,
    uint16_t* __restrict__ nl16, int* __restrict__ row_count, int* __restrict__ row_split,
    int* __restrict__ status, float grid_length,
    unsigned bid, unsigned nbid)
{
    if (need[0] == 0) return;
    __shared__ int first[SNB_WIN_MAX_RANGES];
    __shared__ int pref[SNB_WIN_MAX_RANGES + 1];
    __shared__ float4 cand[SNB_NLC_CAP];
    __shared__ int2 cex[SNB_NLC_CAP];  // exclusion id window of each candidate
    __shared__ uint16_t cslot[SNB_NLC_CAP];
    __shared__ int run_off[SNB_NLC_MAXRUN + 1];
    __shared__ int run_p0[SNB_NLC_MAXRUN];
    __shared__ int run_slot0[SNB_NLC_MAXRUN];
    __shared__ int run_rel[SNB_NLC_MAXRUN];  // (dx + 2) << 6 | (dy + 2) << 3 | (dz + 2)
    __shared__ int grp_r0[SNB_NLC_CAP / 64], grp_r1[SNB_NLC_CAP / 64];  // runs of each group
    __shared__ int nrun_s;
    __shared__ int rowcnt[SNB_NLC_MAXA];   // near pairs (front of the row)
    __shared__ int rowfar[SNB_NLC_MAXA];   // skin pairs (back of the row)
    const int c = bid;
    const int cb = start[c], m = start[c + 1] - cb;
    if (m == 0) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int cz = c % W.nz, cy = (c / W.nz) % W.ny, cx = c / (W.nz * W.ny);
    const int x0 = (cx / W.a) * W.a, y0 = (cy / W.b) * W.b, z0 = (cz / W.s) * W.s;
    if (wave == 0) win_table_wave(W, start, x0, y0, z0, lane, first, pref);
    for (int a = tid; a < SNB_NLC_MAXA; a += blockDim.x)
    {
        rowcnt[a] = 0;
        rowfar[a] = 0;
    }
    __syncthreads();
    if (pref[SNB_WIN_MAX_RANGES] > 65535)
    {
        // the window does not fit 16-bit slots: no rows, flagged
        for (int a = tid; a < m; a += blockDim.x)
        {
            row_count[cb + a] = 0;
            row_split[cb + a] = 0;
        }
        if (tid == 0) atomicOr(status, SNB_FLAG_NL_OVERFLOW);
        return;
    }
    if (wave == 0)
    {
        // lane r builds run r: the own column holds 3 cells (cz .. cz + 2),
        // the 12 others 5 (cz - 2 .. cz + 2); sizes are then scanned
        int size = 0;
        if (lane < SNB_NLC_MAXRUN)
        {
            int col, zu;
            if (lane < 3)
            {
                col = 0;
                zu = cz + lane;
            }
            else
            {
                col = 1 + (lane - 3) / 5;
                zu = cz - 2 + (lane - 3) % 5;
            }
            int dx, dy;
            if (col < 3)
            {
                dx = 0;
                dy = col;
            }
            else
            {
                dx = 1 + (col - 3) / 5;
                dy = (col - 3) % 5 - 2;
            }
            const int u = (cx - x0) + dx, v = (cy - y0) + dy + 2;
            const int base = (snb_wrap(cx + dx, W.nx) * W.ny + snb_wrap(cy + dy, W.ny)) * W.nz;
            // zu unwrapped: the window range and the order of the cells
            const int zw = snb_wrap(zu, W.nz);
            const int k = snb_win_range_of(W, z0, u, v, zu - z0 + 2);
            const int p0 = start[base + zw], p1 = start[base + zw + 1];
            run_p0[lane] = p0;
            run_slot0[lane] = pref[k] + (p0 - first[k]);
            run_rel[lane] = ((dx + 2) << 6) | ((dy + 2) << 3) | (zu - cz + 2);
            size = p1 - p0;
        }
        int incl = size;
        for (int off = 1; off < 64; off <<= 1)
        {
            const int up = __shfl_up(incl, off);
            if (lane >= off) incl += up;
        }
        if (lane < SNB_NLC_MAXRUN) run_off[lane] = incl - size;
        if (lane == SNB_NLC_MAXRUN - 1)
        {
            run_off[SNB_NLC_MAXRUN] = incl;
            nrun_s = SNB_NLC_MAXRUN;
        }
    }
    __syncthreads();
    const int nrun = nrun_s, total = run_off[nrun];
    const uint64_t lower = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int chunk0 = 0; chunk0 < total; chunk0 += SNB_NLC_CAP)
    {
        const int clen = total - chunk0 < SNB_NLC_CAP ? total - chunk0 : SNB_NLC_CAP;
        for (int t = tid; t < clen; t += blockDim.x)
        {
            // the last run starting at or before g (empty runs share offsets)
            const int g = chunk0 + t;
            int lo = 0, hi = nrun - 1;
            while (lo < hi)
            {
                const int mid = (lo + hi + 1) >> 1;
                if (run_off[mid] <= g)
                    lo = mid;
                else
                    hi = mid - 1;
            }
            const int r = lo;
            int off = g - run_off[r];
            float4 q = sorted_crd[run_p0[r] + off];
            cand[t] = q;
            cex[t] = ex_num ? ex_rng[__float_as_int(q.w)] : make_int2(1, 0);
            cslot[t] = (uint16_t)(run_slot0[r] + off);
            if ((t & 63) == 0) grp_r0[t >> 6] = r;
            if ((t & 63) == 63 || t == clen - 1) grp_r1[t >> 6] = r;
        }
        __syncthreads();
        for (int a = wave; a < m && a < SNB_NLC_MAXA; a += (int)(blockDim.x >> 6))
        {
            const int s = cb + a;
            const float4 qi = sorted_crd[s];
            const int id_i = __float_as_int(qi.w);
            const int2 rng_i = ex_num ? ex_rng[id_i] : make_int2(1, 0);
            uint16_t* row = nl16 + (size_t)s * max_nb;
            int count = rowcnt[a], cfar = rowfar[a];
            // stencil cells within reach of atom i: per axis the gap from its
            // wrapped position (the binning's arithmetic) to the cell, a lower
            // bound of the minimum-image distance (the last cell of an axis
            // is wider, which only lengthens the true gap)
            uint64_t reach = ~0ull;
            if (ORTHO)
            {
                const float xw = qi.x - floorf(qi.x / box.c[0]) * box.c[0];
                const float yw = qi.y - floorf(qi.y / box.c[2]) * box.c[2];
                const float zw = qi.z - floorf(qi.z / box.c[5]) * box.c[5];
                auto gap = [&](int d, int c, float w) -> float {
                    float g = d > 0 ? (float)(c + d) * grid_length - w
                                    : (d < 0 ? w - (float)(c + d + 1) * grid_length : 0.0f);
                    return g > 0.0f ? g : 0.0f;
                };
                bool in = false;
                if (lane < nrun)
                {
                    const int rel = run_rel[lane];
                    const float gx = gap((rel >> 6) - 2, cx, xw);
                    const float gy = gap(((rel >> 3) & 7) - 2, cy, yw);
                    const float gz = gap((rel & 7) - 2, cz, zw);
                    in = (gx * gx + gy * gy) + gz * gz < cs2 + 0.05f;
                }
                reach = __ballot(in);
            }
            for (int t0 = 0; t0 < clen; t0 += 64)
            {
                {
                    const int r0 = grp_r0[t0 >> 6], r1 = grp_r1[t0 >> 6];
                    if (!(reach & (((2ull << (r1 - r0)) - 1ull) << r0))) continue;
                }
                const int t = t0 + lane;
                const int g = chunk0 + t;
                bool ok = false, nearp = false;
                // the cell's own atoms come first (g < m): larger sorted index only
                if (t < clen && (g >= m || g > a))
                {
                    const float4 qj = cand[t];
                    const int id_j = __float_as_int(qj.w);
                    float ddx = qi.x - qj.x, ddy = qi.y - qj.y, ddz = qi.z - qj.z;
                    if (ORTHO)
                    {
                        // rint(s) == floorf(s + 0.5) for every pair within
                        // rc + skin (s lies within (rc + skin)/L < 1/2 of an
                        // integer); they differ only half a box away
                        float kx = rintf(ddx * box.rc[0]);
                        float ky = rintf(ddy * box.rc[2]);
                        float kz = rintf(ddz * box.rc[5]);
                        ddx = ddx - kx * box.c[0];
                        ddy = ddy - ky * box.c[2];
                        ddz = ddz - kz * box.c[5];
                    }
                    else
                    {
                        float sx = (ddx * box.rc[0] + ddy * box.rc[1]) + ddz * box.rc[3];
                        float sy = ddy * box.rc[2] + ddz * box.rc[4];
                        float sz = ddz * box.rc[5];
                        float kx = floorf(sx + 0.5f), ky = floorf(sy + 0.5f),
                              kz = floorf(sz + 0.5f);
                        ddx = ddx - ((kx * box.c[0] + ky * box.c[1]) + kz * box.c[3]);
                        ddy = ddy - (ky * box.c[2] + kz * box.c[4]);
                        ddz = ddz - kz * box.c[5];
                    }
                    const float d2 = (ddx * ddx + ddy * ddy) + ddz * ddz;
                    ok = d2 < cs2;
                    nearp = d2 < near2;
                    // the lower id's list decides; its id window rules out
                    // most pairs without touching the list, and decides alone
                    // for a list of consecutive ids (first stored >= 0)
                    const int2 rl = id_i < id_j ? rng_i : cex[t];
                    const int hi = id_i < id_j ? id_j : id_i;
                    const int lo_id = rl.x >= 0 ? rl.x : ~rl.x;
                    if (ok && hi >= lo_id && hi <= rl.y)
                        ok = rl.x < 0 && !excluded_pair(id_i, id_j, ex_start, ex_list, ex_num);
                }
                // near pairs (within rc + skin/4, the ones that can come into
                // range before the next rebuild) from the front, the rest from
                // the back: the pair kernel meets them in that order and whole
                // wavefronts of far pairs skip the force evaluation
                const uint64_t mnear = __ballot(ok && nearp);
                const uint64_t mfar = __ballot(ok && !nearp);
                if (ok)
                {
                    if (nearp)
                    {
                        int pos = count + __popcll(mnear & lower);
                        if (pos < max_nb) row[pos] = cslot[t];
                    }
                    else
                    {
                        int pos = cfar + __popcll(mfar & lower);
                        if (pos < max_nb) row[max_nb - 1 - pos] = cslot[t];
                    }
                }
                count += __popcll(mnear);
                cfar += __popcll(mfar);
            }
            if (lane == 0)
            {
                rowcnt[a] = count;
                rowfar[a] = cfar;
            }
        }
        __syncthreads();
    }
    for (int a = tid; a < m; a += blockDim.x)
    {
        if (a < SNB_NLC_MAXA)
        {
            // an overflowing row (flagged) keeps max_nb valid slots
            int cn = rowcnt[a], cf = rowfar[a];
            row_count[cb + a] = cn + cf < max_nb ? cn + cf : max_nb;
            row_split[cb + a] = cn < max_nb ? cn : max_nb;
            if (cn + cf > max_nb) atomicOr(status, SNB_FLAG_NL_OVERFLOW);
        }
        else
        {
            row_count[cb + a] = 0;
            row_split[cb + a] = 0;
            atomicOr(status, SNB_FLAG_CELL_OVERFLOW);
        }
    }
}

template <bool ORTHO>
__global__ void __launch_bounds__(512) k_nl_build_cell(
    const int* __restrict__ need, SnbWin W, const int* __restrict__ start,
    const float4* __restrict__ sorted_crd, SnbBox box, float cs2,
    const int* __restrict__ ex_start, const int* __restrict__ ex_list,
    const int* __restrict__ ex_num, const int2* __restrict__ ex_rng, int max_nb, float near2,
    uint16_t* __restrict__ nl16, int* __restrict__ row_count, int* __restrict__ row_split,
    int* __restrict__ status, float grid_length)
{
    k_nl_build_cell_body<ORTHO>(
        need, W, start, sorted_crd, box, cs2, ex_start, ex_list, ex_num, ex_rng, max_nb, near2,
        nl16, row_count, row_split, status, grid_length, blockIdx.x, gridDim.x);
}

struct NlBuildCellArgs
{
    const int* need;
    SnbWin W;
    const int* start;
    const float4* sorted_crd;
    SnbBox box;
    float cs2;
    const int* ex_start;
    const int* ex_list;
    const int* ex_num;
    const int2* ex_rng;
    int max_nb;
    float near2;
    uint16_t* nl16;
    int* row_count;
    int* row_split;
    int* status;
    float grid_length;
};

template <bool ORTHO>
__global__ void __launch_bounds__(512) k_nl_build_cell_batch(SnbSegTab t)
{
    unsigned bid, nbid;
    const NlBuildCellArgs& a = snb_seg<NlBuildCellArgs>(t, bid, nbid);
    k_nl_build_cell_body<ORTHO>(
        a.need, a.W, a.start, a.sorted_crd, a.box, a.cs2, a.ex_start, a.ex_list, a.ex_num, a.ex_rng,
        a.max_nb, a.near2, a.nl16, a.row_count, a.row_split, a.status, a.grid_length, bid, nbid);
}

// Number of window atoms of every block (one wavefront per block).
__device__ __forceinline__ void k_win_sizes_body(const int* __restrict__ need, SnbWin W,
                                                  const int* __restrict__ start,
                                                  int* __restrict__ win_size,
    unsigned bid, unsigned nbid)
{
    if (need[0] == 0) return;
    __shared__ int first[SNB_WIN_MAX_RANGES];
    __shared__ int pref[SNB_WIN_MAX_RANGES + 1];
    const int w = bid, lane = threadIdx.x;
    int x0, y0, z0;
    snb_win_block(W, w, x0, y0, z0);
    win_table_wave(W, start, x0, y0, z0, lane, first, pref);
    __syncthreads();
    if (lane == 0) win_size[w] = pref[SNB_WIN_MAX_RANGES];
}

__global__ void __launch_bounds__(64) k_win_sizes(const int* __restrict__ need, SnbWin W,
                                                  const int* __restrict__ start,
                                                  int* __restrict__ win_size)
{
    k_win_sizes_body(need, W, start, win_size, blockIdx.x, gridDim.x);
}

struct WinSizesArgs
{
    const int* need;
    SnbWin W;
    const int* start;
    int* win_size;
};

__global__ void __launch_bounds__(64) k_win_sizes_batch(SnbSegTab t)
{
    unsigned bid, nbid;
    const WinSizesArgs& a = snb_seg<WinSizesArgs>(t, bid, nbid);
    k_win_sizes_body(a.need, a.W, a.start, a.win_size, bid, nbid);
}

// For every window slot of every block, record its stage position in the
// list of the atom it holds (the per-step gather sums these).
__device__ __forceinline__ void k_stage_inverse_body(const int* __restrict__ need, SnbWin W,
                                                      const int* __restrict__ start,
                                                      const int* __restrict__ stage_base,
                                                      int* __restrict__ inv_count,
                                                      int* __restrict__ inv_list, int invmax,
                                                      int* __restrict__ status,
    unsigned bid, unsigned nbid)
{
    if (need[0] == 0) return;
    __shared__ int first[SNB_WIN_MAX_RANGES];
    __shared__ int pref[SNB_WIN_MAX_RANGES + 1];
    int w = bid, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int x0, y0, z0;
    snb_win_block(W, w, x0, y0, z0);
    if (wave == 0) win_table_wave(W, start, x0, y0, z0, lane, first, pref);
    __syncthreads();
    int base = stage_base[w];
    for (int k = wave; k < W.nrange; k += blockDim.x >> 6)
    {
        // ranges beyond nrange are empty, so pref[k + 1] is defined for k < nrange <= 128
        int cnt = pref[k + 1] - pref[k];
        for (int t = lane; t < cnt; t += 64)
        {
            int s = first[k] + t;
            int idx = atomicAdd(&inv_count[s], 1);
            if (idx < invmax)
                inv_list[(size_t)s * invmax + idx] = base + pref[k] + t;
            else
                atomicOr(status, SNB_FLAG_NL_OVERFLOW);
        }
    }
}

__global__ void __launch_bounds__(256) k_stage_inverse(const int* __restrict__ need, SnbWin W,
                                                      const int* __restrict__ start,
                                                      const int* __restrict__ stage_base,
                                                      int* __restrict__ inv_count,
                                                      int* __restrict__ inv_list, int invmax,
                                                      int* __restrict__ status)
{
    k_stage_inverse_body(
        need, W, start, stage_base, inv_count, inv_list, invmax, status, blockIdx.x, gridDim.x);
}

struct StageInverseArgs
{
    const int* need;
    SnbWin W;
    const int* start;
    const int* stage_base;
    int* inv_count;
    int* inv_list;
    int invmax;
    int* status;
};

__global__ void __launch_bounds__(256) k_stage_inverse_batch(SnbSegTab t)
{
    unsigned bid, nbid;
    const StageInverseArgs& a = snb_seg<StageInverseArgs>(t, bid, nbid);
    k_stage_inverse_body(
        a.need, a.W, a.start, a.stage_base, a.inv_count, a.inv_list, a.invmax, a.status, bid, nbid);
}

// slot -> sorted index through the LDS range table (binary search)
__device__ __forceinline__ int win_slot_to_sorted(const SnbWin& W, const int* first,
                                                  const int* pref, int slot)
{
    int lo = 0, hi = W.nrange - 1;
    while (lo < hi)
    {
        int mid = (lo + hi + 1) >> 1;
        if (pref[mid] <= slot)
            lo = mid;
        else
            hi = mid - 1;
    }
    // skip empty ranges that share the prefix
    return first[lo] + (slot - pref[lo]);
}

// Reference-layout rows, one workgroup per window block: the block's window
// table and, when the window fits, the atom id of every window slot are built
// once in LDS; then every internal row of the block's cells sends each pair
// to row min(id) as max(id).  Pairs whose lower id is the row's own atom are
// appended by the wavefront with one counter add per 64 entries.  The others
// go to rows of other window atoms: a first pass counts them per window slot
// in LDS, each slot reserves one range of its row with a single counter add,
// and a second pass places the entries in that range with LDS counters.  A
// window beyond the map takes one counter add per such entry instead.  The
// order within a row follows the appends (rows are sorted afterwards when
// asked for).
#define SNB_EXPORT_MAP_CAP 8192
__global__ void __launch_bounds__(1024) k_nl_export_block(
    const int* __restrict__ need, SnbWin W, const int* __restrict__ start,
    const int* __restrict__ order, const int* __restrict__ row_count,
    const int* __restrict__ row_split, const uint16_t* __restrict__ nl16, int max_nb,
    int* __restrict__ pub_count, int* __restrict__ serial)
{
    if (need[0] == 0) return;
    __shared__ int first[SNB_WIN_MAX_RANGES];
    __shared__ int pref[SNB_WIN_MAX_RANGES + 1];
    __shared__ int wid[SNB_EXPORT_MAP_CAP];  // atom id of each window slot
    __shared__ int fwd[SNB_EXPORT_MAP_CAP];  // appends to each slot's row, then its next position
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nwave = (int)(blockDim.x >> 6);
    int x0, y0, z0;
    snb_win_block(W, blockIdx.x, x0, y0, z0);
    if (wave == 0) win_table_wave(W, start, x0, y0, z0, lane, first, pref);
    __syncthreads();
    const int total = pref[SNB_WIN_MAX_RANGES];
    const bool use_map = total <= SNB_EXPORT_MAP_CAP;
    if (use_map)
        for (int slot = tid; slot < total; slot += blockDim.x)
        {
            wid[slot] = order[win_slot_to_sorted(W, first, pref, slot)];
            fwd[slot] = 0;
        }
    __syncthreads();
    // the block's own cells (the last block of an axis may be partial); the
    // cells (x0 + u, y0 + v, z0 .. z0 + zs - 1) are consecutive
    const int xa = W.nx - x0 < W.a ? W.nx - x0 : W.a;
    const int yb = W.ny - y0 < W.b ? W.ny - y0 : W.b;
    const int zs = W.nz - z0 < W.s ? W.nz - z0 : W.s;
    auto each_row = [&](auto&& visit) {
        for (int u = 0; u < xa; u++)
            for (int v = 0; v < yb; v++)
            {
                const int c0 = ((x0 + u) * W.ny + (y0 + v)) * W.nz + z0;
                const int sb = start[c0], se = start[c0 + zs];
                for (int srow = sb + wave; srow < se; srow += nwave) visit(srow);
            }
    };
    const uint64_t lower = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    if (use_map)
    {
        each_row([&](int srow) {
            const int id_i = order[srow];
            const int cnt = row_count[srow], split = row_split[srow];
            const uint16_t* row = nl16 + (size_t)srow * max_nb;
            for (int k = lane; k < cnt; k += 64)
            {
                const int slot = row[snb_row_entry(k, split, max_nb)];
                if (wid[slot] < id_i) atomicAdd(&fwd[slot], 1);
            }
        });
        __syncthreads();
        for (int slot = tid; slot < total; slot += blockDim.x)
        {
            const int c = fwd[slot];
            if (c > 0) fwd[slot] = atomicAdd(&pub_count[wid[slot]], c);
        }
        __syncthreads();
        each_row([&](int srow) {
            const int id_i = order[srow];
            const int cnt = row_count[srow], split = row_split[srow];
            const uint16_t* row = nl16 + (size_t)srow * max_nb;
            int* own_row = serial + (size_t)id_i * max_nb;
            for (int k0 = 0; k0 < cnt; k0 += 64)
            {
                const int k = k0 + lane;
                int slot = 0, id_j = -1;
                if (k < cnt)
                {
                    slot = row[snb_row_entry(k, split, max_nb)];
                    id_j = wid[slot];
                }
                const bool own = id_j > id_i;
                const uint64_t m = __ballot(own);
                int base = 0;
                if (lane == 0 && m) base = atomicAdd(&pub_count[id_i], (int)__popcll(m));
                base = __shfl(base, 0);
                if (own)
                {
                    const int pos = base + (int)__popcll(m & lower);
                    if (pos < max_nb) own_row[pos] = id_j;
                }
                else if (id_j >= 0)
                {
                    const int pos = atomicAdd(&fwd[slot], 1);
                    if (pos < max_nb) serial[(size_t)id_j * max_nb + pos] = id_i;
                }
            }
        });
        return;
    }
    each_row([&](int srow) {
        const int id_i = order[srow];
        const int cnt = row_count[srow], split = row_split[srow];
        const uint16_t* row = nl16 + (size_t)srow * max_nb;
        int* own_row = serial + (size_t)id_i * max_nb;
        for (int k0 = 0; k0 < cnt; k0 += 64)
        {
            const int k = k0 + lane;
            int id_j = -1;
            if (k < cnt)
                id_j = order[win_slot_to_sorted(W, first, pref,
                                                row[snb_row_entry(k, split, max_nb)])];
            const bool own = id_j > id_i;
            const uint64_t m = __ballot(own);
            int base = 0;
            if (lane == 0 && m) base = atomicAdd(&pub_count[id_i], (int)__popcll(m));
            base = __shfl(base, 0);
            if (own)
            {
                const int pos = base + (int)__popcll(m & lower);
                if (pos < max_nb) own_row[pos] = id_j;
            }
            else if (id_j >= 0)
            {
                const int pos = atomicAdd(&pub_count[id_j], 1);
                if (pos < max_nb) serial[(size_t)id_j * max_nb + pos] = id_i;
            }
        }
    });
}

// ATOM_GROUP::atom_numbers of every public row (rows in append order).
__global__ void k_nl_pub_counts(const int* __restrict__ need, int n,
                                const int* __restrict__ pub_count, int max_nb,
                                snb_atom_group* __restrict__ nl, int* __restrict__ status)
{
    if (need[0] == 0) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int c = pub_count[i];
    if (c > max_nb)
    {
        atomicOr(status, SNB_FLAG_NL_OVERFLOW);
        c = max_nb;
    }
    nl[i].atom_numbers = c;
    nl[i].ghost_numbers = 0;
}

// Sorts each public row ascending (bitonic, one wavefront per row) and sets
// ATOM_GROUP::atom_numbers.
#define SNB_SORT_CAP 4096
__global__ void __launch_bounds__(64) k_nl_pub_finish(const int* __restrict__ need, int n,
                                                      const int* __restrict__ pub_count,
                                                      int max_nb, int* __restrict__ serial,
                                                      snb_atom_group* __restrict__ nl,
                                                      int* __restrict__ status)
{
    if (need[0] == 0) return;
    __shared__ int v[SNB_SORT_CAP];
    int i = blockIdx.x, lane = threadIdx.x;
    int c = pub_count[i];
    if (c > max_nb)
    {
        if (lane == 0) atomicOr(status, SNB_FLAG_NL_OVERFLOW);
        c = max_nb;
    }
    if (lane == 0)
    {
        nl[i].atom_numbers = c;
        nl[i].ghost_numbers = 0;
    }
    if (c <= 1 || c > SNB_SORT_CAP) return;
    int P = 1;
    while (P < c) P <<= 1;
    int* row = serial + (size_t)i * max_nb;
    for (int t = lane; t < P; t += 64) v[t] = t < c ? row[t] : 0x7fffffff;
    __syncthreads();
    for (int k = 2; k <= P; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1)
        {
            for (int t = lane; t < P; t += 64)
            {
                int p = t ^ j;
                if (p > t)
                {
                    int a = v[t], b = v[p];
                    bool up = (t & k) == 0;
                    if ((a > b) == up)
                    {
                        v[t] = b;
                        v[p] = a;
                    }
                }
            }
            __syncthreads();
        }
    for (int t = lane; t < c; t += 64) row[t] = v[t];
}

// The same sort with one wavefront per row and four rows per workgroup, each
// wavefront in its own LDS slab (rows up to SNB_SORT_WAVE_CAP entries); the
// compare-exchange stages need no workgroup barrier because a wavefront
// executes its LDS operations in order.
#define SNB_SORT_WAVE_CAP 2048
__global__ void __launch_bounds__(256) k_nl_pub_finish_wave(const int* __restrict__ need, int n,
                                                           const int* __restrict__ pub_count,
                                                           int max_nb, int* __restrict__ serial,
                                                           snb_atom_group* __restrict__ nl,
                                                           int* __restrict__ status)
{
    if (need[0] == 0) return;
    __shared__ int slab[4][SNB_SORT_WAVE_CAP];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i = blockIdx.x * 4 + wave;
    if (i >= n) return;
    int c = pub_count[i];
    if (c > max_nb)
    {
        if (lane == 0) atomicOr(status, SNB_FLAG_NL_OVERFLOW);
        c = max_nb;
    }
    if (lane == 0)
    {
        nl[i].atom_numbers = c;
        nl[i].ghost_numbers = 0;
    }
    if (c <= 1) return;
    int* v = slab[wave];
    int P = 1;
    while (P < c) P <<= 1;
    int* row = serial + (size_t)i * max_nb;
    for (int t = lane; t < P; t += 64) v[t] = t < c ? row[t] : 0x7fffffff;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    for (int k = 2; k <= P; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1)
        {
            for (int t = lane; t < P; t += 64)
            {
                const int p = t ^ j;
                if (p > t)
                {
                    const int a = v[t], b = v[p];
                    const bool up = (t & k) == 0;
                    if ((a > b) == up)
                    {
                        v[t] = b;
                        v[p] = a;
                    }
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
        }
    for (int t = lane; t < c; t += 64) row[t] = v[t];
}

// Test hook: the internal rows with atom ids, indexed by the owner's id.
__global__ void __launch_bounds__(64) k_nl_copy_internal(int n, SnbWin W,
                                                         const int* __restrict__ start,
                                                         const int* __restrict__ sorted_cell,
                                                         const int* __restrict__ order,
                                                         const int* __restrict__ row_count,
                                                         const int* __restrict__ row_split,
                                                         const uint16_t* __restrict__ nl16,
                                                         int max_nb, int* __restrict__ counts,
                                                         int* __restrict__ rows)
{
    __shared__ int first[SNB_WIN_MAX_RANGES];
    __shared__ int pref[SNB_WIN_MAX_RANGES + 1];
    int s = blockIdx.x, lane = threadIdx.x;
    int cx, cy, cz, x0, y0, z0;
    win_block_of_cell(W, sorted_cell[s], cx, cy, cz, x0, y0, z0);
    win_table_wave(W, start, x0, y0, z0, lane, first, pref);
    __syncthreads();
    int id_i = order[s];
    int cnt = row_count[s], split = row_split[s];
    if (lane == 0) counts[id_i] = cnt;
    for (int k = lane; k < cnt; k += 64)
        rows[(size_t)id_i * max_nb + k] = order[win_slot_to_sorted(
            W, first, pref, nl16[(size_t)s * max_nb + snb_row_entry(k, split, max_nb)])];
}

// Per atom id: first and last entry of its exclusion list (an empty window
// (1, 0) when it has none), the window the reference tests before searching
// the list (neighbor_list.cpp:594-665).
__device__ __forceinline__ void k_ex_range_body(const int* __restrict__ need, int n, const int* __restrict__ ex_start,
                           const int* __restrict__ ex_list, const int* __restrict__ ex_num,
                           int2* __restrict__ ex_rng,
    unsigned bid, unsigned nbid)
{
    if (need[0] == 0) return;
    int i = bid * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int en = ex_num[i];
    int2 r = make_int2(1, 0);
    if (en > 0)
    {
        int es = ex_start[i];
        r = make_int2(ex_list[es], ex_list[es + en - 1]);
        // a list that is exactly first, first + 1, ..., last holds every id of
        // its window, so the window test alone decides; other lists are
        // marked (first stored as ~first) and searched
        bool run = r.y - r.x == en - 1;
        for (int t = 1; run && t < en; t++) run = ex_list[es + t] == r.x + t;
        if (!run) r.x = ~r.x;
    }
    ex_rng[i] = r;
}

__global__ void k_ex_range(const int* __restrict__ need, int n, const int* __restrict__ ex_start,
                           const int* __restrict__ ex_list, const int* __restrict__ ex_num,
                           int2* __restrict__ ex_rng)
{
    k_ex_range_body(need, n, ex_start, ex_list, ex_num, ex_rng, blockIdx.x, gridDim.x);
}

struct ExRangeArgs
{
    const int* need;
    int n;
    const int* ex_start;
    const int* ex_list;
    const int* ex_num;
    int2* ex_rng;
};

__global__ void k_ex_range_batch(SnbSegTab t)
{
    unsigned bid, nbid;
    const ExRangeArgs& a = snb_seg<ExRangeArgs>(t, bid, nbid);
    k_ex_range_body(a.need, a.n, a.ex_start, a.ex_list, a.ex_num, a.ex_rng, bid, nbid);
}

// List entries owned by every block (one wavefront per block).
__device__ __forceinline__ void k_block_work_body(const int* __restrict__ need, SnbWin W,
                                                   const int* __restrict__ start,
                                                   const int* __restrict__ row_count,
                                                   int* __restrict__ work,
    unsigned bid, unsigned nbid)
{
    if (need[0] == 0) return;
    const int w = bid, lane = threadIdx.x;
    int x0, y0, z0;
    snb_win_block(W, w, x0, y0, z0);
    const int z1 = z0 + W.s < W.nz ? z0 + W.s : W.nz;
    int sum = 0;
    for (int col = 0; col < W.a * W.b; col++)
    {
        int ua = col / W.b, vb = col - ua * W.b;
        if (x0 + ua >= W.nx || y0 + vb >= W.ny) continue;
        int colbase = ((x0 + ua) * W.ny + (y0 + vb)) * W.nz;
        for (int si = start[colbase + z0] + lane; si < start[colbase + z1]; si += 64)
            sum += row_count[si];
    }
    for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off);
    if (lane == 0) work[w] = sum;
}

__global__ void __launch_bounds__(64) k_block_work(const int* __restrict__ need, SnbWin W,
                                                   const int* __restrict__ start,
                                                   const int* __restrict__ row_count,
                                                   int* __restrict__ work)
{
    k_block_work_body(need, W, start, row_count, work, blockIdx.x, gridDim.x);
}

struct BlockWorkArgs
{
    const int* need;
    SnbWin W;
    const int* start;
    const int* row_count;
    int* work;
};

__global__ void __launch_bounds__(64) k_block_work_batch(SnbSegTab t)
{
    unsigned bid, nbid;
    const BlockWorkArgs& a = snb_seg<BlockWorkArgs>(t, bid, nbid);
    k_block_work_body(a.need, a.W, a.start, a.row_count, a.work, bid, nbid);
}

// Launch order of the pair kernel's blocks: descending work (ties by block
// index), one workgroup bitonic sort; identity when there are more blocks
// than SNB_ORDER_CAP.
#define SNB_ORDER_CAP 4096
__device__ __forceinline__ void k_block_order_body(const int* __restrict__ need, int nblock,
                                                      const int* __restrict__ work,
                                                      int* __restrict__ perm,
    unsigned bid, unsigned nbid)
{
    if (need[0] == 0) return;
    __shared__ long long key[SNB_ORDER_CAP];
    if (nblock > SNB_ORDER_CAP)
    {
        for (int t = threadIdx.x; t < nblock; t += blockDim.x) perm[t] = t;
        return;
    }
    int P = 1;
    while (P < nblock) P <<= 1;
    // key: -work in the high word, index in the low word; padding sorts last
    for (int t = threadIdx.x; t < P; t += blockDim.x)
        key[t] = t < nblock ? (-(long long)work[t]) * 4294967296ll + (long long)t
                            : 0x7fffffffffffffffll;
    __syncthreads();
    for (int k = 2; k <= P; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1)
        {
            for (int t = threadIdx.x; t < P; t += blockDim.x)
            {
                int q = t ^ j;
                if (q > t)
                {
                    long long a = key[t], b = key[q];
                    bool up = (t & k) == 0;
                    if ((a > b) == up)
                    {
                        key[t] = b;
                        key[q] = a;
                    }
                }
            }
            __syncthreads();
        }
    for (int t = threadIdx.x; t < nblock; t += blockDim.x)
        perm[t] = (int)(key[t] & 0xffffffffll);  // low word: the block index (>= 0)
}

__global__ void __launch_bounds__(1024) k_block_order(const int* __restrict__ need, int nblock,
                                                      const int* __restrict__ work,
                                                      int* __restrict__ perm)
{
    k_block_order_body(need, nblock, work, perm, blockIdx.x, gridDim.x);
}

struct BlockOrderArgs
{
    const int* need;
    int nblock;
    const int* work;
    int* perm;
};

__global__ void __launch_bounds__(1024) k_block_order_batch(SnbSegTab t)
{
    unsigned bid, nbid;
    const BlockOrderArgs& a = snb_seg<BlockOrderArgs>(t, bid, nbid);
    k_block_order_body(a.need, a.nblock, a.work, a.perm, bid, nbid);
}

// Clears p[0 .. n) on rebuild steps only.
__device__ __forceinline__ void k_fill_need_body(const int* __restrict__ need, int* __restrict__ p, int n,
    unsigned bid, unsigned nbid)
{
    if (need[0] == 0) return;
    int i = bid * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void k_fill_need(const int* __restrict__ need, int* __restrict__ p, int n)
{
    k_fill_need_body(need, p, n, blockIdx.x, gridDim.x);
}

struct FillNeedArgs
{
    const int* need;
    int* p;
    int n;
};

__global__ void k_fill_need_batch(SnbSegTab t)
{
    unsigned bid, nbid;
    const FillNeedArgs& a = snb_seg<FillNeedArgs>(t, bid, nbid);
    k_fill_need_body(a.need, a.p, a.n, bid, nbid);
}

__device__ __forceinline__ void k_win_types_body(const int* __restrict__ need, int n, const int* __restrict__ order,
                            const int* __restrict__ type, int* __restrict__ stype,
    unsigned bid, unsigned nbid)
{
    if (need[0] == 0) return;
    int s = bid * blockDim.x + threadIdx.x;
    if (s < n) stype[s] = type[order[s]];
}

__global__ void k_win_types(const int* __restrict__ need, int n, const int* __restrict__ order,
                            const int* __restrict__ type, int* __restrict__ stype)
{
    k_win_types_body(need, n, order, type, stype, blockIdx.x, gridDim.x);
}

struct WinTypesArgs
{
    const int* need;
    int n;
    const int* order;
    const int* type;
    int* stype;
};

__global__ void k_win_types_batch(SnbSegTab t)
{
    unsigned bid, nbid;
    const WinTypesArgs& a = snb_seg<WinTypesArgs>(t, bid, nbid);
    k_win_types_body(a.need, a.n, a.order, a.type, a.stype, bid, nbid);
}

// Called after the cell sort of a rebuild (same device flag).
int snb_launch_nl_window(snb_context* ctx, const SnbBox& box, const int* ex_start,
                         const int* ex_list, const int* ex_num)
{
    SnbNl& L = ctx->nl;
    hipStream_t s = ctx->stream;
    int n = ctx->n;
    const SnbWin& W = L.win;
    SNB_LAUNCH_B(k_win_types, k_win_types_batch, WinTypesArgs, dim3((n + 255) / 256), dim3(256), 0,
                 s, L.need, n, L.order, ctx->type, ctx->stype);
    const bool ortho = box.c[1] == 0.0f && box.c[3] == 0.0f && box.c[4] == 0.0f &&
                       box.rc[1] == 0.0f && box.rc[3] == 0.0f && box.rc[4] == 0.0f;
    if (ex_num)
        SNB_LAUNCH_B(k_ex_range, k_ex_range_batch, ExRangeArgs, dim3((n + 255) / 256), dim3(256), 0,
                     s, L.need, n, ex_start, ex_list, ex_num, L.ex_rng);
    if (ortho)
        SNB_LAUNCH_B(k_nl_build_cell<true>, k_nl_build_cell_batch<true>, NlBuildCellArgs,
                     dim3(L.ncell), dim3(512), 0, s, L.need, W, L.cell_start, L.sorted_crd, box,
                     L.cs2, ex_start, ex_list, ex_num, L.ex_rng, L.max_nb, L.near2, L.nl16,
                     L.row_count, L.row_split, ctx->status, L.grid_length);
    else
        SNB_LAUNCH_B(k_nl_build_cell<false>, k_nl_build_cell_batch<false>, NlBuildCellArgs,
                     dim3(L.ncell), dim3(512), 0, s, L.need, W, L.cell_start, L.sorted_crd, box,
                     L.cs2, ex_start, ex_list, ex_num, L.ex_rng, L.max_nb, L.near2, L.nl16,
                     L.row_count, L.row_split, ctx->status, L.grid_length);
    SNB_LAUNCH_B(k_win_sizes, k_win_sizes_batch, WinSizesArgs, dim3(W.nblock), dim3(64), 0, s,
                 L.need, W, L.cell_start, L.win_size);
    SNB_LAUNCH_B(k_block_work, k_block_work_batch, BlockWorkArgs, dim3(W.nblock), dim3(64), 0, s,
                 L.need, W, L.cell_start, L.row_count, L.block_work);
    SNB_LAUNCH_B(k_block_order, k_block_order_batch, BlockOrderArgs, dim3(1), dim3(1024), 0, s,
                 L.need, W.nblock, L.block_work, L.block_perm);
    SNB_CHECK(hipGetLastError());
    int rc = snb_launch_exclusive_scan(s, L.win_size, L.stage_base, W.nblock);
    if (rc) return rc;
    // the inverse lists persist between rebuilds: clear them on rebuilds only
    SNB_LAUNCH_B(k_fill_need, k_fill_need_batch, FillNeedArgs, dim3((n + 255) / 256), dim3(256), 0,
                 s, L.need, L.inv_count, n);
    SNB_LAUNCH_B(k_stage_inverse, k_stage_inverse_batch, StageInverseArgs, dim3(W.nblock),
                 dim3(256), 0, s, L.need, W, L.cell_start, L.stage_base, L.inv_count, L.inv_list,
                 L.invmax, ctx->status);
    SNB_CHECK(hipGetLastError());
    if (L.export_mode)
    {
        SNB_LAUNCH_B(k_fill_need, k_fill_need_batch, FillNeedArgs, dim3((n + 255) / 256), dim3(256),
                     0, s, L.need, L.pub_count, n);
        SNB_LAUNCH(k_nl_export_block, dim3(W.nblock), dim3(1024), 0, s, L.need, W, L.cell_start,
                   L.order, L.row_count, L.row_split, L.nl16, L.max_nb, L.pub_count, L.serial);
        if (L.export_mode == 1)
            SNB_LAUNCH(k_nl_pub_counts, dim3((n + 255) / 256), dim3(256), 0, s, L.need, n,
                       L.pub_count, L.max_nb, L.dnl, ctx->status);
        else if (L.max_nb <= SNB_SORT_WAVE_CAP)
            SNB_LAUNCH(k_nl_pub_finish_wave, dim3((n + 3) / 4), dim3(256), 0, s, L.need, n,
                       L.pub_count, L.max_nb, L.serial, L.dnl, ctx->status);
        else
            SNB_LAUNCH(k_nl_pub_finish, dim3(n), dim3(64), 0, s, L.need, n, L.pub_count, L.max_nb,
                       L.serial, L.dnl, ctx->status);
        SNB_CHECK(hipGetLastError());
    }
    return SNB_OK;
}

int snb_launch_nl_copy_internal(snb_context* ctx, int* counts, int* rows)
{
    SnbNl& L = ctx->nl;
    if (ctx->n <= 0) return SNB_OK;
    SNB_LAUNCH(k_nl_copy_internal,
               dim3(ctx->n), dim3(64), 0, ctx->stream, ctx->n, L.win, L.cell_start, L.sorted_cell, L.order, L.row_count, L.row_split, L.nl16, L.max_nb, counts, rows);
    SNB_CHECK(hipGetLastError());
    return SNB_OK;
}
