// snb_nl.hip -- cell-list rebuild of the reference's half neighbor list.
//
// Produces the reference's drop-in output (NEIGHBOR_LIST::d_nl: row i of a
// padded int[n][max_neighbor_numbers] table holds every j > i with
// |pbc(r_i - r_j)|^2 < (rc + skin)^2, minus i's exclusions;
// neighbor_list.cpp:279-492, :594-665) with a different construction:
//   1. atoms are counted per cell and scattered into a cell-sorted CSR array
//      (counting sort, no fixed per-cell capacity);
//   2. each cell's slice is ordered by atom id, and the coordinates are
//      packed next to the ids (float4 x, y, z, id) for coalesced reads;
//   3. one wavefront builds one row: it walks the 25 (x, y) stencil columns,
//      each a contiguous run of <= 5 z-cells in the sorted array, tests 64
//      candidates per instruction, and appends survivors with a ballot /
//      popcount prefix -- the wave owns the row, so no atomics, and rows come
//      out in a deterministic order.
// All kernels return early when the device-side rebuild flag is clear, so the
// conditional policy of NEIGHBOR_LIST::Update needs no host round trip.
#include "snb_internal.h"

// Check_Refresh (neighbor_list.cpp:166-180).
__global__ void k_nl_check(int n, const snb_vector* __restrict__ crd,
                           const snb_vector* __restrict__ old, SnbBox box,
                           float permit2, int* __restrict__ need)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    snb_vector a = crd[i], b = old[i];
    float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    float sx = (dx * box.rc[0] + dy * box.rc[1]) + dz * box.rc[3];
    float sy = dy * box.rc[2] + dz * box.rc[4];
    float sz = dz * box.rc[5];
    float kx = floorf(sx + 0.5f), ky = floorf(sy + 0.5f), kz = floorf(sz + 0.5f);
    dx = dx - ((kx * box.c[0] + ky * box.c[1]) + kz * box.c[3]);
    dy = dy - (ky * box.c[2] + kz * box.c[4]);
    dz = dz - kz * box.c[5];
    if ((dx * dx + dy * dy) + dz * dz > permit2) need[0] = 1;
}

__device__ __forceinline__ void k_nl_set_body(int* need, int v,
    unsigned bid, unsigned nbid)
{ need[0] = v; }

__global__ void k_nl_set(int* need, int v)
{
    k_nl_set_body(need, v, blockIdx.x, gridDim.x);
}

struct NlSetArgs
{
    int* need;
    int v;
};

__global__ void k_nl_set_batch(SnbSegTab t)
{
    unsigned bid, nbid;
    const NlSetArgs& a = snb_seg<NlSetArgs>(t, bid, nbid);
    k_nl_set_body(a.need, a.v, bid, nbid);
}

__device__ __forceinline__ void k_nl_clear_counts_body(const int* __restrict__ need, int ncell, int* __restrict__ cnt,
    unsigned bid, unsigned nbid)
{
    if (need[0] == 0) return;
    int i = bid * blockDim.x + threadIdx.x;
    if (i <= ncell) cnt[i] = 0;
}

__global__ void k_nl_clear_counts(const int* __restrict__ need, int ncell, int* __restrict__ cnt)
{
    k_nl_clear_counts_body(need, ncell, cnt, blockIdx.x, gridDim.x);
}

struct NlClearArgs
{
    const int* need;
    int ncell;
    int* cnt;
};

__global__ void k_nl_clear_counts_batch(SnbSegTab t)
{
    unsigned bid, nbid;
    const NlClearArgs& a = snb_seg<NlClearArgs>(t, bid, nbid);
    k_nl_clear_counts_body(a.need, a.ncell, a.cnt, bid, nbid);
}

// Put_Atom_In_Grids (neighbor_list.cpp:206-275): the wrap into the cell and
// the cell index use the reference's float operations; the slot inside the
// cell comes from an atomic counter and is made deterministic later.
__device__ __forceinline__ void k_nl_bin_body(const int* __restrict__ need, int n, const snb_vector* __restrict__ crd,
                         snb_vector* __restrict__ old, int copy_old, SnbBox box,
                         float grid_length, int nx, int ny, int nz, int* __restrict__ atom_cell,
                         int* __restrict__ atom_slot, int* __restrict__ cnt,
    unsigned bid, unsigned nbid)
{
    if (need[0] == 0) return;
    int i = bid * blockDim.x + threadIdx.x;
    if (i >= n) return;
    snb_vector r = crd[i];
    if (copy_old) old[i] = r;
    float k3 = floorf(r.z / box.c[5]);
    r.z -= k3 * box.c[5];
    r.y -= k3 * box.c[4];
    float k2 = floorf(r.y / box.c[2]);
    r.y -= k2 * box.c[2];
    r.x -= k3 * box.c[3] + k2 * box.c[1];
    r.x -= floorf(r.x / box.c[0]) * box.c[0];
    int ix = (int)(r.x / grid_length);
    int iy = (int)(r.y / grid_length);
    int iz = (int)(r.z / grid_length);
    ix = ix < 0 ? 0 : (ix < nx ? ix : nx - 1);
    iy = iy < 0 ? 0 : (iy < ny ? iy : ny - 1);
    iz = iz < 0 ? 0 : (iz < nz ? iz : nz - 1);
    int c = (ix * ny + iy) * nz + iz;
    atom_cell[i] = c;
    atom_slot[i] = atomicAdd(&cnt[c], 1);
}

__global__ void k_nl_bin(const int* __restrict__ need, int n, const snb_vector* __restrict__ crd,
                         snb_vector* __restrict__ old, int copy_old, SnbBox box,
                         float grid_length, int nx, int ny, int nz, int* __restrict__ atom_cell,
                         int* __restrict__ atom_slot, int* __restrict__ cnt)
{
    k_nl_bin_body(
        need, n, crd, old, copy_old, box, grid_length, nx, ny, nz, atom_cell, atom_slot, cnt,
        blockIdx.x, gridDim.x);
}

struct NlBinArgs
{
    const int* need;
    int n;
    const snb_vector* crd;
    snb_vector* old;
    int copy_old;
    SnbBox box;
    float grid_length;
    int nx;
    int ny;
    int nz;
    int* atom_cell;
    int* atom_slot;
    int* cnt;
};

__global__ void k_nl_bin_batch(SnbSegTab t)
{
    unsigned bid, nbid;
    const NlBinArgs& a = snb_seg<NlBinArgs>(t, bid, nbid);
    k_nl_bin_body(
        a.need, a.n, a.crd, a.old, a.copy_old, a.box, a.grid_length, a.nx, a.ny, a.nz, a.atom_cell,
        a.atom_slot, a.cnt, bid, nbid);
}

__device__ __forceinline__ void k_nl_scatter_body(const int* __restrict__ need, int n, const int* __restrict__ atom_cell,
                             const int* __restrict__ atom_slot, const int* __restrict__ start,
                             int* __restrict__ sorted_id,
    unsigned bid, unsigned nbid)
{
    if (need[0] == 0) return;
    int i = bid * blockDim.x + threadIdx.x;
    if (i >= n) return;
    sorted_id[start[atom_cell[i]] + atom_slot[i]] = i;
}

__global__ void k_nl_scatter(const int* __restrict__ need, int n, const int* __restrict__ atom_cell,
                             const int* __restrict__ atom_slot, const int* __restrict__ start,
                             int* __restrict__ sorted_id)
{
    k_nl_scatter_body(need, n, atom_cell, atom_slot, start, sorted_id, blockIdx.x, gridDim.x);
}

struct NlScatterArgs
{
    const int* need;
    int n;
    const int* atom_cell;
    const int* atom_slot;
    const int* start;
    int* sorted_id;
};

__global__ void k_nl_scatter_batch(SnbSegTab t)
{
    unsigned bid, nbid;
    const NlScatterArgs& a = snb_seg<NlScatterArgs>(t, bid, nbid);
    k_nl_scatter_body(a.need, a.n, a.atom_cell, a.atom_slot, a.start, a.sorted_id, bid, nbid);
}

// One wave per cell: rank the cell's atoms by id (O(m^2) with m ~ 20-60) and
// write ids, cell numbers and coordinates in id order.  Also flags cells above
// the reference's per-cell capacity (max_atom_in_grid_numbers) for parity of
// the overflow report; nothing is truncated here.
#define SNB_NL_SORT_MAX 1024
__device__ __forceinline__ void k_nl_cell_sort_body(const int* __restrict__ need, int ncell,
                                                     const int* __restrict__ start,
                                                     const int* __restrict__ sorted_id,
                                                     const snb_vector* __restrict__ crd,
                                                     float4* __restrict__ sorted_crd,
                                                     int* __restrict__ order,
                                                     int* __restrict__ sorted_cell,
                                                     int max_in_cell, int* __restrict__ status,
    unsigned bid, unsigned nbid)
{
    if (need[0] == 0) return;
    __shared__ int ids[SNB_NL_SORT_MAX];
    int c = bid;
    if (c >= ncell) return;
    int b = start[c], m = start[c + 1] - b;
    int lane = threadIdx.x;
    if (m > max_in_cell && lane == 0) atomicOr(status, SNB_FLAG_CELL_OVERFLOW);
    if (m > SNB_NL_SORT_MAX)
    {
        // keep scatter order (still a correct set)
        for (int k = lane; k < m; k += 64)
        {
            int id = sorted_id[b + k];
            snb_vector r = crd[id];
            sorted_crd[b + k] = make_float4(r.x, r.y, r.z, __int_as_float(id));
            order[b + k] = id;
            sorted_cell[b + k] = c;
        }
        return;
    }
    for (int k = lane; k < m; k += 64) ids[k] = sorted_id[b + k];
    __syncthreads();
    for (int k = lane; k < m; k += 64)
    {
        int id = ids[k];
        int rank = 0;
        for (int t = 0; t < m; t++) rank += ids[t] < id;
        snb_vector r = crd[id];
        sorted_crd[b + rank] = make_float4(r.x, r.y, r.z, __int_as_float(id));
        order[b + rank] = id;
        sorted_cell[b + rank] = c;
    }
}

__global__ void __launch_bounds__(64) k_nl_cell_sort(const int* __restrict__ need, int ncell,
                                                     const int* __restrict__ start,
                                                     const int* __restrict__ sorted_id,
                                                     const snb_vector* __restrict__ crd,
                                                     float4* __restrict__ sorted_crd,
                                                     int* __restrict__ order,
                                                     int* __restrict__ sorted_cell,
                                                     int max_in_cell, int* __restrict__ status)
{
    k_nl_cell_sort_body(
        need, ncell, start, sorted_id, crd, sorted_crd, order, sorted_cell, max_in_cell, status,
        blockIdx.x, gridDim.x);
}

struct NlCellSortArgs
{
    const int* need;
    int ncell;
    const int* start;
    const int* sorted_id;
    const snb_vector* crd;
    float4* sorted_crd;
    int* order;
    int* sorted_cell;
    int max_in_cell;
    int* status;
};

__global__ void __launch_bounds__(64) k_nl_cell_sort_batch(SnbSegTab t)
{
    unsigned bid, nbid;
    const NlCellSortArgs& a = snb_seg<NlCellSortArgs>(t, bid, nbid);
    k_nl_cell_sort_body(
        a.need, a.ncell, a.start, a.sorted_id, a.crd, a.sorted_crd, a.order, a.sorted_cell,
        a.max_in_cell, a.status, bid, nbid);
}

// The three axis offset sets of the +-2 stencil: five distinct cells when the
// axis has >= 5 cells, otherwise every cell of the axis once (the reference
// de-duplicates the same way, neighbor_list.cpp:5-49).
__device__ __forceinline__ int stencil_cells(int c, int ncell_axis, int* out)
{
    if (ncell_axis >= 5)
    {
        for (int d = -2; d <= 2; d++)
        {
            int v = c + d;
            v += v < 0 ? ncell_axis : 0;
            v -= v >= ncell_axis ? ncell_axis : 0;
            out[d + 2] = v;
        }
        return 5;
    }
    for (int d = 0; d < ncell_axis; d++) out[d] = d;
    return ncell_axis;
}

// Row builder: one wavefront per atom i (rows in atom order).
__device__ __forceinline__ void k_nl_search_body(
    const int* __restrict__ need, int n, const snb_vector* __restrict__ crd,
    const int* __restrict__ atom_cell, const int* __restrict__ start,
    const float4* __restrict__ sorted_crd, SnbBox box, float cs2, int nx, int ny, int nz,
    const int* __restrict__ ex_start, const int* __restrict__ ex_list,
    const int* __restrict__ ex_num, int max_nb, int* __restrict__ serial,
    snb_atom_group* __restrict__ nl, int* __restrict__ status,
    const int* __restrict__ nbr_cells, const int* __restrict__ nbr_count,
    unsigned bid, unsigned nbid)
{
    if (need[0] == 0) return;
    const int lane = threadIdx.x & 63;
    const int i = bid * 4 + (threadIdx.x >> 6);
    if (i >= n) return;
    snb_vector ri = crd[i];
    int c = atom_cell[i];
    int cz = c % nz, cy = (c / nz) % ny, cx = c / (nz * ny);
    // exclusion window of i (neighbor_list.cpp:611-616)
    int en = ex_num ? ex_num[i] : 0;
    int es = en > 0 ? ex_start[i] : 0;
    int elo = en > 0 ? ex_list[es] : 0;
    int ehi = en > 0 ? ex_list[es + en - 1] : -1;
    int xs[5], ys[5];
    int nxs = stencil_cells(cx, nx, xs);
    int nys = stencil_cells(cy, ny, ys);
    // z runs: [z0, z1] inclusive, at most two after wrapping
    int zr0[2], zr1[2], nzr;
    if (nz >= 5)
    {
        int z0 = cz - 2, z1 = cz + 2;
        if (z0 < 0)
        {
            zr0[0] = 0; zr1[0] = z1; zr0[1] = z0 + nz; zr1[1] = nz - 1; nzr = 2;
        }
        else if (z1 >= nz)
        {
            zr0[0] = z0; zr1[0] = nz - 1; zr0[1] = 0; zr1[1] = z1 - nz; nzr = 2;
        }
        else
        {
            zr0[0] = z0; zr1[0] = z1; nzr = 1;
        }
    }
    else
    {
        zr0[0] = 0; zr1[0] = nz - 1; nzr = 1;
    }
    int* row = serial + (size_t)i * max_nb;
    int count = 0;
    const uint64_t lower = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    // the candidates p0 .. p1 - 1 in cell order: the pairs of i among them
    auto scan = [&](int p0, int p1) {
        for (int base = p0; base < p1; base += 64)
        {
            int p = base + lane;
            bool ok = false;
            int j = 0;
            if (p < p1)
            {
                float4 q = sorted_crd[p];
                j = __float_as_int(q.w);
                if (j > i)
                {
                    float dx = ri.x - q.x, dy = ri.y - q.y, dz = ri.z - q.z;
                    float sx = (dx * box.rc[0] + dy * box.rc[1]) + dz * box.rc[3];
                    float sy = dy * box.rc[2] + dz * box.rc[4];
                    float sz = dz * box.rc[5];
                    float kx = floorf(sx + 0.5f), ky = floorf(sy + 0.5f),
                          kz = floorf(sz + 0.5f);
                    dx = dx - ((kx * box.c[0] + ky * box.c[1]) + kz * box.c[3]);
                    dy = dy - (ky * box.c[2] + kz * box.c[4]);
                    dz = dz - kz * box.c[5];
                    ok = (dx * dx + dy * dy) + dz * dz < cs2;
                    if (ok && j >= elo && j <= ehi)
                    {
                        for (int t = es; t < es + en; t++)
                            if (ex_list[t] == j)
                            {
                                ok = false;
                                break;
                            }
                    }
                }
            }
            uint64_t mask = __ballot(ok);
            if (ok)
            {
                int slot = count + __popcll(mask & lower);
                if (slot < max_nb) row[slot] = j;
            }
            count += __popcll(mask);
        }
    };
    if (nbr_cells)
    {
        // triclinic cell: the cell's neighbour cells from the table
        const int* tab = nbr_cells + (size_t)c * SNB_NL_MAX_NBR;
        const int m = nbr_count[c];
        for (int k = 0; k < m; k++) scan(start[tab[k]], start[tab[k] + 1]);
    }
    else
        for (int a = 0; a < nxs; a++)
            for (int bq = 0; bq < nys; bq++)
            {
                int col = (xs[a] * ny + ys[bq]) * nz;
                for (int zr = 0; zr < nzr; zr++)
                    scan(start[col + zr0[zr]], start[col + zr1[zr] + 1]);
            }
    if (lane == 0)
    {
        nl[i].atom_numbers = count < max_nb ? count : max_nb;
        nl[i].ghost_numbers = 0;
        if (count > max_nb) atomicOr(status, SNB_FLAG_NL_OVERFLOW);
    }
}

__global__ void __launch_bounds__(256) k_nl_search(
    const int* __restrict__ need, int n, const snb_vector* __restrict__ crd,
    const int* __restrict__ atom_cell, const int* __restrict__ start,
    const float4* __restrict__ sorted_crd, SnbBox box, float cs2, int nx, int ny, int nz,
    const int* __restrict__ ex_start, const int* __restrict__ ex_list,
    const int* __restrict__ ex_num, int max_nb, int* __restrict__ serial,
    snb_atom_group* __restrict__ nl, int* __restrict__ status,
    const int* __restrict__ nbr_cells, const int* __restrict__ nbr_count)
{
    k_nl_search_body(
        need, n, crd, atom_cell, start, sorted_crd, box, cs2, nx, ny, nz, ex_start, ex_list, ex_num,
        max_nb, serial, nl, status, nbr_cells, nbr_count, blockIdx.x, gridDim.x);
}

struct NlSearchArgs
{
    const int* need;
    int n;
    const snb_vector* crd;
    const int* atom_cell;
    const int* start;
    const float4* sorted_crd;
    SnbBox box;
    float cs2;
    int nx;
    int ny;
    int nz;
    const int* ex_start;
    const int* ex_list;
    const int* ex_num;
    int max_nb;
    int* serial;
    snb_atom_group* nl;
    int* status;
    const int* nbr_cells;
    const int* nbr_count;
};

__global__ void __launch_bounds__(256) k_nl_search_batch(SnbSegTab t)
{
    unsigned bid, nbid;
    const NlSearchArgs& a = snb_seg<NlSearchArgs>(t, bid, nbid);
    k_nl_search_body(
        a.need, a.n, a.crd, a.atom_cell, a.start, a.sorted_crd, a.box, a.cs2, a.nx, a.ny, a.nz,
        a.ex_start, a.ex_list, a.ex_num, a.max_nb, a.serial, a.nl, a.status, a.nbr_cells,
        a.nbr_count, bid, nbid);
}

int snb_launch_nl_update(snb_context* ctx, const snb_vector* crd, const SnbBox& box, int forced,
                         int step, const int* ex_start, const int* ex_list, const int* ex_num)
{
    SnbNl& L = ctx->nl;
    hipStream_t s = ctx->stream;
    int n = ctx->n;
    snb_prof_begin(ctx, SNB_PROF_NL);
    // NEIGHBOR_LIST::Update policy (neighbor_list.cpp:843-856)
    if (forced)
        SNB_LAUNCH_B(k_nl_set, k_nl_set_batch, NlSetArgs, dim3(1), dim3(1), 0, s, L.need, 1);
    else if (L.refresh_interval <= 0)
    {
        SNB_LAUNCH_B(k_nl_set, k_nl_set_batch, NlSetArgs, dim3(1), dim3(1), 0, s, L.need, 0);
        float permit2 = L.skin * L.skin * L.skin_permit * L.skin_permit;
        SNB_LAUNCH(k_nl_check, dim3((n + 255) / 256), dim3(256), 0, s, n, crd, L.old_crd, box,
                   permit2, L.need);
    }
    else
        SNB_LAUNCH_B(k_nl_set, k_nl_set_batch, NlSetArgs, dim3(1), dim3(1), 0, s, L.need,
                     ((step + 1) % L.refresh_interval == 0) ? 1 : 0);
    SNB_CHECK(hipGetLastError());
    SNB_LAUNCH_B(k_nl_clear_counts, k_nl_clear_counts_batch, NlClearArgs,
                 dim3((L.ncell + 256) / 256), dim3(256), 0, s, L.need, L.ncell, L.cell_count);
    SNB_LAUNCH_B(k_nl_bin, k_nl_bin_batch, NlBinArgs, dim3((n + 255) / 256), dim3(256), 0, s,
                 L.need, n, crd, L.old_crd,
                 L.refresh_interval <= 0 ? 1 : 0, box, L.grid_length, L.nx, L.ny, L.nz, L.atom_cell, L.atom_slot, L.cell_count);
    SNB_CHECK(hipGetLastError());
    int rc = snb_launch_exclusive_scan(s, L.cell_count, L.cell_start, L.ncell);
    if (rc) return rc;
    SNB_LAUNCH_B(k_nl_scatter, k_nl_scatter_batch, NlScatterArgs, dim3((n + 255) / 256), dim3(256),
                 0, s, L.need, n, L.atom_cell, L.atom_slot, L.cell_start, L.sorted_id);
    SNB_LAUNCH_B(k_nl_cell_sort, k_nl_cell_sort_batch, NlCellSortArgs, dim3(L.ncell), dim3(64), 0,
                 s, L.need, L.ncell, L.cell_start, L.sorted_id, crd, L.sorted_crd, L.order,
                 L.sorted_cell, L.max_in_cell, ctx->status);
    if (L.windowed)
    {
        rc = snb_launch_nl_window(ctx, box, ex_start, ex_list, ex_num);
        if (rc) return rc;
    }
    else
        SNB_LAUNCH_B(k_nl_search, k_nl_search_batch, NlSearchArgs, dim3((n + 3) / 4), dim3(256), 0,
                     s, L.need, n, crd, L.atom_cell, L.cell_start, L.sorted_crd, box, L.cs2, L.nx,
                     L.ny, L.nz, ex_start, ex_list, ex_num, L.max_nb, L.serial, L.dnl, ctx->status,
                     L.nbr_cells, L.nbr_count);
    snb_prof_end(ctx, SNB_PROF_NL);
    SNB_CHECK(hipGetLastError());
    return SNB_OK;
}
