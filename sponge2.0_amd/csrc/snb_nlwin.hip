// snb_nlwin.hip -- neighbor-list build for the windowed pair path.
//
// After the cell sort (snb_nl.hip), one wavefront per sorted atom s builds its
// row: the candidates of its forward half-stencil (13 columns, runs of up to
// five z-cells, see snb_window.h) are tested 64 at a time and the survivors
// are stored as 16-bit slots of the owning block's window.  The same pair set
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

// One wavefront (one 64-thread workgroup) per sorted atom s.
__global__ void __launch_bounds__(64) k_nl_build_win(
    const int* __restrict__ need, int n, SnbWin W, const int* __restrict__ start,
    const int* __restrict__ sorted_cell, const float4* __restrict__ sorted_crd, SnbBox box,
    float cs2, const int* __restrict__ ex_start, const int* __restrict__ ex_list,
    const int* __restrict__ ex_num, int max_nb, uint16_t* __restrict__ nl16,
    int* __restrict__ row_count, int* __restrict__ status)
{
    if (need[0] == 0) return;
    __shared__ int first[SNB_WIN_MAX_RANGES];
    __shared__ int pref[SNB_WIN_MAX_RANGES + 1];
    const int s = blockIdx.x, lane = threadIdx.x;
    int cx, cy, cz, x0, y0, z0;
    win_block_of_cell(W, sorted_cell[s], cx, cy, cz, x0, y0, z0);
    win_table_wave(W, start, x0, y0, z0, lane, first, pref);
    __syncthreads();
    if (pref[SNB_WIN_MAX_RANGES] > 65535)
    {
        if (lane == 0)
        {
            atomicOr(status, SNB_FLAG_NL_OVERFLOW);
            row_count[s] = 0;
        }
        return;
    }
    const float4 qi = sorted_crd[s];
    const int id_i = __float_as_int(qi.w);
    uint16_t* row = nl16 + (size_t)s * max_nb;
    const uint64_t lower = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    int count = 0;
    // forward columns: (0,0), (0,1), (0,2), (1,-2..2), (2,-2..2)
    for (int col = 0; col < 13; col++)
    {
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
        int u = (cx - x0) + dx, v = (cy - y0) + dy + 2;
        int ccx = snb_wrap(cx + dx, W.nx), ccy = snb_wrap(cy + dy, W.ny);
        int base = (ccx * W.ny + ccy) * W.nz;
        int zlo = cz + (col == 0 ? 0 : -2), zhi = cz + 2;
        // split the run [zlo, zhi] (unwrapped) at the periodic wrap
        int za[2], zb[2], zu[2], nrun;
        if (zlo < 0)
        {
            za[0] = zlo + W.nz; zb[0] = W.nz - 1; zu[0] = zlo;
            za[1] = 0; zb[1] = zhi; zu[1] = 0;
            nrun = 2;
        }
        else if (zhi >= W.nz)
        {
            za[0] = zlo; zb[0] = W.nz - 1; zu[0] = zlo;
            za[1] = 0; zb[1] = zhi - W.nz; zu[1] = W.nz;
            nrun = 2;
        }
        else
        {
            za[0] = zlo; zb[0] = zhi; zu[0] = zlo;
            nrun = 1;
        }
        for (int r = 0; r < nrun; r++)
        {
            int k = snb_win_range_of(W, z0, u, v, zu[r] - z0 + 2);
            int kfirst = first[k], kpref = pref[k];
            int p0 = start[base + za[r]], p1 = start[base + zb[r] + 1];
            if (col == 0 && r == 0) p0 = s + 1;  // same cell: larger sorted index only
            for (int pb = p0; pb < p1; pb += 64)
            {
                int p = pb + lane;
                bool ok = false;
                if (p < p1)
                {
                    float4 qj = sorted_crd[p];
                    // displacement from the higher id to the lower, as the
                    // reference's row atom sees it (same rounding at the wrap)
                    int id_j = __float_as_int(qj.w);
                    float4 ql = id_i < id_j ? qi : qj, qh = id_i < id_j ? qj : qi;
                    float ddx = ql.x - qh.x, ddy = ql.y - qh.y, ddz = ql.z - qh.z;
                    float sx = (ddx * box.rc[0] + ddy * box.rc[1]) + ddz * box.rc[3];
                    float sy = ddy * box.rc[2] + ddz * box.rc[4];
                    float sz = ddz * box.rc[5];
                    float kx = floorf(sx + 0.5f), ky = floorf(sy + 0.5f), kz = floorf(sz + 0.5f);
                    ddx = ddx - ((kx * box.c[0] + ky * box.c[1]) + kz * box.c[3]);
                    ddy = ddy - (ky * box.c[2] + kz * box.c[4]);
                    ddz = ddz - kz * box.c[5];
                    ok = (ddx * ddx + ddy * ddy) + ddz * ddz < cs2;
                    if (ok) ok = !excluded_pair(id_i, id_j, ex_start, ex_list, ex_num);
                }
                uint64_t mask = __ballot(ok);
                if (ok)
                {
                    int pos = count + __popcll(mask & lower);
                    if (pos < max_nb) row[pos] = (uint16_t)(kpref + (p - kfirst));
                }
                count += __popcll(mask);
            }
        }
    }
    if (lane == 0)
    {
        row_count[s] = count < max_nb ? count : max_nb;
        if (count > max_nb) atomicOr(status, SNB_FLAG_NL_OVERFLOW);
    }
}

// Number of window atoms of every block.
__global__ void k_win_sizes(const int* __restrict__ need, SnbWin W, const int* __restrict__ start,
                            int* __restrict__ win_size)
{
    if (need[0] == 0) return;
    int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= W.nblock) return;
    int x0, y0, z0;
    snb_win_block(W, w, x0, y0, z0);
    int tot = 0;
    for (int k = 0; k < W.nrange; k++)
    {
        int fc, nc;
        snb_win_range_cells(W, x0, y0, z0, k, fc, nc);
        if (nc > 0) tot += start[fc + nc] - start[fc];
    }
    win_size[w] = tot;
}

// For every window slot of every block, record its stage position in the
// list of the atom it holds (the per-step gather sums these).
__global__ void __launch_bounds__(64) k_stage_inverse(const int* __restrict__ need, SnbWin W,
                                                      const int* __restrict__ start,
                                                      const int* __restrict__ stage_base,
                                                      int* __restrict__ inv_count,
                                                      int* __restrict__ inv_list, int invmax,
                                                      int* __restrict__ status)
{
    if (need[0] == 0) return;
    __shared__ int first[SNB_WIN_MAX_RANGES];
    __shared__ int pref[SNB_WIN_MAX_RANGES + 1];
    int w = blockIdx.x, lane = threadIdx.x;
    int x0, y0, z0;
    snb_win_block(W, w, x0, y0, z0);
    win_table_wave(W, start, x0, y0, z0, lane, first, pref);
    __syncthreads();
    int base = stage_base[w];
    for (int k = 0; k < W.nrange; k++)
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

// Reference-layout rows: each internal pair (s, s_j) goes to row min(id) as
// max(id) (atomic append), rows are sorted afterwards.
__global__ void __launch_bounds__(64) k_nl_export(const int* __restrict__ need, int n, SnbWin W,
                                                  const int* __restrict__ start,
                                                  const int* __restrict__ sorted_cell,
                                                  const int* __restrict__ order,
                                                  const int* __restrict__ row_count,
                                                  const uint16_t* __restrict__ nl16, int max_nb,
                                                  int* __restrict__ pub_count,
                                                  int* __restrict__ serial)
{
    if (need[0] == 0) return;
    __shared__ int first[SNB_WIN_MAX_RANGES];
    __shared__ int pref[SNB_WIN_MAX_RANGES + 1];
    int s = blockIdx.x, lane = threadIdx.x;
    int cx, cy, cz, x0, y0, z0;
    win_block_of_cell(W, sorted_cell[s], cx, cy, cz, x0, y0, z0);
    win_table_wave(W, start, x0, y0, z0, lane, first, pref);
    __syncthreads();
    int id_i = order[s];
    int cnt = row_count[s];
    for (int k = lane; k < cnt; k += 64)
    {
        int sj = win_slot_to_sorted(W, first, pref, nl16[(size_t)s * max_nb + k]);
        int id_j = order[sj];
        int lo = id_i < id_j ? id_i : id_j, hi = id_i < id_j ? id_j : id_i;
        int pos = atomicAdd(&pub_count[lo], 1);
        if (pos < max_nb) serial[(size_t)lo * max_nb + pos] = hi;
    }
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

// Test hook: the internal rows with atom ids, indexed by the owner's id.
__global__ void __launch_bounds__(64) k_nl_copy_internal(int n, SnbWin W,
                                                         const int* __restrict__ start,
                                                         const int* __restrict__ sorted_cell,
                                                         const int* __restrict__ order,
                                                         const int* __restrict__ row_count,
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
    int cnt = row_count[s];
    if (lane == 0) counts[id_i] = cnt;
    for (int k = lane; k < cnt; k += 64)
        rows[(size_t)id_i * max_nb + k] =
            order[win_slot_to_sorted(W, first, pref, nl16[(size_t)s * max_nb + k])];
}

// Clears p[0 .. n) on rebuild steps only.
__global__ void k_fill_need(const int* __restrict__ need, int* __restrict__ p, int n)
{
    if (need[0] == 0) return;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void k_win_types(const int* __restrict__ need, int n, const int* __restrict__ order,
                            const int* __restrict__ type, int* __restrict__ stype)
{
    if (need[0] == 0) return;
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) stype[s] = type[order[s]];
}

// Called after the cell sort of a rebuild (same device flag).
int snb_launch_nl_window(snb_context* ctx, const SnbBox& box, const int* ex_start,
                         const int* ex_list, const int* ex_num)
{
    SnbNl& L = ctx->nl;
    hipStream_t s = ctx->stream;
    int n = ctx->n;
    const SnbWin& W = L.win;
    hipLaunchKernelGGL(k_win_types, dim3((n + 255) / 256), dim3(256), 0, s, L.need, n, L.order,
                       ctx->type, ctx->stype);
    hipLaunchKernelGGL(k_nl_build_win, dim3(n), dim3(64), 0, s, L.need, n, W, L.cell_start,
                       L.sorted_cell, L.sorted_crd, box, L.cs2, ex_start, ex_list, ex_num,
                       L.max_nb, L.nl16, L.row_count, ctx->status);
    hipLaunchKernelGGL(k_win_sizes, dim3((W.nblock + 255) / 256), dim3(256), 0, s, L.need, W,
                       L.cell_start, L.win_size);
    SNB_CHECK(hipGetLastError());
    int rc = snb_launch_exclusive_scan(s, L.win_size, L.stage_base, W.nblock);
    if (rc) return rc;
    // the inverse lists persist between rebuilds: clear them on rebuilds only
    hipLaunchKernelGGL(k_fill_need, dim3((n + 255) / 256), dim3(256), 0, s, L.need, L.inv_count,
                       n);
    hipLaunchKernelGGL(k_stage_inverse, dim3(W.nblock), dim3(64), 0, s, L.need, W, L.cell_start,
                       L.stage_base, L.inv_count, L.inv_list, L.invmax, ctx->status);
    SNB_CHECK(hipGetLastError());
    if (L.export_public)
    {
        hipLaunchKernelGGL(k_fill_need, dim3((n + 255) / 256), dim3(256), 0, s, L.need,
                           L.pub_count, n);
        hipLaunchKernelGGL(k_nl_export, dim3(n), dim3(64), 0, s, L.need, n, W, L.cell_start,
                           L.sorted_cell, L.order, L.row_count, L.nl16, L.max_nb, L.pub_count,
                           L.serial);
        hipLaunchKernelGGL(k_nl_pub_finish, dim3(n), dim3(64), 0, s, L.need, n, L.pub_count,
                           L.max_nb, L.serial, L.dnl, ctx->status);
        SNB_CHECK(hipGetLastError());
    }
    return SNB_OK;
}

int snb_launch_nl_copy_internal(snb_context* ctx, int* counts, int* rows)
{
    SnbNl& L = ctx->nl;
    if (ctx->n <= 0) return SNB_OK;
    hipLaunchKernelGGL(k_nl_copy_internal, dim3(ctx->n), dim3(64), 0, ctx->stream, ctx->n, L.win,
                       L.cell_start, L.sorted_cell, L.order, L.row_count, L.nl16, L.max_nb, counts,
                       rows);
    SNB_CHECK(hipGetLastError());
    return SNB_OK;
}
