// snb_window.h -- geometry of the windowed pair path.
//
// Atoms are kept in cell order (cells numbered (x * Ny + y) * Nz + z, atoms of
// a cell ordered by id), so a run of z-cells of one (x, y) column is one
// contiguous range of sorted indices.  A workgroup owns a block of a x b
// columns and S z-cells; every pair is stored once, in the row of the atom
// whose cell precedes the other's by the forward rule on the periodic cell
// displacement (dx, dy, dz):
//     dx > 0, or dx == 0 and dy > 0, or dx == dy == 0 and dz > 0,
//     or the same cell and a larger sorted index,
// with |dx|, |dy|, |dz| <= 2 (unique because every axis has >= 5 cells).
// The j atoms of a block's rows then lie in its window: columns
// x0 + u (u = 0 .. a+1) by y0 - 2 + v (v = 0 .. b+3), z-cells z0 - 2 ..
// z0 + S + 1, each column-run split into at most two ranges at the periodic
// wrap.  A window slot numbers the window's atoms range by range; neighbor
// list entries are 16-bit slots.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define SNB_WIN_HD __host__ __device__ __forceinline__
#else
#define SNB_WIN_HD static inline
#endif

#define SNB_WIN_MAX_RANGES 128
// LDS of the pair kernel besides the per-slot arrays (28 bytes per slot: three
// int64 accumulators and the slot -> atom map): the two range tables and the
// 256-piece Coulomb table (16 bytes a piece)
#define SNB_WIN_FIXED_LDS (4 * (2 * SNB_WIN_MAX_RANGES + 1) + 16 * 256)

struct SnbWin
{
    int nx, ny, nz;     // cell grid
    int a, b, s;        // block: columns a x b, z-cells s
    int nbx, nby, nbz;  // blocks per axis
    int nblock;
    int wu, wv, wz;     // window extents: a+2, b+4, s+4
    int nrange;         // 2 * wu * wv
    int wcap;           // LDS slot capacity of one window
};

// block coordinates of workgroup w
SNB_WIN_HD void snb_win_block(const SnbWin& W, int w, int& x0, int& y0, int& z0)
{
    int bz = w % W.nbz, by = (w / W.nbz) % W.nby, bx = w / (W.nbz * W.nby);
    x0 = bx * W.a;
    y0 = by * W.b;
    z0 = bz * W.s;
}

// block of the cell (x, y, z)
SNB_WIN_HD int snb_win_block_of(const SnbWin& W, int x, int y, int z)
{
    return ((x / W.a) * W.nby + (y / W.b)) * W.nbz + (z / W.s);
}

SNB_WIN_HD int snb_wrap(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

// A row of the internal list keeps its near pairs from the front of its
// max_nb slots and its far (skin) pairs from the back: entry k of a row with
// `split` near pairs.
SNB_WIN_HD int snb_row_entry(int k, int split, int max_nb)
{
    return k < split ? k : max_nb - 1 - (k - split);
}

// Range k = (u * wv + v) * 2 + part of the window of block (x0, y0, z0):
// first cell index and number of cells.  Part 0 starts at the window's first
// z-cell; part 1 is the continuation after the periodic wrap (may be empty).
SNB_WIN_HD void snb_win_range_cells(const SnbWin& W, int x0, int y0, int z0, int k,
                                    int& first_cell, int& ncells)
{
    int part = k & 1, col = k >> 1;
    int u = col / W.wv, v = col - u * W.wv;
    int cx = snb_wrap(x0 + u, W.nx);
    int cy = snb_wrap(y0 - 2 + v, W.ny);
    int zs = z0 - 2, ze = z0 + W.s + 1;  // inclusive, unwrapped
    int base = (cx * W.ny + cy) * W.nz;
    if (zs < 0)
    {
        if (part == 0)
        {
            first_cell = base + zs + W.nz;
            ncells = -zs;
        }
        else
        {
            first_cell = base;
            ncells = ze + 1;
        }
    }
    else if (ze >= W.nz)
    {
        if (part == 0)
        {
            first_cell = base + zs;
            ncells = W.nz - zs;
        }
        else
        {
            first_cell = base;
            ncells = ze - W.nz + 1;
        }
    }
    else
    {
        first_cell = base + zs;
        ncells = part == 0 ? ze - zs + 1 : 0;
    }
}

// Range and position within the window z-run of a cell at window coordinates
// (u, v) and unwrapped window z offset zz (0 .. s+3).
SNB_WIN_HD int snb_win_range_of(const SnbWin& W, int z0, int u, int v, int zz)
{
    int zu = z0 - 2 + zz;  // unwrapped z
    int zs = z0 - 2;
    int part = (zs < 0) ? (zu < 0 ? 0 : 1) : (zu >= W.nz ? 1 : 0);
    return ((u * W.wv) + v) * 2 + part;
}
