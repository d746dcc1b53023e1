// Descriptor of one fused-kernel launch over one advection instance, and the entry
// points shared by the single-grid and batched drivers.
#pragma once
#include <hip/hip_runtime.h>

struct mpd_grid_s;

namespace mpd {

// Tile shapes compiled for the fused kernels (x extent, y extent of the column tile).
#define MPD_FUSED_TILES(X) X(32, 8) X(16, 16) X(64, 8) X(16, 8) X(32, 16)

struct FusedArgs {
    int nx, ny, nz;
    long long np;
    int tiles_x, tiles_y, zc, nchunks;
    // F1
    const double* cin;        // c^n (first pass: the donor field is formed on the fly) or
                              // the pass input field
    const double *U, *V, *W;  // previous pass's face velocities
    double* c1out;            // first pass: the donor field (tile cells)
    double *Ax, *Ay, *Az;     // unlimited antidiffusive faces
    // F2
    const double* cn;         // c^n (limiter bounds)
    const double* cp;         // pass input field
    const double* q;          // emission increment or null
    double* cout;             // pass result
    double *Lx, *Ly, *Lz;     // limited faces or null
    // F2, optional (the adjoint's stored forward): limiter data of the tile cells -- the
    // ratios and which neighbourhood element holds the bounds
    // (code = hc | lc << 8, hc / lc = 2*q + field, q over {self, x-, x+, y-, y+, z-, z+},
    // field 0 = c^n, 1 = pass input; ties go to the later element)
    double *bu, *bd;
    int* code;
    // stores of this launch stream past the caches (non-temporal): set when the arrays a
    // step writes and reads back cannot stay in the 256 MB Infinity Cache
    int nt;
};

// Whether the fused kernels of a grid (or batch) of `cells` cells use streaming stores.
// Measured on MI355X: at 256^3 and 512x512x256 they shorten the step by 3-5 %; at 128^3,
// whose intermediates the Infinity Cache holds, they lengthen it by 15 %.
inline bool fused_streaming(size_t cells) { return cells >= (size_t(1) << 22); }

// Tile counts of an instance for a tile shape (zc must be set).
inline void fused_geometry(FusedArgs& a, int tx, int ty)
{
    a.tiles_x = (a.nx + tx - 1) / tx;
    a.tiles_y = (a.ny + ty - 1) / ty;
    a.nchunks = (a.nz + a.zc - 1) / a.zc;
}

inline int fused_items(const FusedArgs& a) { return a.tiles_x * a.tiles_y * a.nchunks; }

// The fused kernels address a plane with 32-bit byte offsets: planes of up to 2^29 cells.
inline bool fused_plane_fits(long long nx, long long ny) { return nx * ny <= (1LL << 29); }

// One F1 (f1 = true; donor = first pass) or F2 launch over a device table of instance
// descriptors; item0[i] = first work item of instance i; grid workgroups stride over
// the nitems items; nt: the streaming-store kernels (the table's nt fields must agree).
int fused_batch_launch_exact(int tx, int ty, bool f1, bool donor, const FusedArgs* table,
                             const int* item0, int ninst, int nitems, int grid, bool nt,
                             hipStream_t s);
int fused_batch_launch_fast(int tx, int ty, bool f1, bool donor, const FusedArgs* table,
                            const int* item0, int ninst, int nitems, int grid, bool nt,
                            hipStream_t s);
// One batched launch of the one-kernel-per-pass kernel (donor: the first pass; emit: some
// instance's descriptor carries an emission increment) over a device table of instance
// descriptors in 32 x 16 tiles (F12_TILE_X x F12_TILE_Y).
constexpr int F12_TILE_X = 32, F12_TILE_Y = 16;
int fused_batch_launch_f12_exact(const FusedArgs* table, const int* item0, int ninst,
                                 int nitems, bool donor, bool emit, bool nt, hipStream_t s);
int fused_batch_launch_f12_fast(const FusedArgs* table, const int* item0, int ninst,
                                int nitems, bool donor, bool emit, bool nt, hipStream_t s);
// z-chunk length of a grid's one-kernel passes: its own setting, else the cost model's.
int pass_zchunk(const mpd_grid_s* g);
// One exact-arithmetic iord-2 step cin -> cout of a PERIODIC grid that also stores the
// forward's intermediates for the adjoint: the donor field c1, the antidiffusive faces
// A[3] and, with the limiter, its data (see FusedArgs; null pointers when unlimited).
int fused_store_step_exact(mpd_grid_s* g, const double* cin, double* cout, double* c1,
                           double* const A[3], double* bu, double* bd, int* code);
// Resident workgroups per CU of the F1 (donor) or F2 kernel of a tile shape (0: unknown
// shape).
int fused_resident_blocks_exact(int tx, int ty, bool f1);
int fused_resident_blocks_fast(int tx, int ty, bool f1);

}  // namespace mpd
