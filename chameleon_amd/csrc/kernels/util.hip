/*
 * util.hip -- HBM-bound kernels around the Cholesky path: mirror padding
 * and orientation, the dplgsy/dplrnt generators, tile copies and norms.
 */
#include "../internal.h"
#include "rng.h"

#include <algorithm>

namespace chx {
namespace k {

/* Geometry of a device mirror: mt x nt tiles of ldt x wdt elements, tile
 * (m,n) at ((n*mt + m) * ldt * wdt); the valid part of tile (m,n) is
 * min(mb, M - m mb) x min(nb, N - n nb). */
struct Geo {
    int mt, nt, ldt, wdt, mb, nb, M, N;
    __device__ __forceinline__ int vm(int m) const { int r = M - m * mb; return r < mb ? r : mb; }
    __device__ __forceinline__ int vn(int n) const { int r = N - n * nb; return r < nb ? r : nb; }
    __device__ __forceinline__ size_t tile(int m, int n) const { return ((size_t)n * mt + m) * (size_t)ldt * wdt; }
};

static Geo geo(int mt, int nt, int ldt, int wdt, int mb, int nb, int M, int N)
{
    Geo g; g.mt = mt; g.nt = nt; g.ldt = ldt; g.wdt = wdt; g.mb = mb; g.nb = nb; g.M = M; g.N = N;
    return g;
}

/* ------------------------------------------------------------------ */
/* Padding: zero outside the valid part of every tile, 1 on the padded
 * diagonal of diagonal tiles when diag_identity. Grid: one block per tile. */
template <typename T>
__global__ void fix_padding_kernel(T *d, Geo g, int diag_identity)
{
    const int t = blockIdx.x;
    const int m = t % g.mt, n = t / g.mt;
    T *tile = d + g.tile(m, n);
    const int vm = g.vm(m), vn = g.vn(n);
    const bool diag = (m == n);
    const size_t total = (size_t)g.ldt * g.wdt;
    for (size_t e = threadIdx.x; e < total; e += blockDim.x) {
        const int i = (int)(e % g.ldt), j = (int)(e / g.ldt);
        if (i < vm && j < vn) continue;
        tile[e] = (diag && diag_identity && i == j) ? (T)1 : (T)0;
    }
}

void fix_padding(void *d, int mt, int nt, int ldt, int wdt, int mb, int nb, int M, int N, int eltsize,
                 bool diag_identity, hipStream_t s)
{
    if (mt * nt == 0) return;
    Geo g = geo(mt, nt, ldt, wdt, mb, nb, M, N);
    if (eltsize == 8)
        hipLaunchKernelGGL(fix_padding_kernel<double>, dim3(mt * nt), dim3(256), 0, s, (double *)d, g, (int)diag_identity);
    else
        hipLaunchKernelGGL(fix_padding_kernel<float>, dim3(mt * nt), dim3(256), 0, s, (float *)d, g, (int)diag_identity);
}

/* ------------------------------------------------------------------ */
/* In-place transpose of a square tile matrix: tile (m,n) <-> tile (n,m)^T.
 * Grid: (tile pairs m >= n) x (32x32 sub-blocks); each block swaps a
 * 32 x 32 block pair through LDS so both sides are read and written along
 * columns. */
__global__ __launch_bounds__(256) void transpose_kernel(double *d, int nt, int ldt)
{
    __shared__ double S1[32][33];
    __shared__ double S2[32][33];
    const int pair = blockIdx.y;
    int i = (int)((sqrt(8.0 * (double)pair + 1.0) - 1.0) * 0.5);
    while ((i + 1) * (i + 2) / 2 <= pair) ++i;
    while (i * (i + 1) / 2 > pair) --i;
    const int m = i, n = pair - i * (i + 1) / 2;
    const int nsb = (ldt + 31) / 32;
    const int sbi = blockIdx.x % nsb, sbj = blockIdx.x / nsb;
    if (m == n && sbj > sbi) return;
    const size_t ts = (size_t)ldt * ldt;
    double *T1 = d + ((size_t)n * nt + m) * ts;   /* tile (m,n) */
    double *T2 = d + ((size_t)m * nt + n) * ts;   /* tile (n,m) */
    const int r0 = sbi * 32, c0 = sbj * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int jj = ty; jj < 32; jj += 8) {
        const int r = r0 + tx, c = c0 + jj;
        S1[tx][jj] = (r < ldt && c < ldt) ? T1[r + (size_t)c * ldt] : 0.0;
        const int r2 = c0 + tx, c2 = r0 + jj;
        S2[tx][jj] = (r2 < ldt && c2 < ldt) ? T2[r2 + (size_t)c2 * ldt] : 0.0;
    }
    __syncthreads();
    /* T1 block (r0.., c0..) <- (T2 block (c0.., r0..))^T and vice versa */
    for (int jj = ty; jj < 32; jj += 8) {
        const int r = r0 + tx, c = c0 + jj;
        if (r < ldt && c < ldt) T1[r + (size_t)c * ldt] = S2[jj][tx];
        const int r2 = c0 + tx, c2 = r0 + jj;
        if (r2 < ldt && c2 < ldt) T2[r2 + (size_t)c2 * ldt] = S1[jj][tx];
    }
}

void transpose_tiles_inplace(double *d, int nt, int ldt, hipStream_t s)
{
    if (nt == 0) return;
    const int nsb = (ldt + 31) / 32;
    dim3 grid(nsb * nsb, nt * (nt + 1) / 2);
    hipLaunchKernelGGL(transpose_kernel, grid, dim3(256), 0, s, d, nt, ldt);
}

/* ------------------------------------------------------------------ */
/* Generators.  Element (r,c) of the big matrix draws the value
 * 0.5 - x * 2^-64 where x is the state of the LCG x <- a x + 1
 * (a = 6364136223846793005) after pos steps from `seed`:
 *   plgsy: pos = max(r,c) + min(r,c) * bigM, diagonal + bump
 *   plrnt: pos = r + c * bigM
 * (coreblas/include/coreblas/random.h:32-73, core_zplgsy.c:38-112,
 *  core_zplrnt.c:36-54).  One thread fills RUN consecutive rows of one
 * column of one tile: one jump-ahead, then single steps along the column
 * (pos + 1) or, where the position runs across rows, steps of bigM
 * positions with the precomputed affine map (abig, cbig).                */
#define RUN 32

struct GenArgs {
    Geo g;
    int m0, n0, bigM;        /* submatrix origin and the matrix height used for positions */
    int sym;                 /* 1: plgsy, 0: plrnt */
    int uplo_skip;           /* ChamLower / ChamUpper / ChamUpperLower of the *stored* tiles */
    int trans;               /* mirror holds the transpose */
    int diag_identity;
    double bump;
    unsigned long long seed, abig, cbig;
};

__global__ __launch_bounds__(256) void gen_kernel(double *d, GenArgs a)
{
    const Geo &g = a.g;
    const int runs = (g.ldt + RUN - 1) / RUN;
    const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long per_tile = (long long)runs * g.wdt;
    const int t = (int)(gid / per_tile);
    if (t >= g.mt * g.nt) return;
    const int rem = (int)(gid % per_tile);
    const int col = rem / runs, run = rem % runs;
    const int m = t % g.mt, n = t / g.mt;
    /* skip rule of pzplgsy.c:62-69 on the stored orientation */
    const int tm0 = m * g.mb, tn0 = n * g.nb, vm = g.vm(m), vn = g.vn(n);
    if (a.uplo_skip == 122 /* Lower */ && tn0 >= tm0 + vm) return;
    if (a.uplo_skip == 121 /* Upper */ && tm0 >= tn0 + vn) return;
    double *tile = d + g.tile(m, n);
    const int r_begin = run * RUN;
    const int r_end = min(r_begin + RUN, g.ldt);
    unsigned long long x1 = 0, xb = 0;   /* running states: unit steps / bigM steps */
    int next1 = -1, nextb = -1;          /* row at which x1 / xb is valid       */
    for (int r = r_begin; r < r_end; ++r) {
        double v;
        if (r >= vm || col >= vn) {
            v = (a.diag_identity && m == n && r == col) ? 1.0 : 0.0;
        } else {
            /* element (r, col) of this tile is element (hr, hc) of the
             * (sub)matrix in host orientation */
            int hr = tm0 + r, hc = tn0 + col;
            if (a.trans) { int tmp = hr; hr = hc; hc = tmp; }
            const unsigned long long R = (unsigned long long)(hr + a.m0), Cc = (unsigned long long)(hc + a.n0);
            /* along the stored column the row index r varies: in host
             * orientation that is hr (no trans) or hc (trans) */
            bool unit_step;   /* consecutive r -> position + 1 ? */
            unsigned long long pos;
            if (a.sym) {
                if (R >= Cc) { pos = R + Cc * (unsigned long long)a.bigM; unit_step = !a.trans; }
                else         { pos = Cc + R * (unsigned long long)a.bigM; unit_step = a.trans; }
            } else {
                pos = R + Cc * (unsigned long long)a.bigM; unit_step = !a.trans;
            }
            unsigned long long x;
            if (unit_step) {
                if (next1 != r) x1 = rnd_jump(pos, a.seed);
                x = x1; x1 = RND_A * x1 + RND_C; next1 = r + 1;
            } else {
                if (nextb != r) xb = rnd_jump(pos, a.seed);
                x = xb; xb = a.abig * xb + a.cbig; nextb = r + 1;
            }
            v = rnd_val(x);
            if (a.sym && R == Cc) v += a.bump;
        }
        tile[r + (size_t)col * g.ldt] = v;
    }
}

static void affine_pow(unsigned long long n, unsigned long long *A, unsigned long long *C)
{
    unsigned long long a = RND_A, c = RND_C, ra = 1, rc = 0;
    while (n) {
        if (n & 1ULL) { rc = a * rc + c; ra = a * ra; }
        c = c * (a + 1ULL);
        a = a * a;
        n >>= 1;
    }
    *A = ra; *C = rc;
}

static void gen_launch(double *d, GenArgs a, hipStream_t s)
{
    const Geo &g = a.g;
    affine_pow((unsigned long long)a.bigM, &a.abig, &a.cbig);
    const long long runs = (g.ldt + RUN - 1) / RUN;
    const long long total = runs * g.wdt * (long long)g.mt * g.nt;
    if (total == 0) return;
    hipLaunchKernelGGL(gen_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, d, a);
}

void plgsy(double *d, int mt, int nt, int ldt, int wdt, int mb, int nb, int M, int N, double bump, int uplo_skip,
           int bigM, int m0, int n0, unsigned long long seed, int trans, bool diag_identity, hipStream_t s)
{
    GenArgs a;
    a.g = geo(mt, nt, ldt, wdt, mb, nb, M, N);
    a.m0 = m0; a.n0 = n0; a.bigM = bigM; a.sym = 1; a.uplo_skip = uplo_skip; a.trans = trans;
    a.diag_identity = diag_identity; a.bump = bump; a.seed = seed;
    gen_launch(d, a, s);
}

void plrnt(double *d, int mt, int nt, int ldt, int wdt, int mb, int nb, int M, int N, int bigM, int m0, int n0,
           unsigned long long seed, int trans, hipStream_t s)
{
    GenArgs a;
    a.g = geo(mt, nt, ldt, wdt, mb, nb, M, N);
    a.m0 = m0; a.n0 = n0; a.bigM = bigM; a.sym = 0; a.uplo_skip = 123; a.trans = trans;
    a.diag_identity = 0; a.bump = 0.0; a.seed = seed;
    gen_launch(d, a, s);
}

/* ------------------------------------------------------------------ */
__global__ void set_identity_kernel(double *d, int ntiles, int ldt)
{
    const size_t ts = (size_t)ldt * ldt;
    const size_t total = ts * ntiles;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const size_t w = e % ts;
        d[e] = ((w % ldt) == (w / ldt)) ? 1.0 : 0.0;
    }
}

void set_identity(double *d, int ntiles, int ldt, hipStream_t s)
{
    hipLaunchKernelGGL(set_identity_kernel, dim3(1024), dim3(256), 0, s, d, ntiles, ldt);
}

/* ------------------------------------------------------------------ */
/* Norm partials.  For every element of the stored orientation we know its
 * (row, col) in the logical matrix; `uplo` restricts to a stored triangle
 * that represents a symmetric matrix (lansy), ChamUpperLower is a general
 * matrix (lange).  Outputs (device, zeroed by the caller):
 *   colsum[N], rowsum[M], w[0] = max |a|, w[1] = sum a^2                  */
__global__ __launch_bounds__(256) void norm_kernel(const double *d, Geo g, int uplo, int trans, double *colsum,
                                                    double *rowsum, double *w)
{
    /* One workgroup per tile.  Sums are formed per tile row (RI) and per
     * tile column (CJ) over the stored triangle, then added to the logical
     * row/column sums with one atomic per index: element (r, c) adds to
     * rowsum[r] and colsum[c], and, as an off-diagonal element of a
     * symmetric matrix, also to rowsum[c] and colsum[r]. */
    const int t = blockIdx.x;
    const int m = t % g.mt, n = t / g.mt;
    const int vm = g.vm(m), vn = g.vn(n);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const bool sym = uplo != 123;
    /* logical origin of the tile and the roles of its rows/columns */
    const int r0 = trans ? n * g.nb : m * g.mb, c0 = trans ? m * g.mb : n * g.nb;   /* of (i,j) -> (r,c) */
    auto stored = [&](int i, int j) {
        const int r = trans ? r0 + j : r0 + i, c = trans ? c0 + i : c0 + j;
        return uplo == 123 || (uplo == 122 ? r >= c : r <= c);
    };
    auto diagonal = [&](int i, int j) { return (trans ? r0 + j == c0 + i : r0 + i == c0 + j); };
    /* skip tiles outside the stored triangle */
    if (uplo != 123) {
        const int rlo = trans ? r0 : r0, rhi = trans ? r0 + vn - 1 : r0 + vm - 1;
        const int clo = c0, chi = trans ? c0 + vm - 1 : c0 + vn - 1;
        if (uplo == 122 && rhi < clo) return;
        if (uplo == 121 && rlo > chi) return;
    }
    const double *tile = d + g.tile(m, n);
    double mx = 0.0, ss = 0.0;
    /* tile rows: coalesced across threads, one column at a time */
    for (int i = tid; i < vm; i += 256) {
        double s = 0.0, sd = 0.0;
        for (int j = 0; j < vn; ++j) {
            if (!stored(i, j)) continue;
            const double v = fabs(tile[i + (size_t)j * g.ldt]);
            s += v;
            if (diagonal(i, j)) { sd += v; ss += v * v; } else ss += (sym ? 2.0 : 1.0) * v * v;
            mx = nan_max(mx, v);
        }
        /* tile row i is logical row r0 + i (no trans) or column c0 + i (trans) */
        if (!trans) {
            atomicAdd(&rowsum[r0 + i], s);
            if (sym) atomicAdd(&colsum[r0 + i], s - sd);
        } else {
            atomicAdd(&colsum[c0 + i], s);
            if (sym) atomicAdd(&rowsum[c0 + i], s - sd);
        }
    }
    /* tile columns: one wavefront per column, lanes over rows */
    for (int j = wave; j < vn; j += 4) {
        double s = 0.0, sd = 0.0;
        for (int i = lane; i < vm; i += 64) {
            if (!stored(i, j)) continue;
            const double v = fabs(tile[i + (size_t)j * g.ldt]);
            s += v;
            if (diagonal(i, j)) sd += v;
        }
        for (int off = 32; off > 0; off >>= 1) {
            s += __shfl_down(s, off);
            sd += __shfl_down(sd, off);
        }
        if (lane == 0) {
            if (!trans) {
                atomicAdd(&colsum[c0 + j], s);
                if (sym) atomicAdd(&rowsum[c0 + j], s - sd);
            } else {
                atomicAdd(&rowsum[r0 + j], s);
                if (sym) atomicAdd(&colsum[r0 + j], s - sd);
            }
        }
    }
    /* wave reduction then one atomic per wave */
    for (int off = 32; off > 0; off >>= 1) {
        mx = nan_max(mx, __shfl_down(mx, off));
        ss += __shfl_down(ss, off);
    }
    if (lane == 0) {
        atomicAdd(&w[1], ss);
        /* max of non-negative doubles via integer max on the bit pattern */
        atomicMax(reinterpret_cast<unsigned long long *>(&w[0]), (unsigned long long)__double_as_longlong(mx));
    }
}

void norm_partials(const double *d, int mt, int nt, int ldt, int wdt, int mb, int nb, int M, int N, int uplo,
                   int trans, double *colsum, double *rowsum, double *w, hipStream_t s)
{
    if (mt * nt == 0) return;
    Geo g = geo(mt, nt, ldt, wdt, mb, nb, M, N);
    hipLaunchKernelGGL(norm_kernel, dim3(mt * nt), dim3(256), 0, s, d, g, uplo, trans, colsum, rowsum, w);
}

/* ------------------------------------------------------------------ */
/* Tile-wise copy of the `uplo` part between two mirrors of equal geometry
 * (src in orientation src_trans, dst in normal orientation). */
__global__ __launch_bounds__(256) void lacpy_kernel(const double *src, double *dst, Geo g, int uplo)
{
    const int t = blockIdx.y;
    const int m = t % g.mt, n = t / g.mt;
    const size_t ts = (size_t)g.ldt * g.wdt;
    const double *S = src + g.tile(m, n);
    double *D = dst + g.tile(m, n);
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < ts; e += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(e % g.ldt), j = (int)(e / g.ldt);
        const int r = m * g.mb + i, c = n * g.nb + j;
        if (uplo == 122 && r < c) continue;
        if (uplo == 121 && r > c) continue;
        D[e] = S[e];
    }
}

void lacpy_tiles(const double *src, double *dst, int mt, int nt, int ldt, int wdt, int mb, int nb, int M, int N,
                 int uplo, hipStream_t s)
{
    if (mt * nt == 0) return;
    Geo g = geo(mt, nt, ldt, wdt, mb, nb, M, N);
    dim3 grid(16, mt * nt);
    hipLaunchKernelGGL(lacpy_kernel, grid, dim3(256), 0, s, src, dst, g, uplo);
}

} /* namespace k */
} /* namespace chx */

namespace chx {
namespace k {

/* ------------------------------------------------------------------ */
/* Column panel <-> tile mirror.  P is a dense column-major block of
 * `rows` x `cols` doubles (leading dimension rows) holding rows
 * r0 .. r0+rows-1 and columns c0 .. c0+cols-1 of tile column n (global
 * row g = r0 + i lies in tile row g / mb at row g % mb; column c0 + c of
 * the tile column).  to_tiles = 1 scatters P into the tiles, 0 gathers the
 * tiles into P.  One thread per element, consecutive threads down a
 * column on both sides. */
__global__ void panel_tiles_kernel(double *P, double *D, int mt, int ldt, long long tile_elems, int mb, int n, int r0,
                                   int rows, int c0, int cols, int to_tiles)
{
    const long long total = (long long)rows * cols;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(e % rows), c = (int)(e / rows);
        const int g = r0 + i, m = g / mb, ii = g - m * mb;
        double *t = D + ((long long)n * mt + m) * tile_elems + ii + (long long)(c0 + c) * ldt;
        if (to_tiles) *t = P[e];
        else P[e] = *t;
    }
}

void panel_tiles(double *P, double *D, int mt, int ldt, size_t tile_elems, int mb, int n, int r0, int rows, int c0,
                 int cols, int to_tiles, hipStream_t s)
{
    const long long total = (long long)rows * cols;
    if (total <= 0) return;
    const long long blocks = std::min<long long>((total + 255) / 256, 65536);
    hipLaunchKernelGGL(panel_tiles_kernel, dim3((unsigned)blocks), dim3(256), 0, s, P, D, mt, ldt,
                       (long long)tile_elems, mb, n, r0, rows, c0, cols, to_tiles);
}

} /* namespace k */
} /* namespace chx */
