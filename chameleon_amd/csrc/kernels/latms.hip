/*
 * latms.hip -- kernels of the conditioned SPD generator
 * A = Q diag(lambda) Q^T, Q the orthogonal factor of the QR factorisation
 * of a seeded random N x N matrix G (the construction of the reference's
 * chameleon_pzlatms, compute/pzlatms.c:194-239: U <- qr(plrnt), A <- U D U^T;
 * spectrum of coreblas/compute/core_zlatm1.c:172-180, mode 3).
 *
 * Q is formed by shifted CholeskyQR3 (csrc/latms.cpp): C = G^T G, C = R^T R,
 * G <- G R^-1, three passes.  The kernels here are the pieces that the
 * factorisation kernels of the Cholesky path do not already provide:
 *   gauss_kernel          G with N(0,1) entries from the Chameleon LCG
 *   transpose_full_kernel tile matrix transpose, out of place
 *   tile_gemm_nt_kernel   C(m,n) = beta C(m,n) + alpha sum_k A(m,k) B(n,k)^T
 *                         over a rectangle or a lower triangle of tiles
 *   scale_cols_kernel     W = Q diag(s)
 *   symmetrize_kernel     upper triangle <- lower triangle
 *   add_diag_kernel, sumsq_kernel  shift and squared Frobenius norm
 * All matrices are full square tile mirrors: NT x NT tiles of ldt x ldt
 * doubles, tile (m,n) at (n NT + m) ldt^2, valid rows of tile row m are
 * min(mb, N - m mb); padding is the identity on the diagonal, 0 elsewhere.
 */
#include "../internal.h"
#include "mfma.h"
#include "rng.h"

#include <algorithm>

namespace chx {
namespace k {

struct SqGeo {
    int NT, ldt, mb, N;
    __device__ __forceinline__ int valid(int m) const { int r = N - m * mb; return r < mb ? r : mb; }
};

/* ------------------------------------------------------------------ */
/* Normal deviates by the Box-Muller transform of two consecutive values of
 * the Chameleon LCG: element (i, j), i, j < N, uses the states after 2p and
 * 2p + 1 steps from `seed`, p = i + j N;  u = (state >> 11 + 1/2) 2^-53 in
 * (0, 1) for each, G(i, j) = sqrt(-2 ln u1) cos(2 pi u2).
 * One thread per 32-element run of a tile column. */
__device__ __forceinline__ double gauss_from(unsigned long long x1, unsigned long long x2)
{
    const double u1 = ((double)(x1 >> 11) + 0.5) * 0x1p-53;
    const double u2 = ((double)(x2 >> 11) + 0.5) * 0x1p-53;
    return sqrt(-2.0 * log(u1)) * cospi(2.0 * u2);
}

__global__ void gauss_kernel(double *G, SqGeo g, unsigned long long seed)
{
    const int runs = g.ldt / 32 + (g.ldt % 32 ? 1 : 0);
    const long long total = (long long)g.NT * g.NT * g.ldt * runs;
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= total) return;
    const int run = (int)(id % runs);
    long long rest = id / runs;
    const int c = (int)(rest % g.ldt);
    rest /= g.ldt;
    const int m = (int)(rest % g.NT), n = (int)(rest / g.NT);
    const size_t ts = (size_t)g.ldt * g.ldt;
    double *T = G + ((size_t)n * g.NT + m) * ts + (size_t)c * g.ldt;
    const int vm = g.valid(m), vn = g.valid(n);
    const int r0 = run * 32, r1 = min(r0 + 32, g.ldt);
    if (c >= vn) {
        for (int r = r0; r < r1; ++r) T[r] = (m == n && r == c) ? 1.0 : 0.0;
        return;
    }
    const long long gj = (long long)n * g.mb + c;
    unsigned long long x = 0;
    bool have = false;
    for (int r = r0; r < r1; ++r) {
        if (r >= vm) { T[r] = 0.0; continue; }
        if (!have) {
            const unsigned long long p = (unsigned long long)((long long)m * g.mb + r) + (unsigned long long)gj * g.N;
            x = rnd_jump(2ULL * p, seed);
            have = true;
        }
        const unsigned long long x2 = RND_A * x + RND_C;
        T[r] = gauss_from(x, x2);
        x = RND_A * x2 + RND_C;
    }
}

void gauss(double *G, int NT, int ldt, int mb, int N, unsigned long long seed, hipStream_t s)
{
    SqGeo g = {NT, ldt, mb, N};
    const long long total = (long long)NT * NT * ldt * (ldt / 32 + (ldt % 32 ? 1 : 0));
    if (total <= 0) return;
    hipLaunchKernelGGL(gauss_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, G, g, seed);
}

/* ------------------------------------------------------------------ */
/* dst tile (m, n) = src tile (n, m)^T, 32 x 32 blocks through LDS. */
__global__ __launch_bounds__(256) void transpose_full_kernel(const double *src, double *dst, int NT, int ldt)
{
    __shared__ double S[32][33];
    const int nb = ldt / 32 + (ldt % 32 ? 1 : 0);
    const int bpt = nb * nb;
    const long long t = blockIdx.x / bpt;
    const int b = blockIdx.x % bpt;
    const int m = (int)(t % NT), n = (int)(t / NT);
    const int bi = (b % nb) * 32, bj = (b / nb) * 32;     /* block of the dst tile */
    const size_t ts = (size_t)ldt * ldt;
    const double *Src = src + ((size_t)m * NT + n) * ts;  /* tile (n, m) of src */
    double *Dst = dst + ((size_t)n * NT + m) * ts;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int j = ty; j < 32; j += 8) {
        /* src element (bj + tx, bi + j) -> dst element (bi + j, bj + tx) */
        const int r = bj + tx, c = bi + j;
        if (r < ldt && c < ldt) S[j][tx] = Src[r + (size_t)c * ldt];
    }
    __syncthreads();
    for (int j = ty; j < 32; j += 8) {
        const int r = bi + tx, c = bj + j;
        if (r < ldt && c < ldt) Dst[r + (size_t)c * ldt] = S[tx][j];
    }
}

void transpose_full(const double *src, double *dst, int NT, int ldt, hipStream_t s)
{
    const int nb = ldt / 32 + (ldt % 32 ? 1 : 0);
    const long long blocks = (long long)NT * NT * nb * nb;
    if (blocks <= 0) return;
    hipLaunchKernelGGL(transpose_full_kernel, dim3((unsigned)blocks), dim3(256), 0, s, src, dst, NT, ldt);
}

/* ------------------------------------------------------------------ */
/* W(:, j) = Q(:, j) s[j] (s indexed by the padded column index n ldt + c) */
__global__ void scale_cols_kernel(const double *Q, double *W, int NT, int ldt, const double *sc)
{
    const size_t ts = (size_t)ldt * ldt, total = ts * NT * NT;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const size_t t = e / ts;
        const int n = (int)(t / NT), c = (int)((e % ts) / ldt);
        W[e] = Q[e] * sc[(size_t)n * ldt + c];
    }
}

void scale_cols(const double *Q, double *W, int NT, int ldt, const double *sc, hipStream_t s)
{
    const size_t total = (size_t)ldt * ldt * NT * NT;
    if (!total) return;
    const unsigned grid = (unsigned)std::min<size_t>((total + 255) / 256, 65536);
    hipLaunchKernelGGL(scale_cols_kernel, dim3(grid), dim3(256), 0, s, Q, W, NT, ldt, sc);
}

/* A(i, j) <- A(j, i) for i < j: upper tiles from the lower ones, and the
 * strict upper part of the diagonal tiles from their lower part. */
__global__ __launch_bounds__(256) void symmetrize_kernel(double *A, int NT, int ldt)
{
    __shared__ double S[32][33];
    const int nb = ldt / 32 + (ldt % 32 ? 1 : 0);
    const int bpt = nb * nb;
    const long long t = blockIdx.x / bpt;
    const int b = blockIdx.x % bpt;
    const int m = (int)(t % NT), n = (int)(t / NT);        /* destination tile (m, n), m <= n */
    if (m > n) return;
    const int bi = (b % nb) * 32, bj = (b / nb) * 32;
    if (m == n && bi > bj) return;                         /* blocks below the diagonal stay */
    const size_t ts = (size_t)ldt * ldt;
    const double *Src = A + ((size_t)m * NT + n) * ts;     /* tile (n, m) */
    double *Dst = A + ((size_t)n * NT + m) * ts;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int j = ty; j < 32; j += 8) {
        const int r = bj + tx, c = bi + j;
        if (r < ldt && c < ldt) S[j][tx] = Src[r + (size_t)c * ldt];
    }
    __syncthreads();
    for (int j = ty; j < 32; j += 8) {
        const int r = bi + tx, c = bj + j;
        if (r < ldt && c < ldt && (m != n || r < c)) Dst[r + (size_t)c * ldt] = S[tx][j];
    }
}

void symmetrize(double *A, int NT, int ldt, hipStream_t s)
{
    const int nb = ldt / 32 + (ldt % 32 ? 1 : 0);
    const long long blocks = (long long)NT * NT * nb * nb;
    if (blocks <= 0) return;
    hipLaunchKernelGGL(symmetrize_kernel, dim3((unsigned)blocks), dim3(256), 0, s, A, NT, ldt);
}

/* valid diagonal entries += shift */
__global__ void add_diag_kernel(double *A, SqGeo g, double shift)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.N) return;
    const int m = (int)(i / g.mb), r = (int)(i % g.mb);
    A[((size_t)m * g.NT + m) * g.ldt * g.ldt + r + (size_t)r * g.ldt] += shift;
}

void add_diag(double *A, int NT, int ldt, int mb, int N, double shift, hipStream_t s)
{
    SqGeo g = {NT, ldt, mb, N};
    hipLaunchKernelGGL(add_diag_kernel, dim3((N + 255) / 256), dim3(256), 0, s, A, g, shift);
}

/* per-tile sum of squares, one block per tile, summed in a fixed order */
__global__ __launch_bounds__(256) void sumsq_kernel(const double *A, int ldt, double *part)
{
    __shared__ double red[256];
    const size_t ts = (size_t)ldt * ldt;
    const double *T = A + (size_t)blockIdx.x * ts;
    double acc = 0.0;
    for (size_t e = threadIdx.x; e < ts; e += 256) acc += T[e] * T[e];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

void sumsq_tiles(const double *A, int ntiles, int ldt, double *part, hipStream_t s)
{
    if (ntiles <= 0) return;
    hipLaunchKernelGGL(sumsq_kernel, dim3(ntiles), dim3(256), 0, s, A, ldt, part);
}

/* ================================================================== */
/* General tile product C(m,n) = beta C(m,n) + alpha sum_{k0<=k<k1} A(m,k) B(n,k)^T
 * over the tiles r0 <= m < r0+nr, c0 <= n < c0+nc, or (lower) the lower
 * triangle m >= n of the square r0 = c0, nr = nc (diagonal tiles: lower
 * part only).  A, B, C are tile mirrors with amt, bmt, cmt tile rows.
 * Same structure as the trailing update (chol.hip): BM x BN blocks, K in
 * 16-deep LDS slices, two LDS stages, operand fragments one step ahead,
 * XCD-aware block order, operands swapped so results store coalesced.   */
struct TileGemmArgs {
    const double *A, *B;
    double *C;
    int amt, bmt, cmt, ldt;
    int k0, k1;
    int r0, nr, c0, nc, lower;
    double alpha, beta;
    int total;
};

__device__ __forceinline__ int xcd_remap_g(int id, int total)
{
    const int nx = 8;
    int per = total / nx, rem = total % nx;
    int x = id % nx, local = id / nx;
    return (x < rem) ? x * (per + 1) + local : rem * (per + 1) + (x - rem) * per + local;
}

template <int BM, int BN, int WM, int WN>
__global__ __launch_bounds__((BM / WM) * (BN / WN) * 64, 2) void tile_gemm_nt_kernel(TileGemmArgs g)
{
    constexpr int NW = (BM / WM) * (BN / WN);
    constexpr int NTH = NW * 64;
    constexpr int BK = 16;
    constexpr int FM = WM / 16, FN = WN / 16;
    constexpr int LA = (BK * BM) / (2 * NTH);
    constexpr int LB = (BK * BN) / (2 * NTH);
    constexpr int KS = BK / 4;
    static_assert(LA >= 1 && LB >= 1, "tile too small for the thread count");
    __shared__ double As[2][BK * BM];
    __shared__ double Bs[2][BK * BN];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int ldt = g.ldt;
    const int bid = xcd_remap_g(blockIdx.x, g.total);
    const int tbm = ldt / BM, tbn = ldt / BN;
    const int bpt = tbm * tbn;
    const int pair = bid / bpt, b = bid % bpt;
    int m, n;
    if (g.lower) {
        int i = (int)((sqrt(8.0 * (double)pair + 1.0) - 1.0) * 0.5);
        while ((long long)(i + 1) * (i + 2) / 2 <= pair) ++i;
        while ((long long)i * (i + 1) / 2 > pair) --i;
        m = g.r0 + i;
        n = g.c0 + (pair - i * (i + 1) / 2);
    } else {
        m = g.r0 + pair % g.nr;
        n = g.c0 + pair / g.nr;
    }
    const int bi = b % tbm, bj = b / tbm;
    const int row0 = bi * BM, col0 = bj * BN;
    const bool diag = g.lower && m == n;
    if (diag && col0 >= row0 + BM) return;

    const size_t ts = (size_t)ldt * ldt;
    const double *Am = g.A + ((size_t)g.k0 * g.amt + m) * ts + row0;
    const double *Bn = g.B + ((size_t)g.k0 * g.bmt + n) * ts + col0;
    double *C = g.C + ((size_t)n * g.cmt + m) * ts;
    const size_t sa = (size_t)g.amt * ts, sb = (size_t)g.bmt * ts;

    const int wm = (wave % (BM / WM)) * WM, wn = (wave / (BM / WM)) * WN;
    v4d acc[FM][FN];
#pragma unroll
    for (int fm = 0; fm < FM; ++fm)
#pragma unroll
        for (int fn = 0; fn < FN; ++fn) acc[fm][fn] = zero4d();

    v2d ra[LA], rb[LB];
    const int SK = ldt / BK, S = (g.k1 - g.k0) * SK;
    auto gload = [&](int st) {
        const int pi = st / SK, kb = (st - pi * SK) * BK;
        const double *a = Am + pi * sa, *bb = Bn + pi * sb;
#pragma unroll
        for (int it = 0; it < LA; ++it) {
            const int e = (it * NTH + tid) * 2;
            ra[it] = *reinterpret_cast<const v2d *>(a + (e % BM) + (size_t)(kb + e / BM) * ldt);
        }
#pragma unroll
        for (int it = 0; it < LB; ++it) {
            const int e = (it * NTH + tid) * 2;
            rb[it] = *reinterpret_cast<const v2d *>(bb + (e % BN) + (size_t)(kb + e / BN) * ldt);
        }
    };
    auto lstore = [&](int buf) {
#pragma unroll
        for (int it = 0; it < LA; ++it) *reinterpret_cast<v2d *>(&As[buf][(it * NTH + tid) * 2]) = ra[it];
#pragma unroll
        for (int it = 0; it < LB; ++it) *reinterpret_cast<v2d *>(&Bs[buf][(it * NTH + tid) * 2]) = rb[it];
    };
    double fa[2][FM], fb[2][FN];
    auto fload = [&](int buf, int kk, int set) {
#pragma unroll
        for (int fm = 0; fm < FM; ++fm) fa[set][fm] = As[buf][(kk + (lane >> 4)) * BM + wm + fm * 16 + (lane & 15)];
#pragma unroll
        for (int fn = 0; fn < FN; ++fn) fb[set][fn] = Bs[buf][(kk + (lane >> 4)) * BN + wn + fn * 16 + (lane & 15)];
    };
    if (S > 0) {
        gload(0);
        lstore(0);
        if (S > 1) gload(1);
        __syncthreads();
        fload(0, 0, 0);
    }
    for (int st = 0; st < S; ++st) {
        const int cur = st & 1;
        if (st + 1 < S) {
            lstore(cur ^ 1);
            if (st + 2 < S) gload(st + 2);
        }
#pragma unroll
        for (int q = 0; q < KS; ++q) {
            const int set = q & 1;
            if (q + 1 < KS) {
                fload(cur, 4 * (q + 1), set ^ 1);
            } else {
                __syncthreads();
                if (st + 1 < S) fload(cur ^ 1, 0, set ^ 1);
            }
#pragma unroll
            for (int fm = 0; fm < FM; ++fm)
#pragma unroll
                for (int fn = 0; fn < FN; ++fn) acc[fm][fn] = mfma_f64(fb[set][fn], fa[set][fm], acc[fm][fn]);
        }
    }
#pragma unroll
    for (int fm = 0; fm < FM; ++fm)
#pragma unroll
        for (int fn = 0; fn < FN; ++fn)
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int gi = row0 + wm + fm * 16 + (lane & 15);
                const int gj = col0 + wn + fn * 16 + (lane >> 4) + 4 * s;
                if (diag && gi < gj) continue;
                double *pc = C + gi + (size_t)gj * ldt;
                const double v = g.alpha * acc[fm][fn][s];
                *pc = (g.beta == 0.0) ? v : g.beta * (*pc) + v;
            }
}

template <int BM, int BN, int WM, int WN>
static void launch_tile_gemm(TileGemmArgs g, long long pairs, hipStream_t s)
{
    const long long total = pairs * (g.ldt / BM) * (g.ldt / BN);
    if (total <= 0) return;
    g.total = (int)total;
    constexpr int NTH = (BM / WM) * (BN / WN) * 64;
    hipLaunchKernelGGL((tile_gemm_nt_kernel<BM, BN, WM, WN>), dim3((unsigned)total), dim3(NTH), 0, s, g);
}

void tile_gemm_nt(const double *A, int amt, const double *B, int bmt, double *C, int cmt, int ldt, int k0, int k1,
                  int r0, int nr, int c0, int nc, int lower, double alpha, double beta, hipStream_t s)
{
    if (nr <= 0 || nc <= 0) return;
    TileGemmArgs g;
    g.A = A; g.B = B; g.C = C;
    g.amt = amt; g.bmt = bmt; g.cmt = cmt; g.ldt = ldt;
    g.k0 = k0; g.k1 = k1;
    g.r0 = r0; g.nr = nr; g.c0 = c0; g.nc = nc; g.lower = lower;
    g.alpha = alpha; g.beta = beta; g.total = 0;
    const long long pairs = lower ? (long long)nr * (nr + 1) / 2 : (long long)nr * nc;
    if (ldt % 128 == 0)
        launch_tile_gemm<128, 128, 32, 64>(g, pairs, s);
    else if (ldt % 64 == 0)
        launch_tile_gemm<64, 64, 32, 32>(g, pairs, s);
    else if (ldt % 32 == 0)
        launch_tile_gemm<32, 32, 16, 16>(g, pairs, s);
    else
        launch_tile_gemm<16, 16, 16, 16>(g, pairs, s);
}

} /* namespace k */
} /* namespace chx */
