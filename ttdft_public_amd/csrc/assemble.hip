// Hamiltonian assembly on the GPU: H_loc and the projector rows C, truncated
// to CSR (the reference builds both row by row on the host:
// dft/ProjectHamiltonianSparse.cc:330-575 Create_Hloc, :1223-1592
// Compute_nnz_loc, :83-328 Create_Cnloc).
//
// H_loc[(xI,yI,zI),(x,y,z)] =
//     1/2 (DX[xI,x] Y[yI,y] Z[zI,z] + X[xI,x] DY[yI,y] Z[zI,z] + X[xI,x] Y[yI,y] DZ[zI,z])
//   + sum_{p,q,r} G[p,q,r] PX[p][xI][x] PY[q][yI][y] PZ[r][zI][z],
// keeping the entries with |v| > truncation, columns ascending.
//
// The reference forms a dense length-n row per basis function (O(n^2)
// values, 3.75e7 flop per row at rx = ry = rz = 50).  Here the potential part
// of every entry comes out of three matrix products on the matrix cores
// (rocBLAS), with the pairs (xI,x), (yI,y), (zI,z) as matrix indices:
//   A[(x + rx xI) + rx^2 (q + rvy r)] = sum_p PX[p][xI][x] G[p,q,r]
//   B[r][t],  t = (x + rx xI) + rx^2 (y + ry yI) = x + rx g,
//             g = xI + rx (y + ry yI):   B_r = A_r PY^T
//   Cc[(z + rz zI), t] = sum_r PZ[r][zI][z] B[r][t]      (over chunks of g)
// and a hand-written kernel adds the kinetic term, truncates and compacts.
// Compaction is exact and deterministic in two passes:
//   1. count: cnt[row][y][z] = number of kept x for (row, y, z);
//   2. per-row scan in the CSR order (z, then y) -> rel[row][y][z], row
//      lengths -> rowptr;
//   3. write: recompute, entry of x goes to rowptr[row] + rel + (kept x' < x).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <string>
#include <vector>

#include "internal.h"

namespace ttdft {
namespace {

#define TT_BLAS(call)                                                                     \
    do {                                                                                  \
        const rocblas_status st_ = (call);                                                \
        TT_CHECK(st_ == rocblas_status_success, TTDFT_ELIB, "%s: rocBLAS status %d", #call, \
                 (int)st_);                                                               \
    } while (0)

// device buffers freed on scope exit
// A device block from the context's block cache (dev_alloc), released back
// to it when the owner goes out of scope unless handed to a Hamiltonian
// (release), whose ttdft_ham_free returns it.
struct DevBuf {
    void *p = nullptr;
    ttdft_ctx ctx = nullptr;
    ~DevBuf() {
        if (p) dev_release(ctx, p);
    }
    int alloc(ttdft_ctx c, size_t bytes) {
        ctx = c;
        return dev_alloc(c, &p, bytes == 0 ? 8 : bytes);
    }
    template <typename T>
    T *as() const {
        return static_cast<T *>(p);
    }
    template <typename T>
    T *release() {
        T *q = static_cast<T *>(p);
        p = nullptr;
        return q;
    }
};

int upload_to(ttdft_ctx ctx, DevBuf &b, const void *src, size_t bytes) {
    TT_TRY(b.alloc(ctx, bytes));
    if (bytes) TT_HIP(hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    return TTDFT_OK;
}

// an environment variable that lowers a chunk size (smaller workspace)
int64_t env_cap(const char *name, int64_t value) {
    if (const char *env = std::getenv(name)) {
        const long v = std::strtol(env, nullptr, 10);
        if (v > 0 && v < value) return v;
    }
    return value;
}

unsigned blocks_for(int64_t threads, int per_block) {
    int64_t b = (threads + per_block - 1) / per_block;
    return (unsigned)std::max<int64_t>(1, b);
}

// One thread per ((zI,z) pair, group g = (xI, y, yI)); loops over x.  Threads
// of a wavefront take consecutive pairs, so the Cc reads and the cnt writes
// ([row][y][z], z fastest) are coalesced.  WRITE = 0 counts, 1 writes.
template <int WRITE>
__global__ __launch_bounds__(256) void hloc_compact_kernel(
    const double *__restrict__ Cc, int64_t g0, int64_t ng, int rx, int ry, int rz,
    const double *__restrict__ X, const double *__restrict__ DX, const double *__restrict__ Y,
    const double *__restrict__ DY, const double *__restrict__ Z, const double *__restrict__ DZ,
    double trunc, uint16_t *__restrict__ cnt, const int32_t *__restrict__ rel,
    const int64_t *__restrict__ rowptr, int32_t *__restrict__ col, double *__restrict__ val,
    int32_t *__restrict__ flags) {
    const int64_t npair = (int64_t)rz * rz;
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = tid < npair * ng;
    const int64_t t = valid ? tid : 0;  // lanes past the end mirror entry 0 and store nothing
    const int pi = (int)(t % npair);
    const int64_t gl = t / npair;
    const int zI = pi / rz, z = pi - zI * rz;  // Cc row index z + rz zI == pi
    const int64_t g = g0 + gl;
    const int xI = (int)(g % rx);
    const int64_t gy = g / rx;
    const int y = (int)(gy % ry), yI = (int)(gy / ry);
    const int64_t row = xI + (int64_t)rx * (yI + (int64_t)ry * zI);
    const int64_t key = (row * ry + y) * rz + z;
    const double yv = Y[(int64_t)yI * ry + y], dyv = DY[(int64_t)yI * ry + y];
    const double zv = Z[(int64_t)zI * rz + z], dzv = DZ[(int64_t)zI * rz + z];
    const double *xr = X + (int64_t)xI * rx, *dxr = DX + (int64_t)xI * rx;
    const double *cp = Cc + pi + npair * (gl * rx);
    int64_t base = 0;
    int limit = 0;  // entries the count pass found for (row, y, z)
    if (WRITE) {
        base = rowptr[row] + rel[key];
        limit = valid ? cnt[key] : 0;
    }
    const int32_t col0 = (int32_t)((int64_t)rx * (y + (int64_t)ry * z));
    int kept = 0;
    for (int x = 0; x < rx; ++x) {
        // the reference's dyad: 0.5 (DX Y Z + X DY Z + X Y DZ), then + potential
        const double kin = dxr[x] * yv * zv + xr[x] * dyv * zv + xr[x] * yv * dzv;
        const double v = 0.5 * kin + cp[npair * x];
        if (valid && fabs(v) > trunc) {
            if (WRITE && kept < limit) {
                col[base + kept] = col0 + x;
                val[base + kept] = v;
            }
            ++kept;
        }
    }
    if (!WRITE) {
        if (valid) cnt[key] = (uint16_t)kept;
        return;
    }
    // Both passes evaluate the same products, so the counts agree; a
    // difference would leave a slot unwritten and is reported to the host.
    if (kept != limit) flags[0] = 1;
}

// Per row: exclusive scan of cnt[row][y][z] in the CSR order (z outer, y
// inner) -> rel[row][y][z]; row length -> rowlen[row].  One block per row.
__global__ __launch_bounds__(256) void hloc_row_scan_kernel(const uint16_t *__restrict__ cnt,
                                                            int ry, int rz,
                                                            int32_t *__restrict__ rel,
                                                            int64_t *__restrict__ rowlen) {
    extern __shared__ int32_t sc_lds[];  // ry * rz counts in (z, y) order, then 256 partials
    const int64_t row = blockIdx.x;
    const int L = ry * rz;
    int32_t *v = sc_lds;
    int32_t *part = sc_lds + L;
    const uint16_t *c = cnt + row * L;
    for (int e = threadIdx.x; e < L; e += blockDim.x) {
        const int y = e / rz, z = e - y * rz;  // source index e = y rz + z
        v[z * ry + y] = c[e];
    }
    __syncthreads();
    // each thread scans a contiguous range of the (z, y) sequence
    const int per = (L + blockDim.x - 1) / blockDim.x;
    const int s0 = threadIdx.x * per, s1 = min(L, s0 + per);
    int32_t sum = 0;
    for (int s = s0; s < s1; ++s) sum += v[s];
    part[threadIdx.x] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        int32_t acc = 0;
        for (int t = 0; t < (int)blockDim.x; ++t) {
            const int32_t p = part[t];
            part[t] = acc;
            acc += p;
        }
        rowlen[row] = acc;
    }
    __syncthreads();
    int32_t acc = part[threadIdx.x];
    for (int s = s0; s < s1; ++s) {
        const int32_t cv = v[s];
        v[s] = acc;
        acc += cv;
    }
    __syncthreads();
    int32_t *r = rel + row * L;
    for (int e = threadIdx.x; e < L; e += blockDim.x) {
        const int y = e / rz, z = e - y * rz;
        r[e] = v[z * ry + y];
    }
}

// ---- fused product + compaction (matrix cores) -------------------------------
//
// C[pi][t] = sum_r Ae[pi][r] Be[r][t] with the kinetic term folded in as two
// more product terms: Ae[pi][rvz] = Z[zI][z], Ae[pi][rvz+1] = DZ[zI][z],
// Be[rvz][t] = 1/2 (DX[xI][x] Y[yI][y] + X[xI][x] DY[yI][y]),
// Be[rvz+1][t] = 1/2 X[xI][x] Y[yI][y], zero terms up to a multiple of 4.
// A workgroup computes a 64 (pi) x 256 (t) tile that holds Gt whole groups
// g (Gt rx <= 256 columns) on the FP64 16x16x4 MFMA; wave w owns columns
// [64 w, 64 w + 64), all 64 rows, 16 accumulators.  The tile never leaves
// the chip: the epilogue thresholds it, turns the keep bits into per-row
// masks in LDS (wavefront ballots), and from the masks
//   COUNT: cnt[row][y][z] = kept x of segment (pi, g);
//   WRITE: each kept entry goes to rowptr[row] + rel[key] + (kept x' < x).
struct FusedArgs {
    const double *A;  // Ae, (Mp rows) x (K4 + 8 planes), A[pi + Mp k], rows >= M zero
    const double *B;  // Be, (K4 + 8 planes) x ldb, B[k ldb + t], ldb >= ntup + 256
    int64_t ntup, ldb, ngroups, ntiles, nblocks;
    int32_t M, Mp, K4, rx, ry, rz, Gt, mtiles;
    double trunc;
    uint16_t *cnt;
    const int32_t *rel;
    const int64_t *rowptr;
    int32_t *col;
    double *val;
    int32_t *flags;  // mismatch, half-bandwidths x, y, z
};

typedef double fd4_t __attribute__((ext_vector_type(4)));

constexpr int kFusedGt = 16;  // most groups per tile (rx >= 16 uses fewer)

// bits [lo, hi) of a 256-bit row mask held as 16 x 16 bits
__device__ inline int mask_popc(const uint16_t *m, int lo, int hi) {
    int c = 0;
    for (int q = lo >> 4; q <= (hi - 1) >> 4 && lo < hi; ++q) {
        uint32_t bits = m[q];
        const int b0 = q * 16;
        if (lo > b0) bits &= ~((1u << (lo - b0)) - 1u);
        if (hi < b0 + 16) bits &= (1u << (hi - b0)) - 1u;
        c += __popc(bits);
    }
    return c;
}

template <int WRITE>
__global__ __launch_bounds__(256, 2) void hloc_fused_kernel(FusedArgs a) {
    __shared__ uint16_t mask[64][16];
    __shared__ int64_t segbase[64 * kFusedGt];
    __shared__ int32_t segcnt[64 * kFusedGt];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    // consecutive logical tiles on one XCD (dispatch is round-robin over 8)
    const int64_t b = blockIdx.x, per = a.nblocks >> 3;
    const int64_t L = (b & 7) * per + (b >> 3);
    const int mt = (int)(L % a.mtiles);
    const int64_t nt = L / a.mtiles;
    if (nt >= a.ntiles) return;  // whole workgroup: grid padding
    const int64_t g0 = nt * a.Gt;
    const int gcnt = (int)min((int64_t)a.Gt, a.ngroups - g0);
    const int ncols = gcnt * a.rx;
    const int64_t t0 = g0 * a.rx;
    const int pi0 = mt * 64;
    const int li = lane & 15, lk = lane >> 4;
    // A fragment of row tile i: rows pi0 + 16 i + li, term 4 ks + lk; B fragment of
    // column tile j: columns t0 + 64 w + 16 j + li.  Rows past M are zero rows of
    // Ae, columns past the tile's groups lie in Be (padding included) and are
    // masked in the epilogue; the prefetch runs at most 8 terms past K4.
    const double *pa = a.A + pi0 + li + (int64_t)a.Mp * lk;
    const double *pb = a.B + t0 + 64 * w + li + a.ldb * lk;
    const int64_t sa = 4 * (int64_t)a.Mp, sb = 4 * a.ldb;
    const int nks = a.K4 >> 2;
    fd4_t acc[4][4];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) acc[i][j] = fd4_t{0.0, 0.0, 0.0, 0.0};
    double fa0[4], fb0[4], fa1[4], fb1[4];
    auto load = [&](double *fa, double *fb, const double *qa, const double *qb) {
        for (int q = 0; q < 4; ++q) {
            fa[q] = qa[16 * q];
            fb[q] = qb[16 * q];
        }
    };
    auto mma = [&](const double *fa, const double *fb) {
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j)
                acc[i][j] =
                    __builtin_amdgcn_mfma_f64_16x16x4f64(fa[i], fb[j], acc[i][j], 0, 0, 0);
    };
    load(fa0, fb0, pa, pb);
    load(fa1, fb1, pa + sa, pb + sb);
    pa += 2 * sa;
    pb += 2 * sb;
    int ks = 0;
    for (; ks + 2 <= nks; ks += 2) {
        mma(fa0, fb0);
        load(fa0, fb0, pa, pb);
        mma(fa1, fb1);
        load(fa1, fb1, pa + sa, pb + sb);
        pa += 2 * sa;
        pb += 2 * sb;
    }
    if (ks < nks) mma(fa0, fb0);
    // ---- epilogue: keep bits -> row masks
    // lane (li, lk), tile (i, j), element v: row 16 i + lk + 4 v, column 64 w + 16 j + li
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int colj = 64 * w + 16 * j + li;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = 16 * i + lk + 4 * v;
                const bool keep =
                    colj < ncols && pi0 + row < a.M && fabs(acc[i][j][v]) > a.trunc;
                const uint64_t m = __ballot(keep);
                if (lane < 4) mask[16 * i + lane + 4 * v][4 * w + j] =
                                  (uint16_t)(m >> (16 * lane));
            }
        }
    __syncthreads();
    // ---- segments (row, group): counts, and for WRITE the base offsets
    const int nseg = 64 * gcnt;
    for (int s = threadIdx.x; s < nseg; s += 256) {
        const int r = s % 64, gl = s / 64;
        const int pi = pi0 + r;
        if (pi >= a.M) continue;
        const int kept = mask_popc(mask[r], gl * a.rx, gl * a.rx + a.rx);
        const int zI = pi / a.rz, z = pi - zI * a.rz;
        const int g = (int)g0 + gl;  // groups < rx^2 ry^2 < 2^31
        const int gy = g / a.rx, xI = g - gy * a.rx;
        const int yI = gy / a.ry, y = gy - yI * a.ry;
        const int64_t rowid = xI + (int64_t)a.rx * (yI + (int64_t)a.ry * zI);
        const int64_t key = (rowid * a.ry + y) * a.rz + z;
        if (!WRITE) {
            a.cnt[key] = (uint16_t)kept;
        } else {
            const int expect = a.cnt[key];
            if (kept != expect) a.flags[0] = 1;
            segbase[r * kFusedGt + gl] = a.rowptr[rowid] + a.rel[key];
            segcnt[r * kFusedGt + gl] = min(kept, expect);
        }
    }
    if (!WRITE) return;
    __syncthreads();
    // ---- scatter the kept entries (loops unrolled: acc stays in registers)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int colj = 64 * w + 16 * j + li;
        const int gl = colj / a.rx, x = colj - gl * a.rx;
        const int g = (int)g0 + gl;
        const int gy = g / a.rx;
        const int yI = gy / a.ry, y = gy - yI * a.ry;
        const int colx = x + a.rx * y;
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = 16 * i + lk + 4 * v;
                const int pi = pi0 + r;
                const double hv = acc[i][j][v];
                if (colj < ncols && pi < a.M && fabs(hv) > a.trunc) {
                    const int rank = mask_popc(mask[r], gl * a.rx, colj);
                    if (rank < segcnt[r * kFusedGt + gl]) {
                        const int64_t pos = segbase[r * kFusedGt + gl] + rank;
                        a.col[pos] = colx + a.rx * a.ry * (pi % a.rz);
                        a.val[pos] = hv;
                    }
                }
            }
    }
}

// Kronecker half-bandwidths of a CSR matrix in the basis x + rx (y + ry z):
// max |x - x'|, |y - y'|, |z - z'| over its entries.  One wavefront per row,
// one atomic per workgroup and axis.
__global__ __launch_bounds__(256) void band_extent_kernel(const int64_t *__restrict__ rowptr,
                                                          const int32_t *__restrict__ col,
                                                          int64_t n, int rx, int ry,
                                                          int32_t *__restrict__ ext) {
    __shared__ int red[4][3];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t r = (int64_t)blockIdx.x * 4 + w;
    int e0 = -1, e1 = -1, e2 = -1;
    if (r < n) {
        const int rxy = rx * ry;
        const int z = (int)(r / rxy), y = (int)(r % rxy) / rx, x = (int)(r % rx);
        for (int64_t e = rowptr[r] + lane; e < rowptr[r + 1]; e += 64) {
            const int j = col[e];
            const int zj = j / rxy, yj = (j - zj * rxy) / rx, xj = j % rx;
            e0 = max(e0, abs(xj - x));
            e1 = max(e1, abs(yj - y));
            e2 = max(e2, abs(zj - z));
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        e0 = max(e0, __shfl_xor(e0, o));
        e1 = max(e1, __shfl_xor(e1, o));
        e2 = max(e2, __shfl_xor(e2, o));
    }
    if (lane == 0) {
        red[w][0] = e0;
        red[w][1] = e1;
        red[w][2] = e2;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        int m = -1;
        for (int q = 0; q < 4; ++q) m = max(m, red[q][threadIdx.x]);
        if (m >= 0) atomicMax(ext + threadIdx.x, m);
    }
}

// Be kinetic planes (plane stride ldb): t = x + rx g, g = xI + rx (y + ry yI).
__global__ __launch_bounds__(256) void kinetic_planes_kernel(
    const double *__restrict__ X, const double *__restrict__ DX, const double *__restrict__ Y,
    const double *__restrict__ DY, int rx, int ry, int64_t ntup, int64_t ldb,
    double *__restrict__ planes) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntup) return;
    const int x = (int)(t % rx);
    const int64_t g = t / rx;
    const int xI = (int)(g % rx);
    const int64_t gy = g / rx;
    const int y = (int)(gy % ry), yI = (int)(gy / ry);
    const double xv = X[(int64_t)xI * rx + x], dxv = DX[(int64_t)xI * rx + x];
    const double yv = Y[(int64_t)yI * ry + y], dyv = DY[(int64_t)yI * ry + y];
    planes[t] = 0.5 * (dxv * yv + xv * dyv);
    planes[ldb + t] = 0.5 * (xv * yv);
}

// Dense projector rows: D[i][x + rx (y + ry z)] = sum_c W[i][c][x + rx y] MZ[i][z + rz c]
// for the rows of one batch (W = S_i x_0 MX_i x_1 MY_i, padded to rmax terms).
__global__ __launch_bounds__(256) void proj_dense_kernel(const double *__restrict__ W,
                                                         const double *__restrict__ MZ,
                                                         int64_t nrows, int rmax, int rx, int ry,
                                                         int rz, double *__restrict__ D) {
    const int64_t nxy = (int64_t)rx * ry, n = nxy * rz;
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= nrows * n) return;
    const int64_t i = tid / n, j = tid - i * n;
    const int64_t xy = j % nxy;
    const int z = (int)(j / nxy);
    double s = 0.0;
    for (int c = 0; c < rmax; ++c)
        s += W[(i * rmax + c) * nxy + xy] * MZ[(i * rmax + c) * rz + z];
    D[tid] = s;
}

// Ordered compaction of dense rows: block (row i, chunk of 4096 columns).
// COUNT: cnt[i][chunk]; WRITE: entries at rowptr[i] + off[i][chunk] + rank,
// ranks from wavefront ballots and per-wave prefix sums in LDS.
template <int WRITE>
__global__ __launch_bounds__(256) void dense_compact_kernel(
    const double *__restrict__ D, int64_t n, int nchunk, double trunc, int32_t *__restrict__ cnt,
    const int64_t *__restrict__ start, int32_t *__restrict__ col, double *__restrict__ val) {
    __shared__ int32_t wsum[4];
    const int64_t i = blockIdx.x / nchunk;
    const int chunk = (int)(blockIdx.x - i * nchunk);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const double *row = D + i * n;
    const int64_t c0 = (int64_t)chunk * 4096, c1 = std::min<int64_t>(n, c0 + 4096);
    int64_t pos = WRITE ? start[(int64_t)i * nchunk + chunk] : 0;
    int32_t total = 0;
    for (int64_t b = c0; b < c1; b += 256) {
        const int64_t j = b + threadIdx.x;
        const bool keep = j < c1 && fabs(row[j]) > trunc;
        const uint64_t m = __ballot(keep);
        const int rank = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) wsum[w] = __popcll(m);
        __syncthreads();
        int before = 0, all = 0;
        for (int t = 0; t < 4; ++t) {
            if (t < w) before += wsum[t];
            all += wsum[t];
        }
        if (WRITE && keep) {
            col[pos + before + rank] = (int32_t)j;
            val[pos + before + rank] = row[j];
        }
        pos += all;
        total += all;
        __syncthreads();
    }
    if (!WRITE && threadIdx.x == 0) cnt[(int64_t)i * nchunk + chunk] = total;
}

// Row scaling of the projector matrix: D C with D = diag(inv_c_lm) (block per row).
__global__ __launch_bounds__(256) void scale_rows_kernel(const int64_t *__restrict__ rowptr,
                                                         const double *__restrict__ val,
                                                         const double *__restrict__ d,
                                                         double *__restrict__ out64,
                                                         float *__restrict__ out32) {
    const int r = blockIdx.x;
    const double s = d[r];
    for (int64_t e = rowptr[r] + threadIdx.x; e < rowptr[r + 1]; e += blockDim.x) {
        const double v = val[e] * s;
        if (out64) out64[e] = v;
        if (out32) out32[e] = (float)v;
    }
}

// position of column j in the sorted columns [lo, hi), or -1
__device__ inline int64_t find_col(const int32_t *__restrict__ col, int64_t lo, int64_t hi,
                                   int32_t j) {
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        const int32_t c = col[mid];
        if (c == j) return mid;
        if (c < j)
            lo = mid + 1;
        else
            hi = mid;
    }
    return -1;
}

// C^T rows (one thread per basis function j): the projectors whose row of C
// holds column j, in ascending order, found by binary search of each row.
// FILL = 0 counts (ct_len[j]), 1 writes at ct_rowptr[j].
template <int FILL>
__global__ __launch_bounds__(256) void transpose_rows_kernel(
    const int64_t *__restrict__ c_rowptr, const int32_t *__restrict__ c_col,
    const double *__restrict__ c_val, int32_t p, int64_t n, int64_t *__restrict__ ct_len,
    const int64_t *__restrict__ ct_rowptr, int32_t *__restrict__ ct_col,
    double *__restrict__ ct_val64, float *__restrict__ ct_val32) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    int64_t o = FILL ? ct_rowptr[j] : 0;
    for (int32_t r = 0; r < p; ++r) {
        const int64_t e = find_col(c_col, c_rowptr[r], c_rowptr[r + 1], (int32_t)j);
        if (e < 0) continue;
        if (FILL) {
            ct_col[o] = r;
            if (ct_val64) ct_val64[o] = c_val[e];
            if (ct_val32) ct_val32[o] = (float)c_val[e];
        }
        ++o;
    }
    if (!FILL) ct_len[j] = o;
}

__global__ __launch_bounds__(256) void to_f32_kernel(const double *__restrict__ a,
                                                     float *__restrict__ b, int64_t count) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) b[i] = (float)a[i];
}

// ---- device pipelines -------------------------------------------------------

struct HlocIn {
    int32_t rx, ry, rz, rvx, rvy, rvz;
    const double *X, *DX, *Y, *DY, *Z, *DZ, *PX, *PY, *PZ, *G;
    double trunc;
};

// CSR on the device, with a host copy of its row pointer
struct DevCsr {
    DevBuf rowptr, col, val;
    std::vector<int64_t> rp;
    int64_t nnz = 0;
};

int check_hloc(const HlocIn &a) {
    TT_CHECK(a.rx > 0 && a.ry > 0 && a.rz > 0 && a.rvx > 0 && a.rvy > 0 && a.rvz > 0,
             TTDFT_EBADARG, "assemble_hloc: ranks must be > 0");
    TT_CHECK(a.X && a.DX && a.Y && a.DY && a.Z && a.DZ && a.PX && a.PY && a.PZ && a.G,
             TTDFT_EBADARG, "assemble_hloc: NULL input");
    TT_CHECK(a.trunc >= 0.0, TTDFT_EBADARG, "assemble_hloc: truncation < 0");
    TT_CHECK(a.rx <= 65535 && (int64_t)a.rx * a.ry * a.rz < ((int64_t)1 << 31), TTDFT_EBADARG,
             "assemble_hloc: rx = %d or n too large", a.rx);
    TT_CHECK((int64_t)a.rx * a.rx * a.ry * a.ry < ((int64_t)1 << 31), TTDFT_EBADARG,
             "assemble_hloc: rx^2 ry^2 exceeds the 32-bit GEMM dimension");
    TT_CHECK(sizeof(int32_t) * ((size_t)a.ry * a.rz + 256) <= 64 * 1024, TTDFT_EBADARG,
             "assemble_hloc: ry*rz = %d too large", a.ry * a.rz);
    return TTDFT_OK;
}

// H_loc rows.  Always: the row pointer (out.rp, out.nnz).  fill: the device
// CSR (out.rowptr / col / val) and the Kronecker half-bandwidths of the kept
// entries (ext, -1 when there are none).
int hloc_run(ttdft_ctx ctx, const HlocIn &a, bool fill, DevCsr &out, int ext[3]) {
    const int32_t rx = a.rx, ry = a.ry, rz = a.rz, rvx = a.rvx, rvy = a.rvy, rvz = a.rvz;
    const int64_t n = (int64_t)rx * ry * rz;
    const int64_t rx2 = (int64_t)rx * rx, ry2 = (int64_t)ry * ry, rz2 = (int64_t)rz * rz;
    const int64_t ngroups = rx * ry2;   // g = xI + rx (y + ry yI)
    const int64_t ntup = rx * ngroups;  // rx^2 ry^2
    DevBuf dX, dDX, dY, dDY, dZ, dDZ, dPX, dPY, dPZ, dG;
    TT_TRY(upload_to(ctx, dX, a.X, sizeof(double) * rx2));
    TT_TRY(upload_to(ctx, dDX, a.DX, sizeof(double) * rx2));
    TT_TRY(upload_to(ctx, dY, a.Y, sizeof(double) * ry2));
    TT_TRY(upload_to(ctx, dDY, a.DY, sizeof(double) * ry2));
    TT_TRY(upload_to(ctx, dZ, a.Z, sizeof(double) * rz2));
    TT_TRY(upload_to(ctx, dDZ, a.DZ, sizeof(double) * rz2));
    TT_TRY(upload_to(ctx, dPX, a.PX, sizeof(double) * rvx * rx2));
    TT_TRY(upload_to(ctx, dPY, a.PY, sizeof(double) * rvy * ry2));
    TT_TRY(upload_to(ctx, dPZ, a.PZ, sizeof(double) * rvz * rz2));
    TT_TRY(upload_to(ctx, dG, a.G, sizeof(double) * rvx * rvy * rvz));
    // deterministic products: the count and write passes must see equal values
    rocblas_atomics_mode atomics = rocblas_atomics_not_allowed;
    TT_BLAS(rocblas_get_atomics_mode(ctx->blas, &atomics));
    TT_BLAS(rocblas_set_atomics_mode(ctx->blas, rocblas_atomics_not_allowed));
    struct AtomicsRestore {
        rocblas_handle h;
        rocblas_atomics_mode m;
        ~AtomicsRestore() { rocblas_set_atomics_mode(h, m); }
    } restore{ctx->blas, atomics};
    // Matrix-core path (fused product + compaction) for rx <= 256; the
    // chunked library-product path otherwise, or when TTDFT_ASSEMBLE_PATH=gemm.
    const char *path_env = std::getenv("TTDFT_ASSEMBLE_PATH");
    const bool fused = rx <= 256 && !(path_env && std::string(path_env) == "gemm");
    const int K4 = (rvz + 2 + 3) / 4 * 4;  // Veff terms + 2 kinetic terms, multiple of 4
    // A = PX G, B_r = A_r PY^T (B_r are the first rvz planes of Be on the fused path)
    const int64_t mp = (rz2 + 63) / 64 * 64;          // Ae rows, padded to the tile
    const int64_t ldb = fused ? ntup + 256 : ntup;    // Be plane stride, padded
    DevBuf dA, dB, dAe;
    TT_TRY(dA.alloc(ctx, sizeof(double) * rx2 * rvy * rvz));
    TT_TRY(dB.alloc(ctx, sizeof(double) * (fused ? K4 + 8 : rvz) * ldb));
    if (fused) {
        std::vector<double> ae((size_t)mp * (K4 + 8), 0.0);
        for (int r = 0; r < rvz; ++r)
            std::copy(a.PZ + rz2 * r, a.PZ + rz2 * (r + 1), ae.begin() + mp * r);
        std::copy(a.Z, a.Z + rz2, ae.begin() + mp * rvz);
        std::copy(a.DZ, a.DZ + rz2, ae.begin() + mp * (rvz + 1));
        TT_TRY(upload_to(ctx, dAe, ae.data(), sizeof(double) * ae.size()));
        TT_HIP(hipMemsetAsync(dB.p, 0, sizeof(double) * (K4 + 8) * ldb, ctx->stream));
        TT_HIP(hipStreamSynchronize(ctx->stream));  // ae goes out of scope
    }
    const double one = 1.0, zero = 0.0;
    // chunks of groups: Cc holds rz^2 x (rx * gc) doubles (about 1 GB)
    int64_t gc = ((int64_t)1 << 27) / std::max<int64_t>(1, rz2 * rx);
    gc = std::max<int64_t>(1, std::min<int64_t>(gc, ngroups));
    gc = env_cap("TTDFT_ASSEMBLE_GROUPS", gc);  // groups (xI, y, yI) per chunk
    DevBuf dC, dcnt, drel, drowlen, dflag;
    TT_TRY(dflag.alloc(ctx, sizeof(int32_t) * 4));
    const int32_t flag0[4] = {0, -1, -1, -1};  // mismatch, half-bandwidths x, y, z
    TT_HIP(hipMemcpyAsync(dflag.p, flag0, sizeof(flag0), hipMemcpyHostToDevice, ctx->stream));
    if (!fused) TT_TRY(dC.alloc(ctx, sizeof(double) * rz2 * rx * gc));
    TT_TRY(dcnt.alloc(ctx, sizeof(uint16_t) * n * ry * rz));
    TT_TRY(drel.alloc(ctx, sizeof(int32_t) * n * ry * rz));
    TT_TRY(drowlen.alloc(ctx, sizeof(int64_t) * n));
    FusedArgs fa{};
    if (fused) {
        fa.A = dAe.as<double>();
        fa.B = dB.as<double>();
        fa.ntup = ntup;
        fa.ldb = ldb;
        fa.ngroups = ngroups;
        fa.M = (int32_t)rz2;
        fa.Mp = (int32_t)mp;
        fa.K4 = K4;
        fa.rx = rx;
        fa.ry = ry;
        fa.rz = rz;
        fa.Gt = std::min(kFusedGt, 256 / rx);
        fa.mtiles = (int32_t)((rz2 + 63) / 64);
        fa.ntiles = (ngroups + fa.Gt - 1) / fa.Gt;
        fa.nblocks = (fa.mtiles * fa.ntiles + 7) / 8 * 8;
        fa.trunc = a.trunc;
        fa.cnt = dcnt.as<uint16_t>();
        fa.rel = drel.as<int32_t>();
        fa.flags = dflag.as<int32_t>();
        TT_CHECK(fa.nblocks < ((int64_t)1 << 31), TTDFT_EBADARG,
                 "assemble_hloc: %lld workgroups exceed the grid limit",
                 (long long)fa.nblocks);
    }
    auto run_pass = [&](int write, const int64_t *drowptr, int32_t *dcol, double *dval) -> int {
        if (fused) {
            FusedArgs f = fa;
            f.rowptr = drowptr;
            f.col = dcol;
            f.val = dval;
            auto kern = write ? hloc_fused_kernel<1> : hloc_fused_kernel<0>;
            hipLaunchKernelGGL(kern, dim3((unsigned)f.nblocks), dim3(256), 0, ctx->stream, f);
            TT_LAUNCH_CHECK();
            return TTDFT_OK;
        }
        for (int64_t g0 = 0; g0 < ngroups; g0 += gc) {
            const int64_t ng = std::min<int64_t>(gc, ngroups - g0);
            TT_BLAS(rocblas_dgemm(ctx->blas, rocblas_operation_none, rocblas_operation_transpose,
                                  (rocblas_int)rz2, (rocblas_int)(rx * ng), rvz, &one,
                                  dPZ.as<double>(), (rocblas_int)rz2, dB.as<double>() + rx * g0,
                                  (rocblas_int)ldb, &zero, dC.as<double>(), (rocblas_int)rz2));
            const unsigned nb = blocks_for(rz2 * ng, 256);
            auto kern = write ? hloc_compact_kernel<1> : hloc_compact_kernel<0>;
            hipLaunchKernelGGL(kern, dim3(nb), dim3(256), 0, ctx->stream, dC.as<double>(), g0, ng,
                               rx, ry, rz, dX.as<double>(), dDX.as<double>(), dY.as<double>(),
                               dDY.as<double>(), dZ.as<double>(), dDZ.as<double>(), a.trunc,
                               dcnt.as<uint16_t>(), drel.as<int32_t>(), drowptr, dcol, dval,
                               dflag.as<int32_t>());
            TT_LAUNCH_CHECK();
        }
        return TTDFT_OK;
    };
    // algorithmic work of one pass: the three products and the compaction
    // (kinetic term and sum, 7 flop per entry)
    const double n2 = (double)n * (double)n;
    const double flops = 2.0 * rx2 * rvx * rvy * rvz + 2.0 * ntup * rvy * rvz +
                         2.0 * n2 * rvz + 7.0 * n2;  // same count for both paths
    PendingTiming tm;
    TT_TRY(timing_begin(ctx, TTDFT_KSTAT_ASSEMBLY, &tm));
    TT_BLAS(rocblas_dgemm(ctx->blas, rocblas_operation_none, rocblas_operation_none,
                          (rocblas_int)rx2, rvy * rvz, rvx, &one, dPX.as<double>(),
                          (rocblas_int)rx2, dG.as<double>(), rvx, &zero, dA.as<double>(),
                          (rocblas_int)rx2));
    TT_BLAS(rocblas_dgemm_strided_batched(
        ctx->blas, rocblas_operation_none, rocblas_operation_transpose, (rocblas_int)rx2,
        (rocblas_int)ry2, rvy, &one, dA.as<double>(), (rocblas_int)rx2, rx2 * rvy,
        dPY.as<double>(), (rocblas_int)ry2, 0, &zero, dB.as<double>(), (rocblas_int)rx2, ldb,
        rvz));
    if (fused) {
        hipLaunchKernelGGL(kinetic_planes_kernel, dim3(blocks_for(ntup, 256)), dim3(256), 0,
                           ctx->stream, dX.as<double>(), dDX.as<double>(), dY.as<double>(),
                           dDY.as<double>(), rx, ry, ntup, ldb,
                           dB.as<double>() + (int64_t)rvz * ldb);
        TT_LAUNCH_CHECK();
    }
    TT_TRY(run_pass(0, nullptr, nullptr, nullptr));
    const size_t lds = sizeof(int32_t) * ((size_t)ry * rz + 256);
    hipLaunchKernelGGL(hloc_row_scan_kernel, dim3((unsigned)n), dim3(256), lds, ctx->stream,
                       dcnt.as<uint16_t>(), ry, rz, drel.as<int32_t>(), drowlen.as<int64_t>());
    TT_LAUNCH_CHECK();
    TT_TRY(timing_end(ctx, &tm, 14.0 * n * ry * rz, flops));
    std::vector<int64_t> len(n);
    out.rp.assign(n + 1, 0);
    TT_HIP(hipMemcpyAsync(len.data(), drowlen.p, sizeof(int64_t) * n, hipMemcpyDeviceToHost,
                          ctx->stream));
    TT_HIP(hipStreamSynchronize(ctx->stream));
    for (int64_t i = 0; i < n; ++i) out.rp[i + 1] = out.rp[i] + len[i];
    out.nnz = out.rp[n];
    if (!fill) return TTDFT_OK;
    TT_TRY(upload_to(ctx, out.rowptr, out.rp.data(), sizeof(int64_t) * (n + 1)));
    TT_TRY(out.col.alloc(ctx, sizeof(int32_t) * out.nnz));
    TT_TRY(out.val.alloc(ctx, sizeof(double) * out.nnz));
    TT_TRY(timing_begin(ctx, TTDFT_KSTAT_ASSEMBLY, &tm));
    TT_TRY(run_pass(1, out.rowptr.as<int64_t>(), out.col.as<int32_t>(), out.val.as<double>()));
    TT_TRY(timing_end(ctx, &tm, 12.0 * out.nnz + 8.0 * (n + 1), flops));
    hipLaunchKernelGGL(band_extent_kernel, dim3((unsigned)((n + 3) / 4)), dim3(256), 0,
                       ctx->stream, out.rowptr.as<int64_t>(), out.col.as<int32_t>(), n, rx, ry,
                       dflag.as<int32_t>() + 1);
    TT_LAUNCH_CHECK();
    int32_t flags[4];
    TT_HIP(hipMemcpyAsync(flags, dflag.p, sizeof(flags), hipMemcpyDeviceToHost, ctx->stream));
    TT_HIP(hipStreamSynchronize(ctx->stream));
    TT_CHECK(!flags[0], TTDFT_ELIB, "assemble_hloc: count and write passes disagree");
    for (int d = 0; d < 3; ++d) ext[d] = flags[1 + d];
    return TTDFT_OK;
}

int check_proj(int32_t p, const int32_t *ranks, int32_t rx, int32_t ry, int32_t rz,
               const double *cores, const double *mx, const double *my, const double *mz,
               double truncation) {
    TT_CHECK(p >= 0 && rx > 0 && ry > 0 && rz > 0, TTDFT_EBADARG,
             "assemble_projectors: bad sizes");
    TT_CHECK(p == 0 || (ranks && cores && mx && my && mz), TTDFT_EBADARG,
             "assemble_projectors: NULL input");
    TT_CHECK(truncation >= 0.0, TTDFT_EBADARG, "assemble_projectors: truncation < 0");
    TT_CHECK((int64_t)rx * ry * rz < ((int64_t)1 << 31), TTDFT_EBADARG,
             "assemble_projectors: n too large");
    for (int i = 0; i < p; ++i)
        TT_CHECK(ranks[i] > 0, TTDFT_EBADARG, "assemble_projectors: rank of row %d < 1", i);
    return TTDFT_OK;
}

// Projector rows (p > 0): row pointer always, device CSR when fill.
int proj_run(ttdft_ctx ctx, int32_t p, const int32_t *ranks, int32_t rx, int32_t ry, int32_t rz,
             const double *cores, const double *mx, const double *my, const double *mz,
             double truncation, bool fill, DevCsr &out) {
    const int64_t n = (int64_t)rx * ry * rz, nxy = (int64_t)rx * ry;
    int rmax = 0;
    for (int i = 0; i < p; ++i) rmax = std::max(rmax, ranks[i]);
    // host: W_i[c][x + rx y] = sum_{a,b} S_i[a,b,c] MX_i[x,a] MY_i[y,b] (small), and MZ
    // padded to rmax terms; zero terms past a row's rank
    std::vector<double> W((size_t)p * rmax * nxy, 0.0), MZ((size_t)p * rmax * rz, 0.0);
    {
        size_t oc = 0, ox = 0, oy = 0, oz = 0;
        for (int i = 0; i < p; ++i) {
            const int r = ranks[i];
            const double *S = cores + oc, *A = mx + ox, *B = my + oy, *C = mz + oz;
            std::vector<double> u((size_t)r * r * rx);  // [c][b][x]
            for (int c = 0; c < r; ++c)
                for (int b = 0; b < r; ++b)
                    for (int x = 0; x < rx; ++x) {
                        double s = 0.0;
                        for (int a = 0; a < r; ++a)
                            s += A[x + (size_t)rx * a] * S[a + (size_t)r * (b + (size_t)r * c)];
                        u[((size_t)c * r + b) * rx + x] = s;
                    }
            for (int c = 0; c < r; ++c) {
                for (int y = 0; y < ry; ++y)
                    for (int x = 0; x < rx; ++x) {
                        double s = 0.0;
                        for (int b = 0; b < r; ++b)
                            s += B[y + (size_t)ry * b] * u[((size_t)c * r + b) * rx + x];
                        W[((size_t)i * rmax + c) * nxy + x + (size_t)rx * y] = s;
                    }
                for (int z = 0; z < rz; ++z)
                    MZ[((size_t)i * rmax + c) * rz + z] = C[z + (size_t)rz * c];
            }
            oc += (size_t)r * r * r;
            ox += (size_t)rx * r;
            oy += (size_t)ry * r;
            oz += (size_t)rz * r;
        }
    }
    DevBuf dW, dMZ;
    TT_TRY(upload_to(ctx, dW, W.data(), sizeof(double) * W.size()));
    TT_TRY(upload_to(ctx, dMZ, MZ.data(), sizeof(double) * MZ.size()));
    const int nchunk = (int)((n + 4095) / 4096);
    // rows in batches of about 1 GB of dense values
    int64_t rb = std::max<int64_t>(1, std::min<int64_t>(p, ((int64_t)1 << 27) / n));
    rb = env_cap("TTDFT_ASSEMBLE_ROWS", rb);
    DevBuf dD, dcnt, dstart;
    TT_TRY(dD.alloc(ctx, sizeof(double) * rb * n));
    TT_TRY(dcnt.alloc(ctx, sizeof(int32_t) * (size_t)p * nchunk));
    TT_TRY(dstart.alloc(ctx, sizeof(int64_t) * (size_t)p * nchunk));
    auto run_pass = [&](int write, int32_t *dcol, double *dval) -> int {
        for (int64_t i0 = 0; i0 < p; i0 += rb) {
            const int64_t nr = std::min<int64_t>(rb, p - i0);
            hipLaunchKernelGGL(proj_dense_kernel, dim3(blocks_for(nr * n, 256)), dim3(256), 0,
                               ctx->stream, dW.as<double>() + i0 * rmax * nxy,
                               dMZ.as<double>() + i0 * rmax * rz, nr, rmax, rx, ry, rz,
                               dD.as<double>());
            TT_LAUNCH_CHECK();
            const unsigned nb = (unsigned)(nr * nchunk);
            if (write)
                hipLaunchKernelGGL(dense_compact_kernel<1>, dim3(nb), dim3(256), 0, ctx->stream,
                                   dD.as<double>(), n, nchunk, truncation, nullptr,
                                   dstart.as<int64_t>() + i0 * nchunk, dcol, dval);
            else
                hipLaunchKernelGGL(dense_compact_kernel<0>, dim3(nb), dim3(256), 0, ctx->stream,
                                   dD.as<double>(), n, nchunk, truncation,
                                   dcnt.as<int32_t>() + i0 * nchunk, nullptr, nullptr, nullptr);
            TT_LAUNCH_CHECK();
        }
        return TTDFT_OK;
    };
    TT_TRY(run_pass(0, nullptr, nullptr));
    std::vector<int32_t> cnt((size_t)p * nchunk);
    TT_HIP(hipMemcpyAsync(cnt.data(), dcnt.p, sizeof(int32_t) * cnt.size(),
                          hipMemcpyDeviceToHost, ctx->stream));
    TT_HIP(hipStreamSynchronize(ctx->stream));
    std::vector<int64_t> start(cnt.size());
    out.rp.assign(p + 1, 0);
    int64_t acc = 0;
    for (int i = 0; i < p; ++i) {
        for (int c = 0; c < nchunk; ++c) {
            start[(size_t)i * nchunk + c] = acc;
            acc += cnt[(size_t)i * nchunk + c];
        }
        out.rp[i + 1] = acc;
    }
    out.nnz = acc;
    if (!fill) return TTDFT_OK;
    TT_HIP(hipMemcpyAsync(dstart.p, start.data(), sizeof(int64_t) * start.size(),
                          hipMemcpyHostToDevice, ctx->stream));
    TT_TRY(upload_to(ctx, out.rowptr, out.rp.data(), sizeof(int64_t) * (p + 1)));
    TT_TRY(out.col.alloc(ctx, sizeof(int32_t) * acc));
    TT_TRY(out.val.alloc(ctx, sizeof(double) * acc));
    TT_TRY(run_pass(1, out.col.as<int32_t>(), out.val.as<double>()));
    TT_HIP(hipStreamSynchronize(ctx->stream));  // start[] is read by the copy above
    return TTDFT_OK;
}

int download_csr(ttdft_ctx ctx, const DevCsr &c, int64_t rows, int64_t *rowptr, int32_t *col,
                 double *val) {
    std::copy(c.rp.begin(), c.rp.begin() + rows + 1, rowptr);
    if (c.nnz) {
        TT_HIP(hipMemcpyAsync(col, c.col.p, sizeof(int32_t) * c.nnz, hipMemcpyDeviceToHost,
                              ctx->stream));
        TT_HIP(hipMemcpyAsync(val, c.val.p, sizeof(double) * c.nnz, hipMemcpyDeviceToHost,
                              ctx->stream));
    }
    TT_HIP(hipStreamSynchronize(ctx->stream));
    return TTDFT_OK;
}

}  // namespace
}  // namespace ttdft

using namespace ttdft;

extern "C" {

int ttdft_assemble_hloc(ttdft_ctx ctx, int32_t rx, int32_t ry, int32_t rz, int32_t rvx,
                        int32_t rvy, int32_t rvz, const double *X, const double *DX,
                        const double *Y, const double *DY, const double *Z, const double *DZ,
                        const double *PX, const double *PY, const double *PZ, const double *G,
                        double truncation, int64_t *nnz, int64_t *rowptr, int32_t *col,
                        double *val) {
    clear_error();
    TT_CHECK(ctx, TTDFT_EBADARG, "ctx is NULL");
    TT_CHECK(nnz, TTDFT_EBADARG, "assemble_hloc: nnz is NULL");
    const HlocIn a{rx, ry, rz, rvx, rvy, rvz, X, DX, Y, DY, Z, DZ, PX, PY, PZ, G, truncation};
    TT_TRY(check_hloc(a));
    TT_HIP(hipSetDevice(ctx->device));
    const bool fill = rowptr && col && val;
    DevCsr h;
    int ext[3];
    TT_TRY(hloc_run(ctx, a, fill, h, ext));
    *nnz = h.nnz;
    if (!rowptr) return TTDFT_OK;
    if (!fill) {
        std::copy(h.rp.begin(), h.rp.end(), rowptr);
        return TTDFT_OK;
    }
    return download_csr(ctx, h, (int64_t)rx * ry * rz, rowptr, col, val);
}

int ttdft_assemble_projectors(ttdft_ctx ctx, int32_t p, const int32_t *ranks, int32_t rx,
                              int32_t ry, int32_t rz, const double *cores, const double *mx,
                              const double *my, const double *mz, double truncation,
                              int64_t *nnz, int64_t *rowptr, int32_t *col, double *val) {
    clear_error();
    TT_CHECK(ctx, TTDFT_EBADARG, "ctx is NULL");
    TT_CHECK(nnz, TTDFT_EBADARG, "assemble_projectors: nnz is NULL");
    TT_TRY(check_proj(p, ranks, rx, ry, rz, cores, mx, my, mz, truncation));
    *nnz = 0;
    if (p == 0) {
        if (rowptr) rowptr[0] = 0;
        return TTDFT_OK;
    }
    TT_HIP(hipSetDevice(ctx->device));
    const bool fill = rowptr && col && val;
    DevCsr c;
    TT_TRY(proj_run(ctx, p, ranks, rx, ry, rz, cores, mx, my, mz, truncation, fill, c));
    *nnz = c.nnz;
    if (!rowptr) return TTDFT_OK;
    if (!fill) {
        std::copy(c.rp.begin(), c.rp.end(), rowptr);
        return TTDFT_OK;
    }
    return download_csr(ctx, c, p, rowptr, col, val);
}

static int ham_assemble(ttdft_ctx ctx, const HlocIn &a, int32_t p, const int32_t *ranks,
                        const double *cores, const double *mx, const double *my,
                        const double *mz, double c_truncation, const double *inv_c_lm,
                        int copies, ttdft_ham_s *H) {
    const int64_t n = (int64_t)a.rx * a.ry * a.rz;
    DevCsr h;
    int ext[3];
    TT_TRY(hloc_run(ctx, a, true, h, ext));
    H->n = n;
    H->nnz = h.nnz;
    H->copies = copies;
    for (int64_t i = 0; i < n; ++i)
        H->max_row_nnz = std::max<int32_t>(H->max_row_nnz, (int32_t)(h.rp[i + 1] - h.rp[i]));
    H->rowptr = h.rowptr.release<int64_t>();
    H->col = h.col.release<int32_t>();
    H->val64 = h.val.release<double>();
    if (copies & TTDFT_HAM_F32) {
        TT_TRY(dev_alloc(ctx, (void **)&H->val32, sizeof(float) * std::max<int64_t>(1, h.nnz)));
        if (h.nnz) {
            hipLaunchKernelGGL(to_f32_kernel, dim3(blocks_for(h.nnz, 256)), dim3(256), 0,
                               ctx->stream, H->val64, H->val32, h.nnz);
            TT_LAUNCH_CHECK();
        }
    }
    if (ext[0] >= 0) TT_TRY(build_band_device(ctx, H, a.rx, a.ry, a.rz, ext[0], ext[1], ext[2],
                                              copies));
    TT_TRY(maybe_build_dense(ctx, H, copies));
    if (!(copies & TTDFT_HAM_F64)) {
        TT_HIP(hipStreamSynchronize(ctx->stream));
        dev_release(ctx, H->val64);
        H->val64 = nullptr;
    }
    H->p = p;
    if (p == 0) return finish_dense(ctx, H, copies);
    DevCsr c;
    TT_TRY(proj_run(ctx, p, ranks, a.rx, a.ry, a.rz, cores, mx, my, mz, c_truncation, true, c));
    H->nnz_c = c.nnz;
    DevBuf dinv, ctlen;
    TT_TRY(upload_to(ctx, dinv, inv_c_lm, sizeof(double) * p));
    const size_t nc = (size_t)std::max<int64_t>(1, c.nnz);
    if (copies & TTDFT_HAM_F64) {
        TT_TRY(dev_alloc(ctx, (void **)&H->c_val64, sizeof(double) * nc));
        TT_TRY(dev_alloc(ctx, (void **)&H->ct_val64, sizeof(double) * nc));
    }
    if (copies & TTDFT_HAM_F32) {
        TT_TRY(dev_alloc(ctx, (void **)&H->c_val32, sizeof(float) * nc));
        TT_TRY(dev_alloc(ctx, (void **)&H->ct_val32, sizeof(float) * nc));
    }
    TT_TRY(dev_alloc(ctx, (void **)&H->ct_col, sizeof(int32_t) * nc));
    // D C (row-scaled) and C^T (unscaled), ProjectHamiltonianSparse.cc:316-325
    hipLaunchKernelGGL(scale_rows_kernel, dim3((unsigned)p), dim3(256), 0, ctx->stream,
                       c.rowptr.as<int64_t>(), c.val.as<double>(), dinv.as<double>(),
                       H->c_val64, H->c_val32);
    TT_LAUNCH_CHECK();
    TT_TRY(ctlen.alloc(ctx, sizeof(int64_t) * n));
    const unsigned nbj = blocks_for(n, 256);
    hipLaunchKernelGGL(transpose_rows_kernel<0>, dim3(nbj), dim3(256), 0, ctx->stream,
                       c.rowptr.as<int64_t>(), c.col.as<int32_t>(), c.val.as<double>(), p, n,
                       ctlen.as<int64_t>(), nullptr, nullptr, nullptr, nullptr);
    TT_LAUNCH_CHECK();
    std::vector<int64_t> ctp(n + 1, 0);
    TT_HIP(hipMemcpyAsync(ctp.data() + 1, ctlen.p, sizeof(int64_t) * n, hipMemcpyDeviceToHost,
                          ctx->stream));
    TT_HIP(hipStreamSynchronize(ctx->stream));
    for (int64_t j = 0; j < n; ++j) ctp[j + 1] += ctp[j];
    TT_CHECK(ctp[n] == c.nnz, TTDFT_ELIB, "ham_assemble: projector transpose lost entries");
    TT_TRY(dev_alloc(ctx, (void **)&H->ct_rowptr, sizeof(int64_t) * (n + 1)));
    TT_HIP(hipMemcpyAsync(H->ct_rowptr, ctp.data(), sizeof(int64_t) * (n + 1),
                          hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(transpose_rows_kernel<1>, dim3(nbj), dim3(256), 0, ctx->stream,
                       c.rowptr.as<int64_t>(), c.col.as<int32_t>(), c.val.as<double>(), p, n,
                       nullptr, H->ct_rowptr, H->ct_col, H->ct_val64, H->ct_val32);
    TT_LAUNCH_CHECK();
    TT_HIP(hipStreamSynchronize(ctx->stream));  // ctp is read by the copy above
    H->c_rowptr = c.rowptr.release<int64_t>();
    H->c_col = c.col.release<int32_t>();
    return finish_dense(ctx, H, copies);
}

int ttdft_ham_assemble_tucker(ttdft_ctx ctx, int32_t rx, int32_t ry, int32_t rz, int32_t rvx,
                              int32_t rvy, int32_t rvz, const double *X, const double *DX,
                              const double *Y, const double *DY, const double *Z,
                              const double *DZ, const double *PX, const double *PY,
                              const double *PZ, const double *G, double truncation, int32_t p,
                              const int32_t *ranks, const double *cores, const double *mx,
                              const double *my, const double *mz, double c_truncation,
                              const double *inv_c_lm, int copies, ttdft_ham *out) {
    clear_error();
    TT_CHECK(ctx && out, TTDFT_EBADARG, "ham_assemble_tucker: NULL ctx or out");
    *out = nullptr;
    TT_CHECK(copies & (TTDFT_HAM_F64 | TTDFT_HAM_F32), TTDFT_EBADARG,
             "ham_assemble_tucker: no device copy requested");
    const HlocIn a{rx, ry, rz, rvx, rvy, rvz, X, DX, Y, DY, Z, DZ, PX, PY, PZ, G, truncation};
    TT_TRY(check_hloc(a));
    TT_TRY(check_proj(p, ranks, rx, ry, rz, cores, mx, my, mz, c_truncation));
    TT_CHECK(p == 0 || inv_c_lm, TTDFT_EBADARG, "ham_assemble_tucker: inv_c_lm is NULL");
    TT_HIP(hipSetDevice(ctx->device));
    ttdft_ham H = new ttdft_ham_s();
    H->ctx = ctx;
    ctx_attach(ctx);
    const int rc = ham_assemble(ctx, a, p, ranks, cores, mx, my, mz, c_truncation, inv_c_lm,
                                copies, H);
    if (rc != TTDFT_OK) {
        const std::string msg = ttdft_last_error();
        ttdft_ham_free(H);
        set_error("%s", msg.c_str());
        return rc;
    }
    *out = H;
    return TTDFT_OK;
}

}  // extern "C"
