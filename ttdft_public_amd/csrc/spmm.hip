// Hamiltonian application kernels: H X = H_loc X + C^T (D C X), with the
// three-term Chebyshev recurrence fused into the SpMM epilogue.
//
// Data layout: X, Y are basis-major blocks (X[i*ld + j], state j of basis
// function i), so one H_loc row reads whole contiguous state rows of X and a
// 64-wide wavefront loads 64 consecutive states of one basis function in one
// coalesced access.  One wavefront owns one H_loc row and all k states
// (NC = ceil(k/64) accumulator chunks per lane), so H_loc is streamed from
// HBM exactly once per application; the row's (value, column) pairs are read
// 64 at a time with one coalesced load and broadcast with v_readlane.
#include "internal.h"

namespace ttdft {
namespace {

__device__ __forceinline__ double bcast(double v, int l) {
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ float bcast(float v, int l) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}
__device__ __forceinline__ int bcast(int v, int l) { return __builtin_amdgcn_readlane(v, l); }

template <typename T, int NC>
__global__ __launch_bounds__(256) void spmm_fused_kernel(
    int64_t n, const int64_t *__restrict__ rowptr, const int32_t *__restrict__ col,
    const T *__restrict__ val, const T *__restrict__ X, int64_t ldx,
    const int64_t *__restrict__ ct_rowptr, const int32_t *__restrict__ ct_col,
    const T *__restrict__ ct_val, const T *__restrict__ W, int32_t ldw, T *Y, int64_t ldy,
    const T *Z, int64_t ldz, int32_t k, T alpha, T beta, T delta, int use_beta,
    int use_delta) {
    const int lane = threadIdx.x & 63;
    const int64_t row = (int64_t)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (row >= n) return;
    T acc[NC];
#pragma unroll
    for (int t = 0; t < NC; ++t) acc[t] = T(0);

    // U nonzeros per step: their X rows are loaded together before the
    // multiply-adds, so a wave keeps U gathers in flight instead of one (the
    // adds still run in the row's order).  Lanes past the chunk's count hold
    // value 0 and column 0, so the last step reads row 0 and adds zeros.
    constexpr int U = NC <= 2 ? 8 : (NC <= 4 ? 4 : 2);
    const int64_t e0 = rowptr[row], e1 = rowptr[row + 1];
    for (int64_t base = e0; base < e1; base += 64) {
        const int cnt = (int)((e1 - base) < 64 ? (e1 - base) : 64);
        T v = T(0);
        int c = 0;
        if (lane < cnt) {
            v = val[base + lane];
            c = col[base + lane];
        }
        const int steps = (cnt + U - 1) / U;
        for (int s = 0; s < steps; ++s) {
            T a[U], xv[U][NC];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                a[u] = bcast(v, s * U + u);
                const T *xr = X + (int64_t)bcast(c, s * U + u) * ldx + lane;
#pragma unroll
                for (int t = 0; t < NC; ++t) xv[u][t] = lane + 64 * t < k ? xr[64 * t] : T(0);
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
#pragma unroll
                for (int t = 0; t < NC; ++t) acc[t] += a[u] * xv[u][t];
        }
    }
    if (ct_rowptr != nullptr) {
        const int64_t f0 = ct_rowptr[row], f1 = ct_rowptr[row + 1];
        for (int64_t e = f0; e < f1; ++e) {
            const T a = ct_val[e];
            const T *wr = W + (int64_t)ct_col[e] * ldw + lane;
#pragma unroll
            for (int t = 0; t < NC; ++t) {
                if (lane + 64 * t < k) acc[t] += a * wr[64 * t];
            }
        }
    }
    T *yr = Y + row * ldy + lane;
    const T *xr = X + row * ldx + lane;
    const T *zr = Z + row * ldz + lane;
#pragma unroll
    for (int t = 0; t < NC; ++t) {
        if (lane + 64 * t < k) {
            T y = acc[t] * alpha;
            if (use_beta) y = y + beta * xr[64 * t];
            if (use_delta) y = y + delta * zr[64 * t];
            yr[64 * t] = y;
        }
    }
}

// Single-vector H_loc application for the Lanczos bound estimator:
// y = H_loc x + C^T w.  One wavefront per row; lanes stride over the row's
// nonzeros (coalesced value/index loads, gathered x) and the partial sums are
// combined with a wavefront shuffle reduction.
__global__ __launch_bounds__(256) void spmv_kernel(
    int64_t n, const int64_t *__restrict__ rowptr, const int32_t *__restrict__ col,
    const double *__restrict__ val, const double *__restrict__ x,
    const int64_t *__restrict__ ct_rowptr, const int32_t *__restrict__ ct_col,
    const double *__restrict__ ct_val, const double *__restrict__ w, double *__restrict__ y) {
    const int lane = threadIdx.x & 63;
    const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= n) return;
    double s = 0.0;
    for (int64_t e = rowptr[row] + lane; e < rowptr[row + 1]; e += 64) s += val[e] * x[col[e]];
    if (ct_rowptr)
        for (int64_t e = ct_rowptr[row] + lane; e < ct_rowptr[row + 1]; e += 64)
            s += ct_val[e] * w[ct_col[e]];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (lane == 0) y[row] = s;
}

// W[r][j] = sum_e (D C)[r][e] X[col_e][j].  A projector row has a few
// thousand nonzeros and there are only a few hundred rows, so one block per
// row would leave most CUs idle: block (r, panel, split) takes 64 states and
// one of nsplit slices of row r's nonzeros (4 wavefronts over the slice,
// summed through LDS) and writes a partial sum; nl_reduce_kernel adds the
// nsplit partials in a fixed order (deterministic).  X rows are gathered 8 at
// a time so their loads are in flight together.
template <typename T>
__global__ __launch_bounds__(256) void nl_project_kernel(
    int32_t p, int32_t npanel, int32_t nsplit, const int64_t *__restrict__ c_rowptr,
    const int32_t *__restrict__ c_col, const T *__restrict__ c_val, const T *__restrict__ X,
    int64_t ldx, int32_t k, T *__restrict__ Wp) {
    __shared__ T red[4][64];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    int b = blockIdx.x;
    const int split = b % nsplit;
    b /= nsplit;
    const int panel = b % npanel;
    const int r = b / npanel;
    const int c = panel * 64 + lane;
    const int cc = c < k ? c : k - 1;  // clamped: every lane loads a valid element
    const int64_t e0 = c_rowptr[r], len = c_rowptr[r + 1] - e0;
    const int64_t lo = e0 + len * split / nsplit, hi = e0 + len * (split + 1) / nsplit;
    T acc = T(0);
    for (int64_t base = lo + 64 * wave; base < hi; base += 256) {
        const int cnt = (int)((hi - base) < 64 ? (hi - base) : 64);
        T v = T(0);
        int col = 0;
        if (lane < cnt) {
            v = c_val[base + lane];
            col = c_col[base + lane];
        }
        int j = 0;
        for (; j + 8 <= cnt; j += 8) {
            T x[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) x[u] = X[(int64_t)bcast(col, j + u) * ldx + cc];
#pragma unroll
            for (int u = 0; u < 8; ++u) acc += bcast(v, j + u) * x[u];
        }
        for (; j < cnt; ++j) acc += bcast(v, j) * X[(int64_t)bcast(col, j) * ldx + cc];
    }
    red[wave][lane] = acc;
    __syncthreads();
    if (wave == 0 && c < k)
        Wp[((int64_t)split * p + r) * k + c] =
            (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
}

template <typename T>
__global__ void nl_reduce_kernel(int64_t pk, int32_t nsplit, const T *__restrict__ Wp,
                                 T *__restrict__ W) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < pk;
         e += (int64_t)gridDim.x * blockDim.x) {
        T s = Wp[e];
        for (int q = 1; q < nsplit; ++q) s += Wp[(int64_t)q * pk + e];
        W[e] = s;
    }
}

// slices per projector row: about 4096 blocks in all, at least 256 nonzeros a slice
int nl_splits(const ttdft_ham_s *H, int32_t k) {
    if (H->p == 0) return 1;
    const int64_t npanel = (k + 63) / 64;
    int64_t ns = (4096 + H->p * npanel - 1) / (H->p * npanel);
    const int64_t cap = H->nnz_c / H->p / 256;
    if (ns > cap) ns = cap;
    if (ns < 1) ns = 1;
    if (ns > 64) ns = 64;
    return (int)ns;
}

template <typename T>
int nl_launch(ttdft_ctx ctx, const ttdft_ham_s *H, const T *X, int64_t ldx, int32_t k,
              const T *c_val, T *W) {
    const int npanel = (k + 63) / 64;
    const int ns = nl_splits(H, k);
    T *Wp = ns > 1 ? W + (int64_t)H->p * k : W;
    hipLaunchKernelGGL(nl_project_kernel<T>, dim3((unsigned)((int64_t)H->p * npanel * ns)),
                       dim3(256), 0, ctx->stream, H->p, npanel, ns, H->c_rowptr, H->c_col, c_val,
                       X, ldx, k, Wp);
    TT_LAUNCH_CHECK();
    if (ns > 1) {
        const int64_t pk = (int64_t)H->p * k;
        int64_t blocks = (pk + 255) / 256;
        if (blocks > 4096) blocks = 4096;
        hipLaunchKernelGGL(nl_reduce_kernel<T>, dim3((unsigned)blocks), dim3(256), 0, ctx->stream,
                           pk, ns, Wp, W);
        TT_LAUNCH_CHECK();
    }
    return TTDFT_OK;
}

// D[row][col_e] = val_e (D zeroed beforehand, leading dimension ld); one
// wavefront per row
template <typename T>
__global__ __launch_bounds__(256) void csr_to_dense_kernel(int64_t n, int64_t ld,
                                                           const int64_t *__restrict__ rowptr,
                                                           const int32_t *__restrict__ col,
                                                           const T *__restrict__ val,
                                                           T *__restrict__ D) {
    const int lane = threadIdx.x & 63;
    const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= n) return;
    for (int64_t e = rowptr[row] + lane; e < rowptr[row + 1]; e += 64)
        D[row * ld + col[e]] = val[e];
}

// Dense H_loc X with the recurrence epilogue, on matrix cores: a workgroup
// owns 16 rows of Y and 16 JT states; its 8 waves split the inner dimension l
// in eight and their partial tiles are summed through LDS.
//
// Storage: row-major, leading dimension ldh = n rounded up to 32, rows padded
// to a multiple of 16, padding zero.  Per 32-wide step of l, lane (i, q)
// (i = lane % 16, q = lane / 16) loads the 8 consecutive entries
// H[i0 + i][l0 + 8q .. 8q + 7] with wide loads, so each row of the tile is one
// contiguous 32-entry run, and the 8 matching rows of X; MFMA step u
// (u = 0..7) then pairs l = l0 + 8q + u in both operands.  The next step's
// loads are issued before this step's products (two register buffers).  X
// rows past n are read clamped to row n - 1 (finite values times the zero
// padding) and states past k clamped to k - 1 (never stored), so no load is
// predicated.
//
// The product is D[j][i] = sum_l X[l][j] H[i][l]: A operand X^T (16 states x
// 4 l, lane j + 16 kk), B operand H^T (4 l x 16 rows, lane i + 16 kk).  FP64
// storage runs on 16x16x4 FP64 MFMA (element v of lane L is D[L/16 + 4v][L%16])
// and FP32 storage on 16x16x4 FP32 MFMA with FP32 accumulation (element v of
// lane L is D[4 (L/16) + v][L%16]).  Then
// Y[i][j] = alpha (HX[i][j] + sum_e ct[e] W[ct_col_e][j]) (+ beta X[i][j]) (+ delta Z[i][j]).
typedef double dv4 __attribute__((ext_vector_type(4)));
typedef float fv4 __attribute__((ext_vector_type(4)));
constexpr int kDenseWaves = 8;

template <typename T> struct DenseMma;
template <> struct DenseMma<double> {
    typedef dv4 acc_t;
    static __device__ __forceinline__ dv4 step(double a, double b, dv4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int state(int q, int v) { return q + 4 * v; }
};
template <> struct DenseMma<float> {
    typedef fv4 acc_t;
    static __device__ __forceinline__ fv4 step(float a, float b, fv4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int state(int q, int v) { return 4 * q + v; }
};

// Stage depth U (MFMA steps per stage, U consecutive entries per lane) and
// accumulator sets NA per state-tile count: up to 64 states, 8-deep stages and
// two accumulator sets (alternate steps); 80 or 96 states, 4-deep stages and
// one set, which keeps the two stages in registers.
template <int JT> struct DenseCfg {
    static constexpr int U = JT <= 4 ? 8 : 4;
    static constexpr int NA = JT <= 4 ? 2 : 1;
};

template <typename T, int JT> struct DenseStage {
    static constexpr int U = DenseCfg<JT>::U;
    T h[U];
    T x[U][JT];
};

template <typename T, int JT, bool CONTIG>
__device__ __forceinline__ void dense_load(DenseStage<T, JT> &st, const T *hrow, const T *X,
                                           int64_t ldx, int64_t n, const int *cidx, int64_t l0,
                                           int q) {
    constexpr int U = DenseCfg<JT>::U;
    // U consecutive entries (16-64 B), aligned by the padding
    typedef T vU __attribute__((ext_vector_type(U)));
    typedef T v2 __attribute__((ext_vector_type(2)));
    const vU hv = *reinterpret_cast<const vU *>(hrow + l0);
#pragma unroll
    for (int u = 0; u < U; ++u) st.h[u] = hv[u];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        int64_t l = l0 + U * q + u;
        l = l < n ? l : n - 1;
        const T *xr = X + l * ldx;
        if constexpr (CONTIG) {
            // the lane's JT states are consecutive from cidx[0] (even, with
            // ldx even): pairs in one load each
#pragma unroll
            for (int jt = 0; jt < JT; jt += 2) {
                const v2 w = *reinterpret_cast<const v2 *>(xr + cidx[0] + jt);
                st.x[u][jt] = w[0];
                st.x[u][jt + 1] = w[1];
            }
        } else {
#pragma unroll
            for (int jt = 0; jt < JT; ++jt) st.x[u][jt] = xr[cidx[jt]];
        }
    }
}

template <typename T, int JT>
__device__ __forceinline__ void dense_mma(
    const DenseStage<T, JT> &st, typename DenseMma<T>::acc_t (&acc)[DenseCfg<JT>::NA][JT]) {
    constexpr int U = DenseCfg<JT>::U, NA = DenseCfg<JT>::NA;
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
        for (int jt = 0; jt < JT; ++jt)
            acc[u % NA][jt] = DenseMma<T>::step(st.x[u][jt], st.h[u], acc[u % NA][jt]);
}

template <typename T, int JT, bool CONTIG>
__global__ __launch_bounds__(512) void dense_hx_kernel(
    int64_t n, int64_t ldh, int32_t k, const T *__restrict__ H, const T *X, int64_t ldx,
    const int64_t *__restrict__ ct_rowptr, const int32_t *__restrict__ ct_col,
    const T *__restrict__ ct_val, const T *__restrict__ W, T *Y, int64_t ldy, const T *Z,
    int64_t ldz, double alpha, double beta, double delta, int use_beta, int use_delta) {
    typedef typename DenseMma<T>::acc_t acc_t;
    constexpr int U = DenseCfg<JT>::U, NA = DenseCfg<JT>::NA, SW = 4 * U;  // step width in l
    __shared__ T red[kDenseWaves][16 * JT][17];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int li = lane & 15, q = lane >> 4;
    const int64_t i0 = (int64_t)blockIdx.x * 16;
    const int c0 = blockIdx.y * 16 * JT;
    const int64_t part = ((ldh + kDenseWaves - 1) / kDenseWaves + SW - 1) / SW * SW;
    const int64_t lb = wave * part, le = lb + part < ldh ? lb + part : ldh;
    const T *hrow = H + (i0 + li) * ldh + U * q;
    // state of tile jt, tile row i: c0 + 16 jt + i, or with CONTIG (every tile
    // full, ldx even) c0 + JT i + jt, so a lane's JT states are consecutive
    int cidx[JT];
#pragma unroll
    for (int jt = 0; jt < JT; ++jt) {
        const int c = CONTIG ? c0 + JT * li + jt : c0 + 16 * jt + li;
        cidx[jt] = c < k ? c : k - 1;
    }
    acc_t acc[NA][JT];
#pragma unroll
    for (int e = 0; e < NA; ++e)
#pragma unroll
        for (int jt = 0; jt < JT; ++jt) acc[e][jt] = acc_t{0, 0, 0, 0};
    if (lb < le) {
        // two buffers, the loads of step s + 1 in flight during the products
        // of step s (the last pair reloads the final step into the idle
        // buffer; scheduling barriers keep the compiler from sinking the loads
        // into the products)
        DenseStage<T, JT> sa, sb;
        const int64_t nsteps = (le - lb) / SW;
        dense_load<T, JT, CONTIG>(sa, hrow, X, ldx, n, cidx, lb, q);
        int64_t st = 0;
        for (; st + 2 <= nsteps; st += 2) {
            dense_load<T, JT, CONTIG>(sb, hrow, X, ldx, n, cidx, lb + SW * (st + 1), q);
            __builtin_amdgcn_sched_barrier(0);
            dense_mma<T, JT>(sa, acc);
            __builtin_amdgcn_sched_barrier(0);
            const int64_t s2 = st + 2 < nsteps ? st + 2 : nsteps - 1;
            dense_load<T, JT, CONTIG>(sa, hrow, X, ldx, n, cidx, lb + SW * s2, q);
            __builtin_amdgcn_sched_barrier(0);
            dense_mma<T, JT>(sb, acc);
            __builtin_amdgcn_sched_barrier(0);
        }
        if (st < nsteps) dense_mma<T, JT>(sa, acc);
    }
#pragma unroll
    for (int jt = 0; jt < JT; ++jt)
#pragma unroll
        for (int v = 0; v < 4; ++v)
            red[wave][16 * jt + DenseMma<T>::state(q, v)][li] =
                NA == 2 ? acc[0][jt][v] + acc[NA - 1][jt][v] : acc[0][jt][v];
    __syncthreads();
    for (int e = t; e < 16 * 16 * JT; e += 512) {
        const int j = e % (16 * JT), ii = e / (16 * JT);
        const int64_t r = i0 + ii;
        const int c = CONTIG ? c0 + JT * (j % 16) + j / 16 : c0 + j;
        if (r >= n || c >= k) continue;
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < kDenseWaves; ++w) s += (double)red[w][j][ii];
        if (ct_rowptr)
            for (int64_t f = ct_rowptr[r]; f < ct_rowptr[r + 1]; ++f)
                s += (double)ct_val[f] * (double)W[(int64_t)ct_col[f] * k + c];
        double y = s * alpha;
        if (use_beta) y += beta * (double)X[r * ldx + c];
        if (use_delta) y += delta * (double)Z[r * ldz + c];
        Y[r * ldy + c] = (T)y;
    }
}

// Dense H_loc X with X staged through LDS, in two kernels.
// dense_hx_part_kernel: workgroup (p, row block, state block) owns 64 rows (a
// 16-row tile per wave, 4 waves) by 16 JT states over part p of the inner
// dimension l (nparts parts of whole 32-wide steps, so there are enough
// workgroups to fill the chip).  Per step the workgroup copies X[l0 .. l0 +
// 31][its states] into LDS once (coalesced loads of VX consecutive states,
// pairs in FP32 when aligned), so every X value read from L2 serves four
// row tiles; each wave loads its 16 rows of H as before
// (lane (i, q): H[i][l0 + 8q .. 8q + 7]) and MFMA step u pairs l = l0 + 8q + u
// in both operands, the X operand read from LDS (row 8q + u, 16 consecutive
// states per quarter-wave; the row pitch of 16 JT + 18 elements puts the
// quarter-waves' rows 8 apart on distinct banks: 8 pitches are 16 mod 64
// dwords in FP32, 32 in FP64).  One accumulator per state tile.
// The next step's X and H loads are issued before this step's products
// (pinned there by a scheduling barrier, and unconditional so that the wait
// before the products covers only their LDS reads); the X stage alternates
// between two buffers with one barrier per step.  The
// part's sums go to P[p][r][c] in FP64.  Rows past n read row n - 1 (never
// stored); X rows past n meet zero H padding.
// dense_combine_kernel: Y = alpha (sum_p P[p] + C^T W) (+ beta X) (+ delta Z),
// the parts summed in a fixed order.
template <typename T, int JT, int VX>
__global__ __launch_bounds__(256) void dense_hx_part_kernel(int64_t n, int64_t ldh, int32_t k,
                                                            const T *__restrict__ H,
                                                            const T *__restrict__ X,
                                                            int64_t ldx, int32_t nparts,
                                                            double *__restrict__ P) {
    typedef typename DenseMma<T>::acc_t acc_t;
    typedef T xv_t __attribute__((ext_vector_type(VX)));
    constexpr int NS = 16 * JT;                 // states per workgroup
    constexpr int XR = NS + 18;                 // LDS row pitch (elements)
    constexpr int NX = 32 * NS / (256 * VX);    // X vectors per thread per step
    static_assert(32 * NS % (256 * VX) == 0, "every thread stages the same count");
    typedef T v4 __attribute__((ext_vector_type(4)));
    __shared__ __align__(16) T xs[2][32 * XR];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int li = lane & 15, q = lane >> 4;
    const int64_t i0 = (int64_t)blockIdx.x * 64 + wave * 16;
    const int c0 = blockIdx.y * NS;
    const int part = blockIdx.z;
    const int64_t nst = ldh / 32;
    const int64_t sb = nst * part / nparts, se = nst * (part + 1) / nparts;
    const int64_t row = i0 + li < n ? i0 + li : n - 1;
    const T *hrow = H + row * ldh + 8 * q;
    // X vector e of this thread: row idx / (NS / VX) of the step, VX
    // consecutive states from col (the last vector of a ragged k is clamped
    // inside the row: states past k are never stored, and with VX = 2 k is even)
    auto xrow = [&](int e) { return (t + 256 * e) / (NS / VX); };
    auto xcol = [&](int e) { return VX * ((t + 256 * e) % (NS / VX)); };
    xv_t xv[NX];
    v4 ha[2], hb[2];
    auto load_x = [&](int64_t l0) {
#pragma unroll
        for (int e = 0; e < NX; ++e) {
            int64_t l = l0 + xrow(e);
            l = l < n ? l : n - 1;
            const int c = c0 + xcol(e) < k ? c0 + xcol(e) : k - VX;
            xv[e] = *reinterpret_cast<const xv_t *>(X + l * ldx + c);
        }
    };
    auto store_x = [&](int buf) {
#pragma unroll
        for (int e = 0; e < NX; ++e)
            *reinterpret_cast<xv_t *>(&xs[buf][xrow(e) * XR + xcol(e)]) = xv[e];
    };
    auto load_h = [&](int64_t l0, v4 (&h)[2]) {
        h[0] = *reinterpret_cast<const v4 *>(hrow + l0);
        h[1] = *reinterpret_cast<const v4 *>(hrow + l0 + 4);
    };
    // one accumulator per state tile: consecutive products go to different
    // tiles, so a tile's chain has JT - 1 other products between its steps
    acc_t acc[JT];
#pragma unroll
    for (int jt = 0; jt < JT; ++jt) acc[jt] = acc_t{0, 0, 0, 0};
    auto mma = [&](int buf, const v4 (&h)[2]) {
        const T *xb = xs[buf] + 8 * q * XR + li;
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const T hu = h[u >> 2][u & 3];
#pragma unroll
            for (int jt = 0; jt < JT; ++jt)
                acc[jt] = DenseMma<T>::step(xb[u * XR + 16 * jt], hu, acc[jt]);
        }
    };
    if (sb < se) {
        load_x(32 * sb);
        load_h(32 * sb, hb);
        store_x(0);
        // (ha is copied, not loaded, at the loop entry too: the wait the
        // copy needs then stays here and not before every step's products)
        ha[0] = hb[0];
        ha[1] = hb[1];
        __syncthreads();
        int buf = 0;
        for (int64_t st = sb; st < se; ++st) {
            // the last step reloads its own operands (unconditional loads keep
            // the wait before the products to this step's LDS reads)
            const int64_t nx = st + 1 < se ? st + 1 : st;
            load_x(32 * nx);
            load_h(32 * nx, hb);
            // keep the next step's loads ahead of this step's products (the
            // scheduler would otherwise sink the H loads past them into the
            // registers the products free, exposing their latency)
            __builtin_amdgcn_sched_barrier(0);
            mma(buf, ha);
            store_x(buf ^ 1);
            ha[0] = hb[0];
            ha[1] = hb[1];
            __syncthreads();
            buf ^= 1;
        }
    }
    // D[j][i]: element v of lane L holds state DenseMma::state(L/16, v), row L%16
    const int64_t r = i0 + li;
    if (r >= n) return;
    double *Pr = P + ((int64_t)part * n + r) * k;
#pragma unroll
    for (int jt = 0; jt < JT; ++jt)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int c = c0 + 16 * jt + DenseMma<T>::state(q, v);
            if (c < k) Pr[c] = (double)acc[jt][v];
        }
}

template <typename T>
__global__ __launch_bounds__(256) void dense_combine_kernel(
    int64_t n, int32_t k, int32_t nparts, const double *__restrict__ P, const T *X, int64_t ldx,
    const int64_t *__restrict__ ct_rowptr, const int32_t *__restrict__ ct_col,
    const T *__restrict__ ct_val, const T *__restrict__ W, T *Y, int64_t ldy, const T *Z,
    int64_t ldz, double alpha, double beta, double delta, int use_beta, int use_delta) {
    const int64_t total = n * k;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = e / k;
        const int c = (int)(e - r * k);
        double s = P[e];
        for (int p = 1; p < nparts; ++p) s += P[(int64_t)p * total + e];
        if (ct_rowptr)
            for (int64_t f = ct_rowptr[r]; f < ct_rowptr[r + 1]; ++f)
                s += (double)ct_val[f] * (double)W[(int64_t)ct_col[f] * k + c];
        double y = s * alpha;
        if (use_beta) y += beta * (double)X[r * ldx + c];
        if (use_delta) y += delta * (double)Z[r * ldz + c];
        Y[r * ldy + c] = (T)y;
    }
}

__global__ __launch_bounds__(256) void dense_to_f32_kernel(const double *__restrict__ s,
                                                           float *__restrict__ d, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        d[i] = (float)s[i];
}

unsigned grid_blocks(int64_t n) {
    int64_t b = (n + 255) / 256;
    return (unsigned)(b > 65536 ? 65536 : (b < 1 ? 1 : b));
}

template <typename T, int NC>
int spmm_dispatch_nc(ttdft_ctx ctx, const ttdft_ham_s *H, const T *X, int64_t ldx, const T *W,
                     T *Y, int64_t ldy, const T *Z, int64_t ldz, int32_t k, double alpha,
                     double beta, double delta, int use_beta, int use_delta,
                     const T *val, const T *ct_val) {
    const dim3 grid((unsigned)((H->n + 3) / 4));
    hipLaunchKernelGGL((spmm_fused_kernel<T, NC>), grid, dim3(256), 0, ctx->stream, H->n,
                       H->rowptr, H->col, val, X, ldx, H->p > 0 ? H->ct_rowptr : nullptr,
                       H->ct_col, ct_val, W, k, Y, ldy, Z, ldz, k, (T)alpha, (T)beta,
                       (T)delta, use_beta, use_delta);
    TT_LAUNCH_CHECK();
    return TTDFT_OK;
}

#define TT_NC_SWITCH(NCVAL, CALL)                                                 \
    switch (NCVAL) {                                                              \
    case 1: return CALL(1);                                                       \
    case 2: return CALL(2);                                                       \
    case 3: return CALL(3);                                                       \
    case 4: return CALL(4);                                                       \
    case 5: return CALL(5);                                                       \
    case 6: return CALL(6);                                                       \
    case 7: return CALL(7);                                                       \
    case 8: return CALL(8);                                                       \
    case 9: return CALL(9);                                                       \
    case 10: return CALL(10);                                                     \
    case 11: return CALL(11);                                                     \
    case 12: return CALL(12);                                                     \
    case 13: return CALL(13);                                                     \
    case 14: return CALL(14);                                                     \
    case 15: return CALL(15);                                                     \
    case 16: return CALL(16);                                                     \
    default: break;                                                               \
    }

}  // namespace

int64_t nl_workspace(const ttdft_ham_s *H, int32_t k) {
    const int ns = nl_splits(H, k);
    return (int64_t)H->p * k * (ns > 1 ? 1 + ns : 1);
}

int launch_nl_project(ttdft_ctx ctx, const ttdft_ham_s *H, const void *X, int64_t ldx,
                      int32_t k, int dtype, void *W) {
    if (H->p == 0) return TTDFT_OK;
    TT_CHECK(k >= 1, TTDFT_EBADARG, "k=%d < 1", k);
    if (dtype == TTDFT_F64) {
        TT_CHECK(H->c_val64, TTDFT_EBADARG, "no FP64 copy of the projectors uploaded");
        return nl_launch<double>(ctx, H, (const double *)X, ldx, k, H->c_val64, (double *)W);
    }
    TT_CHECK(H->c_val32, TTDFT_EBADARG, "no FP32 copy of the projectors uploaded");
    return nl_launch<float>(ctx, H, (const float *)X, ldx, k, H->c_val32, (float *)W);
}

int launch_spmm_fused(ttdft_ctx ctx, const ttdft_ham_s *H, const void *X, int64_t ldx,
                      const void *W, void *Y, int64_t ldy, const void *Z, int64_t ldz,
                      int32_t k, int dtype, double alpha, double beta, double delta,
                      int use_beta, int use_delta) {
    const int nc = (k + 63) / 64;
    TT_CHECK(nc >= 1 && nc <= 16, TTDFT_EBADARG, "k=%d outside 1..1024", k);
    if (dtype == TTDFT_F64) {
        TT_CHECK(H->val64, TTDFT_EBADARG, "no FP64 copy of H_loc uploaded");
#define CALL(NCV)                                                                        \
    spmm_dispatch_nc<double, NCV>(ctx, H, (const double *)X, ldx, (const double *)W,     \
                                  (double *)Y, ldy, (const double *)Z, ldz, k, alpha,    \
                                  beta, delta, use_beta, use_delta, H->val64, H->ct_val64)
        TT_NC_SWITCH(nc, CALL)
#undef CALL
    } else {
        TT_CHECK(H->val32, TTDFT_EBADARG, "no FP32 copy of H_loc uploaded");
#define CALL(NCV)                                                                        \
    spmm_dispatch_nc<float, NCV>(ctx, H, (const float *)X, ldx, (const float *)W,        \
                                 (float *)Y, ldy, (const float *)Z, ldz, k, alpha, beta, \
                                 delta, use_beta, use_delta, H->val32, H->ct_val32)
        TT_NC_SWITCH(nc, CALL)
#undef CALL
    }
    return TTDFT_EBADARG;
}

int ham_spmv(ttdft_ctx ctx, const ttdft_ham_s *H, const double *x, double *y) {
    TT_CHECK(H->val64, TTDFT_EBADARG, "no FP64 copy of H_loc uploaded");
    double *W = nullptr;
    if (H->p > 0) {
        TT_TRY(ctx->scratch[3].ensure((size_t)nl_workspace(H, 1) * sizeof(double)));
        W = (double *)ctx->scratch[3].ptr;
        TT_TRY(launch_nl_project(ctx, H, x, 1, 1, TTDFT_F64, W));
    }
    hipLaunchKernelGGL(spmv_kernel, dim3((unsigned)((H->n + 3) / 4)), dim3(256), 0, ctx->stream,
                       H->n, H->rowptr, H->col, H->val64, x, H->p > 0 ? H->ct_rowptr : nullptr,
                       H->ct_col, H->ct_val64, W, y);
    TT_LAUNCH_CHECK();
    return TTDFT_OK;
}

void drop_dense(ttdft_ham_s *H) {
    dev_release(H->ctx, H->dense64);
    dev_release(H->ctx, H->dense32);
    H->dense64 = nullptr;
    H->dense32 = nullptr;
    H->dense_nl = false;
    if (H->format == 2) H->format = 0;
}

// Dense storage when H_loc has no Kronecker band storage, at least a fifth of
// its entries are nonzero and the FP64 matrix takes at most 48 GiB (n up to
// about 80,000): the CSR kernel then walks rows of thousands of entries with
// the block's k states on one wavefront, while a matrix product streams the
// n^2 values once at full bandwidth.
constexpr double kDenseMaxBytes = 48.0 * 1024 * 1024 * 1024;
int maybe_build_dense(ttdft_ctx ctx, ttdft_ham_s *H, int copies) {
    (void)copies;
    const int64_t n = H->n;
    if (H->format != 0 || n < 1 || 8.0 * (double)n * (double)n > kDenseMaxBytes ||
        (double)H->nnz < 0.2 * (double)n * n)
        return TTDFT_OK;
    // rows padded to 16 and the leading dimension to 32 (the product's tiles)
    const int64_t ld = (n + 31) / 32 * 32;
    const size_t n2 = (size_t)((n + 15) / 16 * 16) * (size_t)ld;
    const unsigned rows = (unsigned)((n + 3) / 4);
    if (H->val64) {
        TT_TRY(dev_alloc(ctx, (void **)&H->dense64, sizeof(double) * n2));
        TT_HIP(hipMemsetAsync(H->dense64, 0, sizeof(double) * n2, ctx->stream));
        hipLaunchKernelGGL(csr_to_dense_kernel<double>, dim3(rows), dim3(256), 0, ctx->stream, n,
                           ld, H->rowptr, H->col, H->val64, H->dense64);
        TT_LAUNCH_CHECK();
    } else if (H->val32) {
        TT_TRY(dev_alloc(ctx, (void **)&H->dense32, sizeof(float) * n2));
        TT_HIP(hipMemsetAsync(H->dense32, 0, sizeof(float) * n2, ctx->stream));
        hipLaunchKernelGGL(csr_to_dense_kernel<float>, dim3(rows), dim3(256), 0, ctx->stream, n,
                           ld, H->rowptr, H->col, H->val32, H->dense32);
        TT_LAUNCH_CHECK();
    } else {
        return TTDFT_OK;
    }
    H->dense_ld = ld;
    H->dense_nl = false;
    H->format = 2;
    return TTDFT_OK;
}

// Completes the dense storage once the projectors are on the device: the
// nonlocal term C^T D C is added into the FP64 matrix (one rank-p product,
// D C from c_val64 and C from the transposed rows ct_val64), so a product
// with the dense H needs no projection pass; then the FP32 copy is rounded
// from it and the FP64 copy dropped if only FP32 was asked for.  An FP32-only
// build (no FP64 values) keeps the projection pass.
int finish_dense(ttdft_ctx ctx, ttdft_ham_s *H, int copies) {
    if (H->format != 2 || !H->dense64) return TTDFT_OK;
    const int64_t n = H->n, ld = H->dense_ld;
    const size_t n2 = (size_t)((n + 15) / 16 * 16) * (size_t)ld;
    if (H->p > 0 && H->c_val64 && H->ct_val64) {
        const int64_t p = H->p;
        double *A1 = nullptr, *A2 = nullptr;  // D C as p x n rows, C^T as n x p rows
        TT_TRY(dev_alloc(ctx, (void **)&A1, sizeof(double) * 2 * (size_t)p * n));
        A2 = A1 + p * n;
        int rc = TTDFT_OK;
        hipError_t e = hipMemsetAsync(A1, 0, sizeof(double) * 2 * (size_t)p * n, ctx->stream);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(csr_to_dense_kernel<double>, dim3((unsigned)((p + 3) / 4)),
                               dim3(256), 0, ctx->stream, p, n, H->c_rowptr, H->c_col,
                               H->c_val64, A1);
            hipLaunchKernelGGL(csr_to_dense_kernel<double>, dim3((unsigned)((n + 3) / 4)),
                               dim3(256), 0, ctx->stream, n, p, H->ct_rowptr, H->ct_col,
                               H->ct_val64, A2);
            e = hipGetLastError();
        }
        if (e != hipSuccess) {
            set_error("finish_dense: %s", hipGetErrorString(e));
            rc = TTDFT_EHIP;
        } else {
            // column-major view: H(l, i) = dense[i ld + l] += sum_r (D C)(r, l) C(r, i)
            const double one = 1.0;
            const rocblas_status st = rocblas_dgemm(
                ctx->blas, rocblas_operation_none, rocblas_operation_none, (rocblas_int)n,
                (rocblas_int)n, (rocblas_int)p, &one, A1, (rocblas_int)n, A2, (rocblas_int)p,
                &one, H->dense64, (rocblas_int)ld);
            if (st != rocblas_status_success) {
                set_error("finish_dense: rocblas_dgemm status %d", (int)st);
                rc = TTDFT_ELIB;
            }
        }
        dev_release(ctx, A1);  // reused in stream order
        if (rc != TTDFT_OK) return rc;
        H->dense_nl = true;
    }
    if (copies & TTDFT_HAM_F32) {
        TT_TRY(dev_alloc(ctx, (void **)&H->dense32, sizeof(float) * n2));
        hipLaunchKernelGGL(dense_to_f32_kernel, dim3(grid_blocks((int64_t)n2)), dim3(256), 0,
                           ctx->stream, H->dense64, H->dense32, (int64_t)n2);
        TT_LAUNCH_CHECK();
    }
    if (!(copies & TTDFT_HAM_F64)) {
        dev_release(ctx, H->dense64);
        H->dense64 = nullptr;
    }
    return TTDFT_OK;
}

constexpr int kDenseParts = 4;  // most parts of l in the split dense product

// 0: one-kernel dense product; 1: split product, parts picked; 2..4: split
// product with that many parts
int dense_split_form(int dtype) {
    (void)dtype;  // both precisions take the split product by default
    const char *env = std::getenv("TTDFT_DENSE_LDS");
    return env ? std::atoi(env) : 1;
}

template <typename T, int JT>
static void dense_launch(ttdft_ctx ctx, const ttdft_ham_s *H, const T *Hd, const T *ctv,
                         const void *X, int64_t ldx, const void *W, void *Y, int64_t ldy,
                         const void *Z, int64_t ldz, int32_t k, double alpha, double beta,
                         double delta, int use_beta, int use_delta, double *P) {
    const int64_t n = H->n;
    const dim3 grid((unsigned)((n + 15) / 16), (unsigned)((k + 16 * JT - 1) / (16 * JT)));
    const int64_t *ctr = H->p > 0 && !H->dense_nl ? H->ct_rowptr : nullptr;
    // six-state runs per lane, fetched as pairs, when every tile is full and
    // every row of X starts on a pair boundary (even leading dimension and a
    // pair-aligned X: a filter block may start at an odd column)
    // The split product with X staged in LDS, against the one-kernel form:
    // FP32 27,000 x 96 1.39 against 2.54 ms, 27,000 x 48 0.74 against 1.26,
    // 8,000 x 24 0.077 against 0.089; FP64 27,000 x 96 3.02 against 3.26,
    // 27,000 x 48 1.58 against 2.14, 8,000 x 24 0.133 against 0.153.
    // TTDFT_DENSE_LDS=0 turns the split off, 1 picks the parts, 2..4 fixes
    // them.
    const int lds_form = dense_split_form(sizeof(T) == 8 ? TTDFT_F64 : TTDFT_F32);
    if (lds_form >= 1 && P) {
        // parts of l for about 4.5 workgroups per CU, 6 for FP64 at 80-96
        // states (measured best at 27,000: FP64 x 96 3.17 ms at four parts,
        // 3.33 at three; FP64 x 48 1.62 ms at three, 1.74 at four)
        const int64_t wgs = (n + 63) / 64 * grid.y;
        const int64_t target = sizeof(T) == 8 && JT >= 5 ? 1536 : 1152;
        int nparts = (int)((target + wgs - 1) / wgs);
        if (lds_form > 1) nparts = lds_form;
        const int64_t nst = H->dense_ld / 32;
        if (nparts > nst) nparts = (int)nst;
        if (nparts > kDenseParts) nparts = kDenseParts;
        const dim3 g((unsigned)((n + 63) / 64), grid.y, (unsigned)nparts);
        // X staged in pairs when every pair is aligned and whole
        const bool pairs = k % 2 == 0 && ldx % 2 == 0 &&
                           reinterpret_cast<uintptr_t>(X) % (2 * sizeof(T)) == 0;
        if (pairs)
            hipLaunchKernelGGL((dense_hx_part_kernel<T, JT, 2>), g, dim3(256), 0, ctx->stream,
                               n, H->dense_ld, k, Hd, (const T *)X, ldx, nparts, P);
        else
            hipLaunchKernelGGL((dense_hx_part_kernel<T, JT, 1>), g, dim3(256), 0, ctx->stream,
                               n, H->dense_ld, k, Hd, (const T *)X, ldx, nparts, P);
        const int64_t total = n * k;
        const unsigned cb = (unsigned)std::min<int64_t>((total + 255) / 256, 65536);
        hipLaunchKernelGGL((dense_combine_kernel<T>), dim3(cb), dim3(256), 0, ctx->stream, n, k,
                           nparts, P, (const T *)X, ldx, ctr, H->ct_col, ctv, (const T *)W,
                           (T *)Y, ldy, (const T *)Z, ldz, alpha, beta, delta, use_beta,
                           use_delta);
        return;
    }
    if (JT == 6 && k % (16 * JT) == 0 && ldx % 2 == 0 &&
        reinterpret_cast<uintptr_t>(X) % (2 * sizeof(T)) == 0)
        hipLaunchKernelGGL((dense_hx_kernel<T, JT, (JT == 6)>), grid, dim3(64 * kDenseWaves), 0,
                           ctx->stream, n, H->dense_ld, k, Hd, (const T *)X, ldx, ctr, H->ct_col,
                           ctv, (const T *)W, (T *)Y, ldy, (const T *)Z, ldz, alpha, beta, delta,
                           use_beta, use_delta);
    else
        hipLaunchKernelGGL((dense_hx_kernel<T, JT, false>), grid, dim3(64 * kDenseWaves), 0,
                           ctx->stream, n, H->dense_ld, k, Hd, (const T *)X, ldx, ctr, H->ct_col,
                           ctv, (const T *)W, (T *)Y, ldy, (const T *)Z, ldz, alpha, beta, delta,
                           use_beta, use_delta);
}

// H X on the dense storage with the recurrence epilogue in one kernel; the
// state tiles per workgroup follow k (16 to 96 states)
static int launch_dense_spmm(ttdft_ctx ctx, const ttdft_ham_s *H, const void *X, int64_t ldx,
                             const void *W, void *Y, int64_t ldy, const void *Z, int64_t ldz,
                             int32_t k, int dtype, double alpha, double beta, double delta,
                             int use_beta, int use_delta, double *P) {
    // fewest workgroup columns of at most 96 states, then the fewest 16-state
    // tiles that cover k in that many columns (each column streams H once)
    const int ncol = (k + 95) / 96;
    const int jt = (k + 16 * ncol - 1) / (16 * ncol);
#define DENSE_CASE(JTV)                                                                         \
    case JTV:                                                                                   \
        if (dtype == TTDFT_F64)                                                                 \
            dense_launch<double, JTV>(ctx, H, H->dense64, H->ct_val64, X, ldx, W, Y, ldy, Z,    \
                                      ldz, k, alpha, beta, delta, use_beta, use_delta, P);      \
        else                                                                                    \
            dense_launch<float, JTV>(ctx, H, H->dense32, H->ct_val32, X, ldx, W, Y, ldy, Z,     \
                                     ldz, k, alpha, beta, delta, use_beta, use_delta, P);       \
        break;
    switch (jt) {
        DENSE_CASE(1)
        DENSE_CASE(2)
        DENSE_CASE(3)
        DENSE_CASE(4)
        DENSE_CASE(5)
        DENSE_CASE(6)
    default:
        break;
    }
#undef DENSE_CASE
    TT_LAUNCH_CHECK();
    return TTDFT_OK;
}

int ham_step(ttdft_ctx ctx, const ttdft_ham_s *H, const void *X, int64_t ldx, void *Y,
             int64_t ldy, const void *Z, int64_t ldz, int32_t k, int dtype, double alpha,
             double beta, double delta, int use_beta, int use_delta, void *ws,
             size_t ws_bytes) {
    const size_t es = dtype_size(dtype);
    const bool band = H->format == 1 && k >= 16 &&
                      (dtype == TTDFT_F64 ? H->mf64 != nullptr : H->mf32 != nullptr);
    const bool dense = H->format == 2 && (dtype == TTDFT_F64 ? H->dense64 != nullptr
                                                            : H->dense32 != nullptr);
    void *W = nullptr;
    if (H->p > 0 && !(dense && H->dense_nl)) {  // folded into the dense storage otherwise
        const size_t need = (size_t)nl_workspace(H, k) * es;
        if (ws) {
            TT_CHECK(need <= ws_bytes, TTDFT_EBADARG,
                     "the nonlocal projection of %d projectors x %d states needs %zu bytes of "
                     "work space, more than the %zu available",
                     H->p, k, need, ws_bytes);
            W = ws;
        } else {
            TT_TRY(ctx->scratch[3].ensure(need));
            W = ctx->scratch[3].ptr;
        }
        PendingTiming tn;
        TT_TRY(timing_begin(ctx, TTDFT_KSTAT_NLPROJ, &tn));
        TT_TRY(launch_nl_project(ctx, H, X, ldx, k, dtype, W));
        TT_TRY(timing_end(ctx, &tn, (double)H->nnz_c * (es + 4) + (double)H->nnz_c * k * es,
                          2.0 * H->nnz_c * k));
    }
    PendingTiming ts;
    TT_TRY(timing_begin(ctx, TTDFT_KSTAT_SPMM, &ts));
    if (band)
        TT_TRY(launch_band_spmm(ctx, H, X, ldx, W, Y, ldy, Z ? Z : Y, Z ? ldz : ldy, k, dtype,
                                alpha, beta, delta, use_beta, use_delta));
    else if (dense) {
        // partial sums of the split product: after the projections in the
        // caller's work space, else from the context.  When the caller's work
        // space cannot hold them the product runs in one kernel (no partial
        // sums; the same result up to the order of the additions).
        double *P = nullptr;
        if (dense_split_form(dtype) >= 1) {
            const size_t need_p = (size_t)kDenseParts * H->n * k * sizeof(double);
            if (ws) {
                const size_t off = W ? ((size_t)nl_workspace(H, k) * es + 255) / 256 * 256 : 0;
                if (off + need_p <= ws_bytes) P = (double *)((char *)ws + off);
            } else {
                TT_TRY(ctx->scratch[8].ensure(need_p));
                P = (double *)ctx->scratch[8].ptr;
            }
        }
        TT_TRY(launch_dense_spmm(ctx, H, X, ldx, W, Y, ldy, Z ? Z : Y, Z ? ldz : ldy, k, dtype,
                                 alpha, beta, delta, use_beta, use_delta, P));
    }
    else
        TT_TRY(launch_spmm_fused(ctx, H, X, ldx, W, Y, ldy, Z ? Z : Y, Z ? ldz : ldy, k, dtype,
                                 alpha, beta, delta, use_beta, use_delta));
    // algorithmic bytes: H_loc values + indices + row pointers, X panel read once,
    // Y written once, the recurrence operands read once (fused epilogue).
    const double nk = (double)H->n * k * es;
    const double bytes = (double)H->nnz * (es + 4) + 8.0 * (H->n + 1) + nk + nk +
                         (use_delta ? nk : 0.0);
    const double flops = 2.0 * H->nnz * k + 5.0 * H->n * k;
    TT_TRY(timing_end(ctx, &ts, bytes, flops));
    return TTDFT_OK;
}

}  // namespace ttdft
