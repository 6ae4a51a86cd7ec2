// Matrix-core SpMM for H_loc with Kronecker band structure.
//
// In a Tucker product basis the projected Hamiltonian couples (x,y,z) with
// (x',y',z') only for |x-x'| <= bx, |y-y'| <= by, |z-z'| <= bz
// (ProjectHamiltonianSparse.cc:464-548 builds each row as a Kronecker dyad and
// truncates it).  The rows are grouped into 4x4 "bricks" (4 consecutive x, 4
// consecutive y, one z).  For one brick and one neighbouring x-line (y',z'),
// the coupling is a dense 16 x (4 + 2bx) block: 16 brick rows against the x'
// window of that line.  It is padded to 16 x 4KS and stored pre-swizzled in
// the A-operand layout of the 16x16x4 MFMA.  HX for a brick and a 16-state
// tile is then a sum of KS MFMAs per neighbouring line, so the flops of the
// band run on the matrix cores in FP64 or FP32 instead of the VALU.
//
// A workgroup owns one y-brick row (4 y), one z, a chunk of up to 16 x-bricks
// (one wavefront each) and 16*NST states.  For every neighbouring line it
// stages the line's X window (all x-bricks of the chunk plus the 2bx halo,
// 16*NST states) in LDS once; every wave reads its B operands from there.
// The stage is double-buffered in LDS; the global loads of X and of the A
// fragments run two lines ahead through register buffers, so HBM/L2 latency
// hides behind MFMA work.  The nonlocal term C^T W and the Chebyshev three-term
// recurrence are applied in the epilogue, as in the CSR kernel.  Workgroups are
// renumbered so consecutive logical workgroups land on the same XCD (dispatch
// is round-robin over the 8 XCDs), keeping neighbouring lines' X in one L2.
#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <thread>
#include <type_traits>
#include <vector>

#include "internal.h"

namespace ttdft {
namespace {

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ d4_t mfma16x16x4(double a, double b, d4_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f4_t mfma16x16x4(float a, float b, f4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
template <typename T>
struct Acc4;
template <>
struct Acc4<double> {
    typedef d4_t type;
};
template <>
struct Acc4<float> {
    typedef f4_t type;
};

// states per workgroup = 16 * kStateTiles<T>
template <typename T>
constexpr int kStateTiles() { return sizeof(T) == 8 ? 4 : 5; }

template <typename T>
struct BrickArgs {
    const T *mf;
    const T *X;
    int64_t ldx;
    const int64_t *ct_rowptr;
    const int32_t *ct_col;
    const T *ct_val;
    const T *Wnl;
    int32_t ldw;
    T *Y;
    int64_t ldy;
    const T *Z;
    int64_t ldz;
    int32_t k, rx, ry, rz, bx, by, bz;
    int32_t gxn, hyn;  // x-bricks, y-bricks
    int32_t gxb, nxc;  // x-bricks per workgroup, x-chunks
    int32_t nsplit;    // trailing x-bricks whose state tiles are split over two waves
    int32_t nsg;       // state groups
    int64_t nblocks;
    T alpha, beta, delta;
    int use_beta, use_delta;
    int32_t light_last;  // ragged last x-chunk's units after all full ones, per XCD
};

typedef __attribute__((address_space(3))) void lds_void;

// s_waitcnt immediate: vmcnt in bits [3:0] and [15:14], expcnt [6:4], lgkmcnt [11:8].
// Waits until at most n vector-memory operations of this wave are outstanding
// and all LDS operations are done.
#define TT_WAIT_VM_LDS(n)                                                                   \
    do {                                                                                    \
        asm volatile("" ::: "memory");                                                      \
        __builtin_amdgcn_s_waitcnt(((n) & 15) | ((((n) >> 4) & 3) << 14) | (7 << 4) | (0 << 8)); \
    } while (0)
#define TT_WAIT_ALL() __builtin_amdgcn_s_waitcnt(0)

// Workgroup barrier without the memory fence of __syncthreads(): that fence
// would also wait for the line loads still in flight into LDS.  The waits
// the barrier needs are issued explicitly before it; the empty asm keeps the
// compiler from moving memory operations across.
__device__ __forceinline__ void line_barrier() {
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}


// N LDS loads of T from byte offsets OFF0 + i STRIDE past p, completed inside
// the asm block (s_waitcnt lgkmcnt(0)); every output is used by the caller.
// The outputs are early-clobber ("=&v"): a load may return while later loads
// of the block still read the address register, so no output may share it.
// Written as asm because the compiler cannot tell these reads apart from the
// global->LDS loads into the other line buffers and would otherwise wait for
// all of those to finish first.
#define TT_DSR(I) #I
template <typename T, int N, int OFF0, int STRIDE>
__device__ __forceinline__ void lds_read_n(const T *p, T (&v)[N]) {
    static_assert(N >= 1 && N <= 5, "1..5 loads");
    static_assert(OFF0 + (N - 1) * STRIDE < 65536, "ds offset field");
    const unsigned a = (unsigned)(uintptr_t)p;
#define TT_R(i) "ds_read_" TT_OP " %" #i ", %[a] offset:%[o" #i "]\n\t"
    if constexpr (sizeof(T) == 8) {
#define TT_OP "b64"
        if constexpr (N == 1)
            asm volatile(TT_R(0) "s_waitcnt lgkmcnt(0)" : "=&v"(v[0]) : [a] "v"(a), [o0] "i"(OFF0));
        else if constexpr (N == 2)
            asm volatile(TT_R(0) TT_R(1) "s_waitcnt lgkmcnt(0)"
                         : "=&v"(v[0]), "=&v"(v[1])
                         : [a] "v"(a), [o0] "i"(OFF0), [o1] "i"(OFF0 + STRIDE));
        else if constexpr (N == 3)
            asm volatile(TT_R(0) TT_R(1) TT_R(2) "s_waitcnt lgkmcnt(0)"
                         : "=&v"(v[0]), "=&v"(v[1]), "=&v"(v[2])
                         : [a] "v"(a), [o0] "i"(OFF0), [o1] "i"(OFF0 + STRIDE),
                           [o2] "i"(OFF0 + 2 * STRIDE));
        else if constexpr (N == 4)
            asm volatile(TT_R(0) TT_R(1) TT_R(2) TT_R(3) "s_waitcnt lgkmcnt(0)"
                         : "=&v"(v[0]), "=&v"(v[1]), "=&v"(v[2]), "=&v"(v[3])
                         : [a] "v"(a), [o0] "i"(OFF0), [o1] "i"(OFF0 + STRIDE),
                           [o2] "i"(OFF0 + 2 * STRIDE), [o3] "i"(OFF0 + 3 * STRIDE));
        else
            asm volatile(TT_R(0) TT_R(1) TT_R(2) TT_R(3) TT_R(4) "s_waitcnt lgkmcnt(0)"
                         : "=&v"(v[0]), "=&v"(v[1]), "=&v"(v[2]), "=&v"(v[3]), "=&v"(v[4])
                         : [a] "v"(a), [o0] "i"(OFF0), [o1] "i"(OFF0 + STRIDE),
                           [o2] "i"(OFF0 + 2 * STRIDE), [o3] "i"(OFF0 + 3 * STRIDE),
                           [o4] "i"(OFF0 + 4 * STRIDE));
#undef TT_OP
    } else {
#define TT_OP "b32"
        if constexpr (N == 1)
            asm volatile(TT_R(0) "s_waitcnt lgkmcnt(0)" : "=&v"(v[0]) : [a] "v"(a), [o0] "i"(OFF0));
        else if constexpr (N == 2)
            asm volatile(TT_R(0) TT_R(1) "s_waitcnt lgkmcnt(0)"
                         : "=&v"(v[0]), "=&v"(v[1])
                         : [a] "v"(a), [o0] "i"(OFF0), [o1] "i"(OFF0 + STRIDE));
        else if constexpr (N == 3)
            asm volatile(TT_R(0) TT_R(1) TT_R(2) "s_waitcnt lgkmcnt(0)"
                         : "=&v"(v[0]), "=&v"(v[1]), "=&v"(v[2])
                         : [a] "v"(a), [o0] "i"(OFF0), [o1] "i"(OFF0 + STRIDE),
                           [o2] "i"(OFF0 + 2 * STRIDE));
        else if constexpr (N == 4)
            asm volatile(TT_R(0) TT_R(1) TT_R(2) TT_R(3) "s_waitcnt lgkmcnt(0)"
                         : "=&v"(v[0]), "=&v"(v[1]), "=&v"(v[2]), "=&v"(v[3])
                         : [a] "v"(a), [o0] "i"(OFF0), [o1] "i"(OFF0 + STRIDE),
                           [o2] "i"(OFF0 + 2 * STRIDE), [o3] "i"(OFF0 + 3 * STRIDE));
        else
            asm volatile(TT_R(0) TT_R(1) TT_R(2) TT_R(3) TT_R(4) "s_waitcnt lgkmcnt(0)"
                         : "=&v"(v[0]), "=&v"(v[1]), "=&v"(v[2]), "=&v"(v[3]), "=&v"(v[4])
                         : [a] "v"(a), [o0] "i"(OFF0), [o1] "i"(OFF0 + STRIDE),
                           [o2] "i"(OFF0 + 2 * STRIDE), [o3] "i"(OFF0 + 3 * STRIDE),
                           [o4] "i"(OFF0 + 4 * STRIDE));
#undef TT_OP
    }
#undef TT_R
}
#undef TT_DSR

// KS: 4-column MFMA steps covering the x' window (4 KS >= 4 + 2 bx).
// NST: 16-state tiles per workgroup.  DB: bytes per lane of one global->LDS
// load (16 when a row of X is 16-byte aligned, else 4).
// GXM: most x-bricks per workgroup (16, or 8 for half the LDS and two
// workgroups per CU).  G: state groups per workgroup.  G = 1: one wave per
// x-brick, S = 16 NST states.  G > 1: G waves per x-brick, one per group of S
// states, all G S states of a line staged together, so the G waves of a brick
// load the same A fragments at the same line step (the fragments come from
// HBM once instead of once per state group); workgroups then hold exactly GXM
// x-bricks.
// ZB: z planes per workgroup (1 or 2).  With 2, every wave accumulates its
// brick in two consecutive planes: a staged line (y', z') serves both planes
// that lie within the z band of z', so the X staged per matrix-core flop
// halves.  A line only a plane outside the band of z' is skipped for that
// plane; each plane sees its lines in the same order as with ZB = 1, so
// the sums are bitwise the same.
template <typename T, int KS, int NST, int DB, int DEPTH, int GXM, int G, int ZB = 1>
__global__ __launch_bounds__(G > 1 ? 64 * G * GXM : 1024) void brick_spmm_kernel(BrickArgs<T> a) {
    typedef typename Acc4<T>::type acc_t;
    constexpr int S = 16 * NST;                  // states per wave
    constexpr int ST = G * S;                    // states per staged row
    constexpr int ES = sizeof(T);
    constexpr int CH = 64 * DB;                  // bytes one wave loads per instruction
    constexpr int TRMAX = 4 * (GXM - 1) + 4 * KS;  // staged rows at GXM x-bricks
    constexpr int BUF = (TRMAX * ST * ES + CH - 1) / CH * CH;
    // loads per wave per line (a fixed count): G = 1 keeps at least 4 waves
    // and 4 + KS - 1 rows per wave; G > 1 runs G GXM waves
    constexpr int NQ = G == 1 ? (DB == 16 ? 4 : 16) : (BUF / CH + G * GXM - 1) / (G * GXM);
    constexpr int VE = DB >= ES ? DB / ES : 1;   // elements per lane-load
    constexpr int NB = DEPTH + 1;  // LDS buffers: lines i .. i+DEPTH
    static_assert(DEPTH >= 1 && DEPTH <= 3, "1 to 3 lines in flight");
    static_assert(ZB == 1 || ZB == 2, "one or two z planes per workgroup");
    static_assert((4 * (KS - 1) * ST + 16 * (NST - 1)) * ES < 65536, "ds offset field");
    __shared__ __attribute__((aligned(16))) unsigned char lb0[BUF];
    __shared__ __attribute__((aligned(16))) unsigned char lb1[BUF];
    __shared__ __attribute__((aligned(16))) unsigned char lb2[NB >= 3 ? BUF : 16];
    __shared__ __attribute__((aligned(16))) unsigned char lb3[NB == 4 ? BUF : 16];
    const int nw = G == 1 ? a.gxb + a.nsplit : G * a.gxb;
    const int TR = 4 * (a.gxb - 1) + 4 * KS;  // staged x' rows
    const int nch = (TR * ST * ES + CH - 1) / CH;
    const int lane = threadIdx.x & 63;
    const int wl = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int li = lane & 15, lk = lane >> 4;

    const unsigned nb = gridDim.x;  // a multiple of 8
    const unsigned bid = blockIdx.x;
    // Work unit (state group, x-chunk, y-brick row, plane group) of this
    // workgroup.  Dispatch deals workgroups round-robin over the 8 XCDs, so
    // XCD bid % 8 runs positions bid / 8 in order; each XCD gets a contiguous
    // range of units.  With light_last the units of a ragged last x-chunk
    // (fewer x-bricks, less work) come after all full ones on every XCD, so
    // the kernel's last round is made of short workgroups.
    int64_t t;
    int xc;
    if (a.light_last) {
        const int64_t per_chunk = a.nblocks / a.nxc;  // units of one x-chunk
        const int64_t H = per_chunk * (a.nxc - 1);
        const int64_t hs = (H + 7) / 8, ls = (per_chunk + 7) / 8;
        const int64_t p = bid >> 3, xcd = bid & 7u;
        if (p < hs) {
            const int64_t u = xcd * hs + p;
            if (u >= H) return;
            t = u / a.nsg;
            xc = (int)(t % (a.nxc - 1));
            t /= a.nxc - 1;
            t = t * a.nsg + u % a.nsg;  // (sg, rest) packed again below
        } else {
            const int64_t u = xcd * ls + (p - hs);
            if (u >= per_chunk) return;
            xc = a.nxc - 1;
            t = u;
        }
    } else {
        const unsigned L = (bid & 7u) * (nb >> 3) + (bid >> 3);
        if ((int64_t)L >= a.nblocks) return;
        t = L;
        const int64_t q = t / a.nsg;
        xc = (int)(q % a.nxc);
        t = (q / a.nxc) * a.nsg + t % a.nsg;
    }
    // t = sg + nsg (h + hyn zgroup)
    const int sg = (int)(t % a.nsg);
    t /= a.nsg;
    const int h = (int)(t % a.hyn);
    const int z0 = ZB * (int)(t / a.hyn);  // first z plane of the workgroup
    const int s0 = sg * ST;
    // Waves are spread round-robin over the 4 SIMDs and the MFMA work of a
    // line ends at a barrier, so the SIMD with the most tile work sets the
    // pace.  With gxb = 1 or 2 (mod 4) the state tiles of the last one or two
    // x-bricks are split over two waves each (mode 1: lower tiles, mode 2:
    // upper tiles), which evens out the MFMA work per SIMD.
    constexpr int NH = NST - NST / 2;  // tiles of the lower half
    int bl = wl, mode = 0, sgw = 0;
    if (G > 1) {
        bl = wl / G;
        sgw = wl - bl * G;
    } else if (wl >= a.gxb - a.nsplit) {
        const int q = wl - (a.gxb - a.nsplit);
        bl = a.gxb - a.nsplit + (q >> 1);
        mode = 1 + (q & 1);
    }
    const int g0 = xc * a.gxb;
    const int g = g0 + bl;
    const bool active = g < a.gxn;
    const int xw0 = 4 * g0 - a.bx;  // x' of staged row 0
    const int NPY = 2 * a.by + 4;
    const int NP = NPY * (2 * a.bz + 1);

    acc_t acc[ZB][NST];
#pragma unroll
    for (int zi = 0; zi < ZB; ++zi)
#pragma unroll
        for (int st = 0; st < NST; ++st) acc[zi][st] = acc_t{0, 0, 0, 0};

    const int y0 = 4 * h;
    const int zlast = z0 + ZB - 1 < a.rz ? z0 + ZB - 1 : a.rz - 1;  // last plane in the domain
    const int zlo = z0 - a.bz < 0 ? 0 : z0 - a.bz;
    const int zhi = zlast + a.bz > a.rz - 1 ? a.rz - 1 : zlast + a.bz;
    const int ylo = y0 - a.by < 0 ? 0 : y0 - a.by;
    const int yhi = y0 + 3 + a.by > a.ry - 1 ? a.ry - 1 : y0 + 3 + a.by;
    const int ny = yhi - ylo + 1;
    const int nlines = ny * (zhi - zlo + 1);

    // The staged tile of a line is rows x' = xw0 .. xw0 + TR - 1 of that line
    // by states s0 .. s0 + S - 1, row-major.  This wave loads byte chunks
    // q = wl + i nw (i < NQ; past the last chunk it repeats the last one,
    // which writes the same bytes again).  A lane's byte offset inside the
    // line of X is the same for every line: computed once.  Rows outside
    // 0..rx-1 and states past the end load clamped, finite values; the A
    // fragments are zero against those rows and those states are never
    // stored.
    int32_t goff[NQ];
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
        int q = wl + i * nw;
        q = q < nch ? q : nch - 1;
        const int off = q * CH + lane * DB;  // byte offset in the tile
        const int e = off / ES;
        const int row = e / ST, s = e - row * ST;
        const int xp = xw0 + row;
        const int xpc = xp < 0 ? 0 : (xp >= a.rx ? a.rx - 1 : xp);
        int st = s0 + s;
        if (st > (int)a.ldx - VE) st = (int)a.ldx - VE;
        goff[i] = (xpc * (int)a.ldx + st) * ES + (off % ES);
    }
    auto line_of = [&](int li_, int &yy, int &zz) {
        const int lc = li_ < nlines ? li_ : nlines - 1;
        yy = ylo + lc % ny;
        zz = zlo + lc / ny;
    };
    auto fetch = [&](int li_, unsigned char *buf) {
        int yy, zz;
        line_of(li_, yy, zz);
        const char *lbase =
            reinterpret_cast<const char *>(a.X) + ((int64_t)(zz * a.ry + yy) * a.rx) * a.ldx * ES;
#pragma unroll
        for (int i = 0; i < NQ; ++i) {
            int q = wl + i * nw;
            q = q < nch ? q : nch - 1;
#if defined(__HIP_DEVICE_COMPILE__)
            // (the host pass of a .hip file rejects the 16-byte form of this builtin)
            __builtin_amdgcn_global_load_lds(lbase + goff[i], (lds_void *)(buf + q * CH), DB, 0,
                                             0);
#else
            (void)lbase;
            (void)buf;
            (void)q;
#endif
        }
    };
    // Does plane zi of the workgroup take line zz (inside the domain and
    // within the z band)?
    auto plane_on = [&](int zi, int zz) {
        const int dz = zz - (z0 + zi);
        return z0 + zi < a.rz && dz >= -a.bz && dz <= a.bz;
    };
    // A fragments of a line for every plane (wave-private, from HBM/L2).  A
    // plane that does not take the line loads a valid fragment it never uses,
    // so every wave issues the same loads per line.
    auto load_a = [&](int li_, T (&av)[ZB][KS]) {
        int yy, zz;
        line_of(li_, yy, zz);
        const int gc = active ? g : a.gxn - 1;
        const int tt = yy - y0;  // in [-by, 3 + by] by construction
#pragma unroll
        for (int zi = 0; zi < ZB; ++zi) {
            const int zp = z0 + zi < a.rz ? z0 + zi : a.rz - 1;
            const int p = (tt + a.by) + NPY * (plane_on(zi, zz) ? zz - zp + a.bz : 0);
            const T *Ab =
                a.mf + ((((int64_t)(zp * a.hyn + h) * a.gxn + gc) * NP + p) * KS) * 64 + lane;
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) av[zi][ks] = Ab[ks * 64];
        }
    };
    // reads the B operands of k-step ks (compile-time after unrolling is not
    // enough for asm immediates, so dispatch on ks)
    auto read_ks = [&](const T *Xb, auto &b, int ks, auto lo_c, auto hi_c) {
        constexpr int LO = decltype(lo_c)::value, HI = decltype(hi_c)::value;
        constexpr int NT = HI - LO;
        switch (ks) {
        case 1:
            if constexpr (KS > 1) lds_read_n<T, NT, (4 * 1 * ST + 16 * LO) * ES, 16 * ES>(Xb, b[1 < KS ? 1 : 0]);
            break;
        case 2:
            if constexpr (KS > 2) lds_read_n<T, NT, (4 * 2 * ST + 16 * LO) * ES, 16 * ES>(Xb, b[2 < KS ? 2 : 0]);
            break;
        case 3:
            if constexpr (KS > 3) lds_read_n<T, NT, (4 * 3 * ST + 16 * LO) * ES, 16 * ES>(Xb, b[3 < KS ? 3 : 0]);
            break;
        case 4:
            if constexpr (KS > 4) lds_read_n<T, NT, (4 * 4 * ST + 16 * LO) * ES, 16 * ES>(Xb, b[4 < KS ? 4 : 0]);
            break;
        default:
            break;
        }
    };
    // The MFMAs of k-step ks are issued before the B operands of ks+1 are
    // read, so the LDS latency overlaps their execution.
    // With two planes the B operands of a k-step serve both planes' MFMAs.
    // State tiles of this wave that hold a state below k (wave-uniform): in
    // FP64 the tiles past it, padding of the last state group, get no MFMA
    // (4% of the 172-atom product).  The FP32 kernel keeps its loop free of
    // the branches, which cost it more than the skipped work (3%).
    const int tv = (a.k - (s0 + sgw * S) + 15) / 16;
    constexpr bool SKIP = sizeof(T) == 8;
    auto mfma_tiles = [&](const T *Xb, const T (&av)[ZB][KS], const bool (&on)[ZB], auto lo_c,
                          auto hi_c) {
        constexpr int LO = decltype(lo_c)::value, HI = decltype(hi_c)::value;
        constexpr int NT = HI - LO;
        T b[KS][NT];
        lds_read_n<T, NT, 16 * LO * ES, 16 * ES>(Xb, b[0]);
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
#pragma unroll
            for (int zi = 0; zi < ZB; ++zi)
                if (on[zi]) {
#pragma unroll
                    for (int t = 0; t < NT; ++t)
                        if (!SKIP || LO + t < tv)
                            acc[zi][LO + t] = mfma16x16x4(av[zi][ks], b[ks][t], acc[zi][LO + t]);
                }
            if constexpr (true) {
                if (ks + 1 < KS) read_ks(Xb, b, ks + 1, lo_c, hi_c);
            }
        }
    };
    auto compute = [&](const unsigned char *buf, const T (&av)[ZB][KS], int li_) {
        int yy, zz;
        line_of(li_, yy, zz);
        bool on[ZB];
#pragma unroll
        for (int zi = 0; zi < ZB; ++zi) on[zi] = plane_on(zi, zz);
        const T *Xb = reinterpret_cast<const T *>(buf) + (4 * bl + lk) * ST + sgw * S + li;
        if (mode == 0)
            mfma_tiles(Xb, av, on, std::integral_constant<int, 0>(),
                       std::integral_constant<int, NST>());
        else if (mode == 1)
            mfma_tiles(Xb, av, on, std::integral_constant<int, 0>(),
                       std::integral_constant<int, NH>());
        else
            mfma_tiles(Xb, av, on, std::integral_constant<int, NH>(),
                       std::integral_constant<int, NST>());
    };

    // Line i is in LDS buffer i % NB.  While it is multiplied, the loads of
    // lines i+1 .. i+DEPTH (X straight into LDS, A fragments into register
    // sets av[(i+d) % NB]) are in flight: each has DEPTH lines of MFMA work to
    // arrive.  Every wave issues NQ + KS vector-memory operations per line in
    // a fixed order, so before the barrier that ends line i, waiting until at
    // most (DEPTH-1)(NQ+KS) are outstanding (those of lines i+2 .. i+DEPTH)
    // completes this wave's loads of line i+1.  The loop is unrolled by NB so
    // buffers and register sets are addressed statically; buffers are picked
    // by compile-time index so every LDS access is known to touch one buffer.
    auto buf_of = [&](auto bc) -> unsigned char * {
        if constexpr (decltype(bc)::value == 0)
            return lb0;
        else if constexpr (decltype(bc)::value == 1)
            return lb1;
        else if constexpr (decltype(bc)::value == 2)
            return lb2;
        else
            return lb3;
    };
    constexpr int WAIT = (DEPTH - 1) * (NQ + ZB * KS);
    constexpr int PERIOD = NB;
    T av[NB][ZB][KS];
    fetch(0, lb0);
    load_a(0, av[0]);
    if constexpr (DEPTH >= 2) {
        fetch(1, lb1);
        load_a(1, av[1]);
    }
    if constexpr (DEPTH == 3) {
        fetch(2, lb2);
        load_a(2, av[2]);
    }
    TT_WAIT_VM_LDS(WAIT);
    line_barrier();
    auto step = [&](int li_, auto jc) {
        constexpr int J = decltype(jc)::value;
        fetch(li_ + DEPTH, buf_of(std::integral_constant<int, (J + DEPTH) % NB>()));
        load_a(li_ + DEPTH, av[(J + DEPTH) % NB]);
        if (active) compute(buf_of(std::integral_constant<int, J % NB>()), av[J % NB], li_);
        TT_WAIT_VM_LDS(WAIT);
        line_barrier();
    };
    for (int base = 0; base < nlines; base += PERIOD) {
        step(base, std::integral_constant<int, 0>());
        if (base + 1 >= nlines) break;
        step(base + 1, std::integral_constant<int, 1>());
        if constexpr (PERIOD >= 3) {
            if (base + 2 >= nlines) break;
            step(base + 2, std::integral_constant<int, 2>());
        }
        if constexpr (PERIOD >= 4) {
            if (base + 3 >= nlines) break;
            step(base + 3, std::integral_constant<int, 3>());
        }
    }
    TT_WAIT_ALL();  // no load into LDS may outlive the workgroup
    if (!active) return;

    // D layout of the 16x16x4 MFMA: FP64 element v of lane l is D[l/16 + 4v][l%16],
    // FP32 element v is D[4(l/16) + v][l%16].  D rows are brick rows i =
    // (x - 4g) + 4 (y - 4h), D columns are states.
#pragma unroll
    for (int zi = 0; zi < ZB; ++zi) {
    const int z = z0 + zi;
    if (z >= a.rz) break;
#pragma unroll
    for (int st = 0; st < NST; ++st) {
        if ((mode == 1 && st >= NH) || (mode == 2 && st < NH)) continue;
        const int s = s0 + sgw * S + 16 * st + li;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = sizeof(T) == 8 ? (lk + 4 * v) : (4 * lk + v);
            const int x = 4 * g + (i & 3), y = y0 + (i >> 2);
            if (x < a.rx && y < a.ry && s < a.k) {
                const int64_t row = x + (int64_t)a.rx * (y + (int64_t)a.ry * z);
                T val = acc[zi][st][v];
                if (a.ct_rowptr) {
                    const int64_t f0 = a.ct_rowptr[row], f1 = a.ct_rowptr[row + 1];
                    for (int64_t e = f0; e < f1; ++e)
                        val += a.ct_val[e] * a.Wnl[(int64_t)a.ct_col[e] * a.ldw + s];
                }
                T out = val * a.alpha;
                if (a.use_beta) out = out + a.beta * a.X[row * a.ldx + s];
                if (a.use_delta) out = out + a.delta * a.Z[row * a.ldz + s];
                a.Y[row * a.ldy + s] = out;
            }
        }
    }
    }
}

// Workgroup shapes with all state groups co-resident (G > 1):
//   FP64: 3 x-bricks x 5 groups of 64 states, one line in flight (two 61 KB
//         stages of 320 states at KS = 4);
//   FP32: 4 x-bricks x 4 groups of 80 states, one line in flight (two 36 KB
//         stages), two with two planes (launch_db).
// Measured alternatives on the 172-atom product: FP64 with 2 x-bricks and two
// lines in flight 6.48 ms against 5.80.
template <typename T, int V>
struct CoShape;
template <>
struct CoShape<double, 0> {
    static constexpr int GXM = 3, G = 5, DEPTH = 1;
};
template <>
struct CoShape<float, 0> {
    static constexpr int GXM = 4, G = 4, DEPTH = 1;
};

// Does the two-plane form of the state-group shape fit the registers (no
// scratch) at KS?  Checked with the compiler's resource report: up to KS = 4
// in both precisions.
constexpr bool two_planes_fit(int ks) { return ks <= 4; }

template <typename T, int KS, int DB, int ZB>
int launch_db(ttdft_ctx ctx, const BrickArgs<T> &a, int cogroups) {
    constexpr int NST = kStateTiles<T>();
    int64_t blocks = (a.nblocks + 7) / 8 * 8;
    if (a.light_last) {  // 8 XCDs x (full-chunk positions + last-chunk positions)
        const int64_t per_chunk = a.nblocks / a.nxc;
        blocks = 8 * ((per_chunk * (a.nxc - 1) + 7) / 8 + (per_chunk + 7) / 8);
    }
    if (cogroups == 1) {
        typedef CoShape<T, 0> C;
        // FP32 two-plane workgroups keep two lines in flight (three 36 KB
        // stages): 172-atom product 3.45 against 3.54 ms.  With one plane
        // (KS = 5) one line measured faster (500-atom product 37.0 against
        // 39.0 ms).  TTDFT_BRICK_DEPTH=1 keeps one line.
        if constexpr (sizeof(T) == 4 && ZB == 2) {
            const char *env = std::getenv("TTDFT_BRICK_DEPTH");
            if (!(env && env[0] == '1')) {
                hipLaunchKernelGGL((brick_spmm_kernel<T, KS, NST, DB, 2, C::GXM, C::G, ZB>),
                                   dim3((unsigned)blocks), dim3(64 * C::G * C::GXM), 0,
                                   ctx->stream, a);
                TT_LAUNCH_CHECK();
                return TTDFT_OK;
            }
        }
        hipLaunchKernelGGL((brick_spmm_kernel<T, KS, NST, DB, C::DEPTH, C::GXM, C::G, ZB>),
                           dim3((unsigned)blocks), dim3(64 * C::G * C::GXM), 0, ctx->stream, a);
        TT_LAUNCH_CHECK();
        return TTDFT_OK;
    }
    // Two lines in flight.  (Three fit in LDS but measured slower on the
    // 172-atom workload: FP64 8.2 ms against 7.7 ms per H application, FP32
    // 4.47 against 4.24 ms.)
    if constexpr (ZB == 1) {  // the two-plane form exists for state-group workgroups only
        constexpr int DEPTH = 2;
        const int nthr = 64 * (a.gxb + a.nsplit);
        if (a.gxb <= 8)
            hipLaunchKernelGGL((brick_spmm_kernel<T, KS, NST, DB, DEPTH, 8, 1, ZB>),
                               dim3((unsigned)blocks), dim3(nthr), 0, ctx->stream, a);
        else
            hipLaunchKernelGGL((brick_spmm_kernel<T, KS, NST, DB, DEPTH, 16, 1, ZB>),
                               dim3((unsigned)blocks), dim3(nthr), 0, ctx->stream, a);
        TT_LAUNCH_CHECK();
        return TTDFT_OK;
    }
    set_error("band SpMM: two z planes need state-group workgroups");
    return TTDFT_EBADARG;
}

// 16-byte loads need every X row 16-byte aligned
template <typename T>
bool band_wide_loads(const BrickArgs<T> &a) {
    return (a.ldx * (int64_t)sizeof(T)) % 16 == 0 && (reinterpret_cast<uintptr_t>(a.X) & 15) == 0;
}

template <typename T, int KS>
int launch_ks(ttdft_ctx ctx, const BrickArgs<T> &a, int cogroups, int zb) {
    TT_CHECK(a.gxb >= 1 && (cogroups ? CoShape<T, 0>::G * a.gxb : a.gxb + a.nsplit) <= 16,
             TTDFT_EBADARG, "band SpMM: %d x-bricks per chunk", a.gxb);
    TT_CHECK((int64_t)a.rx * a.ldx * (int64_t)sizeof(T) < ((int64_t)1 << 31), TTDFT_EBADARG,
             "band SpMM: one x-line of X exceeds 2 GiB");
    const bool wide = band_wide_loads(a);
    if constexpr (two_planes_fit(KS)) {
        if (zb == 2) {
            TT_CHECK(wide && cogroups, TTDFT_EBADARG, "band SpMM: no two-plane form here");
            return launch_db<T, KS, 16, 2>(ctx, a, cogroups);
        }
    }
    TT_CHECK(zb == 1, TTDFT_EBADARG, "band SpMM: no two-plane form at KS = %d", KS);
    return wide ? launch_db<T, KS, 16, 1>(ctx, a, cogroups)
                : launch_db<T, KS, 4, 1>(ctx, a, cogroups);
}

template <typename T>
int launch_brick(ttdft_ctx ctx, BrickArgs<T> a, int ks) {
    // x-bricks per workgroup: at most TTDFT_BRICK_CHUNK (16 by default; 8 halves
    // the LDS stage so two workgroups share a CU)
    int chunk = sizeof(T) == 8 ? 16 : 8;  // FP32: two workgroups per CU measured faster
    if (const char *env = std::getenv("TTDFT_BRICK_CHUNK")) {
        const int v = std::atoi(env);
        if (v == 8 || v == 16) chunk = v;
    }
    // All state groups of a row block in one workgroup when there are more
    // than G - 1 full groups (172-atom product: FP64 5.80 against 7.43 ms,
    // FP32 3.49 against 4.18 ms).  TTDFT_BRICK_COGROUPS=0 turns it off, 1 on.
    const int S = 16 * kStateTiles<T>();
    int cogroups = a.k > (CoShape<T, 0>::G - 1) * S ? 1 : 0;
    if (const char *env = std::getenv("TTDFT_BRICK_COGROUPS")) {
        const int v = std::atoi(env);
        if (a.k > (CoShape<T, 0>::G - 1) * S && (v == 0 || v == 1)) cogroups = v;
    }
    if (cogroups) {
        a.gxb = CoShape<T, 0>::GXM;
        a.nxc = (a.gxn + a.gxb - 1) / a.gxb;
        a.nsplit = 0;
        a.nsg = (a.k + CoShape<T, 0>::G * S - 1) / (CoShape<T, 0>::G * S);
    } else {
        a.nxc = (a.gxn + chunk - 1) / chunk;
        a.gxb = std::max(4, (a.gxn + a.nxc - 1) / a.nxc);
        if (a.gxb > chunk) a.gxb = chunk;
        a.nxc = (a.gxn + a.gxb - 1) / a.gxb;
        a.nsg = (a.k + S - 1) / S;
        a.nsplit = a.gxb % 4 == 1 ? 1 : (a.gxb % 4 == 2 ? 2 : 0);
    }
    // z planes per workgroup: two for the state-group workgroups (the X
    // stage then serves twice the matrix-core work), else one.
    // TTDFT_BRICK_ZB=1 or 2 overrides.
    // Two planes need the accumulators and A fragments of both in the 128
    // registers a wave has at 15-16 waves per workgroup: that fits the
    // state-group forms (one line in flight) up to KS = 4; beyond it the
    // compiler moves the arrays to scratch memory (20x slower).  The 4-byte
    // load form (unaligned rows) does not fit either.
    const bool zb2_fits = cogroups && two_planes_fit(ks) && band_wide_loads(a);
    int zb = zb2_fits ? 2 : 1;
    if (const char *env = std::getenv("TTDFT_BRICK_ZB")) {
        const int v = std::atoi(env);
        if (v == 1 || (v == 2 && zb2_fits)) zb = v;
    }
    a.nblocks = (int64_t)a.nsg * a.nxc * a.hyn * ((a.rz + zb - 1) / zb);
    // Short workgroups last when the last x-chunk is ragged: the kernel's
    // final round is then made of them (172-atom FP64 product 6.05 -> 5.83 ms;
    // FP32 unchanged).  TTDFT_BRICK_ORDER=0 keeps the plain order.
    int order = 1;
    if (const char *env = std::getenv("TTDFT_BRICK_ORDER")) order = std::atoi(env) == 0 ? 0 : 1;
    a.light_last = (order == 1 && a.nxc >= 2 && a.gxn % a.gxb != 0) ? 1 : 0;
    switch (ks) {
    case 2:
        return launch_ks<T, 2>(ctx, a, cogroups, zb);
    case 3:
        return launch_ks<T, 3>(ctx, a, cogroups, zb);
    case 4:
        return launch_ks<T, 4>(ctx, a, cogroups, zb);
    case 5:
        return launch_ks<T, 5>(ctx, a, cogroups, zb);
    default:
        break;
    }
    set_error("band SpMM: unsupported x half-bandwidth %d", a.bx);
    return TTDFT_EBADARG;
}

}  // namespace

int launch_band_spmm(ttdft_ctx ctx, const ttdft_ham_s *H, const void *X, int64_t ldx,
                     const void *W, void *Y, int64_t ldy, const void *Z, int64_t ldz,
                     int32_t k, int dtype, double alpha, double beta, double delta,
                     int use_beta, int use_delta) {
    if (dtype == TTDFT_F64) {
        TT_CHECK(H->mf64, TTDFT_EBADARG, "no FP64 band copy of H_loc uploaded");
        BrickArgs<double> a{H->mf64, (const double *)X, ldx,
                            H->p > 0 ? H->ct_rowptr : nullptr, H->ct_col, H->ct_val64,
                            (const double *)W, k, (double *)Y, ldy, (const double *)Z, ldz,
                            k, H->rx, H->ry, H->rz, H->bx, H->by, H->bz, H->gxn, H->hyn,
                            H->gxb, H->nxc, 0, 0, 0, alpha, beta, delta, use_beta, use_delta};
        return launch_brick<double>(ctx, a, H->ks);
    }
    TT_CHECK(H->mf32, TTDFT_EBADARG, "no FP32 band copy of H_loc uploaded");
    BrickArgs<float> a{H->mf32, (const float *)X, ldx,
                       H->p > 0 ? H->ct_rowptr : nullptr, H->ct_col, H->ct_val32,
                       (const float *)W, k, (float *)Y, ldy, (const float *)Z, ldz,
                       k, H->rx, H->ry, H->rz, H->bx, H->by, H->bz, H->gxn, H->hyn,
                       H->gxb, H->nxc, 0, 0, 0, (float)alpha, (float)beta, (float)delta,
                       use_beta, use_delta};
    return launch_brick<float>(ctx, a, H->ks);
}

// ---------------------------------------------------------------------------
// Host: detect the band and scatter the CSR values into MFMA A fragments.
//
// Fragment layout: mf[((((z * hyn + h) * gxn + g) * NP + p) * KS + ks) * 64 + lane]
// holds A[i][kk] of the 16x16x4 MFMA (lane = i + 16 kk) for brick (g, h, z),
// neighbouring line p = (y' - 4h + by) + (2by + 4)(z' - z + bz) and x' =
// 4g - bx + 4ks + kk; brick row i = (x - 4g) + 4 (y - 4h).  Zero where H_loc
// has no entry.
namespace {

// Band layout of H_loc for half-bandwidths (bx, by, bz); false: keep CSR.
struct BandShape {
    int ks, gxn, hyn, npy, np;
    int64_t slots;  // A-fragment elements
};
bool band_shape(int32_t rx, int32_t ry, int32_t rz, int bx, int by, int bz, int64_t nnz,
                BandShape *b) {
    const int64_t n = (int64_t)rx * ry * rz;
    const int64_t band_slots = n * (int64_t)(2 * bx + 1) * (2 * by + 1) * (2 * bz + 1);
    const double fill = band_slots > 0 ? (double)nnz / (double)band_slots : 0.0;
    if (bx < 1 || bx > 7 || fill < 0.15 || rx < 4) return false;
    b->ks = (4 + 2 * bx + 3) / 4;
    b->gxn = (rx + 3) / 4;
    b->hyn = (ry + 3) / 4;
    b->npy = 2 * by + 4;
    b->np = b->npy * (2 * bz + 1);
    b->slots = (int64_t)rz * b->hyn * b->gxn * b->np * b->ks * 64;
    return true;
}

void set_band_fields(ttdft_ham_s *H, int32_t rx, int32_t ry, int32_t rz, int bx, int by, int bz,
                     const BandShape &b) {
    const int nxc = (b.gxn + 15) / 16;
    H->format = 1;
    H->rx = rx;
    H->ry = ry;
    H->rz = rz;
    H->bx = bx;
    H->by = by;
    H->bz = bz;
    H->ks = b.ks;
    H->gxn = b.gxn;
    H->hyn = b.hyn;
    H->nxc = nxc;
    H->gxb = std::max(4, (b.gxn + nxc - 1) / nxc);
}

// Fragment slot of entry (row r, column j) of H_loc; the row lies in brick
// (x>>2, y>>2, z), the column in line (y', z') at window offset c.
__host__ __device__ inline int64_t band_slot(int64_t r, int64_t j, int32_t rx, int32_t ry, int bx,
                                             int by, int bz, int ks, int gxn, int hyn, int npy,
                                             int np) {
    const int x = (int)(r % rx), y = (int)((r / rx) % ry), z = (int)(r / ((int64_t)rx * ry));
    const int xj = (int)(j % rx), yj = (int)((j / rx) % ry), zj = (int)(j / ((int64_t)rx * ry));
    const int g = x >> 2, h = y >> 2;
    const int i16 = (x & 3) + 4 * (y & 3);
    const int64_t base = (((int64_t)z * hyn + h) * gxn + g) * np;
    const int p = (yj - 4 * h + by) + npy * (zj - z + bz);
    const int c = xj - 4 * g + bx;  // in [0, 4 ks)
    return ((base + p) * ks + (c >> 2)) * 64 + i16 + 16 * (c & 3);
}

// One wavefront per row of the device CSR; every (row, column) pair of the
// band has its own slot, so the stores never collide.
__global__ __launch_bounds__(256) void band_scatter_kernel(
    const int64_t *__restrict__ rowptr, const int32_t *__restrict__ col,
    const double *__restrict__ val, int64_t n, int32_t rx, int32_t ry, int bx, int by, int bz,
    int ks, int gxn, int hyn, int npy, int np, double *__restrict__ mf) {
    const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= n) return;
    for (int64_t e = rowptr[r] + (threadIdx.x & 63); e < rowptr[r + 1]; e += 64)
        mf[band_slot(r, col[e], rx, ry, bx, by, bz, ks, gxn, hyn, npy, np)] = val[e];
}

__global__ __launch_bounds__(256) void f64_to_f32_kernel(const double *__restrict__ a,
                                                         float *__restrict__ b, int64_t count) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) b[i] = (float)a[i];
}

}  // namespace

int build_band(ttdft_ctx ctx, ttdft_ham_s *H, int32_t rx, int32_t ry, int32_t rz,
               const int64_t *rowptr, const int32_t *colind, const double *val, int copies) {
    const int64_t n = (int64_t)rx * ry * rz;
    unsigned nt = std::thread::hardware_concurrency();
    if (nt == 0) nt = 1;
    if (nt > 32) nt = 32;
    std::vector<int> mx(nt, 0), my(nt, 0), mz(nt, 0);
    {
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nt; ++t)
            th.emplace_back([&, t]() {
                int bx = 0, by = 0, bz = 0;
                for (int64_t i = n * t / nt; i < n * (t + 1) / nt; ++i) {
                    const int x = (int)(i % rx), y = (int)((i / rx) % ry),
                              z = (int)(i / ((int64_t)rx * ry));
                    for (int64_t e = rowptr[i]; e < rowptr[i + 1]; ++e) {
                        const int64_t j = colind[e];
                        const int xj = (int)(j % rx), yj = (int)((j / rx) % ry),
                                  zj = (int)(j / ((int64_t)rx * ry));
                        bx = std::max(bx, std::abs(xj - x));
                        by = std::max(by, std::abs(yj - y));
                        bz = std::max(bz, std::abs(zj - z));
                    }
                }
                mx[t] = bx;
                my[t] = by;
                mz[t] = bz;
            });
        for (auto &x : th) x.join();
    }
    const int bx = *std::max_element(mx.begin(), mx.end());
    const int by = *std::max_element(my.begin(), my.end());
    const int bz = *std::max_element(mz.begin(), mz.end());
    BandShape b;
    if (!band_shape(rx, ry, rz, bx, by, bz, rowptr[n], &b)) return TTDFT_OK;  // keep CSR
    std::vector<double> mf((size_t)b.slots, 0.0);
    {
        // rows of one thread fill distinct brick rows: no two threads write a slot
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nt; ++t)
            th.emplace_back([&, t]() {
                for (int64_t r = n * t / nt; r < n * (t + 1) / nt; ++r)
                    for (int64_t e = rowptr[r]; e < rowptr[r + 1]; ++e)
                        mf[(size_t)band_slot(r, colind[e], rx, ry, bx, by, bz, b.ks, b.gxn,
                                             b.hyn, b.npy, b.np)] += val[e];
            });
        for (auto &x : th) x.join();
    }
    if (copies & TTDFT_HAM_F64) {
        TT_TRY(dev_alloc(ctx, (void **)&H->mf64, sizeof(double) * mf.size()));
        TT_HIP(hipMemcpyAsync(H->mf64, mf.data(), sizeof(double) * mf.size(),
                              hipMemcpyHostToDevice, ctx->stream));
        TT_HIP(hipStreamSynchronize(ctx->stream));
    }
    if (copies & TTDFT_HAM_F32) {
        std::vector<float> m32(mf.size());
        for (size_t q = 0; q < mf.size(); ++q) m32[q] = (float)mf[q];
        TT_TRY(dev_alloc(ctx, (void **)&H->mf32, sizeof(float) * m32.size()));
        TT_HIP(hipMemcpyAsync(H->mf32, m32.data(), sizeof(float) * m32.size(),
                              hipMemcpyHostToDevice, ctx->stream));
        TT_HIP(hipStreamSynchronize(ctx->stream));
    }
    set_band_fields(H, rx, ry, rz, bx, by, bz, b);
    return TTDFT_OK;
}

int build_band_device(ttdft_ctx ctx, ttdft_ham_s *H, int32_t rx, int32_t ry, int32_t rz, int bx,
                      int by, int bz, int copies) {
    const int64_t n = (int64_t)rx * ry * rz;
    BandShape b;
    if (!band_shape(rx, ry, rz, bx, by, bz, H->nnz, &b)) return TTDFT_OK;  // keep CSR
    TT_CHECK(H->val64, TTDFT_EBADARG, "build_band_device: FP64 CSR values missing");
    double *mf = nullptr;
    TT_TRY(dev_alloc(ctx, (void **)&mf, sizeof(double) * b.slots));
    H->mf64 = mf;  // owned by H from here on
    TT_HIP(hipMemsetAsync(mf, 0, sizeof(double) * b.slots, ctx->stream));
    const unsigned nb = (unsigned)((n + 3) / 4);
    hipLaunchKernelGGL(band_scatter_kernel, dim3(nb), dim3(256), 0, ctx->stream, H->rowptr,
                       H->col, H->val64, n, rx, ry, bx, by, bz, b.ks, b.gxn, b.hyn, b.npy, b.np,
                       mf);
    TT_LAUNCH_CHECK();
    if (copies & TTDFT_HAM_F32) {
        TT_TRY(dev_alloc(ctx, (void **)&H->mf32, sizeof(float) * b.slots));
        hipLaunchKernelGGL(f64_to_f32_kernel, dim3((unsigned)((b.slots + 255) / 256)),
                           dim3(256), 0, ctx->stream, mf, H->mf32, b.slots);
        TT_LAUNCH_CHECK();
    }
    if (!(copies & TTDFT_HAM_F64)) {
        // reused by later work on this stream only, after the conversion
        dev_release(ctx, mf);
        H->mf64 = nullptr;
    }
    set_band_fields(H, rx, ry, rz, bx, by, bz, b);
    return TTDFT_OK;
}

}  // namespace ttdft
