// C++ host layer with the reference's ChFSI call surface, over the C ABI of
// libttdft_chfsi.so.  The classes keep the reference's names, method names,
// argument order and meaning; PETSc Mat/Vec are replaced by repo-owned types:
//   CsrMatrix   - host CSR (MPIAIJ / SEQAIJ in the reference)
//   DenseBlock  - device-resident n x k wavefunction block (MATDENSE); stored
//                 basis-major on the device, converted from/to the
//                 reference's column-major layout by from_colmajor/to_colmajor
//   ProjectHamiltonianSparse - holds H_loc, C_nloc (unscaled) and 1/C_lm, the
//                 members the filter reads (dft/ProjectHamiltonianSparse.h:66-97)
// Errors: every failing engine call throws ttdft_host::Error (status + the
// engine's message).  Uncaught, it ends the program through std::terminate,
// which is what the reference does on any CUDA/cuBLAS/cuSPARSE/cuSOLVER
// failure (eigensolver/MultCuda.cu:57-60, MultTm.cu:157-161).
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/ttdft_chfsi.h"

namespace ttdft_host {

class Error : public std::runtime_error {
public:
    Error(int status, const std::string &msg)
        : std::runtime_error("ttdft status " + std::to_string(status) + ": " + msg),
          status_(status) {}
    int status() const { return status_; }

private:
    int status_;
};

void check(int status);

struct CsrMatrix {
    int64_t rows = 0, cols = 0;
    std::vector<int64_t> rowptr;
    std::vector<int32_t> col;
    std::vector<double> val;
    int64_t nnz() const { return rowptr.empty() ? 0 : rowptr.back(); }
};

// One device + one stream (DeviceUtils / DeviceCommUtils device binding).
class DeviceContext {
public:
    explicit DeviceContext(int device = 0);
    ~DeviceContext();
    DeviceContext(const DeviceContext &) = delete;
    DeviceContext &operator=(const DeviceContext &) = delete;
    ttdft_ctx handle() const { return ctx_; }
    void synchronize();

private:
    ttdft_ctx ctx_ = nullptr;
};

class DenseBlock {
public:
    DenseBlock(DeviceContext &ctx, int64_t rows, int32_t cols, int dtype = TTDFT_F64);
    ~DenseBlock();
    DenseBlock(const DenseBlock &) = delete;
    DenseBlock &operator=(const DenseBlock &) = delete;
    int64_t rows() const { return rows_; }
    int32_t cols() const { return cols_; }
    int dtype() const { return dtype_; }
    void *data() const { return ptr_; }
    int64_t ld() const { return cols_; }
    DeviceContext &context() const { return *ctx_; }
    // host column-major rows x cols (PETSc dense layout)
    void from_colmajor(const std::vector<double> &host);
    std::vector<double> to_colmajor() const;

private:
    DeviceContext *ctx_;
    int64_t rows_;
    int32_t cols_;
    int dtype_;
    void *ptr_ = nullptr;
};

// What Create_Hloc builds each row from (dft/ProjectHamiltonianSparse.cc:330-575):
// the 1-D overlap and kinetic matrices of the Tucker basis (MatX[I][J],
// MatDX, ... row-major r x r), the 1-D potential matrices MatPotX[p][I][J]
// (rankVeffX x rx x rx), and the Veff Tucker core, first index fastest.
// The quadrature that forms them from the FEM basis (:577-775) stays with the
// caller.
struct HlocFactors {
    int32_t rankVeffX = 0, rankVeffY = 0, rankVeffZ = 0;
    std::vector<double> MatX, MatDX, MatY, MatDY, MatZ, MatDZ;
    std::vector<double> MatPotX, MatPotY, MatPotZ;
    std::vector<double> coreVeff;
};

// What Create_Cnloc builds each projector row from (:196-278): per row a
// Tucker core (r^3, first index fastest) and column-major factors (rx x r,
// ry x r, rz x r), concatenated over the rows, and the row's 1/C_lm.
struct ProjectorFactors {
    std::vector<int32_t> ranks;
    std::vector<double> cores, factorX, factorY, factorZ;
    std::vector<double> inv_C_lm;
};

struct ProjectHamiltonianSparse {
    CsrMatrix H_loc;                // n x n, symmetric
    CsrMatrix C_nloc;               // p x n projector rows, not row-scaled
    std::vector<double> inv_C_lm;   // p
    int32_t rankTuckerBasisX = 0, rankTuckerBasisY = 0, rankTuckerBasisZ = 0;

    // dft/ProjectHamiltonianSparse.cc:330-575 (with Compute_nnz_loc, :1223-1592):
    // H_loc assembled on the device, entries |h| > truncation, into host CSR
    void Create_Hloc(DeviceContext &ctx, int32_t rankTuckerBasisX, int32_t rankTuckerBasisY,
                     int32_t rankTuckerBasisZ, const HlocFactors &f, double truncation);
    // :83-328: projector rows assembled on the device into host CSR (unscaled)
    void Create_Cnloc(DeviceContext &ctx, const ProjectorFactors &f, double truncation);
};

// eigensolver/Mult.h:30-114
class AX {
public:
    AX(int num_bg, int num_wfns);
    ~AX();
    AX(const AX &) = delete;
    AX &operator=(const AX &) = delete;
    // Uploads H_loc together with the nonlocal projectors of H, so that
    // perform_ax applies the whole operator the filter needs.
    void setup_system(DeviceContext &ctx, const ProjectHamiltonianSparse &H,
                      int copies = TTDFT_HAM_F64);
    // Create_Hloc + Create_Cnloc + setup_system in one step, with H built in
    // device memory (no host CSR, no upload)
    void assemble_system(DeviceContext &ctx, int32_t rankTuckerBasisX, int32_t rankTuckerBasisY,
                         int32_t rankTuckerBasisZ, const HlocFactors &hf, double truncation,
                         const ProjectorFactors &pf, double c_truncation,
                         int copies = TTDFT_HAM_F64);
    // AX = H X (written, unlike the published perform_ax, MultCuda.cu:719-820)
    void perform_ax(DenseBlock &X, DenseBlock &AXout);
    void free_mata_on_device();
    ttdft_ham handle() const { return ham_; }
    DeviceContext *context() const { return ctx_; }

private:
    int num_band_groups_;
    int num_wfns_;
    DeviceContext *ctx_ = nullptr;
    ttdft_ham ham_ = nullptr;
};

// dft/AndersonMixing.h:29-92: Anderson mixing of the nodal and quadrature-grid
// densities, history in device memory.  Densities are device arrays
// (n_nodal and n_grid doubles); jacobian and weights are device grid fields.
class AndersonMixing {
public:
    AndersonMixing(DeviceContext &ctx, int64_t n_grid, int64_t n_nodal, const double *d_jacob,
                   const double *d_weight, double alpha, int maxIterHistory = 40);
    ~AndersonMixing();
    AndersonMixing(const AndersonMixing &) = delete;
    AndersonMixing &operator=(const AndersonMixing &) = delete;
    void computeRhoIn(int scfIter, double *rhoNodalIn, double *rhoGridIn);
    void updateRho(const double *rhoNodalIn, const double *rhoGridIn, const double *rhoNodalOut,
                   const double *rhoGridOut, int scfIter);
    void clearHistory(int scf_iter);
    std::vector<double> mixingConstants() const;

private:
    ttdft_mixer m_ = nullptr;
};

// eigensolver/ChebyShevFilter.h:29-40
class ChebyShevFilter {
public:
    explicit ChebyShevFilter(int degree_convention = TTDFT_DEGREE_REFERENCE)
        : convention_(degree_convention) {}
    void computeFilteredSubspace(ProjectHamiltonianSparse &H, AX &ax, DenseBlock &X,
                                 int basisSize, int m, double a, double b, double a0,
                                 int num_X_blocks);

private:
    int convention_;
};

// eigensolver/MultTm.h:33-53
class MultTM {
public:
    explicit MultTM(DeviceContext &ctx) : ctx_(&ctx) {}
    // result = X^T Y, k x k column-major (resized like MultTm.cu:138)
    void mult(DenseBlock &X, DenseBlock &Y, std::vector<double> &result);
    // Cholesky-QR, X <- X L^{-T}
    void orth(DenseBlock &X);

private:
    DeviceContext *ctx_;
};

// A sequential dense matrix (the MATSEQDENSE Mat of the reference, e.g. PtHP
// and Q in main.cc:1103-1114): host, column-major rows x cols.
struct SeqDenseMatrix {
    int rows = 0, cols = 0;
    std::vector<double> data;
};

// eigensolver/EigenSolver.h:29-45
class EigenSolver {
public:
    // main.cc:672 default-constructs the solver: it then binds device 0 on
    // its first call.  The second form shares a caller's context.
    EigenSolver() = default;
    explicit EigenSolver(DeviceContext &ctx) : ctx_(&ctx) {}
    ~EigenSolver();
    EigenSolver(const EigenSolver &) = delete;
    EigenSolver &operator=(const EigenSolver &) = delete;

    // The reference's signature (EigenSolver.h:33-39): A numGlobalRows x
    // numGlobalRows (lower triangle used); eigVals ascending, eigMat
    // numGlobalRows x numberEigenPairs with column i <-> eigVals[i].
    // `lapack` selects the dense solver in the reference; the engine always
    // uses its dense FP64 solver.
    void computeEigenPairsSeq(SeqDenseMatrix &A, int numberEigenPairs, int numGlobalRows,
                              std::vector<double> &eigVals, SeqDenseMatrix &eigMat, bool lapack);
    // The same with plain column-major vectors on an explicit context.
    void computeEigenPairsSeq(DeviceContext &ctx, const std::vector<double> &A,
                              int numberEigenPairs, int numGlobalRows,
                              std::vector<double> &eigVals, std::vector<double> &eigMat,
                              bool lapack);
    // The reference's signature (EigenSolver.h:41-45): main.cc calls it
    // before AX::setup_system (main.cc:989, 1028 against 1034), so H_loc and
    // the projectors are uploaded (FP64) for this call and freed after it.
    void computeUpperBoundWithLanczos(ProjectHamiltonianSparse &H, int maxLanczosIteration,
                                      double &upperBound, double &lowerBound);
    // On an operator already on the device (no upload).
    void computeUpperBoundWithLanczos(AX &ax, int maxLanczosIteration, double &upperBound,
                                      double &lowerBound, uint64_t seed = 20261202);

private:
    DeviceContext &context();
    DeviceContext *ctx_ = nullptr;
    DeviceContext *own_ = nullptr;  // created by the default form on first use
};

// KSDFTEnergyFunctional::computeFermiEnergy (dft/KSDFTEnergyFunctional.h:46-51)
double computeFermiEnergy(const double *eigenValues, int numberEigenValues, int numberElectrons,
                          double kb, double T, int maxFermiIter = 20);

}  // namespace ttdft_host
