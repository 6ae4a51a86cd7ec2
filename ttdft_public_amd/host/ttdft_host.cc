// Implementation of the reference-shaped C++ host layer over the C ABI.
#include "ttdft_host.h"

#include <algorithm>
#include <cstring>

namespace ttdft_host {

void check(int status) {
    if (status != TTDFT_OK) throw Error(status, ttdft_last_error());
}

// ---- DeviceContext -----------------------------------------------------------
DeviceContext::DeviceContext(int device) { check(ttdft_ctx_create(device, &ctx_)); }
DeviceContext::~DeviceContext() { ttdft_ctx_destroy(ctx_); }
void DeviceContext::synchronize() { check(ttdft_ctx_synchronize(ctx_)); }

// ---- DenseBlock ----------------------------------------------------------------
DenseBlock::DenseBlock(DeviceContext &ctx, int64_t rows, int32_t cols, int dtype)
    : ctx_(&ctx), rows_(rows), cols_(cols), dtype_(dtype) {
    const size_t es = dtype == TTDFT_F32 ? 4 : 8;
    check(ttdft_malloc(ctx.handle(), (size_t)rows * cols * es, &ptr_));
}
DenseBlock::~DenseBlock() {
    if (ptr_) ttdft_free(ctx_->handle(), ptr_);
}
void DenseBlock::from_colmajor(const std::vector<double> &host) {
    if ((int64_t)host.size() != rows_ * cols_) throw Error(TTDFT_EBADARG, "from_colmajor: size");
    check(ttdft_block_upload_colmajor(ctx_->handle(), host.data(), rows_, rows_, cols_, ptr_,
                                      ld(), dtype_));
}
std::vector<double> DenseBlock::to_colmajor() const {
    std::vector<double> host((size_t)rows_ * cols_);
    check(ttdft_block_download_colmajor(ctx_->handle(), ptr_, ld(), dtype_, rows_, cols_,
                                        host.data(), rows_));
    return host;
}

// ---- AX ----------------------------------------------------------------------
AX::AX(int num_bg, int num_wfns) : num_band_groups_(num_bg), num_wfns_(num_wfns) {
    if (num_bg != 1)
        throw Error(TTDFT_EBADARG, "one band group per GPU: the engine runs on one device");
}
AX::~AX() { free_mata_on_device(); }

void AX::setup_system(DeviceContext &ctx, const ProjectHamiltonianSparse &H, int copies) {
    free_mata_on_device();
    ctx_ = &ctx;
    const CsrMatrix &A = H.H_loc;
    const int32_t p = (int32_t)H.C_nloc.rows;
    const int64_t *crp = p ? H.C_nloc.rowptr.data() : nullptr;
    const int32_t *cc = p ? H.C_nloc.col.data() : nullptr;
    const double *cv = p ? H.C_nloc.val.data() : nullptr;
    const double *ic = p ? H.inv_C_lm.data() : nullptr;
    if (H.rankTuckerBasisX > 0)
        check(ttdft_ham_upload_tucker(ctx.handle(), H.rankTuckerBasisX, H.rankTuckerBasisY,
                                      H.rankTuckerBasisZ, A.rowptr.data(), A.col.data(),
                                      A.val.data(), p, crp, cc, cv, ic, copies, &ham_));
    else
        check(ttdft_ham_upload(ctx.handle(), A.rows, A.rowptr.data(), A.col.data(), A.val.data(),
                               p, crp, cc, cv, ic, copies, &ham_));
}

void AX::assemble_system(DeviceContext &ctx, int32_t rx, int32_t ry, int32_t rz,
                         const HlocFactors &f, double truncation, const ProjectorFactors &pf,
                         double c_truncation, int copies) {
    free_mata_on_device();
    ctx_ = &ctx;
    const int32_t p = (int32_t)pf.ranks.size();
    check(ttdft_ham_assemble_tucker(
        ctx.handle(), rx, ry, rz, f.rankVeffX, f.rankVeffY, f.rankVeffZ, f.MatX.data(),
        f.MatDX.data(), f.MatY.data(), f.MatDY.data(), f.MatZ.data(), f.MatDZ.data(),
        f.MatPotX.data(), f.MatPotY.data(), f.MatPotZ.data(), f.coreVeff.data(), truncation, p,
        p ? pf.ranks.data() : nullptr, p ? pf.cores.data() : nullptr,
        p ? pf.factorX.data() : nullptr, p ? pf.factorY.data() : nullptr,
        p ? pf.factorZ.data() : nullptr, c_truncation, p ? pf.inv_C_lm.data() : nullptr, copies,
        &ham_));
}

void AX::perform_ax(DenseBlock &X, DenseBlock &AXout) {
    if (!ham_) throw Error(TTDFT_EBADARG, "perform_ax before setup_system");
    check(ttdft_ham_apply(ctx_->handle(), ham_, X.data(), X.ld(), AXout.data(), AXout.ld(),
                          X.cols(), X.dtype()));
}

void AX::free_mata_on_device() {
    if (ham_) ttdft_ham_free(ham_);
    ham_ = nullptr;
}

// ---- ProjectHamiltonianSparse ------------------------------------------------------
void ProjectHamiltonianSparse::Create_Hloc(DeviceContext &ctx, int32_t rx, int32_t ry, int32_t rz,
                                           const HlocFactors &f, double truncation) {
    const int64_t n = (int64_t)rx * ry * rz;
    auto call = [&](int64_t *nnz, int64_t *rp, int32_t *c, double *v) {
        check(ttdft_assemble_hloc(ctx.handle(), rx, ry, rz, f.rankVeffX, f.rankVeffY,
                                  f.rankVeffZ, f.MatX.data(), f.MatDX.data(), f.MatY.data(),
                                  f.MatDY.data(), f.MatZ.data(), f.MatDZ.data(),
                                  f.MatPotX.data(), f.MatPotY.data(), f.MatPotZ.data(),
                                  f.coreVeff.data(), truncation, nnz, rp, c, v));
    };
    int64_t nnz = 0;
    call(&nnz, nullptr, nullptr, nullptr);  // Compute_nnz_loc
    H_loc.rows = H_loc.cols = n;
    H_loc.rowptr.assign(n + 1, 0);
    H_loc.col.resize(std::max<int64_t>(nnz, 1));
    H_loc.val.resize(std::max<int64_t>(nnz, 1));
    call(&nnz, H_loc.rowptr.data(), H_loc.col.data(), H_loc.val.data());
    H_loc.col.resize(nnz);
    H_loc.val.resize(nnz);
    rankTuckerBasisX = rx;
    rankTuckerBasisY = ry;
    rankTuckerBasisZ = rz;
}

void ProjectHamiltonianSparse::Create_Cnloc(DeviceContext &ctx, const ProjectorFactors &f,
                                            double truncation) {
    const int32_t p = (int32_t)f.ranks.size();
    if (rankTuckerBasisX <= 0) throw Error(TTDFT_EBADARG, "Create_Cnloc before Create_Hloc");
    if ((int32_t)f.inv_C_lm.size() != p) throw Error(TTDFT_EBADARG, "Create_Cnloc: inv_C_lm size");
    auto call = [&](int64_t *nnz, int64_t *rp, int32_t *c, double *v) {
        check(ttdft_assemble_projectors(ctx.handle(), p, p ? f.ranks.data() : nullptr,
                                        rankTuckerBasisX, rankTuckerBasisY, rankTuckerBasisZ,
                                        p ? f.cores.data() : nullptr,
                                        p ? f.factorX.data() : nullptr,
                                        p ? f.factorY.data() : nullptr,
                                        p ? f.factorZ.data() : nullptr, truncation, nnz, rp, c,
                                        v));
    };
    int64_t nnz = 0;
    call(&nnz, nullptr, nullptr, nullptr);
    C_nloc.rows = p;
    C_nloc.cols = (int64_t)rankTuckerBasisX * rankTuckerBasisY * rankTuckerBasisZ;
    C_nloc.rowptr.assign(p + 1, 0);
    C_nloc.col.resize(std::max<int64_t>(nnz, 1));
    C_nloc.val.resize(std::max<int64_t>(nnz, 1));
    call(&nnz, C_nloc.rowptr.data(), C_nloc.col.data(), C_nloc.val.data());
    C_nloc.col.resize(nnz);
    C_nloc.val.resize(nnz);
    inv_C_lm = f.inv_C_lm;
}

// ---- AndersonMixing -----------------------------------------------------------------
AndersonMixing::AndersonMixing(DeviceContext &ctx, int64_t n_grid, int64_t n_nodal,
                               const double *d_jacob, const double *d_weight, double alpha,
                               int maxIterHistory) {
    check(ttdft_mixer_create(ctx.handle(), n_grid, n_nodal, d_jacob, d_weight, alpha,
                             maxIterHistory, &m_));
}
AndersonMixing::~AndersonMixing() { ttdft_mixer_destroy(m_); }
void AndersonMixing::computeRhoIn(int scfIter, double *rhoNodalIn, double *rhoGridIn) {
    check(ttdft_mixer_compute_rho_in(m_, scfIter, rhoNodalIn, rhoGridIn));
}
void AndersonMixing::updateRho(const double *rhoNodalIn, const double *rhoGridIn,
                               const double *rhoNodalOut, const double *rhoGridOut,
                               int scfIter) {
    check(ttdft_mixer_update(m_, rhoNodalIn, rhoGridIn, rhoNodalOut, rhoGridOut, scfIter));
}
void AndersonMixing::clearHistory(int scf_iter) { check(ttdft_mixer_clear_history(m_, scf_iter)); }
std::vector<double> AndersonMixing::mixingConstants() const {
    int32_t n = 0;
    check(ttdft_mixer_constants(m_, nullptr, &n));
    std::vector<double> c(n);
    check(ttdft_mixer_constants(m_, c.data(), &n));
    return c;
}

// ---- ChebyShevFilter -------------------------------------------------------------
void ChebyShevFilter::computeFilteredSubspace(ProjectHamiltonianSparse &H, AX &ax, DenseBlock &X,
                                              int basisSize, int m, double a, double b,
                                              double a0, int num_X_blocks) {
    (void)H;  // the operator was uploaded by ax.setup_system
    if (basisSize != X.rows())
        throw Error(TTDFT_EBADARG, "computeFilteredSubspace: basisSize != rows of X");
    check(ttdft_cheb_filter(ax.context()->handle(), ax.handle(), X.data(), X.ld(), X.cols(), m,
                            a, b, a0, num_X_blocks, X.dtype(), convention_));
}

// ---- MultTM --------------------------------------------------------------------
void MultTM::mult(DenseBlock &X, DenseBlock &Y, std::vector<double> &result) {
    const int k = X.cols();
    if (Y.cols() != k || Y.rows() != X.rows()) throw Error(TTDFT_EBADARG, "mult: shape");
    double *dS = nullptr;
    check(ttdft_malloc(ctx_->handle(), sizeof(double) * k * k, (void **)&dS));
    const int mode = X.dtype() == TTDFT_F64 ? TTDFT_GEMM_F64 : TTDFT_GEMM_BF16X3;
    int rc = ttdft_gram(ctx_->handle(), X.data(), X.ld(), Y.data(), Y.ld(), X.rows(), k,
                        X.dtype(), mode, dS);
    result.resize((size_t)k * k);
    if (rc == TTDFT_OK)
        rc = ttdft_memcpy_d2h(ctx_->handle(), result.data(), dS, sizeof(double) * k * k);
    ttdft_free(ctx_->handle(), dS);
    check(rc);
}

void MultTM::orth(DenseBlock &X) {
    const int mode = X.dtype() == TTDFT_F64 ? TTDFT_GEMM_F64 : TTDFT_GEMM_BF16X3;
    check(ttdft_orth(ctx_->handle(), X.data(), X.ld(), X.rows(), X.cols(), X.dtype(), mode));
}

// ---- EigenSolver -----------------------------------------------------------------
EigenSolver::~EigenSolver() { delete own_; }

DeviceContext &EigenSolver::context() {
    if (!ctx_) {
        own_ = new DeviceContext(0);
        ctx_ = own_;
    }
    return *ctx_;
}

void EigenSolver::computeEigenPairsSeq(SeqDenseMatrix &A, int numberEigenPairs,
                                       int numGlobalRows, std::vector<double> &eigVals,
                                       SeqDenseMatrix &eigMat, bool lapack) {
    if (A.rows != numGlobalRows || A.cols != numGlobalRows)
        throw Error(TTDFT_EBADARG, "computeEigenPairsSeq: A is not numGlobalRows square");
    computeEigenPairsSeq(context(), A.data, numberEigenPairs, numGlobalRows, eigVals,
                         eigMat.data, lapack);
    eigMat.rows = numGlobalRows;
    eigMat.cols = numberEigenPairs;
}

void EigenSolver::computeUpperBoundWithLanczos(ProjectHamiltonianSparse &H,
                                               int maxLanczosIteration, double &upperBound,
                                               double &lowerBound) {
    AX ax(1, 0);
    ax.setup_system(context(), H, TTDFT_HAM_F64);
    computeUpperBoundWithLanczos(ax, maxLanczosIteration, upperBound, lowerBound);
}

void EigenSolver::computeEigenPairsSeq(DeviceContext &ctx, const std::vector<double> &A,
                                       int numberEigenPairs, int numGlobalRows,
                                       std::vector<double> &eigVals,
                                       std::vector<double> &eigMat, bool lapack) {
    (void)lapack;
    const int k = numGlobalRows;
    if ((int64_t)A.size() != (int64_t)k * k || numberEigenPairs > k || numberEigenPairs < 1)
        throw Error(TTDFT_EBADARG, "computeEigenPairsSeq: shape");
    double *d = nullptr;
    check(ttdft_malloc(ctx.handle(), sizeof(double) * (2 * (size_t)k * k + k), (void **)&d));
    double *dA = d, *dQ = d + (size_t)k * k, *dE = dQ + (size_t)k * k;
    std::vector<double> ev(k), Q((size_t)k * k);
    int rc = ttdft_memcpy_h2d(ctx.handle(), dA, A.data(), sizeof(double) * k * k);
    if (rc == TTDFT_OK) rc = ttdft_rr_eig(ctx.handle(), dA, k, dE, dQ);
    if (rc == TTDFT_OK) rc = ttdft_memcpy_d2h(ctx.handle(), ev.data(), dE, sizeof(double) * k);
    if (rc == TTDFT_OK)
        rc = ttdft_memcpy_d2h(ctx.handle(), Q.data(), dQ, sizeof(double) * k * k);
    ttdft_free(ctx.handle(), d);
    check(rc);
    // the reference returns the numberEigenPairs smallest pairs (EPS_SMALLEST_REAL)
    eigVals.assign(ev.begin(), ev.begin() + numberEigenPairs);
    eigMat.assign(Q.begin(), Q.begin() + (size_t)k * numberEigenPairs);
}

void EigenSolver::computeUpperBoundWithLanczos(AX &ax, int maxLanczosIteration, double &upperBound,
                                               double &lowerBound, uint64_t seed) {
    check(ttdft_lanczos_bounds(ax.context()->handle(), ax.handle(), maxLanczosIteration, seed,
                               &upperBound, &lowerBound));
}

double computeFermiEnergy(const double *eigenValues, int numberEigenValues, int numberElectrons,
                          double kb, double T, int maxFermiIter) {
    double ef = 0.0;
    check(ttdft_fermi_occupancy(eigenValues, numberEigenValues, numberElectrons, kb, T,
                                maxFermiIter, &ef, nullptr));
    return ef;
}

}  // namespace ttdft_host
