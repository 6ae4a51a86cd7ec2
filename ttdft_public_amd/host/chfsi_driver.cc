// ChFSI driver over the reference-shaped C++ layer: the control flow of the
// ChFSI block of main.cc:977-1237 (bounds, restart loop of filter / orth /
// Rayleigh-Ritz, stop test, rotation) run for a number of SCF iterations on a
// synthetic Kronecker-banded Hamiltonian held fixed, printing what the
// reference prints (bounds, "%dth e-vals:\t%.18e", Fermi level).
//
//   chfsi_driver [--dims RX RY RZ] [--bands BX BY BZ] [--proj P] [--k K]
//                [--m M] [--nelec NE] [--scf S] [--h H] [--seed S] [--f32]
//                [--assembly FILE [--device-assembly]]
//
//   chfsi_driver --mix IN OUT
//
// --mix runs AndersonMixing (main.cc:698-707, 862-864, 1437-1442) over a
// density sequence and exits.  IN: int64 n_grid n_nodal, int32 nsteps
// max_history, float64 alpha, jacobian[n_grid], weight[n_grid], then per step
// rho_nodal_in rho_grid_in rho_nodal_out rho_grid_out.  Step s sets rho_in by
// computeRhoIn(s) from the given rho_in, then pushes (mixed rho_in, rho_out)
// with updateRho.  OUT: per step the mixed rho_nodal_in, rho_grid_in.
//
// --assembly FILE takes H from assembly inputs instead of the synthetic
// operator: Create_Hloc + Create_Cnloc into host CSR, then setup_system; with
// --device-assembly, AX::assemble_system builds it in device memory.  FILE is
// raw little-endian: int32 rx ry rz rvx rvy rvz p 0, int32 ranks[p], float64
// truncation, c_truncation, then float64 X DX Y DY Z DZ (r x r each),
// PotX PotY PotZ (rv x r x r), core (first index fastest), projector cores,
// factors X, Y, Z (column-major, concatenated), inv_C_lm[p].
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "ttdft_host.h"

using namespace ttdft_host;

namespace {

struct AssemblyInputs {
    int32_t r[3] = {0, 0, 0};
    HlocFactors hf;
    ProjectorFactors pf;
    double truncation = 0.0, c_truncation = 0.0;
};

template <typename T>
void read_into(FILE *f, std::vector<T> &v, size_t count) {
    v.resize(count);
    if (count && fread(v.data(), sizeof(T), count, f) != count)
        throw Error(TTDFT_EBADARG, "assembly file: truncated");
}

AssemblyInputs read_assembly(const char *path) {
    FILE *f = fopen(path, "rb");
    if (!f) throw Error(TTDFT_EBADARG, std::string("cannot open ") + path);
    AssemblyInputs a;
    std::vector<int32_t> hdr;
    std::vector<double> tr;
    try {
        read_into(f, hdr, 8);
        for (int d = 0; d < 3; ++d) a.r[d] = hdr[d];
        a.hf.rankVeffX = hdr[3];
        a.hf.rankVeffY = hdr[4];
        a.hf.rankVeffZ = hdr[5];
        const int p = hdr[6];
        read_into(f, a.pf.ranks, p);
        read_into(f, tr, 2);
        a.truncation = tr[0];
        a.c_truncation = tr[1];
        const size_t rx = a.r[0], ry = a.r[1], rz = a.r[2];
        read_into(f, a.hf.MatX, rx * rx);
        read_into(f, a.hf.MatDX, rx * rx);
        read_into(f, a.hf.MatY, ry * ry);
        read_into(f, a.hf.MatDY, ry * ry);
        read_into(f, a.hf.MatZ, rz * rz);
        read_into(f, a.hf.MatDZ, rz * rz);
        read_into(f, a.hf.MatPotX, (size_t)hdr[3] * rx * rx);
        read_into(f, a.hf.MatPotY, (size_t)hdr[4] * ry * ry);
        read_into(f, a.hf.MatPotZ, (size_t)hdr[5] * rz * rz);
        read_into(f, a.hf.coreVeff, (size_t)hdr[3] * hdr[4] * hdr[5]);
        size_t nc = 0, nr = 0;
        for (int32_t q : a.pf.ranks) {
            nc += (size_t)q * q * q;
            nr += q;
        }
        read_into(f, a.pf.cores, nc);
        read_into(f, a.pf.factorX, rx * nr);
        read_into(f, a.pf.factorY, ry * nr);
        read_into(f, a.pf.factorZ, rz * nr);
        read_into(f, a.pf.inv_C_lm, p);
    } catch (...) {
        fclose(f);
        throw;
    }
    fclose(f);
    return a;
}

int run_mixing(const char *in_path, const char *out_path) {
    FILE *f = fopen(in_path, "rb");
    if (!f) throw Error(TTDFT_EBADARG, std::string("cannot open ") + in_path);
    FILE *o = nullptr;
    try {
        std::vector<int64_t> sizes;
        std::vector<int32_t> steps;
        std::vector<double> alpha, jac, wts;
        read_into(f, sizes, 2);
        read_into(f, steps, 2);
        read_into(f, alpha, 1);
        const int64_t ng = sizes[0], nn = sizes[1];
        read_into(f, jac, ng);
        read_into(f, wts, ng);
        DeviceContext ctx(0);
        double *d[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
        const int64_t len[6] = {ng, ng, nn, nn, ng, nn};  // jac wts nodal_in nodal_out grid_in grid_out
        for (int q = 0; q < 6; ++q)
            check(ttdft_malloc(ctx.handle(), sizeof(double) * std::max<int64_t>(1, len[q]),
                               (void **)&d[q]));
        double *dgrid_out = nullptr;
        check(ttdft_malloc(ctx.handle(), sizeof(double) * ng, (void **)&dgrid_out));
        check(ttdft_memcpy_h2d(ctx.handle(), d[0], jac.data(), sizeof(double) * ng));
        check(ttdft_memcpy_h2d(ctx.handle(), d[1], wts.data(), sizeof(double) * ng));
        o = fopen(out_path, "wb");
        if (!o) throw Error(TTDFT_EBADARG, std::string("cannot open ") + out_path);
        {
            AndersonMixing mixer(ctx, ng, nn, d[0], d[1], alpha[0], steps[1]);
            std::vector<double> ni, gi, no, go;
            for (int s = 0; s < steps[0]; ++s) {
                read_into(f, ni, nn);
                read_into(f, gi, ng);
                read_into(f, no, nn);
                read_into(f, go, ng);
                if (nn) check(ttdft_memcpy_h2d(ctx.handle(), d[2], ni.data(), sizeof(double) * nn));
                check(ttdft_memcpy_h2d(ctx.handle(), d[4], gi.data(), sizeof(double) * ng));
                mixer.computeRhoIn(s, nn ? d[2] : nullptr, d[4]);
                if (nn) check(ttdft_memcpy_d2h(ctx.handle(), ni.data(), d[2], sizeof(double) * nn));
                check(ttdft_memcpy_d2h(ctx.handle(), gi.data(), d[4], sizeof(double) * ng));
                if (fwrite(ni.data(), sizeof(double), nn, o) != (size_t)nn ||
                    fwrite(gi.data(), sizeof(double), ng, o) != (size_t)ng)
                    throw Error(TTDFT_EBADARG, "mix output: write failed");
                if (nn) check(ttdft_memcpy_h2d(ctx.handle(), d[3], no.data(), sizeof(double) * nn));
                check(ttdft_memcpy_h2d(ctx.handle(), dgrid_out, go.data(), sizeof(double) * ng));
                mixer.updateRho(nn ? d[2] : nullptr, d[4], nn ? d[3] : nullptr, dgrid_out, s);
                const std::vector<double> c = mixer.mixingConstants();
                printf("step %d: %zu mixing constants\n", s, c.size());
            }
        }
        for (int q = 0; q < 6; ++q) ttdft_free(ctx.handle(), d[q]);
        ttdft_free(ctx.handle(), dgrid_out);
    } catch (...) {
        fclose(f);
        if (o) fclose(o);
        throw;
    }
    fclose(f);
    fclose(o);
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    int dims[3] = {8, 8, 8}, bands[3] = {2, 2, 3};
    int nproj = 5, k = 16, m = 12, nelec = 19, nscf = 6, f32 = 0;
    double h = 2.0;
    unsigned long long seed = 20261202ULL;
    const char *assembly = nullptr, *mix_in = nullptr, *mix_out = nullptr;
    int device_assembly = 0;
    for (int i = 1; i < argc; ++i) {
        if (!strcmp(argv[i], "--dims") && i + 3 < argc) {
            for (int d = 0; d < 3; ++d) dims[d] = atoi(argv[++i]);
        } else if (!strcmp(argv[i], "--bands") && i + 3 < argc) {
            for (int d = 0; d < 3; ++d) bands[d] = atoi(argv[++i]);
        } else if (!strcmp(argv[i], "--proj") && i + 1 < argc) {
            nproj = atoi(argv[++i]);
        } else if (!strcmp(argv[i], "--k") && i + 1 < argc) {
            k = atoi(argv[++i]);
        } else if (!strcmp(argv[i], "--m") && i + 1 < argc) {
            m = atoi(argv[++i]);
        } else if (!strcmp(argv[i], "--nelec") && i + 1 < argc) {
            nelec = atoi(argv[++i]);
        } else if (!strcmp(argv[i], "--scf") && i + 1 < argc) {
            nscf = atoi(argv[++i]);
        } else if (!strcmp(argv[i], "--h") && i + 1 < argc) {
            h = atof(argv[++i]);
        } else if (!strcmp(argv[i], "--seed") && i + 1 < argc) {
            seed = strtoull(argv[++i], nullptr, 10);
        } else if (!strcmp(argv[i], "--f32")) {
            f32 = 1;
        } else if (!strcmp(argv[i], "--assembly") && i + 1 < argc) {
            assembly = argv[++i];
        } else if (!strcmp(argv[i], "--device-assembly")) {
            device_assembly = 1;
        } else if (!strcmp(argv[i], "--mix") && i + 2 < argc) {
            mix_in = argv[++i];
            mix_out = argv[++i];
        } else {
            fprintf(stderr, "unknown argument %s\n", argv[i]);
            return 2;
        }
    }
    try {
        if (mix_in) return run_mixing(mix_in, mix_out);
        DeviceContext ctx(0);
        ProjectHamiltonianSparse H;
        AssemblyInputs asmin;
        if (assembly) {
            // ---- H from assembly inputs (Create_Hloc / Create_Cnloc) ----
            asmin = read_assembly(assembly);
            for (int d = 0; d < 3; ++d) dims[d] = asmin.r[d];
            if (!device_assembly) {
                H.Create_Hloc(ctx, dims[0], dims[1], dims[2], asmin.hf, asmin.truncation);
                H.Create_Cnloc(ctx, asmin.pf, asmin.c_truncation);
                printf("assembled H: nnz=%lld nnz_c=%lld\n", (long long)H.H_loc.nnz(),
                       (long long)H.C_nloc.nnz());
            }
        } else {
            // ---- synthetic projected Hamiltonian (host CSR) ----
            ttdft_synth_params prm{dims[0], dims[1], dims[2], bands[0], bands[1], bands[2],
                                   nproj,   h,       1e-6,    seed};
            int64_t nnz = 0, nnz_c = 0;
            check(ttdft_synth_hamiltonian(&prm, &nnz, &nnz_c, nullptr, nullptr, nullptr,
                                          nullptr, nullptr, nullptr, nullptr));
            const int64_t n = (int64_t)dims[0] * dims[1] * dims[2];
            H.H_loc.rows = H.H_loc.cols = n;
            H.H_loc.rowptr.resize(n + 1);
            H.H_loc.col.resize(nnz);
            H.H_loc.val.resize(nnz);
            H.C_nloc.rows = nproj;
            H.C_nloc.cols = n;
            H.C_nloc.rowptr.resize(nproj + 1);
            H.C_nloc.col.resize(nnz_c);
            H.C_nloc.val.resize(nnz_c);
            H.inv_C_lm.resize(nproj);
            check(ttdft_synth_hamiltonian(&prm, &nnz, &nnz_c, H.H_loc.rowptr.data(),
                                          H.H_loc.col.data(), H.H_loc.val.data(),
                                          H.C_nloc.rowptr.data(), H.C_nloc.col.data(),
                                          H.C_nloc.val.data(), H.inv_C_lm.data()));
            H.rankTuckerBasisX = dims[0];
            H.rankTuckerBasisY = dims[1];
            H.rankTuckerBasisZ = dims[2];
        }
        const int64_t n = (int64_t)dims[0] * dims[1] * dims[2];

        const int dtype = f32 ? TTDFT_F32 : TTDFT_F64;
        MultTM multtm(ctx);
        EigenSolver eigenSolver(ctx);
        ChebyShevFilter chebyshevFilter;

        // ---- initial subspace: seeded N(0,1), Cholesky-QR (main.cc:807-829) ----
        std::vector<double> g((size_t)n * k), cm((size_t)n * k);
        check(ttdft_rng_normal(5, (int64_t)g.size(), g.data()));
        for (int64_t i = 0; i < n; ++i)
            for (int j = 0; j < k; ++j) cm[i + n * j] = g[i * k + j];
        DenseBlock eig_vecs(ctx, n, k, dtype);
        eig_vecs.from_colmajor(cm);
        multtm.orth(eig_vecs);

        // ---- state of main.cc:730-734, 685-688 ----
        std::vector<double> eig_vals(k, 0.0);
        double upperBoundUnWantedSpectrum = 1e16, error_in_upper_bound = 1e16;
        double lowerBoundWantedSpectrum = 0, lowerBoundUnWantedSpectrum = -1e16;
        double occupied_orbital_energy = 1.0e9;
        int occupied_orbital_number = (int)(nelec / 2.0 + 10);
        if (occupied_orbital_number > k - 1) occupied_orbital_number = k - 1;
        const int maxLanczosIter = 20;

        AX ax(1, k);
        const int copies = dtype == TTDFT_F64 ? TTDFT_HAM_F64 : (TTDFT_HAM_F64 | TTDFT_HAM_F32);
        if (assembly && device_assembly) {
            ax.assemble_system(ctx, dims[0], dims[1], dims[2], asmin.hf, asmin.truncation,
                               asmin.pf, asmin.c_truncation, copies);
            int64_t hn = 0, hnnz = 0;
            int32_t hp = 0;
            check(ttdft_ham_info(ax.handle(), &hn, &hnnz, &hp));
            printf("assembled H on the device: nnz=%lld p=%d\n", (long long)hnnz, hp);
        } else {
            ax.setup_system(ctx, H, copies);
        }
        DenseBlock HP(ctx, n, k, dtype);
        SeqDenseMatrix Q{k, k, std::vector<double>((size_t)k * k)};
        double *dQ = nullptr;
        check(ttdft_malloc(ctx.handle(), sizeof(double) * k * k, (void **)&dQ));

        for (int scfIter = 0; scfIter < nscf; ++scfIter) {
            printf("====================================\n       Entering SCF %d\n", scfIter);
            const int restart_times = scfIter == 0 ? 10 : 3;
            double temp_lower_bound = 0.0;
            if (error_in_upper_bound > 0.01) {
                double current_spectrum;
                if (assembly && device_assembly)  // H exists only on the device
                    eigenSolver.computeUpperBoundWithLanczos(ax, maxLanczosIter,
                                                             current_spectrum, temp_lower_bound);
                else  // main.cc:989-992, the reference's own call
                    eigenSolver.computeUpperBoundWithLanczos(H, maxLanczosIter,
                                                             current_spectrum, temp_lower_bound);
                error_in_upper_bound =
                    (upperBoundUnWantedSpectrum - current_spectrum) / upperBoundUnWantedSpectrum;
                upperBoundUnWantedSpectrum =
                    error_in_upper_bound > 0.01 ? current_spectrum : current_spectrum * 1.05;
            }
            if (scfIter == 0) {
                lowerBoundWantedSpectrum = temp_lower_bound;
                lowerBoundUnWantedSpectrum = 0.0;
            } else {
                lowerBoundWantedSpectrum = eig_vals[0];
                lowerBoundUnWantedSpectrum = eig_vals[k - 1];
            }
            for (int restart_iter = 0; restart_iter < restart_times; ++restart_iter) {
                printf("upperBoundUnWantedSpectrum is %.18e.\n", upperBoundUnWantedSpectrum);
                printf("lowerBoundUnWantedSpectrum is %.18e.\n", lowerBoundUnWantedSpectrum);
                printf("lowerBoundWantedSpectrum is %.18e.\n", lowerBoundWantedSpectrum);
                chebyshevFilter.computeFilteredSubspace(H, ax, eig_vecs, (int)n, m,
                                                        lowerBoundUnWantedSpectrum,
                                                        upperBoundUnWantedSpectrum,
                                                        lowerBoundWantedSpectrum, 1);
                multtm.orth(eig_vecs);
                ax.perform_ax(eig_vecs, HP);
                SeqDenseMatrix PtHP{k, k, {}};
                multtm.mult(eig_vecs, HP, PtHP.data);
                // main.cc:1109-1114
                eigenSolver.computeEigenPairsSeq(PtHP, k, k, eig_vals, Q, true);
                for (int i = 0; i < k; ++i) printf("%dth e-vals:\t%.18e\n", i, eig_vals[i]);
                printf("Projected Eigenvalues of Chebyshev Pass %d\n", restart_iter);
                lowerBoundWantedSpectrum = eig_vals[0];
                lowerBoundUnWantedSpectrum = eig_vals[k - 1];
                const double cur = eig_vals[occupied_orbital_number];
                const double err = std::fabs((cur - occupied_orbital_energy) / occupied_orbital_energy);
                occupied_orbital_energy = cur;
                if (err < 1.0e-3) {
                    printf("the iteration converged at step %d.\n", restart_iter);
                    break;
                }
            }
            // rotation X <- X Q (main.cc:1185-1207)
            check(ttdft_memcpy_h2d(ctx.handle(), dQ, Q.data.data(), sizeof(double) * k * k));
            check(ttdft_rotate(ctx.handle(), eig_vecs.data(), eig_vecs.ld(), n, k, dQ, dtype,
                               dtype == TTDFT_F64 ? TTDFT_GEMM_F64 : TTDFT_GEMM_BF16X3));
            const double ef = computeFermiEnergy(eig_vals.data(), k, nelec, 3.166811429e-06, 500.0);
            printf("0 :Fermi Energy: %.18e\n", ef);
        }
        ttdft_free(ctx.handle(), dQ);
    } catch (const Error &e) {
        fprintf(stderr, "chfsi_driver: %s\n", e.what());
        return 1;
    }
    return 0;
}
