// The reference's Kadanoff-Baym test (test/dyson.cpp:23-119) on the C++ drop-in
// (cntr_amd.hpp): a level eps1 coupled (lam) to a second level eps2, solved as an
// embedding problem with Sigma = lam^2 g(eps2), in the integro-differential form
// (dyson_mat, dyson_start, dyson_timestep) and in the integral form (convolution_timestep,
// vie2_start, vie2_timestep), each compared with the exact 2x2 solution at the reference's
// tolerances: err/Nt < 1e-3 and < 1e-6.  The calls are the reference test's, with
// integration::I<double>(SolverOrder) and the cdmatrix Hamiltonians.
//
// Usage: dyson_embedding [out.bin]
// out.bin receives G of the Dyson form (ret, les, tv, mat as complex doubles, in the
// packed layout) for a parity check against the oracle.  Exit status 0 when both errors
// are within the tolerances.
#include <cmath>
#include <complex>
#include <cstdio>

#include "cntr_amd.hpp"

#define GREEN cntr::herm_matrix<double>
#define CFUNC cntr::function<double>

int main(int argc, char **argv) {
    const int fermion = -1;
    const int Nst = 2;
    const double eps1 = -1.0, eps2 = 1.0, lam = 0.2, mu = -0.1, beta = 10.0;
    const int Ntau = 400, Nt = 200;
    const double dt = 0.05;
    const int SolverOrder = 5;
    const double tol_coarse = 1.0e-3, tol_tight = 1.0e-6;
    const std::complex<double> I(0.0, 1.0);
    int tstp;

    cdmatrix h2x2(Nst, Nst), h1x1(1, 1);
    h2x2(0, 0) = eps1;
    h2x2(1, 1) = eps2;
    h2x2(0, 1) = I * lam;
    h2x2(1, 0) = -I * lam;
    h1x1(0, 0) = eps1;
    CFUNC hfunc(Nt, 1);
    hfunc.set_constant(h1x1);

    // exact Green's function of the 2x2 problem and its (0,0) element
    GREEN G2x2(Nt, Ntau, Nst, fermion);
    cntr::green_from_H(G2x2, mu, h2x2, beta, dt);
    GREEN G_exact(Nt, Ntau, 1, fermion);
    for (tstp = -1; tstp <= Nt; tstp++) G_exact.set_matrixelement(tstp, 0, 0, G2x2, 0, 0);

    // embedding self-energy
    GREEN Sigma(Nt, Ntau, 1, fermion);
    cdmatrix h22(1, 1);
    h22(0, 0) = eps2;
    cntr::green_from_H(Sigma, mu, h22, beta, dt);
    for (tstp = -1; tstp <= Nt; tstp++) Sigma.smul(tstp, lam * lam);

    // integro-differential form
    GREEN G_approx(Nt, Ntau, 1, fermion);
    cntr::dyson_mat(G_approx, Sigma, mu, hfunc, integration::I<double>(SolverOrder), beta);
    cntr::dyson_start(G_approx, mu, hfunc, Sigma, integration::I<double>(SolverOrder), beta, dt);
    for (tstp = SolverOrder + 1; tstp <= Nt; tstp++)
        cntr::dyson_timestep(tstp, G_approx, mu, hfunc, Sigma, integration::I<double>(SolverOrder), beta, dt);
    double err_dyson = 0.0;
    for (tstp = 0; tstp <= Nt; tstp++) err_dyson += cntr::distance_norm2(tstp, G_exact, G_approx);

    if (argc > 1) {
        FILE *f = std::fopen(argv[1], "wb");
        if (!f) return 3;
        const size_t ntri = (size_t)(Nt + 1) * (Nt + 2) / 2, ntv = (size_t)(Nt + 1) * (Ntau + 1);
        std::fwrite(G_approx.retptr(0, 0), sizeof(std::complex<double>), ntri, f);
        std::fwrite(G_approx.lesptr(0, 0), sizeof(std::complex<double>), ntri, f);
        std::fwrite(G_approx.tvptr(0, 0), sizeof(std::complex<double>), ntv, f);
        std::fwrite(G_approx.matptr(0), sizeof(std::complex<double>), Ntau + 1, f);
        std::fclose(f);
    }

    // integral form
    GREEN G_vie(Nt, Ntau, 1, fermion);
    cntr::dyson_mat(G_vie, Sigma, mu, hfunc, integration::I<double>(SolverOrder), beta);
    GREEN G0(Nt, Ntau, 1, fermion), G0xSGM(Nt, Ntau, 1, fermion), SGMxG0(Nt, Ntau, 1, fermion);
    cntr::green_from_H(G0, mu, h1x1, beta, dt);
    for (tstp = 0; tstp <= SolverOrder; tstp++) {
        cntr::convolution_timestep(tstp, G0xSGM, G0, Sigma, integration::I<double>(SolverOrder), beta, dt);
        cntr::convolution_timestep(tstp, SGMxG0, Sigma, G0, integration::I<double>(SolverOrder), beta, dt);
        G0xSGM.smul(tstp, -1);
        SGMxG0.smul(tstp, -1);
    }
    cntr::vie2_start(G_vie, G0xSGM, SGMxG0, G0, integration::I<double>(SolverOrder), beta, dt);
    for (tstp = SolverOrder + 1; tstp <= Nt; tstp++) {
        cntr::convolution_timestep(tstp, G0xSGM, G0, Sigma, integration::I<double>(SolverOrder), beta, dt);
        cntr::convolution_timestep(tstp, SGMxG0, Sigma, G0, integration::I<double>(SolverOrder), beta, dt);
        G0xSGM.smul(tstp, -1);
        SGMxG0.smul(tstp, -1);
        cntr::vie2_timestep(tstp, G_vie, G0xSGM, SGMxG0, G0, integration::I<double>(SolverOrder), beta, dt);
    }
    double err_vie2 = 0.0;
    for (tstp = 0; tstp <= Nt; tstp++) err_vie2 += cntr::distance_norm2(tstp, G_exact, G_vie);

    std::printf("dyson err/Nt %.3e (tol %.0e)  vie2 err/Nt %.3e (tol %.0e)\n", err_dyson / Nt, tol_coarse,
                err_vie2 / Nt, tol_tight);
    return (err_dyson / Nt < tol_coarse && err_vie2 / Nt < tol_tight) ? 0 : 1;
}
