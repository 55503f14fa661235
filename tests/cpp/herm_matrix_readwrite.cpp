// The reference's text read/write test (test/herm_matrix_readwrite.cpp:9-52) on the C++
// drop-in (cntr_amd.hpp): G1 = green_from_H of a 2x2 Hamiltonian, written at precision 10,
// read back whole and up to timestep nt1 = 50, and compared with distance_norm2 summed over
// the timesteps read (< 1e-6).  A third section writes at precision 17 and requires the
// read-back to be exact, and a fourth reads into an object of another shape, which the
// reference resizes.
//
// Usage: herm_matrix_readwrite DIR   (the files are written under DIR)
// Exit status 0 when every section passes.
#include <complex>
#include <cstdio>
#include <string>

#include "cntr_amd.hpp"

#define GREEN cntr::herm_matrix<double>

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s DIR\n", argv[0]);
        return 2;
    }
    const std::string file = std::string(argv[1]) + "/herm_matrix_plain.dat";
    const int prec = 10, size = 2, nt = 100, ntau = 50;
    const double eps = 1e-6, h = 0.01, mu = 0.0, beta = 10.0;
    const double eps1 = -0.4, eps2 = 0.6, lam = 0.1;
    const std::complex<double> I(0.0, 1.0);
    cdmatrix h0(2, 2);
    GREEN G1(nt, ntau, size, -1);
    h0(0, 0) = eps1;
    h0(1, 1) = eps2;
    h0(0, 1) = I * lam;
    h0(1, 0) = -I * lam;
    cntr::green_from_H(G1, mu, h0, beta, h);
    int failed = 0;

    {  // read/write
        GREEN G2(nt, ntau, size, -1);
        double err = 0.0;
        G1.print_to_file(file.c_str(), prec);
        G2.read_from_file(file.c_str());
        for (int tstp = -1; tstp <= nt; tstp++) err += cntr::distance_norm2(tstp, G1, G2);
        std::printf("read/write: err %.3e\n", err);
        failed += !(err < eps);
    }
    {  // read/write up to given time step
        GREEN G2(nt, ntau, size, -1);
        const int nt1 = 50;
        double err = 0.0, rest = 0.0;
        G1.print_to_file(file.c_str(), prec);
        G2.read_from_file(nt1, file.c_str());
        for (int tstp = -1; tstp <= nt1; tstp++) err += cntr::distance_norm2(tstp, G1, G2);
        for (int tstp = nt1 + 1; tstp <= nt; tstp++) rest += std::abs(G2.retptr(tstp, 0)[0]);
        std::printf("read/write up to nt1: err %.3e, timesteps after nt1 untouched: %s\n", err,
                    rest == 0.0 ? "yes" : "no");
        failed += !(err < eps) + (rest != 0.0);
    }
    {  // 17 digits: exact
        GREEN G2(nt, ntau, size, -1);
        double err = 0.0;
        G1.print_to_file(file.c_str(), 17);
        G2.read_from_file(file.c_str());
        for (int tstp = -1; tstp <= nt; tstp++) err += cntr::distance_norm2(tstp, G1, G2);
        std::printf("precision 17: err %.3e\n", err);
        failed += !(err == 0.0);
    }
    {  // another shape: the reader resizes
        GREEN G2(10, 7, 1, +1);
        double err = 0.0;
        G1.print_to_file(file.c_str(), 17);
        G2.read_from_file(file.c_str());
        const bool shape = G2.nt() == nt && G2.ntau() == ntau && G2.size1() == size && G2.sig() == -1;
        if (shape)
            for (int tstp = -1; tstp <= nt; tstp++) err += cntr::distance_norm2(tstp, G1, G2);
        std::printf("resize: shape %s, err %.3e\n", shape ? "ok" : "wrong", err);
        failed += !shape + !(err == 0.0);
    }
    std::printf("%s\n", failed ? "FAILED" : "passed");
    return failed ? 1 : 0;
}
