// A caller of the C++ drop-in (include/cntr_amd.hpp) written against the reference's
// cntr:: API: a two-orbital system coupled to a two-orbital bath by a complex,
// non-Hermitian hopping V, solved as an embedding problem.
//
//   1. G_exact: the system block of green_from_H of the 4x4 Hamiltonian [[Hs, V], [V^+, Hb]].
//   2. Sigma = V g_b V^+ with g_b = green_from_H(Hb), block by block on the host.
//   3. Dyson form: dyson_mat, dyson_start, dyson_timestep (Integrator overloads), with the
//      text checkpoint at precision 17 taken at timestep NT1, read into a second G and
//      continued from there: both runs must agree bit for bit.
//   4. Integral form: G0 = green_from_H(Hs), F = -G0*Sigma, F^cc = -Sigma*G0 by
//      convolution_timestep, vie2_start, vie2_timestep.
//   5. Timestep objects: at every timestep, herm_matrix_timestep copies of G and G_exact
//      give the same distance_norm2 as the herm_matrix pair; Bubble1 into a
//      herm_matrix_timestep equals Bubble1 into a herm_matrix at that timestep; a
//      herm_matrix_timestep_view sum reproduces 2 G.
//   6. green_from_H with a function: a constant H(t) reproduces the constant-H form, and the
//      timestep form at mu = 0 reproduces the whole form's timestep.
//   7. size1 3: a three-level system and a three-level bath against the system block of the
//      exact 6x6 green_from_H, with dyson_mat by conjugate gradients (CNTR_MAT_CG).
//
// Usage: dropin_driver DIR [out.bin].  out.bin receives the Dyson-form G (ret, les, tv, mat
// as complex doubles, packed layout) for the oracle comparison.  Prints the errors and exits
// 0 when every check passes.
#include <cmath>
#include <complex>
#include <cstdio>
#include <string>
#include <vector>

#include "cntr_amd.hpp"

typedef std::complex<double> cplx;
typedef cntr::herm_matrix<double> GREEN;
typedef cntr::herm_matrix_timestep<double> GREEN_TSTP;

static const int NT = 80, NTAU = 300, K = 5, NT1 = 37;
static const double H = 0.05, BETA = 5.0, MU = -0.1;

// z = a b c^+ for 2x2 row-major blocks
static void sandwich(cplx *z, const cplx *a, const cplx *b) {
    cplx t[4];
    for (int r = 0; r < 2; r++)
        for (int c = 0; c < 2; c++) t[2 * r + c] = a[2 * r] * b[c] + a[2 * r + 1] * b[2 + c];
    for (int r = 0; r < 2; r++)
        for (int c = 0; c < 2; c++) z[2 * r + c] = t[2 * r] * std::conj(a[2 * c]) + t[2 * r + 1] * std::conj(a[2 * c + 1]);
}

static double max_diff(const GREEN &a, const GREEN &b, int t) {
    double m = 0.0;
    cdmatrix x, y;
    auto upd = [&]() {
        for (int e = 0; e < x.rows() * x.cols(); e++) m = std::max(m, std::abs(x.data()[e] - y.data()[e]));
    };
    if (t == -1) {
        for (int q = 0; q <= NTAU; q++) {
            a.get_mat(q, x);
            b.get_mat(q, y);
            upd();
        }
        return m;
    }
    for (int j = 0; j <= t; j++) {
        a.get_ret(t, j, x);
        b.get_ret(t, j, y);
        upd();
        a.get_les(j, t, x);
        b.get_les(j, t, y);
        upd();
    }
    for (int q = 0; q <= NTAU; q++) {
        a.get_tv(t, q, x);
        b.get_tv(t, q, y);
        upd();
    }
    return m;
}

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s DIR [out.bin]\n", argv[0]);
        return 2;
    }
    const std::string ckpt = std::string(argv[1]) + "/dropin_G.dat";
    const cplx I(0.0, 1.0);
    cdmatrix hs(2, 2), hb(2, 2), h4(4, 4);
    hs(0, 0) = -1.0;
    hs(0, 1) = 0.3 + 0.2 * I;
    hs(1, 0) = 0.3 - 0.2 * I;
    hs(1, 1) = 0.5;
    hb(0, 0) = 0.8;
    hb(0, 1) = -0.25 * I;
    hb(1, 0) = 0.25 * I;
    hb(1, 1) = -0.4;
    const cplx V[4] = {0.3, 0.1 + 0.2 * I, -0.15 * I, 0.25};
    for (int r = 0; r < 2; r++)
        for (int c = 0; c < 2; c++) {
            h4(r, c) = hs(r, c);
            h4(r + 2, c + 2) = hb(r, c);
            h4(r, c + 2) = V[2 * r + c];
            h4(c + 2, r) = std::conj(V[2 * r + c]);
        }
    int failed = 0;

    // 1. exact system block
    GREEN G4(NT, NTAU, 4, FERMION), Gx(NT, NTAU, 2, FERMION);
    cntr::green_from_H(G4, MU, h4, BETA, H);
    for (int t = -1; t <= NT; t++)
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++) Gx.set_matrixelement(t, a, b, G4, a, b);

    // 2. Sigma = V g_b V^+
    GREEN gb(NT, NTAU, 2, FERMION), Sigma(NT, NTAU, 2, FERMION);
    cntr::green_from_H(gb, MU, hb, BETA, H);
    for (int q = 0; q <= NTAU; q++) sandwich(Sigma.matptr(q), V, gb.matptr(q));
    for (int t = 0; t <= NT; t++) {
        for (int j = 0; j <= t; j++) {
            sandwich(Sigma.retptr(t, j), V, gb.retptr(t, j));
            sandwich(Sigma.lesptr(j, t), V, gb.lesptr(j, t));
        }
        for (int q = 0; q <= NTAU; q++) sandwich(Sigma.tvptr(t, q), V, gb.tvptr(t, q));
    }
    cntr::function<double> hf(NT, 2);
    hf.set_constant(hs);

    // 3. Dyson form, with a checkpoint at NT1 and a resumed second run
    GREEN G(NT, NTAU, 2, FERMION), Gr(NT, NTAU, 2, FERMION);
    cntr::dyson_mat(G, Sigma, MU, hf, integration::I<double>(K), BETA);
    cntr::dyson_start(G, MU, hf, Sigma, integration::I<double>(K), BETA, H);
    double worst = 0.0, dist_pair = 0.0, dist_tstp = 0.0, bubble_diff = 0.0, view_diff = 0.0;
    GREEN Pol(NT, NTAU, 2, BOSON);
    for (int t = K + 1; t <= NT; t++) {
        cntr::dyson_timestep(t, G, MU, hf, Sigma, integration::I<double>(K), BETA, H);
        if (t == NT1) {
            G.print_to_file(ckpt.c_str(), 17);
            Gr.read_from_file(NT1, ckpt.c_str());
        }
        // 5. timestep objects at this timestep
        GREEN_TSTP gt, xt(t, NTAU, 2, FERMION);
        G.get_timestep(t, gt);
        Gx.get_timestep(t, xt);
        dist_pair += cntr::distance_norm2(t, G, Gx);
        dist_tstp += cntr::distance_norm2(t, gt, xt);
        GREEN_TSTP pt(t, NTAU, 2, BOSON);
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++) {
                cntr::Bubble1(t, pt, a, b, gt, a, b, G, b, a);
                cntr::Bubble1(t, Pol, a, b, G, a, b, G, b, a);
            }
        GREEN_TSTP pf;
        Pol.get_timestep(t, pf);
        for (int e = 0; e < pf.total_size() * 4; e++)
            bubble_diff = std::max(bubble_diff, std::abs(pf.retptr(0)[e] - pt.retptr(0)[e]));
        GREEN_TSTP twice(t, NTAU, 2, FERMION);
        cntr::herm_matrix_timestep_view<double> v(twice);
        v.incr_timestep(G, 1.0);
        v.incr_timestep(gt, 1.0);
        for (int e = 0; e < twice.total_size() * 4; e++)
            view_diff = std::max(view_diff, std::abs(twice.retptr(0)[e] - 2.0 * gt.retptr(0)[e]));
    }
    for (int t = -1; t <= NT; t++) worst = std::max(worst, max_diff(G, Gx, t));
    std::printf("dyson: max |G - G_exact| %.3e\n", worst);
    failed += !(worst < 1e-6);
    std::printf("timestep objects: distance %.6e (herm_matrix) vs %.6e (timestep), bubble diff %.1e, "
                "view diff %.1e\n", dist_pair, dist_tstp, bubble_diff, view_diff);
    failed += !(dist_pair == dist_tstp) + !(bubble_diff == 0.0) + !(view_diff == 0.0);

    // resumed run: Matsubara and timesteps 0..NT1 from the file, then the same steps
    for (int t = NT1 + 1; t <= NT; t++)
        cntr::dyson_timestep(t, Gr, MU, hf, Sigma, integration::I<double>(K), BETA, H);
    double resume = 0.0;
    for (int t = -1; t <= NT; t++) resume = std::max(resume, max_diff(G, Gr, t));
    std::printf("resume from the timestep-%d checkpoint: max difference %.1e\n", NT1, resume);
    failed += !(resume == 0.0);

    if (argc > 2) {
        FILE *f = std::fopen(argv[2], "wb");
        if (!f) return 3;
        const size_t ntri = (size_t)(NT + 1) * (NT + 2) / 2 * 4, ntv = (size_t)(NT + 1) * (NTAU + 1) * 4;
        const GREEN &cg = G;
        std::fwrite(cg.retptr(0, 0), sizeof(cplx), ntri, f);
        std::fwrite(cg.lesptr(0, 0), sizeof(cplx), ntri, f);
        std::fwrite(cg.tvptr(0, 0), sizeof(cplx), ntv, f);
        std::fwrite(cg.matptr(0), sizeof(cplx), (size_t)(NTAU + 1) * 4, f);
        std::fclose(f);
    }

    // 4. integral form
    GREEN Gv(NT, NTAU, 2, FERMION), G0(NT, NTAU, 2, FERMION), F(NT, NTAU, 2, FERMION),
        Fcc(NT, NTAU, 2, FERMION);
    cntr::green_from_H(G0, MU, hs, BETA, H);
    auto conv = [&](int t) {
        cntr::convolution_timestep(t, F, G0, Sigma, integration::I<double>(K), BETA, H);
        cntr::convolution_timestep(t, Fcc, Sigma, G0, integration::I<double>(K), BETA, H);
        F.smul(t, -1.0);
        Fcc.smul(t, -1.0);
    };
    conv(-1);
    cntr::vie2_mat(Gv, F, Fcc, G0, BETA, K);
    for (int t = 0; t <= K; t++) conv(t);
    cntr::vie2_start(Gv, F, Fcc, G0, integration::I<double>(K), BETA, H);
    for (int t = K + 1; t <= NT; t++) {
        conv(t);
        cntr::vie2_timestep(t, Gv, F, Fcc, G0, integration::I<double>(K), BETA, H);
    }
    double wv = 0.0;
    for (int t = -1; t <= NT; t++) wv = std::max(wv, max_diff(Gv, Gx, t));
    std::printf("vie2: max |G - G_exact| %.3e\n", wv);
    failed += !(wv < 1e-8);

    // 6. green_from_H with a function: constant H(t) = Hs reproduces the constant-H form;
    //    the timestep form at mu = 0 reproduces the whole form's timestep
    {
        GREEN Ga(NT, NTAU, 2, FERMION), Gb(NT, NTAU, 2, FERMION), Gc(NT, NTAU, 2, FERMION);
        cntr::green_from_H(Ga, MU, hs, BETA, H);
        cntr::green_from_H(Gb, MU, hf, BETA, H);
        double d = 0.0;
        for (int t = -1; t <= NT; t++) d = std::max(d, max_diff(Ga, Gb, t));
        cntr::function<double> hv(NT, 2);
        for (int t = -1; t <= NT; t++) {
            cdmatrix m = hs;
            const double f = 0.3 * std::sin(0.7 * std::max(t, 0) * H);
            m(0, 1) += f;
            m(1, 0) += f;
            hv.set_value(t, m);
        }
        cntr::green_from_H(Gb, 0.0, hv, BETA, H, K, 4);
        cntr::green_from_H(NT1, Gc, 0.0, hv, BETA, H, true, K, 4);
        const double dt = max_diff(Gb, Gc, NT1);
        std::printf("green_from_H(function): constant H vs constant form %.1e, timestep form vs whole %.1e\n",
                    d, dt);
        failed += !(d < 1e-12) + !(dt < 1e-13);
    }

    // 7. size1 3 (the library runs it embedded in 4), Matsubara branch by conjugate gradients
    {
        const int N3 = 40, S3 = 3;
        cdmatrix h3(S3, S3), b3(S3, S3), h6(2 * S3, 2 * S3);
        cplx V3[S3 * S3];
        for (int r = 0; r < S3; r++)
            for (int c = 0; c < S3; c++) {
                h3(r, c) = r == c ? -0.8 + 0.6 * r : cplx(0.2, 0.1 * (c - r));
                b3(r, c) = r == c ? 0.9 - 0.5 * r : cplx(-0.15, 0.05 * (r - c));
                V3[r * S3 + c] = cplx(0.25 - 0.05 * r, 0.07 * c);
            }
        for (int r = 0; r < S3; r++)
            for (int c = 0; c < S3; c++) {
                h6(r, c) = h3(r, c);
                h6(r + S3, c + S3) = b3(r, c);
                h6(r, c + S3) = V3[r * S3 + c];
                h6(c + S3, r) = std::conj(V3[r * S3 + c]);
            }
        GREEN G6(N3, NTAU, 2 * S3, FERMION), X3(N3, NTAU, S3, FERMION), g3(N3, NTAU, S3, FERMION);
        GREEN Sg3(N3, NTAU, S3, FERMION), G3(N3, NTAU, S3, FERMION);
        cntr::green_from_H(G6, MU, h6, BETA, H);
        cntr::green_from_H(g3, MU, b3, BETA, H);
        for (int t = -1; t <= N3; t++)
            for (int a = 0; a < S3; a++)
                for (int b = 0; b < S3; b++) X3.set_matrixelement(t, a, b, G6, a, b);
        // Sigma = V g V^+ block by block
        auto sandwich3 = [&](cplx *z, const cplx *g) {
            cplx t[S3 * S3];
            for (int r = 0; r < S3; r++)
                for (int c = 0; c < S3; c++) {
                    t[r * S3 + c] = 0.0;
                    for (int l = 0; l < S3; l++) t[r * S3 + c] += V3[r * S3 + l] * g[l * S3 + c];
                }
            for (int r = 0; r < S3; r++)
                for (int c = 0; c < S3; c++) {
                    z[r * S3 + c] = 0.0;
                    for (int l = 0; l < S3; l++) z[r * S3 + c] += t[r * S3 + l] * std::conj(V3[c * S3 + l]);
                }
        };
        for (int q = 0; q <= NTAU; q++) sandwich3(Sg3.matptr(q), g3.matptr(q));
        for (int t = 0; t <= N3; t++) {
            for (int j = 0; j <= t; j++) {
                sandwich3(Sg3.retptr(t, j), g3.retptr(t, j));
                sandwich3(Sg3.lesptr(j, t), g3.lesptr(j, t));
            }
            for (int q = 0; q <= NTAU; q++) sandwich3(Sg3.tvptr(t, q), g3.tvptr(t, q));
        }
        cntr::function<double> hf3(N3, S3);
        hf3.set_constant(h3);
        cntr::dyson_mat(G3, Sg3, MU, hf3, integration::I<double>(K), BETA, CNTR_MAT_CG);
        const double wm = max_diff(G3, X3, -1);
        cntr::dyson_start(G3, MU, hf3, Sg3, integration::I<double>(K), BETA, H);
        for (int t = K + 1; t <= N3; t++)
            cntr::dyson_timestep(t, G3, MU, hf3, Sg3, integration::I<double>(K), BETA, H);
        double w3 = 0.0;
        for (int t = 0; t <= N3; t++) w3 = std::max(w3, max_diff(G3, X3, t));
        std::printf("size1 3: Matsubara (conjugate gradients) max |G - G_exact| %.3e, real time %.3e\n", wm, w3);
        failed += !(wm < 1e-10) + !(w3 < 1e-7);
    }

    std::printf("%s\n", failed ? "FAILED" : "passed");
    return failed ? 1 : 0;
}
