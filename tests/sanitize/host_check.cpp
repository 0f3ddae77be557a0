// host_check.cpp -- the library's host-only C++ under the CPU sanitizers
// (AddressSanitizer + UndefinedBehaviorSanitizer, or ThreadSanitizer): the
// per-node stencil drop-ins (weno, WENO_D*, second derivatives and
// curvatures) at every node of small grids, faces and corners included (the
// mirror boundary rules); the threaded row (de)interleaving of the operator
// transfers; and the reference file formats' readers and writers. Every result
// is also checked, so a run proves the code both clean and right.
//
//   host_check <scratch directory>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

#include "../../helmholtz-3d-torrone_amd/csrc/ibim_dropin.hpp"
#include "../../helmholtz-3d-torrone_amd/csrc/ibim_hoststage.hpp"
#include "../../include/ibim.h"

namespace {

int failures = 0;

void expect(bool ok, const char* what) {
    if (!ok) {
        fprintf(stderr, "FAILED: %s\n", what);
        ++failures;
    }
}

// a grid allocated row by row, as the reference's driver does
struct Grid {
    int S;
    double*** g;
    explicit Grid(int s) : S(s) {
        g = new double**[S];
        for (int i = 0; i < S; ++i) {
            g[i] = new double*[S];
            for (int j = 0; j < S; ++j) g[i][j] = new double[S];
        }
    }
    ~Grid() {
        for (int i = 0; i < S; ++i) {
            for (int j = 0; j < S; ++j) delete[] g[i][j];
            delete[] g[i];
        }
        delete[] g;
    }
};

void stencils() {
    const int n = 12, S = n + 1;
    const double h = 4.0 / n;
    Grid u(S), lin(S);
    std::mt19937_64 rng(7);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    for (int i = 0; i < S; ++i)
        for (int j = 0; j < S; ++j)
            for (int k = 0; k < S; ++k) {
                const double y = -2 + i * h, x = -2 + j * h, z = -2 + k * h;
                u.g[i][j][k] = 1.0 - std::sqrt(x * x + y * y + z * z) + 0.01 * U(rng);
                lin.g[i][j][k] = 0.3 * x - 0.2 * y + 0.1 * z;  // linear: one-sided WENO exact inside
            }
    for (int order : {3, 5})
        for (int axis = 0; axis < 3; ++axis)
            for (int i = 0; i < S; ++i)
                for (int j = 0; j < S; ++j)
                    for (int k = 0; k < S; ++k) {
                        double p, m;
                        ibim::dropin_weno_axis(axis, u.g, i, j, k, n, h, &p, &m, order);
                        expect(std::isfinite(p) && std::isfinite(m), "WENO derivative finite");
                        ibim::dropin_weno_axis(axis, lin.g, i, j, k, n, h, &p, &m, order);
                        const int c = axis == 0 ? j : (axis == 1 ? i : k);
                        const double slope = axis == 0 ? 0.3 : (axis == 1 ? -0.2 : 0.1);
                        if (c >= 3 && c <= n - 3)
                            expect(std::fabs(p - slope) < 1e-12 && std::fabs(m - slope) < 1e-12,
                                   "WENO exact on linear data");
                    }
    for (int i = 0; i < S; ++i)
        for (int j = 0; j < S; ++j)
            for (int k = 0; k < S; ++k) {
                double d6[6], mc, gc;
                ibim::dropin_second(u.g, i, j, k, n, h, d6);
                ibim::dropin_curvatures(d6, &mc, &gc);
                expect(std::isfinite(mc) && std::isfinite(gc), "curvatures finite");
                ibim::dropin_second(lin.g, i, j, k, n, h, d6);
                if (i > 1 && i < n - 1 && j > 1 && j < n - 1 && k > 1 && k < n - 1)
                    for (double d : d6) expect(std::fabs(d) < 1e-9, "second derivatives of linear data vanish");
            }
    double f[7] = {0.1, 0.3, -0.2, 0.4, 0.0, -0.1, 0.2}, dp, dm;
    for (int order : {3, 5}) {
        ibim::dropin_weno(&dp, &dm, f, h, order);
        expect(std::isfinite(dp) && std::isfinite(dm), "weno finite");
    }
}

void staging() {
    // rows of several sizes through the threaded (de)interleaving, blocks large
    // enough to use every thread
    std::mt19937_64 rng(11);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    for (int count : {1, 63, 1000, 4099}) {
        const size_t ld = ((size_t)count + 63) / 64 * 64;
        const int nr = 257;
        std::vector<double> re((size_t)nr * count), im((size_t)nr * count);
        for (auto& v : re) v = U(rng);
        for (auto& v : im) v = U(rng);
        std::vector<double> d2((size_t)nr * ld * 2, -7.0);
        std::vector<float> f2((size_t)nr * ld * 2, -7.0f);
        const size_t block = (size_t)64 << 20;  // as many threads as allowed
        ibim::hoststage::for_rows(nr, block, [&](int r) {
            ibim::hoststage::interleave_row(&d2[(size_t)r * ld * 2], &re[(size_t)r * count], &im[(size_t)r * count],
                                            count, ld);
            ibim::hoststage::interleave_row(&f2[(size_t)r * ld * 2], &re[(size_t)r * count], &im[(size_t)r * count],
                                            count, ld);
        });
        std::vector<double> bre(re.size()), bim(im.size()), fre(re.size()), fim(im.size());
        ibim::hoststage::for_rows(nr, block, [&](int r) {
            ibim::hoststage::split_row(&d2[(size_t)r * ld * 2], &bre[(size_t)r * count], &bim[(size_t)r * count],
                                       count);
            ibim::hoststage::split_row(&f2[(size_t)r * ld * 2], &fre[(size_t)r * count], &fim[(size_t)r * count],
                                       count);
        });
        expect(bre == re && bim == im, "complex128 rows round trip");
        bool f_ok = true, pad_ok = true;
        for (size_t e = 0; e < re.size(); ++e)
            f_ok = f_ok && fre[e] == (double)(float)re[e] && fim[e] == (double)(float)im[e];
        for (int r = 0; r < nr; ++r)
            for (size_t j = count; j < ld; ++j)
                pad_ok = pad_ok && d2[((size_t)r * ld + j) * 2] == 0.0 && f2[((size_t)r * ld + j) * 2 + 1] == 0.0f;
        expect(f_ok, "complex64 rows round trip");
        expect(pad_ok, "padding columns zeroed");
    }
}

void formats(const std::string& dir) {
    const int n = 6, S = n + 1, count = 37;
    std::vector<double> level((size_t)S * S * S), back(level.size());
    for (size_t e = 0; e < level.size(); ++e) level[e] = std::sin(0.37 * (double)e) * 1.25;
    const std::string dpath = dir + "/distance6.txt";
    expect(ibim_write_distance_file(dpath.c_str(), n, level.data()) == IBIM_OK, "write distance file");
    expect(ibim_read_distance_file(dpath.c_str(), n, back.data()) == IBIM_OK, "read distance file");
    double worst = 0.0;
    for (size_t e = 0; e < level.size(); ++e) worst = std::fmax(worst, std::fabs(back[e] - level[e]));
    expect(worst < 1e-14, "distance file round trip");
    expect(ibim_read_distance_file(dpath.c_str(), n + 1, back.data()) != IBIM_OK, "short distance file is an error");
    std::vector<double> bre(count), bim(count), rre(count), rim(count);
    for (int p = 0; p < count; ++p) {
        bre[p] = std::cos(0.1 * p);
        bim[p] = -std::sin(0.2 * p);
    }
    const std::string r1 = dir + "/Rel_bN6k2.txt", r2 = dir + "/Rel_bimgN6k2.txt";
    expect(ibim_write_rhs_text(r1.c_str(), r2.c_str(), count, bre.data(), bim.data()) == IBIM_OK, "write rhs");
    expect(ibim_read_rhs_text(r1.c_str(), r2.c_str(), count, rre.data(), rim.data()) == IBIM_OK, "read rhs");
    for (int p = 0; p < count; ++p)
        expect(std::fabs(rre[p] - bre[p]) <= 5e-16 && std::fabs(rim[p] - bim[p]) <= 5e-16, "rhs %.15lf quantum");
    expect(ibim_read_rhs_text(r1.c_str(), r2.c_str(), count + 1, rre.data(), rim.data()) != IBIM_OK,
           "rhs size mismatch is an error");
    const std::string d1 = dir + "/Rel_DensityN6k2.txt", d2 = dir + "/Rel_DensityimgN6k2.txt";
    expect(ibim_write_density_text(d1.c_str(), d2.c_str(), count, bre.data(), bim.data()) == IBIM_OK,
           "write density");
    expect(ibim_read_density_text(d1.c_str(), d2.c_str(), count, rre.data(), rim.data()) == IBIM_OK,
           "read density");
    for (int p = 0; p < count; ++p)
        expect(std::fabs(rre[p] - bre[p]) <= 5e-16 && std::fabs(rim[p] - bim[p]) <= 5e-16, "density round trip");
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 2) {
        fprintf(stderr, "usage: %s scratch_dir\n", argv[0]);
        return 2;
    }
    stencils();
    staging();
    formats(argv[1]);
    if (failures) {
        fprintf(stderr, "%d checks failed\n", failures);
        return 1;
    }
    printf("host_check: stencils, threaded staging and file formats clean\n");
    return 0;
}
