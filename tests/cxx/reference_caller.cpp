// reference_caller.cpp -- a C++ caller of libibim.so written against the
// reference's own headers' prototypes, declared here as those headers declare
// them: without extern "C", so that this program binds to the C++-mangled
// symbols a program built from the reference (its .c files compiled as C++,
// Makefile:1,10-11) would reference. It defines the globals the reference's
// driver defines (large3Dexample2.c:48-55), which override the library's weak
// ones.
//
//   reference_caller <inputs.bin> <outputs.bin>
//
// inputs (little endian): int32 n, size, m; float64 k, eta, delta, eps; a
// (n+1)^3 level set for ASSIGN_GRADIENT, then the combo kernel's level set and
// its five geometry grids gx, gy, gz, mc, gc (layout [i][j][k]); the m x m
// operator (re then im, row-major) and the m right-hand side values (re then
// im).
// outputs: ASSIGN_GRADIENT's gx, gy, gz of the first level set; the combo kernel's
// regcount, closest points, normals, weights and size x size rows (re, im);
// CompBiCG's density (re, im).
#include <cstdio>
#include <cstdlib>
#include <vector>

// BICGSTAB.h:15-17
int CompBiCG(int size, double* density, double* densityimg, double** Kernel, double** Kernelimg, double* b,
             double* bimg);
// distance_application3D.h:12
int ASSIGN_GRADIENT(int N, double H, double*** dist, double*** graddx, double*** graddy, double*** graddz,
                    int weno_mode);
// HelmholtzKernel.h:27-33
int new_HelmholtzKernelCombo(int size, double delta, double epsilon, double*** level, double*** leveldx,
                             double*** leveldy, double*** leveldz, double*** levelmcurv, double*** levelgcurv,
                             double* zxlist, double* zylist, double* zzlist, double* zxstarlist, double* zystarlist,
                             double* zzstarlist, double* dist, double* gradx, double* grady, double* gradz,
                             double* Jacobian, double* regfactor, double* regfactorimg, double* PolyW, double* SineW,
                             double* PolyResult, double* SineResult, double** PolyKernelH, double** PolyKernelHimg,
                             double** SineKernelH, double** SineKernelHimg, double wavenum, double eta);

// the caller-defined globals the kernels read (HelmholtzKernel.h:16-18)
int N, SIZE;
double H, HCUBED, WAVE;

namespace {

// A grid as the reference allocates it: every row separately
// (large3Dexample2.c:378-403).
struct Grid {
    int S;
    double*** g;
    explicit Grid(int s) : S(s) {
        g = (double***)malloc(sizeof(double**) * S);
        for (int i = 0; i < S; ++i) {
            g[i] = (double**)malloc(sizeof(double*) * S);
            for (int j = 0; j < S; ++j) g[i][j] = (double*)calloc(S, sizeof(double));
        }
    }
    ~Grid() {
        for (int i = 0; i < S; ++i) {
            for (int j = 0; j < S; ++j) free(g[i][j]);
            free(g[i]);
        }
        free(g);
    }
    void fill(const double* v) {
        for (int i = 0; i < S; ++i)
            for (int j = 0; j < S; ++j)
                for (int k = 0; k < S; ++k) g[i][j][k] = v[((size_t)i * S + j) * S + k];
    }
    void dump(FILE* f) const {
        for (int i = 0; i < S; ++i)
            for (int j = 0; j < S; ++j) fwrite(g[i][j], sizeof(double), S, f);
    }
};

struct Rows {
    int n;
    double** r;
    explicit Rows(int m) : n(m) {
        r = (double**)malloc(sizeof(double*) * n);
        for (int i = 0; i < n; ++i) r[i] = (double*)calloc(n, sizeof(double));
    }
    ~Rows() {
        for (int i = 0; i < n; ++i) free(r[i]);
        free(r);
    }
};

bool read_all(FILE* f, void* p, size_t bytes) { return fread(p, 1, bytes, f) == bytes; }

}  // namespace

int main(int argc, char** argv) {
    if (argc != 3) {
        fprintf(stderr, "usage: %s inputs.bin outputs.bin\n", argv[0]);
        return 2;
    }
    FILE* in = fopen(argv[1], "rb");
    if (!in) return 2;
    int hdr[3];
    double par[4];
    if (!read_all(in, hdr, sizeof hdr) || !read_all(in, par, sizeof par)) return 2;
    const int n = hdr[0], size = hdr[1], m = hdr[2];
    const int S = n + 1;
    const size_t nodes = (size_t)S * S * S;
    std::vector<double> buf(nodes);
    Grid glevel(S);
    if (!read_all(in, buf.data(), sizeof(double) * nodes)) return 2;
    glevel.fill(buf.data());
    Grid grids[6] = {Grid(S), Grid(S), Grid(S), Grid(S), Grid(S), Grid(S)};
    for (auto& g : grids) {
        if (!read_all(in, buf.data(), sizeof(double) * nodes)) return 2;
        g.fill(buf.data());
    }
    Rows are(m), aim(m);
    for (int i = 0; i < m; ++i)
        if (!read_all(in, are.r[i], sizeof(double) * m)) return 2;
    for (int i = 0; i < m; ++i)
        if (!read_all(in, aim.r[i], sizeof(double) * m)) return 2;
    std::vector<double> bre(m), bim(m);
    if (!read_all(in, bre.data(), sizeof(double) * m) || !read_all(in, bim.data(), sizeof(double) * m)) return 2;
    fclose(in);

    N = n;
    SIZE = S;
    H = 4.0 / n;
    HCUBED = H * H * H;
    WAVE = par[0];

    FILE* out = fopen(argv[2], "wb");
    if (!out) return 2;
    // ASSIGN_GRADIENT with WENO-5 (weno_mode 5, large3Dexample2.c:99)
    Grid gx(S), gy(S), gz(S);
    int rc = ASSIGN_GRADIENT(n, H, glevel.g, gx.g, gy.g, gz.g, 5);
    if (rc != 0) {
        fprintf(stderr, "ASSIGN_GRADIENT returned %d\n", rc);
        return 1;
    }
    gx.dump(out);
    gy.dump(out);
    gz.dump(out);

    // the combo kernel from the supplied geometry grids
    std::vector<std::vector<double>> pts(17, std::vector<double>(size));
    Rows poly(1), polyi(1), K(size), Ki(size);
    double* P[17];
    for (int q = 0; q < 17; ++q) P[q] = pts[q].data();
    const int regcount = new_HelmholtzKernelCombo(
        size, par[2], par[3], grids[0].g, grids[1].g, grids[2].g, grids[3].g, grids[4].g, grids[5].g, P[0], P[1], P[2],
        P[3], P[4], P[5], P[6], P[7], P[8], P[9], P[10], P[11], P[12], P[13], P[14], P[15], P[16], poly.r, polyi.r,
        K.r, Ki.r, par[0], par[1]);
    if (regcount < 0) {
        fprintf(stderr, "new_HelmholtzKernelCombo returned %d\n", regcount);
        return 1;
    }
    const double rcd = (double)regcount;
    fwrite(&rcd, sizeof(double), 1, out);
    // closest points (3..5), normals (7..9), SineResult (16: the weights R)
    for (int q : {3, 4, 5, 7, 8, 9, 16}) fwrite(pts[q].data(), sizeof(double), size, out);
    for (int i = 0; i < size; ++i) fwrite(K.r[i], sizeof(double), size, out);
    for (int i = 0; i < size; ++i) fwrite(Ki.r[i], sizeof(double), size, out);

    // CompBiCG on the supplied operator
    std::vector<double> x(m), xi(m);
    rc = CompBiCG(m, x.data(), xi.data(), are.r, aim.r, bre.data(), bim.data());
    if (rc != 0) {
        fprintf(stderr, "CompBiCG returned %d\n", rc);
        return 1;
    }
    fwrite(x.data(), sizeof(double), m, out);
    fwrite(xi.data(), sizeof(double), m, out);
    fclose(out);
    printf("reference_caller: ASSIGN_GRADIENT, new_HelmholtzKernelCombo (regcount %d), CompBiCG done\n", regcount);
    return 0;
}
