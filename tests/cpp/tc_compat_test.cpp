// tc_compat_test.cpp -- drives the reference's C++ interface
// (include/mdgemm_compat.hpp) the way the reference's own driver
// C++/TCtest.cu:48-179 does: generate with mat(), split, block-Toeplitz
// operand, FP64 tensor-core product, squeeze, then the vector GEMM and the
// host GEMM on the same inputs.  Every array is written as raw float64 into
// the output directory for tests/test_gpu_compat.py to check against the
// oracle; the driver itself prints the reference's "max error" line.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <string>
#include <vector>

#include "mdgemm_compat.hpp"

static void dump(const std::string &dir, const char *name, const std::vector<double> &v)
{
    std::string path = dir + "/" + name + ".bin";
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) {
        std::perror(path.c_str());
        std::exit(2);
    }
    std::fwrite(v.data(), sizeof(double), v.size(), f);
    std::fclose(f);
}

int main(int argc, char **argv)
{
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s outdir n p\n", argv[0]);
        return 2;
    }
    const std::string dir = argv[1];
    const int n = std::atoi(argv[2]);
    const int p = std::atoi(argv[3]);
    const int pp = 4 * p;
    mdgemm_compat_seed(2024);
    std::vector<double> A = mat(n, p, 0, 0);
    std::vector<double> B = mat(n, p, 0, 0);
    std::vector<double> Aq = split4pd(A, p);
    std::vector<double> Bq = split4pd(B, p);
    std::vector<double> B2 = bigB2(Bq, n, pp);
    std::vector<double> C1q = zeros(n, n, pp);
    matmul_tc(Aq, B2, C1q, n, n * pp, n * pp);
    float t_squeeze = 0.f, t_vec = 0.f;
    std::vector<double> C1 = pllsqueeze_old(C1q, p, pp, t_squeeze);
    std::vector<double> C2 = matmulTCnt(A, B, n, std::min(32, n), p, 1, t_vec);
    std::vector<double> C3;
    switch (p) {
    case 2: C3 = matmulddf(A, B, n); break;
    case 4: C3 = matmulqdf(A, B, n); break;
    case 8: C3 = matmulodf(A, B, n); break;
    default: C3 = matmulhdf(A, B, n); break;
    }
    double max_err = 0;
    for (size_t i = 0; i < C1.size(); ++i) max_err = std::fmax(max_err, std::fabs(C1[i] - C2[i]));
    std::printf("n=%d p=%d pp=%d squeeze %.3f ms, vector %.3f ms, max |TC - CUDA| limb diff %g\n", n, p, pp,
                t_squeeze, t_vec, max_err);
    dump(dir, "A", A);
    dump(dir, "B", B);
    dump(dir, "Aq", Aq);
    dump(dir, "Bq", Bq);
    dump(dir, "B2", B2);
    dump(dir, "C1q", C1q);
    dump(dir, "C1", C1);
    dump(dir, "C2", C2);
    dump(dir, "C3", C3);
    dump(dir, "part1", part(A, p, 1));
    dump(dir, "bigA", bigA(A, n, p));
    dump(dir, "bigB", bigB(B, n, p));
    // the reference's other host routines: the limb convolution in plain
    // doubles, and its two renormalisations (in place)
    dump(dir, "mh", matmulhost(A, B, n, p));
    std::vector<double> R1 = A, R2 = bigA(A, n, p);
    renormhost(R1, n, p);
    renormhostbig(R2, n, p);
    dump(dir, "rn", R1);
    dump(dir, "rnbig", R2);
    // the per-(p, pp) squeeze names on the parts the splitters make
    dump(dir, "sq_p", pllsqueeze(C1q, p, pp));
    dump(dir, "sqnt_p", pllntsqueeze(C1q, p, pp));
    if (p == 2) {
        std::vector<double> x8 = split4pd(A, 2), x16 = split8pd(A, 2), x12 = mixsplit2(A);
        dump(dir, "x8", x8);
        dump(dir, "x12", x12);
        dump(dir, "x16", x16);
        dump(dir, "sq_2_8", pllsqueeze_2_8(x8));
        dump(dir, "sq_2_12", pllsqueeze_2_12(x12));
        dump(dir, "sq_2_16", pllsqueeze_2_16(x16));
        dump(dir, "sqnt_2_8", pllntsqueeze_2_8(x8));
        dump(dir, "sqnt_2_12", pllntsqueeze_2_12(x12));
        dump(dir, "sqnt_2_16", pllntsqueeze_2_16(x16));
        dump(dir, "sq_12", pllsqueeze(x12, 2, 12));
    }
    return 0;
}
