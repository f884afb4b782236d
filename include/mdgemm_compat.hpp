// mdgemm_compat.hpp -- the reference's C++ entry points for the multi-double
// GEMM path, same names, signatures and argument meaning, served by
// libmdgemm.so on the GPU.  Host vectors in, host vectors out, exactly as the
// reference's functions take and return them (inputs by value, result by
// value, interleaved limbs, square n x n matrices).
//
// Differences from the reference, all deliberate:
//   * every function checks its arguments and every HIP call and throws
//     std::runtime_error (with mdg_last_error() text) on failure; the
//     reference silently returns zeros for an unsupported p
//     (C++/TCcuda.cu:269-297) and never checks a CUDA call;
//   * device buffers are freed before returning (the reference never frees,
//     C++/TCcuda.cu:255-260);
//   * matmulTCnt ignores nfrag's tiling role (the kernel picks its own
//     tiles) but still requires nfrag to divide n like the reference;
//   * matmul{dd,qd,od,hd}f run on the GPU with the reference's Dekker
//     two_prod, bit-identical to the reference's host loops.
#ifndef MDGEMM_COMPAT_HPP
#define MDGEMM_COMPAT_HPP

#include <vector>

// C++/TCcuda.h:23 -- n x n matrix of positive random p-doubles with leading
// exponents in [expmin, expmax) (the reference's random_pd layout; seeded
// from mdgemm_compat_seed(), advanced per call, instead of rand()).
std::vector<double> mat(int n, int p, int expmin = 0, int expmax = 0);
// C++/TCcuda.h:28 -- m x n matrix of zero p-doubles
std::vector<double> zeros(int m, int n, int p);
// C++/TCcuda.h:33 -- component i of every p-double of x
std::vector<double> part(std::vector<double> x, int p, int i);
// C++/TCcuda.h:36 -- [A_1 .. A_p] stacked row-wise, row-major
std::vector<double> bigA(std::vector<double> A, int n, int p);
// C++/TCcuda.h:45 -- upper block-Toeplitz [B_1..B_p; 0 B_1..; ...], column-major
std::vector<double> bigB(std::vector<double> B, int n, int p);
// C++/TCcuda.h:53 -- per-entry p x p Toeplitz blocks, column-major
std::vector<double> bigB2(std::vector<double> B, int n, int p);

// C++/TCcuda.h:66 -- vector GEMM on the GPU, loop_ct launches timed into t_CUDA (ms)
std::vector<double> matmulTCnt(std::vector<double> A, std::vector<double> B, int n, int nfrag, int p,
                               int loop_ct, float &t_CUDA);

// C++/TCcuda.h:72-75 -- the reference's host GEMMs (here on the GPU, bit-identical)
std::vector<double> matmulddf(std::vector<double> A, std::vector<double> B, int n);
std::vector<double> matmulqdf(std::vector<double> A, std::vector<double> B, int n);
std::vector<double> matmulodf(std::vector<double> A, std::vector<double> B, int n);
std::vector<double> matmulhdf(std::vector<double> A, std::vector<double> B, int n);

// C++/TCcuda.h:69 -- limb convolution of p-part entries in plain doubles (the
// reference's host loop, here on the GPU in the same order: the same bits)
std::vector<double> matmulhost(std::vector<double> A, std::vector<double> B, int n, int p);
// C++/TCcuda.h:78,81 -- the reference's host renormalisation of every entry,
// in place, interleaved (renormhost) or in the bigA layout (renormhostbig)
void renormhost(std::vector<double> &A, int n, int p);
void renormhostbig(std::vector<double> &A, int n, int p);

// C++/TCcuda.h:103 -- fold pp parts into p-doubles, kernel time into t_CUDA (ms)
std::vector<double> pllsqueeze_old(std::vector<double> x, int p, int pp, float &t_CUDA);

// C++/TCcuda.h:86-96 -- the same fold, y = inc_d(...inc_d(0, x_0)..., x_{pp-1}),
// under the reference's per-(p, pp) names: pllsqueeze_2_{8,12,16} (its GPU
// versions) and pllntsqueeze_2_{8,12,16} (its host loops; here on the GPU
// too, the same bits), dispatched by pllsqueeze / pllntsqueeze (C++/TCcuda.cu
// :810-875).  Differences: every entry is folded (the reference's kernel
// skips the last n % 64 entries when n > 64, C++/TCcuda.cu:454-466), and
// pllsqueeze(x, p, pp) serves every p in {2, 4, 8, 16} and pp >= 1 where the
// reference returns a double-double for pp == 12 whatever p is and falls off
// the end of the function for the other pairs; here pp == 12 with p != 2
// and any other unsupported p throw.
std::vector<double> pllsqueeze_2_8(std::vector<double> x);
std::vector<double> pllsqueeze_2_12(std::vector<double> x);
std::vector<double> pllsqueeze_2_16(std::vector<double> x);
std::vector<double> pllntsqueeze_2_8(std::vector<double> x);
std::vector<double> pllntsqueeze_2_12(std::vector<double> x);
std::vector<double> pllntsqueeze_2_16(std::vector<double> x);
std::vector<double> pllsqueeze(std::vector<double> x, int p, int pp);
std::vector<double> pllntsqueeze(std::vector<double> x, int p, int pp);

// C++/floatybits.h:72,87,97 -- the bit splitters
std::vector<double> split4pd(std::vector<double> x, int p);
std::vector<double> split8pd(std::vector<double> x, int p);
std::vector<double> mixsplit2(std::vector<double> x);

// The reference's FP64 tensor-core product C (M x N) += A (M x K) * B
// (K x N column-major), C++/TCtest.cu:21-46 -- with explicit sizes instead
// of the test driver's compile-time M_GLOBAL/N_GLOBAL/K_GLOBAL.
void matmul_tc(const std::vector<double> &A, const std::vector<double> &B, std::vector<double> &C, int M, int N,
               int K);

// Seed of mat(); returns the previous seed.
unsigned long long mdgemm_compat_seed(unsigned long long seed);

#endif
