// ref_split.hip -- the reference's bit splitter and squeeze on the device.
//
// The reference splits on the host through vectors of bits built with
// frexp/pow per bit (C++/floatybits.cu:6-102) and squeezes with a separate
// launch per call (C++/TCcuda.cu:876-958).  Here each p-double is one thread:
// the stored fraction bits come from frexp as an integer, parts are masked
// out of it, and the balance chain runs in registers.  The bit semantics of
// the reference are kept exactly (same values, same signed zeros):
//   * stored fraction = the frexp fraction's bits 2^-1 .. 2^-52 (its 2^-53
//     bit is dropped, floatybits.cu:97-99);
//   * a negative input stores no fraction bits (floatybits.cu:20-28), so it
//     splits into -0.0 parts and a +0.0 remainder;
//   * an exponent field <= 0 is stored as 0 (floatybits.cu:40-49);
//   * each part is rebuilt as sign * (masked fraction * 2^-52) * 2^exp
//     (double_rep, floatybits.cu:92-102);
//   * balance (floatybits.cu:134-148) moves 2^(e_x - threshold) from x to y
//     when their exponents differ by more than threshold - 1.
#include "common.hpp"
#include "md_arith.hpp"

namespace mdg {

struct StoredBits {
    int sign;
    int expo;        // stored exponent field - 1023
    uint64_t frac;   // 52 stored fraction bits, bit 51 <-> 2^-1
};

__device__ __forceinline__ StoredBits stored_bits(double x)
{
    StoredBits b;
    int ex = 0;
    double fr = frexp(x, &ex);
    b.sign = (x >= 0) ? 0 : 1;
    const int field = ex + 1023;
    b.expo = (field > 0 ? field : 0) - 1023;
    b.frac = 0;
    if (fr > 0) b.frac = ((uint64_t)ldexp(fr, 53)) >> 1;
    return b;
}

__device__ __forceinline__ double rep_masked(const StoredBits &b, uint64_t mask)
{
    const double fraction = (double)(b.frac & mask) * 0x1p-52;
    const double sgn = b.sign ? -1.0 : 1.0;
    return sgn * fraction * ldexp(1.0, b.expo);
}

// mask of stored fraction bits [lo, hi), bit 0 <-> 2^-1
__device__ __forceinline__ uint64_t frac_mask(int lo, int hi)
{
    const uint64_t top = (hi >= 52) ? ((1ull << 52) - 1) : (((1ull << hi) - 1) << (52 - hi));
    const uint64_t low = (lo <= 0) ? 0ull : (((1ull << lo) - 1) << (52 - lo));
    return top & ~low;
}

// split4 (floatybits.cu:159-191): 13-bit parts, remainder = value - sum
__device__ __forceinline__ void split4_parts(double x, double *d)
{
    const StoredBits b = stored_bits(x);
    const double d1 = rep_masked(b, frac_mask(0, 13));
    const double d2 = rep_masked(b, frac_mask(13, 26));
    const double d3 = rep_masked(b, frac_mask(26, 39));
    const double full = rep_masked(b, frac_mask(0, 52));
    d[0] = d1;
    d[1] = d2;
    d[2] = d3;
    d[3] = full - (d1 + d2 + d3);
}

// split8 (floatybits.cu:218-301): 7,7,7,7,6,6,12-bit parts, remainder
__device__ __forceinline__ void split8_parts(double x, double *d)
{
    const StoredBits b = stored_bits(x);
    const int cut[8] = {0, 7, 14, 21, 28, 34, 40, 52};
#pragma unroll
    for (int t = 0; t < 7; ++t) d[t] = rep_masked(b, frac_mask(cut[t], cut[t + 1]));
    double s = d[0];
#pragma unroll
    for (int t = 1; t < 7; ++t) s = s + d[t];
    d[7] = rep_masked(b, frac_mask(0, 52)) - s;
}

__device__ __forceinline__ void balance_pair(double &x, double &y, int threshold)
{
    if (x == 0.0) return;
    if (y == 0.0) return;
    int xe, ye;
    frexp(x, &xe);
    frexp(y, &ye);
    if (!(xe - ye > threshold - 1)) return;
    const double bit = ldexp(1.0, xe - threshold);
    x = x - bit;
    y = y + bit;
}

// kind 4: split4pd, kind 8: split8pd, kind 12: mixsplit2 (p == 2).
// Parts are produced limb by limb and streamed through the balance chain:
// `cur` is the part whose pair with the next one is balanced next.
template <int KIND>
__global__ __launch_bounds__(256) void split_kernel(const double *__restrict__ x, long long groups, int p,
                                                    double *__restrict__ out)
{
    const long long gi = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= groups) return;
    const double *xg = x + gi * p;
    const int per = (KIND == 4) ? 4 : (KIND == 8) ? 8 : 6;   // parts per limb (mix: 12 per 2 limbs)
    const int P = per * p;
    double *o = out + gi * P;
    // number of balanced adjacent pairs and their thresholds
    const int nbal = (KIND == 4) ? P - 2 : (KIND == 8) ? P - 3 : 9;
    double cur = 0.0;
    int idx = 0;   // index of the next part to produce
    for (int l = 0; l < p; ++l) {
        double parts[8];
        int np;
        if (KIND == 4) {
            split4_parts(xg[l], parts);
            np = 4;
        } else if (KIND == 8) {
            split8_parts(xg[l], parts);
            np = 8;
        } else {
            if ((l & 1) == 0) {
                split4_parts(xg[l], parts);
                np = 4;
            } else {
                split8_parts(xg[l], parts);
                np = 8;
            }
        }
        for (int q = 0; q < np; ++q) {
            double nxt = parts[q];
            if (idx == 0) {
                cur = nxt;
            } else {
                const int pair = idx - 1;   // balance(part[pair], part[pair+1])
                if (pair < nbal) {
                    int thr;
                    if (KIND == 4)
                        thr = 14;
                    else if (KIND == 8)
                        thr = (pair % 8 < 4) ? 8 : 7;
                    else
                        thr = (pair % 12 < 4) ? 14 : ((pair % 12 < 8) ? 8 : 7);
                    balance_pair(cur, nxt, thr);
                }
                o[pair] = cur;
                cur = nxt;
            }
            ++idx;
        }
    }
    o[P - 1] = cur;
}

template <int L>
__global__ __launch_bounds__(256) void squeeze_kernel(long long nent, int pp, const double *__restrict__ x,
                                                      double *__restrict__ y)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nent) return;
    double acc[L];
#pragma unroll
    for (int l = 0; l < L; ++l) acc[l] = 0.0;
    for (int j = 0; j < pp; ++j) md_inc_d<L>(acc, x[i * pp + j]);
#pragma unroll
    for (int l = 0; l < L; ++l) y[i * L + l] = acc[l];
}

// bigB2 (C++/TCcuda.cu:79-95): column-major (n pp) x (n pp), element
// (row = r pp + j, col = c pp + i) = Bq[r][c][i-j] for j <= i, else 0
__global__ void bigB2_kernel(const double *__restrict__ Bq, int n, int pp, double *__restrict__ out)
{
    const long long total = (long long)n * pp * n * pp;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const int j = (int)(idx % pp);
    long long t = idx / pp;
    const int r = (int)(t % n);
    t /= n;
    const int i = (int)(t % pp);
    const int c = (int)(t / pp);
    out[idx] = (j <= i) ? Bq[((long long)r * n + c) * pp + (i - j)] : 0.0;
}

// matmulhost (C++/TCcuda.cu:304-317): the limb convolution of two n x n
// matrices of p-part entries in plain doubles, each product and each sum
// rounded, in the reference's order -- for entry (r, c) and part i, k
// ascending and within it j = 0 .. i:
//   C[r][c][i] = (...(0 + A[r][0][0] B[0][c][i]) + ...) + A[r][n-1][i] B[n-1][c][0]
// One thread per (r, c, i), i fastest.
__global__ void matmulhost_kernel(const double *__restrict__ A, const double *__restrict__ B, int n, int p,
                                  double *__restrict__ C)
{
    const long long total = (long long)n * n * p;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const int i = (int)(idx % p);
    const long long rc = idx / p;
    const long long r = rc / n, c = rc - r * n;
    double s = 0.0;
    for (int k = 0; k < n; ++k) {
        const double *a = A + (r * n + k) * p;
        const double *b = B + ((long long)k * n + c) * p;
        for (int j = 0; j <= i; ++j) s = s + a[j] * b[i - j];
    }
    C[idx] = s;
}

// renormhost / renormhostbig (C++/TCcuda.cu:374-398), in place: for each
// entry, k = 0 .. p-2 in turn, hi = x_k + x_{k+1}, x_{k+1} = hi - x_k,
// x_k = hi (the reference's own rule: the second value is hi - x_k, not the
// rounding error of the sum).  big = 0: entry (i, j) at A[(i n + j) p + k]
// (interleaved); big = 1: at A[i n p + k n + j] (the bigA layout).
__global__ void renormhost_kernel(double *__restrict__ A, int n, int p, int big)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)n * n) return;
    const long long i = idx / n, j = idx - i * n;
    double *x = big ? A + i * n * p + j : A + idx * p;
    const long long st = big ? n : 1;
    for (int k = 0; k < p - 1; ++k) {
        const double hi = x[k * st] + x[(k + 1) * st];
        const double lo = hi - x[k * st];
        x[k * st] = hi;
        x[(k + 1) * st] = lo;
    }
}

// ------------------------------------------------------------- generator --
__device__ __forceinline__ uint64_t rng_next(uint64_t &s)
{
    s += 0x9E3779B97F4A7C15ULL;
    return mix64(s);
}

__device__ __forceinline__ double rng_u01(uint64_t &s) { return (double)(rng_next(s) >> 11) * 0x1p-53; }

__global__ void fill_random_kernel(int kind, int L, long long count, uint64_t seed, int emin, int emax,
                                   double *__restrict__ out)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    uint64_t s = mix64(seed + 0x632BE59BD9B4E019ULL * ((uint64_t)i + 1));
    double *o = out + i * L;
    if (kind == 0) {
        const double hi = 2.0 * rng_u01(s) - 1.0;
        const double lo = hi * 0x1p-53 * (rng_u01(s) - 0.5);
        double e;
        o[0] = quick_two_sum(hi, lo, e);
        if (L > 1) o[1] = e;
        for (int l = 2; l < L; ++l) o[l] = o[l - 1] * 0x1p-53 * (rng_u01(s) - 0.5);
    } else if (kind == 1) {
        const int span = emax - emin + 1;
        const int ex = emin + (int)(rng_next(s) % (uint64_t)span);
        const double sg = (rng_next(s) & 1) ? -1.0 : 1.0;
        o[0] = sg * ldexp(1.0 + rng_u01(s), ex);
        for (int l = 1; l < L; ++l) o[l] = o[l - 1] * 0x1p-53 * (rng_u01(s) - 0.5);
    } else {
        const int hiexp = (emax == emin) ? emin + 1023
                                         : emin + 1023 + (int)(rng_next(s) % (uint64_t)(emax - emin));
        for (int l = 0; l < L; ++l) {
            const uint64_t f = (rng_next(s) >> 12) | (1ull << 51);
            o[l] = (double)f * 0x1p-52 * ldexp(1.0, hiexp - 52 * l - 1023);
        }
    }
}

// ---------------------------------------------------------------- hosts --
static inline unsigned nblocks(long long nel, int bs) { return (unsigned)((nel + bs - 1) / bs); }

int ref_split(int kind, const double *x, long long count, int p, double *out, hipStream_t s)
{
    const long long groups = count / p;
    if (groups == 0) return MDG_SUCCESS;
    const dim3 grid(nblocks(groups, 256));
    if (kind == 4)
        hipLaunchKernelGGL(split_kernel<4>, grid, dim3(256), 0, s, x, groups, p, out);
    else if (kind == 8)
        hipLaunchKernelGGL(split_kernel<8>, grid, dim3(256), 0, s, x, groups, p, out);
    else
        hipLaunchKernelGGL(split_kernel<12>, grid, dim3(256), 0, s, x, groups, 2, out);
    MDG_LAUNCH_CHECK("split_kernel");
    return MDG_SUCCESS;
}

int squeeze(int L, long long nent, int pp, const double *x, double *y, hipStream_t s)
{
    if (nent == 0) return MDG_SUCCESS;
    const dim3 grid(nblocks(nent, 256));
    switch (L) {
    case 2: hipLaunchKernelGGL(squeeze_kernel<2>, grid, dim3(256), 0, s, nent, pp, x, y); break;
    case 4: hipLaunchKernelGGL(squeeze_kernel<4>, grid, dim3(256), 0, s, nent, pp, x, y); break;
    case 8: hipLaunchKernelGGL(squeeze_kernel<8>, grid, dim3(256), 0, s, nent, pp, x, y); break;
    case 16: hipLaunchKernelGGL(squeeze_kernel<16>, grid, dim3(256), 0, s, nent, pp, x, y); break;
    default: set_error("unsupported limb count L=%d", L); return MDG_ERROR_INVALID_VALUE;
    }
    MDG_LAUNCH_CHECK("squeeze_kernel");
    return MDG_SUCCESS;
}

int bigB2(const double *Bq, int n, int pp, double *out, hipStream_t s)
{
    const long long total = (long long)n * pp * n * pp;
    if (total == 0) return MDG_SUCCESS;
    hipLaunchKernelGGL(bigB2_kernel, dim3(nblocks(total, 256)), dim3(256), 0, s, Bq, n, pp, out);
    MDG_LAUNCH_CHECK("bigB2_kernel");
    return MDG_SUCCESS;
}

int matmulhost(const double *A, const double *B, int n, int p, double *C, hipStream_t s)
{
    const long long total = (long long)n * n * p;
    if (total == 0) return MDG_SUCCESS;
    hipLaunchKernelGGL(matmulhost_kernel, dim3(nblocks(total, 256)), dim3(256), 0, s, A, B, n, p, C);
    MDG_LAUNCH_CHECK("matmulhost_kernel");
    return MDG_SUCCESS;
}

int renormhost(double *A, int n, int p, int big, hipStream_t s)
{
    const long long total = (long long)n * n;
    if (total == 0 || p < 2) return MDG_SUCCESS;
    hipLaunchKernelGGL(renormhost_kernel, dim3(nblocks(total, 256)), dim3(256), 0, s, A, n, p, big);
    MDG_LAUNCH_CHECK("renormhost_kernel");
    return MDG_SUCCESS;
}

int fill_random(int kind, int L, long long count, unsigned long long seed, int emin, int emax, double *out,
                hipStream_t s)
{
    if (count == 0) return MDG_SUCCESS;
    hipLaunchKernelGGL(fill_random_kernel, dim3(nblocks(count, 256)), dim3(256), 0, s, kind, L, count,
                       (uint64_t)seed, emin, emax, out);
    MDG_LAUNCH_CHECK("fill_random_kernel");
    return MDG_SUCCESS;
}

}  // namespace mdg
