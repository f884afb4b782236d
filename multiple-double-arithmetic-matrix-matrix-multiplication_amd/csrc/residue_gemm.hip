// residue_gemm.hip -- multi-double GEMM from INT8 GEMMs of residues
// (Ozaki, Uchino, Imamura, "Ozaki Scheme II: a GEMM-oriented emulation of
// floating-point matrix multiplication using an integer modular technique",
// 2025), extended to the hundreds of bits of quad- to hexa-double.
//
// Like the slice GEMMs, every row of A / column of B gets a scale exponent
// and every entry becomes the fixed-point integer a = floor(x 2^(W-2-E))
// (slices.hpp).  The integer product Z = sum_k a_rk b_kc is then computed
// exactly from its residues modulo pairwise coprime moduli m_1..m_n with
// M = prod m_i > 4 max|Z|:
//   * narrow moduli (m <= 256): the balanced residues of a and b are INT8,
//     and one INT8 x INT8 -> INT32 GEMM gives Z mod m;
//   * wide moduli (odd primes < 64768): the balanced residue is cut into two
//     INT8 digits, r = 256 hi + lo, and three GEMMs (hi*hi, lo*lo, and
//     hi*lo + lo*hi as one GEMM of depth 2k) give Z mod m.
// The slice GEMMs need S(S+1)/2 slice-pair products for S ~ W/8 slices;
// this needs about (2W + log2 k)/8 narrow and 4 GEMM units per 16 further
// bits -- at quad-double 81 INT8 GEMMs of depth k instead of 435.
//
// Reconstruction (per output entry, exact integer arithmetic):
//   Z = sum_i r_i c_i - q M,  c_i = (M/m_i) * ((M/m_i)^-1 mod m_i),
//   q = round(sum_i r_i y_i / m_i)  (in FP64; |Z| <= M/4 makes it exact),
// accumulated in 32-bit words with 64-bit accumulators, carried, and folded
// into L limbs from the most significant word with the TwoSum inc_d cascade
// (C++/TCcuda.cu:876-906 folds its parts the same way).
//
// Error: the only roundings are the truncation of each entry to W bits and
// the final fold, so the normwise error is about L 2^(3-W+spread_a+spread_b);
// W = target + spread_a + spread_b + 3 + log2(L).
#include <climits>
#include <cmath>

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "md_arith.hpp"
#include "slices.hpp"
#include "split_common.hpp"

namespace mdg {

constexpr int RES_P8 = 72;       // 16-bit halfwords of the widest fixed-point integer (36 words)
constexpr int RES_PB = 144;      // bytes of it
constexpr int RES_KC = 130944;   // GEMM depth with exact INT32 sums: RES_KC * 2^14 < 2^31, multiple of 128

// Device view of one modulus set (n = n8 + n16 moduli, narrow ones first).
struct ResTab {
    int n8, n16;
    int nwc;                    // words per CRT constant (incl. a sign word), a multiple of 4
    const double *md;           // m_i
    const double *inv;          // 1 / m_i
    const double *frac;         // y_i / m_i, y_i = (M/m_i)^-1 mod m_i
    const uint32_t *cw;         // c_i = (M/m_i) y_i, [n][nwc] 32-bit words, least significant first
    const uint32_t *Mw;         // M, [nwc] words
    const uint32_t *p8;         // 2^(16h) mod m_i, [n8][RES_P8]
    const int *pb;              // balanced 2^(8b) mod m_j, [n16][RES_PB]
};

// v mod m in [0, m) for an integer-valued double |v| < 2^53: the FP64
// quotient is off by at most one and the remainder fma is exact
__device__ __forceinline__ double mod_pos(double v, double md, double inv)
{
    const double q = floor(v * inv);
    double r = fma(-q, md, v);
    r += (r < 0.0) ? md : 0.0;
    r -= (r >= md) ? md : 0.0;
    return r;
}

// Balanced residues of |N| (sign applied afterwards) for every modulus of t.
// emit8(i, r): narrow modulus i, r in [-128, 127];
// emit16(j, hi, lo): wide modulus j, r = 256 hi + lo, hi, lo in [-128, 127].
template <int NW, typename E8, typename E16>
__device__ __forceinline__ void residues(uint32_t (&N)[NW], const ResTab &t, E8 emit8, E16 emit16)
{
    const bool neg = (int32_t)N[NW - 1] < 0;
    if (neg) {
        uint32_t c = 1;
#pragma unroll
        for (int j = 0; j < NW; ++j) {
            const uint32_t v = ~N[j] + c;
            c = (c && v == 0u) ? 1u : 0u;
            N[j] = v;
        }
    }
    for (int i = 0; i < t.n8; ++i) {
        // halfword (16-bit) x (2^(16h) mod m < 2^8): sums below 2^31
        const uint32_t *pw = t.p8 + i * RES_P8;
        uint32_t acc = 0;
#pragma unroll
        for (int j = 0; j < NW; ++j) acc += (N[j] & 0xFFFFu) * pw[2 * j] + (N[j] >> 16) * pw[2 * j + 1];
        const double md = t.md[i];
        int r = (int)mod_pos((double)acc, md, t.inv[i]);
        const int m = (int)md;
        if (neg && r) r = m - r;
        if (2 * r >= m) r -= m;
        emit8(i, r);
    }
    for (int j = 0; j < t.n16; ++j) {
        // byte x balanced (2^(8b) mod m, |.| < 2^15): |sum| < 2^31
        const int *pw = t.pb + j * RES_PB;
        int acc = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            acc += (int)(N[w] & 0xFFu) * pw[4 * w] + (int)((N[w] >> 8) & 0xFFu) * pw[4 * w + 1] +
                   (int)((N[w] >> 16) & 0xFFu) * pw[4 * w + 2] + (int)(N[w] >> 24) * pw[4 * w + 3];
        }
        const double md = t.md[t.n8 + j];
        int r = (int)mod_pos((double)acc, md, t.inv[t.n8 + j]);
        const int m = (int)md;
        if (neg && r) r = m - r;
        if (2 * r >= m) r -= m;
        const int lo = ((r + 128) & 255) - 128;
        emit16(j, (r - lo) >> 8, lo);
    }
}

// A: thread per entry (r, kk) of the padded mp x kp domain, kk fastest.
// Narrow planes R8[i][r][kk] (mp x kp), wide planes R16[j][r][2kp]: hi at
// kk, lo at kp + kk.
template <int L, int NWU>
__global__ __launch_bounds__(256) void residue_a_kernel(int m, int k, int mp, int kp, const double *__restrict__ A,
                                                        long long lda, const int *__restrict__ E, int W, ResTab t,
                                                        int8_t *__restrict__ R8, int8_t *__restrict__ R16)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)mp * kp) return;
    const long long r = idx / kp;
    const int kk = (int)(idx - r * kp);
    const long long plane = (long long)mp * kp;
    int8_t *o8 = R8 + r * kp + kk;
    int8_t *o16 = R16 + r * 2 * kp + kk;
    if (r >= m || kk >= k || E[r] == ZERO_SCALE) {
        for (int i = 0; i < t.n8; ++i) o8[i * plane] = 0;
        for (int j = 0; j < t.n16; ++j) {
            o16[2 * j * plane] = 0;
            o16[2 * j * plane + kp] = 0;
        }
        return;
    }
    double x[L];
#pragma unroll
    for (int l = 0; l < L; ++l) x[l] = A[(r * lda + kk) * L + l];
    uint32_t N[NWU];
    to_fixed<L, NWU>(x, W - 2 - E[r], N);
    residues<NWU>(
        N, t, [&](int i, int v) { o8[i * plane] = (int8_t)v; },
        [&](int j, int hi, int lo) {
            o16[2 * j * plane] = (int8_t)hi;
            o16[2 * j * plane + kp] = (int8_t)lo;
        });
}

// B: thread per entry (kk, c) of the padded kp x np domain, kk fastest so
// the transposed writes coalesce.  Narrow planes R8[i][c][kk] (np x kp),
// wide planes R16[j][c][3kp]: hi at kk and 2kp + kk, lo at kp + kk, so that
// [kp, 3kp) is (lo, hi) against A's (hi, lo).
template <int L, int NWU>
__global__ __launch_bounds__(256) void residue_b_kernel(int k, int n, int kp, int np, const double *__restrict__ B,
                                                        long long ldb, const int *__restrict__ F, int W, ResTab t,
                                                        int8_t *__restrict__ R8, int8_t *__restrict__ R16)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)np * kp) return;
    const long long c = idx / kp;
    const int kk = (int)(idx - c * kp);
    const long long plane = (long long)np * kp;
    int8_t *o8 = R8 + c * kp + kk;
    int8_t *o16 = R16 + c * 3 * kp + kk;
    if (c >= n || kk >= k || F[c] == ZERO_SCALE) {
        for (int i = 0; i < t.n8; ++i) o8[i * plane] = 0;
        for (int j = 0; j < t.n16; ++j) {
            o16[3 * j * plane] = 0;
            o16[3 * j * plane + kp] = 0;
            o16[3 * j * plane + 2 * kp] = 0;
        }
        return;
    }
    double x[L];
#pragma unroll
    for (int l = 0; l < L; ++l) x[l] = B[((long long)kk * ldb + c) * L + l];
    uint32_t N[NWU];
    to_fixed<L, NWU>(x, W - 2 - F[c], N);
    residues<NWU>(
        N, t, [&](int i, int v) { o8[i * plane] = (int8_t)v; },
        [&](int j, int hi, int lo) {
            o16[3 * j * plane] = (int8_t)hi;
            o16[3 * j * plane + kp] = (int8_t)lo;
            o16[3 * j * plane + 2 * kp] = (int8_t)hi;
        });
}

// C[r][c] from the INT32 planes: narrow D8[i][q] (chunk q of nck1), wide
// Dhh[j][q], Dll[j][q] (nck1 chunks) and Dx[j][q] (nck2 chunks).
template <int L, int NWC>
__global__ __launch_bounds__(256) void residue_reconstruct_kernel(int m, int n, int W, ResTab t, int nck1, int nck2,
                                                                  const int *__restrict__ D8,
                                                                  const int *__restrict__ Dhh,
                                                                  const int *__restrict__ Dll,
                                                                  const int *__restrict__ Dx,
                                                                  const int *__restrict__ E,
                                                                  const int *__restrict__ F, double *__restrict__ C,
                                                                  long long ldc)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)m * n) return;
    const long long r = idx / n, c = idx - r * n;
    double y[L];
#pragma unroll
    for (int l = 0; l < L; ++l) y[l] = 0.0;
    const int er = E[r], fc = F[c];
    if (er != ZERO_SCALE && fc != ZERO_SCALE) {
        const long long mn = (long long)m * n;
        uint64_t acc[NWC];
#pragma unroll
        for (int w = 0; w < NWC; ++w) acc[w] = 0;
        double tq = 0.0;
        auto add = [&](int i, double v) {
            const double res = mod_pos(v, t.md[i], t.inv[i]);
            tq = fma(res, t.frac[i], tq);
            const uint32_t ri = (uint32_t)res;
            const uint32_t *cw = t.cw + (long long)i * NWC;
#pragma unroll
            for (int w = 0; w < NWC; ++w) acc[w] += (uint64_t)ri * cw[w];
        };
        for (int i = 0; i < t.n8; ++i) {
            long long v = 0;
            for (int q = 0; q < nck1; ++q) v += D8[((long long)i * nck1 + q) * mn + idx];
            add(i, (double)v);
        }
        for (int j = 0; j < t.n16; ++j) {
            long long hh = 0, ll = 0, xx = 0;
            for (int q = 0; q < nck1; ++q) {
                hh += Dhh[((long long)j * nck1 + q) * mn + idx];
                ll += Dll[((long long)j * nck1 + q) * mn + idx];
            }
            for (int q = 0; q < nck2; ++q) xx += Dx[((long long)j * nck2 + q) * mn + idx];
            add(t.n8 + j, (double)(hh * 65536 + xx * 256 + ll));
        }
        // Z = sum r_i c_i - q M, carried into 32-bit two's-complement words
        const long long q = (long long)rint(tq);
        uint32_t z[NWC];
        long long carry = 0;
#pragma unroll
        for (int w = 0; w < NWC; ++w) {
            const long long v = (long long)acc[w] - q * (long long)t.Mw[w] + carry;
            z[w] = (uint32_t)v;
            carry = v >> 32;
        }
        const bool neg = (int32_t)z[NWC - 1] < 0;
        if (neg) {
            uint32_t cc = 1;
#pragma unroll
            for (int w = 0; w < NWC; ++w) {
                const uint32_t v = ~z[w] + cc;
                cc = (cc && v == 0u) ? 1u : 0u;
                z[w] = v;
            }
        }
        // fold |Z| 2^sc from the top word down; words more than L*53 + 64
        // bits below the leading one are under the last limb's rounding
        const int sc = er + fc + 4 - 2 * W;
        constexpr int KEEP = (53 * L + 63) / 32 + 2;
        int left = KEEP;
#pragma unroll
        for (int w = NWC - 1; w >= 0; --w) {
            if (z[w] != 0u && left > 0) md_inc_d<L>(y, ldexp((double)z[w], 32 * w + sc));
            if (z[w] != 0u || left < KEEP) --left;
        }
        if (neg) {
#pragma unroll
            for (int l = 0; l < L; ++l) y[l] = -y[l];
        }
    }
    double *o = C + (r * ldc + c) * L;
#pragma unroll
    for (int l = 0; l < L; ++l) o[l] = y[l];
}

// --------------------------------------------------------------- moduli --
// Narrow: every m <= 256 coprime to all larger ones taken (256, 255, 253,
// 251, 247, ...; 49 moduli, 341.9 bits).  Wide: primes below 64768 from the
// top (their balanced residues split into two INT8 digits).
static const std::vector<int> &narrow_moduli()
{
    static const std::vector<int> v = [] {
        std::vector<int> out;
        for (int m = 256; m >= 2; --m) {
            bool ok = true;
            for (int x : out) {
                int a = m, b = x;
                while (b) {
                    const int t = a % b;
                    a = b;
                    b = t;
                }
                if (a != 1) {
                    ok = false;
                    break;
                }
            }
            // below 2^4 a narrow modulus adds fewer bits per GEMM than a wide one
            if (ok && m > 16) out.push_back(m);
        }
        return out;
    }();
    return v;
}

static const std::vector<int> &wide_moduli()
{
    static const std::vector<int> v = [] {
        std::vector<int> out;
        for (int p = 64767; p > 257 && out.size() < 192; p -= 2) {
            bool prime = true;
            for (int d = 3; d * d <= p; d += 2)
                if (p % d == 0) {
                    prime = false;
                    break;
                }
            if (prime) out.push_back(p);
        }
        return out;
    }();
    return v;
}

typedef std::vector<uint32_t> Big;

static void big_mul_small(Big &a, uint32_t f)
{
    uint64_t c = 0;
    for (auto &w : a) {
        const uint64_t t = (uint64_t)w * f + c;
        w = (uint32_t)t;
        c = t >> 32;
    }
    if (c) a.push_back((uint32_t)c);
}

static int big_bitlen(const Big &a)
{
    for (int i = (int)a.size() - 1; i >= 0; --i)
        if (a[i]) return 32 * i + 32 - __builtin_clz(a[i]);
    return 0;
}

static long long inv_mod(long long a, long long m)
{
    long long g = m, x = 0, x1 = 1, aa = a % m;
    while (aa) {
        const long long qq = g / aa;
        long long t = g - qq * aa;
        g = aa;
        aa = t;
        t = x - qq * x1;
        x = x1;
        x1 = t;
    }
    return ((x % m) + m) % m;
}

static std::vector<int> moduli_of(int n8, int n16)
{
    std::vector<int> v(narrow_moduli().begin(), narrow_moduli().begin() + n8);
    v.insert(v.end(), wide_moduli().begin(), wide_moduli().begin() + n16);
    return v;
}

// Cheapest set (n8 + 4 n16 GEMM units) with M >= 2^need; *bits = bit length of M.
static bool choose_moduli(int need, int &n8, int &n16, int *bits)
{
    const auto &a = narrow_moduli();
    const auto &b = wide_moduli();
    std::vector<double> c8(a.size() + 1, 0.0);
    for (size_t i = 0; i < a.size(); ++i) c8[i + 1] = c8[i] + std::log2((double)a[i]);
    int best = INT_MAX;
    double bits16 = 0.0;
    for (size_t j = 0; j <= b.size(); ++j) {
        const double rem = need + 1e-6 - bits16;
        int i8 = -1;
        for (size_t i = 0; i <= a.size(); ++i)
            if (c8[i] >= rem) {
                i8 = (int)i;
                break;
            }
        if (i8 >= 0 && i8 + 4 * (int)j < best) {
            best = i8 + 4 * (int)j;
            n8 = i8;
            n16 = (int)j;
        }
        if (j < b.size()) bits16 += std::log2((double)b[j]);
    }
    if (best == INT_MAX) return false;
    // confirm with exact arithmetic (the log sums carry rounding)
    for (;;) {
        Big M{1u};
        for (int m : moduli_of(n8, n16)) big_mul_small(M, (uint32_t)m);
        *bits = big_bitlen(M);
        if (*bits >= need + 1) return true;
        if (n8 < (int)a.size())
            ++n8;
        else if (n16 < (int)b.size())
            ++n16;
        else
            return false;
    }
}

struct ResTabHost {
    ResTab t{};
    void *dev = nullptr;
};

static std::mutex g_res_mu;
static std::map<std::tuple<int, int, int, int>, ResTabHost *> g_res;

// device tables of a modulus set with CRT constants of nwc words (cached)
static ResTabHost *res_table(int n8, int n16, int nwc)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> lock(g_res_mu);
    const auto key = std::make_tuple(dev, n8, n16, nwc);
    auto it = g_res.find(key);
    if (it != g_res.end()) return it->second;

    const std::vector<int> mods = moduli_of(n8, n16);
    const int n = (int)mods.size();
    std::vector<double> md(n), inv(n), frac(n);
    std::vector<uint32_t> cw((size_t)n * nwc, 0u), Mw(nwc, 0u), p8((size_t)std::max(n8, 1) * RES_P8, 0u);
    std::vector<int> pb((size_t)std::max(n16, 1) * RES_PB, 0);
    Big M{1u};
    for (int m : mods) big_mul_small(M, (uint32_t)m);
    if ((int)M.size() > nwc - 1) return nullptr;
    std::copy(M.begin(), M.end(), Mw.begin());
    for (int i = 0; i < n; ++i) {
        const long long mi = mods[i];
        Big Mi{1u};
        long long rem = 1;
        for (int j = 0; j < n; ++j)
            if (j != i) {
                big_mul_small(Mi, (uint32_t)mods[j]);
                rem = rem * (mods[j] % mi) % mi;
            }
        const long long yi = inv_mod(rem, mi);
        big_mul_small(Mi, (uint32_t)yi);
        std::copy(Mi.begin(), Mi.begin() + std::min((int)Mi.size(), nwc), cw.begin() + (size_t)i * nwc);
        md[i] = (double)mi;
        inv[i] = 1.0 / (double)mi;
        frac[i] = (double)yi / (double)mi;
    }
    for (int i = 0; i < n8; ++i) {
        long long p = 1 % mods[i];
        for (int h = 0; h < RES_P8; ++h) {
            p8[(size_t)i * RES_P8 + h] = (uint32_t)p;
            p = (p << 16) % mods[i];
        }
    }
    for (int j = 0; j < n16; ++j) {
        const long long mj = mods[n8 + j];
        long long p = 1;
        for (int b = 0; b < RES_PB; ++b) {
            pb[(size_t)j * RES_PB + b] = (int)(2 * p >= mj ? p - mj : p);
            p = (p << 8) % mj;
        }
    }
    const size_t bd = sizeof(double) * 3 * n, bc = 4 * cw.size(), bm = 4 * Mw.size(), b8 = 4 * p8.size(),
                 bb = 4 * pb.size();
    char *d = nullptr;
    if (hipMalloc(&d, bd + bc + bm + b8 + bb) != hipSuccess) return nullptr;
    std::vector<char> h(bd + bc + bm + b8 + bb);
    std::copy((char *)md.data(), (char *)md.data() + 8 * n, h.data());
    std::copy((char *)inv.data(), (char *)inv.data() + 8 * n, h.data() + 8 * n);
    std::copy((char *)frac.data(), (char *)frac.data() + 8 * n, h.data() + 16 * n);
    std::copy((char *)cw.data(), (char *)cw.data() + bc, h.data() + bd);
    std::copy((char *)Mw.data(), (char *)Mw.data() + bm, h.data() + bd + bc);
    std::copy((char *)p8.data(), (char *)p8.data() + b8, h.data() + bd + bc + bm);
    std::copy((char *)pb.data(), (char *)pb.data() + bb, h.data() + bd + bc + bm + b8);
    if (hipMemcpy(d, h.data(), h.size(), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(d);
        return nullptr;
    }
    ResTabHost *r = new ResTabHost();
    r->dev = d;
    r->t.n8 = n8;
    r->t.n16 = n16;
    r->t.nwc = nwc;
    r->t.md = (const double *)d;
    r->t.inv = (const double *)(d + 8 * n);
    r->t.frac = (const double *)(d + 16 * n);
    r->t.cw = (const uint32_t *)(d + bd);
    r->t.Mw = (const uint32_t *)(d + bd + bc);
    r->t.p8 = (const uint32_t *)(d + bd + bc + bm);
    r->t.pb = (const int *)(d + bd + bc + bm + b8);
    g_res[key] = r;
    return r;
}

// ------------------------------------------------------------ dispatch --
// fixed-point words the residue kernels build (|N| < 2^(W-2) plus a sign bit)
template <int L>
static int nwu_bucket(int W)
{
    constexpr int NW = FixedWords<L>::NW;
    const int need = (W + 31) / 32;
    int b = (int)round_up(need, 4);
    const int lo = (L == 2) ? 4 : (L == 4) ? 8 : (L == 8) ? 12 : 24;
    return std::min(NW, std::max(lo, b));
}

// CRT constant words per limb count: a multiple of 4 in [lo, hi]
static int nwc_bucket(int L, int words)
{
    const int lo = (L == 2) ? 8 : (L == 4) ? 12 : (L == 8) ? 24 : 40;
    const int step = (L <= 4) ? 4 : 8;
    const int b = (int)round_up(std::max(words, lo), step);
    return b;
}

template <int L, int NWU>
static void launch_residues(int m, int n, int k, int mp, int np, int kp, const double *A, long long lda,
                            const double *B, long long ldb, const int *E, const int *F, int W, const ResTab &t,
                            int8_t *A8, int8_t *A16, int8_t *B8, int8_t *B16, hipStream_t s)
{
    hipLaunchKernelGGL((residue_a_kernel<L, NWU>), dim3(grid_for((long long)mp * kp, 256)), dim3(256), 0, s, m, k,
                       mp, kp, A, lda, E, W, t, A8, A16);
    hipLaunchKernelGGL((residue_b_kernel<L, NWU>), dim3(grid_for((long long)np * kp, 256)), dim3(256), 0, s, k, n,
                       kp, np, B, ldb, F, W, t, B8, B16);
}

template <int L>
static int residues_dispatch(int nwu, int m, int n, int k, int mp, int np, int kp, const double *A, long long lda,
                             const double *B, long long ldb, const int *E, const int *F, int W, const ResTab &t,
                             int8_t *A8, int8_t *A16, int8_t *B8, int8_t *B16, hipStream_t s)
{
#define MDG_RES_CASE(NWU)                                                                                     \
    case NWU:                                                                                                 \
        if constexpr (NWU <= FixedWords<L>::NW) {                                                             \
            launch_residues<L, NWU>(m, n, k, mp, np, kp, A, lda, B, ldb, E, F, W, t, A8, A16, B8, B16, s); \
            return MDG_SUCCESS;                                                                               \
        }                                                                                                     \
        break;
    switch (nwu) {
        MDG_RES_CASE(4)
        MDG_RES_CASE(8)
        MDG_RES_CASE(12)
        MDG_RES_CASE(16)
        MDG_RES_CASE(20)
        MDG_RES_CASE(24)
        MDG_RES_CASE(28)
        MDG_RES_CASE(32)
        MDG_RES_CASE(36)
    }
#undef MDG_RES_CASE
    set_error("no residue kernel for %d fixed-point words", nwu);
    return MDG_ERROR_RANGE;
}

template <int L, int NWC>
static void launch_reconstruct(int m, int n, int W, const ResTab &t, int nck1, int nck2, const int *D8,
                               const int *Dhh, const int *Dll, const int *Dx, const int *E, const int *F, double *C,
                               long long ldc, hipStream_t s)
{
    hipLaunchKernelGGL((residue_reconstruct_kernel<L, NWC>), dim3(grid_for((long long)m * n, 256)), dim3(256), 0, s,
                       m, n, W, t, nck1, nck2, D8, Dhh, Dll, Dx, E, F, C, ldc);
}

template <int L>
static int reconstruct_dispatch(int nwc, int m, int n, int W, const ResTab &t, int nck1, int nck2, const int *D8,
                                const int *Dhh, const int *Dll, const int *Dx, const int *E, const int *F,
                                double *C, long long ldc, hipStream_t s)
{
#define MDG_REC_CASE(NWC)                                                                              \
    case NWC:                                                                                          \
        launch_reconstruct<L, NWC>(m, n, W, t, nck1, nck2, D8, Dhh, Dll, Dx, E, F, C, ldc, s);         \
        return MDG_SUCCESS;
    if constexpr (L == 2) {
        switch (nwc) { MDG_REC_CASE(8) MDG_REC_CASE(12) MDG_REC_CASE(16) MDG_REC_CASE(20) }
    } else if constexpr (L == 4) {
        switch (nwc) { MDG_REC_CASE(12) MDG_REC_CASE(16) MDG_REC_CASE(20) MDG_REC_CASE(24) MDG_REC_CASE(28) }
    } else if constexpr (L == 8) {
        switch (nwc) { MDG_REC_CASE(24) MDG_REC_CASE(32) MDG_REC_CASE(40) MDG_REC_CASE(48) }
    } else {
        switch (nwc) {
            MDG_REC_CASE(40) MDG_REC_CASE(48) MDG_REC_CASE(56) MDG_REC_CASE(64) MDG_REC_CASE(72) MDG_REC_CASE(80)
        }
    }
#undef MDG_REC_CASE
    set_error("no reconstruction kernel for %d CRT words at L=%d", nwc, L);
    return MDG_ERROR_RANGE;
}

// one strided-batched INT8 GEMM per K-chunk: D[b][q] (n x m col-major view of
// the row-major m x n plane) = Bp[b]^T-op * Ap[b] over depth [k0, k0 + kc)
static int res_batched(rocblas_handle h, int batch, int m, int n, int depth, const int8_t *Bp, long long ldb,
                       long long sb, const int8_t *Ap, long long lda, long long sa, int *D, int nck)
{
    const int32_t alpha = 1, beta = 0;
    const long long mn = (long long)m * n;
    for (int q = 0; q * (long long)RES_KC < depth; ++q) {
        const int k0 = q * RES_KC;
        const int kc = std::min(RES_KC, depth - k0);
        int *Dq = D + (long long)q * mn;
        rocblas_status st = rocblas_gemm_strided_batched_ex(
            h, rocblas_operation_transpose, rocblas_operation_none, n, m, kc, &alpha, Bp + k0, rocblas_datatype_i8_r,
            ldb, sb, Ap + k0, rocblas_datatype_i8_r, lda, sa, &beta, Dq, rocblas_datatype_i32_r, n, mn * nck, Dq,
            rocblas_datatype_i32_r, n, mn * nck, batch, rocblas_datatype_i32_r, rocblas_gemm_algo_standard, 0, 0);
        if (st != rocblas_status_success) {
            set_error("rocblas_gemm_strided_batched_ex (int8) failed with status %d", (int)st);
            return MDG_ERROR_HIP;
        }
    }
    return MDG_SUCCESS;
}

template <int L>
static int run_residue_gemm(int m, int n, int k, const double *A, long long lda, const double *B, long long ldb,
                            double *C, long long ldc, int target_bits, int fixed_spread, mdg_gemm_info *info,
                            hipStream_t s)
{
    StageTimer tm(s, profiling_on());
    tm.mark(0);
    MDG_REQUIRE(k <= (1 << 22), MDG_ERROR_RANGE, "the residue GEMM supports k <= 2^22 (k=%d)", k);
    int *E = nullptr, *F = nullptr;
    int spread_a = 0, spread_b = 0;
    int rc = scan_scales(L, m, n, k, A, lda, B, ldb, fixed_spread, s, &E, &F, &spread_a, &spread_b);
    if (rc) return rc;
    tm.mark(1);

    int lgL = 0;
    while ((1 << lgL) < L) ++lgL;
    int lgk = 0;
    while ((1ll << lgk) < k) ++lgk;
    const int W = target_bits + spread_a + spread_b + 3 + lgL;
    MDG_REQUIRE(W <= 32 * FixedWords<L>::NW - 1, MDG_ERROR_RANGE,
                "needs %d fixed-point bits, more than the %d this limb count holds", W, 32 * FixedWords<L>::NW - 1);
    const int need = 2 * W - 1 + lgk;   // M >= 2^need > 4 k 2^(2W-4) >= 4 max|Z|
    int n8 = 0, n16 = 0, mbits = 0;
    MDG_REQUIRE(choose_moduli(need, n8, n16, &mbits), MDG_ERROR_RANGE, "no modulus set reaches %d bits", need);
    const int nwu = nwu_bucket<L>(W);
    const int nwc = nwc_bucket(L, (mbits + 31) / 32 + 1);   // M's words and a sign word
    ResTabHost *th = res_table(n8, n16, nwc);
    MDG_REQUIRE(th, MDG_ERROR_HIP, "could not build the modulus tables (%d + %d moduli)", n8, n16);
    const ResTab &t = th->t;

    const int kp = (int)round_up(k, 128), mp = (int)round_up(m, 128), np = (int)round_up(n, 128);
    const int nck1 = (kp + RES_KC - 1) / RES_KC, nck2 = (2 * kp + RES_KC - 1) / RES_KC;
    const long long mn = (long long)m * n;
    const size_t a8 = (size_t)n8 * mp * kp, a16 = (size_t)n16 * mp * 2 * kp;
    const size_t b8 = (size_t)n8 * np * kp, b16 = (size_t)n16 * np * 3 * kp;
    const size_t nplanes = (size_t)n8 * nck1 + (size_t)n16 * (2 * nck1 + nck2);
    int8_t *Ares = (int8_t *)scratch(s, a8 + a16 + 256, 2);
    int8_t *Bres = (int8_t *)scratch(s, b8 + b16 + 256, 3);
    int *D = (int *)scratch(s, nplanes * mn * 4, 4);
    MDG_REQUIRE(Ares && Bres && D, MDG_ERROR_OUT_OF_MEMORY, "scratch allocation failed");
    int8_t *A8 = Ares, *A16 = Ares + a8, *B8 = Bres, *B16 = Bres + b8;
    int *D8 = D, *Dhh = D + (size_t)n8 * nck1 * mn, *Dll = Dhh + (size_t)n16 * nck1 * mn,
        *Dx = Dll + (size_t)n16 * nck1 * mn;

    rc = residues_dispatch<L>(nwu, m, n, k, mp, np, kp, A, lda, B, ldb, E, F, W, t, A8, A16, B8, B16, s);
    if (rc) return rc;
    MDG_LAUNCH_CHECK("residue kernels");
    tm.mark(2);
    {
        BlasHandle *bh = blas_handle();
        MDG_REQUIRE(bh, MDG_ERROR_HIP, "rocblas_create_handle failed");
        std::lock_guard<std::mutex> lock(bh->mu);
        if (rocblas_set_stream(bh->h, s) != rocblas_status_success) {
            set_error("rocblas_set_stream failed");
            return MDG_ERROR_HIP;
        }
        const long long sa8 = (long long)mp * kp, sb8 = (long long)np * kp;
        const long long sa16 = (long long)mp * 2 * kp, sb16 = (long long)np * 3 * kp;
        if (n8 && (rc = res_batched(bh->h, n8, m, n, kp, B8, kp, sb8, A8, kp, sa8, D8, nck1))) return rc;
        if (n16) {
            // hi*hi, lo*lo, and (hi, lo) . (lo, hi) of depth 2kp
            if ((rc = res_batched(bh->h, n16, m, n, kp, B16, 3ll * kp, sb16, A16, 2ll * kp, sa16, Dhh, nck1)))
                return rc;
            if ((rc = res_batched(bh->h, n16, m, n, kp, B16 + kp, 3ll * kp, sb16, A16 + kp, 2ll * kp, sa16, Dll,
                                  nck1)))
                return rc;
            if ((rc = res_batched(bh->h, n16, m, n, 2 * kp, B16 + kp, 3ll * kp, sb16, A16, 2ll * kp, sa16, Dx,
                                  nck2)))
                return rc;
        }
    }
    tm.mark(3);
    rc = reconstruct_dispatch<L>(nwc, m, n, W, t, nck1, nck2, D8, Dhh, Dll, Dx, E, F, C, ldc, s);
    if (rc) return rc;
    MDG_LAUNCH_CHECK("residue_reconstruct_kernel");
    tm.mark(4);
    float stage[4] = {0.f, 0.f, 0.f, 0.f};
    if (tm.on) MDG_HIP_CHECK(tm.elapsed(stage));
    if (info) {
        info->method = MDG_METHOD_INT8_RESIDUE;
        info->slices = n8 + n16;
        info->slice_bits = 8;
        info->diagonals = n16;
        info->pair_gemms = n8 + 4ll * n16;
        info->total_bits = W;
        info->spread_a = spread_a;
        info->spread_b = spread_b;
        info->flushed = (nck1 > 1 || nck2 > 1) ? 1 : 0;
        info->backend = 1;
        for (int i = 0; i < 4; ++i) info->stage_ms[i] = stage[i];
    }
    return MDG_SUCCESS;
}

int residue_gemm(int L, int m, int n, int k, const double *A, long long lda, const double *B, long long ldb,
                 double *C, long long ldc, int target_bits, int fixed_spread, mdg_gemm_info *info, hipStream_t s)
{
    if (target_bits <= 0) target_bits = 53 * L;
    switch (L) {
    case 2: return run_residue_gemm<2>(m, n, k, A, lda, B, ldb, C, ldc, target_bits, fixed_spread, info, s);
    case 4: return run_residue_gemm<4>(m, n, k, A, lda, B, ldb, C, ldc, target_bits, fixed_spread, info, s);
    case 8: return run_residue_gemm<8>(m, n, k, A, lda, B, ldb, C, ldc, target_bits, fixed_spread, info, s);
    case 16: return run_residue_gemm<16>(m, n, k, A, lda, B, ldb, C, ldc, target_bits, fixed_spread, info, s);
    }
    set_error("unsupported limb count L=%d", L);
    return MDG_ERROR_INVALID_VALUE;
}

}  // namespace mdg
