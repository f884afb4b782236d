// batched.hip -- many independent multi-double GEMMs of uneven shape.
//
// persistent = 0: a loop of per-problem calls on the caller's stream (the
//                 baseline a caller of the reference would write around
//                 matmulTCnt, C++/TCcuda.cu:248-302).
// persistent = 1: the whole batch as one pipeline of stage launches, each
//                 launch covering every problem through a device table of
//                 (problem, block) entries: exponent scans, slicing, ONE
//                 launch of INT8 diagonal-GEMM tiles of all problems (or one
//                 vector-GEMM launch per limb count), reconstruction.  The
//                 pipeline is recorded once into a hipGraph keyed by the
//                 batch's shape signature and replayed on later calls.
// Slice counts come from the caller's spread allowance (no device->host
// read), so nothing in the batch synchronises the stream.  The INT8
// residue method shares the scans and continues with the fused residue
// stages of residue_gemm.hip (residues, one launch of INT8 MFMA tiles of all
// residue planes, CRT reconstruction).  Other methods, and residue batches
// with a problem deeper than one INT32 plane allows, run the per-problem
// loop inside the graph.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "common.hpp"
#include "blkslice.hpp"
#include "i8band.hpp"
#include "i8tile.hpp"
#include "md_arith.hpp"
#include "recon.hpp"
#include "slices.hpp"

namespace mdg {

int vector_gemm(int L, int m, int n, int k, const double *A, long long lda, const double *B, long long ldb,
                double *C, long long ldc, int twoprod, hipStream_t s);
int split_gemm(int L, int m, int n, int k, const double *A, long long lda, const double *B, long long ldb, double *C,
               long long ldc, int method, int target_bits, int fixed_spread, mdg_gemm_info *info, hipStream_t s);
int plan_int8_slices(int L, int target_bits, int spread_a, int spread_b);
struct ResBatch;
int residue_batch_build(int count, const mdg_problem *pr, int *const *E, int *const *F, int target_bits, int spread,
                        ResBatch **out);
int residue_batch_run(const ResBatch *rb, hipStream_t s, hipEvent_t g0, hipEvent_t g1);
int residue_batch_candidates(const ResBatch *rb);
void residue_batch_select(ResBatch *rb, int c);
void residue_batch_free(ResBatch *rb);

constexpr int BZERO_SCALE = -4000;

// one problem of the fused pipeline, as the kernels see it
struct BProb {
    int L, m, n, k, S, kp, mp, np, kmax, nplanes, w;
    long long lda, ldb, ldc;
    const double *A, *B;
    double *C;
    int8_t *Aall, *Ball;
    long long *D;   // band planes (i8band.hpp), INT64
    int *E, *F;
    double *colsum;
    int *flag;     // the problem's flags: bit 0 non-finite input, bit 1 spread above `allow`
    int allow;     // exponent-spread allowance of the slice plan (bits)
};

// one block's work: problem and up to three block coordinates
struct BEnt {
    int prob, x, y, z;
};

__device__ __forceinline__ int bscale_exponent(double s)
{
    if (s == 0.0) return INT_MIN;
    int ex;
    frexp(s * (1.0 + 0x1p-40), &ex);
    return ex;
}

// ---------------------------------------------------------------- scans --
__global__ void b_init_kernel(const BProb *__restrict__ P, const BEnt *__restrict__ T)
{
    const BEnt e = T[blockIdx.x];
    const BProb &p = P[e.prob];
    const int c = e.x * 256 + threadIdx.x;
    if (c < p.n) {
        p.F[c] = INT_MIN;
        p.colsum[c] = 0.0;
    }
    if (e.x == 0 && threadIdx.x == 0) *p.flag = 0;
}

// the measured exponent spread of one row or column (exponent_spread, as
// the single-problem scans measure it)
__device__ __forceinline__ int bspread(int e, double sum0, int k) { return exponent_spread(e, sum0, k); }

template <int L>
__global__ __launch_bounds__(256) void b_scan_rows_kernel(const BProb *__restrict__ P, const BEnt *__restrict__ T)
{
    const BEnt e = T[blockIdx.x];
    const BProb &p = P[e.prob];
    const int lane = threadIdx.x & 63;
    const long long r = (long long)e.x * 4 + (threadIdx.x >> 6);
    if (r >= p.m) return;
    int emax = INT_MIN, bad = 0;
    double sum0 = 0.0;
    for (int kk = lane; kk < p.k; kk += 64) {
        const double *x = p.A + (r * p.lda + kk) * L;
        double s = 0.0;
#pragma unroll
        for (int l = 0; l < L; ++l) {
            bad |= !isfinite(x[l]);
            s += fabs(x[l]);
        }
        emax = max(emax, bscale_exponent(s));
        sum0 += fabs(x[0]);
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        emax = max(emax, __shfl_xor(emax, o));
        sum0 += __shfl_xor(sum0, o);
        bad |= __shfl_xor(bad, o);
    }
    if (lane == 0) {
        p.E[r] = (emax == INT_MIN) ? BZERO_SCALE : emax;
        int f = bad ? 1 : 0;
        if (emax != INT_MIN && sum0 > 0.0 && bspread(emax, sum0, p.k) > p.allow) f |= 2;
        if (f) atomicOr(p.flag, f);
    }
}

template <int L>
__global__ __launch_bounds__(256) void b_scan_cols_kernel(const BProb *__restrict__ P, const BEnt *__restrict__ T)
{
    const BEnt e = T[blockIdx.x];
    const BProb &p = P[e.prob];
    __shared__ int se[4][64];
    __shared__ double ss[4][64];
    const int cl = threadIdx.x & 63, grp = threadIdx.x >> 6;
    const long long c = (long long)e.x * 64 + cl;
    const int k0 = e.y * 256, k1 = min(p.k, k0 + 256);
    int emax = INT_MIN, bad = 0;
    double sum0 = 0.0;
    if (c < p.n) {
        for (int kk = k0 + grp; kk < k1; kk += 4) {
            const double *x = p.B + ((long long)kk * p.ldb + c) * L;
            double s = 0.0;
#pragma unroll
            for (int l = 0; l < L; ++l) {
                bad |= !isfinite(x[l]);
                s += fabs(x[l]);
            }
            emax = max(emax, bscale_exponent(s));
            sum0 += fabs(x[0]);
        }
    }
    se[grp][cl] = emax;
    ss[grp][cl] = sum0;
    if (bad) atomicOr(p.flag, 1);
    __syncthreads();
    if (grp == 0 && c < p.n) {
        const int m4 = max(max(se[0][cl], se[1][cl]), max(se[2][cl], se[3][cl]));
        const double s4 = (ss[0][cl] + ss[1][cl]) + (ss[2][cl] + ss[3][cl]);
        if (m4 != INT_MIN) atomicMax(&p.F[c], m4);
        if (s4 != 0.0) atomicAdd(&p.colsum[c], s4);
    }
}

__global__ void b_finish_cols_kernel(const BProb *__restrict__ P, const BEnt *__restrict__ T)
{
    const BEnt e = T[blockIdx.x];
    const BProb &p = P[e.prob];
    const int c = e.x * 256 + threadIdx.x;
    if (c >= p.n) return;
    if (p.F[c] == INT_MIN) {
        p.F[c] = BZERO_SCALE;
        return;
    }
    if (p.colsum[c] > 0.0 && bspread(p.F[c], p.colsum[c], p.k) > p.allow) atomicOr(p.flag, 2);
}

// Last stage: a problem flagged by the scans (non-finite input, or an
// exponent spread beyond the allowance its slices were sized for) gets C
// filled with NaN instead of a silently inaccurate result.  One block per
// problem.
__global__ void b_guard_kernel(const BProb *__restrict__ P, const BEnt *__restrict__ T)
{
    const BEnt e = T[blockIdx.x];
    const BProb &p = P[e.prob];
    if (*p.flag == 0) return;
    const long long nel = (long long)p.m * p.n * p.L;
    for (long long i = threadIdx.x; i < nel; i += blockDim.x) {
        const long long r = i / ((long long)p.n * p.L), rem = i - r * p.n * p.L;
        p.C[r * p.ldc * p.L + rem] = __builtin_nan("");
    }
}

// --------------------------------------------------------------- slices --
// slices in the blocked layout of the band kernel (blkslice.hpp): entry
// (prob, row tile x, k-block y), blk_rows_per_tile(S) rows (A) or columns (B)
// of 128 k-bytes each; the padding of the plan's slices is zeroed when the
// plan is built and never written again, so row tiles wholly past m (n) get
// no entry and the digits past k are not stored
template <int L, bool TRANS>
__global__ __launch_bounds__(256) void b_slice_kernel(const BProb *__restrict__ P, const BEnt *__restrict__ T)
{
    extern __shared__ __attribute__((aligned(16))) int8_t stage[];   // the stage's largest blk_stage_bytes
    const BEnt e = T[blockIdx.x];
    const BProb &p = P[e.prob];
    with_slice_words<L>(p.S, p.w, [&](auto nw) {
        slice_tile<L, TRANS, int8_t, decltype(nw)::value, true>(TRANS ? p.n : p.m, p.k, p.kp, TRANS ? p.B : p.A,
                                                          TRANS ? p.ldb : p.lda, TRANS ? p.F : p.E, BZERO_SCALE, p.S,
                                                          p.w, e.y, e.x * blk_rows_per_tile(p.S),
                                                          TRANS ? p.Ball : p.Aall, stage);
    });
}

// ------------------------------------------------------- INT8 band tiles --

// Entry (prob, tile row, tile column, item): one band item (i8band.hpp) of
// one 128 x 128 tile of one problem; the item packed as g0 & 0xff | kb0 << 8
// | kb1 << 19 (g0 >= -(G-1) as a signed byte, k-blocks kb0 .. kb1-1 < 2048),
// the band width is the launch's G; prob < 0: an empty entry.  An item runs
// its k-blocks times its slice steps, so every workgroup is deep enough to
// keep its copy pipeline full, however shallow the problem.
template <int G>
__global__ __launch_bounds__(256) void b_band_i8_kernel(const BProb *__restrict__ P, const BEnt *__restrict__ T)
{
    __shared__ __attribute__((aligned(1024))) int8_t lds[bnd_lds_bytes<G, 2>()];
    const BEnt e = T[blockIdx.x];
    if (e.prob < 0) return;
    const BProb &p = P[e.prob];
    const int g0 = (int)(signed char)(e.z & 0xff), kb0 = (e.z >> 8) & 0x7ff, kb1 = (e.z >> 19) & 0x7ff;
    const int nkb = p.kp >> 7, t1 = g0 + G;
    const long long tstride = (long long)nkb * p.S * BND_BLK;
    const int8_t *At = p.Aall + e.x * tstride, *Bt = p.Ball + e.y * tstride;
    const long long row0 = (long long)e.x * BND_TM, col0 = (long long)e.y * BND_TM;
    // 32-row MFMA tiles that hold rows below m: the others only multiply the
    // zero padding of the tile and are left out
    switch ((int)min(4ll, (p.m - row0 + 31) >> 5)) {
    case 1: band_run<G, 2, 1, false, 1>(lds, At, Bt, p.S, nkb, BND_CAP_I8, g0, 0, t1, kb0, kb1, row0, col0, p.m, p.n, p.D); break;
    case 2: band_run<G, 2, 1, false, 2>(lds, At, Bt, p.S, nkb, BND_CAP_I8, g0, 0, t1, kb0, kb1, row0, col0, p.m, p.n, p.D); break;
    case 3: band_run<G, 2, 1, false, 3>(lds, At, Bt, p.S, nkb, BND_CAP_I8, g0, 0, t1, kb0, kb1, row0, col0, p.m, p.n, p.D); break;
    default: band_run<G, 2, 1, false, 4>(lds, At, Bt, p.S, nkb, BND_CAP_I8, g0, 0, t1, kb0, kb1, row0, col0, p.m, p.n, p.D); break;
    }
}

// ------------------------------------------------------- reconstruction --
// the band planes folded as in the single-problem path (recon.hpp
// band_fold), so the fused pipeline and the per-problem loop give identical
// bits
template <int L, int NWB, bool CHUNKED>
__global__ __launch_bounds__(256) void b_reconstruct_kernel(const BProb *__restrict__ P, const BEnt *__restrict__ T)
{
    const BEnt e = T[blockIdx.x];
    const BProb &p = P[e.prob];
    const long long idx = (long long)e.x * 256 + threadIdx.x;
    if (idx >= (long long)p.m * p.n) return;
    const long long r = idx / p.n, c = idx - r * p.n;
    double y[L];
#pragma unroll
    for (int l = 0; l < L; ++l) y[l] = 0.0;
    const int er = p.E[r], fc = p.F[c];
    if (er != BZERO_SCALE && fc != BZERO_SCALE) {
        const long long mn = (long long)p.m * p.n;
        const int S = p.S, nkb = p.kp >> 7;
        // the planes live in global memory (a generic pointer read from the
        // problem table would make every load a flat one)
        const __attribute__((address_space(1))) long long *D =
            (const __attribute__((address_space(1))) long long *)p.D;
        band_fold<L, NWB, 4, CHUNKED>(
            S, er + fc + 4 - 8 * (S + 1),
            [&](int plane) { return (long long)D[(long long)plane * mn + idx]; },
            [&](int q) { return bnd_band_chunks(S, 4, nkb, BND_CAP_I8, q); }, p.nplanes, y);
    }
    double *o = p.C + (r * p.ldc + c) * L;
#pragma unroll
    for (int l = 0; l < L; ++l) o[l] = y[l];
}

// ----------------------------------------------------- fused vector GEMM --
// Entry (prob, tile_m, tile_n): 16 x 16 outputs of one problem, one thread
// per output, the same k-chain as vector_gemm.hip (bit-identical results).
template <int L>
__global__ __launch_bounds__(256) void b_vector_kernel(const BProb *__restrict__ P, const BEnt *__restrict__ T)
{
    constexpr int TM = 16, TN = 16, TK = 16;
    __shared__ double As[L][TM][TK + 1];
    __shared__ double Bs[L][TK][TN + 1];
    const BEnt e = T[blockIdx.x];
    const BProb &p = P[e.prob];
    const int tid = threadIdx.x, tx = tid % TN, ty = tid / TN;
    const long long r0 = (long long)e.x * TM, c0 = (long long)e.y * TN;
    double acc[L];
#pragma unroll
    for (int l = 0; l < L; ++l) acc[l] = 0.0;
    for (int k0 = 0; k0 < p.k; k0 += TK) {
        for (int idx = tid; idx < TM * TK * L; idx += 256) {
            const int i = idx / (TK * L), rem = idx - i * (TK * L), kk = rem / L, l = rem - kk * L;
            const long long gr = r0 + i;
            const int gk = k0 + kk;
            As[l][i][kk] = (gr < p.m && gk < p.k) ? p.A[(gr * p.lda + gk) * L + l] : 0.0;
        }
        for (int idx = tid; idx < TK * TN * L; idx += 256) {
            const int kk = idx / (TN * L), rem = idx - kk * (TN * L), j = rem / L, l = rem - j * L;
            const int gk = k0 + kk;
            const long long gc = c0 + j;
            Bs[l][kk][j] = (gk < p.k && gc < p.n) ? p.B[((long long)gk * p.ldb + gc) * L + l] : 0.0;
        }
        __syncthreads();
        const int kmax = min(TK, p.k - k0);
        for (int kk = 0; kk < kmax; ++kk) {
            double a[L], b[L], q[L];
#pragma unroll
            for (int l = 0; l < L; ++l) {
                a[l] = As[l][ty][kk];
                b[l] = Bs[l][kk][tx];
            }
            md_mul<L, TP_FMA>(a, b, q);
            md_add<L>(acc, q, acc);
        }
        __syncthreads();
    }
    const long long r = r0 + ty, c = c0 + tx;
    if (r < p.m && c < p.n) {
#pragma unroll
        for (int l = 0; l < L; ++l) p.C[(r * p.ldc + c) * L + l] = acc[l];
    }
}

// ------------------------------------------------------------------ plan --
static long long rup(long long x, long long q) { return (x + q - 1) / q * q; }

struct Stage {
    int kind;      // 0 init, 1 scan rows, 2 scan cols, 3 finish, 4 slice A, 5 slice B, 6 gemm, 7 recon,
                   // 8 vector, 9 guard (NaN-fill flagged problems)
    int L;         // limb count of the stage's problems (0: any)
    long long first, count;   // range in the entry table
    int nw = 0;    // reconstruction: the largest slice count of its problems
    bool chunked = false;   // reconstruction: some problem has K-chunk planes
    int tm = 128;  // GEMM: tile rows and columns
    int g = 0;     // GEMM: band width of the band-kernel launch
    int lds = 0;   // slicing: dynamic LDS bytes (the largest any problem of the stage needs)
};

struct Plan {
    std::vector<BProb> probs;
    std::vector<BEnt> ents;
    std::vector<Stage> stages;
    Stage guard{9, 0, 0, 0};    // last stage of slice / residue pipelines
    BProb *dprobs = nullptr;
    BEnt *dents = nullptr;
    void *arena = nullptr;
    int *flags = nullptr;       // per-problem flags (in the arena, or own buffer)
    void *work = nullptr;       // loop-in-graph plans: their own per-problem workspace
    size_t work_bytes = 0;
    size_t bytes = 0;           // device bytes the plan holds
    int count = 0, dev = 0;
    hipStream_t stream = nullptr;
    unsigned long long last_use = 0;
    hipEvent_t gev[2] = {nullptr, nullptr};   // around the GEMM stage (built with profiling on)
    hipGraphExec_t exec = nullptr;
    ResBatch *rb = nullptr;       // INT8 residue stages after the scans
    ~Plan()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        residue_batch_free(rb);
        (void)hipFree(dprobs);
        (void)hipFree(dents);
        (void)hipFree(arena);
        (void)hipFree(work);
        if (flags && !arena) (void)hipFree(flags);
        for (auto &e : gev)
            if (e) (void)hipEventDestroy(e);
    }
};

static int build_plan(Plan &pl, int count, const mdg_problem *pr, int method, int target_bits, int spread)
{
    const bool i8 = (method == MDG_METHOD_INT8_OZAKI);
    const bool res = (method == MDG_METHOD_INT8_RESIDUE);   // scans here, the rest in residue_gemm.hip
    // arena layout
    size_t bytes = 0;
    auto take = [&](size_t b) {
        size_t off = bytes;
        bytes += (b + 255) / 256 * 256;
        return off;
    };
    struct Offs {
        size_t aall, ball, d, e, f, cs;
    };
    std::vector<Offs> offs(count);
    pl.probs.resize(count);
    pl.count = count;
    const size_t off_flags = (i8 || res) ? take((size_t)count * 4) : 0;
    for (int i = 0; i < count; ++i) {
        const mdg_problem &q = pr[i];
        BProb &b = pl.probs[i];
        std::memset(&b, 0, sizeof(b));
        b.L = q.L;
        b.m = q.m;
        b.n = q.n;
        b.k = q.k;
        b.lda = q.lda;
        b.ldb = q.ldb;
        b.ldc = q.ldc;
        b.A = q.A;
        b.B = q.B;
        b.C = q.C;
        if (q.m == 0 || q.n == 0 || q.k == 0) continue;
        if (res) {
            offs[i].e = take((size_t)b.m * 4);
            offs[i].f = take((size_t)b.n * 4);
            offs[i].cs = take((size_t)b.n * 8);
            b.S = -1;   // marks a problem with scale exponents
            continue;
        }
        if (!i8) continue;
        b.w = 8;
        b.S = plan_int8_slices(q.L, target_bits > 0 ? target_bits : 53 * q.L, spread, spread);
        MDG_REQUIRE(b.S > 0 && b.S * 8 <= 32 * (q.L == 2 ? 8 : q.L == 4 ? 12 : q.L == 8 ? 20 : 36) - 1,
                    MDG_ERROR_RANGE, "problem %d: slice plan does not fit", i);
        b.kp = (int)rup(q.k, 128);
        b.mp = (int)rup(q.m, 128);
        b.np = (int)rup(q.n, 128);
        // the band kernel cuts a diagonal only at slice boundaries: one slice
        // pair must fit an INT32 plane (the per-problem loop handles deeper k)
        MDG_REQUIRE(b.kp <= 131071 && b.S < 256, MDG_ERROR_NOT_SUPPORTED, "problem %d too deep for the fused stages", i);
        b.kmax = (131071 / b.kp) * b.kp;
        b.nplanes = bnd_band_planes(b.S, 4, b.kp / 128, BND_CAP_I8);
        offs[i].aall = take((size_t)b.mp * b.S * b.kp);
        offs[i].ball = take((size_t)b.np * b.S * b.kp);
        offs[i].d = take((size_t)b.nplanes * b.m * b.n * 8);
        offs[i].e = take((size_t)b.m * 4);
        offs[i].f = take((size_t)b.n * 4);
        offs[i].cs = take((size_t)b.n * 8);
    }
    if (bytes) {
        MDG_HIP_CHECK(hipMalloc(&pl.arena, bytes));
        pl.bytes += bytes;
        char *base = (char *)pl.arena;
        // zero digits in the slices' padding for every replay (the slicing
        // stage writes only the valid rows and depth)
        if (i8) MDG_HIP_CHECK(hipMemset(pl.arena, 0, bytes));
        if (i8 || res) {
            pl.flags = (int *)(base + off_flags);
            // problems without entries (empty) keep flag 0 on every replay
            MDG_HIP_CHECK(hipMemset(pl.flags, 0, (size_t)count * 4));
            for (int i = 0; i < count; ++i) {
                pl.probs[i].flag = pl.flags + i;
                pl.probs[i].allow = spread;
            }
        }
        for (int i = 0; i < count; ++i) {
            BProb &b = pl.probs[i];
            if (!b.S) continue;
            if (b.S > 0) {
                b.Aall = (int8_t *)(base + offs[i].aall);
                b.Ball = (int8_t *)(base + offs[i].ball);
                b.D = (long long *)(base + offs[i].d);
            }
            b.E = (int *)(base + offs[i].e);
            b.F = (int *)(base + offs[i].f);
            b.colsum = (double *)(base + offs[i].cs);
        }
    }
    // entry tables, one stage per (kind, L)
    auto add_stage = [&](int kind, int L, auto gen) {
        Stage st{kind, L, (long long)pl.ents.size(), 0};
        for (int i = 0; i < count; ++i) {
            const BProb &b = pl.probs[i];
            if (L && b.L != L) continue;
            if (b.m == 0 || b.n == 0 || b.k == 0) continue;
            if ((i8 || res) && !b.S) continue;
            gen(i, b);
        }
        st.count = (long long)pl.ents.size() - st.first;
        if (st.count) pl.stages.push_back(st);
    };
    auto push = [&](int i, int x, int y, int z) { pl.ents.push_back(BEnt{i, x, y, z}); };
    const int Ls[4] = {2, 4, 8, 16};
    if (i8 || res) {
        add_stage(0, 0, [&](int i, const BProb &b) {
            for (int x = 0; x < (b.n + 255) / 256; ++x) push(i, x, 0, 0);
        });
        for (int L : Ls)
            add_stage(1, L, [&](int i, const BProb &b) {
                for (int x = 0; x < (b.m + 3) / 4; ++x) push(i, x, 0, 0);
            });
        for (int L : Ls)
            add_stage(2, L, [&](int i, const BProb &b) {
                for (int x = 0; x < (b.n + 63) / 64; ++x)
                    for (int y = 0; y < (b.k + 255) / 256; ++y) push(i, x, y, 0);
            });
        add_stage(3, 0, [&](int i, const BProb &b) {
            for (int x = 0; x < (b.n + 255) / 256; ++x) push(i, x, 0, 0);
        });
    }
    if (res) {
        std::vector<int *> Es(count, nullptr), Fs(count, nullptr);
        for (int i = 0; i < count; ++i) {
            Es[i] = pl.probs[i].E;
            Fs[i] = pl.probs[i].F;
        }
        const int rc = residue_batch_build(count, pr, Es.data(), Fs.data(), target_bits, spread, &pl.rb);
        if (rc) return rc;
    }
    if (i8) {
        for (int kind : {4, 5})
            for (int L : Ls) {
                int lds = 0;
                const size_t before = pl.stages.size();
                add_stage(kind, L, [&](int i, const BProb &b) {
                    lds = std::max(lds, blk_stage_bytes(b.S));
                    const int rt = blk_rows_per_tile(b.S);
                    for (int x = 0; x < ((kind == 4 ? b.m : b.n) + rt - 1) / rt; ++x)
                        for (int y = 0; y < b.kp / 128; ++y) push(i, x, y, 0);
                });
                if (pl.stages.size() > before) pl.stages.back().lds = lds;
            }
        // GEMM: ONE band-kernel launch over the band items of every problem
        // (band_items: all 4 diagonals wide).  Workgroups go to the 8 XCDs
        // round-robin by their index, so entry 8 j + x runs on XCD x: the
        // problems, heaviest first, are dealt to the XCDs in snake order
        // (0..7, 7..0, ...), and each XCD's list holds its problems in that
        // order, a problem's items heaviest first with each item's tiles next
        // to each other -- the workgroups that read one problem's slices run
        // on one XCD, close in time, and share its L2.  The lists advance at
        // similar rates (problems of similar rank on each XCD); the shorter
        // ones end with empty entries (prob = -1).
        {
            constexpr int NXCD = 8;
            struct PW {
                int prob;
                long long work;
            };
            std::vector<PW> pws;
            std::vector<std::vector<BandWork>> its(count);
            for (int i = 0; i < count; ++i) {
                const BProb &b = pl.probs[i];
                if (b.S <= 0 || b.m == 0 || b.n == 0 || b.k == 0) continue;
                band_items(b.S, 4, b.kp / 128, BND_CAP_I8, its[i]);
                std::stable_sort(its[i].begin(), its[i].end(),
                                 [](const BandWork &x, const BandWork &y) { return x.work > y.work; });
                const long long per_item = (long long)(b.mp / 128) * (b.np / 128);   // item work counts its k-blocks
                long long w = 0;
                for (const BandWork &bw : its[i]) w += bw.work * per_item;
                pws.push_back(PW{i, w});
            }
            std::stable_sort(pws.begin(), pws.end(), [](const PW &x, const PW &y) { return x.work > y.work; });
            std::vector<std::vector<BEnt>> lists(NXCD);
            for (size_t r = 0; r < pws.size(); ++r) {
                const int round = (int)(r / NXCD), pos = (int)(r % NXCD);
                const int x = (round & 1) ? NXCD - 1 - pos : pos;
                const int i = pws[r].prob;
                const BProb &b = pl.probs[i];
                for (const BandWork &bw : its[i]) {
                    const int code = (bw.it.g0 & 0xff) | (bw.it.kb0 << 8) | (bw.it.kb1 << 19);
                    for (int tx = 0; tx < b.mp / 128; ++tx)
                        for (int ty = 0; ty < b.np / 128; ++ty) lists[x].push_back(BEnt{i, tx, ty, code});
                }
            }
            size_t len = 0;
            for (const auto &l : lists) len = std::max(len, l.size());
            if (len) {
                Stage st{6, 0, (long long)pl.ents.size(), 0};
                st.g = 4;
                for (size_t j = 0; j < len; ++j)
                    for (int x = 0; x < NXCD; ++x) pl.ents.push_back(j < lists[x].size() ? lists[x][j] : BEnt{-1, 0, 0, 0});
                st.count = (long long)pl.ents.size() - st.first;
                pl.stages.push_back(st);
            }
        }
        for (int L : Ls) {
            const size_t before = pl.stages.size();
            add_stage(7, L, [&](int i, const BProb &b) {
                const long long nb = ((long long)b.m * b.n + 255) / 256;
                for (long long x = 0; x < nb; ++x) push(i, (int)x, 0, 0);
            });
            if (pl.stages.size() > before)
                for (const BProb &b : pl.probs)
                    if (b.L == L && b.S) {
                        pl.stages.back().nw = std::max(pl.stages.back().nw, b.S);   // the largest slice count
                        pl.stages.back().chunked = pl.stages.back().chunked || b.nplanes > (b.S + 3) / 4;
                    }
        }
    } else if (!res) {
        for (int L : Ls)
            add_stage(8, L, [&](int i, const BProb &b) {
                for (int x = 0; x < (b.m + 15) / 16; ++x)
                    for (int y = 0; y < (b.n + 15) / 16; ++y) push(i, x, y, 0);
            });
    }
    if (i8 || res) {
        pl.guard.first = (long long)pl.ents.size();
        for (int i = 0; i < count; ++i) {
            const BProb &b = pl.probs[i];
            if (b.m == 0 || b.n == 0 || b.k == 0) continue;
            push(i, 0, 0, 0);
        }
        pl.guard.count = (long long)pl.ents.size() - pl.guard.first;
    }
    for (const Stage &st : pl.stages)
        MDG_REQUIRE(st.count <= 2147483647ll, MDG_ERROR_NOT_SUPPORTED, "batch too large for one launch");
    MDG_HIP_CHECK(hipMalloc(&pl.dprobs, sizeof(BProb) * std::max(1, count)));
    MDG_HIP_CHECK(hipMemcpy(pl.dprobs, pl.probs.data(), sizeof(BProb) * count, hipMemcpyHostToDevice));
    MDG_HIP_CHECK(hipMalloc(&pl.dents, sizeof(BEnt) * std::max<size_t>(1, pl.ents.size())));
    pl.bytes += sizeof(BProb) * std::max(1, count) + sizeof(BEnt) * std::max<size_t>(1, pl.ents.size());
    MDG_HIP_CHECK(
        hipMemcpy(pl.dents, pl.ents.data(), sizeof(BEnt) * pl.ents.size(), hipMemcpyHostToDevice));
    return MDG_SUCCESS;
}

static int launch_stage(const Plan &pl, const Stage &st, hipStream_t s)
{
    const dim3 grid((unsigned)st.count), blk(256);
    const BEnt *T = pl.dents + st.first;
    const BProb *P = pl.dprobs;
#define MDG_BY_L(kern)                                                                      \
    switch (st.L) {                                                                         \
    case 2: hipLaunchKernelGGL(kern<2>, grid, blk, 0, s, P, T); break;                      \
    case 4: hipLaunchKernelGGL(kern<4>, grid, blk, 0, s, P, T); break;                      \
    case 8: hipLaunchKernelGGL(kern<8>, grid, blk, 0, s, P, T); break;                      \
    default: hipLaunchKernelGGL(kern<16>, grid, blk, 0, s, P, T); break;                    \
    }
    switch (st.kind) {
    case 0: hipLaunchKernelGGL(b_init_kernel, grid, blk, 0, s, P, T); break;
    case 1: MDG_BY_L(b_scan_rows_kernel); break;
    case 2: MDG_BY_L(b_scan_cols_kernel); break;
    case 3: hipLaunchKernelGGL(b_finish_cols_kernel, grid, blk, 0, s, P, T); break;
    case 9: hipLaunchKernelGGL(b_guard_kernel, grid, blk, 0, s, P, T); break;
    case 4:
        switch (st.L) {
        case 2: hipLaunchKernelGGL((b_slice_kernel<2, false>), grid, blk, st.lds, s, P, T); break;
        case 4: hipLaunchKernelGGL((b_slice_kernel<4, false>), grid, blk, st.lds, s, P, T); break;
        case 8: hipLaunchKernelGGL((b_slice_kernel<8, false>), grid, blk, st.lds, s, P, T); break;
        default: hipLaunchKernelGGL((b_slice_kernel<16, false>), grid, blk, st.lds, s, P, T); break;
        }
        break;
    case 5:
        switch (st.L) {
        case 2: hipLaunchKernelGGL((b_slice_kernel<2, true>), grid, blk, st.lds, s, P, T); break;
        case 4: hipLaunchKernelGGL((b_slice_kernel<4, true>), grid, blk, st.lds, s, P, T); break;
        case 8: hipLaunchKernelGGL((b_slice_kernel<8, true>), grid, blk, st.lds, s, P, T); break;
        default: hipLaunchKernelGGL((b_slice_kernel<16, true>), grid, blk, st.lds, s, P, T); break;
        }
        break;
    case 6: hipLaunchKernelGGL(b_band_i8_kernel<4>, grid, blk, 0, s, P, T); break;
    case 7: {
        auto rec = [&](auto lc) -> int {
            constexpr int L = decltype(lc)::value;
            return with_band_words<L, 4>(st.nw, [&](auto nw) {
                constexpr int NWB = decltype(nw)::value;
                if (st.chunked)
                    hipLaunchKernelGGL((b_reconstruct_kernel<L, NWB, true>), grid, blk, 0, s, P, T);
                else
                    hipLaunchKernelGGL((b_reconstruct_kernel<L, NWB, false>), grid, blk, 0, s, P, T);
            });
        };
        int rc = (st.L == 2)   ? rec(std::integral_constant<int, 2>())
                 : (st.L == 4) ? rec(std::integral_constant<int, 4>())
                 : (st.L == 8) ? rec(std::integral_constant<int, 8>())
                               : rec(std::integral_constant<int, 16>());
        if (rc) return rc;
        break;
    }
    default: MDG_BY_L(b_vector_kernel); break;
    }
#undef MDG_BY_L
    MDG_LAUNCH_CHECK("batched stage");
    return MDG_SUCCESS;
}

static int run_plan(const Plan &pl, hipStream_t s)
{
    bool in_gemm = false;
    for (const Stage &st : pl.stages) {
        if (st.kind == 6 && !in_gemm && pl.gev[0]) MDG_HIP_CHECK(hipEventRecord(pl.gev[0], s));
        if (st.kind != 6 && in_gemm && pl.gev[1]) MDG_HIP_CHECK(hipEventRecord(pl.gev[1], s));
        in_gemm = st.kind == 6;
        int rc = launch_stage(pl, st, s);
        if (rc) return rc;
    }
    if (in_gemm && pl.gev[1]) MDG_HIP_CHECK(hipEventRecord(pl.gev[1], s));
    if (pl.rb) {
        int rc = residue_batch_run(pl.rb, s, pl.gev[0], pl.gev[1]);
        if (rc) return rc;
    }
    return pl.guard.count ? launch_stage(pl, pl.guard, s) : MDG_SUCCESS;
}

// Autotune the residue batch's GEMM tiles: run the whole pipeline once with
// every candidate (128- or 64-row tiles) and keep the fastest; all produce
// identical results.  (The slice pipeline's band stage has one schedule.)
static int tune_plan(Plan &pl, hipStream_t s)
{
    const int nc = pl.rb ? residue_batch_candidates(pl.rb) : 1;
    if (nc <= 1) return run_plan(pl, s);
    hipEvent_t e0, e1;
    MDG_HIP_CHECK(hipEventCreate(&e0));
    MDG_HIP_CHECK(hipEventCreate(&e1));
    float best = 1e30f;
    int pick = 0, rc = MDG_SUCCESS;
    for (int c = 0; c < nc && rc == MDG_SUCCESS; ++c) {
        residue_batch_select(pl.rb, c);
        float ms = 0.f;
        if (hipEventRecord(e0, s) != hipSuccess || (rc = run_plan(pl, s)) || hipEventRecord(e1, s) != hipSuccess ||
            hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
            if (!rc) {
                set_error("timing the batch schedules failed");
                rc = MDG_ERROR_HIP;
            }
            break;
        }
        if (ms < best) {
            best = ms;
            pick = c;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    residue_batch_select(pl.rb, pick);
    return MDG_SUCCESS;
}

__global__ void b_zero_flags_kernel(int *flags, int count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) flags[i] = 0;
}

// The per-problem calls.  flags: each slice / residue problem's guard
// flags (split_gemm.hip).  record: the largest scratch need of one problem
// is written there (an uncaptured sizing pass).  work: a workspace of
// work_bytes that every problem's scratch requests are carved from, from its
// start for each problem (the calls run in stream order): a plan recorded
// into a graph then refers to memory it owns, never to per-stream scratch
// that a later call could grow or release.
static int run_loop(int count, const mdg_problem *pr, int method, int target_bits, int spread, hipStream_t s,
                    int *flags, size_t *record, void *work, size_t work_bytes)
{
    if (flags) {
        hipLaunchKernelGGL(b_zero_flags_kernel, dim3((count + 255) / 256), dim3(256), 0, s, flags, count);
        MDG_LAUNCH_CHECK("b_zero_flags_kernel");
    }
    Arena arena;
    if (work) arena_begin(work, work_bytes, &arena);
    int rc = MDG_SUCCESS;
    size_t need = 0;
    for (int i = 0; i < count && rc == MDG_SUCCESS; ++i) {
        const mdg_problem &p = pr[i];
        if (p.m == 0 || p.n == 0 || p.k == 0) continue;
        size_t cur = 0;
        if (record) scratch_record(&cur);
        if (work) arena_reset();
        set_guard_flag(flags ? flags + i : nullptr);
        rc = (method == MDG_METHOD_VECTOR)
                 ? vector_gemm(p.L, p.m, p.n, p.k, p.A, p.lda, p.B, p.ldb, p.C, p.ldc, MDG_TWOPROD_FMA, s)
                 : split_gemm(p.L, p.m, p.n, p.k, p.A, p.lda, p.B, p.ldb, p.C, p.ldc, method, target_bits, spread,
                              nullptr, s);
        set_guard_flag(nullptr);
        if (record) scratch_record(nullptr);
        need = std::max(need, cur);
    }
    size_t over = 0;
    if (work && arena_overflow(&over) && rc == MDG_SUCCESS) {
        set_error("internal: a batch problem needs %zu bytes of workspace, more than its plan's %zu", over,
                  work_bytes);
        rc = MDG_ERROR_OUT_OF_MEMORY;
    }
    if (work) arena_end();
    if (record) *record = need;
    return rc;
}

// ------------------------------------------------------------ plan cache --
// Plans are cached by batch signature (device, stream, method, precision,
// every problem's shape and pointers): a replay is tied to the buffers it
// was recorded with.  The cache is bounded (least recently used plans are
// freed first) and mdg_release_workspace empties it for the device.
static std::mutex g_plan_mu;
static std::map<std::string, Plan *> g_plans;
static size_t g_plan_max_bytes = size_t(16) << 30;
static int g_plan_max_count = 32;
static unsigned long long g_plan_clock = 0;

// per (device, stream): the flags of the last batched call there
struct LastFlags {
    int *flags = nullptr;
    int count = 0;
    const Plan *plan = nullptr;   // persistent calls: the plan replayed
};
static std::map<std::pair<int, void *>, LastFlags> g_last;
// per (device, stream): the flag buffer of non-persistent calls
struct LoopFlags {
    int *buf = nullptr;
    int cap = 0;
};
static std::map<std::pair<int, void *>, LoopFlags> g_loop_flags;

static void forget_flags_locked(const int *flags, const Plan *plan = nullptr)
{
    for (auto it = g_last.begin(); it != g_last.end();)
        it = (it->second.flags == flags || (plan && it->second.plan == plan)) ? g_last.erase(it) : std::next(it);
}

static void free_plan_locked(std::map<std::string, Plan *>::iterator it)
{
    Plan *pl = it->second;
    (void)hipStreamSynchronize(pl->stream);   // a replay may still run
    forget_flags_locked(pl->flags, pl);
    delete pl;
    g_plans.erase(it);
}

// evict least recently used plans of any device until the cache fits
static void trim_plans_locked(const Plan *keep)
{
    for (;;) {
        size_t bytes = 0;
        for (auto &kv : g_plans) bytes += kv.second->bytes;
        if ((int)g_plans.size() <= g_plan_max_count && bytes <= g_plan_max_bytes) return;
        auto victim = g_plans.end();
        for (auto it = g_plans.begin(); it != g_plans.end(); ++it)
            if (it->second != keep && (victim == g_plans.end() || it->second->last_use < victim->second->last_use))
                victim = it;
        if (victim == g_plans.end()) return;   // only the plan just built is left
        int cur = 0;
        (void)hipGetDevice(&cur);
        if (victim->second->dev != cur) (void)hipSetDevice(victim->second->dev);
        free_plan_locked(victim);
        (void)hipSetDevice(cur);
    }
}

int release_batch_plans()
{
    int dev = 0;
    MDG_HIP_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_plan_mu);
    for (auto it = g_plans.begin(); it != g_plans.end();) {
        auto next = std::next(it);
        if (it->second->dev == dev) free_plan_locked(it);
        it = next;
    }
    for (auto it = g_loop_flags.begin(); it != g_loop_flags.end();) {
        if (it->first.first == dev) {
            (void)hipStreamSynchronize((hipStream_t)it->first.second);
            forget_flags_locked(it->second.buf);
            (void)hipFree(it->second.buf);
            it = g_loop_flags.erase(it);
        } else {
            ++it;
        }
    }
    return MDG_SUCCESS;
}

int set_batch_plan_cache(int max_plans, long long max_bytes)
{
    MDG_REQUIRE(max_plans >= 1 && max_bytes >= 0, MDG_ERROR_INVALID_VALUE, "bad plan cache bounds (%d, %lld)",
                max_plans, max_bytes);
    std::lock_guard<std::mutex> lock(g_plan_mu);
    g_plan_max_count = max_plans;
    g_plan_max_bytes = (size_t)max_bytes;
    trim_plans_locked(nullptr);
    return MDG_SUCCESS;
}

int batched_last_flags(void *stream, int count, int *out)
{
    int dev = 0;
    MDG_HIP_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_plan_mu);
    auto it = g_last.find({dev, stream});
    MDG_REQUIRE(it != g_last.end(), MDG_ERROR_INVALID_VALUE, "no batched call on this stream (or its plan was freed)");
    MDG_REQUIRE(count == it->second.count, MDG_ERROR_INVALID_VALUE,
                "the last batched call on this stream had %d problems, not %d", it->second.count, count);
    if (count == 0) return MDG_SUCCESS;
    if (!it->second.flags) {   // a vector-method batch: nothing to flag
        for (int i = 0; i < count; ++i) out[i] = 0;
        return MDG_SUCCESS;
    }
    MDG_HIP_CHECK(hipMemcpyAsync(out, it->second.flags, sizeof(int) * (size_t)count, hipMemcpyDeviceToHost,
                                 (hipStream_t)stream));
    MDG_HIP_CHECK(hipStreamSynchronize((hipStream_t)stream));
    for (int i = 0; i < count; ++i) {
        if (out[i] & 1) {
            set_error("problem %d: an input limb is Inf or NaN (its C was filled with NaN)", i);
            return MDG_ERROR_NON_FINITE;
        }
        if (out[i] & 2) {
            set_error("problem %d: exponent spread beyond the batch's spread allowance (its C was filled with NaN)",
                      i);
            return MDG_ERROR_RANGE;
        }
    }
    return MDG_SUCCESS;
}

int batched_last_gemm_ms(void *stream, float *ms)
{
    int dev = 0;
    MDG_HIP_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_plan_mu);
    auto it = g_last.find({dev, stream});
    MDG_REQUIRE(it != g_last.end() && it->second.plan && it->second.plan->gev[0], MDG_ERROR_NOT_SUPPORTED,
                "the last batched call on this stream was no fused persistent batch built with profiling on");
    const Plan *pl = it->second.plan;
    MDG_HIP_CHECK(hipEventSynchronize(pl->gev[1]));
    MDG_HIP_CHECK(hipEventElapsedTime(ms, pl->gev[0], pl->gev[1]));
    return MDG_SUCCESS;
}

static std::string signature(int count, const mdg_problem *pr, int method, int target_bits, int spread, hipStream_t s,
                             int dev)
{
    std::string key;
    key.reserve(64 + sizeof(mdg_problem) * (size_t)count);
    key.append((const char *)&dev, sizeof(dev));
    key.append((const char *)&method, sizeof(method));
    key.append((const char *)&target_bits, sizeof(target_bits));
    key.append((const char *)&spread, sizeof(spread));
    key.append((const char *)&s, sizeof(s));
    key.append((const char *)pr, sizeof(mdg_problem) * (size_t)count);
    return key;
}

int batched_gemm(int count, const mdg_problem *pr, int method, int target_bits, int spread_bits, int persistent,
                 hipStream_t s)
{
    const int spread = spread_bits < 0 ? 16 : spread_bits;
    int dev = 0;
    MDG_HIP_CHECK(hipGetDevice(&dev));
    if (!persistent) {
        // flags of the per-problem guards in a small buffer of this stream
        int *flags = nullptr;
        {
            std::lock_guard<std::mutex> lock(g_plan_mu);
            LoopFlags &lf = g_loop_flags[{dev, (void *)s}];
            if (lf.cap < count) {
                if (lf.buf) {
                    MDG_HIP_CHECK(hipStreamSynchronize(s));
                    forget_flags_locked(lf.buf);
                    (void)hipFree(lf.buf);
                    lf.buf = nullptr;
                    lf.cap = 0;
                }
                MDG_HIP_CHECK(hipMalloc(&lf.buf, sizeof(int) * (size_t)std::max(count, 64)));
                lf.cap = std::max(count, 64);
            }
            flags = lf.buf;
            g_last[{dev, (void *)s}] = LastFlags{flags, count};
        }
        return run_loop(count, pr, method, target_bits, spread, s, flags, nullptr, nullptr, 0);
    }
    for (int i = 0; i < count; ++i)
        MDG_REQUIRE(pr[i].k > 0 || pr[i].m == 0 || pr[i].n == 0, MDG_ERROR_NOT_SUPPORTED,
                    "persistent batches need k > 0 (problem %d)", i);
    MDG_REQUIRE(s != nullptr, MDG_ERROR_NOT_SUPPORTED, "persistent batches need a non-null stream");

    const std::string key = signature(count, pr, method, target_bits, spread, s, dev);
    std::lock_guard<std::mutex> lock(g_plan_mu);
    auto it = g_plans.find(key);
    if (it != g_plans.end()) {
        it->second->last_use = ++g_plan_clock;
        MDG_HIP_CHECK(hipGraphLaunch(it->second->exec, s));
        g_last[{dev, (void *)s}] = LastFlags{it->second->flags, count, it->second};
        return MDG_SUCCESS;
    }
    Plan *pl = new Plan();
    pl->dev = dev;
    pl->stream = s;
    pl->count = count;
    bool fused =
        (method == MDG_METHOD_INT8_OZAKI || method == MDG_METHOD_VECTOR || method == MDG_METHOD_INT8_RESIDUE);
    int rc = fused ? build_plan(*pl, count, pr, method, target_bits, spread) : MDG_SUCCESS;
    if (rc == MDG_ERROR_NOT_SUPPORTED && (method == MDG_METHOD_INT8_RESIDUE || method == MDG_METHOD_INT8_OZAKI)) {
        // a problem too deep for the fused residue stages: the per-problem
        // calls inside the graph instead
        delete pl;
        pl = new Plan();
        pl->dev = dev;
        pl->stream = s;
        pl->count = count;
        fused = false;
        rc = MDG_SUCCESS;
    }
    if (rc) {
        delete pl;
        return rc;
    }
    if (fused) {
        if (profiling_on() && method != MDG_METHOD_VECTOR) {
            // HIP events around the GEMM stage, recorded into the graph
            for (auto &ev : pl->gev)
                if (hipEventCreate(&ev) != hipSuccess) ev = nullptr;
            if (!pl->gev[0] || !pl->gev[1]) {
                for (auto &ev : pl->gev)
                    if (ev) (void)hipEventDestroy(ev);
                pl->gev[0] = pl->gev[1] = nullptr;
            }
        }
        // an uncaptured pass sizes every library scratch buffer (capture must
        // not allocate) and chooses the GEMM tile order
        rc = tune_plan(*pl, s);
    } else {
        // the per-problem calls: an uncaptured pass (computing this call's
        // results) records the largest scratch need of one problem; the plan
        // gets a workspace of that size, which the captured calls use
        if (hipMalloc(&pl->flags, sizeof(int) * (size_t)std::max(count, 1)) != hipSuccess) {
            delete pl;
            set_error("flag buffer allocation failed");
            return MDG_ERROR_OUT_OF_MEMORY;
        }
        size_t need = 0;
        rc = run_loop(count, pr, method, target_bits, spread, s, pl->flags, &need, nullptr, 0);
        if (!rc) {
            pl->work_bytes = std::max<size_t>(need, 256);
            if (hipMalloc(&pl->work, pl->work_bytes) != hipSuccess) {
                (void)hipGetLastError();
                set_error("workspace allocation of %zu bytes for the batch plan failed", pl->work_bytes);
                rc = MDG_ERROR_OUT_OF_MEMORY;
            }
            pl->bytes = pl->work_bytes + sizeof(int) * (size_t)count;
        }
    }
    if (rc) {
        delete pl;
        return rc;
    }
    hipError_t e = hipStreamSynchronize(s);
    hipGraph_t graph = nullptr;
    if (e == hipSuccess) e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        delete pl;
        MDG_HIP_CHECK(e);
    }
    rc = fused ? run_plan(*pl, s)
               : run_loop(count, pr, method, target_bits, spread, s, pl->flags, nullptr, pl->work, pl->work_bytes);
    e = hipStreamEndCapture(s, &graph);
    if (rc || e != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        delete pl;
        if (rc) return rc;
        MDG_HIP_CHECK(e);
    }
    e = hipGraphInstantiate(&pl->exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
        delete pl;
        MDG_HIP_CHECK(e);
    }
    pl->last_use = ++g_plan_clock;
    g_plans[key] = pl;
    g_last[{dev, (void *)s}] = LastFlags{pl->flags, count, pl};
    trim_plans_locked(pl);
    return MDG_SUCCESS;   // the uncaptured pass already computed this call
}

}  // namespace mdg
