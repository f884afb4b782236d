// elementwise.hip -- the reference's scalar multi-double operations as
// elementwise device kernels over arrays of interleaved p-doubles:
// add, sub, mul, div, sqr, sqrt, mul by a double, inc by a double,
// power-of-two scaling, abs and negation (C++/{double,quad,octo,hexa}
// _double_functions.cu; the operation order of each is in md_arith.hpp).
//
// One thread per entry; each thread reads its entry before writing it, so
// the result may overwrite an operand.  Every operand pointer that is 16-byte aligned (and
// hipMalloc'd buffers are) is read and written in 16-byte pieces, so a
// wavefront touches one contiguous span per limb pair; other pointers fall
// back to 8-byte accesses.  add/sub/inc/scaling at L <= 4 are HBM-bound
// ((2L + L) * 8 bytes per entry for the binary ops); mul, div and sqrt are
// FP64-VALU-bound from L = 4 on.
#include <cstdint>

#include "common.hpp"
#include "md_arith.hpp"

namespace mdg {

namespace {

template <int L, bool V>
__device__ __forceinline__ void load_entry(const double *p, long long i, double (&x)[L])
{
    if constexpr (V) {
        const double2 *q = reinterpret_cast<const double2 *>(p + i * L);
#pragma unroll
        for (int l = 0; l < L / 2; ++l) {
            const double2 v = q[l];
            x[2 * l] = v.x;
            x[2 * l + 1] = v.y;
        }
    } else {
#pragma unroll
        for (int l = 0; l < L; ++l) x[l] = p[i * L + l];
    }
}

template <int L, bool V>
__device__ __forceinline__ void store_entry(double *p, long long i, const double (&x)[L])
{
    if constexpr (V) {
        double2 *q = reinterpret_cast<double2 *>(p + i * L);
#pragma unroll
        for (int l = 0; l < L / 2; ++l) q[l] = make_double2(x[2 * l], x[2 * l + 1]);
    } else {
#pragma unroll
        for (int l = 0; l < L; ++l) p[i * L + l] = x[l];
    }
}

constexpr bool is_binary(int op) { return op == MDG_OP_ADD || op == MDG_OP_SUB || op == MDG_OP_MUL || op == MDG_OP_DIV; }
constexpr bool takes_double(int op) { return op == MDG_OP_MUL_D || op == MDG_OP_INC_D || op == MDG_OP_MUL_PWR2; }

template <int L, int OP, int TP, bool V>
__global__ __launch_bounds__(256) void elementwise_kernel(long long n, const double *a, const double *b, double *c)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double x[L], y[L], z[L];
    load_entry<L, V>(a, i, x);
    if constexpr (is_binary(OP)) load_entry<L, V>(b, i, y);
    if constexpr (OP == MDG_OP_ADD) {
        md_add<L>(x, y, z);
    } else if constexpr (OP == MDG_OP_SUB) {
        md_sub<L>(x, y, z);
    } else if constexpr (OP == MDG_OP_MUL) {
        md_mul<L, TP>(x, y, z);
    } else if constexpr (OP == MDG_OP_DIV) {
        md_div<L, TP>(x, y, z);
    } else if constexpr (OP == MDG_OP_SQR) {
        md_sqr<L, TP>(x, z);
    } else if constexpr (OP == MDG_OP_SQRT) {
        md_sqrt<L, TP>(x, z);
    } else if constexpr (OP == MDG_OP_MUL_D) {
        md_mul_d<L, TP>(x, b[i], z);
    } else if constexpr (OP == MDG_OP_INC_D) {
#pragma unroll
        for (int l = 0; l < L; ++l) z[l] = x[l];
        md_inc_d<L>(z, b[i]);
    } else if constexpr (OP == MDG_OP_MUL_PWR2) {
        const double d = b[i];
#pragma unroll
        for (int l = 0; l < L; ++l) z[l] = x[l] * d;
    } else if constexpr (OP == MDG_OP_ABS) {
        // the sign of the leading limb decides for every limb (quad :194-212;
        // hexa :1455-1497 copies limb 5 into limb 9, not reproduced)
        const bool neg = x[0] < 0.0;
#pragma unroll
        for (int l = 0; l < L; ++l) z[l] = neg ? -x[l] : x[l];
    } else {
#pragma unroll
        for (int l = 0; l < L; ++l) z[l] = -x[l];
    }
    store_entry<L, V>(c, i, z);
}

constexpr bool uses_twoprod(int op)
{
    return op == MDG_OP_MUL || op == MDG_OP_DIV || op == MDG_OP_SQR || op == MDG_OP_SQRT || op == MDG_OP_MUL_D;
}

template <int L, int OP>
int launch_op(long long n, const double *a, const double *b, double *c, int tp, bool vec, hipStream_t s)
{
    const dim3 grid((unsigned)((n + 255) / 256)), block(256);
    // ops without a product ignore the two_prod flavour: one instantiation
    const int t = uses_twoprod(OP) ? tp : (int)TP_FMA;
    if (t == TP_FMA) {
        if (vec)
            hipLaunchKernelGGL((elementwise_kernel<L, OP, TP_FMA, true>), grid, block, 0, s, n, a, b, c);
        else
            hipLaunchKernelGGL((elementwise_kernel<L, OP, TP_FMA, false>), grid, block, 0, s, n, a, b, c);
    } else {
        if constexpr (uses_twoprod(OP)) {
            if (vec)
                hipLaunchKernelGGL((elementwise_kernel<L, OP, TP_DEKKER, true>), grid, block, 0, s, n, a, b, c);
            else
                hipLaunchKernelGGL((elementwise_kernel<L, OP, TP_DEKKER, false>), grid, block, 0, s, n, a, b, c);
        }
    }
    MDG_LAUNCH_CHECK("elementwise_kernel");
    return MDG_SUCCESS;
}

template <int L>
int launch_l(int op, long long n, const double *a, const double *b, double *c, int tp, bool vec, hipStream_t s)
{
    switch (op) {
    case MDG_OP_ADD: return launch_op<L, MDG_OP_ADD>(n, a, b, c, tp, vec, s);
    case MDG_OP_SUB: return launch_op<L, MDG_OP_SUB>(n, a, b, c, tp, vec, s);
    case MDG_OP_MUL: return launch_op<L, MDG_OP_MUL>(n, a, b, c, tp, vec, s);
    case MDG_OP_DIV: return launch_op<L, MDG_OP_DIV>(n, a, b, c, tp, vec, s);
    case MDG_OP_SQR: return launch_op<L, MDG_OP_SQR>(n, a, b, c, tp, vec, s);
    case MDG_OP_SQRT: return launch_op<L, MDG_OP_SQRT>(n, a, b, c, tp, vec, s);
    case MDG_OP_MUL_D: return launch_op<L, MDG_OP_MUL_D>(n, a, b, c, tp, vec, s);
    case MDG_OP_INC_D: return launch_op<L, MDG_OP_INC_D>(n, a, b, c, tp, vec, s);
    case MDG_OP_MUL_PWR2: return launch_op<L, MDG_OP_MUL_PWR2>(n, a, b, c, tp, vec, s);
    case MDG_OP_ABS: return launch_op<L, MDG_OP_ABS>(n, a, b, c, tp, vec, s);
    case MDG_OP_NEG: return launch_op<L, MDG_OP_NEG>(n, a, b, c, tp, vec, s);
    }
    set_error("unsupported elementwise op %d", op);
    return MDG_ERROR_INVALID_VALUE;
}

bool aligned16(const void *p) { return p == nullptr || ((uintptr_t)p & 15u) == 0; }

}  // namespace

int elementwise(int op, int L, long long n, const double *a, const double *b, double *c, int tp, hipStream_t s)
{
    // the double operand of the _D/_PWR2 ops is read 8 bytes per entry
    const bool vec = aligned16(a) && aligned16(c) && (!is_binary(op) || aligned16(b));
    switch (L) {
    case 2: return launch_l<2>(op, n, a, b, c, tp, vec, s);
    case 4: return launch_l<4>(op, n, a, b, c, tp, vec, s);
    case 8: return launch_l<8>(op, n, a, b, c, tp, vec, s);
    case 16: return launch_l<16>(op, n, a, b, c, tp, vec, s);
    }
    set_error("unsupported limb count L=%d", L);
    return MDG_ERROR_INVALID_VALUE;
}

bool elementwise_binary(int op) { return is_binary(op); }
bool elementwise_takes_double(int op) { return takes_double(op); }

}  // namespace mdg
