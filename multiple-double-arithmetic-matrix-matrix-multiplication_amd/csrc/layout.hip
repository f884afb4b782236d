// layout.hip -- layout utilities for interleaved p-double matrices: limb-
// planar conversion (the reference's part(), C++/TCcuda.cu:34-40, for every
// limb at once) and transposition (the reference is square and row-major
// only; the GEMMs here take leading dimensions but no transposes, so a
// transposed operand is materialised once with mdg_transpose).
//
// All three are pure data movement, HBM-bound at 2 * L * 8 bytes per entry.
// Each block stages a tile in LDS so that both its global reads and its
// global writes are contiguous runs.
#include "common.hpp"

namespace mdg {

namespace {

constexpr int PL_BLOCK = 256;   // entries of one row per planar block

// interleaved (rows x cols, row stride ldx entries) -> planes
// planar[l * rows * cols + r * cols + c]
template <int L>
__global__ __launch_bounds__(PL_BLOCK) void to_planar_kernel(int rows, int cols, const double *__restrict__ x,
                                                              long long ldx, double *__restrict__ planar)
{
    __shared__ double t[PL_BLOCK * L + PL_BLOCK * L / 16];
    const int nbc = (cols + PL_BLOCK - 1) / PL_BLOCK;
    const int r = blockIdx.x / nbc;
    const int c0 = (blockIdx.x % nbc) * PL_BLOCK;
    const int nc = min(PL_BLOCK, cols - c0);
    const double *src = x + ((long long)r * ldx + c0) * L;
    // the run of nc * L doubles is contiguous; one pad double per 16
    for (int i = threadIdx.x; i < nc * L; i += PL_BLOCK) t[i + i / 16] = src[i];
    __syncthreads();
    const long long plane = (long long)rows * cols;
    const long long base = (long long)r * cols + c0;
    if (threadIdx.x < nc) {
#pragma unroll
        for (int l = 0; l < L; ++l) {
            const int i = threadIdx.x * L + l;
            planar[l * plane + base + threadIdx.x] = t[i + i / 16];
        }
    }
}

template <int L>
__global__ __launch_bounds__(PL_BLOCK) void from_planar_kernel(int rows, int cols, const double *__restrict__ planar,
                                                                double *__restrict__ x, long long ldx)
{
    __shared__ double t[PL_BLOCK * L + PL_BLOCK * L / 16];
    const int nbc = (cols + PL_BLOCK - 1) / PL_BLOCK;
    const int r = blockIdx.x / nbc;
    const int c0 = (blockIdx.x % nbc) * PL_BLOCK;
    const int nc = min(PL_BLOCK, cols - c0);
    const long long plane = (long long)rows * cols;
    const long long base = (long long)r * cols + c0;
    if (threadIdx.x < nc) {
#pragma unroll
        for (int l = 0; l < L; ++l) {
            const int i = threadIdx.x * L + l;
            t[i + i / 16] = planar[l * plane + base + threadIdx.x];
        }
    }
    __syncthreads();
    double *dst = x + ((long long)r * ldx + c0) * L;
    for (int i = threadIdx.x; i < nc * L; i += PL_BLOCK) dst[i] = t[i + i / 16];
}

constexpr int TT = 16;   // transpose tile: TT x TT entries

// y (cols x rows, row stride ldy) = x^T, x rows x cols with row stride ldx
template <int L>
__global__ __launch_bounds__(256) void transpose_kernel(int rows, int cols, const double *__restrict__ x, long long ldx,
                                                        double *__restrict__ y, long long ldy)
{
    // tile[tr][tc][l], row stride TT * L + 1 (odd: no bank conflicts on the
    // column-wise reads)
    constexpr int RS = TT * L + 1;
    __shared__ double t[TT * RS];
    const int nbc = (cols + TT - 1) / TT;
    const int r0 = (blockIdx.x / nbc) * TT, c0 = (blockIdx.x % nbc) * TT;
    const int nr = min(TT, rows - r0), nc = min(TT, cols - c0);
    for (int i = threadIdx.x; i < TT * TT * L; i += 256) {
        const int tr = i / (TT * L), rem = i % (TT * L);   // rem = tc * L + l
        if (tr < nr && rem < nc * L) t[tr * RS + rem] = x[((long long)(r0 + tr) * ldx + c0) * L + rem];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < TT * TT * L; i += 256) {
        const int tc = i / (TT * L), rem = i % (TT * L);   // rem = tr * L + l
        const int tr = rem / L, l = rem % L;
        if (tc < nc && tr < nr) y[((long long)(c0 + tc) * ldy + r0) * L + rem] = t[tr * RS + tc * L + l];
    }
}

}  // namespace

int to_planar(int L, int rows, int cols, const double *x, long long ldx, double *planar, hipStream_t s)
{
    const dim3 grid((unsigned)(((cols + PL_BLOCK - 1) / PL_BLOCK) * (long long)rows));
    switch (L) {
    case 2: hipLaunchKernelGGL(to_planar_kernel<2>, grid, dim3(PL_BLOCK), 0, s, rows, cols, x, ldx, planar); break;
    case 4: hipLaunchKernelGGL(to_planar_kernel<4>, grid, dim3(PL_BLOCK), 0, s, rows, cols, x, ldx, planar); break;
    case 8: hipLaunchKernelGGL(to_planar_kernel<8>, grid, dim3(PL_BLOCK), 0, s, rows, cols, x, ldx, planar); break;
    case 16: hipLaunchKernelGGL(to_planar_kernel<16>, grid, dim3(PL_BLOCK), 0, s, rows, cols, x, ldx, planar); break;
    default: set_error("unsupported limb count L=%d", L); return MDG_ERROR_INVALID_VALUE;
    }
    MDG_LAUNCH_CHECK("to_planar_kernel");
    return MDG_SUCCESS;
}

int from_planar(int L, int rows, int cols, const double *planar, double *x, long long ldx, hipStream_t s)
{
    const dim3 grid((unsigned)(((cols + PL_BLOCK - 1) / PL_BLOCK) * (long long)rows));
    switch (L) {
    case 2: hipLaunchKernelGGL(from_planar_kernel<2>, grid, dim3(PL_BLOCK), 0, s, rows, cols, planar, x, ldx); break;
    case 4: hipLaunchKernelGGL(from_planar_kernel<4>, grid, dim3(PL_BLOCK), 0, s, rows, cols, planar, x, ldx); break;
    case 8: hipLaunchKernelGGL(from_planar_kernel<8>, grid, dim3(PL_BLOCK), 0, s, rows, cols, planar, x, ldx); break;
    case 16:
        hipLaunchKernelGGL(from_planar_kernel<16>, grid, dim3(PL_BLOCK), 0, s, rows, cols, planar, x, ldx);
        break;
    default: set_error("unsupported limb count L=%d", L); return MDG_ERROR_INVALID_VALUE;
    }
    MDG_LAUNCH_CHECK("from_planar_kernel");
    return MDG_SUCCESS;
}

int transpose(int L, int rows, int cols, const double *x, long long ldx, double *y, long long ldy, hipStream_t s)
{
    const dim3 grid((unsigned)(((cols + TT - 1) / TT) * (long long)((rows + TT - 1) / TT)));
    switch (L) {
    case 2: hipLaunchKernelGGL(transpose_kernel<2>, grid, dim3(256), 0, s, rows, cols, x, ldx, y, ldy); break;
    case 4: hipLaunchKernelGGL(transpose_kernel<4>, grid, dim3(256), 0, s, rows, cols, x, ldx, y, ldy); break;
    case 8: hipLaunchKernelGGL(transpose_kernel<8>, grid, dim3(256), 0, s, rows, cols, x, ldx, y, ldy); break;
    case 16: hipLaunchKernelGGL(transpose_kernel<16>, grid, dim3(256), 0, s, rows, cols, x, ldx, y, ldy); break;
    default: set_error("unsupported limb count L=%d", L); return MDG_ERROR_INVALID_VALUE;
    }
    MDG_LAUNCH_CHECK("transpose_kernel");
    return MDG_SUCCESS;
}

}  // namespace mdg
