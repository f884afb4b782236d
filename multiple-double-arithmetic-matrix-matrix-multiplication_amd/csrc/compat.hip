// compat.hip -- the reference's vector<double> API (include/mdgemm_compat.hpp)
// over the C ABI.  Host-side plumbing only: copies in, one library call,
// copies out.  Layout utilities (bigA, bigB) run as device kernels too.
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mdgemm_compat.hpp"
#include "common.hpp"

namespace {

void throw_if(int rc, const char *what)
{
    if (rc != MDG_SUCCESS) throw std::runtime_error(std::string(what) + ": " + mdg_last_error());
}

void hip_ok(hipError_t e, const char *what)
{
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// owning device buffer
struct DevBuf {
    double *p = nullptr;
    size_t n = 0;
    explicit DevBuf(size_t count) : n(count)
    {
        hip_ok(hipMalloc(&p, sizeof(double) * (count ? count : 1)), "hipMalloc");
    }
    DevBuf(const std::vector<double> &v) : DevBuf(v.size())
    {
        if (n) hip_ok(hipMemcpy(p, v.data(), sizeof(double) * n, hipMemcpyHostToDevice), "hipMemcpy H2D");
    }
    ~DevBuf() { (void)hipFree(p); }
    std::vector<double> host() const
    {
        std::vector<double> v(n);
        if (n) hip_ok(hipMemcpy(v.data(), p, sizeof(double) * n, hipMemcpyDeviceToHost), "hipMemcpy D2H");
        return v;
    }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
};

unsigned long long g_seed = 0x5EEDull;

void need(bool ok, const char *msg)
{
    if (!ok) throw std::invalid_argument(msg);
}

__global__ void bigA_kernel(const double *A, int n, int p, double *out)
{
    // out row r*p + i, column c  =  A[r][c][i]   (C++/TCcuda.cu:43-53)
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = (long long)n * n * p;
    if (idx >= total) return;
    const int c = (int)(idx % n);
    const long long t = idx / n;
    const int i = (int)(t % p);
    const int r = (int)(t / p);
    out[idx] = A[((long long)r * n + c) * p + i];
}

__global__ void bigB_kernel(const double *B, int n, int p, double *out)
{
    // column-major (n p) x (n p): element (row = j*n + r, col = i*n + c) =
    // B[r][c][i-j] for j <= i  (C++/TCcuda.cu:61-77)
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = (long long)n * p * n * p;
    if (idx >= total) return;
    const int r = (int)(idx % n);
    long long t = idx / n;
    const int j = (int)(t % p);
    t /= p;
    const int c = (int)(t % n);
    const int i = (int)(t / n);
    out[idx] = (j <= i) ? B[((long long)r * n + c) * p + (i - j)] : 0.0;
}

std::vector<double> host_gemm(std::vector<double> &A, std::vector<double> &B, int n, int p)
{
    need(n >= 0, "n must be non-negative");
    need((long long)A.size() == (long long)n * n * p && (long long)B.size() == (long long)n * n * p,
         "A and B must hold n*n*p doubles");
    DevBuf dA(A), dB(B), dC((size_t)n * n * p);
    throw_if(mdg_vector_gemm(p, n, n, n, dA.p, n, dB.p, n, dC.p, n, MDG_TWOPROD_DEKKER, nullptr), "vector GEMM");
    return dC.host();
}

}  // namespace

unsigned long long mdgemm_compat_seed(unsigned long long seed)
{
    unsigned long long old = g_seed;
    g_seed = seed;
    return old;
}

std::vector<double> mat(int n, int p, int expmin, int expmax)
{
    need(n >= 0 && p >= 1, "bad n or p");
    DevBuf d((size_t)n * n * p);
    throw_if(mdg_fill_random(2, p, (long long)n * n, g_seed++, expmin, expmax, d.p, nullptr), "mat");
    return d.host();
}

std::vector<double> zeros(int m, int n, int p) { return std::vector<double>((size_t)m * n * p, 0.0); }

std::vector<double> part(std::vector<double> x, int p, int i)
{
    std::vector<double> xi;
    for (size_t j = (size_t)i; j < x.size(); j += (size_t)p) xi.push_back(x[j]);
    return xi;
}

std::vector<double> bigA(std::vector<double> A, int n, int p)
{
    need((long long)A.size() == (long long)n * n * p, "A must hold n*n*p doubles");
    DevBuf dA(A), out(A.size());
    const long long total = (long long)A.size();
    if (total) hipLaunchKernelGGL(bigA_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, 0, dA.p, n, p, out.p);
    hip_ok(hipGetLastError(), "bigA launch");
    return out.host();
}

std::vector<double> bigB(std::vector<double> B, int n, int p)
{
    need((long long)B.size() == (long long)n * n * p, "B must hold n*n*p doubles");
    const long long total = (long long)n * p * n * p;
    DevBuf dB(B), out((size_t)total);
    if (total) hipLaunchKernelGGL(bigB_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, 0, dB.p, n, p, out.p);
    hip_ok(hipGetLastError(), "bigB launch");
    return out.host();
}

std::vector<double> bigB2(std::vector<double> B, int n, int p)
{
    need((long long)B.size() == (long long)n * n * p, "B must hold n*n*p doubles");
    DevBuf dB(B), out((size_t)n * p * n * p);
    throw_if(mdg_bigB2(dB.p, n, p, out.p, nullptr), "bigB2");
    return out.host();
}

std::vector<double> matmulTCnt(std::vector<double> A, std::vector<double> B, int n, int nfrag, int p, int loop_ct,
                               float &t_CUDA)
{
    need(nfrag >= 1 && n % nfrag == 0, "nfrag must divide n");
    need(p == 2 || p == 4 || p == 8 || p == 16, "p must be 2, 4, 8 or 16");
    need((long long)A.size() == (long long)n * n * p && (long long)B.size() == (long long)n * n * p,
         "A and B must hold n*n*p doubles");
    DevBuf dA(A), dB(B), dC((size_t)n * n * p);
    hip_ok(hipMemset(dC.p, 0, sizeof(double) * dC.n), "hipMemset");
    hipEvent_t t0, t1;
    hip_ok(hipEventCreate(&t0), "hipEventCreate");
    hip_ok(hipEventCreate(&t1), "hipEventCreate");
    hip_ok(hipEventRecord(t0, nullptr), "hipEventRecord");
    for (int l = 0; l < loop_ct; ++l)
        throw_if(mdg_vector_gemm(p, n, n, n, dA.p, n, dB.p, n, dC.p, n, MDG_TWOPROD_FMA, nullptr), "matmulTCnt");
    hip_ok(hipEventRecord(t1, nullptr), "hipEventRecord");
    hip_ok(hipEventSynchronize(t1), "hipEventSynchronize");
    hip_ok(hipEventElapsedTime(&t_CUDA, t0, t1), "hipEventElapsedTime");
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    return dC.host();
}

std::vector<double> matmulddf(std::vector<double> A, std::vector<double> B, int n) { return host_gemm(A, B, n, 2); }
std::vector<double> matmulqdf(std::vector<double> A, std::vector<double> B, int n) { return host_gemm(A, B, n, 4); }
std::vector<double> matmulodf(std::vector<double> A, std::vector<double> B, int n) { return host_gemm(A, B, n, 8); }
std::vector<double> matmulhdf(std::vector<double> A, std::vector<double> B, int n) { return host_gemm(A, B, n, 16); }

std::vector<double> matmulhost(std::vector<double> A, std::vector<double> B, int n, int p)
{
    need(n >= 0 && p >= 1, "bad n or p");
    need((long long)A.size() == (long long)n * n * p && (long long)B.size() == (long long)n * n * p,
         "A and B must hold n*n*p doubles");
    DevBuf dA(A), dB(B), dC((size_t)n * n * p);
    throw_if(mdg_matmulhost(p, n, dA.p, dB.p, dC.p, nullptr), "matmulhost");
    return dC.host();
}

static void renorm_in_place(std::vector<double> &A, int n, int p, int big, const char *what)
{
    need(n >= 0 && p >= 1, "bad n or p");
    need((long long)A.size() == (long long)n * n * p, "A must hold n*n*p doubles");
    DevBuf d(A);
    throw_if(mdg_renormhost(d.p, n, p, big, nullptr), what);
    A = d.host();
}

void renormhost(std::vector<double> &A, int n, int p) { renorm_in_place(A, n, p, 0, "renormhost"); }

void renormhostbig(std::vector<double> &A, int n, int p) { renorm_in_place(A, n, p, 1, "renormhostbig"); }

std::vector<double> pllsqueeze_old(std::vector<double> x, int p, int pp, float &t_CUDA)
{
    need(pp >= 1 && x.size() % (size_t)pp == 0, "x must hold a multiple of pp doubles");
    const long long nent = (long long)(x.size() / pp);
    DevBuf dx(x), dy((size_t)nent * p);
    hipEvent_t t0, t1;
    hip_ok(hipEventCreate(&t0), "hipEventCreate");
    hip_ok(hipEventCreate(&t1), "hipEventCreate");
    hip_ok(hipEventRecord(t0, nullptr), "hipEventRecord");
    throw_if(mdg_squeeze(p, nent, pp, dx.p, dy.p, nullptr), "pllsqueeze_old");
    hip_ok(hipEventRecord(t1, nullptr), "hipEventRecord");
    hip_ok(hipEventSynchronize(t1), "hipEventSynchronize");
    hip_ok(hipEventElapsedTime(&t_CUDA, t0, t1), "hipEventElapsedTime");
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    return dy.host();
}

namespace {

std::vector<double> squeeze_parts(const std::vector<double> &x, int p, int pp, const char *what)
{
    need(p == 2 || p == 4 || p == 8 || p == 16, "p must be 2, 4, 8 or 16");
    need(pp >= 1 && x.size() % (size_t)pp == 0, "x must hold a multiple of pp doubles");
    need(pp != 12 || p == 2, "pp == 12 (mixsplit2 parts) is a double-double split");
    const long long nent = (long long)(x.size() / pp);
    DevBuf dx(x), dy((size_t)nent * p);
    throw_if(mdg_squeeze(p, nent, pp, dx.p, dy.p, nullptr), what);
    return dy.host();
}

}  // namespace

std::vector<double> pllsqueeze_2_8(std::vector<double> x) { return squeeze_parts(x, 2, 8, "pllsqueeze_2_8"); }
std::vector<double> pllsqueeze_2_12(std::vector<double> x) { return squeeze_parts(x, 2, 12, "pllsqueeze_2_12"); }
std::vector<double> pllsqueeze_2_16(std::vector<double> x) { return squeeze_parts(x, 2, 16, "pllsqueeze_2_16"); }
std::vector<double> pllntsqueeze_2_8(std::vector<double> x) { return squeeze_parts(x, 2, 8, "pllntsqueeze_2_8"); }
std::vector<double> pllntsqueeze_2_12(std::vector<double> x)
{
    return squeeze_parts(x, 2, 12, "pllntsqueeze_2_12");
}
std::vector<double> pllntsqueeze_2_16(std::vector<double> x)
{
    return squeeze_parts(x, 2, 16, "pllntsqueeze_2_16");
}
std::vector<double> pllsqueeze(std::vector<double> x, int p, int pp) { return squeeze_parts(x, p, pp, "pllsqueeze"); }
std::vector<double> pllntsqueeze(std::vector<double> x, int p, int pp)
{
    return squeeze_parts(x, p, pp, "pllntsqueeze");
}

std::vector<double> split4pd(std::vector<double> x, int p)
{
    DevBuf dx(x), out(4 * x.size());
    throw_if(mdg_split4pd(dx.p, (long long)x.size(), p, out.p, nullptr), "split4pd");
    return out.host();
}

std::vector<double> split8pd(std::vector<double> x, int p)
{
    DevBuf dx(x), out(8 * x.size());
    throw_if(mdg_split8pd(dx.p, (long long)x.size(), p, out.p, nullptr), "split8pd");
    return out.host();
}

std::vector<double> mixsplit2(std::vector<double> x)
{
    DevBuf dx(x), out(6 * x.size());
    throw_if(mdg_mixsplit2(dx.p, (long long)x.size(), out.p, nullptr), "mixsplit2");
    return out.host();
}

void matmul_tc(const std::vector<double> &A, const std::vector<double> &B, std::vector<double> &C, int M, int N,
               int K)
{
    need((long long)A.size() == (long long)M * K && (long long)B.size() == (long long)K * N &&
             (long long)C.size() == (long long)M * N,
         "sizes must be M*K, K*N and M*N");
    DevBuf dA(A), dB(B), dC(C);
    throw_if(mdg_dgemm_acc(M, N, K, dA.p, dB.p, dC.p, nullptr), "matmul_tc");
    C = dC.host();
}
