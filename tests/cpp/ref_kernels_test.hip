// ref_kernels_test.hip -- launches the exported __global__ kernels
// ddmm/qdmm/odmm/hdmm (include/mdgemm_kernels.hpp) with the reference's own
// launch shape, Gr(n/nfrag, n/nfrag) blocks of Bl(nfrag, nfrag)
// (C++/TCcuda.cu:265-293), on inputs read from files, and writes C back.
//
//     ref_kernels_test p n nfrag A.bin B.bin C.bin
//
// tests/test_gpu_compat.py checks C against the oracle bit for bit.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "mdgemm_kernels.hpp"

static bool ok(hipError_t e, const char *what)
{
    if (e != hipSuccess) std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
    return e == hipSuccess;
}

static bool read_all(const char *path, std::vector<double> &v)
{
    FILE *f = std::fopen(path, "rb");
    if (!f) return false;
    const size_t got = std::fread(v.data(), sizeof(double), v.size(), f);
    std::fclose(f);
    return got == v.size();
}

int main(int argc, char **argv)
{
    if (argc < 7) {
        std::fprintf(stderr, "usage: %s p n nfrag A.bin B.bin C.bin\n", argv[0]);
        return 2;
    }
    const int p = std::atoi(argv[1]), n = std::atoi(argv[2]), nfrag = std::atoi(argv[3]);
    if ((p != 2 && p != 4 && p != 8 && p != 16) || n < 1 || nfrag < 1 || n % nfrag || nfrag * nfrag > 1024) {
        std::fprintf(stderr, "bad arguments\n");
        return 2;
    }
    const size_t cnt = (size_t)n * n * p;
    std::vector<double> A(cnt), B(cnt), C(cnt, 0.0);
    if (!read_all(argv[4], A) || !read_all(argv[5], B)) {
        std::fprintf(stderr, "cannot read the inputs\n");
        return 2;
    }
    double *A_d = nullptr, *B_d = nullptr, *C_d = nullptr;
    if (!ok(hipMalloc(&A_d, cnt * 8), "hipMalloc") || !ok(hipMalloc(&B_d, cnt * 8), "hipMalloc") ||
        !ok(hipMalloc(&C_d, cnt * 8), "hipMalloc"))
        return 1;
    if (!ok(hipMemcpy(A_d, A.data(), cnt * 8, hipMemcpyHostToDevice), "copy A") ||
        !ok(hipMemcpy(B_d, B.data(), cnt * 8, hipMemcpyHostToDevice), "copy B") ||
        !ok(hipMemset(C_d, 0, cnt * 8), "clear C"))
        return 1;
    const int nlen = n / nfrag;
    dim3 Gr(nlen, nlen), Bl(nfrag, nfrag);
    switch (p) {
    case 2: ddmm<<<Gr, Bl>>>(A_d, B_d, C_d, n); break;
    case 4: qdmm<<<Gr, Bl>>>(A_d, B_d, C_d, n); break;
    case 8: odmm<<<Gr, Bl>>>(A_d, B_d, C_d, n); break;
    default: hdmm<<<Gr, Bl>>>(A_d, B_d, C_d, n); break;
    }
    if (!ok(hipGetLastError(), "launch") || !ok(hipDeviceSynchronize(), "kernel")) return 1;
    if (!ok(hipMemcpy(C.data(), C_d, cnt * 8, hipMemcpyDeviceToHost), "copy C")) return 1;
    FILE *f = std::fopen(argv[6], "wb");
    if (!f || std::fwrite(C.data(), sizeof(double), cnt, f) != cnt) {
        std::fprintf(stderr, "cannot write C\n");
        return 2;
    }
    std::fclose(f);
    (void)hipFree(A_d);
    (void)hipFree(B_d);
    (void)hipFree(C_d);
    std::printf("%cdmm n=%d nfrag=%d ok\n", "d?q???o???????h"[p - 2], n, nfrag);
    return 0;
}
