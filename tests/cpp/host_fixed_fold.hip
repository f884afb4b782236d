// host_fixed_fold.hip -- the fixed-point conversion, digit extraction and
// reconstruction folds of the slice GEMMs (csrc/slices.hpp, csrc/recon.hpp),
// compiled for the host so that tests/test_host_fixed_fold.py can check them
// without a GPU.
//
//   host_fixed_fold fixed < cases   per line "L shift S w x_0 .. x_{L-1}"
//                                   (limbs as hexadecimal bit patterns):
//                                   prints N's NW words (least significant
//                                   first), then the S digits of
//                                   fixed_digits and, for w = 8, those of
//                                   fixed_digits8
//   host_fixed_fold fold SEED N     N random cases of diagonal sums: limbs
//                                   of band_fold over band planes (bands of
//                                   4 and of 3, chunked and unchunked) and
//                                   of diag_fold over the same sums; prints
//                                   "mismatch ..." lines and "cases C"
//   host_fixed_fold items           the band kernels' work items (band_items)
//                                   for every slice count, band width, depth
//                                   (to 32767 k-blocks) and chunk capacity
//                                   used: each (band,
//                                   K-chunk) plane written exactly once and
//                                   inside the plane array, each band's
//                                   k-blocks covered once, every chunk's
//                                   INT32 sums within their exact range;
//                                   prints "bad ..." lines and "items I"
#include <hip/hip_runtime.h>
#undef __device__
#define __device__ __attribute__((host, device))
#define __builtin_amdgcn_sched_barrier(x) ((void)0)
#include "../../multiple-double-arithmetic-matrix-matrix-multiplication_amd/csrc/i8band.hpp"
#include "../../multiple-double-arithmetic-matrix-matrix-multiplication_amd/csrc/recon.hpp"
#include "../../multiple-double-arithmetic-matrix-matrix-multiplication_amd/csrc/slices.hpp"

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

namespace mdg {
void set_error(const char *, ...) {}
}
using namespace mdg;

template <int L, int NW>
static void fixed_case(int shift, int S, int w, const double *x)
{
    uint32_t N[NW];
    to_fixed<L, NW>(x, shift, N);
    for (int j = 0; j < NW; ++j) std::printf("%08x ", N[j]);
    std::printf("|");
    std::vector<int> d(S, 0);
    fixed_digits<NW>(N, S, w, [&](int t, int v) { d[t] = v; });
    for (int t = 0; t < S; ++t) std::printf(" %d", d[t]);
    if (w == 8) {
        std::printf(" |");
        uint32_t M[NW];
        std::memcpy(M, N, sizeof(M));
        fixed_digits8<NW>(M, S, [&](int t, int v) { d[t] = v; });
        for (int t = 0; t < S; ++t) std::printf(" %d", d[t]);
    }
    std::printf("\n");
}

static int run_fixed()
{
    char line[4096];
    while (std::fgets(line, sizeof(line), stdin)) {
        char *p = line;
        const int L = (int)std::strtol(p, &p, 10), shift = (int)std::strtol(p, &p, 10);
        const int S = (int)std::strtol(p, &p, 10), w = (int)std::strtol(p, &p, 10);
        double x[16];
        for (int l = 0; l < L; ++l) {
            const unsigned long long b = std::strtoull(p, &p, 16);
            std::memcpy(&x[l], &b, 8);
        }
        switch (L) {
        case 2: fixed_case<2, FixedWords<2>::NW>(shift, S, w, x); break;
        case 4: fixed_case<4, FixedWords<4>::NW>(shift, S, w, x); break;
        case 8: fixed_case<8, FixedWords<8>::NW>(shift, S, w, x); break;
        default: fixed_case<16, FixedWords<16>::NW>(shift, S, w, x); break;
        }
    }
    return 0;
}

// one case: random chunk sums for S slices, bands of G, nkb k-blocks cut
// into K-chunks for a capacity cap
template <int L, int G>
static bool fold_case(std::mt19937_64 &rng, int S, int nkb, int cap, bool big)
{
    const int nb = (S + G - 1) / G;
    std::vector<long long> Dg(S, 0), Pl(bnd_band_planes(S, G, nkb, cap), 0);
    for (int q = 0; q < nb; ++q) {
        const int g0 = S - G * (nb - q);
        for (int c = 0; c < bnd_band_chunks(S, G, nkb, cap, q); ++c) {
            long long P = 0;
            for (int j = 0; j < G; ++j) {
                const int g = g0 + j;
                long long d = 0;
                if (g >= 0) {
                    const long long lim = big ? (1ll << 31) - 1 : (1ll << 20);
                    d = (long long)(rng() % (2 * lim + 1)) - lim;
                    if (rng() % 9 == 0) d = 0;
                    if (rng() % 13 == 0) d = (rng() & 1) ? lim : -lim;
                }
                if (g >= 0) Dg[g] += d;
                P = P * 256 + d;
            }
            Pl[bnd_first_band_plane(S, G, nkb, cap, q) + c] = P;
        }
    }
    const int sc0 = (int)(rng() % 200) - 100 - 8 * (S + 1);
    double y1[L], y2[L];
    const int rc1 = with_band_words<L, G>(S, [&](auto nwb) {
        constexpr int NWB = decltype(nwb)::value;
        const BandPlanes bp = bnd_plane_table(S, G, nkb, cap);
        auto P = [&](int p) { return Pl[p]; };
        auto nch = [&](int q) { return (int)bp.nch[q]; };
        if (bp.total > bp.nb)
            band_fold<L, NWB, G, true>(S, sc0, P, nch, bp.total, y1);
        else
            band_fold<L, NWB, G, false>(S, sc0, P, nch, bp.total, y1);
    });
    const int rc2 = with_diag_words<L>(S, 8, [&](auto nw) {
        constexpr int NW = decltype(nw)::value;
        diag_fold<L, NW, 3>(S, 8, sc0, [&](int g) { return Dg[g]; }, y2);
    });
    if (rc1 || rc2 || std::memcmp(y1, y2, sizeof(y1))) {
        std::printf("mismatch L=%d G=%d S=%d nkb=%d cap=%d rc=%d,%d: %a %a\n", L, G, S, nkb, cap, rc1, rc2, y1[0],
                    y2[0]);
        return false;
    }
    return true;
}

static int run_fold(unsigned long long seed, int n)
{
    std::mt19937_64 rng(seed);
    int cases = 0, bad = 0;
    for (int it = 0; it < n; ++it) {
        // small capacities: many K-chunks per band; the large one: none
        const int cap = (it % 3 == 0) ? 1 + (int)(rng() % 300) : BND_CAP_I8;
        const int nkb = 1 + (int)(rng() % 40);
        const bool big = it & 1;
        bad += !fold_case<2, 4>(rng, 1 + (int)(rng() % 31), nkb, cap, big);
        bad += !fold_case<4, 4>(rng, 1 + (int)(rng() % 47), nkb, cap, big);
        bad += !fold_case<4, 3>(rng, 1 + (int)(rng() % 47), nkb, cap, big);
        bad += !fold_case<8, 4>(rng, 1 + (int)(rng() % 79), nkb, cap, big);
        cases += 4;
        if (it % 8 == 0) {
            bad += !fold_case<16, 4>(rng, 1 + (int)(rng() % 143), nkb, cap, big);
            bad += !fold_case<16, 3>(rng, 1 + (int)(rng() % 143), nkb, cap, big);
            cases += 2;
        }
        if (bad > 10) break;
    }
    std::printf("cases %d\n", cases);
    return bad ? 1 : 0;
}

// every item of band_items(S, G, nkb, cap): the plane the kernel's epilogue
// writes, bnd_first_band_plane(band) + kb0 / bnd_band_kblocks(band), is in
// [0, bnd_band_planes) and taken once; each band's k-blocks are covered once
// by its items; n_t = g0 + G slices times the chunk's k-blocks stay within
// cap (so every INT32 sum of the chunk is exact)
static int run_items()
{
    long long items = 0;
    int bad = 0;
    auto check = [&](int G, int cap, int S, int nkb) {
        const int nb = (S + G - 1) / G, np = bnd_band_planes(S, G, nkb, cap);
        std::vector<int> seen(np, 0);
        std::vector<std::vector<int>> cover(nb, std::vector<int>(nkb, 0));
        std::vector<BandWork> v;
        band_items(S, G, nkb, cap, v);
        for (const BandWork &w : v) {
            const BandItem &it = w.it;
            const int q = (it.g0 - (S - G * nb)) / G;
            const int plane = bnd_first_band_plane(S, G, nkb, cap, q) + it.kb0 / bnd_band_kblocks(S, G, cap, q);
            const bool ok = q >= 0 && q < nb && it.gb == G && it.g0 == S - G * (nb - q) && plane >= 0 && plane < np &&
                            it.kb0 >= 0 && it.kb0 < it.kb1 && it.kb1 <= nkb &&
                            (long long)(it.g0 + G) * (it.kb1 - it.kb0) <= std::max(cap, it.g0 + G);
            if (!ok) {
                if (bad++ < 20)
                    std::printf("bad item G=%d cap=%d S=%d nkb=%d: g0=%d kb %d..%d plane %d of %d\n", G, cap, S, nkb,
                                it.g0, it.kb0, it.kb1, plane, np);
                continue;
            }
            ++seen[plane];
            for (int kb = it.kb0; kb < it.kb1; ++kb) ++cover[q][kb];
            ++items;
        }
        for (int p = 0; p < np; ++p)
            if (seen[p] != 1 && bad++ < 20)
                std::printf("bad plane G=%d cap=%d S=%d nkb=%d: plane %d written %d times\n", G, cap, S, nkb, p,
                            seen[p]);
        for (int q = 0; q < nb; ++q)
            for (int kb = 0; kb < nkb; ++kb)
                if (cover[q][kb] != 1 && bad++ < 20)
                    std::printf("bad cover G=%d cap=%d S=%d nkb=%d: band %d k-block %d\n", G, cap, S, nkb, q, kb);
    };
    std::vector<int> nkbs;
    for (int v = 1; v <= 70; ++v) nkbs.push_back(v);
    for (int v : {127, 128, 129, 255, 256, 511, 1023, 2047}) nkbs.push_back(v);
    for (int G : {3, 4})
        for (int cap : {BND_CAP_I8, BND_CAP_BF16, 40, 7}) {
            for (int S = 1; S < 256; ++S)
                for (int nkb : nkbs) check(G, cap, S, nkb);
            // depths of several INT32 planes per slice pair, up to the
            // 32767 k-blocks a band item can index (the kernels' capacities)
            if (cap == BND_CAP_I8 || cap == BND_CAP_BF16)
                for (int S : {1, 2, 15, 30, 107, 255})
                    for (int nkb : {2048, 3125, 16400, 32767}) check(G, cap, S, nkb);
        }
    std::printf("items %lld\n", items);
    return bad ? 1 : 0;
}

int main(int argc, char **argv)
{
    if (argc >= 2 && std::string(argv[1]) == "fixed") return run_fixed();
    if (argc >= 2 && std::string(argv[1]) == "items") return run_items();
    if (argc >= 4 && std::string(argv[1]) == "fold") return run_fold(std::strtoull(argv[2], nullptr, 10), std::atoi(argv[3]));
    std::fprintf(stderr, "usage: %s fixed < cases | fold SEED N\n", argv[0]);
    return 2;
}
