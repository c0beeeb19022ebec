// Host build of the coupling recursion (gmaster_amd/csrc/coupling_rec.h) for CPU tests.
//   coupling_host LMAX LMAX_W WINDOW_IN OUT_PAIR OUT_WAVE
// WINDOW_IN holds W' (LMAX_W + 1 doubles).  OUT_PAIR / OUT_WAVE receive (L, L, 4) doubles:
// the row sums of every pair l1 <= l2 from couple_pair and from the one-l1-per-wave driver
// (every item of every wave, lanes emulated one by one, window reads bounds-checked);
// entries l2 < l1 are zero.  Exit status 3 when an item decodes outside the triangle, twice,
// or not at all, or a window read leaves its 128 entries.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../gmaster_amd/csrc/coupling_rec.h"

namespace {

struct HostWin {
    const double* wp;
    int lmax_w, d0, base;
    bool bad;
    void stage(int k0) { base = d0 + k0; }
    double at(int t) {
        if (t < 0 || t >= 128) bad = true;
        const int i = base + t;
        return i <= lmax_w ? wp[i] : 0.0;
    }
};

std::vector<double> g_table(int n) {
    std::vector<long double> g(n);
    g[0] = 1.0L;
    for (int p = 1; p < n; ++p) g[p] = g[p - 1] * (2.0L * p - 1.0L) / (2.0L * p);
    return std::vector<double>(g.begin(), g.end());
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 6) return 2;
    const int lmax = atoi(argv[1]), lmax_w = atoi(argv[2]);
    const int L = lmax + 1;
    std::vector<double> wp(lmax_w + 1);
    FILE* f = fopen(argv[3], "rb");
    if (!f || fread(wp.data(), sizeof(double), wp.size(), f) != wp.size()) return 2;
    fclose(f);
    const std::vector<double> g = g_table(L + 1);
    std::vector<double> pair((size_t)L * L * 4, 0.0), wave((size_t)L * L * 4, 0.0);
    for (int l1 = 0; l1 < L; ++l1)
        for (int l2 = l1; l2 < L; ++l2) gm::couple_pair(wp.data(), lmax_w, g.data(), l1, l2, &pair[((size_t)l1 * L + l2) * 4]);
    std::vector<int> seen((size_t)L * L, 0);
    const int64_t items = gm::couple_items(L);
    int status = 0;
    for (int64_t it = 0; it < items; ++it) {
        int l1, j;
        gm::couple_decode(it, L, l1, j);
        if (l1 < 0 || l1 >= L || j < 0 || l1 + 64 * j >= L) {
            status = 3;
            continue;
        }
        for (int lane = 0; lane < 64; ++lane) {
            HostWin w{wp.data(), lmax_w, 64 * j, 0, false};
            double s[4];
            gm::couple_wave_lane(w, lmax_w, g.data(), L, l1, l1 + 64 * j, lane, s);
            if (w.bad) status = 3;
            const int l2 = l1 + 64 * j + lane;
            if (l2 >= L) continue;
            if (seen[(size_t)l1 * L + l2]++) status = 3;
            for (int q = 0; q < 4; ++q) wave[((size_t)l1 * L + l2) * 4 + q] = s[q];
        }
    }
    for (int l1 = 0; l1 < L; ++l1)
        for (int l2 = l1; l2 < L; ++l2)
            if (seen[(size_t)l1 * L + l2] != 1) status = 3;
    FILE* o1 = fopen(argv[4], "wb");
    FILE* o2 = fopen(argv[5], "wb");
    if (!o1 || !o2) return 2;
    fwrite(pair.data(), sizeof(double), pair.size(), o1);
    fwrite(wave.data(), sizeof(double), wave.size(), o2);
    fclose(o1);
    fclose(o2);
    return status;
}
