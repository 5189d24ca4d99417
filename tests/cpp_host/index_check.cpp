// Exhaustive host-side check of the device index math (no GPU): every ghost
// position of random ghost boxes is enumerated exactly once, every read
// (mirror / wrap source) lands in the interior or in storage, and every write
// lands in storage. Built and run by tests/test_index_math.py.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <set>
#include <tuple>

#include "../../opflow_amd/csrc/ghost_index.hpp"

using namespace opf;

static int fail(const char* m, int trial) {
    std::printf("FAIL trial %d: %s\n", trial, m);
    return 1;
}

int main() {
    std::mt19937 rng(1234);
    auto U = [&](int a, int b) { return std::uniform_int_distribution<int>(a, b)(rng); };
    for (int trial = 0; trial < 3000; ++trial) {
        GhostInfo gi{};
        gi.dim = U(1, 3);
        for (int j = 0; j < 3; ++j) {
            if (j >= gi.dim) {
                gi.ls[j] = gi.gs[j] = gi.off[j] = 0;
                gi.le[j] = gi.ge[j] = gi.ext[j] = 1;
                continue;
            }
            gi.per[j] = U(0, 3) == 0;
            gi.corner[j] = U(0, 1);
            int L = U(2, 9);
            gi.ls[j] = U(-2, 2);
            gi.le[j] = gi.ls[j] + L;
            int lo = U(0, 3), hi = U(0, 3);
            int cap = gi.corner[j] && !gi.per[j] ? L - 1 : L;
            if (lo > cap) lo = cap;
            if (hi > cap) hi = cap;
            gi.gs[j] = gi.ls[j] - lo;
            gi.ge[j] = gi.le[j] + hi;
            int pad = std::max(lo, hi) + U(0, 2);
            gi.off[j] = gi.ls[j] - pad;
            gi.ext[j] = L + 2 * pad;
            gi.ngl[j][0] = lo;
            gi.ngl[j][1] = hi;
            for (int s = 0; s < 2; ++s) gi.nobc[j][s] = gi.per[j] ? 0 : (U(0, 5) == 0);
        }
        long long n[3] = {0, 0, 0};
        for (int k = 0; k < gi.dim; ++k) {
            long long c = 1;
            for (int j = 0; j < 3; ++j) {
                if (j < k) c *= gi.le[j] - gi.ls[j];
                else if (j == k) c *= (gi.ls[j] - gi.gs[j]) + (gi.ge[j] - gi.le[j]);
                else c *= gi.ge[j] - gi.gs[j];
            }
            n[k] = c;
        }
        long long tot = n[0] + n[1] + n[2];
        long long storage = (long long)gi.ext[0] * gi.ext[1] * gi.ext[2];
        std::set<std::tuple<int, int, int>> seen;
        for (long long t = 0; t < tot + 64; ++t) {
            int p[3], src[3], mir[3], lay[3], side[3];
            int act = ghost_locate(gi, t, n[0], n[1], n[2], p, src, mir, lay, side);
            {  // the 32-bit index form the kernel launches with gives the same answer
                int p2[3], src2[3], mir2[3], lay2[3], side2[3];
                const int act2 = ghost_locate<uint32_t>(gi, (uint32_t)t, (uint32_t)n[0], (uint32_t)n[1],
                                                        (uint32_t)n[2], p2, src2, mir2, lay2, side2);
                if (act2 != act) return fail("32-bit ghost_locate disagrees", trial);
                for (int j = 0; j < 3 && act != GHOST_SKIP; ++j)
                    if (p2[j] != p[j] || src2[j] != src[j] || (act == GHOST_RULE && (mir2[j] != mir[j] ||
                                                                                      lay2[j] != lay[j] ||
                                                                                      side2[j] != side[j])))
                        return fail("32-bit ghost_locate disagrees", trial);
            }
            if (t >= tot) {
                if (act != GHOST_SKIP) return fail("thread beyond the slabs did work", trial);
                continue;
            }
            bool interior = true, inbox = true;
            for (int j = 0; j < 3; ++j) {
                interior &= gi.ls[j] <= p[j] && p[j] < gi.le[j];
                inbox &= gi.gs[j] <= p[j] && p[j] < gi.ge[j];
            }
            if (interior || !inbox) return fail("enumerated position not a ghost of the box", trial);
            if (!seen.insert({p[0], p[1], p[2]}).second) return fail("ghost position enumerated twice", trial);
            long long lp = ghost_lin(gi, p);
            if (lp < 0 || lp >= storage) return fail("write outside storage", trial);
            if (act == GHOST_COPY) {
                long long ls = ghost_lin(gi, src);
                if (ls < 0 || ls >= storage) return fail("copy source outside storage", trial);
            } else if (act == GHOST_RULE) {
                for (int j = 0; j < 3; ++j)
                    if (mir[j] < gi.ls[j] || mir[j] >= gi.le[j]) return fail("mirror not interior", trial);
                for (int j = 0; j < 3; ++j)
                    if (lay[j] >= 0 && (lay[j] >= gi.ngl[j][side[j]] || lay[j] >= 4))
                        return fail("rule layer out of table", trial);
            }
        }
        // coverage: every ghost of the box was enumerated
        long long box = 1, inner = 1;
        for (int j = 0; j < 3; ++j) {
            box *= gi.ge[j] - gi.gs[j];
            inner *= gi.le[j] - gi.ls[j];
        }
        if ((long long)seen.size() != box - inner) return fail("ghost box not covered", trial);
    }
    std::printf("index checks passed\n");
    return 0;
}
