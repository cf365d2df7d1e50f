// The parallel replay of a sequential floating-point sum (mo-dmft_amd/csrc/moed_osum.h) on the
// host: the workgroup's passes of ordered_sum (moed_fit.hip) emulated thread by thread with the
// header's arithmetic, against the plain loop s = s + t[i] on random and adversarial inputs.
// Usage: osum_check [cases]; exit status 1 on any mismatch.
#include "moed_osum.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>
using namespace moed;
static const int NT = 1024;

// one chunk of n <= NT no terms, as ordered_sum runs it (nseq leading terms one by one)
static double chunk_sum(double s, const double* T, int n, int no, int nseq, int& npass) {
    int pos = 0;
    std::vector<double> kt(NT), excl(NT), incl(NT);
    while (pos < n) {
        if (!sum_normal(s) || pos < nseq) {
            while (pos < n && (pos < nseq || !sum_normal(s))) s = s + T[pos++];
            continue;
        }
        ++npass;
        const int e0 = ilogb(s);
        int ftie = n;
        bool bad = false;
        for (int j = 0; j < NT; ++j) {
            kt[j] = 0.0;
            for (int o = 0; o < no; ++o) {
                const int i = j * no + o;
                if (i >= pos && i < n) {
                    bad = bad || !(T[i] >= 0.0);
                    bool tie;
                    kt[j] = kt[j] + sum_count(T[i], e0, tie);
                    if (tie) ftie = std::min(ftie, i);
                }
            }
        }
        // per-wave inclusive scan, then the waves' totals in wave order (any order of exact adds)
        double woff = 0.0, tot = 0.0;
        for (int w = 0; w < NT / 64; ++w) {
            double c = 0.0;
            for (int l = 0; l < 64; ++l) {
                excl[w * 64 + l] = woff + c;
                c = c + kt[w * 64 + l];
                incl[w * 64 + l] = woff + c;
            }
            woff = woff + c;
        }
        tot = woff;
        const double M = ldexp(s, 52 - e0), thr = (double)kSumLim - M;
        if (!bad && ftie >= n && tot <= thr) {
            s = ldexp(M + tot, e0 - 52);
            pos = n;
            continue;
        }
        if (bad) {
            while (pos < n) s = s + T[pos++];
            continue;
        }
        int cross = n;
        for (int j = 0; j < NT; ++j)
            if (excl[j] <= thr && incl[j] > thr) {
                double c = excl[j];
                for (int o = 0; o < no; ++o) {
                    const int i = j * no + o;
                    if (i >= pos && i < n) {
                        bool tie;
                        c = c + sum_count(T[i], e0, tie);
                        if (c > thr) { cross = i; break; }
                    }
                }
            }
        const int is = std::min(cross, ftie);
        const int j = is / no;
        double c = excl[j];
        for (int o = 0; o < no; ++o) {
            const int i = j * no + o;
            if (i >= pos && i < is) {
                bool tie;
                c = c + sum_count(T[i], e0, tie);
            }
        }
        s = ldexp(M + c, e0 - 52) + T[is];
        pos = is + 1;
    }
    return s;
}

int main(int argc, char** argv) {
    const int ncase = argc > 1 ? atoi(argv[1]) : 20000;
    std::mt19937_64 rng(12345);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    long bad = 0, cases = 0, passes = 0;
    for (int it = 0; it < ncase; ++it) {
        const int no = 1 + (int)(rng() % 4);
        const int n = 1 + (int)(rng() % (3 * NT * no));
        std::vector<double> T(n);
        const int kind = (int)(rng() % 9);
        double s0 = (rng() % 3 == 0) ? 0.0 : ldexp(U(rng) + 0.5, (int)(rng() % 60) - 30);
        for (int i = 0; i < n; ++i) {
            double t;
            switch (kind) {
                case 0: t = U(rng); break;                                                   // uniform
                case 1: t = ldexp(U(rng), -(int)(rng() % 60)); break;                        // wide range
                case 2: t = 1.0 / ((i + 1.0) * (i + 1.0) * (i + 1.0)); break;                // decaying
                case 3: t = (rng() % 5 == 0) ? 0.0 : ldexp(1.0, -(int)(rng() % 80)); break;  // ties
                case 4: t = ldexp((double)(rng() % 7), -(int)(rng() % 8) - 50); break;       // ulp-sized
                case 5: t = (rng() % 50 == 0) ? ldexp(U(rng), 40) : U(rng) * 1e-3; break;    // jumps
                case 6: t = (rng() % 200 == 0) ? -U(rng) : U(rng); break;                    // negatives
                case 7: t = (i == n / 2 && rng() % 2) ? NAN : (i == n / 3 ? INFINITY : U(rng)); break;
                default: t = ldexp(U(rng), -1060); break;                                    // subnormal
            }
            T[i] = t;
        }
        if (kind == 4) s0 = ldexp(1.0, (int)(rng() % 3));
        double ref = s0;
        for (int i = 0; i < n; ++i) ref = ref + T[i];
        const int nseq = (rng() % 2) ? 16 : 0;
        int np = 0;
        double got = s0;
        for (int base = 0; base < n; base += NT * no) {
            const int m = std::min(NT * no, n - base);
            got = chunk_sum(got, T.data() + base, m, no, base == 0 ? nseq : 0, np);
        }
        passes += np;
        ++cases;
        if (memcmp(&ref, &got, 8) != 0 && !(std::isnan(ref) && std::isnan(got))) {
            if (bad < 10) printf("MISMATCH kind %d n %d no %d: %.17g vs %.17g\n", kind, n, no, got, ref);
            ++bad;
        }
    }
    printf("cases %ld mismatches %ld passes/case %.2f\n", cases, bad, (double)passes / cases);
    return bad != 0;
}
