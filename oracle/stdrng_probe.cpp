// stdrng_probe.cpp — prints draws of libstdc++'s std::mt19937_64 and the
// distributions the reference uses (MCIntegrator.hpp:32-34,
// TrialMoveInterface.hpp:120-124, SRRDVecMove.hpp:27), so tests can pin the
// oracle's restatement of them against the library itself.  TEST INFRASTRUCTURE.
//
//   words SEED N          raw engine outputs
//   real SEED N A B       uniform_real_distribution<double>(A, B)
//   int SEED N LO HI      uniform_int_distribution<int>(LO, HI)
//   mixed SEED N NVEC     N x (int(0,NVEC-1), real(-1,1), real(0,1)) from one engine
#include <cstdint>
#include <cstdio>
#include <iostream>
#include <random>
#include <string>

int main()
{
    std::string cmd;
    while (std::cin >> cmd) {
        uint64_t seed; long n;
        std::cin >> seed >> n;
        std::mt19937_64 g(seed);
        if (cmd == "words") {
            for (long i = 0; i < n; ++i) { std::printf("%llu\n", static_cast<unsigned long long>(g())); }
        }
        else if (cmd == "real") {
            double a, b; std::cin >> a >> b;
            std::uniform_real_distribution<double> d(a, b);
            for (long i = 0; i < n; ++i) { std::printf("%a\n", d(g)); }
        }
        else if (cmd == "int") {
            int lo, hi; std::cin >> lo >> hi;
            std::uniform_int_distribution<int> d(lo, hi);
            for (long i = 0; i < n; ++i) { std::printf("%d\n", d(g)); }
        }
        else if (cmd == "mixed") {
            int nvec; std::cin >> nvec;
            std::uniform_int_distribution<int> di(0, nvec - 1);
            std::uniform_real_distribution<double> ds(-1., 1.), du(0., 1.);
            for (long i = 0; i < n; ++i) {
                const int v = di(g);
                const double s = ds(g);
                const double u = du(g);
                std::printf("%d %a %a\n", v, s, u);
            }
        }
        std::printf("end\n");
        std::fflush(stdout);
    }
    return 0;
}
